"""Edit manifests, source fingerprints and cache signatures (host code, no GPU).

This library's manifest code (librawengine_amd.so, driven through
tests/cpp/manifest_tool) against:
  * golden vectors written by the compiled reference
    (tests/golden/edit_manifest_vectors.json, tests/golden/make_edit_vectors.py):
    canonical serialisation of parse(text) or its rejection
    (EditGraph.cpp:175-1461), fingerprint_raw_source / fingerprint_raster_source
    (SourceFingerprint.cpp:61-104), snapshot_legacy_recipe (EditGraph.cpp:1463-1536);
  * FIPS 180-4 / hashlib for SHA-256 and the tile-cache signature rule
    (EditGraph.cpp:1736-1753): SHA-256("librawops.tile-cache.v2\\0" + upstream
    signature + the operation's canonical JSON), where the operation's JSON is
    cut out of the reference's own canonical manifest text.
"""
import hashlib
import json
import os
import subprocess
import tempfile

import pytest

from tests.golden.edit_vector_inputs import LEGACY_CASES, MANIFEST_CASES, RASTER_FINGERPRINTS, RAW_FINGERPRINTS, \
    raster_pixels, raw_samples

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "tests", "cpp", "manifest_tool")
VECTORS = json.load(open(os.path.join(ROOT, "tests", "golden", "edit_manifest_vectors.json")))

pytestmark = pytest.mark.skipif(not os.path.exists(TOOL), reason="tests/cpp/manifest_tool is not built")


def tool(*args):
    p = subprocess.run([TOOL, *map(str, args)], capture_output=True, timeout=60)
    assert p.returncode == 0, p.stderr
    return p.stdout.decode("utf-8")


def tool_on_text(cmd, text):
    with tempfile.NamedTemporaryFile(suffix=".json", delete=False) as f:
        f.write(text if isinstance(text, bytes) else text.encode("utf-8"))
    try:
        return tool(cmd, f.name)
    finally:
        os.unlink(f.name)


def expected_of(v):
    e = v["expected"]
    return ("ERROR", e["error"]) if isinstance(e, dict) else e


@pytest.mark.parametrize("case", range(len(MANIFEST_CASES)), ids=[c[0] for c in MANIFEST_CASES])
def test_manifest_canonical_form_matches_reference(case):
    name, text = MANIFEST_CASES[case]
    vec = VECTORS["manifests"][case]
    assert vec["name"] == name
    got = tool_on_text("canon", text)
    exp = expected_of(vec)
    if isinstance(exp, tuple):
        assert got.startswith("ERROR %d " % exp[1]), got
    else:
        assert got == exp
        # The canonical form is a fixed point.
        assert tool_on_text("canon", got) == got


@pytest.mark.parametrize("k", range(len(RAW_FINGERPRINTS)))
def test_raw_fingerprint_matches_reference(k, tmp_path):
    p = RAW_FINGERPRINTS[k]
    vec = [v for v in VECTORS["fingerprints"] if v["kind"] == "raw"][k]
    path = tmp_path / "raw.bin"
    raw_samples(p).tofile(path)
    got = tool("fingerprint-raw", path, p["width"], p["height"], p["stride"], p["pattern"], p["phase_x"],
               p["phase_y"], *p["area"], *p["black"], *p["white"]).split()
    assert got == [vec["sha256"], vec["sha256"]]


def test_raw_fingerprint_ignores_padding_and_sees_visible_samples(tmp_path):
    p = RAW_FINGERPRINTS[1]
    s = raw_samples(p).reshape(p["height"], p["stride"])
    ax, ay, aw, ah = p["area"]
    args = [p["width"], p["height"], p["stride"], p["pattern"], p["phase_x"], p["phase_y"], *p["area"],
            *p["black"], *p["white"]]

    def digest(a):
        a.tofile(tmp_path / "x.bin")
        return tool("fingerprint-raw", tmp_path / "x.bin", *args).split()[0]

    base = digest(s)
    pad = s.copy()
    pad[:, p["width"]:] ^= 0x5555      # row padding
    pad[:ay, :] ^= 1                   # inactive rows above the area
    assert digest(pad) == base
    vis = s.copy()
    vis[ay + ah - 1, ax + aw - 1] ^= 1  # last visible sample
    assert digest(vis) != base


@pytest.mark.parametrize("k", range(len(RASTER_FINGERPRINTS)))
def test_raster_fingerprint_matches_reference(k, tmp_path):
    p = RASTER_FINGERPRINTS[k]
    vec = [v for v in VECTORS["fingerprints"] if v["kind"] == "raster"][k]
    path = tmp_path / "px.bin"
    raster_pixels(p).tofile(path)
    assert tool("fingerprint-raster", path, p["width"], p["height"], p["stride"], p["space"]) == vec["sha256"]


@pytest.mark.parametrize("k", range(len(LEGACY_CASES)))
def test_legacy_snapshot_matches_reference(k):
    p = LEGACY_CASES[k]
    got = tool("legacy", p["raw"], p["digest"], p["source_id"], p["op_id"], p["era"], *p["gains"], p["stops"],
               p["shoulder"], p["gamma"], p["mode"], p["has_camera"], *p["cam"], p["target"])
    exp = expected_of(VECTORS["legacy"][k])
    if isinstance(exp, tuple):
        assert got.startswith("ERROR %d " % exp[1]), got
    else:
        assert got == exp


def _objects_in_array(text, key):
    """Substrings of the JSON objects in the array under `key` (canonical text)."""
    i = text.index('"%s":[' % key) + len(key) + 4
    out, depth, begin = [], 0, None
    while True:
        ch = text[i]
        if ch == '"':
            i += 1
            while text[i] != '"':
                i += 2 if text[i] == "\\" else 1
        elif ch == "{":
            if depth == 0:
                begin = i
            depth += 1
        elif ch == "}":
            depth -= 1
            if depth == 0:
                out.append(text[begin:i + 1])
        elif ch == "]" and depth == 0:
            return out
        i += 1


@pytest.mark.parametrize("name", ["raw_menon_chain_unordered", "guided_curves_chain", "icc_raster_vendor_op",
                                  "escaped_input"])
def test_cache_signatures_follow_reference_rule(name):
    k = [c[0] for c in MANIFEST_CASES].index(name)
    canon = VECTORS["manifests"][k]["expected"]
    ops = {json.loads(o)["id"]: o for o in _objects_in_array(canon, "operations")}
    srcs = {json.loads(s)["id"]: s for s in _objects_in_array(canon, "sources")}
    tag = b"librawops.tile-cache.v2\x00"
    lines = tool_on_text("signatures", canon).strip().splitlines()
    assert lines
    for line in lines:
        parts = line.split()
        if parts[1] == "-":
            src = json.loads(srcs[parts[0]])
            if src["kind"] == "decoded_bayer_u16" and "demosaic" not in src:
                continue  # canonicalised with the implied bilinear policy; covered by the Bayer case below
            assert parts[2] == hashlib.sha256(tag + srcs[parts[0]].encode("utf-8")).hexdigest()
        else:
            op = json.loads(ops[parts[0]])
            if not op["enabled"]:
                assert parts[2] == parts[1]
                continue
            exp = hashlib.sha256(tag + bytes.fromhex(parts[1]) + ops[parts[0]].encode("utf-8")).hexdigest()
            assert parts[2] == exp, parts[0]


def test_signature_of_implied_bilinear_policy_hashes_canonical_record():
    k = [c[0] for c in MANIFEST_CASES].index("raw_format2_bilinear")
    canon = VECTORS["manifests"][k]["expected"]
    lines = tool_on_text("signatures", canon).strip().splitlines()
    src = [l.split() for l in lines if l.split()[1] == "-"][0]
    rec = json.loads(_objects_in_array(canon, "sources")[0])
    rec["demosaic"] = {"algorithm": "rawengine.bilinear", "processing_version": 1}
    text = json.dumps(rec, sort_keys=True, separators=(",", ":"))
    assert src[2] == hashlib.sha256(b"librawops.tile-cache.v2\x00" + text.encode()).hexdigest()


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref.so")),
                    reason="compiled reference not built here")
def test_golden_vectors_are_current():
    """The committed vectors equal what the compiled reference answers now."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref.so"))
    buf = ctypes.create_string_buffer(1 << 20)
    n = ctypes.c_size_t()
    for (name, text), vec in zip(MANIFEST_CASES, VECTORS["manifests"]):
        rc = lib.refh_manifest_canonical(text.encode("utf-8") if isinstance(text, str) else text, buf,
                                         ctypes.sizeof(buf), ctypes.byref(n))
        exp = expected_of(vec)
        if isinstance(exp, tuple):
            assert rc == exp[1], name
        else:
            assert rc == 0 and buf.value.decode("utf-8") == exp, name
