"""Writes tests/golden/edit_manifest_vectors.json from the compiled reference.

Run here (where /root/reference exists and `make -C oracle ref` has built
oracle/_ref/libref.so):  python tests/golden/make_edit_vectors.py

Each vector holds an input and the reference's answer:
  * manifests: edit-manifest JSON text -> serialize_edit_manifest(
    parse_edit_manifest(text)) (EditGraph.cpp:1405-1461), or the error class;
  * fingerprints: seeded Bayer / raster inputs -> fingerprint_raw_source /
    fingerprint_raster_source (SourceFingerprint.cpp:61-104);
  * legacy: GraphRecipe fields -> serialized snapshot_legacy_recipe
    (EditGraph.cpp:1463-1536).
The inputs are regenerated by tests/test_edit_manifest.py from the same
formulas (edit_vector_inputs below), so only their parameters are stored.
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests.golden.edit_vector_inputs import MANIFEST_CASES, RAW_FINGERPRINTS, RASTER_FINGERPRINTS, LEGACY_CASES, \
    raw_samples, raster_pixels  # noqa: E402


class Bayer(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("row_stride", ctypes.c_uint32),
                ("pattern", ctypes.c_uint32), ("phase_x", ctypes.c_uint32), ("phase_y", ctypes.c_uint32),
                ("ax", ctypes.c_uint32), ("ay", ctypes.c_uint32), ("aw", ctypes.c_uint32), ("ah", ctypes.c_uint32),
                ("black", ctypes.c_uint16 * 4), ("white", ctypes.c_uint16 * 4)]


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref.so"))
    lib.refh_last_error.restype = ctypes.c_char_p
    out = {"source": "compiled reference (oracle/_ref/libref.so)", "manifests": [], "fingerprints": [], "legacy": []}
    buf = ctypes.create_string_buffer(1 << 20)
    n = ctypes.c_size_t()
    for name, text in MANIFEST_CASES:
        rc = lib.refh_manifest_canonical(text.encode("utf-8") if isinstance(text, str) else text, buf,
                                         ctypes.sizeof(buf), ctypes.byref(n))
        exp = buf.value.decode("utf-8") if rc == 0 else {"error": rc}
        out["manifests"].append({"name": name, "expected": exp})
    digest = (ctypes.c_uint8 * 32)()
    for p in RAW_FINGERPRINTS:
        s = raw_samples(p)
        b = Bayer(p["width"], p["height"], p["stride"], p["pattern"], p["phase_x"], p["phase_y"], *p["area"],
                  (ctypes.c_uint16 * 4)(*p["black"]), (ctypes.c_uint16 * 4)(*p["white"]))
        rc = lib.refh_fingerprint_raw(ctypes.byref(b), s.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(s.size),
                                      digest)
        assert rc == 0, lib.refh_last_error()
        out["fingerprints"].append({"kind": "raw", "params": p, "sha256": bytes(digest).hex()})
    for p in RASTER_FINGERPRINTS:
        px = raster_pixels(p)
        rc = lib.refh_fingerprint_raster(ctypes.c_uint32(p["width"]), ctypes.c_uint32(p["height"]),
                                         ctypes.c_uint32(p["stride"]), ctypes.c_int(p["space"]),
                                         px.ctypes.data_as(ctypes.c_void_p), digest)
        assert rc == 0, lib.refh_last_error()
        out["fingerprints"].append({"kind": "raster", "params": p, "sha256": bytes(digest).hex()})
    for p in LEGACY_CASES:
        g = (ctypes.c_float * 3)(*p["gains"])
        cam = (ctypes.c_double * 9)(*p["cam"])
        d = (ctypes.c_uint8 * 32)(*bytes.fromhex(p["digest"]))
        rc = lib.refh_legacy_snapshot(p["raw"], d, p["source_id"].encode(), p["op_id"].encode(), p["era"], g,
                                      ctypes.c_float(p["stops"]), ctypes.c_float(p["shoulder"]),
                                      ctypes.c_float(p["gamma"]), p["mode"], p["has_camera"], cam, p["target"], buf,
                                      ctypes.sizeof(buf), ctypes.byref(n))
        exp = buf.value.decode("utf-8") if rc == 0 else {"error": rc}
        out["legacy"].append({"params": p, "expected": exp})
    with open(os.path.join(HERE, "edit_manifest_vectors.json"), "w") as f:
        json.dump(out, f, indent=1, ensure_ascii=True)
    print("wrote", len(out["manifests"]), "manifest,", len(out["fingerprints"]), "fingerprint and",
          len(out["legacy"]), "legacy vectors")


if __name__ == "__main__":
    main()
