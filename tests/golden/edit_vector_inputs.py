"""Inputs of the edit-manifest golden vectors (tests/golden/edit_manifest_vectors.json).

Shared by the generator (make_edit_vectors.py, run against the compiled
reference) and tests/test_edit_manifest.py (run against this library), so the
JSON file stores only answers.
"""
import json

import numpy as np

SRC = "5a5a5a5a-0000-4000-8000-000000000001"
RAW_SRC = "0a1b2c3d-0000-4000-8000-00000000beef"


def uid(i):
    return "00000000-0000-0000-0000-%012x" % i


def op(i, type_id, upstream, params=None, ind="scene_linear_prophoto_d50", outd="scene_linear_prophoto_d50",
       **extra):
    o = {"id": uid(i), "type": type_id, "schema_version": 1, "processing_version": 2, "enabled": True,
         "input_domain": ind, "output_domain": outd, "parameters": params or {}, "inputs": {"image": upstream},
         "masks": {}, "blend_mode": "normal", "opacity": 1.0}
    o.update(extra)
    return o


def raster_source(sid=SRC, space="linear_prophoto_d50", digest="a5" * 32):
    return {"id": sid, "kind": "scene_linear_raster_f32", "content_sha256": digest, "working_space": space}


def manifest(ops, output, sources=None, fmt=2, pv=2, space="linear_prophoto_d50", **extra):
    m = {"format_version": fmt, "processing_version": pv, "working_space": space,
         "sources": sources or [raster_source()], "operations": ops, "output": output}
    m.update(extra)
    return json.dumps(m, ensure_ascii=False)


ICC = {"profile_sha256": "3a" * 32, "intent": "relative_colorimetric", "black_point_compensation": False,
       "engine": "lcms2-core", "engine_version": "2.19.1"}

_raw_chain = [
    op(6, "rawengine.srgb_encode", uid(5), ind="display_linear_srgb", outd="display_encoded_srgb"),
    op(5, "rawengine.tone_curve", uid(4), {"shoulder": 0.35, "gamma": 1.1}, ind="scene_linear_srgb",
       outd="display_linear_srgb"),
    op(4, "rawengine.working_to_srgb", uid(3), outd="scene_linear_srgb"),
    op(3, "rawengine.camera_to_working", uid(2), {"matrix": [0.7, 0.2, 0.1, 0.25, 0.7, 0.05, 0.02, 0.1, 0.88]},
       ind="camera_linear"),
    op(2, "rawengine.exposure", uid(1), {"stops": 0.75}, ind="camera_linear", outd="camera_linear"),
    op(1, "rawengine.white_balance", RAW_SRC, {"red_gain": 2.0, "green_gain": 1, "blue_gain": 1.5},
       ind="camera_linear", outd="camera_linear"),
]
_raw_source = {"id": RAW_SRC, "kind": "decoded_bayer_u16", "content_sha256": "0f" * 32,
               "demosaic": {"algorithm": "rawengine.menon_base", "processing_version": 1}}

_numbers = {"a": 0.1, "b": -0.0, "c": 1e300, "d": 5e-324, "e": 123456789012345678, "f": 1.0, "g": -2.5e-07,
            "h": 100.0, "i": 1e21, "j": 0.30000000000000004, "k": 1e16, "l": 123456.789, "m": -1, "n": 0,
            "o": 2.2250738585072014e-308, "p": 1.7976931348623157e308, "q": 1e-5, "r": 9007199254740993}
_vendor = lambda params, **kw: op(9, "vendor.future.op", SRC, params, schema_version=7, **kw)

_guided = [op(1, "rawengine.guided_filter_working_y", SRC, {"radius": 8, "epsilon": 0.001}),
           op(2, "rawengine.exposure", uid(1), {"stops": 0.7}),
           op(3, "rawengine.curves_extended", uid(2),
              {"master": [[0, 0], [0.25, 0.3], [1, 1]], "red": [[0, 0], [1, 1]], "green": [[0, 0], [1, 1]],
               "blue": [[0, 0.1], [0.5, 0.45], [1, 1]], "interpolation": "shape_preserving_cubic"})]

MANIFEST_CASES = [
    ("icc_raster_vendor_op", manifest(
        [op(0x10, "rawengine.exposure", SRC, {"stops": 1.25}),
         op(0x20, "vendor.future.denoise", uid(0x10), {"radius": 3, "strategy": "edge-aware"},
            schema_version=7, opacity=0.75, outd="display_encoded_icc",
            future_payload={"labels": ["café", True], "nested": {"key": None}})],
        uid(0x20),
        sources=[{"id": SRC, "kind": "icc_raster_u16", "content_sha256": "a5" * 32,
                  "working_space": "linear_prophoto_d50",
                  "icc_input": dict(ICC, profile_sha256="1f" * 32)}],
        output_profile=ICC, future_manifest_field="kept")),
    ("raw_menon_chain_unordered", manifest(_raw_chain, uid(6), sources=[_raw_source], fmt=3)),
    ("raw_format2_bilinear", manifest(_raw_chain, uid(6), sources=[dict(_raw_source, demosaic=None)], fmt=2)
     .replace(', "demosaic": null', "")),
    ("numbers", manifest([_vendor(_numbers)], uid(9))),
    ("strings", manifest([_vendor({"s": "quote\" backslash\\ slash/ tab\t nl\n ctrl\u0001 del\u007f "
                                        "eé cjk中 emoji\U0001F600"})], uid(9))),
    ("escaped_input", '{"format_version":2,"processing_version":2,"working_space":"linear_prophoto_d50",'
                      '"sources":[{"id":"%s","kind":"scene_linear_raster_f32","content_sha256":"%s",'
                      '"working_space":"linear_prophoto_d50"}],"operations":[{"id":"%s","type":"vendor.x",'
                      '"schema_version":1,"processing_version":2,"enabled":false,"input_domain":'
                      '"scene_linear_prophoto_d50","output_domain":"scene_linear_prophoto_d50","parameters":'
                      '{"u":"\\u00e9\\ud83d\\ude00\\/\\b\\f\\r","E":1E+2,"z":-0.0e5},"inputs":{"image":"%s"},'
                      '"masks":{},"blend_mode":"normal","opacity":1}],"output":"%s"}'
     % (SRC, "a5" * 32, uid(9), SRC, uid(9))),
    ("whitespace", " \r\n\t" + manifest([op(1, "rawengine.exposure", SRC, {"stops": 1})], uid(1)).replace(
        ", ", " ,\n\t ") + "\n "),
    ("format1_raster_migrates", manifest([op(1, "rawengine.exposure", SRC, {"stops": -0.5})], uid(1), fmt=1,
                                         sources=[{"id": SRC, "kind": "scene_linear_raster_f32",
                                                   "content_sha256": "a5" * 32}])),
    ("guided_curves_chain", manifest(_guided, uid(3))),
    ("rec2020_legacy_process", manifest([op(1, "rawengine.exposure", SRC, {"stops": 1},
                                            ind="scene_linear_rec2020_d65", outd="scene_linear_rec2020_d65")],
                                        uid(1), pv=1, space="linear_rec2020_d65",
                                        sources=[raster_source(space="linear_rec2020_d65")]).replace(
        '"processing_version": 2, "enabled"', '"processing_version": 1, "enabled"')),
    ("source_is_output", manifest([], SRC)),
    ("mask_op_format4", manifest([op(1, "rawengine.mask_invert", SRC, {}, ind="coverage", outd="coverage",
                                     inputs={"mask": SRC})], uid(1), fmt=4,
                                 sources=[{"id": SRC, "kind": "coverage_raster_f32", "content_sha256": "a5" * 32}])),
    # rejections
    ("bad_duplicate_key", manifest([], SRC).replace('"output":', '"output": "x", "output":')),
    ("bad_trailing_comma", manifest([], SRC)[:-1] + ",}"),
    ("bad_leading_zero", manifest([_vendor({"v": 1})], uid(9)).replace('"v": 1', '"v": 01')),
    ("bad_nan", manifest([_vendor({"v": 1})], uid(9)).replace('"v": 1', '"v": NaN')),
    ("bad_infinite", manifest([_vendor({"v": 1})], uid(9)).replace('"v": 1', '"v": 1e400')),
    ("bad_int_overflow", manifest([_vendor({"v": 1})], uid(9)).replace('"v": 1', '"v": 9223372036854775808')),
    ("bad_unterminated", manifest([], SRC)[:-3]),
    ("bad_lone_surrogate", manifest([_vendor({"v": "x"})], uid(9)).replace('"v": "x"', '"v": "\\udc00"')),
    ("bad_control_char", manifest([_vendor({"v": "x"})], uid(9)).replace('"v": "x"', '"v": "a\u0002b"')),
    ("bad_depth", manifest([_vendor({"v": 0})], uid(9)).replace('"v": 0', '"v": ' + "[" * 70 + "]" * 70)),
    ("bad_unknown_source_field", manifest([], SRC, sources=[dict(raster_source(), extra=1)])),
    ("bad_cycle", manifest([op(1, "vendor.a", uid(2)), op(2, "vendor.b", uid(1))], uid(2))),
    ("bad_missing_output", manifest([], uid(5))),
    ("bad_uppercase_uuid", manifest([], SRC.upper().replace("0", "A"), sources=[raster_source(
        SRC.upper().replace("0", "A"))])),
    ("bad_opacity", manifest([op(1, "vendor.a", SRC, opacity=1.5)], uid(1))),
    ("bad_exposure", manifest([op(1, "rawengine.exposure", SRC, {"stops": 33})], uid(1))),
    ("bad_wb_gain", manifest([op(1, "rawengine.white_balance", SRC, {"red_gain": 0, "green_gain": 1,
                                                                      "blue_gain": 1})], uid(1))),
    ("bad_guided_radius", manifest([op(1, "rawengine.guided_filter_working_y", SRC,
                                       {"radius": 9, "epsilon": 0.001})], uid(1))),
    ("bad_guided_fraction", manifest([op(1, "rawengine.guided_filter_working_y", SRC,
                                         {"radius": 2.5, "epsilon": 0.001})], uid(1))),
    ("bad_curve_order", manifest([op(3, "rawengine.curves_extended", SRC,
                                     {"master": [[0, 0], [0.5, 0.5], [0.4, 1]], "red": [[0, 0], [1, 1]],
                                      "green": [[0, 0], [1, 1]], "blue": [[0, 0], [1, 1]],
                                      "interpolation": "linear"})], uid(3))),
    ("bad_format3_without_demosaic", manifest(_raw_chain, uid(6), sources=[dict(_raw_source, demosaic=None)],
                                              fmt=3).replace(', "demosaic": null', "")),
    ("bad_format2_with_demosaic", manifest(_raw_chain, uid(6), sources=[_raw_source], fmt=2)),
    ("bad_icc_without_profile", manifest([op(1, "rawengine.icc_display", SRC, ind="display_linear_srgb",
                                             outd="display_encoded_icc")], uid(1))),
    ("bad_format6", manifest([], SRC, fmt=6)),
    ("bad_mask_op_format2", manifest([op(1, "rawengine.mask_invert", SRC, {}, ind="coverage", outd="coverage")],
                                     uid(1))),
    ("bad_not_utf8", manifest([_vendor({"v": "x"})], uid(9)).replace('"x"', '"\udcff"').encode(
        "utf-8", "surrogateescape")),
]

RAW_FINGERPRINTS = [
    dict(width=13, height=9, stride=13, pattern=0, phase_x=0, phase_y=0, area=[0, 0, 13, 9], black=[0] * 4,
         white=[65535] * 4, seed=1),
    dict(width=150, height=98, stride=160, pattern=1, phase_x=1, phase_y=0, area=[3, 2, 141, 90],
         black=[510, 512, 511, 513], white=[16383] * 4, seed=2),
    dict(width=4100, height=3, stride=4104, pattern=2, phase_x=0, phase_y=1, area=[1, 0, 4099, 3],
         black=[64] * 4, white=[4095] * 4, seed=3),
]
RASTER_FINGERPRINTS = [
    dict(width=8, height=6, stride=8, space=0, seed=4),
    dict(width=37, height=5, stride=41, space=1, seed=5),
    dict(width=1500, height=2, stride=1500, space=0, seed=6),
]

LEGACY_CASES = [
    dict(raw=1, digest="0f" * 32, source_id=RAW_SRC, op_id=uid(8), era=0, gains=[1.15, 0.95, 1.3], stops=0.75,
         shoulder=0.35, gamma=1.1, mode=1, has_camera=1, cam=[1, 0, 0, 0, 1, 0, 0, 0, 1], target=1),
    dict(raw=1, digest="0f" * 32, source_id=RAW_SRC, op_id=uid(8), era=1, gains=[2.0, 1.0, 1.5], stops=-1.25,
         shoulder=0.25, gamma=1.0, mode=0, has_camera=0, cam=[0] * 9, target=0),
    dict(raw=0, digest="a5" * 32, source_id=SRC, op_id=uid(7), era=1, gains=[1, 1, 1], stops=0.3, shoulder=0.42,
         gamma=0.9, mode=1, has_camera=0, cam=[0] * 9, target=0),
    dict(raw=1, digest="0f" * 32, source_id=RAW_SRC, op_id=uid(8), era=1, gains=[1, 1, 1], stops=0, shoulder=0.25,
         gamma=1, mode=1, has_camera=0, cam=[0] * 9, target=0),  # rejected: RAW preview without calibration
]


def raw_samples(p):
    n = p["stride"] * p["height"]
    i = np.arange(n, dtype=np.uint64)
    return ((i * np.uint64(2654435761) + np.uint64(p["seed"] * 7919)) % np.uint64(65536)).astype(np.uint16)


def raster_pixels(p):
    n = p["stride"] * p["height"] * 3
    rng = np.random.default_rng(p["seed"])
    v = rng.standard_normal(n).astype(np.float32)
    v[::7] = -0.0
    return v
