"""rawengine_native: the reference's Python entry points (RawEnginePython.cpp)
over the GPU drop-in layer, checked against the CPU oracle.

One-shot render / render_raster with the reference's option keys, the DLPack
device result, RawSession / RasterSession (recipe renders, exported and saved
manifests, tile cache counters, histogram, required source regions, jobs with
priorities, supersession and cancellation, close), error classes.
"""
import json

import numpy as np
import pytest

from oracle import oracle as O

from librawops_amd import rawengine_native as rn

CAM = (0.7, 0.2, 0.1, 0.25, 0.7, 0.05, 0.02, 0.1, 0.88)


def bayer_case(seed=1, w=300, h=200, stride=304, pattern=1, px=1, py=0, area=(5, 3, 288, 190)):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, 16384, (h, stride)).astype(np.uint16)
    opts = dict(pattern=pattern, row_stride_samples=stride, cfa_phase_x=px, cfa_phase_y=py, active_x=area[0],
                active_y=area[1], active_width=area[2], active_height=area[3], black_levels=[510, 512, 511, 513],
                white_level=16383)
    meta = O.Bayer(w, h, stride, pattern, px, py, area, (510, 512, 511, 513), (16383,) * 4)
    return s, opts, meta


def as_image(result):
    w, h, data = result
    return np.frombuffer(data, np.float32).reshape(h, w, 3)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["legacy", "srgb-preview"])
def test_one_shot_render_matches_oracle(mode):
    s, opts, meta = bayer_case()
    opts.update(red_gain=2.0, blue_gain=1.5, exposure_stops=0.25, camera_to_xyz_d50=list(CAM),
                working_space="rec2020-d65", output_mode=mode, tone_shoulder=0.3, tile_size=64)
    got = as_image(rn.render(s.tobytes(), 300, 200, opts))
    exp = O.render_raw(meta, s, O.Chain(wb=(2.0, 1.0, 1.5), exposure_stops=0.25, camera_to_xyz=CAM,
                                        working_space=1, output="bounded" if mode == "legacy" else "srgb",
                                        tone_shoulder=0.3), meta.area)
    assert got.shape == exp.shape
    if mode == "legacy":
        assert got.tobytes() == exp.tobytes()
    else:
        assert np.abs(got - exp).max() <= 1e-5   # powf in the sRGB transfer


@pytest.mark.gpu
def test_one_shot_preview_mip_and_roi():
    s, opts, meta = bayer_case(seed=2)
    opts.update(camera_to_xyz_d50=list(CAM), output_mode="srgb-preview", mip=1, quality="preview", x=10, y=7,
                roi_width=100, roi_height=60)
    got = as_image(rn.render(s, 300, 200, opts))
    exp = O.render_raw(meta, s, O.Chain(camera_to_xyz=CAM, output="srgb"), (10, 7, 100, 60), mip=1)
    assert np.abs(got - exp).max() <= 1e-5


@pytest.mark.gpu
def test_render_dlpack_is_the_device_result():
    torch = pytest.importorskip("torch")
    s, opts, meta = bayer_case(seed=3)
    opts.update(camera_to_xyz_d50=list(CAM))
    w, h, cap = rn.render_dlpack(s, 300, 200, opts)
    t = torch.from_dlpack(cap)
    assert t.is_cuda and t.dtype == torch.float32 and tuple(t.shape) == (h, w, 3)
    ref = as_image(rn.render(s, 300, 200, opts))
    assert t.cpu().numpy().tobytes() == ref.tobytes()


@pytest.mark.gpu
def test_render_raster_matches_oracle():
    rng = np.random.default_rng(4)
    px = rng.lognormal(-2.0, 1.0, (90, 120, 3)).astype(np.float32)
    got = as_image(rn.render_raster(px, 120, 90, dict(working_space="prophoto-d50", exposure_stops=-0.5,
                                                      tone_shoulder=0.42, tone_gamma=0.9)))
    exp = px.copy()
    O.scale(exp, O.exp2f(-0.5))
    O.tone_curve(exp, 0.42, 0.9)
    O.clip(exp)
    assert np.abs(got - exp).max() <= 1e-5


@pytest.mark.gpu
def test_raw_session_recipe_manifest_and_cache():
    s, opts, meta = bayer_case(seed=5)
    md = {k: v for k, v in opts.items()}
    session = rn.RawSession(s, 300, 200, md, cache_bytes=64 << 20, workers=2,
                            demosaic={"algorithm": "rawengine.menon_base", "processing_version": 1})
    info = session.source_info()
    assert info["kind"] == "decoded_bayer_u16" and info["demosaic"]["algorithm"] == "rawengine.menon_base"
    assert info["active_area"] == (5, 3, 288, 190) and len(info["content_sha256"]) == 64
    recipe = dict(red_gain=2.0, blue_gain=1.5, camera_to_xyz_d50=list(CAM), x=20, y=15, roi_width=200,
                  roi_height=120)
    got = as_image(session.render(recipe))
    exp = O.render_raw(meta, s, O.Chain(demosaic="menon", wb=(2.0, 1.0, 1.5), camera_to_xyz=CAM, output="bounded"),
                       (20, 15, 200, 120))
    assert got.tobytes() == exp.tobytes()
    manifest = session.export_manifest(recipe)
    doc = json.loads(manifest)
    assert doc["format_version"] == 3 and doc["sources"][0]["demosaic"]["algorithm"] == "rawengine.menon_base"
    roi = dict(x=20, y=15, roi_width=200, roi_height=120)
    assert as_image(session.render_manifest(manifest, roi)).tobytes() == exp.tobytes()
    before = session.cache_stats()
    again = as_image(session.render_manifest(manifest, roi))
    after = session.cache_stats()
    assert again.tobytes() == exp.tobytes() and after["hits"] > before["hits"] and after["budget_bytes"] == 64 << 20
    session.clear_cache()
    assert session.cache_stats()["entries"] == 0
    regions = session.required_source_regions(manifest, dict(x=20, y=15, roi_width=10, roi_height=10))
    assert regions == {"00000000-0000-0000-0000-000000000001": (14, 9, 22, 22)}   # Menon halo 6
    hist = session.histogram_manifest(manifest, roi, bins=32)
    assert hist["pixel_count"] == 200 * 120
    for c in hist["channels"]:
        assert sum(c["counts"]) + c["underflow"] + c["overflow"] + c["nonfinite"] == 200 * 120
    stats = []
    done = session.analyze_local_manifest(manifest, stats.append, dict(roi, tile_size=64), radius=2)
    assert done == {"completed_tiles": 8, "pixel_count": 200 * 120}
    first = stats[0]
    assert first["bounds"] == (20, 15, 64, 64) and first["radius"] == 2
    mean = np.frombuffer(first["mean_f64"], np.float64).reshape(64, 64, 3)
    lin = as_image(session.render_manifest(manifest, dict(x=18, y=13, roi_width=68, roi_height=68)))
    exp_mean, _, _ = O.local_statistics(lin, (18, 13, 68, 68), (20, 15, 64, 64), (5, 3, 288, 190), 2)
    assert mean.tobytes() == exp_mean.tobytes()
    with pytest.raises(ValueError):
        session.render_manifest(manifest, dict(exposure_stops=1.0))   # edits belong in the manifest
    with pytest.raises(ValueError):
        session.render(dict(pattern=2))                               # sensor metadata is fixed
    session.close()
    with pytest.raises(RuntimeError):
        session.render(recipe)


@pytest.mark.gpu
def test_raw_session_jobs():
    s, opts, meta = bayer_case(seed=6)
    session = rn.RawSession(s, 300, 200, opts, workers=2, max_pending=16)
    recipe = dict(camera_to_xyz_d50=list(CAM), output_mode="srgb-preview", tile_size=64)
    direct = as_image(session.render(recipe))
    job = session.submit(recipe, priority="background")
    assert job.output_kind() == "rgb"
    res = as_image(job.result(timeout=60))
    assert res.tobytes() == direct.tobytes() and job.done()
    prog = job.progress()
    assert prog["completed_tiles"] == prog["total_tiles"] == 5 * 3   # 288x190 active area, 64-pixel tiles
    # A group's newest request always completes with its own viewport.
    # native viewports are sensor-relative: the active area starts at (5, 3)
    jobs = [session.submit_latest("view", dict(recipe, x=5 + 8 * i, y=3 + 4 * i, roi_width=100, roi_height=80))
            for i in range(6)]
    last = as_image(jobs[-1].result(timeout=60))
    assert last.tobytes() == as_image(session.render(dict(recipe, x=45, y=23, roi_width=100, roi_height=80))).tobytes()
    for j in jobs[:-1]:
        try:
            j.result(timeout=60)
        except rn.RenderCancelled:
            pass
    cancelled = session.submit(recipe)
    cancelled.cancel()
    try:
        cancelled.result(timeout=60)
    except rn.RenderCancelled:
        pass
    with pytest.raises(ValueError):
        session.submit(recipe, priority="urgent")
    with pytest.raises(ValueError):
        session.submit_latest("", recipe)
    session.close()
    with pytest.raises(RuntimeError):
        session.submit(recipe)


@pytest.mark.gpu
def test_raster_session_matches_render_raster():
    rng = np.random.default_rng(7)
    px = rng.lognormal(-2.0, 1.0, (64, 80, 3)).astype(np.float32)
    session = rn.RasterSession(px, 80, 64, "rec2020-d65")
    opts = dict(exposure_stops=0.5, output_mode="srgb-preview")
    got = as_image(session.render(opts))
    ref = as_image(rn.render_raster(px, 80, 64, dict(opts, working_space="rec2020-d65")))
    assert got.tobytes() == ref.tobytes()
    assert session.source_info()["working_space"] == "linear_rec2020_d65"
    with pytest.raises(ValueError):
        session.render(dict(working_space="prophoto-d50"))


def test_module_surface_and_errors_without_rendering():
    assert rn.icc_available() is False
    assert issubclass(rn.RenderCancelled, RuntimeError)
    with pytest.raises(ValueError):
        rn.render(np.zeros(10, np.uint16), 5, 3, {"mip": 3})
    with pytest.raises(OverflowError):
        rn.render(np.zeros(15, np.uint16), 5, 3, {"x": 2 ** 40})
    with pytest.raises(ValueError):
        rn.render(np.zeros(10, np.uint16), 5, 3)            # 10 samples for a 5x3 sensor
    with pytest.raises(ValueError):
        rn.render(np.zeros(15, np.uint16), 5, 3, {"mip": 1, "quality": "preview"})   # no calibration
    with pytest.raises(ValueError):
        rn.render_raster(np.zeros(45, np.float32), 5, 3, {"working_space": "prophoto-d50", "crop": (0, 0, 2, 2)})
    with pytest.raises(ValueError):
        rn.RawSession(np.zeros(15, np.uint16), 5, 3, {"red_gain": 2.0})
    with pytest.raises(ValueError):
        rn.RawSession(np.zeros(15, np.uint16), 5, 3, demosaic={"algorithm": "rawengine.ahd", "processing_version": 1})
