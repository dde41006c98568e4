"""Timing fidelity of the oracle's Menon CPU restatement against the compiled
reference (test infrastructure; run where /root/reference and oracle/_ref
exist):

    python -m oracle.fidelity [threads] [rows]

Both render the same band of the 24 MP config-2 frame (6048x4024 BGGR,
phase (1, 0), stride 6080, active 6016x3992, per-site black) through Menon ->
camera->ProPhoto (the config-2 chain), split into `threads` row bands
run concurrently: the reference as concurrent band requests on one const node
chain (its fastest legal host configuration, one OpenMP thread per request),
the restatement as OpenMP row bands (bench.py's cpu_baseline leg).  Prints
both rates and their ratio; the restatement stands in for the reference on
machines without it when the ratio is within 0.8..1.2."""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from librawops_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle import ref as R  # noqa: E402

def port_seconds(om, samples, cam, rect, threads, reps):
    m = O.camera_to_working_f32(cam, 0)

    def once():
        nat = O.menon_bands(om, samples, rect, threads)
        O.mat3(nat, m)
    once()
    t0 = time.perf_counter()
    for _ in range(reps):
        once()
    return (time.perf_counter() - t0) / reps


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    rows = int(sys.argv[2]) if len(sys.argv) > 2 else 512
    if not R.available():
        raise SystemExit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    meta, samples, cam_w = synth.config2_menon(seed=2)
    cam = synth.camera_to_xyz_for(cam_w)
    om = O.Bayer(meta["width"], meta["height"], meta["row_stride_samples"], meta["pattern"], meta["cfa_phase_x"],
                 meta["cfa_phase_y"], meta["active_area"], meta["black_levels"], meta["white_levels"])
    ax, ay, aw, ah = meta["active_area"]
    rect = (ax, ay, aw, min(rows, ah))
    px = rect[2] * rect[3]
    reps = 5
    t_ref = R.time_menon_bands(om, samples, None, 0.0, cam, rect, threads, reps)
    t_port = port_seconds(om, samples, cam, rect, threads, reps)
    out = {"band": f"{rect[2]}x{rect[3]}", "threads": threads, "reference_mpix_s": round(px / t_ref / 1e6, 2),
           "port_mpix_s": round(px / t_port / 1e6, 2), "port_over_reference": round(t_ref / t_port, 3)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
