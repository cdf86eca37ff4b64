"""Autotuner of the SES path's launch shapes on the current GPU.

Measures, on seeded synthetic molecules, and writes the choices to tuning.json beside the
library (which the library reads once per process for the values it uses itself):
- pass1_slab_rows, per voxel size: rows of one x plane per field-pass-1 block (8 to 256), by the
  pass-1 stage time of the workload that uses that voxel size (one launch sequence over it);
- batch_voxels, per voxel size: grid points per batched launch sequence, by the meshes/s of a
  stream of molecules through qs_compute_batch_device;
- serve: worker engines, the grid-point cap of one served batch and the express worker's request
  size cap (0: none), by the p99 latency of a served stream at 0.8 of the plain path's capacity.

usage: python -m quickses_amd.autotune [--out PATH] [--quick]
"""
import argparse
import json
import os
import sys
import time

import numpy as np


def _batch_stage_ms(qs, hipmem, items, reso, stage, reps):
    """Mean time of one stage of a batched launch sequence over `items` at voxel size `reso`."""
    counts = [len(r) for _, r in items]
    dp = hipmem.DeviceArray(np.concatenate([p for p, _ in items]).astype(np.float32))
    dr = hipmem.DeviceArray(np.concatenate([r for _, r in items]).astype(np.float32))
    try:
        qs.compute_batch_device(dp.ptr, dr.ptr, counts, reso, 1)
        qs.set_option("timing", 1)
        acc = 0.0
        for _ in range(reps):
            qs.compute_batch_device(dp.ptr, dr.ptr, counts, reso, 1)
            acc += qs.stage_times()[stage]
        qs.set_option("timing", 0)
    finally:
        dp.free()
        dr.free()
    return acc / reps


def _workload_at(qs, reso, quick):
    """The molecules a voxel size is used for: the 50,000-atom molecule at 0.2 A; the batched
    workload's molecules and frames at 0.5 A; the served stream's requests of that size at 0.3
    and 1.0 A."""
    from .workloads import config4_sets, config5_requests

    if reso == 0.2:
        return [qs.synth_ball(2, 50000)]
    if reso == 0.5:
        mols, frames = config4_sets(qs, 32 if quick else 96, 8 if quick else 24)
        return mols + frames
    reqs = [r for r, v in config5_requests(qs, 48) if v == reso]
    return reqs[: (6 if quick else 16)]


def tune_pass1_rows(qs, hipmem, quick=False):
    """Rows per pass-1 block for each voxel size, by the pass-1 time of that size's workload."""
    out, meas = {}, {}
    for reso in (0.2, 0.3, 0.5, 1.0):
        items = _workload_at(qs, reso, quick)
        times = {}
        for rows in (8, 16, 32, 64, 128, 256):
            qs.set_tuning("pass1_slab_rows", reso, rows)
            times[rows] = _batch_stage_ms(qs, hipmem, items, reso, "pass1", 3 if quick else 6)
        best = min(times, key=times.get)
        qs.set_tuning("pass1_slab_rows", reso, best)
        key = "%g" % reso
        out[key] = best
        meas[key] = {"molecules": len(items), "atoms": int(sum(len(r) for _, r in items)),
                     "pass1_ms": {str(k): round(v, 4) for k, v in times.items()}}
        print("pass1_slab_rows %.1f A (%d molecules): %s -> %d" % (reso, len(items), meas[key]["pass1_ms"], best),
              flush=True)
    return out, meas


def tune_batch_voxels(qs, hipmem, quick=False):
    """Grid points per batched launch sequence at 0.5 A, on molecules and frames of the batched
    workload."""
    from .workloads import chunks_by_voxels, config4_sets

    mols, frames = config4_sets(qs, 64 if quick else 128, 64 if quick else 160)
    items = mols + frames
    counts = [len(r) for _, r in items]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    d_pos = hipmem.DeviceArray(np.concatenate([p for p, _ in items]).astype(np.float32))
    d_rad = hipmem.DeviceArray(np.concatenate([r for _, r in items]).astype(np.float32))
    rates = {}
    try:
        for budget in (250_000_000, 500_000_000, 1_000_000_000, 1_500_000_000, 3_000_000_000):
            bounds, k = [], 0
            for ch in chunks_by_voxels(items, 0.5, budget):
                bounds.append((k, k + len(ch)))
                k += len(ch)
            for a, b in bounds:  # the library's buffers grow to this chunking first
                qs.compute_batch_device(d_pos.ptr + 12 * int(off[a]), d_rad.ptr + 4 * int(off[a]), counts[a:b], 0.5, 1)
            hipmem.synchronize()
            t = time.perf_counter()
            for a, b in bounds:
                qs.compute_batch_device(d_pos.ptr + 12 * int(off[a]), d_rad.ptr + 4 * int(off[a]), counts[a:b], 0.5, 1)
            hipmem.synchronize()
            rates[budget] = len(items) / (time.perf_counter() - t)
            print("batch_voxels %d: %.0f meshes/s (%d launch sequences)" % (budget, rates[budget], len(bounds)), flush=True)
    finally:
        d_pos.free()
        d_rad.free()
    best = max(rates, key=rates.get)
    return {"0.5": best}, {"0.5": {"meshes": len(items), "meshes_per_s": {str(k): round(v, 1) for k, v in rates.items()}}}


def tune_serve(qs, hipmem, quick=False):
    """Worker engines, per-batch grid-point cap and express cap of the server by p99 latency at
    0.8 load."""
    import threading

    from .workloads import config5_requests

    k = 80 if quick else 240
    reqs = config5_requests(qs, k)
    nmax = max(len(r[0][1]) for r in reqs)
    d_pos, d_rad = hipmem.alloc(12 * nmax), hipmem.alloc(4 * nmax)
    try:
        for (pos, rad), reso in reqs[:2]:
            hipmem.upload(d_pos, pos)
            hipmem.upload(d_rad, rad)
            qs.compute_device(d_pos, d_rad, len(rad), reso, 1)
        t = time.perf_counter()
        for (pos, rad), reso in reqs:
            hipmem.upload(d_pos, pos)
            hipmem.upload(d_rad, rad)
            qs.compute_device(d_pos, d_rad, len(rad), reso, 1)
        cap_rate = k / (time.perf_counter() - t)
    finally:
        hipmem.free(d_pos)
        hipmem.free(d_rad)
    lam = 0.8 * cap_rate
    res = {}
    settings = [(w, cap, 0) for w in (1, 2, 3, 4) for cap in (0, 75_000_000, 150_000_000)]
    settings += [(w, cap, 16_000_000) for w in (2, 3, 4) for cap in (0, 75_000_000, 150_000_000)]
    for workers, cap, express in settings:
            p99s, p50s = [], []
            # two arrival streams per setting: a p99 over a short stream is one or two samples
            for rep in range(2):
                arrivals = np.cumsum(np.random.default_rng(5 + rep).exponential(1.0 / lam, size=k))
                lat = []
                with qs.Server(hbm_budget_bytes=48 << 30, result_bytes=32 << 30, workers=workers,
                               max_batch_voxels=cap, express_voxels=express) as srv:
                    w = srv.submit(reqs[0][0][0][:2000], reqs[0][0][1][:2000], 1.0, 1)
                    srv.poll(w, -1)
                    srv.release(w)
                    ids = [None] * k

                    def poller():
                        for i in range(k):
                            while ids[i] is None:
                                time.sleep(0.0002)
                            lat.append(srv.poll(ids[i], -1)["latency_ms"])
                            srv.release(ids[i])

                    th = threading.Thread(target=poller)
                    th.start()
                    start = time.perf_counter()
                    for i, (((pos, rad), reso), ta) in enumerate(zip(reqs, arrivals)):
                        while time.perf_counter() - start < ta:
                            time.sleep(0.0002)
                        ids[i] = srv.submit(pos, rad, reso, 1)
                    th.join()
                p99s.append(float(np.percentile(lat, 99)))
                p50s.append(float(np.percentile(lat, 50)))
            key = (workers, cap, express)
            res[key] = (float(np.mean(p99s)), float(np.mean(p50s)))
            print("serve workers=%d cap=%d express=%d: p50 %.3f ms p99 %.3f ms (mean of 2 streams)"
                  % (workers, cap, express, res[key][1], res[key][0]), flush=True)
    best = min(res, key=lambda key: res[key][0])
    return ({"workers": best[0], "max_batch_voxels": best[1], "express_voxels": best[2]},
            {"arrival_rate_per_s": round(lam, 1), "requests": k, "streams": 2,
             "p50_p99_ms (workers/cap/express)": {"%d/%d/%d" % key: [round(v[1], 3), round(v[0], 3)]
                                                   for key, v in res.items()}})


def main():
    ap = argparse.ArgumentParser()
    here = os.path.dirname(os.path.abspath(__file__))
    ap.add_argument("--out", default=os.path.join(here, "tuning.json"))
    ap.add_argument("--quick", action="store_true", help="fewer repetitions and smaller streams")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(here))
    import quickses_amd as qs
    from quickses_amd import hipmem

    if hipmem.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    qs.load_tuning(None)  # measure from the built-in choices, not an earlier cache
    t0 = time.perf_counter()
    rows, rows_m = tune_pass1_rows(qs, hipmem, args.quick)
    bv, bv_m = tune_batch_voxels(qs, hipmem, args.quick)
    serve, serve_m = tune_serve(qs, hipmem, args.quick)
    out = {"device": "MI355X (gfx950)", "pass1_slab_rows": rows, "batch_voxels": bv, "serve": serve,
           "measurements": {"pass1_slab_rows": rows_m, "batch_voxels": bv_m, "serve": serve_m,
                            "seconds": round(time.perf_counter() - t0, 1)}}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps({k: out[k] for k in ("pass1_slab_rows", "batch_voxels", "serve")}))


if __name__ == "__main__":
    main()
