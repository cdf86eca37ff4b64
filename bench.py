"""Benchmark of the SES mesh path on one MI355X.

Workload (one step): one SES mesh of a 50,000-atom synthetic molecule (ball at 0.05 atoms/A^3,
1.2 A minimum distance, seed 2, radius mix C/O/N/H/S), voxel 0.2 A, probe 1.4 A, one Laplacian
pass, through the device-resident entry point (atoms already in HBM, mesh left in HBM).

value = grid voxels (g^3) processed per second over the timed steps.
Also reported: meshes/s, the host-in/host-out rate of API_computeSES, per-stage GPU times,
the dominant kernel against the HBM roofline, and the CPU oracle on a bounded sample.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(name="plain single-molecule SES, 50,000 atoms, voxel 0.2 A", seed=2, atoms=50000, voxel=0.2, smooth=1)


def algorithmic_bytes(stage, V, N, ncell):
    """Bytes a stage must move at minimum (HBM), per launch."""
    if stage == "pass1":
        return V * 1 + N * 16 + ncell * 8  # label store; atoms and cell ranges read once
    if stage == "pass2":
        return V * (1 + 4)  # label load, field store
    return None


def cpu_baseline(pos, rad, voxel, target_s=12.0):
    from oracle import oracle  # CPU baseline leg only

    t, g, v = oracle.bench_slab(pos, rad, voxel, 4)
    per_plane = t / 4.0
    planes = int(max(4, min(g - 1, target_s / max(per_plane, 1e-6))))
    t, g, v = oracle.bench_slab(pos, rad, voxel, planes)
    return dict(value=v / t, unit="voxels/s", cores=oracle.num_threads(), kind="port",
                sample="CPU oracle (OpenMP, IEEE fp32): grid setup, hashing, cell order, field pass 1 and "
                       "brute-force pass 2 and cube classification on the middle %d of %d x-planes of the same "
                       "molecule (%d voxels, %.1f s)" % (planes, g, v, t))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the SES path runs on a single device and this build has no "
              "multi-GPU mode; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import quickses_amd as qs
    from quickses_amd import hipmem

    if hipmem.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    W = WORKLOAD
    pos, rad = qs.synth_ball(W["seed"], W["atoms"])
    d_pos = hipmem.DeviceArray(pos)
    d_rad = hipmem.DeviceArray(rad)
    qs.set_option("timing", 1)

    def step():
        return qs.compute_device(d_pos.ptr, d_rad.ptr, W["atoms"], W["voxel"], W["smooth"])

    for _ in range(args.warmup):
        step()
    grid = qs.last_grid()
    g, gn = grid["g"], grid["gn"]
    V = g ** 3
    stage_sum = {}
    hipmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nv, nidx, *_ = step()
        for k, v in qs.stage_times().items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    hipmem.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    stage_avg = {k: v / args.steps for k, v in stage_sum.items()}

    # dominant single-kernel stage against the HBM roofline
    kern = {k: v for k, v in stage_avg.items() if algorithmic_bytes(k, V, W["atoms"], gn ** 3) is not None}
    dom = max(kern, key=kern.get)
    abytes = algorithmic_bytes(dom, V, W["atoms"], gn ** 3)
    achieved = abytes / (kern[dom] * 1e-3) / 1e9
    roofline = dict(bound="hbm", kernel=dom, achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 5), traffic=None, avg_ms=round(kern[dom], 4),
                    algorithmic_bytes=abytes)

    # host-in / host-out rate (PCIe included) of the reference-compatible entry point
    qs.set_option("timing", 0)
    t2 = time.perf_counter()
    api_steps = max(1, min(5, args.steps))
    for _ in range(api_steps):
        qs.compute_ses(pos, rad, W["voxel"], W["smooth"])
    api_s = (time.perf_counter() - t2) / api_steps

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pos, rad, W["voxel"])
        cpu["value"] = round(cpu["value"], 1)

    out = {
        "metric": "voxels/s + meshes/s, batched vs plain, 1 MI355X; p50/p99 served latency",
        "value": round(V / (elapsed / args.steps), 1),
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, no downloads)",
        "config": {"workload": W["name"], "path": "plain, device-resident", "atoms": W["atoms"], "voxel_A": W["voxel"],
                   "grid": g, "voxels": V, "laplacian_passes": W["smooth"], "vertices": nv, "triangles": nidx // 3,
                   "parallelism": "single GPU"},
        "meshes_per_s": round(args.steps / elapsed, 3),
        "api_host_to_host_voxels_per_s": round(V / api_s, 1),
        "stage_ms": {k: round(v, 4) for k, v in stage_avg.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
