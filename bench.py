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


def algorithmic_bytes(stage, V, N, ncell, S):
    """Bytes a kernel must move at minimum (HBM) per launch, for the single-kernel stages.

    pass1: writes the outside and shell bit masks and the default sign and zero masks (4 bits
           per voxel), reads the cell-ordered atoms with their thresholds (32 B per atom) and the
           cell starts (4 B per cell).
    pass2: reads the two bit masks and writes the sign and zero bit masks (2 + 2 bits per
           voxel), and the f32 field of the S shell voxels (4 B each; every other voxel's value
           is fixed by the masks: the probe radius outside, -dx inside)."""
    if stage == "pass1":
        return V // 2 + 32 * N + 4 * ncell
    if stage == "pass2":
        return V // 2 + 4 * S
    return None


def shell_voxels(qs, g):
    """Shell voxels of the last single-molecule computation, from its shell mask."""
    import numpy as np
    W = (g + 63) // 64
    words = qs.fetch("shellm", np.uint64, g * g * W).reshape(g * g, W)
    tail = g - 64 * (W - 1)
    if tail < 64:
        words[:, -1] &= np.uint64((1 << tail) - 1)
    return int(np.unpackbits(words.view(np.uint8)).sum())


STAGE_KERNEL = {"pass1": "k_pass1_slab", "pass2": "k_pass2_r"}


def profiled_traffic(stage):
    """HBM bytes per launch of a stage's kernel from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE, each scaled by the same counter on a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_ses_50k_0.2A.txt")
    if stage not in STAGE_KERNEL or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic ") and ("kernel=%s" % STAGE_KERNEL[stage]) in line:
                for tok in line.split():
                    if tok.startswith("bytes_per_launch="):
                        return int(tok.split("=", 1)[1])
    return None


def profiled_valu_rate(stage):
    """VALU wave instructions per CU per cycle of a stage's kernel, from the committed instruction
    counter summary (SQ_INSTS_VALU over GRBM_GUI_ACTIVE / 8 cycles and 256 CUs), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_instructions_ses_50k_0.2A.md")
    if stage not in STAGE_KERNEL or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 8 and cells[0].strip("`").startswith(STAGE_KERNEL[stage]):
                try:
                    return float(cells[7])
                except ValueError:
                    return None
    return None


def copy_bandwidth(qs, hipmem, nbytes=1 << 30, reps=3):
    """GB/s (read + write) of the library's plain copy kernel over nbytes."""
    a, b = hipmem.alloc(nbytes), hipmem.alloc(nbytes)
    try:
        qs.copy_device(b, a, nbytes)
        ms = min(qs.copy_device(b, a, nbytes) for _ in range(reps))
    finally:
        hipmem.free(a)
        hipmem.free(b)
    return 2 * nbytes / (ms * 1e-3) / 1e9


def cpu_baseline(pos, rad, voxel, target_s=25.0):
    from oracle import oracle  # CPU baseline leg only

    t, g, v, _ = oracle.bench_slab(pos, rad, voxel, 4)
    per_plane = t / 4.0
    planes = int(max(4, min(g - 1, target_s / max(per_plane, 1e-6))))
    t, g, v, nt = oracle.bench_slab(pos, rad, voxel, planes)
    return dict(value=v / t, unit="voxels/s", cores=oracle.num_threads(), kind="port",
                sample="CPU oracle (OpenMP for the field passes and closest atom, IEEE fp32), the reference's whole "
                       "path on one slice as its slice loop runs it: grid setup, hashing, cell order, field passes 1 "
                       "and 2 (brute-force window), triangle soup, 0.001 A quantised weld, closest atom, one in-place "
                       "Laplacian pass, degenerate filter, on the middle %d of %d x-planes of the same molecule "
                       "(%d voxels, %d triangles, %.1f s)" % (planes, g, v, nt, t))


def bench_batched(qs, hipmem, voxel=0.5, smooth=1, n_frames=1000, budget=None):
    """meshes/s of the plain path (one call per molecule) against the batched path (one launch
    sequence per chunk of molecules) on the two streams of the batched workload, at two
    boundaries: device-resident (atoms already in HBM, meshes left in HBM) and host arrays in /
    host arrays out, the C ABI's boundary (API_computeSES per molecule, its results in the
    library's host arrays, against qs_compute_batch + qs_batch_fetch_host_async /
    qs_batch_wait_host over about four launch sequences, each sequence's copies to the library's
    host arrays overlapping the next sequence's work).  Each leg runs once untimed first, so that
    the library's grow-only device and host arrays have their steady-state sizes."""
    import numpy as np

    from quickses_amd.workloads import chunks_by_voxels, config4_sets, estimated_voxels

    if budget is None:  # grid points per launch sequence: the tuned value, else 1.5 G
        budget = int(qs.tuning().get("batch_voxels", {}).get("%g" % voxel, 1_500_000_000))
    mols, frames = config4_sets(qs, 256, n_frames)
    link = hipmem.d2h_rate()  # device-to-host rate into pinned memory: the ceiling of host-to-host results
    res = {}
    for name, items in (("molecules_256", mols), ("frames_%d" % n_frames, frames)):
        vox = sum(estimated_voxels(p, r, voxel) for p, r in items)
        chunks = chunks_by_voxels(items, voxel, budget)
        counts = [len(r) for _, r in items]
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        d_pos = hipmem.DeviceArray(np.concatenate([p for p, _ in items]).astype(np.float32))
        d_rad = hipmem.DeviceArray(np.concatenate([r for _, r in items]).astype(np.float32))
        bounds, k = [], 0
        for ch in chunks:
            bounds.append((k, k + len(ch)))
            k += len(ch)
        # warm-up: every launch sequence once, so that the library's grow-only buffers have their
        # steady-state sizes before the timed pass (a stream of batches allocates once)
        for a, b in bounds:
            qs.compute_batch_device(d_pos.ptr + 12 * int(off[a]), d_rad.ptr + 4 * int(off[a]), counts[a:b], voxel, smooth)
        # device-resident: batched
        hipmem.synchronize()
        t0 = time.perf_counter()
        nv_bd = 0
        for a, b in bounds:
            vo, *_ = qs.compute_batch_device(d_pos.ptr + 12 * int(off[a]), d_rad.ptr + 4 * int(off[a]), counts[a:b], voxel,
                                             smooth)
            nv_bd += int(vo[-1])
        hipmem.synchronize()
        tbd = time.perf_counter() - t0
        qs.set_option("timing", 1)
        a, b = bounds[0]
        qs.compute_batch_device(d_pos.ptr + 12 * int(off[a]), d_rad.ptr + 4 * int(off[a]), counts[a:b], voxel, smooth)
        chunk_stage_ms = {k: round(v, 3) for k, v in qs.stage_times().items()}
        qs.set_option("timing", 0)
        # device-resident: plain, one call per molecule
        t0 = time.perf_counter()
        nv_pd = 0
        for m in range(len(items)):
            nv, *_ = qs.compute_device(d_pos.ptr + 12 * int(off[m]), d_rad.ptr + 4 * int(off[m]), counts[m], voxel, smooth)
            nv_pd += nv
        hipmem.synchronize()
        tpd = time.perf_counter() - t0
        d_pos.free()
        d_rad.free()
        # host in, host out (views of the library's host arrays)
        host_chunks = chunks_by_voxels(items, voxel, max(vox // 4, 1))
        for _ in qs.compute_batches_host(host_chunks, voxel, smooth):
            pass
        t0 = time.perf_counter()
        nv_b = nt_b = 0
        for out in qs.compute_batches_host(host_chunks, voxel, smooth):
            nv_b += sum(len(o[0]) for o in out)
            nt_b += sum(len(o[1]) for o in out)
        tb = time.perf_counter() - t0
        d2h = 16 * nv_b + 12 * nt_b  # vertices and atom ids, triangles
        big = max(range(len(items)), key=lambda m: counts[m])
        qs.compute_ses(items[big][0], items[big][1], voxel, smooth, copy=False)
        t0 = time.perf_counter()
        nv_p = 0
        for p, r in items:
            nv_p += len(qs.compute_ses(p, r, voxel, smooth, copy=False)[0])
        tp = time.perf_counter() - t0
        qs.lib().API_freeMesh()
        if not nv_b == nv_p == nv_bd == nv_pd:
            raise RuntimeError("batched and plain meshes differ in size (%d %d %d %d vertices)" % (nv_b, nv_p, nv_bd, nv_pd))
        n = len(items)
        print("batched stream %s: %d meshes, device-resident plain %.3f s batched %.3f s, host-to-host plain %.3f s "
              "batched %.3f s" % (name, n, tpd, tbd, tp, tb), file=sys.stderr, flush=True)
        res[name] = dict(meshes=n, voxels=vox, budget_voxels=budget, launch_sequences=len(chunks), host_launch_sequences=len(host_chunks),
                         vertices=nv_b,
                         first_sequence=dict(meshes=bounds[0][1] - bounds[0][0], stage_ms=chunk_stage_ms),
                         device_resident=dict(plain_meshes_per_s=round(n / tpd, 2), batched_meshes_per_s=round(n / tbd, 2),
                                              batched_voxels_per_s=round(vox / tbd, 1), speedup=round(tpd / tbd, 3)),
                         host_to_host=dict(plain_meshes_per_s=round(n / tp, 2), batched_meshes_per_s=round(n / tb, 2),
                                           batched_voxels_per_s=round(vox / tb, 1), speedup=round(tp / tb, 3),
                                           d2h_bytes=d2h, link_GBps=round(link, 2),
                                           link_bound_meshes_per_s=round(n / (d2h / (link * 1e9)), 2)))
    streams = list(res.values())
    for key in ("device_resident", "host_to_host"):
        tot = sum(v["meshes"] for v in streams)
        tp = sum(v["meshes"] / v[key]["plain_meshes_per_s"] for v in streams)
        tb = sum(v["meshes"] / v[key]["batched_meshes_per_s"] for v in streams)
        res.setdefault("all", {"meshes": tot})[key] = dict(plain_meshes_per_s=round(tot / tp, 2),
                                                           batched_meshes_per_s=round(tot / tb, 2),
                                                           speedup=round(tp / tb, 3))
    res["workload"] = ("256 molecules (500-20,000 atoms, log-uniform) and %d trajectory frames of a 10,000-atom "
                       "molecule, voxel %.1f A, %d Laplacian pass" % (n_frames, voxel, smooth))
    return res


def bench_other_plain(qs, hipmem, reps=5):
    """Plain device-resident meshes of the other single-molecule configurations: config 1 (2,000
    atoms, 0.5 A, 1 pass) and config 3 (1,000,000 atoms on a shell of radii 120 and 186.7 A,
    0.5 A, 2 passes)."""
    out = {}
    for name, kind, seed, n, voxel, smooth in (("config1_2k_0.5A", "ball", 1, 2000, 0.5, 1),
                                               ("config3_1M_shell_0.5A", "shell", 3, 1000000, 0.5, 2)):
        pos, rad = qs.synth_ball(seed, n) if kind == "ball" else qs.synth_shell(seed, n, 120.0, 186.7)
        dp, dr = hipmem.DeviceArray(pos), hipmem.DeviceArray(rad)
        try:
            qs.compute_device(dp.ptr, dr.ptr, n, voxel, smooth)
            hipmem.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                nv, ni, *_ = qs.compute_device(dp.ptr, dr.ptr, n, voxel, smooth)
            hipmem.synchronize()
            dt = (time.perf_counter() - t0) / reps
        finally:
            dp.free()
            dr.free()
        g = qs.last_grid()["g"]
        out[name] = dict(atoms=n, voxel_A=voxel, laplacian_passes=smooth, grid=g, ms_per_mesh=round(1e3 * dt, 3),
                         voxels_per_s=round(g ** 3 / dt, 1), vertices=nv, triangles=ni // 3)
    return out


def _pct(xs, q):
    import numpy as np

    return round(float(np.percentile(np.asarray(xs), q)), 3)


def bench_served(qs, hipmem, k=400, load=0.8, workers=2, budget=48 << 30, arena=32 << 30, max_batch_voxels=0,
                 express_voxels=0):
    """Poisson arrivals at `load` x the plain path's capacity.  Plain path: each request alone,
    atoms copied to HBM and meshed by qs_compute_device (mesh left in HBM), first come first
    served.  Served path: the request server (continuous batching over `workers` engines, HBM
    budget admission, meshes resident in a device arena; with express_voxels > 0 the last worker
    takes only requests of at most that many grid points).  Latency = arrival -> mesh resident on
    the device."""
    import threading

    import numpy as np

    from quickses_amd.workloads import config5_requests

    reqs = config5_requests(qs, k)
    smooth = 1

    # the plain leg copies each request's atoms into device buffers allocated once (as large as the
    # largest request) and meshes it with qs_compute_device, mesh left in HBM: the same boundary
    # as the server's, without an allocation per request
    nmax = max(len(r[0][1]) for r in reqs)
    d_pos, d_rad = hipmem.alloc(12 * nmax), hipmem.alloc(4 * nmax)

    def plain_one(pos, rad, reso):
        hipmem.upload(d_pos, pos)
        hipmem.upload(d_rad, rad)
        qs.compute_device(d_pos, d_rad, len(rad), reso, smooth)

    plain_one(reqs[0][0][0], reqs[0][0][1], reqs[0][1])  # warm-up
    t0 = time.perf_counter()
    for (pos, rad), reso in reqs:
        plain_one(pos, rad, reso)
    t_plain_serial = time.perf_counter() - t0
    capacity = k / t_plain_serial
    lam = load * capacity
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, size=k))

    def wait_until(start, t):
        while True:
            d = start + t - time.perf_counter()
            if d <= 0:
                return
            time.sleep(min(d, 0.0005) if d > 0.0002 else 0)

    # plain path under the same arrivals: one request at a time, first come first served
    lat_plain = []
    start = time.perf_counter()
    for ((pos, rad), reso), ta in zip(reqs, arrivals):
        wait_until(start, ta)
        plain_one(pos, rad, reso)
        lat_plain.append(1e3 * (time.perf_counter() - start - ta))
    wall_plain = time.perf_counter() - start

    # served path
    lat, queue, bsize = [], [], []
    with qs.Server(hbm_budget_bytes=budget, result_bytes=arena, workers=workers, max_batch_voxels=max_batch_voxels,
                   express_voxels=express_voxels) as srv:
        w = srv.submit(reqs[0][0][0][:2000], reqs[0][0][1][:2000], 1.0, smooth)  # warm-up of every engine
        srv.poll(w, -1)
        srv.release(w)
        ids = [None] * k
        got = threading.Event()

        def poller():
            for i in range(k):
                while ids[i] is None:
                    time.sleep(0.0002)
                inf = srv.poll(ids[i], -1)
                lat.append(inf["latency_ms"])
                queue.append(inf["queue_ms"])
                bsize.append(inf["batch_size"])
                srv.release(ids[i])
            got.set()

        th = threading.Thread(target=poller)
        th.start()
        start = time.perf_counter()
        for i, (((pos, rad), reso), ta) in enumerate(zip(reqs, arrivals)):
            wait_until(start, ta)
            ids[i] = srv.submit(pos, rad, reso, smooth)
        th.join()
        wall_served = time.perf_counter() - start
        st = srv.stats()
        # saturated: every request submitted at once; throughput of the server against the plain
        # path's serial capacity
        start = time.perf_counter()
        ids = [srv.submit(pos, rad, reso, smooth) for (pos, rad), reso in reqs]
        lat_sat = []
        for rid in ids:
            lat_sat.append(srv.poll(rid, -1)["latency_ms"])
            srv.release(rid)
        wall_sat = time.perf_counter() - start
    hipmem.free(d_pos)
    hipmem.free(d_rad)
    return dict(
        requests=k, arrival_rate_per_s=round(lam, 2), load=load,
        plain_capacity_meshes_per_s=round(capacity, 2),
        plain=dict(meshes_per_s=round(k / wall_plain, 2), p50_ms=_pct(lat_plain, 50), p99_ms=_pct(lat_plain, 99)),
        served=dict(meshes_per_s=round(k / wall_served, 2), p50_ms=_pct(lat, 50), p99_ms=_pct(lat, 99),
                    queue_p50_ms=_pct(queue, 50), queue_p99_ms=_pct(queue, 99),
                    mean_batch=round(float(np.mean(bsize)), 2),
                    batches=st["batches"] - 1, workers=workers, max_busy_workers=st["max_busy_workers"],
                    hbm_budget_GiB=budget >> 30, max_batch_voxels=max_batch_voxels, express_voxels=express_voxels,
                    express_batches=st["express_batches"]),
        saturated=dict(plain_meshes_per_s=round(capacity, 2), served_meshes_per_s=round(k / wall_sat, 2),
                       speedup=round(wall_sat and (k / wall_sat) / capacity, 3), served_p50_ms=_pct(lat_sat, 50),
                       served_p99_ms=_pct(lat_sat, 99)),
        workload="config-5 stream: atom counts log-uniform 1,000-200,000, voxel {0.3, 0.5, 1.0} A, 1 Laplacian "
                 "pass, Poisson arrivals; latency = arrival -> mesh resident in HBM; plain leg: atoms copied into "
                 "device buffers allocated once, qs_compute_device per request",
    )


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain streams")
    ap.add_argument("--frames", type=int, default=1000, help="trajectory frames of the batched workload")
    ap.add_argument("--no-serve", action="store_true", help="skip the served stream")
    ap.add_argument("--requests", type=int, default=400, help="requests of the served stream")
    ap.add_argument("--serve-workers", type=int, default=None, help="server engines (default: tuned, else 2)")
    ap.add_argument("--serve-cap", type=int, default=None,
                    help="grid points per served batch, 0 = no cap (default: tuned, else 150 M)")
    ap.add_argument("--serve-express", type=int, default=None,
                    help="request size cap of the server's express worker, 0 = none (default: tuned, else 0)")
    ap.add_argument("--no-other", action="store_true", help="skip configs 1 and 3")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the SES path runs on a single device and this build has no "
              "multi-GPU mode; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import quickses_amd as qs
    from quickses_amd import hipmem

    tuned = qs.tuning().get("serve", {})
    if args.serve_workers is None:
        args.serve_workers = int(tuned.get("workers", 2))
    if args.serve_cap is None:
        args.serve_cap = int(tuned.get("max_batch_voxels", 150_000_000))
    if args.serve_express is None:
        args.serve_express = int(tuned.get("express_voxels", 0))

    if hipmem.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    W = WORKLOAD
    pos, rad = qs.synth_ball(W["seed"], W["atoms"])
    d_pos = hipmem.DeviceArray(pos)
    d_rad = hipmem.DeviceArray(rad)

    def step():
        return qs.compute_device(d_pos.ptr, d_rad.ptr, W["atoms"], W["voxel"], W["smooth"])

    qs.set_option("timing", 0)
    for _ in range(args.warmup):
        step()
    grid = qs.last_grid()
    g, gn = grid["g"], grid["gn"]
    V = g ** 3
    hipmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nv, nidx, *_ = step()
    hipmem.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps

    # per-stage times from HIP events on the library's stream, over extra steps
    qs.set_option("timing", 1)
    stage_sum = {}
    n_stage = max(3, min(10, args.steps))
    for _ in range(n_stage):
        step()
        for k, v in qs.stage_times().items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    qs.set_option("timing", 0)
    stage_avg = {k: v / n_stage for k, v in stage_sum.items()}

    # host-in / host-out rate (PCIe included) of the reference-compatible entry point, results in
    # the library's host arrays (two untimed calls first: they grow to the mesh's size once), timed
    # before the copy-bandwidth test below allocates and frees 2 GiB of device memory
    for _ in range(2):
        qs.compute_ses(pos, rad, W["voxel"], W["smooth"], copy=False)
    t2 = time.perf_counter()
    api_steps = max(1, min(5, args.steps))
    for _ in range(api_steps):
        qs.compute_ses(pos, rad, W["voxel"], W["smooth"], copy=False)
    api_s = (time.perf_counter() - t2) / api_steps
    qs.lib().API_freeMesh()

    # per-vertex normals of the last mesh (an extension beyond the reference's outputs): kernel
    # time from HIP events, 12 B written and 2-14 field values read per vertex
    d_nrm = hipmem.alloc(max(nv, 1) * 12)
    try:
        normals_ms = min(qs.vertex_normals_device(d_nrm) for _ in range(3))
    finally:
        hipmem.free(d_nrm)

    # dominant kernel (field pass 2) against the HBM roofline.  achieved = the field pass's
    # algorithmic bytes as stated for a field pass whose two halves meet on chip: 4 B per grid
    # voxel (a dense f32 field over g^3 voxels) in the pass-2 kernel's time.  own_bytes: what this
    # pass-2 kernel must move in its own form (2-bit masks in and out, 4 B per shell voxel).
    S = shell_voxels(qs, g)
    ab = lambda k: algorithmic_bytes(k, V, W["atoms"], gn ** 3, S)
    t2 = stage_avg["pass2"]
    dom = max(("pass1", "pass2"), key=lambda k: stage_avg[k])
    v4 = 4 * V
    achieved = v4 / (t2 * 1e-3) / 1e9
    own = ab("pass2")
    own_gbs = own / (t2 * 1e-3) / 1e9
    copy_gbs = copy_bandwidth(qs, hipmem)
    fp_ms = stage_avg["pass1"] + stage_avg.get("shell_scan", 0.0) + stage_avg["pass2"]
    roofline = dict(bound="hbm", kernel="pass2", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 5), traffic=profiled_traffic("pass2"), avg_ms=round(t2, 4),
                    algorithmic_bytes=v4, bytes_definition="4 B per grid voxel (g^3): the field pass with its two halves "
                                                           "meeting on chip",
                    dominant_stage=dom, valu_wave_instr_per_cu_cycle=profiled_valu_rate("pass2"),
                    own_bytes=dict(bytes=own, definition="2-bit outside/shell masks in, 2-bit sign/zero masks out, "
                                                         "4 B per shell voxel", shell_voxels=S, GBps=round(own_gbs, 2),
                                   frac=round(own_gbs / HBM_PEAK_GBS, 5)),
                    copy_kernel_GBps=round(copy_gbs, 1), frac_of_copy=round(achieved / copy_gbs, 5),
                    field_pass=dict(kernels="pass1 + shell-rank scan + pass2", algorithmic_bytes=v4, ms=round(fp_ms, 4),
                                    achieved=round(v4 / (fp_ms * 1e-3) / 1e9, 2),
                                    frac=round(v4 / (fp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)))


    other = None if args.no_other else bench_other_plain(qs, hipmem)

    batched = None
    if not args.no_batch:
        batched = bench_batched(qs, hipmem, n_frames=args.frames)

    served = None
    if not args.no_serve:
        served = bench_served(qs, hipmem, k=args.requests, workers=args.serve_workers,
                              max_batch_voxels=args.serve_cap, express_voxels=args.serve_express)
        print("served: %s" % json.dumps(served), file=sys.stderr, flush=True)

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
        "normals_ms": round(normals_ms, 4),
        "roofline": roofline,
        "other_configs": other,
        "batched": batched,
        "served": served,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
