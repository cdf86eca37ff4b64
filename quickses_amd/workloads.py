"""Seeded synthetic workloads of the measured configurations, shared by the benchmark and the
workload-scale parity tests.

Every molecule comes from the library's generator (qs_synth.cpp): atoms uniform in a ball at
0.05 atoms/A^3 with 1.2 A minimum distance, radii drawn from the C/O/N/H/S mix.
"""
import numpy as np


def estimated_voxels(pos, rad, voxel):
    """g^3 of the molecule's grid by the grid rule (extent + 2 max radius + 4 probe)."""
    ext = float(np.max(pos.max(axis=0) - pos.min(axis=0)))
    g = int(np.ceil((ext + 2.0 * float(rad.max()) + 4 * 1.4) / voxel))
    return g ** 3


def chunks_by_voxels(mols, voxel, budget):
    """Consecutive groups of molecules whose grids together stay under `budget` voxels."""
    out, cur, acc = [], [], 0
    for pr in mols:
        v = estimated_voxels(pr[0], pr[1], voxel)
        if cur and acc + v > budget:
            out.append(cur)
            cur, acc = [], 0
        cur.append(pr)
        acc += v
    if cur:
        out.append(cur)
    return out


def config4_sets(qs, n_mol=256, n_frames=1000):
    """Batched workload: n_mol molecules (sizes log-uniform 500-20,000, size seed 4, molecule m
    seed 4000+m) and n_frames trajectory frames of one 10,000-atom molecule (seed 4500; frame f
    jittered by N(0, 0.3 A) per coordinate with seed 5000+f)."""
    sizes = qs.synth_sizes(4, n_mol, 500, 20000)
    mols = [qs.synth_ball(4000 + m, int(n)) for m, n in enumerate(sizes)]
    base_pos, base_rad = qs.synth_ball(4500, 10000)
    frames = [(qs.synth_jitter(5000 + f, base_pos, 0.3), base_rad) for f in range(n_frames)]
    return mols, frames


def config5_requests(qs, k):
    """Served workload: k surface requests, atom counts log-uniform 1,000-200,000 (size seed 5,
    request i seed 50000+i), voxel size uniform over {0.3, 0.5, 1.0} A."""
    sizes = qs.synth_sizes(5, k, 1000, 200000)
    rng = np.random.default_rng(5)
    voxels = rng.choice([0.3, 0.5, 1.0], size=k)
    return [(qs.synth_ball(50000 + i, int(n)), float(v)) for i, (n, v) in enumerate(zip(sizes, voxels))]
