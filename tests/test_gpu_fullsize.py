"""Full-size workloads (the benchmark configurations) through the device entry point, checked by
size-independent properties: the unsmoothed mesh is a closed, consistently oriented surface
(every edge is used once in each direction), atom ids are in range, the smoothed vertices
stay near the surface, runs are deterministic, and the host API and a batch give the same
mesh as the device entry point."""
import hashlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run_device(qs, pos, rad, reso, smooth):
    from quickses_amd import hipmem

    dp, dr = hipmem.DeviceArray(pos), hipmem.DeviceArray(rad)
    try:
        nv, ni, pv, pt, pa = qs.compute_device(dp.ptr, dr.ptr, len(rad), reso, smooth)
        V = hipmem.download(pv, (nv, 3), np.float32)
        T = hipmem.download(pt, (ni // 3, 3), np.int32)
        A = hipmem.download(pa, (nv,), np.int32)
        V0 = qs.fetch("vert_unsmoothed", np.float32, 3 * nv).reshape(-1, 3)
        grid = qs.last_grid()
    finally:
        dp.free()
        dr.free()
    return V, T, A, V0, grid


def _check_closed_oriented(T, nv):
    """Every directed edge (u, v), u != v, of the non-degenerate triangles occurs exactly once, and
    so does (v, u): the undirected edges, each tagged with its direction, sort into pairs."""
    keep = (T[:, 0] != T[:, 1]) & (T[:, 1] != T[:, 2]) & (T[:, 0] != T[:, 2])
    t = T[keep].astype(np.uint64)
    u = np.concatenate([t[:, 0], t[:, 1], t[:, 2]])
    v = np.concatenate([t[:, 1], t[:, 2], t[:, 0]])
    key = (np.minimum(u, v) * np.uint64(nv) + np.maximum(u, v)) * np.uint64(2) + (u < v).astype(np.uint64)
    del u, v
    key.sort()
    assert len(key) % 2 == 0
    lo, hi = key[0::2], key[1::2]
    assert np.all(lo % 2 == 0) and np.all(hi == lo + 1), "an edge is not used once in each direction"
    assert np.all(lo[1:] > hi[:-1]), "an edge is used by more than two triangles"
    used = np.count_nonzero(np.bincount(t.ravel().astype(np.int64), minlength=nv))
    chi = used - len(lo) + len(t)
    assert chi % 2 == 0
    return chi


def _digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


CASES = [
    ("config 2: 50,000 atoms, 0.2 A", "ball", 2, 50000, 0.2, 1),
    ("config 5 largest: 200,000 atoms, 0.3 A", "ball", 55, 200000, 0.3, 1),
    ("config 3: 1,000,000-atom shell, 0.5 A", "shell", 3, 1000000, 0.5, 2),
]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("label,kind,seed,n,reso,smooth", CASES)
def test_fullsize_properties(qs, label, kind, seed, n, reso, smooth):
    if kind == "ball":
        pos, rad = qs.synth_ball(seed, n)
    else:
        pos, rad = qs.synth_shell(seed, n, 120.0, 186.7)
    print("case %s" % label, flush=True)
    V, T, A, V0, grid = _run_device(qs, pos, rad, reso, smooth)
    nv = len(V)
    Tall = qs.fetch("tri_all", np.int32, 3 * (4 * nv + 16)).reshape(-1, 3)
    print("  g=%d vertices=%d triangles=%d (all %d)" % (grid["g"], nv, len(T), len(Tall)), flush=True)
    assert nv > 0 and len(T) > 0
    assert T.min() >= 0 and T.max() < nv and Tall.max() < nv
    chi = _check_closed_oriented(Tall, nv)
    # the filtered triangles are the non-degenerate ones, in order
    keep = (Tall[:, 0] != Tall[:, 1]) & (Tall[:, 1] != Tall[:, 2]) & (Tall[:, 0] != Tall[:, 2])
    assert np.array_equal(T, Tall[keep])
    assert A.min() >= 0 and A.max() < n
    # vertices lie in the grid; smoothing moves each by less than a voxel per pass
    o = np.array(grid["origin"], np.float64)
    hi = o + (grid["g"] - 1) * grid["dx"]
    assert np.all(V0 >= o - 1e-3) and np.all(V0 <= hi + 1e-3)
    assert np.abs(V.astype(np.float64) - V0).max() <= smooth * reso * np.sqrt(3)
    # deterministic, and the same through the host API and inside a batch
    first = _digest(V, T, A)
    V2, T2, A2, _, _ = _run_device(qs, pos, rad, reso, smooth)
    assert _digest(V2, T2, A2) == first
    if n <= 200000:
        Vh, Th, Ah = qs.compute_ses(pos, rad, reso, smooth)
        assert _digest(Vh, Th, Ah) == first
        small = qs.synth_ball(seed + 1, 500)
        out = qs.compute_batch([small, (pos, rad)], [0.5, reso], smooth)
        assert _digest(*out[1]) == first
    print("  closed and consistently oriented, Euler characteristic %d" % chi, flush=True)


@pytest.mark.timeout(900)
def test_million_atoms_at_0_2A_in_slabs(qs):
    """The 1,000,000-atom shell at 0.2 A: g = 1,913, 7.0 G grid points (2^32 and more, so it runs
    in the slab form).  The mesh is closed and consistently oriented, and the same bits in one
    slab as in three."""
    pos, rad = qs.synth_shell(3, 1000000, 120.0, 186.7)
    print("case 1,000,000-atom shell at 0.2 A, slabs as planned", flush=True)
    V, T, A, V0, grid = _run_device(qs, pos, rad, 0.2, 1)
    g = grid["g"]
    assert g ** 3 >= 2 ** 32
    assert qs.last_slab_count() >= 1
    nv = len(V)
    Tall = qs.fetch("tri_all", np.int32, 3 * (4 * nv + 16)).reshape(-1, 3)
    print("  g=%d slabs=%d vertices=%d triangles=%d" % (g, qs.last_slab_count(), nv, len(T)), flush=True)
    assert nv > 0 and T.min() >= 0 and T.max() < nv and Tall.max() < nv
    assert A.min() >= 0 and A.max() < len(rad)
    first = _digest(V, T, A)
    del V0
    chi = _check_closed_oriented(Tall, nv)
    del Tall
    print("  closed and consistently oriented, Euler characteristic %d" % chi, flush=True)
    old = qs.get_option("slab_planes")
    qs.set_option("slab_planes", (g + 2) // 3)
    try:
        print("case 1,000,000-atom shell at 0.2 A in 3 slabs", flush=True)
        V3, T3, A3, _, _ = _run_device(qs, pos, rad, 0.2, 1)
        assert qs.last_slab_count() == 3
    finally:
        qs.set_option("slab_planes", old)
    assert _digest(V3, T3, A3) == first
