"""Batched meshing: one sequence of launches over many molecules gives, for every molecule,
exactly the mesh of a call on that molecule alone (and so the oracle's mesh)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_mesh(a, b):
    Va, Ta, Aa = a
    Vb, Tb, Ab = b
    assert Va.shape == Vb.shape, (Va.shape, Vb.shape)
    assert Ta.shape == Tb.shape, (Ta.shape, Tb.shape)
    assert np.array_equal(Ta, Tb)
    assert np.array_equal(Va.view(np.uint32), Vb.view(np.uint32)), np.abs(Va - Vb).max()
    assert np.array_equal(Aa, Ab)


def _mixed(qs):
    mols = []
    for i, n in enumerate([700, 0, 2, 1500, 1, 300, 40]):
        if n >= 2:
            mols.append(qs.synth_ball(700 + i, n))
        else:
            pos, rad = qs.synth_ball(700 + i, 2)
            mols.append((pos[:n], rad[:n]))
    return mols


@pytest.mark.parametrize("smooth", [0, 1, 2])
def test_batch_equals_single(qs, smooth):
    mols = _mixed(qs)
    reso = [0.5, 0.5, 0.3, 0.4, 0.5, 0.2, 1.0]
    print("case batch mixed smooth=%d" % smooth, flush=True)
    out = qs.compute_batch(mols, reso, smooth)
    assert len(out) == len(mols)
    for m, ((pos, rad), r) in enumerate(zip(mols, reso)):
        single = qs.compute_ses(pos, rad, r, smooth)
        _same_mesh(out[m], single)
        if len(rad) < 2:
            assert len(out[m][0]) == 0 and len(out[m][1]) == 0


def test_batch_matches_oracle(qs, orc):
    mols = [qs.synth_ball(720 + i, n) for i, n in enumerate([400, 900, 250])]
    reso = [0.5, 0.3, 0.2]
    print("case batch oracle", flush=True)
    out = qs.compute_batch(mols, reso, 1)
    for m, ((pos, rad), r) in enumerate(zip(mols, reso)):
        _same_mesh(out[m], orc.compute_ses(pos, rad, r, 1))


def test_batch_grids(qs):
    mols = [qs.synth_ball(740 + i, n) for i, n in enumerate([300, 600])]
    qs.compute_batch(mols, [0.5, 0.25], 0)
    grids = [qs.batch_grid(m) for m in range(2)]
    for m, ((pos, rad), r) in enumerate(zip(mols, [0.5, 0.25])):
        qs.compute_ses(pos, rad, r, 0)
        assert qs.last_grid() == grids[m]
    with pytest.raises(qs.QuickSESError):
        qs.batch_grid(5)


def test_trajectory_frames(qs):
    base_pos, rad = qs.synth_ball(750, 1000)
    frames = [(qs.synth_jitter(760 + f, base_pos, 0.3), rad) for f in range(6)]
    print("case frames n=1000 x6", flush=True)
    out = qs.compute_batch(frames, 0.5, 1)
    for m, (pos, r) in enumerate(frames):
        _same_mesh(out[m], qs.compute_ses(pos, r, 0.5, 1))


def test_batch_device_entry(qs):
    from quickses_amd import hipmem

    mols = [qs.synth_ball(780 + i, n) for i, n in enumerate([500, 0, 800])]
    P = np.concatenate([p for p, _ in mols]).astype(np.float32)
    R = np.concatenate([r for _, r in mols]).astype(np.float32)
    counts = [len(r) for _, r in mols]
    host = qs.compute_batch(mols, 0.4, 1)
    dp, dr = hipmem.DeviceArray(P), hipmem.DeviceArray(R)
    try:
        print("case batch device", flush=True)
        vo, to, pv, pt, pa = qs.compute_batch_device(dp.ptr, dr.ptr, counts, 0.4, 1)
        nv, nt = int(vo[-1]), int(to[-1])
        V = hipmem.download(pv, (nv, 3), np.float32)
        T = hipmem.download(pt, (nt, 3), np.int32)
        A = hipmem.download(pa, (nv,), np.int32)
    finally:
        dp.free()
        dr.free()
    for m in range(len(mols)):
        v0, v1, t0, t1 = vo[m], vo[m + 1], to[m], to[m + 1]
        _same_mesh((V[v0:v1], T[t0:t1] - np.int32(v0), A[v0:v1]), host[m])


def test_batch_errors(qs):
    pos, rad = qs.synth_ball(790, 100)
    bad = pos.copy()
    bad[3, 1] = np.nan
    with pytest.raises(qs.QuickSESError):
        qs.compute_batch([(pos, rad), (bad, rad)], 0.5, 1)
    with pytest.raises(qs.QuickSESError):
        qs.compute_batch([(pos, rad)], -1.0, 1)
    assert qs.compute_batch([], 0.5, 1) == []


def test_batch_getters_agree(qs):
    mols = _mixed(qs)
    out = qs.compute_batch(mols, 0.5, 1)
    vo, to = qs.batch_offsets()
    assert len(vo) == len(mols) + 1
    for m in range(len(mols)):
        assert vo[m + 1] - vo[m] == len(out[m][0]) and to[m + 1] - to[m] == len(out[m][1])
        _same_mesh(qs.batch_mesh(m), out[m])


def test_library_host_arrays(qs):
    """Views of the library's (pinned) host arrays equal the copied results: the batch fetch into
    library-owned memory, and API_computeSES's arrays without a copy; a larger batch after a
    smaller one grows the arrays, a smaller one after it reuses them."""
    mols = _mixed(qs)
    big = mols + [qs.synth_ball(790, 6000)]
    for case in (mols, big, mols[:3]):
        print("case host views nmol=%d" % len(case), flush=True)
        ref = qs.compute_batch(case, 0.4, 1)
        views = qs.compute_batch(case, 0.4, 1, copy=False)
        assert len(views) == len(ref)
        for a, b in zip(views, ref):
            _same_mesh(a, b)
    pos, rad = qs.synth_ball(791, 3000)
    ref = qs.compute_ses(pos, rad, 0.3, 1)
    view = qs.compute_ses(pos, rad, 0.3, 1, copy=False)
    _same_mesh(view, ref)
    qs.lib().API_freeMesh()


def test_pipelined_host_fetch(qs):
    """Batches fetched through the two pipelined slots, each batch's copies overlapping the next
    batch's work, equal the batches fetched one at a time; sizes grow, shrink and grow again."""
    mols = _mixed(qs)
    chunks = [mols[:4], mols + [qs.synth_ball(792, 5000)], mols[2:5], mols[:1], mols]
    ref = [qs.compute_batch(ch, 0.4, 1) for ch in chunks]
    print("case pipelined host fetch, %d batches" % len(chunks), flush=True)
    got = 0
    for k, out in enumerate(qs.compute_batches_host(chunks, 0.4, 1)):
        assert len(out) == len(ref[k])
        for a, b in zip(out, ref[k]):
            _same_mesh(a, b)
        got += 1
    assert got == len(chunks)


@pytest.mark.parametrize("opt", ["lim_mode", "atom_ids_sorted", "pass1_direct", "pass1_rows", "pass2_direct"])
def test_batch_options_match_oracle(qs, orc, opt):
    """Each option in a batch: the per-voxel cross-check kernels and the reference's window and
    atom-id conventions give every molecule the oracle's mesh."""
    mols = [qs.synth_ball(800 + i, n) for i, n in enumerate([300, 700])]
    reso = [0.5, 1.0]
    old = qs.get_option(opt)
    qs.set_option(opt, 1)
    try:
        print("case batch option %s" % opt, flush=True)
        out = qs.compute_batch(mols, reso, 1)
    finally:
        qs.set_option(opt, old)
    kw = {}
    if opt == "lim_mode":
        kw["lim_mode"] = 1
    if opt == "atom_ids_sorted":
        kw["atom_ids_sorted"] = True
    for m, ((pos, rad), r) in enumerate(zip(mols, reso)):
        _same_mesh(out[m], orc.compute_ses(pos, rad, r, 1, **kw))
