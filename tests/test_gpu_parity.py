"""GPU path (libquickses.so through the C ABI) against the CPU oracle, same seeded inputs.

Integer outputs (atom order, cell ranges, labels, cube indices, vertex and triangle ids, atom
ids) must be identical; the field and vertex positions too, because both sides evaluate the
same single-precision formulas with one rounding per operation.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ball(qs, seed, n):
    return qs.synth_ball(seed, n)


def _gpu_mesh(qs, pos, rad, reso, smooth, **opts):
    saved = {k: qs.get_option(k) for k in opts}
    try:
        for k, v in opts.items():
            qs.set_option(k, v)
        return qs.compute_ses(pos, rad, reso, smooth)
    finally:
        for k, v in saved.items():
            qs.set_option(k, v)


def _same_mesh(a, b):
    Va, Ta, Aa = a
    Vb, Tb, Ab = b
    assert Va.shape == Vb.shape, (Va.shape, Vb.shape)
    assert Ta.shape == Tb.shape, (Ta.shape, Tb.shape)
    assert np.array_equal(Ta, Tb)
    assert np.array_equal(Va.view(np.uint32), Vb.view(np.uint32)), np.abs(Va - Vb).max()
    assert np.array_equal(Aa, Ab)


@pytest.mark.parametrize("seed,n,reso,smooth", [(1, 2000, 0.5, 1), (11, 300, 0.2, 1), (12, 500, 0.3, 2),
                                                (13, 800, 1.0, 1), (14, 1200, 0.5, 0)])
def test_end_to_end_matches_oracle(qs, orc, seed, n, reso, smooth):
    pos, rad = _ball(qs, seed, n)
    print("case seed=%d n=%d reso=%.2f smooth=%d" % (seed, n, reso, smooth), flush=True)
    gpu = qs.compute_ses(pos, rad, reso, smooth)
    ref = orc.compute_ses(pos, rad, reso, smooth)
    assert len(gpu[0]) > 0
    _same_mesh(gpu, ref)


def test_stages_match_oracle(qs, orc):
    pos, rad = _ball(qs, 1, 2000)
    print("case stages seed=1 n=2000", flush=True)
    qs.compute_ses(pos, rad, 0.5, 1)
    grid = qs.last_grid()
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, 0.5)
    assert (grid["g"], grid["gn"], grid["r"]) == (G.g, G.gn, G.r)
    assert np.array_equal(np.array(grid["origin"], np.float32), G.origin)
    assert np.array_equal(qs.fetch("perm", np.int32, len(rad)), perm)
    cells = qs.fetch("cells", np.int32, 2 * G.gn ** 3).reshape(-1, 2)
    assert np.array_equal(cells[:, 0], cs) and np.array_equal(cells[:, 1], ce)
    assert np.array_equal(qs.fetch("labels", np.uint8, G.g ** 3), lab)
    # the shell mask (the voxels whose field values pass 2 stores), bit z % 64 of word z / 64
    W = (G.g + 63) // 64
    sm = qs.fetch("shellm", np.uint64, G.g * G.g * W).reshape(G.g * G.g, W)
    bits = np.unpackbits(sm.view(np.uint8), bitorder="little").reshape(G.g * G.g, 64 * W)[:, :G.g]
    assert np.array_equal(bits.ravel(), (lab == 1).astype(np.uint8))
    f = qs.fetch("field", np.float32, G.g ** 3)
    assert np.array_equal(f.view(np.uint32), fld.view(np.uint32))


@pytest.mark.parametrize("seed,n,reso", [(31, 1500, 0.5), (32, 800, 0.2), (33, 600, 0.3), (34, 400, 1.0),
                                         (35, 300, 0.15)])
def test_field_matches_oracle(qs, orc, seed, n, reso):
    pos, rad = _ball(qs, seed, n)
    print("case field seed=%d n=%d reso=%.2f" % (seed, n, reso), flush=True)
    qs.compute_ses(pos, rad, reso, 0)
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, reso)
    assert np.array_equal(qs.fetch("labels", np.uint8, G.g ** 3), lab)
    assert np.array_equal(qs.fetch("field", np.float32, G.g ** 3).view(np.uint32), fld.view(np.uint32))


def _lattice():
    """atoms on a 1.5 A lattice: many voxel-atom distances tie exactly at the thresholds"""
    ax = np.arange(-4, 5, dtype=np.float32) * np.float32(1.5)
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    keep = np.linalg.norm(pos, axis=1) < 6.1
    pos = pos[keep]
    rad = np.where(np.arange(len(pos)) % 3 == 0, 1.5, 1.7).astype(np.float32)
    return pos, rad


@pytest.mark.parametrize("reso", [0.5, 0.25, 0.1])
def test_lattice_ties_match_oracle(qs, orc, reso):
    pos, rad = _lattice()
    print("case lattice reso=%.2f" % reso, flush=True)
    gpu = qs.compute_ses(pos, rad, reso, 1)
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, reso)
    assert np.array_equal(qs.fetch("labels", np.uint8, G.g ** 3), lab)
    assert np.array_equal(qs.fetch("field", np.float32, G.g ** 3).view(np.uint32), fld.view(np.uint32))
    _same_mesh(gpu, orc.compute_ses(pos, rad, reso, 1))


@pytest.mark.parametrize("opt", ["pass1_direct", "pass1_rows"])
def test_pass1_slab_equals_other_kernels(qs, opt):
    """The default slab-scatter pass 1 labels every voxel as the per-voxel 27-cell kernel and the
    row-gather kernel do, at a fine and a coarse voxel (many and few rows per slab)."""
    pos, rad = _ball(qs, 36, 1200)
    for reso in (0.5, 0.2, 1.3):
        print("case pass1 %s reso=%.1f" % (opt, reso), flush=True)
        qs.compute_ses(pos, rad, reso, 0)
        g = qs.last_grid()["g"]
        slab = qs.fetch("labels", np.uint8, g ** 3)
        qs.set_option(opt, 1)
        try:
            qs.compute_ses(pos, rad, reso, 0)
        finally:
            qs.set_option(opt, 0)
        assert np.array_equal(qs.fetch("labels", np.uint8, g ** 3), slab)


def test_pass2_tile_equals_direct(qs, orc):
    pos, rad = _ball(qs, 21, 600)
    for reso in (0.2, 0.5):
        G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, reso)
        params = G.params(G.g - 1)
        print("case pass2 reso=%.1f g=%d" % (reso, G.g), flush=True)
        tile = qs.stage_pass2(params, lab)
        qs.set_option("pass2_direct", 1)
        try:
            direct = qs.stage_pass2(params, lab)
        finally:
            qs.set_option("pass2_direct", 0)
        assert np.array_equal(tile.view(np.uint32), fld.view(np.uint32))
        assert np.array_equal(direct.view(np.uint32), fld.view(np.uint32))


@pytest.mark.parametrize("reso", [0.18, 0.16, 1.5])
def test_window_extremes_match_oracle(qs, orc, reso):
    """Window half-widths 8 and 9 (2R + 1 above 16: the unpacked y and x passes of k_pass2_r) and
    1 (the narrowest): field pass 2 from the oracle's labels and the whole mesh, bit for bit."""
    pos, rad = _ball(qs, 23, 120)
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, reso)
    assert G.r in (1, 8, 9)
    print("case wide window r=%d reso=%.2f g=%d" % (G.r, reso, G.g), flush=True)
    got = qs.stage_pass2(G.params(G.g - 1), lab)
    assert np.array_equal(got.view(np.uint32), fld.view(np.uint32))
    _same_mesh(qs.compute_ses(pos, rad, reso, 1), orc.compute_ses(pos, rad, reso, 1))


def test_stage_mesh_matches_oracle(qs, orc):
    pos, rad = _ball(qs, 22, 700)
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, 0.3)
    for lim in (G.g - 1, G.g - 2 * G.r - 1):
        print("case mesh lim=%d" % lim, flush=True)
        V, T = qs.stage_mesh(G.params(lim), fld)
        Vo, To = orc.edge_weld(G, lim, fld)
        assert np.array_equal(T, To)
        assert np.array_equal(V.view(np.uint32), Vo.view(np.uint32))


def test_reference_conventions(qs, orc):
    pos, rad = _ball(qs, 2, 900)
    print("case conventions", flush=True)
    gpu = _gpu_mesh(qs, pos, rad, 0.5, 1, lim_mode=1, atom_ids_sorted=1)
    ref = orc.compute_ses(pos, rad, 0.5, 1, lim_mode=1, atom_ids_sorted=True)
    _same_mesh(gpu, ref)


def test_invert_triangles(qs):
    pos, rad = _ball(qs, 3, 400)
    V, T, A = qs.compute_ses(pos, rad, 0.5, 1)
    V2, T2, A2 = qs.compute_ses(pos, rad, 0.5, 1, invert_triangles=True)
    assert np.array_equal(T2[:, [1, 0, 2]], T)


def test_small_and_invalid_inputs(qs):
    for n in (0, 1):
        pos = np.zeros((n, 3), np.float32)
        rad = np.full(n, 1.5, np.float32)
        V, T, A = qs.compute_ses(pos, rad, 0.5, 1)
        assert len(V) == 0 and len(T) == 0
    pos = np.array([[0, 0, 0], [1.0, 0, 0]], np.float32)
    rad = np.array([1.7, 1.52], np.float32)
    V, T, A = qs.compute_ses(pos, rad, 0.5, 1)
    assert len(T) > 0 and set(np.unique(A)) <= {0, 1}
    bad = pos.copy()
    bad[1, 2] = np.nan
    with pytest.raises(qs.QuickSESError):
        qs.compute_ses(bad, rad, 0.5, 1)
    with pytest.raises(qs.QuickSESError):
        qs.compute_ses(pos, rad, -0.5, 1)


def test_device_entry_point(qs):
    from quickses_amd import hipmem

    pos, rad = _ball(qs, 1, 2000)
    ref = qs.compute_ses(pos, rad, 0.5, 1)
    dp = hipmem.DeviceArray(pos)
    dr = hipmem.DeviceArray(rad)
    nv, ni, pv, pt, pa = qs.compute_device(dp.ptr, dr.ptr, len(rad), 0.5, 1)
    assert nv == len(ref[0]) and ni == 3 * len(ref[1])
    V = hipmem.download(pv, (nv, 3), np.float32)
    T = hipmem.download(pt, (ni // 3, 3), np.int32)
    A = hipmem.download(pa, (nv,), np.int32)
    assert np.array_equal(V, ref[0]) and np.array_equal(T, ref[1]) and np.array_equal(A, ref[2])


def test_repeated_runs_with_stale_tile_marks(qs, orc):
    """The pass-2 tile marks are never cleared: each run marks with its own value (1-255, in
    turn), and a stale mark equal to it only makes pass 2 rewrite a tile without shell voxels.
    Two molecules with different tile layouts alternate over more than two turns of the marks;
    every mesh equals that molecule's first mesh bit for bit, and the first ones equal the
    oracle's."""
    a = qs.synth_ball(41, 700)
    b = qs.synth_ball(42, 1500)
    first = [qs.compute_ses(*a, 0.4, 1), qs.compute_ses(*b, 0.5, 1)]
    _same_mesh(first[0], orc.compute_ses(*a, 0.4, 1))
    _same_mesh(first[1], orc.compute_ses(*b, 0.5, 1))
    print("case 600 alternating runs", flush=True)
    for k in range(600):
        got = qs.compute_ses(*a, 0.4, 1) if k % 2 == 0 else qs.compute_ses(*b, 0.5, 1)
        _same_mesh(got, first[k % 2])


@pytest.mark.parametrize("seed,n,reso,smooth", [(51, 1500, 0.4, 1), (52, 3000, 0.25, 2), (53, 800, 0.6, 0)])
def test_side_stream_runs_equal_one_stream(qs, seed, n, reso, smooth):
    """Without stage timing one molecule's vertex kernel and closest-atom pass run on a second
    stream beside the triangle kernel and smoothing; with it every kernel runs in order on one
    stream.  Both give the same mesh, bit for bit, and repeated runs do too."""
    pos, rad = qs.synth_ball(seed, n)
    old = qs.get_option("timing")
    try:
        qs.set_option("timing", 1)
        ref = qs.compute_ses(pos, rad, reso, smooth)
        qs.set_option("timing", 0)
        print("case side stream seed=%d n=%d reso=%.2f smooth=%d" % (seed, n, reso, smooth), flush=True)
        for _ in range(3):
            _same_mesh(qs.compute_ses(pos, rad, reso, smooth), ref)
    finally:
        qs.set_option("timing", old)
