"""x-slab meshing (the reference's slices, CudaSurf.cu:626-754, with disjoint cube ownership):
a molecule meshed in several slabs gives the whole-grid mesh bit for bit -- vertices, triangles,
atom ids, smoothed positions -- so the seams weld exactly; and the oracle's mesh."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_mesh(a, b, what):
    Va, Ta, Aa = a
    Vb, Tb, Ab = b
    assert Va.shape == Vb.shape and Ta.shape == Tb.shape, (what, Va.shape, Vb.shape, Ta.shape, Tb.shape)
    assert np.array_equal(Ta, Tb), what
    assert np.array_equal(Va.view(np.uint32), Vb.view(np.uint32)), (what, float(np.abs(Va - Vb).max()))
    assert np.array_equal(Aa, Ab), what


def _slabbed(qs, pos, rad, reso, smooth, planes=0, mb=0):
    old = (qs.get_option("slab_planes"), qs.get_option("slab_mb"))
    qs.set_option("slab_planes", planes)
    qs.set_option("slab_mb", mb)
    try:
        out = qs.compute_ses(pos, rad, reso, smooth)
        n = qs.last_slab_count()
        tall = qs.fetch("tri_all", np.int32, 3 * (4 * len(out[0]) + 16)).reshape(-1, 3)
    finally:
        qs.set_option("slab_planes", old[0])
        qs.set_option("slab_mb", old[1])
    return out, n, tall


@pytest.mark.timeout(600)
def test_config2_in_slabs_equals_whole_grid(qs):
    """The 50,000-atom 0.2 A molecule (g = 665) in 4 slabs of 200 planes and in 7 of 100: the
    same mesh as the whole grid, and the same unfiltered triangles."""
    pos, rad = qs.synth_ball(2, 50000)
    print("case config 2 whole grid", flush=True)
    whole = qs.compute_ses(pos, rad, 0.2, 1)
    assert qs.last_slab_count() == 1
    tall = qs.fetch("tri_all", np.int32, 3 * (4 * len(whole[0]) + 16)).reshape(-1, 3)
    for planes, k in ((200, 4), (100, 7)):
        print("case config 2 in slabs of %d planes" % planes, flush=True)
        got, n, tall_s = _slabbed(qs, pos, rad, 0.2, 1, planes)
        assert n == k
        assert np.array_equal(tall_s, tall)
        _same_mesh(got, whole, "config 2, %d slabs" % k)


# (seed, atoms, voxel, smoothing passes, slab planes): window half-widths 1 to 10 (the fixed-R
# pass-2 kernels and, at 10, the direct window), slabs of 1 plane, fewer planes than a pass-2 tile
# (16), tile-straddling and uneven last slabs
CASES = [
    (301, 400, 1.0, 1, 1),
    (302, 700, 0.5, 0, 7),
    (303, 900, 0.5, 2, 16),
    (304, 600, 0.3, 1, 17),
    (305, 500, 0.2, 1, 33),
    (306, 300, 0.15, 2, 25),
    (307, 400, 0.7, 1, 3),
]


@pytest.mark.parametrize("seed,n,reso,smooth,planes", CASES)
def test_slabs_equal_whole_grid_and_oracle(qs, orc, seed, n, reso, smooth, planes):
    pos, rad = qs.synth_ball(seed, n)
    print("case slabs seed=%d n=%d reso=%.2f smooth=%d planes=%d" % (seed, n, reso, smooth, planes), flush=True)
    whole = qs.compute_ses(pos, rad, reso, smooth)
    got, k, _ = _slabbed(qs, pos, rad, reso, smooth, planes)
    g = qs.last_grid()["g"]
    assert k == (g + planes - 1) // planes and k >= 2
    _same_mesh(got, whole, "slabs vs whole grid")
    _same_mesh(got, orc.compute_ses(pos, rad, reso, smooth), "slabs vs oracle")


@pytest.mark.parametrize("opt", ["lim_mode", "atom_ids_sorted", "pass1_direct", "pass1_rows", "pass2_direct"])
def test_slabs_with_options(qs, opt):
    """The cross-check kernels and the reference conventions in slabs equal the whole grid."""
    pos, rad = qs.synth_ball(320, 600)
    old = qs.get_option(opt)
    qs.set_option(opt, 1)
    try:
        print("case slabs option %s" % opt, flush=True)
        whole = qs.compute_ses(pos, rad, 0.4, 1)
        got, k, _ = _slabbed(qs, pos, rad, 0.4, 1, 13)
    finally:
        qs.set_option(opt, old)
    assert k >= 2
    _same_mesh(got, whole, "option %s" % opt)


def test_slabs_from_memory_budget(qs):
    """A budget too small for the whole grid picks the slab size itself."""
    pos, rad = qs.synth_ball(330, 3000)
    whole = qs.compute_ses(pos, rad, 0.25, 1)
    print("case slabs from a 48 MiB budget", flush=True)
    got, k, _ = _slabbed(qs, pos, rad, 0.25, 1, 0, 48)
    assert k >= 2
    _same_mesh(got, whole, "budget slabs")


def test_slab_results_and_limits(qs):
    """After a slab run: the device entry point gives the same mesh, colours work, the field and
    normals are refused; batches are not slabbed."""
    from quickses_amd import hipmem

    pos, rad = qs.synth_ball(340, 800)
    whole = qs.compute_ses(pos, rad, 0.5, 1)
    qs.set_option("slab_planes", 10)
    try:
        dp, dr = hipmem.DeviceArray(pos), hipmem.DeviceArray(rad)
        try:
            nv, ni, pv, pt, pa = qs.compute_device(dp.ptr, dr.ptr, len(rad), 0.5, 1)
            assert qs.last_slab_count() >= 2
            V = hipmem.download(pv, (nv, 3), np.float32)
            T = hipmem.download(pt, (ni // 3, 3), np.int32)
            A = hipmem.download(pa, (nv,), np.int32)
        finally:
            dp.free()
            dr.free()
        _same_mesh((V, T, A), whole, "device entry in slabs")
        tab = np.random.default_rng(5).random((len(rad), 4)).astype(np.float32)
        assert np.array_equal(qs.vertex_colors(tab, nv), tab[A])
        with pytest.raises(qs.QuickSESError):
            qs.vertex_normals(nv)
        with pytest.raises(qs.QuickSESError):
            qs.fetch("field", np.float32, 10)
        other = qs.synth_ball(341, 500)
        out = qs.compute_batch([(pos, rad), other], 0.5, 1)
        assert qs.last_slab_count() == 1
        _same_mesh(out[0], whole, "batch with the slab option")
    finally:
        qs.set_option("slab_planes", 0)
