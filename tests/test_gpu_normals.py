"""Per-vertex normals (unit field gradients) from the GPU against the FP64 oracle, which
evaluates the same definition on the same FP32 field; plain, batch and device entry points."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# float32 gradients, blend and normalisation against float64 on identical field values
TOL = 1e-5


@pytest.mark.parametrize("seed,n,reso", [(21, 800, 0.5), (22, 400, 0.2), (23, 1500, 1.0)])
def test_normals_match_fp64_oracle(qs, orc, seed, n, reso):
    pos, rad = qs.synth_ball(seed, n)
    print("case normals seed=%d n=%d reso=%.1f" % (seed, n, reso), flush=True)
    V = qs.compute_ses(pos, rad, reso, 1)[0]
    got = qs.vertex_normals(len(V)).astype(np.float64)
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, reso)
    ref = orc.vertex_normals_fp64(G, G.g - 1, fld, len(V))
    err = np.abs(got - ref).max(axis=1)
    print("max component error %.3g" % err.max(), flush=True)
    assert err.max() < TOL
    nz = np.linalg.norm(ref, axis=1) > 0
    assert np.abs(np.linalg.norm(got[nz], axis=1) - 1).max() < 1e-5


def test_batch_normals_equal_plain(qs):
    mols = [qs.synth_ball(30 + i, n) for i, n in enumerate([300, 900])]
    reso = [0.5, 0.4]
    print("case batch normals", flush=True)
    qs.compute_batch(mols, reso, 1)
    vo, _ = qs.batch_offsets()
    batch = qs.vertex_normals(vo[-1])
    for m, ((pos, rad), r) in enumerate(zip(mols, reso)):
        V = qs.compute_ses(pos, rad, r, 1)[0]
        assert np.array_equal(qs.vertex_normals(len(V)), batch[vo[m]:vo[m + 1]])


def test_normals_device_entry(qs):
    from quickses_amd import hipmem
    pos, rad = qs.synth_ball(24, 700)
    V = qs.compute_ses(pos, rad, 0.5, 1)[0]
    host = qs.vertex_normals(len(V))
    d = hipmem.alloc(host.nbytes)
    try:
        ms = qs.vertex_normals_device(d)
        out = hipmem.download(d, host.shape, np.float32)
    finally:
        hipmem.free(d)
    assert ms >= 0.0
    assert np.array_equal(out, host)


def test_normals_wrong_count_is_an_error(qs):
    pos, rad = qs.synth_ball(25, 300)
    V = qs.compute_ses(pos, rad, 0.5, 1)[0]
    with pytest.raises(qs.QuickSESError):
        qs.vertex_normals(len(V) + 1)


def test_colors_are_the_closest_atoms_colours(qs):
    """qs_vertex_colors gathers a per-atom RGBA table through the atom ids, plain and batched
    (batch tables concatenated like the atoms)."""
    mols = [qs.synth_ball(40 + i, n) for i, n in enumerate([400, 700])]
    rng = np.random.default_rng(3)
    tabs = [rng.random((len(r), 4)).astype(np.float32) for _, r in mols]
    for (pos, rad), tab in zip(mols, tabs):
        V, T, A = qs.compute_ses(pos, rad, 0.5, 1)
        col = qs.vertex_colors(tab, len(V))
        assert np.array_equal(col, tab[A])
    print("case batch colours", flush=True)
    out = qs.compute_batch(mols, 0.5, 1)
    vo, _ = qs.batch_offsets()
    col = qs.vertex_colors(np.concatenate(tabs), vo[-1])
    for m, (V, T, A) in enumerate(out):
        assert np.array_equal(col[vo[m]:vo[m + 1]], tabs[m][A])
    with pytest.raises(qs.QuickSESError):
        qs.vertex_colors(np.concatenate(tabs)[:-1], vo[-1])


def test_colors_with_sorted_atom_ids(qs):
    """With the reference's cell-sorted atom-id convention (atom_ids_sorted=1) the colours are
    still those of the closest atom in input order: the same colours as in the default
    convention, plain and batched."""
    mols = [qs.synth_ball(45 + i, n) for i, n in enumerate([500, 800])]
    rng = np.random.default_rng(4)
    tabs = [rng.random((len(r), 4)).astype(np.float32) for _, r in mols]
    ref = [tab[qs.compute_ses(pos, rad, 0.5, 1)[2]] for (pos, rad), tab in zip(mols, tabs)]
    old = qs.get_option("atom_ids_sorted")
    qs.set_option("atom_ids_sorted", 1)
    try:
        print("case colours, sorted ids", flush=True)
        for (pos, rad), tab, want in zip(mols, tabs, ref):
            V, T, A = qs.compute_ses(pos, rad, 0.5, 1)
            assert np.array_equal(qs.vertex_colors(tab, len(V)), want)
        qs.compute_batch(mols, 0.5, 1)
        vo, _ = qs.batch_offsets()
        col = qs.vertex_colors(np.concatenate(tabs), vo[-1])
    finally:
        qs.set_option("atom_ids_sorted", old)
    for m in range(len(mols)):
        assert np.array_equal(col[vo[m]:vo[m + 1]], ref[m])
