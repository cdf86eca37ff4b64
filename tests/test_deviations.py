"""Figures of two intentional deviations from the reference, computed on the CPU oracle.

D3, weld.  The reference welds the triangle soup by rounding every coordinate to 0.001 A and
merging equal keys (groupVertices MarchingCubes.cu:853-862, sort/unique/lower_bound
CudaSurf.cu:428-458); the library welds by integer edge id.  The two induce partitions of the
triangle corners; a class of one weld that meets several classes of the other is a difference:
  split  - one edge-id vertex gets several quantised keys: the two directions in which the soup
           computes a shared edge round differently across a key boundary;
  merged - several edge-id vertices get one key: distinct vertices within the 0.001 A quantum.
D5, smoothing.  The reference smooths each slice in place (Gauss-Seidel in its vertex order, the
quantised weld's sorted order; smoothMeshLaplacian SmoothMesh.cpp:96-124); the library runs a
Jacobi pass over the same neighbour multiset.  Reported: the largest and mean distance, in
voxels, between a vertex smoothed both ways (where the welds agree), an upper bound of the two
meshes' Hausdorff distance.
The printed figures are the ones the design notes quote."""
import numpy as np
import pytest


def _ball(qs, seed, n):
    return qs.synth_ball(seed, n)


def _welds(orc, pos, rad, reso):
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, reso)
    lim = G.g - 1
    V, T = orc.edge_weld(G, lim, fld)
    S = orc.soup(G, lim, fld)
    Vq, Tq = orc.quantised_weld(S)
    assert S.shape[0] == 3 * len(T)
    return G, V, T, Vq, Tq


def _partition_diff(V, T, Vq, Tq):
    a = T.reshape(-1).astype(np.int64)
    b = Tq.reshape(-1).astype(np.int64)
    pairs = np.unique(np.stack([a, b], 1), axis=0)
    na = np.bincount(pairs[:, 0], minlength=len(V))
    nb = np.bincount(pairs[:, 1], minlength=len(Vq))
    return a, b, pairs, na, nb


def _half_quantum_axes(G):
    """Axes whose grid planes lie within 4 ulps of a half-quantum boundary (k + 0.5) * 0.001 A:
    the step is a multiple of the quantum, so then every plane of the axis does."""
    out = []
    for c in range(3):
        o = float(G.origin[c])
        frac = abs(o / 0.001 - np.floor(o / 0.001) - 0.5)
        if frac * 0.001 < 4 * np.spacing(np.float32(abs(o) + G.g * float(G.dx))):
            out.append(c)
    return out


# (seed, atoms, voxel): seed 11 at 2,000 atoms puts the x planes on half-quantum boundaries
WELD_CASES = [(11, 300, 0.2), (12, 600, 0.2), (11, 2000, 0.2), (1, 2000, 0.5), (13, 2000, 1.0)]


@pytest.mark.parametrize("seed,n,reso", WELD_CASES)
def test_edge_weld_vs_quantised_weld(qs, orc, seed, n, reso):
    pos, rad = _ball(qs, seed, n)
    G, V, T, Vq, Tq = _welds(orc, pos, rad, reso)
    a, b, pairs, na, nb = _partition_diff(V, T, Vq, Tq)
    split, merged = int((na > 1).sum()), int((nb > 1).sum())
    hq = _half_quantum_axes(G)
    print("D3 seed %d, %d atoms, %.1f A: %d edge-id vertices, %d quantised; split %d (%.2f%%), merged %d (%.2f%%); "
          "half-quantum axes %s" % (seed, n, reso, len(V), len(Vq), split, 100.0 * split / len(V), merged,
                                    100.0 * merged / len(V), hq))
    # the edge-id weld is closed: every edge of the non-degenerate triangles is used an even number
    # of times (twice, or four times where two zero grid points of an exact tangency (the probe a
    # whole number of voxels, as at 0.2 A) share a cube diagonal)
    Tk = T[(T[:, 0] != T[:, 1]) & (T[:, 1] != T[:, 2]) & (T[:, 0] != T[:, 2])]
    e = np.sort(np.concatenate([Tk[:, [0, 1]], Tk[:, [1, 2]], Tk[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert np.all(cnt % 2 == 0)
    print("   edges used 4 times: %d" % int((cnt == 4).sum()))
    if reso != 0.2:
        assert np.all(cnt == 2)
    # merged classes: distinct edge-id vertices within one quantum of each other
    for q in np.nonzero(nb > 1)[0][:2000]:
        vs = pairs[pairs[:, 1] == q, 0]
        assert np.abs(V[vs] - V[vs[0]]).max() <= 0.001 + 1e-6
    # split classes: the keys of one vertex differ by one quantum, only on half-quantum axes
    for v in np.nonzero(na > 1)[0][:2000]:
        ks = Vq[pairs[pairs[:, 0] == v, 1]]
        d = np.abs(ks - ks[0]).max(axis=0)
        assert d.max() <= 0.001 + 1e-5
        assert all(c in hq for c in np.nonzero(d > 0)[0])
    if not hq:
        assert split == 0
    if reso >= 0.5 and not hq:
        assert merged == 0


@pytest.mark.parametrize("reso,n", [(0.2, 300), (0.3, 600), (0.5, 2000), (1.0, 2000)])
def test_gauss_seidel_vs_jacobi(qs, orc, reso, n):
    pos, rad = _ball(qs, 1, n)
    G, V, T, Vq, Tq = _welds(orc, pos, rad, reso)
    a, b, pairs, na, nb = _partition_diff(V, T, Vq, Tq)
    Vj = orc.smooth(V, T, 1, False)  # library: Jacobi, edge-id weld order
    Vg = orc.smooth(Vq, Tq, 1, True)  # reference: in place, quantised weld order
    same = (na[a] == 1) & (nb[b] == 1)
    d = np.linalg.norm(Vg[b].astype(np.float64) - Vj[a].astype(np.float64), axis=1)[same] / reso
    print("D5 %.1f A, %d atoms: Gauss-Seidel vs Jacobi, one pass: max %.3f voxel, mean %.3f voxel (%d vertices)"
          % (reso, n, d.max(), d.mean(), len(np.unique(a[same]))))
    # Jacobi over the quantised weld equals Jacobi over the edge-id weld up to the quantum
    # (the sums differ only in the positions' last 0.001 A)
    Vjq = orc.smooth(Vq, Tq, 1, False)
    dj = np.linalg.norm(Vjq[b].astype(np.float64) - Vj[a].astype(np.float64), axis=1)[same]
    assert dj.max() <= 0.001
    # one in-place pass moves a vertex by less than the umbrella step itself
    assert d.max() < 1.0
