"""Accuracy of the single-precision path against the same mathematics in double precision
(the oracle's FP64 field), per voxel size of the benchmark configurations.

The FP64 mesh is marching cubes on the FP64 field (its own sign pattern, vertices interpolated
in double), then one Jacobi pass in double.  The Hausdorff distance between the FP32 and FP64
meshes is taken over samples (vertices and triangle centroids of each mesh) against the other
mesh's triangles.

Two kinds of ties are reported apart, with the error they cause:
  - pass-1 ties: a threshold test that FP32 and FP64 decide differently (an atom distance
    within an ulp of r + probe or r - dx) changes one voxel's label, and with it the pass-2
    value of every voxel whose (2r+1)^3 window holds that voxel;
  - zero ties: grid points where the field is zero up to rounding (|f64| <= 1e-6 A) and the
    two precisions give it different signs.  They come in patches where the surface is tangent
    to a grid plane, e.g. at 0.2 A, where the probe radius is exactly 7 voxels; marching cubes
    of a field that is zero up to rounding over a patch is ill-conditioned there.
Elsewhere (samples more than one voxel from either kind, two after a Laplacian pass):

  field       max |f32 - f64|                                      <= 1e-4 A
  mesh        Hausdorff distance, unsmoothed and after one pass    <= 0.05 voxel
  area/volume relative difference over the whole mesh              <= 0.1 %

The GPU path is bit-identical to the FP32 oracle (tests/test_gpu_parity.py), so these
figures are the GPU path's figures.
"""
import numpy as np
import pytest

CASES = [
    # (label, seed, atoms, voxel): config 1 in full; the voxel sizes of configs 2, 4 and 5 on
    # molecules the FP64 brute force finishes in seconds (the error is local to the surface)
    ("config 1: 2,000 atoms, 0.5 A", 1, 2000, 0.5),
    ("0.2 A (config 2), 300 atoms", 21, 300, 0.2),
    ("0.3 A (config 5), 600 atoms", 22, 600, 0.3),
    ("1.0 A (config 5), 2,000 atoms", 23, 2000, 1.0),
    # config 3's geometry at its voxel size: a shell of the same density (12-20 A instead of
    # 120-187 A), with an inner surface and the packing's cavities
    ("config 3 shape: 1,300-atom shell (12-20 A), 0.5 A", 24, 1300, 0.5),
]


def _molecule(qs, label, seed, n):
    if "shell" in label:
        return qs.synth_shell(seed, n, 12.0, 20.0)
    return qs.synth_ball(seed, n)


def _area_volume(V, T):
    a, b, c = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1).sum()
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0
    return area, abs(vol)


def _jacobi64(V, T):
    """One pass of v <- mean of the neighbour multiset (each triangle adds its two other corners)."""
    s = np.zeros_like(V)
    n = np.zeros(len(V))
    for i, (j, k) in ((0, (1, 2)), (1, (0, 2)), (2, (0, 1))):
        np.add.at(s, T[:, i], V[T[:, j]] + V[T[:, k]])
        np.add.at(n, T[:, i], 2.0)
    out = V.copy()
    m = n > 0
    out[m] = s[m] / n[m, None]
    return out


def _point_triangle_distance(P, A, B, C):
    """Distance from points P to triangles (A, B, C), row by row (closest point on a triangle,
    region by region)."""
    ab, ac, ap = B - A, C - A, P - A
    d1, d2 = np.einsum("ij,ij->i", ab, ap), np.einsum("ij,ij->i", ac, ap)
    bp = P - B
    d3, d4 = np.einsum("ij,ij->i", ab, bp), np.einsum("ij,ij->i", ac, bp)
    cp = P - C
    d5, d6 = np.einsum("ij,ij->i", ab, cp), np.einsum("ij,ij->i", ac, cp)
    va, vb, vc = d3 * d6 - d5 * d4, d5 * d2 - d1 * d6, d1 * d4 - d3 * d2
    with np.errstate(divide="ignore", invalid="ignore"):
        denom = va + vb + vc
        v = np.where(denom != 0, vb / denom, 0.0)
        w = np.where(denom != 0, vc / denom, 0.0)
        Q = A + ab * v[:, None] + ac * w[:, None]  # interior
        t_ab = np.clip(np.where(d1 - d3 != 0, d1 / (d1 - d3), 0.0), 0, 1)
        t_ac = np.clip(np.where(d2 - d6 != 0, d2 / (d2 - d6), 0.0), 0, 1)
        t_bc = np.clip(np.where((d4 - d3) + (d5 - d6) != 0, (d4 - d3) / ((d4 - d3) + (d5 - d6)), 0.0), 0, 1)
    Q = np.where(((va <= 0) & (d4 - d3 >= 0) & (d5 - d6 >= 0))[:, None], B + (C - B) * t_bc[:, None], Q)
    Q = np.where(((vb <= 0) & (d2 >= 0) & (d6 <= 0))[:, None], A + ac * t_ac[:, None], Q)
    Q = np.where(((vc <= 0) & (d1 >= 0) & (d3 <= 0))[:, None], A + ab * t_ab[:, None], Q)
    Q = np.where(((d6 >= 0) & (d5 <= d6))[:, None], C, Q)
    Q = np.where(((d3 >= 0) & (d4 <= d3))[:, None], B, Q)
    Q = np.where(((d1 <= 0) & (d2 <= 0))[:, None], A, Q)
    return np.linalg.norm(P - Q, axis=1)


def _samples(V, T):
    return np.concatenate([V, (V[T[:, 0]] + V[T[:, 1]] + V[T[:, 2]]) / 3.0])


def _dist_to_mesh(P, V, T, k=12):
    from scipy.spatial import cKDTree

    C = (V[T[:, 0]] + V[T[:, 1]] + V[T[:, 2]]) / 3.0
    _, idx = cKDTree(C).query(P, k=k)
    best = np.full(len(P), np.inf)
    for j in range(k):
        t = T[idx[:, j]]
        best = np.minimum(best, _point_triangle_distance(P, V[t[:, 0]], V[t[:, 1]], V[t[:, 2]]))
    return best


@pytest.mark.parametrize("label,seed,n,reso", CASES)
def test_fp32_path_against_fp64(qs, orc, label, seed, n, reso):
    from scipy.ndimage import maximum_filter

    pos, rad = _molecule(qs, label, seed, n)
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, reso)
    lab64, f64 = orc.field_fp64(G, pos, rad)
    g, r = G.g, G.r
    ties = (lab64 != lab).reshape(g, g, g)
    affected = maximum_filter(ties, size=2 * r + 1, mode="constant", cval=False)
    ok = ~affected.ravel()
    field_err = float(np.abs(fld[ok].astype(np.float64) - f64[ok]).max())
    field_err_all = float(np.abs(fld.astype(np.float64) - f64).max())
    lim = g - 1
    # FP32 mesh (the path's), FP64 mesh (its own sign pattern, double interpolation)
    V32, T32 = orc.edge_weld(G, lim, fld)
    f64_as32 = f64.astype(np.float32)
    assert np.array_equal(f64_as32 < 0, f64 < 0) and np.array_equal(f64_as32 == 0, f64 == 0)
    _, T64 = orc.edge_weld(G, lim, f64_as32)
    V64, _ = orc.vertices_fp64(G, lim, f64_as32, f64, int(T64.max()) + 1 if len(T64) else 0)
    V32 = V32.astype(np.float64)
    S32 = orc.smooth(V32.astype(np.float32), T32, 1, False).astype(np.float64)
    S64 = _jacobi64(V64, T64)
    # samples near a tie are reported apart
    zero_ties = ((fld < 0) != (f64 < 0)) & (np.abs(f64) <= 1e-6)
    near = maximum_filter(affected | zero_ties.reshape(g, g, g), size=3, mode="constant", cval=False)

    # a smoothed vertex averages its neighbours: one more voxel of margin after a pass
    near_s = maximum_filter(near, size=3, mode="constant", cval=False)

    def hausdorff(Va, Ta, Vb, Tb, near):
        Pa, Pb = _samples(Va, Ta), _samples(Vb, Tb)
        da, db = _dist_to_mesh(Pa, Vb, Tb), _dist_to_mesh(Pb, Va, Ta)

        def clear(P):
            ijk = np.clip(np.floor((P - np.asarray(G.origin, np.float64)) / reso).astype(np.int64), 0, g - 1)
            return ~near[ijk[:, 0], ijk[:, 1], ijk[:, 2]]

        ca, cb = clear(Pa), clear(Pb)
        return max(da[ca].max(), db[cb].max()) / reso, max(da.max(), db.max()) / reso

    h0, h0_all = hausdorff(V32, T32, V64, T64, near)
    h1, h1_all = hausdorff(S32, T32, S64, T64, near_s)
    k32, k64 = orc.filter_degenerate(T32), orc.filter_degenerate(T64)
    a32, v32 = _area_volume(S32, k32)
    a64, v64 = _area_volume(S64, k64)
    flips = (fld < 0) != (f64 < 0)
    print("\n%s: g=%d; pass-1 ties %d (voxels whose window holds one: %d of %d, field error there up to %.2e A); "
          "zero ties %d; other sign differences %d; field max|f32-f64| %.2e A; Hausdorff %.2e voxel unsmoothed, "
          "%.2e after one pass (with the tie neighbourhoods: %.2e, %.2e); smoothed area %.2e, volume %.2e relative"
          % (label, g, int(ties.sum()), int(affected.sum()), affected.size, field_err_all, int(zero_ties.sum()),
             int((flips & ~zero_ties & ok).sum()), field_err, h0, h1, h0_all, h1_all, abs(a32 - a64) / a64,
             abs(v32 - v64) / v64))
    assert not (flips & ~zero_ties & ok).any()
    assert field_err <= 1e-4
    assert h0 <= 0.05 and h1 <= 0.05
    assert abs(a32 - a64) / a64 <= 1e-3 and abs(v32 - v64) / v64 <= 1e-3
