"""Pins of the CPU oracle.

The reference ships no tests, fixtures or golden vectors (its CUDA cannot be built here), so the
oracle is pinned by checks that do not share its code: numpy restatements evaluated with numpy's
IEEE float32 arithmetic, brute force over all atoms, closed forms (an isolated atom is a sphere),
properties the case table and the meshes must have, and double-precision evaluations.  Mutation
tests show that a plausible mistake (swapped axes, a wrong sign, a dropped term) fails a pin.
"""
import numpy as np
import pytest

F = np.float32


def _ball(qs, seed, n):
    return qs.synth_ball(seed, n)


# --------------------------------------------------------------------------------------------
# numpy restatements (float32, one rounding per operation)

def np_grid(pos, rad, reso):
    """CudaSurf.cu:149-174, 250-253, 548-576, 629 with numpy float32 scalars."""
    vmin = np.full(3, F(100000.0))
    vmax = np.full(3, F(-100000.0))
    maxr = F(0.0)
    for p, r in zip(pos, rad):
        vmin = np.minimum(vmin, p)
        vmax = np.maximum(vmax, p)
        maxr = max(maxr, r)
    ext = max(vmax[0] - vmin[0], max(vmax[1] - vmin[1], vmax[2] - vmin[2]))
    max_dist = F(F(ext + F(2) * maxr) + F(4) * F(1.4))
    dxn = F(F(1.4) + maxr)
    origin = (vmin - maxr) - F(1.4)
    gn = int(np.ceil(F(max_dist / dxn)))
    g = int(np.ceil(F(max_dist / F(reso))))
    r = int(np.ceil(F(F(1.4) / F(reso))))
    return dict(origin=origin.astype(F), dxn=dxn, gn=gn, g=g, r=r)


def np_hash(pos, origin, dxn, gn, swap_axes=False):
    c = ((pos - origin) / dxn).astype(np.int32)
    if swap_axes:
        c = c[:, ::-1]
    return (gn * gn * c[:, 0] + gn * c[:, 1] + c[:, 2]).astype(np.int32)


def grid_points(G, axis_vals):
    return (G.origin[None, :] + (np.asarray(axis_vals, F)[:, None] * G.dx)).astype(F)


def np_labels_all_atoms(G, pos, rad, idx3):
    """field pass 1 over ALL atoms (no cells) in float32: Kernels.cu:344-392 predicate."""
    dx = G.dx
    out = np.empty(len(idx3), np.uint8)
    P = (G.origin[None, :] + (idx3.astype(F) * dx)).astype(F)
    for k, p in enumerate(P):
        d2 = (pos - p) * (pos - p)
        d = np.sqrt(F(d2[:, 0] + d2[:, 1]) + d2[:, 2]).astype(F)
        if np.any(d < rad - dx):
            out[k] = 2
        elif np.any(d < F(1.4) + rad):
            out[k] = 1
        else:
            out[k] = 0
    return out


# --------------------------------------------------------------------------------------------
# grid, hashing, order, cells

@pytest.mark.parametrize("seed,n,reso", [(1, 2000, 0.5), (5, 300, 0.2), (6, 50, 1.0)])
def test_grid_setup_matches_numpy(qs, orc, seed, n, reso):
    pos, rad = _ball(qs, seed, n)
    G = orc.grid_setup(pos, rad, reso)
    ref = np_grid(pos, rad, reso)
    assert (G.g, G.gn, G.r) == (ref["g"], ref["gn"], ref["r"])
    assert np.array_equal(G.origin, ref["origin"]) and G.dxn == ref["dxn"]


def test_hash_and_cell_order_match_numpy(qs, orc):
    pos, rad = _ball(qs, 7, 3000)
    G = orc.grid_setup(pos, rad, 0.5)
    h = orc.hash_atoms(G, pos)
    assert np.array_equal(h, np_hash(pos, G.origin, G.dxn, G.gn))
    assert h.min() >= 0 and h.max() < G.gn ** 3
    perm, cs, ce, sx = orc.sort_cells(G, pos, rad, h)
    # stable order by cell id == (cell, index) order (thrust::sort CudaSurf.cu:610-612)
    assert np.array_equal(perm, np.argsort(h, kind="stable"))
    assert np.array_equal(perm, sorted(range(len(h)), key=lambda i: (h[i], i)))
    hs = h[perm]
    for c in np.unique(h)[:200]:
        assert cs[c] == np.searchsorted(hs, c, "left") and ce[c] == np.searchsorted(hs, c, "right")
    empty = np.setdiff1d(np.arange(G.gn ** 3), h)
    assert np.all(cs[empty] == orc.EMPTY) and np.all(ce[empty] == orc.EMPTY)
    assert np.array_equal(sx[:, :3], pos[perm]) and np.array_equal(sx[:, 3], rad[perm])


def test_hash_pin_catches_swapped_axes(qs, orc):
    pos, rad = _ball(qs, 7, 500)
    G = orc.grid_setup(pos, rad, 0.5)
    assert not np.array_equal(orc.hash_atoms(G, pos), np_hash(pos, G.origin, G.dxn, G.gn, swap_axes=True))


# --------------------------------------------------------------------------------------------
# field pass 1

def test_pass1_equals_all_atom_evaluation(qs, orc):
    pos, rad = _ball(qs, 8, 400)
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, 0.5)
    g = G.g
    rng = np.random.default_rng(0)
    idx = rng.integers(0, g - 1, size=(4000, 3))
    ref = np_labels_all_atoms(G, pos, rad, idx)
    got = lab.reshape(g, g, g)[idx[:, 0], idx[:, 1], idx[:, 2]]
    assert np.array_equal(got, ref)
    assert set(np.unique(got)) == {0, 1, 2}
    # last layer of each axis is never written: outside
    L = lab.reshape(g, g, g)
    assert np.all(L[g - 1] == 0) and np.all(L[:, g - 1] == 0) and np.all(L[:, :, g - 1] == 0)


def test_pass1_pin_catches_wrong_threshold_sign(qs, orc):
    """a plausible slip, r + dx instead of r - dx for 'inside', fails the all-atom pin"""
    pos, rad = _ball(qs, 8, 400)
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, 0.5)
    g = G.g
    idx = np.random.default_rng(1).integers(0, g - 1, size=(3000, 3))
    got = lab.reshape(g, g, g)[idx[:, 0], idx[:, 1], idx[:, 2]]
    wrong = np_labels_all_atoms(G, pos, rad + F(2) * G.dx, idx)  # r - dx -> r + dx
    assert not np.array_equal(got, wrong)


def test_pass1_fp64_agrees_except_ties(qs, orc):
    pos, rad = _ball(qs, 9, 500)
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, 0.5)
    lab64, f64 = orc.field_fp64(G, pos, rad)
    diff = np.count_nonzero(lab64 != lab)
    assert diff <= 1e-5 * lab.size + 2, diff


# --------------------------------------------------------------------------------------------
# field pass 2

def np_pass2_voxel(G, L, i, j, k):
    """Kernels.cu:478-545 for one voxel, numpy float32 over the clamped window."""
    g, r, dx = G.g, G.r, G.dx
    xs = np.clip(np.arange(i - r, i + r + 1), 0, g - 1)
    ys = np.clip(np.arange(j - r, j + r + 1), 0, g - 1)
    zs = np.clip(np.arange(k - r, k + r + 1), 0, g - 1)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    out = L[X, Y, Z] == 0
    if not out.any():
        return F(-dx)
    q = np.stack([X[out], Y[out], Z[out]], 1)
    P = (G.origin[None, :] + q.astype(F) * dx).astype(F)
    v = (G.origin + np.array([i, j, k], F) * dx).astype(F)
    d2 = (P - v) * (P - v)
    d = np.sqrt(F(d2[:, 0] + d2[:, 1]) + d2[:, 2]).astype(F)
    return F(F(1.4) - d.min())


def test_pass2_equals_numpy_window(qs, orc):
    pos, rad = _ball(qs, 10, 300)
    for reso in (0.5, 0.2):
        G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, reso)
        g = G.g
        L = lab.reshape(g, g, g)
        Fd = fld.reshape(g, g, g)
        shell = np.argwhere(L == 1)
        pick = shell[np.random.default_rng(2).choice(len(shell), 300, replace=False)]
        for i, j, k in pick:
            assert Fd[i, j, k] == np_pass2_voxel(G, L, i, j, k)
        assert np.all(Fd[L == 0] == F(1.4)) and np.all(Fd[L == 2] == -G.dx)


def test_isolated_atom_field_is_signed_distance(qs, orc):
    """An isolated atom's SES is its sphere: near the surface the field is |p - c| - R."""
    pos = np.array([[0, 0, 0], [40, 0, 0]], F)
    rad = np.array([1.7, 1.7], F)
    reso = 0.2
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, reso)
    g = G.g
    Fd = fld.reshape(g, g, g)
    # refined voxels (an outside voxel in the window) near the surface
    idx = np.argwhere((np.abs(Fd) < 0.6) & (Fd != -G.dx))
    P = G.origin[None, :].astype(np.float64) + idx * float(G.dx)
    d = np.min(np.linalg.norm(P[:, None, :] - pos[None].astype(np.float64), axis=2), axis=1)
    err = np.abs(Fd[idx[:, 0], idx[:, 1], idx[:, 2]] - (d - 1.7))
    # the nearest outside voxel lies within one voxel diagonal of the exact sphere R + probe
    assert err.max() < np.sqrt(3) * reso, err.max()


def test_fp32_field_within_1e4_of_fp64(qs, orc):
    pos, rad = _ball(qs, 11, 400)
    for reso in (0.5, 0.3):
        G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, reso)
        lab64, f64 = orc.field_fp64(G, pos, rad)
        same = lab64 == lab
        assert np.abs(fld[same].astype(np.float64) - f64[same]).max() <= 1e-4


# --------------------------------------------------------------------------------------------
# marching cubes table and meshes

EDGE_CORNERS = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def _table():
    import re
    import os

    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "quickses_amd", "csrc",
                            "qs_mc_table.h")).read()
    vals = [int(v, 16) for v in re.findall(r"0x([0-9a-f]{16})ULL", src)]
    assert len(vals) == 256
    rows = []
    for v in vals:
        e = [(v >> (4 * k)) & 0xF for k in range(16)]
        n = e.index(0xF) if 0xF in e else 16
        rows.append(e[:n])
    return rows


def test_case_table_properties():
    rows = _table()
    assert sum(len(r) for r in rows) // 3 == 820
    for c, r in enumerate(rows):
        assert len(r) % 3 == 0 and len(r) <= 15
        crossing = {e for e, (a, b) in enumerate(EDGE_CORNERS) if ((c >> a) & 1) != ((c >> b) & 1)}
        assert set(r) == crossing, c
        for t in range(0, len(r), 3):
            assert len(set(r[t:t + 3])) == 3
    assert rows[0] == [] and rows[255] == []
    assert max(len(r) for r in rows) == 15


CUBE_FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (3, 2, 6, 7), (0, 3, 7, 4), (1, 2, 6, 5)]


def test_triangle_count_closed_form():
    """The cube-count kernel takes a case's triangle count as |S| - 2 f - 4 h (S the negative
    corners, f the faces with four corners in S, h: the two other corners are a body diagonal),
    evaluated on bit planes that hold 64 cubes each.  Here the planes hold all 256 cases at once
    (bit c = case c) and the same adder and subtraction logic must give the table's count for
    every case but 0 and 255; without the h term exactly the four hexagon cases fail."""
    rows = _table()
    count = [len(r) // 3 for r in rows]
    full = (1 << 256) - 1
    plane = [sum(1 << c for c in range(256) if (c >> i) & 1) for i in range(8)]

    def fadd(a, b, c):
        ab = a ^ b
        return ab ^ c, (a & b) | (c & ab)

    a0, a1 = fadd(plane[0], plane[1], plane[2])
    b0, b1 = fadd(plane[3], plane[4], plane[5])
    d0, d1 = plane[6] ^ plane[7], plane[6] & plane[7]
    s0, e1 = fadd(a0, b0, d0)
    t0, t1 = fadd(a1, b1, d1)
    s1, s2 = t0 ^ e1, t1 ^ (t0 & e1)
    faces = []
    for fc in CUBE_FACES:
        m = full
        for i in fc:
            m &= plane[i]
        faces.append(m)
    u0, u1 = fadd(faces[0], faces[1], faces[2])
    v0, v1 = fadd(faces[3], faces[4], faces[5])
    f0, f1 = u0 ^ v0, u1 ^ v1 ^ (u0 & v0)
    anyface = faces[0] | faces[1] | faces[2] | faces[3] | faces[4] | faces[5]
    h = ~s0 & s1 & s2 & ~anyface & full
    k0, k1 = s0, s1 ^ f0
    k2 = (s2 ^ f1 ^ (~s1 & f0 & full) ^ h) & full
    k2_without_h = (s2 ^ f1 ^ (~s1 & f0 & full)) & full

    def at(p, c):
        return (p >> c) & 1

    wrong_without_h = []
    for c in range(1, 255):
        assert at(k0, c) + 2 * at(k1, c) + 4 * at(k2, c) == count[c], c
        if at(k0, c) + 2 * at(k1, c) + 4 * at(k2_without_h, c) != count[c]:
            wrong_without_h.append(c)
        # the same count from the corner sets directly
        S = [i for i in range(8) if (c >> i) & 1]
        f = sum(all(i in S for i in fc) for fc in CUBE_FACES)
        assert len(S) - 2 * f - 4 * (len(S) == 6 and f == 0) == count[c], c
    assert len(wrong_without_h) == 4
    for c in wrong_without_h:
        off = [i for i in range(8) if not (c >> i) & 1]
        assert sorted(off) in ([0, 6], [1, 7], [2, 4], [3, 5]), c


def _mesh_stats(V, T):
    V = V.astype(np.float64)
    a, b, c = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    cr = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(cr, axis=1).sum()
    vol = np.abs(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)
    e = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), axis=1)
    u, cnt = np.unique(e, axis=0, return_counts=True)
    return area, vol, cnt, len(V) - len(u) + len(T)


def test_isolated_atoms_mesh_are_spheres(qs, orc):
    pos = np.array([[0, 0, 0], [40, 0, 0]], F)
    rad = np.array([1.7, 1.7], F)
    V, T, A = orc.compute_ses(pos, rad, 0.1, 0)
    area, vol, cnt, chi = _mesh_stats(V, T)
    assert np.all(cnt == 2)  # watertight
    assert chi == 4  # two spheres
    R = 1.7
    assert abs(area / (2 * 4 * np.pi * R * R) - 1) < 0.03
    assert abs(vol / (2 * 4 / 3 * np.pi * R ** 3) - 1) < 0.03
    assert set(np.unique(A)) == {0, 1}


def test_config1_mesh_is_watertight_and_welds_like_reference(qs, orc):
    pos, rad = _ball(qs, 1, 2000)
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, 0.5)
    lim = G.g - 1
    V, T = orc.edge_weld(G, lim, fld)
    area, vol, cnt, chi = _mesh_stats(V, T)
    assert np.all(cnt == 2) and chi % 2 == 0
    # reference quantised weld on the reference-form soup: same corners, same partition
    S = orc.soup(G, lim, fld)
    Vq, Tq = orc.quantised_weld(S)
    assert S.shape[0] == 3 * len(T)
    a = T.reshape(-1)
    b = Tq.reshape(-1)
    # the two welds induce the same partition of triangle corners
    m1 = {}
    m2 = {}
    bad = 0
    for x, y in zip(a.tolist(), b.tolist()):
        if m1.setdefault(x, y) != y or m2.setdefault(y, x) != x:
            bad += 1
    assert bad == 0
    # and the same positions up to the 0.001 A quantum
    assert np.abs(Vq[b] - V[a]).max() <= 0.0006


def test_closest_atom_equals_brute_force(qs, orc):
    pos, rad = _ball(qs, 12, 500)
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, 0.5)
    V, T = orc.edge_weld(G, G.g - 1, fld)
    got = orc.closest_atom(G, cs, ce, sx, V[:3000])
    S = sx[:, :3]
    for k in range(0, 3000, 7):
        d2 = (S - V[k]) * (S - V[k])
        d = np.sqrt(F(d2[:, 0] + d2[:, 1]) + d2[:, 2]).astype(F)
        assert d[got[k]] == d.min()


# --------------------------------------------------------------------------------------------
# smoothing

def py_smooth(V, T, gauss_seidel):
    """SmoothMesh.cpp:45-63, 96-124 in Python loops with float32 scalars."""
    nb = [[] for _ in range(len(V))]
    for x, y, z in T.tolist():
        nb[x] += [y, z]
        nb[y] += [x, z]
        nb[z] += [x, y]
    out = V.copy()
    src = out if gauss_seidel else V.copy()
    for i in range(len(V)):
        s = np.zeros(3, F)
        for n in nb[i]:
            s = (s + src[n]).astype(F)
        out[i] = (s / F(len(nb[i]))).astype(F)
    return out


def test_smoothing_matches_python_restatement(qs, orc):
    pos = np.array([[0, 0, 0], [2.0, 0.5, 0]], F)
    rad = np.array([1.7, 1.5], F)
    V, T, A = orc.compute_ses(pos, rad, 0.5, 0)
    for gs in (False, True):
        assert np.array_equal(orc.smooth(V, T, 1, gs), py_smooth(V, T, gs))
    # on a closed mesh every edge is in two triangles, so the multiset mean equals the mean over
    # distinct neighbours; a slip in the divisor (incident triangles instead of pushed corners)
    # is caught
    Vj = orc.smooth(V, T, 1, False)
    acc = np.zeros_like(V, dtype=np.float64)
    ntri = np.zeros(len(V))
    for x, y, z in T.tolist():
        acc[x] += V[y] + V[z]; acc[y] += V[x] + V[z]; acc[z] += V[x] + V[y]
        ntri[x] += 1; ntri[y] += 1; ntri[z] += 1
    assert np.allclose(Vj, acc / (2 * ntri[:, None]), atol=1e-4)
    assert not np.allclose(Vj, acc / ntri[:, None], atol=1e-4)


def test_smoothing_fixes_plane_and_keeps_constant(qs, orc):
    # a flat triangulated grid: Laplacian averaging keeps interior vertices of a regular grid
    n = 6
    xs, ys = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    V = np.stack([xs.ravel(), ys.ravel(), np.zeros(n * n)], 1).astype(F)
    T = []
    for i in range(n - 1):
        for j in range(n - 1):
            a, b, c, d = i * n + j, (i + 1) * n + j, (i + 1) * n + j + 1, i * n + j + 1
            T += [[a, b, c], [a, c, d]]
    T = np.array(T, np.int32)
    S = orc.smooth(V, T, 3, False)
    assert np.all(S[:, 2] == 0)


def test_filter_degenerate():
    from oracle import oracle as orc

    T = np.array([[0, 1, 2], [3, 3, 4], [5, 6, 5], [7, 8, 9], [1, 1, 1]], np.int32)
    assert np.array_equal(orc.filter_degenerate(T), np.array([[0, 1, 2], [7, 8, 9]], np.int32))


# --------------------------------------------------------------------------------------------
# per-vertex normals (no reference routine: the library's definition, pinned by mathematics)

def _two_atoms(d=12.0):
    pos = np.array([[0.0, 0.0, 0.0], [d, 0.3, -0.2]], np.float32)
    rad = np.array([1.7, 1.52], np.float32)
    return pos, rad


def test_normals_exact_on_a_linear_field(orc):
    """Central and one-sided differences are exact on an affine field, so every normal is the
    unit coefficient vector (an axis swap or a sign slip fails this)."""
    pos, rad = _two_atoms()
    G = orc.grid_setup(pos, rad, 0.5)
    g = G.g
    a = np.array([0.3, -0.5, 0.8])
    i, j, k = np.meshgrid(np.arange(g), np.arange(g), np.arange(g), indexing="ij")
    fld = (a[0] * i + a[1] * j + a[2] * k - 0.37 * g).astype(np.float32).ravel()
    mesh = orc.edge_weld(G, g - 1, fld)
    nv = len(mesh[0])
    assert nv > 100
    n = orc.vertex_normals_fp64(G, g - 1, fld, nv)
    # float32 field values: the differences carry their rounding, far below this bound
    assert np.abs(n - a / np.linalg.norm(a)).max() < 1e-5
    assert np.abs(np.linalg.norm(n, axis=1) - 1).max() < 1e-12


def test_normals_radial_on_separate_spheres(orc):
    """Two atoms far apart: the SES is their two spheres, whose outward normals are radial.  The
    field is a distance to the nearest outside voxel, so at 0.25 A (7 voxels per radius) the
    lattice tilts the gradient by a few degrees; an axis swap or a sign slip is tens of degrees
    or more off."""
    pos, rad = _two_atoms()
    G, perm, cs, ce, sx, lab, fld = orc.field(pos, rad, 0.25)
    V, T = orc.edge_weld(G, G.g - 1, fld)[:2]
    n = orc.vertex_normals_fp64(G, G.g - 1, fld, len(V))
    c = pos[np.argmin(((V[:, None, :] - pos[None, :, :]) ** 2).sum(-1), axis=1)].astype(np.float64)
    radial = (V - c) / np.linalg.norm(V - c, axis=1, keepdims=True)
    ang = np.degrees(np.arccos(np.clip((n * radial).sum(1), -1, 1)))
    assert np.median(ang) < 5.0 and ang.max() < 15.0, (np.median(ang), ang.max())
    swapped = n[:, [1, 0, 2]]
    assert np.median(np.degrees(np.arccos(np.clip((swapped * radial).sum(1), -1, 1)))) > 15.0
    # a plausible slip, the gradient with its sign flipped, points inwards
    assert np.median(np.degrees(np.arccos(np.clip((-n * radial).sum(1), -1, 1)))) > 170.0
