"""ctypes binding of the CPU oracle (libses_oracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this module.
Each function restates a reference function; the citations are in ses_oracle.h / ses_oracle.c.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libses_oracle.so")
_lib = None

OUT, SHELL, IN = 0, 1, 2
EMPTY = 2147483647 - 1
PROBE = np.float32(1.4)


class _Grid(ctypes.Structure):
    _fields_ = [("n_atoms", ctypes.c_int), ("reso", ctypes.c_float), ("minv", ctypes.c_float * 3),
                ("maxv", ctypes.c_float * 3), ("max_rad", ctypes.c_float), ("max_dist", ctypes.c_float),
                ("dx_n", ctypes.c_float), ("origin", ctypes.c_float * 3), ("gn", ctypes.c_int), ("g", ctypes.c_int),
                ("r", ctypes.c_int)]


class _Mesh(ctypes.Structure):
    _fields_ = [("n_vert", ctypes.c_int64), ("n_tri", ctypes.c_int64), ("vert", ctypes.POINTER(ctypes.c_float)),
                ("tri", ctypes.POINTER(ctypes.c_int)), ("atom_id", ctypes.POINTER(ctypes.c_int))]


class _Opt(ctypes.Structure):
    _fields_ = [("weld", ctypes.c_int), ("smooth_gs", ctypes.c_int), ("lim_mode", ctypes.c_int),
                ("atom_ids_sorted", ctypes.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built (make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        _lib.or_grid_setup.argtypes = [vp, vp, ctypes.c_int, ctypes.c_float, ctypes.POINTER(_Grid)]
        _lib.or_grid_setup.restype = ctypes.c_int
        _lib.or_hash.argtypes = [ctypes.POINTER(_Grid), vp, ctypes.c_int, vp]
        _lib.or_sort_cells.argtypes = [ctypes.POINTER(_Grid), vp, vp, ctypes.c_int, vp, vp, vp, vp, vp]
        _lib.or_pass1.argtypes = [ctypes.POINTER(_Grid), vp, vp, vp, vp]
        _lib.or_pass2.argtypes = [ctypes.POINTER(_Grid), vp, vp]
        _lib.or_mc_classify.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp]
        _lib.or_mc_edge_weld.argtypes = [ctypes.POINTER(_Grid), ctypes.c_int, vp, ctypes.POINTER(_Mesh)]
        _lib.or_mc_soup.argtypes = [ctypes.POINTER(_Grid), ctypes.c_int, vp, ctypes.POINTER(ctypes.POINTER(ctypes.c_float)),
                                    ctypes.POINTER(ctypes.c_int64)]
        _lib.or_quantised_weld.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(_Mesh)]
        _lib.or_closest_atom.argtypes = [ctypes.POINTER(_Grid), vp, vp, vp, vp, ctypes.c_int64, vp]
        _lib.or_smooth.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        _lib.or_filter_degenerate.argtypes = [vp, ctypes.c_int64]
        _lib.or_filter_degenerate.restype = ctypes.c_int64
        _lib.or_compute_ses.argtypes = [vp, vp, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.POINTER(_Opt),
                                        ctypes.POINTER(_Mesh)]
        _lib.or_field_fp64.argtypes = [ctypes.POINTER(_Grid), vp, vp, ctypes.c_int, vp, vp]
        _lib.or_mc_vertices_fp64.argtypes = [ctypes.POINTER(_Grid), ctypes.c_int, vp, vp, vp, vp]
        _lib.or_mc_vertices_fp64.restype = ctypes.c_int64
        _lib.or_vertex_normals_fp64.argtypes = [ctypes.POINTER(_Grid), ctypes.c_int, vp, vp]
        _lib.or_vertex_normals_fp64.restype = ctypes.c_int64
        _lib.or_mesh_free.argtypes = [ctypes.POINTER(_Mesh)]
        _lib.or_num_threads.restype = ctypes.c_int
        _lib.or_bench_slab.argtypes = [vp, vp, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        _lib.or_bench_slab.restype = ctypes.c_double
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class Grid:
    def __init__(self, s):
        self._s = s
        self.g, self.gn, self.r = s.g, s.gn, s.r
        self.origin = np.array(list(s.origin), dtype=np.float32)
        self.dx = np.float32(s.reso)
        self.dxn = np.float32(s.dx_n)
        self.max_dist = np.float32(s.max_dist)
        self.max_rad = np.float32(s.max_rad)

    def params(self, lim):
        return dict(origin=tuple(float(x) for x in self.origin), dx=float(self.dx), dxn=float(self.dxn), g=self.g,
                    gn=self.gn, r=self.r, lim=lim)


def grid_setup(pos, rad, reso):
    pos = np.ascontiguousarray(pos, np.float32)
    rad = np.ascontiguousarray(rad, np.float32)
    s = _Grid()
    rc = lib().or_grid_setup(_p(pos), _p(rad), len(rad), ctypes.c_float(reso), ctypes.byref(s))
    return Grid(s) if rc == 0 else None


def hash_atoms(G, pos):
    pos = np.ascontiguousarray(pos, np.float32)
    out = np.empty(len(pos), np.int32)
    lib().or_hash(ctypes.byref(G._s), _p(pos), len(pos), _p(out))
    return out


def sort_cells(G, pos, rad, hashes):
    pos = np.ascontiguousarray(pos, np.float32)
    rad = np.ascontiguousarray(rad, np.float32)
    n = len(rad)
    nc = G.gn ** 3
    perm = np.empty(n, np.int32)
    cs = np.empty(nc, np.int32)
    ce = np.empty(nc, np.int32)
    sx = np.empty((n, 4), np.float32)
    lib().or_sort_cells(ctypes.byref(G._s), _p(pos), _p(rad), n, _p(np.ascontiguousarray(hashes, np.int32)), _p(perm),
                        _p(cs), _p(ce), _p(sx))
    return perm, cs, ce, sx


def pass1(G, cs, ce, sx):
    lab = np.empty(G.g ** 3, np.uint8)
    lib().or_pass1(ctypes.byref(G._s), _p(cs), _p(ce), _p(np.ascontiguousarray(sx, np.float32)), _p(lab))
    return lab


def pass2(G, labels):
    f = np.empty(G.g ** 3, np.float32)
    lib().or_pass2(ctypes.byref(G._s), _p(np.ascontiguousarray(labels, np.uint8)), _p(f))
    return f


def field(pos, rad, reso):
    """(grid, perm, cs, ce, sx, labels, field) for a molecule."""
    G = grid_setup(pos, rad, reso)
    h = hash_atoms(G, pos)
    perm, cs, ce, sx = sort_cells(G, pos, rad, h)
    lab = pass1(G, cs, ce, sx)
    f = pass2(G, lab)
    return G, perm, cs, ce, sx, lab, f


def mc_classify(g, lim, fld):
    out = np.empty(g ** 3, np.uint8)
    lib().or_mc_classify(g, lim, _p(np.ascontiguousarray(fld, np.float32)), _p(out))
    return out


def _take_mesh(m):
    nv, nt = m.n_vert, m.n_tri
    V = np.ctypeslib.as_array(m.vert, shape=(max(nv, 1) * 3,))[: 3 * nv].reshape(-1, 3).copy() if nv else np.zeros((0, 3), np.float32)
    T = np.ctypeslib.as_array(m.tri, shape=(max(nt, 1) * 3,))[: 3 * nt].reshape(-1, 3).copy() if nt else np.zeros((0, 3), np.int32)
    A = None
    if m.atom_id:
        A = np.ctypeslib.as_array(m.atom_id, shape=(max(nv, 1),))[:nv].copy()
    lib().or_mesh_free(ctypes.byref(m))
    return V, T, A


def edge_weld(G, lim, fld):
    m = _Mesh()
    lib().or_mc_edge_weld(ctypes.byref(G._s), lim, _p(np.ascontiguousarray(fld, np.float32)), ctypes.byref(m))
    V, T, _ = _take_mesh(m)
    return V, T


def soup(G, lim, fld):
    sp = ctypes.POINTER(ctypes.c_float)()
    n = ctypes.c_int64(0)
    lib().or_mc_soup(ctypes.byref(G._s), lim, _p(np.ascontiguousarray(fld, np.float32)), ctypes.byref(sp), ctypes.byref(n))
    out = np.ctypeslib.as_array(sp, shape=(max(n.value, 1) * 3,))[: 3 * n.value].reshape(-1, 3).copy()
    ctypes.CDLL(None).free(sp)
    return out


def quantised_weld(soup_verts):
    s = np.ascontiguousarray(soup_verts, np.float32)
    m = _Mesh()
    lib().or_quantised_weld(_p(s), len(s), ctypes.byref(m))
    V, T, _ = _take_mesh(m)
    return V, T


def closest_atom(G, cs, ce, sx, verts):
    v = np.ascontiguousarray(verts, np.float32)
    out = np.empty(len(v), np.int32)
    lib().or_closest_atom(ctypes.byref(G._s), _p(cs), _p(ce), _p(np.ascontiguousarray(sx, np.float32)), _p(v), len(v), _p(out))
    return out


def smooth(verts, tris, times, gauss_seidel):
    v = np.ascontiguousarray(verts, np.float32).copy()
    t = np.ascontiguousarray(tris, np.int32)
    lib().or_smooth(_p(v), len(v), _p(t), len(t), int(times), int(gauss_seidel))
    return v


def filter_degenerate(tris):
    t = np.ascontiguousarray(tris, np.int32).copy()
    n = lib().or_filter_degenerate(_p(t), len(t))
    return t[:n]


def compute_ses(pos, rad, reso, smooth_times=1, weld="edge", smooth_gs=False, lim_mode=0, atom_ids_sorted=False):
    """Whole pipeline; returns (vertices, triangles, atom_ids) like API_computeSES + getters."""
    pos = np.ascontiguousarray(pos, np.float32)
    rad = np.ascontiguousarray(rad, np.float32)
    o = _Opt(0 if weld == "edge" else 1, int(smooth_gs), int(lim_mode), int(atom_ids_sorted))
    m = _Mesh()
    lib().or_compute_ses(_p(pos), _p(rad), len(rad), ctypes.c_float(reso), int(smooth_times), ctypes.byref(o), ctypes.byref(m))
    return _take_mesh(m)


def field_fp64(G, pos, rad):
    pos = np.ascontiguousarray(pos, np.float32)
    rad = np.ascontiguousarray(rad, np.float32)
    lab = np.empty(G.g ** 3, np.uint8)
    f = np.empty(G.g ** 3, np.float64)
    lib().or_field_fp64(ctypes.byref(G._s), _p(pos), _p(rad), len(rad), _p(lab), _p(f))
    return lab, f


def vertices_fp64(G, lim, fld, f64, nvert):
    """(FP64 positions, grid point) of the vertices of edge_weld(G, lim, fld), placed on their
    edges by the FP64 field f64."""
    out = np.empty((nvert, 3), np.float64)
    point = np.empty(nvert, np.int64)
    n = lib().or_mc_vertices_fp64(ctypes.byref(G._s), lim, _p(np.ascontiguousarray(fld, np.float32)),
                                  _p(np.ascontiguousarray(f64, np.float64)), _p(out), _p(point))
    assert n == nvert
    return out, point


def vertex_normals_fp64(G, lim, fld, nvert):
    """FP64 unit field gradients at the vertices of edge_weld(G, lim, fld), same order."""
    out = np.empty((nvert, 3), np.float64)
    n = lib().or_vertex_normals_fp64(ctypes.byref(G._s), lim, _p(np.ascontiguousarray(fld, np.float32)), _p(out))
    assert n == nvert
    return out


def num_threads():
    return lib().or_num_threads()


def bench_slab(pos, rad, reso, planes):
    """Seconds for the whole path (setup, field passes, soup, quantised weld, closest atom, one
    in-place Laplacian pass, filter) on a slice of `planes` middle x-planes; returns
    (seconds, g, voxels, triangles kept)."""
    pos = np.ascontiguousarray(pos, np.float32)
    rad = np.ascontiguousarray(rad, np.float32)
    g = ctypes.c_int(0)
    v = ctypes.c_int64(0)
    nt = ctypes.c_int64(0)
    t = lib().or_bench_slab(_p(pos), _p(rad), len(rad), ctypes.c_float(reso), int(planes), ctypes.byref(g), ctypes.byref(v),
                            ctypes.byref(nt))
    return t, g.value, v.value, nt.value
