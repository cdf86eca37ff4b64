"""quickses_amd: SES meshes on AMD Instinct MI355X.

Python mirror of the reference's C ABI (CudaSurf.h:75-83) over libquickses.so.  The library
holds every kernel; this module only marshals arrays.  It raises if the library is missing:
there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libquickses.so")
_lib = None

PROBE_RADIUS = 1.4
ELEMENT_RADIUS = {"O": 1.52, "C": 1.70, "N": 1.55, "H": 1.20, "S": 1.80, "P": 1.80, "X": 1.40}  # CudaSurf.h:64-73


class QuickSESError(RuntimeError):
    pass


def lib():
    """Load libquickses.so (built by quickses_amd.build); raises if it is not there."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QuickSESError("libquickses.so not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    f3p = ctypes.POINTER(ctypes.c_float)
    u32p = ctypes.POINTER(ctypes.c_uint)
    L.API_computeSES.argtypes = [ctypes.c_float, f3p, f3p, ctypes.c_uint, f3p, u32p, ctypes.POINTER(ctypes.c_int),
                                 u32p, ctypes.c_int]
    L.API_computeSES.restype = None
    L.API_getTriangles.argtypes = [ctypes.c_bool]
    L.API_getTriangles.restype = ctypes.POINTER(ctypes.c_int)
    L.API_getVertices.argtypes = []
    L.API_getVertices.restype = ctypes.POINTER(ctypes.c_float)
    L.API_getAtomIdPerVert.argtypes = []
    L.API_getAtomIdPerVert.restype = ctypes.POINTER(ctypes.c_int)
    L.API_freeMesh.argtypes = []
    L.API_freeMesh.restype = None
    L.API_lastError.restype = ctypes.c_int
    L.qs_last_error_message.restype = ctypes.c_char_p
    L.qs_set_option.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.qs_set_option.restype = ctypes.c_int
    L.qs_get_option.argtypes = [ctypes.c_char_p]
    L.qs_get_option.restype = ctypes.c_int
    vp = ctypes.c_void_p
    L.qs_compute_device.argtypes = [vp, vp, ctypes.c_uint, ctypes.c_float, ctypes.c_int, u32p, u32p,
                                    ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.qs_compute_device.restype = ctypes.c_int
    L.qs_compute_device_noalloc.argtypes = [vp, vp, ctypes.c_uint, ctypes.c_float, ctypes.c_int, vp, ctypes.c_uint, vp,
                                            ctypes.c_uint, vp, u32p, u32p]
    L.qs_compute_device_noalloc.restype = ctypes.c_int
    L.qs_compute_batch_device.argtypes = [vp, vp, u32p, ctypes.c_int, f3p, ctypes.c_int, u32p, u32p,
                                          ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.qs_compute_batch_device.restype = ctypes.c_int
    L.qs_compute_batch.argtypes = [f3p, f3p, u32p, ctypes.c_int, f3p, ctypes.c_int, u32p, u32p]
    L.qs_compute_batch.restype = ctypes.c_int
    L.qs_batch_mesh.argtypes = [ctypes.c_int, vp, vp, vp]
    L.qs_batch_mesh.restype = ctypes.c_int
    L.qs_batch_fetch.argtypes = [vp, vp, vp]
    L.qs_batch_fetch.restype = ctypes.c_int
    L.qs_batch_fetch_host.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.qs_batch_fetch_host.restype = ctypes.c_int
    L.qs_batch_fetch_host_async.argtypes = [ctypes.c_int]
    L.qs_batch_fetch_host_async.restype = ctypes.c_int
    L.qs_batch_wait_host.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), u32p, u32p]
    L.qs_batch_wait_host.restype = ctypes.c_int
    L.qs_batch_offsets.argtypes = [u32p, u32p]
    L.qs_batch_offsets.restype = ctypes.c_int
    L.qs_copy_device.argtypes = [vp, vp, ctypes.c_uint64, f3p]
    L.qs_copy_device.restype = ctypes.c_int
    L.qs_last_grid.argtypes = [f3p, ctypes.POINTER(ctypes.c_int)]
    L.qs_last_slab_count.argtypes = []
    L.qs_last_slab_count.restype = ctypes.c_int
    L.qs_set_tuning.argtypes = [ctypes.c_char_p, ctypes.c_float, ctypes.c_int]
    L.qs_set_tuning.restype = ctypes.c_int
    L.qs_load_tuning.argtypes = [ctypes.c_char_p]
    L.qs_load_tuning.restype = ctypes.c_int
    L.qs_batch_grid.argtypes = [ctypes.c_int, f3p, ctypes.POINTER(ctypes.c_int)]
    L.qs_batch_grid.restype = ctypes.c_int
    L.qs_stage_times.argtypes = [ctypes.c_char_p, ctypes.c_int, f3p, ctypes.c_int]
    L.qs_stage_times.restype = ctypes.c_int
    L.qs_fetch.argtypes = [ctypes.c_char_p, vp, ctypes.c_int64]
    L.qs_fetch.restype = ctypes.c_int64
    L.qs_stage_pass2.argtypes = [f3p, ctypes.POINTER(ctypes.c_int), vp, vp]
    L.qs_stage_pass2.restype = ctypes.c_int
    L.qs_stage_mesh.argtypes = [f3p, ctypes.POINTER(ctypes.c_int), vp, u32p, u32p, vp, ctypes.c_int64, vp,
                                ctypes.c_int64]
    L.qs_stage_mesh.restype = ctypes.c_int
    L.qs_pdb_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(f3p), ctypes.POINTER(f3p), u32p]
    L.qs_pdb_read.restype = ctypes.c_int
    L.qs_obj_write.argtypes = [ctypes.c_char_p, vp, ctypes.c_uint, vp, ctypes.c_uint]
    L.qs_obj_write.restype = ctypes.c_int
    L.qs_io_error.restype = ctypes.c_char_p
    L.qs_free.argtypes = [vp]
    L.qs_free.restype = None
    i64 = ctypes.c_int64
    L.qs_serve_open.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
    L.qs_serve_open.restype = ctypes.c_int
    L.qs_serve_submit.argtypes = [f3p, f3p, ctypes.c_uint, ctypes.c_float, ctypes.c_int]
    L.qs_serve_submit.restype = i64
    L.qs_serve_poll.argtypes = [i64, ctypes.c_int, u32p, u32p, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp),
                                ctypes.POINTER(ctypes.c_double)]
    L.qs_serve_poll.restype = ctypes.c_int
    L.qs_serve_fetch.argtypes = [i64, vp, vp, vp]
    L.qs_serve_fetch.restype = ctypes.c_int
    L.qs_serve_release.argtypes = [i64]
    L.qs_serve_release.restype = ctypes.c_int
    L.qs_serve_set.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
    L.qs_serve_set.restype = ctypes.c_int
    L.qs_serve_stats.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
    L.qs_serve_stats.restype = ctypes.c_int
    L.qs_vertex_normals.argtypes = [vp, ctypes.c_uint]
    L.qs_vertex_normals.restype = ctypes.c_int
    L.qs_vertex_normals_device.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.qs_vertex_normals_device.restype = ctypes.c_int
    L.qs_vertex_colors.argtypes = [vp, ctypes.c_uint, vp, ctypes.c_uint]
    L.qs_vertex_colors.restype = ctypes.c_int
    L.qs_serve_close.argtypes = []
    L.qs_serve_close.restype = ctypes.c_int
    for name in ("qs_synth_ball", "qs_synth_shell"):
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
    L.qs_synth_ball.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_double, vp, vp]
    L.qs_synth_shell.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 vp, vp]
    L.qs_synth_jitter.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_double, vp, vp]
    L.qs_synth_jitter.restype = ctypes.c_int
    L.qs_synth_sizes.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_double, vp]
    L.qs_synth_sizes.restype = ctypes.c_int
    _lib = L
    return L


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _check(rc=None):
    L = lib()
    err = L.API_lastError()
    if err or (rc is not None and rc != 0):
        raise QuickSESError("quickses error %d: %s" % (err or rc, L.qs_last_error_message().decode()))


def set_option(key, value):
    if lib().qs_set_option(key.encode(), int(value)) != 0:
        raise KeyError(key)


def get_option(key):
    return lib().qs_get_option(key.encode())


def compute_ses(pos, rad, reso=0.5, smooth=1, invert_triangles=False, copy=True):
    """API_computeSES + getters + API_freeMesh (CudaSurf.cu:787-872).

    pos: (N, 3) float32, rad: (N,) float32.  Returns (vertices (NV, 3) float32,
    triangles (NT, 3) int32, atom_ids (NV,) int32).  copy=False returns views of the library's
    host arrays instead of copies, without API_freeMesh: valid until the next computation."""
    L = lib()
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    rad = np.ascontiguousarray(rad, dtype=np.float32).reshape(-1)
    if pos.shape[0] != rad.shape[0]:
        raise ValueError("pos and rad disagree on the atom count")
    nv = ctypes.c_uint(0)
    nt = ctypes.c_uint(0)
    L.API_computeSES(ctypes.c_float(reso), _fptr(pos), _fptr(rad), pos.shape[0], None, ctypes.byref(nv), None,
                     ctypes.byref(nt), int(smooth))
    err = L.API_lastError()
    try:
        if err:
            raise QuickSESError("quickses error %d: %s" % (err, L.qs_last_error_message().decode()))
        tp = L.API_getTriangles(bool(invert_triangles))
        vp = L.API_getVertices()
        ap = L.API_getAtomIdPerVert()
        V = np.ctypeslib.as_array(vp, shape=(max(nv.value, 1) * 3,))[: nv.value * 3].reshape(-1, 3)
        T = np.ctypeslib.as_array(tp, shape=(max(nt.value, 1),))[: nt.value].reshape(-1, 3)
        A = np.ctypeslib.as_array(ap, shape=(max(nv.value, 1),))[: nv.value]
        if not copy:
            return V, T, A
        V, T, A = V.copy(), T.copy(), A.copy()
    finally:
        if copy or err:
            L.API_freeMesh()
    return V, T, A


def compute_device(d_pos, d_rad, n, reso=0.5, smooth=1):
    """qs_compute_device on device pointers; returns (nvert, nidx, d_vert, d_tri, d_atom) with
    device pointers owned by the library (valid until the next call)."""
    L = lib()
    nv = ctypes.c_uint(0)
    ni = ctypes.c_uint(0)
    pv, pt, pa = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    rc = L.qs_compute_device(ctypes.c_void_p(d_pos), ctypes.c_void_p(d_rad), int(n), ctypes.c_float(reso), int(smooth),
                             ctypes.byref(nv), ctypes.byref(ni), ctypes.byref(pv), ctypes.byref(pt), ctypes.byref(pa))
    if rc:
        raise QuickSESError("quickses error %d: %s" % (rc, L.qs_last_error_message().decode()))
    return nv.value, ni.value, pv.value, pt.value, pa.value


def compute_device_noalloc(d_pos, d_rad, n, reso, smooth, d_vert, vert_cap, d_tri, tri_cap, d_atom):
    """qs_compute_device_noalloc: the mesh into the caller's device buffers (capacities in
    vertices and triangles), no device memory allocated; returns (nvert, nidx)."""
    L = lib()
    nv = ctypes.c_uint(0)
    ni = ctypes.c_uint(0)
    rc = L.qs_compute_device_noalloc(ctypes.c_void_p(d_pos), ctypes.c_void_p(d_rad), int(n), ctypes.c_float(reso),
                                     int(smooth), ctypes.c_void_p(d_vert), int(vert_cap), ctypes.c_void_p(d_tri),
                                     int(tri_cap), ctypes.c_void_p(d_atom), ctypes.byref(nv), ctypes.byref(ni))
    if rc:
        raise QuickSESError("quickses error %d: %s" % (rc, L.qs_last_error_message().decode()))
    return nv.value, ni.value


def _batch_desc(counts, reso):
    nmol = len(counts)
    off = np.zeros(nmol + 1, dtype=np.uint32)
    off[1:] = np.cumsum(np.asarray(counts, dtype=np.uint64))
    if np.ndim(reso) == 0:
        reso = np.full(nmol, reso, dtype=np.float32)
    reso = np.ascontiguousarray(reso, dtype=np.float32)
    if reso.shape != (nmol,):
        raise ValueError("one voxel size per molecule")
    return off, reso


def compute_batch(molecules, reso=0.5, smooth=1, copy=True):
    """qs_compute_batch + qs_batch_fetch: molecules is a list of (pos (N,3), rad (N,)) pairs,
    reso a voxel size or one per molecule.  Returns a list of (vertices, triangles, atom_ids),
    each equal to compute_ses on that molecule alone.  copy=False returns views of the library's
    pinned host arrays (qs_batch_fetch_host), valid until the next batch fetch."""
    L = lib()
    pos = [np.ascontiguousarray(p, dtype=np.float32).reshape(-1, 3) for p, _ in molecules]
    rad = [np.ascontiguousarray(r, dtype=np.float32).reshape(-1) for _, r in molecules]
    for p, r in zip(pos, rad):
        if p.shape[0] != r.shape[0]:
            raise ValueError("pos and rad disagree on the atom count")
    off, reso = _batch_desc([p.shape[0] for p in pos], reso)
    nmol = len(pos)
    P = np.ascontiguousarray(np.concatenate(pos) if nmol else np.zeros((0, 3), np.float32))
    R = np.ascontiguousarray(np.concatenate(rad) if nmol else np.zeros(0, np.float32))
    nv = np.zeros(max(nmol, 1), dtype=np.uint32)
    nt = np.zeros(max(nmol, 1), dtype=np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint)
    rc = L.qs_compute_batch(_fptr(P), _fptr(R), off.ctypes.data_as(u32p), nmol, _fptr(reso), int(smooth),
                            nv.ctypes.data_as(u32p), nt.ctypes.data_as(u32p))
    _check(rc)
    vo = np.zeros(nmol + 1, dtype=np.int64)
    to = np.zeros(nmol + 1, dtype=np.int64)
    vo[1:] = np.cumsum(nv[:nmol])
    to[1:] = np.cumsum(nt[:nmol])
    if not copy:
        pv, pt, pa = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(L.qs_batch_fetch_host(ctypes.byref(pv), ctypes.byref(pt), ctypes.byref(pa)))
        nvt, ntt = int(vo[-1]), int(to[-1])
        V = np.ctypeslib.as_array(ctypes.cast(pv, ctypes.POINTER(ctypes.c_float)), shape=(max(nvt, 1) * 3,))[: 3 * nvt]
        T = np.ctypeslib.as_array(ctypes.cast(pt, ctypes.POINTER(ctypes.c_int)), shape=(max(ntt, 1) * 3,))[: 3 * ntt]
        A = np.ctypeslib.as_array(ctypes.cast(pa, ctypes.POINTER(ctypes.c_int)), shape=(max(nvt, 1),))[:nvt]
        V, T = V.reshape(-1, 3), T.reshape(-1, 3)
        return [(V[vo[m]:vo[m + 1]], T[to[m]:to[m + 1]], A[vo[m]:vo[m + 1]]) for m in range(nmol)]
    V = np.empty((int(vo[-1]), 3), dtype=np.float32)
    T = np.empty((int(to[-1]), 3), dtype=np.int32)
    A = np.empty(int(vo[-1]), dtype=np.int32)
    rc = L.qs_batch_fetch(ctypes.c_void_p(V.ctypes.data), ctypes.c_void_p(T.ctypes.data), ctypes.c_void_p(A.ctypes.data))
    _check(rc)
    return [(V[vo[m]:vo[m + 1]], T[to[m]:to[m + 1]], A[vo[m]:vo[m + 1]]) for m in range(nmol)]


def _run_batch(L, molecules, reso, smooth):
    """qs_compute_batch on host arrays; returns the per-molecule vertex and triangle offsets"""
    pos = [np.ascontiguousarray(p, dtype=np.float32).reshape(-1, 3) for p, _ in molecules]
    rad = [np.ascontiguousarray(r, dtype=np.float32).reshape(-1) for _, r in molecules]
    for p, r in zip(pos, rad):
        if p.shape[0] != r.shape[0]:
            raise ValueError("pos and rad disagree on the atom count")
    off, reso = _batch_desc([p.shape[0] for p in pos], reso)
    nmol = len(pos)
    P = np.ascontiguousarray(np.concatenate(pos) if nmol else np.zeros((0, 3), np.float32))
    R = np.ascontiguousarray(np.concatenate(rad) if nmol else np.zeros(0, np.float32))
    nv = np.zeros(max(nmol, 1), dtype=np.uint32)
    nt = np.zeros(max(nmol, 1), dtype=np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint)
    _check(L.qs_compute_batch(_fptr(P), _fptr(R), off.ctypes.data_as(u32p), nmol, _fptr(reso), int(smooth),
                              nv.ctypes.data_as(u32p), nt.ctypes.data_as(u32p)))
    vo = np.zeros(nmol + 1, dtype=np.int64)
    to = np.zeros(nmol + 1, dtype=np.int64)
    vo[1:] = np.cumsum(nv[:nmol])
    to[1:] = np.cumsum(nt[:nmol])
    return vo, to


def _slot_views(L, slot, vo, to):
    vp = ctypes.c_void_p
    pv, pt, pa = vp(), vp(), vp()
    nv, nt = ctypes.c_uint(0), ctypes.c_uint(0)
    _check(L.qs_batch_wait_host(slot, ctypes.byref(pv), ctypes.byref(pt), ctypes.byref(pa), ctypes.byref(nv),
                                ctypes.byref(nt)))
    nvt, ntt = nv.value, nt.value
    V = np.ctypeslib.as_array(ctypes.cast(pv, ctypes.POINTER(ctypes.c_float)), shape=(max(nvt, 1) * 3,))[: 3 * nvt]
    T = np.ctypeslib.as_array(ctypes.cast(pt, ctypes.POINTER(ctypes.c_int)), shape=(max(ntt, 1) * 3,))[: 3 * ntt]
    A = np.ctypeslib.as_array(ctypes.cast(pa, ctypes.POINTER(ctypes.c_int)), shape=(max(nvt, 1),))[:nvt]
    V, T = V.reshape(-1, 3), T.reshape(-1, 3)
    return [(V[vo[m]:vo[m + 1]], T[to[m]:to[m + 1]], A[vo[m]:vo[m + 1]]) for m in range(len(vo) - 1)]


def compute_batches_host(chunks, reso=0.5, smooth=1):
    """Mesh a sequence of batches (each a list of (pos, rad) pairs) with the copies of one batch
    to the host overlapping the next batch's work (qs_batch_fetch_host_async / qs_batch_wait_host).
    Yields, batch by batch, the list of (vertices, triangles, atom_ids) views of the library's
    pinned host arrays; a batch's views stay valid until the generator resumes (copy them to keep
    them)."""
    L = lib()
    prev = None
    for k, ch in enumerate(chunks):
        vo, to = _run_batch(L, ch, reso, smooth)
        _check(L.qs_batch_fetch_host_async(k & 1))
        if prev is not None:
            yield _slot_views(L, *prev)
        prev = (k & 1, vo, to)
    if prev is not None:
        yield _slot_views(L, *prev)


def batch_offsets():
    """(vert_off, tri_off) of the last batch: nmol + 1 entries each (triangles counted in triangles)."""
    L = lib()
    n = L.qs_batch_offsets(None, None)
    if n < 0:
        _check(n)
    vo = np.zeros(n + 1, dtype=np.uint32)
    to = np.zeros(n + 1, dtype=np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint)
    L.qs_batch_offsets(vo.ctypes.data_as(u32p), to.ctypes.data_as(u32p))
    return vo, to


def batch_mesh(m):
    """Mesh of molecule m of the last batch (vertex indices local to it), copied to host."""
    L = lib()
    vo, to = batch_offsets()
    if not 0 <= m < len(vo) - 1:
        raise QuickSESError("no molecule %d in the last batch" % m)
    nv, nt = int(vo[m + 1] - vo[m]), int(to[m + 1] - to[m])
    V = np.empty((nv, 3), dtype=np.float32)
    T = np.empty((nt, 3), dtype=np.int32)
    A = np.empty(nv, dtype=np.int32)
    rc = L.qs_batch_mesh(int(m), ctypes.c_void_p(V.ctypes.data), ctypes.c_void_p(T.ctypes.data),
                         ctypes.c_void_p(A.ctypes.data))
    _check(rc)
    return V, T, A


def compute_batch_device(d_pos, d_rad, counts, reso=0.5, smooth=1):
    """qs_compute_batch_device on device pointers: counts = atoms per molecule (concatenated in
    d_pos / d_rad).  Returns (vert_off, tri_off, d_vert, d_tri, d_atom): per-molecule offsets
    (nmol+1 each, triangles counted in triangles) and library-owned device pointers."""
    L = lib()
    off, reso = _batch_desc(counts, reso)
    nmol = len(counts)
    vo = np.zeros(nmol + 1, dtype=np.uint32)
    to = np.zeros(nmol + 1, dtype=np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint)
    pv, pt, pa = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    rc = L.qs_compute_batch_device(ctypes.c_void_p(d_pos), ctypes.c_void_p(d_rad), off.ctypes.data_as(u32p), nmol,
                                   _fptr(reso), int(smooth), vo.ctypes.data_as(u32p), to.ctypes.data_as(u32p),
                                   ctypes.byref(pv), ctypes.byref(pt), ctypes.byref(pa))
    if rc:
        raise QuickSESError("quickses error %d: %s" % (rc, L.qs_last_error_message().decode()))
    return vo, to, pv.value, pt.value, pa.value


def batch_grid(m):
    """Grid of molecule m of the last batch."""
    L = lib()
    fp = (ctypes.c_float * 5)()
    ip = (ctypes.c_int * 4)()
    if L.qs_batch_grid(int(m), fp, ip) != 0:
        raise QuickSESError("no molecule %d in the last batch" % m)
    return dict(origin=(fp[0], fp[1], fp[2]), dx=fp[3], dxn=fp[4], g=ip[0], gn=ip[1], r=ip[2], lim=ip[3])


CLI_PATH = os.path.join(_HERE, "quickses")


def read_pdb(path):
    """(pos (N, 3) float32, rad (N,) float32) of a PDB file, read as the reference's tool reads it."""
    L = lib()
    pp, rp = ctypes.POINTER(ctypes.c_float)(), ctypes.POINTER(ctypes.c_float)()
    n = ctypes.c_uint(0)
    rc = L.qs_pdb_read(os.fsencode(path), ctypes.byref(pp), ctypes.byref(rp), ctypes.byref(n))
    if rc:
        raise QuickSESError(L.qs_io_error().decode())
    try:
        N = n.value
        pos = np.ctypeslib.as_array(pp, shape=(max(N, 1) * 3,))[: 3 * N].reshape(-1, 3).copy()
        rad = np.ctypeslib.as_array(rp, shape=(max(N, 1),))[:N].copy()
    finally:
        L.qs_free(ctypes.cast(pp, ctypes.c_void_p))
        L.qs_free(ctypes.cast(rp, ctypes.c_void_p))
    return pos, rad


def write_obj(path, vertices, triangles):
    """Write a mesh in the reference tool's OBJ layout."""
    L = lib()
    V = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    T = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
    rc = L.qs_obj_write(os.fsencode(path), ctypes.c_void_p(V.ctypes.data), V.shape[0], ctypes.c_void_p(T.ctypes.data),
                        T.shape[0])
    if rc:
        raise QuickSESError(L.qs_io_error().decode())


class Server:
    """Request server (qs_serve_*): submit surface requests at any time, poll for device-resident
    meshes.  Use as a context manager, or call close()."""

    def __init__(self, hbm_budget_bytes=16 << 30, result_bytes=8 << 30, workers=2, max_batch_voxels=0, express_voxels=0):
        L = lib()
        rc = L.qs_serve_open(int(hbm_budget_bytes), int(result_bytes), int(workers))
        if rc:
            raise QuickSESError("quickses error %d: %s" % (rc, L.qs_last_error_message().decode()))
        self.open = True
        if max_batch_voxels:
            self.set("max_batch_voxels", max_batch_voxels)
        if express_voxels:
            self.set("express_voxels", express_voxels)

    def set(self, key, value):
        rc = lib().qs_serve_set(key.encode(), int(value))
        if rc:
            raise QuickSESError("quickses error %d: %s" % (-rc, lib().qs_last_error_message().decode()))

    def submit(self, pos, rad, reso=0.5, smooth=1):
        L = lib()
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        rad = np.ascontiguousarray(rad, dtype=np.float32).reshape(-1)
        if pos.shape[0] != rad.shape[0]:
            raise ValueError("pos and rad disagree on the atom count")
        rid = L.qs_serve_submit(_fptr(pos), _fptr(rad), pos.shape[0], ctypes.c_float(reso), int(smooth))
        if rid < 0:
            raise QuickSESError("quickses error %d: %s" % (-rid, L.qs_last_error_message().decode()))
        return rid

    def poll(self, rid, wait_ms=-1):
        """None if not finished within wait_ms, else dict(nvert, ntri, d_vert, d_tri, d_atom,
        latency_ms, queue_ms, batch_size)."""
        L = lib()
        nv, nt = ctypes.c_uint(0), ctypes.c_uint(0)
        pv, pt, pa = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        lat = (ctypes.c_double * 3)()
        rc = L.qs_serve_poll(int(rid), int(wait_ms), ctypes.byref(nv), ctypes.byref(nt), ctypes.byref(pv),
                             ctypes.byref(pt), ctypes.byref(pa), lat)
        if rc == 1:
            return None
        if rc:
            raise QuickSESError("quickses error %d: %s" % (-rc, L.qs_last_error_message().decode()))
        return dict(nvert=nv.value, ntri=nt.value, d_vert=pv.value, d_tri=pt.value, d_atom=pa.value,
                    latency_ms=lat[0], queue_ms=lat[1], batch_size=int(lat[2]))

    def fetch(self, rid):
        """(vertices, triangles, atom_ids) on the host; waits for the request."""
        L = lib()
        info = self.poll(rid, -1)
        V = np.empty((info["nvert"], 3), dtype=np.float32)
        T = np.empty((info["ntri"], 3), dtype=np.int32)
        A = np.empty(info["nvert"], dtype=np.int32)
        rc = L.qs_serve_fetch(int(rid), ctypes.c_void_p(V.ctypes.data), ctypes.c_void_p(T.ctypes.data),
                              ctypes.c_void_p(A.ctypes.data))
        _check(rc)
        return V, T, A

    def release(self, rid):
        rc = lib().qs_serve_release(int(rid))
        if rc:
            raise QuickSESError("quickses error %d: %s" % (-rc, lib().qs_last_error_message().decode()))

    def stats(self):
        out = (ctypes.c_uint64 * 8)()
        lib().qs_serve_stats(out, 8)
        keys = ("batches", "requests_batched", "largest_batch", "max_busy_workers", "arena_used", "arena_bytes",
                "pending", "express_batches")
        return {k: int(out[i]) for i, k in enumerate(keys)}

    def close(self):
        if self.open:
            lib().qs_serve_close()
            self.open = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def copy_device(d_dst, d_src, nbytes):
    """Plain device copy kernel (traffic-counter reference); returns its time in ms."""
    L = lib()
    ms = ctypes.c_float(0)
    rc = L.qs_copy_device(ctypes.c_void_p(d_dst), ctypes.c_void_p(d_src), int(nbytes), ctypes.byref(ms))
    _check(rc)
    return ms.value


def vertex_normals(nvert):
    """Unit field gradients (outward) at the last mesh's vertices, in vertex order, as a
    (nvert, 3) float32 array; nvert is the last mesh's vertex count (all molecules of a batch)."""
    out = np.empty((int(nvert), 3), np.float32)
    rc = lib().qs_vertex_normals(ctypes.c_void_p(out.ctypes.data), int(nvert))
    _check(rc)
    return out


def vertex_colors(atom_rgba, nvert):
    """RGBA of each vertex's closest atom: atom_rgba is (natoms, 4) float32 in the atom order of
    the last call (a batch's molecules concatenated); returns (nvert, 4) float32."""
    tab = np.ascontiguousarray(atom_rgba, np.float32).reshape(-1, 4)
    out = np.empty((int(nvert), 4), np.float32)
    rc = lib().qs_vertex_colors(ctypes.c_void_p(tab.ctypes.data), tab.shape[0], ctypes.c_void_p(out.ctypes.data), int(nvert))
    _check(rc)
    return out


def vertex_normals_device(d_out):
    """Normals into a device array (3 floats per vertex); returns the kernel time in ms."""
    ms = ctypes.c_float(0)
    rc = lib().qs_vertex_normals_device(ctypes.c_void_p(d_out), ctypes.byref(ms))
    _check(rc)
    return ms.value


def stage_times():
    """{stage: ms} of the last call (option "timing" must be on)."""
    L = lib()
    names = ctypes.create_string_buffer(4096)
    ms = (ctypes.c_float * 64)()
    k = L.qs_stage_times(names, 4096, ms, 64)
    keys = names.value.decode().split(",")[:k]
    return {key: ms[i] for i, key in enumerate(keys)}


def last_grid():
    L = lib()
    fp = (ctypes.c_float * 5)()
    ip = (ctypes.c_int * 4)()
    L.qs_last_grid(fp, ip)
    return dict(origin=(fp[0], fp[1], fp[2]), dx=fp[3], dxn=fp[4], g=ip[0], gn=ip[1], r=ip[2], lim=ip[3])


TUNING_PATH = os.path.join(_HERE, "tuning.json")


def set_tuning(key, reso, value):
    """qs_set_tuning: one tuned value per voxel size (key "pass1_slab_rows")."""
    if lib().qs_set_tuning(key.encode(), ctypes.c_float(reso), int(value)) != 0:
        raise QuickSESError("tuning %s at %g A not set" % (key, reso))


def load_tuning(path=TUNING_PATH):
    """qs_load_tuning: the library's tuned values from a tuning file (None clears them); returns
    the entries read."""
    n = lib().qs_load_tuning(path.encode() if path else None)
    if n < 0:
        raise QuickSESError(lib().qs_last_error_message().decode())
    return n


def tuning(path=TUNING_PATH):
    """The tuning file's contents as a dict ({} when there is none)."""
    import json

    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def last_slab_count():
    """x-slabs the last computation's grid ran in (1: the whole grid at once)."""
    return int(lib().qs_last_slab_count())


def fetch(what, dtype, count):
    out = np.empty(count, dtype=dtype)
    n = lib().qs_fetch(what.encode(), ctypes.c_void_p(out.ctypes.data), out.nbytes)
    if n < 0:
        raise QuickSESError("fetch of %s failed" % what)
    return out[: n // out.itemsize]


def _grid_params(grid):
    fp = (ctypes.c_float * 5)(grid["origin"][0], grid["origin"][1], grid["origin"][2], grid["dx"], grid["dxn"])
    ip = (ctypes.c_int * 4)(grid["g"], grid["gn"], grid["r"], grid["lim"])
    return fp, ip


def stage_pass2(grid, labels):
    """Field pass 2 on caller-given labels (uint8[g^3])."""
    fp, ip = _grid_params(grid)
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    field = np.empty(labels.shape, dtype=np.float32)
    rc = lib().qs_stage_pass2(fp, ip, ctypes.c_void_p(labels.ctypes.data), ctypes.c_void_p(field.ctypes.data))
    _check(rc)
    return field


def stage_mesh(grid, field, vert_cap=None, tri_cap=None):
    """Marching cubes + edge-id weld on a caller-given field; unsmoothed, unfiltered."""
    fp, ip = _grid_params(grid)
    field = np.ascontiguousarray(field, dtype=np.float32)
    g = grid["g"]
    vert_cap = vert_cap or max(3 * 4 * g * g * 8, 1024)
    tri_cap = tri_cap or max(3 * 8 * g * g * 8, 1024)
    vert = np.empty(vert_cap, dtype=np.float32)
    tri = np.empty(tri_cap, dtype=np.int32)
    nv = ctypes.c_uint(0)
    nt = ctypes.c_uint(0)
    rc = lib().qs_stage_mesh(fp, ip, ctypes.c_void_p(field.ctypes.data), ctypes.byref(nv), ctypes.byref(nt),
                             ctypes.c_void_p(vert.ctypes.data), vert_cap, ctypes.c_void_p(tri.ctypes.data), tri_cap)
    if rc == 5:
        return stage_mesh(grid, field, 3 * nv.value + 3, 3 * nt.value + 3)
    _check(rc)
    return vert[: 3 * nv.value].reshape(-1, 3).copy(), tri[: 3 * nt.value].reshape(-1, 3).copy()


def synth_ball(seed, n, density=0.05, min_dist=1.2):
    """n atoms uniform in a ball at `density` atoms/A^3 with min-distance rejection (seeded)."""
    pos = np.empty((n, 3), dtype=np.float32)
    rad = np.empty(n, dtype=np.float32)
    rc = lib().qs_synth_ball(seed, n, density, min_dist, ctypes.c_void_p(pos.ctypes.data), ctypes.c_void_p(rad.ctypes.data))
    if rc:
        raise QuickSESError("synth_ball failed")
    return pos, rad


def synth_shell(seed, n, r_in, r_out, min_dist=1.2):
    pos = np.empty((n, 3), dtype=np.float32)
    rad = np.empty(n, dtype=np.float32)
    rc = lib().qs_synth_shell(seed, n, r_in, r_out, min_dist, ctypes.c_void_p(pos.ctypes.data),
                              ctypes.c_void_p(rad.ctypes.data))
    if rc:
        raise QuickSESError("synth_shell failed")
    return pos, rad


def synth_jitter(seed, base, sigma=0.3):
    base = np.ascontiguousarray(base, dtype=np.float32)
    out = np.empty_like(base)
    lib().qs_synth_jitter(seed, base.shape[0], sigma, ctypes.c_void_p(base.ctypes.data), ctypes.c_void_p(out.ctypes.data))
    return out


def synth_sizes(seed, count, lo=500, hi=20000):
    out = np.empty(count, dtype=np.int32)
    lib().qs_synth_sizes(seed, count, lo, hi, ctypes.c_void_p(out.ctypes.data))
    return out
