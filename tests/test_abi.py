"""The C-ABI library loads, exports every declared entry point, and its host-only parts work
without a GPU (error returns instead of exits)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "quickses.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", src)
    return sorted(set(n for n in names if n.startswith(("API_", "qs_"))))


def test_header_declares_reference_entry_points():
    names = _declared_functions()
    for ref in ("API_computeSES", "API_getTriangles", "API_getVertices", "API_freeMesh", "API_getAtomIdPerVert"):
        assert ref in names


def test_library_exports_every_declared_symbol(qs):
    L = ctypes.CDLL(qs.LIB_PATH)
    missing = [n for n in _declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_float3_layout():
    class F3(ctypes.Structure):
        _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]

    assert ctypes.sizeof(F3) == 12 and ctypes.alignment(F3) == 4


def test_options_roundtrip(qs):
    for key in ("lim_mode", "atom_ids_sorted", "pass1_direct", "pass1_rows", "pass2_direct", "trace", "timing"):
        old = qs.get_option(key)
        qs.set_option(key, 1)
        assert qs.get_option(key) == 1
        qs.set_option(key, old)
    with pytest.raises(KeyError):
        qs.set_option("no_such_option", 1)


def test_no_gpu_gives_error_not_exit(qs):
    from quickses_amd import hipmem

    if hipmem.device_count() > 0:
        pytest.skip("a GPU is present")
    pos, rad = qs.synth_ball(1, 50)
    with pytest.raises(qs.QuickSESError):
        qs.compute_ses(pos, rad, 0.5, 1)
    # the getters stay valid after a failure
    assert qs.lib().API_lastError() != 0


def test_small_inputs_need_no_gpu(qs):
    # N <= 1 returns an empty mesh before touching the device (CudaSurf.cu:539-543)
    for n in (0, 1):
        V, T, A = qs.compute_ses(np.zeros((n, 3), np.float32), np.full(n, 1.7, np.float32), 0.5, 1)
        assert V.shape == (0, 3) and T.shape == (0, 3) and A.shape == (0,)


def test_synthetic_ball_properties(qs):
    pos, rad = qs.synth_ball(1, 2000)
    pos2, rad2 = qs.synth_ball(1, 2000)
    assert np.array_equal(pos, pos2) and np.array_equal(rad, rad2)
    R = (3 * 2000 / (4 * np.pi * 0.05)) ** (1 / 3)
    assert np.linalg.norm(pos.astype(np.float64), axis=1).max() <= R + 1e-4
    # minimum distance 1.2 A
    from scipy.spatial import cKDTree

    d, _ = cKDTree(pos).query(pos, k=2)
    assert d[:, 1].min() >= 1.2 - 1e-5
    frac = {r: np.mean(np.isclose(rad, r)) for r in (1.7, 1.52, 1.55, 1.2, 1.8)}
    assert abs(frac[1.7] - 0.55) < 0.04 and abs(frac[1.52] - 0.20) < 0.03 and abs(frac[1.2] - 0.09) < 0.03
    assert not np.array_equal(qs.synth_ball(2, 2000)[0], pos)


def test_synthetic_shell_and_sizes(qs):
    pos, rad = qs.synth_shell(3, 3000, 40.0, 55.0)
    r = np.linalg.norm(pos.astype(np.float64), axis=1)
    assert r.min() >= 40.0 - 1e-3 and r.max() <= 55.0 + 1e-3
    s = qs.synth_sizes(4, 256, 500, 20000)
    assert s.min() >= 500 and s.max() <= 20000 and len(np.unique(s)) > 200
    base, _ = qs.synth_ball(4500, 500)
    fr = qs.synth_jitter(5000, base, 0.3)
    dev = (fr - base).astype(np.float64)
    assert abs(dev.std() - 0.3) < 0.02 and abs(dev.mean()) < 0.02


def test_server_without_gpu_returns_errors(qs):
    from quickses_amd import hipmem

    if hipmem.device_count() > 0:
        pytest.skip("a GPU is present")
    L = qs.lib()
    pos, rad = qs.synth_ball(3, 20)
    assert L.qs_serve_submit(qs._fptr(pos), qs._fptr(rad), 20, 0.5, 1) < 0  # no server open
    with pytest.raises(qs.QuickSESError):
        qs.Server(hbm_budget_bytes=1 << 30, result_bytes=1 << 26, workers=1)
    assert L.qs_serve_close() < 0


def test_normals_without_a_mesh_is_an_error(qs):
    """Before any mesh exists the normals entry points return an error (no GPU work)."""
    import quickses_amd as q
    L = q.lib()
    if L.qs_batch_offsets(None, None) >= 0:
        pytest.skip("a mesh exists in this process")
    assert L.qs_vertex_normals(None, 0) != 0
    assert L.qs_vertex_normals_device(None, None) != 0


def test_colors_without_a_mesh_is_an_error(qs):
    import quickses_amd as q
    L = q.lib()
    if L.qs_batch_offsets(None, None) >= 0:
        pytest.skip("a mesh exists in this process")
    assert L.qs_vertex_colors(None, 0, None, 0) != 0


def test_host_fetches_without_a_batch_are_errors(qs):
    """The library-owned host fetches check their state before any GPU work: no batch result,
    a slot other than 0 or 1, a wait on a slot that no fetch started."""
    import ctypes
    import quickses_amd as q
    L = q.lib()
    if L.qs_batch_offsets(None, None) >= 0:
        pytest.skip("a mesh exists in this process")
    vp = ctypes.c_void_p
    pv, pt, pa = vp(), vp(), vp()
    nv, nt = ctypes.c_uint(0), ctypes.c_uint(0)
    assert L.qs_batch_fetch_host(ctypes.byref(pv), ctypes.byref(pt), ctypes.byref(pa)) != 0
    assert L.qs_batch_fetch_host_async(0) != 0
    assert L.qs_batch_wait_host(2, ctypes.byref(pv), ctypes.byref(pt), ctypes.byref(pa), ctypes.byref(nv),
                                ctypes.byref(nt)) != 0
    assert L.qs_batch_wait_host(0, ctypes.byref(pv), ctypes.byref(pt), ctypes.byref(pa), ctypes.byref(nv),
                                ctypes.byref(nt)) != 0
