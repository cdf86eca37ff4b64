"""qs_compute_device_noalloc, through the C ABI: the same mesh as qs_compute_device, bit for bit,
on two calls in a row (the second with a smaller molecule, in the region the first one used);
too small an output capacity and too large a molecule are errors that write nothing."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _reference(qs, hipmem, dp, dr, n, reso, smooth):
    nv, ni, pv, pt, pa = qs.compute_device(dp.ptr, dr.ptr, n, reso, smooth)
    return (hipmem.download(pv, (nv, 3), np.float32), hipmem.download(pt, (ni // 3, 3), np.int32),
            hipmem.download(pa, (nv,), np.int32))


def _noalloc(qs, hipmem, dp, dr, n, reso, smooth, vcap, tcap):
    dv, dt, da = hipmem.alloc(max(vcap, 1) * 12), hipmem.alloc(max(tcap, 1) * 12), hipmem.alloc(max(vcap, 1) * 4)
    try:
        nv, ni = qs.compute_device_noalloc(dp.ptr, dr.ptr, n, reso, smooth, dv, vcap, dt, tcap, da)
        return (hipmem.download(dv, (nv, 3), np.float32), hipmem.download(dt, (ni // 3, 3), np.int32),
                hipmem.download(da, (nv,), np.int32))
    finally:
        for p in (dv, dt, da):
            hipmem.free(p)


def _same(a, b):
    assert a[0].shape == b[0].shape and a[1].shape == b[1].shape
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


def test_noalloc_equals_device_entry_twice(qs):
    from quickses_amd import hipmem

    cases = [(401, 20000, 0.5, 1), (402, 2000, 0.4, 2)]  # the second call smaller than the first
    for seed, n, reso, smooth in cases:
        pos, rad = qs.synth_ball(seed, n)
        dp, dr = hipmem.DeviceArray(pos), hipmem.DeviceArray(rad)
        try:
            print("case noalloc seed=%d n=%d reso=%.1f" % (seed, n, reso), flush=True)
            ref = _reference(qs, hipmem, dp, dr, n, reso, smooth)
            got = _noalloc(qs, hipmem, dp, dr, n, reso, smooth, len(ref[0]) + 100, len(ref[1]) + 100)
        finally:
            dp.free()
            dr.free()
        assert len(ref[1]) > 0
        _same(got, ref)


def test_noalloc_errors(qs):
    from quickses_amd import hipmem

    pos, rad = qs.synth_ball(403, 1500)
    dp, dr = hipmem.DeviceArray(pos), hipmem.DeviceArray(rad)
    try:
        ref = _reference(qs, hipmem, dp, dr, len(rad), 0.5, 1)
        print("case noalloc capacity too small", flush=True)
        with pytest.raises(qs.QuickSESError, match="capacity"):
            _noalloc(qs, hipmem, dp, dr, len(rad), 0.5, 1, len(ref[0]) - 1, len(ref[1]))
        with pytest.raises(qs.QuickSESError, match="capacity"):
            _noalloc(qs, hipmem, dp, dr, len(rad), 0.5, 1, len(ref[0]), len(ref[1]) - 1)
        # exactly enough room
        _same(_noalloc(qs, hipmem, dp, dr, len(rad), 0.5, 1, len(ref[0]), len(ref[1])), ref)
    finally:
        dp.free()
        dr.free()
    # a grid beyond the fixed region: 20,000 atoms at 0.2 A (about 130 M grid points)
    big, brad = qs.synth_ball(404, 20000)
    dp, dr = hipmem.DeviceArray(big), hipmem.DeviceArray(brad)
    try:
        print("case noalloc molecule too large", flush=True)
        with pytest.raises(qs.QuickSESError, match="scratch region"):
            _noalloc(qs, hipmem, dp, dr, len(brad), 0.2, 1, 1 << 20, 1 << 21)
    finally:
        dp.free()
        dr.free()
    # the region is usable again after a rejected call
    pos, rad = qs.synth_ball(405, 800)
    dp, dr = hipmem.DeviceArray(pos), hipmem.DeviceArray(rad)
    try:
        ref = _reference(qs, hipmem, dp, dr, len(rad), 0.5, 1)
        _same(_noalloc(qs, hipmem, dp, dr, len(rad), 0.5, 1, len(ref[0]), len(ref[1])), ref)
    finally:
        dp.free()
        dr.free()
