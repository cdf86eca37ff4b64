"""Request server: every result equals the same request meshed alone, whatever the mix,
arrival pattern, memory budget or number of workers; error paths return errors."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_mesh(a, b):
    Va, Ta, Aa = a
    Vb, Tb, Ab = b
    assert Va.shape == Vb.shape and Ta.shape == Tb.shape
    assert np.array_equal(Ta, Tb)
    assert np.array_equal(Va.view(np.uint32), Vb.view(np.uint32))
    assert np.array_equal(Aa, Ab)


def _requests(qs, k, seed=900):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(np.exp(rng.uniform(np.log(200), np.log(3000))))
        reso = float(rng.choice([0.3, 0.5, 1.0]))
        smooth = int(rng.choice([0, 1, 2]))
        out.append((qs.synth_ball(seed + i, n), reso, smooth))
    pos, rad = qs.synth_ball(seed + 99, 2)
    out.insert(3, ((pos[:1], rad[:1]), 0.5, 1))  # a one-atom request: empty mesh
    return out


# (workers, budget, batch cap, express worker's request cap): the express worker (the last one)
# takes only requests of at most that many grid points
@pytest.mark.parametrize("workers,budget,cap,express", [(2, 8 << 30, 0, 0), (1, 256 << 20, 0, 0), (3, 96 << 20, 0, 0),
                                                        (2, 8 << 30, 3_000_000, 0), (3, 8 << 30, 0, 2_000_000),
                                                        (2, 8 << 30, 3_000_000, 500_000)])
def test_served_results_equal_alone(qs, workers, budget, cap, express):
    reqs = _requests(qs, 14)
    print("case serve workers=%d budget=%d cap=%d express=%d" % (workers, budget, cap, express), flush=True)
    with qs.Server(hbm_budget_bytes=budget, result_bytes=1 << 30, workers=workers, max_batch_voxels=cap,
                   express_voxels=express) as srv:
        ids = []
        for i, ((pos, rad), reso, smooth) in enumerate(reqs):
            ids.append(srv.submit(pos, rad, reso, smooth))
            if i % 4 == 1:
                time.sleep(0.002)  # staggered arrivals: some land while a batch runs
        served = [srv.fetch(r) for r in ids]
        info = [srv.poll(r, 0) for r in ids]
        st = srv.stats()
        for r in ids:
            srv.release(r)
        assert srv.stats()["arena_used"] == 0
    assert st["requests_batched"] == len(reqs)
    if budget < (1 << 30) or cap:
        assert st["batches"] > 1
    for ((pos, rad), reso, smooth), got, inf in zip(reqs, served, info):
        assert inf["latency_ms"] >= inf["queue_ms"] >= 0.0 and inf["batch_size"] >= 1
        _same_mesh(got, qs.compute_ses(pos, rad, reso, smooth))


def test_served_express_lane(qs):
    """Small requests submitted while the other worker meshes a large one are taken by the express
    worker (the large one never is), and every result equals the request meshed alone."""
    big = qs.synth_ball(950, 200_000)  # 0.3 A: 688^3 grid points, a few ms
    small = [qs.synth_ball(951 + i, 300) for i in range(4)]
    with qs.Server(hbm_budget_bytes=16 << 30, result_bytes=2 << 30, workers=2, express_voxels=2_000_000) as srv:
        print("case serve express lane", flush=True)
        ids = [srv.submit(big[0], big[1], 0.3, 1)]
        time.sleep(0.001)  # the normal worker has taken the large request and runs it
        ids += [srv.submit(p, r, 1.0, 1) for p, r in small]
        got = [srv.fetch(r) for r in ids]
        info = [srv.poll(r, 0) for r in ids]
        st = srv.stats()
        for r in ids:
            srv.release(r)
    assert st["express_batches"] >= 1
    assert info[0]["batch_size"] == 1
    for (pos, rad), reso, res in zip([big] + small, [0.3] + [1.0] * 4, got):
        _same_mesh(res, qs.compute_ses(pos, rad, reso, 1))


def test_served_device_pointers(qs):
    from quickses_amd import hipmem

    (pos, rad), reso, smooth = _requests(qs, 1, seed=950)[0]
    with qs.Server(hbm_budget_bytes=1 << 30, result_bytes=256 << 20, workers=1) as srv:
        rid = srv.submit(pos, rad, reso, smooth)
        inf = srv.poll(rid, -1)
        V = hipmem.download(inf["d_vert"], (inf["nvert"], 3), np.float32)
        T = hipmem.download(inf["d_tri"], (inf["ntri"], 3), np.int32)
        A = hipmem.download(inf["d_atom"], (inf["nvert"],), np.int32)
        srv.release(rid)
    _same_mesh((V, T, A), qs.compute_ses(pos, rad, reso, smooth))


def test_served_errors(qs):
    pos, rad = qs.synth_ball(960, 1500)
    assert qs.lib().qs_serve_close() < 0  # no server open
    with qs.Server(hbm_budget_bytes=1 << 30, result_bytes=1 << 20, workers=1) as srv:
        with pytest.raises(qs.QuickSESError):
            qs.Server()  # one server per process
        bad = pos.copy()
        bad[0, 0] = np.inf
        with pytest.raises(qs.QuickSESError):
            srv.submit(bad, rad, 0.5, 1)
        with pytest.raises(qs.QuickSESError):
            srv.submit(pos, rad, 0.0, 1)
        with pytest.raises(qs.QuickSESError):
            srv.set("no_such_setting", 1)
        # a mesh larger than the 1 MiB result arena fails cleanly and frees nothing it did not take
        rid = srv.submit(pos, rad, 0.3, 1)
        with pytest.raises(qs.QuickSESError, match="arena"):
            srv.poll(rid, -1)
        srv.release(rid)
        small, srad = qs.synth_ball(961, 30)
        rid = srv.submit(small, srad, 1.0, 0)
        assert srv.poll(rid, -1)["nvert"] > 0
        srv.release(rid)
        with pytest.raises(qs.QuickSESError):
            srv.release(rid)
