"""The measured workloads at their real sizes, checked against the oracle:
- the 50,000-atom 0.2 A molecule: labels, field and mesh bit for bit;
- the batched workload (256 molecules of 500-20,000 atoms plus trajectory frames of a
  10,000-atom molecule), batched in the benchmark's chunks: molecules across the size range and
  frames bit for bit, every mesh closed and consistently oriented;
- the served stream (1,000-200,000 atoms, voxels 0.3 / 0.5 / 1.0 A, the 200,000-atom 0.3 A
  request included): served results bit for bit against the oracle, the rest against the same
  request alone."""
import numpy as np
import pytest

from test_gpu_fullsize import _check_closed_oriented

pytestmark = pytest.mark.gpu


def _same_mesh(a, b, what):
    Va, Ta, Aa = a
    Vb, Tb, Ab = b
    assert Va.shape == Vb.shape and Ta.shape == Tb.shape, (what, Va.shape, Vb.shape, Ta.shape, Tb.shape)
    assert np.array_equal(Ta, Tb), what
    assert np.array_equal(Va.view(np.uint32), Vb.view(np.uint32)), (what, float(np.abs(Va - Vb).max()))
    assert np.array_equal(Aa, Ab), what


@pytest.mark.timeout(900)
def test_config2_fullsize_matches_oracle(qs, orc):
    """50,000 atoms at 0.2 A (g = 665, 294 M voxels): the oracle's labels, field, welded mesh,
    atom ids, smoothed vertices and filtered triangles, bit for bit."""
    pos, rad = qs.synth_ball(2, 50000)
    print("case config 2 full size: GPU", flush=True)
    V, T, A = qs.compute_ses(pos, rad, 0.2, 1)
    grid = qs.last_grid()
    g = grid["g"]
    lab = qs.fetch("labels", np.uint8, g ** 3)
    fld = qs.fetch("field", np.float32, g ** 3)
    print("case config 2 full size: oracle field (g=%d)" % g, flush=True)
    G, perm, cs, ce, sx, olab, ofld = orc.field(pos, rad, 0.2)
    assert G.g == g
    assert np.array_equal(lab, olab), "labels differ"
    del lab, olab
    assert np.array_equal(fld.view(np.uint32), ofld.view(np.uint32)), "field differs"
    del fld
    print("case config 2 full size: oracle mesh", flush=True)
    Vo, To = orc.edge_weld(G, g - 1, ofld)
    del ofld
    Ao = orc.closest_atom(G, cs, ce, sx, Vo)
    Ao = np.where(Ao >= 0, perm[np.maximum(Ao, 0)], Ao).astype(np.int32)
    Vo = orc.smooth(Vo, To, 1, False)
    To = orc.filter_degenerate(To)
    _same_mesh((V, T, A), (Vo, To, Ao), "config 2")
    print("  %d vertices, %d triangles equal" % (len(V), len(T)), flush=True)


@pytest.mark.timeout(900)
def test_config4_batched_workload(qs, orc):
    """The batched workload in the benchmark's launch sequences (1.5 G voxels each): every mesh
    closed and consistently oriented; 8 molecules across the size range and 4 frames equal the
    oracle's meshes."""
    from quickses_amd.workloads import chunks_by_voxels, config4_sets

    mols, frames = config4_sets(qs, 256, 50)
    for name, items, picks in (("molecules", mols, None), ("frames", frames, [0, 17, 33, 49])):
        chunks = chunks_by_voxels(items, 0.5, 1_500_000_000)
        print("case config 4 %s: %d meshes in %d launch sequences" % (name, len(items), len(chunks)), flush=True)
        out = []
        for ch in chunks:
            out.extend(qs.compute_batch(ch, 0.5, 1))
        assert len(out) == len(items)
        for m, (V, T, A) in enumerate(out):
            assert len(T) > 0 and T.max() < len(V), (name, m)
            assert A.min() >= 0 and A.max() < len(items[m][1]), (name, m)
            _check_closed_oriented(T, len(V))
        if picks is None:
            order = np.argsort([len(r) for _, r in items])
            picks = [int(order[int(round(q * (len(order) - 1)))]) for q in np.linspace(0.0, 1.0, 8)]
        for m in picks:
            pos, rad = items[m]
            print("  %s %d (%d atoms) against the oracle" % (name, m, len(rad)), flush=True)
            _same_mesh(out[m], orc.compute_ses(pos, rad, 0.5, 1), "%s %d" % (name, m))


@pytest.mark.timeout(900)
def test_config5_served_stream_matches_oracle(qs, orc):
    """A served stream of the served workload plus its largest request (200,000 atoms at 0.3 A),
    submitted together so that requests share batches and overlap on the device: the largest,
    the smallest and one request of each voxel size equal the oracle's meshes; every other one
    equals the same request alone."""
    from quickses_amd.workloads import config5_requests

    reqs = config5_requests(qs, 24)
    reqs.append((qs.synth_ball(55, 200000), 0.3))
    print("case config 5: %d requests, %d-%d atoms" % (len(reqs), min(len(r[0][1]) for r in reqs),
                                                      max(len(r[0][1]) for r in reqs)), flush=True)
    with qs.Server(hbm_budget_bytes=48 << 30, result_bytes=16 << 30, workers=2, max_batch_voxels=150_000_000) as srv:
        ids = [srv.submit(pos, rad, reso, 1) for (pos, rad), reso in reqs]
        served = [srv.fetch(r) for r in ids]
        st = srv.stats()
        for r in ids:
            srv.release(r)
    assert st["requests_batched"] == len(reqs)
    sizes = [len(r[0][1]) for r in reqs]
    picks = {int(np.argmax(sizes)), int(np.argmin(sizes))}
    for vox in (0.3, 0.5, 1.0):
        cand = [i for i, r in enumerate(reqs) if r[1] == vox and i not in picks]
        if cand:
            picks.add(cand[len(cand) // 2])
    for i, ((pos, rad), reso) in enumerate(reqs):
        if i in picks:
            print("  request %d (%d atoms, %.1f A) against the oracle" % (i, len(rad), reso), flush=True)
            _same_mesh(served[i], orc.compute_ses(pos, rad, reso, 1), "request %d" % i)
        else:
            _same_mesh(served[i], qs.compute_ses(pos, rad, reso, 1), "request %d" % i)
    assert len(picks) >= 4
