"""Seeded random molecules (sizes, voxel sizes, smoothing passes, a planar and a two-atom case)
meshed by the GPU, alone and as one batch, against the CPU oracle, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_mesh(a, b):
    (Va, Ta, Aa), (Vb, Tb, Ab) = a, b
    assert Va.shape == Vb.shape and Ta.shape == Tb.shape, (Va.shape, Vb.shape, Ta.shape, Tb.shape)
    assert np.array_equal(Ta, Tb)
    assert np.array_equal(Va.view(np.uint32), Vb.view(np.uint32))
    assert np.array_equal(Aa, Ab)


def _cases(qs):
    rng = np.random.default_rng(2024)
    out = []
    for i in range(6):
        n = int(rng.integers(50, 2500))
        reso = float(rng.choice([0.25, 0.3, 0.45, 0.6, 0.9, 1.2]))
        smooth = int(rng.integers(0, 3))
        pos, rad = qs.synth_ball(9000 + i, n)
        out.append(("ball%d" % i, pos, rad, reso, smooth))
    # a flat sheet of atoms (one atom layer thick) and two touching atoms
    g = np.stack(np.meshgrid(np.arange(12), np.arange(9), indexing="ij"), -1).reshape(-1, 2) * 2.9
    sheet = np.concatenate([g, np.zeros((len(g), 1))], 1).astype(np.float32)
    out.append(("sheet", sheet, np.full(len(sheet), 1.7, np.float32), 0.35, 1))
    out.append(("pair", np.array([[0, 0, 0], [2.5, 0.4, -0.3]], np.float32), np.array([1.7, 1.55], np.float32), 0.2, 1))
    return out


def test_random_molecules_match_oracle_alone_and_batched(qs, orc):
    cases = _cases(qs)
    for name, pos, rad, reso, smooth in cases:
        print("case %s n=%d reso=%.2f smooth=%d" % (name, len(rad), reso, smooth), flush=True)
        ref = orc.compute_ses(pos, rad, reso, smooth)
        assert len(ref[0]) > 0
        _same_mesh(qs.compute_ses(pos, rad, reso, smooth), ref)
    # the same molecules as one batch (one smoothing count for all: take 1)
    print("case batch of %d" % len(cases), flush=True)
    out = qs.compute_batch([(p, r) for _, p, r, _, _ in cases], [c[3] for c in cases], 1)
    for (name, pos, rad, reso, _), got in zip(cases, out):
        _same_mesh(got, orc.compute_ses(pos, rad, reso, 1))
