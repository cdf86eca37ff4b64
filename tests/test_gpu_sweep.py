"""Seeded sweep over window half-widths 1 to 10 (voxel sizes 0.14 to 1.6 A), molecule sizes and
smoothing passes: every mesh equals the oracle's bit for bit, alone and in mixed batches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RESOS = [0.16, 0.17, 0.18, 0.19, 0.2, 0.22, 0.25, 0.3, 0.35, 0.45, 0.6, 0.75, 1.0, 1.4, 1.6, 0.14]


def _same(a, b):
    return (a[0].shape == b[0].shape and a[1].shape == b[1].shape and np.array_equal(a[1], b[1])
            and np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and np.array_equal(a[2], b[2]))


def _cases(qs):
    rng = np.random.default_rng(2024)
    out = []
    for i in range(24):
        n = int(rng.integers(50, 1500)) if i % 4 else int(rng.integers(50, 400))
        reso = float(RESOS[i % len(RESOS)])
        if reso < 0.2:
            n = min(n, 500)
        smooth = int(rng.integers(0, 3))
        pos, rad = qs.synth_ball(9000 + i, n)
        out.append((pos, rad, reso, smooth))
    return out


def test_sweep_matches_oracle_alone_and_batched(qs, orc):
    cases = _cases(qs)
    got = []
    for i, (pos, rad, reso, smooth) in enumerate(cases):
        print("case sweep %d n=%d reso=%.2f smooth=%d" % (i, len(rad), reso, smooth), flush=True)
        g = qs.compute_ses(pos, rad, reso, smooth)
        assert _same(g, orc.compute_ses(pos, rad, reso, smooth)), i
        got.append(g)
    for sm in (0, 1, 2):
        idx = [i for i, c in enumerate(cases) if c[3] == sm]
        if not idx:
            continue
        print("case sweep batch smooth=%d molecules=%d" % (sm, len(idx)), flush=True)
        out = qs.compute_batch([(cases[i][0], cases[i][1]) for i in idx], [cases[i][2] for i in idx], sm)
        for k, i in enumerate(idx):
            assert _same(out[k], got[i]), i
