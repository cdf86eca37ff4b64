"""Tuned launch shapes change no result: field pass 1 with any tuned rows per block gives the same
mesh bit for bit, and the tuning file is read as written."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a[0].shape == b[0].shape and a[1].shape == b[1].shape
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_tuned_pass1_rows_keep_the_mesh(qs, tmp_path):
    mols = [(qs.synth_ball(610, 1500), 0.5), (qs.synth_ball(611, 700), 0.2), (qs.synth_ball(612, 900), 1.0)]
    qs.load_tuning(None)
    try:
        ref = [qs.compute_ses(p, r, reso, 1) for (p, r), reso in mols]
        for rows in (1, 3, 8, 64, 256):
            print("case tuned pass-1 rows %d" % rows, flush=True)
            for (p, r), reso in mols:
                qs.set_tuning("pass1_slab_rows", reso, rows)
            for ((p, r), reso), want in zip(mols, ref):
                _same(qs.compute_ses(p, r, reso, 1), want)
        path = tmp_path / "tuning.json"
        path.write_text(json.dumps({"pass1_slab_rows": {"0.5": 4, "0.2": 16}, "serve": {"workers": 2}}))
        qs.load_tuning(None)
        assert qs.load_tuning(str(path)) == 2
        for ((p, r), reso), want in zip(mols, ref):
            _same(qs.compute_ses(p, r, reso, 1), want)
        with pytest.raises(qs.QuickSESError):
            qs.load_tuning(str(tmp_path / "missing.json"))
    finally:
        qs.load_tuning(None)
        qs.load_tuning(qs.TUNING_PATH) if qs.tuning() else None
