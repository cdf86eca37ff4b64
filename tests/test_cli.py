"""Command-line drop-in: PDB reading, OBJ layout and flag handling (host only), and on a GPU
the whole tool against the library API."""
import json
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PDB = os.path.join(HERE, "golden", "mini.pdb")


def _expected_atoms():
    with open(os.path.join(HERE, "golden", "mini_expected.json")) as f:
        e = np.array(json.load(f)["atoms"], dtype=np.float64)
    return e[:, :3].astype(np.float32), e[:, 3].astype(np.float32)


def test_pdb_reader_follows_reference_rules(qs):
    pos, rad = qs.read_pdb(PDB)
    epos, erad = _expected_atoms()
    # ATOM records only, up to ENDMDL; alt-loc B/C dropped; radius by the element's first letter
    assert pos.shape == epos.shape == (42, 3)
    assert np.array_equal(pos, epos) and np.array_equal(rad, erad)


def test_pdb_reader_missing_file(qs, tmp_path):
    with pytest.raises(qs.QuickSESError, match="can not be read"):
        qs.read_pdb(str(tmp_path / "absent.pdb"))


def test_pdb_reader_empty_and_end_first(qs, tmp_path):
    p = tmp_path / "e.pdb"
    p.write_text("END\nATOM      1 CA   GLY A   1       1.000   2.000   3.000  1.00 20.00           C\n")
    pos, rad = qs.read_pdb(str(p))
    assert pos.shape == (0, 3) and rad.shape == (0,)


def test_obj_layout(qs, tmp_path):
    V = np.array([[0.0, 1.0, 2.0], [1.23449, -0.0004, 1e3], [2.5, 2.25, -7.0]], dtype=np.float32)
    T = np.array([[0, 1, 2], [2, 1, 0]], dtype=np.int32)
    out = tmp_path / "m.obj"
    qs.write_obj(str(out), V, T)
    expected = ("v 0.000 1.000 2.000\n"
                "v 1.234 -0.000 1000.000\n"
                "v 2.500 2.250 -7.000\n"
                "\n"
                "f 2 1 3\n"
                "f 2 3 1\n")
    assert out.read_text() == expected


def _cli(*args):
    import quickses_amd as qs

    if not os.path.exists(qs.CLI_PATH):
        pytest.skip("command-line tool not built")
    return subprocess.run([qs.CLI_PATH] + list(args), stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=120)


def test_cli_flags(tmp_path):
    r = _cli("-h")
    assert r.returncode == 0 and "-v[voxel size]" in r.stderr
    r = _cli("-i", PDB)
    assert r.returncode == 255 and "Usage" in r.stderr
    r = _cli("-i", PDB, "-o", str(tmp_path / "x.obj"), "-v", "fine")
    assert r.returncode == 255 and "invalid value" in r.stderr
    r = _cli("-x")
    assert r.returncode == 255
    r = _cli("-i", str(tmp_path / "absent.pdb"), "-o", str(tmp_path / "x.obj"))
    assert r.returncode == 2 and "can not be read" in r.stderr


def test_cli_without_gpu_fails_cleanly(qs, tmp_path):
    from quickses_amd import hipmem

    if hipmem.device_count() > 0:
        pytest.skip("a GPU is present")
    r = _cli("-i", PDB, "-o", str(tmp_path / "x.obj"))
    assert r.returncode == 1 and "SES computation failed" in r.stderr


def _read_obj(path):
    V, F = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("v "):
                V.append([float(t) for t in line.split()[1:]])
            elif line.startswith("f "):
                F.append([int(t) for t in line.split()[1:]])
    return np.array(V, dtype=np.float64).reshape(-1, 3), np.array(F, dtype=np.int64).reshape(-1, 3)


@pytest.mark.gpu
@pytest.mark.parametrize("flags,voxel,smooth", [((), 0.5, 1), (("-v", "0.3", "-l", "2", "-s", "120"), 0.3, 2),
                                                (("-v0.25", "-l0"), 0.25, 0)])
def test_cli_equals_api(qs, tmp_path, flags, voxel, smooth):
    out = tmp_path / "m.obj"
    print("case cli %s" % (flags,), flush=True)
    r = _cli("-i", PDB, "-o", str(out), *flags)
    assert r.returncode == 0, r.stderr
    V, F = _read_obj(out)
    pos, rad = qs.read_pdb(PDB)
    Va, Ta, _ = qs.compute_ses(pos, rad, voxel, smooth)
    assert len(Va) > 0 and V.shape == Va.shape and F.shape == Ta.shape
    assert np.array_equal(V, np.array([[float("%.3f" % c) for c in v] for v in Va]))
    assert np.array_equal(F, Ta[:, [1, 0, 2]] + 1)
