"""The diagnostics comparison tool (counterpart of compare_taylor_green.py):
loaders, interpolation and tolerances, on the literature curve the reference
ships (spectral_Re1600_512.dat) and on synthetic CSVs."""
import os

import numpy as np
import pytest

from spectral_solver_amd import compare

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DAT = os.path.join(GOLDEN, "spectral_Re1600_512.dat")


def write_csv(path, t, e, eps, ens=None):
    cols = ["t", "kinetic_energy", "dissipation"] + (["enstrophy"] if ens is not None else [])
    with open(path, "w") as f:
        f.write(",".join(cols) + "\n")
        for i in range(len(t)):
            row = [t[i], e[i], eps[i]] + ([ens[i]] if ens is not None else [])
            f.write(",".join(repr(float(v)) for v in row) + "\n")


def test_reference_loader_reads_the_literature_curve():
    ref = compare.load_reference(DAT)
    assert ref["t"].shape == (2000,)
    assert ref["t"][0] == 0.0 and abs(ref["t"][-1] - 19.99) < 1e-12
    assert np.all(np.diff(ref["t"]) > 0)
    # first row: E = 1/8, enstrophy = 3/8 (the Taylor-Green closed forms)
    assert ref["kinetic_energy"][0] == 0.125 and ref["enstrophy"][0] == 0.375
    # dissipation is 2 nu enstrophy at Re = 1600 to the file's printed digits
    assert np.max(np.abs(ref["dissipation"] - ref["enstrophy"] / 800.0) / ref["dissipation"]) < 3e-3


def test_reference_loader_rejects_narrow_tables(tmp_path):
    p = tmp_path / "bad.dat"
    p.write_text("0 1 2\n1 2 3\n")
    with pytest.raises(ValueError, match="needs four"):
        compare.load_reference(p)


def test_csv_loader_enstrophy_fallback_and_missing_columns(tmp_path):
    p = tmp_path / "a.csv"
    write_csv(p, [0.0, 1.0], [0.125, 0.12], [4.6875e-4, 5e-4])
    d = compare.load_ours(p, 1600.0)
    assert np.allclose(d["enstrophy"], np.array([4.6875e-4, 5e-4]) * 800.0)
    q = tmp_path / "b.csv"
    q.write_text("t,kinetic_energy\n0,0.125\n")
    with pytest.raises(ValueError, match="dissipation"):
        compare.load_ours(q, 1600.0)
    one = tmp_path / "c.csv"
    write_csv(one, [0.0], [0.125], [4.6875e-4], [0.375])
    assert compare.load_ours(one, 1600.0)["t"].shape == (1,)


def test_identical_series_and_planted_deviation(tmp_path):
    ref = compare.load_reference(DAT)
    sel = slice(0, 2000, 50)  # every 0.5 time units, exactly on reference rows
    p = tmp_path / "same.csv"
    write_csv(p, ref["t"][sel], ref["kinetic_energy"][sel], ref["dissipation"][sel], ref["enstrophy"][sel])
    devs = compare.compare(compare.load_ours(p, 1600.0), ref)
    assert all(d.max_rel < 1e-15 and d.points == 40 for d in devs)
    # a 2% error planted in the dissipation at one time is found and located
    eps = ref["dissipation"][sel].copy()
    eps[18] *= 1.02
    q = tmp_path / "planted.csv"
    write_csv(q, ref["t"][sel], ref["kinetic_energy"][sel], eps, ref["enstrophy"][sel])
    devs = {d.quantity: d for d in compare.compare(compare.load_ours(q, 1600.0), ref)}
    assert abs(devs["dissipation"].max_rel - 0.02) < 1e-12
    assert devs["dissipation"].at_t == ref["t"][sel][18]
    assert devs["kinetic_energy"].max_rel < 1e-15


def test_interpolation_between_reference_rows(tmp_path):
    # a linear reference is interpolated exactly at off-grid times
    rp = tmp_path / "lin.dat"
    t = np.linspace(0.0, 2.0, 21)
    np.savetxt(rp, np.column_stack([t, 1 + t, 2 + t, 3 + t]))
    ts = np.array([0.05, 0.333, 1.999, 2.5])  # the last lies outside and is skipped
    p = tmp_path / "o.csv"
    write_csv(p, ts, 1 + ts, 2 + ts, 3 + ts)
    devs = compare.compare(compare.load_ours(p, 1600.0), compare.load_reference(rp))
    assert all(d.max_rel < 1e-15 and d.points == 3 for d in devs)
    devs = compare.compare(compare.load_ours(p, 1600.0), compare.load_reference(rp), t_max=1.0)
    assert all(d.points == 2 for d in devs)


def test_command_line_exit_codes(tmp_path, capsys):
    ref = compare.load_reference(DAT)
    sel = slice(0, 2000, 100)
    p = tmp_path / "x.csv"
    eps = ref["dissipation"][sel] * 1.01
    write_csv(p, ref["t"][sel], ref["kinetic_energy"][sel], eps)
    assert compare.main(["--ours", str(p), "--reference", DAT, "--rtol-dissipation", "2e-2",
                         "--rtol-enstrophy", "5e-2"]) == 0
    assert compare.main(["--ours", str(p), "--reference", DAT, "--rtol-dissipation", "5e-3"]) == 1
    assert "FAIL" in capsys.readouterr().out
    assert compare.main(["--ours", str(tmp_path / "missing.csv"), "--reference", DAT]) == 2
