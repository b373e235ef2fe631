"""Numeric comparison of Taylor-Green diagnostics against a reference series.

The counterpart of the reference's plotting script compare_taylor_green.py:
the same two loaders (a whitespace ``.dat`` with columns ``t E epsilon
enstrophy``, compare_taylor_green.py:35-44, and the solver's CSV with named
columns and an enstrophy fallback of dissipation / (2 nu) for files without
that column, :47-65), but instead of drawing the curves it interpolates the
reference series at the solver's output times and reports the relative
deviation of each quantity, with a pass/fail against tolerances.

    python -m spectral_solver_amd.compare --ours out/taylor_green_diagnostics.csv \\
        --reference spectral_Re1600_512.dat --rtol-energy 1e-2 --rtol-dissipation 5e-2

exits 0 when every quantity is within its tolerance, 1 otherwise, 2 on
unreadable input.
"""
from __future__ import annotations

import argparse
import dataclasses
import sys
from pathlib import Path

import numpy as np

QUANTITIES = ("kinetic_energy", "dissipation", "enstrophy")


DAT_COLUMNS = ("t",) + QUANTITIES  # column order of the whitespace table
CSV_REQUIRED = ("t", "kinetic_energy", "dissipation")


def load_reference(path) -> dict:
    """Whitespace table whose first four columns are t, E, epsilon and
    enstrophy (compare_taylor_green.py:35-44); further columns are ignored."""
    table = np.loadtxt(path, ndmin=2)
    if table.shape[1] < len(DAT_COLUMNS):
        raise ValueError(f"{path}: {table.shape[1]} columns per row, the table needs four "
                         f"({' '.join(DAT_COLUMNS)})")
    return {name: table[:, i].copy() for i, name in enumerate(DAT_COLUMNS)}


def load_ours(path, reynolds: float) -> dict:
    """The solver's CSV (header row of names). t, kinetic_energy and
    dissipation are required; enstrophy falls back to dissipation / (2 nu)
    with nu = 1 / reynolds (compare_taylor_green.py:47-65)."""
    rec = np.genfromtxt(path, delimiter=",", names=True, ndmin=1)
    have = rec.dtype.names or ()
    absent = [name for name in CSV_REQUIRED if name not in have]
    if absent:
        raise ValueError(f"{path}: header lacks {' and '.join(absent)}")
    col = lambda name: np.atleast_1d(np.asarray(rec[name], dtype=np.float64))
    out = {name: col(name) for name in CSV_REQUIRED}
    # dissipation = 2 nu enstrophy in the solver's output
    out["enstrophy"] = col("enstrophy") if "enstrophy" in have else out["dissipation"] * (0.5 * reynolds)
    return out


@dataclasses.dataclass
class Deviation:
    quantity: str
    max_rel: float  # max over compared times of |ours - ref| / |ref|
    at_t: float  # time of that maximum
    rms_rel: float
    points: int


def compare(ours: dict, ref: dict, t_min: float | None = None, t_max: float | None = None) -> list:
    """Relative deviations of ours from the reference, the reference linearly
    interpolated at our output times inside its own time range (and inside
    [t_min, t_max] when given)."""
    rt = np.asarray(ref["t"], dtype=np.float64)
    if rt.size < 1 or np.any(np.diff(rt) <= 0):
        raise ValueError("reference times must be strictly increasing")
    t = np.asarray(ours["t"], dtype=np.float64)
    sel = (t >= rt[0] - 1e-12) & (t <= rt[-1] + 1e-12)
    if t_min is not None:
        sel &= t >= t_min
    if t_max is not None:
        sel &= t <= t_max
    if not np.any(sel):
        raise ValueError("no output time of ours lies inside the reference's time range")
    ts = t[sel]
    res = []
    for q in QUANTITIES:
        r = np.interp(ts, rt, np.asarray(ref[q], dtype=np.float64))
        rel = np.abs(np.asarray(ours[q], dtype=np.float64)[sel] - r) / np.abs(r)
        i = int(np.argmax(rel))
        res.append(Deviation(q, float(rel[i]), float(ts[i]), float(np.sqrt(np.mean(rel * rel))), int(ts.size)))
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--ours", type=Path, default=Path("output_cpp/taylor_green_diagnostics.csv"),
                    help="CSV written by the Taylor-Green driver")
    ap.add_argument("--reference", type=Path, default=Path("spectral_Re1600_512.dat"),
                    help="whitespace .dat with columns: t E epsilon enstrophy")
    ap.add_argument("--re", type=float, default=1600.0, help="Reynolds number for CSVs without enstrophy")
    ap.add_argument("--t-min", type=float, default=None)
    ap.add_argument("--t-max", type=float, default=None)
    ap.add_argument("--rtol-energy", type=float, default=1e-2)
    ap.add_argument("--rtol-dissipation", type=float, default=5e-2)
    ap.add_argument("--rtol-enstrophy", type=float, default=5e-2)
    a = ap.parse_args(argv)
    try:
        ref = load_reference(a.reference)
        ours = load_ours(a.ours, a.re)
        devs = compare(ours, ref, a.t_min, a.t_max)
    except (OSError, ValueError) as e:
        print(f"error: {e}", file=sys.stderr)
        return 2
    tol = {"kinetic_energy": a.rtol_energy, "dissipation": a.rtol_dissipation, "enstrophy": a.rtol_enstrophy}
    ok = True
    for d in devs:
        good = d.max_rel <= tol[d.quantity]
        ok &= good
        print(f"{d.quantity:15s} max rel {d.max_rel:.3e} at t={d.at_t:g}  rms rel {d.rms_rel:.3e}  "
              f"({d.points} times, tol {tol[d.quantity]:g}) {'ok' if good else 'FAIL'}")
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
