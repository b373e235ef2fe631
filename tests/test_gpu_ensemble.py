"""GPU parity of the batched ensemble: every member against its own oracle
solve (the reference's RK3 step, src/ava_taylor_green_impl.hpp:976-1039, from
the member's random initial field and viscosity), FP64 to 1e-12 relative L2,
FP32 against FP64, graph against direct launches, and the 64-member
configuration (16 each of 32^3, 48^3, 64^3, 96^3, FP32, nu = 1e-2, 100 steps).
"""
import time

import numpy as np
import pytest

import oracle
import spectral_solver_amd as sph

pytestmark = pytest.mark.gpu


def member_oracle(n, nu, seed, energy=0.5):
    o = oracle.TaylorGreenOracle(n, 1.0 / nu, initialise=False)
    u = oracle.random_solenoidal_field(n, seed, energy)
    for c in range(3):
        o.set_u_hat(c, oracle.half_to_c2c(u[c]))
    return o


def member_error(ens, j, o):
    num = den = 0.0
    for c in range(3):
        ref = oracle.c2c_to_half(o.u_hat(c))
        num += np.linalg.norm(ens.u_hat(j, c) - ref) ** 2
        den += np.linalg.norm(ref) ** 2
    return np.sqrt(num / den)


def cfl_dt(ens, cfl=0.5):
    return np.array([cfl * (2 * np.pi / n) / ens.max_velocity(j) for j, n in enumerate(ens.sizes)])


def test_initial_field_on_device_is_the_host_field():
    print("case ensemble ic", flush=True)
    ens = sph.Ensemble([16, 24, 16], [0.01, 0.02, 0.03], seeds=[5, 6, 7], precision="fp64")
    for j, (n, s) in enumerate(zip(ens.sizes, [5, 6, 7])):
        ref = sph.random_solenoidal_field(n, s)
        for c in range(3):
            assert np.array_equal(ens.u_hat(j, c), ref[c])
        d = ens.diagnostics(j)
        assert abs(d.kinetic_energy - 0.5) < 1e-13
        assert d.max_spectral_divergence < 1e-10


def test_fp64_members_match_their_oracle_solves():
    """Uneven sizes, per-member viscosity and dt (changing between steps)."""
    sizes = [24, 32, 24, 48, 32]
    nus = [0.01, 0.02, 0.005, 0.01, 1.0 / 64]
    seeds = [1000 + j for j in range(len(sizes))]
    print("case ensemble fp64", sizes, flush=True)
    ens = sph.Ensemble(sizes, nus, seeds, precision="fp64")
    orcs = [member_oracle(n, nu, s) for n, nu, s in zip(sizes, nus, seeds)]
    dts = [np.array([2e-3, 3e-3, 1e-3, 2e-3, 2.5e-3]), np.array([1e-3, 3e-3, 1.5e-3, 2e-3, 2e-3])]
    for k in range(3):
        dt = dts[k % 2]
        ens.step_rk3(dt)
        for o, d in zip(orcs, dt):
            o.step_rk3(float(d))
    for j, o in enumerate(orcs):
        err = member_error(ens, j, o)
        assert err < 1e-12, (j, err)
        dg, do = ens.diagnostics(j, 0.0), o.diagnostics(0.0)
        for key in ("kinetic_energy", "enstrophy", "max_velocity"):
            assert abs(getattr(dg, key) - do[key]) <= 1e-12 * abs(do[key]), (j, key)
        assert abs(ens.max_velocity(j) - o.max_velocity()) <= 1e-12 * o.max_velocity()


@pytest.mark.parametrize("fused", ["0", "1", "2"])
def test_graph_and_direct_launches_agree_bitwise(monkeypatch, fused):
    # the launch structure (per-size chains, one launch per pass or per
    # stage, whose tiles round differently in FP32) is forced: left to the
    # timing at creation, two ensembles may choose differently; one launch
    # per stage is never captured, so there both run directly
    print("case ensemble graph, launch structure:", fused, flush=True)
    monkeypatch.setenv("SPH_ENSEMBLE_FUSED", fused)
    sizes, nus, seeds = [32, 48, 32], [0.01, 0.01, 0.02], [1, 2, 3]
    a = sph.Ensemble(sizes, nus, seeds, precision="fp32")
    b = sph.Ensemble(sizes, nus, seeds, precision="fp32")
    b.set_graph(False)
    for _ in range(3):
        a.step_rk3([4e-3, 3e-3, 4e-3])
        b.step_rk3([4e-3, 3e-3, 4e-3])
    assert a.launch_mode()[0] == b.launch_mode()[0] == int(fused)
    for j in range(3):
        for c in range(3):
            assert np.array_equal(a.u_hat(j, c), b.u_hat(j, c))


def test_time_passes_leaves_state_unchanged():
    ens = sph.Ensemble([32, 48], [0.01, 0.01], precision="fp32")
    ens.step_rk3([1e-2, 1e-2])
    before = [ens.u_hat(j, c) for j in range(2) for c in range(3)]
    t = ens.time_passes(2)
    assert all(ms > 0 and by > 0 for ms, by in t.values())
    after = [ens.u_hat(j, c) for j in range(2) for c in range(3)]
    for x, y in zip(before, after):
        assert np.array_equal(x, y)


def test_fp32_members_against_fp64_oracle():
    sizes = [32, 48, 64, 96]
    print("case ensemble fp32", sizes, flush=True)
    ens = sph.Ensemble(sizes, 1e-2, [1000 + j for j in range(4)], precision="fp32")
    dt = cfl_dt(ens)
    orcs = [member_oracle(n, 1e-2, 1000 + j) for j, n in enumerate(sizes)]
    for _ in range(5):
        ens.step_rk3(dt)
        for o, d in zip(orcs, dt):
            o.step_rk3(float(d))
    for j, o in enumerate(orcs):
        err = member_error(ens, j, o)
        assert err < 2e-6, (j, err)


def test_config_64_members_100_steps():
    """16 each of 32^3, 48^3, 64^3, 96^3; FP32; nu = 1e-2; seeds 1000 + j;
    dt fixed per member from CFL 0.5 at t = 0; 100 RK3 steps."""
    sizes = [32] * 16 + [48] * 16 + [64] * 16 + [96] * 16
    seeds = [1000 + j for j in range(64)]
    print("case ensemble config 64 members", flush=True)
    t0 = time.time()
    ens = sph.Ensemble(sizes, 1e-2, seeds, precision="fp32")
    dt = cfl_dt(ens)
    e0 = [ens.diagnostics(j).kinetic_energy for j in range(64)]
    print(f"  created, dt and E0 of 64 members: {time.time() - t0:.1f} s", flush=True)
    for _ in range(100):
        ens.step_rk3(dt)
    ens.synchronize()
    print(f"  100 steps done: {time.time() - t0:.1f} s", flush=True)
    for j in range(64):
        d = ens.diagnostics(j)
        assert 0.0 < d.kinetic_energy < e0[j], j
        # projection leaves |k . u_hat| at FP32 rounding of |k| |u_hat|
        n = sizes[j]
        sg = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n).astype(float)
        k2 = np.arange(n // 2 + 1)[:, None, None] ** 2 + sg[None, :, None] ** 2 + sg[None, None, :] ** 2
        mag = np.sqrt(k2 * sum(np.abs(ens.u_hat(j, c)) ** 2 for c in range(3)))
        assert d.max_spectral_divergence <= 1e-5 * mag.max(), j
    print(f"  member checks done: {time.time() - t0:.1f} s", flush=True)
    # the first 32^3 and 48^3 members after 100 steps against the FP64 oracle
    for j in (0, 16):
        o = member_oracle(sizes[j], 1e-2, seeds[j])
        for k in range(100):
            o.step_rk3(float(dt[j]))
            if k % 20 == 19:
                print(f"  oracle member {j}: {k + 1} steps, {time.time() - t0:.1f} s", flush=True)
        err = member_error(ens, j, o)
        assert err < 2e-5, (j, err)
        de = abs(ens.diagnostics(j).dissipation - o.diagnostics(0.0)["dissipation"])
        assert de <= 1e-4 * o.diagnostics(0.0)["dissipation"], j


@pytest.mark.parametrize("mode", ["1", "2"])
@pytest.mark.parametrize("layout", ["0", "2"])
def test_one_launch_per_pass_matches_oracle(monkeypatch, layout, mode):
    """Every fused size (16, 32, 48, 64, 96, 128) in one launch per pass
    (mode 1) or per stage (mode 2), plus a size outside that set (24) on its
    own chain beside them; FP64 members against their oracle solves. Layout
    2 puts the groups of 64, 96 and 128 in the ky-blocked layout (the y
    tiles then move pairs between it and the x pass's fields)."""
    monkeypatch.setenv("SPH_ENSEMBLE_FUSED", mode)
    monkeypatch.setenv("SPH_LAYOUT", layout)
    sizes = [128, 16, 24, 32, 48, 64, 96, 32]
    nus = [0.01, 0.02, 0.01, 0.005, 0.01, 0.02, 0.01, 0.03]
    seeds = [2000 + j for j in range(len(sizes))]
    print("case ensemble launch structure", mode, "fp64 layout", layout, sizes, flush=True)
    ens = sph.Ensemble(sizes, nus, seeds, precision="fp64")
    assert ens.launch_mode()[0] == int(mode)
    orcs = [member_oracle(n, nu, s) for n, nu, s in zip(sizes, nus, seeds)]
    dt = [1e-3 * (1 + 0.1 * j) for j in range(len(sizes))]
    for _ in range(3):
        ens.step_rk3(dt)
        for o, d in zip(orcs, dt):
            o.step_rk3(float(d))
    for j, o in enumerate(orcs):
        err = member_error(ens, j, o)
        print(f"  member {j} ({sizes[j]}^3): rel L2 vs oracle {err:.2e}", flush=True)
        assert err < 1e-12, (j, err)


@pytest.mark.parametrize("fused", ["1", "2"])
@pytest.mark.parametrize("layout", ["0", "2"])
@pytest.mark.parametrize("precision,tol", [("fp32", 1e-6), ("fp64", 1e-14)])
def test_one_launch_per_pass_agrees_with_per_size_chains(monkeypatch, precision, tol, layout, fused):
    monkeypatch.setenv("SPH_LAYOUT", layout)
    sizes = [32, 48, 64, 96] * 2
    seeds = [3000 + j for j in range(len(sizes))]
    print("case ensemble launch structures", fused, "vs 0", precision, "layout", layout, flush=True)
    states = []
    for mode in (fused, "0"):
        monkeypatch.setenv("SPH_ENSEMBLE_FUSED", mode)
        ens = sph.Ensemble(sizes, 1e-2, seeds, precision=precision)
        assert ens.launch_mode()[0] == int(mode)
        dt = cfl_dt(ens)
        for _ in range(4):
            ens.step_rk3(dt)
        states.append([[ens.u_hat(j, c) for c in range(3)] for j in range(len(sizes))])
        ens.close()
    for j in range(len(sizes)):
        num = sum(np.linalg.norm(states[0][j][c] - states[1][j][c]) ** 2 for c in range(3))
        den = sum(np.linalg.norm(states[1][j][c]) ** 2 for c in range(3))
        assert np.sqrt(num / den) <= tol, (j, np.sqrt(num / den))


def test_launch_mode_reports_every_timing():
    ens = sph.Ensemble([32, 48, 64, 96], 1e-2, [1, 2, 3, 4], precision="fp32")
    mode, ms = ens.launch_mode()
    print("  launch structure:", mode, "step ms per-size chains / one launch per pass / per stage:", ms, flush=True)
    assert all(x > 0 for x in ms)
    assert mode == int(np.argmin(ms))
