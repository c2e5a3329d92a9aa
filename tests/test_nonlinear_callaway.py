"""Nonlinear Callaway scattering (scattering.callaway_formulation = 'nonlinear')
on the host, in the oracle and in the device's cell routine built for the CPU.

Against the reference (tests/golden/callaway_nl_c1cn.npz, written from the
reference built from its sources):
- rates: PhononMaterial::mode_callaway_rates at temperatures inside and at the
  ends of the table, and its refusal outside; from the host's tables through
  the oracle and through the device's rate bracket, bit for bit;
- single cells: nonlinear_callaway_evolve on 36 cells of 1 to 2000 carriers
  (three temperatures, dt 1 and 4 ps, occupations 5 % and 40 % off Bose): the
  oracle's long-double restatement to 1e-13 relative, and the device's FP64
  cell routine (pmc_nlcallaway.hpp, one emulated wavefront with the DPP
  summation order) to 1e-12 relative;
- NonlinearCallawayCollision::apply on 6000 carriers in 20 cells with local and
  fixed rate temperatures: occupations and the lifetime residual.

Independently of the reference, on the oracle and the device routine: a cell
keeps its energy; the normal channel alone (zero resistive rates) also keeps
its crystal momentum; a displaced Bose population is a fixed point of the
normal channel; an equilibrium population is a fixed point of the whole step;
flipping the sign of every wavevector changes nothing; and a planted mistake
(the momentum constraints dropped) moves the result far outside the tolerance.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle_ctypes as O
from phonomc_amd import runtime as rt

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HBAR, KB = 6.582119569e-4, 8.617333262145e-5


@pytest.fixture(scope="module")
def fix():
    return np.load(os.path.join(HERE, "golden", "callaway_nl_c1cn.npz"))


@pytest.fixture(scope="module")
def problem():
    return rt.Problem(rt.config_text("c1cn_nonlinear.toml"), input_directory=rt.CONFIGS, mode_sampling=1)


class Tables:
    def __init__(self, problem):
        s, m, g, i, rough, res = problem.views()
        nl = problem.callaway_nonlinear()
        self.na, self.nt = m.n_active, nl.n_temps
        self.omega = rt.arr(m.omega, self.na)
        self.neq = rt.arr(m.neq_background, self.na)
        self.omega_neq = np.ascontiguousarray(np.stack([self.omega, self.neq], 1).reshape(-1))
        self.q = np.ascontiguousarray(rt.arr(nl.wavevector, 3 * self.na))
        self.temps = np.ascontiguousarray(rt.arr(nl.temps, self.nt))
        self.gn = np.ascontiguousarray(rt.arr(nl.gamma_normal, self.nt * self.na))
        self.gu = np.ascontiguousarray(rt.arr(nl.gamma_umklapp, self.nt * self.na))
        self.ge = np.ascontiguousarray(rt.arr(nl.gamma_extra, self.nt * self.na))


@pytest.fixture(scope="module")
def tables(problem):
    return Tables(problem)


@pytest.fixture(scope="module")
def device_routine(tmp_path_factory):
    """The device's cell routine (phonomc_amd/csrc/pmc_nlcallaway.hpp) built
    for the host by g++ (tests/cpu_build/nlcallaway_host.cpp)."""
    out = str(tmp_path_factory.mktemp("nlh") / "libnlh.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "phonomc_amd", "csrc"),
                    os.path.join(HERE, "cpu_build", "nlcallaway_host.cpp"), "-o", out], check=True)
    L = C.CDLL(out)
    d, P, I = C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.nlh_cell.argtypes = [C.c_int, I, P, P, P, C.c_int, C.c_int, P, P, P, P, d, d, d, d, P]
    L.nlh_rates.argtypes = [C.c_int, I, C.c_int, C.c_int, P, P, P, P, d, P, P]
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def device_cell(L, t, active, occ, T, rate_T, dt, weight=1.0, q=None):
    """Runs the device routine on one cell; returns (status, occupations, [energy after, residual])."""
    a = np.ascontiguousarray(active, dtype=np.int32)
    n = np.array(occ, dtype=np.float64)
    out = np.zeros(2)
    qq = t.q if q is None else np.ascontiguousarray(q, dtype=np.float64)
    rc = L.nlh_cell(len(a), _ip(a), _dp(n), _dp(t.omega_neq), _dp(qq), t.na, t.nt, _dp(t.temps), _dp(t.gn), _dp(t.gu),
                    _dp(t.ge), T, rate_T, dt, weight, _dp(out))
    return rc, n, out


def device_rates(L, t, T, active):
    a = np.ascontiguousarray(active, dtype=np.int32)
    rn, rr = np.zeros(len(a)), np.zeros(len(a))
    rc = L.nlh_rates(len(a), _ip(a), t.na, t.nt, _dp(t.temps), _dp(t.gn), _dp(t.gu), _dp(t.ge), T, _dp(rn), _dp(rr))
    return rc, rn, rr


def bose(omega, T):
    return 1.0 / np.expm1(HBAR * omega / (KB * T))


def test_host_tables_select_nonlinear_conservative_transport(problem):
    s, m, g, i, rough, res = problem.views()
    assert problem.callaway() is None and problem.callaway_nonlinear() is not None
    assert s.conservative_walls == 1 and s.refill_policy == 1
    # the bank holds the stationary modes: every positive frequency of the
    # (odd-mesh, uncompleted) table
    mv = problem.material()
    f = rt.arr(mv.frequency_thz, mv.qpoints * mv.branches)
    assert m.n_active == int(np.count_nonzero(f > 0))
    # Bose (not linearized) reservoir occupations
    for r in res:
        occ = rt.arr(r.occupation, m.n_active)
        w = rt.arr(m.omega, m.n_active)
        np.testing.assert_allclose(occ, bose(w, r.temperature), rtol=1e-12)


def test_rates_match_reference(fix, problem, tables, device_routine):
    act = fix["rates_active"]
    for k, T in enumerate(fix["rates_t"]):
        st, rn, rr = O.nl_callaway_rates(problem, float(T), act)
        assert st == 0
        assert np.array_equal(rn, fix["rates_normal"][k]) and np.array_equal(rr, fix["rates_resistive"][k])
        st, dn, dr = device_rates(device_routine, tables, float(T), act)
        assert st == 0
        assert np.array_equal(dn, fix["rates_normal"][k]) and np.array_equal(dr, fix["rates_resistive"][k])
    for T in fix["rates_outside"]:
        assert O.nl_callaway_rates(problem, float(T), act[:3])[0] == 7
        assert device_rates(device_routine, tables, float(T), act[:3])[0] == 7


def cells(fix):
    for k in range(int(fix["cells"])):
        yield (fix[f"cell{k}_active"], fix[f"cell{k}_occ"], float(fix[f"cell{k}_t"]), float(fix[f"cell{k}_dt"]),
               fix[f"cell{k}_normal"], fix[f"cell{k}_resistive"], fix[f"cell{k}_after"])


def test_oracle_cells_match_reference(fix, tables):
    worst = 0.0
    for a, occ, T, dt, rn, rr, ref in cells(fix):
        rc, out = O.nl_callaway_evolve(tables.omega[a], tables.q.reshape(-1, 3)[a], rn, rr, occ, T, dt)
        assert rc == 0
        worst = max(worst, float(np.max(np.abs(out - ref) / np.maximum(ref, 1e-300))))
    assert worst <= 1e-13, worst


def test_device_cell_routine_matches_reference(fix, tables, device_routine):
    """FP64 device arithmetic (the reference is long double): 1e-12 relative."""
    worst, moved = 0.0, 0.0
    for a, occ, T, dt, rn, rr, ref in cells(fix):
        rc, out, _ = device_cell(device_routine, tables, a, occ, T, T, dt)
        assert rc == 0
        worst = max(worst, float(np.max(np.abs(out - ref) / np.maximum(ref, 1e-300))))
        moved = max(moved, float(np.max(np.abs(ref - occ) / np.maximum(occ, 1e-300))))
    assert worst <= 1e-12, worst
    assert moved > 0.1  # the cells do relax


def test_oracle_apply_matches_reference(fix, problem):
    for tag in ("a", "b"):
        st, occ, res = O.nl_callaway_apply(problem, fix[f"{tag}_cell_t"], fix[f"{tag}_active"], fix[f"{tag}_cell"],
                                           fix[f"{tag}_occ"], 1.0, int(fix[f"{tag}_local"]),
                                           float(fix[f"{tag}_lifetime_t"]))
        assert st == 0
        ref = fix[f"{tag}_after"]
        assert np.max(np.abs(occ - ref) / np.maximum(ref, 1e-300)) <= 1e-13
        # the residual is a sum of roundoff-sized differences: compare on the
        # scale of the energy the collision moved
        hw = HBAR * rt.arr(problem.views()[1].omega, problem.views()[1].n_active)[fix[f"{tag}_active"]]
        moved = np.sum(np.abs(hw * (ref - fix[f"{tag}_occ"])))
        assert abs(res - float(fix[f"{tag}_residual"])) <= 1e-12 * moved


def test_device_routine_apply_matches_reference(fix, problem, tables, device_routine):
    """The per-cell driver of the device kernel: carriers of a cell in index
    order, rates at the cell or the fixed lifetime temperature."""
    w = problem.views()[1].energy_weight
    for tag in ("a", "b"):
        act, cell, occ = fix[f"{tag}_active"], fix[f"{tag}_cell"], fix[f"{tag}_occ"].copy()
        ct, local, lt = fix[f"{tag}_cell_t"], int(fix[f"{tag}_local"]), float(fix[f"{tag}_lifetime_t"])
        residual = 0.0
        for c in range(len(ct)):
            idx = np.nonzero(cell == c)[0]
            rc, out, sums = device_cell(device_routine, tables, act[idx], occ[idx], float(ct[c]),
                                        float(ct[c]) if local else lt, 1.0, w)
            assert rc == 0
            occ[idx] = out
            residual += sums[1]
        ref = fix[f"{tag}_after"]
        assert np.max(np.abs(occ - ref) / np.maximum(ref, 1e-300)) <= 1e-12
        hw = HBAR * tables.omega[act]
        moved = np.sum(np.abs(hw * (ref - fix[f"{tag}_occ"]))) * w
        assert abs(residual - float(fix[f"{tag}_residual"])) <= 1e-11 * moved


def random_cell(tables, rng, n, T, spread):
    a = rng.integers(0, tables.na, n).astype(np.int32)
    occ = np.maximum(bose(tables.omega[a], T) * (1.0 + spread * rng.standard_normal(n)), 0.0)
    return a, occ


def rates_at(problem, T, a):
    st, rn, rr = O.nl_callaway_rates(problem, T, a)
    assert st == 0
    return rn, rr


def test_conservation_and_fixed_points(problem, tables, device_routine):
    rng = np.random.default_rng(11)
    q = tables.q.reshape(-1, 3)
    T = 320.0
    a, occ = random_cell(tables, rng, 700, T, 0.3)
    rn, rr = rates_at(problem, T, a)
    hw = HBAR * tables.omega[a]
    for name, run in (("oracle", lambda o, r, s=rn: O.nl_callaway_evolve(tables.omega[a], q[a], s, r, o, T, 2.0)),
                      ("device", None)):
        if run is None:
            def run(o, r, s=rn):
                # the device routine reads the tables: emulate zero resistive
                # rates by a zero table
                t2 = Tables.__new__(Tables)
                t2.__dict__.update(tables.__dict__)
                if not r.any():
                    t2.gu = np.zeros_like(tables.gu)
                    t2.ge = np.zeros_like(tables.ge)
                rc, out, _ = device_cell(device_routine, t2, a, o, T, T, 2.0)
                return rc, out
        rc, out = run(occ, rr)
        assert rc == 0, name
        # energy is conserved by every substep
        assert abs(np.sum(hw * (out - occ))) <= 1e-12 * np.sum(hw * occ), name
        # normal scattering alone also keeps the crystal momentum sum q n
        rc, out_n = run(occ, np.zeros_like(rr))
        assert rc == 0
        p0, p1 = q[a].T @ occ, q[a].T @ out_n
        assert np.max(np.abs(p1 - p0)) <= 1e-11 * np.sum(np.abs(q[a]).T @ occ), name
        assert abs(np.sum(hw * (out_n - occ))) <= 1e-12 * np.sum(hw * occ), name
        # a displaced Bose population is a fixed point of normal scattering
        u = np.array([1.5e-3, -0.7e-3, 0.4e-3])  # drift in phi units per 1/A
        disp = 1.0 / np.expm1(HBAR * tables.omega[a] / (KB * T) - q[a] @ u)
        rc, out_d = run(disp, np.zeros_like(rr))
        assert rc == 0
        assert np.max(np.abs(out_d - disp) / disp) <= 1e-11, name
        # and relaxes under resistive scattering
        rc, out_r = run(disp, rr)
        assert np.max(np.abs(out_r - disp) / disp) > 1e-4, name
        # equilibrium is a fixed point of the whole step
        eq = bose(tables.omega[a], T)
        rc, out_e = run(eq, rr)
        assert rc == 0 and np.max(np.abs(out_e - eq) / eq) <= 1e-12, name


def test_wavevector_sign_changes_nothing(problem, tables, device_routine):
    rng = np.random.default_rng(12)
    q = tables.q.reshape(-1, 3)
    T = 280.0
    a, occ = random_cell(tables, rng, 500, T, 0.2)
    rn, rr = rates_at(problem, T, a)
    rc, o1 = O.nl_callaway_evolve(tables.omega[a], q[a], rn, rr, occ, T, 1.0)
    rc2, o2 = O.nl_callaway_evolve(tables.omega[a], -q[a], rn, rr, occ, T, 1.0)
    assert rc == rc2 == 0
    assert np.max(np.abs(o1 - o2) / o1) <= 1e-13
    rc3, o3, _ = device_cell(device_routine, tables, a, occ, T, T, 1.0, q=-tables.q)
    assert rc3 == 0 and np.max(np.abs(o3 - o1) / o1) <= 1e-12


def test_planted_mistake_is_caught(fix, tables, device_routine):
    """Dropping the crystal-momentum constraints of the normal channel (zero
    wavevectors) changes the cells by far more than the tolerance."""
    worst = 0.0
    for a, occ, T, dt, rn, rr, ref in cells(fix):
        if len(a) < 40:
            continue
        rc, out = O.nl_callaway_evolve(tables.omega[a], np.zeros((len(a), 3)), rn, rr, occ, T, dt)
        assert rc == 0
        worst = max(worst, float(np.max(np.abs(out - ref) / np.maximum(ref, 1e-300))))
    assert worst > 1e-6, worst


def test_oracle_step_closes_energy_and_keeps_walls_conservative():
    """Whole oracle steps of the nonlinear in-plane film: strict-elastic walls,
    fixed-flux reservoirs, Bose occupations; the ledger closes every step."""
    p = rt.Problem(rt.config_text("c2cn_nonlinear.toml", **{"simulation.particle_count": 30000}),
                   input_directory=rt.CONFIGS, mode_sampling=1)
    sim = O.Sim(p, threads=4)
    sim.initialize()
    for _ in range(6):
        sim.step(1)
        L = sim.ledger()
        assert abs(L["step_balance_residual_ev"]) <= 1e-12 * L["thermal_energy_ev"]
    L = sim.ledger()
    assert L["rough_events"] > 0 and L["total_injected"] > 0
    occ = sim.carriers()["occupation"]
    assert np.all(occ >= 0.0)
    T, _ = sim.cells()
    assert np.all((T > 295.0) & (T < 305.0))
