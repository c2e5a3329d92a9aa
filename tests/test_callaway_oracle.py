"""Linearized Callaway scattering on the host and in the oracle.

- The host's CallawayData (capacities, wavevectors, N and resistive rates at
  T_bg over the active bank with stationary modes) equals the reference's
  PhononMaterial::callaway_data bit for bit (SHA-256 of every array,
  tests/golden/callaway_c1c.npz), on the odd 15^3 table and on the even 16^3
  table, whose zone-boundary q points the equal-weight first-BZ completion
  turns into eight copies each (complete_callaway_bz,
  src/material/BzQuadrature.cpp:11-161; tests/golden/callaway_c1c16.npz).
- The oracle's CarrierCollisionOperator restatement reproduces the reference's
  occupations on three random carrier sets (dt 1, 3 and 12 ps; the last takes
  two Taylor substeps).
- Independently of the reference: on a small cell the oracle equals
  scipy.linalg.expm of the explicit generator; normal scattering conserves
  energy and crystal momentum, resistive scattering conserves energy only.
"""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest
import scipy.linalg

from oracle import oracle_ctypes as O
from phonomc_amd import runtime as rt

HERE = os.path.dirname(os.path.abspath(__file__))
HBAR = 6.582119569e-4


TABLES = {"si15": ("synthetic:si15:callaway", "callaway_c1c.npz"),
          "si16": ("synthetic:si16:callaway", "callaway_c1c16.npz")}


def problem_of(table):
    spec = TABLES[table][0]
    return rt.Problem(rt.config_text("c1c_callaway.toml", **{"io.material_folder": f'"{spec}"'}), mode_sampling=1)


@pytest.fixture(scope="module")
def c1c():
    return problem_of("si15")


@pytest.fixture(scope="module", params=sorted(TABLES))
def table(request):
    return problem_of(request.param), np.load(os.path.join(HERE, "golden", TABLES[request.param][1]))


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(HERE, "golden", "callaway_c1c.npz"))


def host_table(problem):
    _, m, _, _, _, _ = problem.views()
    na = m.n_active
    cw = problem.callaway()
    return {"omega": rt.arr(m.omega, na), "capacity": rt.arr(cw.capacity, na),
            "wavevector": rt.arr(cw.wavevector, 3 * na).reshape(na, 3),
            "normal_rate": rt.arr(cw.normal_rate, na), "resistive_rate": rt.arr(cw.resistive_rate, na)}


def test_callaway_table_matches_reference(table):
    problem, golden = table
    t = host_table(problem)
    assert len(t["omega"]) == int(golden["active_count"])
    for k, v in t.items():
        assert hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() == str(golden[f"digest_{k}"]), k
        np.testing.assert_array_equal(v[golden["rows"]], golden[f"rows_{k}"])
    # Stationary modes are active under Callaway: the optical branches at Gamma
    # (15^3), and on the even mesh also the zone-boundary modes of the sine
    # dispersion, each in all of its first-BZ copies.
    _, m, _, _, _, _ = problem.views()
    v = rt.arr(m.velocity, 3 * m.n_active).reshape(-1, 3)
    copies = m.n_flat // (int(problem.material().qpoints) * int(problem.material().branches))
    stationary = int(np.sum(np.linalg.norm(v, axis=1) <= 1e-12))
    assert stationary == 3 if copies == 1 else (stationary > 3 * copies and stationary % copies == 0)


def test_even_mesh_completion_counts():
    """16^3 mesh on a cubic reciprocal lattice: q points with one, two or three
    components at the zone boundary (1/2) have 2, 4 or 8 shortest
    representatives, so every point is kept 8 times (lcm), the representatives
    of each in equal numbers."""
    p = problem_of("si16")
    _, m, _, _, _, _ = p.views()
    raw = p.material()
    assert m.n_flat == 8 * raw.qpoints * raw.branches
    assert m.n_active == 8 * (raw.qpoints * raw.branches) - 3 * 8  # acoustic Gamma modes stay inactive


@pytest.mark.parametrize("tag", ["a", "b", "c"])
def test_callaway_collisions_match_reference(table, tag):
    problem, golden = table
    occ, residual = O.callaway_apply(problem, golden[f"{tag}_active"], golden[f"{tag}_cell"], golden[f"{tag}_occ"],
                                     float(golden[f"{tag}_dt"]))
    np.testing.assert_allclose(occ, golden[f"{tag}_occ_after"], rtol=1e-14, atol=0)
    assert residual == pytest.approx(float(golden[f"{tag}_residual"]), abs=1e-9)
    # The collision changed the occupations (not a no-op).
    assert np.abs(occ - golden[f"{tag}_occ"]).max() > 1e-6


def custom_apply(problem, active, occ, dt, normal, resistive):
    """orc_callaway_apply on one cell with substituted rate tables."""
    _, m, _, _, _, _ = problem.views()
    cw = problem.callaway()
    na = m.n_active
    cap = rt.arr(cw.capacity, na)
    q = rt.arr(cw.wavevector, 3 * na)
    t = rt.CallawayTable()
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (cap, q, normal, resistive)]
    t.capacity, t.wavevector, t.normal_rate, t.resistive_rate = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in keep]
    t.heat_capacity_density = cw.heat_capacity_density
    t.background_energy_density = cw.background_energy_density
    act = np.ascontiguousarray(active, dtype=np.int32)
    cell = np.zeros(len(act), dtype=np.int32)
    out = np.array(occ, dtype=np.float64)
    status = C.c_int()
    O.lib().orc_callaway_apply(C.byref(m), C.byref(t), 1, len(act), act.ctypes.data_as(C.POINTER(C.c_int32)),
                               cell.ctypes.data_as(C.POINTER(C.c_int32)), out.ctypes.data_as(C.POINTER(C.c_double)),
                               dt, m.energy_weight, C.byref(status))
    assert status.value == 0
    return out


def small_cell(problem, n, seed):
    _, m, _, _, _, _ = problem.views()
    na = m.n_active
    rng = np.random.default_rng(seed)
    act = rng.choice(na, n, replace=False).astype(np.int32)
    neq = rt.arr(m.neq_background, na)
    occ = neq[act] * (1.0 + 0.05 * rng.normal(size=n))
    return act, occ


def test_callaway_equals_matrix_exponential(c1c):
    """exp(dt L) by scipy on the explicit generator of a 9-carrier cell:
    L = -diag(g_N + g_R) + U_N U_N^T + U_R U_R^T in x = e / sqrt(C), U spanning
    sqrt(g) times the energy and crystal-momentum vectors (orthonormalized by QR)."""
    t = host_table(c1c)
    _, m, _, _, _, _ = c1c.views()
    neq = rt.arr(m.neq_background, m.n_active)
    act, occ = small_cell(c1c, 9, 3)
    gn, gr = t["normal_rate"] * 3.0, t["resistive_rate"] * 3.0  # faster rates: more Taylor terms
    dt = 2.5
    w, cap, q = t["omega"][act], t["capacity"][act], t["wavevector"][act]
    sc = np.sqrt(cap)
    B = np.stack([sc, sc * q[:, 0] / w, sc * q[:, 1] / w, sc * q[:, 2] / w], axis=1)

    def gains(rate, basis):
        Q, R = np.linalg.qr(np.sqrt(rate[act])[:, None] * basis)
        keep = np.abs(np.diag(R)) > 1e-12 * np.abs(R).max()
        return np.sqrt(rate[act])[:, None] * Q[:, keep]

    UN, UR = gains(gn, B), gains(gr, B[:, :1])
    L = -np.diag(gn[act] + gr[act]) + UN @ UN.T + UR @ UR.T
    e = HBAR * w * (occ - neq[act])
    x = scipy.linalg.expm(dt * L) @ (e / sc)
    want = neq[act] + (x * sc) / (HBAR * w)
    got = custom_apply(c1c, act, occ, dt, gn, gr)
    np.testing.assert_allclose(got - neq[act], want - neq[act], rtol=1e-10, atol=1e-13)


def test_callaway_conservation_laws(c1c):
    """Normal scattering alone keeps energy and crystal momentum; resistive
    scattering alone keeps energy and relaxes the momentum."""
    t = host_table(c1c)
    _, m, _, _, _, _ = c1c.views()
    neq = rt.arr(m.neq_background, m.n_active)
    act, occ = small_cell(c1c, 400, 5)
    w, q = t["omega"][act], t["wavevector"][act]
    zero = np.zeros_like(t["normal_rate"])

    def moments(o):
        e = HBAR * w * (o - neq[act])
        return e.sum(), (e[:, None] * q / w[:, None]).sum(axis=0)

    e0, p0 = moments(occ)
    scale = np.abs(HBAR * w * (occ - neq[act])).sum()
    n_only = custom_apply(c1c, act, occ, 5.0, t["normal_rate"] * 10, zero)
    e1, p1 = moments(n_only)
    assert abs(e1 - e0) < 1e-12 * scale
    assert np.abs(p1 - p0).max() < 1e-12 * np.abs(p0).max()
    assert np.abs(n_only - occ).max() > 1e-4  # it did scatter
    r_only = custom_apply(c1c, act, occ, 5.0, zero, t["resistive_rate"] * 10)
    e2, p2 = moments(r_only)
    assert abs(e2 - e0) < 1e-12 * scale
    assert np.abs(p2).max() < 0.9 * np.abs(p0).max()


def test_callaway_inputs_rejected_with_messages():
    base = rt.config_text("c1c_callaway.toml")
    nonlinear = base.replace('model = "callaway"', 'model = "callaway"\ncallaway_formulation = "nonlinear"')
    assert rt.Problem(nonlinear).callaway_nonlinear() is not None
    with pytest.raises(RuntimeError, match="gamma_N and gamma_U"):
        rt.Problem(nonlinear.replace("synthetic:si15:callaway", "synthetic:si15"))
    with pytest.raises(RuntimeError, match="gamma_N and gamma_U"):
        rt.Problem(base.replace("synthetic:si15:callaway", "synthetic:si15"))
