"""Newton inversion of the exact Bose-Einstein crystal energy per cell
(pmc_setup.temperature_inversion = 1) against the closed form.

The reference inverts a 0.1 K lookup table of PhononMaterial::crystal_energy_density
(src/material/Thermodynamics.cpp:151-224) by linear interpolation.  With
inversion 1 the device solves sum_a hw_a n_BE(T) / norm + zp = e per cell by
Newton from the table's temperature.  Checked here against a bisection of the
same sum in numpy (to 1e-12 K) on the device's own cell energies, and against
the table: the two agree within the table's interpolation error, and they do
differ (the inversion is really taken)."""
import numpy as np
import pytest

from phonomc_amd import runtime as rt

KB, HBAR = 8.617333262145e-5, 6.582119569e-4


def exact_temperature(omega, norm, zp, e):
    """Bisection of sum hw n_BE(T) / norm + zp = e (n_BE as the reference writes it)."""
    hw = HBAR * omega[omega > 0]

    def energy(t):
        x = hw / (KB * t)
        n = np.where(x > 700.0, 0.0, 1.0 / np.maximum(np.exp(np.minimum(x, 700.0)) - 1.0, 1e-12))
        return float(np.sum(hw * n)) / norm + zp

    lo, hi = 1.0, 2000.0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if energy(mid) < e:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,n", [("c1_cross_plane.toml", 20000), ("c1s_source.toml", 20000)])
def test_newton_inversion_matches_exact_sum(cfg, n):
    p = rt.Problem(rt.config_text(cfg, **{"simulation.particle_count": n}), mode_sampling=1)
    s, m, g, i, rough, res = p.views()
    omega = rt.arr(m.omega, m.n_active)
    newton = rt.Context(p, inversion=1)
    table = rt.Context(p, inversion=0)
    for ctx in (newton, table):
        ctx.step(5)
    tn, en = newton.cells()
    tt, _ = table.cells()
    exact = np.array([exact_temperature(omega, m.energy_normalization, m.zero_point_density, e) for e in en])
    lut = np.array([rt.lib().pmcs_problem_temperature_from_energy(p.h, e) for e in en])
    print(f"{cfg}: max |T_newton - T_exact| = {np.abs(tn - exact).max():.2e} K, "
          f"max |T_newton - T_table| = {np.abs(tn - lut).max():.2e} K")
    assert np.abs(tn - exact).max() < 1e-9
    assert np.abs(tn - lut).max() < 2e-3          # 0.1 K table: interpolation error ~ 1e-4 K
    assert np.abs(tn - lut).max() > 1e-9          # the Newton path is the one taken
    # same carriers and transport; only the relaxation temperature differs
    assert np.abs(tn - tt).max() < 5e-3
    led_n, led_t = newton.ledger(), table.ledger()
    assert led_n["total_absorbed"] == led_t["total_absorbed"]
    assert abs(led_n["step_balance_residual_ev"]) <= 1e-12 * led_n["thermal_energy_ev"]


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", ["c1c_callaway.toml", "c1l_layered.toml"])
def test_newton_inversion_rejected_where_it_does_not_apply(cfg):
    p = rt.Problem(rt.config_text(cfg, **{"simulation.particle_count": 2000}), input_directory=rt.CONFIGS,
                   mode_sampling=1)
    with pytest.raises(RuntimeError, match="pmc error 9"):
        rt.Context(p, inversion=1)


def test_unknown_inversion_value_is_an_argument_error():
    """pmc_create validates temperature_inversion before touching the device."""
    p = rt.Problem(rt.config_text("c1_cross_plane.toml", **{"simulation.particle_count": 2000}))
    s = p.views()[0]
    s.temperature_inversion = 2
    import ctypes as C
    h = C.c_void_p()
    assert rt.lib().pmc_create(C.byref(s), C.byref(h)) == 1
