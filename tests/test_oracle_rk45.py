"""Pins of the oracle's RK45 (RK45Strategy.cpp:101-520 restated in
oracle/icarion_oracle.cpp) against mathematics the tableau must satisfy:

  * one Dormand-Prince step of a linear autonomous system y' = A y is the
    stability polynomial R(hA) y0, R(z) = sum_{j<=5} z^j/j! + z^6/600 (the
    published function of DP5's propagated solution), and the embedded
    solution is Rhat(hA) y0 with Rhat(z) = sum_{j<=4} z^j/j! + 1097 z^5/120000
    + 161 z^6/120000 + z^7/24000; the error estimate is therefore the closed
    form max_i |(Rhat - R)(hA) y0|_i / (atol + rtol |y0_i|);
  * the error estimate of a smooth problem scales as h^5;
  * a constant acceleration is integrated exactly, its error estimate sits at
    the 1e-10 floor and the next step grows by the max_step_increase clamp;
  * the PI controller's closed forms (accept, reject, clamps);
  * rejected trials end in a step that equals a fresh step of the accepted
    size; the FSAL derivative is the acceleration at the accepted state;
  * the engine time after a step is the max of the ion times, or the next
    birth time, or T + dt (SimulationEngineUtils.cpp:49-83).
The linear system is an ion on the Orbitrap axis: z'' = -(q k / m) z."""
import math

import numpy as np
import pytest

from icarion_amd import _abi
from icarion_amd import workloads as W
from icarion_amd.ensemble import Ensemble

Q = 1.602176634e-19
AMU = 1.66053906660e-27


def R_poly(Z):
    I = np.eye(Z.shape[0])
    out, term = I.copy(), I.copy()
    for j in range(1, 6):
        term = term @ Z / j
        out = out + term
    return out + np.linalg.matrix_power(Z, 6) / 600.0


def Rhat_poly(Z):
    I = np.eye(Z.shape[0])
    out, term = I.copy(), I.copy()
    for j in range(1, 5):
        term = term @ Z / j
        out = out + term
    Z5 = np.linalg.matrix_power(Z, 5)
    return out + Z5 * (1097 / 120000) + Z5 @ Z * (161 / 120000) + Z5 @ Z @ Z / 24000.0


def axis_domain():
    return W.orbitrap_domain()


def system(m_amu=300.0):
    """State (x, y, z, vx, vy, vz) of an ion on the Orbitrap axis (x = y = 0):
    the field there is (0, 0, -k z), so z'' = -w^2 z with w^2 = q k / m."""
    w2 = Q * W.orbitrap_k(axis_domain()) / (m_amu * AMU)
    A = np.zeros((6, 6))
    A[0, 3] = A[1, 4] = A[2, 5] = 1.0
    A[5, 2] = -w2
    return A, w2


def step(cfg, y0, h, m_amu=300.0, state=None):
    from oracle import binding as O
    d = axis_domain()
    ph = W.physics(_abi.INTEGRATOR_RK45)
    return O.rk45_push_one(d, ph, cfg, W.species(), m_amu * AMU, Q, y0[:3], y0[3:], 0.0, h,
                           state=state)


Y0 = np.array([0.0, 0.0, 1.5e-3, 0.0, 0.0, 2000.0])


def test_dp5_step_is_the_stability_polynomial(oracle):
    cfg = _abi.rk45_defaults(atol=1.0, rtol=1.0)  # accept the trial step
    A, w2 = system()
    h = 0.05 / math.sqrt(w2)
    p, v, used, _, rej, _ = step(cfg, Y0, h)
    assert used == h and rej == 0
    ref = R_poly(h * A) @ Y0
    np.testing.assert_allclose(np.concatenate([p, v]), ref, rtol=1e-12, atol=1e-18)


@pytest.mark.parametrize("hw", [0.02, 0.1, 0.3])
def test_error_estimate_closed_form(oracle, hw):
    from oracle import binding as O
    cfg = _abi.rk45_defaults(atol=1e-6, rtol=1e-3)   # loose enough to accept hw = 0.3
    A, w2 = system()
    h = hw / math.sqrt(w2)
    y4 = R_poly(h * A) @ Y0
    y5 = Rhat_poly(h * A) @ Y0
    scale = cfg.atol + cfg.rtol * np.abs(Y0)
    err_ref = max(np.max(np.abs(y5 - y4) / scale), 1e-10)
    # the oracle's estimate from its own stages, through one accepted trial;
    # y5 - y4 carries rounding of a few ulp of |y|
    _, _, used, dt_next, rej, st = step(cfg, Y0, h)
    assert rej == 0 and used == h
    noise = np.max(8 * np.finfo(float).eps * np.abs(Y0) / scale)
    assert abs(st["last_error"] - err_ref) <= 1e-6 * err_ref + noise
    # and the estimator itself on the closed-form states
    assert O.rk45_error(cfg, y4, y5, Y0) == pytest.approx(err_ref, rel=1e-12)


def test_error_estimate_scales_as_h5(oracle):
    cfg = _abi.rk45_defaults(atol=1e-12, rtol=1e-12, max_step_factor=1e9)
    A, w2 = system()
    errs = []
    for hw in (0.08, 0.04):
        h = hw / math.sqrt(w2)
        y4 = R_poly(h * A) @ Y0
        y5 = Rhat_poly(h * A) @ Y0
        errs.append(np.max(np.abs(y5 - y4)))
    assert 28 < errs[0] / errs[1] < 36


def test_constant_field_exact_and_step_growth(oracle):
    from oracle import binding as O
    d = W.ims_domain(EN_Td=10.0)
    ph = W.physics(_abi.INTEGRATOR_RK45)
    cfg = _abi.rk45_defaults()
    m = 100 * AMU
    a = Q * d.dc_axial_V / d.length_m / m
    p0, v0 = np.array([1e-4, 2e-4, 3e-4]), np.array([10.0, -5.0, 3.0])
    h = 1e-10
    p, v, used, dt_next, rej, st = O.rk45_push_one(d, ph, cfg, W.species(), m, Q, p0, v0, 0.0, h)
    assert used == h and rej == 0
    np.testing.assert_allclose(p, p0 + v0 * h + np.array([0, 0, 0.5 * a * h * h]), rtol=1e-13)
    np.testing.assert_allclose(v, v0 + np.array([0, 0, a * h]), rtol=1e-13)
    assert st["last_error"] < 1e-6
    assert dt_next == pytest.approx(5.0 * h, rel=1e-15)  # max_step_increase clamp
    np.testing.assert_allclose(st["k1a"], [0.0, 0.0, a], rtol=1e-15)  # FSAL = a at the end
    assert st["fsal"] == 1


def test_controller_closed_forms(oracle):
    from oracle import binding as O
    c = _abi.rk45_defaults()
    alpha = 0.2 - 0.04 * 0.75
    # accepted: PI law
    f = 0.9 * (1.0 / 0.5) ** 0.04 * (1.0 / 0.5) ** alpha
    assert O.rk45_new_step(c, 1e-9, 0.5, 1e-15, 1e-6, 1.0) == pytest.approx(1e-9 * f, rel=1e-15)
    f = 0.9 * (0.2 / 0.8) ** 0.04 * (1.0 / 0.8) ** alpha
    assert O.rk45_new_step(c, 1e-9, 0.8, 1e-15, 1e-6, 0.2) == pytest.approx(1e-9 * f, rel=1e-15)
    # rejected: safety * err^-1/5
    assert O.rk45_new_step(c, 1e-9, 4.0, 1e-15, 1e-6, 1.0) == pytest.approx(
        1e-9 * 0.9 * 4.0 ** -0.2, rel=1e-15)
    # clamps: growth 5, shrink 0.1, absolute limits
    assert O.rk45_new_step(c, 1e-9, 1e-10, 1e-15, 1e-6, 1.0) == pytest.approx(5e-9, rel=1e-15)
    assert O.rk45_new_step(c, 1e-9, 1e12, 1e-15, 1e-6, 1.0) == pytest.approx(1e-10, rel=1e-15)
    assert O.rk45_new_step(c, 1e-9, 1e-10, 1e-15, 2e-9, 1.0) == 2e-9
    assert O.rk45_new_step(c, 1e-9, 1e12, 5e-10, 1e-6, 1.0) == 5e-10


def test_rejected_trials_end_in_a_fresh_step_of_the_accepted_size(oracle):
    cfg = _abi.rk45_defaults()
    A, w2 = system()
    h = 0.8 / math.sqrt(w2)
    p, v, used, _, rej, st = step(cfg, Y0, h)
    assert rej >= 1 and used < h
    ref = R_poly(used * A) @ Y0
    np.testing.assert_allclose(np.concatenate([p, v]), ref, rtol=1e-11, atol=1e-15)
    assert st["last_error"] <= 1.0
    # the FSAL derivative is the acceleration at the accepted state
    np.testing.assert_allclose(st["k1a"], (A @ np.concatenate([p, v]))[3:], rtol=1e-12)


def test_fsal_step_equals_a_fresh_step(oracle):
    cfg = _abi.rk45_defaults(atol=1.0, rtol=1.0)
    A, w2 = system()
    h = 0.05 / math.sqrt(w2)
    p1, v1, _, _, _, st = step(cfg, Y0, h)
    y1 = np.concatenate([p1, v1])
    pa, va, _, _, _, _ = step(cfg, y1, h, state=st)        # k1 from the FSAL slot
    pb, vb, _, _, _, _ = step(cfg, y1, h, state=None)      # k1 evaluated
    np.testing.assert_allclose(np.concatenate([pa, va]), np.concatenate([pb, vb]), rtol=1e-14,
                               atol=1e-20)


def _ions(n, m_amu=300.0):
    e = Ensemble(n)
    rng = np.random.default_rng(4)
    e.x[:] = rng.uniform(-1e-3, 1e-3, n)
    e.y[:] = rng.uniform(-1e-3, 1e-3, n)
    e.z[:] = rng.uniform(0, 1e-3, n)
    e.vx[:] = rng.normal(0, 30, n)
    e.vy[:] = rng.normal(0, 30, n)
    e.mass_kg[:] = m_amu * AMU
    e.charge_C[:] = Q
    return e


def test_engine_time_is_the_max_of_the_ion_times(oracle):
    d = W.quadrupole_domain()
    ph = W.physics(_abi.INTEGRATOR_RK45)
    cfg = _abi.rk45_defaults()
    e = _ions(16)
    e.dt_s[:] = np.linspace(1e-9, 4e-9, 16)   # per-ion dt
    st, state = oracle.run_rk45(d, ph, cfg, [W.species()], e, 0.0, 1e-9, 1)
    assert st.ion_steps == 16
    assert st.t_end_s == state["t"].max()
    np.testing.assert_allclose(state["t"], np.linspace(1e-9, 4e-9, 16), rtol=1e-15)
    st2, state = oracle.run_rk45(d, ph, cfg, [W.species()], e, st.t_end_s, 1e-9, 3, state=state)
    assert st2.t_end_s == state["t"].max()


def test_engine_time_jumps_to_the_next_birth(oracle):
    d = W.quadrupole_domain()
    ph = W.physics(_abi.INTEGRATOR_RK45)
    cfg = _abi.rk45_defaults()
    e = _ions(4)
    e.born[:] = 0
    e.active[:] = 0
    e.birth_time_s = np.array([5e-6, 7e-6, 9e-6, 11e-6])
    st, state = oracle.run_rk45(d, ph, cfg, [W.species()], e, 0.0, 1e-9, 1)
    assert st.ion_steps == 0 and st.t_end_s == 5e-6
    st, state = oracle.run_rk45(d, ph, cfg, [W.species()], e, st.t_end_s, 1e-9, 1, state=state)
    # ion 0 is born at T = 5 us and steps from its own time (0): max t_i < T,
    # so the engine time jumps to the next birth
    assert st.ion_steps == 1 and e.born.tolist() == [1, 0, 0, 0]
    assert st.t_end_s == 7e-6


def test_engine_time_fallback_dt(oracle):
    d = W.quadrupole_domain()
    ph = W.physics(_abi.INTEGRATOR_RK45)
    e = _ions(3)
    e.active[:] = 0
    st, _ = oracle.run_rk45(d, ph, _abi.rk45_defaults(), [W.species()], e, 1e-6, 2e-9, 1)
    assert st.ion_steps == 0 and st.t_end_s == 1e-6 + 2e-9
