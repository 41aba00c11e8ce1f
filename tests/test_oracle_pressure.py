"""Time-varying gas pressure in the oracle: SimulationEngine::process_timestep
calls update_dynamic_environments at the engine time of each step start
(SimulationEngine.cpp:380, SimulationEngineLifecycle.cpp:34-47), which sets
P(t) from the domain's pressure_Pa_waveform and derives n = P/(kB T) and a
mixture's n_i = x_i n (EnvironmentConfig.h:132-175).  Collision rates and the
Friction / Langevin damping coefficients of that step follow n.

Pinned here by:
  * a CONSTANT waveform at the descriptor's pressure is the constant run, bit
    for bit (HSS, Friction with OU kicks);
  * a CONSTANT waveform at 3 P0 is the constant run of a descriptor whose n and
    component densities n_i are tripled (single gas, mixture, Friction): a loop
    that scaled n but not the n_i, or not the damping, fails it;
  * heavy ions at rest under a pressure pulse: the mean number of events per
    ion is sum_k E[1 - exp(-n(t_k) sigma |v| dt)] over the Maxwell speed law,
    met to 4 standard errors; the law of the constant pressure misses by far;
  * Friction drift under a pressure ramp ends at the terminal velocity
    K(n_end) E with K proportional to 1 / n(t_end): half the constant
    pressure's drift for a ramp to 2 P0;
  * a waveform reaching P <= 0 is rejected, as the reference throws.
"""
import math

import numpy as np
import pytest

from icarion_amd import _abi
from icarion_amd import workloads as W
from test_oracle_substeps import expected_events, resting_heavy_ions

KB = 1.380649e-23


def const_wave(p):
    return _abi.waveform(_abi.WAVEFORM_CONSTANT, (p,))


def assert_same(a, b):
    for f in ("x", "y", "z", "vx", "vy", "vz", "n_collisions", "active"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)


def scaled_domain(d, s):
    d2 = type(d).from_buffer_copy(d)
    d2.pressure_Pa = d.pressure_Pa * s
    d2.gas_density_m3 = d.gas_density_m3 * s
    return d2


def close_fraction(a, b, rtol=1e-9):
    ok = np.ones(a.n, bool)
    for f in ("x", "y", "z"):
        ok &= np.abs(getattr(a, f) - getattr(b, f)) <= rtol * np.maximum(np.abs(getattr(b, f)), 1e-6)
    for f in ("vx", "vy", "vz"):
        ok &= np.abs(getattr(a, f) - getattr(b, f)) <= rtol * np.maximum(np.abs(getattr(b, f)), 1.0)
    return ok.mean()


@pytest.mark.parametrize("collision", [_abi.COLLISION_HSS, _abi.COLLISION_FRICTION])
def test_constant_waveform_at_the_descriptor_pressure_is_the_constant_run(oracle, collision):
    w = W.ims_hss_workload(n=512, seed=21)
    ph = W.physics(_abi.INTEGRATOR_RK4, collision, ou=1, seed=21)
    sp = [W.species(sigma_hss=100e-20, ccs=100e-20, K0=2.1)]
    a, b = w["ensemble"].copy(), w["ensemble"].copy()
    sa = oracle.run(w["domain"], ph, sp, a, 0.0, w["dt"], 400)
    sb = oracle.run(w["domain"], ph, sp, b, 0.0, w["dt"], 400, pressure=const_wave(w["domain"].pressure_Pa))
    assert_same(a, b)
    assert sa.collisions == sb.collisions


@pytest.mark.parametrize("case", ["hss", "friction_ou", "langevin", "mixture"])
def test_tripled_pressure_is_the_tripled_density(oracle, case):
    s = 3.0
    if case == "mixture":
        w = W.ims_mixture_workload(n=1024, seed=22)
        comps, sigma = w["mixture"]
        comps3 = [(n * s, m, part) for (n, m, part) in comps]
        ph = w["physics"]
        sp = w["species"]
        mix, mix3 = oracle.GasMixture(comps, sigma), oracle.GasMixture(comps3, sigma)
    else:
        w = W.ims_hss_workload(n=1024, seed=22)
        coll = {"hss": _abi.COLLISION_HSS, "friction_ou": _abi.COLLISION_FRICTION,
                "langevin": _abi.COLLISION_LANGEVIN}[case]
        ph = W.physics(_abi.INTEGRATOR_RK4, coll, ou=1, seed=22)
        sp = [W.species(sigma_hss=100e-20, ccs=100e-20, K0=2.1)]
        mix = mix3 = None
    d = w["domain"]
    a, b = w["ensemble"].copy(), w["ensemble"].copy()
    sa = oracle.run(d, ph, sp, a, 0.0, w["dt"], 600, mixture=mix, pressure=const_wave(s * d.pressure_Pa))
    sb = oracle.run(scaled_domain(d, s), ph, sp, b, 0.0, w["dt"], 600, mixture=mix3)
    # equal up to the rounding of x_i n against s n_i (a collision test within
    # rounding of its threshold can flip)
    assert close_fraction(a, b) >= 0.99
    if case in ("hss", "mixture"):
        assert sa.collisions == pytest.approx(sb.collisions, rel=2e-3)
        # and not the constant pressure's
        c = w["ensemble"].copy()
        sc = oracle.run(d, ph, sp, c, 0.0, w["dt"], 600, mixture=mix)
        assert sa.collisions > 1.1 * sc.collisions


@pytest.mark.parametrize("rng_kind", ["philox", "mt19937_64"])
def test_event_law_under_a_pressure_pulse(oracle, rng_kind):
    P0 = 101325.0
    dom = W.ims_domain(origin_z=-5e-3, length=20e-3, radius=10e-3, axial_V=0.0, gas="He",
                       pressure_Pa=P0, temperature_K=300.0)
    ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, seed=5)
    sp = [W.species(sigma_hss=100e-20, ccs=100e-20)]
    dt, n_steps, n_ions = 3e-11, 100, 4096
    # 4 P0 for the steps 30..69 (the step times sit half a step inside)
    pulse = _abi.waveform(_abi.WAVEFORM_PULSED, (P0, 4 * P0, 29.5 * dt, 40 * dt))
    e = resting_heavy_ions(n_ions)
    kind = oracle.RNG_PHILOX if rng_kind == "philox" else oracle.RNG_MT19937_64
    st = oracle.run(dom, ph, sp, e, 0.0, dt, n_steps, rng_kind=kind, pressure=pulse)
    assert st.ion_steps == n_ions * n_steps
    per_ion = e.n_collisions.astype(np.float64)
    mean, se = per_ion.mean(), per_ion.std(ddof=1) / math.sqrt(n_ions)
    n0 = P0 / (KB * 300.0)
    m_gas = dom.gas_mass_kg

    def law(n):
        return expected_events(n, 100e-20, dt, 300.0, m_gas, 1)

    want = 60 * law(n0) + 40 * law(4 * n0)
    assert abs(mean - want) < 4 * se, (mean, want, se)
    assert abs(mean - 100 * law(n0)) > 20 * se


def test_friction_drift_follows_a_pressure_ramp(oracle):
    K0, T, P0 = 2.0, 300.0, 101325.0
    d = W.ims_domain(axial_V=100.0, gas="N2", pressure_Pa=P0, temperature_K=T)
    E = 100.0 / d.length_m
    ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_FRICTION)
    dt, n_steps = 1e-10, 2000
    ramp = _abi.waveform(_abi.WAVEFORM_LINEAR, (P0, 2 * P0, 20e-9, 100e-9), clamp=1)
    a = W.gaussian_cloud(4, 100.0, 1e-6, temperature_K=0.0, seed=1)
    b = a.copy()
    oracle.run(d, ph, [W.species(K0=K0)], a, 0.0, dt, n_steps)
    oracle.run(d, ph, [W.species(K0=K0)], b, 0.0, dt, n_steps, pressure=ramp)
    # K = K0 (n_L / n) sqrt(T0 / T): the drift at 2 P0 is half that at P0
    n_L = 2.6867811e25
    K_end = K0 * 1e-4 * (n_L / (2 * P0 / (KB * T))) * math.sqrt(273.15 / T)
    np.testing.assert_allclose(b.vz, K_end * E, rtol=1e-9)
    np.testing.assert_allclose(b.vz, 0.5 * a.vz, rtol=1e-9)


def test_non_positive_pressure_is_rejected(oracle):
    w = W.ims_hss_workload(n=16, seed=23)
    # falls through zero at t = 50 dt
    ramp = _abi.waveform(_abi.WAVEFORM_LINEAR, (1000.0, -1000.0, 0.0, 100 * w["dt"]), clamp=1)
    with pytest.raises(RuntimeError):
        oracle.run(w["domain"], w["physics"], w["species"], w["ensemble"].copy(), 0.0, w["dt"], 80,
                   pressure=ramp)
    # ending before the sign change: runs
    oracle.run(w["domain"], w["physics"], w["species"], w["ensemble"].copy(), 0.0, w["dt"], 40, pressure=ramp)


def test_several_domains_take_their_own_pressure(oracle):
    """A domain's pressure waveform changes that domain's rates only: with a
    CONSTANT waveform of 3 P0 on domain 0 the run is the one whose domain 0
    has a tripled density."""
    w = W.multi_domain_workload(n=1024, seed=24, collision=_abi.COLLISION_HSS)
    doms = w["domains"]
    pw = [const_wave(3 * doms[0].pressure_Pa), None, None]
    a, b = w["ensemble"].copy(), w["ensemble"].copy()
    da, db = np.zeros(a.n, np.int32), np.zeros(b.n, np.int32)
    sa = oracle.run_multi(doms, w["physics"], w["species"], a, da, 0.0, w["dt"], 3000,
                          end_apertures=w.get("apertures"), pressures=pw)
    doms3 = [scaled_domain(doms[0], 3.0)] + list(doms[1:])
    sb = oracle.run_multi(doms3, w["physics"], w["species"], b, db, 0.0, w["dt"], 3000,
                          end_apertures=w.get("apertures"))
    assert close_fraction(a, b) >= 0.99
    assert (da == db).mean() >= 0.99
    assert sa.collisions == pytest.approx(sb.collisions, rel=3e-3)
    c, dc = w["ensemble"].copy(), np.zeros(a.n, np.int32)
    sc = oracle.run_multi(doms, w["physics"], w["species"], c, dc, 0.0, w["dt"], 3000,
                          end_apertures=w.get("apertures"))
    assert sa.collisions > 1.2 * sc.collisions


def test_rk45_takes_the_pressure_at_each_engine_step(oracle):
    """RK45: a CONSTANT waveform at P0 is the constant run bit for bit, one at
    3 P0 the run of the tripled density (HSS and Friction + OU), and a
    pressure that falls through zero is rejected."""
    w = W.ims_hss_workload(n=256, seed=31)
    cfg = _abi.rk45_defaults()
    for coll in (_abi.COLLISION_HSS, _abi.COLLISION_FRICTION):
        ph = W.physics(_abi.INTEGRATOR_RK45, coll, ou=1, seed=31)
        sp = [W.species(sigma_hss=100e-20, ccs=100e-20, K0=2.1)]
        d = w["domain"]
        a, b, c = w["ensemble"].copy(), w["ensemble"].copy(), w["ensemble"].copy()
        sa, _ = oracle.run_rk45(d, ph, cfg, sp, a, 0.0, w["dt"], 40)
        sb, _ = oracle.run_rk45(d, ph, cfg, sp, b, 0.0, w["dt"], 40,
                                field=oracle._with_pressure(None, const_wave(d.pressure_Pa)))
        assert_same(a, b)
        sc, _ = oracle.run_rk45(d, ph, cfg, sp, c, 0.0, w["dt"], 40,
                                field=oracle._with_pressure(None, const_wave(3 * d.pressure_Pa)))
        e3 = w["ensemble"].copy()
        s3, _ = oracle.run_rk45(scaled_domain(d, 3.0), ph, cfg, sp, e3, 0.0, w["dt"], 40)
        assert close_fraction(c, e3) >= 0.99
        if coll == _abi.COLLISION_HSS:
            assert sc.collisions == pytest.approx(s3.collisions, rel=3e-3)
            assert sc.collisions != sa.collisions
    ramp = _abi.waveform(_abi.WAVEFORM_LINEAR, (1000.0, -1000.0, 0.0, 100 * w["dt"]), clamp=1)
    with pytest.raises(RuntimeError):
        oracle.run_rk45(w["domain"], w["physics"], cfg, w["species"], w["ensemble"].copy(), 0.0, w["dt"], 80,
                        field=oracle._with_pressure(None, ramp))
