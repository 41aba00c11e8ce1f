"""Collision substeps in the oracle: the engine's CPU collision path
(handle_collisions_cpu, SimulationEngineCollisionRuntime.cpp:26-33, 84-191)
runs the collision stage N times per engine step with collision_dt = dt / N on
the pushed state; every substep is an event of its own, and all draw from the
ion's one random stream.  N is collision_subcycles_per_step, or
max(it, collision_max_events_per_step) in multi-event mode.

Pinned here by:
  * 0 and 1 substeps are the one-stage loop, bit for bit;
  * a heavy ion at rest in a Maxwellian gas: |v_rel| is the neutral's speed,
    so the mean number of events per ion-step is N E[1 - exp(-n sigma |v| dt/N)]
    over the Maxwell speed law with scale sqrt(kT/m_n); the oracle meets it to
    4 standard errors for N = 1, 4 and 16, and the laws of a loop that forgot
    to divide dt (N E[1 - exp(-n sigma |v| dt)]) or to run the substeps
    (E[1 - exp(-n sigma |v| dt)]) miss by far more;
  * with the reference's mt19937_64 streams (the engine's rng_by_ion_), the
    loop gives the same law.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from icarion_amd import _abi
from icarion_amd import workloads as W
from icarion_amd.ensemble import Ensemble

KB = 1.380649e-23
AMU = 1.66053906660e-27
E_CHARGE = 1.602176634e-19


def resting_heavy_ions(n, mass_amu=1e7):
    e = Ensemble(n)
    e.mass_kg[:] = mass_amu * AMU
    e.charge_C[:] = E_CHARGE
    e.z[:] = 1e-3
    return e


def expected_events(n_gas, sigma, dt, T, m_gas, N):
    """N E[1 - exp(-n sigma v dt / N)] for v ~ Maxwell speed, scale sqrt(kT/m)."""
    s = math.sqrt(KB * T / m_gas)
    a = n_gas * sigma * dt / N

    def f(v):
        return math.sqrt(2.0 / math.pi) * v * v / s ** 3 * math.exp(-v * v / (2 * s * s)) * (1.0 - math.exp(-a * v))

    val, _ = integrate.quad(f, 0.0, 12 * s, limit=200)
    return N * val


def setup(substeps, seed=5):
    dom = W.ims_domain(origin_z=-5e-3, length=20e-3, radius=10e-3, axial_V=0.0, gas="He",
                       pressure_Pa=101325.0, temperature_K=300.0)
    ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, seed=seed, substeps=substeps)
    sp = [W.species(sigma_hss=100e-20, ccs=100e-20)]
    return dom, ph, sp


def test_zero_and_one_substep_are_the_one_stage_loop(oracle):
    w = W.ims_hss_workload(n=512, n_steps=300, seed=9)
    outs = []
    for N in (0, 1):
        ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, seed=9, substeps=N)
        e = w["ensemble"].copy()
        st = oracle.run(w["domain"], ph, w["species"], e, 0.0, w["dt"], 300)
        outs.append((e, st.collisions))
    (a, ca), (b, cb) = outs
    for f in ("x", "y", "z", "vx", "vy", "vz", "n_collisions", "active"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    assert ca == cb > 0


@pytest.mark.parametrize("N", [1, 4, 16])
@pytest.mark.parametrize("rng_kind", ["philox", "mt19937_64"])
def test_substep_event_law_for_an_ion_at_rest(oracle, N, rng_kind):
    dom, ph, sp = setup(N)
    dt = 3e-11   # n sigma <|v|> dt ~ 0.9: one stage and N stages differ strongly
    n_ions, n_steps = 4096, 100
    e = resting_heavy_ions(n_ions)
    kind = oracle.RNG_PHILOX if rng_kind == "philox" else oracle.RNG_MT19937_64
    st = oracle.run(dom, ph, sp, e, 0.0, dt, n_steps, rng_kind=kind)
    assert st.ion_steps == n_ions * n_steps
    per_ion = e.n_collisions.astype(np.float64) / n_steps
    mean = per_ion.mean()
    se = per_ion.std(ddof=1) / math.sqrt(n_ions)
    n_gas, T, m_gas = dom.gas_density_m3, dom.temperature_K, dom.gas_mass_kg
    want = expected_events(n_gas, 100e-20, dt, T, m_gas, N)
    assert abs(mean - want) < 4 * se, (mean, want, se)
    if N > 1:
        undivided = N * expected_events(n_gas, 100e-20, dt, T, m_gas, 1)
        one_stage = expected_events(n_gas, 100e-20, dt, T, m_gas, 1)
        assert abs(mean - undivided) > 20 * se
        assert abs(mean - one_stage) > 20 * se
