"""The oracle's EHSS gas-mixture path (CPU).

EHSSCollisionHandler::handle_collision's mixture path (EHSSCollisionHandler.cpp:
311-424) is restated in the oracle; the handler needs spdlog and the species
database, so it cannot be compiled here.  Pinned by:
  * its loop structure, bit for bit:
    - a component that does not take part draws nothing;
    - one that takes part draws its neutral even at zero density;
  * its collision probability per stage against an independent Monte-Carlo
    evaluation of E[1 - exp(-sum_c n_c sigma_c |v - v_c| dt)] over Maxwellian
    neutrals of each component.
The partner choice and the collision itself reuse the single-gas pieces (the
online geometry kernel is pinned bit for bit against the reference's own
CollisionKernels.cpp); the device path is compared with this restatement ion
by ion on the GPU.
"""
import numpy as np
import pytest

from icarion_amd import workloads as W
from icarion_amd.constants import AMU_TO_KG, BOLTZMANN, GASES


def _setup():
    rng = np.random.default_rng(5)
    atoms = np.concatenate([rng.normal(0, 3e-10, (9, 3)), rng.uniform(1.5e-10, 2e-10, (9, 1))], 1)
    d = W.ims_domain(EN_Td=20.0, gas="He", pressure_Pa=101325.0)
    return d, atoms


def test_loop_structure_bitexact(oracle):
    d, atoms = _setup()
    n = d.gas_density_m3
    mN2, mHe = GASES["N2"][0], GASES["He"][0]
    sig = 150e-20
    base = oracle.GasMixture([(n, mN2, 1)], np.array([[sig]]))
    skip = oracle.GasMixture([(0.3 * n, mHe, 0), (n, mN2, 1)], np.array([[sig, sig]]))
    empty = oracle.GasMixture([(0.0, mHe, 1), (n, mN2, 1)], np.array([[sig, sig]]))
    m = 300 * AMU_TO_KG
    dt = 5e-10
    same = differ = hits = 0
    for seed in range(1, 400):
        v = np.array([120.0, -40.0, 900.0])
        h0, v0 = oracle.mt_ehss_mixture(seed, d, base, 0, atoms, m, v, dt)
        h1, v1 = oracle.mt_ehss_mixture(seed, d, skip, 0, atoms, m, v, dt)
        h2, v2 = oracle.mt_ehss_mixture(seed, d, empty, 0, atoms, m, v, dt)
        # a component that does not take part draws nothing: the same stream, bit for bit
        assert h0 == h1 and np.array_equal(v0, v1)
        hits += h0
        if h0 and h2:
            if np.array_equal(v0, v2):
                same += 1
            else:
                differ += 1
    # one that takes part draws its neutral (6 uniforms) even at n = 0, so the
    # streams shift and the outcomes differ
    assert hits > 50 and differ > 10 * max(same, 1)


def test_collision_probability_law(oracle):
    d, atoms = _setup()
    n = d.gas_density_m3
    mN2, mHe = GASES["N2"][0], GASES["He"][0]
    sig = np.array([[220e-20, 90e-20]])
    mix = oracle.GasMixture([(0.6 * n, mN2, 1), (0.4 * n, mHe, 1)], sig)
    m = 200 * AMU_TO_KG
    dt = 2e-10
    v = np.array([300.0, 0.0, 1500.0])
    N = 40000
    hits = sum(oracle.mt_ehss_mixture(seed, d, mix, 0, atoms, m, v, dt)[0] for seed in range(1, N + 1))
    # independent Monte Carlo of the law: k = sum_c n_c sigma_c |v - v_c|, P = 1 - exp(-k dt)
    rng = np.random.default_rng(17)
    M = 400000
    k = np.zeros(M)
    for (nc, mc), s in zip(((0.6 * n, mN2), (0.4 * n, mHe)), sig[0]):
        vn = rng.normal(0.0, np.sqrt(BOLTZMANN * d.temperature_K / mc), (M, 3))
        k += nc * s * np.linalg.norm(v - vn, axis=1)
    p = np.mean(1.0 - np.exp(-k * dt))
    se = np.sqrt(p * (1 - p) / N)
    assert abs(hits / N - p) < 4 * se, (hits / N, p)
