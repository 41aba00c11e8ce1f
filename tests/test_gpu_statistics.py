"""Engine-level statistics of the device against the oracle driven by the
reference's own random numbers: std::mt19937_64 per ion seeded rng_seed + i,
uniform_real_distribution and normal_distribution (PhysicsRng,
src/core/physics/collisions/CollisionTypes.cpp:8-20; RngUtils.cpp:51-53).  The
device draws Philox4x32-10 streams instead, so the two runs are independent
samples of the same process and agree only statistically.  Each comparison
states its bar: a relative bound (the 0.5 % mobility gate) and, where the
sample allows, 4 standard errors of the difference, plus two-sample KS tests
(p > 1e-3).

Workloads (icarion_amd.workloads): the IMS drift tube (10k ions, HSS in N2 at
1 atm, 10 Td, RK4 dt 0.1 ns; 5e4 steps = 5 us), the RF quadrupole (Boris,
Langevin + OU in He at 1 Pa; 20k ions x 2e4 steps) and the field-array drift
(HSS in N2 at 300 Pa on a 24 x 20 x 40 grid of the same Gaussian-sum field;
20k ions x 4000 steps)."""
import math

import numpy as np
import pytest
from scipy import stats

from icarion_amd import _abi
from icarion_amd import workloads as W

pytestmark = pytest.mark.gpu


def gpu_run(w, ens, n_steps, precision):
    from icarion_amd import gpu
    ph = w["physics"]
    ph = W.physics(ph.integrator, ph.collision_model, ph.ou_thermal_kicks, precision, ph.rng_seed,
                   ph.ehss_mode)
    with gpu.Handle() as h:
        h.set_domain(w["domain"], ph)
        h.set_species(w["species"], w.get("atoms"))
        if w.get("field") is not None:
            h.upload_field_array(*w["field"])
        h.upload_state(ens)
        st = h.step(0.0, w["dt"], n_steps, 0)
        h.download_state(ens)
    return st


def mt_run(oracle, w, ens, n_steps):
    field = oracle.Field(*w["field"]) if w.get("field") is not None else None
    return oracle.run(w["domain"], w["physics"], w["species"], ens, 0.0, w["dt"], n_steps,
                      field=field, rng_kind=oracle.RNG_MT19937_64)


def mean_se(x):
    x = np.asarray(x, float)
    return x.mean(), x.std(ddof=1) / math.sqrt(x.size)


def assert_same_mean(a, b, rel, what):
    """|mean(a) / mean(b) - 1| within rel and within 4 standard errors of the
    difference."""
    ma, sa = mean_se(a)
    mb, sb = mean_se(b)
    d = abs(ma - mb)
    s = math.hypot(sa, sb)
    print(f"{what}: {ma:.6e} vs {mb:.6e} (rel. diff {d / abs(mb):.3e}, sigma of the diff "
          f"{s / abs(mb):.3e})", flush=True)
    assert d <= rel * abs(mb), what
    assert d <= 4 * s + 1e-12 * abs(mb), what


@pytest.mark.parametrize("precision", [_abi.PRECISION_FP64, _abi.PRECISION_FP32_KAHAN])
def test_ims_mobility_arrival_survival_against_reference_rng(oracle, precision):
    """IMS drift tube: mobility K from the mean drift (0.5 % gate and 4 sigma),
    arrival-time mean and width at the detector plane (z = +0.5 mm, first
    crossing), survival fraction, and KS tests on the drift and the arrival
    times."""
    steps = 50_000
    w = W.ims_hss_workload(n=10_000, seed=123)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    z0 = w["ensemble"].z.copy()
    sg = gpu_run(w, g, steps, precision)
    so = mt_run(oracle, w, o, steps)
    T = steps * w["dt"]
    E = w["domain"].dc_axial_V / w["domain"].length_m
    assert g.alive.mean() == o.alive.mean() == 1.0   # the tube holds every ion
    Kg = (g.z - z0)[g.alive] / (T * E)
    Ko = (o.z - z0)[o.alive] / (T * E)
    assert_same_mean(Kg, Ko, 0.005, "mobility K [m^2/Vs]")
    ag = g.arrival_time_s[np.isfinite(g.arrival_time_s)]
    ao = o.arrival_time_s[np.isfinite(o.arrival_time_s)]
    # arrived fraction: binomial, 4 sigma
    pg, po = ag.size / g.n, ao.size / o.n
    assert abs(pg - po) <= 4 * math.sqrt(po * (1 - po) * 2 / o.n), (pg, po)
    assert_same_mean(ag, ao, 0.02, "arrival-time mean [s]")
    wg, wo = ag.std(ddof=1), ao.std(ddof=1)
    # standard error of a sample standard deviation ~ sigma / sqrt(2 (n - 1))
    sw = math.hypot(wg / math.sqrt(2 * (ag.size - 1)), wo / math.sqrt(2 * (ao.size - 1)))
    print(f"arrival-time width: {wg:.4e} vs {wo:.4e} s", flush=True)
    assert abs(wg - wo) <= max(4 * sw, 0.0) and abs(wg / wo - 1) < 0.05
    assert stats.ks_2samp(Kg, Ko).pvalue > 1e-3
    assert stats.ks_2samp(ag, ao).pvalue > 1e-3
    assert_same_mean(g.n_collisions, o.n_collisions, 0.005, "collisions per ion")
    assert sg.ion_steps == so.ion_steps


@pytest.mark.parametrize("precision", [_abi.PRECISION_FP64, _abi.PRECISION_FP32_KAHAN])
def test_quadrupole_transmission_against_reference_rng(oracle, precision):
    """RF quadrupole (U = 0): survival overall and per m/z bin (4 binomial
    sigma of the difference + 0.005), and a KS test on the survivors'
    radial positions."""
    steps = 20_000
    w = W.quadrupole_workload(n=20_000, U=0.0, seed=12)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, steps, precision)
    mt_run(oracle, w, o, steps)
    n = g.n
    tg, to = g.alive.mean(), o.alive.mean()
    print(f"transmission {tg:.4f} vs {to:.4f}", flush=True)
    assert 0.05 < to < 0.95   # both fates occur: the comparison has content
    assert abs(tg - to) <= 4 * math.sqrt(2 * to * (1 - to) / n) + 0.005
    m = w["ensemble"].mass_kg / 1.66053906660e-27
    bins = np.geomspace(100, 2000, 9)
    for lo, hi in zip(bins[:-1], bins[1:]):
        sel = (m >= lo) & (m < hi)
        a, b = g.alive[sel].mean(), o.alive[sel].mean()
        err = 4 * math.sqrt(max(b * (1 - b), 1e-4) * 2 / sel.sum()) + 0.005
        assert abs(a - b) < err, (lo, hi, a, b)
    rg = np.hypot(g.x, g.y)[g.alive]
    ro = np.hypot(o.x, o.y)[o.alive]
    assert stats.ks_2samp(rg, ro).pvalue > 1e-3


@pytest.mark.parametrize("precision", [_abi.PRECISION_FP64, _abi.PRECISION_FP32_KAHAN])
def test_field_array_drift_survival_against_reference_rng(oracle, precision):
    """Field-array drift with HSS: survival (4 binomial sigma + 0.002), mean
    drift (1 % and 4 sigma), drift spread (3 %), collisions per ion (0.5 %) and
    a KS test on the final z."""
    steps = 4000
    w = W.field_array_workload(n=20_000, grid=(24, 20, 40), seed=4)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    z0 = w["ensemble"].z.copy()
    gpu_run(w, g, steps, precision)
    mt_run(oracle, w, o, steps)
    sg, so = g.alive.mean(), o.alive.mean()
    assert abs(sg - so) <= 4 * math.sqrt(2 * max(so * (1 - so), 1e-4) / g.n) + 0.002
    dzg, dzo = (g.z - z0)[g.alive], (o.z - z0)[o.alive]
    assert_same_mean(dzg, dzo, 0.01, "mean drift [m]")
    assert dzg.std() == pytest.approx(dzo.std(), rel=0.03)
    assert_same_mean(g.n_collisions, o.n_collisions, 0.005, "collisions per ion")
    assert stats.ks_2samp(g.z[g.alive], o.z[o.alive]).pvalue > 1e-3
