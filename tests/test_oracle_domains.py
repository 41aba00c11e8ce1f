"""The oracle's multi-domain engine loop (CPU).

The reference's DomainManager cannot be compiled in this image: its
constructor builds AnalyticalFieldModel, which needs spdlog.  So the
restatement is pinned here by properties and by hand-derived cases:
  * find_domain_index returns a domain that contains the point, and -1 only
    when none does (checked against each domain's own containment test);
  * the returned domain is the unique one when only one contains the point;
  * on a shared face (both domains contain it), the z-sorted span search
    returns the lower domain; with an Orbitrap, the linear scan returns the
    first containing domain in config order;
  * the bridge keeps an ion that left through the end face into a gap alive
    for one step with the next domain's index, and it dies at the next step
    start (death time = that step's t);
  * leaving the last domain kills at the step's start time;
  * the handoff moves the index on within 1e-9 m of the end face, inside the
    end aperture only;
  * one domain through the multi-domain loop is the single-domain loop, bit
    for bit;
  * each domain's waveforms act in that domain only: while every ion stays in
    the first domain, the multi-domain run with waveforms on all three domains
    is the single-domain run with the first domain's waveforms, bit for bit,
    and differs from the run without them;
  * each domain's gas mixture acts in that domain only, in the same way;
  * each domain's species table (cross sections, CCS and K0 for its gas) acts
    in that domain only, in the same way, for HSS and for Friction damping
    (K0 enters the damping force).
"""
import ctypes as C

import numpy as np
import pytest

from icarion_amd import _abi
from icarion_amd import workloads as W
from icarion_amd.ensemble import Ensemble


def contains(oracle, d, p):
    return bool(oracle.load().oracle_contains(C.byref(d), oracle.vec(*p)))


def test_find_domain_properties(oracle):
    w = W.multi_domain_workload(n=8)
    doms = w["domains"]
    rng = np.random.default_rng(2)
    n_unique = n_none = n_multi = 0
    for _ in range(4000):
        p = (rng.uniform(-4e-3, 4e-3), rng.uniform(-4e-3, 4e-3), rng.uniform(-2e-3, 22e-3))
        r = oracle.find_domain(doms, p, w["apertures"])
        inside = [k for k, d in enumerate(doms) if contains(oracle, d, p)]
        if r < 0:
            assert inside == []
            n_none += 1
        else:
            assert r in inside
            if len(inside) == 1:
                n_unique += 1
            else:
                n_multi += 1
    assert n_unique > 1000 and n_none > 500
    # shared faces (z = 5 mm and 15 mm are in both neighbours within 1e-12 m)
    for z, lower in ((5e-3, 0), (15e-3, 1)):
        p = (0.1e-3, -0.2e-3, z)
        inside = [k for k, d in enumerate(doms) if contains(oracle, d, p)]
        assert inside == [lower, lower + 1]
        assert oracle.find_domain(doms, p, w["apertures"]) == lower


def test_find_domain_with_orbitrap_is_config_order(oracle):
    orb = W.orbitrap_domain()
    big = W.ims_domain(origin_z=-30e-3, length=60e-3, radius=30e-3)
    p = None
    rng = np.random.default_rng(3)
    for _ in range(10000):
        q = (rng.uniform(-15e-3, 15e-3), rng.uniform(-15e-3, 15e-3), rng.uniform(-10e-3, 10e-3))
        if contains(oracle, orb, q):
            p = q
            break
    assert p is not None and contains(oracle, big, p)
    assert oracle.find_domain([orb, big], p) == 0
    assert oracle.find_domain([big, orb], p) == 0


def _one_ion(z, vz=0.0, r=0.0, mass_amu=200.0):
    e = Ensemble(1)
    e.x[:] = r
    e.y[:] = 0.0
    e.z[:] = z
    e.vx[:] = 0.0
    e.vy[:] = 0.0
    e.vz[:] = vz
    e.mass_kg[:] = mass_amu * W.AMU_TO_KG
    e.charge_C[:] = W.ELEM_CHARGE_C
    return e


def _gap_domains():
    d0 = W.ims_domain(origin_z=0.0, length=5e-3, radius=3e-3, axial_V=0.0, gas="N2", pressure_Pa=10.0)
    d1 = W.ims_domain(origin_z=6e-3, length=5e-3, radius=3e-3, axial_V=0.0, gas="N2", pressure_Pa=10.0)
    return [d0, d1]


def test_bridge_keeps_ion_one_step(oracle):
    doms = _gap_domains()
    ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE)
    sp = [W.species()]
    dt = 1e-9
    # 4.999 mm -> 5.499 mm in one step: in the gap between 5 and 6 mm
    e = _one_ion(4.999e-3, vz=5e5)
    di = np.full(1, -1, np.int32)
    oracle.run_multi(doms, ph, sp, e, di, 0.0, dt, 1, end_apertures=[2e-3, 0.0])
    assert e.active[0] == 1 and di[0] == 1
    assert e.z[0] == pytest.approx(5.499e-3, rel=1e-12)
    oracle.run_multi(doms, ph, sp, e, di, dt, dt, 1, step_offset=1, end_apertures=[2e-3, 0.0])
    assert e.active[0] == 0 and e.death_time_s[0] == dt
    # leaving the last domain: no bridge, death at the step's start time
    e = _one_ion(10.999e-3, vz=5e5)
    di = np.full(1, -1, np.int32)
    oracle.run_multi(doms, ph, sp, e, di, 3e-9, dt, 1)
    assert e.active[0] == 0 and e.death_time_s[0] == 3e-9 and di[0] == 1


def test_handoff_at_the_end_face(oracle):
    d0 = W.ims_domain(origin_z=0.0, length=5e-3, radius=3e-3, axial_V=0.0, gas="N2", pressure_Pa=10.0)
    d1 = W.ims_domain(origin_z=5e-3, length=5e-3, radius=1e-3, axial_V=0.0, gas="N2", pressure_Pa=10.0)
    ph = W.physics(_abi.INTEGRATOR_BORIS, _abi.COLLISION_NONE)
    sp = [W.species()]
    for r, ap, expect in ((0.5e-3, 1.5e-3, 1),    # inside the aperture
                          (2.0e-3, 1.5e-3, 0),    # outside it
                          (2.0e-3, 0.0, 1)):      # aperture 0 = the radius (3 mm)
        e = _one_ion(5e-3 - 0.5e-9, r=r)
        di = np.full(1, -1, np.int32)
        oracle.run_multi([d0, d1], ph, sp, e, di, 0.0, 1e-9, 3, end_apertures=[ap, 0.0])
        assert e.active[0] == 1 and di[0] == expect, (r, ap)
    # 2e-9 m before the face: outside the tolerance, no handoff
    e = _one_ion(5e-3 - 2e-9, r=0.5e-3)
    di = np.full(1, -1, np.int32)
    oracle.run_multi([d0, d1], ph, sp, e, di, 0.0, 1e-9, 3, end_apertures=[1.5e-3, 0.0])
    assert di[0] == 0


@pytest.mark.parametrize("collision", [_abi.COLLISION_NONE, _abi.COLLISION_HSS, _abi.COLLISION_LANGEVIN])
def test_one_domain_is_the_single_domain_loop(oracle, collision):
    w = W.multi_domain_workload(n=512, collision=collision)
    d0 = w["domains"][0]
    a, b = w["ensemble"].copy(), w["ensemble"].copy()
    di = np.full(a.n, -1, np.int32)
    oracle.run_multi([d0], w["physics"], w["species"], a, di, 0.0, w["dt"], 2500)
    oracle.run(d0, w["physics"], w["species"], b, 0.0, w["dt"], 2500)
    for f in ("x", "y", "z", "vx", "vy", "vz", "active", "death_time_s", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    alive = a.active == 1
    assert (di[alive] == 0).all()


def domain_waveforms():
    """Waveforms for the three-domain workload: an axial ramp in the first IMS
    region, a modulated RF amplitude and a frequency ramp in the quadrupole,
    pulsed axial voltage in the last region."""
    return [
        {_abi.SLOT_DC_AXIAL: _abi.waveform(_abi.WAVEFORM_LINEAR, (20.0, 60.0, 0.0, 1e-7), 1)},
        {_abi.SLOT_RF_VOLTAGE: _abi.waveform(_abi.WAVEFORM_SINUSOIDAL, (200.0, 20.0, 1e5, 0.0)),
         _abi.SLOT_RF_FREQUENCY: _abi.waveform(_abi.WAVEFORM_LINEAR, (3.0e6, 3.2e6, 0.0, 6e-6), 1)},
        {_abi.SLOT_DC_AXIAL: _abi.waveform(_abi.WAVEFORM_PULSED, (10.0, 30.0, 2e-6, 1e-6))},
    ]


def test_per_domain_waveforms_act_in_their_domain(oracle):
    w = W.multi_domain_workload(n=256, seed=60)
    wfs = domain_waveforms()
    n_steps = 200   # 0.3 mm of drift: no ion reaches the first boundary at 5 mm
    a, b, c = w["ensemble"].copy(), w["ensemble"].copy(), w["ensemble"].copy()
    di = np.zeros(a.n, np.int32)
    oracle.run_multi(w["domains"], w["physics"], w["species"], a, di, 0.0, w["dt"], n_steps,
                     end_apertures=w["apertures"], waveforms=wfs)
    single = dict(domain=w["domains"][0], physics=w["physics"], species=w["species"], dt=w["dt"])
    oracle.run(single["domain"], w["physics"], w["species"], b, 0.0, w["dt"], n_steps,
               field=oracle.Waveforms(wfs[0]))
    oracle.run(single["domain"], w["physics"], w["species"], c, 0.0, w["dt"], n_steps)
    assert (di[a.active == 1] == 0).all()
    for f in ("x", "y", "z", "vx", "vy", "vz", "active", "death_time_s"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    live = a.active == 1
    assert live.sum() > 100 and (a.vz[live] != c.vz[live]).all()


def domain_mixtures(n_species=1):
    """Air-like mixtures at 10 Pa / 300 K for the first and last regions of the
    three-domain workload (the quadrupole keeps its single gas)."""
    n = 10.0 / (1.380649e-23 * 300.0)
    amu = 1.66053906660e-27
    comps = [(0.78 * n, 28.0 * amu, 1), (0.21 * n, 32.0 * amu, 1), (0.01 * n, 40.0 * amu, 1)]
    sigma = np.tile([150e-20, 160e-20, 120e-20], (n_species, 1))
    return [(comps, sigma), None, (comps[:2], sigma[:, :2].copy())]


def test_per_domain_mixtures_act_in_their_domain(oracle):
    w = W.multi_domain_workload(n=256, seed=63, collision=_abi.COLLISION_HSS)
    mixes = domain_mixtures()
    n_steps = 200   # no ion reaches the first boundary
    a, b, c = w["ensemble"].copy(), w["ensemble"].copy(), w["ensemble"].copy()
    di = np.zeros(a.n, np.int32)
    sa = oracle.run_multi(w["domains"], w["physics"], w["species"], a, di, 0.0, w["dt"], n_steps,
                          end_apertures=w["apertures"], mixtures=mixes)
    sb = oracle.run(w["domains"][0], w["physics"], w["species"], b, 0.0, w["dt"], n_steps,
                    mixture=oracle.GasMixture(*mixes[0]))
    sc = oracle.run(w["domains"][0], w["physics"], w["species"], c, 0.0, w["dt"], n_steps)
    for f in ("x", "y", "z", "vx", "vy", "vz", "active", "death_time_s", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    assert sa.collisions == sb.collisions > 0
    assert sa.collisions != sc.collisions


def domain_species_tables():
    """Per-domain rows of the three-domain workload's one species: its cross
    section and K0 as if resolved for N2 at 10 Pa, He at 1 Pa and N2 again
    (the values are the test's own; the handlers' lookup is tested with the
    configuration helpers)."""
    return [[W.species(sigma_hss=240e-20, ccs=240e-20, K0=1.4)],
            [W.species(sigma_hss=90e-20, ccs=90e-20, K0=5.5)],
            [W.species(sigma_hss=210e-20, ccs=210e-20, K0=1.7)]]


@pytest.mark.parametrize("collision,integrator", [(_abi.COLLISION_HSS, _abi.INTEGRATOR_RK4),
                                                  (_abi.COLLISION_FRICTION, _abi.INTEGRATOR_BORIS)])
def test_per_domain_species_act_in_their_domain(oracle, collision, integrator):
    w = W.multi_domain_workload(n=256, seed=64, collision=collision, integrator=integrator)
    if collision == _abi.COLLISION_FRICTION:
        w["physics"].ou_thermal_kicks = 1
    tabs = domain_species_tables()
    n_steps = 200   # no ion reaches the first boundary
    a, b, c = w["ensemble"].copy(), w["ensemble"].copy(), w["ensemble"].copy()
    di = np.zeros(a.n, np.int32)
    oracle.run_multi(w["domains"], w["physics"], w["species"], a, di, 0.0, w["dt"], n_steps,
                     end_apertures=w["apertures"], domain_species=tabs)
    oracle.run(w["domains"][0], w["physics"], tabs[0], b, 0.0, w["dt"], n_steps)
    oracle.run(w["domains"][0], w["physics"], w["species"], c, 0.0, w["dt"], n_steps)
    for f in ("x", "y", "z", "vx", "vy", "vz", "active", "death_time_s", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    assert not np.array_equal(a.vz, c.vz)
    if collision == _abi.COLLISION_HSS:
        # 240 against 150 A^2: more collisions with the domain's row
        assert a.n_collisions.sum() > 1.3 * c.n_collisions.sum() > 0


def test_per_domain_species_equal_to_the_table_change_nothing(oracle):
    w = W.multi_domain_workload(n=256, seed=65, collision=_abi.COLLISION_HSS)
    a, b = w["ensemble"].copy(), w["ensemble"].copy()
    da, db = np.zeros(a.n, np.int32), np.zeros(b.n, np.int32)
    sa = oracle.run_multi(w["domains"], w["physics"], w["species"], a, da, 0.0, w["dt"], 4000,
                          end_apertures=w["apertures"], domain_species=[w["species"], None, w["species"]])
    sb = oracle.run_multi(w["domains"], w["physics"], w["species"], b, db, 0.0, w["dt"], 4000,
                          end_apertures=w["apertures"])
    for f in ("x", "y", "z", "vx", "vy", "vz", "active", "death_time_s", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    np.testing.assert_array_equal(da, db)
    assert sa.collisions == sb.collisions > 0


def field_array_domain_workload(n=256, seed=66, collision=_abi.COLLISION_HSS, integrator=_abi.INTEGRATOR_RK4):
    """The three-domain workload with its middle region (5-15 mm) a drift
    region of radius 2 mm in He at 1 Pa whose field comes from a field array:
    E = (-g x, -g y, E0 (1 + 0.3 sin(2 pi (z - 5 mm) / 10 mm))) on a uniform
    24 x 24 x 41 grid.  Returns the workload and the per-domain fields
    [None, (E, origin, spacing), None]."""
    w = W.multi_domain_workload(n=n, seed=seed, collision=collision, integrator=integrator)
    d1 = W.ims_domain(origin_z=5e-3, length=10e-3, radius=2e-3, axial_V=15.0, gas="He", pressure_Pa=1.0)
    d1.use_field_array = 1
    nx, ny, nz = 24, 24, 41
    origin = (-2.3e-3, -2.3e-3, 5e-3)
    spacing = (4.6e-3 / (nx - 1), 4.6e-3 / (ny - 1), 10e-3 / (nz - 1))
    x = origin[0] + spacing[0] * np.arange(nx)
    y = origin[1] + spacing[1] * np.arange(ny)
    z = origin[2] + spacing[2] * np.arange(nz)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    g, E0 = 2e5, 1500.0
    E = np.stack([-g * X, -g * Y, E0 * (1.0 + 0.3 * np.sin(2 * np.pi * (Z - 5e-3) / 10e-3))], axis=-1)
    w["domains"][1] = d1
    w["apertures"] = [1.5e-3, 0.0, 0.0]
    return w, [None, (E, origin, spacing), None]


def test_per_domain_field_array_acts_in_its_domain(oracle):
    """A domain's field array through the multi-domain loop is the
    single-domain loop on that array, bit for bit, while every ion stays in it."""
    w, fields = field_array_domain_workload(n=128)
    E, origin, spacing = fields[1]
    e = w["ensemble"]
    e.z[:] = 9e-3 + 0.2e-3 * np.sin(np.arange(e.n))   # every ion in the middle domain
    e.x[:] *= 0.5
    e.y[:] *= 0.5
    a, b = e.copy(), e.copy()
    di = np.ones(a.n, np.int32)
    n_steps = 150   # 1.5 km/s x 150 ns: no ion reaches a boundary
    sa = oracle.run_multi(w["domains"], w["physics"], w["species"], a, di, 0.0, w["dt"], n_steps,
                          end_apertures=w["apertures"], fields=[None, oracle.Field(E, origin, spacing), None])
    sb = oracle.run(w["domains"][1], w["physics"], w["species"], b, 0.0, w["dt"], n_steps,
                    field=oracle.Field(E, origin, spacing))
    for f in ("x", "y", "z", "vx", "vy", "vz", "active", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    assert (di == 1).all()
    assert sa.ion_steps == sb.ion_steps == a.n * n_steps
    # the array's focusing pulls |x| in: the analytic drift field has none
    c = e.copy()
    oracle.run(W.ims_domain(origin_z=5e-3, length=10e-3, radius=2e-3, axial_V=15.0, gas="He",
                            pressure_Pa=1.0), w["physics"], w["species"], c, 0.0, w["dt"], n_steps)
    assert not np.array_equal(a.vx, c.vx)
