"""Several domains on the device (icarion_gpu_set_domains) against the oracle's
multi-domain engine loop.

Identical Philox streams on both sides, so ions are compared one by one:
  * collisionless runs: every ion to 1e-9 relative, the same domain index,
    liveness and death time;
  * stochastic runs (HSS, Langevin + OU): >= 99 % of ions to 1e-9 (a
    collision decision within rounding of its threshold can flip), the
    collision count within 0.3 %.
The hand-derived bridge and handoff cases of the oracle's own tests
(tests/test_oracle_domains.py) run on the device and must agree exactly.
FP32 + Kahan runs are held to the FP64 oracle by path agreement and phase-space
error (no gas) and by statistics (with gas).
"""
import numpy as np
import pytest
from scipy import stats

from icarion_amd import _abi
from icarion_amd import workloads as W
from test_gpu_parity import compare_ions
from test_oracle_domains import (_gap_domains, _one_ion, domain_mixtures, domain_species_tables, domain_waveforms,
                                 field_array_domain_workload)

pytestmark = pytest.mark.gpu


def gpu_multi(w, ens, n_steps, t0=0.0, step_offset=0, physics=None, domains=None, apertures=None,
              steps_per_launch=0, h=None):
    from icarion_amd import gpu
    own = h is None
    if own:
        h = gpu.Handle(steps_per_launch=steps_per_launch)
    try:
        if own:
            h.set_domains(domains or w["domains"], physics or w["physics"],
                          apertures if apertures is not None else w.get("apertures"))
            for k, slots in enumerate(w.get("waveforms") or []):
                if slots:
                    h.set_domain_waveforms(k, slots)
            h.set_species(w["species"], w.get("atoms"))
            for k, mx in enumerate(w.get("mixtures") or []):
                if mx:
                    h.set_domain_gas_mixture(k, *mx)
            for k, rows in enumerate(w.get("domain_species") or []):
                if rows:
                    h.set_domain_species(k, rows)
            for k, fld in enumerate(w.get("fields") or []):
                if fld is not None:
                    h.upload_domain_field_array(k, *fld)
            for k, pw in enumerate(w.get("pressures") or []):
                if pw is not None:
                    h.set_domain_pressure_waveform(k, pw)
            for k, fw in enumerate(w.get("flows") or []):
                if fw is not None:
                    h.set_domain_gas_flow(k, *fw)
            h.upload_state(ens)
        st = h.step(t0, w["dt"], n_steps, step_offset)
        h.download_state(ens)
        dom = h.download_domains(ens.n)
    finally:
        if own:
            h.close()
    return st, dom


def oracle_multi(oracle, w, ens, n_steps, t0=0.0, step_offset=0, physics=None, domains=None,
                 apertures=None, dom=None):
    di = np.zeros(ens.n, np.int32) if dom is None else dom
    st = oracle.run_multi(domains or w["domains"], physics or w["physics"], w["species"], ens, di,
                          t0, w["dt"], n_steps, step_offset=step_offset, atoms=w.get("atoms"),
                          mixtures=w.get("mixtures"), domain_species=w.get("domain_species"),
                          fields=[None if f is None else oracle.Field(*f) for f in w["fields"]]
                          if w.get("fields") else None,
                          end_apertures=apertures if apertures is not None else w.get("apertures"),
                          waveforms=w.get("waveforms"), pressures=w.get("pressures"), flows=w.get("flows"))
    return st, di


@pytest.mark.parametrize("integrator,collision", [
    (_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE),
    (_abi.INTEGRATOR_BORIS, _abi.COLLISION_NONE),
    (_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS),
    (_abi.INTEGRATOR_BORIS, _abi.COLLISION_LANGEVIN),
])
def test_three_domains_ion_by_ion(oracle, integrator, collision):
    w = W.multi_domain_workload(n=2048, seed=52, integrator=integrator, collision=collision)
    if collision == _abi.COLLISION_LANGEVIN:
        w["physics"].ou_thermal_kicks = 1
    print(f"domains integ={integrator} coll={collision}", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg, dg = gpu_multi(w, g, 6000)
    so, do = oracle_multi(oracle, w, o, 6000)
    ok = compare_ions(g, o, rtol=1e-9)
    ok &= dg == do
    ok &= g.death_time_s == o.death_time_s
    if collision == _abi.COLLISION_NONE:
        assert ok.all(), np.flatnonzero(~ok)[:10]
        assert sg.ion_steps == so.ion_steps
    else:
        assert ok.mean() >= 0.99
        if collision == _abi.COLLISION_HSS:
            assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    assert sg.alive == pytest.approx(so.alive, rel=0.02, abs=5)
    # the run crosses every domain: live ions sit in several, dead ones died in several
    assert len(set(do[o.active == 1].tolist())) >= 2 or len(set(do.tolist())) >= 3


@pytest.mark.parametrize("integrator,collision", [
    (_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE),
    (_abi.INTEGRATOR_BORIS, _abi.COLLISION_HSS),
])
def test_per_domain_waveforms_ion_by_ion(oracle, integrator, collision):
    """Each domain with its own waveforms (an axial ramp in the first IMS
    region, a modulated RF amplitude and a frequency ramp in the quadrupole, a
    pulsed axial voltage in the last region) against the oracle with the same
    waveforms per domain: ion by ion as the constant-voltage runs are held."""
    w = W.multi_domain_workload(n=2048, seed=61, integrator=integrator, collision=collision)
    w["waveforms"] = domain_waveforms()
    print(f"domain waveforms integ={integrator} coll={collision}", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg, dg = gpu_multi(w, g, 6000)
    so, do = oracle_multi(oracle, w, o, 6000)
    ok = compare_ions(g, o, rtol=1e-9) & (dg == do) & (g.death_time_s == o.death_time_s)
    if collision == _abi.COLLISION_NONE:
        assert ok.all(), np.flatnonzero(~ok)[:10]
        assert sg.ion_steps == so.ion_steps
    else:
        assert ok.mean() >= 0.99
        assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    # the waveforms change the outcome: the constant-voltage run differs
    c = w["ensemble"].copy()
    gpu_multi(dict(w, waveforms=None), c, 6000)
    assert (c.z != g.z).mean() > 0.5
    # clearing a domain's waveforms restores its constants
    from icarion_amd import gpu
    with gpu.Handle() as h:
        h.set_domains(w["domains"], w["physics"], w["apertures"])
        for k, slots in enumerate(w["waveforms"]):
            h.set_domain_waveforms(k, slots)
            h.set_domain_waveforms(k, None)
        h.set_species(w["species"])
        e = w["ensemble"].copy()
        h.upload_state(e)
        h.step(0.0, w["dt"], 6000)
        h.download_state(e)
    np.testing.assert_array_equal(e.z, c.z)


@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS])
def test_ehss_three_domains_ion_by_ion(oracle, integrator):
    """Geometry EHSS over three domains (N2, He, N2): each ion collides with
    its step domain's gas (mass, density, temperature, contact radius), as the
    oracle's multi-domain loop does with the same clusters (1, 5 and 23 atoms,
    FP32 coordinates on both sides)."""
    from test_gpu_parity import ehss_workload
    e = ehss_workload(n=2048, seed=62)
    w = W.multi_domain_workload(n=2048, seed=62, integrator=integrator, collision=_abi.COLLISION_EHSS)
    w["species"], w["atoms"] = e["species"], e["atoms"]
    w["ensemble"].species[:] = e["ensemble"].species
    print(f"domains EHSS integ={integrator}", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg, dg = gpu_multi(w, g, 6000)
    so, do = oracle_multi(oracle, w, o, 6000)
    ok = compare_ions(g, o, rtol=1e-9) & (dg == do) & (g.death_time_s == o.death_time_s)
    assert ok.mean() >= 0.99, ok.mean()
    assert so.collisions > 2 * g.n
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    # collisions happen in every gas: ions that reached the last region collided there
    assert len(set(do[o.active == 1].tolist())) >= 2 or len(set(do.tolist())) >= 3


@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS])
def test_per_domain_gas_mixtures_ion_by_ion(oracle, integrator):
    """HSS with an air-like mixture in the first region, the quadrupole's
    single gas and a two-gas mixture in the last region: each ion collides
    with its step domain's gas, ion by ion against the oracle."""
    w = W.multi_domain_workload(n=2048, seed=64, integrator=integrator, collision=_abi.COLLISION_HSS)
    w["mixtures"] = domain_mixtures()
    print(f"domain mixtures integ={integrator}", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg, dg = gpu_multi(w, g, 6000)
    so, do = oracle_multi(oracle, w, o, 6000)
    ok = compare_ions(g, o, rtol=1e-9) & (dg == do) & (g.death_time_s == o.death_time_s)
    assert ok.mean() >= 0.99, ok.mean()
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    # the mixtures change the run: the single-gas run collides differently
    c = w["ensemble"].copy()
    sc, _ = gpu_multi(dict(w, mixtures=None), c, 6000)
    assert sc.collisions != sg.collisions


def test_ehss_per_domain_gas_mixtures_ion_by_ion(oracle):
    """Geometry EHSS with the per-domain mixtures (air-like first region, the
    quadrupole's single He, a two-gas last region): each lane takes the EHSS
    handler's mixture path with its step domain's components, or the
    single-gas path in the quadrupole, ion by ion against the oracle."""
    from test_gpu_parity import ehss_workload
    e = ehss_workload(n=2048, seed=65)
    w = W.multi_domain_workload(n=2048, seed=65, collision=_abi.COLLISION_EHSS)
    w["species"], w["atoms"] = e["species"], e["atoms"]
    w["ensemble"].species[:] = e["ensemble"].species
    w["mixtures"] = domain_mixtures(n_species=3)
    print("domain EHSS mixtures", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg, dg = gpu_multi(w, g, 6000)
    so, do = oracle_multi(oracle, w, o, 6000)
    ok = compare_ions(g, o, rtol=1e-9) & (dg == do) & (g.death_time_s == o.death_time_s)
    assert ok.mean() >= 0.99, ok.mean()
    assert so.collisions > 2 * g.n
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)


@pytest.mark.parametrize("integrator,collision", [
    (_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS),
    (_abi.INTEGRATOR_BORIS, _abi.COLLISION_FRICTION),
])
def test_per_domain_species_ion_by_ion(oracle, integrator, collision):
    """Each domain with its own species table (cross section, CCS, K0 for its
    gas): an ion collides and is damped with its step domain's row."""
    w = W.multi_domain_workload(n=2048, seed=58, integrator=integrator, collision=collision)
    if collision == _abi.COLLISION_FRICTION:
        w["physics"].ou_thermal_kicks = 1
    w["domain_species"] = domain_species_tables()
    print(f"per-domain species integ={integrator} coll={collision}", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg, dg = gpu_multi(w, g, 6000)
    so, do = oracle_multi(oracle, w, o, 6000)
    ok = compare_ions(g, o, rtol=1e-9) & (dg == do) & (g.death_time_s == o.death_time_s)
    assert ok.mean() >= 0.99
    if collision == _abi.COLLISION_HSS:
        assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    assert sg.alive == pytest.approx(so.alive, rel=0.02, abs=5)
    # the per-domain rows matter: the same run on the handle's table differs
    w2 = dict(w)
    w2["domain_species"] = None
    g2 = w["ensemble"].copy()
    gpu_multi(w2, g2, 6000)
    assert (compare_ions(g2, o, rtol=1e-9)).mean() < 0.5


@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4])
def test_fp32_kahan_per_domain_species_statistics(oracle, integrator):
    w = W.multi_domain_workload(n=8192, seed=59, integrator=integrator, collision=_abi.COLLISION_HSS)
    w["domain_species"] = domain_species_tables()
    print("per-domain species FP32", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    ph = W.physics(integrator, _abi.COLLISION_HSS, precision=_abi.PRECISION_FP32_KAHAN, seed=59)
    sg, dg = gpu_multi(w, g, 4000, physics=ph)
    so, do = oracle_multi(oracle, w, o, 4000)
    assert sg.collisions == pytest.approx(so.collisions, rel=0.02)
    for d in range(3):
        assert (dg == d).sum() == pytest.approx((do == d).sum(), rel=0.05, abs=30)
    assert sg.alive == pytest.approx(so.alive, rel=0.03, abs=30)


@pytest.mark.parametrize("integrator,collision", [
    (_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS),
    (_abi.INTEGRATOR_BORIS, _abi.COLLISION_NONE),
])
def test_per_domain_field_array_ion_by_ion(oracle, integrator, collision):
    """The middle domain's field from a field array, the others analytic."""
    w, fields = field_array_domain_workload(n=2048, seed=67, collision=collision, integrator=integrator)
    w["fields"] = fields
    print(f"per-domain field array integ={integrator} coll={collision}", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg, dg = gpu_multi(w, g, 6000)
    so, do = oracle_multi(oracle, w, o, 6000)
    ok = compare_ions(g, o, rtol=1e-9) & (dg == do) & (g.death_time_s == o.death_time_s)
    print(f"  match {ok.mean():.4f} alive {sg.alive}/{so.alive}", flush=True)
    if collision == _abi.COLLISION_NONE:
        assert ok.all(), np.flatnonzero(~ok)[:10]
        assert sg.ion_steps == so.ion_steps
    else:
        assert ok.mean() >= 0.99
        assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    # many ions have reached the array's domain
    assert (do == 1).sum() > 100


def test_fp32_kahan_per_domain_field_array_collisionless(oracle):
    """FP32 + Kahan through the field-array domain against the FP64 oracle, no
    gas: nearly every ion takes the oracle's path, positions agree to 1e-6 of
    the array domain's radius at the median."""
    w, fields = field_array_domain_workload(n=4096, seed=69, collision=_abi.COLLISION_NONE)
    w["fields"] = fields
    ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE, precision=_abi.PRECISION_FP32_KAHAN, seed=69)
    print("per-domain field array FP32 collisionless", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    # 4000 steps: the leading ions are a few mm into the array domain
    sg, dg = gpu_multi(w, g, 4000, physics=ph)
    so, do = oracle_multi(oracle, w, o, 4000)
    same = (g.active == o.active) & (dg == do) & (g.death_time_s == o.death_time_s)
    both = g.alive & o.alive & (do == 1)
    dp = np.sqrt((g.x - o.x) ** 2 + (g.y - o.y) ** 2 + (g.z - o.z) ** 2)[both] / 2e-3
    print(f"  same path {same.mean():.4f}; in the array domain {both.sum()}", flush=True)
    assert same.mean() > 0.995
    assert both.sum() > 100
    print(f"  pos err/R median {np.median(dp):.2e} p99 {np.quantile(dp, 0.99):.2e}", flush=True)
    assert np.median(dp) < 1e-6 and np.quantile(dp, 0.99) < 1e-5


def test_fp32_kahan_per_domain_field_array_statistics(oracle):
    """With gas: collision count within 2 %, the ions per domain index and the
    survivors within four standard errors of the difference of two binomial
    samples."""
    w, fields = field_array_domain_workload(n=8192, seed=68, collision=_abi.COLLISION_HSS)
    w["fields"] = fields
    print("per-domain field array FP32", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, precision=_abi.PRECISION_FP32_KAHAN, seed=68)
    sg, dg = gpu_multi(w, g, 5000, physics=ph)
    so, do = oracle_multi(oracle, w, o, 5000)
    assert sg.collisions == pytest.approx(so.collisions, rel=0.02)
    n = g.n

    def within(a, b, what):
        p = b / n
        tol = 4.0 * np.sqrt(2.0 * n * p * (1.0 - p)) + 5
        print(f"  {what}: {a} vs {b} (tolerance {tol:.0f})", flush=True)
        assert abs(a - b) <= tol, what

    for d in range(3):
        within(int((dg == d).sum()), int((do == d).sum()), f"domain {d}")
    within(int(sg.alive), int(so.alive), "alive")


def test_domain_field_array_arguments():
    from icarion_amd import gpu
    w, fields = field_array_domain_workload(n=16)
    E, origin, spacing = fields[1]
    with gpu.Handle() as h:
        h.set_domains(w["domains"], w["physics"], w["apertures"])
        h.set_species(w["species"])
        h.upload_state(w["ensemble"].copy())
        with pytest.raises(gpu.IcarionError):   # the domain needs its array
            h.step(0.0, w["dt"], 4)
        with pytest.raises(gpu.IcarionError):   # domain 0 has use_field_array = 0
            h.upload_domain_field_array(0, E, origin, spacing)
        with pytest.raises(gpu.IcarionError):
            h.upload_domain_field_array(5, E, origin, spacing)
        h.upload_domain_field_array(1, E.astype(np.float32), origin, spacing)
        h.step(0.0, w["dt"], 4)


def test_domain_species_arguments():
    from icarion_amd import gpu
    w = W.multi_domain_workload(n=16, collision=_abi.COLLISION_HSS)
    with gpu.Handle() as h:
        h.set_domains(w["domains"], w["physics"], w["apertures"])
        h.set_species(w["species"])
        with pytest.raises(gpu.IcarionError):
            h.set_domain_species(3, w["species"])
        with pytest.raises(gpu.IcarionError):
            h.set_domain_species(0, w["species"] * 2)
        with pytest.raises(gpu.IcarionError):
            h.set_domain_species(0, [W.species(sigma_hss=1e-18, atom_offset=0, atom_count=4)])
        h.set_domain_species(1, domain_species_tables()[1])
        h.set_domain_species(1, [])
        h.upload_state(w["ensemble"].copy())
        h.step(0.0, w["dt"], 10)


def test_bridge_and_handoff_cases_match_the_oracle(oracle):
    w = {"species": [W.species()], "dt": 1e-9}
    ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE)
    # bridge into a gap: alive one step with the next index, dead at the next start
    for n_steps in (1, 2):
        g, o = _one_ion(4.999e-3, vz=5e5), _one_ion(4.999e-3, vz=5e5)
        _, dg = gpu_multi(w, g, n_steps, physics=ph, domains=_gap_domains(), apertures=[2e-3, 0.0])
        _, do = oracle_multi(oracle, w, o, n_steps, physics=ph, domains=_gap_domains(),
                             apertures=[2e-3, 0.0])
        assert (g.active[0], dg[0], g.death_time_s[0]) == (o.active[0], do[0], o.death_time_s[0])
        assert g.z[0] == o.z[0]
    # handoff within 1e-9 m of the end face, inside / outside the aperture
    d0 = W.ims_domain(origin_z=0.0, length=5e-3, radius=3e-3, axial_V=0.0, gas="N2", pressure_Pa=10.0)
    d1 = W.ims_domain(origin_z=5e-3, length=5e-3, radius=1e-3, axial_V=0.0, gas="N2", pressure_Pa=10.0)
    phb = W.physics(_abi.INTEGRATOR_BORIS, _abi.COLLISION_NONE)
    for r, ap, expect in ((0.5e-3, 1.5e-3, 1), (2.0e-3, 1.5e-3, 0), (2.0e-3, 0.0, 1)):
        g = _one_ion(5e-3 - 0.5e-9, r=r)
        _, dg = gpu_multi(w, g, 3, physics=phb, domains=[d0, d1], apertures=[ap, 0.0])
        assert g.active[0] == 1 and dg[0] == expect, (r, ap)


def test_split_calls_and_launch_shape():
    w = W.multi_domain_workload(n=1024, seed=53, collision=_abi.COLLISION_HSS)
    a, b, c = w["ensemble"].copy(), w["ensemble"].copy(), w["ensemble"].copy()
    _, da = gpu_multi(w, a, 3000)
    from icarion_amd import gpu
    with gpu.Handle() as h:
        h.set_domains(w["domains"], w["physics"], w["apertures"])
        h.set_species(w["species"])
        h.upload_state(b)
        st = h.step(0.0, w["dt"], 1300, 0)
        st = h.step(st.t_end_s, w["dt"], 1700, 1300)
        h.download_state(b)
        db = h.download_domains(b.n)
    _, dc = gpu_multi(w, c, 3000, steps_per_launch=97)
    for f in ("x", "y", "z", "vx", "vy", "vz", "active", "death_time_s", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
        np.testing.assert_array_equal(getattr(a, f), getattr(c, f), err_msg=f)
    np.testing.assert_array_equal(da, db)
    np.testing.assert_array_equal(da, dc)


def test_one_domain_equals_set_domain():
    from test_gpu_parity import gpu_run
    w = W.multi_domain_workload(n=1024, seed=54, collision=_abi.COLLISION_HSS)
    single = dict(w, domain=w["domains"][0])
    a, b = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(single, a, 2000)
    gpu_multi(w, b, 2000, domains=[w["domains"][0]], apertures=[0.0])
    for f in ("x", "y", "z", "vx", "vy", "vz", "active", "death_time_s", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)


def test_unsupported_combinations():
    from icarion_amd import gpu
    w = W.multi_domain_workload(n=64, seed=55)
    p = w["physics"]
    cases = [
        W.physics(_abi.INTEGRATOR_RK45, p.collision_model),
        W.physics(p.integrator, _abi.COLLISION_EHSS, precision=_abi.PRECISION_FP32_KAHAN),
        W.physics(p.integrator, _abi.COLLISION_EHSS, ehss_mode=_abi.EHSS_SURROGATE,
                  precision=_abi.PRECISION_FP32_KAHAN),
    ]
    for ph in cases:
        with gpu.Handle() as h:
            h.set_domains(w["domains"], ph, w["apertures"])
            h.set_species(w["species"])
            h.upload_state(w["ensemble"].copy())
            with pytest.raises(gpu.IcarionError) as ei:
                h.step(0.0, w["dt"], 5)
            assert ei.value.code == _abi.ERR_UNSUPPORTED
    with gpu.Handle() as h:
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_domains(w["domains"] * 3, p)
        assert ei.value.code == _abi.ERR_UNSUPPORTED
    # handle-level waveforms with several domains; a domain index out of range
    with gpu.Handle() as h:
        h.set_domains(w["domains"], p, w["apertures"])
        h.set_species(w["species"])
        h.upload_state(w["ensemble"].copy())
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_domain_waveforms(3, domain_waveforms()[0])
        assert ei.value.code == _abi.ERR_INVALID
        h.set_waveforms(domain_waveforms()[0])
        with pytest.raises(gpu.IcarionError) as ei:
            h.step(0.0, w["dt"], 5)
        assert ei.value.code == _abi.ERR_UNSUPPORTED
        h.set_waveforms(None)
        # a per-domain mixture needs HSS; a handle-level one is not combined
        h.set_domain_gas_mixture(0, *domain_mixtures()[0])
        with pytest.raises(gpu.IcarionError) as ei:
            h.step(0.0, w["dt"], 5)   # the workload's physics: no collisions
        assert ei.value.code == _abi.ERR_UNSUPPORTED


@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS])
def test_fp32_kahan_three_domains_collisionless(oracle, integrator):
    """FP32 + Kahan over three domains against the FP64 oracle, no gas.  The
    domain lookup, bridge and handoff run in FP64 on the (hi, lo) position, so
    nearly every ion takes the oracle's path (liveness, death step, domain
    index); positions of ions alive on both sides agree to 1e-6 of the
    quadrupole radius at the median and 1e-5 at the 99th percentile (all
    masses sit at q <= 0.55, far from the stability edge)."""
    w = W.multi_domain_workload(n=4096, seed=57, integrator=integrator)
    ph = W.physics(integrator, _abi.COLLISION_NONE, precision=_abi.PRECISION_FP32_KAHAN, seed=57)
    print(f"domains fp32 integ={integrator}", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg, dg = gpu_multi(w, g, 6000, physics=ph)
    so, do = oracle_multi(oracle, w, o, 6000)
    same = (g.active == o.active) & (dg == do) & (g.death_time_s == o.death_time_s)
    assert same.mean() > 0.995, same.mean()
    assert sg.ion_steps == pytest.approx(so.ion_steps, rel=2e-3)
    both = g.alive & o.alive
    assert both.sum() > 200
    r0 = 2e-3
    dp = np.sqrt((g.x - o.x) ** 2 + (g.y - o.y) ** 2 + (g.z - o.z) ** 2)[both] / r0
    q50, q99 = np.quantile(dp, 0.5), np.quantile(dp, 0.99)
    print(f"pos err/r0 median {q50:.2e} p99 {q99:.2e}")
    assert q50 < 1e-6 and q99 < 1e-5


@pytest.mark.parametrize("integrator,collision", [
    (_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS),
    (_abi.INTEGRATOR_BORIS, _abi.COLLISION_LANGEVIN),
])
def test_fp32_kahan_three_domains_statistics(oracle, integrator, collision):
    """FP32 + Kahan with gas over three domains against the FP64 oracle: the
    surviving fraction in each domain and overall within four binomial standard
    errors (+0.5 %), the mean collision count within 0.5 %, and the axial
    distribution of the survivors by a two-sample KS test."""
    w = W.multi_domain_workload(n=20000, seed=58, integrator=integrator, collision=collision)
    ph = W.physics(integrator, collision, precision=_abi.PRECISION_FP32_KAHAN, seed=58)
    if collision == _abi.COLLISION_LANGEVIN:
        ph.ou_thermal_kicks = 1
        w["physics"].ou_thermal_kicks = 1
    print(f"domains fp32 integ={integrator} coll={collision}", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg, dg = gpu_multi(w, g, 6000, physics=ph)
    so, do = oracle_multi(oracle, w, o, 6000)
    n = g.n
    for d in (None, 0, 1, 2):
        sel_g = g.alive if d is None else g.alive & (dg == d)
        sel_o = o.alive if d is None else o.alive & (do == d)
        fg, fo = sel_g.mean(), sel_o.mean()
        err = 4 * np.sqrt(max(fo * (1 - fo), 1e-4) / n) + 0.005
        assert abs(fg - fo) < err, (d, fg, fo)
    if collision == _abi.COLLISION_HSS:
        assert g.n_collisions.mean() == pytest.approx(o.n_collisions.mean(), rel=5e-3)
    assert stats.ks_2samp(g.z[g.alive], o.z[o.alive]).pvalue > 1e-3


def test_snapshots_over_several_domains():
    """Trajectory snapshots of a three-domain run: every snapshot (positions,
    velocities, liveness, domain indices) equals segmented stepping to that
    step bit for bit, and the domain indices move through the instrument."""
    from icarion_amd import gpu
    w = W.multi_domain_workload(n=3000, seed=56)
    dt, n_steps = w["dt"], 9000
    with gpu.Handle() as h:
        h.set_domains(w["domains"], w["physics"], w["apertures"])
        h.set_species(w["species"])
        h.upload_state(w["ensemble"].copy())
        snaps, _, _ = h.step_snapshots(0.0, dt, n_steps, 300 * dt)
    assert len(snaps) >= 29
    # late snapshots: few ions still move, so only their records cross PCIe
    assert ((snaps[-1]["flags"] & 3) == 3).mean() < 1 / 8
    with gpu.Handle() as h:
        h.set_domains(w["domains"], w["physics"], w["apertures"])
        h.set_species(w["species"])
        e = w["ensemble"].copy()
        h.upload_state(e)
        t, prev = 0.0, 0
        for s in snaps:
            st = h.step(t, dt, s["step"] + 1 - prev, prev)
            h.download_state(e)
            d = h.download_domains(e.n)
            np.testing.assert_array_equal(s["positions"], np.stack([e.x, e.y, e.z], axis=1))
            np.testing.assert_array_equal(s["velocities"], np.stack([e.vx, e.vy, e.vz], axis=1))
            np.testing.assert_array_equal(s["flags"] & 1, e.active)
            np.testing.assert_array_equal(s["domain"], d)
            t, prev = st.t_end_s, s["step"] + 1
    seen = set()
    for s in snaps:
        seen |= set(s["domain"].tolist())
    assert seen == {0, 1, 2}
