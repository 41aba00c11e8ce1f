"""Parity of the gfx950 kernels (through the C ABI) with the CPU oracle.

FP64 plain path: identical Philox streams on both sides, so trajectories are
compared ion by ion.  Tolerances:
  * collisionless runs: relative 1e-9 over the whole run (the per-step bound of
    the north-star criterion, held cumulatively);
  * stochastic runs: >= 99 % of ions match to 1e-9 (a collision decision that
    sits within rounding of its threshold can flip and send one ion on another
    path), and the ensemble statistics match to 0.2 %.
FP32 + Kahan fast path: against the FP64 oracle, relative 2e-5 on collisionless
runs and statistical agreement (mobility 1 %, KS test) on stochastic runs.
"""
import math

import numpy as np
import pytest
from scipy import stats

from icarion_amd import _abi
from icarion_amd import workloads as W

pytestmark = pytest.mark.gpu


def gpu_run(w, ens, n_steps, t0=0.0, step_offset=0, precision=None, steps_per_launch=0,
            block=0):
    from icarion_amd import gpu
    phys = w["physics"]
    if precision is not None:
        phys = W.physics(phys.integrator, phys.collision_model, phys.ou_thermal_kicks, precision,
                         phys.rng_seed, phys.ehss_mode, phys.collision_substeps)
    with gpu.Handle(steps_per_launch=steps_per_launch, block_size=block) as h:
        h.set_domain(w["domain"], phys)
        h.set_species(w["species"], w.get("atoms"))
        if w.get("field") is not None:
            E, origin, spacing = w["field"]
            h.upload_field_array(E, origin, spacing)
        if w.get("terms") is not None:
            h.upload_field_terms(w["axes"], w["terms"])
        if w.get("mixture") is not None:
            h.set_gas_mixture(*w["mixture"])
            if w.get("mixture_radii") is not None:
                h.set_gas_radii(w["mixture_radii"])
        if w.get("ehss_offline") is not None:
            h.set_ehss_offline(w["ehss_offline"])
        if w.get("ehss_orient") is not None:
            h.set_ehss_orientations(w["ehss_orient"])
        if w.get("waveforms") is not None:
            h.set_waveforms(w["waveforms"])
        if w.get("pressure") is not None:
            h.set_pressure_waveform(w["pressure"])
        if w.get("flow") is not None:
            h.set_gas_flow(*w["flow"])
        h.upload_state(ens)
        st = h.step(t0, w["dt"], n_steps, step_offset)
        h.download_state(ens)
    return st


def oracle_run(oracle, w, ens, n_steps, t0=0.0, step_offset=0):
    field = None
    if w.get("field") is not None:
        field = oracle.Field(*w["field"])
    if w.get("terms") is not None:
        field = oracle.FieldTerms(w["axes"], w["terms"], w.get("waveforms"))
    elif w.get("waveforms") is not None and w.get("field") is None:
        field = oracle.Waveforms(w["waveforms"])
    mixture = (oracle.GasMixture(*w["mixture"], radii=w.get("mixture_radii"))
               if w.get("mixture") is not None else None)
    off = oracle.EhssOffline(w["ehss_offline"]) if w.get("ehss_offline") is not None else None
    ori = oracle.EhssOrientations(w["ehss_orient"]) if w.get("ehss_orient") is not None else None
    return oracle.run(w["domain"], w["physics"], w["species"], ens, t0, w["dt"], n_steps,
                      step_offset=step_offset, atoms=w.get("atoms"), field=field, mixture=mixture,
                      ehss_offline=off, ehss_orient=ori, pressure=w.get("pressure"), flow=w.get("flow"))


def rel_close(a, b, rtol, scale):
    return np.abs(a - b) <= rtol * np.maximum(np.abs(b), scale)


def compare_ions(g, o, rtol=1e-9, pos_scale=1e-6, vel_scale=1.0):
    ok = np.ones(g.n, bool)
    for f in ("x", "y", "z"):
        ok &= rel_close(getattr(g, f), getattr(o, f), rtol, pos_scale)
    for f in ("vx", "vy", "vz"):
        ok &= rel_close(getattr(g, f), getattr(o, f), rtol, vel_scale)
    ok &= g.active == o.active
    return ok


def small(w, n):
    w = dict(w)
    w["ensemble"] = w["ensemble"].subset(np.arange(n))
    return w


# ------------------------------------------------------------------ collisionless
@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS])
def test_quadrupole_collisionless_fp64(oracle, integrator):
    w = W.quadrupole_workload(n=4096, U=20.0, seed=3)
    w["physics"] = W.physics(integrator, _abi.COLLISION_NONE, seed=3)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    st = gpu_run(w, g, 3000, t0=1e-7)
    oracle_run(oracle, w, o, 3000, t0=1e-7)
    ok = compare_ions(g, o, rtol=1e-9)
    assert ok.all(), np.nonzero(~ok)[0][:10]
    np.testing.assert_array_equal(g.death_time_s, o.death_time_s)
    assert st.ion_steps > 0 and 0 < st.alive < 4096


def test_orbitrap_collisionless_fp64(oracle):
    w = W.orbitrap_workload(n=2048, seed=5)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 5000)
    oracle_run(oracle, w, o, 5000)
    ok = compare_ions(g, o, rtol=1e-9)
    assert ok.all()
    assert g.alive.mean() > 0.9


def test_ims_collisionless_single_step_bitwise_close(oracle):
    w = W.ims_hss_workload(n=4096)
    w["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 1)
    oracle_run(oracle, w, o, 1)
    assert compare_ions(g, o, rtol=1e-14).all()


# ------------------------------------------------------------------ stochastic FP64
def test_ims_hss_fp64_ion_by_ion(oracle):
    w = W.ims_hss_workload(n=4096, seed=123)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    z0 = w["ensemble"].z.copy()
    sg = gpu_run(w, g, 3000)
    so = oracle_run(oracle, w, o, 3000)
    ok = compare_ions(g, o, rtol=1e-9)
    assert ok.mean() >= 0.99, ok.mean()
    assert (g.n_collisions == o.n_collisions).mean() >= 0.99
    assert sg.ion_steps == so.ion_steps
    assert sg.collisions == pytest.approx(so.collisions, rel=2e-3)
    dzg = (g.z - z0)[g.alive].mean()
    dzo = (o.z - z0)[o.alive].mean()
    assert dzg == pytest.approx(dzo, rel=2e-3)


def test_quadrupole_langevin_ou_fp64_ion_by_ion(oracle):
    w = W.quadrupole_workload(n=4096, U=10.0, seed=9)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 3000)
    oracle_run(oracle, w, o, 3000)
    ok = compare_ions(g, o, rtol=1e-9)
    assert ok.mean() >= 0.99, ok.mean()
    assert (g.active == o.active).mean() >= 0.995


def test_friction_fp64(oracle):
    w = W.ims_hss_workload(n=2048)
    w["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_FRICTION, ou=1, seed=4)
    w["species"] = [W.species(K0=2.1, ccs=100e-20)]
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 2000)
    oracle_run(oracle, w, o, 2000)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99


def ehss_workload(n=1024, seed=31, atom_dtype=np.float32):
    rng = np.random.default_rng(seed)
    clusters = []
    for nat in (1, 5, 23):
        pos = rng.normal(0, 3e-10, (nat, 3))
        pos -= pos.mean(0)
        clusters.append(np.concatenate([pos, rng.uniform(1.5e-10, 2e-10, (nat, 1))], 1))
    # the device keeps atoms in float; the oracle reads the same float values
    atoms = np.concatenate(clusters).astype(atom_dtype)
    offs = np.cumsum([0] + [len(c) for c in clusters])[:-1]
    sp = [W.species(sigma_ehss=s, atom_offset=int(a), atom_count=len(c))
          for s, a, c in zip((40e-20, 90e-20, 200e-20), offs, clusters)]
    w = W.ims_hss_workload(n=n, seed=seed)
    w["domain"] = W.ims_domain(EN_Td=20.0, gas="He", pressure_Pa=101325.0)
    w["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_EHSS, seed=seed)
    w["species"] = sp
    w["atoms"] = atoms
    w["ensemble"].species[:] = rng.integers(0, 3, n)
    return w


@pytest.mark.skip(reason="a GPU memory fault whose cause was not found came up in this test's "
                         "run with a double atom table on the device; that device table is "
                         "switched off (atoms are held in float) until the cause is shown")
def test_ehss_fp64_double_atom_table(oracle):
    w = ehss_workload(atom_dtype=np.float64)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg = gpu_run(w, g, 1000)
    so = oracle_run(oracle, w, o, 1000)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)


def test_double_atom_table_declined():
    """A double atom table is declined (UNSUPPORTED), not rounded silently."""
    from icarion_amd import gpu
    w = ehss_workload(n=64, atom_dtype=np.float64)
    with gpu.Handle(device=0) as h:
        h.set_domain(w["domain"], w["physics"])
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_species(w["species"], w["atoms"])
        assert ei.value.code == _abi.ERR_UNSUPPORTED


def test_ehss_fp64_ion_by_ion(oracle):
    w = ehss_workload()
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg = gpu_run(w, g, 1000)
    so = oracle_run(oracle, w, o, 1000)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)


@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS])
def test_ehss_offline_sets_fp64_ion_by_ion(oracle, integrator):
    """Offline EHSS sample sets for two of three species (orientation-resolved
    sigma_eff and mu samples), the online geometry for the third."""
    w = ehss_workload(n=2048, seed=35)
    w["physics"] = W.physics(integrator, _abi.COLLISION_EHSS, seed=35)
    rng = np.random.default_rng(35)
    w["ehss_offline"] = [
        (rng.uniform(30e-20, 60e-20, 16), np.sqrt(rng.uniform(0.0, 1.0, (16, 50)))),
        None,
        (rng.uniform(150e-20, 260e-20, 40), rng.uniform(-0.2, 1.0, (40, 7))),
    ]
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg = gpu_run(w, g, 1000)
    so = oracle_run(oracle, w, o, 1000)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    # the sets change the result: without them every species uses the geometry
    plain = dict(w)
    plain.pop("ehss_offline")
    o2 = w["ensemble"].copy()
    so2 = oracle_run(oracle, plain, o2, 1000)
    assert compare_ions(g, o2, rtol=1e-9).mean() < 0.9


def orientation_sets(rng, n_species, with_samples):
    """Per species: None, or random (unnormalised) quaternions with areas."""
    sets = []
    for s_ in range(n_species):
        if not with_samples[s_]:
            sets.append(None)
            continue
        n = int(rng.integers(3, 40))
        sets.append((rng.normal(size=(n, 4)), rng.uniform(40e-20, 250e-20, n)))
    return sets


@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS])
def test_ehss_orientation_samples_fp64_ion_by_ion(oracle, integrator):
    """EHSS orientation samples (stored quaternions + areas) for two of three
    species, the online geometry for the third; an offline set for one of the
    two takes precedence over its orientation samples."""
    w = ehss_workload(n=2048, seed=37)
    w["physics"] = W.physics(integrator, _abi.COLLISION_EHSS, seed=37)
    rng = np.random.default_rng(37)
    w["ehss_orient"] = orientation_sets(rng, 3, (True, False, True))
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg = gpu_run(w, g, 1000)
    so = oracle_run(oracle, w, o, 1000)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    plain = dict(w)
    plain.pop("ehss_orient")
    o2 = w["ensemble"].copy()
    oracle_run(oracle, plain, o2, 1000)
    assert compare_ions(g, o2, rtol=1e-9).mean() < 0.9
    # precedence: with an offline set for species 0 as well, species 0 follows
    # the offline set exactly as without orientation samples
    w3 = dict(w)
    w3["ehss_offline"] = [(rng.uniform(30e-20, 60e-20, 16), np.sqrt(rng.uniform(0.0, 1.0, (16, 50)))),
                          None, None]
    w4 = dict(w3)
    w4["ehss_orient"] = [None, None, w["ehss_orient"][2]]
    a, b = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w3, a, 600)
    gpu_run(w4, b, 600)
    for f in ("x", "y", "z", "vx", "vy", "vz", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)


@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS])
def test_ehss_mixture_fp64_ion_by_ion(oracle, integrator):
    """EHSS over a four-component gas mixture: one component does not take
    part, one takes part at zero density (its neutral is still drawn), two
    collide with their own cross sections and contact radii; offline sets are
    not used with a mixture."""
    w = ehss_workload(n=2048, seed=38)
    w["physics"] = W.physics(integrator, _abi.COLLISION_EHSS, seed=38)
    n = w["domain"].gas_density_m3
    comps = [(0.5 * n, 28.0134 * W.AMU_TO_KG, 1), (0.2 * n, 39.948 * W.AMU_TO_KG, 0),
             (0.0, 44.01 * W.AMU_TO_KG, 1), (0.5 * n, 4.0026 * W.AMU_TO_KG, 1)]
    sig = np.array([[60e-20, 50e-20, 70e-20, 40e-20], [130e-20, 110e-20, 150e-20, 90e-20],
                    [260e-20, 240e-20, 300e-20, 200e-20]])
    w["mixture"] = (comps, sig)
    w["mixture_radii"] = [1.8e-10, 1.7e-10, 2.0e-10, 0.0]
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg = gpu_run(w, g, 1000)
    so = oracle_run(oracle, w, o, 1000)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    # offline sets are ignored with a mixture: bit for bit the same run
    w2 = dict(w)
    rng = np.random.default_rng(38)
    w2["ehss_offline"] = [(rng.uniform(30e-20, 60e-20, 16), np.sqrt(rng.uniform(0.0, 1.0, (16, 50)))),
                          None, None]
    a = w["ensemble"].copy()
    gpu_run(w2, a, 1000)
    for f in ("x", "y", "z", "vx", "vy", "vz", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(g, f), err_msg=f)


def test_ehss_mixture_controls():
    from icarion_amd import gpu
    w = ehss_workload(n=64, seed=39)
    n = w["domain"].gas_density_m3
    with gpu.Handle() as h:
        h.set_domain(w["domain"], w["physics"])
        h.set_species(w["species"], w["atoms"])
        h.set_gas_mixture([(n, 28.0134 * W.AMU_TO_KG, 1)], np.full((3, 1), 100e-20))
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_gas_radii([1e-10, 1e-10])
        assert ei.value.code == _abi.ERR_INVALID
        h.set_gas_radii([1.8e-10])
        h.set_ehss_orientations([(np.array([[1.0, 0, 0, 0]]), np.array([1e-18])), None, None])
        h.upload_state(w["ensemble"].copy())
        with pytest.raises(gpu.IcarionError) as ei:
            h.step(0.0, w["dt"], 5)
        assert ei.value.code == _abi.ERR_UNSUPPORTED
        h.set_ehss_orientations([])
        st = h.step(0.0, w["dt"], 5)
        assert st.ion_steps > 0


def test_ehss_offline_controls():
    from icarion_amd import gpu
    w = ehss_workload(n=64, seed=36)
    with gpu.Handle() as h:
        h.set_domain(w["domain"], w["physics"])
        h.set_species(w["species"], w["atoms"])
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_ehss_offline([None, None])
        assert ei.value.code == _abi.ERR_INVALID
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_ehss_offline([None, None, None])
        assert ei.value.code == _abi.ERR_INVALID
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_ehss_offline([(np.array([1e-18, np.nan]), np.ones((2, 3))), None, None])
        assert ei.value.code == _abi.ERR_INVALID
        h.set_ehss_offline([(np.array([1e-18]), np.ones((1, 3))), None, None])
        h.set_ehss_offline([])


def test_field_array_fp64(oracle):
    w = W.field_array_workload(n=4096, grid=(24, 20, 40), seed=2)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 1500)
    oracle_run(oracle, w, o, 1500)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99


def test_field_array_fp32_collisionless_against_fp64_oracle(oracle):
    """The uniform single-grid FP32 + Kahan kernel (step_f32_kernel<RK4,
    FIELD_ARRAY, ...>, the field-array workload's fast path) without gas: ion by
    ion against the FP64 oracle's trilinear interpolation over 3000 steps.
    Errors as vector norms: position against 1 mm (median 1e-6, max 1e-5),
    velocity against the ion's speed (median 1e-6, max 2e-5); the same fate for
    99.9 % of ions.  (A CPU float32 emulation of the kernel gives medians
    2.7e-7 / 1.1e-7 and maxima 1.3e-6 / 4.6e-6 on this workload.)"""
    w = W.field_array_workload(n=4096, grid=(24, 20, 40), seed=2)
    w["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE, seed=2)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 3000, precision=_abi.PRECISION_FP32_KAHAN)
    oracle_run(oracle, w, o, 3000)
    assert (g.active == o.active).mean() >= 0.999
    both = g.alive & o.alive
    dp = np.linalg.norm(g.pos - o.pos, axis=1)[both] / 1e-3
    speed = np.maximum(np.linalg.norm(o.vel, axis=1), 10.0)
    dv = (np.linalg.norm(g.vel - o.vel, axis=1) / speed)[both]
    print("field array FP32 vs FP64 oracle: position/1mm median %.2e max %.2e; velocity/speed "
          "median %.2e max %.2e" % (np.median(dp), dp.max(), np.median(dv), dv.max()))
    assert np.median(dp) < 1e-6 and dp.max() < 1e-5
    assert np.median(dv) < 1e-6 and dv.max() < 2e-5
    # the ions crossed cells: the test sees the interpolation, not one cell
    assert np.abs(g.z - w["ensemble"].z).max() > 50e-3 / 39


def test_field_array_fp32_hss_statistics_against_fp64_oracle(oracle):
    """The uniform-grid FP32 + Kahan kernel with HSS in N2 (the field-array
    workload's kernel, step_f32_kernel<RK4, FIELD_ARRAY, HSS>) against the FP64
    oracle: survival 0.5 %, mean drift 1 %, drift spread 3 %, collisions per
    ion 0.5 %, and a two-sample KS test on the final z (p > 1e-3)."""
    w = W.field_array_workload(n=20000, grid=(24, 20, 40), seed=4)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    z0 = w["ensemble"].z.copy()
    sg = gpu_run(w, g, 4000, precision=_abi.PRECISION_FP32_KAHAN)
    so = oracle_run(oracle, w, o, 4000)
    assert abs(g.alive.mean() - o.alive.mean()) < 0.005
    dzg = (g.z - z0)[g.alive]
    dzo = (o.z - z0)[o.alive]
    assert abs(dzo.mean()) > 5 * dzo.std() / math.sqrt(dzo.size)  # a real drift to compare
    assert dzg.mean() == pytest.approx(dzo.mean(), rel=0.01)
    assert dzg.std() == pytest.approx(dzo.std(), rel=0.03)
    assert stats.ks_2samp(g.z[g.alive], o.z[o.alive]).pvalue > 1e-3
    assert sg.collisions / sg.ion_steps == pytest.approx(so.collisions / so.ion_steps, rel=0.005)


@pytest.mark.parametrize("precision", [_abi.PRECISION_FP64, _abi.PRECISION_FP32_KAHAN])
def test_field_array_brick_order_changes_nothing(precision):
    """Field-array clouds of 65536 ions or more run each launch in grid-brick
    order (lane t takes ion order[t]); the allocation-free entry point keeps
    lane = ion.  Both give the same state bit for bit, over launches that re-sort
    a cloud that moves, with dead ions (sorted last) and ions born during the
    run."""
    from icarion_amd import gpu
    w = W.field_array_workload(n=100_000, grid=(40, 36, 90), seed=6)
    ph = w["physics"]
    ph = W.physics(ph.integrator, ph.collision_model, precision=precision, seed=6)
    e0 = w["ensemble"].copy()
    e0.active[::10] = 0
    e0.birth_time_s = np.zeros(e0.n)
    e0.birth_time_s[5::20] = 300 * w["dt"]
    e0.born[5::20] = 0
    res = []
    for noalloc in (False, True):
        e = e0.copy()
        with gpu.Handle(steps_per_launch=150) as h:
            h.set_domain(w["domain"], ph)
            h.set_species(w["species"])
            h.upload_field_array(*w["field"])
            h.upload_state(e)
            st = (h.step_noalloc if noalloc else h.step)(0.0, w["dt"], 600, 0)
            h.download_state(e)
        res.append((e, st))
    (a, sa), (b, sb) = res
    for f in ("x", "y", "z", "vx", "vy", "vz", "active", "death_time_s", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    assert sa.ion_steps == sb.ion_steps and sa.collisions == sb.collisions
    assert sa.alive == 90_000 and (a.born != 0).all()


@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS])
def test_field_terms_nonuniform_collisionless_fp64(oracle, integrator):
    """Composite field (DC + RF + constant terms) on non-uniform axes: ion by
    ion against the oracle's axis search and CompositeFieldProvider sum."""
    w = W.field_terms_workload(n=2048, seed=3, integrator=integrator)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 1500, t0=2e-8)
    oracle_run(oracle, w, o, 1500, t0=2e-8)
    assert compare_ions(g, o, rtol=1e-9).all()
    # the ions moved through the grid: several cells along z
    assert np.abs(g.z - w["ensemble"].z).max() > 1e-4


def test_field_terms_hss_fp64(oracle):
    w = W.field_terms_workload(n=2048, seed=4, collision=_abi.COLLISION_HSS)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg = gpu_run(w, g, 1500)
    so = oracle_run(oracle, w, o, 1500)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)


def test_field_terms_fp32_against_fp64_oracle(oracle):
    w = W.field_terms_workload(n=2048, seed=5, n_terms=2)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 1500, precision=_abi.PRECISION_FP32_KAHAN)
    oracle_run(oracle, w, o, 1500)
    assert compare_ions(g, o, rtol=2e-5, pos_scale=1e-3, vel_scale=10.0).mean() >= 0.99


def test_field_terms_single_uniform_term_equals_field_array():
    """One term on uniform axes is the plain field array (unscaled, whatever
    its kind): bit for bit the same trajectories as icarion_gpu_upload_field_array."""
    w = W.field_terms_workload(n=1024, seed=6, uniform=True, n_terms=1,
                               collision=_abi.COLLISION_HSS)
    w["terms"] = [(w["terms"][0][0], _abi.FIELD_SCALE_RF, 0.0, 3e6, 1.0)]
    xs, ys, zs = w["axes"]
    plain = dict(w)
    plain.pop("terms")
    plain["field"] = (w["terms"][0][0], (xs[0], ys[0], zs[0]),
                      ((xs[-1] - xs[0]) / (len(xs) - 1), (ys[-1] - ys[0]) / (len(ys) - 1),
                       (zs[-1] - zs[0]) / (len(zs) - 1)))
    a, b = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, a, 800)
    gpu_run(plain, b, 800)
    for f in ("x", "y", "z", "vx", "vy", "vz", "active", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)


def test_field_terms_reject_bad_input():
    from icarion_amd import gpu
    w = W.field_terms_workload(n=16, seed=7)
    xs, ys, zs = w["axes"]
    with gpu.Handle() as h:
        h.set_domain(w["domain"], w["physics"])
        bad = xs.copy()
        bad[3], bad[4] = bad[4], bad[3]
        with pytest.raises(gpu.IcarionError) as ei:
            h.upload_field_terms((bad, ys, zs), w["terms"])
        assert ei.value.code == _abi.ERR_INVALID
        with pytest.raises(gpu.IcarionError) as ei:
            h.upload_field_terms((xs, ys, zs), w["terms"] * 2)
        assert ei.value.code == _abi.ERR_UNSUPPORTED
        bad_kind = [(w["terms"][0][0], 9, 0.0, 0.0, 0.0)]
        with pytest.raises(gpu.IcarionError) as ei:
            h.upload_field_terms((xs, ys, zs), bad_kind)
        assert ei.value.code == _abi.ERR_INVALID


def composite_waveforms(source=_abi.WAVEFORM_LIBRARY):
    """An axial DC ramp, an amplitude-modulated RF voltage and a ramped RF
    frequency for the composite field-array terms."""
    return {
        _abi.SLOT_DC_AXIAL: _abi.waveform(_abi.WAVEFORM_LINEAR, (200.0, 40.0, 0.0, 6e-8),
                                          source=source),
        _abi.SLOT_RF_VOLTAGE: _abi.waveform(_abi.WAVEFORM_SINUSOIDAL, (150.0, 60.0, 3e7, 0.2)),
        _abi.SLOT_RF_FREQUENCY: _abi.waveform(_abi.WAVEFORM_LINEAR, (1e6, 3e6, 1e-8, 5e-8)),
    }


def field_terms_waveform_workload(n, seed, **kw):
    w = W.field_terms_workload(n=n, seed=seed, **kw)
    E1 = w["terms"][1][0]
    w["terms"][1] = (E1, _abi.FIELD_SCALE_RF, 0.0, 0.0, 0.3)   # f from the frequency waveform
    w["waveforms"] = composite_waveforms()
    return w


@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS])
def test_field_terms_waveforms_fp64(oracle, integrator):
    """Composite field-array terms scaled by voltage waveforms (the host's scale
    table at every stage time) against the oracle evaluating
    CompositeFieldProvider's scales at every force call, ion by ion."""
    w = field_terms_waveform_workload(2048, 13, integrator=integrator)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 1500, t0=2e-9)
    oracle_run(oracle, w, o, 1500, t0=2e-9)
    assert compare_ions(g, o, rtol=1e-9).all()
    plain = dict(w)
    plain.pop("waveforms")
    c = w["ensemble"].copy()
    gpu_run(plain, c, 1500, t0=2e-9)
    assert compare_ions(g, c, rtol=1e-6).mean() < 0.5


def test_field_terms_waveforms_hss_and_fp32(oracle):
    w = field_terms_waveform_workload(2048, 14, collision=_abi.COLLISION_HSS)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg = gpu_run(w, g, 1500)
    so = oracle_run(oracle, w, o, 1500)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    w2 = field_terms_waveform_workload(2048, 15)
    g2, o2 = w2["ensemble"].copy(), w2["ensemble"].copy()
    gpu_run(w2, g2, 1500, precision=_abi.PRECISION_FP32_KAHAN)
    oracle_run(oracle, w2, o2, 1500)
    assert compare_ions(g2, o2, rtol=2e-5, pos_scale=1e-3, vel_scale=10.0).mean() >= 0.99


def test_field_terms_inline_waveform_scales_to_zero():
    """An inline waveform on a composite's DC scale counts as 0
    (CompositeFieldProvider::eval_voltage looks up library references only):
    bit for bit the run whose DC term has the constant scale 0."""
    w = field_terms_waveform_workload(1024, 16)
    w["waveforms"] = composite_waveforms(source=_abi.WAVEFORM_INLINE)
    z = dict(w)
    z["terms"] = list(w["terms"])
    z["terms"][0] = (w["terms"][0][0], _abi.FIELD_SCALE_CONSTANT, 0.0, 0.0, 0.0)
    z["waveforms"] = {k: v for k, v in composite_waveforms().items() if k != _abi.SLOT_DC_AXIAL}
    a, b = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, a, 600)
    gpu_run(z, b, 600)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)


@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS])
def test_hss_mixture_fp64_ion_by_ion(oracle, integrator):
    """HSS over a four-component gas mixture (one component outside the
    collisions), two ion species with per-gas cross sections."""
    w = W.ims_mixture_workload(n=2048, seed=31, integrator=integrator)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg = gpu_run(w, g, 2000)
    so = oracle_run(oracle, w, o, 2000)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    assert so.collisions > 0.3 * so.ion_steps


def test_hss_mixture_fp32_statistics(oracle):
    w = W.ims_mixture_workload(n=8192, seed=32)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    z0 = w["ensemble"].z.copy()
    sg = gpu_run(w, g, 4000, precision=_abi.PRECISION_FP32_KAHAN)
    so = oracle_run(oracle, w, o, 4000)
    assert sg.collisions == pytest.approx(so.collisions, rel=5e-3)
    for s in (0, 1):
        a = w["ensemble"].species == s
        dz_g, dz_o = (g.z - z0)[a], (o.z - z0)[a]
        se = math.sqrt(dz_g.var() / a.sum() + dz_o.var() / a.sum())
        assert abs(dz_g.mean() - dz_o.mean()) < 4 * se + 0.01 * abs(dz_o.mean())


def test_hss_mixture_controls():
    """Mixture set-up errors, and the empty mixture (no component taking part)
    has no collisions; clearing it returns to the single gas."""
    from icarion_amd import gpu
    w = W.ims_mixture_workload(n=512, seed=33)
    comps, sigma = w["mixture"]
    with gpu.Handle() as h:
        h.set_domain(w["domain"], w["physics"])
        h.set_species(w["species"])
        bad = sigma.copy()
        bad[1, 2] = 0.0
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_gas_mixture(comps, bad)
        assert ei.value.code == _abi.ERR_INVALID
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_gas_mixture(comps + comps[:1], np.concatenate([sigma, sigma[:, :1]], 1))
        assert ei.value.code == _abi.ERR_UNSUPPORTED
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_gas_mixture(comps, sigma[:1])
        assert ei.value.code == _abi.ERR_INVALID
        h.set_gas_mixture([(c[0], c[1], 0) for c in comps], sigma)
        h.upload_state(w["ensemble"].copy())
        st = h.step(0.0, w["dt"], 200)
        assert st.collisions == 0 and st.ion_steps == 200 * 512
        h.set_gas_mixture([], None)
        h.upload_state(w["ensemble"].copy())
        st = h.step(0.0, w["dt"], 200)
        assert st.collisions > 0


def quadrupole_waveforms():
    """RF amplitude modulated (sinusoidal), RF frequency ramped (linear), a DC
    quad pulse and an axial DC ramp (arbitrary, cubic)."""
    return {
        _abi.SLOT_RF_VOLTAGE: _abi.waveform(_abi.WAVEFORM_SINUSOIDAL, (450.0, 120.0, 2e5, 0.3)),
        _abi.SLOT_RF_FREQUENCY: _abi.waveform(_abi.WAVEFORM_LINEAR, (1.0e6, 1.3e6, 0.5e-6, 2.5e-6)),
        _abi.SLOT_DC_QUAD: _abi.waveform(_abi.WAVEFORM_PULSED, (0.0, 35.0, 1.0e-6, 0.7e-6)),
        _abi.SLOT_DC_AXIAL: _abi.waveform(_abi.WAVEFORM_ARBITRARY, (), 1, _abi.INTERP_CUBIC,
                                          [(0.0, 0.0), (0.8e-6, 4.0), (1.7e-6, -3.0), (3e-6, 6.0)]),
    }


@pytest.mark.parametrize("integrator", [_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS])
def test_waveforms_quadrupole_fp64_ion_by_ion(oracle, integrator):
    """Time-varying voltages (four waveform types) evaluated by the host into the
    step table, against the oracle evaluating them at every force call."""
    w = W.quadrupole_workload(n=2048, U=10.0, seed=41)
    w["physics"] = W.physics(integrator, _abi.COLLISION_NONE, seed=41)
    w["waveforms"] = quadrupole_waveforms()
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 3000)
    oracle_run(oracle, w, o, 3000)
    assert compare_ions(g, o, rtol=1e-9).all()
    # the waveforms matter: constant voltages give other trajectories
    plain = dict(w)
    plain.pop("waveforms")
    o2 = w["ensemble"].copy()
    oracle_run(oracle, plain, o2, 3000)
    assert compare_ions(g, o2, rtol=1e-6).mean() < 0.5


def test_waveforms_tims_ramp_hss(oracle):
    """A TIMS-like linear ramp of the axial voltage with HSS collisions (FP64
    ion by ion) and the FP32 path against the FP64 oracle."""
    w = W.ims_hss_workload(n=2048, seed=42)
    w["domain"].instrument = _abi.INSTRUMENT_TIMS
    w["waveforms"] = {_abi.SLOT_DC_AXIAL: _abi.waveform(
        _abi.WAVEFORM_LINEAR, (w["domain"].dc_axial_V, 0.2 * w["domain"].dc_axial_V, 0.0, 2e-7))}
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg = gpu_run(w, g, 2000)
    so = oracle_run(oracle, w, o, 2000)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99
    assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    w2 = W.quadrupole_workload(n=2048, U=10.0, seed=43)
    w2["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE, seed=43)
    w2["waveforms"] = quadrupole_waveforms()
    g2, o2 = w2["ensemble"].copy(), w2["ensemble"].copy()
    gpu_run(w2, g2, 3000, precision=_abi.PRECISION_FP32_KAHAN)
    oracle_run(oracle, w2, o2, 3000)
    assert compare_ions(g2, o2, rtol=2e-5, pos_scale=1e-3, vel_scale=10.0).mean() >= 0.99


def test_waveforms_constant_equals_constant_and_limits():
    from icarion_amd import gpu
    w = W.quadrupole_workload(n=512, U=10.0, seed=44)
    w["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE, seed=44)
    d = w["domain"]
    a, b = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, a, 500)
    wc = dict(w)
    wc["waveforms"] = {_abi.SLOT_RF_VOLTAGE: _abi.waveform(_abi.WAVEFORM_CONSTANT, (d.rf_voltage_V,)),
                       _abi.SLOT_DC_QUAD: _abi.waveform(_abi.WAVEFORM_CONSTANT, (d.dc_quad_V,))}
    gpu_run(wc, b, 500)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    with gpu.Handle() as h:
        h.set_domain(d, W.physics(_abi.INTEGRATOR_RK45, seed=44))
        h.set_species(w["species"])
        h.set_waveforms(quadrupole_waveforms())
        h.upload_state(w["ensemble"].copy())
        st = h.step(0.0, 1e-9, 5)   # RK45 evaluates the waveforms in the kernel
        assert st.ion_steps > 0
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_waveforms({_abi.SLOT_DC_AXIAL: _abi.waveform(
                _abi.WAVEFORM_ARBITRARY, (), 1, 0, [(1e-6, 1.0), (0.5e-6, 2.0)])})
        assert ei.value.code == _abi.ERR_INVALID
        h.set_waveforms(None)


def test_birth_times_fp64(oracle):
    w = W.ims_hss_workload(n=2048)
    e = w["ensemble"]
    rng = np.random.default_rng(1)
    e.born[:] = 0
    e.active[:] = 0
    e.birth_time_s = rng.uniform(0, 1e-7, e.n)
    g, o = e.copy(), e.copy()
    gpu_run(w, g, 2000)
    oracle_run(oracle, w, o, 2000)
    assert compare_ions(g, o, rtol=1e-9).mean() >= 0.99
    assert g.born.all()


def test_multi_call_equals_single_call():
    """Resident stepping split over calls (and the per-step step_batch shape)
    equals one call: same kernels, same RNG counters, same time accumulation."""
    from icarion_amd import gpu
    w = W.ims_hss_workload(n=1000)
    a, b, c = w["ensemble"].copy(), w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, a, 40)
    with gpu.Handle() as h:
        h.set_domain(w["domain"], w["physics"])
        h.set_species(w["species"])
        h.upload_state(b)
        t = 0.0
        for k in range(4):
            st = h.step(t, w["dt"], 10, 10 * k)
            t = st.t_end_s
        h.download_state(b)
        t = 0.0
        for k in range(40):
            st = h.step_batch(c, t, w["dt"], k)
            t = st.t_end_s
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))
        np.testing.assert_array_equal(getattr(a, f), getattr(c, f))


def test_launch_shape_invariance():
    w = W.quadrupole_workload(n=3000, seed=2)
    ref = w["ensemble"].copy()
    gpu_run(w, ref, 500)
    for spl, blk in ((7, 64), (128, 128), (500, 192)):
        e = w["ensemble"].copy()
        gpu_run(w, e, 500, steps_per_launch=spl, block=blk)
        np.testing.assert_array_equal(e.x, ref.x)
        np.testing.assert_array_equal(e.vz, ref.vz)


# ------------------------------------------------------------------ FP32 + Kahan
def test_fp32_kahan_collisionless_quadrupole(oracle):
    """FP32 + Kahan vs the FP64 oracle, collisionless RF quadrupole (Boris), 5 RF
    periods.  Error measured on the phase-space scale of the trap: position
    against r0, velocity against r0 * Omega.  Ions within 5 % of the Mathieu
    stability edge (q = 0.908), whose orbits amplify any perturbation, are held
    to the survival check only."""
    w = W.quadrupole_workload(n=4096, U=0.0, seed=4)
    w["physics"] = W.physics(_abi.INTEGRATOR_BORIS, _abi.COLLISION_NONE, seed=4)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 5000, precision=_abi.PRECISION_FP32_KAHAN)
    oracle_run(oracle, w, o, 5000)
    assert (g.active == o.active).mean() > 0.999
    r0, Om = 4e-3, 2 * math.pi * 1e6
    qM = 4 * 1.602176634e-19 * 500.0 / (w["ensemble"].mass_kg * r0 * r0 * Om * Om)
    sel = g.alive & o.alive & (np.abs(qM - 0.908) > 0.05 * 0.908)
    dp = np.sqrt((g.x - o.x) ** 2 + (g.y - o.y) ** 2 + (g.z - o.z) ** 2)[sel] / r0
    dv = np.sqrt((g.vx - o.vx) ** 2 + (g.vy - o.vy) ** 2 + (g.vz - o.vz) ** 2)[sel] / (r0 * Om)
    q = [np.quantile(dp, x) for x in (0.5, 0.99, 1.0)] + [np.quantile(dv, x) for x in (0.5, 0.99, 1.0)]
    msg = "pos err/r0 (median, p99, max) %.2e %.2e %.2e; vel err (median, p99, max) %.2e %.2e %.2e" % tuple(q)
    print(msg)
    assert q[0] < 1e-6 and q[1] < 1e-5 and q[3] < 1e-6 and q[4] < 1e-5, msg


def test_fp32_kahan_orbitrap_phase(oracle):
    w = W.orbitrap_workload(n=1024, seed=8)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 20000, precision=_abi.PRECISION_FP32_KAHAN)
    oracle_run(oracle, w, o, 20000)
    both = g.alive & o.alive
    dz = np.abs(g.z - o.z)[both]
    assert np.median(dz) < 2e-8   # 20 us of axial motion, ~1e-5 of the amplitude


def test_fp32_kahan_ims_hss_statistics(oracle):
    w = W.ims_hss_workload(n=20000, seed=123)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    z0 = w["ensemble"].z.copy()
    gpu_run(w, g, 5000, precision=_abi.PRECISION_FP32_KAHAN)
    oracle_run(oracle, w, o, 5000)
    dzg = (g.z - z0)[g.alive]
    dzo = (o.z - z0)[o.alive]
    assert dzg.mean() == pytest.approx(dzo.mean(), rel=0.01)
    assert dzg.std() == pytest.approx(dzo.std(), rel=0.03)
    assert stats.ks_2samp(dzg, dzo).pvalue > 1e-3
    assert g.n_collisions.mean() == pytest.approx(o.n_collisions.mean(), rel=0.005)


def test_fp32_kahan_quadrupole_transmission(oracle):
    w = W.quadrupole_workload(n=20000, U=0.0, seed=12)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 20000, precision=_abi.PRECISION_FP32_KAHAN)
    oracle_run(oracle, w, o, 20000)
    bins = np.geomspace(100, 2000, 9)
    m = w["ensemble"].mass_kg / 1.66053906660e-27
    for lo, hi in zip(bins[:-1], bins[1:]):
        sel = (m >= lo) & (m < hi)
        tg, to = g.alive[sel].mean(), o.alive[sel].mean()
        err = 4 * math.sqrt(max(to * (1 - to), 1e-4) / sel.sum()) + 0.01
        assert abs(tg - to) < err, (lo, hi, tg, to)


@pytest.mark.parametrize("precision", [_abi.PRECISION_FP64, _abi.PRECISION_FP32_KAHAN])
def test_orbitrap_radial_waveform(oracle, precision):
    """An Orbitrap whose DC radial voltage is a waveform (a 3.0 -> 3.5 kV ramp
    over the first microsecond) plus waveforms on slots the Orbitrap never
    reads: the field takes V(0) (AnalyticalFieldModel.cpp:198), so the run
    equals the constant-3.0 kV run bit for bit; in FP64 every ion also matches
    the oracle with the same waveforms to 1e-9."""
    import copy
    w = W.orbitrap_workload(n=1024, seed=32)
    w["waveforms"] = {
        _abi.SLOT_DC_RADIAL: _abi.waveform(_abi.WAVEFORM_LINEAR, (3000.0, 3500.0, 0.0, 1e-6)),
        _abi.SLOT_DC_AXIAL: _abi.waveform(_abi.WAVEFORM_CONSTANT, (900.0,)),
        _abi.SLOT_RF_VOLTAGE: _abi.waveform(_abi.WAVEFORM_SINUSOIDAL, (300.0, 0.0, 1e6, 0.0)),
    }
    c = dict(w, domain=copy.copy(w["domain"]))
    c["domain"].dc_radial_V = 3000.0
    c.pop("waveforms")
    g, k = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, g, 3000, precision=precision)
    gpu_run(c, k, 3000, precision=precision)
    for f in ("x", "y", "z", "vx", "vy", "vz", "active"):
        np.testing.assert_array_equal(getattr(g, f), getattr(k, f), err_msg=f)
    if precision == _abi.PRECISION_FP64:
        o = w["ensemble"].copy()
        oracle_run(oracle, w, o, 3000)
        assert compare_ions(g, o, rtol=1e-9).all()


# ------------------------------------------------------------------ API errors
def test_error_paths():
    from icarion_amd import gpu
    w = W.ims_hss_workload(n=16)
    with gpu.Handle() as h:
        with pytest.raises(gpu.IcarionError) as ei:
            h.step(0.0, 1e-10, 1)
        assert ei.value.code == _abi.ERR_STATE
        h.set_domain(w["domain"], w["physics"])
        h.set_species(w["species"])
        bad = w["ensemble"].copy()
        bad.species[3] = 5
        with pytest.raises(gpu.IcarionError) as ei:
            h.upload_state(bad)
        assert ei.value.code == _abi.ERR_INVALID
        ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_EHSS,
                       precision=_abi.PRECISION_FP32_KAHAN)
        h.set_domain(w["domain"], ph)
        h.upload_state(w["ensemble"])
        with pytest.raises(gpu.IcarionError) as ei:
            h.step(0.0, 1e-10, 1)
        assert ei.value.code == _abi.ERR_UNSUPPORTED
        with pytest.raises(gpu.IcarionError) as ei:
            h.step(0.0, -1.0, 1)
        assert ei.value.code == _abi.ERR_INVALID


def test_empty_ensemble():
    from icarion_amd import gpu
    from icarion_amd.ensemble import Ensemble
    w = W.ims_hss_workload(n=16)
    with gpu.Handle() as h:
        h.set_domain(w["domain"], w["physics"])
        h.set_species(w["species"])
        e = Ensemble(0)
        h.upload_state(e)
        st = h.step(0.0, 1e-10, 10)
        assert st.ion_steps == 0 and st.alive == 0
        assert st.t_end_s == pytest.approx(1e-9)
