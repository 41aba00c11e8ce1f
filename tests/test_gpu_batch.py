"""Batched clouds in one persistent launch (icarion_gpu_batch_*): each cloud's
result must be the one a handle of its own gives (same kernel bodies, same
random streams) - to 1e-9 relative, the per-step bar of the single-cloud
parity tests, since the batch kernel is a separate function the compiler
contracts on its own (differences seen: ~1e-13) - and bit for bit the same
whatever the tile / launch shape; the autotuner only picks among launch
shapes.  Single-cloud runs are pinned against the oracle
(test_gpu_parity.py), which carries that parity over to the batch."""
import os

import numpy as np
import pytest

from icarion_amd import _abi
from icarion_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "death_time_s", "arrival_time_s", "n_collisions",
          "active", "born")


def mixed_clouds(seed=3):
    """Twelve small clouds over every batch kind and both precisions."""
    rng = np.random.default_rng(seed)
    clouds = []
    for k in range(12):
        n = int(rng.integers(100, 3000))
        prec = _abi.PRECISION_FP64 if k % 2 == 0 else _abi.PRECISION_FP32_KAHAN
        kind = ("ims", "quadrupole", "orbitrap", "quadrupole_rk4_friction")[k % 4]
        if kind == "ims":
            w = W.ims_hss_workload(n=n, seed=100 + k, precision=prec)
            steps = int(rng.integers(50, 400))
        elif kind == "quadrupole":
            w = W.quadrupole_workload(n=n, U=10.0, seed=100 + k, precision=prec)
            steps = int(rng.integers(50, 400))
        elif kind == "orbitrap":
            w = W.orbitrap_workload(n=n, seed=100 + k, precision=prec)
            steps = int(rng.integers(200, 800))
        else:
            w = W.quadrupole_workload(n=n, U=5.0, seed=100 + k, precision=prec)
            w["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_FRICTION, ou=1,
                                     precision=prec, seed=100 + k)
            w["species"] = [W.species(K0=2.0)]
            steps = int(rng.integers(50, 400))
        clouds.append(dict(kind=kind, domain=w["domain"], physics=w["physics"],
                           species=w["species"], ensemble=w["ensemble"].copy(),
                           t0=1e-7 * k, dt=w["dt"], n_steps=steps, step_offset=7 * k))
    return clouds


def copy_clouds(clouds):
    out = []
    for c in clouds:
        d = dict(c)
        d["ensemble"] = c["ensemble"].copy()
        out.append(d)
    return out


def single_runs(clouds, params=None):
    from icarion_amd import gpu
    stats = []
    for c in clouds:
        with gpu.Handle() as h:
            h.set_domain(c["domain"], c["physics"])
            h.set_species(c["species"], c.get("atoms"))
            if (c["physics"].collision_model == _abi.COLLISION_EHSS
                    and c["physics"].ehss_mode == _abi.EHSS_SURROGATE):
                h.surrogate_set_params(params)
            if c.get("waveforms"):
                h.set_waveforms(c["waveforms"])
            if c.get("mixture"):
                h.set_gas_mixture(*c["mixture"])
            h.upload_state(c["ensemble"])
            stats.append(h.step(c["t0"], c["dt"], c["n_steps"], c["step_offset"]))
            h.download_state(c["ensemble"])
    return stats


def batch_runs(clouds, fp64=None, fp32=None, autotune=0, params=None):
    from icarion_amd import gpu
    with gpu.Handle() as h:
        if params is not None:
            h.surrogate_set_params(params)
        h.batch_upload(clouds)
        tuned = h.batch_autotune(autotune) if autotune else None
        stats, ms = h.batch_run(fp64, fp32)
        h.batch_download(clouds)
    return stats, ms, tuned


def assert_same(a, b):
    for ca, cb in zip(a, b):
        for f in FIELDS:
            np.testing.assert_array_equal(getattr(ca["ensemble"], f), getattr(cb["ensemble"], f),
                                          err_msg=f"{ca['kind']} {f}")


def close_fraction(a, b, rtol=1e-9):
    """Fraction of ions whose state agrees to rtol (positions on a 1 um scale,
    velocities on 1 m/s) with identical discrete state."""
    ok = np.ones(a.n, bool)
    for f, scale in (("x", 1e-6), ("y", 1e-6), ("z", 1e-6), ("vx", 1.0), ("vy", 1.0),
                     ("vz", 1.0)):
        u, v = getattr(a, f), getattr(b, f)
        ok &= np.abs(u - v) <= rtol * np.maximum(np.abs(v), scale)
    for f in ("active", "born", "n_collisions"):
        ok &= getattr(a, f) == getattr(b, f)
    return ok.mean()


def test_batch_equals_one_handle_per_cloud():
    ref = mixed_clouds()
    bat = copy_clouds(ref)
    s_ref = single_runs(ref)
    s_bat, ms, _ = batch_runs(bat)
    for ca, cb, a, b in zip(bat, ref, s_bat, s_ref):
        frac = close_fraction(ca["ensemble"], cb["ensemble"])
        stochastic = ca["physics"].collision_model != _abi.COLLISION_NONE
        assert frac >= (0.995 if stochastic else 1.0), (ca["kind"], frac)
        assert a.ion_steps == pytest.approx(b.ion_steps, rel=1e-3)
        assert a.collisions == pytest.approx(b.collisions, rel=5e-3, abs=2)
        assert a.t_end_s == b.t_end_s
        assert 0.0 < a.done_ms <= ms * 1.05 + 0.05


def rtol_of(cloud):
    """Batch against a handle of its own: 1e-9 in FP64 (contraction differs
    between the two compiled kernels at ~1e-13); 1e-5 in FP32 + Kahan, where
    the same differences are single-precision roundings."""
    return 1e-9 if cloud["physics"].precision == _abi.PRECISION_FP64 else 1e-5


def test_batch_clouds_differing_only_in_the_axial_field():
    """Clouds with the same engine times and RF inputs but different axial
    voltages or lengths: the step tables carry E_z, so each cloud must still
    get its own field (each equals a handle of its own).  Collisionless, so
    that the comparison is deterministic in both precisions."""
    import copy
    ref = []
    V0 = W.ims_hss_workload(n=8)["domain"].dc_axial_V
    for V, L in ((V0, 20e-3), (2.0 * V0, 20e-3), (V0, 28e-3)):
        for prec in (_abi.PRECISION_FP64, _abi.PRECISION_FP32_KAHAN):
            w = W.ims_hss_workload(n=800, seed=200, precision=prec)
            d = copy.copy(w["domain"])
            d.dc_axial_V, d.length_m = V, L
            ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE, precision=prec, seed=200)
            ref.append(dict(kind=f"ims V={V:.0f} L={L} prec={prec}", domain=d, physics=ph,
                            species=w["species"], ensemble=w["ensemble"].copy(), t0=0.0, dt=w["dt"],
                            n_steps=300, step_offset=0))
    bat = copy_clouds(ref)
    single_runs(ref)
    batch_runs(bat)
    fr = [close_fraction(ca["ensemble"], cb["ensemble"], rtol_of(ca)) for ca, cb in zip(bat, ref)]
    print("close fractions", [(c["kind"], f) for c, f in zip(bat, fr)])
    assert min(fr) == 1.0
    # and the fields differ (same start, same random streams): twice the
    # voltage drifts further, the longer region less far
    z = [c["ensemble"].z.mean() for c in bat]
    assert z[2] > z[0] > z[4] and z[3] > z[1] > z[5]


def waveform_clouds():
    """An IMS cloud with an axial ramp, a quadrupole with a modulated RF
    amplitude and a frequency ramp, an Orbitrap with a radial ramp (read at
    t = 0), and the same IMS cloud without waveforms on the same engine times."""
    ims = W.ims_hss_workload(n=1500, seed=301)
    quad = W.quadrupole_workload(n=1500, U=10.0, seed=302)
    orbi = W.orbitrap_workload(n=800, seed=303)
    wf_ims = {_abi.SLOT_DC_AXIAL: _abi.waveform(_abi.WAVEFORM_LINEAR, (10.0, 40.0, 0.0, 2e-8))}
    wf_quad = {_abi.SLOT_RF_VOLTAGE: _abi.waveform(_abi.WAVEFORM_SINUSOIDAL, (500.0, 50.0, 2e5, 0.0)),
               _abi.SLOT_RF_FREQUENCY: _abi.waveform(_abi.WAVEFORM_LINEAR, (1.0e6, 1.2e6, 0.0, 4e-7))}
    wf_orbi = {_abi.SLOT_DC_RADIAL: _abi.waveform(_abi.WAVEFORM_LINEAR, (3000.0, 3500.0, 0.0, 1e-6))}
    out = []
    for kind, w, wf, steps in (("ims_wf", ims, wf_ims, 300), ("quad_wf", quad, wf_quad, 400),
                               ("orbi_wf", orbi, wf_orbi, 600), ("ims_plain", ims, None, 300)):
        out.append(dict(kind=kind, domain=w["domain"], physics=w["physics"], species=w["species"],
                        ensemble=w["ensemble"].copy(), t0=0.0, dt=w["dt"], n_steps=steps, step_offset=0,
                        waveforms=wf))
    return out


def test_batch_waveform_clouds_equal_one_handle_per_cloud():
    ref = waveform_clouds()
    bat = copy_clouds(ref)
    single_runs(ref)
    batch_runs(bat)
    for ca, cb in zip(bat, ref):
        stochastic = ca["physics"].collision_model != _abi.COLLISION_NONE
        frac = close_fraction(ca["ensemble"], cb["ensemble"], rtol_of(ca))
        print(ca["kind"], frac)
        assert frac >= (0.995 if stochastic else 1.0), (ca["kind"], frac)
    # the ramped IMS cloud and the plain one share engine times but not fields
    a, b = bat[0]["ensemble"], bat[3]["ensemble"]
    assert (a.z != b.z).mean() > 0.9


def test_batch_gas_mixture_clouds_equal_one_handle_per_cloud():
    """HSS gas-mixture clouds (IMS in a three-gas mixture, RK4 and Boris, FP64
    and FP32 + Kahan) in one batch with plain clouds: each FP64 cloud equals a
    handle of its own with the same mixture ion by ion (the 0.995 bar of the
    stochastic clouds); each FP32 cloud agrees in its collision count and
    drift, since single-precision roundings of two compiled kernels flip a few
    of the ~500 collision decisions per ion.  The plain clouds run in the same
    launch through the kernel variant that holds the mixture kinds."""
    ref = []
    for k, (integ, prec) in enumerate(((_abi.INTEGRATOR_RK4, _abi.PRECISION_FP64),
                                       (_abi.INTEGRATOR_BORIS, _abi.PRECISION_FP64),
                                       (_abi.INTEGRATOR_RK4, _abi.PRECISION_FP32_KAHAN))):
        w = W.ims_mixture_workload(n=1500, seed=400 + k, precision=prec)
        w["physics"] = W.physics(integ, _abi.COLLISION_HSS, precision=prec, seed=400 + k)
        ref.append(dict(kind=f"mixture integ={integ} prec={prec}", domain=w["domain"], physics=w["physics"],
                        species=w["species"], ensemble=w["ensemble"].copy(), t0=0.0, dt=w["dt"],
                        n_steps=300, step_offset=0, mixture=w["mixture"]))
    ref += mixed_clouds(seed=12)[:4]
    bat = copy_clouds(ref)
    s_ref = single_runs(ref)
    s_bat, _, _ = batch_runs(bat)
    for ca, cb, a, b in zip(bat, ref, s_bat, s_ref):
        fp64 = ca["physics"].precision == _abi.PRECISION_FP64
        stochastic = ca["physics"].collision_model != _abi.COLLISION_NONE
        if fp64 or not stochastic:
            frac = close_fraction(ca["ensemble"], cb["ensemble"], rtol_of(ca))
            print(ca["kind"], frac)
            assert frac >= (0.995 if stochastic else 1.0), (ca["kind"], frac)
        else:
            assert a.collisions == pytest.approx(b.collisions, rel=5e-3), ca["kind"]
            za, zb = ca["ensemble"].z.mean(), cb["ensemble"].z.mean()
            z0 = ca["ensemble"].z.std()
            assert abs(za - zb) < 0.02 * z0, ca["kind"]
    assert s_bat[0].collisions > 100 * 1500


def test_batch_ehss_geometry_clouds_equal_one_handle_per_cloud():
    """Geometry EHSS clouds (FP64, RK4 and Boris; 1-, 5- and 23-atom clusters
    in each cloud's own atom table) in one batch with plain clouds: each equals
    a handle of its own with the same species and atoms, ion by ion at the
    0.995 bar of the stochastic clouds."""
    from test_gpu_parity import ehss_workload
    ref = []
    for k, integ in enumerate((_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS)):
        w = ehss_workload(n=1200, seed=500 + k)
        w["physics"] = W.physics(integ, _abi.COLLISION_EHSS, seed=500 + k)
        ref.append(dict(kind=f"ehss integ={integ}", domain=w["domain"], physics=w["physics"],
                        species=w["species"], atoms=w["atoms"], ensemble=w["ensemble"].copy(), t0=0.0,
                        dt=w["dt"], n_steps=400, step_offset=0))
    ref += mixed_clouds(seed=13)[:3]
    bat = copy_clouds(ref)
    s_ref = single_runs(ref)
    s_bat, _, _ = batch_runs(bat)
    for ca, cb, a, b in zip(bat, ref, s_bat, s_ref):
        stochastic = ca["physics"].collision_model != _abi.COLLISION_NONE
        if ca["physics"].precision == _abi.PRECISION_FP64 or not stochastic:
            frac = close_fraction(ca["ensemble"], cb["ensemble"], rtol_of(ca))
            print(ca["kind"], frac)
            assert frac >= (0.995 if stochastic else 1.0), (ca["kind"], frac)
        assert a.collisions == pytest.approx(b.collisions, rel=5e-3, abs=2), ca["kind"]
    assert s_bat[0].collisions > 1200 * 50


@pytest.mark.parametrize("shape", [(64, 1), (128, 8), (256, 16)])
def test_batch_result_does_not_depend_on_the_launch_shape(shape):
    base = mixed_clouds(seed=9)
    other = copy_clouds(base)
    batch_runs(base, fp64=(256, 4), fp32=(256, 4))
    batch_runs(other, fp64=shape, fp32=shape)
    assert_same(other, base)


def test_batch_autotune_picks_a_candidate_and_keeps_results():
    base = mixed_clouds(seed=11)
    tuned = copy_clouds(base)
    batch_runs(base)
    _, _, (t64, t32, r64, r32) = batch_runs(tuned, autotune=100)
    for b, g in (t64, t32):
        assert b in (64, 128, 256) and g in (1, 2, 4, 8, 16) and b * g <= 2048
    assert r64 > 0 and r32 > 0
    assert_same(tuned, base)


def test_batch_rejects_what_it_does_not_cover():
    from icarion_amd import gpu
    w = W.ims_hss_workload(n=64)
    bad = [dict(domain=w["domain"], physics=W.physics(_abi.INTEGRATOR_RK45), species=w["species"],
                ensemble=w["ensemble"], dt=1e-10, n_steps=10)]
    with gpu.Handle() as h:
        with pytest.raises(gpu.IcarionError) as ei:
            h.batch_upload(bad)
        assert ei.value.code == _abi.ERR_UNSUPPORTED
        bad[0]["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_EHSS,
                                      precision=_abi.PRECISION_FP32_KAHAN)
        with pytest.raises(gpu.IcarionError) as ei:
            h.batch_upload(bad)
        assert ei.value.code == _abi.ERR_UNSUPPORTED
        # a field-array domain needs the cloud's array; with a gas mixture it is
        # not covered
        wf = W.field_array_workload(n=64, grid=(8, 8, 8), seed=1)
        fa = dict(domain=wf["domain"], physics=wf["physics"], species=wf["species"],
                  ensemble=wf["ensemble"], dt=wf["dt"], n_steps=10)
        with pytest.raises(gpu.IcarionError) as ei:
            h.batch_upload([fa])
        assert ei.value.code == _abi.ERR_INVALID
        fa["field"] = wf["field"]
        fa["mixture"] = ([(1e22, 4.65e-26, True)], np.full((len(wf["species"]), 1), 1e-18))
        with pytest.raises(gpu.IcarionError) as ei:
            h.batch_upload([fa])
        assert ei.value.code == _abi.ERR_UNSUPPORTED
        with pytest.raises(gpu.IcarionError) as ei:
            h.batch_run()
        assert ei.value.code == _abi.ERR_STATE


SURROGATE_WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                 "icarion_amd", "data", "surrogate_he_20td_clusters.npy")


def surrogate_clouds(seed=4):
    """Three EHSS-surrogate clouds (RK4 and Boris; 16 random clusters from the
    device generator) next to one ordinary cloud of each precision, so that all
    three classes of the batch run in one launch."""
    from icarion_amd import gpu
    rng = np.random.default_rng(seed)
    counts = W.ehss_cluster_counts(16, seed)
    w0 = W.ehss_surrogate_workload(n=16, seed=seed)
    with gpu.Handle() as h:
        h.set_domain(w0["domain"], w0["physics"])
        _, sigma, feat = h.synth_clusters(counts, seed=seed)
    clouds = []
    for k, integ in enumerate((_abi.INTEGRATOR_RK4, _abi.INTEGRATOR_BORIS, _abi.INTEGRATOR_RK4)):
        n = int(rng.integers(100, 2000))
        w = W.ehss_surrogate_cloud(n, sigma, feat, counts, seed=50 + k, integrator=integ)
        clouds.append(dict(kind=f"ehss_surrogate_{k}", domain=w["domain"], physics=w["physics"],
                           species=w["species"], ensemble=w["ensemble"].copy(),
                           ensemble0=w["ensemble"].copy(), t0=1e-9 * k,
                           dt=w["dt"], n_steps=int(rng.integers(200, 600)), step_offset=11 * k))
    return clouds + mixed_clouds(seed)[:2]


def test_batch_surrogate_clouds_equal_one_handle_per_cloud():
    """The FP32 surrogate clouds run the single-cloud kernel's per-ion body,
    but the compiler fuses multiply-adds differently in the two kernels, so
    their FP32 trajectories differ by rounding, which the stochastic
    collisions amplify: they are compared as ensembles (steps, collision
    counts, mean drift); the ordinary clouds of the same batch ion by ion."""
    params = np.load(SURROGATE_WEIGHTS)
    ref = surrogate_clouds()
    bat = copy_clouds(ref)
    s_ref = single_runs(ref, params)
    s_bat, ms, _ = batch_runs(bat, params=params)
    for ca, cb, a, b in zip(bat, ref, s_bat, s_ref):
        assert a.t_end_s == b.t_end_s
        assert 0.0 < a.done_ms <= ms * 1.05 + 0.05
        if ca["kind"].startswith("ehss_surrogate"):
            g, o, e0 = ca["ensemble"], cb["ensemble"], ca["ensemble0"]
            assert a.ion_steps == b.ion_steps
            assert a.collisions == pytest.approx(b.collisions, rel=0.02)
            dz_g, dz_o = g.z - e0.z, o.z - e0.z
            se = np.sqrt(dz_g.var() / g.n + dz_o.var() / o.n)
            assert abs(dz_g.mean() - dz_o.mean()) < 4 * se, ca["kind"]
        else:
            frac = close_fraction(ca["ensemble"], cb["ensemble"])
            assert frac >= 0.995, (ca["kind"], frac)
            assert a.ion_steps == pytest.approx(b.ion_steps, rel=1e-3)
            assert a.collisions == pytest.approx(b.collisions, rel=5e-3, abs=2)
    # the surrogate clouds collide (about 0.6 events per ion-step at 1 atm He)
    for c, a in zip(bat[:3], s_bat[:3]):
        assert a.collisions > 0.3 * a.ion_steps, c["kind"]


def test_batch_surrogate_results_do_not_depend_on_the_other_classes_shapes():
    params = np.load(SURROGATE_WEIGHTS)
    base = surrogate_clouds(seed=6)
    other = copy_clouds(base)
    batch_runs(base, fp64=(256, 4), fp32=(256, 4), params=params)
    batch_runs(other, fp64=(64, 2), fp32=(128, 16), params=params)
    assert_same(other, base)


def test_batch_surrogate_needs_parameters():
    from icarion_amd import gpu
    clouds = surrogate_clouds(seed=7)[:1]
    with gpu.Handle() as h:
        h.batch_upload(clouds)
        with pytest.raises(gpu.IcarionError) as ei:
            h.batch_run()
        assert ei.value.code == _abi.ERR_STATE
        bad = dict(clouds[0])
        bad["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_EHSS,
                                   precision=_abi.PRECISION_FP64, ehss_mode=_abi.EHSS_SURROGATE)
        with pytest.raises(gpu.IcarionError) as ei:
            h.batch_upload([bad])
        assert ei.value.code == _abi.ERR_UNSUPPORTED
