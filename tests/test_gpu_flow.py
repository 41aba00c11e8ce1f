"""The AxialParabolic gas flow on the device against the oracle: the gas
velocity at each ion's position (EnvironmentConfig::gas_velocity_at,
EnvironmentConfig.h:96-129) in the damping force at every force evaluation
and for the neutrals of every collision substep at the pushed position
(SimulationEngineCollisionRuntime.cpp:122-130).  The oracle's restatement is
pinned on the CPU (tests/test_oracle_flow.py), against the reference's own
tests of the profile among others.

FP64 runs are compared ion by ion (>= 99 % of ions to 1e-9), FP32 + Kahan by
statistics; the Friction relaxation to the local flow is checked against its
closed form in both precisions."""
import math

import numpy as np
import pytest

from icarion_amd import _abi
from icarion_amd import workloads as W
from test_gpu_domains import gpu_multi, oracle_multi
from test_gpu_parity import compare_ions, gpu_run, oracle_run
from test_oracle_flow import radial_cloud, rotated

pytestmark = pytest.mark.gpu

FLOW = (_abi.FLOW_AXIAL_PARABOLIC, 400.0)


@pytest.mark.parametrize("case", ["hss", "hss_substeps", "friction_ou", "langevin_boris", "mixture"])
def test_parabolic_flow_fp64_ion_by_ion(oracle, case):
    if case == "mixture":
        w = W.ims_mixture_workload(n=2048, seed=111)
    else:
        w = W.ims_hss_workload(n=2048, seed=112)
        if case == "hss_substeps":
            w["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, seed=112, substeps=3)
        elif case == "friction_ou":
            w["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_FRICTION, ou=1, seed=112)
            w["species"] = [W.species(K0=2.1, ccs=100e-20)]
        elif case == "langevin_boris":
            w["physics"] = W.physics(_abi.INTEGRATOR_BORIS, _abi.COLLISION_LANGEVIN, ou=1, seed=112)
    # the same drift tube rotated (local z along global x), the cloud carried
    # with it: p = origin + R_l2g (p_old - origin_old), v = R_l2g v_old
    d = w["domain"]
    o_old = np.array(d.origin_m)
    rotated(d, origin=(-5e-3, 1e-3, 0.0))
    e = w["ensemble"]
    L = np.array(d.rot_l2g).reshape(3, 3)
    P = L @ (np.stack([e.x, e.y, e.z]) - o_old[:, None]) + np.array(d.origin_m)[:, None]
    V = L @ np.stack([e.vx, e.vy, e.vz])
    e.x[:], e.y[:], e.z[:] = P
    e.vx[:], e.vy[:], e.vz[:] = V
    w["flow"] = FLOW
    print(f"flow FP64 {case}", flush=True)
    g, o = e.copy(), e.copy()
    sg = gpu_run(w, g, 1200)
    so = oracle_run(oracle, w, o, 1200)
    ok = compare_ions(g, o, rtol=1e-9)
    print(f"  match {ok.mean():.4f} collisions {sg.collisions}/{so.collisions}", flush=True)
    assert ok.mean() >= 0.99
    assert sg.ion_steps == so.ion_steps
    if case in ("hss", "hss_substeps", "mixture"):
        assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    # and not the run without the flow
    w0 = dict(w)
    w0["flow"] = None
    g0 = e.copy()
    gpu_run(w0, g0, 1200)
    assert abs(g.vx.mean() - g0.vx.mean()) > 50.0


@pytest.mark.parametrize("precision", [_abi.PRECISION_FP64, _abi.PRECISION_FP32_KAHAN])
def test_friction_relaxes_to_the_local_flow(precision):
    d = rotated(W.ims_domain(length=20e-3, radius=10e-3, axial_V=0.0, gas="N2", pressure_Pa=101325.0))
    R, u_max = d.radius_m, 500.0
    e, r = radial_cloud(d, (0.0, 0.3 * R, 0.6 * R, 0.9 * R), 16, 100.0)
    w = dict(domain=d, physics=W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_FRICTION, precision=precision),
             species=[W.species(K0=2.0)], dt=1e-10, flow=(_abi.FLOW_AXIAL_PARABOLIC, u_max))
    gpu_run(w, e, 2000)
    want = u_max * (1.0 - (r / R) ** 2)
    tol = 1e-9 if precision == _abi.PRECISION_FP64 else 1e-5
    np.testing.assert_allclose(e.vx, want, rtol=tol, atol=tol * u_max)
    np.testing.assert_allclose(e.vz, 0.0, atol=tol * u_max)


def test_parabolic_flow_fp32_statistics(oracle):
    d = rotated(W.ims_domain(length=20e-3, radius=10e-3, axial_V=0.0, gas="N2", pressure_Pa=101325.0))
    R, u_max = d.radius_m, 500.0
    per = 4096
    e, r = radial_cloud(d, (0.0, 0.5 * R), per, 28.0, seed=4)
    w = dict(domain=d, physics=W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, seed=4),
             species=[W.species(sigma_hss=50e-20, ccs=50e-20)], dt=1e-10, flow=(_abi.FLOW_AXIAL_PARABOLIC, u_max))
    g, o = e.copy(), e.copy()
    sg = gpu_run(w, g, 1500, precision=_abi.PRECISION_FP32_KAHAN)
    so = oracle_run(oracle, w, o, 1500)
    assert sg.collisions == pytest.approx(so.collisions, rel=0.01)
    for k in range(2):
        sel = slice(k * per, (k + 1) * per)
        se = math.hypot(g.vx[sel].std() / math.sqrt(per), o.vx[sel].std() / math.sqrt(per))
        print(f"  radius {k}: {g.vx[sel].mean():.2f} vs {o.vx[sel].mean():.2f} (se {se:.2f})", flush=True)
        assert abs(g.vx[sel].mean() - o.vx[sel].mean()) < 4 * se


@pytest.mark.parametrize("precision,collision", [(_abi.PRECISION_FP64, _abi.COLLISION_LANGEVIN),
                                                 (_abi.PRECISION_FP64, _abi.COLLISION_HSS),
                                                 (_abi.PRECISION_FP32_KAHAN, _abi.COLLISION_LANGEVIN)])
def test_several_domains_with_their_own_flow(oracle, precision, collision):
    """Flows on the two IMS regions of the three-domain instrument (none in the
    quadrupole): FP64 ion by ion with domain indices, FP32 by per-domain
    survival and mean axial velocity."""
    n = 2048 if precision == _abi.PRECISION_FP64 else 16384
    w = W.multi_domain_workload(n=n, seed=117, collision=collision)
    w["physics"] = W.physics(_abi.INTEGRATOR_RK4, collision, ou=1 if collision == _abi.COLLISION_LANGEVIN else 0,
                             precision=precision, seed=117)
    w["flows"] = [(_abi.FLOW_AXIAL_PARABOLIC, 800.0), None, (_abi.FLOW_AXIAL_PARABOLIC, -300.0)]
    print(f"flow domains precision={precision} coll={collision}", flush=True)
    g, o = w["ensemble"].copy(), w["ensemble"].copy()
    sg, dg = gpu_multi(w, g, 4000)
    so, do = oracle_multi(oracle, w, o, 4000)
    if precision == _abi.PRECISION_FP64:
        ok = compare_ions(g, o, rtol=1e-9) & (dg == do)
        print(f"  match {ok.mean():.4f}", flush=True)
        assert ok.mean() >= 0.99
        if collision == _abi.COLLISION_HSS:
            assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    else:
        for d in (None, 0, 1, 2):
            sel_g = g.alive if d is None else g.alive & (dg == d)
            sel_o = o.alive if d is None else o.alive & (do == d)
            fg, fo = sel_g.mean(), sel_o.mean()
            err = 4 * np.sqrt(max(fo * (1 - fo), 1e-4) / g.n) + 0.005
            assert abs(fg - fo) < err, (d, fg, fo)
    # the flows change the run
    w0 = dict(w)
    w0["flows"] = None
    g0 = w["ensemble"].copy()
    gpu_multi(w0, g0, 4000)
    assert np.abs(g.vz - g0.vz).max() > 10.0


@pytest.mark.parametrize("collision", [_abi.COLLISION_HSS, _abi.COLLISION_FRICTION])
def test_rk45_parabolic_flow(oracle, collision):
    """RK45: the damping force at each stage position and the neutrals at the
    pushed position take the local flow, in the rotated drift tube.  HSS is
    compared ion by ion with the oracle's RK45 loop; Friction + OU at 1 atm
    holds the adaptive step at the stiff damping's stability edge, where
    rounding flips accept / reject decisions, so it is held to the oracle's
    mean velocity along the flow and its spread."""
    from test_gpu_rk45 import gpu_rk45, match, oracle_rk45
    d = rotated(W.ims_domain(length=20e-3, radius=10e-3, axial_V=50.0, gas="N2", pressure_Pa=101325.0))
    R = d.radius_m
    e, _ = radial_cloud(d, (0.0, 0.3 * R, 0.6 * R), 512, 40.0, seed=8)
    ph = W.physics(_abi.INTEGRATOR_RK45, collision, ou=1, seed=8)
    w = dict(domain=d, species=[W.species(sigma_hss=60e-20, ccs=60e-20, K0=2.5)], dt=1e-10, flow=FLOW)
    print(f"flow RK45 coll={collision}", flush=True)
    g, o = e.copy(), e.copy()
    sg = gpu_rk45(w, ph, g, 80, w["dt"])
    so = oracle_rk45(oracle, w, ph, o, 80, w["dt"])
    if collision == _abi.COLLISION_HSS:
        assert match(g, o, rtol=1e-9, dt_rtol=1e-9).mean() >= 0.99
        assert sg.collisions == pytest.approx(so.collisions, rel=3e-3)
    else:
        se = math.hypot(g.vx.std() / math.sqrt(g.n), o.vx.std() / math.sqrt(o.n))
        print(f"  vx {g.vx.mean():.3f} vs {o.vx.mean():.3f} (se {se:.3f})", flush=True)
        assert abs(g.vx.mean() - o.vx.mean()) < 4 * se
        assert g.vx.std() == pytest.approx(o.vx.std(), rel=0.1)
    g0 = e.copy()
    w0 = dict(w)
    w0["flow"] = None
    gpu_rk45(w0, ph, g0, 80, w["dt"])
    assert abs(g.vx.mean() - g0.vx.mean()) > 20.0


def test_flow_arguments_and_unsupported():
    from icarion_amd import gpu
    w = W.ims_hss_workload(n=64, seed=113)
    with gpu.Handle() as h:
        with pytest.raises(gpu.IcarionError) as ei:
            h.set_gas_flow(7, 1.0)
        assert ei.value.code == _abi.ERR_INVALID
        with pytest.raises(gpu.IcarionError):
            h.set_gas_flow(_abi.FLOW_AXIAL_PARABOLIC, math.inf)
    # clearing restores the constant run bit for bit
    a, b = w["ensemble"].copy(), w["ensemble"].copy()
    gpu_run(w, a, 300)
    with gpu.Handle() as h:
        h.set_domain(w["domain"], w["physics"])
        h.set_species(w["species"])
        h.set_gas_flow(_abi.FLOW_AXIAL_PARABOLIC, 300.0)
        h.set_gas_flow(_abi.FLOW_CONSTANT)
        h.upload_state(b)
        h.step(0.0, w["dt"], 300)
        h.download_state(b)
    for f in ("x", "y", "z", "vx", "vy", "vz", "n_collisions"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    # the geometry EHSS kernel and the surrogate keep a constant flow
    cases = []
    from test_gpu_parity import ehss_workload
    we = ehss_workload(n=64, seed=114)
    cases.append((we, we["physics"], None))
    ws = W.ehss_surrogate_workload(n=256, seed=115)
    cases.append((ws, ws["physics"], "surrogate"))
    for wk, ph, kind in cases:
        with gpu.Handle() as h:
            h.set_domain(wk["domain"], ph)
            if kind == "surrogate":
                h.synth_clusters(wk["counts"], seed=115)
                h.surrogate_set_params(np.zeros(_abi.SURROGATE_NPARAMS, np.float32))
            else:
                h.set_species(wk["species"], wk.get("atoms"))
            h.set_gas_flow(*FLOW)
            h.upload_state(wk["ensemble"].copy())
            with pytest.raises(gpu.IcarionError) as ei:
                h.step(0.0, wk["dt"], 4)
            assert ei.value.code == _abi.ERR_UNSUPPORTED
    wm = W.multi_domain_workload(n=64, seed=116)
    with gpu.Handle() as h:
        h.set_domains(wm["domains"], wm["physics"], wm.get("apertures"))
        h.set_species(wm["species"])
        h.set_gas_flow(*FLOW)
        h.upload_state(wm["ensemble"].copy())
        with pytest.raises(gpu.IcarionError) as ei:
            h.step(0.0, wm["dt"], 4)
        assert ei.value.code == _abi.ERR_UNSUPPORTED
