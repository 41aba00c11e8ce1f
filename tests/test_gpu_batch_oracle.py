"""Batched clouds (one persistent launch, icarion_gpu_batch_*) against the FP64
CPU oracle directly, cloud by cloud: the same engine loop and Philox streams
(keyed by each cloud's seed, ion index and step), so FP64 clouds are compared
ion by ion and FP32 + Kahan clouds by path agreement and statistics.

Bars:
  * FP64, collisionless: every ion to 1e-9 relative;
  * FP64, stochastic: >= 99 % of ions to 1e-9 with equal collision counts (a
    collision test within rounding of its threshold can send one ion on
    another path);
  * FP32 + Kahan, collisionless (Orbitrap): the same fate for >= 99.9 % of
    ions, position within 2e-5 of the 9 mm orbit and velocity within 2e-5 of
    the speed (vector norms);
  * FP32 + Kahan, stochastic: the FP32 collision and kick code draws its
    normals from both Box-Muller branches, so its random numbers map to other
    events than the FP64 code's and the paths are independent samples:
    survival fraction, mean drift and collisions per ion within 4 standard
    errors of the difference (plus 1 % of the value), and a KS test on the
    final z (p > 1e-3).
The long Orbitrap test runs the mixed workload's full 1e6 steps and states the
FP32 + Kahan error against the FP64 oracle (position and orbital phase)."""
import math

import numpy as np
import pytest
from scipy import stats

from icarion_amd import _abi
from icarion_amd import workloads as W

pytestmark = pytest.mark.gpu


def mixed_clouds(seed=17):
    """Sixteen clouds over the batch kinds and both precisions: IMS (HSS, RK4),
    quadrupole (Langevin + OU, Boris), Orbitrap (collisionless RK4) and a
    friction + OU quadrupole (RK4)."""
    rng = np.random.default_rng(seed)
    clouds = []
    for k in range(16):
        n = int(rng.integers(200, 4000))
        prec = _abi.PRECISION_FP64 if (k // 4) % 2 == 0 else _abi.PRECISION_FP32_KAHAN
        kind = ("ims", "quadrupole", "orbitrap", "quadrupole_rk4_friction")[k % 4]
        s = 300 + k
        if kind == "ims":
            w = W.ims_hss_workload(n=n, seed=s, precision=prec)
            steps = int(rng.integers(300, 1500))
        elif kind == "quadrupole":
            w = W.quadrupole_workload(n=n, U=10.0, seed=s, precision=prec)
            steps = int(rng.integers(300, 1500))
        elif kind == "orbitrap":
            w = W.orbitrap_workload(n=n, seed=s, precision=prec)
            steps = int(rng.integers(1000, 4000))
        else:
            w = W.quadrupole_workload(n=n, U=5.0, seed=s, precision=prec)
            w["physics"] = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_FRICTION, ou=1,
                                     precision=prec, seed=s)
            w["species"] = [W.species(K0=2.0)]
            steps = int(rng.integers(300, 1500))
        clouds.append(dict(kind=kind, domain=w["domain"], physics=w["physics"],
                           species=w["species"], ensemble=w["ensemble"].copy(),
                           t0=1e-7 * k, dt=w["dt"], n_steps=steps, step_offset=11 * k))
    return clouds


def batch_run(clouds):
    from icarion_amd import gpu
    with gpu.Handle() as h:
        h.batch_upload(clouds)
        stats, _ = h.batch_run()
        h.batch_download(clouds)
    return stats


def oracle_cloud(oracle, c, ens):
    """The oracle's FP64 engine loop over one cloud's initial state."""
    ph = c["physics"]
    ph64 = W.physics(ph.integrator, ph.collision_model, ph.ou_thermal_kicks, _abi.PRECISION_FP64,
                     ph.rng_seed, ph.ehss_mode)
    field = oracle.Field(*c["field"]) if c.get("field") is not None else None
    return oracle.run(c["domain"], ph64, c["species"], ens, c["t0"], c["dt"], c["n_steps"],
                      step_offset=c["step_offset"], field=field)


def agree(g, o, rtol, pos_scale, vel_scale):
    ok = g.active == o.active
    for f in ("x", "y", "z"):
        a, b = getattr(g, f), getattr(o, f)
        ok &= np.abs(a - b) <= rtol * np.maximum(np.abs(b), pos_scale)
    for f in ("vx", "vy", "vz"):
        a, b = getattr(g, f), getattr(o, f)
        ok &= np.abs(a - b) <= rtol * np.maximum(np.abs(b), vel_scale)
    return ok


def close_mean(a, b, what):
    """mean(a) and mean(b) within 4 standard errors of their difference plus
    1 % of mean(b)."""
    ma, mb = a.mean(), b.mean()
    se = math.hypot(a.std(ddof=1) / math.sqrt(a.size), b.std(ddof=1) / math.sqrt(b.size))
    print(f"{what}: {ma:.5e} vs {mb:.5e} (sigma of the difference {se:.2e})", flush=True)
    assert abs(ma - mb) <= 4 * se + 0.01 * abs(mb), what


def test_mixed_batch_against_oracle(oracle):
    clouds = mixed_clouds()
    init = [c["ensemble"].copy() for c in clouds]
    cloud_stats = batch_run(clouds)
    for c, e0, st in zip(clouds, init, cloud_stats):
        o = e0.copy()
        so = oracle_cloud(oracle, c, o)
        g = c["ensemble"]
        fp64 = c["physics"].precision == _abi.PRECISION_FP64
        stochastic = c["physics"].collision_model != _abi.COLLISION_NONE
        tag = f"{c['kind']} {'fp64' if fp64 else 'fp32'} n={g.n} steps={c['n_steps']}"
        assert st.ion_steps == pytest.approx(so.ion_steps, rel=0.01 if stochastic else 0), tag
        if fp64:
            ok = agree(g, o, 1e-9, 1e-6, 1.0) & (g.n_collisions == o.n_collisions)
            print(f"{tag}: {ok.mean():.4f} of ions equal the oracle to 1e-9", flush=True)
            assert ok.mean() >= (0.99 if stochastic else 1.0), tag
            if not stochastic:
                np.testing.assert_array_equal(g.death_time_s, o.death_time_s)
        elif not stochastic:
            # vector errors: position against the 9 mm orbit, velocity against the speed
            both = g.alive & o.alive
            dp = (np.linalg.norm(g.pos - o.pos, axis=1) / 9e-3)[both]
            dv = (np.linalg.norm(g.vel - o.vel, axis=1) / np.linalg.norm(o.vel, axis=1))[both]
            print(f"{tag}: position/9mm max {dp.max():.2e}, velocity/speed max {dv.max():.2e} "
                  f"against the FP64 oracle", flush=True)
            assert (g.active == o.active).mean() >= 0.999, tag
            assert dp.max() < 2e-5 and dv.max() < 2e-5, tag
        else:
            z0 = e0.z
            sg, so_ = g.alive.mean(), o.alive.mean()
            s_err = 4 * math.sqrt(2 * max(so_ * (1 - so_), 1.0 / g.n) / g.n) + 0.01
            print(f"{tag}: survival {sg:.4f} vs {so_:.4f}", flush=True)
            assert abs(sg - so_) <= s_err, tag
            close_mean(((g.z - z0)[g.alive]), ((o.z - z0)[o.alive]), tag + " drift")
            close_mean(g.n_collisions.astype(float), o.n_collisions.astype(float),
                       tag + " collisions")
            assert stats.ks_2samp(g.z[g.alive], o.z[o.alive]).pvalue > 1e-3, tag


def test_batch_orbitrap_1e6_steps_against_fp64_oracle(oracle):
    """One Orbitrap cloud of 192 ions at the batched workload's 1e6 steps
    (1 ms, dt 1 ns), in FP64 and in FP32 + Kahan (FP64 time and RF phase) as two
    clouds of one launch, against the FP64 oracle.  Stated errors (position
    against 9 mm, velocity against the ion's speed, azimuth in rad): FP64
    within 2e-9 at the maximum (measured 9.7e-10 / 9.0e-10 / 9.7e-10); FP32 +
    Kahan within 1e-3 at the median and 1e-2 at the maximum (measured median
    4.7e-4 / 4.4e-4 / 4.7e-4, maximum 4.0e-3 / 3.6e-3 / 3.9e-3).  The FP32
    phase error grows linearly with the number of orbits, about 1.5e-7 of the
    orbital phase: the rounding of the force constants to float is a
    systematic frequency offset that compensated summation of the state does
    not remove."""
    w = W.orbitrap_workload(n=192, seed=41)
    clouds = []
    for prec in (_abi.PRECISION_FP64, _abi.PRECISION_FP32_KAHAN):
        ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_NONE, precision=prec, seed=41)
        clouds.append(dict(kind="orbitrap", domain=w["domain"], physics=ph, species=w["species"],
                           ensemble=w["ensemble"].copy(), t0=0.0, dt=w["dt"],
                           n_steps=1_000_000, step_offset=0))
    print("orbitrap 192 ions x 1e6 steps: batch (FP64 + FP32 clouds), then the oracle",
          flush=True)
    cloud_stats = batch_run(clouds)
    o = w["ensemble"].copy()
    so = oracle.run(w["domain"], w["physics"], w["species"], o, 0.0, w["dt"], 1_000_000)
    assert so.alive == 192
    r_scale = 9e-3
    v_scale = np.sqrt(o.vx ** 2 + o.vy ** 2 + o.vz ** 2)
    res = {}
    for c, st in zip(clouds, cloud_stats):
        g = c["ensemble"]
        assert st.ion_steps == so.ion_steps and st.alive == 192
        dp = np.sqrt((g.x - o.x) ** 2 + (g.y - o.y) ** 2 + (g.z - o.z) ** 2) / r_scale
        dv = np.sqrt((g.vx - o.vx) ** 2 + (g.vy - o.vy) ** 2 + (g.vz - o.vz) ** 2) / v_scale
        dphi = np.abs(np.angle(np.exp(1j * (np.arctan2(g.y, g.x) - np.arctan2(o.y, o.x)))))
        name = "fp64" if c["physics"].precision == _abi.PRECISION_FP64 else "fp32"
        res[name] = (np.median(dp), dp.max(), np.median(dv), dv.max(), np.median(dphi),
                     dphi.max())
        print(f"orbitrap 1e6 steps {name} vs FP64 oracle: position/9mm median {res[name][0]:.2e} "
              f"max {res[name][1]:.2e}; velocity median {res[name][2]:.2e} max "
              f"{res[name][3]:.2e}; azimuth median {res[name][4]:.2e} max {res[name][5]:.2e} rad",
              flush=True)
    assert res["fp64"][1] < 2e-9 and res["fp64"][3] < 2e-9 and res["fp64"][5] < 2e-9
    assert res["fp32"][0] < 1e-3 and res["fp32"][1] < 1e-2
    assert res["fp32"][2] < 1e-3 and res["fp32"][3] < 1e-2
    assert res["fp32"][4] < 1e-3 and res["fp32"][5] < 1e-2


def field_array_clouds(seed=23):
    """Uniform-grid field-array clouds (IMS in N2 at 300 Pa on a 24 x 20 x 40
    grid, FP32 nodes) in both precisions - RK4 + HSS, Boris collisionless, RK4
    collisionless - two of them sharing one array, next to an analytic IMS
    cloud and a quadrupole cloud, in one launch."""
    rng = np.random.default_rng(seed)
    wa = W.field_array_workload(n=3000, grid=(24, 20, 40), seed=2)
    wb = W.field_array_workload(n=2500, grid=(24, 20, 40), seed=5)
    plan = [("fa_rk4_hss", wa, _abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, _abi.PRECISION_FP64),
            ("fa_boris", wa, _abi.INTEGRATOR_BORIS, _abi.COLLISION_NONE, _abi.PRECISION_FP64),
            ("fa_rk4_hss", wb, _abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, _abi.PRECISION_FP32_KAHAN),
            ("fa_rk4", wb, _abi.INTEGRATOR_RK4, _abi.COLLISION_NONE, _abi.PRECISION_FP32_KAHAN),
            ("fa_rk4", wa, _abi.INTEGRATOR_RK4, _abi.COLLISION_NONE, _abi.PRECISION_FP64)]
    clouds = []
    for k, (kind, w, integ, coll, prec) in enumerate(plan):
        ph = W.physics(integ, coll, precision=prec, seed=70 + k)
        n = int(rng.integers(1000, w["ensemble"].n))
        clouds.append(dict(kind=kind, domain=w["domain"], physics=ph, species=w["species"],
                           ensemble=w["ensemble"].subset(np.arange(n)), field=w["field"],
                           t0=2e-8 * k, dt=w["dt"], n_steps=int(rng.integers(800, 2000)),
                           step_offset=13 * k))
    for prec in (_abi.PRECISION_FP64, _abi.PRECISION_FP32_KAHAN):
        w = W.ims_hss_workload(n=900, seed=9, precision=prec)
        clouds.append(dict(kind="ims", domain=w["domain"], physics=w["physics"], species=w["species"],
                           ensemble=w["ensemble"].copy(), t0=0.0, dt=w["dt"], n_steps=700,
                           step_offset=0))
    return clouds


def test_field_array_clouds_in_the_batch_against_oracle(oracle):
    """Field-array clouds of the batch against the FP64 oracle's trilinear
    interpolation on the same nodes: FP64 ion by ion (1e-9; >= 99 % with HSS),
    FP32 + Kahan collisionless by vector errors (position against 1 mm: median
    1e-6, max 1e-5; velocity against the speed: median 1e-6, max 2e-5; the same
    fate for >= 99.9 % of ions), FP32 + Kahan with HSS by statistics."""
    clouds = field_array_clouds()
    init = [c["ensemble"].copy() for c in clouds]
    print("batch with field-array clouds: " + ", ".join(
        f"{c['kind']}/{c['physics'].precision} n={c['ensemble'].n} steps={c['n_steps']}" for c in clouds),
        flush=True)
    cloud_stats = batch_run(clouds)
    for c, e0, st in zip(clouds, init, cloud_stats):
        o = e0.copy()
        so = oracle_cloud(oracle, c, o)
        g = c["ensemble"]
        fp64 = c["physics"].precision == _abi.PRECISION_FP64
        stochastic = c["physics"].collision_model != _abi.COLLISION_NONE
        tag = f"{c['kind']} {'fp64' if fp64 else 'fp32'} n={g.n} steps={c['n_steps']}"
        assert st.ion_steps == pytest.approx(so.ion_steps, rel=0.01 if stochastic else 1e-3), tag
        if fp64:
            ok = agree(g, o, 1e-9, 1e-6, 1.0) & (g.n_collisions == o.n_collisions)
            print(f"{tag}: {ok.mean():.4f} of ions equal the oracle to 1e-9", flush=True)
            assert ok.mean() >= (0.99 if stochastic else 1.0), tag
        elif not stochastic:
            assert (g.active == o.active).mean() >= 0.999, tag
            both = g.alive & o.alive
            dp = (np.linalg.norm(g.pos - o.pos, axis=1) / 1e-3)[both]
            speed = np.maximum(np.linalg.norm(o.vel, axis=1), 10.0)
            dv = (np.linalg.norm(g.vel - o.vel, axis=1) / speed)[both]
            print(f"{tag}: position/1mm median {np.median(dp):.2e} max {dp.max():.2e}; "
                  f"velocity/speed median {np.median(dv):.2e} max {dv.max():.2e}", flush=True)
            assert np.median(dp) < 1e-6 and dp.max() < 1e-5, tag
            assert np.median(dv) < 1e-6 and dv.max() < 2e-5, tag
        else:
            z0 = e0.z
            sg, so_ = g.alive.mean(), o.alive.mean()
            s_err = 4 * math.sqrt(2 * max(so_ * (1 - so_), 1.0 / g.n) / g.n) + 0.01
            print(f"{tag}: survival {sg:.4f} vs {so_:.4f}", flush=True)
            assert abs(sg - so_) <= s_err, tag
            close_mean(((g.z - z0)[g.alive]), ((o.z - z0)[o.alive]), tag + " drift")
            close_mean(g.n_collisions.astype(float), o.n_collisions.astype(float),
                       tag + " collisions")
            assert stats.ks_2samp(g.z[g.alive], o.z[o.alive]).pvalue > 1e-3, tag
