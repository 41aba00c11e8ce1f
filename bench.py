"""Benchmark of the batch ion-step path on one MI355X.

Headline workload (the EHSS surrogate configuration): 1M ions in He at 1 atm /
300 K and E/N = 20 Td, each ion its own random rigid cluster of 20-500 atoms
(radii 1.5-2.0 A, FP32 coordinates, mass 12 Da per atom), RK4 with dt = 2 ps.
Two arithmetic paths run over the same clusters and initial cloud:

  * plain: the EHSS geometry method (random orientation, first-contact ray test
    over the atoms) in FP64;
  * fast (the headline value): FP32 + Kahan push with the scattering normal
    drawn by the MLP surrogate (6->128->128->3, BF16 weights, FP32
    accumulation) - trained before the timed region on the GPU, by the
    library's own kernels, on 10^7 first-contact samples from the EHSS
    precompute of 10^4 of the clusters.

A "step" is one engine time step of the whole cloud.  The ensemble is uploaded
before the timed region (inputs resident in HBM); the timed region is exactly
--steps engine steps, bracketed by device synchronisation
(icarion_gpu_synchronize: stream + device).  The plain path is timed on
--plain-steps steps of the same cloud (its rate per ion-step does not depend on
the step count).  The mobility check runs both paths over a long drift of a
smaller cloud and reports K of each, their relative difference and the
statistical error of one K.  The CPU baseline is the oracle's plain EHSS (FP64,
OpenMP) on the host cores, timed on a bounded sample of the same workload.

torch is not loaded: it bundles a second HIP/HSA runtime that cannot share the
device with the library's in one process.

--workload quadrupole runs the RF-quadrupole configuration instead (1M ions,
Boris, Langevin + OU in He at 1 Pa, FP64 vs FP32 + Kahan, transmission per m/z).

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_PEAK_FP64_TFLOPS = 78.6
MI355X_PEAK_FP32_TFLOPS = 157.3
MI355X_PEAK_BF16_TFLOPS = 2500.0

# Algorithmic FLOPs per ion-step (FMA = 2, add/mul/div = 1; sqrt, log, exp,
# sin/cos and the integer Philox rounds not counted), counted from
# icarion_amd/csrc/step_kernels.hip.
#   quadrupole (Boris + quadrupole RF field + Langevin damping + OU kick):
#     force 26, Boris update 18, OU kick 25, uniforms 8, cylinder test 9 (FP64);
#     the FP32 path adds 36 for the Kahan compensation of six components.
#   EHSS surrogate kernel, per ion-step: RK4 in a uniform field 4 stages x
#     (3 velocity + 3 position) FMAs + combination 24 + Kahan 6 x 4 = 84;
#     neutral draw 10, |v_rel| 5, probability 6, cylinder test 5 (twice) = 115.
#     Per collision: the MLP 2 (8*128 + 128*128 + 128*3) + 2*128 + 2*128 + 3 =
#     36 099 (the first layer's K is padded to 8 in the matrix cores), frame +
#     reflection + centre of mass 60.
FLOPS_PER_ION_STEP = {"fp64": 86, "fp32": 122}
EHSS_FLOPS_STEP = 115
# HBM bytes per ion per launch of step_surrogate_kernel, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE (calibrated on a 1 GiB copy) at 1M ions x 2000 steps:
# 141.1 MB read + 70.0 MB written (state round trip + each ion's species row).
EHSS_TRAFFIC_BYTES_PER_ION_LAUNCH = 211.1
EHSS_FLOPS_COLLISION = 2 * (8 * 128 + 128 * 128 + 128 * 3) + 2 * 128 + 2 * 128 + 3 + 60
# Field-array FP32 kernel (RK4 + trilinear + HSS), per ion-step: per stage the
# cell location 9, corner weights 19, 8 corners x 3 components FMA 48, q/m 3
# (x 4 stages = 316); RK4 stage and final combinations 78; Kahan 24; HSS 40;
# two containment tests 18.
FA_FLOPS_PER_ION_STEP = 476
# HBM bytes per ion-step of that kernel in grid-brick order (rocprofv3
# FETCH_SIZE + WRITE_SIZE, copy-calibrated, 4M ions x 2000 steps)
FA_TRAFFIC_BYTES_PER_ION_STEP = 1.06


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("ehss", "quadrupole", "batched", "batched_ehss",
                                           "field_array", "ims", "snapshots", "domains"),
                    default="ehss")
    ap.add_argument("--steps", type=int, default=None,
                    help="timed engine steps (default: ehss 4000, quadrupole 100000)")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--ions", type=int, default=1_000_000)
    ap.add_argument("--steps-per-launch", type=int, default=0)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=100,
                    help="ehss: engine steps of the timed plain FP64 run")
    ap.add_argument("--train-clusters", type=int, default=10_000)
    ap.add_argument("--train-samples", type=int, default=3000,
                    help="first-contact samples per training cluster")
    ap.add_argument("--train-schedule", default="30:2e-3,10:5e-4,10:1.25e-4",
                    help="training stages epochs:learning-rate, comma separated")
    ap.add_argument("--train-epochs", type=int, default=None,
                    help="one constant-rate stage of this many epochs instead of --train-schedule")
    ap.add_argument("--params-out", default=None,
                    help="ehss: save the trained surrogate parameters (.npy)")
    ap.add_argument("--params-in", default=None,
                    help="ehss: load surrogate parameters saved by --params-out instead of "
                         "training (for profiler runs)")
    ap.add_argument("--k-ions", type=int, default=262_144)
    ap.add_argument("--k-steps", type=int, default=40_000)
    ap.add_argument("--no-k-check", action="store_true")
    ap.add_argument("--cpu-sample-ions", type=int, default=None)
    ap.add_argument("--cpu-sample-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--clouds", type=int, default=256, help="batched: number of clouds")
    ap.add_argument("--steps-scale", type=float, default=1.0,
                    help="batched: factor on every cloud's step count (1 = the full workload)")
    ap.add_argument("--autotune-steps", type=int, default=300,
                    help="batched: steps per cloud in each autotuner trial (0 = no tuning)")
    ap.add_argument("--fp64", action="store_true", help="batched: FP64 clouds instead of FP32")
    ap.add_argument("--quad-u", type=float, default=0.0,
                    help="quadrupole / snapshots: DC voltage U (V)")
    ap.add_argument("--write-every", type=int, default=100,
                    help="snapshots: write interval in steps (the reference's default 100)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter "
                         "calibration for rocprofv3 --pmc); loads torch, not the library")
    return ap.parse_args()


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def copy_probe(mib):
    """A plain copy kernel of known size: rocprofv3 FETCH_SIZE / WRITE_SIZE of
    this run calibrate the counters (expected mib MiB read and written)."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4}))


def timed_steps(h, ens, dt, steps, warmup):
    """Warm up, re-upload the initial state, then time exactly `steps` steps."""
    if warmup:
        h.upload_state(ens)
        h.step(0.0, dt, warmup, 0)
    h.upload_state(ens)
    h.synchronize()
    t0 = time.perf_counter()
    st = h.step(0.0, dt, steps, 0)
    h.synchronize()
    return st, time.perf_counter() - t0


# ---------------------------------------------------------------- EHSS surrogate
def ehss_handle(w, precision, mode, spl=0, params=None, n_clusters=None):
    """A handle holding the workload's clusters [0, n_clusters) (all by default);
    cluster c depends only on (seed, c, counts[c])."""
    from icarion_amd import _abi, gpu
    from icarion_amd import workloads as W
    ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_EHSS, precision=precision,
                   seed=w["physics"].rng_seed, ehss_mode=mode)
    h = gpu.Handle(device=0, steps_per_launch=spl)
    h.set_domain(w["domain"], ph)
    _, sig, feat = h.synth_clusters(w["counts"][:n_clusters], seed=w["cluster_seed"])
    if params is not None:
        h.surrogate_set_params(params)
    return h, sig, feat


def train_schedule(args):
    """[(epochs, learning rate), ...] from --train-epochs or --train-schedule."""
    if args.train_epochs is not None:
        return [(args.train_epochs, 2e-3)]
    out = []
    for part in args.train_schedule.split(","):
        e, lr = part.split(":")
        out.append((int(e), float(lr)))
    return out


def train_surrogate(w, args):
    from icarion_amd import _abi
    from icarion_amd import surrogate as S
    n_tr = min(args.train_clusters, len(w["counts"]))
    h, _, feat = ehss_handle(w, _abi.PRECISION_FP32_KAHAN, _abi.EHSS_SURROGATE, n_clusters=n_tr)
    with h:
        t0 = time.perf_counter()
        tm = {}
        schedule = train_schedule(args)
        losses, X, _ = S.train_on_device(h, n_tr, samples_per_cluster=args.train_samples,
                                         features=feat, schedule=schedule, batch=8192, seed=2,
                                         timings=tm)
        params = h.surrogate_get_params()
    return params, dict(clusters=n_tr, samples=int(X.shape[0]), schedule=schedule,
                        epochs=sum(e for e, _ in schedule),
                        batch=8192, loss_first=float(losses[0]), loss_last=float(losses[-1]),
                        seconds=time.perf_counter() - t0, **tm)


def k_of(ens, z0, T, E):
    a = ens.alive
    dz = ens.z[a] - z0[a]
    return dz.mean() / (T * E), dz.std() / math.sqrt(a.sum()) / (T * E)


def mobility_check(args, params):
    """K of the plain FP64 method and of FP32 + surrogate over the same long drift."""
    from icarion_amd import _abi
    from icarion_amd import workloads as W
    w = W.ehss_surrogate_workload(n=args.k_ions, n_steps=args.k_steps, seed=21)
    w["cluster_seed"] = 21
    z0 = w["ensemble"].z.copy()
    T = args.k_steps * w["dt"]
    E = w["domain"].dc_axial_V / w["domain"].length_m
    out = {}
    for name, prec, mode, p in (("plain", _abi.PRECISION_FP64, _abi.EHSS_GEOMETRY, None),
                                ("surrogate", _abi.PRECISION_FP32_KAHAN, _abi.EHSS_SURROGATE,
                                 params)):
        e = w["ensemble"].copy()
        h, _, _ = ehss_handle(w, prec, mode, params=p)
        with h:
            h.upload_state(e)
            h.synchronize()
            t0 = time.perf_counter()
            st = h.step(0.0, w["dt"], args.k_steps, 0)
            el = time.perf_counter() - t0
            h.download_state(e)
        out[name] = (k_of(e, z0, T, E), st, el)
        log(f"mobility check {name}: K {out[name][0][0]:.5e} in {el:.1f} s")
    (kp, se), sp, tp = out["plain"]
    (ks, ses), ss, ts = out["surrogate"]
    # the two runs are independent samples: the difference carries both errors
    sd = math.hypot(se, ses) / kp
    return {"ions": args.k_ions, "steps": args.k_steps, "drift_time_s": T,
            "K_plain_m2_Vs": kp, "K_surrogate_m2_Vs": ks, "rel_error": abs(ks / kp - 1),
            "stat_error_one_K": se / kp, "stat_error_difference": sd,
            "rel_error_upper_95": abs(ks / kp - 1) + 1.96 * sd,
            "collisions_plain": sp.collisions, "collisions_surrogate": ss.collisions,
            "rate_plain": sp.ion_steps / tp, "rate_surrogate": ss.ion_steps / ts}


def angle_law_check(w, params, n_clusters=64, n_samples=20000):
    """Scattering-angle law of the surrogate against plain EHSS on clusters
    outside the training set (device clusters, device EHSS precompute, matrix-
    core inference): per-cluster two-sample KS distance of mu = cos(contact
    normal, axis), next to that of two independent EHSS samples and of the
    hard-sphere law."""
    from scipy import stats
    from icarion_amd import _abi, gpu
    from icarion_amd import surrogate as S
    from icarion_amd import workloads as W
    seed = 9999
    counts = W.ehss_cluster_counts(n_clusters, seed=seed)
    rng = np.random.default_rng(seed)
    with gpu.Handle(device=0) as h:
        h.set_domain(w["domain"], W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_EHSS))
        _, _, feat = h.synth_clusters(counts, seed=seed)
        mu1 = h.ehss_precompute(0, n_clusters, n_samples, 0, 1, seed=31)[1][..., 0]
        mu2 = h.ehss_precompute(0, n_clusters, n_samples, 0, 1, seed=32)[1][..., 0]
        h.surrogate_set_params(params)
        X = np.zeros((n_clusters, n_samples, 6), np.float32)
        psi = rng.uniform(0.0, 1.0, X.shape[:2])
        X[..., 0] = rng.uniform(0.0, 1.0, X.shape[:2])
        X[..., 1], X[..., 2] = np.cos(2 * np.pi * psi), np.sin(2 * np.pi * psi)
        X[..., 3:] = feat[:, None, :]
        mu_s = S.mu_from_output(h.surrogate_infer(X.reshape(-1, 6))).reshape(mu1.shape)
    sur = [stats.ks_2samp(mu_s[c], mu1[c]) for c in range(n_clusters)]
    d_null = [stats.ks_2samp(mu2[c], mu1[c]).statistic for c in range(n_clusters)]
    d_hs = [stats.ks_2samp(np.sqrt(rng.uniform(0.0, 1.0, n_samples)), mu1[c]).statistic
            for c in range(n_clusters)]
    return {"held_out_clusters": n_clusters, "samples_per_cluster": n_samples,
            "ks_D_surrogate_vs_ehss": float(np.mean([r.statistic for r in sur])),
            "ks_D_ehss_vs_ehss": float(np.mean(d_null)),
            "ks_D_hard_sphere_vs_ehss": float(np.mean(d_hs)),
            "clusters_with_p_below_0.01": int(sum(r.pvalue < 0.01 for r in sur))}


def cpu_baseline_ehss(w, n_ions, n_steps):
    """The oracle's plain EHSS (FP64) on the host cores over the first n_ions
    ions of the workload (their clusters regenerated on the device)."""
    from oracle import binding
    from icarion_amd import _abi, gpu
    from icarion_amd import workloads as W
    counts = w["counts"][:n_ions]
    with gpu.Handle(device=0) as h:
        h.set_domain(w["domain"], W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_EHSS))
        atoms, sig, _ = h.synth_clusters(counts, seed=w["cluster_seed"], want_atoms=True)
    offs = np.zeros(n_ions, np.int64)
    offs[1:] = np.cumsum(counts[:-1])
    species = [W.species(sigma_ehss=float(s), atom_offset=int(o), atom_count=int(c))
               for s, o, c in zip(sig, offs, counts)]
    ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_EHSS, seed=w["physics"].rng_seed)
    e = w["ensemble"].subset(np.arange(n_ions))
    threads = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    st = binding.run(w["domain"], ph, species, e, 0.0, w["dt"], n_steps, atoms=atoms,
                     threads=threads)
    el = time.perf_counter() - t0
    return dict(value=st.ion_steps / el, unit="ion-steps/s", cores=threads, kind="port",
                sample=f"{n_ions} ions x {n_steps} steps of the same workload, plain EHSS "
                       f"(FP64 oracle, OpenMP over {threads} threads, {el:.1f} s)")


def bench_ehss(args):
    from icarion_amd import _abi
    from icarion_amd import workloads as W
    steps = args.steps if args.steps is not None else 4000
    w = W.ehss_surrogate_workload(n=args.ions, n_steps=steps, seed=21)
    w["cluster_seed"] = 21
    if args.params_in:
        params = np.load(args.params_in)
        train = {"loaded_from": os.path.basename(args.params_in)}
        log(f"{w['name']} {args.ions} ions x {steps} steps; surrogate loaded")
    else:
        log(f"{w['name']} {args.ions} ions x {steps} steps; training the surrogate")
        params, train = train_surrogate(w, args)
        log(f"trained in {train['seconds']:.1f} s (loss {train['loss_first']:.4f} -> "
            f"{train['loss_last']:.4f})")
    if args.params_out:
        np.save(args.params_out, params)

    h, _, _ = ehss_handle(w, _abi.PRECISION_FP32_KAHAN, _abi.EHSS_SURROGATE,
                          args.steps_per_launch, params)
    with h:
        fs, fwall = timed_steps(h, w["ensemble"].copy(), w["dt"], steps, args.warmup)
    spl = args.steps_per_launch or 2000
    launches = math.ceil(steps / spl)
    value = fs.ion_steps / fwall
    log(f"fp32+surrogate {fwall:.3f} s, {value:.3e} ion-steps/s")
    avg_launch_s = fs.device_ms * 1e-3 / launches
    # the MLP's bf16 matrix-core work is >99 % of the algorithmic FLOPs, so the
    # kernel is priced against the dense bf16 peak
    flops = EHSS_FLOPS_STEP * fs.ion_steps + EHSS_FLOPS_COLLISION * fs.collisions
    achieved = flops / launches / avg_launch_s / 1e12
    res = {
        "metric": "ion-steps/s on 1 MI355X, plain FP64 vs FP32+surrogate at <=0.5% K error",
        "value": value,
        "unit": "ion-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": fwall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (Kahan-compensated state; bf16 MLP weights, f32 accumulation; "
                 "f64 time)",
        "data": "synthetic",
        "config": {
            "workload": "ehss_surrogate_he_20td",
            "ions": args.ions, "clusters": "one per ion, 20-500 atoms, radii 1.5-2.0 A",
            "gas": "He 101325 Pa 300 K", "E_over_N_Td": 20.0, "dt_s": w["dt"],
            "integrator": "rk4", "engine_steps": steps,
            "mlp": "6-128-128-3 bf16/f32", "training": train,
        },
        "roofline": {
            "bound": "mfma", "achieved": achieved, "peak": MI355X_PEAK_BF16_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / MI355X_PEAK_BF16_TFLOPS,
            "traffic": EHSS_TRAFFIC_BYTES_PER_ION_LAUNCH * args.ions,
            "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, copy-calibrated "
                              "(profiles/step_surrogate_kernel_ehss_he_20td_1M.md)",
            "kernel": "step_surrogate_kernel<RK4, linear field>",
            "avg_launch_ms": avg_launch_s * 1e3, "launches": launches,
            "flops_per_ion_step": EHSS_FLOPS_STEP,
            "flops_per_collision": EHSS_FLOPS_COLLISION,
            "collisions_per_ion_step": fs.collisions / max(fs.ion_steps, 1),
        },
        "alive_fraction": fs.alive / args.ions,
    }
    if not args.skip_fp64:
        h, _, _ = ehss_handle(w, _abi.PRECISION_FP64, _abi.EHSS_GEOMETRY, args.steps_per_launch)
        with h:
            ps, pwall = timed_steps(h, w["ensemble"].copy(), w["dt"], args.plain_steps,
                                    min(args.warmup, args.plain_steps))
        pv = ps.ion_steps / pwall
        log(f"fp64 plain EHSS {pwall:.3f} s, {pv:.3e} ion-steps/s")
        res["fp64_plain"] = {"value": pv, "unit": "ion-steps/s", "steps": args.plain_steps,
                             "wall_s": pwall,
                             "collisions_per_ion_step": ps.collisions / max(ps.ion_steps, 1)}
        res["surrogate_over_plain"] = value / pv
    if not args.no_k_check:
        res["mobility_check"] = mc = mobility_check(args, params)
        log(f"K rel. error {mc['rel_error']:.4%} (statistical error of the difference "
            f"{mc['stat_error_difference']:.3%}, 95 % upper bound {mc['rel_error_upper_95']:.3%})")
    if not args.no_k_check:
        res["angle_law_check"] = al = angle_law_check(w, params)
        log(f"angle law KS D: surrogate {al['ks_D_surrogate_vs_ehss']:.4f}, two EHSS samples "
            f"{al['ks_D_ehss_vs_ehss']:.4f}, hard sphere {al['ks_D_hard_sphere_vs_ehss']:.4f}")
    if not args.no_cpu:
        n = args.cpu_sample_ions or 8192
        k = args.cpu_sample_steps or 16000
        res["cpu_baseline"] = cpu_baseline_ehss(w, min(n, args.ions), k)
    return res


# ---------------------------------------------------------------- quadrupole
def run_quadrupole_path(w, precision, steps, warmup, spl):
    from icarion_amd import gpu
    from icarion_amd import workloads as W
    ph = W.physics(w["physics"].integrator, w["physics"].collision_model,
                   w["physics"].ou_thermal_kicks, precision, w["physics"].rng_seed)
    ens = w["ensemble"].copy()
    with gpu.Handle(device=0, steps_per_launch=spl) as h:
        h.set_domain(w["domain"], ph)
        h.set_species(w["species"])
        st, wall = timed_steps(h, ens, w["dt"], steps, warmup)
        h.download_state(ens)
    launches = math.ceil(steps / (spl or 2000))
    return dict(ens=ens, stats=st, wall=wall, launches=launches)


def transmission(ens, masses_amu):
    bins = np.geomspace(100, 2000, 9)
    out = []
    for lo, hi in zip(bins[:-1], bins[1:]):
        sel = (masses_amu >= lo) & (masses_amu < hi)
        out.append(float(ens.alive[sel].mean()))
    return bins, out


def cpu_baseline_quadrupole(w, n_ions, n_steps):
    from oracle import binding
    threads = len(os.sched_getaffinity(0))
    e = w["ensemble"].subset(np.arange(n_ions))
    t0 = time.perf_counter()
    st = binding.run(w["domain"], w["physics"], w["species"], e, 0.0, w["dt"], n_steps,
                     threads=threads)
    el = time.perf_counter() - t0
    return dict(value=st.ion_steps / el, unit="ion-steps/s", cores=threads, kind="port",
                sample=f"{n_ions} ions x {n_steps} steps of the same workload "
                       f"(FP64 oracle, OpenMP over {threads} threads, {el:.1f} s)")


def bench_quadrupole(args):
    from icarion_amd import _abi
    from icarion_amd import workloads as W
    steps = args.steps if args.steps is not None else 100_000
    w = W.quadrupole_workload(n=args.ions, n_steps=steps)
    masses = w["ensemble"].mass_kg / 1.66053906660e-27
    log(f"{w['name']} {args.ions} ions x {steps} steps")
    fast = run_quadrupole_path(w, _abi.PRECISION_FP32_KAHAN, steps, args.warmup,
                               args.steps_per_launch)
    log(f"fp32+kahan {fast['wall']:.3f} s")
    fs = fast["stats"]
    value = fs.ion_steps / fast["wall"]
    avg_launch_s = fs.device_ms * 1e-3 / fast["launches"]
    achieved = FLOPS_PER_ION_STEP["fp32"] * (fs.ion_steps / fast["launches"]) / avg_launch_s / 1e12
    res = {
        "metric": "ion-steps/s on 1 MI355X, plain FP64 vs FP32+surrogate at <=0.5% K error",
        "value": value, "unit": "ion-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": fast["wall"] * 1e3 / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (Kahan-compensated state; f64 time and RF phase)",
        "data": "synthetic",
        "config": {
            "workload": "rf_quadrupole_boris_langevin_ou",
            "ions": args.ions, "dt_s": w["dt"], "engine_steps": steps,
            "masses_amu": "log-uniform 100-2000", "r0_m": 4e-3, "V_rf": 500.0,
            "f_rf_Hz": 1e6, "U_dc": 0.0, "gas": "He 1 Pa 300 K",
        },
        "roofline": {
            "bound": "mfma", "achieved": achieved, "peak": MI355X_PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / MI355X_PEAK_FP32_TFLOPS, "traffic": None,
            "kernel": "step_f32_kernel<Boris, linear field, damping>",
            "avg_launch_ms": avg_launch_s * 1e3, "launches": fast["launches"],
            "flops_per_ion_step": FLOPS_PER_ION_STEP["fp32"],
        },
        "alive_fraction": fs.alive / args.ions,
    }
    if not args.skip_fp64:
        plain = run_quadrupole_path(w, _abi.PRECISION_FP64, steps, args.warmup,
                                    args.steps_per_launch)
        log(f"fp64 {plain['wall']:.3f} s")
        ps = plain["stats"]
        pv = ps.ion_steps / plain["wall"]
        p_avg = ps.device_ms * 1e-3 / plain["launches"]
        p_ach = FLOPS_PER_ION_STEP["fp64"] * (ps.ion_steps / plain["launches"]) / p_avg / 1e12
        bins, tf = transmission(fast["ens"], masses)
        _, tp = transmission(plain["ens"], masses)
        res["fp64_plain"] = {"value": pv, "unit": "ion-steps/s", "wall_s": plain["wall"],
                             "achieved_tflops": p_ach, "peak_tflops": MI355X_PEAK_FP64_TFLOPS,
                             "frac": p_ach / MI355X_PEAK_FP64_TFLOPS}
        res["fast_over_plain"] = value / pv
        res["transmission_bins_amu"] = [float(b) for b in bins]
        res["transmission_fp32"] = tf
        res["transmission_fp64"] = tp
        res["transmission_max_abs_diff"] = float(np.max(np.abs(np.array(tf) - np.array(tp))))
    if not args.no_cpu:
        n_cpu = min(args.cpu_sample_ions or 100_000, args.ions)
        res["cpu_baseline"] = cpu_baseline_quadrupole(w, n_cpu,
                                                      min(args.cpu_sample_steps or 20000, steps))
    return res


# ---------------------------------------------------------------- batched clouds
# Algorithmic FLOPs per ion-step of the FP32 + Kahan bodies, counted from
# step_kernels.hip as above: IMS RK4 + HSS 84 push + 31 collision test + ~20 x
# the collision fraction; quadrupole Boris + Langevin + OU 122; Orbitrap RK4 4 x
# (13 field + 3 accel + 9 stage) + 30 combination + 24 Kahan = 154.
BATCH_FLOPS = {"ims": 130, "quadrupole": 122, "orbitrap": 154}


def cpu_baseline_batched(clouds, stats, sample_ions, sample_steps):
    """The batch on the CPU is the sum of per-cloud oracle runs: one cloud of
    each kind is timed on a sample (its first sample_ions ions x sample_steps
    steps), and every cloud's ion-steps are priced at its kind's rate."""
    from oracle import binding
    threads = len(os.sched_getaffinity(0))
    rate, parts = {}, []
    for c in clouds:
        k = c["kind"]
        if k in rate:
            continue
        n = min(sample_ions, c["ensemble"].n)
        e = c["ensemble"].subset(np.arange(n))
        steps = min(sample_steps, c["n_steps"])
        t0 = time.perf_counter()
        st = binding.run(c["domain"], c["physics"], c["species"], e, 0.0, c["dt"], steps, threads=threads)
        el = time.perf_counter() - t0
        rate[k] = st.ion_steps / el
        parts.append(f"{k} {n} ions x {steps} steps {el:.1f} s ({rate[k]:.3e} ion-steps/s)")
    total = sum(st.ion_steps for st in stats)
    cpu_s = sum(st.ion_steps / rate[c["kind"]] for c, st in zip(clouds, stats))
    return dict(value=total / cpu_s, unit="ion-steps/s", cores=threads, kind="port",
                sample="sum of per-cloud FP64 oracle runs, each kind timed on one cloud's sample "
                       f"(OpenMP over {threads} threads): " + "; ".join(parts)
                       + f"; whole batch {cpu_s:.0f} CPU-seconds")


def bench_batched(args):
    from icarion_amd import _abi, gpu
    from icarion_amd import workloads as W
    prec = _abi.PRECISION_FP64 if args.fp64 else _abi.PRECISION_FP32_KAHAN
    steps = {k: max(1, int(round(v * args.steps_scale)))
             for k, v in {"ims": 100_000, "quadrupole": 100_000, "orbitrap": 1_000_000}.items()}
    t0 = time.perf_counter()
    clouds = W.batched_workload(n_clouds=args.clouds, precision=prec, steps=steps)
    n_ions = sum(c["ensemble"].n for c in clouds)
    log(f"batched: {args.clouds} clouds, {n_ions} ions, steps {steps} "
        f"(built in {time.perf_counter() - t0:.1f} s)")
    with gpu.Handle(device=0) as h:
        h.batch_upload(clouds)
        tuning = None
        if args.autotune_steps:
            t1 = time.perf_counter()
            t64, t32, r64, r32 = h.batch_autotune(args.autotune_steps)
            tuning = {"fp64": t64, "fp32": t32, "trial_steps": args.autotune_steps,
                      "trial_rate": r64 if args.fp64 else r32,
                      "seconds": time.perf_counter() - t1}
            log(f"autotuned {t64 if args.fp64 else t32} in {tuning['seconds']:.1f} s")
        h.synchronize()
        t1 = time.perf_counter()
        stats, dev_ms = h.batch_run()
        h.synchronize()
        wall = time.perf_counter() - t1
    total = sum(st.ion_steps for st in stats)
    value = total / wall
    log(f"batched: {total:.3e} ion-steps in {wall:.2f} s -> {value:.3e} ion-steps/s")
    done = np.array([st.done_ms for st in stats])
    kinds = [c["kind"] for c in clouds]
    flops = sum(BATCH_FLOPS[k] * st.ion_steps for k, st in zip(kinds, stats))
    peak = MI355X_PEAK_FP64_TFLOPS if args.fp64 else MI355X_PEAK_FP32_TFLOPS
    achieved = flops / (dev_ms * 1e-3) / 1e12
    per_kind = {}
    for k in ("ims", "quadrupole", "orbitrap"):
        sel = [i for i, kk in enumerate(kinds) if kk == k]
        per_kind[k] = {"clouds": len(sel), "ions": int(sum(clouds[i]["ensemble"].n for i in sel)),
                       "steps": steps[k],
                       "ion_steps": int(sum(stats[i].ion_steps for i in sel)),
                       "alive_fraction": float(sum(stats[i].alive for i in sel)
                                               / max(1, sum(clouds[i]["ensemble"].n for i in sel))),
                       "done_ms_max": float(done[sel].max()) if sel else 0.0,
                       "done_ms_median": float(np.median(done[sel])) if sel else 0.0}
    res = {
        "metric": "ion-steps/s on 1 MI355X, plain FP64 vs FP32+surrogate at <=0.5% K error",
        "value": value, "unit": "ion-steps/s", "n_gpus": 1, "steps": max(steps.values()),
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.fp64 else "f32 (Kahan-compensated state; f64 time and RF phase)",
        "data": "synthetic",
        "config": {"workload": "batched_mixed_instruments", "clouds": args.clouds,
                   "ions": n_ions, "sizes": "log-uniform 1e3-2e5", "steps_per_kind": steps,
                   "launch": "one graph: reset + persistent tile-queue kernel",
                   "tuning": tuning},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "batch_kernel<%s>" % ("FP64" if args.fp64 else "FP32"),
                     "device_ms": dev_ms, "flops_per_ion_step": BATCH_FLOPS},
        "per_kind": per_kind,
        "completion_ms": {"min": float(done.min()), "median": float(np.median(done)),
                          "max": float(done.max())},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_batched(clouds, stats, args.cpu_sample_ions or 16384,
                                                   args.cpu_sample_steps or 2000)
    return res


def bench_batched_ehss(args):
    """Many EHSS clouds (sizes log-uniform 1e3-2e5, each with its own table of 16
    random clusters) through the batch's FP32 + surrogate class in one graph
    launch, against the plain FP64 EHSS rate on the 1M-ion EHSS workload."""
    from icarion_amd import _abi, gpu
    from icarion_amd import workloads as W
    steps = args.steps if args.steps is not None else 4000
    params = np.load(args.params_in or os.path.join(ROOT, "icarion_amd", "data",
                                                    "surrogate_he_20td_clusters.npy"))
    n_pool = 4096
    pool = W.ehss_surrogate_workload(n=n_pool, seed=21)
    pool["cluster_seed"] = 21
    h, sig, feat = ehss_handle(pool, _abi.PRECISION_FP32_KAHAN, _abi.EHSS_SURROGATE)
    h.close()
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(math.log(1e3), math.log(2e5), args.clouds)).astype(int)
    clouds = []
    for c, n in enumerate(sizes):
        idx = rng.choice(n_pool, 16, replace=False)
        integ = _abi.INTEGRATOR_RK4 if c % 2 == 0 else _abi.INTEGRATOR_BORIS
        w = W.ehss_surrogate_cloud(int(n), sig[idx], feat[idx], pool["counts"][idx],
                                   seed=5000 + c, integrator=integ, n_steps=steps)
        clouds.append(dict(kind="ehss_surrogate", domain=w["domain"], physics=w["physics"],
                           species=w["species"], ensemble=w["ensemble"], t0=0.0, dt=w["dt"],
                           n_steps=steps, step_offset=0))
    n_ions = int(sizes.sum())
    log(f"batched EHSS: {args.clouds} clouds, {n_ions} ions x {steps} steps")
    with gpu.Handle(device=0) as h:
        h.surrogate_set_params(params)
        h.batch_upload(clouds)
        h.synchronize()
        t1 = time.perf_counter()
        stats, dev_ms = h.batch_run()
        h.synchronize()
        wall = time.perf_counter() - t1
    total = sum(st.ion_steps for st in stats)
    coll = sum(st.collisions for st in stats)
    value = total / wall
    log(f"batched EHSS: {total:.3e} ion-steps in {wall:.2f} s -> {value:.3e} ion-steps/s")
    achieved = (EHSS_FLOPS_STEP * total + EHSS_FLOPS_COLLISION * coll) / (dev_ms * 1e-3) / 1e12
    done = np.array([st.done_ms for st in stats])
    res = {
        "metric": "ion-steps/s on 1 MI355X, plain FP64 vs FP32+surrogate at <=0.5% K error",
        "value": value, "unit": "ion-steps/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": wall * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (Kahan-compensated state; bf16 MLP weights, f32 accumulation; f64 time)",
        "data": "synthetic",
        "config": {"workload": "batched_ehss_surrogate_he_20td", "clouds": args.clouds,
                   "ions": n_ions, "sizes": "log-uniform 1e3-2e5",
                   "clusters_per_cloud": 16, "integrators": "RK4 / Boris alternating",
                   "launch": "one graph: reset + persistent surrogate tile-queue kernel",
                   "weights": "committed (icarion_amd/data/surrogate_he_20td_clusters.npy)"
                   if not args.params_in else os.path.basename(args.params_in)},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MI355X_PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / MI355X_PEAK_BF16_TFLOPS,
                     "traffic": None, "kernel": "batch_surrogate_kernel", "device_ms": dev_ms,
                     "collisions_per_ion_step": coll / max(total, 1)},
        "completion_ms": {"min": float(done.min()), "median": float(np.median(done)),
                          "max": float(done.max())},
        "alive_fraction": float(sum(st.alive for st in stats) / n_ions),
    }
    if not args.skip_fp64:
        w = W.ehss_surrogate_workload(n=args.ions, n_steps=args.plain_steps, seed=21)
        w["cluster_seed"] = 21
        h, _, _ = ehss_handle(w, _abi.PRECISION_FP64, _abi.EHSS_GEOMETRY, args.steps_per_launch)
        with h:
            ps, pwall = timed_steps(h, w["ensemble"].copy(), w["dt"], args.plain_steps,
                                    min(args.warmup, args.plain_steps))
        pv = ps.ion_steps / pwall
        log(f"fp64 plain EHSS (1M-ion workload) {pv:.3e} ion-steps/s")
        res["fp64_plain"] = {"value": pv, "unit": "ion-steps/s", "steps": args.plain_steps,
                             "workload": "ehss_he_20td, %d ions" % args.ions}
        res["surrogate_batched_over_plain"] = value / pv
    if not args.no_cpu:
        # the same physics (plain EHSS of random clusters in He, RK4) as the CPU
        # would run it, timed on the EHSS workload's first ions
        w = W.ehss_surrogate_workload(n=args.ions, n_steps=args.plain_steps, seed=21)
        w["cluster_seed"] = 21
        res["cpu_baseline"] = cpu_baseline_ehss(w, min(args.cpu_sample_ions or 8192, args.ions),
                                                args.cpu_sample_steps or 16000)
    return res


# ---------------------------------------------------------------- field array (C3)
def cpu_baseline_field_array(w, n_ions, n_steps):
    """The oracle (FP64, trilinear interpolation on the same 256^3 grid) over the
    first n_ions ions of the workload."""
    from oracle import binding
    threads = len(os.sched_getaffinity(0))
    E, origin, spacing = w["field"]
    field = binding.Field(E, origin, spacing)
    e = w["ensemble"].subset(np.arange(n_ions))
    t0 = time.perf_counter()
    st = binding.run(w["domain"], w["physics"], w["species"], e, 0.0, w["dt"], n_steps, field=field,
                     threads=threads)
    el = time.perf_counter() - t0
    return dict(value=st.ion_steps / el, unit="ion-steps/s", cores=threads, kind="port",
                sample=f"{n_ions} ions x {n_steps} steps of the same workload "
                       f"(FP64 oracle, OpenMP over {threads} threads, {el:.1f} s)")


def bench_field_array(args):
    from icarion_amd import _abi, gpu
    from icarion_amd import workloads as W
    steps = args.steps if args.steps is not None else 20_000
    t0 = time.perf_counter()
    w = W.field_array_workload(n=args.ions if args.ions != 1_000_000 else 4_000_000,
                               n_steps=steps)
    n = w["ensemble"].n
    log(f"field array: {n} ions x {steps} steps, grid built in {time.perf_counter() - t0:.1f} s")
    out = {}
    for name, prec in (("fp32", _abi.PRECISION_FP32_KAHAN), ("fp64", _abi.PRECISION_FP64)):
        if name == "fp64" and args.skip_fp64:
            continue
        ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, precision=prec,
                       seed=w["physics"].rng_seed)
        e = w["ensemble"].copy()
        with gpu.Handle(device=0, steps_per_launch=args.steps_per_launch) as h:
            h.set_domain(w["domain"], ph)
            h.set_species(w["species"])
            E, origin, spacing = w["field"]
            h.upload_field_array(E, origin, spacing)
            st, wall = timed_steps(h, e, w["dt"], steps, min(args.warmup, steps))
            h.download_state(e)
        out[name] = dict(stats=st, wall=wall, ens=e)
        log(f"{name}: {wall:.2f} s, {st.ion_steps / wall:.3e} ion-steps/s")
    fs = out["fp32"]["stats"]
    value = fs.ion_steps / out["fp32"]["wall"]
    spl = args.steps_per_launch or 2000
    launches = math.ceil(steps / spl)
    avg = fs.device_ms * 1e-3 / launches
    # the launches run in grid-brick order, so the corner gathers (4 stages x 8
    # corners x 16 B per ion-step) are served by the caches: rocprofv3 counts
    # 1.06 B of HBM traffic per ion-step (FETCH_SIZE + WRITE_SIZE, copy-
    # calibrated) and 99 % L2 hits; the kernel is bound by vector issue, priced
    # against the FP32 vector peak
    tflops = FA_FLOPS_PER_ION_STEP * fs.ion_steps / launches / avg / 1e12
    gbytes = 4 * 8 * 16 * fs.ion_steps / launches / avg / 1e9
    res = {
        "metric": "ion-steps/s on 1 MI355X, plain FP64 vs FP32+surrogate at <=0.5% K error",
        "value": value, "unit": "ion-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": min(args.warmup, steps), "ms_per_step": out["fp32"]["wall"] * 1e3 / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (Kahan-compensated state; f64 time)", "data": "synthetic",
        "config": {"workload": "field_array_hss_rk4", "ions": n, "grid": [256, 256, 256],
                   "extent_mm": [10, 10, 50], "gas": "N2 300 Pa 300 K", "dt_s": w["dt"]},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": MI355X_PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": tflops / MI355X_PEAK_FP32_TFLOPS,
                     "traffic": FA_TRAFFIC_BYTES_PER_ION_STEP * fs.ion_steps / launches,
                     "kernel": "step_f32_kernel<RK4, field array, HSS>",
                     "avg_launch_ms": avg * 1e3, "launches": launches,
                     "flops_per_ion_step": FA_FLOPS_PER_ION_STEP,
                     "gather_rate_GB_s": gbytes,
                     "note": "vector FP32 (not matrix) peak; gather_rate = 512 B of corner "
                             "gathers per ion-step, served by L1/L2 in grid-brick order"},
        "alive_fraction": fs.alive / n,
    }
    if "fp64" in out:
        ps = out["fp64"]["stats"]
        pv = ps.ion_steps / out["fp64"]["wall"]
        g, o = out["fp32"]["ens"], out["fp64"]["ens"]
        z0 = w["ensemble"].z
        res["fp64_plain"] = {"value": pv, "unit": "ion-steps/s"}
        res["fast_over_plain"] = value / pv
        res["fp32_vs_fp64"] = {
            "alive": [float(g.alive.mean()), float(o.alive.mean())],
            "mean_dz_m": [float((g.z - z0)[g.alive].mean()), float((o.z - z0)[o.alive].mean())]}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_field_array(w, min(args.cpu_sample_ions or 65536, n),
                                                       min(args.cpu_sample_steps or 2000, steps))
    return res


# ---------------------------------------------------------------- IMS (C1)
def bench_ims(args):
    """C1 at full size on the GPU (FP64 and FP32) and on the host (the oracle,
    all cores): mobility K, arrival-time mean and width at the detector plane
    and survival, compared."""
    from oracle import binding
    from icarion_amd import _abi, gpu
    from icarion_amd import workloads as W
    steps = args.steps if args.steps is not None else 100_000
    n = args.ions if args.ions != 1_000_000 else 10_000
    w = W.ims_hss_workload(n=n, n_steps=steps)
    z0 = w["ensemble"].z.copy()
    E = w["domain"].dc_axial_V / w["domain"].length_m
    T = steps * w["dt"]

    def summary(e, st, wall):
        a = e.alive
        arr = e.arrival_time_s[np.isfinite(e.arrival_time_s)]
        return {"ion_steps_per_s": st.ion_steps / wall, "K_m2_Vs": float((e.z - z0)[a].mean() / (T * E)),
                "arrival_mean_s": float(arr.mean()) if arr.size else None,
                "arrival_std_s": float(arr.std()) if arr.size else None,
                "arrived_fraction": float(arr.size / e.n), "survival": float(a.mean()),
                "collisions_per_ion_step": st.collisions / max(st.ion_steps, 1)}
    out = {}
    for name, prec in (("fp32", _abi.PRECISION_FP32_KAHAN), ("fp64", _abi.PRECISION_FP64)):
        ph = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, precision=prec,
                       seed=w["physics"].rng_seed)
        e = w["ensemble"].copy()
        with gpu.Handle(device=0) as h:
            h.set_domain(w["domain"], ph)
            h.set_species(w["species"])
            st, wall = timed_steps(h, e, w["dt"], steps, 0)
            h.download_state(e)
        out[name] = summary(e, st, wall)
        log(f"ims {name}: {out[name]}")
    e = w["ensemble"].copy()
    threads = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    st = binding.run(w["domain"], w["physics"], w["species"], e, 0.0, w["dt"], steps,
                     threads=threads)
    wall = time.perf_counter() - t0
    out["oracle"] = summary(e, st, wall)
    log(f"ims oracle: {out['oracle']}")
    f, o = out["fp32"], out["oracle"]
    return {
        "metric": "ion-steps/s on 1 MI355X, plain FP64 vs FP32+surrogate at <=0.5% K error",
        "value": f["ion_steps_per_s"], "unit": "ion-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (Kahan-compensated state; f64 time)",
        "data": "synthetic",
        "config": {"workload": "ims_hss_rk4_10td", "ions": n, "dt_s": w["dt"],
                   "gas": "N2 101325 Pa 300 K", "E_over_N_Td": 10.0},
        "results": out,
        "fp32_vs_oracle": {"K_rel": abs(f["K_m2_Vs"] / o["K_m2_Vs"] - 1),
                           "arrival_mean_rel": (abs(f["arrival_mean_s"] / o["arrival_mean_s"] - 1)
                                                if o["arrival_mean_s"] else None),
                           "survival_abs": abs(f["survival"] - o["survival"])},
        "cpu_baseline": {"value": o["ion_steps_per_s"], "unit": "ion-steps/s", "cores": threads,
                         "kind": "port", "sample": f"the whole workload ({n} ions x {steps} "
                                                    f"steps), FP64 oracle, {wall:.1f} s"},
    }


def bench_snapshots(args):
    """Stepping with trajectory snapshots every --write-every steps (the engine's
    sampling of OutputManager) on the quadrupole cloud, FP32 + Kahan.  Four runs
    of the same steps from the same state: no snapshots; snapshots through
    icarion_gpu_step_snapshots (packed on the device, copied out on a second stream
    while stepping goes on), consumed by a copy into a host buffer (the value);
    the same with a consumer that keeps nothing (the device and PCIe share); and
    the resident-mode loop without overlap (step to the snapshot, download the
    state, copy it into the buffer)."""
    from icarion_amd import _abi, gpu
    from icarion_amd import workloads as W
    steps = args.steps if args.steps is not None else 20_000
    w = W.quadrupole_workload(n=args.ions, n_steps=steps, U=args.quad_u)
    ph = W.physics(w["physics"].integrator, w["physics"].collision_model,
                   w["physics"].ou_thermal_kicks, _abi.PRECISION_FP32_KAHAN, w["physics"].rng_seed)
    dt, n = w["dt"], args.ions
    interval = args.write_every * dt
    ens = w["ensemble"].copy()
    log(f"snapshots: {n} ions x {steps} steps, every {args.write_every} steps")
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    spec = np.empty(n, np.uint32)
    flags = np.empty(n, np.uint8)
    snap_steps = []

    def consume(snap):
        np.copyto(pos, snap["positions"])
        np.copyto(vel, snap["velocities"])
        np.copyto(spec, snap["species"])
        np.copyto(flags, snap["flags"])
        snap_steps.append(snap["step"])

    with gpu.Handle(device=0) as h:
        h.set_domain(w["domain"], ph)
        h.set_species(w["species"])
        h.upload_state(ens)
        h.step(0.0, dt, args.warmup, 0)
        # no snapshots
        h.upload_state(ens)
        h.synchronize()
        t0 = time.perf_counter()
        st_plain = h.step(0.0, dt, steps, 0)
        h.synchronize()
        wall_plain = time.perf_counter() - t0
        log(f"no snapshots {wall_plain:.3f} s")
        # snapshots, overlapped
        h.upload_state(ens)
        h.synchronize()
        t0 = time.perf_counter()
        _, _, st_snap = h.step_snapshots(0.0, dt, steps, interval, on_snapshot=consume, copy=False)
        h.synchronize()
        wall_snap = time.perf_counter() - t0
        log(f"snapshots {wall_snap:.3f} s ({len(snap_steps)} snapshots)")
        # snapshots handed to a consumer that keeps nothing: the device + PCIe share
        h.upload_state(ens)
        h.synchronize()
        t0 = time.perf_counter()
        _, _, st_noop = h.step_snapshots(0.0, dt, steps, interval, on_snapshot=lambda s: None,
                                         copy=False)
        h.synchronize()
        wall_noop = time.perf_counter() - t0
        log(f"snapshots, no-op consumer {wall_noop:.3f} s")
        # resident-mode loop: step to each snapshot, download, consume
        e = ens.copy()
        h.upload_state(ens)
        h.synchronize()
        t0 = time.perf_counter()
        t, prev = 0.0, 0
        for k in snap_steps:
            st = h.step(t, dt, k + 1 - prev, prev)
            h.download_state(e)
            np.copyto(pos[:, 0], e.x)
            np.copyto(pos[:, 1], e.y)
            np.copyto(pos[:, 2], e.z)
            np.copyto(vel[:, 0], e.vx)
            np.copyto(vel[:, 1], e.vy)
            np.copyto(vel[:, 2], e.vz)
            t, prev = st.t_end_s, k + 1
        if prev < steps:
            h.step(t, dt, steps - prev, prev)
        h.synchronize()
        wall_loop = time.perf_counter() - t0
        log(f"download loop {wall_loop:.3f} s")
    value = st_snap.ion_steps / wall_snap
    launches_plain = math.ceil(steps / 2000)
    avg_launch_s = st_plain.device_ms * 1e-3 / launches_plain
    achieved = FLOPS_PER_ION_STEP["fp32"] * (st_plain.ion_steps / launches_plain) / avg_launch_s / 1e12
    snap_bytes = 57 * n
    return {
        "metric": "ion-steps/s on 1 MI355X with trajectory snapshots",
        "value": value, "unit": "ion-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": wall_snap * 1e3 / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (Kahan-compensated state; f64 time and RF phase)", "data": "synthetic",
        "config": {"workload": "rf_quadrupole_boris_langevin_ou_snapshots", "ions": n, "U_dc": args.quad_u,
                   "dt_s": dt, "engine_steps": steps, "write_every_steps": args.write_every,
                   "snapshots": len(snap_steps), "snapshot_bytes": snap_bytes},
        "no_snapshots": {"value": st_plain.ion_steps / wall_plain, "wall_s": wall_plain},
        "noop_consumer": {"value": st_noop.ion_steps / wall_noop, "wall_s": wall_noop,
                          "overhead": wall_noop / wall_plain - 1.0},
        "download_loop": {"value": st_plain.ion_steps / wall_loop, "wall_s": wall_loop},
        "snapshot_overhead": wall_snap / wall_plain - 1.0,
        "download_loop_overhead": wall_loop / wall_plain - 1.0,
        "snapshot_rate_GBps": len(snap_steps) * snap_bytes / wall_snap / 1e9,
        "alive_fraction_end": st_snap.alive / n,
        "roofline": {
            "bound": "mfma", "achieved": achieved, "peak": MI355X_PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / MI355X_PEAK_FP32_TFLOPS, "traffic": None,
            "kernel": "step_f32_kernel<Boris, linear field, damping> (run without snapshots)",
        },
    }


def bench_domains(args):
    """The three-domain instrument (IMS region -> RF quadrupole -> IMS region,
    HSS in each domain's gas, RK4 dt 1 ns) with 1M ions in FP64: ion-steps/s of
    the multi-domain kernel (a domain lookup per ion and step, the domain's
    field and gas per lane) and the oracle's multi-domain loop on the host;
    the FP32 + Kahan kernel on the same run is reported beside it."""
    from icarion_amd import _abi, gpu
    from icarion_amd import workloads as W
    steps = args.steps if args.steps is not None else 10_000
    w = W.multi_domain_workload(n=args.ions, collision=_abi.COLLISION_HSS)
    log(f"domains: {args.ions} ions x {steps} steps, 3 domains")
    ens = w["ensemble"].copy()
    with gpu.Handle(device=0, steps_per_launch=args.steps_per_launch) as h:
        h.set_domains(w["domains"], w["physics"], w["apertures"])
        h.set_species(w["species"])
        st, wall = timed_steps(h, ens, w["dt"], steps, args.warmup)
        h.download_state(ens)
        dom = h.download_domains(ens.n)
    value = st.ion_steps / wall
    launches = math.ceil(steps / (args.steps_per_launch or 2000))
    res = {
        "metric": "ion-steps/s on 1 MI355X, three-domain instrument (FP64)",
        "value": value, "unit": "ion-steps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "three_domains_ims_quadrupole_ims_hss", "ions": args.ions,
                   "dt_s": w["dt"], "engine_steps": steps},
        "collisions": st.collisions, "alive_fraction": st.alive / args.ions,
        "alive_by_domain": np.bincount(dom[(ens.active == 1) & (ens.born == 1)], minlength=3).tolist(),
        "avg_launch_ms": st.device_ms / launches,
    }
    # the same run with FP32 + Kahan numerics (domain decisions in FP64)
    ph32 = W.physics(_abi.INTEGRATOR_RK4, _abi.COLLISION_HSS, precision=_abi.PRECISION_FP32_KAHAN,
                     seed=w["physics"].rng_seed)
    e32 = w["ensemble"].copy()
    with gpu.Handle(device=0, steps_per_launch=args.steps_per_launch) as h:
        h.set_domains(w["domains"], ph32, w["apertures"])
        h.set_species(w["species"])
        st32, wall32 = timed_steps(h, e32, w["dt"], steps, args.warmup)
    res["fp32_kahan"] = {"value": st32.ion_steps / wall32, "unit": "ion-steps/s",
                         "ms_per_step": wall32 * 1e3 / steps, "alive_fraction": st32.alive / args.ions,
                         "avg_launch_ms": st32.device_ms / launches}
    log(f"domains: FP64 {value:.3e}, FP32 + Kahan {st32.ion_steps / wall32:.3e} ion-steps/s")
    if not args.no_cpu:
        from oracle import binding
        threads = len(os.sched_getaffinity(0))
        n_cpu = min(args.cpu_sample_ions or 200_000, args.ions)
        k = min(args.cpu_sample_steps or 5000, steps)
        e = w["ensemble"].subset(np.arange(n_cpu))
        di = np.full(n_cpu, -1, np.int32)
        t0 = time.perf_counter()
        so = binding.run_multi(w["domains"], w["physics"], w["species"], e, di, 0.0, w["dt"], k,
                               end_apertures=w["apertures"], threads=threads)
        el = time.perf_counter() - t0
        res["cpu_baseline"] = dict(value=so.ion_steps / el, unit="ion-steps/s", cores=threads, kind="port",
                                   sample=f"{n_cpu} ions x {k} steps of the same workload (FP64 oracle, "
                                          f"OpenMP over {threads} threads, {el:.1f} s)")
    return res


def main():
    args = parse()
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); "
              "multi-GPU sharding is not implemented", file=sys.stderr)
        sys.exit(2)
    from icarion_amd import gpu
    if gpu.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    res = {"ehss": bench_ehss, "quadrupole": bench_quadrupole, "batched": bench_batched,
           "batched_ehss": bench_batched_ehss, "field_array": bench_field_array,
           "ims": bench_ims, "snapshots": bench_snapshots, "domains": bench_domains}[args.workload](args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
