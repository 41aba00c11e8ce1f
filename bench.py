"""Benchmark of the batch ion-step path on one MI355X.

Workload (one GPU, the quadrupole configuration): 1M ions, masses log-uniform
over 100-2000 Da, z = +1, RF quadrupole r0 = 4 mm, V = 500 V at 1 MHz, U = 0,
Langevin damping + OU thermal kicks in He at 1 Pa / 300 K, Gaussian cloud
(sigma 0.2 mm), Maxwellian 300 K velocities, Boris push, dt = 1e-9 s.

A "step" is one engine time step of the whole cloud.  The ensemble is uploaded
before the timed region (inputs resident in HBM); the timed region is exactly
--steps engine steps, bracketed by device synchronisation
(icarion_gpu_synchronize: stream + device).  torch is not loaded: it bundles a
second HIP/HSA runtime that cannot share the device with the library's in one
process.  Both arithmetic
paths run: the plain FP64 method and the FP32 + Kahan fast path (the headline
value); their transmission per m/z bin is compared.  The CPU baseline is the
oracle (same algorithm, FP64) on the host cores, timed on a bounded sample.

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

# Algorithmic FLOPs per ion-step (FMA = 2, add/mul/div = 1; sqrt, log, exp,
# sin/cos and the integer Philox rounds not counted), counted from
# icarion_amd/csrc/step_kernels.hip for the Boris + quadrupole RF field +
# Langevin damping + OU kick kernels: force 26, Boris update 18, OU kick 25,
# uniforms 8, cylinder test 9 (FP64); the FP32 path adds 36 for the Kahan
# compensation of six components and the kick.
FLOPS_PER_ION_STEP = {"fp64": 86, "fp32": 122}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100_000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--ions", type=int, default=1_000_000)
    ap.add_argument("--steps-per-launch", type=int, default=0)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--cpu-sample-ions", type=int, default=100_000)
    ap.add_argument("--cpu-sample-steps", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter "
                         "calibration for rocprofv3 --pmc); loads torch, not the library")
    return ap.parse_args()


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


def run_path(w, precision, steps, warmup, spl):
    from icarion_amd import gpu
    from icarion_amd import workloads as W
    ph = W.physics(w["physics"].integrator, w["physics"].collision_model,
                   w["physics"].ou_thermal_kicks, precision, w["physics"].rng_seed)
    ens = w["ensemble"].copy()
    with gpu.Handle(device=0, steps_per_launch=spl) as h:
        h.set_domain(w["domain"], ph)
        h.set_species(w["species"])
        if warmup:
            h.upload_state(ens)
            h.step(0.0, w["dt"], warmup, 0)
        h.upload_state(ens)  # timed run starts from the initial state
        h.synchronize()
        t0 = time.perf_counter()
        st = h.step(0.0, w["dt"], steps, 0)
        h.synchronize()
        wall = time.perf_counter() - t0
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


def cpu_baseline(w, n_ions, n_steps):
    from oracle import binding
    threads = len(os.sched_getaffinity(0))
    sub = dict(w)
    e = w["ensemble"].subset(np.arange(n_ions))
    t0 = time.perf_counter()
    st = binding.run(w["domain"], w["physics"], w["species"], e, 0.0, w["dt"], n_steps,
                     threads=threads)
    el = time.perf_counter() - t0
    del sub
    return dict(value=st.ion_steps / el, unit="ion-steps/s", cores=threads, kind="port",
                sample=f"{n_ions} ions x {n_steps} steps of the same workload "
                       f"(FP64 oracle, OpenMP over {threads} threads, {el:.1f} s)")


def main():
    args = parse()
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); "
              "multi-GPU sharding is not part of this build", file=sys.stderr)
        sys.exit(2)
    from icarion_amd import _abi, gpu
    from icarion_amd import workloads as W
    if gpu.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    w = W.quadrupole_workload(n=args.ions, n_steps=args.steps)
    masses = w["ensemble"].mass_kg / 1.66053906660e-27
    print(f"bench: {w['name']} {args.ions} ions x {args.steps} steps", file=sys.stderr,
          flush=True)
    fast = run_path(w, _abi.PRECISION_FP32_KAHAN, args.steps, args.warmup, args.steps_per_launch)
    print(f"bench: fp32+kahan {fast['wall']:.3f} s", file=sys.stderr, flush=True)
    plain = None
    if not args.skip_fp64:
        plain = run_path(w, _abi.PRECISION_FP64, args.steps, args.warmup, args.steps_per_launch)
        print(f"bench: fp64 {plain['wall']:.3f} s", file=sys.stderr, flush=True)

    fs = fast["stats"]
    value = fs.ion_steps / fast["wall"]
    avg_launch_s = fs.device_ms * 1e-3 / fast["launches"]
    achieved = FLOPS_PER_ION_STEP["fp32"] * (fs.ion_steps / fast["launches"]) / avg_launch_s / 1e12
    res = {
        "metric": "ion-steps/s on 1 MI355X, plain FP64 vs FP32+surrogate at <=0.5% K error",
        "value": value,
        "unit": "ion-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fast["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (Kahan-compensated state; f64 time and RF phase)",
        "data": "synthetic",
        "config": {
            "workload": "rf_quadrupole_boris_langevin_ou",
            "ions": args.ions, "dt_s": w["dt"], "engine_steps": args.steps,
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
    if plain is not None:
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
        n_cpu = min(args.cpu_sample_ions, args.ions)
        res["cpu_baseline"] = cpu_baseline(w, n_cpu, min(args.cpu_sample_steps, args.steps))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
