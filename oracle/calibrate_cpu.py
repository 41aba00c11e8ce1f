"""Calibration of the CPU oracle against the reference's own compiled hot-path
code (oracle/_ref/ref_harness, built from the reference sources by
`make ref`): the same collisionless RF-quadrupole push, one thread each.

  reference: RK4Strategy::step_batch (RK4Strategy.cpp:172-361) or
             BorisStrategy::step (BorisStrategy.cpp:18-98) over an IonEnsemble,
             with the analytic quadrupole field as a registry force;
  oracle:    oracle_run's engine loop (push + boundary test per ion-step).

Prints ion-steps/s of each and their ratio.  Test infrastructure: it needs the
harness, so it runs where the reference sources are, not on the GPU host.

usage: python oracle/calibrate_cpu.py [n_ions] [n_steps]
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from icarion_amd import _abi  # noqa: E402
from icarion_amd import workloads as W  # noqa: E402
from oracle import binding  # noqa: E402


def harness_input(w, integrator, n_steps):
    d = w["domain"]
    e = w["ensemble"]
    head = (f"push {integrator} {e.n} {n_steps} {w['dt']!r} 0.0 {d.radius_m!r} {d.rf_voltage_V!r} "
            f"{d.rf_frequency_Hz!r} {d.dc_quad_V!r} {d.dc_axial_V!r} {d.length_m!r}")
    rows = []
    for i in range(e.n):
        rows.append(" ".join(repr(float(v)) for v in (e.x[i], e.y[i], e.z[i], e.vx[i], e.vy[i], e.vz[i],
                                                        e.mass_kg[i], e.charge_C[i])))
    return head + "\n" + "\n".join(rows) + "\n"


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 2000
    n_steps = int(sys.argv[2]) if len(sys.argv) > 2 else 1000
    if not os.path.exists(binding.REF_HARNESS):
        sys.exit("oracle/_ref/ref_harness not built (make -C oracle ref)")
    for integrator, name in ((_abi.INTEGRATOR_RK4, "RK4"), (_abi.INTEGRATOR_BORIS, "Boris")):
        # 1000 Da at U = 0 (Mathieu q ~ 0.3): every ion stays inside, so both
        # sides push every ion at every step
        w = W.quadrupole_workload(n=n, U=0.0, seed=7, ou=0)
        w["ensemble"].mass_kg[:] = 1000.0 * 1.66053906660e-27
        w["physics"] = W.physics(integrator, _abi.COLLISION_NONE, seed=7)
        text = harness_input(w, integrator, n_steps)
        t0 = time.perf_counter()
        subprocess.run([binding.REF_HARNESS], input=text, capture_output=True, text=True, check=True)
        t_ref = time.perf_counter() - t0
        e = w["ensemble"].copy()
        t0 = time.perf_counter()
        st = binding.run(w["domain"], w["physics"], w["species"], e, 0.0, w["dt"], n_steps, threads=1)
        t_or = time.perf_counter() - t0
        if st.ion_steps != n * n_steps:
            sys.exit(f"{name}: {n * n_steps - st.ion_steps} ion-steps lost to the boundary")
        ref_rate = n * n_steps / t_ref
        or_rate = st.ion_steps / t_or
        print(f"{name}: reference {ref_rate:.3e} ion-steps/s ({t_ref:.2f} s), oracle {or_rate:.3e} "
              f"ion-steps/s ({t_or:.2f} s), oracle / reference {or_rate / ref_rate:.2f}", flush=True)


if __name__ == "__main__":
    main()
