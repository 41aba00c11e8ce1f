"""Trajectory snapshots taken while stepping on the device (icarion_gpu_step_snapshots).

The snapshot steps and times must equal the engine's schedule (oracle/sampling.py)
exactly. Each snapshot must hold the state after its step:
  * bit for bit, the state a plain icarion_gpu_step run reaches over the same
    segments (FP64 and FP32);
  * for FP64, the oracle's state after the same steps, to the parity bar of the
    plain path: 1e-9 relative; on a stochastic run, >= 99 % of ions within it.
"""
import numpy as np
import pytest

from icarion_amd import _abi
from icarion_amd import workloads as W
from oracle import sampling
from test_gpu_parity import gpu_run, oracle_run

pytestmark = pytest.mark.gpu


def _handle(w, phys=None):
    from icarion_amd import gpu
    h = gpu.Handle()
    h.set_domain(w["domain"], phys or w["physics"])
    h.set_species(w["species"], w.get("atoms"))
    return h


def _fp32(w):
    p = w["physics"]
    return W.physics(p.integrator, p.collision_model, p.ou_thermal_kicks,
                     _abi.PRECISION_FP32_KAHAN, p.rng_seed, p.ehss_mode)


def _close(a, b, rtol, scale):
    return np.abs(a - b) <= rtol * np.maximum(np.abs(b), scale)


def _ens_match(s, e, rtol=1e-9):
    pos = np.stack([e.x, e.y, e.z], axis=1)
    vel = np.stack([e.vx, e.vy, e.vz], axis=1)
    ok = np.all(_close(s["positions"], pos, rtol, 1e-6), axis=1)
    ok &= np.all(_close(s["velocities"], vel, rtol, 1.0), axis=1)
    ok &= (s["flags"] & 1).astype(bool) == e.active.astype(bool)
    return ok


@pytest.mark.parametrize("case", ["quadrupole", "ims_hss"])
def test_snapshots_match_schedule_and_oracle(oracle, case):
    if case == "quadrupole":
        w = W.quadrupole_workload(n=600, seed=4)
        need = 1.0
    else:
        w = W.ims_hss_workload(n=600, seed=9)
        need = 0.99
    dt, n_steps = w["dt"], 230
    interval = 37.3 * dt
    print(f"snapshots {case}: n=600 steps={n_steps}", flush=True)
    expect, nxt = sampling.snapshot_steps(0.0, dt, n_steps, interval)
    h = _handle(w)
    try:
        g = w["ensemble"].copy()
        h.upload_state(g)
        snaps, nxt_g, st = h.step_snapshots(0.0, dt, n_steps, interval)
        h.download_state(g)
    finally:
        h.close()
    assert [(s["step"], s["t"]) for s in snaps] == expect
    assert [s["index"] for s in snaps] == list(range(len(expect)))
    assert nxt_g == nxt
    assert st.ion_steps > 0
    np.testing.assert_array_equal(snaps[0]["species"], w["ensemble"].species)
    # the oracle, stopped at every snapshot step
    o = w["ensemble"].copy()
    t, prev = 0.0, 0
    for s in snaps:
        oracle_steps = s["step"] + 1 - prev
        oracle_run(oracle, w, o, oracle_steps, t0=t, step_offset=prev)
        frac = _ens_match(s, o).mean()
        assert frac >= need, (s["index"], frac)
        t, prev = s["t"], s["step"] + 1
    # the run's end state is the plain path's
    ref = w["ensemble"].copy()
    gpu_run(w, ref, n_steps)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(getattr(g, f), getattr(ref, f))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_snapshots_equal_segmented_steps(precision):
    w = W.ims_hss_workload(n=3000, seed=12)
    phys = _fp32(w) if precision == "fp32" else w["physics"]
    dt, n_steps = w["dt"], 300
    interval = 50 * dt
    print(f"snapshots vs segments {precision}", flush=True)
    h = _handle(w, phys)
    try:
        h.upload_state(w["ensemble"].copy())
        snaps, _, _ = h.step_snapshots(0.0, dt, n_steps, interval)
    finally:
        h.close()
    h = _handle(w, phys)
    try:
        e = w["ensemble"].copy()
        h.upload_state(e)
        t, prev = 0.0, 0
        for s in snaps:
            st = h.step(t, dt, s["step"] + 1 - prev, prev)
            assert st.t_end_s == s["t"]
            h.download_state(e)
            np.testing.assert_array_equal(s["positions"], np.stack([e.x, e.y, e.z], axis=1))
            np.testing.assert_array_equal(s["velocities"], np.stack([e.vx, e.vy, e.vz], axis=1))
            np.testing.assert_array_equal(s["flags"] & 1, e.active)
            t, prev = s["t"], s["step"] + 1
    finally:
        h.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_snapshots_with_deaths_and_births(precision):
    """Every ion leaves a quadrupole whose DC term (U = 200 V, 2U/V = 0.8) makes
    all masses unstable, and a quarter are born during the run: after the first
    snapshot only changed ions cross PCIe, and every snapshot must still equal
    the full state of segmented stepping."""
    w = W.quadrupole_workload(n=4000, U=200.0, seed=8)
    phys = _fp32(w) if precision == "fp32" else w["physics"]
    e0 = w["ensemble"].copy()
    rng = np.random.default_rng(3)
    late = rng.random(e0.n) < 0.25
    e0.born[late] = 0
    e0.active[late] = 0
    e0.birth_time_s = np.where(late, rng.uniform(0.0, 2e-6, e0.n), 0.0)
    dt, n_steps = w["dt"], 10000
    print(f"snapshots with deaths and births {precision}", flush=True)
    h = _handle(w, phys)
    try:
        h.upload_state(e0.copy())
        snaps, _, _ = h.step_snapshots(0.0, dt, n_steps, 100 * dt)
    finally:
        h.close()
    alive = [(s["flags"] & 3 == 3).mean() for s in snaps]
    assert alive[-1] < 1 / 16, alive[-1]      # the late snapshots take the compact branch
    assert any(((s["flags"] & 2) == 0).any() for s in snaps[1:])   # births under way
    h = _handle(w, phys)
    try:
        e = e0.copy()
        h.upload_state(e)
        t, prev = 0.0, 0
        for s in snaps:
            st = h.step(t, dt, s["step"] + 1 - prev, prev)
            h.download_state(e)
            np.testing.assert_array_equal(s["positions"], np.stack([e.x, e.y, e.z], axis=1))
            np.testing.assert_array_equal(s["velocities"], np.stack([e.vx, e.vy, e.vz], axis=1))
            np.testing.assert_array_equal(s["flags"] & 1, e.active)
            np.testing.assert_array_equal((s["flags"] >> 1) & 1, e.born)
            np.testing.assert_array_equal(s["species"], e.species)
            assert not s["domain"].any()   # one domain: index 0 throughout
            t, prev = st.t_end_s, s["step"] + 1
    finally:
        h.close()


def test_snapshots_continue_across_calls():
    w = W.quadrupole_workload(n=500, seed=6)
    dt = w["dt"]
    interval = 23.5 * dt
    h = _handle(w)
    try:
        h.upload_state(w["ensemble"].copy())
        one, n1, st1 = h.step_snapshots(0.0, dt, 300, interval)
    finally:
        h.close()
    h = _handle(w)
    try:
        h.upload_state(w["ensemble"].copy())
        a, na, sta = h.step_snapshots(0.0, dt, 170, interval)
        b, nb, _ = h.step_snapshots(sta.t_end_s, dt, 130, interval, step_offset=170,
                                    write_first=False, next_write=na)
    finally:
        h.close()
    both = a + b
    assert [(s["step"], s["t"]) for s in both] == [(s["step"], s["t"]) for s in one]
    assert nb == n1
    for x, y in zip(both, one):
        np.testing.assert_array_equal(x["positions"], y["positions"])
        np.testing.assert_array_equal(x["velocities"], y["velocities"])


def test_snapshot_callback_stop_and_errors():
    from icarion_amd import gpu
    w = W.quadrupole_workload(n=200, seed=1)
    dt = w["dt"]
    seen = []
    h = _handle(w)
    try:
        h.upload_state(w["ensemble"].copy())
        with pytest.raises(gpu.IcarionError) as ei:
            h.step_snapshots(0.0, dt, 400, 10 * dt, on_snapshot=seen.append, stop_after=2)
        assert ei.value.code == _abi.ERR_STATE
        assert [s["index"] for s in seen] == [0, 1, 2]
        with pytest.raises(gpu.IcarionError) as ei:
            h.step_snapshots(0.0, dt, 10, 0.0)
        assert ei.value.code == _abi.ERR_INVALID
        # the handle still steps normally afterwards
        st = h.step(0.0, dt, 10, 0)
        assert st.ion_steps > 0
    finally:
        h.close()


def test_snapshots_without_ions():
    w = W.quadrupole_workload(n=10, seed=1)
    empty = w["ensemble"].copy().subset(np.zeros(0, dtype=np.int64))
    dt = w["dt"]
    expect, nxt = sampling.snapshot_steps(0.0, dt, 100, 7 * dt)
    h = _handle(w)
    try:
        h.upload_state(empty)
        snaps, nxt_g, _ = h.step_snapshots(0.0, dt, 100, 7 * dt)
    finally:
        h.close()
    assert [(s["step"], s["t"]) for s in snaps] == expect and nxt_g == nxt
    assert all(s["positions"].shape == (0, 3) for s in snaps)


def test_snapshots_reject_rk45():
    from icarion_amd import gpu
    w = W.quadrupole_workload(n=100, seed=1)
    p = w["physics"]
    phys = W.physics(_abi.INTEGRATOR_RK45, p.collision_model, p.ou_thermal_kicks, p.precision,
                     p.rng_seed, p.ehss_mode)
    h = _handle(w, phys)
    try:
        h.upload_state(w["ensemble"].copy())
        with pytest.raises(gpu.IcarionError) as ei:
            h.step_snapshots(0.0, w["dt"], 10, 2 * w["dt"])
        assert ei.value.code == _abi.ERR_UNSUPPORTED
    finally:
        h.close()


def test_snapshots_into_the_trajectory_file(tmp_path):
    """The device snapshot path writing the reference's trajectory file: the
    library's C callback (icarion_h5_on_snapshot) handed straight to
    icarion_gpu_step_snapshots, and the Python consumer, give the same
    /trajectory datasets as the in-memory TrajectoryBuffer of the same run;
    /ions/death_time_s takes the final death times."""
    import ctypes as C
    import json
    import os
    import shutil
    import subprocess
    from icarion_amd import build as B
    from icarion_amd import trajectory as TR
    if B.hdf5_dir() is None or not os.path.exists(TR.H5_LIB_PATH):
        pytest.skip("no HDF5 C library / libicarion_h5.so")
    d5 = B.hdf5_dir()
    dump = str(tmp_path / "h5_dump")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run([shutil.which("g++") or "g++", "-O1", "-std=c++17", "-I", os.path.join(d5, "include"),
                    os.path.join(root, "tests", "cpp", "h5_dump.cpp"), "-L", os.path.join(d5, "lib"),
                    "-lhdf5", f"-Wl,-rpath,{os.path.join(d5, 'lib')}", "-o", dump], check=True)
    # U = 200 V makes every mass unstable: ions die during the run, so the
    # death times written at the end are not all -1
    w = W.quadrupole_workload(n=700, U=200.0, seed=6)
    dt, n_steps, interval = w["dt"], 2000, 100 * w["dt"]
    print("trajectory file: quadrupole n=700 U=200 steps=2000", flush=True)
    # (1) Python consumers: a TrajectoryBuffer and the file, fed the same snapshots
    buf = TR.TrajectoryBuffer(buffer_max=4)
    e1 = w["ensemble"].copy()
    p1 = str(tmp_path / "py.h5")
    with _handle(w) as h, TR.H5TrajectoryFile(p1, e1, batch=4) as f:
        h.upload_state(e1)

        def both(s):
            buf(s)
            f(s)
        h.step_snapshots(0.0, dt, n_steps, interval, on_snapshot=both, copy=False)
        h.download_state(e1)
        f.death_times(e1.death_time_s)
    # (2) the C callback of the writer passed to the library directly
    e2 = w["ensemble"].copy()
    p2 = str(tmp_path / "c.h5")
    lib5 = TR.load_h5()
    f2 = TR.H5TrajectoryFile(p2, e2, batch=4)
    fn = C.cast(lib5.icarion_h5_on_snapshot, _abi.SnapshotFn)
    with _handle(w) as h:
        h.upload_state(e2)
        nxt = C.c_double(interval)
        st = _abi.StepStats()
        rc = h.lib.icarion_gpu_step_snapshots(h.h, 0.0, dt, n_steps, 0, interval, 1, C.byref(nxt), fn,
                                              f2.h, C.byref(st))
        assert rc == 0
        h.download_state(e2)
    f2.death_times(e2.death_time_s)
    f2.close()
    ref = buf.datasets()
    assert ref["time"].size >= 10 and (e1.active == 0).any()
    for p in (p1, p2):
        d = json.loads(subprocess.run([dump, p], check=True, capture_output=True, text=True).stdout)
        for k in ("time", "positions", "velocities", "domain_indices", "species_id_indices"):
            np.testing.assert_array_equal(np.array(d["/trajectory/" + k]["data"]).reshape(ref[k].shape),
                                          ref[k], err_msg=f"{p} {k}")
        np.testing.assert_array_equal(d["/ions/death_time_s"]["data"], e1.death_time_s)
        np.testing.assert_array_equal(d["/ions/initial_pos_x"]["data"], w["ensemble"].x)
