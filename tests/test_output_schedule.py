"""Trajectory sampling schedule and snapshot buffer (CPU).

The schedule restatement (oracle/sampling.py) is pinned against the
reference's own OutputManager tests (tests/integrator/test_output_manager.cpp:133-188).
It is also pinned against the closed form for a dt that is exact in binary: steps
0, W, 2W, ... when the interval is W steps.
"""
import numpy as np
import pytest

from icarion_amd.trajectory import TrajectoryBuffer
from oracle import sampling


def test_time_trigger_known_answers():
    # test_output_manager.cpp:133-163 ("Time-based write trigger")
    m = sampling.OutputSchedule(1e-4)
    m.log_step(0.0)
    assert not m.should_write(5e-5)
    assert m.should_write(1.5e-4)
    m.log_step(1.5e-4)
    assert not m.should_write(2.0e-4)
    assert m.should_write(2.5e-4)


def test_full_buffer_known_answers():
    # test_output_manager.cpp:165-188: the sample is anchored to the logged time
    m = sampling.OutputSchedule(1e-4)
    m.log_step(0.0)
    assert m.should_write(1e-4)
    m.log_step(1e-4)
    assert not m.should_write(1.5e-4)
    assert m.should_write(2e-4)


@pytest.mark.parametrize("W", [1, 3, 10, 64])
def test_exact_dt_closed_form(W):
    dt = 2.0 ** -20
    snaps, nxt = sampling.snapshot_steps(0.0, dt, 200, W * dt)
    steps = [s for s, _ in snaps]
    # step 0 always; then every step whose end time reaches the last sample + W dt
    assert steps == [0] + list(range(W, 200, W))
    for s, t in snaps:
        assert t == (s + 1) * dt
    assert nxt == (steps[-1] + 1) * dt + W * dt


def test_first_flag_and_continuation():
    dt = 1.7e-9
    one, n1 = sampling.snapshot_steps(0.0, dt, 300, 23.5 * dt)
    a, na = sampling.snapshot_steps(0.0, dt, 170, 23.5 * dt)
    t = 0.0
    for _ in range(170):
        t = t + dt
    b, nb = sampling.snapshot_steps(t, dt, 130, 23.5 * dt, step_offset=170, first=False,
                                    next_write=na)
    assert a + b == one and nb == n1
    # without the first-step rule the first sample is the first step ending at or
    # after the interval: t = 24 dt >= 23.5 dt, step 23
    c, _ = sampling.snapshot_steps(0.0, dt, 300, 23.5 * dt, first=False)
    assert c[0][0] == 23
    # with it, step 0 is logged at dt, so the next is the first t >= 24.5 dt
    assert one[0][0] == 0 and one[1][0] == 24


def test_schedule_not_step_modulo():
    # the interval is W dt in the reference's config (SimulationEngine.cpp:109), but
    # the schedule is time based: with dt = 0.1 the accumulated time drifts, so the
    # snapshot steps are not every W-th step on the nose
    dt = 0.1
    snaps, _ = sampling.snapshot_steps(0.0, dt, 100, 3 * dt)
    steps = [s for s, _ in snaps]
    t = 0.0
    nxt = 3 * dt
    expect = []
    for k in range(100):
        t = t + dt
        if k == 0 or t >= nxt:
            expect.append(k)
            nxt = t + 3 * dt
    assert steps == expect
    assert steps != [0] + list(range(3, 100, 3))


def _snap(t, n, seed):
    r = np.random.default_rng(seed)
    return {"index": 0, "step": 0, "t": t, "positions": r.normal(size=(n, 3)),
            "velocities": r.normal(size=(n, 3)), "species": r.integers(0, 3, n).astype(np.uint32),
            "flags": np.ones(n, np.uint8)}


def test_trajectory_buffer_layout_and_flush(tmp_path):
    batches = []
    buf = TrajectoryBuffer(buffer_max=5, sink=batches.append)
    snaps = [_snap(k * 1e-6, 7, k) for k in range(6)]
    for s in snaps[:5]:
        buf.add(s)
    assert batches == []            # buffer reaches buffer_max before a flush
    buf.add(snaps[5])               # the next snapshot flushes the five first
    assert len(batches) == 1 and batches[0]["time"].shape == (5,)
    b = batches[0]
    assert b["positions"].shape == (5, 7, 3) and b["species_id_indices"].dtype == np.uint32
    assert b["domain_indices"].dtype == np.int32 and not b["domain_indices"].any()
    np.testing.assert_array_equal(b["positions"][2], snaps[2]["positions"])
    buf.flush()
    assert len(batches) == 2 and batches[1]["time"].tolist() == [snaps[5]["t"]]
    with pytest.raises(RuntimeError):
        buf.add(_snap(1.0, 8, 0))


def test_trajectory_buffer_npz(tmp_path):
    buf = TrajectoryBuffer(buffer_max=2)
    snaps = [_snap(k * 1e-6, 4, k) for k in range(5)]
    for s in snaps:
        buf(s)
    p = tmp_path / "traj.npz"
    buf.save_npz(p)
    with np.load(p) as z:
        assert sorted(z.files) == sorted("trajectory/" + k for k in (
            "time", "positions", "velocities", "domain_indices", "species_id_indices"))
        np.testing.assert_array_equal(z["trajectory/time"], [s["t"] for s in snaps])
        np.testing.assert_array_equal(z["trajectory/velocities"][4], snaps[4]["velocities"])


def test_trajectory_buffer_keeps_copies_of_transient_snapshots():
    """A consumer of step_snapshots(copy=False) receives views of buffers the
    library reuses for the next snapshot; the buffer must hold its own copies."""
    buf = TrajectoryBuffer(buffer_max=10)
    scratch = _snap(0.0, 6, 1)
    first = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in scratch.items()}
    buf(scratch)
    for k in ("positions", "velocities", "species"):
        scratch[k][...] = 0                       # the library overwrites its buffer
    d = buf.datasets()
    np.testing.assert_array_equal(d["positions"][0], first["positions"])
    np.testing.assert_array_equal(d["velocities"][0], first["velocities"])
    np.testing.assert_array_equal(d["species_id_indices"][0], first["species"])
