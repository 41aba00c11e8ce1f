"""The trajectory file writer (libicarion_h5.so, include/icarion_h5.h) on the
CPU: synthetic snapshots written through H5TrajectoryFile and read back with a
small HDF5 C reader (tests/cpp/h5_dump.cpp).  Checks the reference's layout
(src/core/io/hdf5Writer.cpp: /ions per write_ion_metadata :1656-1730,
/trajectory per append_trajectory_batch_flat :530-709, update_death_times
:711-729): dataset names, shapes, element types, chunk shapes
(min(first batch, 100) along T; time min(first batch, 1000)), the deflate level,
and every value, across several batches and a final partial one."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from icarion_amd import build as B
from icarion_amd import trajectory as TR
from icarion_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(B.hdf5_dir() is None, reason="no HDF5 C library")


@pytest.fixture(scope="module")
def h5_dump(tmp_path_factory):
    if not os.path.exists(TR.H5_LIB_PATH):
        B.build_h5()
    d = B.hdf5_dir()
    exe = str(tmp_path_factory.mktemp("h5") / "h5_dump")
    subprocess.run([shutil.which("g++") or "g++", "-O1", "-std=c++17", "-I", os.path.join(d, "include"),
                    os.path.join(ROOT, "tests", "cpp", "h5_dump.cpp"), "-L", os.path.join(d, "lib"),
                    "-lhdf5", f"-Wl,-rpath,{os.path.join(d, 'lib')}", "-o", exe], check=True)

    def dump(path):
        return json.loads(subprocess.run([exe, path], check=True, capture_output=True, text=True).stdout)
    return dump


def snapshots(ens, T, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(T):
        out.append(dict(t=1e-9 * (k + 1), positions=rng.normal(size=(ens.n, 3)),
                        velocities=rng.normal(size=(ens.n, 3)),
                        species=rng.integers(0, 3, ens.n).astype(np.uint32),
                        domain=rng.integers(-1, 2, ens.n).astype(np.int32)))
    return out


def test_layout_and_values_over_several_batches(tmp_path, h5_dump):
    ens = W.gaussian_cloud(37, np.array([100.0, 200.0, 300.0])[np.arange(37) % 3], 1e-3, seed=3)
    ens.species[:] = np.arange(37) % 3
    ens.birth_time_s = np.linspace(0, 1e-6, 37)
    pool = ["H3O+", "26DTBPH+", "CaffeineH+"]
    snaps = snapshots(ens, 23, seed=4)
    path = str(tmp_path / "traj.h5")
    with TR.H5TrajectoryFile(path, ens, species_pool=pool, deflate=2, batch=10) as f:
        for s in snaps:
            f(s)
        assert f.rows() == 23
        death = np.where(np.arange(37) % 4 == 0, 5e-7, -1.0)
        f.death_times(death)
    d = h5_dump(path)
    n, T = 37, 23
    # /ions
    for k, arr in (("initial_pos_x", ens.x), ("initial_pos_y", ens.y), ("initial_pos_z", ens.z),
                   ("initial_vel_x", ens.vx), ("initial_vel_y", ens.vy), ("initial_vel_z", ens.vz),
                   ("charge_C", ens.charge_C), ("birth_time_s", ens.birth_time_s),
                   ("death_time_s", death)):
        ds = d["/ions/" + k]
        assert ds["shape"] == [n] and ds["class"] == "float" and ds["size"] == 8, k
        np.testing.assert_array_equal(np.array(ds["data"]), arr, err_msg=k)
    assert d["/ions/initial_species_id"]["data"] == [pool[s] for s in ens.species]
    # /trajectory
    tr = {k[len("/trajectory/"):]: v for k, v in d.items() if k.startswith("/trajectory/")}
    assert set(tr) == {"time", "positions", "velocities", "domain_indices", "species_id_indices",
                       "species_ids"}
    assert tr["time"]["shape"] == [T] and tr["time"]["chunk"] == [10]
    np.testing.assert_array_equal(tr["time"]["data"], [s["t"] for s in snaps])
    for k in ("positions", "velocities"):
        assert tr[k]["shape"] == [T, n, 3] and tr[k]["chunk"] == [10, n, 3] and tr[k]["size"] == 8
        np.testing.assert_array_equal(np.array(tr[k]["data"]).reshape(T, n, 3),
                                      np.stack([s[k] for s in snaps]))
    assert tr["domain_indices"]["class"] == "int" and tr["domain_indices"]["size"] == 4
    assert tr["domain_indices"]["signed"] == 1
    np.testing.assert_array_equal(np.array(tr["domain_indices"]["data"]).reshape(T, n),
                                  np.stack([s["domain"] for s in snaps]))
    assert tr["species_id_indices"]["signed"] == 0 and tr["species_id_indices"]["chunk"] == [10, n]
    np.testing.assert_array_equal(np.array(tr["species_id_indices"]["data"]).reshape(T, n),
                                  np.stack([s["species"] for s in snaps]))
    assert tr["species_ids"]["data"] == [pool[i] for s in snaps for i in s["species"]]
    for v in tr.values():
        assert v["deflate"] == 2


def test_defaults_time_per_ion_and_errors(tmp_path, h5_dump):
    ens = W.gaussian_cloud(5, 100.0, 1e-3, seed=1)
    path = str(tmp_path / "t.h5")
    f = TR.H5TrajectoryFile(path, ens, deflate=0)
    for k in range(3):
        f.append(1e-9 * k, np.zeros((5, 3)) + k, np.ones((5, 3)), np.zeros(5, np.uint32),
                 time_per_ion=np.full(5, 1e-9 * k))
    with pytest.raises(RuntimeError, match="time_per_ion"):
        f.append(1.0, np.zeros((5, 3)), np.zeros((5, 3)), np.zeros(5, np.uint32))
    with pytest.raises(RuntimeError, match="one per ion"):
        f.death_times(np.zeros(4))
    f.close()
    d = h5_dump(path)
    # a single first batch of 3 rows: chunks of 3 along T, no filter, no species
    # pool: names are the indices and there is no species_ids dataset
    assert d["/trajectory/positions"]["chunk"] == [3, 5, 3]
    assert d["/trajectory/positions"]["deflate"] == -1
    assert "/trajectory/species_ids" not in d
    assert d["/ions/initial_species_id"]["data"] == ["0"] * 5
    np.testing.assert_array_equal(np.array(d["/trajectory/time_per_ion"]["data"]).reshape(3, 5),
                                  np.repeat(1e-9 * np.arange(3), 5).reshape(3, 5))
    np.testing.assert_array_equal(d["/trajectory/domain_indices"]["data"], [0] * 15)
    np.testing.assert_array_equal(d["/ions/death_time_s"]["data"], [-1.0] * 5)


def test_snapshot_consumer_of_the_device_run_layout(tmp_path, h5_dump):
    """TrajectoryBuffer and H5TrajectoryFile fed the same snapshot dicts hold
    the same /trajectory datasets."""
    ens = W.gaussian_cloud(11, 150.0, 1e-3, seed=2)
    snaps = snapshots(ens, 7, seed=8)
    buf = TR.TrajectoryBuffer(buffer_max=3)
    path = str(tmp_path / "b.h5")
    with TR.H5TrajectoryFile(path, ens, batch=3) as f:
        for s in snaps:
            buf(s)
            f(s)
    ref = buf.datasets()
    d = h5_dump(path)
    for k in ("time", "positions", "velocities", "domain_indices", "species_id_indices"):
        np.testing.assert_array_equal(np.array(d["/trajectory/" + k]["data"]).reshape(ref[k].shape), ref[k])
