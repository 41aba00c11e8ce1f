"""icarion_gpu_step_noalloc (the step table and launch counters in the
library's fixed per-device scratch array) against icarion_gpu_step (the
handle's own buffers), through the C ABI: the same kernels on the same inputs,
so the ion state and the counters must agree bit for bit.  Two calls in a row
share the scratch; the second has fewer ions and fewer steps, so anything the
first call left in the scratch (counters, table rows past the second call's
end) would show up in its results."""
import numpy as np
import pytest

from icarion_amd import _abi
from icarion_amd import workloads as W

pytestmark = pytest.mark.gpu

STATE = ("x", "y", "z", "vx", "vy", "vz", "death_time_s", "arrival_time_s", "active", "born",
         "n_collisions")


def setup(h, w):
    h.set_domain(w["domain"], w["physics"])
    h.set_species(w["species"], w.get("atoms"))


def run(w, ens, n_steps, step_offset, noalloc):
    from icarion_amd import gpu
    e = ens.copy()
    with gpu.Handle(steps_per_launch=700) as h:
        setup(h, w)
        h.upload_state(e)
        f = h.step_noalloc if noalloc else h.step
        st = f(0.0, w["dt"], n_steps, step_offset)
        h.download_state(e)
    return e, st


def same(a, b, st_a, st_b):
    for f in STATE:
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    for f in ("ion_steps", "collisions", "alive", "arrived", "t_end_s"):
        assert getattr(st_a, f) == getattr(st_b, f), f


CASES = [
    ("ims_hss_fp64", lambda: W.ims_hss_workload(n=3000, seed=123)),
    ("ims_hss_fp32", lambda: W.ims_hss_workload(n=3000, seed=5,
                                                precision=_abi.PRECISION_FP32_KAHAN)),
    ("quadrupole_boris_langevin_fp32",
     lambda: W.quadrupole_workload(n=3000, seed=9, precision=_abi.PRECISION_FP32_KAHAN)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_step_two_calls(name, make):
    from icarion_amd import gpu
    w = make()
    big = w["ensemble"]
    small = big.subset(np.arange(0, big.n, 3))
    # one process-wide scratch per device: the calls below run back to back
    # on the same array
    for ens, n_steps, off in ((big, 1500, 0), (small, 400, 77)):
        print(f"noalloc {name}: {ens.n} ions x {n_steps} steps", flush=True)
        e_ref, st_ref = run(w, ens, n_steps, off, noalloc=False)
        e_na, st_na = run(w, ens, n_steps, off, noalloc=True)
        assert st_ref.ion_steps > 0
        same(e_ref, e_na, st_ref, st_na)
    # a second call on one handle continues the first exactly, as icarion_gpu_step does
    e1 = big.copy()
    e2 = big.copy()
    with gpu.Handle() as h:
        setup(h, w)
        h.upload_state(e1)
        h.step(0.0, w["dt"], 300, 0)
        st1 = h.step(300 * w["dt"], w["dt"], 200, 300)
        h.download_state(e1)
    with gpu.Handle() as h:
        setup(h, w)
        h.upload_state(e2)
        h.step_noalloc(0.0, w["dt"], 300, 0)
        st2 = h.step_noalloc(300 * w["dt"], w["dt"], 200, 300)
        h.download_state(e2)
    same(e1, e2, st1, st2)


def test_noalloc_limits():
    from icarion_amd import gpu
    w = W.ims_hss_workload(n=256, seed=1)
    with gpu.Handle() as h:
        setup(h, w)
        h.upload_state(w["ensemble"].copy())
        with pytest.raises(gpu.IcarionError) as ei:
            h.step_noalloc(0.0, w["dt"], gpu.NOALLOC_MAX_STEPS + 1, 0)
        assert ei.value.code == _abi.ERR_INVALID
        st = h.step_noalloc(0.0, w["dt"], gpu.NOALLOC_MAX_STEPS, 0)
        assert st.ion_steps > 0
        ph = W.physics(_abi.INTEGRATOR_RK45, _abi.COLLISION_HSS, seed=1)
        h.set_domain(w["domain"], ph)
        with pytest.raises(gpu.IcarionError) as ei:
            h.step_noalloc(0.0, w["dt"], 10, 0)
        assert ei.value.code == _abi.ERR_UNSUPPORTED
