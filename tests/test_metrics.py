"""Transport metrics as the reference's analysis defines them (CPU).

find_arrival_times follows analysis/trajectory.py:250-273; the cases below are
worked by hand from that definition.  drift_mobility / reduced_mobility are
checked against their closed forms (analysis/ims_mobility.py:160-171,
analysis/physics.py:79-88), and the histogram fit against a sample of known
mean and width.
"""
import numpy as np
import pytest

from icarion_amd import metrics


def test_find_arrival_times_hand_cases():
    L, R, tol = 10e-3, 2e-3, 0.01            # z_target = 10 mm - 0.1 mm, r <= 2.02 mm
    times = np.array([0.0, 1.0, 2.0, 3.0, 4.0]) * 1e-6
    pos = np.zeros((5, 4, 3))
    pos[:, :, 2] = np.array([0.0, 2e-3, 5e-3, 8e-3, 9e-3])[:, None]
    pos[2:, 0, 2] = [9.9e-3, 11e-3, 12e-3]     # ion 0: first at the target at snapshot 2
    pos[:, 2, 2] = 10.5e-3                      # ion 2 past the plane ...
    pos[:, 2, 0] = 2.5e-3                       # ... but outside the radius throughout
    pos[1, 3, 2], pos[2, 3, 2] = 10e-3, 3e-3    # ion 3 reaches at 1 and comes back
    pos[4, 3, 0] = 2.01e-3
    arr, mask, z_target = metrics.find_arrival_times(times, pos, 0.0, L, R, tol)
    assert z_target == pytest.approx(9.9e-3)
    assert mask.tolist() == [True, False, False, True]
    assert arr.tolist() == [2e-6, 1e-6]
    # an explicit plane is used as it is, no tolerance
    arr2, mask2, z2 = metrics.find_arrival_times(times, pos, 0.0, L, R, tol, arrival_plane_z_m=11e-3)
    assert z2 == 11e-3 and mask2.tolist() == [True, False, False, False] and arr2.tolist() == [3e-6]


def test_mobility_closed_forms():
    K = metrics.drift_mobility(0.05, 2.5e-3, 400.0)
    assert K == pytest.approx(0.05 / 2.5e-3 / 400.0, rel=1e-15)
    K0 = metrics.reduced_mobility(K, 80000.0, 310.0)
    assert K0 == pytest.approx(K * (80000.0 / 101325.0) * (273.15 / 310.0), rel=1e-15)


def test_histogram_edges_and_gaussian_fit():
    assert metrics.histogram_bin_edges([2.0], 4) == pytest.approx(np.linspace(1.98, 2.02 + 1e-12, 5))
    rng = np.random.default_rng(4)
    x = rng.normal(3e-4, 2e-5, 200000)
    fit, centers, hist = metrics.fit_gaussian_histogram(x, 60)
    assert hist.sum() == x.size and len(centers) == 60
    assert fit["mu"] == pytest.approx(3e-4, rel=2e-3)
    assert abs(fit["sigma"]) == pytest.approx(2e-5, rel=1e-2)
    assert metrics.fit_gaussian_histogram([1.0, 2.0], 10)[0] is None
