"""Arrival-time statistics and histogram on the device.

The counts must equal numpy's np.histogram over the same edges exactly,
including times that lie on an edge or one ulp either side of it (the bin
rule the reference's analysis uses through numpy); the statistics must
equal numpy's count, min and max exactly, and its mean and std to rounding.
"""
import numpy as np
import pytest

from icarion_amd import metrics
from icarion_amd import workloads as W

pytestmark = pytest.mark.gpu


def _handle(w):
    from icarion_amd import gpu
    h = gpu.Handle()
    h.set_domain(w["domain"], w["physics"])
    h.set_species(w["species"])
    return h


def test_arrival_statistics_and_histogram_after_a_run():
    w = W.ims_hss_workload(n=60000, seed=61)
    e = w["ensemble"].copy()
    # a plane just ahead of the cloud's centre: a few per cent of the ions
    # cross it in 4000 steps by drift and diffusion
    w["domain"].detector_z_m = float(np.mean(e.z)) + 5e-6
    h = _handle(w)
    try:
        h.upload_state(e)
        h.step(0.0, w["dt"], 4000, 0)
        h.download_state(e)
        st = h.arrival_stats()
        ar = e.arrival_time_s[np.isfinite(e.arrival_time_s)]
        assert st.count == ar.size > 1000
        assert (st.min_s, st.max_s) == (ar.min(), ar.max())
        assert st.mean_s == pytest.approx(ar.mean(), rel=1e-12)
        assert st.std_s == pytest.approx(ar.std(), rel=1e-6)
        for bins in (1, 7, 64, 1000):
            edges = np.linspace(ar.min(), ar.max(), bins + 1)
            got = h.arrival_histogram(edges[0], edges[-1], bins)
            np.testing.assert_array_equal(got, np.histogram(ar, bins=edges)[0])
        # the device distribution bins exactly as the reference's analysis
        # does; its Gaussian fit is the host's fit of those counts (this
        # skewed crossing-time distribution makes the fit itself ill-posed,
        # so only the counts it starts from are compared)
        dist = metrics.device_arrival_distribution(h, 50)
        _, _, hist = metrics.fit_gaussian_histogram(ar, 50)
        np.testing.assert_array_equal(dist["counts"], hist)
        np.testing.assert_array_equal(dist["edges"], metrics.histogram_bin_edges(ar, 50))
    finally:
        h.close()


def test_histogram_bin_rule_on_edges():
    w = W.ims_hss_workload(n=64, seed=62)
    e = w["ensemble"].copy()
    lo, hi, bins = 1.0e-4, 3.7e-4, 9
    edges = np.linspace(lo, hi, bins + 1)
    vals = np.concatenate([edges, np.nextafter(edges, -np.inf), np.nextafter(edges, np.inf)])
    vals = np.concatenate([vals, [np.nan] * (e.n - vals.size)])[: e.n]
    e.arrival_time_s = vals.copy()
    h = _handle(w)
    try:
        h.upload_state(e)
        got = h.arrival_histogram(lo, hi, bins)
        fin = vals[np.isfinite(vals)]
        np.testing.assert_array_equal(got, np.histogram(fin, bins=edges)[0])
        st = h.arrival_stats()
        assert st.count == fin.size and st.min_s == fin.min() and st.max_s == fin.max()
        from icarion_amd import gpu
        with pytest.raises(gpu.IcarionError):
            h.arrival_histogram(hi, lo, bins)
        with pytest.raises(gpu.IcarionError):
            h.arrival_histogram(lo, hi, 0)
    finally:
        h.close()
