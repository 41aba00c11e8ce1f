"""Transport metrics of a run, as the reference's analysis scripts define them.

  find_arrival_times      analysis/trajectory.py:250-273: the first snapshot at
                          which an ion is at or past the end plane (less
                          tol_frac * L, or an explicit plane) within the radius
                          (plus tol_frac * R);
  drift_mobility          analysis/ims_mobility.py:160-171: v = L / t_arrival,
                          K = v / E;
  reduced_mobility        analysis/physics.py:79-88: K0 = K (P / P0) (T0 / T),
                          P0 = 101325 Pa, T0 = 273.15 K;
  histogram_bin_edges     analysis/physics.py:96-103;
  fit_gaussian_histogram  analysis/physics.py:119-145 (scipy curve_fit on the
                          counts, start at (max count, mean, std));
  device_arrival_distribution
                          the same histogram and fit over the in-kernel arrival
                          times of a resident handle: edges from the device's
                          min / max, counts binned on the device.
"""
import numpy as np

STANDARD_PRESSURE_PA = 101325.0
STANDARD_TEMPERATURE_K = 273.15


def find_arrival_times(times, positions, z_origin_m, length_m, radius_m, tol_frac,
                       arrival_plane_z_m=None):
    """Snapshot-granular first arrival times.

    times [T], positions [T, N, 3] (the /trajectory datasets).  Returns
    (arrival times of the ions that arrived, mask of those ions, z_target)."""
    times = np.asarray(times, dtype=float)
    positions = np.asarray(positions, dtype=float)
    tol_z = length_m * tol_frac
    tol_r = radius_m * tol_frac
    z = positions[:, :, 2]
    r = np.sqrt(positions[:, :, 0] ** 2 + positions[:, :, 1] ** 2)
    z_end = z_origin_m + length_m if arrival_plane_z_m is None else arrival_plane_z_m
    z_target = z_end - tol_z if arrival_plane_z_m is None else z_end
    reached = (z >= z_target) & (r <= radius_m + tol_r)
    any_reached = reached.any(axis=0)
    first = np.argmax(reached, axis=0)
    arrivals = np.where(any_reached, times[first], np.nan)
    mask = ~np.isnan(arrivals)
    return arrivals[mask], mask, z_target


def drift_mobility(length_m, arrival_s, field_V_m):
    """K = (L / t) / E in m^2/(V s); L may be per ion."""
    return (np.asarray(length_m, dtype=float) / np.asarray(arrival_s, dtype=float)) / field_V_m


def reduced_mobility(mobility_m2_Vs, pressure_pa, temperature_K, p0_pa=STANDARD_PRESSURE_PA,
                     t0_K=STANDARD_TEMPERATURE_K):
    return np.asarray(mobility_m2_Vs, dtype=float) * (pressure_pa / p0_pa) * (t0_K / temperature_K)


def histogram_bin_edges(values, bins):
    values = np.asarray(values, dtype=float)
    if values.size == 0:
        raise ValueError("Cannot build histogram edges for empty input.")
    if values.size > 1:
        return np.linspace(values.min(), values.max(), bins + 1)
    return np.linspace(values.min() * 0.99, values.max() * 1.01 + 1e-12, bins + 1)


def gaussian(x, amplitude, mean, sigma):
    return amplitude * np.exp(-0.5 * ((x - mean) / sigma) ** 2)


def _fit_counts(centers, hist, mean0, sigma0):
    from scipy.optimize import curve_fit
    if hist.size < 3 or hist.max() == 0:
        return None
    try:
        params, _ = curve_fit(gaussian, centers, hist, p0=[float(hist.max()), mean0, sigma0], maxfev=10000)
        return {"a": float(params[0]), "mu": float(params[1]), "sigma": float(params[2])}
    except Exception:
        return None


def fit_gaussian_histogram(values, bins):
    """(fit {a, mu, sigma} or None, bin centres, counts)."""
    values = np.asarray(values, dtype=float)
    if values.size < 3:
        return None, np.array([]), np.array([])
    edges = histogram_bin_edges(values, bins)
    hist, _ = np.histogram(values, bins=edges)
    centers = 0.5 * (edges[:-1] + edges[1:])
    return _fit_counts(centers, hist, float(values.mean()), float(values.std())), centers, hist


def device_arrival_distribution(handle, bins):
    """The arrival-time distribution of a resident handle, binned on the device.

    Returns dict(count, mean, std, min, max, edges, counts, fit); the edges are
    histogram_bin_edges of the arrival times (from the device's min / max)."""
    st = handle.arrival_stats()
    out = {"count": int(st.count), "mean": st.mean_s, "std": st.std_s, "min": st.min_s, "max": st.max_s,
           "edges": None, "counts": None, "fit": None}
    if st.count == 0:
        return out
    if st.count > 1 and st.max_s > st.min_s:
        edges = np.linspace(st.min_s, st.max_s, bins + 1)
    else:
        edges = np.linspace(st.min_s * 0.99, st.max_s * 1.01 + 1e-12, bins + 1)
    counts = handle.arrival_histogram(float(edges[0]), float(edges[-1]), bins).astype(np.int64)
    out["edges"], out["counts"] = edges, counts
    if st.count >= 3:
        out["fit"] = _fit_counts(0.5 * (edges[:-1] + edges[1:]), counts, st.mean_s, st.std_s)
    return out
