"""EHSS scattering surrogate: training-set construction and training driver.

The offline EHSS precompute (icarion_gpu_ehss_precompute; restating
tools/ehss_offline_precompute.cpp:438-517) gives, per cluster, first-contact
samples mu = -n.axis, each under a fresh random orientation and impact point -
the distribution the runtime EHSS draws from.  The azimuth of the contact
normal about the relative velocity is uniform (orientations are uniform), so the
whole scattering law of a cluster is the distribution of mu.  The surrogate
learns its quantile function: sorting a cluster's S samples gives pairs
(u_i = (i + 1/2)/S, mu_(i)); the network input is (u, cos psi, sin psi, f1, f2,
f3) with a random azimuth psi and the cluster features, the target the contact
normal (sqrt(1 - mu^2) cos psi, sqrt(1 - mu^2) sin psi, -mu) in the collision
frame.
"""
import numpy as np

from . import _abi

N_IN = 6
N_OUT = 3


def training_set(mu, features, seed=0):
    """mu: (clusters, S) precompute samples; features: (clusters, 3).
    Returns X (clusters*S, 6), T (clusters*S, 3) float32."""
    mu = np.asarray(mu, dtype=np.float64)
    good = np.isfinite(mu).all(axis=1)
    mu = np.sort(mu[good], axis=1)
    f = np.asarray(features, dtype=np.float32)[good]
    nc, S = mu.shape
    rng = np.random.default_rng(seed)
    u = (np.arange(S) + 0.5) / S
    psi = rng.uniform(0.0, 1.0, (nc, S))
    c, s = np.cos(2 * np.pi * psi), np.sin(2 * np.pi * psi)
    st = np.sqrt(np.clip(1.0 - mu * mu, 0.0, 1.0))
    X = np.empty((nc, S, N_IN), np.float32)
    X[..., 0] = u[None, :]
    X[..., 1] = c
    X[..., 2] = s
    X[..., 3:6] = f[:, None, :]
    T = np.stack([st * c, st * s, -mu], -1).astype(np.float32)
    return X.reshape(-1, N_IN), T.reshape(-1, N_OUT)


def mu_from_output(Y):
    """mu = -n_z / |n| of network outputs (clipped to [0, 1]), as the step kernel does."""
    Y = np.asarray(Y, dtype=np.float64)
    n = np.linalg.norm(Y, axis=-1)
    mu = np.where(n > 0, -Y[..., 2] / np.where(n > 0, n, 1), 0.0)
    return np.clip(mu, 0.0, 1.0)


def train_on_device(handle, n_clusters, samples_per_cluster=1000, features=None, epochs=20,
                    batch=8192, lr=1e-3, seed=0, max_tries=100000, timings=None, schedule=None):
    """Precompute + build the training set + train on the device.  Clusters
    [0, n_clusters) of the handle's species table are the training clusters.
    schedule, when given, is a list of (epochs, learning rate) stages run one
    after the other (each stage starts Adam's moments afresh) in place of
    (epochs, lr): a decaying rate lets the network settle, which removes a
    systematic bias of the learned angle law (a constant rate leaves it
    fluctuating about the optimum).  Returns (losses of every epoch, X, T);
    timings (a dict), when given, receives the seconds of the precompute, the
    training-set construction and the training."""
    import time
    t0 = time.perf_counter()
    _, mu = handle.ehss_precompute(0, n_clusters, samples_per_cluster, 0, 1, max_tries=max_tries,
                                   seed=seed)
    t1 = time.perf_counter()
    X, T = training_set(mu[..., 0], features[:n_clusters], seed=seed + 1)
    t2 = time.perf_counter()
    losses = []
    for ep, rate in (schedule or [(epochs, lr)]):
        losses += list(handle.surrogate_train(X, T, epochs=ep, batch=batch, lr=rate, seed=seed))
    losses = np.array(losses)
    t3 = time.perf_counter()
    if timings is not None:
        timings.update(precompute_s=t1 - t0, training_set_s=t2 - t1, train_s=t3 - t2)
    return losses, X, T


def params_count():
    return _abi.SURROGATE_NPARAMS
