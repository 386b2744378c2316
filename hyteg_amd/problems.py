"""Synthetic, seeded problem data of the benchmark configurations.

* cube:  unit cube split into 6 macro-tetrahedra, RHS uniform[-1,1] on inner DoFs.
* shell: thick spherical shell r in [0.55, 1], ntan = 5, 3 radial layers
  (1920 macro-tetrahedra); coefficient
      k(x) = exp( sum_{m=1}^{16} a_m sin(omega_m . x + phi_m) ),
  a_m ~ N(0, 0.5^2), omega_m ~ N(0, 3^2 I), phi_m ~ U[0, 2 pi); RHS uniform[-1,1].
* Stokes on the shell: viscosity eta(r) = exp(ln(1e4) (1 - r) / 0.45) (10^4
  contrast between the outer and the inner surface); buoyancy T(x) x/|x| from a
  temperature anomaly of 12 Gaussian blobs (sigma 0.05, amplitude 1) whose
  centres are uniform in the shell volume, seeded.
"""
import numpy as np

SHELL = dict(ntan=5, nrad=3, rmin=0.55, rmax=1.0)


def shell_coefficient(seed=0, modes=16, amplitude=0.5, frequency=3.0):
    rng = np.random.default_rng(seed)
    a = rng.normal(0.0, amplitude, modes)
    omega = rng.normal(0.0, frequency, (modes, 3))
    phi = rng.uniform(0.0, 2.0 * np.pi, modes)
    return a, omega, phi


def k_values(a, omega, phi, x):
    x = np.atleast_2d(x)
    return np.exp((np.asarray(a)[None, :] * np.sin(x @ np.asarray(omega).T + np.asarray(phi)[None, :])).sum(axis=1))


def canonical_uniform(n, seed, lo=-1.0, hi=1.0):
    """value of canonical DoF i: SplitMix64(seed + (i+1) * golden) -> [lo, hi) (the
    device generator of hg_function_interpolate_random)"""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(n, dtype=np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return lo + (hi - lo) * ((z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0))


def uniform_rhs(flags, seed=0, inner=1):
    """uniform[-1,1] on DoFs of type `inner`, 0 elsewhere"""
    rng = np.random.default_rng(seed)
    return np.where(flags == inner, rng.uniform(-1.0, 1.0, len(flags)), 0.0)


STOKES_VISCOSITY = (np.log(1e4) / 0.45, -np.log(1e4) / 0.45)  # eta = exp(c0 + c1 r)


def blob_centres(seed=0, n=12, rmin=0.55, rmax=1.0):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = (rmin ** 3 + rng.uniform(size=n) * (rmax ** 3 - rmin ** 3)) ** (1.0 / 3.0)
    return d * r[:, None]


def buoyancy(x, seed=0, sigma=0.05, amplitude=1.0, chunk=1 << 22):
    """nodal buoyancy T(x) x/|x| at points x (n x 3), T = sum of the Gaussian blobs"""
    c = blob_centres(seed)
    out = np.empty_like(x)
    for a in range(0, len(x), chunk):
        xs = x[a:a + chunk]
        T = np.zeros(len(xs))
        for cb in c:
            d2 = ((xs - cb) ** 2).sum(axis=1)
            near = d2 < 36.0 * sigma * sigma
            T[near] += amplitude * np.exp(-d2[near] / (2.0 * sigma * sigma))
        out[a:a + chunk] = xs * (T / np.linalg.norm(xs, axis=1))[:, None]
    return out
