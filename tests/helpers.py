"""Shared test helpers: seeded periodic fields, padded-layout conversion and
normwise relative errors."""
import numpy as np

from compactles_hip.workloads import fourier_field  # noqa: F401

H = 4


def halo_for(n):
    return tuple(H if x > 1 else 0 for x in n)


def padded_shape(n):
    h = halo_for(n)
    return tuple(n[d] + 2 * h[d] for d in range(3))


def pad(a, n):
    """Interior array -> padded Fortran array with zero halos."""
    h = halo_for(n)
    out = np.zeros(padded_shape(n), order="F")
    out[h[0]:h[0] + n[0], h[1]:h[1] + n[1], h[2]:h[2] + n[2]] = a
    return out


def interior(a, n):
    h = halo_for(n)
    return a[h[0]:h[0] + n[0], h[1]:h[1] + n[1], h[2]:h[2] + n[2]]


def relerr(a, ref):
    """Normwise relative error ||a - ref||_inf / ||ref||_inf."""
    den = np.abs(ref).max()
    return np.abs(a - ref).max() / (den if den > 0 else 1.0)


def random_field(n, seed=0):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal(n))
