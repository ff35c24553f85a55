"""Shared test helpers: seeded periodic fields, padded-layout conversion and
normwise relative errors."""
import numpy as np

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


def fourier_field(n, kmax, seed, slope=-5.0 / 3.0):
    """Periodic field sum_{1<=|k|<=kmax} |k|^slope cos(k.x + phi_k) on [0,2pi)^3,
    phases from numpy.random.default_rng(seed) drawn in lexicographic (kx,ky,kz)
    order over the half space kx>0 | (kx==0 & ky>0) | (kx==ky==0 & kz>0);
    synthesised with an inverse real FFT and normalised to unit RMS."""
    rng = np.random.default_rng(seed)
    F = np.zeros((n[0], n[1], n[2] // 2 + 1), dtype=complex)
    for kx in range(0, kmax + 1):
        for ky in range(-kmax, kmax + 1):
            for kz in range(-kmax, kmax + 1):
                if not (kx > 0 or (kx == 0 and ky > 0) or (kx == 0 and ky == 0 and kz > 0)):
                    continue
                kk = kx * kx + ky * ky + kz * kz
                if kk == 0 or kk > kmax * kmax:
                    continue
                amp = kk ** (slope / 2)
                ph = rng.uniform(0, 2 * np.pi)
                c = 0.5 * amp * np.exp(1j * ph)
                # place the pair (k, -k) in the rfft half-spectrum (z is the halved axis)
                if kz > 0:
                    F[kx % n[0], ky % n[1], kz] += c
                elif kz < 0:
                    F[(-kx) % n[0], (-ky) % n[1], -kz] += np.conj(c)
                else:
                    F[kx % n[0], ky % n[1], 0] += c
                    F[(-kx) % n[0], (-ky) % n[1], 0] += np.conj(c)
    f = np.fft.irfftn(F, s=n, axes=(0, 1, 2)) * (n[0] * n[1] * n[2])
    f /= np.sqrt(np.mean(f * f))
    return np.asfortranarray(f)
