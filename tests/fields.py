"""Test-input generators and comparison helpers (numpy; no oracle arithmetic)."""
import numpy as np


def random_su3_field(n_links, seed):
    """Haar-like random SU(3) links as the (n_links, 18) AoS layout."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n_links, 3, 3)) + 1j * rng.standard_normal((n_links, 3, 3))
    q, r = np.linalg.qr(a)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    q = q / (det ** (1.0 / 3.0))[:, None, None]
    out = np.empty((n_links, 18))
    out[:, 0::2] = q.reshape(n_links, 9).real
    out[:, 1::2] = q.reshape(n_links, 9).imag
    return out


def gaussian_field(n_sites, ncomp, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n_sites, 2 * ncomp))


def site_rel_err(got, ref):
    """max over sites of |got(x) - ref(x)|_2 / |ref(x)|_2 (floor 1e-300)."""
    num = np.sqrt(((got - ref) ** 2).sum(axis=1))
    den = np.maximum(np.sqrt((ref ** 2).sum(axis=1)), 1e-300)
    return float((num / den).max())


def rel(a, b):
    return abs(a - b) / abs(b)
