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


def half_links(g):
    """links as the F16_12 format stores them: rows 0, 1 rounded to FP16, row 2 = conj(row0 x row1)"""
    u = g.reshape(-1, 9, 2).copy()
    u[:, :6] = u[:, :6].astype(np.float16).astype(np.float64)
    c = u[:, :, 0] + 1j * u[:, :, 1]
    r0, r1 = c[:, 0:3], c[:, 3:6]
    c[:, 6:9] = np.conj(np.cross(r0, r1))
    out = np.empty_like(g)
    out[:, 0::2] = c.real
    out[:, 1::2] = c.imag
    return out


def field_rel(a, b):
    """whole-field relative 2-norm difference"""
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
