"""ctypes bindings of the CPU oracle (oracle/cl2_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  Nothing in the
cl2qcd_amd package imports it.

Host arrays are numpy float64 in the reference's AoS layouts:
  gauge  (4*V, 18)   link = mu + 4*site, 3x3 row-major (re, im)
  spinor (Veo, 24)   component 3*spin + colour, (re, im)
  su3vec (Veo, 6)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libcl2_oracle.so")

EVEN, ODD = 0, 1
WILSON, TWISTEDMASS = 0, 1
OP_AEE, OP_AEE_AND_GAMMA5, OP_AEE_MINUS, OP_AEE_MINUS_AND_GAMMA5 = 0, 1, 2, 3
OP_QPLUS_EO, OP_QMINUS_EO, OP_QPLUSQMINUS_EO = 4, 5, 6
MTM_SITEDIAG, MTM_INVERSE, MTM_SITEDIAG_MINUS, MTM_INVERSE_MINUS = 0, 1, 2, 3
MTM_SITEDIAG_AND_GAMMA5, MTM_SITEDIAG_MINUS_AND_GAMMA5 = 4, 5
OK, STUCK, DID_NOT_SOLVE = 0, 3, 4


class Lattice(ctypes.Structure):
    _fields_ = [
        ("ns", ctypes.c_int),
        ("nt", ctypes.c_int),
        ("theta_s", ctypes.c_double),
        ("theta_t", ctypes.c_double),
        ("use_cp_re", ctypes.c_int),
        ("cp_re", ctypes.c_double),
        ("use_cp_im", ctypes.c_int),
        ("cp_im", ctypes.c_double),
    ]


def lattice(ns, nt, theta_s=0.0, theta_t=0.0, cp_re=None, cp_im=None):
    return Lattice(ns, nt, theta_s, theta_t, int(cp_re is not None), cp_re or 0.0,
                   int(cp_im is not None), cp_im or 0.0)


_lib = None
_P = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
        L = ctypes.POINTER(Lattice)
        d, i, l = ctypes.c_double, ctypes.c_int, ctypes.c_long
        sig = {
            "orc_volume": (l, [L]),
            "orc_veo": (l, [L]),
            "orc_even_site": (i, [L, l, ctypes.POINTER(l)]),
            "orc_odd_site": (i, [L, l, ctypes.POINTER(l)]),
            "orc_fill_doubles_drand48": (None, [_P, l, i]),
            "orc_gauge_cold": (None, [L, _P]),
            "orc_ildg_read": (i, [ctypes.c_char_p, L, _P, ctypes.c_char_p, i]),
            "orc_plaquette": (d, [L, _P]),
            "orc_convert_to_eo": (None, [L, _P, _P, _P]),
            "orc_convert_from_eo": (None, [L, _P, _P, _P]),
            "orc_dslash_eo": (None, [L, _P, _P, _P, i, d]),
            "orc_dslash_eo_sites": (None, [L, _P, _P, _P, i, d, ctypes.POINTER(l), l]),
            "orc_m_tm": (None, [L, _P, _P, d, i]),
            "orc_dslash_minv_eo": (None, [L, _P, _P, _P, i, d, d, i]),
            "orc_gamma5_eo": (None, [L, _P]),
            "orc_saxpy_and_gamma5_eo": (None, [L, _P, _P, _P, d, d]),
            "orc_m_full": (None, [L, _P, _P, _P, d, d, i, i]),
            "orc_apply_eo": (i, [L, i, _P, _P, _P, i, d, d, i]),
            "orc_saxpy": (None, [_P, _P, _P, d, d, l, i]),
            "orc_sax": (None, [_P, _P, d, d, l, i]),
            "orc_saxsbypz": (None, [_P, _P, _P, _P, d, d, d, d, l, i]),
            "orc_dot": (None, [_P, _P, l, i, _P, _P]),
            "orc_norm2": (d, [_P, l, i]),
            "orc_dks_eo": (None, [L, _P, _P, _P, i]),
            "orc_mdagm_eo": (None, [L, _P, _P, _P, d, i]),
            "orc_stagg_saxpby": (None, [_P, _P, _P, d, d, d, d, l]),
            "orc_cg_eo": (i, [L, i, _P, i, d, d, i, _P, _P, d, i, i, i, i, ctypes.POINTER(i)]),
            "orc_bicgstab_eo": (i, [L, i, _P, i, d, d, i, _P, _P, d, i, i, i, ctypes.POINTER(i)]),
            "orc_cgm": (i, [L, _P, d, i, ctypes.POINTER(_P), _P, _P, i, d, i, i, i, ctypes.POINTER(i)]),
            "orc_set_threads": (None, [i]),
            "orc_get_threads": (i, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def volume(lat):
    return lat.ns ** 3 * lat.nt


def veo(lat):
    return volume(lat) // 2


def eo_sites(lat, parity):
    """Lattice site of every eo index for the given parity (0 even, 1 odd)."""
    f = lib().orc_even_site if parity == 0 else lib().orc_odd_site
    out = np.empty(veo(lat), dtype=np.int64)
    s = ctypes.c_long()
    for i in range(out.size):
        f(ctypes.byref(lat), i, ctypes.byref(s))
        out[i] = s.value
    return out


def set_threads(n):
    lib().orc_set_threads(int(n))


def threads():
    return lib().orc_get_threads()


def fill_drand48(n_doubles, seed):
    out = np.empty(n_doubles, dtype=np.float64)
    lib().orc_fill_doubles_drand48(_p(out), n_doubles, seed)
    return out


def pseudo_randomize_spinor(n_sites, seed):
    return fill_drand48(24 * n_sites, seed).reshape(n_sites, 24)


def pseudo_randomize_su3vec(n_sites, seed):
    return fill_drand48(6 * n_sites, seed).reshape(n_sites, 6)


def gauge_cold(lat):
    g = np.empty((4 * volume(lat), 18))
    lib().orc_gauge_cold(ctypes.byref(lat), _p(g))
    return g


def ildg_read(path, lat):
    g = np.empty((4 * volume(lat), 18))
    err = ctypes.create_string_buffer(256)
    rc = lib().orc_ildg_read(path.encode(), ctypes.byref(lat), _p(g), err, 256)
    if rc:
        raise RuntimeError(err.value.decode())
    return g


def plaquette(lat, g):
    return lib().orc_plaquette(ctypes.byref(lat), _p(g))


def convert_to_eo(lat, full):
    e = np.empty((veo(lat), 24))
    o = np.empty((veo(lat), 24))
    lib().orc_convert_to_eo(ctypes.byref(lat), _p(full), _p(e), _p(o))
    return e, o


def dslash_eo(lat, inp, g, evenodd, kappa):
    out = np.empty_like(inp)
    lib().orc_dslash_eo(ctypes.byref(lat), _p(out), _p(inp), _p(g), evenodd, kappa)
    return out


def dslash_eo_sites(lat, inp, g, evenodd, kappa, eo_idx):
    """dslash_eo at the given output eo indices only: returns (len(eo_idx), 24)."""
    idx = np.ascontiguousarray(eo_idx, dtype=np.int64)
    out = np.empty((idx.size, 24))
    lib().orc_dslash_eo_sites(ctypes.byref(lat), _p(out), _p(inp), _p(g), evenodd, kappa,
                              idx.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), idx.size)
    return out


def m_tm(lat, inp, mubar, variant):
    out = np.empty_like(inp)
    lib().orc_m_tm(ctypes.byref(lat), _p(out), _p(inp), mubar, variant)
    return out


def dslash_minv_eo(lat, inp, g, evenodd, kappa, mubar, minus=False):
    out = np.empty_like(inp)
    lib().orc_dslash_minv_eo(ctypes.byref(lat), _p(out), _p(inp), _p(g), evenodd, kappa, mubar, int(minus))
    return out


def gamma5_eo(lat, inp):
    out = inp.copy()
    lib().orc_gamma5_eo(ctypes.byref(lat), _p(out))
    return out


def saxpy_and_gamma5_eo(lat, x, y, alpha):
    out = np.empty_like(x)
    a = complex(alpha)
    lib().orc_saxpy_and_gamma5_eo(ctypes.byref(lat), _p(out), _p(x), _p(y), a.real, a.imag)
    return out


def m_full(lat, inp, g, kappa, mubar=0.0, action=WILSON, minus=False):
    out = np.empty_like(inp)
    lib().orc_m_full(ctypes.byref(lat), _p(out), _p(inp), _p(g), kappa, mubar, action, int(minus))
    return out


def apply_eo(lat, op, inp, g, action, kappa, mubar=0.0, merged=False):
    out = np.empty_like(inp)
    rc = lib().orc_apply_eo(ctypes.byref(lat), op, _p(out), _p(inp), _p(g), action, kappa, mubar, int(merged))
    assert rc == 0
    return out


def _ncomp(a):
    return a.shape[-1] // 2


def saxpy(x, y, alpha):
    """Wilson saxpy: y - alpha x."""
    out = np.empty_like(x)
    a = complex(alpha)
    lib().orc_saxpy(_p(out), _p(x), _p(y), a.real, a.imag, x.shape[0], _ncomp(x))
    return out


def sax(x, alpha):
    out = np.empty_like(x)
    a = complex(alpha)
    lib().orc_sax(_p(out), _p(x), a.real, a.imag, x.shape[0], _ncomp(x))
    return out


def saxsbypz(x, y, z, alpha, beta):
    out = np.empty_like(x)
    a, b = complex(alpha), complex(beta)
    lib().orc_saxsbypz(_p(out), _p(x), _p(y), _p(z), a.real, a.imag, b.real, b.imag, x.shape[0], _ncomp(x))
    return out


def dot(x, y):
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().orc_dot(_p(x), _p(y), x.shape[0], _ncomp(x), ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def norm2(x):
    return lib().orc_norm2(_p(x), x.shape[0], _ncomp(x))


def dks_eo(lat, inp, g, evenodd):
    out = np.empty_like(inp)
    lib().orc_dks_eo(ctypes.byref(lat), _p(out), _p(inp), _p(g), evenodd)
    return out


def mdagm_eo(lat, inp, g, mass, upper_left=EVEN):
    out = np.empty_like(inp)
    lib().orc_mdagm_eo(ctypes.byref(lat), _p(out), _p(inp), _p(g), mass, upper_left)
    return out


def stagg_saxpby(x, y, alpha, beta):
    out = np.empty_like(x)
    a, b = complex(alpha), complex(beta)
    lib().orc_stagg_saxpby(_p(out), _p(x), _p(y), a.real, a.imag, b.real, b.imag, x.shape[0])
    return out


def cg_eo(lat, op, g, b, x0, action, kappa, mubar=0.0, prec=1e-23, cgmax=1000, iter_refresh=100,
          check_every=10, merged_fermion=False, merged_spinor=False):
    x = x0.copy()
    it = ctypes.c_int()
    rc = lib().orc_cg_eo(ctypes.byref(lat), op, _p(g), action, kappa, mubar, int(merged_fermion), _p(x), _p(b),
                         prec, cgmax, iter_refresh, check_every, int(merged_spinor), ctypes.byref(it))
    return rc, it.value, x


def bicgstab_eo(lat, op, g, b, x0, action, kappa, mubar=0.0, prec=1e-23, cgmax=1000, iter_refresh=100,
                save=False, merged_fermion=False):
    x = x0.copy()
    it = ctypes.c_int()
    rc = lib().orc_bicgstab_eo(ctypes.byref(lat), op, _p(g), action, kappa, mubar, int(merged_fermion), _p(x),
                               _p(b), prec, cgmax, iter_refresh, int(save), ctypes.byref(it))
    return rc, it.value, x


def cgm(lat, g, b, sigma, mass, prec=1e-23, cgmax=1000, check_every=10, upper_left=EVEN, merged_spinor=False):
    xs = [np.zeros_like(b) for _ in sigma]
    arr = (_P * len(xs))(*[_p(x) for x in xs])
    sig = np.ascontiguousarray(sigma, dtype=np.float64)
    it = ctypes.c_int()
    rc = lib().orc_cgm(ctypes.byref(lat), _p(g), mass, upper_left, arr, _p(b), _p(sig), len(xs), prec, cgmax,
                       check_every, int(merged_spinor), ctypes.byref(it))
    return rc, it.value, xs
