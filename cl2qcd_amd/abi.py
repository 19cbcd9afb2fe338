"""ctypes bindings of the C ABI in include/cl2qcd.h (libcl2qcd_hip.so).

This module is plumbing for Python callers (tests, bench.py, the serving loop);
the kernels and all arithmetic live in the HIP library.  Importing it fails
loudly when the library has not been built: there is no fallback path.
"""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcl2qcd_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "cl2qcd.h")

OK, E_ARG, E_HIP, E_STUCK, E_NOT_CONVERGED, E_UNSUPPORTED = 0, 1, 2, 3, 4, 5
EVEN, ODD = 0, 1
F64, F32, F16 = 0, 1, 2
GAUGE_F64_18, GAUGE_F64_12, GAUGE_F32_12, GAUGE_F16_12 = 0, 1, 2, 3
FIELD_SPINOR_EO, FIELD_SU3VEC_EO, FIELD_GAUGE = 0, 1, 2
ACTION_WILSON, ACTION_TWISTEDMASS = 0, 1
OP_AEE, OP_AEE_AND_GAMMA5, OP_AEE_MINUS, OP_AEE_MINUS_AND_GAMMA5 = 0, 1, 2, 3
OP_QPLUS_EO, OP_QMINUS_EO, OP_QPLUSQMINUS_EO = 4, 5, 6
MTM_SITEDIAGONAL, MTM_INVERSE_SITEDIAGONAL, MTM_SITEDIAGONAL_MINUS = 0, 1, 2
MTM_INVERSE_SITEDIAGONAL_MINUS, MTM_SITEDIAGONAL_AND_GAMMA5, MTM_SITEDIAGONAL_MINUS_AND_GAMMA5 = 3, 4, 5
SCALAR_VALUE, SCALAR_DEV_COMPLEX, SCALAR_DEV_REAL, SCALAR_DEV_VECTOR = 0, 1, 2, 3
SCALAR_OP_PRODUCT, SCALAR_OP_RATIO, SCALAR_OP_SUM, SCALAR_OP_SUBTRACTION = 0, 1, 2, 3


class Cl2Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cl2 error {code}: {msg}")
        self.code = code


class LatticeStruct(ctypes.Structure):
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


class ScalarStruct(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("re", ctypes.c_double),
        ("im", ctypes.c_double),
        ("dev", ctypes.c_void_p),
        ("index", ctypes.c_int),
    ]


_lib = None
_vp, _dp, _i, _d = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_SP = ctypes.POINTER(ScalarStruct)

# name -> (restype, argtypes); every entry point declared in include/cl2qcd.h
SIGNATURES = {
    "cl2_ctx_create": (_i, [ctypes.POINTER(_vp), _i, ctypes.POINTER(LatticeStruct)]),
    "cl2_ctx_destroy": (_i, [_vp]),
    "cl2_last_error": (ctypes.c_char_p, [_vp]),
    "cl2_sync": (_i, [_vp]),
    "cl2_get_stream": (_vp, [_vp]),
    "cl2_eo_sites": (ctypes.c_long, [_vp]),
    "cl2_field_bytes": (ctypes.c_size_t, [_vp, _i, _i]),
    "cl2_alloc": (_i, [_vp, ctypes.POINTER(_vp), ctypes.c_size_t]),
    "cl2_free": (_i, [_vp, _vp]),
    "cl2_copy": (_i, [_vp, _vp, _vp, ctypes.c_size_t]),
    "cl2_spinor_import": (_i, [_vp, _vp, _dp, _i]),
    "cl2_spinor_export": (_i, [_vp, _dp, _vp, _i]),
    "cl2_su3vec_import": (_i, [_vp, _vp, _dp, _i]),
    "cl2_su3vec_export": (_i, [_vp, _dp, _vp, _i]),
    "cl2_gauge_import": (_i, [_vp, _vp, _dp, _i]),
    "cl2_gauge_set_random": (_i, [_vp, _vp, _i, ctypes.c_uint64, _d]),
    "cl2_gauge_heatbath": (_i, [_vp, _vp, _d, _i, _i, ctypes.c_uint64]),
    "cl2_gauge_overrelax": (_i, [_vp, _vp, _i]),
    "cl2_gauge_plaquette": (_i, [_vp, _vp, ctypes.POINTER(ctypes.c_double)]),
    "cl2_gauge_convert": (_i, [_vp, _vp, _i, _vp]),
    "cl2_gauge_export": (_i, [_vp, _dp, _vp, _i]),
    "cl2_scalar_store": (_i, [_vp, _dp, _dp, _i]),
    "cl2_scalar_get": (_i, [_vp, _dp, _dp, _i]),
    "cl2_dslash_eo": (_i, [_vp, _vp, _vp, _vp, _i, _i, _d]),
    "cl2_dslash_AND_M_tm_inverse_sitediagonal_eo": (_i, [_vp, _vp, _vp, _vp, _i, _i, _d, _d, _i]),
    "cl2_dslash_AND_sitediagonal_saxpy_eo": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _d, _i]),
    "cl2_M_tm_sitediagonal_eo": (_i, [_vp, _vp, _vp, _i, _d, _i]),
    "cl2_gamma5_eo": (_i, [_vp, _vp, _i]),
    "cl2_saxpy_AND_gamma5_eo": (_i, [_vp, _vp, _vp, _vp, _i, _SP]),
    "cl2_fermionmatrix_eo": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _d, _d, _vp]),
    "cl2_fermionmatrix_eo_dot": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _d, _d, _vp, _dp]),
    "cl2_fermionmatrix_eo_cg_step": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _d, _d, _vp, _dp, _dp, _dp, _dp]),
    "cl2_fermionmatrix_eo_mrhs_cg_step": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _d, _d, ctypes.POINTER(_d),
                                               ctypes.POINTER(_d), _vp, _i, _dp, _dp, _dp, _dp, _dp]),
    "cl2_saxpy_eo": (_i, [_vp, _vp, _vp, _vp, _i, _SP]),
    "cl2_sax_eo": (_i, [_vp, _vp, _vp, _i, _SP]),
    "cl2_saxsbypz_eo": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _SP, _SP]),
    "cl2_scalar_product_eo": (_i, [_vp, _dp, _vp, _vp, _i]),
    "cl2_scalar_product_eo_noalloc": (_i, [_vp, _dp, _vp, _vp, _i]),
    "cl2_squarenorm_eo": (_i, [_vp, _dp, _vp, _i]),
    "cl2_saxpy_AND_squarenorm_eo": (_i, [_vp, _vp, _vp, _vp, _i, _SP, _dp]),
    "cl2_saxpy_AND_cgnorm_eo": (_i, [_vp, _vp, _vp, _vp, _i, _SP, _dp]),
    "cl2_cg_update_xp": (_i, [_vp, _i, _vp, _vp, _vp, _i, _SP, _SP]),
    "cl2_fermionmatrix_eo_mrhs": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _d, _d, _vp, _i]),
    "cl2_fermionmatrix_eo_mrhs_params": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, ctypes.POINTER(_d), ctypes.POINTER(_d),
                                              _vp, _i]),
    "cl2_dslash_eo_mrhs": (_i, [_vp, _vp, _vp, _vp, _i, _i, _d, _i]),
    "cl2_block_insert_eo": (_i, [_vp, _vp, _vp, _i, _i, _i]),
    "cl2_block_extract_eo": (_i, [_vp, _vp, _vp, _i, _i, _i]),
    "cl2_block_axpy_eo": (_i, [_vp, _vp, _vp, _vp, _i, _i, _dp, _d, _dp]),
    "cl2_block_dot_eo": (_i, [_vp, _dp, _vp, _vp, _i, _i]),
    "cl2_block_cg_update_xp_eo": (_i, [_vp, _vp, _vp, _vp, _i, _i, _dp, _dp]),
    "cl2_D_KS_eo_mrhs": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i]),
    "cl2_MdagM_eo_mrhs": (_i, [_vp, _vp, _vp, _vp, _i, _d, _i, ctypes.POINTER(_d), _vp, _i]),
    "cl2_block_insert_stagg_eo": (_i, [_vp, _vp, _vp, _i, _i, _i]),
    "cl2_block_extract_stagg_eo": (_i, [_vp, _vp, _vp, _i, _i, _i]),
    "cl2_block_axpy_stagg_eo": (_i, [_vp, _vp, _vp, _vp, _i, _i, _dp, _d, _dp]),
    "cl2_block_dot_stagg_eo": (_i, [_vp, _dp, _vp, _vp, _i, _i]),
    "cl2_block_squarenorm_stagg_eo": (_i, [_vp, _dp, _vp, _i, _i]),
    "cl2_block_convert_stagg_eo": (_i, [_vp, _vp, _i, _vp, _i, _i]),
    "cl2_block_set_zero_stagg_eo": (_i, [_vp, _vp, _i, _i]),
    "cl2_block_cg_update_xp_stagg_eo": (_i, [_vp, _vp, _vp, _vp, _i, _i, _dp, _dp]),
    "cl2_block_squarenorm_eo": (_i, [_vp, _dp, _vp, _i, _i]),
    "cl2_block_convert_eo": (_i, [_vp, _vp, _i, _vp, _i, _i]),
    "cl2_block_set_zero_eo": (_i, [_vp, _vp, _i, _i]),
    "cl2_block_cg_alpha": (_i, [_vp, _dp, _dp, _dp, _dp, _i]),
    "cl2_block_cg_beta": (_i, [_vp, _dp, _dp, _dp, _dp, _i]),
    "cl2_autotune_dslash": (_i, [_vp, _vp, _i, _i, ctypes.POINTER(_d), ctypes.POINTER(_i), ctypes.POINTER(_i),
                                 ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "cl2_autotune_dks": (_i, [_vp, _vp, _i, ctypes.POINTER(_d), ctypes.POINTER(_i), ctypes.POINTER(_i),
                              ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "cl2_graph_begin": (_i, [_vp]),
    "cl2_graph_end": (_i, [_vp, ctypes.POINTER(_vp)]),
    "cl2_graph_launch": (_i, [_vp, _vp, _i]),
    "cl2_graph_destroy": (_i, [_vp]),
    "cl2_graph_node_count": (_i, [_vp, _vp, _vp]),
    "cl2_cg_beta": (_i, [_vp, _dp, _dp, _dp]),
    "cl2_scalar_copy": (_i, [_vp, _dp, _dp, _i]),
    "cl2_slice_squarenorm_eo": (_i, [_vp, _dp, _vp, _i, _i, _i]),
    "cl2_gaussian_eo": (_i, [_vp, _i, _vp, _i, ctypes.c_uint64]),
    "cl2_multishift_update_stagg_eo": (_i, [_vp, ctypes.POINTER(_vp), _i, ctypes.POINTER(_vp), _vp, _i, _i, _dp, _dp,
                                            _dp, _dp]),
    "cl2_set_zero_eo": (_i, [_vp, _vp, _i]),
    "cl2_set_cold_eo": (_i, [_vp, _vp, _i]),
    "cl2_convert_precision_eo": (_i, [_vp, _i, _vp, _i, _vp, _i]),
    "cl2_D_KS_eo": (_i, [_vp, _vp, _vp, _vp, _i, _i]),
    "cl2_MdagM_eo": (_i, [_vp, _vp, _vp, _vp, _i, _d, _i, _vp]),
    "cl2_saxpy_stagg_eo": (_i, [_vp, _vp, _vp, _vp, _i, _SP]),
    "cl2_saxpby_stagg_eo": (_i, [_vp, _vp, _vp, _vp, _i, _SP, _SP]),
    "cl2_sax_stagg_eo": (_i, [_vp, _vp, _vp, _i, _SP]),
    "cl2_saxpbypz_stagg_eo": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _SP, _SP]),
    "cl2_scalar_product_stagg_eo": (_i, [_vp, _dp, _vp, _vp, _i]),
    "cl2_scalar_product_real_part_stagg_eo": (_i, [_vp, _dp, _vp, _vp, _i]),
    "cl2_squarenorm_stagg_eo": (_i, [_vp, _dp, _vp, _i]),
    "cl2_sax_vec_AND_squarenorm_stagg_eo": (_i, [_vp, _dp, _dp, _i, _vp, _i]),
    "cl2_set_zero_stagg_eo": (_i, [_vp, _vp, _i]),
    "cl2_set_cold_stagg_eo": (_i, [_vp, _vp, _i]),
    "cl2_complex_op": (_i, [_vp, _i, _dp, _dp, _dp]),
    "cl2_real_op": (_i, [_vp, _i, _dp, _dp, _dp]),
    "cl2_real_get_vector_element": (_i, [_vp, _dp, _dp, _i]),
    "cl2_real_set_vector_element": (_i, [_vp, _dp, _dp, _i]),
    "cl2_update_zeta_cgm": (_i, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i]),
    "cl2_update_beta_cgm": (_i, [_vp, _dp, _dp, _dp, _dp, _i]),
    "cl2_update_alpha_cgm": (_i, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _i]),
    "cl2_cgm_beta_step": (_i, [_vp, _dp, _dp, _dp, _dp]),
    "cl2_MdagM_eo_cgm_step": (_i, [_vp, _vp, _vp, _vp, _i, _d, _d, _i, _vp, _dp, _dp, _dp, _dp, _dp]),
    "cl2_MdagM_eo_mrhs_cg_step": (_i, [_vp, _vp, _vp, _vp, _i, _d, _i, ctypes.POINTER(_d), _vp, _i, _dp, _dp, _dp,
                                       _dp, _dp]),
    "cl2_cgm_alpha_step": (_i, [_vp, _i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "cl2_tune_set": (_i, [_vp, ctypes.c_char_p, ctypes.c_long]),
    "cl2_tune_get": (ctypes.c_long, [_vp, ctypes.c_char_p]),
    "cl2_timer_start": (_i, [_vp]),
    "cl2_timer_stop": (_i, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "cl2_reference_read_write_size": (ctypes.c_uint64, [_vp, ctypes.c_char_p]),
    "cl2_reference_flop_size": (ctypes.c_uint64, [_vp, ctypes.c_char_p]),
    "cl2_algorithmic_bytes": (ctypes.c_uint64, [_vp, ctypes.c_char_p, _i]),
}


def header_symbols(path=HEADER_PATH):
    """Function names declared in include/cl2qcd.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cl2_\w+)\s*\(", text)))


def load():
    """Load libcl2qcd_hip.so and bind every C-ABI entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build the HIP library first (python -c "
                          "'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def scalar(value=0.0, dev=None, kind=None, index=0):
    """Build a cl2_scalar: a host value, or a device scalar / vector element."""
    if dev is None:
        v = complex(value)
        return ScalarStruct(SCALAR_VALUE, v.real, v.imag, None, 0)
    return ScalarStruct(kind if kind is not None else SCALAR_DEV_COMPLEX, 0.0, 0.0, int(dev), index)


class DeviceBuffer:
    """A device allocation owned by a Context (freed with it or on close())."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        ctx.check(ctx.lib.cl2_alloc(ctx.handle, ctypes.byref(p), self.nbytes))
        self.ptr = p.value
        ctx._buffers.add(self)

    def close(self):
        if self.ptr is not None and self.ctx.handle:
            self.ctx.check(self.ctx.lib.cl2_free(self.ctx.handle, self.ptr))
            self.ctx._buffers.discard(self)
        self.ptr = None


class Field(DeviceBuffer):
    """An eo field (Wilson spinor or staggered su3vec) in the device SoA layout."""

    def __init__(self, ctx, kind=FIELD_SPINOR_EO, prec=F64):
        self.kind, self.prec = kind, prec
        self.ncomp = 12 if kind == FIELD_SPINOR_EO else 3
        super().__init__(ctx, ctx.lib.cl2_field_bytes(ctx.handle, kind, prec))

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64).reshape(-1, 2 * self.ncomp)
        assert host.shape[0] == self.ctx.veo
        f = self.ctx.lib.cl2_spinor_import if self.ncomp == 12 else self.ctx.lib.cl2_su3vec_import
        self.ctx.check(f(self.ctx.handle, self.ptr, host.ctypes.data, self.prec))
        return self

    def download(self):
        out = np.empty((self.ctx.veo, 2 * self.ncomp))
        f = self.ctx.lib.cl2_spinor_export if self.ncomp == 12 else self.ctx.lib.cl2_su3vec_export
        self.ctx.check(f(self.ctx.handle, out.ctypes.data, self.ptr, self.prec))
        return out


class Gauge(DeviceBuffer):
    """A device gauge field: from a host AoS array, or a random/cold start on the device."""

    def __init__(self, ctx, host_aos=None, fmt=GAUGE_F64_18, seed=None, epsilon=0.0):
        self.fmt = fmt
        super().__init__(ctx, ctx.lib.cl2_field_bytes(ctx.handle, FIELD_GAUGE, fmt))
        if host_aos is not None:
            host = np.ascontiguousarray(host_aos, dtype=np.float64)
            assert host.size == 4 * 2 * ctx.veo * 18
            ctx.check(ctx.lib.cl2_gauge_import(ctx.handle, self.ptr, host.ctypes.data, fmt))
        else:
            ctx.check(ctx.lib.cl2_gauge_set_random(ctx.handle, self.ptr, fmt, seed or 0, epsilon))

    def download(self):
        out = np.empty((8 * self.ctx.veo, 18))
        self.ctx.check(self.ctx.lib.cl2_gauge_export(self.ctx.handle, out.ctypes.data, self.ptr, self.fmt))
        return out

    def heatbath(self, beta, sweeps, overrelax=0, seed=0):
        assert self.fmt == GAUGE_F64_18
        self.ctx.check(self.ctx.lib.cl2_gauge_heatbath(self.ctx.handle, self.ptr, beta, sweeps, overrelax, seed))

    def plaquette(self):
        p = ctypes.c_double()
        self.ctx.check(self.ctx.lib.cl2_gauge_plaquette(self.ctx.handle, self.ptr, ctypes.byref(p)))
        return p.value

    def converted(self, fmt):
        """A copy of this F64_18 field in another device format."""
        assert self.fmt == GAUGE_F64_18
        g = Gauge.__new__(Gauge)
        g.fmt = fmt
        DeviceBuffer.__init__(g, self.ctx, self.ctx.lib.cl2_field_bytes(self.ctx.handle, FIELD_GAUGE, fmt))
        self.ctx.check(self.ctx.lib.cl2_gauge_convert(self.ctx.handle, g.ptr, fmt, self.ptr))
        return g


class DeviceScalars(DeviceBuffer):
    """n device doubles (complex scalars take two)."""

    def __init__(self, ctx, n):
        self.n = n
        super().__init__(ctx, 8 * n)

    def get(self):
        out = np.empty(self.n)
        self.ctx.check(self.ctx.lib.cl2_scalar_get(self.ctx.handle, out.ctypes.data, self.ptr, self.n))
        return out

    def store(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        assert v.size <= self.n
        self.ctx.check(self.ctx.lib.cl2_scalar_store(self.ctx.handle, self.ptr, v.ctypes.data, v.size))

    def at(self, k):
        return self.ptr + 8 * k


class Context:
    """One GPU, one lattice, one HIP stream (cl2_ctx)."""

    def __init__(self, ns, nt, theta_s=0.0, theta_t=0.0, cp_re=None, cp_im=None, device=0):
        self.lib = load()
        self.ns, self.nt = ns, nt
        lat = LatticeStruct(ns, nt, theta_s, theta_t, int(cp_re is not None), cp_re or 0.0,
                            int(cp_im is not None), cp_im or 0.0)
        h = ctypes.c_void_p()
        rc = self.lib.cl2_ctx_create(ctypes.byref(h), device, ctypes.byref(lat))
        if rc:
            raise Cl2Error(rc, self.lib.cl2_last_error(None).decode())
        self.handle = h.value
        self.veo = self.lib.cl2_eo_sites(self.handle)
        self._buffers = set()

    def check(self, rc):
        if rc:
            raise Cl2Error(rc, self.lib.cl2_last_error(self.handle).decode())
        return rc

    def __getattr__(self, name):
        # ctx.dslash_eo(...) -> cl2_dslash_eo(handle, ...), raising on error
        if name.startswith("_") or "lib" not in self.__dict__:
            raise AttributeError(name)
        f = getattr(self.lib, "cl2_" + name)

        def call(*args):
            return self.check(f(self.handle, *args))

        return call

    def spinor(self, prec=F64):
        return Field(self, FIELD_SPINOR_EO, prec)

    def su3vec(self, prec=F64):
        return Field(self, FIELD_SU3VEC_EO, prec)

    def gauge(self, host_aos=None, fmt=GAUGE_F64_18, seed=None, epsilon=0.0):
        return Gauge(self, host_aos, fmt, seed, epsilon)

    def scalars(self, n):
        return DeviceScalars(self, n)

    def buffer(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def close(self):
        if self.handle:
            for b in list(self._buffers):
                b.close()
            self.lib.cl2_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
