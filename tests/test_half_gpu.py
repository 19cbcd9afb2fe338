"""FP16 Wilson spinors (one FP32 scale per site) and F16_12 links on the GPU.

Bars:
* storage: every component of an imported field is within half an FP16 ulp of
  the site's largest entry (2^-12 * max, plus FP32 rounding) of the FP64 value;
* arithmetic on already-quantised inputs (the dequantised field, links rounded
  to FP16 with row 2 rebuilt in FP64) matches the FP64 oracle to 1e-3 relative
  per site: the hopping term rounds each spin projection and link row 2 to FP16
  and accumulates its half products in FP32, plus one output quantisation;
* against the exact FP64 inputs, the whole-field relative error stays below 3e-3;
* reductions of FP16 fields accumulate in FP64 from the dequantised values, so
  they match numpy on those values to 1e-12.
"""
import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import abi
from fields import half_links, random_su3_field, site_rel_err

pytestmark = pytest.mark.gpu


def field_rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_storage_roundtrip():
    with abi.Context(4, 8) as ctx:
        src = np.random.default_rng(3).standard_normal((ctx.veo, 24))
        src[5] *= 1e-6  # sites of very different scale keep their own precision
        src[7] = 0.0
        f = ctx.spinor(abi.F16).upload(src)
        back = f.download()
        m = np.abs(src).max(axis=1, keepdims=True)
        assert np.all(np.abs(back - src) <= (2.0 ** -12 + 1e-6) * m)
        assert np.all(back[7] == 0.0)
        assert ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F16) == 52 * ctx.veo
        assert ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_GAUGE, abi.GAUGE_F16_12) == 8 * 24 * ctx.veo
        # conversions FP64 <-> FP16 <-> FP32 give the same quantised values
        d64, d32, h = ctx.spinor().upload(src), ctx.spinor(abi.F32), ctx.spinor(abi.F16)
        ctx.convert_precision_eo(abi.FIELD_SPINOR_EO, h.ptr, abi.F16, d64.ptr, abi.F64)
        ctx.convert_precision_eo(abi.FIELD_SPINOR_EO, d32.ptr, abi.F32, h.ptr, abi.F16)
        assert np.array_equal(h.download(), back)
        assert np.allclose(d32.download(), back, rtol=0, atol=1e-6 * m.max())
        ctx.set_cold_eo(h.ptr, abi.F16)
        assert np.allclose(h.download(), np.tile([1.0 / np.sqrt(12 * 2 * ctx.veo), 0.0], 12), rtol=1e-3)


@pytest.mark.parametrize("evenodd", [abi.EVEN, abi.ODD])
@pytest.mark.parametrize("case", [dict(ns=8, nt=8, theta_t=1.0), dict(ns=4, nt=6, theta_s=0.3, theta_t=1.0, cp_re=0.11),
                                  dict(ns=16, nt=32, theta_t=1.0), dict(ns=6, nt=10, theta_t=1.0)],
                         ids=["8x8", "4x6cp", "16x32", "6x10padded"])
def test_dslash_half(case, evenodd):
    lat = O.lattice(case["ns"], case["nt"], **{k: v for k, v in case.items() if k not in ("ns", "nt")})
    with abi.Context(**case) as ctx:
        g = random_su3_field(4 * O.volume(lat), 11)
        src = np.random.default_rng(12).standard_normal((O.veo(lat), 24))
        a, b = ctx.spinor(abi.F16).upload(src), ctx.spinor(abi.F16)
        gf = ctx.gauge(g, abi.GAUGE_F16_12)
        ctx.dslash_eo(b.ptr, a.ptr, gf.ptr, abi.GAUGE_F16_12, evenodd, 0.16)
        out = b.download()
        ref_q = O.dslash_eo(lat, a.download(), half_links(g), evenodd, 0.16)
        assert site_rel_err(out, ref_q) < 1e-3
        assert field_rel(out, O.dslash_eo(lat, src, g, evenodd, 0.16)) < 3e-3
        # the device conversion F64_18 -> F16_12 stores the same links as the host import
        g64 = ctx.gauge(g)
        gc = ctx.gauge(fmt=abi.GAUGE_F16_12)
        ctx.gauge_convert(gc.ptr, abi.GAUGE_F16_12, g64.ptr)
        ctx.dslash_eo(b.ptr, a.ptr, gc.ptr, abi.GAUGE_F16_12, evenodd, 0.16)
        assert np.array_equal(b.download(), out)


@pytest.mark.parametrize("op", [abi.OP_AEE, abi.OP_QPLUS_EO, abi.OP_QPLUSQMINUS_EO])
@pytest.mark.parametrize("action", [abi.ACTION_WILSON, abi.ACTION_TWISTEDMASS])
def test_fermionmatrix_half(op, action):
    lat = O.lattice(8, 8, theta_t=1.0)
    with abi.Context(8, 8, theta_t=1.0) as ctx:
        g = random_su3_field(4 * O.volume(lat), 21)
        src = np.random.default_rng(22).standard_normal((O.veo(lat), 24))
        a, b = ctx.spinor(abi.F16).upload(src), ctx.spinor(abi.F16)
        tmp = ctx.buffer(2 * ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F16))
        gf = ctx.gauge(g, abi.GAUGE_F16_12)
        mubar = 0.0016 if action == abi.ACTION_TWISTEDMASS else 0.0
        ctx.fermionmatrix_eo(op, b.ptr, a.ptr, gf.ptr, abi.GAUGE_F16_12, action, 0.16, mubar, tmp.ptr)
        act = O.TWISTEDMASS if action == abi.ACTION_TWISTEDMASS else O.WILSON
        ref = O.apply_eo(lat, op, src, g, act, 0.16, mubar)
        assert field_rel(b.download(), ref) < 5e-3


def test_blas_half():
    with abi.Context(8, 8) as ctx:
        rng = np.random.default_rng(31)
        X, Y, Z = (rng.standard_normal((ctx.veo, 24)) for _ in range(3))
        x, y, z, o = (ctx.spinor(abi.F16) for _ in range(4))
        x.upload(X), y.upload(Y), z.upload(Z)
        Xq, Yq, Zq = x.download(), y.download(), z.download()
        a, b = 0.5 - 0.25j, 1.5 + 2j
        ctx.saxpy_eo(o.ptr, x.ptr, y.ptr, abi.F16, abi.scalar(a))
        assert site_rel_err(o.download(), O.saxpy(Xq, Yq, a)) < 1e-3
        ctx.saxsbypz_eo(o.ptr, x.ptr, y.ptr, z.ptr, abi.F16, abi.scalar(a), abi.scalar(b))
        assert site_rel_err(o.download(), O.saxsbypz(Xq, Yq, Zq, a, b)) < 1e-3
        ctx.sax_eo(o.ptr, x.ptr, abi.F16, abi.scalar(a))
        assert site_rel_err(o.download(), O.saxpy(Xq, np.zeros_like(Xq), -a)) < 1e-3
        res = ctx.scalars(2)
        ctx.scalar_product_eo(res.ptr, x.ptr, y.ptr, abi.F16)
        d = res.get()
        ref = O.dot(Xq, Yq)
        assert abs(complex(d[0], d[1]) - ref) < 1e-12 * np.sqrt(O.norm2(Xq) * O.norm2(Yq))
        ctx.squarenorm_eo(res.ptr, x.ptr, abi.F16)
        assert abs(res.get()[0] - O.norm2(Xq)) < 1e-12 * O.norm2(Xq)
        ctx.saxpy_AND_squarenorm_eo(o.ptr, x.ptr, y.ptr, abi.F16, abi.scalar(a), res.ptr)
        oq = o.download()
        assert abs(res.get()[0] - O.norm2(O.saxpy(Xq, Yq, a))) < 1e-3 * O.norm2(oq)
        ctx.gamma5_eo(x.ptr, abi.F16)
        assert np.array_equal(x.download(), O.gamma5_eo(O.lattice(8, 8), Xq))
        ctx.M_tm_sitediagonal_eo(o.ptr, y.ptr, abi.F16, 0.37, abi.MTM_INVERSE_SITEDIAGONAL)
        assert site_rel_err(o.download(), O.m_tm(O.lattice(8, 8), Yq, 0.37, abi.MTM_INVERSE_SITEDIAGONAL)) < 1e-3


def test_half_rejected_for_staggered():
    with abi.Context(4, 4) as ctx:
        assert ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SU3VEC_EO, abi.F16) == 0
        v, w = ctx.su3vec(), ctx.su3vec()
        g = ctx.gauge(fmt=abi.GAUGE_F16_12)
        with pytest.raises(abi.Cl2Error):
            ctx.D_KS_eo(w.ptr, v.ptr, g.ptr, abi.GAUGE_F16_12, abi.EVEN)
        res = ctx.scalars(1)
        with pytest.raises(abi.Cl2Error):
            ctx.squarenorm_stagg_eo(res.ptr, v.ptr, abi.F16)


@pytest.mark.parametrize("prec", [abi.F64, abi.F32, abi.F16])
def test_saxpy_and_cgnorm(prec):
    """r <- r - alpha v in place, with (|r_new|^2, Re <r_new, r_old>, |r_new|^2, 0) from the same pass."""
    with abi.Context(8, 8) as ctx:
        rng = np.random.default_rng(41)
        R, Vv = rng.standard_normal((ctx.veo, 24)), rng.standard_normal((ctx.veo, 24))
        r, v = ctx.spinor(prec).upload(R), ctx.spinor(prec).upload(Vv)
        Rq, Vq = r.download(), v.download()
        res = ctx.scalars(4)
        a = 0.3 + 0.1j
        ctx.saxpy_AND_cgnorm_eo(r.ptr, v.ptr, r.ptr, prec, abi.scalar(a), res.ptr)
        new = O.saxpy(Vq, Rq, a)
        out = res.get()
        tol = 1e-13 if prec == abi.F64 else 1e-6
        assert abs(out[0] - O.norm2(new)) < tol * O.norm2(new)
        assert abs(out[1] - O.dot(new, Rq).real) < tol * O.norm2(new)
        assert out[2] == out[0] and out[3] == 0.0
        assert site_rel_err(r.download(), new) < (1e-15 if prec == abi.F64 else 1e-3)
