"""The fused pieces of the reduced-precision CG step, through the C ABI.

* cl2_fermionmatrix_eo_dot: the field is bit-identical to cl2_fermionmatrix_eo's, and <in, out> (summed from
  |Qminus_eo in|^2 by the kernel that writes Qminus_eo in) agrees with scalar_product(in, out) to 1e-12 in
  FP64 and with the oracle's <in, QplusQminus_eo in> to 1e-12; FP32 to 1e-5; FP16 (the intermediate field is
  stored in half precision and re-read) to 1e-2.
* cl2_fermionmatrix_eo_cg_step: the same bits as the unfused operator / ratio / saxpy_AND_cgnorm sequence
  (FP64, FP32), FP16 to 5e-3, and the oracle's step in FP64 to 1e-12.
* cl2_cg_update_xp: bit-identical to the two saxpy calls it replaces (Wilson FP64/FP32/FP16) and to the
  oracle's x - a p, r - b p (staggered); the FP32 BLAS kernels' unaligned path (staggered fields) agrees bit
  for bit with the aligned one, and FP32 Wilson spinors off a 16-byte boundary are refused.
"""
import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import abi
from fields import random_su3_field, site_rel_err

pytestmark = pytest.mark.gpu

KAPPA, MUBAR = 0.16, 2 * 0.16 * 0.005


@pytest.mark.parametrize("fmt,prec,tol", [(abi.GAUGE_F64_18, abi.F64, 1e-12), (abi.GAUGE_F64_12, abi.F64, 1e-12),
                                          (abi.GAUGE_F32_12, abi.F32, 1e-5), (abi.GAUGE_F16_12, abi.F16, 1e-2)])
@pytest.mark.parametrize("action", [abi.ACTION_TWISTEDMASS, abi.ACTION_WILSON])
def test_fermionmatrix_eo_dot(fmt, prec, tol, action):
    lat = O.lattice(8, 8, theta_t=1.0)
    mubar = MUBAR if action == abi.ACTION_TWISTEDMASS else 0.0
    with abi.Context(8, 8, theta_t=1.0) as ctx:
        g = random_su3_field(4 * O.volume(lat), 31)
        src = np.random.default_rng(32).standard_normal((O.veo(lat), 24))
        g64 = ctx.gauge(g)
        gf = g64 if fmt == abi.GAUGE_F64_18 else g64.converted(fmt)
        a, b, c = ctx.spinor(prec).upload(src), ctx.spinor(prec), ctx.spinor(prec)
        tmp = ctx.buffer(2 * ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, prec))
        res = ctx.scalars(4)
        ctx.fermionmatrix_eo(abi.OP_QPLUSQMINUS_EO, b.ptr, a.ptr, gf.ptr, fmt, action, KAPPA, mubar, tmp.ptr)
        ctx.fermionmatrix_eo_dot(abi.OP_QPLUSQMINUS_EO, c.ptr, a.ptr, gf.ptr, fmt, action, KAPPA, mubar, tmp.ptr,
                                 res.at(0))
        assert np.array_equal(b.download(), c.download())
        ctx.scalar_product_eo(res.at(2), a.ptr, c.ptr, prec)
        d = res.get()
        assert d[1] == 0.0
        assert abs(d[0] - d[2]) < tol * abs(d[2]) and abs(d[3]) < tol * abs(d[2])
        if prec == abi.F64:
            act = O.TWISTEDMASS if action == abi.ACTION_TWISTEDMASS else O.WILSON
            ref = O.dot(src, O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, src, g, act, KAPPA, mubar))
            assert abs(d[0] - ref.real) < 1e-12 * abs(ref) and abs(ref.imag) < 1e-12 * abs(ref)
        with pytest.raises(abi.Cl2Error):
            ctx.fermionmatrix_eo_dot(abi.OP_QPLUS_EO, c.ptr, a.ptr, gf.ptr, fmt, action, KAPPA, mubar, tmp.ptr,
                                     res.at(0))


@pytest.mark.parametrize("ns,nt", [(8, 8), (6, 10)])  # 6^3x10: 1080 eo sites, partial blocks and link tiles
@pytest.mark.parametrize("fmt,prec", [(abi.GAUGE_F64_18, abi.F64), (abi.GAUGE_F32_12, abi.F32),
                                      (abi.GAUGE_F16_12, abi.F16)])
def test_fermionmatrix_eo_cg_step(fmt, prec, ns, nt):
    """pAp, alpha, r <- r - alpha A p and its norms from the fused step against the unfused sequence (operator
    with <p, A p>, ratio, saxpy_AND_cgnorm): the same bits for FP64/FP32 fields and scalars (norms summed in
    another grouping: 1e-12); FP16 stores A p in half precision in the unfused sequence only: 5e-3.  FP64
    also against the oracle."""
    lat = O.lattice(ns, nt, theta_t=1.0)
    with abi.Context(ns, nt, theta_t=1.0) as ctx:
        g = random_su3_field(4 * O.volume(lat), 41)
        rng = np.random.default_rng(42)
        P, R = rng.standard_normal((O.veo(lat), 24)), rng.standard_normal((O.veo(lat), 24))
        g64 = ctx.gauge(g)
        gf = g64 if fmt == abi.GAUGE_F64_18 else g64.converted(fmt)
        p, r1, r2f = ctx.spinor(prec).upload(P), ctx.spinor(prec).upload(R), ctx.spinor(prec).upload(R)
        v = ctx.spinor(prec)
        tmp = ctx.buffer(2 * ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, prec))
        s = ctx.scalars(24)  # r2 0..1, pAp 2..3, alpha 4..5, norms 6..9 | unfused: pAp 12..13, alpha 14..15, norms 16..19
        r2 = O.norm2(R)
        s.store([r2, 0.0] + [0.0] * 22)
        ctx.fermionmatrix_eo_dot(abi.OP_QPLUSQMINUS_EO, v.ptr, p.ptr, gf.ptr, fmt, abi.ACTION_TWISTEDMASS, KAPPA, MUBAR,
                                 tmp.ptr, s.at(12))
        ctx.complex_op(abi.SCALAR_OP_RATIO, s.at(14), s.at(0), s.at(12))
        ctx.saxpy_AND_cgnorm_eo(r1.ptr, v.ptr, r1.ptr, prec, abi.scalar(dev=s.at(14), kind=abi.SCALAR_DEV_COMPLEX),
                                s.at(16))
        ctx.fermionmatrix_eo_cg_step(abi.OP_QPLUSQMINUS_EO, r2f.ptr, p.ptr, gf.ptr, fmt, abi.ACTION_TWISTEDMASS, KAPPA,
                                     MUBAR, tmp.ptr, s.at(0), s.at(2), s.at(4), s.at(6))
        d = s.get()
        assert np.array_equal(d[2:6], d[12:16])
        got, ref = r2f.download(), r1.download()
        if prec == abi.F16:
            assert site_rel_err(got, ref) < 5e-3
            assert abs(d[6] - d[16]) < 1e-2 * d[16] and abs(d[7] - d[17]) < 1e-2 * d[16]
        else:
            assert np.array_equal(got, ref)
            assert abs(d[6] - d[16]) < 1e-12 * d[16] and abs(d[7] - d[17]) < 1e-12 * d[16]
        assert d[8] == d[6] and d[9] == 0.0
        if prec == abi.F64:
            Ap = O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, P, g, O.TWISTEDMASS, KAPPA, MUBAR)
            alpha = r2 / O.dot(P, Ap)
            new = O.saxpy(Ap, R, alpha)
            assert abs(complex(d[4], d[5]) - alpha) < 1e-12 * abs(alpha)
            assert site_rel_err(got, new) < 1e-12
            assert abs(d[6] - O.norm2(new)) < 1e-12 * O.norm2(new)
            assert abs(d[7] - O.dot(new, R).real) < 1e-12 * O.norm2(new)
        with pytest.raises(abi.Cl2Error):  # r aliasing p
            ctx.fermionmatrix_eo_cg_step(abi.OP_QPLUSQMINUS_EO, p.ptr, p.ptr, gf.ptr, fmt, abi.ACTION_TWISTEDMASS,
                                         KAPPA, MUBAR, tmp.ptr, s.at(0), s.at(2), s.at(4), s.at(6))


@pytest.mark.parametrize("prec", [abi.F64, abi.F32, abi.F16])
def test_cg_update_xp_wilson(prec):
    with abi.Context(8, 4) as ctx:
        rng = np.random.default_rng(51)
        X, P, R = (rng.standard_normal((ctx.veo, 24)) for _ in range(3))
        s = ctx.scalars(4)
        ma, mb = -0.31 + 0.02j, -0.77 - 0.1j
        s.store([ma.real, ma.imag, mb.real, mb.imag])
        sa = abi.scalar(dev=s.at(0), kind=abi.SCALAR_DEV_COMPLEX)
        sb = abi.scalar(dev=s.at(2), kind=abi.SCALAR_DEV_COMPLEX)
        x1, p1, r = ctx.spinor(prec).upload(X), ctx.spinor(prec).upload(P), ctx.spinor(prec).upload(R)
        x2, p2 = ctx.spinor(prec).upload(X), ctx.spinor(prec).upload(P)
        ctx.saxpy_eo(x1.ptr, p1.ptr, x1.ptr, prec, sa)  # x - ma p
        ctx.saxpy_eo(p1.ptr, p1.ptr, r.ptr, prec, sb)   # r - mb p
        ctx.cg_update_xp(abi.FIELD_SPINOR_EO, x2.ptr, p2.ptr, r.ptr, prec, sa, sb)
        assert np.array_equal(x1.download(), x2.download()) and np.array_equal(p1.download(), p2.download())
        if prec == abi.F64:
            assert site_rel_err(x2.download(), O.saxpy(P, X, ma)) < 1e-15
            assert site_rel_err(p2.download(), O.saxpy(P, R, mb)) < 1e-15
        with pytest.raises(abi.Cl2Error):
            ctx.cg_update_xp(abi.FIELD_SPINOR_EO, x2.ptr, x2.ptr, r.ptr, prec, sa, sb)


@pytest.mark.parametrize("prec", [abi.F64, abi.F32])
def test_cg_update_xp_staggered(prec):
    with abi.Context(8, 4) as ctx:
        rng = np.random.default_rng(52)
        X, P, R = (rng.standard_normal((ctx.veo, 6)) for _ in range(3))
        s = ctx.scalars(2)
        s.store([-0.4, -1.3])
        sa = abi.scalar(dev=s.at(0), kind=abi.SCALAR_DEV_REAL)
        sb = abi.scalar(dev=s.at(1), kind=abi.SCALAR_DEV_REAL)
        x, p, r = (ctx.su3vec(prec).upload(v) for v in (X, P, R))
        Xq, Pq, Rq = x.download(), p.download(), r.download()
        ctx.cg_update_xp(abi.FIELD_SU3VEC_EO, x.ptr, p.ptr, r.ptr, prec, sa, sb)
        tol = 1e-15 if prec == abi.F64 else 1e-6
        assert site_rel_err(x.download(), Xq + 0.4 * Pq) < tol
        assert site_rel_err(p.download(), Rq + 1.3 * Pq) < tol


def test_blas_fp32_unaligned_matches_aligned():
    """FP32 staggered operands 8 bytes off a 16-byte boundary take the one-complex-per-lane path: same bits.
    FP32 Wilson spinors are stored as 16-byte component pairs: 8 bytes off, the entry points refuse them."""
    with abi.Context(8, 4) as ctx:
        rng = np.random.default_rng(53)
        X, Y = rng.standard_normal((ctx.veo, 6)), rng.standard_normal((ctx.veo, 6))
        nb = ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SU3VEC_EO, abi.F32)
        bx, by, bo = ctx.buffer(nb + 16), ctx.buffer(nb + 16), ctx.buffer(nb + 16)
        ux, uy, uo = bx.ptr + 8, by.ptr + 8, bo.ptr + 8
        for ptr, v in ((ux, X), (uy, Y)):
            h = np.ascontiguousarray(v, dtype=np.float64)
            ctx.su3vec_import(ptr, h.ctypes.data, abi.F32)
        fx, fy, fo = ctx.su3vec(abi.F32).upload(X), ctx.su3vec(abi.F32).upload(Y), ctx.su3vec(abi.F32)
        a = abi.scalar(0.3 - 0.7j)
        ctx.saxpy_stagg_eo(fo.ptr, fx.ptr, fy.ptr, abi.F32, a)
        ctx.saxpy_stagg_eo(uo, ux, uy, abi.F32, a)
        out = np.empty((ctx.veo, 6))
        ctx.su3vec_export(out.ctypes.data, uo, abi.F32)
        assert np.array_equal(out, fo.download())
        res = ctx.scalars(4)
        ctx.scalar_product_stagg_eo(res.at(0), fx.ptr, fy.ptr, abi.F32)
        ctx.scalar_product_stagg_eo(res.at(2), ux, uy, abi.F32)
        d = res.get()
        assert abs(d[0] - d[2]) < 1e-12 * abs(d[0]) and abs(d[1] - d[3]) < 1e-12 * abs(d[0])
        W = np.ascontiguousarray(rng.standard_normal((ctx.veo, 24)))
        bw = ctx.buffer(ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F32) + 16)
        with pytest.raises(abi.Cl2Error, match="16-byte aligned"):
            ctx.spinor_import(bw.ptr + 8, W.ctypes.data, abi.F32)
        g = ctx.gauge(fmt=abi.GAUGE_F32_12, seed=1, epsilon=0.3)
        fw = ctx.spinor(abi.F32).upload(W)
        with pytest.raises(abi.Cl2Error, match="16-byte aligned"):
            ctx.dslash_eo(bw.ptr + 8, fw.ptr, g.ptr, abi.GAUGE_F32_12, abi.EVEN, 0.125)
