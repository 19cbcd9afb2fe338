"""Multi-right-hand-side staggered operators, staggered block BLAS and the lockstep shifted solver on the GPU.

* slot r of D_KS_eo_mrhs and of MdagM_eo_mrhs without shifts equals the single-field kernels bit for bit
  (same arithmetic in the same order); with a shift per right-hand side, MdagM_eo_mrhs equals the
  single-field MdagM_eo plus sigma_r x to rounding, and the FP64 oracle to 1e-12;
* the staggered block BLAS agree with numpy per right-hand side, and the fused x/p update with the two
  block axpys it replaces bit for bit;
* the lockstep solve of shifted systems meets |b - (MdagM + sigma_r) x_r| / |b| <= 1e-10 for every shift
  with the FP64 oracle operator and agrees with one-shift-at-a-time solves.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import abi
from conftest import CONF_00200
from fields import random_su3_field, site_rel_err

pytestmark = pytest.mark.gpu


def make_block(ctx, vecs, prec):
    n = len(vecs)
    blk = ctx.buffer(n * ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SU3VEC_EO, prec))
    for r, v in enumerate(vecs):
        f = ctx.su3vec(prec).upload(v)
        ctx.block_insert_stagg_eo(blk.ptr, f.ptr, prec, n, r)
        f.close()
    return blk


def slot(ctx, blk, prec, n, r):
    f = ctx.su3vec(prec)
    ctx.block_extract_stagg_eo(f.ptr, blk.ptr, prec, n, r)
    out = f.download()
    f.close()
    return out


@pytest.mark.parametrize("nrhs", [2, 12, 16])
@pytest.mark.parametrize("fmt", [abi.GAUGE_F64_18, abi.GAUGE_F32_12])
def test_dks_and_mdagm_mrhs_match_single(fmt, nrhs):
    kw = dict(theta_s=0.3, theta_t=1.0, cp_im=0.2)
    lat = O.lattice(8, 4, **kw)
    prec = abi.F32 if fmt == abi.GAUGE_F32_12 else abi.F64
    mass = 0.05
    with abi.Context(8, 4, **kw) as ctx:
        g = random_su3_field(4 * O.volume(lat), 61)
        g64 = ctx.gauge(g)
        gf = g64 if fmt == abi.GAUGE_F64_18 else g64.converted(fmt)
        rng = np.random.default_rng(nrhs)
        srcs = [rng.standard_normal((ctx.veo, 6)) for _ in range(nrhs)]
        blk = make_block(ctx, srcs, prec)
        out, tmp = ctx.buffer(blk.nbytes), ctx.buffer(blk.nbytes)
        a, b, t = ctx.su3vec(prec), ctx.su3vec(prec), ctx.su3vec(prec)
        for evenodd in (abi.EVEN, abi.ODD):
            ctx.D_KS_eo_mrhs(out.ptr, blk.ptr, gf.ptr, fmt, evenodd, nrhs)
            for r in (0, nrhs // 2, nrhs - 1):
                a.upload(srcs[r])
                ctx.D_KS_eo(b.ptr, a.ptr, gf.ptr, fmt, evenodd)
                assert np.array_equal(slot(ctx, out, prec, nrhs, r), b.download())
        ctx.MdagM_eo_mrhs(out.ptr, blk.ptr, gf.ptr, fmt, mass, abi.EVEN, None, tmp.ptr, nrhs)
        for r in (0, nrhs - 1):
            a.upload(srcs[r])
            ctx.MdagM_eo(b.ptr, a.ptr, gf.ptr, fmt, mass, abi.EVEN, t.ptr)
            assert np.array_equal(slot(ctx, out, prec, nrhs, r), b.download())
        sigma = list(10.0 ** rng.uniform(-4, 0, nrhs))
        ctx.MdagM_eo_mrhs(out.ptr, blk.ptr, gf.ptr, fmt, mass, abi.EVEN, (ctypes.c_double * nrhs)(*sigma), tmp.ptr,
                          nrhs)
        for r in range(nrhs):
            a.upload(srcs[r])
            ctx.MdagM_eo(b.ptr, a.ptr, gf.ptr, fmt, mass, abi.EVEN, t.ptr)
            ref = b.download() + sigma[r] * a.download()
            got = slot(ctx, out, prec, nrhs, r)
            assert site_rel_err(got, ref) < (1e-14 if prec == abi.F64 else 1e-6)
            if prec == abi.F64:
                assert site_rel_err(got, O.mdagm_eo(lat, srcs[r], g, mass) + sigma[r] * srcs[r]) < 1e-12
        with pytest.raises(abi.Cl2Error):
            ctx.D_KS_eo_mrhs(out.ptr, out.ptr, gf.ptr, fmt, abi.EVEN, nrhs)


@pytest.mark.parametrize("nrhs", [3, 12, 16])
@pytest.mark.parametrize("prec", [abi.F64, abi.F32])
def test_stagg_block_blas(prec, nrhs):
    with abi.Context(4, 8) as ctx:
        rng = np.random.default_rng(7)
        X = [rng.standard_normal((ctx.veo, 6)) for _ in range(nrhs)]
        Y = [rng.standard_normal((ctx.veo, 6)) for _ in range(nrhs)]
        R = [rng.standard_normal((ctx.veo, 6)) for _ in range(nrhs)]
        bx, by, br = make_block(ctx, X, prec), make_block(ctx, Y, prec), make_block(ctx, R, prec)
        X = [slot(ctx, bx, prec, nrhs, r) for r in range(nrhs)]  # as stored
        Y = [slot(ctx, by, prec, nrhs, r) for r in range(nrhs)]
        coef = rng.standard_normal(2 * nrhs)
        cdev, cdev2, norms, res = (ctx.scalars(2 * nrhs) for _ in range(4))
        cdev.store(coef)
        cdev2.store(rng.standard_normal(2 * nrhs))
        tol = 1e-13 if prec == abi.F64 else 1e-6
        cx = lambda a: a[:, 0::2] + 1j * a[:, 1::2]
        ctx.block_dot_stagg_eo(res.ptr, bx.ptr, by.ptr, prec, nrhs)
        d = res.get()
        for r in range(nrhs):
            ref = np.vdot(cx(X[r]), cx(Y[r]))
            assert abs(complex(d[2 * r], d[2 * r + 1]) - ref) < 1e-12 * np.sqrt(O.norm2(X[r]) * O.norm2(Y[r]))
        ctx.block_squarenorm_stagg_eo(res.ptr, bx.ptr, prec, nrhs)
        d = res.get()
        for r in range(nrhs):
            assert abs(d[r] - O.norm2(X[r])) < 1e-12 * O.norm2(X[r])
        ctx.block_axpy_stagg_eo(by.ptr, bx.ptr, by.ptr, prec, nrhs, cdev.ptr, -1.0, norms.ptr)
        nm = norms.get()
        for r in range(nrhs):
            c = complex(coef[2 * r], coef[2 * r + 1])
            new = cx(Y[r]) - c * cx(X[r])
            got = cx(slot(ctx, by, prec, nrhs, r))
            assert np.abs(got - new).max() < 10 * tol * np.abs(new).max()
            n2 = float(np.vdot(new, new).real)
            assert abs(nm[2 * r] - n2) < 10 * tol * n2
            assert abs(nm[2 * r + 1] - float(np.vdot(new, cx(Y[r])).real)) < 10 * tol * n2
        # fused x/p update against the two block axpys
        x1, p1 = make_block(ctx, Y, prec), make_block(ctx, X, prec)
        x2, p2 = make_block(ctx, Y, prec), make_block(ctx, X, prec)
        ctx.block_axpy_stagg_eo(x1.ptr, p1.ptr, x1.ptr, prec, nrhs, cdev.ptr, 1.0, None)
        ctx.block_axpy_stagg_eo(p1.ptr, p1.ptr, br.ptr, prec, nrhs, cdev2.ptr, 1.0, None)
        ctx.block_cg_update_xp_stagg_eo(x2.ptr, p2.ptr, br.ptr, prec, nrhs, cdev.ptr, cdev2.ptr)
        for r in range(nrhs):
            assert np.array_equal(slot(ctx, x1, prec, nrhs, r), slot(ctx, x2, prec, nrhs, r))
            assert np.array_equal(slot(ctx, p1, prec, nrhs, r), slot(ctx, p2, prec, nrhs, r))


@pytest.mark.parametrize("inner", ["f32", "f64"])
def test_stagg_block_solver(inner):
    from cl2qcd_amd import _host as H
    l = H.Lattice(["--nTime=4", "--fermionAction=rooted_stagg", "--mass=0.05"])
    gf = H.Gaugefield(l, CONF_00200)
    A = H.MdagM_eo(l)
    b = H.Staggeredfield_eo(l)
    H.pseudo_randomize(b, 19)
    sigma = [0.0, 0.001, 0.01, 0.1]
    xs = [H.Staggeredfield_eo(l) for _ in sigma]
    for x in xs:
        x.set_zero()
    prec = H.F32 if inner == "f32" else H.F64
    it, per = H.cg_reliable_update_stagg_block(xs, A, gf, sigma, b, l, 1e-10, l.staggered_parameters(), prec)
    assert it > 0 and all(p["true_relative_residual"] <= 1e-10 for p in per)
    g = O.ildg_read(CONF_00200, O.lattice(4, 4))
    lat = O.lattice(4, 4)
    bh = b.get_host()
    for s, x in zip(sigma, xs):
        xh = x.get_host()
        r = O.mdagm_eo(lat, xh, g, 0.05) + s * xh - bh
        assert np.sqrt(O.norm2(r) / O.norm2(bh)) <= 1.01e-10
        one = H.Staggeredfield_eo(l)
        one.set_zero()
        H.cg_reliable_update_stagg(one, A, gf, s, b, l, 1e-10, l.staggered_parameters(), prec)
        oh = one.get_host()
        assert np.sqrt(((oh - xh) ** 2).sum() / (oh ** 2).sum()) < 1e-6


@pytest.mark.parametrize("upper_left", [abi.EVEN, abi.ODD])
@pytest.mark.parametrize("nrhs", [2, 8, 16])
@pytest.mark.parametrize("fmt", [abi.GAUGE_F64_18, abi.GAUGE_F32_12])
def test_mdagm_mrhs_cg_step_matches_unfused_sequence(fmt, nrhs, upper_left):
    """cl2_MdagM_eo_mrhs_cg_step against the sequence it replaces (cl2_MdagM_eo_mrhs, cl2_block_dot_stagg_eo,
    cl2_block_cg_alpha, the residual cl2_block_axpy_stagg_eo with norms), with a shift per right-hand side,
    boundary phases and an imaginary chemical potential: pAp from (m^2 + sigma) |p|^2 + |D p|^2 equals <p, A p>
    to rounding (D_KS_eo anti-hermitian), and so do alpha, the updated residual and its norms; inactive
    right-hand sides get alpha 0 and keep their residual."""
    kw = dict(theta_s=0.3, theta_t=1.0, cp_im=0.2)
    lat = O.lattice(8, 4, **kw)
    prec = abi.F32 if fmt == abi.GAUGE_F32_12 else abi.F64
    tol = 1e-12 if prec == abi.F64 else 2e-5
    mass = 0.05
    with abi.Context(8, 4, **kw) as ctx:
        g = random_su3_field(4 * O.volume(lat), 67)
        g64 = ctx.gauge(g)
        gf = g64 if fmt == abi.GAUGE_F64_18 else g64.converted(fmt)
        rng = np.random.default_rng(100 + nrhs)
        P = make_block(ctx, [rng.standard_normal((ctx.veo, 6)) for _ in range(nrhs)], prec)
        R0 = [rng.standard_normal((ctx.veo, 6)) for _ in range(nrhs)]
        Ra, Rb = make_block(ctx, R0, prec), make_block(ctx, R0, prec)
        V, tmp = ctx.buffer(P.nbytes), ctx.buffer(2 * P.nbytes)
        sigma = np.sort(rng.uniform(1e-3, 1.0, nrhs))
        sig = sigma.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        r2v = rng.uniform(0.5, 2.0, nrhs)
        act = np.ones(nrhs)
        act[nrhs - 1] = 0.
        sc = ctx.scalars(2 * nrhs + 2 * nrhs + 2 * nrhs + nrhs + nrhs)
        sd = ctx.scalars(2 * nrhs + 2 * nrhs + 2 * nrhs + nrhs + nrhs)
        for s in (sc, sd):
            s.store(np.concatenate([np.zeros(6 * nrhs), r2v, act]))
        pap, alpha, norms, r2, active = (lambda s: (s.at(0), s.at(2 * nrhs), s.at(4 * nrhs), s.at(6 * nrhs),
                                                    s.at(7 * nrhs)))(sc)
        ctx.MdagM_eo_mrhs_cg_step(Ra.ptr, P.ptr, gf.ptr, fmt, mass, upper_left, sig, tmp.ptr, nrhs, r2, active, pap,
                                  alpha, norms)
        pap2, alpha2, norms2, r22, active2 = (lambda s: (s.at(0), s.at(2 * nrhs), s.at(4 * nrhs), s.at(6 * nrhs),
                                                         s.at(7 * nrhs)))(sd)
        ctx.MdagM_eo_mrhs(V.ptr, P.ptr, gf.ptr, fmt, mass, upper_left, sig, tmp.ptr, nrhs)
        ctx.block_dot_stagg_eo(pap2, P.ptr, V.ptr, prec, nrhs)
        ctx.block_cg_alpha(alpha2, r22, pap2, active2, nrhs)
        ctx.block_axpy_stagg_eo(Rb.ptr, V.ptr, Rb.ptr, prec, nrhs, alpha2, -1.0, norms2)
        a, b = sc.get(), sd.get()
        for k in range(nrhs):
            assert abs(a[2 * k] - b[2 * k]) <= tol * abs(b[2 * k]), ("pAp", k)
            assert a[2 * k + 1] == 0.0
            assert abs(a[2 * nrhs + 2 * k] - b[2 * nrhs + 2 * k]) <= tol * abs(b[2 * nrhs + 2 * k]) + 1e-300, ("alpha", k)
            assert abs(a[4 * nrhs + 2 * k] - b[4 * nrhs + 2 * k]) <= tol * abs(b[4 * nrhs + 2 * k]), ("|r|^2", k)
            assert abs(a[4 * nrhs + 2 * k + 1] - b[4 * nrhs + 2 * k + 1]) <= tol * abs(b[4 * nrhs + 2 * k]), ("Re<>", k)
            ra, rb = slot(ctx, Ra, prec, nrhs, k), slot(ctx, Rb, prec, nrhs, k)
            assert site_rel_err(ra, rb) <= (1e-12 if prec == abi.F64 else 1e-5), ("r", k)
        # the inactive right-hand side: alpha 0, residual unchanged
        assert a[2 * nrhs + 2 * (nrhs - 1)] == 0.0
        assert np.array_equal(slot(ctx, Ra, prec, nrhs, nrhs - 1), slot(ctx, Rb, prec, nrhs, nrhs - 1))
        # and pAp against the FP64 oracle's <p, (m^2 + sigma - D D) p> for the first right-hand side
        p0 = slot(ctx, P, prec, nrhs, 0)
        ref = O.dot(p0, O.mdagm_eo(lat, p0, g, mass, upper_left) + sigma[0] * p0).real
        assert abs(a[0] - ref) <= (1e-12 if prec == abi.F64 else 1e-5) * ref
