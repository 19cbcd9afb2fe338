"""Multi-right-hand-side operators, block BLAS and the lockstep block CG on the GPU.

Every right-hand side of a block launch runs the same arithmetic as the
single-field kernel, so slot r of the multi-RHS result must equal the
single-field result for source r bit for bit; block reductions are checked per right-hand side against
numpy; the block solver's solutions meet |b - A x| / |b| <= 1e-10 in the FP64
oracle and agree with single-right-hand-side solves.
"""
import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import abi
from fields import half_links, random_su3_field, site_rel_err

pytestmark = pytest.mark.gpu


def make_block(ctx, srcs, prec):
    n = len(srcs)
    blk = ctx.buffer(n * ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, prec))
    for r, s in enumerate(srcs):
        f = ctx.spinor(prec).upload(s)
        ctx.block_insert_eo(blk.ptr, f.ptr, prec, n, r)
        f.close()
    return blk


def block_slot(ctx, blk, prec, n, r):
    f = ctx.spinor(prec)
    ctx.block_extract_eo(f.ptr, blk.ptr, prec, n, r)
    out = f.download()
    f.close()
    return out


@pytest.mark.parametrize("lds", [1, 2, 0, -1, -2])
@pytest.mark.parametrize("nrhs", [2, 12, 16])
@pytest.mark.parametrize("fmt", [abi.GAUGE_F64_18, abi.GAUGE_F64_12, abi.GAUGE_F32_12])
def test_dslash_mrhs_matches_single(nrhs, fmt, lds):
    """Slot r of the multi-RHS result is the single-field result for source r bit for bit, with the links
    staged in LDS per site (lds=1, the default; lds=2 with twice the sites per block) or loaded by every
    right-hand-side thread (lds=0); -1 / -2: lds 1 / 2 with the 64-bit-offset gathers ("mrhs.off32" 0) that
    FP64 fields above 4 GiB per component take."""
    lat = O.lattice(8, 8, theta_t=1.0)
    prec = abi.F32 if fmt == abi.GAUGE_F32_12 else abi.F64
    with abi.Context(8, 8, theta_t=1.0) as ctx:
        ctx.check(ctx.lib.cl2_tune_set(ctx.handle, b"mrhs.lds", abs(lds)))
        ctx.check(ctx.lib.cl2_tune_set(ctx.handle, b"mrhs.off32", 0 if lds < 0 else 1))
        g = random_su3_field(4 * O.volume(lat), 5)
        gf = ctx.gauge(g, fmt)
        rng = np.random.default_rng(nrhs)
        srcs = [rng.standard_normal((ctx.veo, 24)) for _ in range(nrhs)]
        blk = make_block(ctx, srcs, prec)
        out = ctx.buffer(blk.nbytes)
        for evenodd in (abi.EVEN, abi.ODD):
            ctx.dslash_eo_mrhs(out.ptr, blk.ptr, gf.ptr, fmt, evenodd, 0.16, nrhs)
            a, b = ctx.spinor(prec), ctx.spinor(prec)
            for r in range(nrhs):
                a.upload(srcs[r])
                ctx.dslash_eo(b.ptr, a.ptr, gf.ptr, fmt, evenodd, 0.16)
                single, multi = b.download(), block_slot(ctx, out, prec, nrhs, r)
                assert np.array_equal(single, multi)
            if prec == abi.F64 and evenodd == abi.EVEN:
                ref = O.dslash_eo(lat, srcs[nrhs - 1], g, evenodd, 0.16)
                assert site_rel_err(block_slot(ctx, out, prec, nrhs, nrhs - 1), ref) < 1e-12


@pytest.mark.parametrize("phases", [dict(theta_t=1.0), dict(theta_t=1.0, theta_s=0.3, cp_re=0.11)],
                         ids=["real-spatial-weights", "complex-weights"])
@pytest.mark.parametrize("nrhs", [2, 12, 16])
def test_dslash_mrhs_f16_matches_single(nrhs, phases):
    """FP16 blocks (one FP16 field of Veo nrhs sites with per-site norms) and FP16 links: slot r of the
    multi-RHS result equals the single-field FP16 kernel's result for source r bit for bit, both parities,
    with real spatial hop weights and with a spatial boundary phase and chemical potential (complex weights
    in every direction)."""
    lat = O.lattice(8, 8, **phases)
    with abi.Context(8, 8, **phases) as ctx:
        g = random_su3_field(4 * O.volume(lat), 5)
        gf = ctx.gauge(g, abi.GAUGE_F16_12)
        rng = np.random.default_rng(nrhs)
        srcs = [rng.standard_normal((ctx.veo, 24)) for _ in range(nrhs)]
        blk = make_block(ctx, srcs, abi.F16)
        out = ctx.buffer(blk.nbytes)
        a, b = ctx.spinor(abi.F16), ctx.spinor(abi.F16)
        for evenodd in (abi.EVEN, abi.ODD):
            ctx.dslash_eo_mrhs(out.ptr, blk.ptr, gf.ptr, abi.GAUGE_F16_12, evenodd, 0.16, nrhs)
            for r in range(nrhs):
                a.upload(srcs[r])
                ctx.dslash_eo(b.ptr, a.ptr, gf.ptr, abi.GAUGE_F16_12, evenodd, 0.16)
                assert np.array_equal(b.download(), block_slot(ctx, out, abi.F16, nrhs, r))
            # and the last slot against exact arithmetic on the stored values and links
            ref = O.dslash_eo(lat, a.download(), half_links(g), evenodd, 0.16)
            assert site_rel_err(block_slot(ctx, out, abi.F16, nrhs, nrhs - 1), ref) < 1e-3


@pytest.mark.parametrize("action", [abi.ACTION_WILSON, abi.ACTION_TWISTEDMASS])
def test_fermionmatrix_mrhs(action):
    lat = O.lattice(8, 8, theta_t=1.0)
    nrhs = 4
    with abi.Context(8, 8, theta_t=1.0) as ctx:
        g = random_su3_field(4 * O.volume(lat), 6)
        gf = ctx.gauge(g)
        rng = np.random.default_rng(7)
        srcs = [rng.standard_normal((ctx.veo, 24)) for _ in range(nrhs)]
        blk = make_block(ctx, srcs, abi.F64)
        out = ctx.buffer(blk.nbytes)
        tmp = ctx.buffer(2 * blk.nbytes)
        mubar = 0.0016 if action == abi.ACTION_TWISTEDMASS else 0.0
        act = O.TWISTEDMASS if action == abi.ACTION_TWISTEDMASS else O.WILSON
        ctx.fermionmatrix_eo_mrhs(abi.OP_QPLUSQMINUS_EO, out.ptr, blk.ptr, gf.ptr, abi.GAUGE_F64_18, action, 0.16,
                                  mubar, tmp.ptr, nrhs)
        for r in range(nrhs):
            ref = O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, srcs[r], g, act, 0.16, mubar)
            assert site_rel_err(block_slot(ctx, out, abi.F64, nrhs, r), ref) < 1e-12


@pytest.mark.parametrize("nrhs", [1, 3, 12])
@pytest.mark.parametrize("prec", [abi.F64, abi.F32, abi.F16])
def test_block_blas(prec, nrhs):
    """Block dot, norm and axpy (with its norms) per right-hand side against numpy on the stored values:
    FP64 to 1e-13, FP32 to 1e-6; FP16 outputs are requantised per site (1e-3), their norms are summed before
    that (1e-5)."""
    with abi.Context(4, 8) as ctx:
        rng = np.random.default_rng(8)
        X = [rng.standard_normal((ctx.veo, 24)) for _ in range(nrhs)]
        Y = [rng.standard_normal((ctx.veo, 24)) for _ in range(nrhs)]
        bx, by = make_block(ctx, X, prec), make_block(ctx, Y, prec)
        if prec != abi.F64:  # the values the block holds
            X = [block_slot(ctx, bx, prec, nrhs, r) for r in range(nrhs)]
            Y = [block_slot(ctx, by, prec, nrhs, r) for r in range(nrhs)]
        coef = rng.standard_normal(2 * nrhs)
        cdev, norms, res = ctx.scalars(2 * nrhs), ctx.scalars(2 * nrhs), ctx.scalars(2 * nrhs)
        cdev.store(coef)
        tol = {abi.F64: 1e-13, abi.F32: 1e-6, abi.F16: 1e-3}[prec]
        ntol = 1e-5 if prec == abi.F16 else tol
        ctx.block_dot_eo(res.ptr, bx.ptr, by.ptr, prec, nrhs)
        d = res.get()
        for r in range(nrhs):
            ref = O.dot(X[r], Y[r])
            assert abs(complex(d[2 * r], d[2 * r + 1]) - ref) < 1e-12 * np.sqrt(O.norm2(X[r]) * O.norm2(Y[r]))
        ctx.block_squarenorm_eo(res.ptr, bx.ptr, prec, nrhs)
        d = res.get()
        for r in range(nrhs):
            assert abs(d[r] - O.norm2(X[r])) < 1e-12 * O.norm2(X[r])
        # y <- y - c_r x in place, with (|y_new|^2, Re <y_new, y_old>)
        ctx.block_axpy_eo(by.ptr, bx.ptr, by.ptr, prec, nrhs, cdev.ptr, -1.0, norms.ptr)
        nm = norms.get()
        for r in range(nrhs):
            c = complex(coef[2 * r], coef[2 * r + 1])
            new = O.saxpy(X[r], Y[r], c)
            assert site_rel_err(block_slot(ctx, by, prec, nrhs, r), new) < tol
            assert abs(nm[2 * r] - O.norm2(new)) < ntol * O.norm2(new)
            assert abs(nm[2 * r + 1] - O.dot(new, Y[r]).real) < ntol * O.norm2(new)


@pytest.mark.parametrize("nrhs", [1, 3, 12])
@pytest.mark.parametrize("prec", [abi.F64, abi.F32, abi.F16])
def test_block_cg_update_xp(prec, nrhs):
    """x += alpha_r p, p = r + beta_r p in one pass: the same bits as the two block axpys it replaces."""
    with abi.Context(4, 8) as ctx:
        rng = np.random.default_rng(9)
        X, P, R = ([rng.standard_normal((ctx.veo, 24)) for _ in range(nrhs)] for _ in range(3))
        x1, p1, r = make_block(ctx, X, prec), make_block(ctx, P, prec), make_block(ctx, R, prec)
        x2, p2 = make_block(ctx, X, prec), make_block(ctx, P, prec)
        al, be = ctx.scalars(2 * nrhs), ctx.scalars(2 * nrhs)
        al.store(rng.standard_normal(2 * nrhs))
        be.store(rng.standard_normal(2 * nrhs))
        ctx.block_axpy_eo(x1.ptr, p1.ptr, x1.ptr, prec, nrhs, al.ptr, 1.0, None)
        ctx.block_axpy_eo(p1.ptr, p1.ptr, r.ptr, prec, nrhs, be.ptr, 1.0, None)
        ctx.block_cg_update_xp_eo(x2.ptr, p2.ptr, r.ptr, prec, nrhs, al.ptr, be.ptr)
        for k in range(nrhs):
            assert np.array_equal(block_slot(ctx, x1, prec, nrhs, k), block_slot(ctx, x2, prec, nrhs, k))
            assert np.array_equal(block_slot(ctx, p1, prec, nrhs, k), block_slot(ctx, p2, prec, nrhs, k))
        if prec == abi.F64:
            a, b = al.get(), be.get()
            for k in range(nrhs):
                ca, cb = complex(a[2 * k], a[2 * k + 1]), complex(b[2 * k], b[2 * k + 1])
                assert site_rel_err(block_slot(ctx, x2, prec, nrhs, k), O.saxpy(P[k], X[k], -ca)) < 1e-15
                assert site_rel_err(block_slot(ctx, p2, prec, nrhs, k), O.saxpy(P[k], R[k], -cb)) < 1e-15


@pytest.mark.parametrize("inner", ["f32", "f16"])
def test_block_cg_matches_single_solves(inner):
    """The lockstep block solve with an FP32 or FP16 (per-site norm) inner loop: every right-hand side meets
    1e-10 in the oracle's FP64 operator and agrees with its own single solve (same inner precision) to 1e-6."""
    from cl2qcd_amd import _host as H
    prec = H.F16 if inner == "f16" else H.F32
    l = H.Lattice(["--nSpace=8", "--nTime=8", "--fermionAction=twistedmass", "--kappa=0.15", "--mu=0.02",
                   "--beta=5.7", "--thetaFermionTemporal=1"])
    gf = H.Gaugefield(l, False)
    for _ in range(3):
        H.su3heatbath(gf, l, 1)
    nrhs = 4
    rng = np.random.default_rng(9)
    bs, xs = [], []
    for r in range(nrhs):
        b = H.Spinorfield_eo(l)
        b.import_(rng.standard_normal((l.eo_sites, 24)))
        bs.append(b)
        x = H.Spinorfield_eo(l)
        x.zero()
        xs.append(x)
    A = H.QplusQminus_eo(l)
    it, per = H.cg_reliable_update_block(xs, A, gf, bs, l, 1e-10, l.wilson_parameters(), prec)
    lat = O.lattice(8, 8, theta_t=1.0)
    g = gf.get_host()
    for r in range(nrhs):
        assert per[r]["true_relative_residual"] <= 1e-10
        bh, xh = bs[r].get_host(), xs[r].get_host()
        res = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, xh, g, O.TWISTEDMASS, 0.15, l.mubar), bh, 1.0)
        assert np.sqrt(O.norm2(res) / O.norm2(bh)) <= 1.01e-10
        xs1 = H.Spinorfield_eo(l)
        xs1.zero()
        H.cg_reliable_update(xs1, A, gf, bs[r], l, 1e-10, l.wilson_parameters(), prec)
        x1 = xs1.get_host()
        assert np.sqrt(((xh - x1) ** 2).sum() / (x1 ** 2).sum()) < 1e-6
    assert it <= 1.5 * max(p["iterations"] for p in per)


@pytest.mark.parametrize("op", [abi.OP_AEE, abi.OP_QPLUSQMINUS_EO])
@pytest.mark.parametrize("fmt", [abi.GAUGE_F64_18, abi.GAUGE_F32_12])
def test_fermionmatrix_mrhs_per_rhs_params(op, fmt):
    """One block launch with kappa and mubar per right-hand side: slot r equals the single-field operator with
    that right-hand side's parameters (the hop sum is scaled by kappa after summing instead of inside each hop
    weight, so to rounding) and, in FP64, the oracle."""
    import ctypes
    lat = O.lattice(8, 8, theta_t=1.0)
    prec = abi.F32 if fmt == abi.GAUGE_F32_12 else abi.F64
    kappas = [0.15, 0.155, 0.16, 0.145]
    mubars = [2 * k * m for k, m in zip(kappas, [0.003, 0.02, 0.01, 0.0])]
    nrhs = len(kappas)
    with abi.Context(8, 8, theta_t=1.0) as ctx:
        g = random_su3_field(4 * O.volume(lat), 12)
        gf = ctx.gauge(g, fmt)
        rng = np.random.default_rng(13)
        srcs = [rng.standard_normal((ctx.veo, 24)) for _ in range(nrhs)]
        blk = make_block(ctx, srcs, prec)
        out = ctx.buffer(blk.nbytes)
        tmp = ctx.buffer(2 * blk.nbytes)
        ka, ma = (ctypes.c_double * nrhs)(*kappas), (ctypes.c_double * nrhs)(*mubars)
        ctx.fermionmatrix_eo_mrhs_params(op, out.ptr, blk.ptr, gf.ptr, fmt, abi.ACTION_TWISTEDMASS, ka, ma, tmp.ptr,
                                         nrhs)
        a, b = ctx.spinor(prec), ctx.spinor(prec)
        stmp = ctx.buffer(2 * ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F64))
        for r in range(nrhs):
            a.upload(srcs[r])
            ctx.fermionmatrix_eo(op, b.ptr, a.ptr, gf.ptr, fmt, abi.ACTION_TWISTEDMASS, kappas[r], mubars[r], stmp.ptr)
            got = block_slot(ctx, out, prec, nrhs, r)
            assert site_rel_err(got, b.download()) < (1e-14 if prec == abi.F64 else 1e-6)
            if prec == abi.F64:
                ref = O.apply_eo(lat, op, srcs[r], g, O.TWISTEDMASS, kappas[r], mubars[r])
                assert site_rel_err(got, ref) < 1e-12


def test_block_cg_per_rhs_params():
    """The lockstep block solver with a different (kappa, mu) per right-hand side: every solution meets
    |b - A_r x| / |b| <= 1e-10 with its own operator in the FP64 oracle and agrees with a single solve."""
    from cl2qcd_amd import _host as H
    l = H.Lattice(["--nSpace=8", "--nTime=8", "--fermionAction=twistedmass", "--kappa=0.15", "--mu=0.02",
                   "--beta=5.7", "--thetaFermionTemporal=1"])
    gf = H.Gaugefield(l, False)
    for _ in range(3):
        H.su3heatbath(gf, l, 1)
    kappas = [0.15, 0.155, 0.145, 0.152]
    mus = [0.02, 0.005, 0.01, 0.015]
    mubars = [2 * k * m for k, m in zip(kappas, mus)]
    rng = np.random.default_rng(14)
    bs, xs = [], []
    for r in range(4):
        b = H.Spinorfield_eo(l)
        b.import_(rng.standard_normal((l.eo_sites, 24)))
        bs.append(b)
        x = H.Spinorfield_eo(l)
        x.zero()
        xs.append(x)
    A = H.QplusQminus_eo(l)
    it, per = H.cg_reliable_update_block_params(xs, A, gf, bs, l, 1e-10, kappas, mubars)
    lat = O.lattice(8, 8, theta_t=1.0)
    g = gf.get_host()
    for r in range(4):
        assert per[r]["true_relative_residual"] <= 1e-10
        bh, xh = bs[r].get_host(), xs[r].get_host()
        res = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, xh, g, O.TWISTEDMASS, kappas[r], mubars[r]), bh, 1.0)
        assert np.sqrt(O.norm2(res) / O.norm2(bh)) <= 1.01e-10
    x1 = H.Spinorfield_eo(l)
    x1.zero()
    H.cg_reliable_update(x1, A, gf, bs[1], l, 1e-10, H.WilsonAdditionalParameters(kappas[1], mubars[1]))
    assert np.sqrt(((xs[1].get_host() - x1.get_host()) ** 2).sum() / (x1.get_host() ** 2).sum()) < 1e-6


@pytest.mark.parametrize("ns,nt", [(8, 8), (6, 10)])  # 6^3x10: a partial last block of staged sites
@pytest.mark.parametrize("per_rhs", [False, True])
@pytest.mark.parametrize("nrhs", [4, 12])
@pytest.mark.parametrize("fmt", [abi.GAUGE_F64_18, abi.GAUGE_F32_12, abi.GAUGE_F16_12, "F64_18 64-bit offsets"])
def test_mrhs_cg_step(fmt, nrhs, per_rhs, ns, nt):
    """The fused lockstep CG step (pAp per right-hand side from |Qminus p_r|^2, alpha, r_r -= alpha_r A p_r
    with its norms, A p never stored) against the unfused block sequence (operator, block dot, alpha, block
    axpy with norms): FP64 to 1e-12, FP32 to 1e-5 (pAp is summed from another field, so alpha and everything
    after it differ by rounding), FP16 to 5e-3 (the unfused sequence also stores A p in half precision); an
    inactive right-hand side keeps r and gets alpha 0; FP64 against the oracle."""
    import ctypes
    off32 = 0 if isinstance(fmt, str) else 1
    fmt = abi.GAUGE_F64_18 if isinstance(fmt, str) else fmt
    lat = O.lattice(ns, nt, theta_t=1.0)
    prec = {abi.GAUGE_F32_12: abi.F32, abi.GAUGE_F16_12: abi.F16}.get(fmt, abi.F64)
    tol = {abi.F64: 1e-12, abi.F32: 1e-5, abi.F16: 5e-3}[prec]
    rng = np.random.default_rng(nrhs)
    kappas = list(0.15 + 0.01 * rng.random(nrhs)) if per_rhs else [0.16] * nrhs
    mubars = [2 * k * m for k, m in zip(kappas, 0.003 + 0.017 * rng.random(nrhs))] if per_rhs else \
        [2 * 0.16 * 0.005] * nrhs
    with abi.Context(ns, nt, theta_t=1.0) as ctx:
        ctx.check(ctx.lib.cl2_tune_set(ctx.handle, b"mrhs.off32", off32))
        g = random_su3_field(4 * O.volume(lat), 14)
        gf = ctx.gauge(g, fmt)
        P = [rng.standard_normal((ctx.veo, 24)) for _ in range(nrhs)]
        R = [rng.standard_normal((ctx.veo, 24)) for _ in range(nrhs)]
        p, r1, r2b = make_block(ctx, P, prec), make_block(ctx, R, prec), make_block(ctx, R, prec)
        v, tmp = ctx.buffer(p.nbytes), ctx.buffer(2 * p.nbytes)
        n2 = 2 * nrhs
        sc = ctx.scalars(14 * nrhs)  # r2 | active | fused pAp, alpha, norms | unfused pAp, alpha, norms
        r2 = [O.norm2(x) for x in R]
        active = [1.0] * nrhs
        active[1] = 0.0
        sc.store(r2 + active + [0.0] * (12 * nrhs))
        at = lambda k: sc.at(2 * nrhs + k)
        ka, ma = (ctypes.c_double * nrhs)(*kappas), (ctypes.c_double * nrhs)(*mubars)
        if per_rhs:
            ctx.fermionmatrix_eo_mrhs_params(abi.OP_QPLUSQMINUS_EO, v.ptr, p.ptr, gf.ptr, fmt, abi.ACTION_TWISTEDMASS,
                                             ka, ma, tmp.ptr, nrhs)
        else:
            ctx.fermionmatrix_eo_mrhs(abi.OP_QPLUSQMINUS_EO, v.ptr, p.ptr, gf.ptr, fmt, abi.ACTION_TWISTEDMASS,
                                      kappas[0], mubars[0], tmp.ptr, nrhs)
        ctx.block_dot_eo(at(3 * n2), p.ptr, v.ptr, prec, nrhs)
        ctx.block_cg_alpha(at(4 * n2), sc.at(0), at(3 * n2), sc.at(nrhs), nrhs)
        ctx.block_axpy_eo(r1.ptr, v.ptr, r1.ptr, prec, nrhs, at(4 * n2), -1.0, at(5 * n2))
        ctx.fermionmatrix_eo_mrhs_cg_step(abi.OP_QPLUSQMINUS_EO, r2b.ptr, p.ptr, gf.ptr, fmt, abi.ACTION_TWISTEDMASS,
                                          kappas[0], mubars[0], ka if per_rhs else None, ma if per_rhs else None,
                                          tmp.ptr, nrhs, sc.at(0), sc.at(nrhs), at(0), at(n2), at(2 * n2))
        d = sc.get()[2 * nrhs:]
        fp, fa, fn = d[0:n2], d[n2:2 * n2], d[2 * n2:3 * n2]
        up, ua, un = d[3 * n2:4 * n2], d[4 * n2:5 * n2], d[5 * n2:6 * n2]
        for k in range(nrhs):
            assert abs(fp[2 * k] - up[2 * k]) < tol * up[2 * k] and fp[2 * k + 1] == 0.0
            got, ref = block_slot(ctx, r2b, prec, nrhs, k), block_slot(ctx, r1, prec, nrhs, k)
            if k == 1:
                assert fa[2] == 0.0 and fa[3] == 0.0 and np.array_equal(got, ref)
                continue
            assert abs(complex(fa[2 * k], fa[2 * k + 1]) - complex(ua[2 * k], ua[2 * k + 1])) < tol * abs(ua[2 * k])
            assert site_rel_err(got, ref) < 10 * tol
            assert abs(fn[2 * k] - un[2 * k]) < 10 * tol * un[2 * k]
            assert abs(fn[2 * k + 1] - un[2 * k + 1]) < 10 * tol * un[2 * k]
            if prec == abi.F64:
                Ap = O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, P[k], g, O.TWISTEDMASS, kappas[k], mubars[k])
                alpha = r2[k] / O.dot(P[k], Ap)
                assert site_rel_err(got, O.saxpy(Ap, R[k], alpha)) < 1e-12
