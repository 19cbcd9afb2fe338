"""Fused multi-shift update, staggered reliable-update CG and the mixed-precision
multi-shift solver on the GPU.

* the fused update equals, per shift, the reference's saxpy / saxpby pair
  (solver_shifted.cpp:184-198) computed in numpy, and leaves inactive shifts alone;
* every shift's solution from cg_reliable_update_stagg and cg_m_mixed meets
  |b - (MdagM + sigma_k) x_k| / |b| <= 1e-10 with the FP64 oracle operator, and
  cg_m_mixed agrees with the FP64 reference-algorithm cg_m to 1e-8.
"""
import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import abi
from conftest import CONF_00200

pytestmark = pytest.mark.gpu

SIGMA = [0.0002065381736724, 0.00302707751065980, 0.0200732678058145, 0.12517586269872370, 1.0029328743375700]


@pytest.mark.parametrize("x_prec,prec", [(abi.F64, abi.F64), (abi.F32, abi.F32), (abi.F64, abi.F32)])
def test_multishift_update(x_prec, prec):
    import ctypes
    n = 5
    with abi.Context(4, 8) as ctx:
        rng = np.random.default_rng(1)
        R = rng.standard_normal((ctx.veo, 6))
        X = [rng.standard_normal((ctx.veo, 6)) for _ in range(n)]
        P = [rng.standard_normal((ctx.veo, 6)) for _ in range(n)]
        r = ctx.su3vec(prec).upload(R)
        xs = [ctx.su3vec(x_prec).upload(x) for x in X]
        ps = [ctx.su3vec(prec).upload(p) for p in P]
        if prec == abi.F32:  # the values the device holds
            R = r.download()
            X = [x.download() for x in xs]
            P = [p.download() for p in ps]
        beta, zeta, alpha = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
        active = np.array([1.0, 0.0, 1.0, 1.0, 0.0])
        dev = ctx.scalars(4 * n)
        dev.store(np.concatenate([beta, zeta, alpha, active]))
        xp = (ctypes.c_void_p * n)(*[x.ptr for x in xs])
        pp = (ctypes.c_void_p * n)(*[p.ptr for p in ps])
        ctx.multishift_update_stagg_eo(xp, x_prec, pp, r.ptr, prec, n, dev.at(0), dev.at(n), dev.at(2 * n), dev.at(3 * n))
        tol = 1e-15 if prec == abi.F64 else 1e-6
        for k in range(n):
            xn, pn = xs[k].download(), ps[k].download()
            # FP64 x with FP32 p: the product is formed in FP64 from the FP32 value of p
            if active[k]:
                xr = beta[k] * P[k] + X[k]
                pr = zeta[k] * R + alpha[k] * P[k]
            else:
                xr, pr = X[k], P[k]
            assert np.abs(xn - xr).max() <= tol * max(1.0, np.abs(xr).max())
            assert np.abs(pn - pr).max() <= tol * max(1.0, np.abs(pr).max())


def residuals(l, gf_host, xs, b, sigma, mass):
    lat = O.lattice(l.nspace, l.ntime)
    bh = b.get_host()
    out = []
    for s, x in zip(sigma, xs):
        xh = x.get_host()
        ax = O.mdagm_eo(lat, xh, gf_host, mass) + s * xh
        out.append(np.sqrt(O.norm2(ax - bh) / O.norm2(bh)))
    return out


@pytest.mark.parametrize("inner", ["f32", "f64"])
def test_cg_reliable_update_stagg(inner):
    from cl2qcd_amd import _host as H
    l = H.Lattice(["--nTime=4", "--fermionAction=rooted_stagg", "--mass=0.05"])
    gf = H.Gaugefield(l, CONF_00200)
    A = H.MdagM_eo(l)
    b, x = H.Staggeredfield_eo(l), H.Staggeredfield_eo(l)
    H.pseudo_randomize(b, 17)
    x.set_zero()
    st = H.cg_reliable_update_stagg(x, A, gf, 0.01, b, l, 1e-10, l.staggered_parameters(),
                                    H.F32 if inner == "f32" else H.F64)
    assert st["true_relative_residual"] <= 1e-10
    g = O.ildg_read(CONF_00200, O.lattice(4, 4))
    assert residuals(l, g, [x], b, [0.01], 0.05)[0] <= 1.01e-10


def test_cg_m_mixed_against_reference_cgm():
    from cl2qcd_amd import _host as H
    l = H.Lattice(["--nTime=4", "--fermionAction=rooted_stagg", "--mass=0.567", "--solverMaxIterations=5000"])
    gf = H.Gaugefield(l, CONF_00200)
    A = H.MdagM_eo(l)
    b = H.Staggeredfield_eo(l)
    H.pseudo_randomize(b, 13)
    xs = [H.Staggeredfield_eo(l) for _ in SIGMA]
    total, ms, per = H.cg_m_mixed(xs, A, gf, SIGMA, b, l, 1e-10, l.staggered_parameters())
    assert ms > 0 and all(p["true_relative_residual"] <= 1e-10 for p in per)
    g = O.ildg_read(CONF_00200, O.lattice(4, 4))
    assert max(residuals(l, g, xs, b, SIGMA, 0.567)) <= 1.01e-10
    ref = [H.Staggeredfield_eo(l) for _ in SIGMA]
    H.cg_m(ref, A, gf, SIGMA, b, l, 1e-24 * H.squarenorm(b), l.staggered_parameters())
    for x, r in zip(xs, ref):
        xh, rh = x.get_host(), r.get_host()
        assert np.sqrt(((xh - rh) ** 2).sum() / (rh ** 2).sum()) < 1e-8


def test_cg_m_mixed_graph_replays_match_eager():
    """The recorded CG-M iteration, replayed ten at a time between convergence checks, gives the same
    solutions bit for bit as the same iteration run eagerly."""
    from cl2qcd_amd import _host as H
    l = H.Lattice(["--nTime=4", "--fermionAction=rooted_stagg", "--mass=0.567", "--solverMaxIterations=5000"])
    gf = H.Gaugefield(l, CONF_00200)
    A = H.MdagM_eo(l)
    b = H.Staggeredfield_eo(l)
    H.pseudo_randomize(b, 14)
    out = []
    for use_graph in (True, False, True):
        xs = [H.Staggeredfield_eo(l) for _ in SIGMA]
        total, ms, per = H.cg_m_mixed(xs, A, gf, SIGMA, b, l, 1e-10, l.staggered_parameters(), use_graph=use_graph)
        assert all(p["true_relative_residual"] <= 1e-10 for p in per)
        out.append((total, ms, [x.get_host() for x in xs]))
    for total, ms, xs in out[1:]:
        assert (total, ms) == out[0][:2] and all(np.array_equal(a, b) for a, b in zip(xs, out[0][2]))


@pytest.mark.parametrize("n", [1, 5, 16, 64])
def test_fused_cgm_scalar_steps_equal_reference_sequence(n):
    """cl2_cgm_beta_step and cl2_cgm_alpha_step give, bit for bit, the values of the reference's per-iteration
    scalar sequence (solver_shifted.cpp:136-198): copy, ratio and subtraction for beta; copy and ratio for
    alpha; update_{zeta,beta,alpha}_cgm; and the copies zeta_prev = zeta, zeta = zeta_foll, rr_old = rr_new.
    Three iterations in a row, so every updated value feeds the next step."""
    rng = np.random.default_rng(n)
    # layout: 6 vectors of n (zeta, zeta_prev, zeta_foll, beta_vec, alpha_vec, sigma), then scalars
    V = ["zeta", "zeta_prev", "zeta_foll", "beta_vec", "alpha_vec", "sigma"]
    S = ["beta", "beta_prev", "alpha", "alpha_prev", "rr_old", "rr_new", "pap", "zero"]
    init = np.concatenate([rng.uniform(0.5, 1.0, n), rng.uniform(0.5, 1.0, n), np.zeros(3 * n),
                           np.sort(rng.uniform(1e-4, 1.0, n)), [-0.9, 0.0, 0.3, 0.0, 2.0, 0.0, 0.0, 0.0]])
    with abi.Context(4, 4) as ctx:
        fused, ref = ctx.scalars(len(init)), ctx.scalars(len(init))
        fused.store(init)
        ref.store(init)

        def at(d, name):
            return d.at(V.index(name) * n) if name in V else d.at(6 * n + S.index(name))

        for it in range(3):
            rr_new, pap = 1.1 * 0.8 ** it, 1.7 + 0.1 * it
            for d in (fused, ref):  # the two reductions of the iteration write their results
                vals = d.get()
                vals[6 * n + S.index("rr_new")] = rr_new
                vals[6 * n + S.index("pap")] = pap
                d.store(vals)
            ctx.cgm_beta_step(at(fused, "beta"), at(fused, "beta_prev"), at(fused, "rr_old"), at(fused, "pap"))
            ctx.cgm_alpha_step(n, at(fused, "alpha"), at(fused, "alpha_prev"), at(fused, "rr_new"),
                               at(fused, "rr_old"), at(fused, "beta"), at(fused, "beta_prev"), at(fused, "sigma"),
                               at(fused, "zeta"), at(fused, "zeta_prev"), at(fused, "zeta_foll"),
                               at(fused, "beta_vec"), at(fused, "alpha_vec"))
            r = lambda name: at(ref, name)  # noqa: E731
            ctx.scalar_copy(r("beta_prev"), r("beta"), 1)
            ctx.real_op(abi.SCALAR_OP_RATIO, r("beta"), r("rr_old"), r("pap"))
            ctx.real_op(abi.SCALAR_OP_SUBTRACTION, r("beta"), r("zero"), r("beta"))
            ctx.scalar_copy(r("alpha_prev"), r("alpha"), 1)
            ctx.real_op(abi.SCALAR_OP_RATIO, r("alpha"), r("rr_new"), r("rr_old"))
            ctx.update_zeta_cgm(r("zeta_foll"), r("zeta"), r("zeta_prev"), r("beta_prev"), r("beta"),
                                r("alpha_prev"), r("sigma"), n)
            ctx.update_beta_cgm(r("beta_vec"), r("beta"), r("zeta_foll"), r("zeta"), n)
            ctx.update_alpha_cgm(r("alpha_vec"), r("alpha"), r("zeta_foll"), r("beta_vec"), r("zeta"), r("beta"), n)
            ctx.scalar_copy(r("zeta_prev"), r("zeta"), n)
            ctx.scalar_copy(r("zeta"), r("zeta_foll"), n)
            ctx.scalar_copy(r("rr_old"), r("rr_new"), 1)
            a, b = fused.get(), ref.get()
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (it, np.flatnonzero(a != b))
        # the same three steps in numpy from the closed forms (operations_real.cl:20-39)
        z, zp, sig = init[:n].copy(), init[n:2 * n].copy(), init[5 * n:6 * n]
        beta, alpha, rr_old = -0.9, 0.3, 2.0
        for it in range(3):
            rr_new, pap = 1.1 * 0.8 ** it, 1.7 + 0.1 * it
            beta_prev, beta = beta, -(rr_old / pap)
            alpha_prev, alpha = alpha, rr_new / rr_old
            zf = (z * zp * beta_prev) / (beta * alpha_prev * (zp - z) + zp * beta_prev * (1 - sig * beta))
            bv = -beta * zf / z
            av = -alpha * zf * bv / (z * beta)
            zp, z, rr_old = z, zf, rr_new
        got = a[:6 * n].reshape(6, n)
        for want, row in ((z, 0), (zp, 1), (zf, 2), (bv, 3), (av, 4)):
            assert np.abs(got[row] - want).max() <= 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("upper_left", [abi.EVEN, abi.ODD])
@pytest.mark.parametrize("sigma", [0.0, 0.3])
@pytest.mark.parametrize("fmt", [abi.GAUGE_F64_18, abi.GAUGE_F32_12])
def test_mdagm_cgm_step_matches_unfused_sequence(fmt, sigma, upper_left):
    """cl2_MdagM_eo_cgm_step against the sequence it replaces in a CG-M iteration (solver_shifted.cpp:136-198):
    MdagM_eo (+ sigma, as the reduced-precision CG on a shifted system uses it), <p, A p>, cl2_cgm_beta_step,
    r = beta A p + r and |r|^2, with boundary phases and an imaginary chemical potential; pAp =
    (m^2 + sigma) |p|^2 + |D p|^2 agrees with <p, A p> to rounding (D_KS_eo anti-hermitian), and so do beta, the
    updated r and |r|^2; beta_prev takes the old beta exactly."""
    from fields import random_su3_field, site_rel_err
    kw = dict(theta_s=0.3, theta_t=1.0, cp_im=0.2)
    lat = O.lattice(8, 4, **kw)
    prec = abi.F32 if fmt == abi.GAUGE_F32_12 else abi.F64
    tol = 1e-12 if prec == abi.F64 else 2e-5
    mass = 0.05
    with abi.Context(8, 4, **kw) as ctx:
        g = random_su3_field(4 * O.volume(lat), 71)
        g64 = ctx.gauge(g)
        gf = g64 if fmt == abi.GAUGE_F64_18 else g64.converted(fmt)
        rng = np.random.default_rng(5)
        pv, rv = rng.standard_normal((ctx.veo, 6)), rng.standard_normal((ctx.veo, 6))
        p = ctx.su3vec(prec).upload(pv)
        ra, rb = ctx.su3vec(prec).upload(rv), ctx.su3vec(prec).upload(rv)
        v, t1, t2 = ctx.su3vec(prec), ctx.su3vec(prec), ctx.su3vec(prec)
        # scalars: rr_old, pap, beta, beta_prev, rr_new
        sa, sb = ctx.scalars(5), ctx.scalars(5)
        for s in (sa, sb):
            s.store([1.7, 0.0, -0.4, 0.0, 0.0])
        ctx.MdagM_eo_cgm_step(ra.ptr, p.ptr, gf.ptr, fmt, mass, sigma, upper_left, t1.ptr, sa.at(0), sa.at(1),
                              sa.at(2), sa.at(3), sa.at(4))
        ctx.MdagM_eo(v.ptr, p.ptr, gf.ptr, fmt, mass, upper_left, t2.ptr)
        if sigma:
            ctx.saxpy_stagg_eo(v.ptr, p.ptr, v.ptr, prec, abi.scalar(sigma))
        ctx.scalar_product_real_part_stagg_eo(sb.at(1), p.ptr, v.ptr, prec)
        ctx.cgm_beta_step(sb.at(2), sb.at(3), sb.at(0), sb.at(1))
        ctx.saxpy_stagg_eo(rb.ptr, v.ptr, rb.ptr, prec, abi.scalar(dev=sb.at(2), kind=abi.SCALAR_DEV_REAL))
        ctx.squarenorm_stagg_eo(sb.at(4), rb.ptr, prec)
        a, b = sa.get(), sb.get()
        assert a[0] == b[0] == 1.7 and a[3] == b[3] == -0.4
        for k, name in ((1, "pap"), (2, "beta"), (4, "|r|^2")):
            assert abs(a[k] - b[k]) <= tol * abs(b[k]), (name, a[k], b[k])
        assert site_rel_err(ra.download(), rb.download()) <= (1e-12 if prec == abi.F64 else 1e-5)
        # pap against the oracle's <p, MdagM p> on the stored p
        ps = p.download()
        ref = O.dot(ps, O.mdagm_eo(lat, ps, g, mass, upper_left) + sigma * ps).real
        assert abs(a[1] - ref) <= (1e-12 if prec == abi.F64 else 1e-5) * ref
