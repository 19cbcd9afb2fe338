"""Parity at the sizes and parameters of the workloads the performance numbers are quoted on.

* 8^4 plain CPU-reference workload exactly as written: a configuration from this build's su3heatbath at
  beta = 5.7 (cold start, 50 sweeps, fixed seed), twisted mass kappa = 0.15, mu = 0.01, antiperiodic in t,
  eo CG on QplusQminus_eo to |r|^2/|b|^2 <= 1e-20 from a seeded unit-variance Gaussian source, FP64 -- run
  on the oracle (the reference's CG restated, physics/algorithms/solvers/cg.cpp:277-400) and on the C++
  drop-in over the HIP kernels: the same iteration count within the 10-iteration check granularity, both
  true residuals below the bound (recomputed by the oracle), the same solution to 1e-9.
* 32^3x64, heatbath beta = 5.7: dslash_eo per site against the oracle on every output site, for every
  link format and every register-budget variant of the hopping kernel, in the autotuned launch shape too
  (FP64 18-real <= 1e-13, FP64 12-real <= 1e-12 relative per site; FP32 <= 2e-5; FP16 against the
  oracle on the quantised inputs <= 1e-3 and on the exact inputs <= 3e-3 over the field); FP64 eo-CG on
  QplusQminus_eo (kappa = 0.16, mu = 0.005) to |r|/|b| <= 1e-10 with the true residual recomputed by the
  oracle.
* 48^3x96, heatbath beta = 5.7: dslash_eo on a sample of output sites that covers the t-slices at both
  ends and the middle of the lattice, every z-plane on a boundary of the work order's z-chunks (one per
  XCD), the first and last x-rows of those planes in every t-slice, the first and last 64-site link tiles
  and uniformly drawn sites -- <= 1e-13 per site for 18-real links (the field is 6.1 GB, so link offsets
  pass 4 GiB), <= 1e-12 for 12-real links.
* The operators' byte and flop models against the reference's formulas
  (hardware/code/fermions.cpp:811-975, spinors.cpp:969-972,1085-1088, fermions_staggered.cpp:189-193,
  259-263, spinors_staggered.cpp:1712-1715,1853-1856; physics/fermionmatrix/fermionmatrix.cpp:264-676,
  fermionmatrix_stagg.cpp:62-104), restated here from the reference's constituent counts.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import abi
from fields import field_rel, half_links, site_rel_err

pytestmark = pytest.mark.gpu

BETA = 5.7


def heatbath_ctx(ns, nt, seed, sweeps=50):
    """abi context (antiperiodic in t) with a beta 5.7 heatbath configuration: 50 x (heatbath + 1
    overrelaxation) sweeps from a cold start, as the solver benchmarks make it"""
    ctx = abi.Context(ns, nt, theta_t=1.0)
    gf = ctx.gauge(seed=0, epsilon=0.0)  # cold start: every link the unit matrix
    gf.heatbath(BETA, sweeps, overrelax=1, seed=seed)
    return ctx, gf


def test_config1_twisted_mass_cg_8x8():
    from cl2qcd_amd import _host as H
    seed = 20261201
    l = H.Lattice(["--nSpace=8", "--nTime=8", f"--beta={BETA}", f"--hostSeed={seed}",
                   "--fermionAction=twistedmass", "--kappa=0.15", "--mu=0.01", "--thetaFermionTemporal=1",
                   "--solverMaxIterations=5000"])
    gf = H.Gaugefield(l, False)
    for _ in range(50):
        H.su3heatbath(gf, l, 0)
    g = np.ascontiguousarray(gf.get_host().reshape(-1, 18))
    lat = O.lattice(8, 8, theta_t=1.0)
    plaq = O.plaquette(lat, g)
    assert abs(H.measurePlaquette(gf) - plaq) < 1e-12
    assert 0.5 < plaq < 0.6  # beta 5.7 equilibrium plaquette is about 0.549
    mubar = l.mubar
    assert abs(mubar - 2 * 0.15 * 0.01) < 1e-15
    bh = np.random.default_rng(seed).standard_normal((l.eo_sites, 24))
    b, x = H.Spinorfield_eo(l), H.Spinorfield_eo(l)
    b.import_(bh)
    x.cold()
    x0 = x.get_host()
    prec = 1e-20 * O.norm2(bh)  # the reference's absolute criterion |r|^2 < prec
    it = H.cg(x, H.QplusQminus_eo(l), gf, b, l, prec, l.wilson_parameters())
    rc, it_o, xo = O.cg_eo(lat, O.OP_QPLUSQMINUS_EO, g, bh, x0, O.TWISTEDMASS, 0.15, mubar, prec=prec, cgmax=5000)
    assert rc == O.OK
    assert abs(it - it_o) <= 10, (it, it_o)
    xs = x.get_host()
    assert site_rel_err(xs, xo) < 1e-9
    for sol in (xs, xo):
        r = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, sol, g, O.TWISTEDMASS, 0.15, mubar), bh, 1.0)
        assert O.norm2(r) / O.norm2(bh) < 2e-20


@pytest.fixture(scope="module")
def conf32():
    ctx, gf = heatbath_ctx(32, 64, seed=20261203)
    g = gf.download()
    lat = O.lattice(32, 64, theta_t=1.0)
    yield lat, ctx, gf, g
    ctx.close()


def set_shape(ctx, occ, block=256, zplanes=-1, xcd=1):
    for key, v in (("hop.block", block), ("hop.occupancy", occ), ("hop.zplanes", zplanes), ("hop.xcd", xcd)):
        ctx.check(ctx.lib.cl2_tune_set(ctx.handle, key.encode(), v))


@pytest.mark.parametrize("fmt,prec", [(abi.GAUGE_F64_18, abi.F64), (abi.GAUGE_F64_12, abi.F64),
                                      (abi.GAUGE_F32_12, abi.F32), (abi.GAUGE_F16_12, abi.F16)],
                         ids=["f64_18", "f64_12", "f32_12", "f16_12"])
def test_dslash_32x64_every_site_every_variant(conf32, fmt, prec):
    lat, ctx, gf64, g = conf32
    assert abs(gf64.plaquette() - O.plaquette(lat, g)) < 1e-12
    gf = gf64 if fmt == abi.GAUGE_F64_18 else gf64.converted(fmt)
    src = np.random.default_rng(32).standard_normal((ctx.veo, 24))
    a, b = ctx.spinor(prec).upload(src), ctx.spinor(prec)
    src_q = a.download()  # what the kernel reads (quantised for FP32 / FP16)
    g_q = half_links(g) if fmt == abi.GAUGE_F16_12 else g
    refs = {eo: O.dslash_eo(lat, src_q, g_q, eo, 0.16) for eo in (abi.EVEN, abi.ODD)}
    exact = {eo: O.dslash_eo(lat, src, g, eo, 0.16) for eo in (abi.EVEN, abi.ODD)}

    def check(tag):
        for eo in (abi.EVEN, abi.ODD):
            ctx.dslash_eo(b.ptr, a.ptr, gf.ptr, fmt, eo, 0.16)
            out = b.download()
            if fmt == abi.GAUGE_F64_18:
                assert site_rel_err(out, refs[eo]) < 1e-13, tag
            elif fmt == abi.GAUGE_F64_12:
                assert site_rel_err(out, refs[eo]) < 1e-12, tag
            elif fmt == abi.GAUGE_F32_12:
                assert site_rel_err(out, exact[eo]) < 2e-5, tag
            else:
                assert site_rel_err(out, refs[eo]) < 1e-3, tag
                assert field_rel(out, exact[eo]) < 3e-3, tag

    try:
        for occ in (0, 1, 2, 4):
            set_shape(ctx, occ)
            check(f"occupancy {occ}")
        set_shape(ctx, -1)
        ms, zp, xcd, rb, blk, occ = (ctypes.c_double(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(),
                                     ctypes.c_int(), ctypes.c_int())
        ctx.autotune_dslash(gf.ptr, fmt, 1, ctypes.byref(ms), ctypes.byref(zp), ctypes.byref(xcd), ctypes.byref(rb),
                            ctypes.byref(blk), ctypes.byref(occ))
        check(f"autotuned zplanes {zp.value} xcd {xcd.value} block {blk.value} occupancy {occ.value}")
    finally:
        set_shape(ctx, -1)
        for f in (a, b):
            f.close()
        if gf is not gf64:
            gf.close()


def test_fp64_cg_32x64_true_residual(conf32):
    """the reference's CG (drop-in) on QplusQminus_eo at 32^3x64 to |r|/|b| <= 1e-10; the true residual is
    recomputed by the oracle's operator from the downloaded solution"""
    from cl2qcd_amd import _host as H
    lat, _, _, g = conf32
    l = H.Lattice(["--nSpace=32", "--nTime=64", "--fermionAction=twistedmass", "--kappa=0.16", "--mu=0.005",
                   "--thetaFermionTemporal=1", "--solverMaxIterations=20000"])
    gf = H.Gaugefield(l, False)
    gf.import_(g)
    bh = np.random.default_rng(64).standard_normal((l.eo_sites, 24))
    b, x = H.Spinorfield_eo(l), H.Spinorfield_eo(l)
    b.import_(bh)
    x.zero()
    b2 = O.norm2(bh)
    it = H.cg(x, H.QplusQminus_eo(l), gf, b, l, 1e-20 * b2, l.wilson_parameters())
    assert 50 < it < 20000
    xs = x.get_host()
    r = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, xs, g, O.TWISTEDMASS, 0.16, l.mubar), bh, 1.0)
    assert np.sqrt(O.norm2(r) / b2) <= 1e-10


def sample_sites_48x96(ns=48, nt=96, zplanes=6, seed=4896):
    """eo indices of the sampled output sites (see the module docstring)"""
    half = ns // 2
    veo = ns ** 3 * nt // 2

    def eo(x0, y, z, t):  # x0 even: the eo index of the pair {x0, x0 + 1}
        return (x0 + ns * (y + ns * (z + ns * np.asarray(t, dtype=np.int64)))) // 2

    zb = sorted({z for c in range(0, ns, zplanes) for z in (c, c + zplanes - 1)})
    idx = []
    xs, ys = np.meshgrid(np.arange(0, ns, 2), np.arange(ns), indexing="ij")
    for t in (0, 1, nt // 2 - 1, nt // 2, nt - 2, nt - 1):
        for z in zb:
            idx.append(eo(xs.ravel(), ys.ravel(), z, t))
    x0 = np.arange(0, ns, 2)
    for t in range(nt):
        for z in zb:
            for y in (0, ns - 1):
                idx.append(eo(x0, y, z, t))
    idx.append(np.arange(64))
    idx.append(np.arange(veo - 64, veo))
    idx.append(np.random.default_rng(seed).integers(0, veo, 50000))
    out = np.unique(np.concatenate(idx))
    assert out.min() >= 0 and out.max() < veo and half * ns * 2 > 0
    return out


def test_dslash_48x96_sampled_sites():
    ctx, gf = heatbath_ctx(48, 96, seed=20261202)
    try:
        lat = O.lattice(48, 96, theta_t=1.0)
        assert ctx.veo == 5308416
        g = gf.download()
        sites = sample_sites_48x96()
        src = np.random.default_rng(96).standard_normal((ctx.veo, 24))
        a, b = ctx.spinor().upload(src), ctx.spinor()
        gf12 = gf.converted(abi.GAUGE_F64_12)
        for eo in (abi.EVEN, abi.ODD):
            ref = O.dslash_eo_sites(lat, src, g, eo, 0.16, sites)
            for fmt, field, tol in ((abi.GAUGE_F64_18, gf, 1e-13), (abi.GAUGE_F64_12, gf12, 1e-12)):
                ctx.dslash_eo(b.ptr, a.ptr, field.ptr, fmt, eo, 0.16)
                out = b.download()[sites]
                assert site_rel_err(out, ref) < tol, (fmt, eo)
    finally:
        ctx.close()


def test_operator_metrics_match_reference_models():
    from cl2qcd_amd import _host as H
    NC, NDIM, D, R, C, CM = 3, 4, 8, 9, 2, 6  # colours, dimensions, FP64 bytes, 18-real links, complex, complex mult
    su3_vec = (CM * NC + (NC - 1) * 2) * NC  # flopUtilities.hpp:42-47
    dslash_f = NDIM * 2 * (su3_vec * 2 + 6 * NC * 2 + 2 * NC * CM)
    dslash_b = (C * 12 * (2 * 4 + 1) + C * 2 * 4 * R) * D
    diag_f, diag_inv_f, diag_b = NC * NDIM * CM, NC * NDIM * CM + NC * NDIM * 2, 48 * D
    g5_f, g5_b = NDIM * NC, 48 * D
    saxpy_f, saxpy_b = NDIM * NC * (CM + 2), C * D * (12 * (2 + 1) + 1)
    g5saxpy_f, g5saxpy_b = NDIM * NC * (1 + CM + 2), C * D * (12 * (2 + 1) + 1)
    dks_f, dks_b = NDIM * (2 * su3_vec + NC * 2) + 3 * NC * 2, (C * NC * (2 * 4 + 1) + C * 2 * 4 * R) * D
    saxpby_f = NC * (2 * CM + 2)
    assert (dslash_f, dslash_b, dks_f, dks_b) == (1632, 2880, 570, 1584)
    for action in ("wilson", "twistedmass"):
        tm = action == "twistedmass"
        for merged in ("false", "true"):
            l = H.Lattice(["--nTime=4", f"--fermionAction={action}", f"--useKernelMergingFermionMatrix={merged}"])
            S = l.eo_sites

            def aee(g5):
                f = 2 * dslash_f + (diag_f + diag_inv_f if tm else 0) + (g5saxpy_f if g5 else saxpy_f)
                b = 2 * dslash_b + (2 * diag_b if tm else 0) + (g5saxpy_b if g5 else saxpy_b)
                return f, b

            q = (aee(False)[0] + g5_f, aee(False)[1] + g5_b)  # Qplus_eo / Qminus_eo: Aee + gamma5_eo
            want = {"Aee": aee(False), "Aee_minus": aee(False), "Aee_AND_gamma5_eo": aee(True),
                    "Aee_minus_AND_gamma5_eo": aee(True), "Qplus_eo": q, "Qminus_eo": q,
                    "QplusQminus_eo": (2 * q[0], 2 * q[1])}
            for name, (f, b) in want.items():
                op = getattr(H, name)(l)
                assert op.get_flops() == S * f, (action, merged, name)
                assert op.get_read_write_size() == S * b, (action, merged, name)
    l = H.Lattice(["--nTime=4", "--fermionAction=rooted_stagg"])
    S = l.eo_sites
    assert H.D_KS_eo(l, 0).get_flops() == S * dks_f
    assert H.D_KS_eo(l, 0).get_read_write_size() == S * dks_b
    assert H.MdagM_eo(l).get_flops() == S * (2 * dks_f + saxpby_f)
    with abi.Context(4, 4) as ctx:
        S = ctx.veo
        lib = ctx.lib
        rw = {"dslash_eo": dslash_b, "dslash_AND_M_tm_inverse_sitediagonal_eo": dslash_b, "gamma5_eo": g5_b,
              "M_tm_sitediagonal": diag_b, "M_tm_inverse_sitediagonal_minus": diag_b,
              "M_tm_sitediagonal_AND_gamma5_eo": diag_b, "saxpy_AND_gamma5_eo": g5saxpy_b, "saxpy_eoprec": saxpy_b,
              "D_KS_eo": dks_b, "saxpby_cplx_staggered_eoprec": C * D * (NC * (2 + 1) + 2)}
        fl = {"dslash_eo": dslash_f, "dslash_AND_M_tm_inverse_sitediagonal_eo": dslash_f + diag_inv_f,
              "gamma5_eo": g5_f, "M_tm_sitediagonal": diag_f, "M_tm_inverse_sitediagonal_minus": diag_inv_f,
              "M_tm_sitediagonal_AND_gamma5_eo": diag_f + g5_f, "saxpy_AND_gamma5_eo": g5saxpy_f,
              "saxpy_eoprec": saxpy_f, "D_KS_eo": dks_f, "saxpby_cplx_staggered_eoprec": saxpby_f}
        for k, v in rw.items():
            assert lib.cl2_reference_read_write_size(ctx.handle, k.encode()) == S * v, k
        for k, v in fl.items():
            assert lib.cl2_reference_flop_size(ctx.handle, k.encode()) == S * v, k
