"""GPU parity of the Wilson / twisted-mass kernels against the oracle, through the C ABI.

FP64 bar: <= 1e-13 relative per site (|out(x) - ref(x)|_2 / |ref(x)|_2 over the
24 reals).  12-real links reconstruct the third row, which adds rounding of the
links themselves: <= 1e-12.  FP32 (12-real): <= 2e-5 per site.
"""
import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import abi
from conftest import CONF_00200
from fields import random_su3_field, site_rel_err, rel

pytestmark = pytest.mark.gpu

KAPPA = 0.125
MUBAR = 2 * 0.125 * 0.006
TOL64 = 1e-13


def make(ns, nt, **kw):
    lat = O.lattice(ns, nt, **kw)
    ctx = abi.Context(ns, nt, **kw)
    return lat, ctx


def src_eo(lat, seed):
    return O.convert_to_eo(lat, O.pseudo_randomize_spinor(O.volume(lat), seed))


def test_dslash_golden_conf00200():
    lat, ctx = make(4, 4)
    with ctx:
        g = O.ildg_read(CONF_00200, lat)
        gf = ctx.gauge(g)
        sf1, _ = src_eo(lat, 28)
        a, b = ctx.spinor().upload(sf1), ctx.spinor()
        ctx.dslash_eo(b.ptr, a.ptr, gf.ptr, abi.GAUGE_F64_18, abi.EVEN, KAPPA)
        out = b.download()
        ref = O.dslash_eo(lat, sf1, g, O.EVEN, KAPPA)
        assert site_rel_err(out, ref) < TOL64
        assert rel(O.norm2(out), 251.84231257415126) < 1e-12  # explicit_test.cpp:358
        ctx.dslash_eo(a.ptr, b.ptr, gf.ptr, abi.GAUGE_F64_18, abi.ODD, KAPPA)
        out2 = a.download()
        assert site_rel_err(out2, O.dslash_eo(lat, ref, g, O.ODD, KAPPA)) < TOL64
        assert rel(O.norm2(out2), 75.926255640020059) < 1e-12  # explicit_test.cpp:360


def test_dslash_golden_cold():
    lat, ctx = make(4, 16)
    with ctx:
        g = O.gauge_cold(lat)
        gf = ctx.gauge(g)
        sf1, _ = src_eo(lat, 13)
        a, b = ctx.spinor().upload(sf1), ctx.spinor()
        ctx.dslash_eo(b.ptr, a.ptr, gf.ptr, abi.GAUGE_F64_18, abi.EVEN, KAPPA)
        assert rel(O.norm2(b.download()), 3311.2698428285048) < 1e-12  # explicit_test.cpp:333
        ctx.dslash_eo(a.ptr, b.ptr, gf.ptr, abi.GAUGE_F64_18, abi.ODD, KAPPA)
        assert rel(O.norm2(a.download()), 3146.1039504225546) < 1e-12  # explicit_test.cpp:335


CASES = [
    dict(ns=8, nt=8, theta_s=0.0, theta_t=1.0),
    dict(ns=4, nt=6, theta_s=0.3, theta_t=1.0, cp_re=0.11),
    dict(ns=6, nt=4, theta_s=0.0, theta_t=0.5, cp_im=0.27),
    dict(ns=2, nt=4, theta_s=1.0, theta_t=1.0, cp_re=0.05, cp_im=-0.3),
    dict(ns=16, nt=32, theta_s=0.0, theta_t=1.0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, (c["ns"], c["nt"]))))
@pytest.mark.parametrize("evenodd", [abi.EVEN, abi.ODD])
def test_dslash_random_su3(case, evenodd):
    lat, ctx = make(**case)
    with ctx:
        g = random_su3_field(4 * O.volume(lat), 100 + case["ns"])
        src = np.random.default_rng(case["nt"]).standard_normal((O.veo(lat), 24))
        ref = O.dslash_eo(lat, src, g, evenodd, 0.16)
        a, b = ctx.spinor().upload(src), ctx.spinor()
        for fmt, tol in ((abi.GAUGE_F64_18, TOL64), (abi.GAUGE_F64_12, 1e-12)):
            gf = ctx.gauge(g, fmt)
            ctx.dslash_eo(b.ptr, a.ptr, gf.ptr, fmt, evenodd, 0.16)
            assert site_rel_err(b.download(), ref) < tol, f"format {fmt}"
            gf.close()
        # FP32 spinors and links
        a32, b32 = ctx.spinor(abi.F32).upload(src), ctx.spinor(abi.F32)
        gf = ctx.gauge(g, abi.GAUGE_F32_12)
        ctx.dslash_eo(b32.ptr, a32.ptr, gf.ptr, abi.GAUGE_F32_12, evenodd, 0.16)
        assert site_rel_err(b32.download(), ref) < 2e-5


@pytest.mark.parametrize("minus", [0, 1])
def test_dslash_and_minv(minus):
    lat, ctx = make(8, 8, theta_t=1.0)
    with ctx:
        g = random_su3_field(4 * O.volume(lat), 3)
        src = np.random.default_rng(4).standard_normal((O.veo(lat), 24))
        gf = ctx.gauge(g)
        a, b = ctx.spinor().upload(src), ctx.spinor()
        for evenodd in (abi.EVEN, abi.ODD):
            ctx.dslash_AND_M_tm_inverse_sitediagonal_eo(b.ptr, a.ptr, gf.ptr, abi.GAUGE_F64_18, evenodd, 0.16,
                                                        0.0016, minus)
            ref = O.dslash_minv_eo(lat, src, g, evenodd, 0.16, 0.0016, bool(minus))
            assert site_rel_err(b.download(), ref) < TOL64


@pytest.mark.parametrize("variant", range(6))
@pytest.mark.parametrize("prec", [abi.F64, abi.F32])
def test_m_tm_sitediagonal(variant, prec):
    lat, ctx = make(4, 8)
    with ctx:
        src = np.random.default_rng(variant).standard_normal((O.veo(lat), 24))
        a, b = ctx.spinor(prec).upload(src), ctx.spinor(prec)
        ctx.M_tm_sitediagonal_eo(b.ptr, a.ptr, prec, 0.37, variant)
        ref = O.m_tm(lat, src, 0.37, variant)
        assert site_rel_err(b.download(), ref) < (1e-15 if prec == abi.F64 else 1e-6)


def test_gamma5_and_saxpy_gamma5():
    lat, ctx = make(4, 4)
    with ctx:
        rng = np.random.default_rng(7)
        x, y = rng.standard_normal((O.veo(lat), 24)), rng.standard_normal((O.veo(lat), 24))
        fx, fy, fo = ctx.spinor().upload(x), ctx.spinor().upload(y), ctx.spinor()
        ctx.gamma5_eo(fx.ptr, abi.F64)
        assert np.array_equal(fx.download(), O.gamma5_eo(lat, x))
        fx.upload(x)
        al = abi.scalar(0.3 - 1.2j)
        ctx.saxpy_AND_gamma5_eo(fo.ptr, fx.ptr, fy.ptr, abi.F64, al)
        assert site_rel_err(fo.download(), O.saxpy_and_gamma5_eo(lat, x, y, 0.3 - 1.2j)) < 1e-15


def test_dslash_aee_epilogue():
    lat, ctx = make(6, 4, theta_t=1.0)
    with ctx:
        g = random_su3_field(4 * O.volume(lat), 9)
        rng = np.random.default_rng(10)
        src, x = rng.standard_normal((O.veo(lat), 24)), rng.standard_normal((O.veo(lat), 24))
        gf = ctx.gauge(g)
        fs, fx, fo = ctx.spinor().upload(src), ctx.spinor().upload(x), ctx.spinor()
        f1, f2 = 1 + 0.2j, 1 - 0.2j
        for g5 in (0, 1):
            ctx.dslash_AND_sitediagonal_saxpy_eo(fo.ptr, fs.ptr, fx.ptr, gf.ptr, abi.GAUGE_F64_18, abi.EVEN, 0.15,
                                                 f1.real, f1.imag, f2.real, f2.imag, g5)
            ref = O.saxpy(O.dslash_eo(lat, src, g, O.EVEN, 0.15), O.m_tm(lat, x, 0.2, O.MTM_SITEDIAG), 1.0)
            if g5:
                ref = O.gamma5_eo(lat, ref)
            assert site_rel_err(fo.download(), ref) < TOL64


FM_REFS = {  # physics/fermionmatrix/fermionmatrix_test.cpp:68-90 (Wilson action)
    abi.OP_AEE: (5, (1111.6772283004893, 1290.5580533222192, 1114.183875802898, 1304.0219632505139)),
    abi.OP_AEE_MINUS: (6, (1115.2304524177298, 1291.4471557813199, 1125.22036086961, 1315.8354179702137)),
    abi.OP_QPLUS_EO: (7, (1120.7679957612427, 1303.6316057267845, 1095.6614513886482, 1285.3471211115243)),
    abi.OP_QMINUS_EO: (8, (1100.7016006189097, 1282.8216593368254, 1111.2138438131915, 1309.970927161924)),
    abi.OP_QPLUSQMINUS_EO: (9, (1286.5408969350369, 1931.6496938488942, 1350.5870046327818, 2182.2775949059878)),
}


@pytest.mark.parametrize("op", sorted(FM_REFS))
def test_fermionmatrix_eo_golden(op):
    seed, refs = FM_REFS[op]
    got = []
    for ns, nt, conf in ((4, 16, False), (4, 4, True)):
        lat, ctx = make(ns, nt)
        with ctx:
            g = O.ildg_read(CONF_00200, lat) if conf else O.gauge_cold(lat)
            gf = ctx.gauge(g)
            sf1, _ = src_eo(lat, seed)
            a, b = ctx.spinor().upload(sf1), ctx.spinor()
            tmp = ctx.buffer(2 * ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F64))
            ctx.fermionmatrix_eo(op, b.ptr, a.ptr, gf.ptr, abi.GAUGE_F64_18, abi.ACTION_WILSON, KAPPA, 0.0, tmp.ptr)
            r1 = O.apply_eo(lat, op, sf1, g, O.WILSON, KAPPA)
            o1 = b.download()
            assert site_rel_err(o1, r1) < TOL64
            ctx.fermionmatrix_eo(op, a.ptr, b.ptr, gf.ptr, abi.GAUGE_F64_18, abi.ACTION_WILSON, KAPPA, 0.0, tmp.ptr)
            o2 = a.download()
            assert site_rel_err(o2, O.apply_eo(lat, op, r1, g, O.WILSON, KAPPA)) < TOL64
            got += [O.norm2(o1), O.norm2(o2)]
    for a, b in zip(got, refs):
        assert rel(a, b) < 1e-12


@pytest.mark.parametrize("op", [abi.OP_AEE, abi.OP_AEE_AND_GAMMA5, abi.OP_AEE_MINUS, abi.OP_AEE_MINUS_AND_GAMMA5,
                                abi.OP_QPLUS_EO, abi.OP_QMINUS_EO, abi.OP_QPLUSQMINUS_EO])
def test_fermionmatrix_eo_twisted_mass(op):
    lat, ctx = make(8, 8, theta_t=1.0)
    with ctx:
        g = random_su3_field(4 * O.volume(lat), 21)
        src = np.random.default_rng(22).standard_normal((O.veo(lat), 24))
        gf = ctx.gauge(g)
        a, b = ctx.spinor().upload(src), ctx.spinor()
        tmp = ctx.buffer(2 * ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F64))
        ctx.fermionmatrix_eo(op, b.ptr, a.ptr, gf.ptr, abi.GAUGE_F64_18, abi.ACTION_TWISTEDMASS, 0.16, 0.0016,
                             tmp.ptr)
        ref = O.apply_eo(lat, op, src, g, O.TWISTEDMASS, 0.16, 0.0016)
        assert site_rel_err(b.download(), ref) < TOL64


def test_rejects_bad_arguments():
    lat, ctx = make(4, 4)
    with ctx:
        g = O.gauge_cold(lat)
        gf = ctx.gauge(g)
        a = ctx.spinor()
        with pytest.raises(abi.Cl2Error) as e:
            ctx.dslash_eo(a.ptr, a.ptr, gf.ptr, abi.GAUGE_F64_18, abi.EVEN, 0.1)
        assert e.value.code == abi.E_ARG
        with pytest.raises(abi.Cl2Error):
            ctx.dslash_eo(a.ptr, gf.ptr, gf.ptr, abi.GAUGE_F64_18, 7, 0.1)
        with pytest.raises(abi.Cl2Error):
            ctx.dslash_eo(a.ptr, gf.ptr, gf.ptr, 9, abi.EVEN, 0.1)
    with pytest.raises(abi.Cl2Error):
        abi.Context(5, 4)


@pytest.mark.parametrize("fmt,prec,nrhs", [(abi.GAUGE_F64_18, abi.F64, 1), (abi.GAUGE_F32_12, abi.F32, 1),
                                           (abi.GAUGE_F16_12, abi.F16, 1), (abi.GAUGE_F32_12, abi.F32, 4)])
def test_autotune_keeps_results(fmt, prec, nrhs):
    """The autotuned launch shape changes which thread runs which site and how many hops a thread has in
    flight, not the arithmetic, so results stay bit identical."""
    import ctypes
    lat, ctx = make(8, 8, theta_t=1.0)
    with ctx:
        g = random_su3_field(4 * O.volume(lat), 77)
        gf = ctx.gauge(g, fmt)
        fb = ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, prec)
        src = np.random.default_rng(78).standard_normal((ctx.veo, 24))
        single = ctx.spinor(prec).upload(src)
        a, b = ctx.buffer(nrhs * fb), ctx.buffer(nrhs * fb)
        for r in range(nrhs):
            if nrhs == 1:
                ctx.copy(a.ptr, single.ptr, fb)
            else:
                ctx.block_insert_eo(a.ptr, single.ptr, prec, nrhs, r)

        outf = ctx.spinor(prec)

        def run():
            if nrhs == 1:
                ctx.dslash_eo(b.ptr, a.ptr, gf.ptr, fmt, abi.EVEN, 0.16)
            else:
                ctx.dslash_eo_mrhs(b.ptr, a.ptr, gf.ptr, fmt, abi.EVEN, 0.16, nrhs)
            res = []
            for r in range(nrhs):
                if nrhs == 1:
                    ctx.copy(outf.ptr, b.ptr, fb)
                else:
                    ctx.block_extract_eo(outf.ptr, b.ptr, prec, nrhs, r)
                res.append(outf.download())
            return np.stack(res)

        before = run()
        ms, zp, xcd, rb, blk, occ = (ctypes.c_double(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(),
                                     ctypes.c_int())
        ctx.autotune_dslash(gf.ptr, fmt, nrhs, ctypes.byref(ms), ctypes.byref(zp), ctypes.byref(xcd),
                            ctypes.byref(rb), ctypes.byref(blk), ctypes.byref(occ))
        assert ms.value > 0 and zp.value >= -1 and xcd.value in (0, 1) and rb.value in (-2, -1, 1, 2, 4)
        assert blk.value in (128, 256, 512) and occ.value in (0, 1, 2, 4)
        assert np.array_equal(run(), before)


@pytest.mark.parametrize("fmt,prec", [(abi.GAUGE_F64_18, abi.F64), (abi.GAUGE_F64_12, abi.F64),
                                      (abi.GAUGE_F32_12, abi.F32), (abi.GAUGE_F16_12, abi.F16)])
def test_every_launch_variant_bit_identical(fmt, prec):
    """Every block size x register budget x work order of the single-field hopping kernel gives the same
    bits (the arithmetic helpers fix their rounding), with the Aee epilogue as well as the plain hop; and the
    FP64 result agrees with the oracle."""
    lat, ctx = make(8, 8, theta_t=1.0)
    with ctx:
        g = random_su3_field(4 * O.volume(lat), 91)
        gf = ctx.gauge(g, fmt)
        src = np.random.default_rng(92).standard_normal((ctx.veo, 24))
        a, b, c = ctx.spinor(prec).upload(src), ctx.spinor(prec), ctx.spinor(prec)
        tmp = ctx.buffer(2 * ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F64))
        results = {}
        for block in (128, 256, 512):
            for occ in (0, 1, 2, 4):
                for zp, xcd in ((-1, 1), (0, 0), (2, 1)):
                    for key, v in (("hop.block", block), ("hop.occupancy", occ), ("hop.zplanes", zp), ("hop.xcd", xcd)):
                        ctx.check(ctx.lib.cl2_tune_set(ctx.handle, key.encode(), v))
                    ctx.dslash_eo(b.ptr, a.ptr, gf.ptr, fmt, abi.ODD, 0.16)
                    ctx.fermionmatrix_eo(abi.OP_AEE, c.ptr, a.ptr, gf.ptr, fmt, abi.ACTION_TWISTEDMASS, 0.16, 0.0016,
                                         tmp.ptr)
                    results[(block, occ, zp, xcd)] = (b.download(), c.download())
        ref = results[(256, 0, -1, 1)]
        for k, v in results.items():
            assert np.array_equal(v[0], ref[0]) and np.array_equal(v[1], ref[1]), k
        if prec == abi.F64:
            assert site_rel_err(ref[0], O.dslash_eo(lat, src, g, O.ODD, 0.16)) < 1e-13


def test_padded_link_tiles():
    """6^3x10 has 1080 eo sites, not a multiple of the 64-site link tile: the padded layout must round-trip
    through import/export in every format, feed the hopping kernels, and agree between the heatbath's own
    link addressing and the export (plaquette of the exported field)."""
    lat, ctx = make(6, 10, theta_t=1.0)
    with ctx:
        assert ctx.veo % 64 != 0
        g = random_su3_field(4 * O.volume(lat), 93)
        src = np.random.default_rng(94).standard_normal((ctx.veo, 24))
        for fmt, prec, tol in ((abi.GAUGE_F64_18, abi.F64, 1e-13), (abi.GAUGE_F64_12, abi.F64, 1e-13),
                               (abi.GAUGE_F32_12, abi.F32, 1e-5)):
            gf = ctx.gauge(g, fmt)
            back = gf.download()
            assert np.abs(back - g).max() < (1e-14 if prec == abi.F64 else 1e-6)
            a, b = ctx.spinor(prec).upload(src), ctx.spinor(prec)
            for evenodd in (abi.EVEN, abi.ODD):
                ctx.dslash_eo(b.ptr, a.ptr, gf.ptr, fmt, evenodd, 0.16)
                assert site_rel_err(b.download(), O.dslash_eo(lat, src, g, evenodd, 0.16)) < tol
            if fmt == abi.GAUGE_F32_12:  # 12 right-hand sides: slot r equals the single-field result bit for bit
                nrhs, fb = 12, ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, prec)
                blk, out = ctx.buffer(nrhs * fb), ctx.buffer(nrhs * fb)
                srcs = [np.random.default_rng(100 + r).standard_normal((ctx.veo, 24)) for r in range(nrhs)]
                for r in range(nrhs):
                    a.upload(srcs[r])
                    ctx.block_insert_eo(blk.ptr, a.ptr, prec, nrhs, r)
                ctx.dslash_eo_mrhs(out.ptr, blk.ptr, gf.ptr, fmt, abi.ODD, 0.16, nrhs)
                for r in (0, 7, 11):
                    a.upload(srcs[r])
                    ctx.dslash_eo(b.ptr, a.ptr, gf.ptr, fmt, abi.ODD, 0.16)
                    ctx.block_extract_eo(a.ptr, out.ptr, prec, nrhs, r)
                    assert np.array_equal(a.download(), b.download())
        gf = ctx.gauge(g)
        assert abs(gf.plaquette() - O.plaquette(lat, g)) < 1e-12
        gf.heatbath(5.7, 2, overrelax=1, seed=3)
        assert abs(gf.plaquette() - O.plaquette(lat, gf.download())) < 1e-12


@pytest.mark.parametrize("fmt,prec", [(abi.GAUGE_F64_18, abi.F64), (abi.GAUGE_F32_12, abi.F32)])
def test_staggered_launch_variants_bit_identical(fmt, prec):
    """D_KS_eo gives the same bits for every block size x register budget x work order, the autotuner keeps
    them, and the FP64 result agrees with the oracle."""
    import ctypes
    lat, ctx = make(8, 8, theta_t=1.0)
    with ctx:
        g = random_su3_field(4 * O.volume(lat), 95)
        gf = ctx.gauge(g, fmt)
        src = np.random.default_rng(96).standard_normal((ctx.veo, 6))
        a, b = ctx.su3vec(prec).upload(src), ctx.su3vec(prec)
        results = {}
        for block in (128, 256, 512):
            for occ in (0, 1, 2, 4):
                for zp, xcd in ((-1, 1), (0, 0), (2, 1)):
                    for key, v in (("hop.block", block), ("hop.occupancy", occ), ("hop.zplanes", zp), ("hop.xcd", xcd)):
                        ctx.check(ctx.lib.cl2_tune_set(ctx.handle, key.encode(), v))
                    ctx.D_KS_eo(b.ptr, a.ptr, gf.ptr, fmt, abi.ODD)
                    results[(block, occ, zp, xcd)] = b.download()
        ref = results[(256, 0, -1, 1)]
        for k, v in results.items():
            assert np.array_equal(v, ref), k
        ctx.check(ctx.lib.cl2_tune_set(ctx.handle, b"hop.occupancy", -1))
        ms, zp, xcd, blk, occ = ctypes.c_double(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ctx.autotune_dks(gf.ptr, fmt, ctypes.byref(ms), ctypes.byref(zp), ctypes.byref(xcd), ctypes.byref(blk),
                         ctypes.byref(occ))
        assert ms.value > 0 and blk.value in (128, 256, 512) and occ.value in (0, 1, 4)
        ctx.D_KS_eo(b.ptr, a.ptr, gf.ptr, fmt, abi.ODD)
        assert np.array_equal(b.download(), ref)
        if prec == abi.F64:
            assert site_rel_err(ref, O.dks_eo(lat, src, g, O.ODD)) < 1e-13
