"""Pins the CPU oracle (oracle/cl2_oracle.cpp) before it is trusted as the checker.

Each function of the oracle is checked against the reference's own golden
numbers (known-answer tests copied from its Boost tests, with the same inputs:
drand48 fills, cold links, the committed ILDG configuration conf.00200) and,
where those leave room, against identities the mathematics fixes (gamma5
hermiticity, gauge covariance, anti-hermiticity of D_KS, M^-1 M = 1, the
solver's own residual).  These run on the CPU only.
"""
import numpy as np
import pytest

import oracle as O
from conftest import CONF_00200
from fields import random_su3_field, rel

KAPPA = 0.125  # meta/parametersFermion.cpp:109 default
MUBAR = 2 * 0.125 * 0.006  # meta::get_mubar, meta/util.cpp:56-58, with default mu


@pytest.fixture(scope="module")
def cold16():
    lat = O.lattice(4, 16)
    return lat, O.gauge_cold(lat)


@pytest.fixture(scope="module")
def conf4():
    lat = O.lattice(4, 4)
    return lat, O.ildg_read(CONF_00200, lat)


def src_eo(lat, seed):
    return O.convert_to_eo(lat, O.pseudo_randomize_spinor(O.volume(lat), seed))


def test_conf00200_plaquette(conf4):
    # xlf-info record of the file: plaquette = 0.571077
    lat, g = conf4
    assert abs(O.plaquette(lat, g) - 0.571077) < 1e-6


def test_drand48_fill_is_posix():
    # POSIX drand48 after srand48(0): first value 0.170828036106 (X0 = 0x1234ABCD330E)
    v = O.fill_drand48(2, 0)
    assert abs(v[0] - 0.17082803610628972) < 1e-15


# physics/fermionmatrix/explicit_test.cpp:338-361
def test_dslash_golden_cold(cold16):
    lat, g = cold16
    sf1, _ = src_eo(lat, 13)
    sf2 = O.dslash_eo(lat, sf1, g, O.EVEN, KAPPA)
    assert rel(O.norm2(sf2), 3311.2698428285048) < 1e-12
    sf1 = O.dslash_eo(lat, sf2, g, O.ODD, KAPPA)
    assert rel(O.norm2(sf1), 3146.1039504225546) < 1e-12


def test_dslash_golden_conf(conf4):
    lat, g = conf4
    sf1, _ = src_eo(lat, 28)
    sf2 = O.dslash_eo(lat, sf1, g, O.EVEN, KAPPA)
    assert rel(O.norm2(sf2), 251.84231257415126) < 1e-12
    sf1 = O.dslash_eo(lat, sf2, g, O.ODD, KAPPA)
    assert rel(O.norm2(sf1), 75.926255640020059) < 1e-12


# explicit_test.cpp:38-196: non-eo M_wilson / M_tm_plus / M_tm_minus, built here from the eo hop
@pytest.mark.parametrize("case,seed,action,minus,refs", [
    ("cold", 1, O.WILSON, False, (2610.3804893063798, 4356.332327032359)),
    ("conf", 2, O.WILSON, False, (2655.7059719467552, 4487.2227568771214)),
    ("cold", 3, O.TWISTEDMASS, False, (2573.5343424130424, 4301.7240841150415)),
    ("conf", 4, O.TWISTEDMASS, False, (2536.3274644928938, 4202.4585412814668)),
    ("cold", 5, O.TWISTEDMASS, True, (2580.3455858426491, 4323.4437436445069)),
    ("conf", 6, O.TWISTEDMASS, True, (2648.2135270529998, 4457.8676622761632)),
])
def test_m_full_golden(case, seed, action, minus, refs, cold16, conf4):
    lat, g = cold16 if case == "cold" else conf4
    sf1 = O.pseudo_randomize_spinor(O.volume(lat), seed)
    sf2 = O.m_full(lat, sf1, g, KAPPA, MUBAR, action, minus)
    assert rel(O.norm2(sf2), refs[0]) < 1e-12
    sf1 = O.m_full(lat, sf2, g, KAPPA, MUBAR, action, minus)
    assert rel(O.norm2(sf1), refs[1]) < 1e-12


# explicit_test.cpp:199-305
@pytest.mark.parametrize("variant,seed,refs", [
    (O.MTM_INVERSE, 5, (4133.0112721801961, 4133.0019729257556)),
    (O.MTM_SITEDIAG, 7, (4142.9167782423428, 4142.9260998050931)),
    (O.MTM_INVERSE_MINUS, 9, (4095.9297157806914, 4095.9204999595668)),
    (O.MTM_SITEDIAG_MINUS, 11, (4099.5308263836096, 4099.5400503279689)),
])
def test_m_tm_golden(variant, seed, refs, cold16):
    lat, _ = cold16
    sf1, _ = src_eo(lat, seed)
    sf2 = O.m_tm(lat, sf1, MUBAR, variant)
    assert rel(O.norm2(sf2), refs[0]) < 1e-12
    sf1 = O.m_tm(lat, sf2, MUBAR, variant)
    assert rel(O.norm2(sf1), refs[1]) < 1e-12


def test_m_tm_inverse_identity(cold16):
    lat, _ = cold16
    x, _ = src_eo(lat, 3)
    for fwd, inv in ((O.MTM_SITEDIAG, O.MTM_INVERSE), (O.MTM_SITEDIAG_MINUS, O.MTM_INVERSE_MINUS)):
        y = O.m_tm(lat, O.m_tm(lat, x, 0.37, fwd), 0.37, inv)
        assert np.abs(y - x).max() < 1e-15
    # gamma5-merged variants equal gamma5 of the plain ones
    for plain, merged in ((O.MTM_SITEDIAG, O.MTM_SITEDIAG_AND_GAMMA5),
                          (O.MTM_SITEDIAG_MINUS, O.MTM_SITEDIAG_MINUS_AND_GAMMA5)):
        a = O.gamma5_eo(lat, O.m_tm(lat, x, 0.37, plain))
        b = O.m_tm(lat, x, 0.37, merged)
        assert np.abs(a - b).max() < 1e-15


def test_dslash_minv_is_m_inverse_after_dslash():
    """The fused dslash_AND_M_tm_inverse_sitediagonal{,_minus}_eo restatement (its own factor table) equals
    M_tm_inverse_sitediagonal{,_minus} applied to dslash_eo, both pinned by the reference's goldens above
    (fermionmatrix_eo_dslash_AND_M_tm_inverse_sitediagonal{,_minus}.cl:30-62)."""
    lat = O.lattice(4, 6, theta_t=1.0, cp_re=0.05)
    g = random_su3_field(4 * O.volume(lat), 17)
    src = np.random.default_rng(18).standard_normal((O.veo(lat), 24))
    for evenodd in (O.EVEN, O.ODD):
        d = O.dslash_eo(lat, src, g, evenodd, 0.16)
        for minus, variant in ((False, O.MTM_INVERSE), (True, O.MTM_INVERSE_MINUS)):
            fused = O.dslash_minv_eo(lat, src, g, evenodd, 0.16, 0.0016, minus)
            assert np.abs(fused - O.m_tm(lat, d, 0.0016, variant)).max() <= 1e-15 * np.abs(d).max()
            # and it is not the non-inverse or the other sign
            other = O.MTM_INVERSE_MINUS if not minus else O.MTM_INVERSE
            assert np.abs(fused - O.m_tm(lat, d, 0.0016, other)).max() > 1e-6 * np.abs(d).max()


def test_dslash_eo_sites_matches_full():
    lat = O.lattice(6, 4, theta_t=1.0, theta_s=0.3)
    g = random_su3_field(4 * O.volume(lat), 19)
    src = np.random.default_rng(20).standard_normal((O.veo(lat), 24))
    sites = np.random.default_rng(21).choice(O.veo(lat), 50, replace=False)
    for evenodd in (O.EVEN, O.ODD):
        full = O.dslash_eo(lat, src, g, evenodd, 0.13)
        assert np.array_equal(O.dslash_eo_sites(lat, src, g, evenodd, 0.13, sites), full[sites])


# physics/fermionmatrix/fermionmatrix_test.cpp:68-133 (Wilson action default)
@pytest.mark.parametrize("op,seed,refs", [
    (O.OP_AEE, 5, (1111.6772283004893, 1290.5580533222192, 1114.183875802898, 1304.0219632505139)),
    (O.OP_AEE_MINUS, 6, (1115.2304524177298, 1291.4471557813199, 1125.22036086961, 1315.8354179702137)),
    (O.OP_QPLUS_EO, 7, (1120.7679957612427, 1303.6316057267845, 1095.6614513886482, 1285.3471211115243)),
    (O.OP_QMINUS_EO, 8, (1100.7016006189097, 1282.8216593368254, 1111.2138438131915, 1309.970927161924)),
    (O.OP_QPLUSQMINUS_EO, 9, (1286.5408969350369, 1931.6496938488942, 1350.5870046327818, 2182.2775949059878)),
])
@pytest.mark.parametrize("merged", [False, True])
def test_fermionmatrix_eo_golden(op, seed, refs, merged, cold16, conf4):
    got = []
    for lat, g in (cold16, conf4):
        sf1, _ = src_eo(lat, seed)
        sf2 = O.apply_eo(lat, op, sf1, g, O.WILSON, KAPPA, merged=merged)
        sf1 = O.apply_eo(lat, op, sf2, g, O.WILSON, KAPPA, merged=merged)
        got += [O.norm2(sf2), O.norm2(sf1)]
    for a, b in zip(got, refs):
        assert rel(a, b) < 1e-12


def test_dslash_gamma5_hermiticity():
    # <phi_o, D_EVEN psi_e> = <gamma5 D_ODD gamma5 phi_o, psi_e> with BC phases and imaginary mu
    lat = O.lattice(4, 6, theta_s=0.3, theta_t=1.0, cp_im=0.21)
    g = random_su3_field(4 * O.volume(lat), 7)
    rng = np.random.default_rng(1)
    phi = rng.standard_normal((O.veo(lat), 24))
    psi = rng.standard_normal((O.veo(lat), 24))
    lhs = O.dot(phi, O.dslash_eo(lat, psi, g, O.EVEN, 0.17))
    rhs = O.dot(O.gamma5_eo(lat, O.dslash_eo(lat, O.gamma5_eo(lat, phi), g, O.ODD, 0.17)), psi)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # a real chemical potential maps mu -> -mu under the adjoint
    lp = O.lattice(4, 6, theta_t=1.0, cp_re=0.13)
    lm = O.lattice(4, 6, theta_t=1.0, cp_re=-0.13)
    lhs = O.dot(phi, O.dslash_eo(lp, psi, g, O.EVEN, 0.17))
    rhs = O.dot(O.gamma5_eo(lm, O.dslash_eo(lm, O.gamma5_eo(lm, phi), g, O.ODD, 0.17)), psi)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_qplus_is_qminus_adjoint():
    # <phi, Qplus psi> = <Qminus phi, psi> (twisted mass and Wilson), so <p, QplusQminus p> = |Qminus p|^2:
    # the identity behind the fused <p, A p> of the reliable-update CG
    lat = O.lattice(4, 6, theta_s=0.3, theta_t=1.0)
    g = random_su3_field(4 * O.volume(lat), 9)
    rng = np.random.default_rng(3)
    phi = rng.standard_normal((O.veo(lat), 24))
    psi = rng.standard_normal((O.veo(lat), 24))
    for action, mubar in ((O.TWISTEDMASS, 2 * 0.16 * 0.02), (O.WILSON, 0.0)):
        lhs = O.dot(phi, O.apply_eo(lat, O.OP_QPLUS_EO, psi, g, action, 0.16, mubar))
        rhs = O.dot(O.apply_eo(lat, O.OP_QMINUS_EO, phi, g, action, 0.16, mubar), psi)
        assert abs(lhs - rhs) < 1e-12 * abs(lhs)
        pAp = O.dot(psi, O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, psi, g, action, 0.16, mubar))
        qm = O.norm2(O.apply_eo(lat, O.OP_QMINUS_EO, psi, g, action, 0.16, mubar))
        assert abs(pAp - qm) < 1e-12 * qm
    # a wrong sign of mu in one factor breaks it
    bad = O.dot(phi, O.apply_eo(lat, O.OP_QPLUS_EO, psi, g, O.TWISTEDMASS, 0.16, 0.0064))
    good = O.dot(O.apply_eo(lat, O.OP_QPLUS_EO, phi, g, O.TWISTEDMASS, 0.16, 0.0064), psi)
    assert abs(bad - good) > 1e-6 * abs(bad)


def _gauge_transform(lat, g, gmat):
    """U_mu(x) -> G(x) U_mu(x) G(x+mu)^dagger on the AoS layout."""
    ns, nt = lat.ns, lat.nt
    V = O.volume(lat)
    U = (g[:, 0::2] + 1j * g[:, 1::2]).reshape(V, 4, 3, 3)
    s = np.arange(V)
    x, y, z, t = s % ns, (s // ns) % ns, (s // ns ** 2) % ns, s // ns ** 3
    up = [((x + y * ns + z * ns ** 2 + ((t + 1) % nt) * ns ** 3)),
          (((x + 1) % ns) + y * ns + z * ns ** 2 + t * ns ** 3),
          (x + ((y + 1) % ns) * ns + z * ns ** 2 + t * ns ** 3),
          (x + y * ns + ((z + 1) % ns) * ns ** 2 + t * ns ** 3)]
    out = np.empty_like(U)
    for mu in range(4):
        out[:, mu] = gmat @ U[:, mu] @ np.conj(np.transpose(gmat[up[mu]], (0, 2, 1)))
    o = out.reshape(4 * V, 9)
    r = np.empty((4 * V, 18))
    r[:, 0::2], r[:, 1::2] = o.real, o.imag
    return r


def _apply_site_matrix(lat, field, gmat, parity, ncomp):
    """psi(x) -> G(x) psi(x) on every colour vector of an eo field of the given parity."""
    out = field.copy()
    for i, site in enumerate(O.eo_sites(lat, parity)):
        G = gmat[site]
        v = (field[i, 0::2] + 1j * field[i, 1::2]).reshape(ncomp // 3, 3)
        w = (v @ G.T).reshape(-1)
        out[i, 0::2], out[i, 1::2] = w.real, w.imag
    return out


def test_dslash_and_dks_gauge_covariance():
    # D[U^G](G psi) = G D[U] psi pins which link each hop uses and where the dagger goes
    lat = O.lattice(2, 4, theta_t=1.0)
    V = O.volume(lat)
    g = random_su3_field(4 * V, 11)
    gm = random_su3_field(V, 12)
    G = (gm[:, 0::2] + 1j * gm[:, 1::2]).reshape(V, 3, 3)
    gt = _gauge_transform(lat, g, G)
    rng = np.random.default_rng(2)
    for evenodd, in_par, out_par in ((O.EVEN, 0, 1), (O.ODD, 1, 0)):
        psi = rng.standard_normal((O.veo(lat), 24))
        lhs = O.dslash_eo(lat, _apply_site_matrix(lat, psi, G, in_par, 12), gt, evenodd, 0.2)
        rhs = _apply_site_matrix(lat, O.dslash_eo(lat, psi, g, evenodd, 0.2), G, out_par, 12)
        assert np.abs(lhs - rhs).max() < 1e-13
    for evenodd in (O.EVEN, O.ODD):
        chi = rng.standard_normal((O.veo(lat), 6))
        out_par, in_par = evenodd, 1 - evenodd
        lhs = O.dks_eo(lat, _apply_site_matrix(lat, chi, G, in_par, 3), gt, evenodd)
        rhs = _apply_site_matrix(lat, O.dks_eo(lat, chi, g, evenodd), G, out_par, 3)
        assert np.abs(lhs - rhs).max() < 1e-13


# physics/fermionmatrix/explicit_stagg_test.cpp:40-116, fermionmatrix_stagg_test.cpp:54-66
def test_dks_golden():
    lat8 = O.lattice(8, 8)
    g8 = O.gauge_cold(lat8)
    a = O.pseudo_randomize_su3vec(O.veo(lat8), 13)
    b = O.pseudo_randomize_su3vec(O.veo(lat8), 31)
    assert rel(O.norm2(O.dks_eo(lat8, a, g8, O.EVEN)), 2030.1639500272767691) < 1e-12
    assert rel(O.norm2(O.dks_eo(lat8, b, g8, O.ODD)), 2076.7437224316167885) < 1e-12
    assert rel(O.norm2(O.mdagm_eo(lat8, a, g8, 1.0, O.ODD)), 12670.438549218517437) < 1e-12
    assert rel(O.norm2(O.mdagm_eo(lat8, b, g8, 1.0, O.EVEN)), 12925.274439556447760) < 1e-12
    lat4 = O.lattice(4, 4)
    g4 = O.ildg_read(CONF_00200, lat4)
    a = O.pseudo_randomize_su3vec(O.veo(lat4), 123)
    b = O.pseudo_randomize_su3vec(O.veo(lat4), 321)
    assert rel(O.norm2(O.dks_eo(lat4, a, g4, O.EVEN)), 547.69039343718509372) < 1e-12
    assert rel(O.norm2(O.dks_eo(lat4, b, g4, O.ODD)), 536.10645183266251479) < 1e-12
    assert rel(O.norm2(O.mdagm_eo(lat4, a, g4, 1.0, O.ODD)), 2980.0540282922056576) < 1e-12
    assert rel(O.norm2(O.mdagm_eo(lat4, b, g4, 1.0, O.EVEN)), 2883.0056841550949684) < 1e-12


def test_dks_antihermitian():
    lat = O.lattice(4, 4, theta_s=0.25, theta_t=1.0, cp_im=0.3)
    g = random_su3_field(4 * O.volume(lat), 5)
    rng = np.random.default_rng(3)
    phi = rng.standard_normal((O.veo(lat), 6))
    psi = rng.standard_normal((O.veo(lat), 6))
    lhs = O.dot(phi, O.dks_eo(lat, psi, g, O.EVEN))
    rhs = -O.dot(O.dks_eo(lat, phi, g, O.ODD), psi)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_blas_closed_forms():
    rng = np.random.default_rng(9)
    x, y, z = (rng.standard_normal((50, 24)) for _ in range(3))
    cx = lambda a: a[:, 0::2] + 1j * a[:, 1::2]
    a, b = 0.3 - 0.7j, -1.1 + 0.2j
    assert np.allclose(cx(O.saxpy(x, y, a)), cx(y) - a * cx(x), rtol=0, atol=1e-14)
    assert np.allclose(cx(O.sax(x, a)), a * cx(x), rtol=0, atol=1e-14)
    assert np.allclose(cx(O.saxsbypz(x, y, z, a, b)), a * cx(x) + b * cx(y) + cx(z), rtol=0, atol=1e-14)
    assert abs(O.dot(x, y) - np.vdot(cx(x), cx(y))) < 1e-12
    assert abs(O.norm2(x) - np.vdot(cx(x), cx(x)).real) < 1e-12
    xs, ys = x[:, :6].copy(), y[:, :6].copy()
    assert np.allclose(cx(O.stagg_saxpby(xs, ys, a, b)), a * cx(xs) + b * cx(ys), rtol=0, atol=1e-14)


def test_cg_eo_true_residual(conf4):
    lat, g = conf4
    b, _ = src_eo(lat, 21)
    x0 = np.zeros_like(b)
    prec = 1e-20 * O.norm2(b)
    for action, mubar in ((O.WILSON, 0.0), (O.TWISTEDMASS, 0.05)):
        rc, it, x = O.cg_eo(lat, O.OP_QPLUSQMINUS_EO, g, b, x0, action, 0.15, mubar, prec=prec)
        assert rc == O.OK and it % 10 == 0
        r = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, x, g, action, 0.15, mubar), b, 1.0)
        assert O.norm2(r) / O.norm2(b) < 1e-18


@pytest.mark.parametrize("save", [False, True])
def test_bicgstab_eo_true_residual(conf4, save):
    lat, g = conf4
    b, _ = src_eo(lat, 22)
    x0 = np.zeros_like(b)
    prec = 1e-20 * O.norm2(b)
    rc, it, x = O.bicgstab_eo(lat, O.OP_AEE, g, b, x0, O.TWISTEDMASS, 0.15, 0.05, prec=prec, save=save)
    assert rc == O.OK
    r = O.saxpy(O.apply_eo(lat, O.OP_AEE, x, g, O.TWISTEDMASS, 0.15, 0.05), b, 1.0)
    assert O.norm2(r) / O.norm2(b) < 1e-18


def test_solver_failure_modes(conf4):
    lat, g = conf4
    b, _ = src_eo(lat, 23)
    x0 = np.zeros_like(b)
    rc, it, _ = O.cg_eo(lat, O.OP_QPLUSQMINUS_EO, g, b, x0, O.WILSON, 0.15, prec=1e-60, cgmax=20)
    assert rc == O.DID_NOT_SOLVE and it == 20
    bn = b.copy()
    bn[0, 0] = np.nan
    rc, it, _ = O.cg_eo(lat, O.OP_QPLUSQMINUS_EO, g, bn, x0, O.WILSON, 0.15, prec=1e-20)
    assert rc == O.STUCK and it == 0


# physics/algorithms/solver_shifted_test.cpp:42-116 and 258-312
def test_cgm_golden(conf4):
    lat, g = conf4
    sigma = [0.0002065381736724, 0.00302707751065980, 0.0200732678058145, 0.12517586269872370, 1.0029328743375700]
    b = O.pseudo_randomize_su3vec(O.veo(lat), 13)
    rc, it, xs = O.cgm(lat, g, b, sigma, 0.567, prec=1e-23)
    assert rc == O.OK
    for s, x in zip(sigma, xs):
        ax = O.stagg_saxpby(x, O.mdagm_eo(lat, x, g, 0.567), s, 1.0)
        assert rel(O.norm2(ax), O.norm2(b)) < 1e-10
    bc = np.zeros((O.veo(lat), 6))
    bc[:, 0::2] = 1.0 / np.sqrt(3 * O.volume(lat))
    rc, it, xs = O.cgm(lat, g, bc, [0.0], 0.01, prec=1e-24)
    assert rc == O.OK and rel(O.norm2(xs[0]), 9.0597433493689383255) < 1e-10
    b = O.pseudo_randomize_su3vec(O.veo(lat), 123)
    rc, it, xs = O.cgm(lat, g, b, [0.0], 0.01, prec=1e-24)
    assert rc == O.OK and rel(O.norm2(xs[0]), 3790.3193634090343949) < 1e-10
