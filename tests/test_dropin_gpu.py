"""The C++ drop-in (physics::lattices / fermionmatrix / algorithms::solvers) on the GPU.

These follow the reference's own Boost tests step by step -- same parameters,
same drand48 fills, same ILDG configuration, same golden numbers -- and add
parity against the oracle where the reference tests only check a norm:
physics/fermionmatrix/explicit_test.cpp, fermionmatrix_test.cpp,
explicit_stagg_test.cpp, fermionmatrix_stagg_test.cpp,
physics/algorithms/solver_shifted_test.cpp; cg and bicgstab (which the
reference has no unit test for) against the oracle's restatement.
"""
import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import _host as H
from conftest import CONF_00200
from fields import rel, site_rel_err

pytestmark = pytest.mark.gpu


def lattice(*args):
    return H.Lattice(list(args))


def conf(l):
    return H.Gaugefield(l, CONF_00200)


def cold(l):
    return H.Gaugefield(l, False)


def eo_pair(l, seed):
    src = H.Spinorfield(l)
    H.pseudo_randomize(src, seed)
    sf1, sf2 = H.Spinorfield_eo(l), H.Spinorfield_eo(l)
    H.convert_to_eoprec(sf1, sf2, src)
    return sf1, sf2


# explicit_test.cpp:308-362
def test_dslash():
    l = lattice("--nTime=16")
    gf = cold(l)
    sf1, sf2 = eo_pair(l, 13)
    H.dslash(sf2, gf, sf1, H.EVEN, l.kappa)
    assert rel(H.squarenorm(sf2), 3311.2698428285048) < 1e-12
    H.dslash(sf1, gf, sf2, H.ODD, l.kappa)
    assert rel(H.squarenorm(sf1), 3146.1039504225546) < 1e-12
    l = lattice("--nTime=4")
    gf = conf(l)
    sf1, sf2 = eo_pair(l, 28)
    H.dslash(sf2, gf, sf1, H.EVEN, l.kappa)
    assert rel(H.squarenorm(sf2), 251.84231257415126) < 1e-12
    H.dslash(sf1, gf, sf2, H.ODD, l.kappa)
    assert rel(H.squarenorm(sf1), 75.926255640020059) < 1e-12


# explicit_test.cpp:199-305
@pytest.mark.parametrize("fn,seed,refs", [
    ("M_tm_inverse_sitediagonal", 5, (4133.0112721801961, 4133.0019729257556)),
    ("M_tm_sitediagonal", 7, (4142.9167782423428, 4142.9260998050931)),
    ("M_tm_inverse_sitediagonal_minus", 9, (4095.9297157806914, 4095.9204999595668)),
    ("M_tm_sitediagonal_minus", 11, (4099.5308263836096, 4099.5400503279689)),
])
def test_m_tm(fn, seed, refs):
    l = lattice("--nTime=16", "--fermionAction=twistedmass")
    sf1, sf2 = eo_pair(l, seed)
    getattr(H, fn)(sf2, sf1, l.mubar)
    assert rel(H.squarenorm(sf2), refs[0]) < 1e-12
    getattr(H, fn)(sf1, sf2, l.mubar)
    assert rel(H.squarenorm(sf1), refs[1]) < 1e-12


# fermionmatrix_test.cpp:68-133
@pytest.mark.parametrize("name,seed,refs", [
    ("Aee", 5, (1111.6772283004893, 1290.5580533222192, 1114.183875802898, 1304.0219632505139)),
    ("Aee_minus", 6, (1115.2304524177298, 1291.4471557813199, 1125.22036086961, 1315.8354179702137)),
    ("Qplus_eo", 7, (1120.7679957612427, 1303.6316057267845, 1095.6614513886482, 1285.3471211115243)),
    ("Qminus_eo", 8, (1100.7016006189097, 1282.8216593368254, 1111.2138438131915, 1309.970927161924)),
    ("QplusQminus_eo", 9, (1286.5408969350369, 1931.6496938488942, 1350.5870046327818, 2182.2775949059878)),
])
def test_fermionmatrix(name, seed, refs):
    got = []
    for args, gauge in ((["--nTime=16"], cold), (["--nTime=4"], conf)):
        l = lattice(*args)
        matrix = getattr(H, name)(l)
        gf = gauge(l)
        sf1, sf2 = eo_pair(l, seed)
        matrix(sf2, gf, sf1, l.wilson_parameters())
        got.append(H.squarenorm(sf2))
        matrix(sf1, gf, sf2, l.wilson_parameters())
        got.append(H.squarenorm(sf1))
    for a, b in zip(got, refs):
        assert rel(a, b) < 1e-12
    assert getattr(H, name)(lattice()).isHermitian() == (name == "QplusQminus_eo")


# explicit_stagg_test.cpp:40-116 and fermionmatrix_stagg_test.cpp:54-66
def test_staggered_operators():
    l = lattice("--nSpace=8", "--fermionAction=rooted_stagg", "--mass=1.", "--nDevices=1")
    gf = cold(l)
    sf1, sf2, out = H.Staggeredfield_eo(l), H.Staggeredfield_eo(l), H.Staggeredfield_eo(l)
    H.pseudo_randomize(sf1, 13)
    H.pseudo_randomize(sf2, 31)
    H.DKS_eo(out, gf, sf1, H.EVEN)
    assert rel(H.squarenorm(out), 2030.1639500272767691) < 1e-12
    H.D_KS_eo(l, H.ODD)(out, gf, sf2)
    assert rel(H.squarenorm(out), 2076.7437224316167885) < 1e-12
    H.MdagM_eo(l, H.ODD)(out, gf, sf1, l.staggered_parameters())
    assert rel(H.squarenorm(out), 12670.438549218517437) < 1e-12
    H.MdagM_eo(l)(out, gf, sf2, l.staggered_parameters())
    assert rel(H.squarenorm(out), 12925.274439556447760) < 1e-12
    with pytest.raises(ValueError):
        H.MdagM_eo(l)(out, gf, sf2)  # MdagM_eo needs its additional parameters
    l = lattice("--nTime=4", "--fermionAction=rooted_stagg", "--mass=1.")
    gf = conf(l)
    sf1, out = H.Staggeredfield_eo(l), H.Staggeredfield_eo(l)
    H.pseudo_randomize(sf1, 123)
    H.DKS_eo(out, gf, sf1, H.EVEN)
    assert rel(H.squarenorm(out), 547.69039343718509372) < 1e-12
    H.MdagM_eo(l, H.ODD)(out, gf, sf1, l.staggered_parameters())
    assert rel(H.squarenorm(out), 2980.0540282922056576) < 1e-12


# solver_shifted_test.cpp:42-116
@pytest.mark.parametrize("merged", [False, True])
def test_cgm_1(merged):
    args = ["--nTime=4", "--fermionAction=rooted_stagg", "--mass=0.567", "--nDevices=1"]
    if merged:
        args.append("--useKernelMergingSpinor=true")
    l = lattice(*args)
    sigma = [0.0002065381736724, 0.00302707751065980, 0.0200732678058145, 0.12517586269872370, 1.0029328743375700]
    matrix = H.MdagM_eo(l)
    gf = conf(l)
    b = H.Staggeredfield_eo(l)
    out = [H.Staggeredfield_eo(l) for _ in sigma]
    H.pseudo_randomize(b, 13)
    it = H.cg_m(out, matrix, gf, sigma, b, l, 1e-23, l.staggered_parameters())
    sqnorm_b = H.squarenorm(b)
    aux = H.Staggeredfield_eo(l)
    for s, x in zip(sigma, out):
        matrix(aux, gf, x, l.staggered_parameters())
        H.saxpy(aux, s, x, aux)
        assert rel(H.squarenorm(aux), sqnorm_b) < 1e-10
    # the oracle's CG-M runs the same recurrence: same iteration count and solutions
    lat = O.lattice(4, 4)
    g = O.ildg_read(CONF_00200, lat)
    rc, it_o, xs = O.cgm(lat, g, O.pseudo_randomize_su3vec(O.veo(lat), 13), sigma, 0.567, prec=1e-23)
    assert rc == O.OK and abs(it - it_o) <= 10  # reductions sum in another order
    for x, xo in zip(out, xs):
        assert site_rel_err(x.get_host(), xo) < 1e-10


# solver_shifted_test.cpp:258-312
def test_cgm_4():
    l = lattice("--nTime=4", "--fermionAction=rooted_stagg", "--mass=0.01", "--nDevices=1")
    matrix = H.MdagM_eo(l)
    gf = conf(l)
    b = H.Staggeredfield_eo(l)
    out = [H.Staggeredfield_eo(l)]
    b.set_cold()
    H.cg_m(out, matrix, gf, [0.0], b, l, 1e-24, l.staggered_parameters())
    assert rel(H.squarenorm(out[0]), 9.0597433493689383255) < 1e-10
    H.pseudo_randomize(b, 123)
    H.cg_m(out, matrix, gf, [0.0], b, l, 1e-24, l.staggered_parameters())
    assert rel(H.squarenorm(out[0]), 3790.3193634090343949) < 1e-10


@pytest.mark.parametrize("action,op", [("wilson", "QplusQminus_eo"), ("twistedmass", "QplusQminus_eo")])
@pytest.mark.parametrize("merged", [False, True])
def test_cg_against_oracle(action, op, merged):
    args = ["--nTime=4", f"--fermionAction={action}", "--kappa=0.15", "--mu=0.02", "--solverMaxIterations=2000"]
    if merged:
        args.append("--useKernelMergingSpinor=true")
    l = lattice(*args)
    gf = conf(l)
    b, x = eo_pair(l, 31)
    x.cold()
    bh = b.get_host()
    prec = 1e-20 * H.squarenorm(b)
    it = H.cg(x, getattr(H, op)(l), gf, b, l, prec, l.wilson_parameters())
    lat = O.lattice(4, 4)
    g = O.ildg_read(CONF_00200, lat)
    act = O.WILSON if action == "wilson" else O.TWISTEDMASS
    x0 = np.zeros_like(bh)
    x0[:, 0::2] = 1.0 / np.sqrt(12 * O.volume(lat))  # Spinorfield_eo::cold
    rc, it_o, xo = O.cg_eo(lat, O.OP_QPLUSQMINUS_EO, g, bh, x0, act, 0.15, l.mubar, prec=prec, cgmax=2000)
    assert rc == O.OK and abs(it - it_o) <= 10  # same check-every-10 granularity
    xs = x.get_host()
    assert site_rel_err(xs, xo) < 1e-9
    r = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, xs, g, act, 0.15, l.mubar), bh, 1.0)
    assert O.norm2(r) < 2 * prec


@pytest.mark.parametrize("solver", ["bicgstab", "bicgstab_save"])
def test_bicgstab_against_oracle(solver):
    l = lattice("--nTime=4", "--fermionAction=twistedmass", "--kappa=0.15", "--mu=0.02", f"--solver={solver}")
    gf = conf(l)
    b, x = eo_pair(l, 32)
    x.cold()
    bh = b.get_host()
    prec = 1e-20 * H.squarenorm(b)
    it = H.bicgstab(x, H.Aee(l), gf, b, l, prec, l.wilson_parameters())
    lat = O.lattice(4, 4)
    g = O.ildg_read(CONF_00200, lat)
    x0 = np.zeros_like(bh)
    x0[:, 0::2] = 1.0 / np.sqrt(12 * O.volume(lat))
    rc, it_o, xo = O.bicgstab_eo(lat, O.OP_AEE, g, bh, x0, O.TWISTEDMASS, 0.15, l.mubar, prec=prec,
                                 save=(solver == "bicgstab_save"))
    assert rc == O.OK and abs(it - it_o) <= 1
    xs = x.get_host()
    assert site_rel_err(xs, xo) < 1e-8
    r = O.saxpy(O.apply_eo(lat, O.OP_AEE, xs, g, O.TWISTEDMASS, 0.15, l.mubar), bh, 1.0)
    assert O.norm2(r) < 2 * prec


def test_solver_exceptions():
    l = lattice("--nTime=4", "--solverMaxIterations=20")
    gf = conf(l)
    b, x = eo_pair(l, 33)
    with pytest.raises(H.SolverDidNotSolve) as e:
        H.cg(x, H.QplusQminus_eo(l), gf, b, l, 1e-60, l.wilson_parameters())
    assert e.value.args[1] == 20
    bh = b.get_host()
    bh[0, 0] = np.nan
    b.import_(bh)
    x.zero()
    with pytest.raises(H.SolverStuck) as e:
        H.cg(x, H.QplusQminus_eo(l), gf, b, l, 1e-20, l.wilson_parameters())
    assert e.value.args[1] == 0
    with pytest.raises(H.SolverStuck):
        H.bicgstab(x, H.Aee(l), gf, b, l, 1e-20, l.wilson_parameters())


def test_blas_dropin():
    l = lattice("--nTime=4")
    x, y, z, out = (H.Spinorfield_eo(l) for _ in range(4))
    for f, s in ((x, 1), (y, 2), (z, 3)):
        H.pseudo_randomize(f, s)
    X, Y, Z = x.get_host(), y.get_host(), z.get_host()
    a, b = 0.5 - 0.25j, 1.5 + 2j
    H.saxpy(out, a, x, y)
    assert site_rel_err(out.get_host(), O.saxpy(X, Y, a)) < 1e-15
    H.saxsbypz(out, a, x, b, y, z)
    assert site_rel_err(out.get_host(), O.saxsbypz(X, Y, Z, a, b)) < 1e-15
    assert abs(H.scalar_product(x, y) - O.dot(X, Y)) < 1e-12 * abs(O.dot(X, Y))
    assert rel(H.squarenorm(x), O.norm2(X)) < 1e-13
    out.cold()
    assert rel(H.squarenorm(out), 0.5) < 1e-14  # |phi|^2 of the whole lattice is 1, eo half


def test_convert_precision():
    l = lattice("--nTime=4")
    x, back = H.Spinorfield_eo(l), H.Spinorfield_eo(l)
    lo = H.Spinorfield_eo(l, H.F32)
    H.pseudo_randomize(x, 41)
    H.convert_precision(lo, x)
    H.convert_precision(back, lo)
    X = x.get_host()
    assert np.array_equal(back.get_host(), X.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("action,op", [("wilson", "QplusQminus_eo"), ("twistedmass", "QplusQminus_eo"),
                                       ("twistedmass", "Aee")])
def test_cg_reliable_update(action, op):
    """FP32 inner CG with FP64 reliable updates reaches |b - A x| / |b| <= 1e-10 in FP64 (checked with
    the oracle's FP64 operator) and agrees with the FP64 CG solution.  Aee is not hermitian; CG on it is
    not expected to converge and must raise a solver exception instead of returning garbage."""
    l = lattice("--nSpace=8", "--nTime=8", f"--fermionAction={action}", "--kappa=0.15", "--mu=0.02",
                "--beta=5.7")
    gf = H.Gaugefield(l, False)
    for _ in range(3):
        H.su3heatbath(gf, l, 1)
    b, _ = eo_pair(l, 51)
    x = H.Spinorfield_eo(l)
    A = getattr(H, op)(l)
    if op == "Aee":
        with pytest.raises((H.SolverDidNotSolve, H.SolverStuck)):
            H.cg_reliable_update(x, A, gf, b, l, 1e-10, l.wilson_parameters(), max_iterations=300)
        return
    st = H.cg_reliable_update(x, A, gf, b, l, 1e-10, l.wilson_parameters())
    assert st["true_relative_residual"] <= 1e-10 and st["reliable_updates"] >= 2
    lat = O.lattice(8, 8)
    g = gf.get_host()
    act = O.WILSON if action == "wilson" else O.TWISTEDMASS
    bh, xh = b.get_host(), x.get_host()
    r = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, xh, g, act, 0.15, l.mubar), bh, 1.0)
    assert np.sqrt(O.norm2(r) / O.norm2(bh)) <= 1.01e-10
    x64 = H.Spinorfield_eo(l)
    x64.zero()
    H.cg(x64, A, gf, b, l, 1e-24 * H.squarenorm(b), l.wilson_parameters())
    x64h = x64.get_host()
    assert np.sqrt(((xh - x64h) ** 2).sum() / (x64h ** 2).sum()) < 1e-6


@pytest.mark.parametrize("solver", ["single", "block", "stagg"])
def test_reliable_update_iteration_limit_keeps_the_correction(solver):
    """Stopped by the iteration limit before any reliable update (delta 1e-6, 20 iterations), the
    reduced-precision solvers still add the correction gathered so far to the FP64 solution: x (zero at the
    start) reduces the true residual, recomputed by the oracle's FP64 operator, and SolverDidNotSolve is
    raised."""
    l = lattice("--nSpace=8", "--nTime=8", "--fermionAction=twistedmass", "--kappa=0.15", "--mu=0.02",
                "--beta=5.7")
    gf = H.Gaugefield(l, False)
    for _ in range(3):
        H.su3heatbath(gf, l, 1)
    g = gf.get_host()
    lat = O.lattice(8, 8)
    opts = dict(delta=1e-6, max_iterations=20)
    if solver == "stagg":
        ls = lattice("--nSpace=8", "--nTime=8", "--fermionAction=rooted_stagg", "--mass=0.1")
        gs = H.Gaugefield(ls, False)
        gs.import_(g)
        bh = np.random.default_rng(53).standard_normal((ls.eo_sites, 6))
        b, x = H.Staggeredfield_eo(ls), H.Staggeredfield_eo(ls)
        b.import_(bh)
        x.set_zero()
        with pytest.raises(H.SolverDidNotSolve):
            H.cg_reliable_update_stagg(x, H.MdagM_eo(ls), gs, 0.0, b, ls, 1e-10, ls.staggered_parameters(), **opts)
        xs = [x.get_host()]
        res = [O.norm2(O.stagg_saxpby(O.mdagm_eo(lat, xs[0], g, 0.1), bh, -1.0, 1.0)) / O.norm2(bh)]
    else:
        A = H.QplusQminus_eo(l)
        bs = [eo_pair(l, 61 + r)[0] for r in range(2 if solver == "block" else 1)]
        xs_f = [H.Spinorfield_eo(l) for _ in bs]
        for x in xs_f:
            x.zero()
        with pytest.raises(H.SolverDidNotSolve):
            if solver == "block":
                H.cg_reliable_update_block(xs_f, A, gf, bs, l, 1e-10, l.wilson_parameters(), **opts)
            else:
                H.cg_reliable_update(xs_f[0], A, gf, bs[0], l, 1e-10, l.wilson_parameters(), **opts)
        xs, res = [], []
        for x, b in zip(xs_f, bs):
            xh, bh = x.get_host(), b.get_host()
            xs.append(xh)
            r = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, xh, g, O.TWISTEDMASS, 0.15, l.mubar), bh, 1.0)
            res.append(O.norm2(r) / O.norm2(bh))
    for xh, rr in zip(xs, res):
        assert np.abs(xh).max() > 0, "the correction of the last iterations was dropped"
        assert rr < 0.5, rr


def test_cg_reliable_update_half():
    """FP16 (per-site scale) inner CG with FP64 reliable updates reaches |b - A x| / |b| <= 1e-10 in FP64."""
    l = lattice("--nSpace=8", "--nTime=8", "--fermionAction=twistedmass", "--kappa=0.15", "--mu=0.02", "--beta=5.7")
    gf = H.Gaugefield(l, False)
    for _ in range(3):
        H.su3heatbath(gf, l, 1)
    b, _ = eo_pair(l, 52)
    x = H.Spinorfield_eo(l)
    A = H.QplusQminus_eo(l)
    st = H.cg_reliable_update(x, A, gf, b, l, 1e-10, l.wilson_parameters(), H.F16)
    assert st["true_relative_residual"] <= 1e-10 and st["reliable_updates"] >= 3
    lat = O.lattice(8, 8)
    bh, xh = b.get_host(), x.get_host()
    r = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, xh, gf.get_host(), O.TWISTEDMASS, 0.15, l.mubar), bh, 1.0)
    assert np.sqrt(O.norm2(r) / O.norm2(bh)) <= 1.01e-10


@pytest.mark.parametrize("prec", ["F32", "F16"])
def test_reliable_update_graph_replays_match_eager(prec):
    """The recorded inner iteration, replayed once per iteration with eager reliable updates and host reads
    in between, gives bit-identical solutions and iteration counts to the same iteration run eagerly."""
    l = lattice("--nSpace=8", "--nTime=8", "--fermionAction=twistedmass", "--kappa=0.15", "--mu=0.02", "--beta=5.7")
    gf = H.Gaugefield(l, False)
    for _ in range(3):
        H.su3heatbath(gf, l, 1)
    b, _ = eo_pair(l, 61)
    A = H.QplusQminus_eo(l)
    out = []
    for use_graph in (True, False, True):
        x = H.Spinorfield_eo(l)
        x.zero()
        st = H.cg_reliable_update(x, A, gf, b, l, 1e-10, l.wilson_parameters(), getattr(H, prec),
                                  use_graph=use_graph)
        assert st["true_relative_residual"] <= 1e-10
        # the captured iteration is kernel launches only
        assert st["graph_nodes"] == st["graph_kernel_nodes"] and (st["graph_nodes"] > 0) == use_graph
        out.append((st["iterations"], x.get_host()))
    for it, xh in out[1:]:
        assert it == out[0][0] and np.array_equal(xh, out[0][1])


def test_block_and_staggered_graph_replays_match_eager():
    l = lattice("--nSpace=8", "--nTime=8", "--fermionAction=twistedmass", "--kappa=0.15", "--mu=0.02", "--beta=5.7")
    gf = H.Gaugefield(l, False)
    for _ in range(3):
        H.su3heatbath(gf, l, 1)
    rng = np.random.default_rng(62)
    bs = []
    for _ in range(4):
        b = H.Spinorfield_eo(l)
        b.import_(rng.standard_normal((l.eo_sites, 24)))
        bs.append(b)
    A = H.QplusQminus_eo(l)
    res = []
    for use_graph in (True, False):
        xs = [H.Spinorfield_eo(l) for _ in bs]
        for x in xs:
            x.zero()
        it, per = H.cg_reliable_update_block(xs, A, gf, bs, l, 1e-10, l.wilson_parameters(), use_graph=use_graph)
        assert max(p["true_relative_residual"] for p in per) <= 1e-10
        assert all(p["graph_nodes"] == p["graph_kernel_nodes"] and (p["graph_nodes"] > 0) == use_graph for p in per)
        res.append((it, [x.get_host() for x in xs]))
    assert res[0][0] == res[1][0] and all(np.array_equal(a, b) for a, b in zip(res[0][1], res[1][1]))
    ls = lattice("--nTime=4", "--fermionAction=rooted_stagg", "--mass=0.05")
    g2 = conf(ls)
    M = H.MdagM_eo(ls)
    bst = H.Staggeredfield_eo(ls)
    H.pseudo_randomize(bst, 63)
    sres = []
    for use_graph in (True, False):
        x = H.Staggeredfield_eo(ls)
        x.set_zero()
        st = H.cg_reliable_update_stagg(x, M, g2, 0.01, bst, ls, 1e-10, ls.staggered_parameters(),
                                        use_graph=use_graph)
        assert st["true_relative_residual"] <= 1e-10
        assert st["graph_nodes"] == st["graph_kernel_nodes"] and (st["graph_nodes"] > 0) == use_graph
        sres.append((st["iterations"], x.get_host()))
    assert sres[0][0] == sres[1][0] and np.array_equal(sres[0][1], sres[1][1])
