"""Point-source inversion and the pseudoscalar correlator against the reference's
inverter tests (executables/tests/CMakeLists.txt Inverter1-28: inverter_test_input_1
with inverter_test_ref; inverter_test_input_2 with inverter_test_ref2 for the default
z direction and inverter_test_ref3 for --correlatorDirection=0; the Wilson action
(--fermionAction=wilson) on input_1 with inverter_test_ref4 and on input_2 with
inverter_test_ref5) at the reference's 0.1 % tolerance (testInverter.py) and at 1e-8,
for its default BiCGStab on Aee, bicgstab_save and solver=cg (BiCGStab first, CG on
QplusQminus_eo only if BiCGStab throws), and for the 12 sources solved as one lockstep
block (FP32 or FP16 inner, FP64 reliable updates).  On the cold 4^4 configuration with periodic
boundaries t and z are equivalent, so inverter_test_ref holds for both.  The eo solution
is checked against the oracle's unpreconditioned operator, M phi = eta.

The CG fallback is exercised where BiCGStab really breaks down: the Wilson operator on
conf.00200 at kappa = 0.2, where the oracle's BiCGStab (the reference's bicgstab_fast)
stops with SolverStuck and CG on QplusQminus_eo from its last iterate converges.

Volume, time-slice and z-slice sources (physics/sources.cpp:63-110) are checked for their
content distributions and support, and the stochastic z correlator against the reference
kernel's own double sum over the slice sites and the source sites, done by brute force on
the downloaded fields (ocl_kernel/fermionobservables_correlators_stochastic.cl:78-133).
"""
import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import _host as H
from conftest import CONF_00200

pytestmark = pytest.mark.gpu

REF1 = [6.35347927764573e-03, 1.21141739833823e-04, 2.40482229187376e-06, 1.21141739833823e-04]
REF2 = [6.40008946516344e-03, 1.17492646085198e-04, 2.50278817534017e-06, 1.18464343697740e-04]
REF3 = [6.39838341066878e-03, 1.18922612884654e-04, 2.48500477806187e-06, 1.18758214790220e-04]
REF4 = [1.83008474327911e-02, 3.14850262023767e-03, 7.81590667980677e-04, 3.14850262023767e-03]
REF5 = [2.23946028664745e-02, 2.05485346957700e-03, 6.96566865496375e-04, 2.17148162775779e-03]
ARGS = ["--nSpace=4", "--nTime=4", "--kappa=0.15", "--mu=4.", "--fermionAction=twistedmass"]
WILSON = ["--nSpace=4", "--nTime=4", "--kappa=0.15", "--mu=4.", "--fermionAction=wilson"]

GOLDEN_CASES = [(ARGS, None, 0, REF1), (ARGS, None, 3, REF1), (ARGS, "conf", 3, REF2), (ARGS, "conf", 0, REF3),
                (WILSON, None, 3, REF4), (WILSON, "conf", 3, REF5)]
GOLDEN_IDS = ["tm-cold-t", "tm-cold-z", "tm-conf00200-z", "tm-conf00200-t", "wilson-cold-z", "wilson-conf00200-z"]


@pytest.mark.parametrize("solver", ["bicgstab", "bicgstab_save", "cg"])
@pytest.mark.parametrize("args,conf,direction,ref", GOLDEN_CASES, ids=GOLDEN_IDS)
def test_ps_correlator_goldens(solver, args, conf, direction, ref):
    l = H.Lattice(args + [f"--solver={solver}", f"--correlatorDirection={direction}"])
    gf = H.Gaugefield(l, CONF_00200) if conf else H.Gaugefield(l, False)
    c = H.point_source_ps_correlator(gf, l)
    assert len(c) == 4
    for got, want in zip(c, ref):
        assert abs(got - want) <= 1e-3 * want  # testInverter.py tolerance
        assert abs(got - want) <= 1e-8 * want  # what FP64 actually reaches


@pytest.mark.parametrize("inner", ["f32", "f16"])
@pytest.mark.parametrize("args,conf,direction,ref", GOLDEN_CASES, ids=GOLDEN_IDS)
def test_ps_correlator_goldens_block(args, conf, direction, ref, inner):
    """the 12 point sources as one 12-RHS lockstep block (FP32 or FP16 inner loop, each source to 1e-10 in
    FP64)"""
    l = H.Lattice(args + [f"--correlatorDirection={direction}"])
    gf = H.Gaugefield(l, CONF_00200) if conf else H.Gaugefield(l, False)
    c = H.point_source_ps_correlator(gf, l, block=True, inner_prec=H.F16 if inner == "f16" else H.F32)
    for got, want in zip(c, ref):
        assert abs(got - want) <= 1e-8 * want


def test_cg_fallback_after_bicgstab_breakdown():
    args = ["--nSpace=4", "--nTime=4", "--kappa=0.2", "--fermionAction=wilson", "--solver=cg"]
    l = H.Lattice(args)
    gf = H.Gaugefield(l, CONF_00200)
    lat = O.lattice(4, 4)
    g = O.ildg_read(CONF_00200, lat)
    eta = np.zeros((O.volume(lat), 24))
    eta[0, 0] = 1.0
    full = H.Spinorfield(l)
    full.import_(eta)
    ee, eo = H.Spinorfield_eo(l), H.Spinorfield_eo(l)
    H.convert_to_eoprec(ee, eo, full)
    pe, po = H.Spinorfield_eo(l), H.Spinorfield_eo(l)
    st = H.invert_eo(pe, po, gf, ee, eo, l, 1e-24)
    assert st["cg_fallback"] and st["bicgstab_iterations"] > 0 and st["iterations"] > 0
    H.convert_from_eoprec(full, pe, po)
    r = O.m_full(lat, full.get_host(), g, 0.2, 0.0, O.WILSON) - eta
    assert np.sqrt((r ** 2).sum()) < 1e-10
    # the same input without the fallback (BiCGStab alone) fails the way the reference's does
    l2 = H.Lattice(args[:-1] + ["--solver=bicgstab"])
    with pytest.raises(H.SolverException):
        H.invert_eo(pe, po, H.Gaugefield(l2, CONF_00200), ee, eo, l2, 1e-24)


def test_cg_mode_keeps_bicgstab_when_it_converges():
    l = H.Lattice(ARGS + ["--solver=cg"])
    gf = H.Gaugefield(l, CONF_00200)
    full = H.Spinorfield(l)
    eta = np.zeros((256, 24))
    eta[3, 7] = 1.0
    full.import_(eta)
    ee, eo = H.Spinorfield_eo(l), H.Spinorfield_eo(l)
    H.convert_to_eoprec(ee, eo, full)
    pe, po = H.Spinorfield_eo(l), H.Spinorfield_eo(l)
    st = H.invert_eo(pe, po, gf, ee, eo, l, 1e-24)
    assert not st["cg_fallback"] and st["iterations"] == st["bicgstab_iterations"] > 0


@pytest.mark.parametrize("content", ["one", "z4", "gaussian", "z2"])
@pytest.mark.parametrize("stype", ["volume", "timeslice", "zslice"])
def test_random_sources(stype, content):
    l = H.Lattice(["--nSpace=4", "--nTime=6", f"--sourceType={stype}", f"--sourceContent={content}",
                   "--sourceT=2", "--sourceZ=1", "--hostSeed=7"])
    # site = x + 4 (y + 4 (z + 4 t)): the site axis reshapes to [t, z, y, x]
    srcs = [np.asarray(s).reshape(6, 4, 4, 4, 24) for s in H.create_sources(l, 2)]
    for s in srcs:
        keep = np.zeros((6, 4, 4, 4), dtype=bool)
        if stype == "volume":
            keep[...] = True
        elif stype == "timeslice":
            keep[2] = True
        else:
            keep[:, 1] = True
        assert np.all(s[~keep] == 0.0)
        v = s[keep].reshape(-1, 12, 2)
        if content == "one":
            assert np.all(v[..., 0] == 1.0) and np.all(v[..., 1] == 0.0)
        elif content == "z4":
            assert np.allclose(np.abs(v), 1 / np.sqrt(2), rtol=0, atol=1e-15)
        elif content == "z2":
            assert np.all(np.abs(v[..., 0]) == 1.0) and np.all(v[..., 1] == 0.0)
        else:
            assert abs(v.mean()) < 0.1 and abs(v.var() - 1.0) < 0.1
    if content != "one":  # fresh draws per source
        assert not np.array_equal(srcs[0], srcs[1])


@pytest.mark.parametrize("stype", ["volume", "zslice", "timeslice"])
def test_stochastic_z_correlator_brute_force(stype):
    """C(z) = (2 kappa)^2 / (NS^2 NT) sum_k sum_{x in slice z} sum_{y} sum_{a,b} |phi_k,a(x) conj(b_k,b(y))|^2 as
    the reference's kernel writes it, by brute force over the downloaded fields"""
    ns, nt, nsrc = 4, 4, 2
    args = ARGS + [f"--sourceType={stype}", "--sourceContent=z4", f"--nSources={nsrc}", "--sourceZ=2",
                   "--sourceT=1", "--hostSeed=11", "--correlatorDirection=3"]
    l = H.Lattice(args)
    gf = H.Gaugefield(l, CONF_00200)
    c = H.ps_correlator(gf, l)
    # the same sources (same seed, same draws) and their solutions
    l2 = H.Lattice(args)
    srcs = H.create_sources(l2, nsrc)
    want = np.zeros(ns)
    for s in srcs:
        full = H.Spinorfield(l2)
        full.import_(np.asarray(s))
        ee, eo = H.Spinorfield_eo(l2), H.Spinorfield_eo(l2)
        H.convert_to_eoprec(ee, eo, full)
        pe, po = H.Spinorfield_eo(l2), H.Spinorfield_eo(l2)
        H.invert_eo(pe, po, gf, ee, eo, l2, 1e-23)
        H.convert_from_eoprec(full, pe, po)
        phi = full.get_host().reshape(nt, ns, ns, ns, 12, 2)  # [t, z, y, x, component, re/im]
        phi = phi[..., 0] + 1j * phi[..., 1]
        b = np.asarray(s).reshape(-1, 12, 2)
        b = b[..., 0] + 1j * b[..., 1]
        bb = np.conj(b).ravel()
        for z in range(ns):
            px = phi[:, z].reshape(-1, 12)
            want[z] += sum(np.sum(np.abs(np.outer(px[i], bb)) ** 2) for i in range(px.shape[0]))
    want *= (2 * 0.15) ** 2 / (ns * ns * nt)
    assert np.allclose(c, want, rtol=1e-9, atol=0)


def test_ps_correlator_point_equals_point_source_ps_correlator():
    l = H.Lattice(ARGS + ["--correlatorDirection=3"])
    gf = H.Gaugefield(l, CONF_00200)
    a = np.array(H.ps_correlator(gf, l))
    b = np.array(H.point_source_ps_correlator(gf, l))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("source", [(0, 0, 0, 0), (1, 2, 3, 1)], ids=["even", "odd"])
def test_eo_inversion_solves_full_operator(source):
    x, y, z, t = source
    l = H.Lattice(ARGS)
    gf = H.Gaugefield(l, CONF_00200)
    lat = O.lattice(4, 4)
    vol = O.volume(lat)
    eta = np.zeros((vol, 24))
    site = x + 4 * (y + 4 * (z + 4 * t))
    eta[site, 4] = 1.0
    full = H.Spinorfield(l)
    full.import_(eta)
    ee, eo = H.Spinorfield_eo(l), H.Spinorfield_eo(l)
    H.convert_to_eoprec(ee, eo, full)
    pe, po = H.Spinorfield_eo(l), H.Spinorfield_eo(l)
    H.invert_eo(pe, po, gf, ee, eo, l, 1e-26)
    H.convert_from_eoprec(full, pe, po)
    phi = full.get_host()
    g = O.ildg_read(CONF_00200, lat)
    r = O.m_full(lat, phi, g, 0.15, 2 * 0.15 * 4.0, O.TWISTEDMASS) - eta
    assert np.sqrt((r ** 2).sum()) < 1e-11
