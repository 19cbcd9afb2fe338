"""Point-source inversion and the pseudoscalar correlator against the reference's
inverter tests (executables/tests/CMakeLists.txt Inverter1-20: inverter_test_input_1
with inverter_test_ref; inverter_test_input_2 with inverter_test_ref2 for the default
z direction and inverter_test_ref3 for --correlatorDirection=0) at the reference's
0.1 % tolerance (testInverter.py) and at 1e-8, for its default BiCGStab on Aee and for
solver=cg (CG on QplusQminus_eo, then Qminus).  On the cold 4^4 configuration with
periodic boundaries t and z are equivalent, so inverter_test_ref holds for both.  The
eo solution is checked against the oracle's unpreconditioned operator, M phi = eta.
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
ARGS = ["--nSpace=4", "--nTime=4", "--kappa=0.15", "--mu=4.", "--fermionAction=twistedmass"]


@pytest.mark.parametrize("solver", ["bicgstab", "cg"])
@pytest.mark.parametrize("conf,direction,ref", [(None, 0, REF1), (None, 3, REF1), ("conf", 3, REF2), ("conf", 0, REF3)],
                         ids=["cold-t", "cold-z", "conf00200-z", "conf00200-t"])
def test_ps_correlator_goldens(solver, conf, direction, ref):
    l = H.Lattice(ARGS + [f"--solver={solver}", f"--correlatorDirection={direction}"])
    gf = H.Gaugefield(l, CONF_00200) if conf else H.Gaugefield(l, False)
    c = H.point_source_ps_correlator(gf, l)
    assert len(c) == 4
    for got, want in zip(c, ref):
        assert abs(got - want) <= 1e-3 * want  # testInverter.py tolerance
        assert abs(got - want) <= 1e-8 * want  # what FP64 actually reaches


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
