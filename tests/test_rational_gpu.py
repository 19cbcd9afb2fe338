"""Rational-approximation shifts through the multi-shift solvers on the GPU.

The reference's cgm_2 and cgm_3 tests (physics/algorithms/solver_shifted_test.cpp:117-257) take the poles b_k
of Rational_Approximation(8, 1, 2, 1e-5, 1) and (16, 1, 2, 1e-5, 1) as the shifts of cg_m on MdagM_eo
(conf.00200, source pseudo_randomize seed 123) at mass 1.01335 / prec 1e-12 and mass 0.01 / prec 1e-24,
and compare |x_k|^2 with the reference code's values to 1e-8 percent (1e-10 relative).  Here the shifts come
from this build's own Remez exchange, so the goldens pin the shifts and the solver together.

The mixed-precision multi-shift solver (FP32 CG-M, FP64 refinement of every shift) is held to what it
promises on the same shifts: every shift's FP64 residual |b - (MdagM + s_k) x_k| <= 1e-10 |b|, recomputed
with the oracle's operator, and |x_k|^2 within the bound that residual allows around the reference values,
2 cond(MdagM + s_k) 1e-10 with cond <= (m^2 + 16 + s_k) / (m^2 + s_k) (D_KS_eo is a sum of 8 hops of
weight 1/2, so |D_eo D_oe| <= 16).
"""
import numpy as np
import pytest

import oracle as O
from conftest import CONF_00200
from fields import rel

pytestmark = pytest.mark.gpu

CGM2 = [59.877118728264179026, 59.875307563978779513, 59.865246895125224569, 59.811193256571264953,
        59.521693061471694364, 57.988143162101970063, 50.327004662008008040, 22.236536652925686042]
CGM3 = [3790.2414703421331978, 3789.5266994555458950, 3787.5285290870838253, 3782.6617780789301833,
        3771.0928645597095965, 3743.8629449816535271, 3680.7256977108900173, 3539.0807487819220114,
        3243.3813294554852291, 2708.4856234685380514, 1947.5005180873681638, 1157.0234695785266013,
        559.56186864729193076, 215.02535309652577666, 58.529059535207736076, 6.2407847688851161294]


def residuals(l, g, xs, b, sigma, mass):
    lat = O.lattice(l.nspace, l.ntime)
    bh = b.get_host()
    out = []
    for s, x in zip(sigma, xs):
        xh = x.get_host()
        ax = O.mdagm_eo(lat, xh, g, mass) + s * xh
        out.append(np.sqrt(O.norm2(ax - bh) / O.norm2(bh)))
    return out


CASES = [
    # degree, mass, prec, goldens, merged spinor kernels (cgm_2 runs both)
    (8, 1.01335, 1e-12, CGM2, False),
    (8, 1.01335, 1e-12, CGM2, True),
    (16, 0.01, 1e-24, CGM3, False),
]


@pytest.mark.parametrize("degree,mass,prec,golden,merged", CASES, ids=["cgm_2", "cgm_2_merged", "cgm_3"])
def test_cg_m_remez_shifts_golden(degree, mass, prec, golden, merged):
    from cl2qcd_amd import _host as H
    args = ["--nTime=4", "--fermionAction=rooted_stagg", f"--mass={mass}", "--nDevices=1"]
    if merged:
        args.append("--useKernelMergingSpinor=true")
    l = H.Lattice(args)
    sigma = H.Rational_Approximation(degree, 1, 2, 1e-5, 1.0).b()
    assert len(sigma) == degree and all(s > 0 for s in sigma)
    gf = H.Gaugefield(l, CONF_00200)
    b = H.Staggeredfield_eo(l)
    H.pseudo_randomize(b, 123)
    out = [H.Staggeredfield_eo(l) for _ in sigma]
    it = H.cg_m(out, H.MdagM_eo(l), gf, sigma, b, l, prec, l.staggered_parameters())
    assert it > 0
    for k, (x, ref) in enumerate(zip(out, golden)):
        assert rel(H.squarenorm(x), ref) < 1e-10, k


@pytest.mark.parametrize("degree,mass,golden", [(8, 1.01335, CGM2), (16, 0.01, CGM3)], ids=["cgm_2", "cgm_3"])
def test_cg_m_mixed_remez_shifts(degree, mass, golden):
    from cl2qcd_amd import _host as H
    l = H.Lattice(["--nTime=4", "--fermionAction=rooted_stagg", f"--mass={mass}", "--solverMaxIterations=20000"])
    sigma = H.Rational_Approximation(degree, 1, 2, 1e-5, 1.0).b()
    gf = H.Gaugefield(l, CONF_00200)
    b = H.Staggeredfield_eo(l)
    H.pseudo_randomize(b, 123)
    xs = [H.Staggeredfield_eo(l) for _ in sigma]
    total, ms, per = H.cg_m_mixed(xs, H.MdagM_eo(l), gf, sigma, b, l, 1e-10, l.staggered_parameters())
    assert all(p["true_relative_residual"] <= 1e-10 for p in per)
    g = O.ildg_read(CONF_00200, O.lattice(4, 4))
    assert max(residuals(l, g, xs, b, sigma, mass)) <= 1.01e-10
    for k, (x, s, ref) in enumerate(zip(xs, sigma, golden)):
        cond = (mass ** 2 + 16 + s) / (mass ** 2 + s)
        assert rel(H.squarenorm(x), ref) <= 2 * cond * 1.01e-10, k
