"""Rational approximations for the multi-shift solver's shifts (host side, no GPU).

Pinned by the reference's own tests (physics/algorithms/rational_approximation_test.cpp; BOOST_CHECK_CLOSE
tolerances are percentages, so 1e-6 there is 1e-8 relative here):
* `coefficients` (:110-135): degree 5, x^(-1/2) on [1e-3, 1] -- a0, a_k, b_k and the error (Lüscher's
  coefficients; the b_k are the shifts the multi-shift tests use);
* `errors` (:80-108): the minimax errors of degree 6, 12 and 24 on [1e-5, 1], Lüscher's values, to 1e-10;
* `rescale` (:137-212): degree 15, x^(1/4) on [1e-5, 1] rescaled to two eigenvalue intervals;
and by what the mathematics fixes: the relative error equioscillates (its extremes are +-error, with the
2d + 2 alternations of a minimax solution), which also shows that the reference's saved degree-6 file
(tests/golden/rational_approximation_test_input, read here for its format) is slightly off the minimax
solution: its coefficients are ours times (1 + 2.5e-7), its b_k equal ours.
"""
import os

import numpy as np
import pytest

from conftest import ROOT

H = pytest.importorskip("cl2qcd_amd._host")

GOLDEN = os.path.join(ROOT, "tests", "golden", "rational_approximation_test_input")


def rel(a, b):
    return abs(a - b) / abs(b)


def rel_err_curve(r, expo, lo, hi, n=200001):
    x = np.exp(np.linspace(np.log(lo), np.log(hi), n))
    approx = r.a0() + sum(a / (x + b) for a, b in zip(r.a(), r.b()))
    return approx / x ** expo - 1


def test_coefficients_golden():
    r = H.Rational_Approximation(5, 1, 2, 0.001, 1)
    res = [0.0194661374678695, 0.03561358805557574, 0.0821572765515051, 0.21113622523593300, 0.7946025292155642]
    pol = [0.0002065381736724, 0.00302707751065980, 0.0200732678058145, 0.12517586269872370, 1.0029328743375700]
    assert r.order() == 5 and r.exponent() == -0.5
    assert rel(r.a0(), 0.2801782460422457) < 1e-8
    for a, b, ra, rb in zip(r.a(), r.b(), res, pol):
        assert rel(a, ra) < 1e-8 and rel(b, rb) < 1e-8
    assert rel(r.error(), 5.3847952988591471e-05) < 1e-8


@pytest.mark.parametrize("d,err", [(6, 5.0287212083799089e-04), (12, 1.2615414383737566e-07),
                                   (24, 7.9394311806793910e-15)])
def test_errors_against_luescher(d, err):
    r = H.Rational_Approximation(d, 1, 2, 1e-5, 1)
    assert rel(r.error(), err) < 1e-10
    if d <= 12:  # double-precision coefficients resolve the error curve to ~1e-13
        e = rel_err_curve(r, -0.5, 1e-5, 1.0)
        assert abs(e.max() - r.error()) < 1e-6 * r.error() and abs(-e.min() - r.error()) < 1e-6 * r.error()
        sign_changes = np.count_nonzero(np.diff(np.sign(e[np.abs(e) > 0.5 * r.error()])))
        assert sign_changes == 2 * d + 1  # 2d + 2 alternating extrema


def test_rescale_golden():
    r = H.Rational_Approximation(15, 1, 4, 1e-5, 1, False)
    assert r.exponent() == 0.25
    a0_ref = 3.78396627036665123
    a_ref = [-2.722986658932525683e-07, -1.2631475639728917521e-06, -4.360435755914117958e-06,
             -1.4160691433260296987e-05, -4.5207926211912900696e-05, -0.00014352921651252466598,
             -0.00045497933178524781715, -0.0014431546521933168083, -0.0045926953908840342788,
             -0.014747783330565073998, -0.048456946841957317107, -0.16880046472141346792,
             -0.68431552061715394952, -4.2198332136416603078, -117.03837887995429412]
    b_ref = [9.2907369101588763806e-06, 5.0627210699159516975e-05, 0.00016198783558680096995,
             0.00044474583457844121164, 0.0011563267279465730842, 0.0029447969579766914219,
             0.007440491639209977949, 0.018751086281155571189, 0.047271072978110562079,
             0.11959352757092110708, 0.30563294009891067704, 0.80189441104810432748,
             2.2562393220051499831, 7.7746421837770229857, 59.252309299420609534]
    a0_cons = 3.8304052181004228927
    a_cons = [-2.8942286130576286221e-07, -1.3425838169907005556e-06, -4.6346528686679587346e-06,
              -1.5051222595004849162e-05, -4.8050941840019680289e-05, -0.00015255541700051565286,
              -0.00048359186633654911497, -0.0015339111096058541221, -0.0048815187425726713766,
              -0.015675235262161982264, -0.051504285410779379606, -0.17941591204330820108,
              -0.72735044574402074602, -4.4852081772741634325, -124.39863554566063897]
    b_cons = [9.7552862512145591676e-06, 5.3158639325027219212e-05, 0.00017008744523117717558,
              0.00046698372446689161283, 0.0012141446195419417376, 0.003092040766471044512,
              0.007812526228236587808, 0.019688665814428581158, 0.04963469020409881638,
              0.12557336479641045823, 0.32091499816392937694, 0.84199021010590602287,
              2.3690543226274733968, 8.1633847494467222106, 62.215004455600926292]
    c = r.rescale(0.3485318319429664, 5.2827906935473500)
    assert c["order"] == 15 and rel(c["a0"], a0_ref) < 5e-7
    for a, b, ra, rb in zip(c["a"], c["b"], a_ref, b_ref):
        assert rel(a, ra) < 5e-7 and rel(b, rb) < 5e-7
    c = r.rescale(0.321489, 5.546930228224717)
    assert rel(c["a0"], a0_cons) < 5e-7
    for a, b, ra, rb in zip(c["a"], c["b"], a_cons, b_cons):
        assert rel(a, ra) < 2e-6 and rel(b, rb) < 2e-6
    with pytest.raises(Exception):  # [1e-5, 1] must cover min/max
        r.rescale(1e-6, 1.0)


def test_reference_file_format_and_minimax(tmp_path):
    f = H.Rational_Approximation(GOLDEN)
    r = H.Rational_Approximation(6, 1, 2, 1e-5, 1)
    assert f.order() == 6 and f.exponent() == -0.5 and rel(f.error(), r.error()) < 1e-11
    for b1, b2 in zip(f.b(), r.b()):
        assert rel(b1, b2) < 1e-9
    ratio = [a1 / a2 for a1, a2 in zip(f.a(), r.a())] + [f.a0() / r.a0()]
    assert max(ratio) - min(ratio) < 1e-9 and abs(ratio[0] - 1 - 2.53e-7) < 1e-8
    # ours equioscillates, the file's does not
    e = rel_err_curve(r, -0.5, 1e-5, 1.0)
    assert abs(e.max() + e.min()) < 1e-9 * r.error()
    e = rel_err_curve(f, -0.5, 1e-5, 1.0)
    assert abs(e.max() + e.min()) > 1e-4 * f.error()
    # save and read back
    p = str(tmp_path / "approx.txt")
    r.save(p)
    g = H.Rational_Approximation(p)
    assert g.order() == 6 and rel(g.a0(), r.a0()) < 1e-15 and rel(g.error(), r.error()) < 1e-15
    for x, y in zip(g.a() + g.b(), r.a() + r.b()):
        assert rel(x, y) < 1e-15


def test_rejects_bad_arguments():
    for args in ((0, 1, 2, 1e-3, 1), (4, 2, 2, 1e-3, 1), (4, 1, 2, 1.0, 0.5), (4, -1, 2, 1e-3, 1)):
        with pytest.raises(Exception):
            H.Rational_Approximation(*args)
