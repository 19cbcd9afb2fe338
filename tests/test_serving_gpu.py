"""Inversion serving engine on the GPU.

* every answer meets |b - A x| / |b| <= 1e-10 in FP64 (recomputed by the engine and, for
  explicit sources, by the oracle's operator);
* the same request gives the same solution alone and in a concurrent mix (merged into
  another request's multi-RHS solve, overlapping other solves on other streams) to
  1e-6 relative -- both answers meet the 1e-10 residual, and 1e-6 bounds cond(A) x 1e-10
  at these lattices;
* compatible requests (one action on one configuration, any masses) are merged, incompatible ones overlap in
  time, a tiny HBM budget
  serialises the solves, and the FP64 one-at-a-time mode agrees with the mixed mode.
"""
import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import abi

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_gaussian_source():
    with abi.Context(8, 8) as ctx:
        a, b, c = ctx.spinor(), ctx.spinor(), ctx.spinor()
        ctx.gaussian_eo(abi.FIELD_SPINOR_EO, a.ptr, abi.F64, 7)
        ctx.gaussian_eo(abi.FIELD_SPINOR_EO, b.ptr, abi.F64, 7)
        ctx.gaussian_eo(abi.FIELD_SPINOR_EO, c.ptr, abi.F64, 8)
        x = a.download()
        assert np.array_equal(x, b.download()) and not np.array_equal(x, c.download())
        assert abs(x.mean()) < 0.02 and abs(x.var() - 1.0) < 0.02
        v = ctx.su3vec(abi.F32)
        ctx.gaussian_eo(abi.FIELD_SU3VEC_EO, v.ptr, abi.F32, 9)
        y = v.download()
        assert abs(y.mean()) < 0.03 and abs(y.var() - 1.0) < 0.03


def solve_spans(results):
    """(start, end) of every solve, in start order: a solve runs from its first admission to its last answer
    (requests joined into a running lockstep solve share its span)"""
    spans = {}
    for m in results:
        a, f = spans.get(m["solve"], (m["admitted"], m["finished"]))
        spans[m["solve"]] = (min(a, m["admitted"]), max(f, m["finished"]))
    return sorted(spans.values())


def make_engine(H, **kw):
    e = H.Engine(**kw)
    c0 = e.load_config(8, 8, 5.7, 5, 11)
    c1 = e.load_config(6, 12, 5.7, 5, 12)
    e.warmup()
    return e, c0, c1


def requests(e, c0, c1):
    rng = np.random.default_rng(5)
    out = []
    # with two workers the first two run at once and the rest queue, so the three twisted-mass requests on
    # config c0 are merged when a worker frees up, one of them with its own kappa and mu, and so are the two
    # staggered requests on config c1, of different masses
    for cfg, action, kappa, mu, mass, n in ((c1, "twistedmass", 0.155, 0.01, 0, 1), (c0, "staggered", 0, 0, 0.05, 2),
                                             (c0, "twistedmass", 0.15, 0.02, 0, 2), (c1, "staggered", 0, 0, 0.1, 1),
                                             (c0, "twistedmass", 0.155, 0.005, 0, 3), (c0, "twistedmass", 0.15, 0.02, 0, 1),
                                             (c1, "staggered", 0, 0, 0.07, 2)):
        w = 24 if action == "twistedmass" else 6
        srcs = [rng.standard_normal((e.eo_sites(cfg), w)) for _ in range(n)]
        out.append(dict(config=cfg, action=action, kappa=kappa, mu=mu, mass=mass, sources=srcs))
    return out


def test_same_alone_and_in_mix():
    from cl2qcd_amd import _host as H
    e, c0, c1 = make_engine(H, workers=2)
    reqs = requests(e, c0, c1)
    alone = []
    for r in reqs:
        res = e.wait(e.submit(**r))
        assert res["ok"], res["error"]
        assert max(res["true_rel_residual"]) <= 1e-10
        alone.append(res)
    ids = [e.submit(**r) for r in reqs]
    mix = [e.wait(i) for i in ids]
    for r, a, m in zip(reqs, alone, mix):
        assert m["ok"], m["error"]
        assert max(m["true_rel_residual"]) <= 1e-10
        for xa, xm in zip(a["solutions"], m["solutions"]):
            assert rel(xm, xa) < 1e-6
    # the twisted-mass requests on config 0 were merged into one solve despite their different masses
    merged = [m["batch_requests"] for r, m in zip(reqs, mix) if r["config"] == c0 and r["action"] == "twistedmass"]
    assert max(merged) >= 2
    merged = [m["batch_requests"] for r, m in zip(reqs, mix) if r["config"] == c1 and r["action"] == "staggered"]
    assert max(merged) >= 2
    # solves overlapped in time
    spans = solve_spans(mix)
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))
    # the oracle agrees that the first answer solves its system
    lat = O.lattice(8, 8, theta_t=1.0)
    l = H.Lattice(["--nSpace=8", "--nTime=8", "--beta=5.7", "--hostSeed=11", "--thetaFermionTemporal=1"])
    gf = H.Gaugefield(l, False)  # the engine's config 0, made again from the same seed
    for _ in range(5):
        H.su3heatbath(gf, l, 1)
    g = gf.get_host()
    x = mix[2]["solutions"][0]
    b = reqs[2]["sources"][0]
    r = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, x, g, O.TWISTEDMASS, 0.15, 2 * 0.15 * 0.02), b, 1.0)
    assert np.sqrt(O.norm2(r) / O.norm2(b)) <= 1.01e-10


def test_budget_serialises_and_fp64_mode_agrees():
    from cl2qcd_amd import _host as H
    e, c0, c1 = make_engine(H, workers=4, hbm_budget_bytes=1)
    reqs = requests(e, c0, c1)
    ids = [e.submit(**r) for r in reqs]
    res = [e.wait(i) for i in ids]
    spans = solve_spans(res)
    assert all(spans[k + 1][0] >= spans[k][1] for k in range(len(spans) - 1))
    f, d0, d1 = make_engine(H, workers=2, mixed=False)
    ids = [f.submit(**r) for r in reqs]
    ref = [f.wait(i) for i in ids]
    for a, b in zip(res, ref):
        assert a["ok"] and b["ok"], (a["error"], b["error"])
        assert b["batch_requests"] == 1 and max(b["true_rel_residual"]) <= 1e-10
        for xa, xb in zip(a["solutions"], b["solutions"]):
            assert rel(xa, xb) < 1e-6


def test_gaussian_requests_and_errors():
    from cl2qcd_amd import _host as H
    e, c0, c1 = make_engine(H, workers=2)
    i = e.submit(config=c1, action="twistedmass", kappa=0.15, mu=0.01, gaussian_seeds=[1, 2, 3],
                 return_solutions=False)
    j = e.submit(config=c0, action="staggered", mass=0.05, gaussian_seeds=[4])
    a, b = e.wait(i), e.wait(j)
    assert a["ok"] and len(a["solutions"]) == 0 and len(a["true_rel_residual"]) == 3
    assert max(a["true_rel_residual"]) <= 1e-10
    assert b["ok"] and b["solutions"][0].shape == (e.eo_sites(c0), 6)
    with pytest.raises(Exception):
        e.submit(config=c0, action="twistedmass", sources=[np.zeros((5, 24))])
    e.drain()


def test_merge_respects_budget_next_to_running_solve():
    """A budget that fits one running 1-RHS solve and a 2-RHS merged solve: queued compatible 1-RHS requests
    behind a 1-RHS head join it only while the merged solve still fits next to the running one."""
    from cl2qcd_amd import _host as H
    veo = 8 * 8 * 8 * 8 // 2
    per_rhs = veo * (6 * 192 + 6 * 96 + 2 * 192)  # the engine's twisted-mass work fields per padded RHS
    budget = 3 * per_rhs
    e = H.Engine(workers=2, hbm_budget_bytes=budget)
    c0 = e.load_config(8, 8, 5.7, 5, 11)
    e.warmup()
    assert e.eo_sites(c0) == veo
    first = e.submit(config=c0, action="twistedmass", kappa=0.16, mu=0.003, gaussian_seeds=[1])
    rest = [e.submit(config=c0, action="twistedmass", kappa=0.15, mu=0.01, gaussian_seeds=[10 + k],
                     return_solutions=False) for k in range(8)]
    res = [e.wait(i) for i in [first] + rest]
    assert all(r["ok"] and max(r["true_rel_residual"]) <= 1e-10 for r in res)
    assert e.peak_work_bytes() <= budget
    # the merged solves stayed within the 2 RHS that fit (12 would have fitted the RHS limit)
    assert max(r["batch_rhs"] for r in res) <= 2


@pytest.mark.parametrize("inner", ["f32", "f16"])
def test_running_block_takes_queued_requests_into_free_slots(inner):
    """Continuous batching: with one worker busy on a 3-RHS twisted-mass solve (padded to 4 slots), a 1-RHS
    request of another quark mass on the same configuration joins the running solve's free slot at its next
    reliable update instead of waiting for the worker; two more join slots that converged.  Every answer meets
    1e-10 (the oracle recomputes one) and equals the same request solved alone to 1e-6; the first request's
    answer is published before the solve ends.  With an FP32 and with an FP16 (per-site norm) inner loop."""
    from cl2qcd_amd import _host as H
    e = H.Engine(workers=1, inner_precision=H.F16 if inner == "f16" else H.F32)
    c0 = e.load_config(8, 8, 5.7, 5, 11)
    e.warmup()
    rng = np.random.default_rng(9)
    mk = lambda kappa, mu, n: dict(config=c0, action="twistedmass", kappa=kappa, mu=mu,
                                   sources=[rng.standard_normal((e.eo_sites(c0), 24)) for _ in range(n)])
    reqs = [mk(0.15, 0.02, 3), mk(0.155, 0.008, 1), mk(0.152, 0.012, 2), mk(0.15, 0.003, 1)]
    alone = []
    for r in reqs:
        a = e.wait(e.submit(**r))
        assert a["ok"] and not a["joined_running"]
        alone.append(a)
    s0 = e.stats()
    ids = [e.submit(**r) for r in reqs]
    res = [e.wait(i) for i in ids]
    s1 = e.stats()
    assert all(r["ok"] for r in res), [r["error"] for r in res]
    assert all(max(r["true_rel_residual"]) <= 1e-10 for r in res)
    joined = [r for r in res[1:] if r["joined_running"]]
    assert joined, "no request joined the running solve"
    assert all(r["solve"] == res[0]["solve"] and r["admitted"] > res[0]["admitted"] for r in joined)
    assert s1["requests_joined"] - s0["requests_joined"] == len(joined)
    assert s1["solves_started"] - s0["solves_started"] == 1 + (len(reqs) - 1 - len(joined))
    for a, m in zip(alone, res):
        for xa, xm in zip(a["solutions"], m["solutions"]):
            assert rel(xm, xa) < 1e-6
    # the first request was answered before its solve had finished the joined ones
    assert res[0]["finished"] < max(r["finished"] for r in joined)
    lat = O.lattice(8, 8, theta_t=1.0)
    l = H.Lattice(["--nSpace=8", "--nTime=8", "--beta=5.7", "--hostSeed=11", "--thetaFermionTemporal=1"])
    gf = H.Gaugefield(l, False)
    for _ in range(5):
        H.su3heatbath(gf, l, 1)
    g = gf.get_host()
    k = res.index(joined[0])
    x, b = res[k]["solutions"][0], reqs[k]["sources"][0]
    kap, mu = reqs[k]["kappa"], reqs[k]["mu"]
    r = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, x, g, O.TWISTEDMASS, kap, 2 * kap * mu), b, 1.0)
    assert np.sqrt(O.norm2(r) / O.norm2(b)) <= 1.01e-10
