#!/usr/bin/env python3
"""Benchmark of the eo hopping kernel on one MI355X.

A step is one FP64 dslash_eo application (18-real links) over a 48^3x96
lattice, alternating EVEN/ODD as the reference's dslash benchmark does
(executables/dslashBenchmark.cpp:42-50).  Inputs are synthetic and resident in
HBM before timing: random SU(3) links from the device generator and a seeded
Gaussian spinor.  `value` is algorithmic GB/s: compulsory bytes per application
(8 links + input + output spinor per eo site) times steps over the timed
region, measured with HIP events on the context stream.

Also reported: the same for 12-real FP64, FP32 and FP16 (per-site scale) links, the reference's own
byte convention (2880 B/site), the CPU oracle timed on host cores
(cpu_baseline), and the HBM traffic per launch measured by rocprofv3 counters
(read from profiles/ when the counter summary for this workload is there).

The "cg" object is the solver half of the metric: on a heatbath beta 5.7
configuration generated on the device, the reference's FP64 eo-CG on
QplusQminus_eo (twisted mass, kappa 0.16, mu 0.005) against the reliable-update
CG with a reduced-precision inner loop, both to |r|/|b| <= 1e-10, with the true
residual recomputed in FP64 and the distance between the two solutions.

The other objects: "batch" (12 point sources at 32^3x64, one lockstep block solve
against 12 sequential solves), "multishift" (16-shift staggered CG-M at 48^3x12,
fused mixed solver against sequential solves and the reference's FP64 CG-M) and
"serving" (the inversion engine's capacity with a batch of requests queued at once,
its p50/p99 latency under a Poisson stream at 70 % of that capacity, and its
throughput against the reference's FP64 solvers one request at a time).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "eo-Dslash GB/s & %HBM peak; mixed vs FP64 CG time-to-1e-10 at 48³×96"
HBM_PEAK_GBPS = 8000.0
COUNTERS_FILE = os.path.join(ROOT, "profiles", "dslash_eo_f64_18_48x96_counters.txt")
USE_GRAPH = True  # replay the captured inner iteration of the mixed-precision solvers


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(ctx, fn, steps):
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timer_start()
    for k in range(steps):
        fn(k)
    ms = ctypes.c_double()
    ctx.check(ctx.lib.cl2_timer_stop(ctx.handle, ctypes.byref(ms)))
    ctx.sync()
    return ms.value, (time.perf_counter() - t0) * 1e3


def dslash_leg(ctx, abi, fmt, prec, steps, warmup, src, gauge18_ptr, autotune=True):
    """FP64 / FP32 / FP16 hopping applications on the heatbath configuration whose 18-real links sit at
    gauge18_ptr (the other link formats are converted from it on the device)"""
    g = None
    if fmt == abi.GAUGE_F64_18:
        gptr = gauge18_ptr
    else:
        g = ctx.buffer(ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_GAUGE, fmt))
        ctx.check(ctx.lib.cl2_gauge_convert(ctx.handle, g.ptr, fmt, ctypes.c_void_p(gauge18_ptr)))
        gptr = g.ptr
    a, b = ctx.spinor(prec).upload(src), ctx.spinor(prec)
    kappa = 0.16
    # launch shape from the autotuner (z-planes per slab, XCD mapping, block size), timed on this field;
    # without it the library's default shape
    # without the autotuner: the library default (-1 = the format's register budget)
    tuned = [ctypes.c_double(), ctypes.c_int(-1), ctypes.c_int(1), ctypes.c_int(1), ctypes.c_int(256), ctypes.c_int(-1)]
    if autotune:
        ctx.check(ctx.lib.cl2_autotune_dslash(ctx.handle, ctypes.c_void_p(gptr), fmt, 1,
                                              *[ctypes.byref(t) for t in tuned]))

    def step(k):
        if k % 2 == 0:
            ctx.dslash_eo(b.ptr, a.ptr, gptr, fmt, abi.EVEN, kappa)
        else:
            ctx.dslash_eo(a.ptr, b.ptr, gptr, fmt, abi.ODD, kappa)

    for k in range(warmup):
        step(k)
    ev_ms, wall_ms = time_steps(ctx, step, steps)
    out = b.download() if steps % 2 else a.download()
    finite = bool(np.isfinite(out).all())
    a.close()
    b.close()
    if g is not None:
        g.close()
    alg = ctx.lib.cl2_algorithmic_bytes(ctx.handle, b"dslash_eo", fmt)
    return dict(ms=ev_ms / steps, wall_ms=wall_ms / steps, alg_bytes=alg, GBps=alg * steps / (ev_ms * 1e6),
                finite=finite, tuned={"zplanes": tuned[1].value, "xcd": tuned[2].value, "block": tuned[4].value,
                                      "occupancy": tuned[5].value, "autotuned": autotune})


def wilson_lattice(ns, nt, seed, extra=()):
    from cl2qcd_amd import _host as H

    return H.Lattice([f"--nSpace={ns}", f"--nTime={nt}", "--beta=5.7", "--fermionAction=twistedmass",
                      "--kappa=0.16", "--mu=0.005", "--thetaFermionTemporal=1", "--solverMaxIterations=200000",
                      f"--hostSeed={seed}", *extra])


def make_config(ns, nt, sweeps, overrelax, seed):
    """The heatbath beta 5.7 configuration of the benchmark lattice, made on the device by the drop-in's
    su3heatbath from a cold start (seeded): the links of the dslash legs and of the solver legs."""
    from cl2qcd_amd import _host as H

    l = wilson_lattice(ns, nt, seed)
    gf = H.Gaugefield(l, False)
    t0 = time.perf_counter()
    for _ in range(sweeps):
        H.su3heatbath(gf, l, overrelax)
    plaq = H.measurePlaquette(gf)
    l.synchronize()
    hb_s = time.perf_counter() - t0
    log(f"heatbath: {sweeps} x (1 + {overrelax} OR) sweeps in {hb_s:.1f} s, plaquette {plaq:.6f}")
    desc = f"beta 5.7, cold start, {sweeps} sweeps x (heatbath + {overrelax} overrelaxation), seed {seed}"
    return l, gf, plaq, desc


def cg_leg(ns, nt, l, gf, plaq, desc, seed):
    """FP64 eo-CG (reference algorithm) vs FP32-inner reliable-update CG on QplusQminus_eo, both to
    |r|/|b| <= 1e-10, through the C++ drop-in.  Twisted mass kappa 0.16, mu 0.005, antiperiodic in t, on a
    heatbath beta 5.7 configuration made on the device from a cold start (seeded), Gaussian source.  The
    plain FP64 CG runs with the reference's 18-real links and with 12-real links (useReconstruct12)."""
    from cl2qcd_amd import _host as H

    b = H.Spinorfield_eo(l)
    b.import_(np.random.default_rng(seed).standard_normal((l.eo_sites, 24)))
    b2 = H.squarenorm(b)
    qq, ap = H.QplusQminus_eo(l), l.wilson_parameters()
    r = H.Spinorfield_eo(l)
    # launch shapes of the hopping kernel for the three link formats (FP64 operator of the plain CG and of the
    # reliable updates, FP32 and FP16 inner loops), autotuned on this configuration before anything is timed
    tuned_ms = {name: round(H.autotune_dslash(gf, l, prec, 1), 4)
                for name, prec in (("f64", H.F64), ("f32", H.F32), ("f16", H.F16))}
    log(f"autotuned hopping kernel (ms per application): {tuned_ms}")

    def true_rel_residual(x):
        qq(r, gf, x, ap)
        H.saxpy(r, 1.0, r, b)  # r = b - A x
        return float(np.sqrt(H.squarenorm(r) / b2))

    out = {"lattice": f"{ns}^3x{nt}", "heatbath": desc, "plaquette": round(plaq, 6), "operator": "QplusQminus_eo twisted mass kappa 0.16 mu 0.005",
           "target_rel_residual": 1e-10, "autotuned_dslash_ms": tuned_ms}
    x64 = H.Spinorfield_eo(l)
    x64.zero()
    l.synchronize()
    t0 = time.perf_counter()
    it = H.cg(x64, qq, gf, b, l, 1e-20 * b2, ap)
    l.synchronize()
    dt = time.perf_counter() - t0
    out["fp64"] = dict(iterations=it, seconds=round(dt, 3), ms_per_iteration=round(1e3 * dt / max(it, 1), 3),
                       true_rel_residual=true_rel_residual(x64))
    log(f"FP64 CG: {it} iterations, {dt:.2f} s, true |r|/|b| {out['fp64']['true_rel_residual']:.3e}")
    x64n = H.squarenorm(x64)
    d = H.Spinorfield_eo(l)
    # the same plain FP64 CG with 12-real links (row 2 rebuilt in the kernel): a second FP64 baseline
    l12 = wilson_lattice(ns, nt, seed, ("--useReconstruct12=true",))
    gf12 = H.Gaugefield(l12, False)
    gf12.convert_from(gf)
    H.autotune_dslash(gf12, l12, H.F64, 1)
    b12, x12 = H.Spinorfield_eo(l12), H.Spinorfield_eo(l12)
    b12.import_(b.get_host())
    x12.zero()
    l12.synchronize()
    t0 = time.perf_counter()
    it12 = H.cg(x12, H.QplusQminus_eo(l12), gf12, b12, l12, 1e-20 * b2, l12.wilson_parameters())
    l12.synchronize()
    dt12 = time.perf_counter() - t0
    x12h = x12.get_host()
    del x12, b12, gf12, l12
    x12d = H.Spinorfield_eo(l)
    x12d.import_(x12h)
    del x12h
    H.saxpy(d, 1.0, x64, x12d)
    out["fp64_12real_links"] = dict(iterations=it12, seconds=round(dt12, 3),
                                    ms_per_iteration=round(1e3 * dt12 / max(it12, 1), 3),
                                    true_rel_residual=true_rel_residual(x12d),
                                    rel_diff_to_fp64_solution=float(np.sqrt(H.squarenorm(d) / x64n)))
    del x12d
    log(f"FP64 CG, 12-real links: {it12} iterations, {dt12:.2f} s")
    for name, prec in (("fp32", H.F32), ("fp16", H.F16)):
        x = H.Spinorfield_eo(l)
        x.zero()
        l.synchronize()
        t0 = time.perf_counter()
        st = H.cg_reliable_update(x, qq, gf, b, l, 1e-10, ap, prec, 0.1, use_graph=USE_GRAPH)
        l.synchronize()
        dt = time.perf_counter() - t0
        H.saxpy(d, 1.0, x64, x)
        out[name] = dict(iterations=st["iterations"], reliable_updates=st["reliable_updates"],
                         seconds=round(dt, 3), ms_per_iteration=round(1e3 * dt / max(st["iterations"], 1), 3),
                         true_rel_residual=true_rel_residual(x),
                         rel_diff_to_fp64_solution=float(np.sqrt(H.squarenorm(d) / x64n)),
                         speedup_vs_fp64=round(out["fp64"]["seconds"] / dt, 3),
                         speedup_vs_fp64_12real_links=round(out["fp64_12real_links"]["seconds"] / dt, 3))
        log(f"{name}-inner CG: {st['iterations']} iterations ({st['reliable_updates']} reliable updates), "
            f"{dt:.2f} s, true |r|/|b| {out[name]['true_rel_residual']:.3e}, "
            f"speedup {out[name]['speedup_vs_fp64']:.2f}x")
    return out


def batch_leg(ns, nt, sweeps, seed, nrhs=12):
    """12 point sources (every spin x colour at the origin) on QplusQminus_eo, twisted mass kappa 0.16,
    mu 0.005, heatbath beta 5.7: one lockstep block solve with the fused multi-RHS operator against 12
    single-source solves, both with FP64 reliable updates to |r|/|b| <= 1e-10, for an FP32 and for an FP16
    (per-site norm) inner loop.  Also the per-RHS cost of one hopping application with 12 right-hand sides
    against one with a single field, FP32 and FP16."""
    from cl2qcd_amd import _host as H, abi

    l = H.Lattice([f"--nSpace={ns}", f"--nTime={nt}", "--beta=5.7", "--fermionAction=twistedmass",
                   "--kappa=0.16", "--mu=0.005", "--thetaFermionTemporal=1", "--solverMaxIterations=200000",
                   f"--hostSeed={seed}"])
    gf = H.Gaugefield(l, False)
    for _ in range(sweeps):
        H.su3heatbath(gf, l, 1)
    qq, ap = H.QplusQminus_eo(l), l.wilson_parameters()
    # launch shapes of the FP32 hopping kernel for one and for nrhs right-hand sides, and of the single-field
    # FP16 kernel, autotuned on this lattice
    H.autotune_dslash(gf, l, H.F32, 1)
    H.autotune_dslash(gf, l, H.F32, nrhs)
    H.autotune_dslash(gf, l, H.F16, 1)
    veo = l.eo_sites
    bs, xs = [], []
    for r in range(nrhs):
        src = np.zeros((veo, 24))
        src[0, 2 * r] = 1.0  # eo site 0 is the origin (even); component r = 3 * spin + colour
        b = H.Spinorfield_eo(l)
        b.import_(src)
        bs.append(b)
        x = H.Spinorfield_eo(l)
        x.zero()
        xs.append(x)

    def solves(prec):
        for x in xs:
            x.zero()
        l.synchronize()
        t0 = time.perf_counter()
        seq_it = []
        for r in range(nrhs):
            st = H.cg_reliable_update(xs[r], qq, gf, bs[r], l, 1e-10, ap, prec, 0.1, use_graph=USE_GRAPH)
            seq_it.append(st["iterations"])
        l.synchronize()
        t_seq = time.perf_counter() - t0
        for x in xs:
            x.zero()
        l.synchronize()
        t0 = time.perf_counter()
        it, per = H.cg_reliable_update_block(xs, qq, gf, bs, l, 1e-10, ap, prec, 0.1, use_graph=USE_GRAPH)
        l.synchronize()
        t_blk = time.perf_counter() - t0
        worst = max(p["true_relative_residual"] for p in per)
        name = "FP16" if prec == H.F16 else "FP32"
        log(f"12 point sources {ns}^3x{nt}, {name} inner: sequential {t_seq:.2f} s ({sum(seq_it)} iterations), "
            f"block {t_blk:.2f} s ({it} lockstep iterations, {per[0]['reliable_updates']} reliable updates), "
            f"worst true |r|/|b| {worst:.2e}")
        return t_seq, t_blk, {"sequential": {"seconds": round(t_seq, 3), "per_rhs_seconds": round(t_seq / nrhs, 4),
                                             "iterations": seq_it},
                              "block": {"seconds": round(t_blk, 3), "per_rhs_seconds": round(t_blk / nrhs, 4),
                                        "lockstep_iterations": it, "iterations": [p["iterations"] for p in per],
                                        "reliable_updates": per[0]["reliable_updates"],
                                        "worst_true_rel_residual": worst},
                              "speedup_per_rhs": round(t_seq / t_blk, 3)}

    t_seq32, t_blk32, f32 = solves(H.F32)
    t_seq16, t_blk16, f16 = solves(H.F16)
    f16["block_speedup_vs_fp32_sequential"] = round(t_seq32 / t_blk16, 3)
    # hopping term alone: nrhs right-hand sides in one launch vs one field, FP32 and FP16 (random links)
    ctx = abi.Context(ns, nt, theta_t=1.0)

    def rate(fn, reps=50):
        fn()
        ctx.sync()
        ctx.timer_start()
        for _ in range(reps):
            fn()
        ms = ctypes.c_double()
        ctx.check(ctx.lib.cl2_timer_stop(ctx.handle, ctypes.byref(ms)))
        return ms.value / reps

    def hop_rates(fmt, prec, link_b, site_b):
        g = ctx.gauge(fmt=fmt, seed=3, epsilon=0.35)
        one_b = ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, prec)
        a1, b1 = ctx.buffer(one_b), ctx.buffer(one_b)
        an, bn = ctx.buffer(nrhs * one_b), ctx.buffer(nrhs * one_b)
        ctx.set_cold_eo(a1.ptr, prec)
        ctx.block_set_zero_eo(an.ptr, prec, nrhs)
        t1 = rate(lambda: ctx.dslash_eo(b1.ptr, a1.ptr, g.ptr, fmt, abi.EVEN, 0.16))
        tn = rate(lambda: ctx.dslash_eo_mrhs(bn.ptr, an.ptr, g.ptr, fmt, abi.EVEN, 0.16, nrhs))
        alg_n = veo * (8 * link_b + 2 * nrhs * site_b)
        for buf in (a1, b1, an, bn, g):
            buf.close()
        name = "FP16" if prec == abi.F16 else "FP32"
        log(f"{name} hopping term: 1 field {t1:.4f} ms, {nrhs} fields {tn:.4f} ms ({tn / nrhs:.4f} ms per field, "
            f"{alg_n / (tn * 1e6):.0f} GB/s algorithmic)")
        return {"single": round(t1, 5), f"mrhs{nrhs}": round(tn, 5), f"mrhs{nrhs}_per_rhs": round(tn / nrhs, 5),
                f"mrhs{nrhs}_alg_GBps": round(alg_n / (tn * 1e6), 1)}

    d32 = hop_rates(abi.GAUGE_F32_12, abi.F32, 48, 96)
    d16 = hop_rates(abi.GAUGE_F16_12, abi.F16, 24, 52)
    ctx.close()
    return {"lattice": f"{ns}^3x{nt}", "sources": f"{nrhs} point sources (spin x colour) at the origin",
            "operator": "QplusQminus_eo twisted mass kappa 0.16 mu 0.005, FP32 inner, FP64 reliable updates",
            **f32, "dslash_f32_ms": d32,
            "fp16_inner": {"operator": "the same with an FP16 (per-site norm) inner loop", **f16},
            "dslash_f16_ms": d16}


def mrhs_shapes_leg(ns=32, nt=64, nrhs=12, reps=30):
    """The FP32 12-RHS hopping kernel alone at 32^3x64 in every work order (z-planes per slab x XCD mapping):
    time per launch from HIP events, after one copy of an FP64 field for counter calibration.  Run under
    rocprofv3 --pmc, the dispatches come in the printed order (one warmup and `reps` timed per shape)."""
    from cl2qcd_amd import abi

    ctx = abi.Context(ns, nt, theta_t=1.0)
    g = ctx.gauge(fmt=abi.GAUGE_F32_12, seed=3, epsilon=0.35)
    one = ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F32)
    an, bn = ctx.buffer(nrhs * one), ctx.buffer(nrhs * one)
    ctx.block_set_zero_eo(an.ptr, abi.F32, nrhs)
    f64, dst = ctx.spinor(), ctx.spinor()
    nb = ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F64)
    ctx.check(ctx.lib.cl2_copy(ctx.handle, dst.ptr, f64.ptr, nb))
    out = []
    for zp in (1, 2, 4, 8, 16, 32, 0):
        for xcd in (1, 0):
            ctx.check(ctx.lib.cl2_tune_set(ctx.handle, b"hop.zplanes", zp))
            ctx.check(ctx.lib.cl2_tune_set(ctx.handle, b"hop.xcd", xcd))

            def f():
                ctx.dslash_eo_mrhs(bn.ptr, an.ptr, g.ptr, abi.GAUGE_F32_12, abi.EVEN, 0.16, nrhs)

            f()
            ms, _ = time_steps(ctx, lambda k: f(), reps)
            out.append({"zplanes": zp, "xcd": xcd, "ms": round(ms / reps, 5)})
            log(f"12-RHS FP32 hopping kernel, zplanes {zp:2d} xcd {xcd}: {ms / reps:.4f} ms")
    ctx.close()
    return out


def multishift_leg(ns, nt, sweeps, seed, nshift=16):
    """Staggered MdagM_eo (mass 0.01) with 16 shifts sigma_k = 10^(-4 + 4k/15) on a heatbath beta 5.9
    configuration, Gaussian source, every shift to |b - (A + sigma_k) x_k| <= 1e-10 |b| in FP64:
    the fused multi-shift solver (FP32 CG-M, then FP64 reliable-update refinement per shift) against
    16 independent FP32-inner reliable-update solves and against the reference's FP64 CG-M."""
    from cl2qcd_amd import _host as H

    l = H.Lattice([f"--nSpace={ns}", f"--nTime={nt}", "--beta=5.9", "--fermionAction=rooted_stagg",
                   "--mass=0.01", "--solverMaxIterations=200000", f"--hostSeed={seed}"])
    gf = H.Gaugefield(l, False)
    for _ in range(sweeps):
        H.su3heatbath(gf, l, 1)
    plaq = H.measurePlaquette(gf)
    A, ap = H.MdagM_eo(l), l.staggered_parameters()
    # launch shapes of the staggered hopping kernel for the FP32 inner loop and the FP64 operator, autotuned here
    H.autotune_dks(gf, l, H.F32)
    H.autotune_dks(gf, l, H.F64)
    sigma = [10.0 ** (-4 + 4 * k / (nshift - 1)) for k in range(nshift)]
    b = H.Staggeredfield_eo(l)
    b.import_(np.random.default_rng(seed).standard_normal((l.eo_sites, 6)))
    b2 = H.squarenorm(b)
    xs = [H.Staggeredfield_eo(l) for _ in sigma]
    l.synchronize()
    t0 = time.perf_counter()
    total, ms_it, per = H.cg_m_mixed(xs, A, gf, sigma, b, l, 1e-10, ap, H.F32, 1e-7, use_graph=USE_GRAPH)
    l.synchronize()
    t_ms = time.perf_counter() - t0
    worst_ms = max(p["true_relative_residual"] for p in per)
    seq = [H.Staggeredfield_eo(l) for _ in sigma]
    l.synchronize()
    t0 = time.perf_counter()
    seq_it = []
    for s_, x in zip(sigma, seq):
        x.set_zero()
        seq_it.append(H.cg_reliable_update_stagg(x, A, gf, s_, b, l, 1e-10, ap, H.F32, use_graph=USE_GRAPH)["iterations"])
    l.synchronize()
    t_seq = time.perf_counter() - t0
    ref = [H.Staggeredfield_eo(l) for _ in sigma]
    l.synchronize()
    t0 = time.perf_counter()
    ref_it = H.cg_m(ref, A, gf, sigma, b, l, 1e-20 * b2, ap)
    l.synchronize()
    t_ref = time.perf_counter() - t0
    log(f"16-shift CG-M {ns}^3x{nt} (plaquette {plaq:.5f}): fused mixed {t_ms:.2f} s ({ms_it} multi-shift + "
        f"{total - ms_it} refinement iterations, worst |r|/|b| {worst_ms:.2e}); 16 sequential mixed solves "
        f"{t_seq:.2f} s ({sum(seq_it)} iterations); reference FP64 CG-M {t_ref:.2f} s ({ref_it} iterations)")
    return {"lattice": f"{ns}^3x{nt}", "plaquette": round(plaq, 6),
            "heatbath": f"beta 5.9, cold start, {sweeps} sweeps x (heatbath + 1 overrelaxation), seed {seed}",
            "operator": "staggered MdagM_eo mass 0.01", "shifts": "16, log-spaced in [1e-4, 1]",
            "fused_mixed": {"seconds": round(t_ms, 3), "per_shift_seconds": round(t_ms / nshift, 4),
                            "multishift_iterations": ms_it, "refinement_iterations": total - ms_it,
                            "worst_true_rel_residual": worst_ms},
            "sequential_mixed": {"seconds": round(t_seq, 3), "per_shift_seconds": round(t_seq / nshift, 4),
                                 "iterations": sum(seq_it)},
            "reference_fp64_cgm": {"seconds": round(t_ref, 3), "iterations": ref_it,
                                   "stop": "|zeta_k r|^2 < 1e-20 |b|^2 for every shift"},
            "speedup_vs_sequential": round(t_seq / t_ms, 3), "speedup_vs_reference_cgm": round(t_ref / t_ms, 3)}


SERVE_SIZES = ((16, 32), (24, 48), (32, 64))
SERVE_CONFIGS = {(16, 32): 3, (24, 48): 3, (32, 64): 2}  # 8 resident configurations


def serve_requests(n, seed, cfgs):
    """n inversion requests: lattice size uniform over 16^3x32, 24^3x48, 32^3x64, one of that size's resident
    configurations, twisted mass or staggered with equal probability, kappa uniform in [0.15, 0.16], mu in
    [0.003, 0.02], mass in [0.005, 0.05], 1-12 Gaussian right-hand sides generated on the device,
    |r|/|b| <= 1e-10.  Queued twisted-mass requests on one configuration are merged into one multi-RHS solve
    whatever their kappa and mu."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        size = SERVE_SIZES[rng.integers(3)]
        cfg = cfgs[size][rng.integers(len(cfgs[size]))]
        k = int(rng.integers(1, 13))
        req = dict(config=cfg, gaussian_seeds=[int(v) for v in rng.integers(0, 2 ** 62, size=k)], rel_tol=1e-10,
                   return_solutions=False)
        if rng.random() < 0.5:
            req.update(action="twistedmass", kappa=float(rng.uniform(0.15, 0.16)), mu=float(rng.uniform(0.003, 0.02)))
        else:
            req.update(action="staggered", mass=float(rng.uniform(0.005, 0.05)))
        out.append(req)
    return out


def serve_run(engine, reqs, times=None):
    """Submit reqs at the given arrival times (seconds from now; None: all at once) and wait for all."""
    engine.reset_peaks()
    s0 = engine.stats()
    start = engine.now()
    ids = []
    for k, req in enumerate(reqs):
        if times is not None:
            delay = start + times[k] - engine.now()
            if delay > 0:
                time.sleep(delay)
        ids.append(engine.submit(**req))
    res = [engine.wait(i) for i in ids]
    end = max(r["finished"] for r in res)
    lat = np.array([r["finished"] - r["arrival"] for r in res])
    ok = all(r["ok"] for r in res)
    worst = max(max(r["true_rel_residual"]) for r in res) if ok else None
    rhs = sum(len(r["true_rel_residual"]) for r in res)
    s1 = engine.stats()
    per_solve = {}
    for r in res:
        per_solve[r["solve"]] = per_solve.get(r["solve"], 0) + 1
    return {"requests": len(res), "right_hand_sides": rhs, "all_ok": ok, "worst_true_rel_residual": worst,
            "solves": len(per_solve), "solves_carrying_several_requests": sum(v > 1 for v in per_solve.values()),
            "requests_merged_at_admission": s1["requests_merged"] - s0["requests_merged"],
            "requests_joined_running_solves": s1["requests_joined"] - s0["requests_joined"],
            "max_queue_depth": s1["max_queue_depth"],
            "throughput_req_per_s": round(len(res) / (end - start), 3),
            "throughput_rhs_per_s": round(rhs / (end - start), 3),
            "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
            "makespan_s": round(end - start, 3)}


def serving_leg(n, load, seed, n_baseline, rate=None, inner="f16", n_capacity=60, n_burst=100, burst_rate=20.0):
    """Continuous-batching engine (4 workers / HIP streams, mixed solvers with an FP16 (or FP32) inner loop for
    twisted mass and FP32 for staggered, HBM budget 160 GB):
    its capacity with the first n_capacity requests queued at once; latency under a Poisson stream of n
    requests offered at `load` times that capacity (or at `rate`); a bounded run of n_burst requests arriving
    as a Poisson stream at the workload's own burst_rate (20/s) with the queue depth, admissions and joins it
    causes; and the plain FP64 solvers one request at a time on the first n_baseline requests against the
    engine on those same requests."""
    from cl2qcd_amd import _host as H

    def engine(**kw):
        e = H.Engine(hbm_budget_bytes=160 << 30, **kw)
        cfgs, k = {}, 0
        for size in SERVE_SIZES:
            cfgs[size] = []
            for _ in range(SERVE_CONFIGS[size]):
                cfgs[size].append(e.load_config(size[0], size[1], 5.7, 20, seed + k))
                k += 1
        e.warmup()
        return e, cfgs

    e, cfgs = engine(workers=4, mixed=True, inner_precision=H.F16 if inner == "f16" else H.F32)
    reqs = serve_requests(max(n, n_capacity, n_burst, n_baseline), seed, cfgs)
    log(f"serving: {n_capacity} requests queued at once ...")
    closed = serve_run(e, reqs[:n_capacity])
    log(f"  capacity: {closed}")
    offered = rate if rate else load * closed["throughput_req_per_s"]
    times = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / offered, size=n))
    log(f"serving: {n} requests as a Poisson stream at {offered:.3f}/s ...")
    stream = serve_run(e, reqs[:n], times)
    log(f"  stream: {stream}")
    times = np.cumsum(np.random.default_rng(seed + 2).exponential(1.0 / burst_rate, size=n_burst))
    log(f"serving: {n_burst} requests as a Poisson stream at {burst_rate:.1f}/s ...")
    burst = serve_run(e, reqs[:n_burst], times)
    log(f"  burst: {burst}")
    first = serve_run(e, reqs[:n_baseline])
    del e
    f, _ = engine(workers=1, mixed=False)
    base = serve_run(f, reqs[:n_baseline])
    del f
    log(f"  first {n_baseline}, queued at once: engine {first}; FP64 one at a time {base}")
    return {"requests": f"seed {seed}; sizes 16^3x32 / 24^3x48 / 32^3x64, twisted mass or staggered, 1-12 "
                        "Gaussian RHS, |r|/|b| <= 1e-10; 8 resident heatbath beta 5.7 configurations",
            "inner_precision": {"twisted_mass": inner, "staggered": "f32"},
            "capacity_all_queued": closed,
            "poisson_stream": {"offered_req_per_s": round(offered, 3),
                               "load": round(offered / closed["throughput_req_per_s"], 3), **stream},
            "burst_at_workload_rate": {"offered_req_per_s": burst_rate, **burst},
            "first_requests_all_queued": {"n": n_baseline, "engine": first, "fp64_one_at_a_time": base,
                                          "throughput_ratio": round(first["throughput_req_per_s"] /
                                                                    base["throughput_req_per_s"], 3)}}


def read_counters():
    """HBM bytes per launch of the FP64 dslash from the committed rocprofv3 counter summary."""
    if not os.path.exists(COUNTERS_FILE):
        return None
    vals = {}
    for line in open(COUNTERS_FILE):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def host_cpu():
    """core count and model of the host this runs on"""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def config1_leg(seed=20261201, sweeps=50):
    """The CPU-runnable workload exactly as written: 8^4, beta 5.7 heatbath (50 sweeps from a cold start,
    the drop-in's su3heatbath, seeded), twisted mass kappa 0.15, mu 0.01, antiperiodic in t, FP64 eo CG on
    QplusQminus_eo to |r|^2/|b|^2 <= 1e-20 from a seeded unit-variance Gaussian source.  Timed three ways on
    the same inputs: the drop-in on the GPU, and the oracle (the reference's CG restated in C++ with
    OpenMP, the reference's CPU execution mode) on all host cores and on one core."""
    from cl2qcd_amd import _host as H
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    l = H.Lattice(["--nSpace=8", "--nTime=8", "--beta=5.7", f"--hostSeed={seed}", "--fermionAction=twistedmass",
                   "--kappa=0.15", "--mu=0.01", "--thetaFermionTemporal=1", "--solverMaxIterations=5000"])
    gf = H.Gaugefield(l, False)
    for _ in range(sweeps):
        H.su3heatbath(gf, l, 0)
    g = np.ascontiguousarray(gf.get_host().reshape(-1, 18))
    bh = np.random.default_rng(seed).standard_normal((l.eo_sites, 24))
    b, x = H.Spinorfield_eo(l), H.Spinorfield_eo(l)
    b.import_(bh)
    x.cold()
    x0 = x.get_host()
    prec = 1e-20 * float((bh ** 2).sum())
    qq, ap = H.QplusQminus_eo(l), l.wilson_parameters()
    H.cg(x, qq, gf, b, l, prec, ap)  # warm: kernels loaded, launch shapes chosen
    x.import_(x0)
    l.synchronize()
    t0 = time.perf_counter()
    it = H.cg(x, qq, gf, b, l, prec, ap)
    l.synchronize()
    dt = time.perf_counter() - t0
    lat = O.lattice(8, 8, theta_t=1.0)
    mubar = l.mubar

    def true_rel(sol):
        r = O.saxpy(O.apply_eo(lat, O.OP_QPLUSQMINUS_EO, sol, g, O.TWISTEDMASS, 0.15, mubar), bh, 1.0)
        return float(np.sqrt(O.norm2(r) / O.norm2(bh)))

    out = {"workload": "8^4, beta 5.7 heatbath (50 sweeps, cold start, seeded), twisted mass kappa 0.15 mu 0.01, "
                       "antiperiodic t, FP64 eo CG on QplusQminus_eo to |r|^2/|b|^2 <= 1e-20, Gaussian source",
           "plaquette": round(float(H.measurePlaquette(gf)), 6),
           "gpu_dropin": {"iterations": it, "seconds": round(dt, 4), "ms_per_iteration": round(1e3 * dt / it, 4),
                          "true_rel_residual": true_rel(x.get_host())}}
    all_cores = O.threads()
    for key, n in (("oracle_all_cores", all_cores), ("oracle_one_core", 1)):
        O.set_threads(n)
        O.cg_eo(lat, O.OP_QPLUSQMINUS_EO, g, bh, x0, O.TWISTEDMASS, 0.15, mubar, prec=prec, cgmax=5000)  # warm
        t0 = time.perf_counter()
        rc, it_o, xo = O.cg_eo(lat, O.OP_QPLUSQMINUS_EO, g, bh, x0, O.TWISTEDMASS, 0.15, mubar, prec=prec,
                               cgmax=5000)
        dto = time.perf_counter() - t0
        out[key] = {"threads": n, "iterations": it_o, "converged": rc == O.OK, "seconds": round(dto, 4),
                    "ms_per_iteration": round(1e3 * dto / max(it_o, 1), 4), "true_rel_residual": true_rel(xo)}
    O.set_threads(all_cores)
    out["gpu_speedup_vs_oracle_all_cores"] = round(out["oracle_all_cores"]["seconds"] / dt, 2)
    log(f"config 1 (8^4): GPU drop-in {it} iterations {dt * 1e3:.1f} ms; oracle {out['oracle_all_cores']['seconds']:.3f} s "
        f"on {all_cores} threads, {out['oracle_one_core']['seconds']:.3f} s on 1")
    return out


def cpu_baseline(seconds=12.0):
    """The oracle's dslash_eo on host cores over a bounded sample (24^3x48, FP64)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    lat = O.lattice(24, 48, theta_t=1.0)
    rng = np.random.default_rng(7)
    g = O.gauge_cold(lat)
    g += 0.01 * rng.standard_normal(g.shape)  # values do not affect the timing
    src = rng.standard_normal((O.veo(lat), 24))
    O.dslash_eo(lat, src, g, O.EVEN, 0.16)  # warm caches and the thread pool
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        O.dslash_eo(lat, src, g, O.EVEN if n % 2 == 0 else O.ODD, 0.16)
        n += 1
    dt = time.perf_counter() - t0
    per_app = 1536 * O.veo(lat)
    return dict(value=round(per_app * n / dt / 1e9, 3), unit="GB/s", cores=O.threads(), kind="port",
                sample=f"oracle dslash_eo FP64 18-real, 24^3x48, {n} applications in {dt:.1f} s, "
                       f"{O.threads()} OpenMP threads")


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--ns", type=int, default=48)
    ap.add_argument("--nt", type=int, default=96)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only-f64", action="store_true", help="time only the FP64 18-real leg")
    ap.add_argument("--no-cg", action="store_true", help="skip the FP64 vs mixed-precision CG comparison")
    ap.add_argument("--hb-sweeps", type=int, default=50, help="heatbath sweeps for the CG configuration")
    ap.add_argument("--no-batch", action="store_true", help="skip the 12-source block solve at 32^3x64")
    ap.add_argument("--mrhs-shapes", action="store_true",
                    help="only time the 12-RHS FP32 hopping kernel at 32^3x64 in every work order (--steps launches "
                         "each), for profiles")
    ap.add_argument("--no-config1", action="store_true",
                    help="skip the 8^4 CPU-reference workload (GPU drop-in and oracle CG)")
    ap.add_argument("--no-multishift", action="store_true", help="skip the 16-shift staggered solve at 48^3x12")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving run")
    ap.add_argument("--serve-requests", type=int, default=300, help="requests of the serving Poisson stream")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="Poisson arrival rate as a fraction of the engine's measured capacity")
    ap.add_argument("--serve-rate", type=float, default=None, help="fixed arrival rate (requests per second)")
    ap.add_argument("--serve-baseline", type=int, default=10, help="requests also run by the FP64 baseline")
    ap.add_argument("--serve-inner", choices=("f32", "f16"), default="f16",
                    help="inner precision of the engine's twisted-mass solves (single and lockstep blocks; "
                         "staggered solves run FP32 inner)")
    ap.add_argument("--hb-overrelax", type=int, default=1, help="overrelaxation sweeps per heatbath sweep")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the hopping kernels in the library's default launch shape (for profiles whose "
                         "per-kernel statistics should hold the timed launches only)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the solvers' inner iterations kernel by kernel instead of replaying a hipGraph")
    args = ap.parse_args()
    if args.mrhs_shapes:
        print(json.dumps({"mrhs_shapes_f32_12rhs_32x64": mrhs_shapes_leg(reps=args.steps)}), flush=True)
        return
    global USE_GRAPH
    USE_GRAPH = not args.no_graph
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU path is "
              "designed but not built)", file=sys.stderr)
        sys.exit(2)

    from cl2qcd_amd import abi

    # the benchmark configuration: its 18-real links are the dslash legs' input and the solver legs' gauge field
    cfg_seed = 20261202
    l_cfg, gf_cfg, plaq_cfg, desc_cfg = make_config(args.ns, args.nt, args.hb_sweeps, args.hb_overrelax, cfg_seed)
    ctx = abi.Context(args.ns, args.nt, theta_t=1.0)
    src = np.random.default_rng(1).standard_normal((ctx.veo, 24))
    # counter calibration for rocprofv3 runs of this script: one pure read and one
    # pure write of exactly 12 * Veo * 16 bytes (k_reduce1 and k_fill)
    cal = ctx.spinor()
    res = ctx.scalars(1)
    ctx.squarenorm_eo(res.ptr, cal.ptr, abi.F64)
    ctx.set_zero_eo(cal.ptr, abi.F64)
    ctx.sync()
    # achievable streaming rate on this GPU: the library's copy kernel on one FP64 spinor field
    dst = ctx.spinor()
    nbytes = ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F64)
    for _ in range(2):
        ctx.check(ctx.lib.cl2_copy(ctx.handle, dst.ptr, cal.ptr, nbytes))
    copy_ms, _ = time_steps(ctx, lambda k: ctx.check(ctx.lib.cl2_copy(ctx.handle, dst.ptr, cal.ptr, nbytes)), 20)
    copy_GBps = 2 * nbytes * 20 / (copy_ms * 1e6)
    dst.close()
    cal.close()
    res.close()
    log(f"copy kernel: {copy_GBps:.0f} GB/s (read + write of one FP64 field)")
    log(f"lattice {args.ns}^3x{args.nt}, {ctx.veo} eo sites; warmup {args.warmup}, steps {args.steps}")
    f64 = dslash_leg(ctx, abi, abi.GAUGE_F64_18, abi.F64, args.steps, args.warmup, src, gf_cfg.device_ptr,
                     not args.no_autotune)
    log(f"dslash F64/18: {f64['ms']:.4f} ms/app, {f64['GBps']:.0f} GB/s")
    extra = {}
    if not args.only_f64:
        for name, fmt, prec in (("f64_12", abi.GAUGE_F64_12, abi.F64), ("f32_12", abi.GAUGE_F32_12, abi.F32),
                                ("f16_12", abi.GAUGE_F16_12, abi.F16)):
            r = dslash_leg(ctx, abi, fmt, prec, max(args.steps // 4, 2), args.warmup, src, gf_cfg.device_ptr,
                           not args.no_autotune)
            log(f"dslash {name}: {r['ms']:.4f} ms/app, {r['GBps']:.0f} GB/s")
            extra[f"dslash_{name}"] = dict(ms_per_app=round(r["ms"], 5), GBps=round(r["GBps"], 1),
                                           pct_hbm_peak=round(100 * r["GBps"] / HBM_PEAK_GBPS, 1),
                                           launch_shape=r["tuned"])
    ref_bytes = ctx.lib.cl2_reference_read_write_size(ctx.handle, b"dslash_eo")
    veo = ctx.veo
    ctx.close()

    cg = None
    if not args.no_cg:
        cg = cg_leg(args.ns, args.nt, l_cfg, gf_cfg, plaq_cfg, desc_cfg, cfg_seed)
    del gf_cfg, l_cfg

    config1 = None
    if not args.no_config1:
        config1 = config1_leg()

    batch = None
    if not args.no_batch:
        batch = batch_leg(32, 64, args.hb_sweeps, 20261203)

    multishift = None
    if not args.no_multishift:
        multishift = multishift_leg(48, 12, args.hb_sweeps, 20261204)

    serving = None
    if not args.no_serving:
        serving = serving_leg(args.serve_requests, args.serve_load, 20261205, args.serve_baseline, args.serve_rate,
                              args.serve_inner)

    traffic = read_counters()
    achieved = f64["alg_bytes"] / (f64["ms"] * 1e6)
    line = {
        "metric": METRIC,
        "value": round(f64["GBps"], 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f64["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic: heatbath configuration made on the device ({desc_cfg}, plaquette {plaq_cfg:.6f}), "
                "seeded Gaussian spinor",
        "config": {"workload": f"dslash_eo FP64 18-real links, {args.ns}^3x{args.nt}, kappa 0.16, antiperiodic t",
                   "lattice": f"{args.ns}^3x{args.nt}", "eo_sites": veo,
                   "bytes_per_site": 1536},
        "pct_hbm_peak": round(100 * f64["GBps"] / HBM_PEAK_GBPS, 1),
        "copy_GBps": round(copy_GBps, 1),
        "frac_of_copy": round(f64["GBps"] / copy_GBps, 4),
        "wall_ms_per_step": round(f64["wall_ms"], 5),
        "cl2qcd_convention_GBps": round(ref_bytes / (f64["ms"] * 1e6), 1),
        "output_finite": f64["finite"],
        "launch_shape": f64["tuned"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic},
        **extra,
        "cg": cg,
        "config1": config1,
        "batch": batch,
        "multishift": multishift,
        "serving": serving,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on host cores) ...")
        line["cpu_baseline"] = cpu_baseline()
        line["cpu_baseline"]["host"] = host_cpu()
        if config1 is not None:
            line["cpu_baseline"]["cg_config1"] = {k: config1[k] for k in ("oracle_all_cores", "oracle_one_core")}
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
