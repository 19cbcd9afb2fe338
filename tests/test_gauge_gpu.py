"""Gauge-field generation on the GPU: plaquette, overrelaxation and heatbath.

The CG benchmark configurations are thermalised on the device with this
heatbath, so it is pinned here:
* plaquette: 1 on a cold field, the ILDG header value 0.571077 (and the oracle's
  sum) on conf.00200;
* overrelaxation is microcanonical: it changes the links but keeps the Wilson
  action (the plaquette sum) to rounding -- a wrong staple breaks this;
* heatbath, strong coupling: <P> = beta/18 + O(beta^2) (beta^2 term <= 0.002 at
  beta = 0.6), which catches a wrong coupling factor in the sampler;
* heatbath at beta = 6.0 on 16^4 from a cold start, 20 x (heatbath + 4
  overrelaxation) sweeps: plaquette 0.593678 within 0.3 %, the reference's own
  test (executables/tests/heatbath_test_input_1, heatbath_test_ref, testHeatbath.py);
* links stay in SU(3).
"""
import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import abi
from conftest import CONF_00200

pytestmark = pytest.mark.gpu


def su3_defect(links):
    u = (links[:, 0::2] + 1j * links[:, 1::2]).reshape(-1, 3, 3)
    unit = np.abs(u @ np.conj(np.transpose(u, (0, 2, 1))) - np.eye(3)).max()
    det = np.abs(np.linalg.det(u) - 1.0).max()
    return max(unit, det)


def test_plaquette_cold_and_conf00200():
    with abi.Context(4, 4) as ctx:
        assert abs(ctx.gauge().plaquette() - 1.0) < 1e-14
        lat = O.lattice(4, 4)
        g = O.ildg_read(CONF_00200, lat)
        p = ctx.gauge(g).plaquette()
        assert abs(p - O.plaquette(lat, g)) < 1e-13
        assert abs(p - 0.571077) < 1e-6


def test_overrelaxation_keeps_action():
    with abi.Context(8, 8) as ctx:
        gf = ctx.gauge(seed=5, epsilon=10.0)  # hot start
        gf.heatbath(5.7, 2)
        p0, u0 = gf.plaquette(), gf.download()
        ctx.gauge_overrelax(gf.ptr, 3)
        p1, u1 = gf.plaquette(), gf.download()
        assert np.abs(u1 - u0).max() > 0.1
        assert abs(p1 - p0) < 1e-12
        assert su3_defect(u1) < 1e-13


def test_heatbath_strong_coupling():
    beta = 0.6
    with abi.Context(8, 8) as ctx:
        gf = ctx.gauge()
        gf.heatbath(beta, 10, seed=7)
        ps = []
        for s in range(40):
            gf.heatbath(beta, 1, seed=100 + s)
            ps.append(gf.plaquette())
        assert abs(np.mean(ps) - beta / 18) < 0.003
        assert su3_defect(gf.download()) < 1e-13


def test_heatbath_beta6_reference():
    with abi.Context(16, 16) as ctx:
        gf = ctx.gauge()
        gf.heatbath(6.0, 20, overrelax=4, seed=1)
        p = gf.plaquette()
        assert abs(p - 0.593678) < 0.003 * 0.593678, p
        assert su3_defect(gf.download()) < 1e-13


def test_heatbath_seeded():
    with abi.Context(4, 4) as ctx:
        a, b, c = ctx.gauge(), ctx.gauge(), ctx.gauge()
        a.heatbath(5.7, 2, overrelax=1, seed=3)
        b.heatbath(5.7, 2, overrelax=1, seed=3)
        c.heatbath(5.7, 2, overrelax=1, seed=4)
        assert np.array_equal(a.download(), b.download())
        assert not np.array_equal(a.download(), c.download())


def test_dropin_heatbath():
    from cl2qcd_amd import _host as H
    l = H.Lattice(["--nSpace=8", "--nTime=8", "--beta=5.7"])
    gf = H.Gaugefield(l, False)
    assert abs(H.measurePlaquette(gf) - 1.0) < 1e-14
    for _ in range(3):
        H.su3heatbath(gf, l, 2)
    p = H.measurePlaquette(gf)
    assert 0.4 < p < 0.8
    assert abs(H.measurePlaquetteWithoutNormalization(gf) - p * 18 * 8 ** 4) < 1e-8 * p * 18 * 8 ** 4
    H.overrelax(gf, l, 2)
    assert abs(H.measurePlaquette(gf) - p) < 1e-12
    l12 = H.Lattice(["--nTime=4", "--useReconstruct12=true"])
    with pytest.raises(Exception):
        H.su3heatbath(H.Gaugefield(l12, False), l12)


def test_save_and_reload_ildg(tmp_path):
    """Gaugefield.save (physics/lattices/gaugefield.cpp:113-127) of a heatbath configuration, read back by the
    ILDG constructor: bit-identical links and the same plaquette; the serving engine checkpoints a resident
    configuration the same way and loads it back as a new resident one."""
    from cl2qcd_amd import _host as H
    l = H.Lattice(["--nSpace=8", "--nTime=8", "--beta=5.7", "--hostSeed=77"])
    gf = H.Gaugefield(l, False)
    for _ in range(5):
        H.su3heatbath(gf, l, 1)
    path = str(tmp_path / "conf.hb")
    gf.save(path, 5)
    back = H.Gaugefield(l, path)
    assert np.array_equal(back.get_host(), gf.get_host())
    assert H.measurePlaquette(back) == H.measurePlaquette(gf)
    links, info = H.ildg_read(path, 8, 8)
    assert "trajectory nr = 5" in info["xlf_info"] and info["has_checksum"]
    lat = O.lattice(8, 8)
    assert abs(O.plaquette(lat, np.ascontiguousarray(links.reshape(-1, 18))) - H.measurePlaquette(gf)) < 1e-12
    e = H.Engine(workers=1)
    c = e.load_config(8, 8, 5.7, 3, 11)
    ck = str(tmp_path / "conf.engine")
    e.save_config(c, ck)
    c2 = e.load_config_file(8, 8, ck)
    assert e.plaquette(c2) == e.plaquette(c)
    a = e.wait(e.submit(config=c, action="twistedmass", kappa=0.15, mu=0.01, gaussian_seeds=[3]))
    b = e.wait(e.submit(config=c2, action="twistedmass", kappa=0.15, mu=0.01, gaussian_seeds=[3]))
    assert a["ok"] and b["ok"] and np.array_equal(a["solutions"][0], b["solutions"][0])
