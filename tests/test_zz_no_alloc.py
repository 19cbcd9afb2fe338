"""cl2_scalar_product_eo_noalloc, the variant of cl2_scalar_product_eo that allocates nothing (its partial sums
live in a library-owned device array), called through the C ABI: it gives the same complex result bit for bit
as cl2_scalar_product_eo on two calls in a row, the second on a smaller field, which must not see the first
call's partial sums; FP16 fields and null pointers are rejected."""
import ctypes

import numpy as np
import pytest

from cl2qcd_amd import abi

pytestmark = pytest.mark.gpu


def _both(ctx, x, y, prec):
    res = ctx.scalars(4)
    ctx.scalar_product_eo(res.ptr, x.ptr, y.ptr, prec)
    ctx.scalar_product_eo_noalloc(res.at(2), x.ptr, y.ptr, prec)
    v = res.get()
    res.close()
    return v[:2], v[2:]


@pytest.mark.parametrize("prec", [abi.F64, abi.F32], ids=["f64", "f32"])
def test_noalloc_matches_scalar_product_on_shrinking_fields(prec):
    rng = np.random.default_rng(7)
    for ns, nt in ((16, 32), (8, 8)):  # the second call is on a field 32 times smaller
        with abi.Context(ns, nt) as ctx:
            xh = rng.standard_normal((ctx.veo, 24))
            yh = rng.standard_normal((ctx.veo, 24))
            x, y = ctx.spinor(prec).upload(xh), ctx.spinor(prec).upload(yh)
            ref, got = _both(ctx, x, y, prec)
            assert np.array_equal(ref, got), (ns, nt, ref, got)
            # and the value itself: sum conj(x) y of what the kernels read
            xq, yq = x.download(), y.download()
            xc = xq[:, 0::2] + 1j * xq[:, 1::2]
            yc = yq[:, 0::2] + 1j * yq[:, 1::2]
            exact = np.vdot(xc.ravel(), yc.ravel())
            assert abs(complex(*got) - exact) <= 1e-12 * np.sqrt(np.vdot(xc, xc).real * np.vdot(yc, yc).real)


def test_noalloc_rejects_what_it_does_not_serve():
    with abi.Context(4, 4) as ctx:
        x = ctx.spinor(abi.F64)
        res = ctx.scalars(2)
        with pytest.raises(abi.Cl2Error):
            ctx.scalar_product_eo_noalloc(res.ptr, x.ptr, x.ptr, abi.F16)
        with pytest.raises(abi.Cl2Error):
            ctx.scalar_product_eo_noalloc(res.ptr, x.ptr, None, abi.F64)
        assert ctx.lib.cl2_scalar_product_eo_noalloc(None, ctypes.c_void_p(res.ptr), x.ptr, x.ptr, abi.F64) != 0
