"""The device layouts of Wilson spinor fields that C-ABI callers may read directly, checked byte for byte.

A field uploaded from the reference's AoS order (site-major, 12 complex components per eo site) must sit on
the device as the header documents: FP64 double2[12][Veo]; FP32 float4[6][Veo], components 2q and 2q+1 of a
site in one 16-byte lane; a block of nrhs FP32 fields float4[6][Veo][nrhs]; and precision conversion maps
one layout onto the other element by element.  The raw bytes are read back with cl2_scalar_get.
"""
import numpy as np
import pytest

from cl2qcd_amd import abi

pytestmark = pytest.mark.gpu


def raw(ctx, ptr, nbytes):
    """the device bytes at ptr as a uint8 array"""
    assert nbytes % 8 == 0
    out = np.empty(nbytes // 8)
    ctx.scalar_get(out.ctypes.data, ptr, nbytes // 8)
    return out.view(np.uint8)


def aos(veo, seed):
    return np.random.default_rng(seed).standard_normal((veo, 24))


def test_fp64_and_fp32_spinor_layouts():
    with abi.Context(8, 4) as ctx:
        V = ctx.veo
        X = aos(V, 5)
        cx = X[:, 0::2] + 1j * X[:, 1::2]  # [site][component]
        f64 = ctx.spinor(abi.F64).upload(X)
        d64 = raw(ctx, f64.ptr, 12 * V * 16).view(np.float64).reshape(12, V, 2)
        assert np.array_equal(d64[..., 0], cx.real.T) and np.array_equal(d64[..., 1], cx.imag.T)
        f32 = ctx.spinor(abi.F32).upload(X)
        d32 = raw(ctx, f32.ptr, 12 * V * 8).view(np.float32).reshape(6, V, 4)
        c32 = cx.astype(np.complex64)
        for q in range(6):
            assert np.array_equal(d32[q, :, 0], c32[:, 2 * q].real) and np.array_equal(d32[q, :, 1], c32[:, 2 * q].imag)
            assert np.array_equal(d32[q, :, 2], c32[:, 2 * q + 1].real)
            assert np.array_equal(d32[q, :, 3], c32[:, 2 * q + 1].imag)
        # FP64 -> FP32 on the device gives the same bytes as the FP32 upload, and back again the rounded values
        g32 = ctx.spinor(abi.F32)
        ctx.convert_precision_eo(abi.FIELD_SPINOR_EO, g32.ptr, abi.F32, f64.ptr, abi.F64)
        assert np.array_equal(raw(ctx, g32.ptr, 12 * V * 8), raw(ctx, f32.ptr, 12 * V * 8))
        g64 = ctx.spinor(abi.F64)
        ctx.convert_precision_eo(abi.FIELD_SPINOR_EO, g64.ptr, abi.F64, g32.ptr, abi.F32)
        back = raw(ctx, g64.ptr, 12 * V * 16).view(np.float64).reshape(12, V, 2)
        assert np.array_equal(back[..., 0], c32.real.T.astype(np.float64))
        assert np.array_equal(back[..., 1], c32.imag.T.astype(np.float64))


@pytest.mark.parametrize("nrhs", [3, 12])
def test_fp32_block_layout(nrhs):
    with abi.Context(8, 4) as ctx:
        V = ctx.veo
        one = ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F32)
        blk = ctx.buffer(nrhs * one)
        cxs = []
        for r in range(nrhs):
            X = aos(V, 100 + r)
            cxs.append((X[:, 0::2] + 1j * X[:, 1::2]).astype(np.complex64))
            f = ctx.spinor(abi.F32).upload(X)
            ctx.block_insert_eo(blk.ptr, f.ptr, abi.F32, nrhs, r)
        d = raw(ctx, blk.ptr, nrhs * one).view(np.float32).reshape(6, V, nrhs, 4)
        for r in range(nrhs):
            for q in range(6):
                assert np.array_equal(d[q, :, r, 0], cxs[r][:, 2 * q].real)
                assert np.array_equal(d[q, :, r, 3], cxs[r][:, 2 * q + 1].imag)
        # an FP64 block converted to FP32 lands in the same layout
        one64 = ctx.lib.cl2_field_bytes(ctx.handle, abi.FIELD_SPINOR_EO, abi.F64)
        b64, b32 = ctx.buffer(nrhs * one64), ctx.buffer(nrhs * one)
        for r in range(nrhs):
            X = aos(V, 100 + r)
            f = ctx.spinor(abi.F64).upload(X)
            ctx.block_insert_eo(b64.ptr, f.ptr, abi.F64, nrhs, r)
        ctx.block_convert_eo(b32.ptr, abi.F32, b64.ptr, abi.F64, nrhs)
        assert np.array_equal(raw(ctx, b32.ptr, nrhs * one), raw(ctx, blk.ptr, nrhs * one))
