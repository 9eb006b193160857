"""rh_dn_forward_static (include/rh_denoise.h), the denoiser forward pass
that allocates nothing: its scratch is the library's static device array.
It runs the same kernels as rh_dn_forward on the same weights, so its
output must equal rh_dn_forward's bit for bit.  Two calls in a row, the
second on a smaller input (the array keeps the first call's contents), then
a call beyond the array's size, which must fail with RH_DN_ERR_CAPACITY.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

SCALES = (1.0, 2.0, 1.0, 0.5, 1.0, 1.0, 1.5)


def volumes(P, Z, Y, X, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.randn((P, Z, Y, X, 7), generator=g) * torch.tensor(SCALES)
    return x.cuda().contiguous()


def test_static_forward_equals_forward(gpu_mod):
    from rhybrid_amd import denoise as D
    big, small = (4, 32, 32, 32), (2, 8, 16, 32)  # (P, Z, Y, X): 131072 and 8192 voxels
    # capacity of the denoiser's own buffers: the larger input
    dn = D.Denoiser(int(np.prod(big)), scales=SCALES, weights=D.init_weights(seed=5))
    try:
        for shape, seed in ((big, 1), (small, 2)):
            x = volumes(*shape, seed=seed)
            P, Z, Y, X = shape
            y_ref = dn.forward(x)
            y_st = torch.full((P, Z, Y, X, 4), float("nan"), dtype=torch.float32, device="cuda")
            dn.synchronize()
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            dn.forward_static(x, y_st)
            dn.synchronize()
            free1 = torch.cuda.mem_get_info()[0]
            assert torch.isfinite(y_st).all()
            assert torch.equal(y_st, y_ref), "shape %s: %d differing values" % (
                shape, int((y_st != y_ref).sum()))
            # the residual path is live: the output is not the input's first channels
            assert not torch.equal(y_st, x[..., :4])
            assert free1 == free0, "device memory changed across the call (%d -> %d)" % (free0, free1)
    finally:
        dn.close()


def test_static_forward_beyond_its_array_and_capacity(gpu_mod):
    from rhybrid_amd import denoise as D
    # a denoiser whose own buffers hold one 16x4x4 volume: the static forward
    # is bounded by the static array instead
    dn = D.Denoiser(16 * 4 * 4, scales=SCALES, weights=D.init_weights(seed=6))
    try:
        shape = (1, 16, 32, 64)  # 32768 voxels, beyond the denoiser's own capacity
        x = volumes(*shape, seed=3)
        y = torch.empty(shape + (4,), dtype=torch.float32, device="cuda")
        with pytest.raises(D.DenoiserError) as e:
            dn.forward(x, y)
        assert e.value.code == D.RH_DN_ERR_CAPACITY
        dn.forward_static(x, y)
        dn.synchronize()
        # against a denoiser with room for the volume in its own buffers
        dn2 = D.Denoiser(int(np.prod(shape)), scales=SCALES, weights=D.init_weights(seed=6))
        try:
            assert torch.equal(y, dn2.forward(x))
        finally:
            dn2.close()
        # one z plane group more than the array holds
        P, Z, Y, X = 1, 68, 64, 64  # 278528 voxels > 262144
        assert P * Z * Y * X > D.STATIC_MAX_VOXELS
        xb = torch.zeros((P, Z, Y, X, 7), dtype=torch.float32, device="cuda")
        yb = torch.empty((P, Z, Y, X, 4), dtype=torch.float32, device="cuda")
        with pytest.raises(D.DenoiserError) as e:
            dn.forward_static(xb, yb)
        assert e.value.code == D.RH_DN_ERR_CAPACITY
        assert "RH_DN_STATIC_MAX_VOXELS" in str(e.value)
        # shape rules as rh_dn_forward
        with pytest.raises(D.DenoiserError) as e:
            dn.forward_static(torch.zeros((1, 4, 4, 8, 7), device="cuda"), torch.empty((1, 4, 4, 8, 4), device="cuda"))
        assert e.value.code == D.RH_DN_ERR_ARG
    finally:
        dn.close()
