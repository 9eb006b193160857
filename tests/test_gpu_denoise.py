"""Learned moment denoiser (include/rh_denoise.h) against a plain PyTorch
FP32 CPU reference of the same network.

The device computes with BF16 operands and FP32 accumulation and stores BF16
activations; the reference rounds inputs, weights and (for the forward
check) activations to BF16 the same way and accumulates in FP32, so the
remaining differences are summation order and the odd BF16 rounding flip
that it causes.  Tolerances: forward output within 1e-2 (relative L2 of the
convolution's contribution), loss within 1e-3, weight gradients within 3e-2
per layer (the device also rounds the backward signals to BF16).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import torch.nn.functional as F  # noqa: E402

pytestmark = pytest.mark.gpu

SCALES = (1.0, 2.0, 1.0, 0.5, 1.0, 1.0, 1.5)


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def reference(x, w, scales, round_act=True, requires_grad=False):
    from rhybrid_amd.denoise import split_weights
    s = torch.tensor(scales, dtype=torch.float32)
    h = bf(x / s).permute(0, 4, 1, 2, 3)
    Ws = [bf(torch.from_numpy(np.ascontiguousarray(t))) for t in split_weights(w)]
    if requires_grad:
        for t in Ws:
            t.requires_grad_(True)
    for i in range(3):
        h = F.conv3d(h, Ws[i], padding=1).relu()
        if round_act:
            h = bf(h)
    out = F.conv3d(h, Ws[3], padding=1).permute(0, 2, 3, 4, 1)
    return x[..., :4] + s[:4] * out, Ws


def volumes(P, Z, Y, X, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.randn((P, Z, Y, X, 7), generator=g) * torch.tensor(SCALES)
    t = x[..., :4] + 0.3 * torch.randn((P, Z, Y, X, 4), generator=g) * torch.tensor(SCALES[:4])
    return x, t


@pytest.fixture(scope="module")
def dn_mod(gpu_mod):
    from rhybrid_amd import denoise
    return denoise


@pytest.mark.parametrize("shape", [(2, 8, 16, 32), (1, 4, 4, 16)])
def test_forward_matches_reference(dn_mod, shape):
    P, Z, Y, X = shape
    x, _ = volumes(P, Z, Y, X, seed=1)
    w = dn_mod.init_weights(seed=3)
    d = dn_mod.Denoiser(P * Z * Y * X, scales=SCALES, weights=w)
    y = d.forward(x.cuda()).cpu()
    d.close()
    ref, _ = reference(x, w, SCALES)
    conv_part = (ref - x[..., :4]).norm()
    err = (y - ref).norm() / conv_part
    print("forward rel err %.3e (|conv| %.3e)" % (err, conv_part))
    assert conv_part > 1.0
    assert err < 1e-2


def test_gradient_matches_autograd(dn_mod):
    P, Z, Y, X = 2, 8, 8, 16
    x, t = volumes(P, Z, Y, X, seed=2)
    w = dn_mod.init_weights(seed=5)
    d = dn_mod.Denoiser(P * Z * Y * X, scales=SCALES, weights=w)
    g, loss = d.gradient(x.cuda(), t.cuda())
    d.close()
    s = torch.tensor(SCALES[:4])
    y, Ws = reference(x, w, SCALES, round_act=False, requires_grad=True)
    ref_loss = (((y - t) / s) ** 2).mean()
    ref_loss.backward()
    print("loss gpu %.6e ref %.6e" % (loss, ref_loss.item()))
    assert abs(loss - ref_loss.item()) <= 1e-3 * ref_loss.item()
    off = 0
    for i, Wt in enumerate(Ws):
        n = Wt.numel()
        gr = Wt.grad.numpy().ravel()
        gg = g[off:off + n]
        err = np.linalg.norm(gg - gr) / np.linalg.norm(gr)
        print("layer %d grad rel err %.3e" % (i + 1, err))
        assert err < 3e-2
        off += n


def test_training_reduces_loss(dn_mod):
    """Denoise a smooth field from a noisy copy: 40 Adam steps reduce the loss."""
    P, Z, Y, X = 2, 8, 16, 16
    g = torch.Generator().manual_seed(7)
    zz, yy, xx = torch.meshgrid(torch.arange(Z), torch.arange(Y), torch.arange(X), indexing="ij")
    clean = torch.stack([torch.sin(0.3 * xx + 0.2 * c) * torch.cos(0.25 * yy) + 0.1 * zz / Z for c in range(7)], -1)
    clean = clean.expand(P, Z, Y, X, 7).contiguous().float()
    x = clean + 0.5 * torch.randn(clean.shape, generator=g)
    t = clean[..., :4].contiguous()
    d = dn_mod.Denoiser(P * Z * Y * X, scales=[1.0] * 7, weights=dn_mod.init_weights(seed=1, gain=0.5))
    xc, tc = x.cuda(), t.cuda()
    losses = [d.train_step(xc, tc, lr=2e-3) for _ in range(40)]
    d.close()
    print("loss %.4f -> %.4f" % (losses[0], losses[-1]))
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.6 * losses[0]


def test_weights_roundtrip_and_errors(dn_mod):
    w = dn_mod.init_weights(seed=9)
    d = dn_mod.Denoiser(16 * 4 * 4, weights=w)
    np.testing.assert_array_equal(d.weights(), w)
    x = torch.zeros((1, 4, 4, 16, 7), device="cuda")
    with pytest.raises(dn_mod.DenoiserError) as e:
        d.forward(torch.zeros((1, 4, 4, 8, 7), device="cuda"))   # X not a multiple of 16
    assert e.value.code == dn_mod.RH_DN_ERR_ARG
    with pytest.raises(dn_mod.DenoiserError) as e:
        d.forward(torch.zeros((2, 4, 4, 16, 7), device="cuda"))  # beyond the capacity
    assert e.value.code == dn_mod.RH_DN_ERR_CAPACITY
    with pytest.raises(dn_mod.DenoiserError):
        d.set_scales([1, 1, 1, 0, 1, 1, 1])
    y = d.forward(x)
    assert torch.count_nonzero(y).item() == 0   # zero input, no biases: zero output
    d.close()
