"""Learned moment denoiser (include/rh_denoise.h) against a plain PyTorch
FP32 CPU reference of the same network.

The device computes with BF16 operands and FP32 accumulation and stores BF16
activations; the reference rounds inputs, weights and (for the forward
check) activations to BF16 the same way and accumulates in FP32, so the
remaining differences are summation order and the odd BF16 rounding flip
that it causes.  Tolerances: forward output within 1e-2 (relative L2 of the
convolution's contribution); weight gradients within 5e-3 per layer of a
CPU emulation with the device's BF16 rounding points, and within 6e-2 of an
all-FP32 autograd backward; loss within 1e-5 and 1e-3 of the two.
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


def emulated_gradient(x, t, w, scales):
    """The device's backward pass restated with torch ops on the CPU: BF16
    rounding of the input, the weights, the forward activations and every
    backward signal at the points where the device stores BF16, FP32
    accumulation everywhere else."""
    from rhybrid_amd.denoise import split_weights
    s = torch.tensor(scales, dtype=torch.float32)
    Ws = [bf(torch.from_numpy(np.ascontiguousarray(t_))) for t_ in split_weights(w)]
    xn = bf(x / s).permute(0, 4, 1, 2, 3)
    A = [xn]
    for i in range(3):
        A.append(bf(F.conv3d(A[-1], Ws[i], padding=1).relu()))
    out = F.conv3d(A[3], Ws[3], padding=1).permute(0, 2, 3, 4, 1)
    y = x[..., :4] + s[:4] * out
    d = (y - t) / s[:4]
    loss = (d.double() ** 2).mean().item()
    G = [None] * 4
    G[3] = bf(2.0 / d.numel() * d).permute(0, 4, 1, 2, 3)
    for l in (3, 2, 1):
        G[l - 1] = bf(F.conv_transpose3d(G[l], Ws[l], padding=1) * (A[l] > 0))
    grads = [torch.nn.grad.conv3d_weight(A[l], Ws[l].shape, G[l], padding=1) for l in range(4)]
    return grads, loss


def test_gradient_matches_emulation_and_autograd(dn_mod):
    P, Z, Y, X = 2, 8, 8, 16
    x, t = volumes(P, Z, Y, X, seed=2)
    w = dn_mod.init_weights(seed=5)
    d = dn_mod.Denoiser(P * Z * Y * X, scales=SCALES, weights=w)
    g, loss = d.gradient(x.cuda(), t.cuda())
    d.close()
    emu, emu_loss = emulated_gradient(x, t, w, SCALES)
    s = torch.tensor(SCALES[:4])
    y, Ws = reference(x, w, SCALES, round_act=False, requires_grad=True)
    ref_loss = (((y - t) / s) ** 2).mean()
    ref_loss.backward()
    print("loss gpu %.6e emulated %.6e fp32 %.6e" % (loss, emu_loss, ref_loss.item()))
    assert abs(loss - emu_loss) <= 1e-5 * emu_loss
    assert abs(loss - ref_loss.item()) <= 1e-3 * ref_loss.item()
    off = 0
    for i, Wt in enumerate(Ws):
        n = Wt.numel()
        gg = g[off:off + n]
        ge = emu[i].numpy().ravel()
        gr = Wt.grad.numpy().ravel()
        e_emu = np.linalg.norm(gg - ge) / np.linalg.norm(ge)
        e_ref = np.linalg.norm(gg - gr) / np.linalg.norm(gr)
        print("layer %d grad rel err vs emulation %.3e, vs FP32 autograd %.3e" % (i + 1, e_emu, e_ref))
        assert e_emu < 5e-3      # same rounding points: summation order only
        assert e_ref < 6e-2      # BF16 backward signals vs an all-FP32 backward
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


def binomial_filter_periodic(rho, shape):
    """[1,2,1]/4 along each axis with periodic wrap (the 16^3 box below)."""
    r = rho.reshape(shape)
    for ax in range(3):
        r = 0.25 * np.roll(r, 1, axis=ax) + 0.5 * r + 0.25 * np.roll(r, -1, axis=ax)
    return r.reshape(-1)


def test_denoise_moments_in_context(dn_mod, gpu_mod):
    """rh_denoise_moments: with zero weights the output is the deposit, and
    exactly the cells whose deposit departs from its binomial filter by more
    than the bound are counted as fallbacks (moments unchanged either way);
    with a loose bound the moments become the denoiser's output on the volume
    (n, Ji, cellB); with a zero bound every cell keeps its moments."""
    from rhybrid_amd import problem as P
    pr = P.config_periodic(n=16, ppc=8)
    ctx = gpu_mod.Context(pr, precision=gpu_mod.RH_FP32)
    ctx.step(1)
    ctx.deposit()
    nc = pr.grid.nc
    rho0, ji0, b0 = ctx.field("cellRhoQi"), ctx.field("cellJi"), ctx.field("cellB")
    scales = [3e6 * P.QE, 3e6 * P.QE * 4e5, 3e6 * P.QE * 4e5, 3e6 * P.QE * 4e5, 3e-9, 3e-9, 3e-9]
    d = dn_mod.Denoiser(nc, scales=scales, weights=np.zeros(dn_mod.N_WEIGHTS, np.float32))
    bound = 0.2
    f = binomial_filter_periodic(rho0, (16, 16, 16))
    dev = np.abs(rho0 - f) / np.abs(f)
    clear = np.abs(dev - bound) > 1e-4
    assert 0 < (dev > bound).sum() < nc
    ctx.denoise_moments(d, bound)
    np.testing.assert_array_equal(ctx.field("cellRhoQi"), rho0)
    np.testing.assert_array_equal(ctx.field("cellJi"), ji0)
    nfb = ctx.counters()["denoise_fallback"]
    assert (dev[clear] > bound).sum() <= nfb <= (dev[clear] > bound).sum() + (~clear).sum()
    # random weights, loose bound: the denoiser's own output
    d.set_weights(dn_mod.init_weights(seed=4, gain=0.05))
    vol = np.concatenate([rho0[:, None], ji0.reshape(-1, 3), b0.reshape(-1, 3)], 1).astype(np.float32)
    x = torch.from_numpy(vol).reshape(1, 16, 16, 16, 7).cuda()
    want = d.forward(x).cpu().numpy().reshape(-1, 4)
    ctx.denoise_moments(d, 1e30)
    np.testing.assert_allclose(ctx.field("cellRhoQi"), np.maximum(want[:, 0], pr.params.min_rho), rtol=1e-6)
    np.testing.assert_allclose(ctx.field("cellJi").reshape(-1, 3), want[:, 1:], rtol=1e-5,
                               atol=1e-6 * np.abs(want[:, 1:]).max())
    assert ctx.counters()["denoise_fallback"] == nfb
    # zero bound: every cell keeps its moments and is counted
    rho1, ji1 = ctx.field("cellRhoQi"), ctx.field("cellJi")
    ctx.denoise_moments(d, 0.0)
    np.testing.assert_array_equal(ctx.field("cellRhoQi"), rho1)
    np.testing.assert_array_equal(ctx.field("cellJi"), ji1)
    assert ctx.counters()["denoise_fallback"] == nfb + nc
    d.close()
    ctx.close()


def test_training_on_moment_pairs_beats_raw_low_ppc(dn_mod):
    """Pairs from the 32^3 periodic box at 64 ppc (target) and 1/8 subsamples
    (8 ppc input): after training on four snapshots, the denoised moments of
    a held-out snapshot are closer to the 64-ppc deposit than the raw 8-ppc
    deposit is, for n and for J."""
    from rhybrid_amd import denoise_data as DD
    from rhybrid_amd import problem as P
    pr = P.config_periodic(n=32, ppc=64)
    X, T = DD.make_pairs(pr, subsample=8, snapshots=5, steps_between=2, seed=3)
    scales = DD.channel_scales(pr)
    raw = DD.rel_error(X[4:, ..., :4], T[4:], scales)
    d = dn_mod.Denoiser(4 * 32 ** 3, scales=scales, weights=dn_mod.init_weights(seed=2, gain=0.1))
    xt, tt = torch.from_numpy(X[:4]).cuda(), torch.from_numpy(T[:4]).cuda()
    losses = [d.train_step(xt, tt, lr=2e-3) for _ in range(150)]
    y = d.forward(torch.from_numpy(X[4:]).cuda()).cpu().numpy()
    d.close()
    den = DD.rel_error(y, T[4:], scales)
    print("loss %.4f -> %.4f; held-out n error raw %.4f denoised %.4f; J raw %.4f denoised %.4f"
          % (losses[0], losses[-1], raw[0], den[0], raw[1], den[1]))
    assert den[0] < 0.8 * raw[0]
    assert den[1] < 0.8 * raw[1]


def test_denoised_8ppc_in_the_loop_beats_plain_8ppc(dn_mod):
    """The trained denoiser attached to an 8-ppc FP32 run (applied before
    every field update): after 20 steps its n and J are closer to the FP64
    256-ppc run than the plain 8-ppc run's, and within 1.5x of the plain
    64-ppc run's error; the FP32 256-ppc run from the same particles matches
    the FP64 one to 1e-5."""
    from rhybrid_amd import denoise_loop as DL
    r = DL.compare(nx=64, ny=32, nz=32, steps=20, train_iters=400)
    runs = r["runs"]
    print({k: v for k, v in runs.items()})
    d8, p8, p64 = runs["fp32_8_denoised"], runs["fp32_8"], runs["fp32_64"]
    assert runs["fp32_256"]["n_error"] < 1e-5 and runs["fp32_256"]["J_error"] < 1e-5
    assert d8["n_error"] < p8["n_error"] and d8["J_error"] < p8["J_error"]
    assert d8["n_error"] < 1.5 * p64["n_error"] and d8["J_error"] < 1.5 * p64["J_error"]
    assert 0.0 <= d8["fallback_fraction"] < 0.5
