"""Design surrogate kernels (BF16 MFMA GEMM, im2col/col2im, head, Adam).

GPU checks compare against float64 CPU references computed from the same
bf16-rounded operands: products of bf16 values are exact in FP32, so the GEMM
differs from the float64 result only by FP32 summation (relative 1e-5 here);
backward passes round dZ to bf16 before the GEMMs (relative 2e-2).
"""
import ctypes
import math

import numpy as np
import pytest

from ffip_amd import _lib
from ffip_amd import surrogate as SG


def test_output_extents_and_crop():
    assert SG.conv_out((21, 101, 101), (2, 1, 1)) == (11, 101, 101)
    shape = (21, 101, 101)
    for _ in range(4):
        shape = SG.conv_out(shape, (2, 1, 1))
    assert shape == (2, 101, 101)
    # the crop (40..60) covers x, y in [-10, 10] of the [-50, 50] design nodes
    xs = np.linspace(-50, 50, 101)
    assert xs[40] == -10 and xs[60] == 10


def test_capi_rejects_bad_arguments():
    L = _lib.lib()
    assert L.ffip_sg_gemm(0, 0, 4, 4, 4, None, 4, None, 4, None, 4, 0, 1, None) == -1  # FFIP_EINVAL
    assert b"GEMM" in L.ffip_last_error()
    h = (ctypes.c_int64 * 9)(1, 2, 10, 10, 8, 5, 5, 6, 6)  # crop 5..11 outside 10
    assert L.ffip_sg_head(ctypes.cast(h, ctypes.c_void_p), ctypes.c_void_p(1), ctypes.c_void_p(1), None) != 0
    assert b"crop" in L.ffip_last_error()


def test_no_cpu_path():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(_lib.FfipError):
        SG.Surrogate()


# ---------------------------------------------------------------- GPU
def _bf16_round(x):
    import torch
    return torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("ksplit", [1, 3])
def test_gemm_matches_float64(ta, tb, ksplit):
    import torch
    rng = np.random.default_rng(5 + 2 * ta + tb + 10 * ksplit)
    M, N, K = 200, 40, 100  # none a multiple of the 128 x 32 x 32 tile
    A = _bf16_round(rng.standard_normal((K, M) if ta else (M, K)))
    B = _bf16_round(rng.standard_normal((N, K) if tb else (K, N)))
    C0 = rng.standard_normal((M, N))
    Ad = torch.tensor(A, dtype=torch.bfloat16, device="cuda")
    Bd = torch.tensor(B, dtype=torch.bfloat16, device="cuda")
    Cd = torch.tensor(C0, dtype=torch.float32, device="cuda")
    SG.gemm(Ad, Bd, Cd, M, N, K, ta=bool(ta), tb=bool(tb), accumulate=True, ksplit=ksplit)
    ref = (A.T if ta else A) @ (B.T if tb else B) + C0.astype(np.float32)
    got = Cd.cpu().numpy().astype(np.float64)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-5, err
    # overwrite mode
    SG.gemm(Ad, Bd, Cd, M, N, K, ta=bool(ta), tb=bool(tb), ksplit=ksplit)
    ref = (A.T if ta else A) @ (B.T if tb else B)
    err = np.linalg.norm(Cd.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err < 1e-5, err


@pytest.mark.gpu
def test_split_k_gemm_replays_in_a_graph():
    """The split-K GEMM clears its output with a kernel, so a caller can capture
    it in a HIP graph: every replay, with eager work between replays, gives the
    eager result."""
    import torch
    rng = np.random.default_rng(17)
    M, N, K = 200, 40, 100
    B = _bf16_round(rng.standard_normal((K, N)))
    Ad = torch.tensor(_bf16_round(rng.standard_normal((M, K))), dtype=torch.bfloat16, device="cuda")
    Bd = torch.tensor(B, dtype=torch.bfloat16, device="cuda")
    Cd = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        SG.gemm(Ad, Bd, Cd, M, N, K, ksplit=3)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        SG.gemm(Ad, Bd, Cd, M, N, K, ksplit=3)
    for rep in range(3):
        A = _bf16_round(rng.standard_normal((M, K)))
        Ad.copy_(torch.tensor(A, dtype=torch.bfloat16))
        Cd.fill_(float(rep + 3))  # stale values the replay must clear
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        ref = A @ B
        err = np.linalg.norm(Cd.cpu().numpy() - ref) / np.linalg.norm(ref)
        assert err < 1e-5, (rep, err)


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0)])
@pytest.mark.parametrize("accumulate", [0, 1])
def test_gemm_workspace_matches_float64_and_repeats(ta, tb, accumulate):
    """Split-K with a workspace: the partial tiles are added in a fixed order, so
    two calls give the same bits, and the product equals float64 to 1e-5."""
    import torch
    rng = np.random.default_rng(11 + ta + 2 * accumulate)
    M, N, K, ks = 200, 40, 1000, 7
    A = _bf16_round(rng.standard_normal((K, M) if ta else (M, K)))
    B = _bf16_round(rng.standard_normal((K, N)))
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    Ad = torch.tensor(A, dtype=torch.bfloat16, device="cuda")
    Bd = torch.tensor(B, dtype=torch.bfloat16, device="cuda")
    work = torch.empty(ks * M * N, dtype=torch.float32, device="cuda")
    out = []
    for _ in range(2):
        Cd = torch.tensor(C0, device="cuda")
        SG.gemm(Ad, Bd, Cd, M, N, K, ta=bool(ta), accumulate=bool(accumulate), ksplit=ks, work=work)
        out.append(Cd.cpu().numpy())
    assert np.array_equal(out[0], out[1])
    ref = (A.T if ta else A) @ B + (C0 if accumulate else 0)
    err = np.linalg.norm(out[0] - ref) / np.linalg.norm(ref)
    assert err < 1e-5, err
    small = torch.empty(M * N, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.FfipError):
        SG.gemm(Ad, Bd, torch.tensor(C0, device="cuda"), M, N, K, ta=bool(ta), ksplit=ks, work=small)


@pytest.mark.gpu
def test_training_repeats_bit_for_bit():
    """Two trainings from the same seed give the same weights: the weight
    gradients' split-K tiles are added in a fixed order."""
    rng = np.random.default_rng(3)
    rhos = rng.random((8, 5, 48, 48)).astype(np.float32)
    t = np.stack([np.stack([r[:, 10:19, 10:19].mean(0) * (k + 1) for k in range(4)], -1) for r in rhos])
    weights = []
    for _ in range(2):
        net = SG.Surrogate(seed=1, shape=(5, 48, 48), crop=(10, 10, 9, 9), width=16)
        net.fit(rhos, t, epochs=2, batch=8, lr=3e-3)
        weights.append([L.w.cpu().numpy().copy() for L in net.layers])
    for a, b in zip(*weights):
        assert np.array_equal(a, b)


def _torch_conv_ref(x, w, b, stride, dil, relu):
    """float64 CPU reference: x [B, z, y, x, C], w [27*C, Co]."""
    import torch
    Bn, Z, Y, X, C = x.shape
    Co = w.shape[1]
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    wt = torch.tensor(w.reshape(3, 3, 3, C, Co).transpose(4, 3, 0, 1, 2).copy(), dtype=torch.float64,
                      requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
    z = torch.nn.functional.conv3d(xt.permute(0, 4, 1, 2, 3), wt, bt, stride=stride, padding=dil, dilation=dil)
    a = torch.relu(z) if relu else z
    return xt, wt, bt, a.permute(0, 2, 3, 4, 1)


@pytest.mark.gpu
@pytest.mark.parametrize("stride,dil,relu,cin", [((2, 1, 1), (1, 2, 2), True, 3), ((1, 1, 1), (1, 1, 1), False, 1),
                                                 ((2, 1, 1), (1, 4, 4), True, 32)])
def test_conv_forward_and_backward(stride, dil, relu, cin):
    import torch
    rng = np.random.default_rng(11)
    Bn, shape, cout = 2, (7, 13, 11), 32
    layer = SG.Conv3(cin, cout, stride, dil, relu, rng)
    layer.b.copy_(torch.tensor(rng.standard_normal(cout) * 0.1, dtype=torch.float32))
    x = _bf16_round(rng.standard_normal((Bn,) + shape + (cin,)))
    w = layer.wb.to(torch.float64).cpu().numpy()
    b = layer.b.cpu().numpy().astype(np.float64)
    xd = torch.tensor(x, dtype=torch.bfloat16, device="cuda").contiguous()
    a, out = layer.forward(xd, Bn, shape, keep=True)
    assert out == SG.conv_out(shape, stride)
    xt, wt, bt, ref = _torch_conv_ref(x, w, b, stride, dil, relu)
    got = a.to(torch.float64).cpu().numpy().reshape(ref.shape)
    err = np.linalg.norm(got - ref.detach().numpy()) / np.linalg.norm(ref.detach().numpy())
    assert err < 1e-2, err  # output rounded to bf16
    # backward with a random upstream gradient
    dA = rng.standard_normal(tuple(ref.shape))
    ref.backward(torch.tensor(dA))
    dX = layer.backward(torch.tensor(dA.reshape(-1, cout), dtype=torch.float32, device="cuda"))
    gx = dX.cpu().numpy().astype(np.float64)
    gw = layer.gw.cpu().numpy().astype(np.float64)
    gb = layer.gb.cpu().numpy().astype(np.float64)
    rx = xt.grad.numpy()
    rw = wt.grad.numpy().transpose(2, 3, 4, 1, 0).reshape(27 * cin, cout)
    rb = bt.grad.numpy()
    for g, r, name in ((gx, rx, "dX"), (gw, rw, "dW"), (gb, rb, "db")):
        e = np.linalg.norm(g - r) / np.linalg.norm(r)
        assert e < 2e-2, (name, e)


@pytest.mark.gpu
def test_head_and_adam():
    import torch
    rng = np.random.default_rng(3)
    Bn, nz, ny, nx, C = 2, 3, 9, 10, 8
    A = _bf16_round(rng.standard_normal((Bn, nz, ny, nx, C)))
    Ad = torch.tensor(A, dtype=torch.bfloat16, device="cuda")
    h = [Bn, nz, ny, nx, C, 2, 3, 5, 4]
    hp, keep = SG._arr(h)
    out = torch.empty((Bn, 5, 4, C), dtype=torch.float32, device="cuda")
    assert _lib.lib().ffip_sg_head(hp, SG._vp(Ad), SG._vp(out), SG._stream()) == 0
    ref = A.mean(axis=1)[:, 2:7, 3:7, :]
    assert np.allclose(out.cpu().numpy(), ref, rtol=1e-6, atol=1e-6)
    dout = rng.standard_normal((Bn, 5, 4, C))
    dA = torch.empty((Bn, nz, ny, nx, C), dtype=torch.float32, device="cuda")
    assert _lib.lib().ffip_sg_head_back(hp, SG._vp(torch.tensor(dout, dtype=torch.float32, device="cuda")),
                                        SG._vp(dA), SG._stream()) == 0
    want = np.zeros_like(A)
    want[:, :, 2:7, 3:7, :] = dout[:, None] / nz
    assert np.allclose(dA.cpu().numpy(), want, rtol=1e-6, atol=1e-7)
    # Adam: one step from zero moments moves every weight by lr * sign(g) (up to eps)
    w = torch.tensor(rng.standard_normal(100), dtype=torch.float32, device="cuda")
    g = torch.tensor(rng.standard_normal(100), dtype=torch.float32, device="cuda")
    m, v = torch.zeros_like(w), torch.zeros_like(w)
    wb = torch.empty(100, dtype=torch.bfloat16, device="cuda")
    w0 = w.cpu().numpy().copy()
    hpar = (ctypes.c_float * 4)(1e-2, 0.9, 0.999, 1e-8)
    assert _lib.lib().ffip_sg_adam(100, SG._vp(w), SG._vp(g), SG._vp(m), SG._vp(v),
                                   ctypes.cast(hpar, ctypes.c_void_p), 1, SG._vp(wb), SG._stream()) == 0
    step = w0 - w.cpu().numpy()
    assert np.allclose(step, 1e-2 * np.sign(g.cpu().numpy()), rtol=1e-4, atol=1e-7)
    assert np.allclose(wb.to(torch.float32).cpu().numpy(), w.cpu().numpy(), rtol=1e-2)


@pytest.mark.gpu
def test_training_reduces_loss_and_input_gradient():
    rng = np.random.default_rng(0)
    net = SG.Surrogate(seed=1, shape=(5, 24, 24), crop=(6, 6, 9, 9), width=16)
    rhos = rng.random((16, 5, 24, 24)).astype(np.float32)
    # a smooth target that depends on the density near the crop
    t = np.stack([np.stack([r[:, 6:15, 6:15].mean(0) * (k + 1) for k in range(4)], -1) for r in rhos])
    hist = net.fit(rhos, t, epochs=25, batch=4, lr=3e-3)
    assert hist[-1] < 0.5 * hist[0], hist
    # surrogate objective gradient vs a central difference on one density node
    F, g = net.objective(rhos[0], fidx=1)
    q = np.unravel_index(np.argmax(np.abs(g)), g.shape)
    h = 0.05
    rp, rm = rhos[0].copy(), rhos[0].copy()
    rp[q] += h
    rm[q] -= h
    fd = (net.values(rp[None], fidx=1)[0] - net.values(rm[None], fidx=1)[0]) / (2 * h)
    # activations are bf16, so this checks the wiring (sign, magnitude), not digits
    assert abs(fd - g[q]) < 0.3 * abs(g[q]) + 1e-6, (fd, g[q])


def test_mirror_augment_flips_densities_and_targets():
    rng = np.random.default_rng(2)
    r = rng.random((2, 3, 5, 7))
    t = rng.random((2, 4, 6, 3))
    R, T = SG.mirror_augment(r, t)
    assert R.shape == (8, 3, 5, 7) and T.shape == (8, 4, 6, 3)
    assert np.array_equal(R[2], r[0][..., ::-1]) and np.array_equal(T[2], t[0][:, ::-1, :])   # x
    assert np.array_equal(R[4], r[0][:, ::-1, :]) and np.array_equal(T[4], t[0][::-1, :, :])  # y
    assert np.array_equal(R[7], r[1][:, ::-1, ::-1]) and np.array_equal(T[7], t[1][::-1, ::-1, :])


@pytest.mark.gpu
def test_design_workload_is_mirror_symmetric():
    """The premise of mirror_augment: the solver's patch intensities for a
    density flipped along x (y) are the original ones flipped along x (y)."""
    from ffip_amd.workloads import blurred_density
    p, mon = SG.design_problem(steps=1200, precision="fp64")
    rho = blurred_density(p.shape, seed=1001, radius=3, sigma=1.5)
    out = []
    for r in (rho, rho[..., ::-1], rho[:, ::-1, :]):
        p.objective(np.ascontiguousarray(r))
        out.append(SG.to_intensity(SG.patch_fields(mon)).astype(np.float64))
    base = out[0]
    assert np.linalg.norm(base) > 0
    for got, want in ((out[1], base[:, ::-1, :]), (out[2], base[::-1, :, :])):
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err < 1e-9, err
