"""Adjoint design gradient on the GPU (a10/a11).

* The device gather (ffip_adjoint_gradient) equals the reference's formula
  Re(2 pi i f E_fwd E_adj eps'/norm), scattered with transpose, evaluated in
  NumPy on the same DFT data (points at their exact positions).
* The "discrete" normalisation matches central finite differences of F (step
  1e-4 in rho) on seeded voxels and on a uniform shift of rho, once the fields
  have decayed (eps 1..4, 3000 steps); the reference's continuous normalisation
  differs from it by the factor (1 - exp(-i w dt)) / (i w dt), about 0.4 % here.
* FP32 storage gives the FP64 gradient within 1e-3 relative L2.
* torch.autograd: rho.grad equals the problem's gradient.
"""
import math

import numpy as np
import pytest

import ffip_amd as ffip
from ffip_amd import adjoint as A

pytestmark = pytest.mark.gpu


def _mini(precision="fp64", mode="discrete", steps=3000):
    # source sheet at z = -10, design box z in [-3, 3], objective patch at z = 8,
    # CPML from |z| = 12: nothing overlaps and the patch is clear of the CPML
    return A.c3_problem(grid=(40, 40, 40), pml_cells=8, design=(10, 10, 6), patch=(6, 6), patch_z=8.0,
                        eps_hi=4.0, steps=steps, precision=precision, mode=mode, source_z_offset=10)


def _rho(p, seed=3):
    return A_blur(p.shape, seed)


def A_blur(shape, seed):
    from ffip_amd.workloads import blurred_density
    return blurred_density(shape, seed=seed, radius=2, sigma=1.0)


def test_gpu_gradient_equals_reference_formula():
    p = _mini(mode="reference")
    rho = _rho(p)
    p.objective(rho)
    g_gpu = p.gradient()
    vol = p.vol
    total = 0
    for i in range(3):
        k_f = p.fwd.dft_fields.index(vol.forward_dfts[i])
        k_a = p.adj.dft_fields.index(vol.adjoint_dfts[i])
        f = p.fwd.engine.dft(k_f)
        a = p.adj.engine.dft(k_a)
        ef = (f["real"] + 1j * f["imag"])[0]
        ea = (a["real"] + 1j * a["imag"])[0]
        Z, Y, X = np.meshgrid(f["z"], f["y"], f["x"], indexing="ij")
        c, s = vol.center, vol.size
        inside = ((np.abs(X - c.x) < s.x / 2) & (np.abs(Y - c.y) < s.y / 2) & (np.abs(Z - c.z) < s.z / 2))
        se = np.real(2j * math.pi * p.frequency * ef * ea * vol.epsilon_der(0.0) / p.norm)[inside]
        pts = np.stack([X[inside], Y[inside], Z[inside], se], axis=1)
        total = total + ffip.transpose(vol.x, vol.y, vol.z, pts)
    err = np.linalg.norm(g_gpu - total) / np.linalg.norm(total)
    print("gpu gather vs numpy transpose: rel L2 %.3e" % err)
    assert err < 1e-10


def test_gradient_matches_central_differences():
    p = _mini()
    rho = _rho(p)
    F0, g = p.value_and_grad(rho)
    pref = _mini(mode="reference")
    pref.objective(rho)
    g_ref = pref.gradient()
    rng = np.random.default_rng(33)
    # face nodes that no E point strictly inside the box weights have gradient 0
    touched = np.flatnonzero(np.abs(g).ravel() > 1e-2 * np.abs(g).max())
    flat = rng.choice(touched, size=6, replace=False)
    h = 1e-4
    errs, errs_ref = [], []
    for q in flat:
        idx = np.unravel_index(q, rho.shape)
        rp, rm = rho.copy(), rho.copy()
        rp[idx] += h
        rm[idx] -= h
        fd = (p.objective(rp) - p.objective(rm)) / (2 * h)
        errs.append(abs(g[idx] - fd) / abs(fd))
        errs_ref.append(abs(g_ref[idx] - fd) / abs(fd))
    fd_u = (p.objective(rho + h) - p.objective(rho - h)) / (2 * h)
    err_u = abs(g.sum() - fd_u) / abs(fd_u)
    print("F=%.6e; adjoint vs FD relative error: discrete max %.2e (uniform %.2e), reference-normalised max %.2e"
          % (F0, max(errs), err_u, max(errs_ref)))
    assert max(errs) < 2e-3
    assert err_u < 2e-3
    # the continuous normalisation is off by the discrete-derivative factor
    assert 1e-3 < max(errs_ref) < 2e-2


def test_fp32_gradient_close_to_fp64():
    rho = None
    gs = []
    for prec in ("fp64", "fp32"):
        p = _mini(precision=prec)
        rho = _rho(p) if rho is None else rho
        gs.append(p.value_and_grad(rho)[1])
    err = np.linalg.norm(gs[1] - gs[0]) / np.linalg.norm(gs[0])
    print("fp32 vs fp64 gradient rel L2 %.3e" % err)
    assert err < 1e-3


def test_torch_autograd_path():
    import torch
    p = _mini(steps=400)
    rho_np = _rho(p)
    rho = torch.tensor(rho_np, requires_grad=True)
    F = A.design_objective(rho, p)
    F.backward()
    assert rho.grad is not None and rho.grad.shape == rho.shape
    p2 = _mini(steps=400)
    g = p2.value_and_grad(rho_np)[1]
    assert np.allclose(rho.grad.numpy(), g, rtol=1e-12, atol=0)
    assert float(F) == pytest.approx(p2.F, rel=1e-14)


def test_full_size_gradient_matches_central_differences():
    """The inverse-design grid at full size (200 x 200 x 100 cells, a
    100 x 100 x 20 design box, 3000 + 3000 steps, FP64 storage) with eps 1..2:
    the adjoint gradient against central differences of F (step 1e-4 in rho)
    on 64 voxels drawn with seed 33 among those whose gradient exceeds 1 % of
    the largest.  Tolerance: 5e-3 relative per voxel (measured: largest
    1.8e-3, median 2.1e-4).

    The frequency-domain adjoint (the reference's method) is the derivative of
    F only once the fields have decayed inside the window.  With eps up to 12
    (the benchmark's design) the box rings far longer: F changes by 2.3 %
    between 3000 and 6000 steps and by 0.9 % more up to 12000, and per-voxel
    differences from central differences stay at 10-20 % even at 12000 steps;
    with eps up to 4 they are 0.1-10 % at 3000 and 6000 steps.  With eps up to
    2 the fields have decayed by step 3000."""
    from ffip_amd.workloads import blurred_density
    p = A.c3_problem(precision="fp64", eps_hi=2.0)
    rho = blurred_density(p.shape, seed=3, radius=3, sigma=1.5)
    F0, g = p.value_and_grad(rho)
    rng = np.random.default_rng(33)
    touched = np.flatnonzero(np.abs(g).ravel() > 1e-2 * np.abs(g).max())
    h = 1e-4
    errs = []
    for q in rng.choice(touched, size=64, replace=False):
        idx = np.unravel_index(q, rho.shape)
        rp, rm = rho.copy(), rho.copy()
        rp[idx] += h
        rm[idx] -= h
        fd = (p.objective(rp) - p.objective(rm)) / (2 * h)
        errs.append(abs(g[idx] - fd) / abs(fd))
    errs = np.array(errs)
    print("F=%.6e; 64 voxels: adjoint vs central differences, relative error max %.2e, median %.2e"
          % (F0, errs.max(), np.median(errs)))
    assert errs.max() < 5e-3
