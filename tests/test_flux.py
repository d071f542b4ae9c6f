"""Flux helpers of the front end (Flux_Region, Flux_Box, Adjoint_Flux;
ffip/simulation.py:94-250, 585-645) on GPU runs.

* A plane wave in a column periodic in x and y: the flux through a plane equals
  |E|^2 A cos(k dx / 2) / eta.  With the DFT time stamps of E ((n+1) dt) and H
  ((n+1/2) dt) the two fields of a Yee plane wave are in phase and of equal
  size at their own lattice points (the update's dispersion relation); the H
  monitor is read half a cell from its lattice, where linear interpolation of
  the wave gives the factor cos(k dx / 2).  The values are DFT spectra (no
  time-harmonic factor 1/2).
* Flux_Box: in a column carrying an Ey plane wave along x, boxes around the
  source sheet see the same outflow and a box beside it none.
* Adjoint_Flux: the adjoint gradient of a flux objective with respect to the
  design density matches central finite differences.
"""
import math

import numpy as np
import pytest

import ffip_amd as ffip
from ffip_amd import adjoint as A

pytestmark = pytest.mark.gpu
V = ffip.Vector3


def _discrete_k(f, dt=0.5):
    return 2 * math.asin(math.sin(2 * math.pi * f * dt / 2) / dt)


def test_plane_wave_flux_closed_form():
    P, nz = 8, 120
    freqs = [0.03, 0.05, 0.07]
    src = ffip.Source(ffip.Gaussian1(0.05, cutoff=8), center=V(0, 0, -40), size=V(P, P, 0), field_component="Ex")
    sim = ffip.Simulation(size=V(P, P, nz), resolution=1, pmls=[ffip.PML(16, direction="z")], sources=[src],
                          periodic=[1, 1, 0])
    plane = sim.add_flux_region(center=V(0, 0, -20), size=V(P - 2, P - 2, 0), frequency=freqs, weight=1)
    sim.run(ffip.run_until_time(800))
    flux = plane.get_values()
    e1, h2 = plane.get_fields()
    area = (P - 2) ** 2
    for fi, f in enumerate(freqs):
        e, h = e1[fi].mean(), h2[fi].mean()
        assert np.std(e1[fi]) <= 1e-6 * abs(e)  # one plane wave across the section
        k = _discrete_k(f)
        ratio = flux[fi] / (area * abs(e) ** 2)
        print("f=%.2f: flux / (|E|^2 A) = %.6f, cos(k/2) = %.6f, H/E = %.6f%+.6fi"
              % (f, ratio, math.cos(k / 2), (h / e).real, (h / e).imag))
        assert abs(ratio - math.cos(k / 2)) <= 1e-3
        assert abs(h / e - math.cos(k / 2)) <= 1e-3
        # the factor is measurable: without it the flux is off by k^2/8
        assert abs(1 - math.cos(k / 2)) > 1e-3


def test_flux_box_conservation():
    """Closed boxes in a column periodic in y and z, CPML on x, around and
    beside an Ey sheet at x = -20 (every face clear of the CPML, |x| > 44).  The reference's flux planes take the one
    term Re(E1 conj(H2)) of the Poynting vector (E1, H2 the components after the
    plane's normal in cyclic order: Ey Hz on x faces), so a box sum is the net
    outflow only for fields of that polarization, which an Ey plane wave along
    x is.  Boxes around the sheet give the same outflow whatever their size."""
    P = 8
    freqs = [0.03, 0.05, 0.07]
    src = ffip.Source(ffip.Gaussian1(0.05, cutoff=8), center=V(-20, 0, 0), size=V(0, P, P), field_component="Ey")
    sim = ffip.Simulation(size=V(120, P, P), resolution=1, pmls=[ffip.PML(16, direction="x")], sources=[src],
                          periodic=[0, 1, 1])
    around = sim.add_flux_box(center=V(-20, 0, 0), size=V(10, P - 2, P - 2), frequency=freqs)
    wider = sim.add_flux_box(center=V(-15, 0, 0), size=V(30, P - 2, P - 2), frequency=freqs)
    sim.run(ffip.run_until_time(800))
    a, b = around.get_values(), wider.get_values()
    print("outflow around the sheet: %s, wider box %s" % (a, b))
    assert np.all(a > 0)
    assert np.all(np.abs(b - a) <= 1e-3 * a)


def test_flux_box_source_free():
    """A box beside the sheet of the column above: what flows in flows out."""
    P = 8
    freqs = [0.03, 0.05, 0.07]
    src = ffip.Source(ffip.Gaussian1(0.05, cutoff=8), center=V(-20, 0, 0), size=V(0, P, P), field_component="Ey")
    sim = ffip.Simulation(size=V(120, P, P), resolution=1, pmls=[ffip.PML(16, direction="x")], sources=[src],
                          periodic=[0, 1, 1])
    box = sim.add_flux_box(center=V(20, 0, 0), size=V(20, P - 2, P - 2), frequency=freqs)
    sim.run(ffip.run_until_time(800))
    net = box.get_values()
    inflow = box.get_flux_region("x", "negative").get_values()
    print("box beside the sheet: net %s, inflow %s" % (net, inflow))
    assert np.all(np.abs(inflow) > 0)
    assert np.all(np.abs(net) <= 1e-3 * np.abs(inflow))


def _flux_problem(precision="fp64"):
    """Ex sheet at z = -10, a design box z in [-3, 3] (eps 1..4), the flux plane
    z = 8 (20 x 20 cells) as the objective at f = 0.1, CPML from |z| = 12."""
    f0 = 0.1
    start = 1.0 / (2 * f0)
    size = V(40, 40, 40)
    fwd_src = ffip.Source(ffip.Gaussian1(f0, start_time=start), center=V(0, 0, -10), size=V(40, 40, 0),
                          field_component="Ex")
    fwd = ffip.Simulation(size=size, resolution=1, pmls=[ffip.PML(8)], sources=[fwd_src], precision=precision)
    adj = ffip.Simulation(size=size, resolution=1, pmls=[ffip.PML(8)], sources=[], precision=precision)
    flux = fwd.add_flux_region(center=V(0, 0, 8), size=V(20, 20, 0), frequency=[f0], weight=1)
    fn = ffip.Gaussian1(f0, start_time=start)
    obj = ffip.Adjoint_Flux(adj, fwd, fn, fluxes=[flux], frequency=f0)
    dt = fwd.courant * fwd.dx
    norm = A.pulse_spectrum(fn, f0, dt)
    vol = ffip.Adjoint_Volume(adj, fwd, frequency=f0, center=V(), size=V(10, 10, 6), dim=V(11, 11, 7),
                              medium1=ffip.Medium(epsilon=4), medium2=ffip.Medium(), norm=norm)
    return dict(fwd=fwd, adj=adj, obj=obj, vol=vol, f0=f0, dt=dt, stop=ffip.run_until_time(1500))


def _objective(p, rho):
    p["vol"].density = rho
    p["fwd"].run(p["stop"])
    return float(np.real(p["obj"].eval_functionals_and_set_sources()[0]))


def test_adjoint_flux_gradient_matches_central_differences():
    p = _flux_problem()
    rho = A_blur((7, 11, 11))
    F0 = _objective(p, rho)
    p["adj"].run(p["stop"])
    g = p["vol"].get_sensitivity_gpu(factor=A.derivative_factor(p["f0"], p["dt"]))
    assert abs(F0) > 0
    rng = np.random.default_rng(33)
    touched = np.flatnonzero(np.abs(g).ravel() > 1e-2 * np.abs(g).max())
    h = 1e-4
    errs = []
    for q in rng.choice(touched, size=5, replace=False):
        idx = np.unravel_index(q, rho.shape)
        rp, rm = rho.copy(), rho.copy()
        rp[idx] += h
        rm[idx] -= h
        fd = (_objective(p, rp) - _objective(p, rm)) / (2 * h)
        errs.append(abs(g[idx] - fd) / abs(fd))
        print("voxel %s: adjoint %.6e, central difference %.6e" % (idx, g[idx], fd))
    print("F = %.6e; adjoint vs central differences: max relative error %.2e" % (F0, max(errs)))
    assert max(errs) < 1e-2


def A_blur(shape, seed=5):
    from ffip_amd.workloads import blurred_density
    return blurred_density(shape, seed=seed, radius=2, sigma=1.0)
