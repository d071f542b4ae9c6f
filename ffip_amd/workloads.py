"""Synthetic workloads of the FDTD hot path, written as the JSON documents ffip's
Python front end produces (ffip/simulation.py:362-402), so that the same
document drives this engine and the reference.

Units follow the reference's normalised convention: resolution 1 (dx = 1),
c = 1, Courant 0.5 (dt = 0.5).  Random draws use numpy ``default_rng(seed)``.
"""
from __future__ import annotations

import math

import numpy as np

VACUUM = {"epsilon": 1.0, "mu": 1.0, "electric conductivity": 0.0, "magnetic conductivity": 0.0,
          "electric susceptibility": [], "magnetic susceptibility": []}


def medium(eps=1.0, mu=1.0, e_sus=(), e_cond=0.0):
    return {"epsilon": float(eps), "mu": float(mu), "electric conductivity": float(e_cond),
            "magnetic conductivity": 0.0, "electric susceptibility": list(e_sus), "magnetic susceptibility": []}


def lorentz(f, gamma, amp):
    return {"type": "Lorentz", "frequency": float(f), "gamma": float(gamma), "amplitude": float(amp)}


def drude(f, gamma, amp):
    return {"type": "Drude", "frequency": float(f), "gamma": float(gamma), "amplitude": float(amp)}


def gaussian1(f, cutoff=4.0, start=0.0):
    return {"type": "Gaussian1", "frequency": float(f), "cutoff": float(cutoff), "start time": float(start)}


def sigma_max_for(order, thickness, R=1e-6):
    """sigma_max giving normal-incidence reflection R for a polynomial CPML in a
    unit-impedance medium: R = exp(-2 sigma_max d / (m + 1))."""
    return -(order + 1) * math.log(R) / (2 * thickness)


def pml(thickness, order=3, R=1e-6, k_max=1.0):
    return {"thickness": float(thickness), "k max": float(k_max),
            "sigma max": sigma_max_for(order, thickness, R), "order": int(order)}


def dft(center, size, freqs, comp, group):
    return {"type": "volume fields dft", "center": [float(v) for v in center], "size": [float(v) for v in size],
            "frequency": [float(f) for f in freqs], "field component": comp, "output group": group}


def base(size, pml_cells, precision="fp64"):
    return {"courant number": 0.5, "resolution": 1.0, "size": [float(s) for s in size],
            "default material": dict(VACUUM), "progress interval": 4.0,
            "input file": "input.h5", "fields output file": "output.h5",
            "PML": [pml(pml_cells)] if pml_cells else [], "precision": precision}


def vacuum_dipole(n=64, steps=1000, pml_cells=8, precision="fp64", freqs=(0.08, 0.10, 0.12), plane_z=10):
    """Plain FP64 vacuum pulse: n^3 cells, 8-cell CPML (m = 3, R = 1e-6), Ez point
    dipole at the grid centre driven by Gaussian1(f = 0.1), one Ez DFT plane
    `plane_z` cells from the source at 3 frequencies."""
    cfg = base((n, n, n), pml_cells, precision)
    cfg["sources"] = [{"type": "volume source", "center": [0.0, 0.0, 0.0], "size": [0.0, 0.0, 0.0],
                       "amplitude": 1.0, "field component": "Ez", "function": gaussian1(0.1)}]
    cfg["fields output"] = [dft((0, 0, plane_z), (n, n, 0), freqs, "Ez", "volume fields dft 0")]
    cfg["stop condition"] = {"type": "time", "time": steps * 0.5}
    return cfg


VACUUM_DIPOLE_PROBES = [((0.0, 0.0, 5.0), "Ez"), ((10.0, 0.0, 0.0), "Ez"), ((0.0, 20.0, -10.0), "Ez")]


# Drude plasma angular frequency 1.0 rad per unit time.  The reference's
# `frequency` field is multiplied by 2*pi (src/medium.cpp:104-108); reading
# "omega_p = 1.0" as that field would give omega_p*dt = pi at Courant 0.5, for
# which the reference's explicit ADE update (src/structure.cpp:56-85) diverges.
DRUDE_WP = 1.0


def dispersive_spheres(n=256, nspheres=32, steps=4000, pml_cells=12, nfreq=16, seed=2, precision="fp32",
                       source_offset=20, comp="Ez"):
    """Dispersive scattering: n^3 grid, CPML, a planar Gaussian current sheet
    (component `comp`) at x = -n/2 + source_offset; `nspheres` seeded spheres (radius U[8,24] cells
    scaled by n/256, centres uniform in the interior), the first half 2-pole
    Lorentz dielectric (eps_inf = 2, f0 in U[0.05,0.2], gamma in U[0.005,0.02],
    delta-eps in U[0.5,2]), the second half Drude metal (omega_p = 1.0 rad per
    unit time, gamma = 0.01, amplitude 1); 2 DFT planes of `comp` at x = -n/4,
    +n/4 with `nfreq` frequencies in [0.05, 0.15].

    An x-directed current in an x-normal sheet launches no wave (its field has
    no curl away from the sheet edges), so the sheet and the monitors default to
    Ez, which launches a plane wave along x; the cost per step is the same."""
    rng = np.random.default_rng(seed)
    s = n / 256.0
    r = rng.uniform(8 * s, 24 * s, size=nspheres)
    c = rng.uniform(12 * s + r[:, None], n - 12 * s - r[:, None], size=(nspheres, 3)) - n / 2
    nl = nspheres // 2
    f0 = rng.uniform(0.05, 0.2, size=(nl, 2))
    g0 = rng.uniform(0.005, 0.02, size=(nl, 2))
    a0 = rng.uniform(0.5, 2.0, size=(nl, 2))
    cfg = base((n, n, n), pml_cells, precision)
    geoms = []
    for q in range(nspheres):
        if q < nl:
            m = medium(2.0, e_sus=[lorentz(f0[q, p], g0[q, p], a0[q, p]) for p in range(2)])
        else:
            m = medium(1.0, e_sus=[drude(DRUDE_WP / (2 * math.pi), 0.01, 1.0)])
        geoms.append({"type": "sphere", "center": [float(v) for v in c[q]], "radius": float(r[q]), "medium": m})
    cfg["geometry"] = geoms
    cfg["sources"] = [{"type": "volume source", "center": [-n / 2 + source_offset, 0.0, 0.0],
                       "size": [0.0, float(n), float(n)], "amplitude": 1.0, "field component": comp,
                       "function": gaussian1(0.1)}]
    freqs = list(np.linspace(0.05, 0.15, nfreq))
    x0 = n / 4
    cfg["fields output"] = [dft((-x0, 0, 0), (0, n, n), freqs, comp, "volume fields dft 0"),
                            dft((x0, 0, 0), (0, n, n), freqs, comp, "volume fields dft 1")]
    cfg["stop condition"] = {"type": "time", "time": steps * 0.5}
    return cfg


def periodic_plane_wave(nz=64, nxy=4, steps=400, pml_cells=8, comp="Ex", freqs=(0.05, 0.08), planes=(-10, 10),
                        precision="fp64", eps=1.0, e_sus=()):
    """Plane wave along z in a box periodic in x and y (CPML on z only): a
    Gaussian1 current sheet at z = -nz/2 + pml_cells + 4 and two DFT planes."""
    cfg = base((nxy, nxy, nz), 0, precision)
    p = pml(pml_cells)
    p["direction"] = "z"
    cfg["PML"] = [p]
    cfg["periodic"] = [1, 1, 0]
    cfg["default material"] = medium(eps, e_sus=e_sus)
    zs = -nz / 2 + pml_cells + 4
    cfg["sources"] = [{"type": "volume source", "center": [0.0, 0.0, zs], "size": [float(nxy), float(nxy), 0.0],
                       "amplitude": 1.0, "field component": comp, "function": gaussian1(0.1)}]
    cfg["fields output"] = [dft((0, 0, z), (nxy, nxy, 0), freqs, comp, "volume fields dft %d" % q)
                            for q, z in enumerate(planes)]
    cfg["stop condition"] = {"type": "time", "time": steps * 0.5}
    return cfg


def blurred_density(shape, seed=3, radius=3, sigma=1.5):
    """rho ~ U[0,1] (shape = (nz, ny, nx)) blurred by a separable, normalised
    Gaussian of the given radius and sigma with zero padding ("same" size)."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.0, 1.0, size=shape)
    t = np.arange(-radius, radius + 1, dtype=float)
    k = np.exp(-0.5 * (t / sigma) ** 2)
    k /= k.sum()
    for ax in range(3):
        n = rho.shape[ax]
        pad = [(0, 0)] * 3
        pad[ax] = (radius, radius)
        p = np.pad(rho, pad)
        out = np.zeros_like(rho)
        for q, wq in enumerate(k):
            sl = [slice(None)] * 3
            sl[ax] = slice(q, q + n)
            out += wq * p[tuple(sl)]
        rho = out
    return rho


def design_box(grid=(200, 200, 100), pml_cells=10, design=(100, 100, 20), patch=(20, 20), patch_z=30.0,
               freqs=(0.1,), seed=3, precision="fp64", source_z_offset=15, eps_hi=12.0, steps=3000):
    """Adjoint workload: grid with CPML, Ex planar source at z = -gz/2 +
    source_z_offset, a two-medium design box (eps 1 .. eps_hi, density on
    design+1 nodes seeded and blurred), and the objective patch DFT (Ez) at z =
    patch_z.  Returns (config, datasets)."""
    gx, gy, gz = grid
    cfg = base(grid, pml_cells, precision)
    dim = [design[0] + 1, design[1] + 1, design[2] + 1]
    rho = blurred_density((dim[2], dim[1], dim[0]), seed=seed)
    cfg["geometry"] = [{"type": "two medium box", "center": [0.0, 0.0, 0.0], "size": [float(v) for v in design],
                        "dimension": [float(v) for v in dim], "density dataset": "two medium box adjoint",
                        "medium1": medium(eps_hi), "medium2": medium(1.0)}]
    cfg["sources"] = [{"type": "volume source", "center": [0.0, 0.0, -gz / 2 + source_z_offset],
                       "size": [float(gx), float(gy), 0.0], "amplitude": 1.0, "field component": "Ex",
                       "function": gaussian1(0.1)}]
    cfg["fields output"] = [dft((0, 0, patch_z), (patch[0], patch[1], 0), freqs, "Ez", "volume fields dft 0")]
    cfg["stop condition"] = {"type": "time", "time": steps * 0.5}
    return cfg, {"two medium box adjoint": rho.ravel()}


def voxel_mix(n=512, voxel=8, seed=5, pml_cells=16, steps=10000, precision="fp32", nfreq=0, nplanes=8,
              lorentz_frac=0.3, drude_frac=0.1, source_offset=24):
    """Large single-GPU workload: n^3 cells, CPML, a Gaussian1 Ez current sheet in
    the plane x = -n/2 + source_offset, and a seeded mix of voxel-sized boxes in vacuum:
    `lorentz_frac` of the voxel^3-cell voxels hold a 3-pole Lorentz dielectric
    drawn from a palette of 8 (eps_inf = 2, f0 in U[0.05,0.2], gamma in
    U[0.005,0.02], delta-eps in U[0.5,2]), `drude_frac` a Drude metal from a
    palette of 4 (omega_p in U[0.8,1.2] rad per unit time, gamma = 0.01).  Voxels
    are box geometries a quarter cell larger than the voxel on every side, so the
    points on shared faces belong to one of them (first in list order) instead of
    to neither.  With nfreq > 0, `nplanes` Ez DFT planes at half-integer z carry
    nfreq frequencies in [0.05, 0.15].  28 distinct poles: within the
    reference's 31-pole pool.

    The sheet is normal to x and its current points along z, so it launches a
    plane wave along x that crosses every z-normal monitor plane (a current
    normal to its own sheet has no curl away from the sheet edges and radiates
    nothing)."""
    rng = np.random.default_rng(seed)
    f0 = rng.uniform(0.05, 0.2, size=(8, 3))
    g0 = rng.uniform(0.005, 0.02, size=(8, 3))
    a0 = rng.uniform(0.5, 2.0, size=(8, 3))
    wp = rng.uniform(0.8, 1.2, size=4)
    lor = [medium(2.0, e_sus=[lorentz(f0[q, p], g0[q, p], a0[q, p]) for p in range(3)]) for q in range(8)]
    dru = [medium(1.0, e_sus=[drude(wp[q] / (2 * math.pi), 0.01, 1.0)]) for q in range(4)]
    nv = n // voxel
    u = rng.random(nv ** 3)
    pick = rng.integers(0, 8, size=nv ** 3)
    pick_d = rng.integers(0, 4, size=nv ** 3)
    geoms = []
    half = voxel / 2 + 0.25
    for q in range(nv ** 3):
        if u[q] < lorentz_frac:
            m = lor[pick[q]]
        elif u[q] < lorentz_frac + drude_frac:
            m = dru[pick_d[q]]
        else:
            continue
        iz, r = divmod(q, nv * nv)
        iy, ix = divmod(r, nv)
        c = [(ix + 0.5) * voxel - n / 2, (iy + 0.5) * voxel - n / 2, (iz + 0.5) * voxel - n / 2]
        geoms.append({"type": "box", "center": c, "size": [2 * half] * 3, "medium": m})
    cfg = base((n, n, n), pml_cells, precision)
    cfg["geometry"] = geoms
    cfg["sources"] = [{"type": "volume source", "center": [-n / 2 + source_offset, 0.0, 0.0],
                       "size": [0.0, float(n), float(n)], "amplitude": 1.0, "field component": "Ez",
                       "function": gaussian1(0.1)}]
    if nfreq > 0:
        freqs = list(np.linspace(0.05, 0.15, nfreq))
        zs = np.linspace(-n / 4, n / 4, nplanes)
        cfg["fields output"] = [dft((0, 0, float(np.floor(z)) + 0.5), (n, n, 0), freqs, "Ez", "volume fields dft %d" % q)
                                for q, z in enumerate(zs)]
    cfg["stop condition"] = {"type": "time", "time": steps * 0.5}
    return cfg
