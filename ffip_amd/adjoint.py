"""Adjoint-method design gradients on the GPU behind torch.autograd.

The reference computes a design gradient from two full solves and NumPy
post-processing (ffip/simulation.py:648-915; scripts/optimizations/
disk_adjoint.py:239-291).  ``DesignProblem`` runs the same two solves on the GPU
(forward with the physical sources; adjoint with inhomogeneous sources set from
conj(dF/dE), encoded as amplitude and time shift, src/simulation.cpp:316-333)
and forms the gradient on the device from the two runs' DFT sums
(ffip_adjoint_gradient), without storing per-step fields.  ``DesignObjective``
exposes F(rho) as a torch.autograd.Function.

Two normalisations of the sensitivity Re(c E_fwd E_adj eps' / S):
  "reference": c = 2 pi i f and S = sum_t exp(-2 pi i f t) g(t) over
               t = arange(start, end, dt), as the reference's scripts do
               (scripts/sensitivity tests/check_sensitivity.py:43-44);
  "discrete":  c = (1 - exp(-i w dt)) / dt and S = sum_n g((n+1/2) dt)
               exp(-i w (n+1) dt), the pulse samples at the times the engine
               injects them and the DFT time stamps of E.  This is the exact
               derivative of the time-stepped system (a change of eps at a point
               acts as the D source -d eps (E^{n+1} - E^n)) once the fields
               have decayed.
"""
from __future__ import annotations

import math

import numpy as np

from .geom import Medium, Vector3
from .simulation import Adjoint_Source, Adjoint_Volume, Simulation, field_match_objective, run_until_time
from .source import Gaussian1


def pulse_spectrum(function: Gaussian1, frequency: float, dt: float, mode: str = "discrete") -> complex:
    """Spectrum S of the adjoint pulse at `frequency` (see the module notes)."""
    w = 2 * math.pi * frequency
    if mode == "reference":
        t = np.arange(function.start_time, function.end_time, dt)
        return complex(np.sum(np.exp(-1j * w * t) * function(t)))
    n = np.arange(0, int(math.ceil(function.end_time / dt)) + 2)
    t = (n + 0.5) * dt
    act = (t > function.start_time) & (t < function.end_time)
    return complex(np.sum(function(t[act]) * np.exp(-1j * w * (n[act] + 1) * dt)))


def derivative_factor(frequency: float, dt: float, mode: str = "discrete") -> complex:
    """c / (2 pi i f): 1 for the reference's continuous factor, else the ratio of
    the time-discrete derivative (1 - exp(-i w dt)) / dt to 2 pi i f."""
    if mode == "reference":
        return 1.0
    w = 2 * math.pi * frequency
    return complex((1 - np.exp(-1j * w * dt)) / dt / (1j * w))


class DesignProblem:
    """Forward/adjoint pair for a two-medium design box and a field objective.

    size, resolution, pmls, sources: as for Simulation (sources are the forward
    excitation); design: dict(center, size, dim, medium1, medium2);
    objective: dict(center, size, dim, component, obj_val (complex array or 0),
    weight); frequency: design frequency; function: pulse of the adjoint sources
    (give it a start time of at least 1/(2 f) so that phase delays stay after t = 0);
    stop_time; precision "fp64" | "fp32"; mode "discrete" | "reference".
    """

    def __init__(self, size, resolution, pmls, sources, design, objective, frequency, function, stop_time,
                 precision="fp64", mode="discrete"):
        self.frequency = float(frequency)
        self.function = function
        self.stop = run_until_time(stop_time)
        self.mode = mode
        self.fwd = Simulation(size=size, resolution=resolution, pmls=pmls, sources=sources, precision=precision)
        self.adj = Simulation(size=size, resolution=resolution, pmls=pmls, sources=[], precision=precision)
        dims = objective["dim"]
        shape = (int(dims.z), int(dims.y), int(dims.x))
        obj_val = objective.get("obj_val", 0)
        if np.isscalar(obj_val):
            obj_val = np.full(shape, obj_val, dtype=complex)
        self.udf = field_match_objective(obj_val, objective.get("component", "Ez"), weight=objective.get("weight", 1))
        self.adj_src = Adjoint_Source(self.adj, self.fwd, function, frequency=frequency, center=objective["center"],
                                      size=objective["size"], dim=dims, udfs=[self.udf])
        self.dt = self.fwd.courant * self.fwd.dx
        self.norm = pulse_spectrum(function, frequency, self.dt, mode)
        self.factor = derivative_factor(frequency, self.dt, mode)
        self.vol = Adjoint_Volume(self.adj, self.fwd, frequency=frequency, center=design["center"],
                                  size=design["size"], dim=design["dim"], density=None,
                                  medium1=design.get("medium1", Medium()), medium2=design.get("medium2", Medium()),
                                  norm=self.norm)
        self.F = None

    @property
    def shape(self):
        return self.vol.shape

    def objective(self, rho):
        """Forward solve at density rho; returns F and prepares the adjoint sources."""
        self.vol.density = np.asarray(rho, dtype=float).reshape(self.shape)
        self.fwd.run(self.stop)
        self.F = float(np.real(self.adj_src.eval_functionals_and_set_sources()))
        return self.F

    def gradient(self):
        """Adjoint solve and the gradient dF/drho on the density nodes."""
        if self.F is None:
            raise RuntimeError("call objective() first")
        self.adj.run(self.stop)
        return self.vol.get_sensitivity_gpu(norm=self.norm, factor=self.factor)

    def value_and_grad(self, rho):
        F = self.objective(rho)
        return F, self.gradient()


def _torch_function():
    import torch

    class DesignObjective(torch.autograd.Function):
        """F(rho) of a DesignProblem; backward runs the adjoint solve."""

        @staticmethod
        def forward(ctx, rho, problem):
            F = problem.objective(rho.detach().cpu().numpy().astype(np.float64))
            ctx.problem = problem
            ctx.meta = (rho.dtype, rho.device)
            return torch.tensor(F, dtype=rho.dtype, device=rho.device)

        @staticmethod
        def backward(ctx, grad_out):
            g = ctx.problem.gradient()
            dtype, device = ctx.meta
            return grad_out * torch.from_numpy(g).to(dtype=dtype, device=device), None

    return DesignObjective


def design_objective(rho, problem):
    """F(rho) with an adjoint backward: ``design_objective(rho, problem).backward()``."""
    return _torch_function().apply(rho, problem)


def c3_problem(grid=(200, 200, 100), pml_cells=10, design=(100, 100, 20), patch=(20, 20), patch_z=30.0,
               frequency=0.1, eps_hi=12.0, steps=3000, precision="fp64", mode="discrete", source_z_offset=15):
    """The adjoint workload: grid with CPML, Ex planar source at z = -gz/2 +
    source_z_offset, a design box of (design + 1) density nodes (eps 1 .. eps_hi),
    F = sum |DFT Ez|^2 over the patch nodes at z = patch_z, `steps` steps per solve."""
    gx, gy, gz = grid
    start = 1.0 / (2 * frequency)  # leaves room for the adjoint phase delays
    fwd_fn = Gaussian1(frequency, start_time=start)
    src = __import__("ffip_amd.source", fromlist=["Source"]).Source(
        fwd_fn, center=Vector3(0, 0, -gz / 2 + source_z_offset), size=Vector3(gx, gy, 0), field_component="Ex")
    from .simulation import PML
    des = dict(center=Vector3(), size=Vector3(*design), dim=Vector3(design[0] + 1, design[1] + 1, design[2] + 1),
               medium1=Medium(epsilon=eps_hi), medium2=Medium())
    obj = dict(center=Vector3(0, 0, patch_z), size=Vector3(patch[0], patch[1], 0),
               dim=Vector3(patch[0] + 1, patch[1] + 1, 1), component="Ez", obj_val=0, weight=2)
    return DesignProblem(size=Vector3(gx, gy, gz), resolution=1, pmls=[PML(pml_cells)], sources=[src], design=des,
                         objective=obj, frequency=frequency, function=Gaussian1(frequency, start_time=start),
                         stop_time=steps * 0.5, precision=precision, mode=mode)
