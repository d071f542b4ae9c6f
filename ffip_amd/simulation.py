"""Simulation front end on the MI355X engine.

Scripts written for the reference's Python package use the names of
ffip/simulation.py: Simulation and its output objects (DFT monitors, flux planes
and boxes, geometry-info requests), the stop conditions, PML and Symmetry, and
the adjoint helpers (Adjoint_Source, Adjoint_Flux, Adjoint_Volume,
field_match_objective).  Those names, their arguments and defaults, the
attributes scripts read, the JSON document (src/simulation.cpp:58-188) and the
output group names are kept.  How a run happens is this package's own: the
reference writes config.json and an HDF5 input file, starts
``mpirun -np N run_sim_json`` and reads an HDF5 output file back
(ffip/simulation.py:404-431); here the same document and arrays go in-process
to libffip_hip.so through its C ABI, one GPU holds the whole grid, and the
results are read back from device memory into the same objects.  ``np`` and
``decomposition`` are accepted and have no effect.

Output objects read their results from a mapping of group name to arrays
("real", "imag", "x", "y", "z", "f" for a DFT group; "x", "y", "z" for a
geometry-info group), which is what an HDF5 output file of run_sim_json holds
too, so ``read`` accepts either.
"""
from __future__ import annotations

from copy import deepcopy
from math import pi

import numpy as np
from scipy.interpolate import RegularGridInterpolator

from ._lib import LibSim, transpose
from .geom import Medium, Two_Medium_Box, Vector3, cmp_shape, metadata_to_pts
from .source import Inhom_Source

__all__ = ["PML", "Adjoint_Flux", "Adjoint_Source", "Adjoint_Volume", "Fields_DFT", "Flux_Box", "Flux_Region",
           "Geom_Info", "Simulation", "Symmetry", "direction_dict", "field_match_objective", "run_until_dft",
           "run_until_fields_decay", "run_until_time", "side_dict"]

direction_dict = {"x": 0, "y": 1, "z": 2}
side_dict = {"positive": 1, "negative": -1}
_AXES = "xyz"


def _vec_or_zero(v):
    return v.copy() if v is not None else Vector3()


# ------------------------------------------------------------ JSON specifications
def _jsonable(value):
    return value.get_json() if hasattr(value, "get_json") else value


class _Spec:
    """An object written as one JSON object: "type" (when the class has one)
    and then one entry per (key, attribute) pair of ``json_keys``; attributes
    that are None are left out when listed in ``json_optional``."""

    json_type = None
    json_keys = ()
    json_optional = ()

    def get_json(self):
        out = {} if self.json_type is None else {"type": self.json_type}
        for key, attr in self.json_keys:
            value = getattr(self, attr)
            if value is None and attr in self.json_optional:
                continue
            out[key] = _jsonable(value)
        return out


class run_until_time(_Spec):
    """Stop once time_step * dt reaches `time` (src/simulation.cpp:768-780)."""

    json_type = "time"
    json_keys = (("time", "time"),)

    def __init__(self, time=0):
        self.time = float(time)


class run_until_dft(_Spec):
    """After the sources turn off, run windows of `time_interval_examined` until
    the spread (max - min) / max of a registered DFT monitor's norm at
    `frequency` over one window is at most `var` (src/simulation.cpp:782-842)."""

    json_type = "dft"
    json_keys = (("center", "center"), ("size", "size"), ("field component", "field_component"),
                 ("time interval examined", "time_interval_examined"), ("variation", "var"),
                 ("frequency", "frequency"))

    def __init__(self, center=None, size=None, frequency=1, field_component="Ex", time_interval_examined=1,
                 var=1e-2):
        self.center, self.size = _vec_or_zero(center), _vec_or_zero(size)
        self.frequency = float(frequency)
        self.field_component = str(field_component)
        self.time_interval_examined = float(time_interval_examined)
        self.var = float(var)


class run_until_fields_decay(_Spec):
    """After the sources turn off, run windows of `time_interval_examined` until
    a window's largest squared field at `position` is at most `decayed_by`
    times the largest seen so far (src/simulation.cpp:844-901)."""

    json_type = "decay"
    json_keys = (("position", "position"), ("field component", "field_component"),
                 ("time interval examined", "time_interval_examined"), ("decayed by", "decayed_by"))

    def __init__(self, position=None, field_component="Ez", time_interval_examined=1, decayed_by=1e-3):
        self.position = _vec_or_zero(position)
        self.field_component = str(field_component)
        self.time_interval_examined = float(time_interval_examined)
        self.decayed_by = float(decayed_by)


class Symmetry(_Spec):
    """Mirror plane through the origin normal to `direction`: phase_factor +1
    makes it a magnetic wall (PMC), -1 an electric wall (PEC); the grid keeps
    the positive half (src/simulation.cpp:583-653)."""

    json_keys = (("direction", "direction"), ("phase factor", "phase_factor"))

    def __init__(self, direction="x", phase_factor=1):
        self.direction = direction
        self.phase_factor = int(phase_factor)


class PML(_Spec):
    """CPML layer of `thickness` on both sides of every axis, or on one
    `direction` / `side`.  The grading is sigma_max (x/d)^order and
    1 + (k_max - 1)(x/d)^order.  As in the reference's front end, a sigma_max
    given here is not used: Simulation sets 0.8 (order + 1) resolution."""

    json_keys = (("thickness", "thickness"), ("direction", "direction"), ("side", "side"),
                 ("k max", "k_max"), ("order", "order"), ("sigma max", "sigma_max"))
    json_optional = ("direction", "side")

    def __init__(self, thickness, direction=None, side=None, k_max=1, order=3, sigma_max=None):
        self.thickness = float(thickness)
        self.direction, self.side = direction, side
        self.k_max = float(k_max)
        self.order = int(order)
        self.sigma_max = None


# ------------------------------------------------------------------ DFT monitors
def _widened(freqs, axes, values, dx):
    """Axes and values with every singleton axis doubled so that linear
    interpolation works on it: a single frequency f becomes (0.9 f, 1.1 f), a
    single coordinate c becomes (c - dx/2, c + dx/2), and the values are
    repeated along that axis."""
    freqs = freqs * np.array([0.9, 1.1]) if freqs.size == 1 else freqs
    axes = [a + np.array([-dx / 2, dx / 2]) if a.size == 1 else a for a in axes]
    for i in range(values.ndim):
        if values.shape[i] == 1:
            values = np.repeat(values, 2, axis=i)
    return freqs, axes, values


class Fields_DFT:
    """A volume DFT monitor of one field component: complex sums
    sum_n F(t_n) exp(-2 pi i f t_n) at its frequencies over the points of its
    box, read back as ``values`` [f, z, y, x] with the axes ``frequency``, ``z``,
    ``y``, ``x``.  Output group "volume fields dft <suffix>"."""

    def __init__(self, center, size, frequency, field_component, suffix="0", degenerate=False, dx=0):
        if not degenerate and dx == 0:
            raise ValueError("a monitor read for interpolation needs the grid spacing dx")
        self.suffix = str(suffix)
        self.group_name = "volume fields dft %s" % self.suffix
        self.center, self.size = center.copy(), size.copy()
        self.degenerate = bool(degenerate)
        self.dx = float(dx)
        self.frequency = np.sort(np.asarray(frequency, dtype=float))
        self.frequency_config = self.frequency.tolist()
        self.field_component = field_component

    def get_json(self):
        return {"type": "volume fields dft", "center": self.center.get_json(), "size": self.size.get_json(),
                "frequency": self.frequency_config, "field component": self.field_component,
                "output group": self.group_name}

    def read(self, groups):
        g = groups[self.group_name]
        freqs = np.array(g["f"])
        axes = [np.array(g[a]) for a in _AXES]
        values = np.asarray(g["real"]) + 1j * np.asarray(g["imag"])
        if not self.degenerate:
            freqs, axes, values = _widened(freqs, axes, values, self.dx)
        self.frequency = freqs
        self.x, self.y, self.z = axes
        self.v = values

    @property
    def values(self):
        return self.v

    def get_metadata(self):
        return self.frequency, self.z, self.y, self.x

    def get_interpolant(self, substract=None, **kwargs):
        """Interpolant over (f, z, y, x) of the values, or of the values minus
        those of another monitor (`substract`) with the same axes."""
        if substract is not None and not isinstance(substract, Fields_DFT):
            raise TypeError("substract must be a Fields_DFT, not %s" % type(substract).__name__)
        data = self.values if substract is None else self.values - substract.values
        return RegularGridInterpolator(self.get_metadata(), data, **kwargs)

    def __call__(self, f, z, y, x, substract=None, **kwargs):
        return self.get_interpolant(substract, **kwargs)(metadata_to_pts(f, z, y, x))


# ------------------------------------------------------------------------- flux
def _trapezoid(n):
    w = np.ones(int(n))
    if w.size > 1:
        w[0] = w[-1] = 0.5
    return w


class Flux_Region:
    """Flux Re(E1 conj(H2)) through a plane normal to `direction` (E1, H2 the
    components next in cyclic order), integrated with the trapezoid rule on a
    lattice of spacing dx / scale and multiplied by `weight`.  Registers one E
    and one H monitor on the plane.  As in the reference
    (ffip/simulation.py:186-200) this is one term of the normal Poynting
    component: the other, -Re(E2 conj(H1)), is left out, so the value is the
    flux of fields polarized with E along E1."""

    def __init__(self, sim, center, size, frequency, weight):
        extended = [size[a] != 0 for a in range(3)]
        if sum(extended) != 2:
            raise NameError("a flux region must be a plane (two nonzero extents), got %d nonzero"
                            % sum(extended))
        self.direction = extended.index(False)
        self.weight = np.array(weight, copy=True)
        self.center, self.size = center.copy(), size.copy()
        self.frequency = np.sort(np.asarray(frequency))
        self.dx = sim.dx
        self._e1_dft = sim.add_dft_fields(center=center, size=size, frequency=frequency, field_component=self.e1_str)
        self._h2_dft = sim.add_dft_fields(center=center, size=size, frequency=frequency, field_component=self.h2_str)

    @property
    def e1_str(self):
        return "E" + _AXES[(self.direction + 1) % 3]

    @property
    def h2_str(self):
        return "H" + _AXES[(self.direction + 2) % 3]

    @property
    def e1_dft(self):
        return self._e1_dft

    @property
    def h2_dft(self):
        return self._h2_dft

    def get_len(self, scale=1):
        """Integration nodes per axis: size / dx * scale rounded, plus one."""
        return (self.size / self.dx * scale).round() + Vector3(1, 1, 1)

    def get_metadata(self, scale=1):
        n = self.get_len(scale=scale)
        lo, hi = self.center - self.size / 2, self.center + self.size / 2
        xs = [np.linspace(lo[a], hi[a], int(n[a])) for a in range(3)]
        return self.frequency, xs[2], xs[1], xs[0]

    def get_pts(self, scale=1):
        return metadata_to_pts(*self.get_metadata(scale=scale))

    def get_int_weights(self, scale=1):
        """Trapezoid weights [z, y, x] times the cell measure and `weight`."""
        n = self.get_len(scale=scale)
        measure = 1.0
        for a in range(3):
            if n[a] > 1:
                measure *= self.size[a] / (n[a] - 1)
        wz, wy, wx = (_trapezoid(n[a]) for a in (2, 1, 0))
        return wz[:, None, None] * wy[None, :, None] * wx[None, None, :] * measure * self.weight

    def get_fields(self, scale=1, substract=None, **kwargs):
        """(E1, H2) on the integration lattice, [f, z, y, x] each."""
        if substract is not None and not isinstance(substract, Flux_Region):
            raise TypeError("substract must be a Flux_Region, not %s" % type(substract).__name__)
        pts = self.get_pts(scale=scale)
        fields = []
        for mine, other in ((self.e1_dft, "e1_dft"), (self.h2_dft, "h2_dft")):
            base = getattr(substract, other) if substract is not None else None
            fields.append(mine.get_interpolant(substract=base, **kwargs)(pts))
        return tuple(fields)

    def get_values(self, scale=1, substract=None, **kwargs):
        """Flux at each frequency."""
        e1, h2 = self.get_fields(scale=scale, substract=substract, **kwargs)
        density = np.real(e1 * np.conj(h2)) * self.get_int_weights(scale=scale)
        return density.sum(axis=(1, 2, 3))

    def get_interpolant(self, scale=1, substract=None):
        return RegularGridInterpolator((self.frequency,), self.get_values(scale=scale, substract=substract))


class Flux_Box:
    """Outward flux through the six faces of a box: flux regions ordered
    x-, x+, y-, y+, z-, z+ with weights -1 and +1."""

    def __init__(self, sim, center, size, frequency):
        self.frequency = np.sort(np.asarray(frequency))
        self._flux_regions = []
        for a in range(3):
            for side in (-1, 1):
                face_center = center.copy()
                face_center[a] = center[a] + side * size[a] / 2
                face_size = size.copy()
                face_size[a] = 0.0
                self._flux_regions.append(sim.add_flux_region(center=face_center, size=face_size,
                                                              frequency=frequency, weight=side))

    @property
    def flux_regions(self):
        return self._flux_regions

    def get_flux_region(self, direction="x", side="negative"):
        return self._flux_regions[2 * direction_dict[direction] + (side_dict[side] > 0)]

    def get_values(self, scale=1, substract=None, **kwargs):
        if substract is not None and not isinstance(substract, Flux_Box):
            raise TypeError("substract must be a Flux_Box, not %s" % type(substract).__name__)
        total = 0
        for i, face in enumerate(self._flux_regions):
            other = substract.flux_regions[i] if substract is not None else None
            total = total + face.get_values(scale=scale, substract=other, **kwargs)
        return total

    def get_interpolant(self, scale=1, substract=None, **kwargs):
        return RegularGridInterpolator((self.frequency,), self.get_values(scale=scale, substract=substract, **kwargs))


class Geom_Info(_Spec):
    """Request for the coordinates of the points of one field component that a
    geometry (by its index in the simulation's list) owns
    (src/simulation.cpp:190-233).  Output group
    "geometry info request <index> <component>"."""

    json_keys = (("field component", "field_component"), ("geometry index", "index"), ("output group", "gname"))

    def __init__(self, index=0, field_component="Ex"):
        self.field_component = str(field_component)
        self.index = int(index)
        self.gname = "geometry info request %d %s" % (self.index, self.field_component)

    def read(self, groups):
        g = groups[self.gname]
        self._x, self._y, self._z = (np.asarray(g[a]) for a in _AXES)

    @property
    def x(self):
        return self._x

    @property
    def y(self):
        return self._y

    @property
    def z(self):
        return self._z


# ------------------------------------------------------------------- simulation
class _Arrays:
    """The input store: what objects' ``dump`` writes, by dataset name."""

    def __init__(self):
        self.data = {}

    def create_dataset(self, name, data):
        if name in self.data:
            raise ValueError("two objects write the dataset %r" % name)
        self.data[name] = np.asarray(data, dtype=np.float64).ravel()


def _check_file_name(value):
    if not isinstance(value, str):
        raise ValueError("file names are strings, not %s" % type(value).__name__)
    return value


class Simulation:
    """A simulation set-up.  ``run`` solves it on the GPU and reads the results
    into its output objects; the engine of the last run stays available as
    ``engine`` until the next run.

    Extra keyword ``precision``: "fp64" (default, the reference's arithmetic) or
    "fp32" (fields and polarizations stored in FP32 with FP64 arithmetic,
    coefficients and DFT sums)."""

    def __init__(self, size, resolution, pmls=[], geometry=[], sources=[], symmetry=[], default_material=Medium(),
                 periodic=False, progress_interval=4, courant=0.5, input_file="empty.h5",
                 fields_output_file="output.h5", precision="fp64"):
        print("Simulation created\n")
        self.size = size.copy()
        self.resolution = float(resolution)
        self.geometry = deepcopy(list(geometry))
        self.sources = deepcopy(list(sources))
        self.symmetry = deepcopy(list(symmetry))
        self.default_material = deepcopy(default_material)
        self.periodic = deepcopy(periodic)
        self.progress_interval = float(progress_interval)
        self.courant = float(courant)
        self.pmls = deepcopy(list(pmls))
        self.input_file = input_file
        self.fields_output_file = fields_output_file
        self.precision = precision
        self.dft_fields, self.flux_regions, self.flux_boxes, self.geom_info_requests = [], [], [], []
        self.engine = None
        self._geom_info_cache = (None, {})
        for layer in self.pmls:
            if layer.sigma_max is None:
                layer.sigma_max = 0.8 * (1 + layer.order) * self.resolution

    input_file = property(lambda self: self._input_file,
                          lambda self, v: setattr(self, "_input_file", _check_file_name(v)))
    fields_output_file = property(lambda self: self._fields_output_file,
                                  lambda self, v: setattr(self, "_fields_output_file", _check_file_name(v)))

    @property
    def dx(self):
        return 1 / self.resolution

    # -- the document and its arrays
    def dump(self, store):
        for obj in list(self.geometry) + list(self.sources):
            obj.dump(store)

    def get_json(self, stop_condition, decomposition=None):
        doc = {"courant number": self.courant, "resolution": self.resolution, "size": self.size.get_json(),
               "default material": self.default_material.get_json(), "progress interval": self.progress_interval,
               "input file": self.input_file, "fields output file": self.fields_output_file,
               "stop condition": stop_condition.get_json(), "precision": self.precision}
        if decomposition is not None:
            doc["decomposition"] = decomposition
        # (the reference's front end writes self.periodic under "symmetry",
        # ffip/simulation.py:380, which cannot serialise; the symmetry planes go here)
        optional = (("symmetry", self.symmetry), ("PML", self.pmls), ("geometry", self.geometry),
                    ("sources", self.sources), ("fields output", self.dft_fields),
                    ("geometry info request", self.geom_info_requests))
        for key, items in optional:
            if items:
                doc[key] = [item.get_json() for item in items]
        if self.periodic:
            doc["periodic"] = self.periodic
        # keep the reference's key order for the documents a user inspects
        order = ["courant number", "resolution", "size", "default material", "progress interval", "input file",
                 "fields output file", "stop condition", "precision", "decomposition", "symmetry", "PML",
                 "periodic", "geometry", "sources", "fields output", "geometry info request"]
        return {k: doc[k] for k in order if k in doc}

    def _inputs(self):
        store = _Arrays()
        self.dump(store)
        return store.data

    # -- running
    def run(self, stop_condition, decomposition=None, np=1, skip=False, pop=False, **kwargs):
        """Solve to the stop condition and read the results into the output
        objects.  ``skip=True`` re-reads the last run's results without solving."""
        if not skip or self.engine is None:
            doc = self.get_json(stop_condition=stop_condition, decomposition=decomposition)
            if self.engine is not None:
                # free the previous engine's device memory before the new one allocates
                self.engine.close()
                self.engine = None
            self.engine = LibSim(doc, self._inputs())
            self.engine.run(doc["stop condition"])
            self._last_doc = doc
        self._read_results()

    def _read_results(self):
        groups = {}
        for k in range(self.engine.dft_count()):
            out = self.engine.dft(k)
            groups[out["group"]] = out
        groups.update(self._geometry_info())
        for item in list(self.dft_fields) + list(self.geom_info_requests):
            item.read(groups)

    def _geometry_info(self):
        # which points a geometry owns depends on the document only (shapes and
        # grid, not the arrays), so an unchanged document reuses the answer
        key = repr(sorted(self._last_doc.items())) if self.geom_info_requests else None
        cached_key, cache = self._geom_info_cache
        if key != cached_key:
            cache = {}
            self._geom_info_cache = (key, cache)
        for k, req in enumerate(self.geom_info_requests):
            if k not in cache:
                cache[k] = dict(zip(_AXES, self.engine.geom_info(k)))
        return {req.gname: cache[k] for k, req in enumerate(self.geom_info_requests)}

    # -- building up
    def add_geometry(self, geom):
        """Insert `geom` first: it then wins over every geometry given before."""
        self.geometry.insert(0, geom)

    def add_source(self, src):
        self.sources.insert(0, src)

    def request_geom_info(self, geom, field_component="Ex"):
        for i, g in enumerate(self.geometry):
            if g is geom:
                req = Geom_Info(index=i, field_component=field_component)
                self.geom_info_requests.append(req)
                return req
        raise ValueError("the geometry is not part of this simulation")

    def add_dft_fields(self, center=None, size=None, frequency=[], field_component="Ex"):
        mon = Fields_DFT(_vec_or_zero(center), _vec_or_zero(size), frequency, field_component,
                         str(len(self.dft_fields)), degenerate=False, dx=self.dx)
        self.dft_fields.append(mon)
        return mon

    def add_flux_region(self, center=None, size=None, frequency=[], weight=1):
        region = Flux_Region(self, center=_vec_or_zero(center), size=_vec_or_zero(size), frequency=frequency,
                             weight=weight)
        self.flux_regions.append(region)
        return region

    def add_flux_box(self, center=None, size=None, frequency=[]):
        box = Flux_Box(self, center=_vec_or_zero(center), size=_vec_or_zero(size), frequency=frequency)
        self.flux_boxes.append(box)
        return box


# --------------------------------------------------------------- objectives
class _ComplexMatch:
    """F = sum w |E - E0|^2 / 2, conj(dF/dE) = w conj(E - E0)."""

    def __init__(self, target, weight):
        self.target, self.weight = target, weight

    def fun(self, e):
        return np.sum(0.5 * np.ravel(self.weight) * np.abs(np.ravel(e) - np.ravel(self.target)) ** 2)

    def fprime(self, e):
        return (np.conj(e - self.target) * self.weight,)


class _ModulusMatch(_ComplexMatch):
    """F = sum w (|E| - A0)^2 / 2, conj(dF/dE) = w (1 - A0 / |E|) conj(E)."""

    def fun(self, e):
        return np.sum(0.5 * np.ravel(self.weight) * (np.abs(np.ravel(e)) - np.ravel(self.target)) ** 2)

    def fprime(self, e):
        return ((1 - self.target / np.abs(e)) * np.conj(e) * self.weight,)


class _VectorModulusMatch(_ComplexMatch):
    """F = sum w (|E| - A0)^2 / 2 with |E| the length of (Ex, Ey, Ez)."""

    @staticmethod
    def _length(ex, ey, ez):
        return np.sqrt(np.abs(ex) ** 2 + np.abs(ey) ** 2 + np.abs(ez) ** 2)

    def fun(self, ex, ey, ez):
        r = self._length(ex, ey, ez)
        return np.sum(0.5 * (np.ravel(r) - np.ravel(self.target)) ** 2 * np.ravel(self.weight))

    def fprime(self, ex, ey, ez):
        factor = (1 - self.target / self._length(ex, ey, ez)) * self.weight
        return tuple(factor * np.conj(c) for c in (ex, ey, ez))


def field_match_objective(obj_val, component="Ex", weight=1):
    """(components, fprime, fun) of a field-matching objective: `fun` maps the
    components' DFT values to F and `fprime` to conj(dF/dE) per component, the
    adjoint source amplitudes.  `component`: "Ex", "Ey", "Ez" (complex match),
    "|Ex|", "|Ey|", "|Ez|" (modulus match) or "|E|" (vector length)."""
    if component in ("Ex", "Ey", "Ez"):
        m, comps = _ComplexMatch(obj_val, weight), [component]
    elif component in ("|Ex|", "|Ey|", "|Ez|"):
        m, comps = _ModulusMatch(obj_val, weight), [component[1:3]]
    elif component == "|E|":
        m, comps = _VectorModulusMatch(obj_val, weight), ["Ex", "Ey", "Ez"]
    else:
        raise ValueError("objective component %r is not one of Ex, Ey, Ez, |Ex|, |Ey|, |Ez|, |E|" % (component,))
    return comps, m.fprime, m.fun


# ------------------------------------------------------------- adjoint sources
class Adjoint_Flux:
    """Adjoint sources of a flux objective F = sum of the fluxes' values: on
    each flux plane, an E1 source of amplitude conj(H2) w and an H2 source of
    amplitude -conj(E1) w (w the integration weights), on the plane's
    integration lattice."""

    def __init__(self, adjoint_simulation, forward_simulation, function, fluxes=[], frequency=1.0, scale=1):
        self.adjoint_simulation = adjoint_simulation
        self.forward_simulation = forward_simulation
        self.frequency = float(frequency)
        self.scale = float(scale)
        self.fluxes = fluxes
        self.e_adjoint_sources, self.m_adjoint_sources = [], []
        for i, flux in enumerate(fluxes):
            if flux.frequency.size > 1:
                raise ValueError("a flux in an adjoint flux objective must have one frequency, flux %d has %d"
                                 % (i, flux.frequency.size))
            for group, comp in ((self.e_adjoint_sources, flux.e1_str), (self.m_adjoint_sources, flux.h2_str)):
                src = Inhom_Source(function=function, frequency=frequency, amplitude=None, center=flux.center,
                                   size=flux.size, dim=flux.get_len(scale=scale), field_component=comp,
                                   suffix="flux %d %s adjoint" % (i, comp))
                group.append(src)
                adjoint_simulation.add_source(src)

    def eval_functionals_and_set_sources(self):
        """F from the forward run's monitors; sets the adjoint amplitudes."""
        total = 0
        for flux, e_src, m_src in zip(self.fluxes, self.e_adjoint_sources, self.m_adjoint_sources):
            total = total + flux.get_values(scale=self.scale)
            e1, h2 = flux.get_fields(scale=self.scale)
            w = flux.get_int_weights(scale=self.scale)
            e_src.amplitude = np.conj(h2[0]) * w
            m_src.amplitude = -np.conj(e1[0]) * w
        return total


class Adjoint_Source:
    """Objective on the forward DFT fields over a patch (`dim` nodes of a box)
    and the matching adjoint sources.  `functionals`: [component, target]
    pairs (field matches); `udfs`: (components, fprime, fun) triples such as
    ``field_match_objective`` returns."""

    comp_map = {"Ex": ["Ex"], "Ey": ["Ey"], "Ez": ["Ez"], "|Ex|": ["Ex"], "|Ey|": ["Ey"], "|Ez|": ["Ez"],
                "|E|": ["Ex", "Ey", "Ez"]}

    def __init__(self, adjoint_simulation, forward_simulation, function, frequency=1.0, center=None, size=None,
                 dim=None, functionals=[], udfs=[]):
        self.adjoint_simulation = adjoint_simulation
        self.forward_simulation = forward_simulation
        self.frequency = float(frequency)
        self.center, self.size = _vec_or_zero(center), _vec_or_zero(size)
        self.dim = _vec_or_zero(dim).round()
        self.function = deepcopy(function)
        self.functionals = deepcopy(functionals)
        self.udfs = list(udfs)
        self.forward_dfts, self.adjoint_sources, self.forward_fields = {}, {}, {}
        lo = self.center - self.size / 2
        self.x, self.y, self.z = (np.linspace(lo[a], lo[a] + self.size[a], int(self.dim[a])) for a in range(3))
        for comp, target in functionals:
            if not cmp_shape(self.shape, np.shape(target)):
                raise ValueError("objective target of shape %s does not match the patch %s"
                                 % (np.shape(target), self.shape))
            self.udfs.append(field_match_objective(target, comp, weight=1))
        for comps, fprime, fun in self.udfs:
            if not callable(fprime) or not callable(fun):
                raise ValueError("an objective needs callable fun and fprime")
            for comp in comps:
                self.add_component(comp)

    def add_component(self, component="Ex"):
        """A forward monitor and an adjoint source of `component` on the patch."""
        if component in self.forward_dfts:
            return
        self.forward_dfts[component] = self.forward_simulation.add_dft_fields(
            center=self.center, size=self.size, frequency=[self.frequency], field_component=component)
        src = Inhom_Source(function=self.function, frequency=self.frequency, amplitude=None, center=self.center,
                           size=self.size, dim=self.dim, field_component=component, suffix=component + " adjoint")
        self.adjoint_sources[component] = src
        self.adjoint_simulation.add_source(src)

    @property
    def shape(self):
        return (int(self.dim.z), int(self.dim.y), int(self.dim.x))

    @property
    def numel(self):
        return int(self.dim.prod())

    def eval_functionals_and_set_sources(self):
        """F from the forward run's monitors; sets the adjoint amplitudes to
        the sum over objectives of conj(dF/dE)."""
        for comp, mon in self.forward_dfts.items():
            self.forward_fields[comp] = mon(self.frequency, self.z, self.y, self.x).reshape(self.shape)
        amps = {comp: np.zeros(self.shape, dtype=complex) for comp in self.forward_dfts}
        total = 0
        for comps, fprime, fun in self.udfs:
            args = [self.forward_fields[c] for c in comps]
            total = total + fun(*args)
            for c, d in zip(comps, fprime(*args)):
                amps[c] = amps[c] + d
        for comp, a in amps.items():
            self.adjoint_sources[comp].amplitude = a
        return total


class Adjoint_Volume:
    """A two-medium design box shared by a forward and an adjoint simulation,
    E monitors over it in both, and the sensitivity
        dF/deps = Re(2 pi i f E_fwd E_adj eps'(rho) / norm)
    mapped onto the density nodes with the adjoint of trilinear interpolation.
    Only the linear two-medium design (epsilon_fun=None) is provided."""

    _COMPS = ("Ex", "Ey", "Ez")

    def __init__(self, adjoint_simulation, forward_simulation, frequency=1.0, center=None, size=None, dim=None,
                 density=None, medium1=Medium(), medium2=Medium(), epsilon_fun=None, epsilon_der=None, e_sus=[],
                 norm=1):
        if epsilon_fun is not None:
            raise NotImplementedError("only the two-medium design (epsilon_fun=None) is provided")
        self.adjoint_simulation = adjoint_simulation
        self.forward_simulation = forward_simulation
        self.frequency = float(frequency)
        self.center, self.size = _vec_or_zero(center), _vec_or_zero(size)
        self.dim = _vec_or_zero(dim).round()
        self.norm = complex(norm)
        self.medium1, self.medium2 = deepcopy(medium1), deepcopy(medium2)
        e1, e2 = medium1.get_epsilon(self.frequency), medium2.get_epsilon(self.frequency)
        self.epsilon_fun = lambda rho: e1 * rho + (1 - rho) * e2
        self.epsilon_der = lambda rho: e1 - e2
        self.geom = Two_Medium_Box(size=self.size, center=self.center, dim=self.dim, density=density,
                                   medium1=medium1, medium2=medium2, suffix="adjoint")
        for sim in (adjoint_simulation, forward_simulation):
            sim.add_geometry(self.geom)
        self.forward_dfts = [forward_simulation.add_dft_fields(center=self.center, size=self.size,
                                                               frequency=[self.frequency], field_component=c)
                             for c in self._COMPS]
        self.adjoint_dfts = [adjoint_simulation.add_dft_fields(center=self.center, size=self.size,
                                                               frequency=[self.frequency], field_component=c)
                             for c in self._COMPS]
        self.geom_infos = [forward_simulation.request_geom_info(geom=self.geom, field_component=c)
                           for c in self._COMPS]

    def forward_dft(self, field_component="Ex"):
        return self.forward_dfts[self._COMPS.index(field_component)]

    def adjoint_dft(self, field_component="Ex"):
        return self.adjoint_dfts[self._COMPS.index(field_component)]

    def get_sensitivity(self):
        """Sensitivity at the E points the design box owns (their coordinates
        from the geometry-info requests, which the reference truncates to
        integers, src/simulation.cpp:201-211, and which are used as given),
        scattered onto the density nodes."""
        rho_at = RegularGridInterpolator((self.z, self.y, self.x), self.density, bounds_error=True)
        self.rho, self.forward_fields, self.adjoint_fields, self.se, self.se_transposed = [], [], [], [], []
        total = 0
        for fwd, adj, info in zip(self.forward_dfts, self.adjoint_dfts, self.geom_infos):
            pts = np.stack((np.full(info.z.shape, self.frequency), info.z, info.y, info.x), axis=-1)
            rho = rho_at(pts[:, 1:])
            ef = fwd.get_interpolant(method="linear", bounds_error=True)(pts)
            ea = adj.get_interpolant(method="linear", bounds_error=True)(pts)
            se = np.real(2j * pi * self.frequency * ef * ea * self.epsilon_der(rho) / self.norm)
            spread = transpose(self.x, self.y, self.z, np.stack((info.x, info.y, info.z, se), axis=-1))
            for lst, v in ((self.rho, rho), (self.forward_fields, ef), (self.adjoint_fields, ea), (self.se, se),
                           (self.se_transposed, spread)):
                lst.append(v)
            total = total + spread
        return np.reshape(total, self.shape)

    def get_sensitivity_gpu(self, norm=None, factor=1.0):
        """get_sensitivity formed on the GPU from the two simulations' device
        DFT sums (ffip_adjoint_gradient): the E points at their exact positions
        and the transpose as a deterministic gather.  ``norm`` defaults to the
        construction's; ``factor`` multiplies eps' (it replaces 2 pi i f by
        the time-discrete derivative)."""
        from ._lib import adjoint_gradient
        fe, ae = self.forward_simulation.engine, self.adjoint_simulation.engine
        if fe is None or ae is None:
            raise RuntimeError("run the forward and the adjoint simulation first")
        fwd_out = [self.forward_simulation.dft_fields.index(d) for d in self.forward_dfts]
        adj_out = [self.adjoint_simulation.dft_fields.index(d) for d in self.adjoint_dfts]
        eps_der = complex(self.epsilon_der(0.0)) * complex(factor)
        return adjoint_gradient(fe, ae, fwd_out, adj_out, self.frequency, eps_der,
                                complex(self.norm if norm is None else norm), self.center, self.size, self.dim)

    def get_sensitivity2(self):
        """The sensitivity evaluated at the density nodes themselves."""
        pts = metadata_to_pts(self.frequency, self.z, self.y, self.x)
        acc = 0
        for fwd, adj in zip(self.forward_dfts, self.adjoint_dfts):
            ef = fwd.get_interpolant(method="linear", bounds_error=True)(pts)
            ea = adj.get_interpolant(method="linear", bounds_error=True)(pts)
            acc = acc + 2j * pi * self.frequency * ef * ea / self.norm
        return np.real(np.reshape(acc, self.shape) * self.epsilon_der(self.density))

    @property
    def x(self):
        return self.geom.x

    @property
    def y(self):
        return self.geom.y

    @property
    def z(self):
        return self.geom.z

    @property
    def p1(self):
        return self.center - self.size / 2

    @property
    def density(self):
        return self.geom.density

    @density.setter
    def density(self, value):
        self.geom.density = value

    @property
    def numel(self):
        return int(self.dim.prod())

    @property
    def shape(self):
        return self.geom.shape
