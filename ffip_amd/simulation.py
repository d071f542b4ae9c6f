"""Simulation front end (the public names of ffip/simulation.py) on the MI355X
engine.

The reference's ``Simulation.run`` writes config.json and an input HDF5 file,
runs ``mpirun -np N run_sim_json`` and reads the DFT groups back from the output
HDF5 file (ffip/simulation.py:404-431).  Here the same JSON document and the same
datasets go in-process to libffip_hip.so through its C ABI, and results are read
back into the same objects.  ``np`` and ``decomposition`` are accepted and
ignored: one GPU holds the whole grid.
"""
from __future__ import annotations

import json
from copy import copy, deepcopy
from math import pi

import numpy as np
from scipy.interpolate import RegularGridInterpolator

from ._lib import LibSim, transpose
from .geom import Medium, Two_Medium_Box, Vector3, cmp_shape
from .source import Inhom_Source

direction_dict = {"x": 0, "y": 1, "z": 2}
side_dict = {"positive": 1, "negative": -1}


class _DatasetStore:
    """Stands in for the input HDF5 file: objects' dump() writes named arrays."""

    def __init__(self):
        self.data = {}

    def create_dataset(self, name, data):
        self.data[name] = np.asarray(data, dtype=np.float64).ravel()


class Fields_DFT:
    """One DFT output group (ffip/simulation.py:17-91)."""

    def __init__(self, center, size, frequency, field_component, suffix="0", degenerate=False, dx=0):
        self.suffix = str(suffix)
        self.group_name = "volume fields dft %s" % suffix
        self.center = center.copy()
        self.size = size.copy()
        self.degenerate = bool(degenerate)
        self.dx = float(dx)
        if not self.degenerate and dx == 0:
            raise ValueError("for non degenerate dft, dx must be provided")
        self.frequency = np.sort(np.asarray(frequency, dtype=float))
        self.frequency_config = self.frequency.tolist()
        self.field_component = field_component

    def get_json(self):
        return {"type": "volume fields dft", "center": self.center.get_json(), "size": self.size.get_json(),
                "frequency": self.frequency_config, "field component": self.field_component,
                "output group": self.group_name}

    @property
    def values(self):
        return self.v

    def read(self, groups):
        g = groups[self.group_name]
        self.frequency = np.array(g["f"])
        self.x, self.y, self.z = np.array(g["x"]), np.array(g["y"]), np.array(g["z"])
        self.v = g["real"] + 1j * g["imag"]
        if not self.degenerate:
            # singleton axes are widened so that linear interpolation works on them
            if self.frequency.size == 1:
                self.frequency = self.frequency * np.array([0.9, 1.1])
            if self.x.size == 1:
                self.x = self.x + np.array([-self.dx / 2, self.dx / 2])
            if self.y.size == 1:
                self.y = self.y + np.array([-self.dx / 2, self.dx / 2])
            if self.z.size == 1:
                self.z = self.z + np.array([-self.dx / 2, self.dx / 2])
            for i in range(4):
                if self.v.shape[i] == 1:
                    self.v = np.concatenate((self.v, self.v), axis=i)

    def __call__(self, f, z, y, x, substract=None, **kwargs):
        interp = self.get_interpolant(substract, **kwargs)
        return interp(np.stack(np.meshgrid(f, z, y, x, indexing="ij"), axis=-1))

    def get_interpolant(self, substract=None, **kwargs):
        if substract is not None and not isinstance(substract, Fields_DFT):
            raise TypeError("Substract is not an instance of Fields_DFT")
        vals = self.values - substract.values if substract is not None else self.values
        return RegularGridInterpolator((self.frequency, self.z, self.y, self.x), vals, **kwargs)

    def get_metadata(self):
        return self.frequency, self.z, self.y, self.x


class Flux_Region:
    """Poynting flux through a plane from an E and an H DFT (ffip/simulation.py:94-210)."""

    def __init__(self, sim, center, size, frequency, weight):
        dim = (size.x != 0) + (size.y != 0) + (size.z != 0)
        if dim != 2:
            raise NameError("Flux region dimension %d is invalid" % dim)
        self.direction = 0 if size.x == 0 else 1 if size.y == 0 else 2
        self.weight = np.array(weight, copy=True)
        self.center = center.copy()
        self.size = size.copy()
        self.frequency = np.sort(np.asarray(frequency))
        self.dx = sim.dx
        self._e1_dft = sim.add_dft_fields(center=center, size=size, frequency=frequency, field_component=self.e1_str)
        self._h2_dft = sim.add_dft_fields(center=center, size=size, frequency=frequency, field_component=self.h2_str)

    @property
    def e1_str(self):
        return "E" + "xyz"[(self.direction + 1) % 3]

    @property
    def h2_str(self):
        return "H" + "xyz"[(self.direction + 2) % 3]

    @property
    def e1_dft(self):
        return self._e1_dft

    @property
    def h2_dft(self):
        return self._h2_dft

    def get_fields(self, scale=1, substract=None, **kwargs):
        if substract is not None and not isinstance(substract, Flux_Region):
            raise TypeError("Substract is not an instance of Flux_Region")
        e1 = self.e1_dft.get_interpolant(substract=substract.e1_dft if substract is not None else None, **kwargs)
        h2 = self.h2_dft.get_interpolant(substract=substract.h2_dft if substract is not None else None, **kwargs)
        pts = self.get_pts(scale=scale)
        return e1(pts), h2(pts)

    def get_len(self, scale=1):
        return (self.size / self.dx * scale).round() + Vector3(1, 1, 1)

    def get_metadata(self, scale=1):
        n = self.get_len(scale=scale)
        p1 = self.center - self.size / 2
        p2 = self.center + self.size / 2
        return (self.frequency, np.linspace(p1.z, p2.z, int(n.z)), np.linspace(p1.y, p2.y, int(n.y)),
                np.linspace(p1.x, p2.x, int(n.x)))

    def get_pts(self, scale=1):
        return np.stack(np.meshgrid(*self.get_metadata(scale=scale), indexing="ij"), axis=-1)

    def get_int_weights(self, scale=1):
        n = self.get_len(scale=scale)
        ws = [np.ones((int(n.x),)), np.ones((int(n.y),)), np.ones((int(n.z),))]
        meas = 1
        for a, w in zip("xyz", ws):
            if w.size > 1:
                meas *= getattr(self.size, a) / (w.size - 1)
                w[0] = w[-1] = 0.5
        Z, Y, X = np.meshgrid(ws[2], ws[1], ws[0], indexing="ij")
        return X * Y * Z * meas * self.weight

    def get_values(self, scale=1, substract=None, **kwargs):
        e1, h2 = self.get_fields(scale=scale, substract=substract, **kwargs)
        flux = np.real(e1 * np.conj(h2)) * self.get_int_weights(scale=scale)
        return np.sum(flux, axis=(1, 2, 3))

    def get_interpolant(self, scale=1, substract=None):
        return RegularGridInterpolator((self.frequency,), self.get_values(scale=scale, substract=substract))


class Flux_Box:
    """Six signed flux planes around a box (ffip/simulation.py:212-250)."""

    def __init__(self, sim, center, size, frequency):
        self._flux_regions = []
        self.frequency = np.sort(np.asarray(frequency))
        for d in range(3):
            for side in (-1, 1):
                p1 = center - size / 2
                p2 = center + size / 2
                if side == -1:
                    p2[d] = p1[d]
                else:
                    p1[d] = p2[d]
                self._flux_regions.append(sim.add_flux_region(center=(p1 + p2) / 2, size=p2 - p1,
                                                              frequency=frequency, weight=side))

    @property
    def flux_regions(self):
        return self._flux_regions

    def get_flux_region(self, direction="x", side="negative"):
        return self.flux_regions[direction_dict[direction] * 2 + (0 if side_dict[side] == -1 else 1)]

    def get_values(self, scale=1, substract=None, **kwargs):
        if substract is not None and not isinstance(substract, Flux_Box):
            raise TypeError("Substract is not an instance of Flux_Box")
        return sum(self.flux_regions[i].get_values(
            scale=scale, substract=substract.flux_regions[i] if substract is not None else None, **kwargs)
            for i in range(6))

    def get_interpolant(self, scale=1, substract=None, **kwargs):
        return RegularGridInterpolator((self.frequency,), self.get_values(scale=scale, substract=substract, **kwargs))


class Geom_Info:
    """Coordinates of the points of a component inside one geometry
    (ffip/simulation.py:253-283, src/simulation.cpp:190-233)."""

    def __init__(self, index=0, field_component="Ex"):
        self.field_component = str(field_component)
        self.index = int(index)
        self.gname = "geometry info request %d %s" % (self.index, self.field_component)

    def read(self, groups):
        g = groups[self.gname]
        self._x, self._y, self._z = g["x"], g["y"], g["z"]

    def get_json(self):
        return {"field component": self.field_component, "geometry index": self.index, "output group": self.gname}

    @property
    def x(self):
        return self._x

    @property
    def y(self):
        return self._y

    @property
    def z(self):
        return self._z


class Simulation:
    """A simulation set-up; ``run`` executes it on the GPU (ffip/simulation.py:286-464).

    Extra keyword ``precision``: "fp64" (default, the reference's arithmetic) or
    "fp32" (fields and polarizations stored in FP32, FP64 arithmetic and DFT sums;
    E/H monitors gather 16 steps in FP32 partial sums before adding into them).
    """

    def __init__(self, size, resolution, pmls=[], geometry=[], sources=[], symmetry=[], default_material=Medium(),
                 periodic=False, progress_interval=4, courant=0.5, input_file="empty.h5",
                 fields_output_file="output.h5", precision="fp64"):
        print("Simulation created\n")
        self.size = size.copy()
        self.resolution = float(resolution)
        self.geometry = deepcopy(geometry)
        self.sources = deepcopy(sources)
        self.symmetry = deepcopy(symmetry)
        self.default_material = deepcopy(default_material)
        self.periodic = deepcopy(periodic)
        self.progress_interval = float(progress_interval)
        self.courant = float(courant)
        self.pmls = deepcopy(pmls)
        self.input_file = copy(input_file)
        self.fields_output_file = copy(fields_output_file)
        self.precision = precision
        self.dft_fields = []
        self.flux_regions = []
        self.flux_boxes = []
        self.geom_info_requests = []
        self.engine = None
        self._geom_key, self._geom_cache = None, {}
        for p in self.pmls:
            if p.sigma_max is None:
                p.sigma_max = 0.8 * (1 + p.order) * resolution

    @property
    def input_file(self):
        return self._input_file

    @input_file.setter
    def input_file(self, val):
        if not isinstance(val, str):
            raise ValueError("file name is not a string")
        self._input_file = val

    @property
    def fields_output_file(self):
        return self._fields_output_file

    @fields_output_file.setter
    def fields_output_file(self, val):
        if not isinstance(val, str):
            raise ValueError("file name is not a string")
        self._fields_output_file = val

    @property
    def dx(self):
        return 1 / self.resolution

    def dump(self, store):
        for g in self.geometry or []:
            g.dump(store)
        for s in self.sources or []:
            s.dump(store)

    def get_json(self, stop_condition, decomposition=None):
        res = {"courant number": self.courant, "resolution": self.resolution, "size": self.size.get_json(),
               "default material": self.default_material.get_json(), "progress interval": self.progress_interval,
               "input file": self.input_file, "fields output file": self.fields_output_file,
               "stop condition": stop_condition.get_json(), "precision": self.precision}
        if decomposition is not None:
            res["decomposition"] = decomposition
        if self.symmetry:
            # the reference serialises self.periodic here (ffip/simulation.py:380),
            # which fails for any symmetry list; the symmetry objects are meant
            res["symmetry"] = [item.get_json() for item in self.symmetry]
        if self.pmls:
            res["PML"] = [item.get_json() for item in self.pmls]
        if self.periodic:
            res["periodic"] = self.periodic
        if self.geometry:
            res["geometry"] = [item.get_json() for item in self.geometry]
        if self.sources:
            res["sources"] = [item.get_json() for item in self.sources]
        if self.dft_fields:
            res["fields output"] = [item.get_json() for item in self.dft_fields]
        if self.geom_info_requests:
            res["geometry info request"] = [item.get_json() for item in self.geom_info_requests]
        return res

    def run(self, stop_condition, decomposition=None, np=1, skip=False, pop=False, **kwargs):
        """Run to the stop condition and read the DFT and geometry-info results
        into the registered objects.  ``skip=True`` re-reads the previous run."""
        if not skip or self.engine is None:
            config = self.get_json(stop_condition=stop_condition, decomposition=decomposition)
            store = _DatasetStore()
            self.dump(store)
            if self.engine is not None:
                # the previous run's results have been read; releasing its device
                # memory first lets the new engine reuse the blocks
                self.engine.close()
                self.engine = None
            self.engine = LibSim(config, store.data)
            self.engine.run(config["stop condition"])
            key = json.dumps(config, sort_keys=True)
            if key != self._geom_key:
                self._geom_key, self._geom_cache = key, {}
        groups = {}
        for k in range(self.engine.dft_count()):
            d = self.engine.dft(k)
            groups[d["group"]] = d
        for k, req in enumerate(self.geom_info_requests):
            # which points a geometry owns depends on the document only (shapes and
            # grid, not the datasets), so an unchanged document reuses them
            if k not in self._geom_cache:
                self._geom_cache[k] = self.engine.geom_info(k)
            x, y, z = self._geom_cache[k]
            groups[req.gname] = {"x": x, "y": y, "z": z}
        for item in self.dft_fields:
            item.read(groups)
        for item in self.geom_info_requests:
            item.read(groups)

    def add_geometry(self, geom):
        self.geometry.insert(0, geom)

    def add_source(self, src):
        self.sources.insert(0, src)

    def request_geom_info(self, geom, field_component="Ex"):
        for i in range(len(self.geometry)):
            if self.geometry[i] is geom:
                req = Geom_Info(index=i, field_component=field_component)
                self.geom_info_requests.append(req)
                return req
        raise ValueError("Geometry non-existed")

    def add_dft_fields(self, center=Vector3(), size=Vector3(), frequency=[], field_component="Ex"):
        res = Fields_DFT(center, size, frequency, field_component, str(len(self.dft_fields)), degenerate=False,
                         dx=self.dx)
        self.dft_fields.append(res)
        return res

    def add_flux_region(self, center=Vector3(), size=Vector3(), frequency=[], weight=1):
        res = Flux_Region(self, center=center, size=size, frequency=frequency, weight=weight)
        self.flux_regions.append(res)
        return res

    def add_flux_box(self, center=Vector3(), size=Vector3(), frequency=[]):
        res = Flux_Box(self, center=center, size=size, frequency=frequency)
        self.flux_boxes.append(res)
        return res


class Symmetry:
    def __init__(self, direction="x", phase_factor=1):
        self.direction = direction
        self.phase_factor = int(phase_factor)

    def get_json(self):
        return {"direction": self.direction, "phase factor": self.phase_factor}


class run_until_time:
    def __init__(self, time=0):
        self.time = float(time)

    def get_json(self):
        return {"type": "time", "time": self.time}


class run_until_dft:
    def __init__(self, center=Vector3(), size=Vector3(), frequency=1, field_component="Ex",
                 time_interval_examined=1, var=1e-2):
        self.center, self.size = center.copy(), size.copy()
        self.field_component = str(field_component)
        self.time_interval_examined = float(time_interval_examined)
        self.var = float(var)
        self.frequency = float(frequency)

    def get_json(self):
        return {"type": "dft", "center": self.center.get_json(), "size": self.size.get_json(),
                "field component": self.field_component, "time interval examined": self.time_interval_examined,
                "variation": self.var, "frequency": self.frequency}


class run_until_fields_decay:
    def __init__(self, position=Vector3(), field_component="Ez", time_interval_examined=1, decayed_by=1e-3):
        self.position = position.copy()
        self.field_component = str(field_component)
        self.time_interval_examined = float(time_interval_examined)
        self.decayed_by = float(decayed_by)

    def get_json(self):
        return {"type": "decay", "position": self.position.get_json(), "field component": self.field_component,
                "time interval examined": self.time_interval_examined, "decayed by": self.decayed_by}


class PML:
    """CPML layer.  As in the reference (ffip/simulation.py:519-542) the sigma_max
    argument is not kept: Simulation sets 0.8 (order + 1) resolution."""

    def __init__(self, thickness, direction=None, side=None, k_max=1, order=3, sigma_max=None):
        self.thickness = float(thickness)
        self.direction, self.side = direction, side
        self.k_max = float(k_max)
        self.order = int(order)
        self.sigma_max = None

    def get_json(self):
        res = {"thickness": self.thickness}
        if self.direction is not None:
            res["direction"] = self.direction
        if self.side is not None:
            res["side"] = self.side
        res["k max"] = self.k_max if self.k_max is not None else 1.0
        res["order"] = self.order if self.order is not None else 3
        res["sigma max"] = self.sigma_max
        return res


def field_match_objective(obj_val, component="Ex", weight=1):
    """(components, conj(dF/dE) function, F function) of a field-matching
    objective (ffip/simulation.py:544-582)."""
    if component in ["Ex", "Ey", "Ez"]:
        def fun1(val):
            return np.sum(0.5 * np.ravel(weight) * np.abs(val.ravel() - obj_val.ravel()) ** 2)

        def fpconj1(val):
            return (np.conj(val - obj_val) * weight,)

        return [component], fpconj1, fun1
    if component in ["|Ex|", "|Ey|", "|Ez|"]:
        def fun2(val):
            return np.sum(0.5 * np.ravel(weight) * (np.abs(val).ravel() - obj_val.ravel()) ** 2)

        def fpconj2(val):
            return ((1 - obj_val / np.abs(val)) * np.conj(val) * weight,)

        return [component[1:-1]], fpconj2, fun2
    if component == "|E|":
        def fun3(ex, ey, ez):
            val = np.sqrt(np.abs(ex) ** 2 + np.abs(ey) ** 2 + np.abs(ez) ** 2)
            return np.sum(0.5 * (np.ravel(val) - np.ravel(obj_val)) ** 2 * np.ravel(weight))

        def fpconj3(ex, ey, ez):
            val = np.sqrt(np.abs(ex) ** 2 + np.abs(ey) ** 2 + np.abs(ez) ** 2)
            r = (1 - obj_val / val) * weight
            return (r * np.conj(ex), r * np.conj(ey), r * np.conj(ez))

        return ["Ex", "Ey", "Ez"], fpconj3, fun3
    raise ValueError("unknown objective component %r" % component)


class Adjoint_Flux:
    """Adjoint sources of a flux objective (ffip/simulation.py:585-645)."""

    def __init__(self, adjoint_simulation, forward_simulation, function, fluxes=[], frequency=1.0, scale=1):
        self.adjoint_simulation = adjoint_simulation
        self.forward_simulation = forward_simulation
        self.frequency = float(frequency)
        self.scale = float(scale)
        self.fluxes = fluxes
        self.e_adjoint_sources, self.m_adjoint_sources = [], []
        for i, flux in enumerate(fluxes):
            if flux.frequency.size > 1:
                raise ValueError("Fluxes used in Adjoint_Flux cannot have more than 1 frequency")
            for lst, comp in ((self.e_adjoint_sources, flux.e1_str), (self.m_adjoint_sources, flux.h2_str)):
                lst.append(Inhom_Source(function=function, frequency=frequency, amplitude=None, center=flux.center,
                                        size=flux.size, dim=flux.get_len(scale=scale), field_component=comp,
                                        suffix="flux %d %s adjoint" % (i, comp)))
            adjoint_simulation.add_source(self.e_adjoint_sources[-1])
            adjoint_simulation.add_source(self.m_adjoint_sources[-1])

    def eval_functionals_and_set_sources(self):
        res = 0
        for i, flux in enumerate(self.fluxes):
            res += flux.get_values(scale=self.scale)
            e1, h2 = flux.get_fields(scale=self.scale)
            w = flux.get_int_weights(scale=self.scale)
            self.e_adjoint_sources[i].amplitude = np.conj(h2[0, ...]) * w
            self.m_adjoint_sources[i].amplitude = -np.conj(e1[0, ...]) * w
        return res


class Adjoint_Source:
    """Objective on a patch of forward DFT fields and the matching adjoint
    sources (ffip/simulation.py:648-757)."""

    comp_map = {"Ex": ["Ex"], "Ey": ["Ey"], "Ez": ["Ez"], "|Ex|": ["Ex"], "|Ey|": ["Ey"], "|Ez|": ["Ez"],
                "|E|": ["Ex", "Ey", "Ez"]}

    def __init__(self, adjoint_simulation, forward_simulation, function, frequency=1.0, center=Vector3(),
                 size=Vector3(), dim=Vector3(), functionals=[], udfs=[]):
        self.adjoint_simulation = adjoint_simulation
        self.forward_simulation = forward_simulation
        self.frequency = float(frequency)
        self.center, self.size = center.copy(), size.copy()
        self.dim = dim.round()
        self.functionals = deepcopy(functionals)
        self.udfs = list(udfs)
        self.function = deepcopy(function)
        self.forward_dfts, self.adjoint_sources, self.forward_fields = {}, {}, {}
        p1 = self.center - self.size / 2
        self.x = np.linspace(p1.x, p1.x + self.size.x, int(dim.x))
        self.y = np.linspace(p1.y, p1.y + self.size.y, int(dim.y))
        self.z = np.linspace(p1.z, p1.z + self.size.z, int(dim.z))
        for func in functionals:
            if not cmp_shape(self.shape, func[1].shape):
                raise ValueError("objective values does not match the shape")
            self.udfs.append(field_match_objective(func[1], func[0], weight=1))
        for components, fprime, fun in self.udfs:
            if not callable(fprime):
                raise ValueError("The derivative function is not callable")
            if not callable(fun):
                raise ValueError("The objective function is not callable")
            for c in components:
                self.add_component(c)

    def add_component(self, component="Ex"):
        if component not in self.forward_dfts:
            self.forward_dfts[component] = self.forward_simulation.add_dft_fields(
                center=self.center, size=self.size, frequency=[self.frequency], field_component=component)
            self.adjoint_sources[component] = Inhom_Source(
                function=self.function, frequency=self.frequency, amplitude=None, center=self.center, size=self.size,
                dim=self.dim, field_component=component, suffix=component + " adjoint")
            self.adjoint_simulation.add_source(self.adjoint_sources[component])

    @property
    def shape(self):
        return (int(self.dim.z), int(self.dim.y), int(self.dim.x))

    def eval_functionals_and_set_sources(self):
        for c in self.forward_dfts:
            self.adjoint_sources[c].amplitude = np.zeros(self.shape, dtype=complex)
            self.forward_fields[c] = self.forward_dfts[c](self.frequency, self.z, self.y, self.x).reshape(self.shape)
        res = 0
        for components, fprime, fun in self.udfs:
            args = [self.forward_fields[c] for c in components]
            res = res + fun(*args)
            der = fprime(*args)
            for i, c in enumerate(components):
                self.adjoint_sources[c].amplitude += der[i]
        return res

    @property
    def numel(self):
        return int(self.dim.prod())


class Adjoint_Volume:
    """Design volume: forward/adjoint DFTs over it and the sensitivity
    Re(2 pi i f E_fwd E_adj eps'(rho) / norm) mapped onto the density nodes
    (ffip/simulation.py:759-915).  Only the two-medium (linear) design is
    provided."""

    def __init__(self, adjoint_simulation, forward_simulation, frequency=1.0, center=Vector3(), size=Vector3(),
                 dim=Vector3(), density=None, medium1=Medium(), medium2=Medium(), epsilon_fun=None,
                 epsilon_der=None, e_sus=[], norm=1):
        if epsilon_fun is not None:
            raise NotImplementedError("only the two-medium design (epsilon_fun=None) is supported")
        self.adjoint_simulation = adjoint_simulation
        self.forward_simulation = forward_simulation
        self.frequency = float(frequency)
        self.center, self.size = center.copy(), size.copy()
        self.dim = dim.round()
        self.norm = complex(norm)
        e1 = medium1.get_epsilon(self.frequency)
        e2 = medium2.get_epsilon(self.frequency)
        self.epsilon_fun = lambda rho: e1 * rho + (1 - rho) * e2
        self.epsilon_der = lambda rho: e1 - e2
        self.medium1, self.medium2 = deepcopy(medium1), deepcopy(medium2)
        self.geom = Two_Medium_Box(size=size, center=center, dim=dim, density=density, medium1=medium1,
                                   medium2=medium2, suffix="adjoint")
        adjoint_simulation.add_geometry(self.geom)
        forward_simulation.add_geometry(self.geom)
        comps = ["Ex", "Ey", "Ez"]
        self.forward_dfts = [forward_simulation.add_dft_fields(center=center, size=size, frequency=[frequency],
                                                               field_component=c) for c in comps]
        self.adjoint_dfts = [adjoint_simulation.add_dft_fields(center=center, size=size, frequency=[frequency],
                                                               field_component=c) for c in comps]
        self.geom_infos = [forward_simulation.request_geom_info(geom=self.geom, field_component=c) for c in comps]

    def forward_dft(self, field_component="Ex"):
        return self.forward_dfts["xyz".index(field_component[1])]

    def adjoint_dft(self, field_component="Ex"):
        return self.adjoint_dfts["xyz".index(field_component[1])]

    def get_sensitivity(self):
        """Sensitivity sampled at the E points inside the design box and
        scattered onto the density nodes with the adjoint of trilinear
        interpolation.  The points are those of the geometry-info requests,
        whose coordinates the reference truncates to integers
        (src/simulation.cpp:201-211); they are used as given."""
        pts = [np.stack((self.frequency * np.ones(g.z.shape), g.z, g.y, g.x), axis=-1) for g in self.geom_infos]
        rho_interp = RegularGridInterpolator((self.z, self.y, self.x), self.density, bounds_error=True)
        self.rho = [rho_interp(np.stack((g.z, g.y, g.x), axis=-1)) for g in self.geom_infos]
        self.forward_fields = [self.forward_dfts[i].get_interpolant(method="linear", bounds_error=True)(pts[i])
                               for i in range(3)]
        self.adjoint_fields = [self.adjoint_dfts[i].get_interpolant(method="linear", bounds_error=True)(pts[i])
                               for i in range(3)]
        self.se, self.se_transposed = [], []
        res = 0
        for i in range(3):
            self.se.append(np.real(2j * pi * self.frequency * self.forward_fields[i] * self.adjoint_fields[i] *
                                   self.epsilon_der(self.rho[i]) / self.norm))
            g = self.geom_infos[i]
            self.se_transposed.append(transpose(self.x, self.y, self.z, np.stack((g.x, g.y, g.z, self.se[i]), axis=-1)))
            res = res + self.se_transposed[i]
        return np.reshape(res, self.shape)

    def get_sensitivity_gpu(self, norm=None, factor=1.0):
        """get_sensitivity computed on the GPU from the two simulations' device
        DFT sums (ffip_adjoint_gradient): the E points are taken at their exact
        positions and the transpose is a deterministic gather.  ``norm`` defaults
        to the norm given at construction; ``factor`` multiplies eps' (used to
        replace 2 pi i f by the time-discrete derivative factor)."""
        from ._lib import adjoint_gradient
        fe, ae = self.forward_simulation.engine, self.adjoint_simulation.engine
        if fe is None or ae is None:
            raise RuntimeError("run the forward and adjoint simulations first")
        fwd_out = [self.forward_simulation.dft_fields.index(d) for d in self.forward_dfts]
        adj_out = [self.adjoint_simulation.dft_fields.index(d) for d in self.adjoint_dfts]
        eps_der = complex(self.epsilon_der(0.0)) * complex(factor)
        return adjoint_gradient(fe, ae, fwd_out, adj_out, self.frequency, eps_der,
                                complex(self.norm if norm is None else norm), self.center, self.size, self.dim)

    def get_sensitivity2(self):
        pts = np.stack(np.meshgrid(self.frequency, self.z, self.y, self.x, indexing="ij"), axis=-1)
        res = 0
        for i in range(3):
            f = self.forward_dfts[i].get_interpolant(method="linear", bounds_error=True)(pts)
            a = self.adjoint_dfts[i].get_interpolant(method="linear", bounds_error=True)(pts)
            res = res + 2j * pi * self.frequency * f * a / self.norm
        return np.real(np.reshape(res, self.shape) * self.epsilon_der(self.density))

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
    def density(self, val):
        self.geom.density = val

    @property
    def numel(self):
        return int(self.dim.prod())

    @property
    def shape(self):
        return self.geom.shape
