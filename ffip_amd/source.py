"""Excitations of the front end.

The names, constructor arguments, attributes, JSON keys and dataset names are
those of ffip/source.py, which scripts use; the JSON is what
Simulation::set_volume_source / set_inhom_source read
(src/simulation.cpp:235-336).  An inhomogeneous source is a sampled array on a
box and shares the node layout of the density boxes (``geom.NodeGrid``).
"""
from __future__ import annotations

from copy import deepcopy

import numpy as np

from .geom import NodeGrid, Vector3

__all__ = ["Gaussian1", "Inhom_Source", "Source", "Source_Function"]


class Source_Function:
    """A pulse active on (start_time, end_time)."""

    start_time = 0.0
    end_time = 0.0

    @property
    def duration(self):
        return self.end_time - self.start_time


class Gaussian1(Source_Function):
    """First derivative of a Gaussian, g(t) = u exp(-u^2 / 2) with
    u = (t - t_c) / width, width = 1 / (2 pi frequency), active for
    2 cutoff widths from `start_time` and centred in that window
    (src/utility.cpp:260-264, src/fields.cpp:17-33)."""

    def __init__(self, frequency, cutoff=4, start_time=0):
        self.frequency = float(frequency)
        self.cutoff = float(cutoff)
        self.start_time = float(start_time)
        # evaluated in the engine's operation order, so that the window ends
        # agree to the last bit
        self.width = 1 / (2 * np.pi * self.frequency)
        self.end_time = self.start_time + self.width * self.cutoff * 2
        self.offset_time = (self.start_time + self.end_time) / 2

    def __call__(self, t):
        u = (np.asarray(t) - self.offset_time) / self.width
        return u * np.exp(-0.5 * u * u)

    def get_json(self):
        return {"type": "Gaussian1", "frequency": self.frequency, "cutoff": self.cutoff,
                "start time": self.start_time}


def _field_name(component):
    name = str(component)
    if len(name) != 2 or name[0] not in "EH" or name[1] not in "xyz":
        raise ValueError("field component must be one of Ex, Ey, Ez, Hx, Hy, Hz, not %r" % (component,))
    return name


class Source:
    """Uniform current density `amplitude` over a box (a point, line, sheet or
    volume), sampled by one dipole per cell along each extended axis."""

    def __init__(self, function, center=None, size=None, amplitude=1, field_component="Ex"):
        self.function = deepcopy(function)
        self.center = center.copy() if center is not None else Vector3()
        self.size = size.copy() if size is not None else Vector3()
        self.amplitude = float(amplitude)
        self.field_component = _field_name(field_component)

    def dump(self, store):
        """A uniform source carries no arrays."""

    def get_json(self):
        return {"type": "volume source", "center": self.center.get_json(), "size": self.size.get_json(),
                "amplitude": self.amplitude, "field component": self.field_component,
                "function": self.function.get_json()}


class Inhom_Source(NodeGrid):
    """Dipoles on the `dim` nodes of a box with complex amplitudes a: each
    dipole's pulse is scaled by |a| and delayed by arg(a) / (2 pi frequency)
    (src/simulation.cpp:316-333).  The amplitudes are given like a density (an
    array [z, y, x], a function of the node points, or None for zeros) and go to
    the datasets "inhom source amplitude <suffix>" and
    "inhom source phase <suffix>"."""

    _grid_array = "amplitude"
    _grid_dtype = complex

    def __init__(self, function, amplitude=None, frequency=1, center=None, size=None, dim=None,
                 field_component="Ex", suffix="0"):
        self._place(center or Vector3(), size or Vector3(), dim or Vector3())
        self.function = deepcopy(function)
        self.frequency = float(frequency)
        self.field_component = _field_name(field_component)
        self.amplitude = self._initial(amplitude)
        self.amplitude_dataset_name = "inhom source amplitude %s" % suffix
        self.phase_dataset_name = "inhom source phase %s" % suffix

    @property
    def amplitude(self):
        return self._amp

    @amplitude.setter
    def amplitude(self, value):
        self._amp = self._checked(value)

    def dump(self, store):
        a = self.amplitude.ravel()
        store.create_dataset(self.amplitude_dataset_name, data=np.abs(a))
        store.create_dataset(self.phase_dataset_name, data=np.angle(a))

    def get_json(self):
        return {"type": "inhomogeneous source", "center": self.center.get_json(), "size": self.size.get_json(),
                "dimension": self.dim.get_json(), "field component": self.field_component,
                "amplitude dataset": self.amplitude_dataset_name, "phase dataset": self.phase_dataset_name,
                "frequency": self.frequency, "function": self.function.get_json()}
