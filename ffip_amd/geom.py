"""Geometry, media and grid helpers of the front end.

Scripts written against the reference's Python package reach these objects by
the names of ffip/geom.py.  What they fix, and what is kept here, is the
interface: class and function names, constructor arguments and defaults, the
attributes scripts read, the JSON each object writes for Simulation::init
(src/simulation.cpp:58-100, 402-518), the dataset names of the input store, and
the closed forms of the material responses.  The implementation is this
package's own:

* a vector is three slots combined component-wise by one helper;
* a susceptibility is the rational response sigma N(iw) / D(iw) with its
  numerator and denominator coefficients, plus the recurrence coefficients the
  engine's update uses, from which the time-discrete response follows;
* every object that carries a sampled array over a box (density boxes here,
  inhomogeneous sources in ``source``) shares ``NodeGrid``: the node axes, the
  array's shape and its validation;
* JSON documents are built from per-class key tables.
"""
from __future__ import annotations

import math
import operator
from copy import deepcopy
from numbers import Number

import numpy as np

__all__ = ["Box", "DeybeSusceptibility", "DrudeSusceptibility", "General_Medium_Box", "GeometricObject",
           "LorentzianSusceptibility", "Medium", "NodeGrid", "Sphere", "Susceptibility", "Two_Medium_Box",
           "Vector3", "cmp_shape", "getgrid", "metadata_to_pts", "planar_polygon", "sph2cart"]


# ---------------------------------------------------------------- small helpers
def cmp_shape(shape1=(), shape2=()):
    """Whether two shapes agree in rank and in every extent."""
    a, b = tuple(shape1), tuple(shape2)
    return len(a) == len(b) and all(p == q for p, q in zip(a, b))


def sph2cart(az, el, r):
    """Cartesian (x, y, z) of the point at azimuth `az`, elevation `el`, radius `r`."""
    rho = r * np.cos(el)  # distance from the z axis
    return rho * np.cos(az), rho * np.sin(az), r * np.sin(el)


def metadata_to_pts(*axes):
    """Points of the ij-indexed grid spanned by `axes`, coordinates on the last axis."""
    return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)


def _axis(center, width, count):
    """`count` equally spaced nodes from center - width/2 to center + width/2."""
    return np.linspace(center - width / 2, center + width / 2, int(count))


def getgrid(center=None, size=None, dim=None):
    """Node axes (z, y, x) of a box sampled at `dim` nodes per axis, end points
    included (the layout of every density and amplitude array)."""
    center = Vector3() if center is None else center
    size = Vector3() if size is None else size
    dim = Vector3() if dim is None else dim
    return tuple(_axis(center[a], size[a], dim[a]) for a in (2, 1, 0))


# ---------------------------------------------------------------------- vectors
class Vector3:
    """A 3-vector of (usually real) numbers with the arithmetic scripts use.

    ``+`` and ``-`` broadcast a number over the components; ``*`` is the dot
    product with a vector and scaling by a number; ``/`` divides component-wise
    by a vector or by a number.  Integer arguments are stored as floats."""

    __slots__ = ("x", "y", "z")

    def __init__(self, x=0.0, y=0.0, z=0.0):
        self.x, self.y, self.z = (float(v) if type(v) is int else v for v in (x, y, z))

    # -- component access
    def _parts(self):
        return (self.x, self.y, self.z)

    def __getitem__(self, i):
        if i not in (0, 1, 2):
            raise IndexError("Vector3 has components 0, 1 and 2, not %r" % (i,))
        return self._parts()[i]

    def __setitem__(self, i, value):
        if i not in (0, 1, 2):
            raise IndexError("Vector3 has components 0, 1 and 2, not %r" % (i,))
        setattr(self, "xyz"[i], value)

    def __iter__(self):
        return iter(self._parts())

    def __array__(self, dtype=None, copy=None):
        return np.array(self._parts(), dtype=dtype)

    def __repr__(self):
        return "Vector3(%r, %r, %r)" % self._parts()

    def __eq__(self, other):
        return isinstance(other, Vector3) and self._parts() == other._parts()

    __hash__ = None

    # -- arithmetic
    def _combine(self, other, op, what):
        if isinstance(other, Number):
            return Vector3(*(op(a, other) for a in self._parts()))
        if isinstance(other, Vector3):
            return Vector3(*(op(a, b) for a, b in zip(self._parts(), other._parts())))
        raise TypeError("cannot %s a Vector3 and a %s" % (what, type(other).__name__))

    def __add__(self, other):
        return self._combine(other, operator.add, "add")

    def __sub__(self, other):
        return self._combine(other, operator.sub, "subtract")

    def __mul__(self, other):
        if isinstance(other, Vector3):
            return self.dot(other)
        if isinstance(other, Number):
            return self.scale(other)
        raise TypeError("cannot multiply a Vector3 by a %s" % type(other).__name__)

    def __rmul__(self, other):
        if isinstance(other, Number):
            return self.scale(other)
        raise TypeError("cannot multiply a %s by a Vector3" % type(other).__name__)

    def __truediv__(self, other):
        return self._combine(other, operator.truediv, "divide")

    def scale(self, s):
        return Vector3(*(a * s for a in self._parts()))

    def conj(self):
        return Vector3(*(a.conjugate() for a in self._parts()))

    def dot(self, v):
        return sum(a * b for a, b in zip(self._parts(), v._parts()))

    def cdot(self, v):
        return self.conj().dot(v)

    def cross(self, v):
        (a1, a2, a3), (b1, b2, b3) = self._parts(), v._parts()
        return Vector3(a2 * b3 - a3 * b2, a3 * b1 - a1 * b3, a1 * b2 - a2 * b1)

    def norm(self):
        return math.sqrt(abs(self.dot(self)))

    def unit(self):
        return self.scale(1 / self.norm())

    def prod(self):
        return self.x * self.y * self.z

    def close(self, v, tol=1.0e-7):
        return all(abs(a - b) <= tol for a, b in zip(self._parts(), v._parts()))

    def rotate(self, axis, theta):
        """Rotation by `theta` about `axis` (right-handed), Rodrigues' formula
        v cos t + (k x v) sin t + k (k . v)(1 - cos t)."""
        k = axis.unit()
        c, s = math.cos(theta), math.sin(theta)
        return self.scale(c) + k.cross(self).scale(s) + k.scale(k.dot(self) * (1 - c))

    # -- rounding and copies
    def ceil(self):
        return Vector3(*(math.ceil(a) for a in self._parts()))

    def floor(self):
        return Vector3(*(math.floor(a) for a in self._parts()))

    def round(self):
        return Vector3(*(round(a) for a in self._parts()))

    def copy(self):
        return Vector3(*self._parts())

    def get_json(self):
        return list(self._parts())


# ------------------------------------------------------------------ materials
def _json_of(obj, table):
    """JSON object {key: getattr(obj, attribute)} for a (key, attribute) table."""
    return {key: getattr(obj, attr) for key, attr in table}


class Susceptibility:
    """A polarization pole: chi(w) = sigma (a0 + a1 iw) / (b0 + b1 iw - b2 w^2).

    The engine integrates it with the recurrence of src/medium.cpp:24-31,
        P^{n+1} = c1 P^n + c2 P^{n-1} + sigma c3 E^n,
    whose response at z = exp(iw dt) is sigma c3 z / (z^2 - c1 z - c2)
    (``get_dis_epsilon``)."""

    json_type = None  # "type" value of the JSON object
    json_keys = ()    # (key, attribute) pairs besides type and amplitude

    def __init__(self, sigma=1.0):
        self.sigma = sigma
        self.build_eqs()

    def build_eqs(self, a0=0.0, a1=0.0, b0=0.0, b1=0.0, b2=0.0):
        """Set the response's numerator (a0, a1) and denominator (b0, b1, b2)."""
        self.a0, self.a1 = float(a0), float(a1)
        self.b0, self.b1, self.b2 = float(b0), float(b1), float(b2)

    def _coefficients(self):
        return (self.a0, self.a1, self.b0, self.b1, self.b2)

    def __eq__(self, other):
        return isinstance(other, Susceptibility) and self._coefficients() == other._coefficients()

    __hash__ = None

    def get_epsilon(self, frequency):
        iw = 2j * np.pi * frequency
        return self.sigma * (self.a0 + self.a1 * iw) / (self.b0 + self.b1 * iw + self.b2 * iw * iw)

    def coefficients(self, dt):
        """(c1, c2, c3) of the engine's recurrence for time step dt."""
        d = 2 * self.b2 + self.b1 * dt
        return ((4 * self.b2 - 2 * self.b0 * dt * dt) / d, (self.b1 * dt - 2 * self.b2) / d,
                2 * self.a0 * dt * dt / d)

    def get_dis_epsilon(self, frequency, dt):
        """Susceptibility of the time-stepped recurrence at frequency `frequency`."""
        c1, c2, c3 = self.coefficients(dt)
        z = np.exp(2j * np.pi * frequency * dt)
        return self.sigma * c3 * z / (z * z - c1 * z - c2)

    def get_json(self):
        out = {"type": self.json_type}
        out.update(_json_of(self, self.json_keys))
        out["amplitude"] = self.sigma
        return out


class LorentzianSusceptibility(Susceptibility):
    """Lorentz pole at `frequency` with damping `gamma` (both in cycles per unit
    time): chi = sigma w0^2 / (w0^2 + i g w - w^2), w0 = 2 pi frequency,
    g = 2 pi gamma."""

    json_type = "Lorentz"
    json_keys = (("frequency", "frequency"), ("gamma", "gamma"))

    def __init__(self, frequency=0.0, gamma=0.0, **kwargs):
        super().__init__(**kwargs)
        self.frequency, self.gamma = float(frequency), float(gamma)
        w0, g = 2 * np.pi * self.frequency, 2 * np.pi * self.gamma
        self.build_eqs(a0=w0 * w0, b0=w0 * w0, b1=g, b2=1)


class DrudeSusceptibility(Susceptibility):
    """Drude pole: chi = sigma wp^2 / (i g w - w^2), wp = 2 pi frequency,
    g = 2 pi gamma."""

    json_type = "Drude"
    json_keys = (("frequency", "frequency"), ("gamma", "gamma"))

    def __init__(self, frequency=0.0, gamma=0.0, **kwargs):
        super().__init__(**kwargs)
        self.frequency, self.gamma = float(frequency), float(gamma)
        wp, g = 2 * np.pi * self.frequency, 2 * np.pi * self.gamma
        self.build_eqs(a0=wp * wp, b1=g, b2=1)


class DeybeSusceptibility(Susceptibility):
    """Debye relaxation: chi = sigma / (1 + i w tau).  (The class and its JSON
    type keep the reference's spelling "Deybe".)  With b2 = 0 the engine's
    second-order recurrence has c2 = 1 and c1 < 0, a root outside the unit
    circle: Debye media grow without bound in the reference and here alike."""

    json_type = "Deybe"
    json_keys = (("tau", "tau"),)

    def __init__(self, tau=0.0, **kwargs):
        super().__init__(**kwargs)
        self.tau = float(tau)
        self.build_eqs(a0=1, b0=1, b1=self.tau)


class Medium:
    """Non-dispersive epsilon and mu, electric and magnetic conductivities, and
    lists of electric and magnetic susceptibilities."""

    json_keys = (("epsilon", "epsilon"), ("mu", "mu"), ("electric conductivity", "E_conductivity"),
                 ("magnetic conductivity", "M_conductivity"))

    def __init__(self, epsilon=1, mu=1, E_conductivity=0, M_conductivity=0, E_susceptibilities=(),
                 M_susceptibilities=()):
        self.epsilon, self.mu = float(epsilon), float(mu)
        self.E_conductivity, self.M_conductivity = float(E_conductivity), float(M_conductivity)
        self.E_susceptibilities = [deepcopy(s) for s in E_susceptibilities]
        self.M_susceptibilities = [deepcopy(s) for s in M_susceptibilities]

    def get_imp(self):
        """Wave impedance of the non-dispersive part."""
        return math.sqrt(self.mu / self.epsilon)

    def get_c(self):
        """Wave speed of the non-dispersive part."""
        return 1 / math.sqrt(self.mu * self.epsilon)

    def get_epsilon(self, frequency):
        """Complex permittivity at `frequency` (exp(iwt) convention)."""
        eps = self.epsilon - 1j * self.E_conductivity / (2 * np.pi * frequency)
        return eps + sum(s.get_epsilon(frequency) for s in self.E_susceptibilities)

    def get_dis_epsilon(self, frequency, dt):
        """Permittivity of the time-stepped update at `frequency`: the
        conductivity term of the centred update is 2 dt z / (z^2 - 1) sigma_e."""
        z = np.exp(2j * np.pi * frequency * dt)
        eps = self.epsilon + self.E_conductivity * 2 * dt * z / (z * z - 1)
        return eps + sum(s.get_dis_epsilon(frequency, dt) for s in self.E_susceptibilities)

    def get_json(self):
        out = _json_of(self, self.json_keys)
        out["electric susceptibility"] = [s.get_json() for s in self.E_susceptibilities]
        out["magnetic susceptibility"] = [s.get_json() for s in self.M_susceptibilities]
        return out


# ------------------------------------------------------------------- shapes
class GeometricObject:
    """A shape filled with one medium (`material`), placed at `center`."""

    def __init__(self, material=None, center=None):
        self.material = deepcopy(material) if material is not None else Medium()
        self.center = center.copy() if center is not None else Vector3()

    def shift(self, vec):
        """A copy moved by `vec`."""
        moved = deepcopy(self)
        moved.center = moved.center + vec
        return moved

    __add__ = shift

    def dump(self, store):
        """Uniform shapes carry no arrays."""

    def get_json(self):
        out = {"type": self.json_type, "center": self.center.get_json()}
        out.update(self._shape_json())
        out["medium"] = self.material.get_json()
        return out


class Sphere(GeometricObject):
    """Sphere of `radius` (a negative radius is an empty sphere)."""

    json_type = "sphere"

    def __init__(self, radius, **kwargs):
        super().__init__(**kwargs)
        self.radius = radius

    @property
    def radius(self):
        return self._radius

    @radius.setter
    def radius(self, value):
        self._radius = max(float(value), 0.0)

    def _shape_json(self):
        return {"radius": self.radius}


class Box(GeometricObject):
    """Axis-aligned box; a size with any extent <= 0 makes it empty (size 0)."""

    json_type = "box"

    def __init__(self, size=None, **kwargs):
        super().__init__(**kwargs)
        self.size = size if size is not None else Vector3()

    @property
    def size(self):
        return self._size

    @size.setter
    def size(self, value):
        self._size = value.copy() if min(value.x, value.y, value.z) > 0 else Vector3()

    def _shape_json(self):
        return {"size": self.size.get_json()}


# ------------------------------------------------------- sampled arrays on boxes
class NodeGrid:
    """An array sampled on the nodes of a box: `dim` nodes per axis from
    center - size/2 to center + size/2, stored [z, y, x].

    Subclasses name the array through ``_grid_array`` (for example "density")
    and may give it as an array of that shape, a function of the node points
    (an array [..., (z, y, x)] -> values), or None for zeros."""

    _grid_array = "values"
    _grid_dtype = float

    def _place(self, center, size, dim):
        self.center = center.copy()
        self.size = size.copy()
        self.dim = dim.round()

    def _initial(self, given):
        if callable(given):
            return given(metadata_to_pts(self.z, self.y, self.x))
        if isinstance(given, np.ndarray):
            return given
        if given is None:
            return np.zeros(self.shape, dtype=self._grid_dtype)
        raise ValueError("%s must be an array, a function of the node coordinates or None, not %s"
                         % (self._grid_array, type(given).__name__))

    def _checked(self, value):
        value = np.asarray(value)
        if not cmp_shape(self.shape, value.shape):
            raise ValueError("%s has shape %s; the %d x %d x %d nodes need %s" % (
                self._grid_array, value.shape, self.shape[2], self.shape[1], self.shape[0], self.shape))
        return np.array(value, copy=True)

    @property
    def x(self):
        return _axis(self.center.x, self.size.x, self.dim.x)

    @property
    def y(self):
        return _axis(self.center.y, self.size.y, self.dim.y)

    @property
    def z(self):
        return _axis(self.center.z, self.size.z, self.dim.z)

    @property
    def dimension(self):
        return self.dim

    @property
    def shape(self):
        return (int(self.dim.z), int(self.dim.y), int(self.dim.x))

    @property
    def numel(self):
        return int(self.dim.prod())


class _DensityBox(NodeGrid):
    _grid_array = "density"

    @property
    def density(self):
        return self._density

    @density.setter
    def density(self, value):
        self._density = self._checked(value)


class Two_Medium_Box(_DensityBox):
    """medium1 where the density is 1, medium2 where it is 0, linear in between;
    the density is interpolated trilinearly between the nodes
    (src/geometry.cpp:49-76).  Dataset: "two medium box <suffix>"."""

    def __init__(self, size=None, center=None, dim=None, density=None, medium1=None, medium2=None, suffix="0"):
        self._place(center or Vector3(), size or Vector3(), dim or Vector3())
        self.medium1 = deepcopy(medium1) if medium1 is not None else Medium()
        self.medium2 = deepcopy(medium2) if medium2 is not None else Medium()
        self.density = self._initial(density)
        self.dataset_name = "two medium box %s" % suffix

    def dump(self, store):
        store.create_dataset(self.dataset_name, data=self.density.ravel())

    def get_json(self):
        return {"type": "two medium box", "center": self.center.get_json(), "size": self.size.get_json(),
                "dimension": self.dim.get_json(), "density dataset": self.dataset_name,
                "medium1": self.medium1.get_json(), "medium2": self.medium2.get_json()}


class General_Medium_Box(_DensityBox):
    """A density box whose medium follows a table sampled along the density
    (src/geometry.cpp:79-107): epsilon_inf = mu_inf = 1 and two electric poles
    whose amplitudes a1(rho), a2(rho) make 1 + a1 chi1(f) + a2 chi2(f) equal
    `epsilon_fun(rho)` at `frequency` (chi_i the unit-amplitude responses): a
    2 x 2 real system in Re and Im.  The table holds TABLE densities from 0 to 1.
    Datasets: "general medium box <suffix> " + density, density function,
    epsilon, mu, e sus amp, m sus amp."""

    TABLE = 1000

    def __init__(self, size=None, center=None, dim=None, frequency=1.0, density=None, epsilon_fun=None, e_sus=(),
                 suffix="0"):
        if len(e_sus) != 2:
            raise ValueError("a general medium box takes exactly two electric susceptibilities, got %d" % len(e_sus))
        self._place(center or Vector3(), size or Vector3(), dim or Vector3())
        self.frequency = float(frequency)
        self.e_sus = [deepcopy(s) for s in e_sus]
        self.medium = Medium(epsilon=1, E_susceptibilities=self.e_sus)
        self.epsilon_fun = epsilon_fun
        # columns: unit-amplitude responses of the two poles, rows: Re, Im
        chi = [s.get_epsilon(self.frequency) / s.sigma for s in self.e_sus]
        self.mat = np.array([[c.real for c in chi], [c.imag for c in chi]])
        self.invmat = np.linalg.inv(self.mat)
        self.density = self._initial(density)
        stem = "general medium box %s " % suffix
        self.density_dataset = stem + "density"
        self.density_fun_dataset = stem + "density function"
        self.epsilon_dataset = stem + "epsilon"
        self.mu_dataset = stem + "mu"
        self.e_sus_amp_dataset = stem + "e sus amp"
        self.m_sus_amp_dataset = stem + "m sus amp"

    def amplitude_table(self):
        """(densities [TABLE], pole amplitudes [TABLE, 2])."""
        rho = np.linspace(0.0, 1.0, self.TABLE)
        eps = np.asarray(self.epsilon_fun(rho), dtype=complex)
        amps = np.linalg.solve(self.mat, np.stack((eps.real - 1, eps.imag)))
        return rho, amps.T

    def dump(self, store):
        rho, amps = self.amplitude_table()
        store.create_dataset(self.density_dataset, data=self.density.ravel())
        store.create_dataset(self.density_fun_dataset, data=rho)
        store.create_dataset(self.epsilon_dataset, data=np.ones(rho.size))
        store.create_dataset(self.mu_dataset, data=np.ones(rho.size))
        store.create_dataset(self.e_sus_amp_dataset, data=amps.ravel())
        store.create_dataset(self.m_sus_amp_dataset, data=np.zeros(0))

    def get_json(self):
        return {"type": "general medium box", "center": self.center.get_json(), "size": self.size.get_json(),
                "dimension": self.dim.get_json(), "density dataset": self.density_dataset,
                "density function dataset": self.density_fun_dataset, "epsilon dataset": self.epsilon_dataset,
                "mu dataset": self.mu_dataset,
                "electric susceptibility amplitudes dataset": self.e_sus_amp_dataset,
                "magnetic susceptibility amplitudes dataset": self.m_sus_amp_dataset,
                "medium": self.medium.get_json()}


# ----------------------------------------------------------- polygon prisms
def planar_polygon(polygon, center=None, height=np.inf):
    """Density function of a prism: 1 for points strictly inside `polygon` in the
    x-y plane (coordinates relative to `center`) and within height/2 of center.z,
    else 0.  The function takes points [..., (z, y, x)].  `polygon` is a vertex
    array (n, 2) or a shapely-style object: its ``exterior`` (or that object's
    ``coords``) gives the vertices, and ``geoms`` a union of polygons."""
    from .polygon import check_inside
    center = Vector3() if center is None else center

    def vertices(g):
        ext = getattr(g, "exterior", None)
        if ext is not None:
            return np.asarray(getattr(ext, "coords", ext), dtype=float)
        return np.asarray(g, dtype=float)

    rings = [vertices(g) for g in (polygon.geoms if hasattr(polygon, "geoms") else [polygon])]

    def density(req_pts):
        req_pts = np.asarray(req_pts, dtype=float)
        xy = np.stack((req_pts[..., 2] - center.x, req_pts[..., 1] - center.y), axis=-1)
        inside = np.zeros(req_pts.shape[:-1], dtype=bool)
        for ring in rings:
            inside |= check_inside(xy, ring)
        inside &= np.abs(req_pts[..., 0] - center.z) < height / 2
        return inside.astype(int)

    return density
