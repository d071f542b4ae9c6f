"""Geometry and media of the front end (the public names of ffip/geom.py).

Objects serialise to the same JSON the reference's Simulation::init reads
(src/simulation.cpp:58-100, 402-518), so a configuration written through this
module runs unchanged on the reference and on the MI355X engine.
"""
from __future__ import annotations

import math
from copy import deepcopy
from numbers import Number

import numpy as np


def cmp_shape(shape1=(), shape2=()):
    """True when two shapes have the same rank and extents (ffip/geom.py:12-20)."""
    return len(shape1) == len(shape2) and all(a == b for a, b in zip(shape1, shape2))


def metadata_to_pts(*args):
    """Stack an ij-meshgrid of the given axes into points (ffip/geom.py:28-30)."""
    return np.stack(np.meshgrid(*args, indexing="ij"), axis=-1)


def sph2cart(az, el, r):
    """Cartesian (x, y, z) of azimuth az, elevation el, radius r (ffip/geom.py:22-27)."""
    horiz = r * np.cos(el)
    return horiz * np.cos(az), horiz * np.sin(az), r * np.sin(el)


class Vector3:
    """3-vector with the arithmetic of ffip.Vector3 (ffip/geom.py:32-170)."""

    def __init__(self, x=0.0, y=0.0, z=0.0):
        self.x = float(x) if type(x) is int else x
        self.y = float(y) if type(y) is int else y
        self.z = float(z) if type(z) is int else z

    def _other(self, other):
        return Vector3(other, other, other) if isinstance(other, Number) else other

    def __eq__(self, other):
        return self.x == other.x and self.y == other.y and self.z == other.z

    def __add__(self, other):
        o = self._other(other)
        return Vector3(self.x + o.x, self.y + o.y, self.z + o.z)

    def __sub__(self, other):
        o = self._other(other)
        return Vector3(self.x - o.x, self.y - o.y, self.z - o.z)

    def __mul__(self, other):
        if type(other) is Vector3:
            return self.dot(other)
        if isinstance(other, Number):
            return self.scale(other)
        raise TypeError("No operation known for 'Vector3 * {}'".format(type(other)))

    def __rmul__(self, other):
        if isinstance(other, Number):
            return self.scale(other)
        raise TypeError("No operation known for '{} * Vector3'".format(type(other)))

    def __truediv__(self, other):
        if type(other) is Vector3:
            return Vector3(self.x / other.x, self.y / other.y, self.z / other.z)
        if isinstance(other, Number):
            return Vector3(self.x / other, self.y / other, self.z / other)
        raise TypeError("No operation known for 'Vector3 / {}'".format(type(other)))

    def __getitem__(self, i):
        if i in (0, 1, 2):
            return (self.x, self.y, self.z)[i]
        raise IndexError("No value at index {}".format(i))

    def __setitem__(self, i, v):
        if i == 0:
            self.x = v
        elif i == 1:
            self.y = v
        elif i == 2:
            self.z = v
        else:
            raise IndexError("No value at index {}".format(i))

    def __repr__(self):
        return "Vector3<{}, {}, {}>".format(self.x, self.y, self.z)

    def __array__(self, dtype=None, copy=None):
        return np.array([self.x, self.y, self.z], dtype=dtype)

    def conj(self):
        return Vector3(self.x.conjugate(), self.y.conjugate(), self.z.conjugate())

    def scale(self, s):
        return Vector3(self.x * s, self.y * s, self.z * s)

    def dot(self, v):
        return self.x * v.x + self.y * v.y + self.z * v.z

    def cdot(self, v):
        return self.conj().dot(v)

    def cross(self, v):
        return Vector3(self.y * v.z - self.z * v.y, self.z * v.x - self.x * v.z, self.x * v.y - self.y * v.x)

    def norm(self):
        return math.sqrt(abs(self.dot(self)))

    def unit(self):
        return self.scale(1 / self.norm())

    def prod(self):
        return self.x * self.y * self.z

    def close(self, v, tol=1.0e-7):
        return abs(self.x - v.x) <= tol and abs(self.y - v.y) <= tol and abs(self.z - v.z) <= tol

    def rotate(self, axis, theta):
        u = axis.unit()
        vpar = u.scale(u.dot(self))
        vcross = u.cross(self)
        vperp = self - vpar
        return vpar + (vperp.scale(math.cos(theta)) + vcross.scale(math.sin(theta)))

    def get_json(self):
        return [self.x, self.y, self.z]

    def ceil(self):
        return Vector3(math.ceil(self.x), math.ceil(self.y), math.ceil(self.z))

    def floor(self):
        return Vector3(math.floor(self.x), math.floor(self.y), math.floor(self.z))

    def round(self):
        return Vector3(round(self.x), round(self.y), round(self.z))

    def copy(self):
        return Vector3(self.x, self.y, self.z)


class Susceptibility:
    """sigma * (a0 + a1 jw) / (b0 + b1 jw - b2 w^2) (ffip/geom.py:229-262)."""

    def __init__(self, sigma=1.0):
        self.sigma = sigma

    def build_eqs(self, a0=0.0, a1=0.0, b0=0.0, b1=0.0, b2=0.0):
        self.a0, self.a1, self.b0, self.b1, self.b2 = (float(v) for v in (a0, a1, b0, b1, b2))

    def __eq__(self, other):
        return (self.a0, self.a1, self.b0, self.b1, self.b2) == (other.a0, other.a1, other.b0, other.b1, other.b2)

    def get_epsilon(self, frequency):
        w = 2 * np.pi * frequency
        return self.sigma * (self.a0 + self.a1 * 1j * w) / (self.b0 + 1j * w * self.b1 - w * w * self.b2)

    def coefficients(self, dt):
        """(c1, c2, c3) of the discrete recurrence (src/medium.cpp:24-31)."""
        c0 = 2 * self.b2 + self.b1 * dt
        return ((4 * self.b2 - 2 * self.b0 * dt * dt) / c0, (-2 * self.b2 + self.b1 * dt) / c0,
                (2 * self.a0 * dt * dt) / c0)

    def get_dis_epsilon(self, frequency, dt):
        """Susceptibility of the time-discrete update at z = exp(i w dt)."""
        c1, c2, c3 = self.coefficients(dt)
        z = np.exp(1j * 2 * np.pi * frequency * dt)
        return self.sigma * c3 * z / (z ** 2 - c1 * z - c2)


class LorentzianSusceptibility(Susceptibility):
    def __init__(self, frequency=0.0, gamma=0.0, **kwargs):
        w, g = 2 * np.pi * frequency, 2 * np.pi * gamma
        super().__init__(**kwargs)
        self.build_eqs(a0=w ** 2, b0=w ** 2, b1=g, b2=1)
        self.frequency, self.gamma = float(frequency), float(gamma)

    def get_json(self):
        return {"type": "Lorentz", "frequency": self.frequency, "gamma": self.gamma, "amplitude": self.sigma}


class DrudeSusceptibility(Susceptibility):
    def __init__(self, frequency=0.0, gamma=0.0, **kwargs):
        w, g = 2 * np.pi * frequency, 2 * np.pi * gamma
        super().__init__(**kwargs)
        self.build_eqs(a0=w ** 2, b1=g, b2=1)
        self.frequency, self.gamma = float(frequency), float(gamma)

    def get_json(self):
        return {"type": "Drude", "frequency": self.frequency, "gamma": self.gamma, "amplitude": self.sigma}


class DeybeSusceptibility(Susceptibility):
    """Debye pole.  Note: the reference's second-order recurrence with b2 = 0
    (c2 = 1, c1 < 0) has a root outside the unit circle, so Debye media diverge
    in the reference and here alike."""

    def __init__(self, tau=0.0, **kwargs):
        super().__init__(**kwargs)
        self.build_eqs(a0=1, b0=1, b1=tau)
        self.tau = float(tau)

    def get_json(self):
        return {"type": "Deybe", "tau": self.tau, "amplitude": self.sigma}


class Medium:
    """epsilon, mu, conductivities and susceptibilities (ffip/geom.py:173-227)."""

    def __init__(self, epsilon=1, mu=1, E_conductivity=0, M_conductivity=0, E_susceptibilities=(),
                 M_susceptibilities=()):
        self.epsilon, self.mu = float(epsilon), float(mu)
        self.E_susceptibilities = deepcopy(list(E_susceptibilities))
        self.M_susceptibilities = deepcopy(list(M_susceptibilities))
        self.E_conductivity, self.M_conductivity = float(E_conductivity), float(M_conductivity)

    def get_imp(self):
        return math.sqrt(self.mu / self.epsilon)

    def get_c(self):
        return 1 / math.sqrt(self.mu * self.epsilon)

    def get_json(self):
        return {"epsilon": self.epsilon, "mu": self.mu, "electric conductivity": self.E_conductivity,
                "magnetic conductivity": self.M_conductivity,
                "electric susceptibility": [s.get_json() for s in self.E_susceptibilities],
                "magnetic susceptibility": [s.get_json() for s in self.M_susceptibilities]}

    def get_epsilon(self, frequency):
        res = self.epsilon - 1j * self.E_conductivity / (2 * np.pi * frequency)
        for s in self.E_susceptibilities:
            res += s.get_epsilon(frequency)
        return res

    def get_dis_epsilon(self, frequency, dt):
        z = np.exp(1j * 2 * np.pi * frequency * dt)
        res = self.epsilon + 2 * dt * z / (z ** 2 - 1) * self.E_conductivity
        for s in self.E_susceptibilities:
            res += s.get_dis_epsilon(frequency, dt)
        return res


class GeometricObject:
    def __init__(self, material=Medium(), center=Vector3()):
        self.material = deepcopy(material)
        self.center = center.copy()

    def shift(self, vec):
        c = deepcopy(self)
        c.center += vec
        return c

    def dump(self, store):
        pass


class Sphere(GeometricObject):
    def __init__(self, radius, **kwargs):
        self.radius = radius
        super().__init__(**kwargs)

    @property
    def radius(self):
        return self._radius

    @radius.setter
    def radius(self, val):
        self._radius = float(val) if val > 0 else 0.0

    def get_json(self):
        return {"type": "sphere", "center": self.center.get_json(), "radius": self.radius,
                "medium": self.material.get_json()}


class Box(GeometricObject):
    def __init__(self, size=Vector3(), **kwargs):
        self.size = size
        super().__init__(**kwargs)

    @property
    def size(self):
        return self._size

    @size.setter
    def size(self, val):
        # any non-positive extent collapses the box (ffip/geom.py:363-364)
        self._size = val.copy() if (val.x > 0 and val.y > 0 and val.z > 0) else Vector3()

    def get_json(self):
        return {"type": "box", "center": self.center.get_json(), "size": self.size.get_json(),
                "medium": self.material.get_json()}


class Two_Medium_Box:
    """medium1 * rho + medium2 * (1 - rho) over linspace(box) nodes
    (ffip/geom.py:375-444, src/geometry.cpp:49-76)."""

    def __init__(self, size=Vector3(), center=Vector3(), dim=Vector3(), density=None, medium1=Medium(),
                 medium2=Medium(), suffix="0"):
        self.size = size.copy()
        self.center = center.copy()
        self.dim = dim.round()
        self.medium1 = deepcopy(medium1)
        self.medium2 = deepcopy(medium2)
        if callable(density):
            self.density = density(np.stack(np.meshgrid(self.z, self.y, self.x, indexing="ij"), axis=-1))
        elif isinstance(density, np.ndarray):
            self.density = density
        elif density is None:
            self.density = np.zeros(self.shape, dtype=float)
        else:
            raise ValueError("amplitude input is not numpy array or a function")
        self.dataset_name = "two medium box %s" % suffix

    @property
    def density(self):
        return self._density

    @density.setter
    def density(self, val):
        if not cmp_shape(self.shape, val.shape):
            raise ValueError("density size is not compatible with the box dimension")
        self._density = np.array(val, copy=True)

    @property
    def x(self):
        return np.linspace(self.center.x - self.size.x / 2, self.center.x + self.size.x / 2, int(self.dim.x))

    @property
    def y(self):
        return np.linspace(self.center.y - self.size.y / 2, self.center.y + self.size.y / 2, int(self.dim.y))

    @property
    def z(self):
        return np.linspace(self.center.z - self.size.z / 2, self.center.z + self.size.z / 2, int(self.dim.z))

    @property
    def dimension(self):
        return self.dim

    @property
    def shape(self):
        return (int(self.dim.z), int(self.dim.y), int(self.dim.x))

    @property
    def numel(self):
        return int(self.dim.prod())

    def dump(self, store):
        store.create_dataset(self.dataset_name, data=self.density.ravel())

    def get_json(self):
        return {"type": "two medium box", "center": self.center.get_json(), "size": self.size.get_json(),
                "dimension": self.dim.get_json(), "density dataset": self.dataset_name,
                "medium1": self.medium1.get_json(), "medium2": self.medium2.get_json()}


class General_Medium_Box(Two_Medium_Box):
    """A density box whose medium is read from a table of media sampled along the
    density (ffip/geom.py:446-558; src/geometry.cpp:79-107).

    The medium is epsilon_inf = 1 with the two given electric susceptibilities.
    `epsilon_fun(rho)` gives the complex permittivity wanted at `frequency` for
    density rho in [0, 1].  The two susceptibility amplitudes that reproduce it
    solve a 2 x 2 real system: Re and Im of epsilon - 1 against Re and Im of each
    susceptibility's unit-amplitude response at `frequency`.  The table holds 1000
    densities linspace(0, 1) and their amplitudes.  The engine interpolates the
    density at each point, then the table linearly."""

    TABLE = 1000

    def __init__(self, size=Vector3(), center=Vector3(), dim=Vector3(), frequency=1.0, density=None,
                 epsilon_fun=None, e_sus=(), suffix="0"):
        if len(e_sus) != 2:
            raise ValueError("Currently only support two poles")
        super().__init__(size=size, center=center, dim=dim, density=density, suffix=suffix)
        self.frequency = float(frequency)
        self.e_sus = deepcopy(list(e_sus))
        self.medium = Medium(epsilon=1, E_susceptibilities=self.e_sus)
        self.epsilon_fun = deepcopy(epsilon_fun)
        # unit-amplitude responses (get_epsilon carries the amplitude sigma)
        unit = [s.get_epsilon(self.frequency) / s.sigma for s in self.e_sus]
        self.mat = np.array([[np.real(unit[0]), np.real(unit[1])], [np.imag(unit[0]), np.imag(unit[1])]])
        self.invmat = np.linalg.inv(self.mat)
        prefix = "general medium box %s " % suffix
        self.density_dataset = prefix + "density"
        self.density_fun_dataset = prefix + "density function"
        self.epsilon_dataset = prefix + "epsilon"
        self.mu_dataset = prefix + "mu"
        self.e_sus_amp_dataset = prefix + "e sus amp"
        self.m_sus_amp_dataset = prefix + "m sus amp"

    def amplitude_table(self):
        """(densities, amplitudes [n, 2]) of the sampled media."""
        rho = np.linspace(0, 1, self.TABLE)
        eps = np.asarray(self.epsilon_fun(rho), dtype=complex)
        amps = self.invmat @ np.stack((np.real(eps) - 1, np.imag(eps)))
        return rho, amps.T

    def dump(self, store):
        rho, amps = self.amplitude_table()
        store.create_dataset(self.density_dataset, data=self.density.ravel())
        store.create_dataset(self.density_fun_dataset, data=rho)
        store.create_dataset(self.epsilon_dataset, data=np.ones(rho.size))
        store.create_dataset(self.mu_dataset, data=np.ones(rho.size))
        store.create_dataset(self.e_sus_amp_dataset, data=amps.ravel())
        store.create_dataset(self.m_sus_amp_dataset, data=np.ones(0))

    def get_json(self):
        return {"type": "general medium box", "center": self.center.get_json(), "size": self.size.get_json(),
                "dimension": self.dim.get_json(), "density dataset": self.density_dataset,
                "density function dataset": self.density_fun_dataset, "epsilon dataset": self.epsilon_dataset,
                "mu dataset": self.mu_dataset,
                "electric susceptibility amplitudes dataset": self.e_sus_amp_dataset,
                "magnetic susceptibility amplitudes dataset": self.m_sus_amp_dataset,
                "medium": self.medium.get_json()}


def getgrid(center=None, size=None, dim=None):
    """(z, y, x) node axes linspace(center - size/2, center + size/2, dim)
    (ffip/geom.py:561-567)."""
    center = Vector3() if center is None else center
    size = Vector3() if size is None else size
    dim = Vector3() if dim is None else dim
    axes = [np.linspace(c - w / 2, c + w / 2, int(n)) for c, w, n in
            ((center.x, size.x, dim.x), (center.y, size.y, dim.y), (center.z, size.z, dim.z))]
    return axes[2], axes[1], axes[0]


def planar_polygon(polygon, center=None, height=np.inf):
    """Density function of a prism (ffip/geom.py:569-590): 1 where a point
    (req_pts[..., :] = (z, y, x)) lies strictly inside the polygon in the x-y plane,
    taken relative to `center`, and within height/2 of center.z; else 0.
    `polygon` is a vertex array (n, 2), or a shapely-style object: `.exterior`
    (its coordinates are used) or `.geoms` (a union of such polygons)."""
    from .polygon import check_inside
    center = Vector3() if center is None else center

    def vertices(g):
        ext = getattr(g, "exterior", None)
        if ext is not None:
            return np.asarray(getattr(ext, "coords", ext), dtype=float)
        return np.asarray(g, dtype=float)

    parts = list(polygon.geoms) if hasattr(polygon, "geoms") else [polygon]

    def density(req_pts):
        req_pts = np.asarray(req_pts, dtype=float)
        xy = np.stack((req_pts[..., 2] - center.x, req_pts[..., 1] - center.y), axis=-1)
        inside = np.zeros(req_pts.shape[:-1], dtype=bool)
        for g in parts:
            inside |= check_inside(xy, vertices(g))
        inside &= np.abs(req_pts[..., 0] - center.z) < height / 2
        return inside.astype(int)

    return density
