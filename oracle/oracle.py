"""CPU oracle for the ffip FDTD hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by
the ``cpu_baseline`` leg of ``bench.py`` -- always as the checker, never as the thing
measured or shipped.  The product path (``ffip_amd``) never imports this module.

This module restates, in numpy/Python, the set-up half of the reference's
``Simulation::init`` (src/simulation.cpp:127-188) on the reference's own doubled Yee
grid, and drives the FP64 hot loop restated in C (oracle/ffip_oracle.c, built into
oracle/liboracle.so by oracle/Makefile).  Every function cites the reference
file:line it follows.  Arithmetic is written operation-by-operation in the same
order as the reference so that set-up quantities (PML grading, source weights,
material interpolation) are bit-identical to what the reference computes.

Parity pins for every function here are in tests/test_oracle_pins.py.
"""
from __future__ import annotations

import ctypes
import math
import os
from fractions import Fraction

import numpy as np

PI = 3.141592653589793e+00  # src/utility.cpp:5

# Coord_Type bit patterns, include/utility.hpp:48-64
CTYPE = {
    "Ex": 0b001, "Ey": 0b010, "Ez": 0b100,
    "Hx": 0b110, "Hy": 0b101, "Hz": 0b011,
    "Dx": 0b1001, "Dy": 0b1010, "Dz": 0b1100,
    "Bx": 0b1110, "By": 0b1101, "Bz": 0b1011,
}
EH_ORDER = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]

BC_NONE, BC_PERIOD, BC_PEC, BC_PMC = "None", "Period", "PEC", "PMC"


def count_bits(x: int) -> int:
    """count_hex_bits, include/utility.hpp:72-76."""
    return (x & 1) + ((x & 2) >> 1) + ((x & 4) >> 2) + ((x & 8) >> 3)


def is_e_point(ct: int) -> bool:
    """include/utility.hpp:79-83."""
    return count_bits(ct & 0b111) == 1


def is_eh_point(ct: int) -> bool:
    """include/utility.hpp:93-97."""
    return 0 < ct < 0b111


def dir_of_ctype(ct: int) -> int:
    """get_dir_from_ctype, include/utility.hpp:118-128."""
    c = ct & 0b111
    if count_bits(c) == 2:
        c = 0b111 ^ c
    return {0b001: 0, 0b010: 1, 0b100: 2}.get(c, 3)


def cround(x: float) -> float:
    """std::round: half away from zero (Python's round() is half-to-even)."""
    return math.copysign(math.floor(abs(x) + 0.5), x)


# ----------------------------------------------------------------------------
# Grid utilities: include/utility.hpp:478-552, src/utility.cpp:28-257
# ----------------------------------------------------------------------------

def nearest_int(x: float, S: int, T: int) -> int:
    """get_nearest_int<S,T>(double), include/utility.hpp:487-495."""
    xi = int(math.floor(x)) if S == -1 else int(math.ceil(x))
    return xi + S if ((xi & 1) ^ T) else xi


def nearest_point(p, ct: int, S: int):
    """get_nearest_point<S>, include/utility.hpp:498-534."""
    return (nearest_int(p[0], S, 1 if ct & 1 else 0),
            nearest_int(p[1], S, 1 if ct & 2 else 0),
            nearest_int(p[2], S, 1 if ct & 4 else 0))


def component_interior(p1, p2, ct):
    """get_component_interior, include/utility.hpp:537-543."""
    return nearest_point(p1, ct, +1), nearest_point(p2, ct, -1)


def component_closure(p1, p2, ct):
    """get_component_closure, include/utility.hpp:546-552."""
    return nearest_point(p1, ct, -1), nearest_point(p2, ct, +1)


def interp_weights(sx, sy, sz):
    """Yee3::get_interp_weights, src/utility.cpp:231-241 (x fastest)."""
    return [(1 - sx) * (1 - sy) * (1 - sz),
            sx * (1 - sy) * (1 - sz),
            (1 - sx) * sy * (1 - sz),
            sx * sy * (1 - sz),
            (1 - sx) * (1 - sy) * sz,
            sx * (1 - sy) * sz,
            (1 - sx) * sy * sz,
            sx * sy * sz]


def yee_coords(p1, p2, stride):
    """Yee_Iterator coordinates (src/utility.cpp:28-125): z outer, x fastest.

    ``len = (p2 - p1) / stride + 1`` with C++ truncating division; empty if any
    length is <= 0.  Returns an (n, 3) int64 array."""
    lens = []
    for a in range(3):
        d = p2[a] - p1[a]
        q = int(d / stride)  # truncation toward zero
        lens.append(q + 1)
    if min(lens) <= 0:
        return np.zeros((0, 3), dtype=np.int64)
    z = p1[2] + stride * np.arange(lens[2], dtype=np.int64)
    y = p1[1] + stride * np.arange(lens[1], dtype=np.int64)
    x = p1[0] + stride * np.arange(lens[0], dtype=np.int64)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


class Yee3:
    """Doubled grid with one ghost shell, src/utility.cpp:128-229."""

    def __init__(self, ghost_p1, ghost_p2):
        self.ghost_p1 = tuple(ghost_p1)
        self.ghost_p2 = tuple(ghost_p2)
        self.dim = tuple(ghost_p2[a] - ghost_p1[a] + 1 for a in range(3))
        self.grid_p1 = tuple(g + 1 for g in ghost_p1)
        self.grid_p2 = tuple(g - 1 for g in ghost_p2)
        self.stride = (1, self.dim[0], self.dim[0] * self.dim[1])

    def size(self):
        return self.dim[0] * self.dim[1] * self.dim[2]

    def index(self, x, y, z):
        """get_index_from_coord, src/utility.cpp:206-210 (vectorised)."""
        return ((np.asarray(x, dtype=np.int64) - self.ghost_p1[0]) * self.stride[0]
                + (np.asarray(y, dtype=np.int64) - self.ghost_p1[1]) * self.stride[1]
                + (np.asarray(z, dtype=np.int64) - self.ghost_p1[2]) * self.stride[2])

    def is_inside(self, p):
        return all(self.grid_p1[a] <= p[a] <= self.grid_p2[a] for a in range(3))

    def base_point(self, pos, ct):
        return nearest_point(pos, ct, -1)

    def interp_terms(self, pos, ct):
        """(index, weight) terms of Yee3::interp, include/utility.hpp:983-1022."""
        base = self.base_point(pos, ct)
        s = [(pos[a] - base[a]) / 2 for a in range(3)]
        w = interp_weights(*s)
        top = [base[a] + 2 for a in range(3)]
        inside_all = all(self.grid_p1[a] <= base[a] and top[a] <= self.grid_p2[a] for a in range(3))
        intersects = all(base[a] <= self.grid_p2[a] and self.grid_p1[a] <= top[a] for a in range(3))
        terms = []
        wi = 0
        for k in (0, 2):
            for j in (0, 2):
                for i in (0, 2):
                    p = (base[0] + i, base[1] + j, base[2] + k)
                    if inside_all or (intersects and self.is_inside(p)):
                        terms.append((int(self.index(*p)), w[wi]))
                    wi += 1
        return terms


def gaussian1(t, width):
    """src/utility.cpp:260-264."""
    arg = t / width
    return arg * math.exp(-arg * arg / 2)


def linspace(s, e, dim):
    """src/utility.cpp:281-291."""
    if dim <= 1:
        return [float(s)]
    stride = (e - s) / (dim - 1)
    return [s + i * stride for i in range(dim)]


# ----------------------------------------------------------------------------
# Media: src/medium.cpp, src/simulation.cpp:58-100
# ----------------------------------------------------------------------------

def sus_lorentz(frequency, gamma):
    """make_Lorentz_susceptibility, src/medium.cpp:98-102."""
    omega = 2 * PI * frequency
    return (omega * omega, 0.0, omega * omega, 2 * PI * gamma, 1.0)


def sus_drude(frequency, gamma):
    """make_Drude_susceptibility, src/medium.cpp:104-108."""
    omega = 2 * PI * frequency
    return (omega * omega, 0.0, 0.0, 2 * PI * gamma, 1.0)


def sus_debye(tau):
    """make_Deybe_susceptibility, src/medium.cpp:110-113."""
    return (1.0, 0.0, 1.0, float(tau), 0.0)


SUS_COND = (1.0, 0.0, 0.0, 1.0, 0.0)  # make_conductivity_susceptibility, src/medium.cpp:115-118


def abstract_sus(sus, dt):
    """Abstract_Susceptibility, src/medium.cpp:24-31 -> (c1, c2, c3)."""
    a0, a1, b0, b1, b2 = sus
    c0 = 2 * b2 + b1 * dt
    c1 = (4 * b2 - 2 * b0 * dt * dt) / c0
    c2 = (-2 * b2 + b1 * dt) / c0
    c3 = (2 * a0 * dt * dt) / c0
    return c1, c2, c3


def json2sus(j):
    """json2susceptibility, src/simulation.cpp:90-101."""
    t = j["type"]
    if t == "Lorentz":
        return sus_lorentz(float(j["frequency"]), float(j["gamma"]))
    if t == "Drude":
        return sus_drude(float(j["frequency"]), float(j["gamma"]))
    if t == "Deybe":
        return sus_debye(float(j["tau"]))
    raise ValueError("Unknonw susceptibility")


class Medium:
    """Medium and json2medium, src/medium.cpp:34-53, src/simulation.cpp:58-83."""

    def __init__(self, eps=1.0, mu=1.0):
        self.eps = float(eps)
        self.mu = float(mu)
        self.e_sus, self.e_amp, self.m_sus, self.m_amp = [], [], [], []

    @staticmethod
    def from_json(j):
        m = Medium(float(j["epsilon"]), float(j["mu"]))
        ec = float(j["electric conductivity"])
        if ec != 0:
            m.e_sus.append(SUS_COND)
            m.e_amp.append(ec)
        mc = float(j["magnetic conductivity"])
        if mc != 0:
            m.m_sus.append(SUS_COND)
            m.m_amp.append(mc)
        for s in j.get("electric susceptibility", []) or []:
            m.e_sus.append(json2sus(s))
            m.e_amp.append(float(s["amplitude"]))
        for s in j.get("magnetic susceptibility", []) or []:
            m.m_sus.append(json2sus(s))
            m.m_amp.append(float(s["amplitude"]))
        return m


class AbstractMedium:
    """Abstract_Medium, include/medium.hpp:42-54."""

    def __init__(self, eps, mu, e_amp, m_amp):
        self.eps, self.mu = eps, mu
        self.e_amp = np.asarray(e_amp, dtype=float)
        self.m_amp = np.asarray(m_amp, dtype=float)


# ----------------------------------------------------------------------------
# Geometry (doubled-grid units), src/geometry.cpp:13-76
# ----------------------------------------------------------------------------

def _interp_level(data, n, stride, xq, rest, q, offset=None):
    """interpn<N>::operator() with zero padding (include/utility.hpp:676-700,
    770-794), vectorised over points; ``rest``/``q`` ordered (z, y, x)."""
    npts = xq.shape[0]
    if offset is None:
        offset = np.zeros(npts, dtype=np.int64)

    def base(off):
        if not rest:
            return data[off]
        return _interp_level(data, rest[0], int(np.prod(rest[1:])) if len(rest) > 1 else 1,
                             q[1], rest[1:], q[1:], off)

    if n == 1:
        return base(offset)
    out = np.zeros(npts)
    lo = xq <= 0
    lo_in = lo & (xq > -1)
    hi = (~lo) & (xq >= n - 1)
    hi_in = hi & (xq < n)
    mid = (~lo) & (~hi)
    if lo_in.any():
        out[lo_in] = (xq[lo_in] + 1) * base(offset)[lo_in]
    if hi_in.any():
        out[hi_in] = (n - xq[hi_in]) * base(offset + stride * (n - 1))[hi_in]
    if mid.any():
        idx = np.zeros(npts, dtype=np.int64)
        idx[mid] = xq[mid].astype(np.int64)
        tx = xq - idx
        safe_hi = np.where(mid, offset + stride * (idx + 1), offset)
        safe_lo = np.where(mid, offset + stride * idx, offset)
        v = tx * base(safe_hi) + (1 - tx) * base(safe_lo)
        out[mid] = v[mid]
    return out


class Box:
    """Box, src/geometry.cpp:29-46: strict interior of [c - s/2, c + s/2]."""

    def __init__(self, center, size, medium):
        self.p1 = [center[a] - size[a] / 2 for a in range(3)]
        self.p2 = [center[a] + size[a] / 2 for a in range(3)]
        self.lo = [min(self.p1[a], self.p2[a]) for a in range(3)]
        self.hi = [max(self.p1[a], self.p2[a]) for a in range(3)]
        self.medium = medium

    def bounds(self):
        return self.lo, self.hi

    def inside(self, X):
        m = np.ones(X.shape[0], dtype=bool)
        for a in range(3):
            m &= (X[:, a] > self.lo[a]) & (X[:, a] < self.hi[a])
        return m

    def media(self, X):
        n = X.shape[0]
        return (np.full(n, self.medium.eps), np.full(n, self.medium.mu),
                np.tile(self.medium.e_amp, (n, 1)), np.tile(self.medium.m_amp, (n, 1)))


class Sphere(Box):
    """Sphere, src/geometry.cpp:13-27: squared distance < r*r (CGAL exact
    predicate; ties near the surface are resolved with exact rationals)."""

    def __init__(self, center, radius, medium):
        self.c = [float(v) for v in center]
        self.r2 = radius * radius
        self.medium = medium
        rr = abs(radius)
        self.lo = [v - rr - 1 for v in self.c]
        self.hi = [v + rr + 1 for v in self.c]

    def inside(self, X):
        d2 = ((X[:, 0] - self.c[0]) ** 2 + (X[:, 1] - self.c[1]) ** 2 + (X[:, 2] - self.c[2]) ** 2)
        m = d2 < self.r2
        close = np.abs(d2 - self.r2) <= 1e-9 * max(self.r2, 1.0)
        if close.any():
            r2 = Fraction(self.r2)
            c = [Fraction(v) for v in self.c]
            for q in np.nonzero(close)[0]:
                s = sum((Fraction(int(X[q, a])) - c[a]) ** 2 for a in range(3))
                m[q] = s < r2
        return m


class TwoMediumBox(Box):
    """Two_Medium_Box, src/geometry.cpp:49-76: m1*rho + m2*(1 - rho) with rho the
    zero-padded trilinear interpolant of the density over linspace(box) nodes."""

    def __init__(self, center, size, dim, m1, m2, rho):
        super().__init__(center, size, None)
        self.dim = [int(d) for d in dim]
        self.m1, self.m2 = m1, m2
        self.rho = np.asarray(rho, dtype=float).ravel()
        if self.rho.size != self.dim[0] * self.dim[1] * self.dim[2]:
            raise ValueError("density dataset size does not match dimension")

    def rho_at(self, X):
        qz = (X[:, 2] - self.p1[2]) / (self.p2[2] - self.p1[2]) * (self.dim[2] - 1)
        qy = (X[:, 1] - self.p1[1]) / (self.p2[1] - self.p1[1]) * (self.dim[1] - 1)
        qx = (X[:, 0] - self.p1[0]) / (self.p2[0] - self.p1[0]) * (self.dim[0] - 1)
        dims = (self.dim[2], self.dim[1], self.dim[0])
        return _interp_level(self.rho, dims[0], dims[1] * dims[2], qz, dims[1:], (qz, qy, qx))

    def media(self, X):
        v = self.rho_at(X)
        one_m = 1 - v
        eps = self.m1.eps * v + self.m2.eps * one_m
        mu = self.m1.mu * v + self.m2.mu * one_m
        ea = self.m1.e_amp[None, :] * v[:, None] + self.m2.e_amp[None, :] * one_m[:, None]
        ma = self.m1.m_amp[None, :] * v[:, None] + self.m2.m_amp[None, :] * one_m[:, None]
        return eps, mu, ea, ma


class GeneralMediumBox(TwoMediumBox):
    """General_Medium_Box, src/geometry.cpp:79-107 (media built as
    src/simulation.cpp:439-514): the density interpolant at the point (as
    Two_Medium_Box) is placed on the sampled density function, taken as uniformly
    spaced between its first and last samples,
        q = (rho - f[0]) / (f[-1] - f[0]) * (len(f) - 1),
    and the medium is the zero-padded linear interpolation (interpn<1>,
    include/utility.hpp:753-801) of the sampled media at q, field by field."""

    def __init__(self, center, size, dim, rho, rho_fun, media):
        super().__init__(center, size, dim, None, None, rho)
        self.rho_fun = np.asarray(rho_fun, dtype=float).ravel()
        if self.rho_fun.size == 0 or len(media) != self.rho_fun.size:
            raise ValueError("density function and media tables differ in size")
        self.tab_eps = np.array([m.eps for m in media], dtype=float)
        self.tab_mu = np.array([m.mu for m in media], dtype=float)
        self.tab_ea = np.array([m.e_amp for m in media], dtype=float).reshape(len(media), -1)
        self.tab_ma = np.array([m.m_amp for m in media], dtype=float).reshape(len(media), -1)

    def media(self, X):
        v = self.rho_at(X)
        f = self.rho_fun
        n = f.size
        q = (v - f[0]) / (f[-1] - f[0]) * (n - 1)

        def lin(tab):
            return _interp_level(tab, n, 1, q, (), (q,))

        ea = np.stack([lin(self.tab_ea[:, p]) for p in range(self.tab_ea.shape[1])], axis=1) \
            if self.tab_ea.shape[1] else np.zeros((q.size, 0))
        ma = np.stack([lin(self.tab_ma[:, p]) for p in range(self.tab_ma.shape[1])], axis=1) \
            if self.tab_ma.shape[1] else np.zeros((q.size, 0))
        return lin(self.tab_eps), lin(self.tab_mu), ea, ma


# ----------------------------------------------------------------------------
# The simulation set-up (src/simulation.cpp:127-188) and the C driver
# ----------------------------------------------------------------------------

_LIB = None


_LIB_NAME = "liboracle.so"


def use_timing_build():
    """Run the C loop from oracle/liboracle_timing.so (-O3, x86-64-v3, contraction
    on) instead of the parity build.  For the CPU-baseline timing only: call it
    before the first simulation of the process; results then differ from the
    parity build's by rounding."""
    global _LIB_NAME
    if _LIB is not None and _LIB_NAME != "liboracle_timing.so":
        raise RuntimeError("the parity build of the oracle is already loaded in this process")
    _LIB_NAME = "liboracle_timing.so"


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), _LIB_NAME)
        if not os.path.exists(path):
            raise RuntimeError("oracle/%s is not built; run `make -C oracle`" % _LIB_NAME)
        _LIB = ctypes.CDLL(path)
        _LIB.osim_run_steps.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        _LIB.osim_run_until_time.argtypes = [ctypes.c_void_p, ctypes.c_double]
        _LIB.osim_run_until_time.restype = ctypes.c_int64
        _LIB.osim_sizeof.restype = ctypes.c_size_t
        _LIB.osim_set_threads.argtypes = [ctypes.c_int]
        _LIB.osim_get_threads.restype = ctypes.c_int
    return _LIB


def set_threads(n):
    """Threads of the C hot loop's per-point loops (same bits at any count)."""
    _lib().osim_set_threads(int(n))


def get_threads():
    return int(_lib().osim_get_threads())


P = ctypes.POINTER
c_i64p = P(ctypes.c_int64)
c_dp = P(ctypes.c_double)


class _Pml(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("idx", c_i64p),
                ("b1", c_dp), ("c1", c_dp), ("ik1", c_dp), ("b2", c_dp), ("c2", c_dp), ("ik2", c_dp),
                ("phi1", c_dp), ("phi2", c_dp)]


class _Curl(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("idx", c_i64p)]


class _Dip(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("idx", c_i64p), ("amp", c_dp), ("start", c_dp),
                ("end", c_dp), ("width", c_dp), ("max_end", ctypes.c_double)]


class _Group(ctypes.Structure):
    _fields_ = [("npts", ctypes.c_int64), ("npoles", ctypes.c_int),
                ("c1", c_dp), ("c2", c_dp), ("c3", c_dp), ("idx", c_i64p), ("ginf", c_dp),
                ("amp", c_dp), ("p", c_dp), ("p1", c_dp)]


class _Unit(ctypes.Structure):
    _fields_ = [("is_db", ctypes.c_int), ("nf", ctypes.c_int), ("n", ctypes.c_int64),
                ("idx", c_i64p), ("freqs", c_dp), ("re", c_dp), ("im", c_dp), ("buf", c_dp)]


class _Face(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("src", c_i64p), ("dst", c_i64p), ("mult", ctypes.c_double),
                ("buf", c_dp)]


class _Sim(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("inv_dx", ctypes.c_double), ("stride", ctypes.c_int64 * 3),
                ("eh", c_dp), ("accdb", c_dp),
                ("pml", (_Pml * 3) * 2), ("curl", (_Curl * 3) * 2), ("dip", _Dip * 2),
                ("ngroups", ctypes.c_int * 2), ("groups", P(_Group) * 2),
                ("nunits", ctypes.c_int * 2), ("units", P(_Unit) * 2),
                ("sync", (_Face * 2) * 3),
                ("nprobes", ctypes.c_int), ("probe_off", c_i64p), ("probe_idx", c_i64p),
                ("probe_w", c_dp), ("trace", c_dp), ("trace_cap", ctypes.c_int64),
                ("time_step", ctypes.c_int64)]


class OracleSim:
    """Restatement of ``Simulation::init`` (src/simulation.cpp:127-188) for one
    rank, producing the reference's point lists on its doubled grid.

    ``config``   -- the JSON dict ``ffip.Simulation.get_json`` produces
                    (schema: src/simulation.cpp, keys listed in the reference's
                    ``Simulation::init`` and helpers).
    ``datasets`` -- name -> flat float64 array, replacing the input HDF5 file.
    ``probes``   -- list of (position (physical), component) sampled after each
                    full step with ``Simulation::at`` semantics.
    """

    def __init__(self, config, datasets=None, probes=(), trace_steps=0, strict_pool=True):
        self.cfg = config
        self.strict_pool = strict_pool
        self.datasets = datasets or {}
        self._keep = []
        sc = float(config.get("courant number", 0.5))
        self.resolution = float(config["resolution"])
        self.dx = 1 / self.resolution
        self.dt = self.dx * sc
        self.dx2 = self.dx / 2
        self._set_size(config)
        self._read_bcs(config)
        self._set_grid()
        self.bg = Medium.from_json(config["default material"]) if "default material" in config else Medium(1, 1)
        self.e_pool, self.m_pool = [], []
        self._add_to_pool([self.bg])
        self._set_pmls(config.get("PML", []))
        self._set_geometry(config.get("geometry", []))
        self.turnoff_time = 0.0
        self.dip = {0: [], 1: []}
        for src in config.get("sources", []) or []:
            self._set_source(src)
        self.units = {0: [], 1: []}
        self.outputs = []
        for fo in config.get("fields output", []) or []:
            self._register_dft(fo)
        for half in (0, 1):
            for u in self.units[half]:
                u["freqs"] = sorted(set(u["freqs"]))
        self.probes = list(probes)
        self.trace_steps = int(trace_steps)
        self._build_c()

    # -- src/simulation.cpp:570-581
    def _set_size(self, cfg):
        self.sim_size = [float(v) for v in cfg["size"]]
        span = [int(cround(self.sim_size[a] / self.dx) * 2) for a in range(3)]
        self.sim_p1 = [-(s // 2) for s in span]
        self.sim_p2 = [s // 2 for s in span]

    # -- src/simulation.cpp:583-653
    def _read_bcs(self, cfg):
        self.bcs = [[BC_NONE, BC_NONE] for _ in range(3)]
        for s in cfg.get("symmetry", []) or []:
            d = {"x": 0, "y": 1, "z": 2}[s["direction"]]
            self.sim_p1[d] = 0
            self.bcs[d][0] = BC_PMC if int(s["phase factor"]) == 1 else BC_PEC
        pd = cfg.get("periodic")
        if pd:
            for i in range(3):
                if int(pd[i]):
                    if self.bcs[i][0] in (BC_PMC, BC_PEC):
                        self.bcs[i][1] = self.bcs[i][0]
                    else:
                        self.bcs[i][0] = self.bcs[i][1] = BC_PERIOD

    # -- src/simulation.cpp:525-568 (np = 1)
    def _set_grid(self):
        self.sim_dim = [self.sim_p2[a] - self.sim_p1[a] + 1 for a in range(3)]
        self.grid_p1 = tuple(self.sim_p1)
        self.grid_p2 = tuple(self.sim_p2)
        self.grid = Yee3([p - 1 for p in self.grid_p1], [p + 1 for p in self.grid_p2])
        self.eh = np.zeros(self.grid.size())
        self.accdb = np.zeros(self.grid.size())

    # -- src/structure.cpp:95-124
    def _add_to_pool(self, media):
        for m in media:
            for s in m.e_sus:
                if s not in self.e_pool:
                    self.e_pool.append(s)
            for s in m.m_sus:
                if s not in self.m_pool:
                    self.m_pool.append(s)
        if self.strict_pool and (len(self.e_pool) > 31 or len(self.m_pool) > 31):
            raise ValueError("more than 31 distinct susceptibilities (src/structure.cpp:261)")

    # -- src/structure.cpp:126-153
    def _abstract(self, m):
        ea = np.zeros(len(self.e_pool))
        ma = np.zeros(len(self.m_pool))
        for s, a in zip(m.e_sus, m.e_amp):
            ea[self.e_pool.index(s)] = a
        for s, a in zip(m.m_sus, m.m_amp):
            ma[self.m_pool.index(s)] = a
        return AbstractMedium(m.eps, m.mu, ea, ma)

    # -- src/simulation.cpp:655-726, src/fields.cpp:485-506
    def _set_pmls(self, pmls):
        dt = self.dt
        k = [[1.0] * self.sim_dim[a] for a in range(3)]
        b = [[0.0] * self.sim_dim[a] for a in range(3)]
        c = [[0.0] * self.sim_dim[a] for a in range(3)]
        for p in pmls or []:
            d = float(p["thickness"])
            kmax = float(p["k max"])
            smax = float(p["sigma max"])
            m = int(p["order"])
            dir0, dir1 = 0, 2
            side0, side1 = -1, 1
            if "direction" in p:
                dir0 = dir1 = {"x": 0, "y": 1, "z": 2}[p["direction"]]
            if "side" in p:
                side0 = side1 = {"positive": 1, "negative": -1}[p["side"]]

            def sigma(x):
                return math.pow(x / d, m) * smax

            def kk(x):
                return 1 + (kmax - 1) * math.pow(x / d, m)

            def bb(x):
                return math.exp(-sigma(x) / kk(x) * dt)

            def cc(x):
                return 1 / kk(x) * (bb(x) - 1)

            for di in range(dir0, dir1 + 1):
                for side in range(side0, side1 + 1, 2):
                    L = self.sim_size[di]
                    for i in range(self.sim_p1[di], self.sim_p2[di] + 1):
                        if side == 1 and (i * self.dx2 >= L / 2 - d):
                            x = i * self.dx2 - (L / 2 - d)
                        elif side == -1 and (i * self.dx2 <= -L / 2 + d):
                            x = -L / 2 + d - i * self.dx2
                        else:
                            continue
                        ix = i - self.sim_p1[di]
                        k[di][ix] = kk(x)
                        b[di][ix] = bb(x)
                        c[di][ix] = cc(x)
        self.pml_k, self.pml_b, self.pml_c = [np.array(v) for v in k], [np.array(v) for v in b], [np.array(v) for v in c]
        # Fields::init -> PML_init_helper<Ex..Hz>, include/fields.hpp:233-262
        self.pml_lists = {0: [None] * 3, 1: [None] * 3}
        self.curl_lists = {0: [None] * 3, 1: [None] * 3}
        for name in EH_ORDER:
            ct = CTYPE[name]
            x3 = dir_of_ctype(ct)
            x1, x2 = (x3 + 1) % 3, (x3 + 2) % 3
            half = 0 if is_e_point(ct) else 1
            p1, p2 = component_interior(self.grid.grid_p1, self.grid.grid_p2, ct)
            X = yee_coords(p1, p2, 2)
            idx = self.grid.index(X[:, 0], X[:, 1], X[:, 2])
            i1 = X[:, x1] - self.sim_p1[x1]
            i2 = X[:, x2] - self.sim_p1[x2]
            cx1 = self.pml_c[x1][i1]
            cx2 = self.pml_c[x2][i2]
            is_pml = (cx1 != 0) | (cx2 != 0)
            self.pml_lists[half][x3] = dict(
                idx=idx[is_pml],
                b1=self.pml_b[x1][i1][is_pml], c1=cx1[is_pml], ik1=1 / self.pml_k[x1][i1][is_pml],
                b2=self.pml_b[x2][i2][is_pml], c2=cx2[is_pml], ik2=1 / self.pml_k[x2][i2][is_pml],
            )
            self.curl_lists[half][x3] = dict(idx=idx[~is_pml])

    # -- src/simulation.cpp:359-518, src/structure.cpp:155-249
    def _set_geometry(self, geoms):
        mats = []
        for g in geoms or []:
            if g["type"] != "two medium box":
                mats.append(Medium.from_json(g["medium"]))
            else:
                mats.append(Medium.from_json(g["medium1"]))
                mats.append(Medium.from_json(g["medium2"]))
        self._add_to_pool(mats)
        dx2 = self.dx2
        glist = []
        for g in geoms or []:
            t = g["type"]
            if t == "box":
                glist.append(Box([v / dx2 for v in g["center"]], [v / dx2 for v in g["size"]],
                                 self._abstract(Medium.from_json(g["medium"]))))
            elif t == "sphere":
                glist.append(Sphere([v / dx2 for v in g["center"]], float(g["radius"]) / dx2,
                                    self._abstract(Medium.from_json(g["medium"]))))
            elif t == "two medium box":
                rho = self.datasets[g["density dataset"]]
                glist.append(TwoMediumBox([v / dx2 for v in g["center"]], [v / dx2 for v in g["size"]],
                                          [int(v) for v in g["dimension"]],
                                          self._abstract(Medium.from_json(g["medium1"])),
                                          self._abstract(Medium.from_json(g["medium2"])), rho))
            elif t == "general medium box":
                # the medium with each sample's permittivity, permeability and
                # amplitudes (src/simulation.cpp:498-511)
                base = Medium.from_json(g["medium"])
                eps = np.asarray(self.datasets[g["epsilon dataset"]], dtype=float).ravel()
                mu = np.asarray(self.datasets[g["mu dataset"]], dtype=float).ravel()
                ea = np.asarray(self.datasets[g["electric susceptibility amplitudes dataset"]], dtype=float).ravel()
                ma = np.asarray(self.datasets[g["magnetic susceptibility amplitudes dataset"]], dtype=float).ravel()
                ne, nm = len(base.e_amp), len(base.m_amp)
                media = []
                for i in range(eps.size):
                    m = Medium(eps[i], mu[i])
                    m.e_sus, m.m_sus = list(base.e_sus), list(base.m_sus)
                    m.e_amp = [float(ea[i * ne + j]) for j in range(ne)]
                    m.m_amp = [float(ma[i * nm + j]) for j in range(nm)]
                    media.append(self._abstract(m))
                glist.append(GeneralMediumBox([v / dx2 for v in g["center"]], [v / dx2 for v in g["size"]],
                                              [int(v) for v in g["dimension"]], self.datasets[g["density dataset"]],
                                              self.datasets[g["density function dataset"]], media))
            else:
                raise ValueError("geometry type %r is not supported" % t)
        self.geoms = glist
        bg = self._abstract(self.bg)
        e_c = [abstract_sus(s, self.dt) for s in self.e_pool]
        m_c = [abstract_sus(s, self.dt) for s in self.m_pool]
        self.groups = {0: {}, 1: {}}
        self.geom_index = {}
        for name in EH_ORDER:
            ct = CTYPE[name]
            half = 0 if is_e_point(ct) else 1
            p1, p2 = component_interior(self.grid.grid_p1, self.grid.grid_p2, ct)
            dims = [(p2[a] - p1[a]) // 2 + 1 for a in range(3)]
            if min(dims) <= 0:
                continue
            # first geometry (list order) strictly containing each lattice point,
            # tested only inside each geometry's bounding box
            gidx = np.full((dims[2], dims[1], dims[0]), -1, dtype=np.int32)
            for gi, g in enumerate(glist):
                lo, hi = g.bounds()
                r0 = [max(0, int(math.floor((lo[a] - p1[a]) / 2))) for a in range(3)]
                r1 = [min(dims[a] - 1, int(math.ceil((hi[a] - p1[a]) / 2))) for a in range(3)]
                if any(r0[a] > r1[a] for a in range(3)):
                    continue
                sub = gidx[r0[2]:r1[2] + 1, r0[1]:r1[1] + 1, r0[0]:r1[0] + 1]
                zz = p1[2] + 2 * np.arange(r0[2], r1[2] + 1)
                yy = p1[1] + 2 * np.arange(r0[1], r1[1] + 1)
                xx = p1[0] + 2 * np.arange(r0[0], r1[0] + 1)
                Z, Y, X = np.meshgrid(zz, yy, xx, indexing="ij")
                P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(float)
                free = sub.ravel() == -1
                ins = np.zeros(P.shape[0], dtype=bool)
                ins[free] = g.inside(P[free])
                flat = sub.reshape(-1).copy()
                flat[ins] = gi
                sub[...] = flat.reshape(sub.shape)
            gflat = gidx.ravel()
            self.geom_index[name] = (p1, dims, gidx)
            Xall = yee_coords(p1, p2, 2)
            idx = self.grid.index(Xall[:, 0], Xall[:, 1], Xall[:, 2])
            coefs = e_c if half == 0 else m_c
            npool = len(self.e_pool) if half == 0 else len(self.m_pool)

            def add(sel_idx, ginf, amps):
                nz = np.abs(amps) > 1e-4  # get_non_zeros_from_array, src/structure.cpp:252-260
                mask = np.zeros(amps.shape[0], dtype=np.int64)
                for i in range(npool):
                    mask |= nz[:, i].astype(np.int64) << i
                for mv in np.unique(mask):
                    sel = mask == mv
                    poles = [i for i in range(npool) if (int(mv) >> i) & 1]
                    grp = self.groups[half].setdefault(int(mv), dict(poles=poles, idx=[], ginf=[], amp=[]))
                    grp["idx"].append(sel_idx[sel])
                    grp["ginf"].append(ginf[sel])
                    grp["amp"].append(amps[sel][:, poles] if poles else np.zeros((int(sel.sum()), 0)))

            for src_g in [-1] + list(range(len(glist))):
                sel = gflat == src_g
                m = int(sel.sum())
                if m == 0:
                    continue
                g = glist[src_g] if src_g >= 0 else None
                if isinstance(g, TwoMediumBox):
                    eps, mu, ea, ma = g.media(Xall[sel].astype(float))
                else:
                    med = bg if g is None else g.medium
                    eps, mu = np.full(m, med.eps), np.full(m, med.mu)
                    ea = np.tile(med.e_amp, (m, 1))
                    ma = np.tile(med.m_amp, (m, 1))
                add(idx[sel], eps if half == 0 else mu, ea if half == 0 else ma)
        for half in (0, 1):
            coefs = e_c if half == 0 else m_c
            for mv, grp in self.groups[half].items():
                grp["idx"] = np.concatenate(grp["idx"])
                grp["ginf"] = np.concatenate(grp["ginf"])
                grp["amp"] = np.ascontiguousarray(np.concatenate(grp["amp"], axis=0))
                grp["c"] = np.array([coefs[i] for i in grp["poles"]]).reshape(-1, 3)

    # -- src/fields.cpp:183-207
    def geometry_info(self, comp, geom):
        """Simulation::geometry_info_request (src/simulation.cpp:190-233): the
        points of component `comp` whose first containing geometry is `geom`, in
        Yee_Iterator order (x fastest), as coordinate * dx/2 stored in the
        reference's std::vector<int> (truncated towards zero) and written as doubles."""
        if comp not in self.geom_index:
            return np.zeros(0), np.zeros(0), np.zeros(0)
        p1, dims, gidx = self.geom_index[comp]
        k, j, i = np.nonzero(gidx == geom)  # C order: z slowest, x fastest
        out = []
        for a, n in ((0, i), (1, j), (2, k)):
            c = (p1[a] + 2 * n).astype(np.float64) * self.dx2
            out.append(np.trunc(c))
        return tuple(out)

    def _add_dipole_g1(self, pos, ct, amp, start, freq, cutoff):
        amp = -amp
        base = self.grid.base_point(pos, ct)
        s = [(pos[a] - base[a]) / 2 for a in range(3)]
        w = interp_weights(*s)
        half = 0 if is_e_point(ct) else 1
        width = 1 / (2 * PI * freq)
        end = start + 2 * width * cutoff
        if end > self.turnoff_time:
            self.turnoff_time = end
        wi = 0
        for k in (0, 2):
            for j in (0, 2):
                for i in (0, 2):
                    p = (base[0] + i, base[1] + j, base[2] + k)
                    if self.grid.is_inside(p) and w[wi] > 1e-4:
                        self.dip[half].append((int(self.grid.index(*p)), amp * w[wi], start, end, width))
                    wi += 1

    # -- src/simulation.cpp:235-357
    def _set_source(self, src):
        t = src["type"]
        dx2 = self.dx2
        center = [float(v) / dx2 for v in src["center"]]
        size = [float(v) / dx2 for v in src["size"]]
        ct = CTYPE[src["field component"]]
        fn = src["function"]
        if fn["type"] != "Gaussian1":
            raise ValueError("only Gaussian1 source functions are supported (Gaussian2 is defective "
                             "in the reference: src/fields.cpp:229-230)")
        freq = float(fn["frequency"])
        cutoff = float(fn["cutoff"])
        stime = float(fn["start time"])
        p1 = [center[a] - size[a] / 2 for a in range(3)]
        if t == "volume source":
            amp = float(src["amplitude"])
            count = [int(cround(size[a] / 2)) for a in range(3)]
            dlen = [0.0] * 3
            for a in range(3):
                if count[a] != 0:
                    dlen[a] = size[a] / count[a]
                    amp *= dlen[a] / 2
            for q in yee_coords((0, 0, 0), count, 1):
                p = [p1[a] + int(q[a]) * dlen[a] for a in range(3)]
                self._add_dipole_g1(p, ct, amp, stime, freq, cutoff)
        elif t == "inhomogeneous source":
            dim = [int(v) for v in src["dimension"]]
            fcen = float(src["frequency"])
            amp = np.asarray(self.datasets[src["amplitude dataset"]], dtype=float).ravel()
            phase = np.asarray(self.datasets[src["phase dataset"]], dtype=float).ravel()
            dlen = [0.0] * 3
            for a in range(3):
                if dim[a] > 1:
                    dlen[a] = size[a] / (dim[a] - 1)
            for index, q in enumerate(yee_coords((0, 0, 0), [d - 1 for d in dim], 1)):
                p = [p1[a] + int(q[a]) * dlen[a] for a in range(3)]
                self._add_dipole_g1(p, ct, float(amp[index]), stime - float(phase[index]) / (2 * PI * fcen),
                                    freq, cutoff)
        else:
            raise ValueError("Unknown source type %r" % t)

    # -- src/simulation.cpp:103-118, src/dft.cpp:253-271, 163-180
    def _register_dft(self, fo):
        if fo["type"] != "volume fields dft":
            raise ValueError("Unable to read fields_output configuration")
        dx2 = self.dx2
        center = [float(v) / dx2 for v in fo["center"]]
        size = [float(v) / dx2 for v in fo["size"]]
        p1 = [center[a] - size[a] / 2 for a in range(3)]
        p2 = [center[a] + size[a] / 2 for a in range(3)]
        ct = CTYPE[fo["field component"]]
        q1, q2 = component_closure(p1, p2, ct)
        half = 0 if is_e_point(ct) else 1
        unit = None
        for u in self.units[half]:
            if u["ct"] == ct and u["p1"] == q1 and u["p2"] == q2:
                unit = u
        if unit is None:
            g1, g2 = self.grid.grid_p1, self.grid.grid_p2
            if all(q1[a] <= g2[a] and g1[a] <= q2[a] for a in range(3)):
                i1 = [max(q1[a], g1[a]) for a in range(3)]
                i2 = [min(q2[a], g2[a]) for a in range(3)]
                l1, l2 = component_interior(i1, i2, ct)
            else:
                l1 = l2 = None
            unit = dict(ct=ct, p1=q1, p2=q2, l1=l1, l2=l2, freqs=[])
            self.units[half].append(unit)
        unit["freqs"] += [float(f) for f in fo["frequency"]]
        self.outputs.append(dict(group=fo["output group"], unit=unit, half=half,
                                 freqs=[float(f) for f in fo["frequency"]], p1=q1, p2=q2))

    # -- boundary copies, src/fields.cpp:327-482 (single rank)
    def _face(self, axis, coord):
        p1 = list(self.grid.grid_p1)
        p2 = list(self.grid.grid_p2)
        p1[axis] = p2[axis] = coord
        X = yee_coords(p1, p2, 1)
        return self.grid.index(X[:, 0], X[:, 1], X[:, 2])

    def _build_sync(self):
        ops = [[None, None] for _ in range(3)]
        g1, g2 = self.grid.grid_p1, self.grid.grid_p2
        for a in range(3):
            neg, pos = self.bcs[a]
            src_neg = self._face(a, g1[a] + 1)
            src_pos = self._face(a, g2[a] - 1)
            dst_neg = self._face(a, g1[a] - 1)
            dst_pos = self._face(a, g2[a] + 1)
            if neg == BC_PERIOD:
                ops[a][0] = (src_pos, dst_neg, 1.0)
            elif neg in (BC_PEC, BC_PMC):
                phase = -1 if neg == BC_PEC else 1
                ops[a][0] = (src_neg, dst_neg, float(phase if (g1[a] & 1) else -phase))
            if pos == BC_PERIOD:
                ops[a][1] = (src_neg, dst_pos, 1.0)
            elif pos in (BC_PEC, BC_PMC):
                phase = -1 if pos == BC_PEC else 1
                ops[a][1] = (src_pos, dst_pos, float(phase if (g2[a] & 1) else -phase))
        return ops

    # ------------------------------------------------------------------ C side
    def _ptr(self, arr, dtype):
        a = np.ascontiguousarray(arr, dtype=dtype)
        self._keep.append(a)
        return a.ctypes.data_as(c_i64p if dtype == np.int64 else c_dp), a

    def _build_c(self):
        s = _Sim()
        s.dt = self.dt
        s.inv_dx = 1 / self.dx
        for a in range(3):
            s.stride[a] = self.grid.stride[a]
        s.eh = self.eh.ctypes.data_as(c_dp)
        s.accdb = self.accdb.ctypes.data_as(c_dp)
        self.phis = {}
        for half in (0, 1):
            for d in range(3):
                L = self.pml_lists[half][d]
                n = L["idx"].size
                pl = _Pml()
                pl.n = n
                pl.idx, _ = self._ptr(L["idx"], np.int64)
                for key in ("b1", "c1", "ik1", "b2", "c2", "ik2"):
                    setattr(pl, key, self._ptr(L[key], np.float64)[0])
                pl.phi1, a1 = self._ptr(np.zeros(n), np.float64)
                pl.phi2, a2 = self._ptr(np.zeros(n), np.float64)
                self.phis[(half, d)] = (a1, a2)
                s.pml[half][d] = pl
                cl = _Curl()
                cl.n = self.curl_lists[half][d]["idx"].size
                cl.idx, _ = self._ptr(self.curl_lists[half][d]["idx"], np.int64)
                s.curl[half][d] = cl
            dl = _Dip()
            D = self.dip[half]
            dl.n = len(D)
            arr = np.array(D, dtype=float).reshape(-1, 5)
            dl.idx, _ = self._ptr(arr[:, 0].astype(np.int64), np.int64)
            dl.amp, _ = self._ptr(arr[:, 1], np.float64)
            dl.start, _ = self._ptr(arr[:, 2], np.float64)
            dl.end, _ = self._ptr(arr[:, 3], np.float64)
            dl.width, _ = self._ptr(arr[:, 4], np.float64)
            dl.max_end = float(arr[:, 3].max()) if len(D) else 0.0
            s.dip[half] = dl
            groups = list(self.groups[half].values())
            garr = (_Group * max(1, len(groups)))()
            self.state = getattr(self, "state", {})
            for gi, g in enumerate(groups):
                G = garr[gi]
                npts = g["idx"].size
                npol = len(g["poles"])
                G.npts = npts
                G.npoles = npol
                G.c1, _ = self._ptr(g["c"][:, 0], np.float64)
                G.c2, _ = self._ptr(g["c"][:, 1], np.float64)
                G.c3, _ = self._ptr(g["c"][:, 2], np.float64)
                G.idx, _ = self._ptr(g["idx"], np.int64)
                G.ginf, _ = self._ptr(g["ginf"], np.float64)
                G.amp, _ = self._ptr(g["amp"].ravel(), np.float64)
                G.p, _ = self._ptr(np.zeros(npts * npol), np.float64)
                G.p1, _ = self._ptr(np.zeros(npts * npol), np.float64)
            self._keep.append(garr)
            s.ngroups[half] = len(groups)
            s.groups[half] = ctypes.cast(garr, P(_Group))
            units = self.units[half]
            uarr = (_Unit * max(1, len(units)))()
            for ui, u in enumerate(units):
                U = uarr[ui]
                U.is_db = 0 if is_eh_point(u["ct"]) else 1
                U.nf = len(u["freqs"])
                if u["l1"] is None:
                    X = np.zeros((0, 3), dtype=np.int64)
                else:
                    X = yee_coords(u["l1"], u["l2"], 2)
                u["local_coords"] = X
                n = X.shape[0]
                U.n = n
                U.idx, _ = self._ptr(self.grid.index(X[:, 0], X[:, 1], X[:, 2]), np.int64)
                U.freqs, _ = self._ptr(np.array(u["freqs"]), np.float64)
                U.re, u["re"] = self._ptr(np.zeros(U.nf * n), np.float64)
                U.im, u["im"] = self._ptr(np.zeros(U.nf * n), np.float64)
                U.buf, _ = self._ptr(np.zeros(n), np.float64)
            self._keep.append(uarr)
            s.nunits[half] = len(units)
            s.units[half] = ctypes.cast(uarr, P(_Unit))
        ops = self._build_sync()
        for a in range(3):
            for side in range(2):
                F = _Face()
                op = ops[a][side]
                if op is None:
                    F.n = 0
                else:
                    F.n = op[0].size
                    F.src, _ = self._ptr(op[0], np.int64)
                    F.dst, _ = self._ptr(op[1], np.int64)
                    F.mult = op[2]
                    F.buf, _ = self._ptr(np.zeros(op[0].size), np.float64)
                s.sync[a][side] = F
        offs, idxs, ws = [0], [], []
        for pos, comp in self.probes:
            terms = self.grid.interp_terms([float(v) / self.dx2 for v in pos], CTYPE[comp])
            for i, w in terms:
                idxs.append(i)
                ws.append(w)
            offs.append(len(idxs))
        s.nprobes = len(self.probes)
        s.probe_off, _ = self._ptr(np.array(offs), np.int64)
        s.probe_idx, _ = self._ptr(np.array(idxs, dtype=np.int64), np.int64)
        s.probe_w, _ = self._ptr(np.array(ws, dtype=float), np.float64)
        self.trace = np.zeros((max(1, self.trace_steps), max(1, len(self.probes))))
        s.trace = self.trace.ctypes.data_as(c_dp)
        s.trace_cap = self.trace_steps
        s.time_step = 0
        self._sim = s

    # ------------------------------------------------------------------ running
    @property
    def time_step(self):
        return int(self._sim.time_step)

    def run_steps(self, n):
        _lib().osim_run_steps(ctypes.byref(self._sim), int(n))

    def run_until_time(self, t):
        """Simulation::run_until_time, src/simulation.cpp:768-780."""
        return int(_lib().osim_run_until_time(ctypes.byref(self._sim), float(t)))

    def run(self, stop):
        """Simulation::run dispatch, src/simulation.cpp:903-949."""
        t0 = self.time_step
        if stop["type"] == "time":
            self.run_until_time(float(stop["time"]))
        elif stop["type"] == "decay":
            self.run_until_fields_decayed(float(stop["time interval examined"]), stop["position"],
                                          stop["field component"], float(stop["decayed by"]))
        else:
            self.run_until_dft(stop["center"], stop["size"], stop["field component"], float(stop["frequency"]),
                               float(stop["time interval examined"]), float(stop["variation"]))
        return self.time_step - t0

    def run_until_dft(self, center, size, comp, freq, interval, var):
        """src/simulation.cpp:782-842 (single rank: the all-reduce is the identity)."""
        dx2 = self.dx2
        p1 = [(float(center[a]) - float(size[a]) / 2) / dx2 for a in range(3)]
        p2 = [(float(center[a]) + float(size[a]) / 2) / dx2 for a in range(3)]
        ct = CTYPE[comp]
        q1, q2 = component_closure(p1, p2, ct)
        half = 0 if is_e_point(ct) else 1
        unit = None
        for u in self.units[half]:
            if u["ct"] == ct and u["p1"] == q1 and u["p2"] == q2:
                unit = u
        if unit is None:
            raise ValueError("dft is not registered")
        if freq not in unit["freqs"]:
            raise ValueError("frequency is not registered")
        fi = unit["freqs"].index(freq)
        n = unit["local_coords"].shape[0]

        def norm():  # DFT_Unit::get_local_norm, src/dft.cpp:93-107
            re = unit["re"][fi * n:(fi + 1) * n]
            im = unit["im"][fi * n:(fi + 1) * n]
            return float(np.sum(re * re + im * im))

        turnoff = self.turnoff_time
        while self.time_step * self.dt < turnoff:
            self.run_steps(1)
        extension = 1
        self.stop_log = []
        while True:
            mx, mn = np.finfo(float).tiny, np.finfo(float).max
            while self.time_step * self.dt < turnoff + extension * interval:
                self.run_steps(1)
                v = norm()
                mx, mn = max(mx, v), min(mn, v)
            extension += 1
            cur = (mx - mn) / mx
            self.stop_log.append((mx, mn, cur))
            if not cur > var:
                break

    def run_until_fields_decayed(self, interval, pos, comp, decay_by):
        """src/simulation.cpp:844-901."""
        terms = self.grid.interp_terms([float(v) / self.dx2 for v in pos], CTYPE[comp])

        def value():
            return sum(self.eh[i] * w for i, w in terms)

        turnoff = self.turnoff_time
        max_sqr = 0.0
        while self.time_step * self.dt < turnoff:
            self.run_steps(1)
            max_sqr = max(max_sqr, value() ** 2)
        extension = 1
        self.stop_log = []
        while True:
            cur = 0.0
            while self.time_step * self.dt < turnoff + extension * interval:
                self.run_steps(1)
                cur = max(cur, value() ** 2)
            extension += 1
            max_sqr = max(max_sqr, cur)
            decay = cur / max_sqr
            self.stop_log.append((cur, max_sqr, decay))
            if not decay > decay_by:
                break

    def dft_output(self, k):
        """Volume_Fields_DFT::output -> output_fields_collective (src/dft.cpp:273-320):
        real/imag [nf, nz, ny, nx] over the closure (zeros outside the grid), plus
        x, y, z, f."""
        o = self.outputs[k]
        u = o["unit"]
        ct = u["ct"]
        dims = [(o["p2"][a] - o["p1"][a]) // 2 + 1 for a in range(3)]
        nf = len(o["freqs"])
        re = np.zeros((nf, dims[2], dims[1], dims[0]))
        im = np.zeros_like(re)
        if u["l1"] is not None:
            ld = [(u["l2"][a] - u["l1"][a]) // 2 + 1 for a in range(3)]
            off = [(u["l1"][a] - o["p1"][a]) // 2 for a in range(3)]
            n = ld[0] * ld[1] * ld[2]
            for fi, f in enumerate(o["freqs"]):
                j = u["freqs"].index(f)
                sl = (fi, slice(off[2], off[2] + ld[2]), slice(off[1], off[1] + ld[1]), slice(off[0], off[0] + ld[0]))
                re[sl] = u["re"][j * n:(j + 1) * n].reshape(ld[2], ld[1], ld[0])
                im[sl] = u["im"][j * n:(j + 1) * n].reshape(ld[2], ld[1], ld[0])
        dx2 = self.dx2
        x = linspace(o["p1"][0] * dx2, o["p2"][0] * dx2, dims[0])
        y = linspace(o["p1"][1] * dx2, o["p2"][1] * dx2, dims[1])
        z = linspace(o["p1"][2] * dx2, o["p2"][2] * dx2, dims[2])
        return dict(real=re, imag=im, x=np.array(x), y=np.array(y), z=np.array(z), f=np.array(o["freqs"]))

    def field(self, comp):
        """Raw component values on its lattice inside the grid, [nz, ny, nx]."""
        ct = CTYPE[comp]
        p1, p2 = component_interior(self.grid.grid_p1, self.grid.grid_p2, ct)
        X = yee_coords(p1, p2, 2)
        dims = [(p2[a] - p1[a]) // 2 + 1 for a in range(3)]
        return self.eh[self.grid.index(X[:, 0], X[:, 1], X[:, 2])].reshape(dims[2], dims[1], dims[0])


# ----------------------------------------------------------------------------
# Density filters of filtered-design scripts (app/ffip.cpp:222-306)
# ----------------------------------------------------------------------------

def to_convolve(p1, p2):
    """TO_convolve, app/ffip.cpp:222-262: window-normalised correlation; loops and
    summation order as there (both filter axes indexed with the first radius,
    app/ffip.cpp:248, so only square filters are defined)."""
    p1 = np.asarray(p1, dtype=float)
    p2 = np.asarray(p2, dtype=float)
    r0, r1 = p2.shape[0] // 2, p2.shape[1] // 2
    n0, n1 = p1.shape
    out = np.zeros_like(p1)
    for i in range(n0):
        for j in range(n1):
            w = 0.0
            s = 0.0
            for m in range(max(i - r0, 0), min(i + r0 + 1, n0)):
                for n in range(max(j - r1, 0), min(j + r1 + 1, n1)):
                    w += p2[m - i + r0, n - j + r0]
                    s += p1[m, n] * p2[m - i + r0, n - j + r0]
            out[i, j] = s / w
    return out


def to_convolve_transpose(p1, p2):
    """TO_convolve_transpose, app/ffip.cpp:264-306: the scatter of p1(i, j) times
    the filter over the window of (i, j), divided by that window's weight."""
    p1 = np.asarray(p1, dtype=float)
    p2 = np.asarray(p2, dtype=float)
    r0, r1 = p2.shape[0] // 2, p2.shape[1] // 2
    n0, n1 = p1.shape
    out = np.zeros_like(p1)
    for i in range(n0):
        for j in range(n1):
            lo0, lo1 = max(i - r0, 0), max(j - r1, 0)
            hi0, hi1 = min(i + r0 + 1, n0), min(j + r1 + 1, n1)
            w = 0.0
            for m in range(lo0, hi0):
                for n in range(lo1, hi1):
                    w += p2[m - i + r0, n - j + r0]
            for m in range(lo0, hi0):
                for n in range(lo1, hi1):
                    out[m, n] += p1[i, j] * p2[m - i + r0, n - j + r0] / w
    return out
