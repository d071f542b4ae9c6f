"""ctypes binding of libffip_hip.so (include/ffip_hip.h).

The library is the product path: there is no CPU fallback.  Loading fails loudly
when the shared object has not been built, and every failing call raises
``FfipError`` with the library's message.
"""
from __future__ import annotations

import ctypes
import json
import os
from typing import Dict, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libffip_hip.so")

_lib = None


class FfipError(RuntimeError):
    pass


_SIGS = {
    "ffip_last_error": ([], ctypes.c_char_p),
    "ffip_sim_create": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "ffip_sim_set_dataset": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t], ctypes.c_int),
    "ffip_sim_add_probe": ([ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_char_p],
                           ctypes.c_int),
    "ffip_sim_finalize_setup": ([ctypes.c_void_p], ctypes.c_int),
    "ffip_sim_run": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "ffip_sim_step": ([ctypes.c_void_p, ctypes.c_int64], ctypes.c_int),
    "ffip_sim_time_step": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "ffip_sim_dft_count": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "ffip_sim_dft_dims": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "ffip_sim_dft_group": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t], ctypes.c_int),
    "ffip_sim_get_dft": ([ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 6, ctypes.c_int),
    "ffip_sim_set_trace_capacity": ([ctypes.c_void_p, ctypes.c_int64], ctypes.c_int),
    "ffip_sim_get_trace": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                            ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "ffip_sim_probe": ([ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_char_p,
                        ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "ffip_sim_component_dims": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p], ctypes.c_int),
    "ffip_sim_get_component": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p], ctypes.c_int),
    "ffip_sim_geom_info_count": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "ffip_sim_get_geom_info": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
                               ctypes.c_int),
    "ffip_sim_step_timed": ([ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "ffip_sim_info": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "ffip_sim_traffic": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "ffip_sim_plan": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)],
                      ctypes.c_int),
    "ffip_adjoint_gradient": ([ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_void_p] * 2 + [ctypes.c_double] +
                              [ctypes.c_void_p] * 6, ctypes.c_int),
    "ffip_transpose": ([ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                        ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p], ctypes.c_int),
    "ffip_sim_destroy": ([ctypes.c_void_p], None),
    "ffip_to_convolve_noalloc": ([ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                  ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "ffip_stream_bandwidth": ([ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p],
                              ctypes.c_int),
    "ffip_to_convolve": ([ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                          ctypes.c_int64, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    # surrogate kernels (device pointers, stream as void*)
    "ffip_sg_gemm": ([ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                      ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                      ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "ffip_sg_gemm_ws": ([ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                         ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                         ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p], ctypes.c_int),
    "ffip_sg_im2col": ([ctypes.c_void_p] * 4, ctypes.c_int),
    "ffip_sg_col2im": ([ctypes.c_void_p] * 4, ctypes.c_int),
    "ffip_sg_bias_act": ([ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "ffip_sg_act_back": ([ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                          ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "ffip_sg_colsum": ([ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                        ctypes.c_void_p], ctypes.c_int),
    "ffip_sg_to_bf16": ([ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "ffip_sg_head": ([ctypes.c_void_p] * 4, ctypes.c_int),
    "ffip_sg_head_back": ([ctypes.c_void_p] * 4, ctypes.c_int),
    "ffip_sg_adam": ([ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
}


def lib():
    """Load libffip_hip.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FfipError("libffip_hip.so is not built (run `python -m ffip_amd.build`); "
                            "ffip_amd has no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().ffip_last_error()
        raise FfipError("ffip error %d: %s" % (rc, msg.decode() if msg else "?"))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def stream_bandwidth(reads: int, writes: int, bytes_per_array: int = 1 << 30, reps: int = 10):
    """(mean, best) GB/s of a plain streaming kernel reading `reads` and writing
    `writes` arrays of `bytes_per_array` bytes on the current device."""
    out = np.zeros(2)
    _check(lib().ffip_stream_bandwidth(int(reads), int(writes), int(bytes_per_array), int(reps), _ptr(out)))
    return float(out[0]), float(out[1])


def transpose(gx, gy, gz, pt):
    """Adjoint of zero-padded trilinear interpolation onto the grid (gx, gy, gz):
    pt[:, :3] are positions, pt[:, 3] values; returns [nz, ny, nx].  Same
    contract as the reference's _ffip.transpose (app/ffip.cpp:308-352): KeyError
    for bad grid arrays, ValueError when pt is not (n, 4)."""
    gx, gy, gz = (np.ascontiguousarray(np.asarray(g, dtype=np.float64)) for g in (gx, gy, gz))
    pt = np.ascontiguousarray(np.asarray(pt, dtype=np.float64))
    if gx.ndim != 1 or gy.ndim != 1 or gz.ndim != 1 or pt.ndim != 2:
        raise KeyError("some dimensions are not right")
    if gx.size < 1 or gy.size < 1 or gz.size < 1:
        raise KeyError("some grid size is 0")
    if pt.shape[1] != 4:
        raise ValueError("shape(1) of pts is not 4")
    out = np.zeros((gz.size, gy.size, gx.size))
    _check(lib().ffip_transpose(_ptr(gx), gx.size, _ptr(gy), gy.size, _ptr(gz), gz.size, _ptr(pt), pt.shape[0],
                                _ptr(out)))
    return out


def _to_convolve(in1, in2, transpose):
    a = np.ascontiguousarray(np.asarray(in1, dtype=np.float64))
    f = np.ascontiguousarray(np.asarray(in2, dtype=np.float64))
    if a.ndim != 2 or f.ndim != 2:
        raise KeyError("dimensions are not 2")
    if f.shape[0] % 2 == 0 or f.shape[1] % 2 == 0:
        raise KeyError("filter shape is not odd")
    if f.shape[0] != f.shape[1]:
        raise KeyError("filter is not square (the reference indexes both axes with the first radius)")
    out = np.empty_like(a)
    _check(lib().ffip_to_convolve(_ptr(a), a.shape[0], a.shape[1], _ptr(f), f.shape[0], f.shape[1], int(transpose),
                                  _ptr(out)))
    return out


def TO_convolve(in1, in2):
    """Window-normalised 2-D correlation of in1 with the odd square filter in2
    (the reference's _ffip.TO_convolve, app/ffip.cpp:222-262), on the GPU."""
    return _to_convolve(in1, in2, False)


def TO_convolve_transpose(in1, in2):
    """Adjoint of TO_convolve (app/ffip.cpp:264-306), on the GPU."""
    return _to_convolve(in1, in2, True)


def adjoint_gradient(fwd: "LibSim", adj: "LibSim", fwd_out, adj_out, freq, eps_prime, norm, center, size, dim):
    """ffip_adjoint_gradient: gradient on the density nodes [nz, ny, nx]."""
    fo = np.ascontiguousarray(fwd_out, dtype=np.int32)
    ao = np.ascontiguousarray(adj_out, dtype=np.int32)
    ep = np.array([complex(eps_prime).real, complex(eps_prime).imag])
    nm = np.array([complex(norm).real, complex(norm).imag])
    c = np.array([center[0], center[1], center[2]], dtype=np.float64)
    s = np.array([size[0], size[1], size[2]], dtype=np.float64)
    d = np.array([int(round(dim[0])), int(round(dim[1])), int(round(dim[2]))], dtype=np.int64)
    out = np.zeros((d[2], d[1], d[0]))
    _check(lib().ffip_adjoint_gradient(fwd._h, adj._h, _ptr(fo), _ptr(ao), float(freq), _ptr(ep), _ptr(nm), _ptr(c),
                                       _ptr(s), _ptr(d), _ptr(out)))
    return out


def geometry_info(config, datasets: Optional[Dict[str, np.ndarray]] = None):
    """Host-only answers to the configuration's geometry info requests
    (ffip_sim_get_geom_info before finalize; needs no GPU): [(x, y, z), ...]."""
    L = lib()
    text = config if isinstance(config, str) else json.dumps(config)
    cfg = json.loads(text)
    h = ctypes.c_void_p()
    _check(L.ffip_sim_create(text.encode(), ctypes.byref(h)))
    try:
        for name, arr in (datasets or {}).items():
            a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel())
            _check(L.ffip_sim_set_dataset(h, name.encode(), _ptr(a), a.size))
        out = []
        for k in range(len(cfg.get("geometry info request", []))):
            n = ctypes.c_int64()
            _check(L.ffip_sim_geom_info_count(h, k, ctypes.byref(n)))
            x, y, z = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value)
            _check(L.ffip_sim_get_geom_info(h, k, _ptr(x), _ptr(y), _ptr(z)))
            out.append((x, y, z))
        return out
    finally:
        L.ffip_sim_destroy(h)


def plan(config, datasets: Optional[Dict[str, np.ndarray]] = None, probes: Sequence = ()):
    """Host-only set-up summary (ffip_sim_plan); needs no GPU."""
    L = lib()
    text = config if isinstance(config, str) else json.dumps(config)
    h = ctypes.c_void_p()
    _check(L.ffip_sim_create(text.encode(), ctypes.byref(h)))
    try:
        for name, arr in (datasets or {}).items():
            a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel())
            _check(L.ffip_sim_set_dataset(h, name.encode(), _ptr(a), a.size))
        for pos, comp in probes:
            _check(L.ffip_sim_add_probe(h, float(pos[0]), float(pos[1]), float(pos[2]), comp.encode()))
        need = ctypes.c_size_t()
        _check(L.ffip_sim_plan(h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(L.ffip_sim_plan(h, buf, need.value, None))
        return json.loads(buf.value.decode())
    finally:
        L.ffip_sim_destroy(h)


class LibSim:
    """One simulation handle of the C ABI."""

    def __init__(self, config, datasets: Optional[Dict[str, np.ndarray]] = None,
                 probes: Sequence = (), trace_steps: int = 0):
        L = lib()
        text = config if isinstance(config, str) else json.dumps(config)
        h = ctypes.c_void_p()
        _check(L.ffip_sim_create(text.encode(), ctypes.byref(h)))
        self._h = h
        for name, arr in (datasets or {}).items():
            a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel())
            _check(L.ffip_sim_set_dataset(self._h, name.encode(), _ptr(a), a.size))
        for pos, comp in probes:
            _check(L.ffip_sim_add_probe(self._h, float(pos[0]), float(pos[1]), float(pos[2]), comp.encode()))
        if trace_steps:
            _check(L.ffip_sim_set_trace_capacity(self._h, int(trace_steps)))
        _check(L.ffip_sim_finalize_setup(self._h))
        self.nprobes = len(probes)

    def close(self):
        if getattr(self, "_h", None):
            lib().ffip_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stepping
    def run(self, stop) -> int:
        n = ctypes.c_int64()
        text = stop if isinstance(stop, str) else json.dumps(stop)
        _check(lib().ffip_sim_run(self._h, text.encode(), ctypes.byref(n)))
        return int(n.value)

    def step(self, n: int):
        _check(lib().ffip_sim_step(self._h, int(n)))

    @property
    def time_step(self) -> int:
        n = ctypes.c_int64()
        _check(lib().ffip_sim_time_step(self._h, ctypes.byref(n)))
        return int(n.value)

    def step_timed(self, n: int, per_kernel: bool = True):
        out = np.zeros(3)
        _check(lib().ffip_sim_step_timed(self._h, int(n), 1 if per_kernel else 0, _ptr(out)))
        return dict(total_ms=out[0], e_ms=out[1], h_ms=out[2])

    def info(self):
        out = np.zeros(6)
        _check(lib().ffip_sim_info(self._h, _ptr(out)))
        return dict(cells=(out[0], out[1], out[2]), device_bytes=out[3], dt=out[4], dx=out[5])

    @property
    def fused(self) -> bool:
        """Whether each step runs as one fused kernel (traffic(2) is then defined)."""
        out = np.zeros(7)
        return lib().ffip_sim_traffic(self._h, 2, _ptr(out)) == 0

    def traffic(self, half: int):
        out = np.zeros(7)
        _check(lib().ffip_sim_traffic(self._h, int(half), _ptr(out)))
        keys = ("fields", "material", "polarization", "cpml", "dft", "total", "points")
        return dict(zip(keys, out.tolist()))

    # -- results
    def dft_count(self) -> int:
        n = ctypes.c_int()
        _check(lib().ffip_sim_dft_count(self._h, ctypes.byref(n)))
        return n.value

    def dft(self, k: int):
        dims = np.zeros(4, dtype=np.int64)
        _check(lib().ffip_sim_dft_dims(self._h, int(k), _ptr(dims)))
        nf, nz, ny, nx = (int(v) for v in dims)
        re = np.zeros((nf, nz, ny, nx))
        im = np.zeros_like(re)
        x, y, z, f = np.zeros(nx), np.zeros(ny), np.zeros(nz), np.zeros(nf)
        _check(lib().ffip_sim_get_dft(self._h, int(k), _ptr(re), _ptr(im), _ptr(x), _ptr(y), _ptr(z), _ptr(f)))
        buf = ctypes.create_string_buffer(256)
        _check(lib().ffip_sim_dft_group(self._h, int(k), buf, 256))
        return dict(real=re, imag=im, x=x, y=y, z=z, f=f, group=buf.value.decode())

    def trace(self):
        n = ctypes.c_int64()
        p = ctypes.c_int()
        _check(lib().ffip_sim_get_trace(self._h, None, 0, ctypes.byref(n), ctypes.byref(p)))
        out = np.zeros((int(n.value), max(1, p.value)))
        if n.value and p.value:
            _check(lib().ffip_sim_get_trace(self._h, _ptr(out), int(n.value), None, None))
        return out[:, :p.value]

    def probe(self, pos, comp: str) -> float:
        v = ctypes.c_double()
        _check(lib().ffip_sim_probe(self._h, float(pos[0]), float(pos[1]), float(pos[2]), comp.encode(),
                                    ctypes.byref(v)))
        return v.value

    def component(self, comp: str) -> np.ndarray:
        dims = np.zeros(3, dtype=np.int64)
        _check(lib().ffip_sim_component_dims(self._h, comp.encode(), _ptr(dims)))
        out = np.zeros(tuple(int(d) for d in dims))
        _check(lib().ffip_sim_get_component(self._h, comp.encode(), _ptr(out)))
        return out

    def geom_info(self, k: int):
        n = ctypes.c_int64()
        _check(lib().ffip_sim_geom_info_count(self._h, int(k), ctypes.byref(n)))
        x, y, z = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value)
        _check(lib().ffip_sim_get_geom_info(self._h, int(k), _ptr(x), _ptr(y), _ptr(z)))
        return x, y, z
