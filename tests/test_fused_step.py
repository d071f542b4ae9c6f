"""The fused full-step kernel against the two half-step kernels.

Both paths compute every value with the same expressions in the same order
(the fused kernel forms the magnetic half of a neighbouring tile's rim itself),
so FP64 storage must give the same fields, DFT monitors and probe traces to
1e-13 relative L2: what remains is the compiler fusing a multiply and an add in
one path and not in the other.  With FP32 storage such a last-bit difference
can flip the rounding of a stored value, so FP32 is held to 1e-6.  The cases cover the
polarization layouts (full grid, compact), per-point media, the CPML shell,
odd grids and z segments shorter than the grid, and the FP64 fused path is also
checked against the oracle."""
import os

import numpy as np
import pytest

from ffip_amd import workloads as W
from ffip_amd._lib import LibSim
from oracle.oracle import OracleSim

from .conftest import rel_l2

pytestmark = pytest.mark.gpu
COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


# the fused kernel's planes are staged by direct-to-LDS loads one or two planes
# ahead of the plane being computed (FFIP_FDEPTH)
KERNELS = {"depth1": ("1", "1"), "depth2": ("1", "2")}


def _run(cfg, ds, steps, fused, monkeypatch, lzf=None, probes=(), kernel="depth1"):
    """lzf = (planes per half-step tile, planes per fused segment), or None for
    the defaults."""
    kind, depth = KERNELS[kernel]
    monkeypatch.setenv("FFIP_FUSED", kind if fused else "0")
    monkeypatch.setenv("FFIP_FDEPTH", depth)
    if lzf:
        monkeypatch.setenv("FFIP_LZ", str(lzf[0]))
        monkeypatch.setenv("FFIP_LZF", str(lzf[1]))
    else:
        monkeypatch.delenv("FFIP_LZ", raising=False)
        monkeypatch.delenv("FFIP_LZF", raising=False)
    sim = LibSim(cfg, ds, probes=probes, trace_steps=steps + 1)
    assert sim.fused == fused
    if fused:
        assert sim.traffic(2)["points"] > 0  # some tiles run the fused kernel
    sim.step(steps)
    out = {c: sim.component(c) for c in COMPS}
    out["dft"] = [sim.dft(k) for k in range(sim.dft_count())]
    out["trace"] = sim.trace() if probes else None
    sim.close()
    return out


def _compare(a, b, tol, label):
    worst = 0.0
    scale = max(np.linalg.norm(a[c]) for c in COMPS)
    assert scale > 0
    for c in COMPS:
        e = float(np.linalg.norm(a[c] - b[c]) / scale)
        worst = max(worst, e)
        assert e <= tol, (label, c, e)
    for k, (x, y) in enumerate(zip(a["dft"], b["dft"])):
        for part in ("real", "imag"):
            e = rel_l2(y[part], x[part])
            worst = max(worst, e)
            assert e <= tol, (label, "dft", k, part, e)
    if a["trace"] is not None:
        e = rel_l2(b["trace"], a["trace"])
        worst = max(worst, e)
        assert e <= tol, (label, "trace", e)
    print("%s: fused vs half-step kernels, worst rel L2 %.3e" % (label, worst))


def _widen(cfg, size):
    """The same configuration on a grid wide enough in x for 32-point tiles clear
    of the CPML (the fused kernel's interior)."""
    cfg = dict(cfg)
    cfg["size"] = [float(v) for v in size]
    return cfg


def _cases():
    d = W.design_box(grid=(72, 30, 26), pml_cells=5, design=(30, 12, 6), patch=(6, 6), patch_z=7.0,
                     source_z_offset=7)
    return {
        "dipole": (_widen(W.vacuum_dipole(n=72, steps=60, pml_cells=5, plane_z=4), (72, 26, 30)), None),
        "spheres": (_widen(W.dispersive_spheres(n=72, nspheres=6, steps=60, pml_cells=5, nfreq=3,
                                                precision="fp64"), (72, 28, 28)), None),
        "voxels": (_widen(W.voxel_mix(n=72, voxel=4, pml_cells=5, nfreq=3, nplanes=3, precision="fp64",
                                      source_offset=8), (72, 24, 28)), None),
        "design": d,
        # 64 cells along x without CPML there: the fused box would reach the
        # packed last line of the rows and stops one tile short of it
        "tail64": _tail64(),
    }


def _tail64():
    cfg = _widen(W.vacuum_dipole(n=64, steps=60, pml_cells=5, plane_z=4), (64, 26, 30))
    cfg["PML"] = [dict(W.pml(5), direction="y"), dict(W.pml(5), direction="z")]
    return cfg, None


FUSED_CASES = _cases()


@pytest.mark.parametrize("kernel", sorted(KERNELS))
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name", sorted(FUSED_CASES))
def test_fused_equals_half_step_kernels(name, prec, kernel, monkeypatch):
    cfg, ds = FUSED_CASES[name]
    cfg = dict(cfg)
    cfg["precision"] = prec
    probes = [((1.0, 0.5, 0.0), "Ez"), ((-2.5, 1.0, 3.0), "Hy")]
    print("case %s %s %s" % (name, prec, kernel), flush=True)
    a = _run(cfg, ds, 70, False, monkeypatch, probes=probes)
    for lzf in (None, (3, 6), (2, 2)):
        b = _run(cfg, ds, 70, True, monkeypatch, lzf=lzf, probes=probes, kernel=kernel)
        _compare(a, b, 1e-13 if prec == "fp64" else 1e-6, "%s %s %s lz=%s" % (name, prec, kernel, lzf))


@pytest.mark.parametrize("name", ["spheres", "voxels", "design", "tail64"])
@pytest.mark.parametrize("kernel", sorted(KERNELS))
def test_fused_fp64_matches_oracle(name, kernel, monkeypatch):
    cfg, ds = FUSED_CASES[name]
    kind, depth = KERNELS[kernel]
    monkeypatch.setenv("FFIP_FUSED", kind)
    monkeypatch.setenv("FFIP_FDEPTH", depth)
    monkeypatch.setenv("FFIP_LZ", "3")
    monkeypatch.setenv("FFIP_LZF", "6")
    sim = LibSim(cfg, ds)
    assert sim.fused and sim.traffic(2)["points"] > 0
    sim.step(50)
    ora = OracleSim(cfg, ds, strict_pool=False)
    ora.run_steps(50)
    refs = {c: ora.field(c) for c in COMPS}
    scale = max(np.linalg.norm(r) for r in refs.values())
    for c in COMPS:
        assert float(np.linalg.norm(sim.component(c) - refs[c]) / scale) <= 1e-10, c
    for k in range(len(cfg.get("fields output", []))):
        a, b = sim.dft(k), ora.dft_output(k)
        for part in ("real", "imag"):
            assert rel_l2(a[part], b[part]) <= 1e-10, (k, part)


def test_fused_scope():
    """Periodic faces and magnetic media keep the two half-step kernels."""
    cfg = W.periodic_plane_wave(nz=40, nxy=4, steps=10)
    sim = LibSim(cfg)
    assert not sim.fused
    sim.close()
