"""HIP engine (through the C ABI) against the oracle and the golden fixture.

Bars: FP64 storage agrees with the FP64 oracle within 1e-10 relative L2 on every
DFT monitor and probe trace (the FP64 parity bar); axes and frequencies are
bit-identical.  FP32 storage (FP64 arithmetic, coefficients and DFT sums) is held
to the per-case tolerances written below and its error is printed."""
import math
import os

import numpy as np
import pytest

from ffip_amd import workloads as W
from ffip_amd._lib import LibSim
from oracle.oracle import OracleSim

from .conftest import rel_l2
from .test_setup_parity import CASES

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vacuum_dipole_64.npz")
TOL64 = 1e-10


def _run_gpu(cfg, ds=None, probes=(), steps=None):
    cap = steps if steps is not None else int(math.ceil(cfg["stop condition"]["time"] / 0.5)) + 1
    sim = LibSim(cfg, ds, probes=probes, trace_steps=cap)
    n = sim.run(cfg["stop condition"]) if steps is None else (sim.step(steps) or steps)
    return sim, n


def _compare(cfg, ds=None, probes=(), steps=None, tol=TOL64, label=""):
    gpu, n = _run_gpu(cfg, ds, probes, steps)
    ora = OracleSim(cfg, ds, probes=probes, trace_steps=n, strict_pool=False)
    if steps is None:
        ora.run(cfg["stop condition"])
    else:
        ora.run_steps(steps)
    assert gpu.time_step == ora.time_step == n
    worst = 0.0
    for k in range(len(cfg.get("fields output", []))):
        a, b = gpu.dft(k), ora.dft_output(k)
        assert a["real"].shape == b["real"].shape
        for ax in "xyzf":
            assert np.array_equal(a[ax], b[ax]), (k, ax)
        for part in ("real", "imag"):
            e = rel_l2(a[part], b[part])
            worst = max(worst, e)
            assert e <= tol, (label, k, part, e)
    if probes:
        e = rel_l2(gpu.trace(), ora.trace[:n])
        worst = max(worst, e)
        assert e <= tol, (label, "trace", e)
    # raw fields, relative to the largest component norm: a component that is
    # zero by symmetry holds only rounding noise, which no relative test can hold
    refs = {c: ora.field(c) for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
    scale = max(np.linalg.norm(r) for r in refs.values())
    for comp, ref in refs.items():
        if scale > 0:
            e = float(np.linalg.norm(gpu.component(comp) - ref) / scale)
            worst = max(worst, e)
            assert e <= tol, (label, comp, e)
    print("%s: %d steps, worst rel L2 %.3e" % (label, n, worst))
    return worst


def test_vacuum_dipole_64_matches_golden_fp64():
    g = np.load(GOLDEN)
    cfg = W.vacuum_dipole()
    sim, n = _run_gpu(cfg, probes=W.VACUUM_DIPOLE_PROBES, steps=None)
    assert n == int(g["steps"]) == 1000
    d = sim.dft(0)
    for ax in "xyzf":
        assert np.array_equal(d[ax], g[ax])
    assert rel_l2(d["real"], g["real"]) <= TOL64
    assert rel_l2(d["imag"], g["imag"]) <= TOL64
    assert rel_l2(sim.trace(), g["trace"]) <= TOL64


def test_vacuum_dipole_64_fp32_error():
    g = np.load(GOLDEN)
    cfg = W.vacuum_dipole(precision="fp32")
    sim, n = _run_gpu(cfg, probes=W.VACUUM_DIPOLE_PROBES)
    d = sim.dft(0)
    e = max(rel_l2(d["real"], g["real"]), rel_l2(d["imag"], g["imag"]))
    t = rel_l2(sim.trace(), g["trace"])
    print("fp32 vs fp64 oracle: DFT %.2e, trace %.2e" % (e, t))
    assert e <= 1e-6 and t <= 1e-5


@pytest.mark.parametrize("name", sorted(CASES))
def test_setup_cases_fp64(name):
    cfg, ds, probes = CASES[name]
    _compare(cfg, ds, probes, steps=150, label=name)


@pytest.mark.parametrize("name", ["spheres", "design", "voxels", "odd_grid"])
def test_cpml_split_kernels_match_oracle(name, monkeypatch):
    """Tiles clear of every CPML slab run kernels compiled without the CPML
    paths (and, without per-point media, without that path too); large grids use
    them by default.  Forced here on small cases, against the oracle."""
    monkeypatch.setenv("FFIP_CPML_SPLIT", "1")
    cfg, ds, probes = CASES[name]
    _compare(cfg, ds, probes, steps=150, label=name + "-split")


def _widened(cfg, nx):
    """The same configuration on a grid nx cells wide."""
    cfg = dict(cfg)
    cfg["size"] = [float(nx)] + [float(v) for v in cfg["size"][1:]]
    return cfg


def _magnetic(nx=20):
    cfg = W.base((nx, 18, 16), 4)
    m = W.medium(2.5, e_sus=[W.lorentz(0.09, 0.02, 0.8)], e_cond=0.02)
    m["mu"] = 1.5
    m["magnetic conductivity"] = 0.03
    m["magnetic susceptibility"] = [W.lorentz(0.11, 0.03, 0.4)]
    cfg["geometry"] = [{"type": "box", "center": [1, 0, 0], "size": [8, 6, 6], "medium": m},
                       {"type": "sphere", "center": [-4, 2, 1], "radius": 2.5, "medium": W.medium(4.0)}]
    cfg["sources"] = [{"type": "volume source", "center": [-2.0, 0.5, 0.0], "size": [0, 0, 0], "amplitude": 1.0,
                       "field component": "Ey", "function": W.gaussian1(0.1)},
                      {"type": "volume source", "center": [3.0, 0.0, 1.5], "size": [0, 2, 0], "amplitude": 0.5,
                       "field component": "Hz", "function": W.gaussian1(0.08)}]
    return cfg


def _tail_cases():
    c = {"spheres64": (W.dispersive_spheres(n=64, nspheres=6, pml_cells=6, nfreq=3, precision="fp64"), None, [])}
    for name in ("design", "general_box", "odd_grid", "voxels"):
        cfg, ds, probes = CASES[name]
        c[name + "32"] = (_widened(cfg, 32) if name != "voxels" else cfg, ds, probes)
    cfg = _magnetic(32)
    f = [0.07, 0.1]
    # flux-density and field monitors across the whole x range, boundary faces included
    cfg["fields output"] = [W.dft((0, 0, 0), (32, 4, 0), f, comp, "t%d" % q)
                            for q, comp in enumerate(["Ey", "Dy", "Bz", "Hz", "Dx", "Hx"])]
    c["magnetic32"] = (cfg, None, [((15.5, 0.5, 0.0), "Ey"), ((16.0, 1.0, 0.5), "Hz")])
    # ghost faces on the packed last line: a mirror plane at the low x face with
    # periodic y (x from 0 to 64: 32 cells), and periodic x and y
    c["symmetry64"] = (_widened(CASES["symmetry"][0], 64), None, [((15.5, 1.0, 0.5), "Ey")])
    cfg = dict(CASES["periodic"][0])
    cfg["size"] = [32.0, 32.0] + [float(v) for v in cfg["size"][2:]]
    c["periodic32"] = (cfg, None, [])
    return c


TAIL_CASES = _tail_cases()


@pytest.mark.parametrize("name", sorted(TAIL_CASES))
def test_power_of_two_grids_match_oracle(name):
    """Grids of 2^k cells along x have 2^k + 1 lattice points there, so the last
    32-point tile column holds one point per row (the layout of the benchmark
    grids), and the last two x indices of 15 rows share a packed line: FP64
    against the oracle, with poles in both polarization layouts, per-point
    media, magnetic media, flux-density monitors reaching the boundary faces,
    mirror and periodic ghost faces."""
    cfg, ds, probes = TAIL_CASES[name]
    _compare(cfg, ds, probes, steps=150, label=name)


@pytest.mark.parametrize("name", sorted(TAIL_CASES))
def test_power_of_two_grids_fp32_storage(name):
    """The packed last line in FP32 storage (FP64 arithmetic) against the FP64
    oracle, within the FP32-storage tolerance of the small-grid cases."""
    cfg, ds, probes = TAIL_CASES[name]
    _compare(dict(cfg, precision="fp32"), ds, probes, steps=150, tol=1e-5, label=name + "-fp32")


@pytest.mark.parametrize("name", ["voxels32", "spheres64", "magnetic32"])
def test_twelve_plane_segments(name, monkeypatch):
    """The 12-plane z segments that large grids use (and the polarization plan
    ordered for them), forced on grids the oracle runs: FP64 against the oracle."""
    monkeypatch.setenv("FFIP_LZ", "12")
    cfg, ds, probes = TAIL_CASES[name]
    _compare(cfg, ds, probes, steps=150, label=name + "-lz12")


@pytest.mark.parametrize("split", ["0", "1"])
def test_dispersive_default_medium(split, monkeypatch):
    """A dispersive default medium with pole-free boxes in it: points outside the
    box of non-default media take the default's material word without reading
    it, and those points carry poles (compact or full polarization storage)."""
    monkeypatch.setenv("FFIP_CPML_SPLIT", split)
    cfg = W.base((18, 16, 14), 3)
    cfg["default material"] = W.medium(1.5, e_sus=[W.lorentz(0.08, 0.01, 0.6)])
    cfg["geometry"] = [{"type": "box", "center": [1.0, 0.5, 0.0], "size": [4.0, 3.0, 3.0], "medium": W.medium(3.0)},
                       {"type": "sphere", "center": [-3.0, -2.0, 1.0], "radius": 2.2,
                        "medium": W.medium(2.0, e_sus=[W.drude(0.15, 0.02, 0.5)])}]
    cfg["sources"] = [{"type": "volume source", "center": [0.0, 0.0, -3.0], "size": [6.0, 6.0, 0.0],
                       "amplitude": 1.0, "field component": "Ex", "function": W.gaussian1(0.1)}]
    cfg["fields output"] = [W.dft((0, 0, 2.5), (12, 10, 0), [0.07, 0.1], "Ex", "a")]
    cfg["stop condition"] = {"type": "time", "time": 100.0}
    _compare(cfg, steps=200, label="dispersive-default-split%s" % split)


def test_dispersive_spheres_fp64_and_fp32():
    cfg = W.dispersive_spheres(n=48, nspheres=6, steps=600, pml_cells=6, nfreq=4, precision="fp64")
    _compare(cfg, label="spheres48-fp64")
    cfg["precision"] = "fp32"
    worst = _compare(cfg, tol=1e-4, label="spheres48-fp32")
    assert worst > 0


def test_magnetic_media_and_flux_density_monitors():
    """mu != 1, magnetic conductivity and a magnetic Lorentz pole run the ADE on
    the H half step; D and B monitors sample eps_inf*E + sum P (B likewise)."""
    cfg = _magnetic()
    f = [0.07, 0.1]
    cfg["fields output"] = [W.dft((1, 0, 0), (6, 4, 0), f, c, "g%d" % q)
                            for q, c in enumerate(["Ey", "Dy", "Bz", "Hz", "Dx", "Hx"])]
    probes = [((1.0, 0.5, 0.0), "Ey"), ((1.5, 1.0, 0.5), "Hz")]
    _compare(cfg, probes=probes, steps=300, label="magnetic")
    cfg["precision"] = "fp32"
    _compare(cfg, probes=probes, steps=300, tol=1e-4, label="magnetic-fp32")


def test_many_frequency_monitors_match_oracle():
    """Monitors with 9 frequencies (three chunks of the general DFT kernel, the
    last one partial) on planes normal to x, y and z, each spanning several
    256-point blocks, against the oracle."""
    freqs = list(np.linspace(0.05, 0.15, 9))
    cfg = W.vacuum_dipole(n=32, steps=120, pml_cells=6, freqs=freqs, plane_z=5)
    cfg["fields output"] += [W.dft((3, 0, 0), (0, 32, 32), freqs, "Ez", "volume fields dft 1"),
                             W.dft((0, -2.5, 0), (32, 0, 32), freqs, "Ex", "volume fields dft 2")]
    _compare(cfg, label="9-frequency monitors")


@pytest.mark.parametrize("nf", [9, 1])
def test_dft_sample_batches_are_flushed(nf):
    """E/H monitors gather 16 steps of samples before adding them into the FP64
    sums.  Read after 75 steps (11 samples pending) and after 120 (8 pending),
    both storage types must equal the oracle: a lost or doubled batch would miss
    or repeat steps of a pulse that is still passing."""
    freqs = list(np.linspace(0.05, 0.15, nf)) if nf > 1 else [0.1]
    cfg = W.vacuum_dipole(n=32, steps=120, pml_cells=6, freqs=freqs, plane_z=5)
    cfg["fields output"] += [W.dft((3, 0, 0), (0, 32, 32), freqs, "Ez", "volume fields dft 1")]
    ora = OracleSim(cfg)
    ora.run({"type": "time", "time": 37.5})
    ref = [[ora.dft_output(k) for k in range(2)]]
    ora.run(cfg["stop condition"])
    ref.append([ora.dft_output(k) for k in range(2)])
    for prec, tol in (("fp64", TOL64), ("fp32", 1e-5)):
        cfg["precision"] = prec
        sim = LibSim(cfg)
        sim.run({"type": "time", "time": 37.5})
        got = [[sim.dft(k) for k in range(2)]]
        sim.run(cfg["stop condition"])
        got.append([sim.dft(k) for k in range(2)])
        sim.close()
        for stage in range(2):
            for k in range(2):
                a, b = got[stage][k], ref[stage][k]
                assert np.abs(b["real"]).max() > 0
                for part in ("real", "imag"):
                    e = rel_l2(a[part], b[part])
                    assert e <= tol, (prec, nf, stage, k, part, e)


def test_monitors_at_sources_volumes_and_past_the_grid():
    """E/H monitors read after 77 steps (13 samples pending in the ring): a plane
    through a current sheet and one through a dipole (the samples are the values
    after the point sources), an H plane, a volume and a plane reaching past the
    grid, against the oracle in both storage types."""
    freqs = [0.07, 0.1, 0.13]
    cfg = W.vacuum_dipole(n=32, steps=90, pml_cells=6, freqs=freqs, plane_z=5)
    cfg["sources"].append({"type": "volume source", "center": [0.0, 0.0, -3.0], "size": [10.0, 8.0, 0.0],
                           "amplitude": 0.7, "field component": "Ex", "function": W.gaussian1(0.08)})
    cfg["fields output"] += [W.dft((0, 0, -3), (12, 10, 0), freqs, "Ex", "sheet"),
                             W.dft((0, 0, 0), (10, 10, 0), freqs, "Ez", "dipole plane"),
                             W.dft((0, 0, 4), (16, 16, 0), freqs, "Hy", "h plane"),
                             W.dft((2, 1, 0), (6, 5, 4), freqs, "Ey", "volume"),
                             W.dft((0, 0, 2), (40, 40, 0), freqs, "Ez", "past the grid")]
    ora = OracleSim(cfg)
    ora.run_steps(77)
    for prec, tol in (("fp64", TOL64), ("fp32", 1e-5)):
        sim = LibSim(dict(cfg, precision=prec))
        sim.step(77)
        for k in range(sim.dft_count()):
            a, b = sim.dft(k), ora.dft_output(k)
            assert np.abs(b["real"]).max() > 0
            for part in ("real", "imag"):
                e = rel_l2(a[part], b[part])
                assert e <= tol, (prec, k, part, e)
        sim.close()


def test_continuation_equals_single_run():
    cfg = W.vacuum_dipole(n=20, steps=100, pml_cells=4, plane_z=3)
    a = LibSim(cfg)
    a.run({"type": "time", "time": 20.0})
    a.run({"type": "time", "time": 50.0})
    b = LibSim(cfg)
    b.run({"type": "time", "time": 50.0})
    assert a.time_step == b.time_step == 100
    assert np.array_equal(a.dft(0)["real"], b.dft(0)["real"])
    assert np.array_equal(a.component("Hx"), b.component("Hx"))


def test_fields_and_probe_match_oracle():
    cfg = CASES["odd_grid"][0]
    gpu, n = _run_gpu(cfg, steps=80)
    ora = OracleSim(cfg)
    ora.run_steps(80)
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        assert rel_l2(gpu.component(comp), ora.field(comp)) <= TOL64, comp
    pos = (0.7, -1.3, 0.4)
    want = sum(ora.eh[i] * w for i, w in ora.grid.interp_terms([v / ora.dx2 for v in pos], 4))
    assert gpu.probe(pos, "Ez") == pytest.approx(want, rel=1e-10, abs=1e-14)


def test_monitor_outside_grid_is_zero_with_closure_dims():
    cfg = CASES["odd_grid"][0]
    gpu, n = _run_gpu(cfg, steps=10)
    d = gpu.dft(3)
    # a 1-cell box at x = 40 (doubled 80): Ex closure is x in {79, 81}, y and z in {-2, 0, 2}
    assert d["real"].shape == (1, 3, 3, 2) and not d["real"].any()
    assert np.array_equal(d["x"], [39.5, 40.5]) and np.array_equal(d["y"], [-1.0, 0.0, 1.0])


def test_full_size_dispersive_fp32_against_fp64():
    """Full 256^3 dispersive workload, 4000 steps: FP32 storage against FP64
    storage for every monitor (the FP64 engine equals the oracle at 1e-10 on the
    smaller cases above)."""
    cfg = W.dispersive_spheres(n=256, precision="fp64")
    a = LibSim(cfg)
    a.run(cfg["stop condition"])
    cfg["precision"] = "fp32"
    b = LibSim(cfg)
    b.run(cfg["stop condition"])
    for k in range(2):
        da, db = a.dft(k), b.dft(k)
        e = max(rel_l2(db["real"], da["real"]), rel_l2(db["imag"], da["imag"]))
        print("monitor %d: fp32 vs fp64 rel L2 %.3e, |DFT| max %.3e" % (k, e, np.abs(da["real"]).max()))
        assert np.abs(da["real"]).max() > 0
        assert e <= 1e-4


def test_voxel_mix_64_frequency_planes_match_oracle():
    """The large-run recipe (three-pole Lorentz and Drude voxels, compact
    polarization, an Ez sheet normal to x) at a size the oracle runs, with its
    monitor set: 8 z-normal Ez planes at half-integer z x 64 frequencies in
    [0.05, 0.15].  FP64 against the oracle at the parity bar; the planes carry
    the pulse (their largest |DFT| is printed)."""
    cfg = W.voxel_mix(n=32, voxel=4, pml_cells=5, nfreq=64, nplanes=8, precision="fp64", source_offset=7)
    cfg["stop condition"] = {"type": "time", "time": 60.0}
    worst = _compare(cfg, label="voxel mix 32^3, 8 planes x 64 freqs")
    ora = OracleSim(cfg, strict_pool=False)
    ora.run_steps(120)
    peak = max(np.abs(ora.dft_output(k)["real"]).max() for k in range(8))
    print("largest |DFT| %.3e, worst rel L2 %.3e" % (peak, worst))
    assert peak > 0.1


def _fp32_vs_fp64(cfg, steps, nmon):
    out = {}
    for prec in ("fp64", "fp32"):
        cfg["precision"] = prec
        sim = LibSim(cfg)
        sim.step(steps)
        out[prec] = [sim.dft(k) for k in range(nmon)]
        sim.close()
    errs, peaks = [], []
    for k in range(nmon):
        a, b = out["fp64"][k], out["fp32"][k]
        peaks.append(float(np.abs(a["real"] + 1j * a["imag"]).max()))
        errs.append(max(rel_l2(b[part], a[part]) for part in ("real", "imag")))
        print("monitor %d: |DFT| max %.3e, fp32 vs fp64 rel L2 %.3e" % (k, peaks[-1], errs[-1]))
    return errs, peaks


def test_full_size_voxel_mix_512_fp32_against_fp64():
    """The 512^3 voxel mix at its real shape with two Ez DFT planes x 4
    frequencies, 300 steps: FP32 storage against FP64 storage on monitors that
    the pulse has reached (|DFT| of order one)."""
    cfg = W.voxel_mix(n=512, precision="fp64", nfreq=4, nplanes=2)
    errs, peaks = _fp32_vs_fp64(cfg, 300, 2)
    assert min(peaks) > 0.5
    assert max(errs) <= 1e-4


def test_full_size_voxel_mix_640_eight_planes_64_freqs_fp32_against_fp64():
    """The large single-GPU run at its real shape: 640^3 voxel mix, 16-cell CPML,
    8 Ez planes x 64 frequencies, 300 steps: FP32 storage (FP64 sums) against
    FP64 storage on every monitor, each carrying the pulse."""
    cfg = W.voxel_mix(n=640, precision="fp64", nfreq=64, nplanes=8)
    errs, peaks = _fp32_vs_fp64(cfg, 300, 8)
    assert min(peaks) > 0.5
    assert max(errs) <= 1e-4
