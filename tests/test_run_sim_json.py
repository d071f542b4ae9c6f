"""run_sim_json, the file drop-in (app/run_sim_json.cpp:6-25; files as written
and read by ffip/simulation.py:404-431).

The HDF5 files are written and read by h5py in the image's second interpreter
(/opt/conda/bin/python3.9), an HDF5 implementation independent of the
executable's own reader and writer; the tests skip when it is absent.

CPU: the executable reads the input file's datasets exactly (its host-only
set-up plan equals the in-process plan built from the same arrays), creates a
missing input file as HighFive's ReadOnly | Create does, and reports errors
with a non-zero status.  GPU: a full run's output file holds the groups the
reference writes, with the same shapes and axes, and its values equal the
in-process run's bit for bit."""
import json
import os
import subprocess

import numpy as np
import pytest

import ffip_amd as ffip
from ffip_amd import _lib
from ffip_amd.simulation import _Arrays

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "ffip_amd", "run_sim_json")
H5PY = "/opt/conda/bin/python3.9"


def _have_h5py():
    if not os.path.exists(H5PY):
        return False
    r = subprocess.run([H5PY, "-c", "import h5py"], capture_output=True)
    return r.returncode == 0


needs_h5 = pytest.mark.skipif(not _have_h5py(), reason="no h5py interpreter in this image")
needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="run_sim_json not built")


def _h5(mode, src, dst):
    subprocess.run([H5PY, os.path.join(HERE, "h5tool.py"), mode, src, dst], check=True, capture_output=True)


def _sim(precision="fp64"):
    """24^3 grid: a two-medium box (density dataset), an inhomogeneous source
    (amplitude and phase datasets), a DFT plane and a geometry-info request."""
    rng = np.random.default_rng(7)
    box = ffip.Two_Medium_Box(size=ffip.Vector3(8, 8, 4), center=ffip.Vector3(0, 0, 1), dim=ffip.Vector3(5, 5, 3),
                              density=rng.random((3, 5, 5)), medium1=ffip.Medium(epsilon=4), medium2=ffip.Medium())
    amp = rng.normal(size=(1, 5, 5)) + 1j * rng.normal(size=(1, 5, 5))
    src = ffip.Inhom_Source(ffip.Gaussian1(0.1), amplitude=amp, frequency=0.1, center=ffip.Vector3(0, 0, -6),
                            size=ffip.Vector3(6, 6, 0), dim=ffip.Vector3(5, 5, 1), field_component="Ex")
    sim = ffip.Simulation(size=ffip.Vector3(24, 24, 24), resolution=1, pmls=[ffip.PML(4)], geometry=[box],
                          sources=[src], precision=precision)
    sim.add_dft_fields(center=ffip.Vector3(z=5), size=ffip.Vector3(10, 10, 0), frequency=[0.1, 0.08],
                       field_component="Ex")
    sim.request_geom_info(sim.geometry[0], "Ey")
    return sim


def _stage(tmp_path, sim, stop):
    """What ffip's Simulation.run leaves before it starts mpirun: config.json and
    the input file, both in the working directory."""
    cfg = sim.get_json(stop_condition=stop)
    store = _Arrays()
    sim.dump(store)
    with open(tmp_path / "config.json", "w") as f:
        json.dump(cfg, f, indent=4)
    np.savez(tmp_path / "in.npz", **store.data)
    _h5("write", str(tmp_path / "in.npz"), str(tmp_path / cfg["input file"]))
    return cfg, store.data


def _run(tmp_path, *args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([EXE, *args], cwd=tmp_path, capture_output=True, text=True, env=e, timeout=300)


@needs_exe
def test_usage_and_missing_config(tmp_path):
    r = _run(tmp_path, "--help")
    assert r.returncode == 0 and "config.json" in r.stdout
    r = _run(tmp_path)
    assert r.returncode == 1 and "config.json" in r.stderr


@needs_exe
def test_other_mpi_ranks_exit_at_once(tmp_path):
    # no config.json: rank 0 would fail, rank 1 must not even look
    r = _run(tmp_path, env={"OMPI_COMM_WORLD_RANK": "1"})
    assert r.returncode == 0 and "rank 1" in r.stdout


@needs_exe
@needs_h5
def test_plan_reads_input_datasets_exactly(tmp_path):
    sim = _sim()
    cfg, data = _stage(tmp_path, sim, ffip.run_until_time(30))
    r = _run(tmp_path, "--plan")
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    want = _lib.plan(cfg, data)
    assert got == want
    # the plan lists each dipole's amplitude, so it depends on the file's values
    data2 = dict(data)
    name = cfg["sources"][0]["amplitude dataset"]
    data2[name] = data[name] * 0.5
    assert _lib.plan(cfg, data2) != got


@needs_exe
@needs_h5
def test_missing_input_file_is_created_and_datasets_reported(tmp_path):
    sim = _sim()
    cfg = sim.get_json(stop_condition=ffip.run_until_time(30))
    with open(tmp_path / "config.json", "w") as f:
        json.dump(cfg, f)
    r = _run(tmp_path, "--plan")
    assert r.returncode == 1
    assert "was not provided" in r.stderr
    assert (tmp_path / cfg["input file"]).exists()  # ReadOnly | Create, src/simulation.cpp:151-153
    _h5("read", str(tmp_path / cfg["input file"]), str(tmp_path / "empty.npz"))
    assert len(np.load(tmp_path / "empty.npz").files) == 0


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
@needs_exe
@needs_h5
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_output_file_equals_in_process_run(tmp_path, precision):
    sim = _sim(precision)
    stop = ffip.run_until_time(60)
    cfg, data = _stage(tmp_path, sim, stop)
    print("run_sim_json %s" % precision, flush=True)
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Last time step = 120" in r.stdout
    _h5("read", str(tmp_path / cfg["fields output file"]), str(tmp_path / "out.npz"))
    out = np.load(tmp_path / "out.npz")
    # in-process run of the same document and arrays
    ref = _lib.LibSim(cfg, data)
    ref.run(cfg["stop condition"])
    d = ref.dft(0)
    g = "volume fields dft 0"
    assert tuple(out["__shape__/%s/real" % g]) == d["real"].shape
    assert str(out["__dtype__/%s/real" % g]) == "float64"
    for key in ("real", "imag", "x", "y", "z", "f"):
        assert np.array_equal(out["%s/%s" % (g, key)], d[key]), key
    assert list(out["%s/f" % g]) == [0.08, 0.1]
    gi = "geometry info request 0 Ey"
    x, y, z = ref.geom_info(0)
    assert x.size > 0
    for key, v in (("x", x), ("y", y), ("z", z)):
        assert np.array_equal(out["%s/%s" % (gi, key)], v)
    ref.close()
    # the front end reads the file's groups the way it reads the h5py groups
    sim2 = _sim(precision)
    mon = sim2.dft_fields[0]
    groups = {g: {k: out["%s/%s" % (g, k)] for k in ("real", "imag", "x", "y", "z", "f")}}
    mon.read(groups)
    # (read widens singleton axes by repeating them)
    assert np.array_equal(mon.values.real[tuple(slice(0, n) for n in d["real"].shape)], d["real"])


@pytest.mark.gpu
@needs_exe
@needs_h5
def test_dft_stop_condition_through_file(tmp_path):
    sim = _sim()
    stop = ffip.run_until_dft(center=ffip.Vector3(z=5), size=ffip.Vector3(10, 10, 0), field_component="Ex",
                              time_interval_examined=20, var=1e-2, frequency=0.1)
    cfg, data = _stage(tmp_path, sim, stop)
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = _lib.LibSim(cfg, data)
    n = ref.run(cfg["stop condition"])
    assert "Last time step = %d" % n in r.stdout
    _h5("read", str(tmp_path / cfg["fields output file"]), str(tmp_path / "out.npz"))
    out = np.load(tmp_path / "out.npz")
    assert np.array_equal(out["volume fields dft 0/real"], ref.dft(0)["real"])
    ref.close()
