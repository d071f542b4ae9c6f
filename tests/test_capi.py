"""C-ABI of libffip_hip.so without a GPU: every symbol include/ffip_hip.h declares
is exported, and invalid input is rejected with an error code and message (the
reference aborts the MPI rank instead, src/simulation.cpp / app/run_sim_json.cpp)."""
import ctypes
import json
import os
import re

import pytest

from ffip_amd import workloads as W
from ffip_amd._lib import LIB_PATH, FfipError, LibSim, lib, plan

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "ffip_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ffip_\w+)\s*\(", text)))


def _gpu_present():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH)
    names = _declared()
    assert "ffip_sim_create" in names and "ffip_sim_get_dft" in names and len(names) >= 20
    L = ctypes.CDLL(LIB_PATH)
    for n in names:
        assert hasattr(L, n), n


def _create(cfg):
    h = ctypes.c_void_p()
    rc = lib().ffip_sim_create(json.dumps(cfg).encode() if not isinstance(cfg, str) else cfg.encode(),
                               ctypes.byref(h))
    return rc, h


def test_bad_json_is_einval():
    rc, h = _create("{not json")
    assert rc == -1 and b"JSON" in lib().ffip_last_error()


@pytest.mark.parametrize("mutate,msg", [
    (lambda c: c["sources"][0]["function"].update(type="Gaussian2"), "Gaussian1"),
    (lambda c: c.update(geometry=[{"type": "cylinder", "center": [0, 0, 0], "medium": W.medium()}]), "cylinder"),
    (lambda c: c["sources"][0].update({"field component": "Qx"}), "unknown field component"),
    (lambda c: c.update(precision="fp16"), "precision"),
    (lambda c: c.pop("resolution"), "resolution"),
    (lambda c: c["default material"].pop("mu"), "mu"),
    (lambda c: c.update(geometry=[{"type": "two medium box", "center": [0, 0, 0], "size": [2, 2, 2],
                                   "dimension": [2, 2, 2], "density dataset": "missing",
                                   "medium1": W.medium(2), "medium2": W.medium()}]), "missing"),
    (lambda c: c["fields output"][0].update(type="flux"), "fields_output"),
])
def test_invalid_configuration_is_rejected(mutate, msg):
    cfg = W.vacuum_dipole(n=8, steps=1, pml_cells=2, plane_z=1)
    mutate(cfg)
    with pytest.raises(FfipError) as e:
        plan(cfg)
    assert "error -1" in str(e.value) and msg in str(e.value)


def test_state_errors():
    rc, h = _create(W.vacuum_dipole(n=8, steps=1, pml_cells=2, plane_z=1))
    assert rc == 0
    try:
        out = ctypes.c_int64()
        assert lib().ffip_sim_step(h, 1) == -2 and b"finalize" in lib().ffip_last_error()
        assert lib().ffip_sim_time_step(h, ctypes.byref(out)) == -2
        assert lib().ffip_sim_step(h, -3) == -1
    finally:
        lib().ffip_sim_destroy(h)


@pytest.mark.skipif(_gpu_present(), reason="checks the no-device error path")
def test_finalize_without_device_is_edevice():
    with pytest.raises(FfipError) as e:
        LibSim(W.vacuum_dipole(n=8, steps=1, pml_cells=2, plane_z=1))
    assert "error -3" in str(e.value)


def test_plan_is_host_only_and_consistent():
    p = plan(W.vacuum_dipole(n=16, pml_cells=4, plane_z=3))
    assert p["n"] == [18, 18, 18] and p["M"] == [17, 17, 17]
    assert len(p["dipoles"]) == 2 and p["dipoles"][0][4] == -0.5
    # an 8-cell-thick PML on both sides covers the same slab positions on each axis
    assert [len(p["slab"][a][0]) for a in range(3)] == [len(p["slab"][0][0])] * 3
