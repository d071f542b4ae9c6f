"""Host set-up of libffip_hip.so (ffip_sim_plan, no GPU) against the oracle's
restatement of Simulation::init: index mapping, PML classification and grading,
source splitting, DFT closures, materials and probe weights must agree exactly
(integer work bit-exact, set-up doubles bit-identical)."""
import math

import numpy as np
import pytest

from ffip_amd import workloads as W
from ffip_amd._lib import geometry_info, plan
from oracle import oracle as O

EH = O.EH_ORDER


def _coords(sim, idx):
    g = sim.grid
    idx = np.asarray(idx, dtype=np.int64)
    x = idx % g.dim[0] + g.ghost_p1[0]
    y = (idx // g.dim[0]) % g.dim[1] + g.ghost_p1[1]
    z = idx // (g.dim[0] * g.dim[1]) + g.ghost_p1[2]
    return np.stack([x, y, z], axis=1)


def _comp_of(X):
    ct = (X[:, 0] & 1) | ((X[:, 1] & 1) << 1) | ((X[:, 2] & 1) << 2)
    lut = {O.CTYPE[n] & 7: i for i, n in enumerate(EH)}
    return np.array([lut[int(c)] for c in ct])


def _ix(P, X):
    M = np.array(P["M"])
    return (X + M[None, :]) >> 1


def _cases():
    c = {}
    c["vacuum_dipole"] = (W.vacuum_dipole(n=24, pml_cells=5, plane_z=4), None,
                          [((0, 0, 3), "Ez"), ((4.3, -1.2, 0.7), "Hx"), ((11.9, 11.9, 11.9), "Ey")])
    c["spheres"] = (W.dispersive_spheres(n=48, nspheres=6, pml_cells=6, nfreq=3, precision="fp64"), None, [])
    c["periodic"] = (W.periodic_plane_wave(nz=40, nxy=4), None, [])
    cfg = W.base((15, 13, 11), 0)  # odd cell counts on every axis
    cfg["PML"] = [dict(W.pml(3), direction="x", side="negative"), dict(W.pml(2, order=2, k_max=3.0), direction="z")]
    cfg["geometry"] = [{"type": "box", "center": [1.0, -0.5, 0.25], "size": [6.0, 4.0, 5.0],
                        "medium": W.medium(3.0, e_sus=[W.lorentz(0.1, 0.02, 0.7)], e_cond=0.05)},
                       {"type": "sphere", "center": [-3.1, 2.2, 0.4], "radius": 3.3, "medium": W.medium(2.0)}]
    cfg["sources"] = [{"type": "volume source", "center": [0.3, -1.1, 0.6], "size": [4.0, 0.0, 3.0],
                       "amplitude": 1.7, "field component": "Hy", "function": W.gaussian1(0.08, 5.0, 1.5)},
                      {"type": "volume source", "center": [-6.9, 0.0, 0.0], "size": [0.0, 0.0, 0.0],
                       "amplitude": 1.0, "field component": "Ex", "function": W.gaussian1(0.1)}]
    cfg["fields output"] = [W.dft((0, 0, 0), (20, 20, 3.3), [0.1, 0.05], "Ey", "a"),
                            W.dft((0, 0, 0), (20, 20, 3.3), [0.07, 0.1], "Ey", "b"),
                            W.dft((2.5, -1, 0.5), (1, 0, 0), [0.1], "Hz", "c"),
                            W.dft((40, 0, 0), (1, 1, 1), [0.1], "Ex", "outside")]
    c["odd_grid"] = (cfg, None, [((0.2, 0.3, -0.4), "Ez"), ((-7.5, 6.5, 5.5), "Hz")])
    rng = np.random.default_rng(7)
    cfg, ds = W.design_box(grid=(24, 22, 20), pml_cells=4, design=(8, 6, 4), patch=(6, 6), patch_z=6.0)
    dim = [9, 7, 5]
    amp = rng.uniform(0.1, 1.0, size=dim[0] * dim[1])
    ph = rng.uniform(-np.pi, np.pi, size=dim[0] * dim[1])
    cfg["sources"].append({"type": "inhomogeneous source", "center": [0.25, 0.0, 6.0], "size": [8.0, 6.0, 0.0],
                           "dimension": [dim[0], dim[1], 1], "amplitude dataset": "amp", "phase dataset": "ph",
                           "field component": "Ez", "frequency": 0.1, "function": W.gaussian1(0.1)})
    ds = dict(ds, amp=amp, ph=ph)
    c["design"] = (cfg, ds, [])
    cfg = W.base((10, 10, 8), 2)
    cfg["symmetry"] = [{"direction": "x", "phase factor": -1}]
    cfg["periodic"] = [0, 1, 0]
    cfg["sources"] = [{"type": "volume source", "center": [2.0, 0.0, 0.0], "size": [0.0, 0.0, 0.0],
                       "amplitude": 1.0, "field component": "Ey", "function": W.gaussian1(0.1)}]
    c["symmetry"] = (cfg, None, [])
    # the large-grid recipe at a size the oracle runs: ~200 overlapping voxel
    # boxes, 3-pole Lorentz and Drude palettes, z-normal DFT planes
    c["voxels"] = (W.voxel_mix(n=32, voxel=4, pml_cells=5, nfreq=2, nplanes=2, precision="fp64", source_offset=7),
                   None, [((1.5, -2.0, 3.0), "Ez")])
    # General_Medium_Box through the front end: two Lorentz poles whose amplitudes
    # follow a 1000-entry table over the density
    from ffip_amd.geom import General_Medium_Box, LorentzianSusceptibility, Vector3
    from ffip_amd.simulation import _Arrays
    gmb = General_Medium_Box(size=Vector3(8, 6, 4), center=Vector3(0.5, 0, 0), dim=Vector3(9, 7, 5), frequency=0.1,
                             density=np.random.default_rng(11).uniform(0, 1, (5, 7, 9)),
                             epsilon_fun=lambda r: 1 + 2.5 * r + 0.2j * r * (1 - r),
                             e_sus=[LorentzianSusceptibility(frequency=0.08, gamma=0.01),
                                    LorentzianSusceptibility(frequency=0.2, gamma=0.02)])
    cfg = W.base((18, 14, 12), 3)
    cfg["geometry"] = [gmb.get_json()]
    cfg["sources"] = [{"type": "volume source", "center": [-6.0, 0.0, 0.0], "size": [0.0, 10.0, 8.0],
                       "amplitude": 1.0, "field component": "Ez", "function": W.gaussian1(0.1)}]
    cfg["fields output"] = [W.dft((6, 0, 0), (0, 10, 8), [0.08, 0.1], "Ez", "a")]
    store = _Arrays()
    gmb.dump(store)
    c["general_box"] = (cfg, store.data, [((0.5, 0.5, 0.5), "Ez")])
    return c


CASES = _cases()


@pytest.fixture(scope="module", params=sorted(CASES))
def pair(request):
    cfg, ds, probes = CASES[request.param]
    p = plan(cfg, ds, probes)
    o = O.OracleSim(cfg, ds, probes=probes, trace_steps=1, strict_pool=False)
    return request.param, cfg, p, o


def test_layout_matches_interior(pair):
    name, cfg, p, o = pair
    assert p["P1"] == list(o.sim_p1) and p["P2"] == list(o.sim_p2)
    for ci, n in enumerate(EH):
        q1, q2 = O.component_interior(o.grid.grid_p1, o.grid.grid_p2, O.CTYPE[n])
        lo = _ix(p, np.array([q1]))[0]
        hi = _ix(p, np.array([q2]))[0]
        for a in range(3):
            if q1[a] <= q2[a]:
                assert p["lo"][ci][a] == lo[a] and p["hi"][ci][a] == hi[a], (n, a)
            else:
                assert p["hi"][ci][a] < p["lo"][ci][a]


def test_dipoles_bit_identical(pair):
    name, cfg, p, o = pair
    want = []
    for half in (0, 1):
        for idx, amp, st, en, wd in o.dip[half]:
            X = _coords(o, [idx])
            c = int(_comp_of(X)[0])
            i, j, k = (int(v) for v in _ix(p, X)[0])
            want.append([c, i, j, k, amp, st, en, wd])
    got = p["dipoles"]
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g[:4] == w[:4]
        assert g[4:] == w[4:]  # doubles parsed from %.17g are exact


def test_pml_classification_and_grading(pair):
    name, cfg, p, o = pair
    for half in (0, 1):
        par = half  # electric points sit on even derivative-axis parity, magnetic on odd
        tables = []
        for a in range(3):
            tab = {row[0]: row[1:] for row in p["slab"][a][par]}
            tables.append(tab)
        for x3 in range(3):
            L = o.pml_lists[half][x3]
            x1, x2 = (x3 + 1) % 3, (x3 + 2) % 3
            X = _coords(o, L["idx"])
            I = _ix(p, X)
            for q in range(X.shape[0]):
                i1, i2 = int(I[q, x1]), int(I[q, x2])
                in1, in2 = i1 in tables[x1], i2 in tables[x2]
                assert in1 or in2
                if in1:
                    assert tables[x1][i1] == [L["b1"][q], L["c1"][q], L["ik1"][q]]
                else:
                    assert L["c1"][q] == 0
                if in2:
                    assert tables[x2][i2] == [L["b2"][q], L["c2"][q], L["ik2"][q]]
                else:
                    assert L["c2"][q] == 0
            # plain points are in no slab of their derivative axes
            Xc = _coords(o, o.curl_lists[half][x3]["idx"])
            Ic = _ix(p, Xc)
            for q in range(0, Xc.shape[0], max(1, Xc.shape[0] // 500)):
                assert int(Ic[q, x1]) not in tables[x1] and int(Ic[q, x2]) not in tables[x2]


def test_dft_units(pair):
    name, cfg, p, o = pair
    mine = p["units"]
    theirs = o.units[0] + o.units[1]
    assert len(mine) == len(theirs)
    for m, t in zip(mine, theirs):
        assert m["nf"] == len(t["freqs"])
        if t["l1"] is None:
            assert m["has_local"] == 0
            continue
        lo = _ix(p, np.array([t["l1"]]))[0]
        hi = _ix(p, np.array([t["l2"]]))[0]
        assert m["box"] == [int(lo[0]), int(hi[0]), int(lo[1]), int(hi[1]), int(lo[2]), int(hi[2])]


def test_materials(pair):
    name, cfg, p, o = pair
    for half in (0, 1):
        M = p["materials"][half]
        table = M["table"]
        for c in range(3):
            comp = half * 3 + c
            want = {}
            for grp in o.groups[half].values():
                X = _coords(o, grp["idx"])
                sel = _comp_of(X) == comp
                for g in np.unique(grp["ginf"][sel]):
                    key = (float(g), len(grp["poles"]))
                    want[key] = want.get(key, 0) + int(np.sum(grp["ginf"][sel] == g))
            got = {}
            if M["uniform"]:
                got[(table[0][0], table[0][1])] = sum(want.values())
            else:
                for e, cnt in enumerate(M["hist"][c][:255]):
                    if cnt:
                        key = (table[e][0], table[e][1])
                        got[key] = got.get(key, 0) + cnt
                dense = M["hist"][c][255]
                if dense:
                    # two-medium points: compare their count with the oracle's non-default points
                    rest = {k: v for k, v in want.items() if k not in got}
                    assert sum(rest.values()) >= dense - sum(v for k, v in got.items() if k in want and v != want[k])
                    continue
            assert got == want, (name, half, c)


def test_probe_weights(pair):
    name, cfg, p, o = pair
    for pi, (pos, comp) in enumerate(CASES[name][2]):
        terms = o.grid.interp_terms([v / o.dx2 for v in pos], O.CTYPE[comp])
        mine = p["probe_terms"][pi]
        assert len(mine) == len(terms)
        for (ci, lin, w), (idx, wo) in zip(mine, terms):
            X = _coords(o, [idx])
            assert ci == int(_comp_of(X)[0])
            assert w == wo


def _geom_info_case(resolution):
    """Overlapping box, sphere and two-medium box; requests for every geometry
    on E and H components."""
    cfg = W.base((12, 11, 10), 2)
    cfg["resolution"] = resolution
    cfg["geometry"] = [{"type": "box", "center": [1.0, -0.5, 0.25], "size": [5.0, 4.0, 3.5],
                        "medium": W.medium(3.0)},
                       {"type": "sphere", "center": [-1.3, 1.2, 0.4], "radius": 2.6, "medium": W.medium(2.0)},
                       {"type": "two medium box", "center": [0.0, 0.0, -1.0], "size": [6.0, 5.0, 3.0],
                        "dimension": [4, 3, 2], "density dataset": "rho", "medium1": W.medium(4.0),
                        "medium2": W.medium(1.0)}]
    reqs = []
    for g in range(3):
        for comp in ("Ex", "Ey", "Ez", "Hx", "Hz"):
            reqs.append({"field component": comp, "geometry index": g,
                         "output group": "geometry info request %d %s" % (g, comp)})
    cfg["geometry info request"] = reqs
    ds = {"rho": np.random.default_rng(4).random(24)}
    return cfg, ds


@pytest.mark.parametrize("resolution", [1, 2])
def test_geometry_info_matches_oracle(resolution):
    cfg, ds = _geom_info_case(resolution)
    got = geometry_info(cfg, ds)
    o = O.OracleSim(cfg, ds, trace_steps=1, strict_pool=False)
    total = 0
    for r, (x, y, z) in zip(cfg["geometry info request"], got):
        wx, wy, wz = o.geometry_info(r["field component"], r["geometry index"])
        assert np.array_equal(x, wx) and np.array_equal(y, wy) and np.array_equal(z, wz), r
        total += x.size
    assert total > 0
    # points owned by an earlier geometry are not reported for a later one
    counts = [g[0].size for g in got]
    assert counts[0] > 0 and counts[5] > 0 and counts[10] > 0


def test_material_box_holds_every_geometry_point(pair):
    """The half-step kernels read material words only inside the material box
    and take the default medium elsewhere: every point a geometry owns (oracle
    restatement of the first-containing-geometry rule) must lie inside it."""
    name, cfg, p, o = pair
    for half in (0, 1):
        M = p["materials"][half]
        if M["uniform"]:
            continue
        box = np.array(M["mbox"])
        for n in EH[3 * half:3 * half + 3]:
            if n not in o.geom_index:
                continue
            p1, dims, gidx = o.geom_index[n]
            k, j, i = np.nonzero(gidx >= 0)
            X = np.stack([p1[0] + 2 * i, p1[1] + 2 * j, p1[2] + 2 * k], axis=1)
            q = _ix(p, X)
            for a in range(3):
                assert np.all(q[:, a] >= box[a, 0]) and np.all(q[:, a] <= box[a, 1]), (name, n, a)
