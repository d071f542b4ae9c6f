"""Pins for the CPU oracle (oracle/oracle.py + oracle/ffip_oracle.c).

The reference's own tests hold only the Yee3 indexing known answers
(test/utility_test.cpp:8-19); every other part of the oracle is pinned here by
mathematics that does not go through the restated formula:
  * a1 grid indexing          -> the reference's known answers
  * a4 source pulse + split   -> Gaussian1 closed form at the first E update
  * a7 DFT time stamps/sign   -> direct summation of a recorded time series
  * a2 curl / Yee staggering  -> discrete plane-wave dispersion relation
  * a3 CPML                   -> measured normal-incidence reflection
  * a5 ADE coefficients       -> discrete susceptibility closed form D/E = eps(z)
  * a6 density interpolation  -> textbook trilinear interpolation (scipy)
  * PEC/PMC ghost faces        -> a half grid with a mirror plane equals the full
                                  grid driven by mirror-symmetric sources
  * volume-source amplitude    -> discrete 1-D Green's function of a current sheet
                                  and the area scaling of a stretched sheet
  * stop-condition windows     -> replay of the window logic from a recorded series
Each of the last three was checked against copies of the oracle with a slip
planted (multiplier sign, dropped or doubled spacing factor, norm read before
the step, running maximum not folded): every slip fails its pin.
"""
import math

import numpy as np
import pytest

from ffip_amd import workloads as W
from oracle import oracle as O

PI = math.pi


def test_yee3_known_answers():
    # test/utility_test.cpp:8-19
    g = O.Yee3((-2, -2, -3), (3, 5, 2))
    assert g.size() == 288
    assert g.grid_p1 == (-1, -1, -2) and g.grid_p2 == (2, 4, 1)
    assert g.stride == (1, 6, 48)
    assert g.base_point((1.2, 3, 4), O.CTYPE["Ex"]) == (1, 2, 4)
    assert g.base_point((1.2, 4, 1.1), O.CTYPE["Hx"]) == (0, 3, 1)
    assert int(g.index(1, 1, 1)) == 213


def test_ctype_predicates():
    # test/utility_test.cpp:46-77
    assert O.count_bits(3) == 2 and O.count_bits(14) == 3 and O.count_bits(7) == 3
    assert O.dir_of_ctype(O.CTYPE["Ex"]) == 0 and O.dir_of_ctype(O.CTYPE["Hy"]) == 1
    assert O.dir_of_ctype(O.CTYPE["Dz"]) == 2
    assert O.is_e_point(O.CTYPE["Ex"]) and not O.is_e_point(O.CTYPE["Hy"])
    assert not O.is_e_point(O.CTYPE["Bx"]) and O.is_e_point(O.CTYPE["Dz"])
    assert O.is_eh_point(O.CTYPE["Hz"]) and not O.is_eh_point(O.CTYPE["Bx"])


def test_closure_and_interior():
    ez = O.CTYPE["Ez"]
    # a plane at z = 10 cells (doubled 20, even) is between Ez layers 19 and 21
    assert O.component_closure((-64, -64, 20), (64, 64, 20), ez) == ((-64, -64, 19), (64, 64, 21))
    assert O.component_interior((-64, -64, -64), (64, 64, 64), ez) == ((-64, -64, -63), (64, 64, 63))


def test_interp_partition_of_unity():
    g = O.Yee3((-6, -6, -6), (6, 6, 6))
    for pos, ct in (((0.3, -1.7, 2.2), "Ex"), ((1.0, 1.0, 1.0), "Hz"), ((-3.9, 2.5, 0.1), "Ey")):
        w = sum(t[1] for t in g.interp_terms(pos, O.CTYPE[ct]))
        assert abs(w - 1.0) < 1e-15
    # far outside: no terms (the reference returns T{})
    assert g.interp_terms((-20, -20, -20), O.CTYPE["Ex"]) == []


def _tiny_dipole(n=8, comp="Ez"):
    cfg = W.base((n, n, n), 0)
    cfg["sources"] = [{"type": "volume source", "center": [0.0, 0.0, 0.0], "size": [0.0, 0.0, 0.0],
                       "amplitude": 1.0, "field component": comp, "function": W.gaussian1(0.1)}]
    return cfg


def test_source_pulse_first_step():
    """After one step the only nonzero E values are the two split Ez dipoles:
    E = dt * (-amp) * w * G1(t - t_c, width) at t = dt/2, w = 1/2 each,
    width = 1/(2 pi f), t_c = 4 width."""
    cfg = _tiny_dipole()
    sim = O.OracleSim(cfg)
    sim.run_steps(1)
    ez = sim.field("Ez")  # [nz, ny, nx], z = -7..7 odd
    dt = 0.5
    width = 1 / (2 * PI * 0.1)
    arg = (dt / 2 - 4 * width) / width
    expect = dt * (-1.0) * 0.5 * arg * math.exp(-arg * arg / 2)
    nz = ez.shape[0]
    c = ez.shape[1] // 2
    assert ez[nz // 2, c, c] == pytest.approx(expect, rel=1e-14)
    assert ez[nz // 2 - 1, c, c] == pytest.approx(expect, rel=1e-14)
    others = ez.copy()
    others[nz // 2, c, c] = 0
    others[nz // 2 - 1, c, c] = 0
    assert np.abs(others).max() == 0
    assert np.abs(sim.field("Ex")).max() == 0


def test_dft_is_direct_sum_of_samples():
    """A one-point DFT monitor equals sum_n x_n exp(-2 pi i f t_n) with
    t_n = (n+1) dt for E and (n+1/2) dt for H (src/simulation.cpp:748-766)."""
    n = 12
    cfg = W.base((n, n, n), 3)
    cfg["sources"] = _tiny_dipole()["sources"]
    freqs = [0.05, 0.11]
    # Ez lattice point (0, 0, 0.5) and Hx lattice point (0, 0.5, 0.5)
    cfg["fields output"] = [W.dft((0, 0, 0.5), (0, 0, 0), freqs, "Ez", "g0"),
                            W.dft((0, 0.5, 0.5), (0, 0, 0), freqs, "Hx", "g1")]
    steps = 160
    sim = O.OracleSim(cfg, probes=[((0, 0, 0.5), "Ez"), ((0, 0.5, 0.5), "Hx")], trace_steps=steps)
    sim.run_steps(steps)
    dt = sim.dt
    t = np.arange(steps)
    for k, toff in ((0, 1.0), (1, 0.5)):
        out = sim.dft_output(k)
        x = sim.trace[:, k]
        for fi, f in enumerate(freqs):
            direct = np.sum(x * np.exp(-2j * PI * f * (t + toff) * dt))
            got = out["real"][fi].ravel()[0] + 1j * out["imag"][fi].ravel()[0]
            assert abs(got - direct) <= 1e-12 * max(abs(direct), 1e-30)
            assert abs(direct) > 0


def _column(nz, steps, freqs, planes, eps=1.0, e_sus=(), extra_dft=()):
    cfg = W.periodic_plane_wave(nz=nz, nxy=2, steps=steps, freqs=freqs, planes=planes, eps=eps, e_sus=e_sus)
    cfg["fields output"] += list(extra_dft)
    return cfg


def _discrete_k(f, dt=0.5, dx=1.0):
    w = 2 * PI * f
    return 2 / dx * math.asin(dx / dt * math.sin(w * dt / 2))


def test_plane_wave_dispersion():
    """Phase advance between two planes 40 cells apart follows the Yee
    dispersion relation sin(w dt/2)/dt = sin(k dx/2)/dx (c = 1).  The residual
    (~1e-3 rad, independent of the separation) comes from the finite CPML and
    time window; the continuous relation k = w is off by 0.3-0.9 rad here."""
    freqs = (0.1, 0.14)
    cfg = W.periodic_plane_wave(nz=128, nxy=2, steps=1400, freqs=freqs, planes=(-20, 20), pml_cells=16)
    sim = O.OracleSim(cfg)
    sim.run(cfg["stop condition"])
    a, b = sim.dft_output(0), sim.dft_output(1)
    for fi, f in enumerate(freqs):
        za = a["real"][fi].mean() + 1j * a["imag"][fi].mean()
        zb = b["real"][fi].mean() + 1j * b["imag"][fi].mean()
        ratio = zb / za
        k = _discrete_k(f)
        err = abs(np.angle(ratio * np.exp(1j * k * 40)))
        err_cont = abs(np.angle(ratio * np.exp(1j * 2 * PI * f * 40)))
        assert err < 3e-3, (f, err)
        assert err_cont > 0.3  # the continuous relation is measurably wrong
        assert abs(abs(ratio) - 1) < 2e-3


def _column_monitor(nz, pml_cells=8, R=1e-6, steps=800, freqs=(0.03, 0.05, 0.07)):
    """DFT 10 cells past a band-limited sheet source (Gaussian1 f = 0.05, cutoff
    8: no content near the Yee band edge, where waves barely move) in a column
    periodic in x, y with CPML on z."""
    cfg = W.periodic_plane_wave(nz=nz, nxy=2, steps=steps, freqs=freqs, planes=(0,), pml_cells=pml_cells)
    cfg["PML"][0]["sigma max"] = W.sigma_max_for(3, pml_cells, R) if R < 1 else 0.0
    cfg["sources"][0]["function"] = W.gaussian1(0.05, cutoff=8)
    zs = -nz / 2 + pml_cells + 4
    cfg["sources"][0]["center"] = [0.0, 0.0, zs]
    cfg["fields output"] = [W.dft((0, 0, zs + 10), (2, 2, 0), freqs, "Ex", "m")]
    sim = O.OracleSim(cfg)
    sim.run(cfg["stop condition"])
    d = sim.dft_output(0)
    return d["real"].mean(axis=(1, 2, 3)) + 1j * d["imag"].mean(axis=(1, 2, 3))


def test_cpml_reflection():
    """Normal-incidence reflection of the CPML, measured against a column long
    enough that nothing returns within the window: small for the R = 1e-6
    design grading, order one without the PML, and larger for a weak grading."""
    ref = _column_monitor(400)
    r_good = np.abs(_column_monitor(80) - ref) / np.abs(ref)
    r_weak = np.abs(_column_monitor(80, R=1e-2) - ref) / np.abs(ref)
    r_none = np.abs(_column_monitor(80, R=1.0) - ref) / np.abs(ref)
    assert np.all(r_good < 2e-4), r_good
    assert np.all(r_none > 0.5), r_none
    assert r_weak[0] > 5 * r_good[0], (r_weak, r_good)


def test_ade_discrete_susceptibility():
    """D(w) / E(w) at one plane equals eps_inf + sum_k amp_k c3_k z /
    (z^2 - c1_k z - c2_k), z = exp(i w dt) -- the closed form of
    ffip/geom.py:252-260 -- once the fields have decayed.  Two Lorentz poles:
    a Drude pole (z = 1) keeps any net charge of the discrete pulse as a
    constant polarization that a finite DFT window does not reject, and the
    reference's Debye recurrence (c2 = 1, c1 < 0) has a root outside the unit
    circle, so it diverges."""
    freqs = (0.06, 0.1, 0.14)
    sus = [W.lorentz(0.12, 0.05, 1.5), W.lorentz(0.07, 0.03, 0.6)]
    cfg = _column(nz=64, steps=3000, freqs=freqs, planes=(0,), eps=2.0, e_sus=sus,
                  extra_dft=[W.dft((0, 0, 0), (2, 2, 0), freqs, "Dx", "d")])
    sim = O.OracleSim(cfg)
    sim.run(cfg["stop condition"])
    e, d = sim.dft_output(0), sim.dft_output(1)
    dt = sim.dt
    for fi, f in enumerate(freqs):
        E = e["real"][fi].mean() + 1j * e["imag"][fi].mean()
        D = d["real"][fi].mean() + 1j * d["imag"][fi].mean()
        z = np.exp(1j * 2 * PI * f * dt)
        expect = 2.0
        shifted = 2.0  # P driven by E^{n+1} instead of E^n: a plausible slip
        for s in sus:
            a0, a1, b0, b1, b2 = O.sus_lorentz(s["frequency"], s["gamma"])
            c0 = 2 * b2 + b1 * dt
            c1 = (4 * b2 - 2 * b0 * dt * dt) / c0
            c2 = (-2 * b2 + b1 * dt) / c0
            c3 = 2 * a0 * dt * dt / c0
            expect += s["amplitude"] * c3 * z / (z ** 2 - c1 * z - c2)
            shifted += s["amplitude"] * c3 * z * z / (z ** 2 - c1 * z - c2)
        # residual ~3e-5: the window ends before the slowest Yee modes decay
        assert abs(D / E - expect) / abs(expect) < 1e-4, (f, D / E, expect)
        assert abs(shifted - expect) / abs(expect) > 1e-2


def test_density_interpolation_matches_trilinear():
    from scipy.interpolate import RegularGridInterpolator
    rng = np.random.default_rng(0)
    dims = (4, 5, 3)  # (nz, ny, nx)
    rho = rng.uniform(size=dims)
    box = O.TwoMediumBox([0, 0, 0], [6, 8, 4], [3, 5, 4], O.AbstractMedium(2, 1, [], []),
                         O.AbstractMedium(1, 1, [], []), rho.ravel())
    X = rng.uniform(-1.99, 1.99, size=(50, 3)) * np.array([1.5, 2.0, 1.0])
    got = box.rho_at(X)
    zs = np.linspace(box.p1[2], box.p2[2], dims[0])
    ys = np.linspace(box.p1[1], box.p2[1], dims[1])
    xs = np.linspace(box.p1[0], box.p2[0], dims[2])
    ref = RegularGridInterpolator((zs, ys, xs), rho)(np.stack([X[:, 2], X[:, 1], X[:, 0]], axis=1))
    assert np.max(np.abs(got - ref)) < 1e-13
    # zero padding: half a node spacing outside the first node along x gives half the edge value
    q = np.array([[box.p1[0] - (box.p2[0] - box.p1[0]) / (dims[2] - 1) / 2, box.p1[1], box.p1[2]]])
    assert box.rho_at(q)[0] == pytest.approx(0.5 * rho[0, 0, 0], rel=1e-12)


def test_two_medium_box_media():
    """Strictly inside: eps = eps1*rho + eps2*(1-rho); on the surface: default."""
    dims = [3, 3, 3]
    cfg = W.base((8, 8, 8), 0)
    cfg["geometry"] = [{"type": "two medium box", "center": [0, 0, 0], "size": [4, 4, 4], "dimension": dims,
                        "density dataset": "rho", "medium1": W.medium(12.0), "medium2": W.medium(1.0)}]
    sim = O.OracleSim(cfg, {"rho": np.full(27, 0.5)})
    g = list(sim.groups[0].values())
    assert len(g) == 1
    ginf = g[0]["ginf"]
    assert set(np.unique(ginf)) == {1.0, 6.5}
    # Ex points strictly inside |x|<2, |y|<2, |z|<2 (doubled |.|<4): x odd in {-3..3} (4),
    # y, z even in {-2, 0, 2} (3 each) -> 36 points per E component
    assert int(np.sum(ginf == 6.5)) == 3 * 36


def test_general_medium_box_reduces_to_two_medium_box():
    """A General_Medium_Box whose table holds two media at densities 0 and 1 is
    the Two_Medium_Box of those media: the table position equals the density and
    the linear interpolation m1*rho + m2*(1 - rho) is the same arithmetic, so the
    oracle's general box must equal its (pinned) two-medium box bit for bit, for
    densities in [0, 1] including both ends."""
    rng = np.random.default_rng(5)
    dims = [4, 3, 5]
    rho = rng.uniform(0, 1, size=dims[0] * dims[1] * dims[2])
    rho[:3] = [0.0, 1.0, 0.5]
    m1 = O.AbstractMedium(7.0, 1.5, [0.3, -0.2], [0.1])
    m2 = O.AbstractMedium(1.0, 1.0, [0.0, 0.05], [0.0])
    tmb = O.TwoMediumBox([0, 0, 0], [6, 4, 8], dims, m1, m2, rho)
    gmb = O.GeneralMediumBox([0, 0, 0], [6, 4, 8], dims, rho, [0.0, 1.0], [m2, m1])
    X = np.concatenate([rng.uniform(-1, 1, size=(200, 3)) * np.array([2.9, 1.9, 3.9]),
                        np.array([[-3.0, -2.0, -4.0], [3.0, 2.0, 4.0], [0.0, 0.0, 0.0]])])
    for a, b in zip(gmb.media(X), tmb.media(X)):
        assert np.array_equal(a, b)


def test_general_medium_box_table_reproduces_the_permittivity():
    """The front end's amplitude table: 1 + sum_i a_i(rho) chi_i(f) = epsilon_fun(rho)
    at the design frequency (chi_i the unit-amplitude responses), to rounding."""
    from ffip_amd.geom import General_Medium_Box, LorentzianSusceptibility, Vector3
    sus = [LorentzianSusceptibility(frequency=0.08, gamma=0.01), LorentzianSusceptibility(frequency=0.2, gamma=0.02)]
    fun = lambda r: 1 + 2.5 * r + 0.2j * r * (1 - r)  # noqa: E731
    g = General_Medium_Box(size=Vector3(2, 2, 2), dim=Vector3(2, 2, 2), frequency=0.1, epsilon_fun=fun, e_sus=sus)
    rho, amps = g.amplitude_table()
    chi = np.array([s.get_epsilon(0.1) / s.sigma for s in sus])
    assert rho.size == 1000 and amps.shape == (1000, 2)
    assert np.max(np.abs(1 + amps @ chi - fun(rho))) < 1e-12
    with pytest.raises(ValueError):
        General_Medium_Box(e_sus=sus[:1], epsilon_fun=fun)


def test_threads_give_the_same_bits():
    """The C loop's OpenMP split writes distinct points per phase: any thread count
    reproduces the single-thread run bit for bit (CPML, curl, two-pole media,
    Drude, DFT monitors)."""
    cfg = W.dispersive_spheres(n=40, nspheres=4, steps=30, nfreq=3, pml_cells=6, precision="fp64")
    out = []
    prev = O.get_threads()
    for n in (1, 4):
        O.set_threads(n)
        s = O.OracleSim(cfg, None, strict_pool=False)
        s.run(cfg["stop condition"])
        out.append(([s.dft_output(k)["real"].copy() for k in range(2)],
                    [s.field(c).copy() for c in ("Ex", "Ez", "Hy")]))
    O.set_threads(prev)
    for a, b in zip(out[0][0] + out[0][1], out[1][0] + out[1][1]):
        assert np.array_equal(a, b)
    assert any(np.abs(f).max() > 0 for f in out[0][1])


# --------------------------------------------------------------- mirror planes
def _lattice(sim, comp):
    """Doubled-grid coordinates (x, y, z arrays) of `comp` inside the grid, the
    axes of ``sim.field(comp)`` reversed."""
    p1, p2 = O.component_interior(sim.grid.grid_p1, sim.grid.grid_p2, O.CTYPE[comp])
    return [np.arange(p1[a], p2[a] + 1, 2) for a in range(3)]


def _dipoles(cfg, dips):
    cfg["sources"] = [{"type": "volume source", "center": list(map(float, c)), "size": [0.0, 0.0, 0.0],
                       "amplitude": float(a), "field component": comp, "function": W.gaussian1(0.15)}
                      for c, comp, a in dips]
    return cfg


class _FlipSide(O.OracleSim):
    """The oracle with one mirror face's multiplier negated: the slip a pin
    must catch."""
    flip = (0, 0)

    def _build_sync(self):
        ops = super()._build_sync()
        a, side = self.flip
        src, dst, mult = ops[a][side]
        ops[a][side] = (src, dst, -mult)
        return ops


# (axis, size, periodic, full-grid dipoles, half-grid dipoles, phase)
_MIRRORS = {
    # Ez on the plane x = 0: Ez, Ey, Hx even in x -> tangential H vanishes (PMC)
    "x-even": (0, (12, 10, 10), None, [((0, 0.3, 0.25), "Ez", 1)], [((0, 0.3, 0.25), "Ez", 1)], 1),
    # Ez pair odd in x -> tangential E vanishes (PEC)
    "x-odd": (0, (12, 10, 10), None, [((2.25, 0.3, 0.25), "Ez", 1), ((-2.25, 0.3, 0.25), "Ez", -1)],
              [((2.25, 0.3, 0.25), "Ez", 1)], -1),
    # Ez pointing through the plane z = 0 and even in z: the mirror image of a
    # polar current flips its normal part, so this is the PEC case
    "z-normal": (2, (10, 10, 12), None, [((0.3, 0.25, 0), "Ez", 1)], [((0.3, 0.25, 0), "Ez", 1)], -1),
    # Ex on y = 0, even in y: tangential, PMC
    "y-even": (1, (10, 12, 10), None, [((0.25, 0, 0.3), "Ex", 1)], [((0.25, 0, 0.3), "Ex", 1)], 1),
}


def _mirror_case(axis, size, periodic, full_dips, half_dips, phase, steps=40, sim_cls=None):
    full = _dipoles(W.base(size, 0), full_dips)
    half = _dipoles(W.base(size, 0), half_dips)
    if periodic:
        full["periodic"] = list(periodic)
        half["periodic"] = list(periodic)
    half["symmetry"] = [{"direction": "xyz"[axis], "phase factor": phase}]
    a, b = O.OracleSim(full), (sim_cls or O.OracleSim)(half)
    a.run_steps(steps)
    b.run_steps(steps)
    worst, scale = 0.0, 0.0
    for comp in O.EH_ORDER:
        fa, fb = a.field(comp), b.field(comp)
        la, lb = _lattice(a, comp), _lattice(b, comp)
        # the half grid's lattice is the full grid's at coordinates >= 0 on `axis`
        keep = la[axis] >= 0
        assert np.array_equal(la[axis][keep], lb[axis])
        sel = [slice(None)] * 3
        sel[2 - axis] = keep
        fa = fa[tuple(sel)]
        assert fa.shape == fb.shape
        worst = max(worst, float(np.abs(fa - fb).max()))
        scale = max(scale, float(np.abs(fa).max()))
    return worst, scale


@pytest.mark.parametrize("name", list(_MIRRORS))
def test_mirror_plane_equals_full_grid(name):
    """PEC/PMC ghost faces (src/fields.cpp:464-482, multiplier phase or -phase by
    the parity of the face coordinate): a half grid with a `symmetry` plane at 0
    reproduces, on its half, the full grid driven by the mirror-symmetric
    sources, and the other phase does not."""
    axis, size, per, fd, hd, phase = _MIRRORS[name]
    worst, scale = _mirror_case(axis, size, per, fd, hd, phase)
    assert scale > 1e-3
    assert worst <= 1e-13 * scale, (worst, scale)
    wrong, _ = _mirror_case(axis, size, per, fd, hd, -phase)
    assert wrong > 1e-2 * scale


def test_mirror_plane_at_odd_coordinate():
    """The positive side of a periodic mirrored axis mirrors about its last
    plane, here at an odd doubled coordinate (13 cells), where the even-lattice
    components (Ey, Ez, Hx) are the ones copied: their multiplier is +phase.  The
    full grid is a periodic 13-cell box holding the mirror image; a slip that
    negates either face's multiplier fails."""
    fd = [((0, 0.3, 0.25), "Ez", 1)]
    args = (0, (13, 10, 10), (1, 0, 0), fd, fd, 1)
    worst, scale = _mirror_case(*args, steps=60)
    assert scale > 1e-3
    assert worst <= 1e-13 * scale, (worst, scale)
    for side in (0, 1):
        cls = type("Flip%d" % side, (_FlipSide,), {"flip": (0, side)})
        wrong, _ = _mirror_case(*args, steps=60, sim_cls=cls)
        assert wrong > 1e-2 * scale, side


# ----------------------------------------------------------- sheet amplitudes
def _sheet_column(P, size, nz=120, pml_cells=16, steps=1600, zmon=20, freqs=(0.03, 0.05, 0.07)):
    """An Ex sheet of the given (x, y) size in a column periodic in x, y with
    period P cells and CPML on z; returns the cross-section mean of the Ex DFT
    on a plane zmon cells past the sheet over one period of distinct points
    (the (0, 0) Floquet mode), and the oracle."""
    cfg = W.periodic_plane_wave(nz=nz, nxy=P, steps=steps, freqs=freqs, planes=(0,), pml_cells=pml_cells)
    zs = -nz / 2 + pml_cells + 4
    src = cfg["sources"][0]
    src["center"] = [0.0, 0.0, zs]
    src["size"] = [float(size[0]), float(size[1]), 0.0]
    src["function"] = W.gaussian1(0.05, cutoff=8)
    cfg["fields output"] = [W.dft((0, 0, zs + zmon), (P, P, 0), freqs, "Ex", "m")]
    sim = O.OracleSim(cfg)
    sim.run(cfg["stop condition"])
    d = sim.dft_output(0)
    # x: the P odd-lattice points inside (the closure adds one outside at each
    # end); y: P + 1 even-lattice points, the last a copy of the first across
    # the periodic boundary
    xs = (d["x"] > -P / 2) & (d["x"] < P / 2)
    ys = (d["y"] >= -P / 2) & (d["y"] < P / 2)
    re = d["real"][:, 0][:, ys][:, :, xs]
    im = d["imag"][:, 0][:, ys][:, :, xs]
    assert re.shape[1:] == (P, P)
    return re.mean(axis=(1, 2)) + 1j * im.mean(axis=(1, 2)), sim


def test_sheet_amplitude_is_a_current_density():
    """Volume-source amplitudes are current densities (src/simulation.cpp:252-262).

    1. Absolute: a 6 x 6-cell sheet (7 x 7 dipoles one cell apart, each split
       half-and-half between two Ex points) in a 10 x 10 periodic column holds
       a mean current of 0.49 per lattice point.  Its (0, 0) mode must equal the
       discrete 1-D Green's function of a uniform sheet of that density,
           E(w) = -J(w) dx^2 sin(w dt/2) / (dt sin(k dx)) exp(-i k d),
       with k from the Yee dispersion relation and J(w) the DFT of the Gaussian1
       samples at t = (n + 1/2) dt.  (Derived from the two 1-D Yee updates with
       a current in the E update; its continuum limit is E = -eta J_s / 2.)
       Scaling by the dipole spacing in doubled-grid units (2, not 1) would
       make this 4x too large.
    2. Relative: stretching the sheet to 6.4 x 6.4 cells keeps 7 x 7 dipoles but
       spreads them 6.4/6 cells apart; at the same density the radiated mode
       grows with the area, (6.4/6)^2, not with the dipole count (ratio 1)."""
    freqs = (0.03, 0.05, 0.07)
    b, sim = _sheet_column(10, (6, 6), freqs=freqs)
    a, _ = _sheet_column(10, (6.4, 6.4), freqs=freqs)
    dt = sim.dt
    width = 1 / (2 * PI * 0.05)
    t = (np.arange(1600) + 0.5) * dt
    on = (t > 0) & (t < 2 * width * 8)
    arg = (t - 8 * width) / width
    g = np.where(on, arg * np.exp(-arg * arg / 2), 0.0)
    for fi, f in enumerate(freqs):
        w = 2 * PI * f
        J = np.sum(g * np.exp(-1j * w * t))
        k = _discrete_k(f)
        expect = -0.49 * J * math.sin(w * dt / 2) / (dt * math.sin(k)) * np.exp(-1j * k * 20)
        assert abs(b[fi] - expect) <= 2e-3 * abs(expect), (f, b[fi], expect)
        ratio = a[fi] / b[fi]
        assert abs(ratio - (6.4 / 6) ** 2) <= 1e-9, (f, ratio)


# ------------------------------------------------------------ stop conditions
def _windows(n_turnoff_time, dt, interval, n_max):
    """Step counts closing each stop-condition window: window 0 ends at the
    first step count n with n dt >= T, window k at the first with
    n dt >= T + k * interval (the loop tests the count before stepping)."""
    ends = []
    k = 0
    n = 0
    while n <= n_max:
        lim = n_turnoff_time + k * interval
        while not n * dt >= lim:
            n += 1
        ends.append(n)
        k += 1
        if n >= n_max:
            break
    return ends


def _differs(slipped, got):
    """A replay with a slipped window differs from the true one in its window
    count or in some window's (max, min) by more than rounding."""
    if len(slipped) != len(got):
        return True
    a, b = np.array(slipped)[:, :2], np.array(got)[:, :2]
    return bool(np.abs(a - b).max() > 1e-6 * np.abs(b).max())


def _stop_cfg(interval, var_or_decay, kind, n=16, at=(1.0, 1.0, 0.5)):
    """Ez dipole at the centre of an n^3 box with CPML; the stop condition
    watches the Ez lattice point `at`."""
    cfg = W.base((n, n, n), 4 if n < 24 else 6)
    cfg["sources"] = [{"type": "volume source", "center": [0.0, 0.0, 0.0], "size": [0.0, 0.0, 0.0],
                       "amplitude": 1.0, "field component": "Ez", "function": W.gaussian1(0.1)}]
    cfg["fields output"] = [W.dft(at, (0, 0, 0), [0.1], "Ez", "m")]
    if kind == "dft":
        cfg["stop condition"] = {"type": "dft", "center": list(at), "size": [0.0, 0.0, 0.0],
                                 "field component": "Ez", "frequency": 0.1,
                                 "time interval examined": interval, "variation": var_or_decay}
    else:
        cfg["stop condition"] = {"type": "decay", "position": list(at), "field component": "Ez",
                                 "time interval examined": interval, "decayed by": var_or_decay}
    return cfg


def _probe_trace(cfg, nsteps, at=(1.0, 1.0, 0.5)):
    """Ez after every step at the monitored lattice point, from the C loop's probe
    recorder (the DFT and probe paths are pinned above)."""
    sim = O.OracleSim(cfg, probes=[(at, "Ez")], trace_steps=nsteps)
    sim.run_steps(nsteps)
    return sim.trace[:, 0].copy(), sim


def test_dft_stop_condition_windows():
    """run_until_dft (src/simulation.cpp:782-842) replayed from a recorded
    series: after the source turns off, windows of `interval` close at the first
    step count reaching T + k*interval; each window's spread (max - min) / max of
    the monitor's DFT norm |sum_{m<=n} x_m exp(-i w m dt)|^2, taken after every
    step of the window, decides whether to go on.  Windows shifted by one sample,
    or norms read before the step, give other numbers."""
    interval, var = 10.0, 5e-3
    cfg = _stop_cfg(interval, var, "dft")
    ref = O.OracleSim(cfg)
    taken = ref.run(cfg["stop condition"])
    log = ref.stop_log
    x, sim = _probe_trace(cfg, taken + 200)
    dt = sim.dt
    n = np.arange(1, x.size + 1)
    S = np.cumsum(x * np.exp(-2j * PI * 0.1 * n * dt))
    norm = np.abs(S) ** 2  # norm[i]: after step i + 1
    T = 2 * (1 / (2 * PI * 0.1)) * 4.0  # Gaussian1 end time: 2 width cutoff
    assert ref.turnoff_time == pytest.approx(T, rel=1e-15)
    ends = _windows(T, dt, interval, x.size - 1)

    def replay(shift):
        out = []
        for k in range(1, len(ends)):
            w = norm[ends[k - 1] + shift:ends[k] + shift]
            mx, mn = w.max(), w.min()
            out.append((mx, mn, (mx - mn) / mx))
            if not out[-1][2] > var:
                return out, ends[k]
        raise AssertionError("series too short")

    got, stop = replay(0)
    assert len(got) >= 3  # several windows exercised
    assert taken == stop
    # the spread is a difference of nearly equal norms: compare it absolutely
    np.testing.assert_allclose(np.array(log)[:, :2], np.array(got)[:, :2], rtol=1e-10)
    np.testing.assert_allclose(np.array(log)[:, 2], np.array(got)[:, 2], rtol=0, atol=1e-12)
    for shift in (-1, 1):
        assert _differs(replay(shift)[0], got)


def test_decay_stop_condition_windows():
    """run_until_fields_decayed (src/simulation.cpp:844-901) replayed from a
    recorded series: the running maximum of x^2 starts over the source-on steps;
    each window's maximum is folded into it, and the run stops at the first
    window whose maximum is at most `decayed by` times the running maximum.  The
    point watched is 10 cells from the source, so the pulse peaks there after
    the source has turned off and the folding decides when the run stops."""
    interval, decay_by = 4.0, 1e-3
    at = (10.0, 1.0, 0.5)
    cfg = _stop_cfg(interval, decay_by, "decay", n=32, at=at)
    ref = O.OracleSim(cfg)
    taken = ref.run(cfg["stop condition"])
    log = ref.stop_log
    x, sim = _probe_trace(cfg, taken + 200, at)
    dt = sim.dt
    sq = x * x
    T = 2 * (1 / (2 * PI * 0.1)) * 4.0
    ends = _windows(T, dt, interval, x.size - 1)

    def replay(shift):
        top = sq[:ends[0] + shift].max()
        out = []
        for k in range(1, len(ends)):
            cur = sq[ends[k - 1] + shift:ends[k] + shift].max()
            top = max(top, cur)
            out.append((cur, top, cur / top))
            if not out[-1][2] > decay_by:
                return out, ends[k]
        raise AssertionError("series too short")

    got, stop = replay(0)
    assert len(got) >= 3
    assert taken == stop
    # the first window's peak exceeds the source-on peak: folding matters
    assert sq[:ends[0]].max() < got[0][0]
    np.testing.assert_allclose(np.array(log), np.array(got), rtol=1e-10)
    for shift in (-1, 1):
        assert _differs(replay(shift)[0], got)
