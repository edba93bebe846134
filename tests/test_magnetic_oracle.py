"""The oracle's uniform-magnetic-field transport pinned by mathematics.

The reference's headers for this part (UniformMagneticField.hpp,
CurvedBoundaryAccess.hpp, StableQuartic.hpp, EnvironmentLeptonBoundary.hpp)
include Kokkos_Macros.hpp, which is not in the tree, so they cannot be compiled
here; the restatement is checked against closed forms, numpy's polynomial
roots and the trajectory itself:

* root finders (StableQuartic.inl, PolynomialRoots.inl): roots of polynomials
  built from known roots, and numpy.roots on random depressed quartics;
* maximum_uniform_magnetic_step: the gyroradius p_perp / (0.2998 |z| B) from an
  independent cross product, the limit R sin(2 delta), the Linear cases;
* advance_uniform_magnetic_field: the leapfrog kick angle atan(z 0.2998 L B / p)
  for u perpendicular to B, the half-kick displacement, time L / (beta c), the
  bending sense of the charge;
* lepton_boundary: the distance returned puts the particle, moved along the
  leapfrog parabola x0 + L u + (z 0.2998 / 2p) L^2 (u x B) (the curve of
  advance_uniform_magnetic_field), on the sphere it names, and a scan of the
  same parabola finds no earlier crossing; lepton_observation_distance: its
  quadratic has SI coefficients below quadratic_roots' 1e-12, so it is the
  straight-line distance.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib
from c8mi import em

RIGIDITY = 0.2997925146834389  # GeV / (T m)
C = 299792458.0
ME = 0.51099895e-3  # GeV (any mass works for the closed forms)


def poly_from_roots(roots):
    """x^4 + p x^2 + q x + r with the given four roots (their sum must be 0)."""
    c = np.poly(roots)
    assert abs(c[1]) < 1e-9 * max(1.0, np.max(np.abs(roots)))
    return c[2], c[3], c[4]


def test_quartic_roots_from_known_roots():
    rng = np.random.default_rng(3)
    for _ in range(400):
        a, b, c = rng.uniform(-10, 10, 3) * 10.0 ** rng.integers(-3, 4)
        roots = np.array([a, b, c, -(a + b + c)])
        p, q, r = poly_from_roots(roots)
        got = oracle_lib.quartic_roots(p, q, r)
        scale = np.max(np.abs(roots))
        want = np.sort(roots)
        # distinct roots come back ascending; near-double roots may merge
        if np.min(np.diff(want)) > 1e-3 * scale:
            assert len(got) == 4
            assert np.allclose(got, want, rtol=0, atol=1e-8 * scale), (roots, got)
        for x in got:
            assert np.min(np.abs(want - x)) < 1e-6 * scale


def test_quartic_roots_match_numpy_real_roots():
    rng = np.random.default_rng(4)
    checked = 0
    for _ in range(2000):
        p, q, r = rng.normal(size=3) * 10.0 ** rng.integers(-2, 3, 3)
        if rng.random() < 0.2:
            q = 0.0
        got = oracle_lib.quartic_roots(p, q, r)
        z = np.roots([1.0, 0.0, p, q, r])
        scale = max(1.0, np.max(np.abs(z)))
        real = np.sort(z[np.abs(z.imag) < 1e-7 * scale].real)
        if len(real) and np.min(np.abs(z.imag[np.abs(z.imag) > 0] if np.any(z.imag) else [1.0])) < 1e-4 * scale:
            continue  # nearly double roots: the real count is ill-conditioned
        if any(not np.isfinite(x) for x in got):
            continue  # the incomplete-roots signal of a non-positive resolvent
        merged = [x for k, x in enumerate(real) if k == 0 or x - real[k - 1] > 1e-6 * scale]
        assert len(got) == len(merged), (p, q, r, got, real)
        assert np.allclose(got, merged, rtol=0, atol=1e-7 * scale)
        for x in got:  # residual of the polynomial
            assert abs(x ** 4 + p * x * x + q * x + r) <= 1e-9 * max(1.0, scale ** 4)
        checked += 1
    assert checked > 1500


def test_quartic_roots_biquadratic_and_zero():
    assert oracle_lib.quartic_roots(0.0, 0.0, 0.0) == [0.0]
    got = oracle_lib.quartic_roots(-5.0, 0.0, 4.0)  # (x^2 - 1)(x^2 - 4)
    assert np.allclose(got, [-2.0, -1.0, 1.0, 2.0], rtol=0, atol=1e-15)
    assert oracle_lib.quartic_roots(5.0, 0.0, 4.0) == []  # (x^2 + 1)(x^2 + 4)
    got = oracle_lib.quartic_roots(-1.0, 0.0, 0.0)  # x^2 (x^2 - 1)
    assert np.allclose(got, [-1.0, 0.0, 1.0], rtol=0, atol=1e-15)


def test_quadratic_roots_closed_form():
    rng = np.random.default_rng(5)
    for _ in range(500):
        a, b, c = rng.normal(size=3) * 10.0 ** rng.integers(-3, 4, 3)
        got = sorted(oracle_lib.quadratic_roots(a, b, c))
        disc = b * b - 4 * a * c
        if disc < 0:
            assert got == []
            continue
        want = sorted([(-b - np.sqrt(disc)) / (2 * a), (-b + np.sqrt(disc)) / (2 * a)])
        scale = max(abs(want[0]), abs(want[1]))
        assert len(got) == 2 and np.allclose(got, want, rtol=0, atol=1e-9 * scale)
    # |a| < 1e-12: the linear root; |c| < 1e-12: the linear root and 0
    assert oracle_lib.quadratic_roots(1e-13, 2.0, 3.0) == [-1.5]
    assert sorted(oracle_lib.quadratic_roots(2.0, 3.0, 1e-13)) == [-1.5, 0.0]


def lepton(energy, direction, position=(0.0, 0.0, 0.0), node=1, time=0.0):
    p = em.particle(11, energy, position, direction, node=node)
    p["time_s"] = time
    return p


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def test_step_limit_gyroradius_and_limit():
    rng = np.random.default_rng(6)
    for _ in range(300):
        u = unit(rng.normal(size=3))
        b = rng.normal(size=3) * 10.0 ** rng.uniform(-6, 0)
        e = 10.0 ** rng.uniform(-2.5, 2)
        charge = rng.choice([-1.0, 1.0, 2.0])
        delta = rng.uniform(0.01, 1.5)
        st, dist, radius = oracle_lib.magnetic_step_limit(lepton(e, u), ME, charge, b, delta)
        pmag = np.sqrt((e - ME) * (e + ME))
        p_perp = pmag * np.linalg.norm(np.cross(u, b)) / np.linalg.norm(b)
        want_r = p_perp / (RIGIDITY * abs(charge) * np.linalg.norm(b))
        if want_r > 1e9:
            assert st == 1 and dist == np.inf
            continue
        assert st == 0
        assert radius == pytest.approx(want_r, rel=1e-9)
        assert dist == pytest.approx(want_r * np.sin(2 * delta), rel=1e-12)


def test_step_limit_linear_and_invalid_cases():
    u = (0.0, 0.0, 1.0)
    assert oracle_lib.magnetic_step_limit(lepton(1.0, u), ME, 1.0, (0, 0, 0))[0] == 1  # no field
    assert oracle_lib.magnetic_step_limit(lepton(1.0, u), ME, 0.0, (0, 1e-4, 0))[0] == 1  # neutral
    assert oracle_lib.magnetic_step_limit(lepton(1.0, u), ME, 1.0, (0, 0, 1e-4))[0] == 1  # parallel
    assert oracle_lib.magnetic_step_limit(lepton(1.0, u), ME, 1.0, (1e-13, 0, 0))[0] == 1  # R > 1e9 m
    assert oracle_lib.magnetic_step_limit(lepton(ME, u), ME, 1.0, (1e-4, 0, 0))[0] == 2  # at rest
    assert oracle_lib.magnetic_step_limit(lepton(1.0, (0, 0, 1.1)), ME, 1.0, (1e-4, 0, 0))[0] == 2
    assert oracle_lib.magnetic_step_limit(lepton(1.0, u), ME, 1.0, (np.nan, 0, 0))[0] == 2
    assert oracle_lib.magnetic_step_limit(lepton(1.0, u), ME, 1.0, (1e-4, 0, 0), np.pi / 2)[0] == 2


def test_advance_is_the_leapfrog_step():
    rng = np.random.default_rng(7)
    for _ in range(300):
        u = unit(rng.normal(size=3))
        b = np.cross(u, unit(rng.normal(size=3)))  # perpendicular to u
        b *= 10.0 ** rng.uniform(-5, -2) / np.linalg.norm(b)
        e = 10.0 ** rng.uniform(-2.5, 1)
        charge = rng.choice([-1.0, 1.0])
        length = 10.0 ** rng.uniform(-1, 2)
        x0 = rng.normal(size=3) * 100
        st, end = oracle_lib.magnetic_advance(lepton(e, u, x0, time=1e-6), ME, charge, b, length)
        assert st == 0
        pmag = np.sqrt((e - ME) * (e + ME))
        coef = charge * RIGIDITY * length / pmag
        kick = coef * np.cross(u, b)
        d = np.asarray(end["direction"])
        assert np.linalg.norm(d) == pytest.approx(1.0, abs=1e-14)
        angle = np.arctan2(np.linalg.norm(np.cross(u, d)), np.dot(u, d))
        assert angle == pytest.approx(np.arctan(abs(coef) * np.linalg.norm(b)), rel=1e-6, abs=1e-12)
        assert np.sign(np.dot(d, np.cross(u, b))) == np.sign(charge)  # bending sense
        assert np.allclose(end["position_m"], x0 + length * (u + 0.5 * kick), rtol=0, atol=1e-9 * (1 + length))
        assert end["time_s"] == pytest.approx(1e-6 + length / (pmag / e * C), rel=1e-14)
    st, end = oracle_lib.magnetic_advance(lepton(1.0, (0, 0, 1)), ME, 1.0, (0, 0, 1e-3), 5.0)
    assert st == 1 and np.allclose(end["position_m"], (0, 0, 5.0))  # along B: Linear


class Tree:
    """universe > sphere A (R 100 m at the origin) > sphere B (R 20 m at
    (40, 0, 0)), homogeneous media with a uniform field; observation plane
    z = -60 m looking up."""

    def __init__(self, field, observation=True):
        self.models = (em.EnvModel * 2)()
        self.nodes = (em.EnvNode * 3)()
        self.children = (ctypes.c_int32 * 2)(1, 2)
        for k in range(2):
            m = self.models[k]
            m.density_kind = em.DENSITY_HOMOGENEOUS
            m.reference_density_g_per_cm3 = 1e-3
            m.medium_id = 1
            m.axis[2] = 1.0
            for a in range(3):
                m.magnetic_field_t[a] = field[a]
        n = self.nodes
        n[0].shape_kind, n[0].parent, n[0].model, n[0].child_begin, n[0].child_count = em.SHAPE_UNIVERSE, -1, -1, 0, 1
        n[1].shape_kind, n[1].parent, n[1].model, n[1].child_begin, n[1].child_count = em.SHAPE_SPHERE, 0, 0, 1, 1
        n[1].radius_m = 100.0
        n[2].shape_kind, n[2].parent, n[2].model, n[2].child_count = em.SHAPE_SPHERE, 1, 1, 0
        n[2].center_m[0], n[2].radius_m = 40.0, 20.0
        s = self.struct = em.Environment()
        s.nodes, s.node_count, s.children, s.child_index_count = n, 3, self.children, 2
        s.models, s.model_count = self.models, 2
        s.observation_enabled = 1 if observation else 0
        s.observation_point_m[2] = -60.0
        s.observation_normal[2] = 1.0
        s.maximum_magnetic_deflection_rad = 0.2
        self.centers = {1: np.zeros(3), 2: np.array([40.0, 0.0, 0.0])}
        self.radii = {1: 100.0, 2: 20.0}


def parabola(x0, u, b, e, charge, length):
    pmag = np.sqrt((e - ME) * (e + ME))
    k = charge * RIGIDITY / pmag
    return x0[None, :] + length[:, None] * u[None, :] + 0.5 * k * length[:, None] ** 2 * np.cross(u, b)[None, :]


def test_curved_boundary_lies_on_the_trajectory():
    rng = np.random.default_rng(8)
    hits = {"exit": 0, "child": 0, "limit": 0, "observation": 0, "curved": 0}
    for _ in range(1500):
        b = rng.normal(size=3) * 10.0 ** rng.uniform(-4, -2)
        tree = Tree(b)
        x0 = unit(rng.normal(size=3)) * rng.uniform(0, 95)
        if np.linalg.norm(x0 - tree.centers[2]) < 21:
            continue
        u = unit(rng.normal(size=3))
        e = 10.0 ** rng.uniform(-2.5, -0.5)
        charge = rng.choice([-1.0, 1.0])
        p = lepton(e, u, x0, node=1)
        st, dist, nxt, obs = oracle_lib.lepton_boundary(tree, p, ME, charge, b)
        assert st in (0, 1)
        _, limit, radius = oracle_lib.magnetic_step_limit(p, ME, charge, b, 0.2)
        if nxt == 1:
            assert dist == pytest.approx(limit, rel=1e-15)  # limit / speed * speed
            hits["limit"] += 1
        else:
            node = 1 if nxt == 0 else nxt
            hits["exit" if nxt == 0 else "child"] += 1
            c, rr = tree.centers[node], tree.radii[node]
            # the curved branch runs unless the surface is far beyond R or R > 1000 x its distance
            d0 = abs(np.linalg.norm(x0 - c) - rr)
            straight = st == 1 or radius > 1000 * d0
            end = (x0 + dist * u) if straight else parabola(x0, u, b, e, charge, np.array([dist]))[0]
            assert abs(np.linalg.norm(end - c) - rr) < 1e-7 * rr, (st, dist, nxt)
            if not straight:  # no earlier crossing of either sphere along the same parabola
                hits["curved"] += 1
                ls = np.linspace(0, dist, 4001)[:-1]
                pts = parabola(x0, u, b, e, charge, ls)
                inside_a = np.linalg.norm(pts, axis=1) < 100 * (1 + 1e-9)
                outside_b = np.linalg.norm(pts - tree.centers[2], axis=1) > 20 * (1 - 1e-9)
                assert inside_a.all() and outside_b.all()
        # The plane's quadratic has SI coefficients (charge in C, momentum in
        # kg m/s) far below quadratic_roots' absolute epsilon 1e-12, so the
        # reference's branch reduces to the linear root: the straight-line
        # distance along the start direction.
        ahead = -x0[2] - 60.0  # n . (p0 - x0) with n = +z, p0 = (0, 0, -60)
        if u[2] != 0.0 and ahead / u[2] > 0:
            hits["observation"] += 1
            assert obs == pytest.approx(ahead / u[2], rel=1e-12)
        else:
            assert obs == np.inf
    assert min(hits.values()) > 10, hits


def test_zero_field_boundary_is_straight():
    tree = Tree((0.0, 0.0, 0.0))
    p = lepton(0.1, (1.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    st, dist, nxt, obs = oracle_lib.lepton_boundary(tree, p, ME, -1.0, (0.0, 0.0, 0.0))
    assert st == 1 and nxt == 2 and dist == pytest.approx(20.0, rel=1e-15)
    assert obs == np.inf  # parallel to the plane
    p = lepton(0.1, (0.0, 0.0, -1.0), (0.0, 0.0, 0.0))
    st, dist, nxt, obs = oracle_lib.lepton_boundary(tree, p, ME, -1.0, (0.0, 0.0, 0.0))
    assert nxt == 0 and dist == pytest.approx(100.0, rel=1e-15) and obs == pytest.approx(60.0, rel=1e-15)


def test_oracle_shower_in_a_field_closes_and_separates_charges():
    """A 10 GeV photon shower started 3 km above ground in a horizontal field
    (B along +y): energy closes to 1e-9 E0, and the e+- steps bend the way
    q u x B says: for downgoing particles u x B points along +x, so positron
    steps (positive charge-signed track weight) drift to +x and electron steps
    to -x; without the field neither drifts. A 0.05 T field makes the magnetic
    limit end most e+- steps."""
    from c8mi import airtables

    tables = airtables.build_air_tables()
    prim = em.particle(22, 10.0, (0.0, 0.0, 3000.0), (0.0, 0.0, -1.0), history_id=1)
    drift = {}
    for b in (0.0, 1.0e-3, 5.0e-2):
        atm = em.Atmosphere(magnetic_field_t=(0.0, b, 0.0))
        cfg = em.make_config(airtables, seed=31, atmosphere=atm, start=(0.0, 0.0, 3000.0))
        r = oracle_lib.em_run_shower(cfg, tables, atm, prim, tracks=True)
        led = em.decode_counters(r["counters"], cfg.struct.profile_energy_scale)
        assert abs(em.ledger_residual(led)) <= 1e-9 * 10.0
        t = r["tracks"]
        dx, w = t[3] - t[0], t[8]
        pos, ele = w > 0, w < 0
        se = np.hypot(dx[pos].std() / np.sqrt(pos.sum()), dx[ele].std() / np.sqrt(ele.sum()))
        drift[b] = (dx[pos].mean(), dx[ele].mean(), se)
        limits = led["lepton_limits"]
        if b == 5.0e-2:
            assert limits[6] > sum(limits) / 2
    p0, e0, se0 = drift[0.0]
    assert abs(p0 - e0) < 4 * se0
    p1, e1, se1 = drift[1.0e-3]
    assert p1 > 4 * se1 and e1 < -4 * se1
