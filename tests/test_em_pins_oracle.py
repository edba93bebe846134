"""Three parts of the EM restatement pinned against what the physics or the
geometry fixes, each by a check a plausible slip fails.

* The continuous-loss step limit (LeptonContinuousStep.inl:83-120,
  LeptonContinuousStep.hpp:18): a lepton step that ends at the continuous limit
  loses exactly 10% of its total energy (or stops at the table's minimum), and
  the grammage it crossed, integrated in closed form through the exponential
  layer, equals range(E0) - range(E1) of the native range table (a query pinned
  bit for bit against the reference code). A wrong fraction (0.1 for 0.9), a
  kinetic-energy limit or a grammage/arclength slip fails.
* The profile projection (ProfileProjection.inl:12-29 with the axis table):
  project_profile_grammage of points on and off vertical and inclined axes
  equals the slant depth along the axis, in closed form through the
  exponential layers (vertical) or by adaptive quadrature of the layered
  density (inclined, spherical Earth). A wrong sign, unit, start offset or
  direction normalisation fails.
* The ledger residual of a thinned shower: each thinned two-child vertex books
  the transport/native electron mass difference dm of its e+- children at the
  parent's weight in the thinning term and at the kept children's weights in
  the mass-convention term, so the closure residual is exactly
  sum over thinned vertices of dm * (sum of kept e+- weights - n(e+-) * w),
  predicted from the oracle's vertex log and compared with the ledger.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy import integrate

import oracle_lib
from c8mi import airtables, em
from c8mi import tables as native

DOWN = (0.0, 0.0, -1.0)
CONTINUOUS = 1
TRANSPORT_ELECTRON_MASS_GEV = 5.109989e-4
DM = TRANSPORT_ELECTRON_MASS_GEV - airtables.ELECTRON_MASS_MEV / 1000.0


@pytest.fixture(scope="module")
def air():
    return airtables.build_air_tables()


def vertical_grammage(z_low, z_high, layers=em.USSTDBK):
    """g/cm^2 between heights z_low < z_high (m) in the layered atmosphere."""
    total, lower = 0.0, -math.inf
    for k, (top, b, c) in enumerate(layers):
        lo, hi = max(z_low, lower), min(z_high, top)
        if hi > lo:
            if k == len(layers) - 1:
                total += b / c * (hi - lo) * 100.0
            else:
                total += b * (math.exp(-lo * 100.0 / c) - math.exp(-hi * 100.0 / c))
        lower = top
    return total


def native_range(tables, pid, e_mev):
    q = np.zeros(1, dtype=native.QUERY_DTYPE)
    q["kind"], q["pdg_id"], q["energy_mev"] = native.QUERY["range"], pid, e_mev
    r = oracle_lib.native_queries(tables, q)
    assert r["status"][0] == 0
    return float(r["value"][0])


def test_continuous_loss_limit(air):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=17, atmosphere=atm)
    one_round = em.hybrid_config(minimum_batch=1 << 40, expansion_rounds=1)
    rng = np.random.default_rng(4)
    checked = 0
    for i in range(400):
        pid = 11 if i % 2 else -11
        e0 = float(np.exp(rng.uniform(np.log(0.02), np.log(20.0))))
        z0 = float(rng.uniform(12000.0, 36000.0))
        p = em.particle(pid, e0, (0.0, 0.0, z0), DOWN, history_id=1)
        want = oracle_lib.em_run_hybrid(cfg, air, atm, p, shower_id=100 + i, hybrid=one_round, prefix_only=True)
        led = em.decode_counters(want["host"]["counters"], cfg.struct.profile_energy_scale)
        if led["lepton_limits"][CONTINUOUS] != 1 or led["steps"] != 1:
            continue
        cont = want["staged"]
        assert cont.size == 1 and cont["pid"][0] == pid and cont["step_id"][0] == 1
        e1 = float(cont["energy_gev"][0])
        minimum = 0.6  # MeV: the lower end of the e+- utility tables
        target = max(0.9 * e0 * 1000.0, minimum) / 1000.0
        assert abs(e1 - target) <= 1e-9 * e0, (e0, e1)
        # the deposit is the loss, at the particle's weight
        assert abs(led["continuous_deposit"] - (e0 - e1)) <= 1e-9 * e0
        # the grammage crossed (vertical chord through the layer, closed form)
        # is the range difference of the native table
        track = want["host"]["tracks"]
        assert track.shape[1] == 1
        z_start, z_end = track[2, 0], track[5, 0]
        assert z_start == z0 and z_end < z_start
        crossed = vertical_grammage(z_end, z_start)
        expected = native_range(air, pid, e0 * 1000.0) - native_range(air, pid, e1 * 1000.0)
        assert abs(crossed - expected) <= 1e-6 * expected, (crossed, expected)
        checked += 1
    assert checked >= 40, checked


def project(cfg, point):
    lib = oracle_lib.load()
    lib.oracle_em_project_grammage.restype = ctypes.c_double
    p = (ctypes.c_double * 3)(*point)
    return lib.oracle_em_project_grammage(ctypes.byref(cfg.struct), p)


def test_profile_projection_vertical_closed_form():
    atm = em.Atmosphere()
    z0 = 112750.0
    cfg = em.make_config(airtables, seed=1, atmosphere=atm, start=(0.0, 0.0, z0), direction=DOWN)
    rng = np.random.default_rng(9)
    for z in np.concatenate([rng.uniform(0.0, z0, 300), [z0, z0 - 5.0, 6999.0, 7001.0, 100000.0]]):
        want = vertical_grammage(z, z0)
        # any point of the plane through the axis point projects alike
        for offset in ((0.0, 0.0), (350.0, -120.0)):
            got = project(cfg, (offset[0], offset[1], z))
            assert abs(got - want) <= 1e-6 * want + 2e-6, (z, got, want)


def test_profile_projection_inclined_quadrature():
    atm = em.Atmosphere()
    zenith = math.radians(55.0)
    d = (math.sin(zenith), 0.0, -math.cos(zenith))
    start = (0.0, 0.0, 60000.0)
    cfg = em.make_config(airtables, seed=1, atmosphere=atm, start=start, direction=d)
    center = np.array([0.0, 0.0, -em.EARTH_RADIUS_M])
    tops = [top for top, _, _ in em.USSTDBK]

    def height(s):
        p = np.array(start) + s * np.array(d)
        return np.linalg.norm(p - center) - em.EARTH_RADIUS_M

    # layer crossings along the axis, for the quadrature's break points
    s_axis = np.linspace(0.0, 60000.0 / math.cos(zenith), 20001)
    h_axis = np.array([height(s) for s in s_axis])
    breaks = []
    for top in tops:
        idx = np.nonzero(np.diff(np.sign(h_axis - top)))[0]
        breaks.extend(s_axis[idx].tolist())
    rho = lambda s: float(atm.density(np.array(start) + s * np.array(d))) * 100.0
    for s in (1234.5, 20000.0, 47000.25, 80000.0, 100100.0):
        inner = [b for b in breaks if 0.0 < b < s]
        want = integrate.quad(rho, 0.0, s, points=inner or None, limit=400, epsabs=1e-10, epsrel=1e-11)[0]
        got = project(cfg, tuple(np.array(start) + s * np.array(d)))
        assert abs(got - want) <= 1e-6 * want + 2e-6, (s, got, want)


@pytest.mark.parametrize("seed,energy", [(5, 30.0), (6, 100.0), (7, 10.0)])
def test_thinned_ledger_residual_is_the_mass_convention_term(air, seed, energy):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, thinning=(0.5, 50.0))
    prim = em.particle(22, energy, (0.0, 0.0, 112750.0), DOWN, history_id=1)
    vertices, _ = oracle_lib.em_vertices(cfg, air, atm, prim)
    want = oracle_lib.em_run_shower(cfg, air, atm, prim)
    led = em.decode_counters(want["counters"], cfg.struct.profile_energy_scale)
    assert sum(led["fallback_reasons"]) == 0
    thinned = vertices[(vertices["kind"] == 1) & (vertices["thin_status"] != 0)]
    assert thinned.size > 100
    predicted = 0.0
    for r in thinned:
        leptons = [k for k in range(2) if abs(r["child_pid"][k]) == 11]
        kept = sum(r["child_weight"][k] for k in leptons if (r["keep"] >> k) & 1)
        predicted += DM * (kept - len(leptons) * r["parent_weight"])
    residual = em.ledger_residual(led)
    assert abs(predicted) > 1e-9
    # fixed-point rounding of the ledger (2^-40 GeV per booking) is the rest
    assert abs(residual - predicted) <= 3e-10 + 0.02 * abs(predicted), (residual, predicted)
    # the closure without thinning is exact to the rounding
    plain = em.make_config(airtables, seed=seed, atmosphere=atm)
    unthinned = em.decode_counters(oracle_lib.em_run_shower(plain, air, atm, prim)["counters"],
                                   plain.struct.profile_energy_scale)
    assert abs(em.ledger_residual(unthinned)) <= 1e-9 * energy
