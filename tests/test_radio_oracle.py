"""Pins the FP64 radio restatement (oracle/radio_oracle.c) on the CPU.

Each check is chosen so that a plausible slip in the restatement — a dropped
factor, a wrong sign, the naive half-bin rounding, a missing fraction or branch —
fails against an independent statement of the physics or of the reference's
documented edge semantics:

* the half-bin rule of observer_nearest_bin (RadioProjection.inl:75-82);
* the exact Doppler factor in the cancelling case (RadioProjection.inl:113-143);
* propagate(): uniform tables reproduce n R / c; edge branches
  (RadioProjection.inl:170-238);
* separate_same_bin_endpoints always yields adjacent bins (RadioProjection.inl:303-336);
* the CoREAS endpoint field equals the Lienard-Wiechert radiation term
  q/(4 pi eps0 c) * f_s * e x (e x beta) / ((1 - n beta.e) R) (endpoint formalism,
  James et al., Phys. Rev. E 84, 056602 (2011));
* the ZHS potential integrates to q/(4 pi eps0 c) * beta_perp * dt * f_s / R over
  the bins a segment spans, whatever the binning;
* the ZHS potential's leading edge, differentiated as the output layer does
  (TimeDomainObserver.inl:116-119), reproduces the CoREAS start field;
* the table builder equals a sequential cumulative sum.
"""
import math

import numpy as np
import pytest

import c8mi
import oracle_lib
from radio_cases import (C, ELEMENTARY_CHARGE, all_reference_cases, case_observers, case_track,
                         em_constant, reference_case, uniform_table)


def test_nearest_bin_half_rule():
    below_half = np.nextafter(0.5, 0.0)
    # floor(x + 0.5) would round nextafter(0.5, 0) + 0.5 up to 1
    assert math.floor(below_half + 0.5) == 1.0
    assert oracle_lib.nearest_bin(below_half) == 0.0
    assert oracle_lib.nearest_bin(0.0) == 0.0
    assert oracle_lib.nearest_bin(0.5) == 1.0
    assert oracle_lib.nearest_bin(1.5) == 2.0
    assert oracle_lib.nearest_bin(np.nextafter(2.5, 0.0)) == 2.0
    assert oracle_lib.nearest_bin(-0.5) == 0.0
    assert oracle_lib.nearest_bin(-0.51) == -1.0
    assert oracle_lib.nearest_bin(1234.75) == 1235.0


def test_doppler_ordinary_and_exact_cancellation():
    beta = np.array([0.3, 0.4, 0.5])
    emit = np.array([1.0, 2.0, 2.0]) / 3.0
    for variant in (False, True):
        value = oracle_lib.doppler(1.0003, beta, emit, variant)
        assert value == pytest.approx(1.0 - 1.0003 * float(beta @ emit), rel=1e-15)
    # (1 - 2^-30)(1 + 2^-30) = 1 - 2^-60 rounds to 1 in double: the ordinary
    # expression cancels to 0; the exact answer is 2^-60.
    beta = np.array([1.0 - 2.0 ** -30, 0.0, 0.0])
    emit = np.array([1.0 + 2.0 ** -30, 0.0, 0.0])
    assert 1.0 - 1.0 * float(beta @ emit) == 0.0
    for variant in (False, True):
        assert oracle_lib.doppler(1.0, beta, emit, variant) == 2.0 ** -60
    # Split across components: still exact.
    beta = np.array([0.5 - 2.0 ** -31, 0.5, 0.0])
    emit = np.array([1.0 + 2.0 ** -30, 1.0, 0.0])
    exact = 1.0 - ((0.5 - 2.0 ** -31) * (1.0 + 2.0 ** -30) + 0.5)
    assert oracle_lib.doppler(1.0, beta, emit, True) == pytest.approx(2.0 ** -61, rel=1e-12)
    assert abs(exact) < 1e-15


def test_propagate_uniform_table_matches_homogeneous():
    rng = np.random.default_rng(7)
    table = uniform_table(1.0003, 1.0)
    homogeneous = c8mi.Propagation(homogeneous_index=1.0003)
    for _ in range(200):
        src = np.array([rng.uniform(-500, 500), rng.uniform(-500, 500), rng.uniform(10.2, 1900.0)])
        dst = np.array([rng.uniform(-300, 300), rng.uniform(-300, 300), 0.0])
        t_tab, n_s, n_d, dist, emit = oracle_lib.propagate(table, src, dst)
        t_hom, n_s2, n_d2, dist2, emit2 = oracle_lib.propagate(homogeneous, src, dst)
        r = float(np.linalg.norm(dst - src))
        assert dist == pytest.approx(r, rel=1e-15)
        assert np.allclose(emit, (dst - src) / r, rtol=0, atol=1e-15)
        assert t_hom == pytest.approx(1.0003 * r / C, rel=1e-15)
        # integrated refractivity over the height difference / height = n - 1
        assert t_tab == pytest.approx(t_hom, rel=1e-12)
        assert n_s == pytest.approx(1.0003, rel=1e-15) and n_d == pytest.approx(1.0003, rel=1e-15)


def test_propagate_edge_branches():
    table = uniform_table(1.0003, 1.0)
    dst = np.array([100.0, 2.0, 0.0])
    # exact lower edge: both indices shadowed to 1 (RadioProjection.inl:213-220)
    _, n_s, n_d, _, _ = oracle_lib.propagate(table, np.array([3.0, 4.0, -1000.0]), dst)
    assert (n_s, n_d) == (1.0, 1.0)
    # below the table: source extrapolated, destination shadowed to 1
    _, n_s, n_d, _, _ = oracle_lib.propagate(table, np.array([3.0, 4.0, -1001.0]), dst)
    assert n_s == pytest.approx(1.0003, rel=1e-15) and n_d == 1.0
    # above the table: upper extrapolation, destination kept
    t, n_s, n_d, dist, _ = oracle_lib.propagate(table, np.array([3.0, 4.0, 2500.0]), dst)
    assert n_s == pytest.approx(1.0003, rel=1e-15) and n_d == pytest.approx(1.0003, rel=1e-15)
    assert np.isfinite(t) and t > dist / C


@pytest.mark.parametrize("start,end", [(10.2, 10.3), (10.6, 10.7), (10.6, 11.3), (10.9, 11.45),
                                       (11.3, 10.6), (10.3, 10.2), (11.45, 10.9)])
@pytest.mark.parametrize("signed_order", [False, True])
def test_separate_same_bin_gives_adjacent_bins(start, end, signed_order):
    grid = 1e-9
    s0, e0 = start * grid, end * grid
    assert oracle_lib.nearest_bin(s0 / grid) == oracle_lib.nearest_bin(e0 / grid)
    s, e = oracle_lib.separate_same_bin(s0, e0, grid, signed_order)
    sb, eb = oracle_lib.nearest_bin(s / grid), oracle_lib.nearest_bin(e / grid)
    assert abs(sb - eb) == 1.0
    # exactly one endpoint moved by exactly one grid step
    moved = sorted([abs(s - s0), abs(e - e0)])
    assert moved[0] == 0.0
    assert moved[1] == pytest.approx(grid, rel=1e-9)
    if signed_order:
        # the signed form keeps the time order of the two endpoints
        assert (e >= s) == (e0 >= s0)


def test_separate_leaves_distinct_bins_alone():
    grid = 1e-9
    s, e = oracle_lib.separate_same_bin(10.3 * grid, 10.8 * grid, grid, False)
    assert (s, e) == (10.3 * grid, 10.8 * grid)


def _single(track, observers, propagation, **kw):
    return oracle_lib.radio_project(propagation, observers, track, **kw)


def test_coreas_start_field_is_lienard_wiechert():
    propagation = c8mi.Propagation(homogeneous_index=1.0)
    start, end, beta = np.array([10.0, -5.0, 1000.0]), np.array([10.6, -5.3, 999.2]), 0.95
    track = case_track(start, end, beta, 0.0, 11, 2.5)
    obs = case_observers(8.0e-6, rate=2.0e9, positions=((300.0, 40.0, 0.0),))
    out = _single(track, obs, propagation, zhs=False)
    field = out["coreas_float"]
    r_vec = np.asarray(obs[0]["position_m"]) - start
    r = np.linalg.norm(r_vec)
    e_hat = r_vec / r
    duration = track[7, 0] - track[6, 0]
    beta_vec = (end - start) / (C * duration)
    charge = -ELEMENTARY_CHARGE * 2.5
    expected = charge * em_constant() * 2.0e9 * np.cross(e_hat, np.cross(e_hat, beta_vec)) / (
        (1.0 - beta_vec @ e_hat) * r)
    bin_index = int(round((track[6, 0] + r / C) * 2.0e9))
    assert np.allclose(field[:, bin_index], expected, rtol=1e-12, atol=0)
    nonzero = np.flatnonzero(np.abs(field).sum(axis=0))
    assert len(nonzero) == 2 and bin_index in nonzero
    assert out["counters"]["coreas_contributions"] == 2


def test_zhs_potential_integral_is_binning_independent():
    propagation = c8mi.Propagation(homogeneous_index=1.0)
    start, end, beta = np.array([0.0, 0.0, 1000.0]), np.array([1.5, 0.5, 999.0]), 0.9
    track = case_track(start, end, beta, 0.0, -11, 1.0)
    mid = 0.5 * (start + end)
    duration = track[7, 0] - track[6, 0]
    beta_vec = (end - start) / (C * duration)
    for rate in (1.0e9, 5.0e9, 2.0e10):
        obs = case_observers(8.0e-6, rate=rate, positions=((600.0, 50.0, 0.0),))
        out = _single(track, obs, propagation, coreas=False)
        r_vec = np.asarray(obs[0]["position_m"]) - mid
        r = np.linalg.norm(r_vec)
        e_hat = r_vec / r
        beta_perp = np.cross(e_hat, np.cross(beta_vec, e_hat))
        expected = ELEMENTARY_CHARGE * em_constant() * beta_perp * duration * rate / r
        assert np.allclose(out["zhs_float"].sum(axis=1), expected, rtol=1e-12, atol=0)
        # fixed-point decode agrees within the quantisation of the scale
        decoded = oracle_lib.radio_decode(out["zhs_fixed"], obs, potential=True)
        assert np.allclose(decoded, out["zhs_float"], rtol=0,
                           atol=4 * 2.0 ** -62 / rate)


def test_zhs_leading_edge_reproduces_coreas_start_field():
    propagation = c8mi.Propagation(homogeneous_index=1.0)
    start, end, beta = np.array([0.0, 0.0, 2000.0]), np.array([0.8, 0.0, 1999.4]), 0.9
    track = case_track(start, end, beta, 0.0, 11, 1.0)
    rate = 4.0e10
    obs = case_observers(1.0e-5, rate=rate, positions=((-900.0, 100.0, 0.0),))
    out = _single(track, obs, propagation)
    coreas, zhs = out["coreas_float"], out["zhs_float"]
    start_bin = int(np.flatnonzero(np.abs(coreas).sum(axis=0))[0])
    nz = np.flatnonzero(np.abs(zhs).sum(axis=0))
    assert len(nz) > 8
    # The product's output step (c8mi_radio_zhs_electric_field): across the
    # rising edge E = -(A[i+1] - A[i]) f_s is minus one full bin's potential
    # times f_s, which must equal the CoREAS start field; the falling edge the
    # CoREAS end field.
    field, times = c8mi.zhs_electric_field(obs, zhs)
    assert field.shape == (3, zhs.shape[1] - 1)
    full = zhs[:, nz[len(nz) // 2]]
    edge = np.flatnonzero(np.abs(field).sum(axis=0))
    e_start = field[:, edge[0]:edge[0] + 2].sum(axis=1)
    assert np.allclose(coreas[:, start_bin], e_start, rtol=3e-3, atol=0)
    assert np.allclose(e_start, -full * rate, rtol=3e-3, atol=0)
    end_bin = int(np.flatnonzero(np.abs(coreas).sum(axis=0))[-1])
    e_end = field[:, edge[-1] - 1:edge[-1] + 1].sum(axis=1)
    assert np.allclose(coreas[:, end_bin], e_end, rtol=3e-3, atol=0)
    # midpoint times: half a bin after each axis time
    axis = obs[0]["start_time_s"] + np.arange(zhs.shape[1]) / rate
    assert np.allclose(times, 0.5 * (axis[1:] + axis[:-1]), rtol=0, atol=1e-18)
    # between the edges the field vanishes (a constant potential)
    inner = field[:, edge[0] + 2:edge[-1] - 1]
    assert np.abs(inner).max() <= 1e-9 * np.abs(e_start).max()


def test_charge_and_weight_linearity():
    propagation = uniform_table(1.0003)
    obs = case_observers(2.0e-6)
    a = case_track((3., 4., 100.), (4., 6., 99.), .8, 0., 11, 1.0)
    b = case_track((3., 4., 100.), (4., 6., 99.), .8, 0., -11, 37.0)
    oa = _single(a, obs, propagation)
    ob = _single(b, obs, propagation)
    for key in ("coreas_float", "zhs_float"):
        assert np.allclose(ob[key], -37.0 * oa[key], rtol=1e-14, atol=0)


def test_fraunhofer_subdivision_count():
    propagation = c8mi.Propagation(homogeneous_index=1.0)
    start, end, beta = np.array([0.0, 0.0, 60.0]), np.array([5.0, 0.0, 58.0]), 0.99
    track = case_track(start, end, beta, 0.0, 11, 1.0)
    rate = 5.0e9
    obs = case_observers(2.0e-6, rate=rate, positions=((0.0, 80.0, 0.0),))
    out = _single(track, obs, propagation, coreas=False)
    mid = 0.5 * (start + end)
    r_vec = np.asarray(obs[0]["position_m"]) - mid
    r = np.linalg.norm(r_vec)
    u = (end - start) / np.linalg.norm(end - start)
    cos_t = u @ (r_vec / r)
    length = np.linalg.norm(end - start)
    fraunhofer = (1 - cos_t ** 2) * length ** 2 / r / (C / rate) * 2 * math.pi
    assert fraunhofer > 1
    assert out["counters"]["zhs_subtracks"] == math.ceil(math.sqrt(fraunhofer) + 1)
    refined = _single(track, obs, propagation, coreas=False, refinement=4)
    assert refined["counters"]["zhs_subtracks"] == 4 * math.ceil(math.sqrt(fraunhofer) + 1)
    # the refined projection approximates the same time integral
    assert np.allclose(refined["zhs_float"].sum(axis=1), out["zhs_float"].sum(axis=1),
                       rtol=2e-2, atol=0)


def test_flat_table_builder_is_running_sum():
    rng = np.random.default_rng(3)
    refractivity = 3e-4 * np.exp(-np.arange(5000) / 8000.0) + rng.uniform(0, 1e-9, 5000)
    for step in (1.0, 0.5, 2.0):
        integrated, slopes = oracle_lib.build_flat_table(refractivity, 1.0 / step)
        expected = np.cumsum(refractivity * (1.0 / step))
        assert np.array_equal(integrated, expected)
        assert slopes[0] == (refractivity[10] - refractivity[0]) / 10.0
        assert slopes[3] == (integrated[-1] - integrated[-11]) / 10.0
        prop = c8mi.Propagation(refractivity=refractivity, minimum_height_m=-1000.0,
                                maximum_height_m=-1000.0 + 4999 * step, step_m=step)
        assert np.array_equal(prop.integrated, integrated)


@pytest.mark.parametrize("name,step", all_reference_cases())
def test_reference_cases_fixed_matches_floating(name, step):
    track, obs, propagation = reference_case(name, step)
    out = _single(track, obs, propagation)
    for key, potential in (("coreas", False), ("zhs", True)):
        decoded = oracle_lib.radio_decode(out[f"{key}_fixed"], obs, potential=potential)
        scale = (2.0 ** 62) * (obs["sample_rate_hz"][0] if potential else 1.0)
        assert np.allclose(decoded, out[f"{key}_float"], rtol=0, atol=2.0 / scale * 1.0)
    assert out["counters"]["valid_tracks"] == 1
    assert out["counters"]["fixed_point_overflows"] == 0


def test_threaded_oracle_matches_serial():
    from radio_cases import random_tracks, window_observers

    tracks = random_tracks(600, 11)
    obs = window_observers(tracks, ((0., 0., 0.), (150., -40., 0.), (-300., 250., 0.)), 2.0e9, 1024)
    propagation = uniform_table(1.0003)
    serial = oracle_lib.radio_project(propagation, obs, tracks, threads=1)
    threaded = oracle_lib.radio_project(propagation, obs, tracks, threads=4)
    for key in ("coreas_fixed", "zhs_fixed"):
        assert np.array_equal(serial[key], threaded[key])
    assert serial["counters"] == threaded["counters"]
