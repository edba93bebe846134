"""GPU parity of the HIP radio projection against the FP64 oracle.

Exact mode (exact_fp64 = 1): waveforms are compared as the int64 fixed-point
sums the reference accumulates by default (RadioConfig.hpp:30,
RadioProjection.inl:245-301). The kernels keep the reference's FP64 operation
order, so the bar is bit equality of every bin and equality of the
contribution counters.

Fast mode (the default): FP64 with hardware reciprocal / reciprocal-square-root
refinement. Its decoded waveforms are compared with the unquantised FP64
oracle; tolerance per observer and algorithm: relative L2 error <= 1e-6 (the
north-star bar is 1e-3) and every bin within 1e-6 of that observer's peak.

Floating mode (FP64 atomics, order-dependent rounding) is compared with a
tolerance of 1e-10 of the peak. At the full 2^20-track size the checks are
size-independent: superposition, permutation invariance, and the field error
against the FP64 oracle on a subset of antennas.
"""
import numpy as np
import pytest

import c8mi
import oracle_lib
from c8mi import workloads
from radio_cases import (all_reference_cases, random_tracks, reference_case, uniform_table,
                         window_observers)

pytestmark = pytest.mark.gpu


def gpu_project(tracks, observers, propagation, coreas=True, zhs=True, refinement=1,
                deterministic=True, tuning=None, zhs_observers=None, limit=1.0, exact=True):
    acc = c8mi.RadioAccumulator(observers, propagation, coreas=coreas, zhs=zhs,
                                deterministic=deterministic, zhs_refinement=refinement,
                                zhs_observers=zhs_observers, fixed_point_limit=limit,
                                exact=exact)
    if tuning:
        acc.set_tuning(**tuning)
    n = tracks.shape[1]
    dev = c8mi.DeviceArray.from_host(np.ascontiguousarray(tracks))
    acc.accumulate_tracks(dev, n, n)
    if deterministic:
        out = acc.download_fixed()
    else:
        out = acc.download()
    dev.free()
    acc.close()
    return out


def assert_same(gpu, ref, coreas=True, zhs=True):
    c, z, counters = gpu
    if coreas:
        assert np.array_equal(c, ref["coreas_fixed"])
        assert counters["coreas_contributions"] == ref["counters"]["coreas_contributions"]
    if zhs:
        assert np.array_equal(z, ref["zhs_fixed"])
        assert counters["zhs_contributions"] == ref["counters"]["zhs_contributions"]
        assert counters["zhs_subtracks"] == ref["counters"]["zhs_subtracks"]
    assert counters["valid_tracks"] == ref["counters"]["valid_tracks"]
    assert counters["fixed_point_overflows"] == ref["counters"]["fixed_point_overflows"] == 0


@pytest.mark.parametrize("name,step", all_reference_cases())
def test_reference_cases_bit_exact(name, step):
    print(f"case {name} step {step}", flush=True)
    track, obs, propagation = reference_case(name, step)
    ref = oracle_lib.radio_project(propagation, obs, track)
    assert_same(gpu_project(track, obs, propagation), ref)


POSITIONS = ((0., 0., 0.), (150., -40., 0.), (-300., 250., 0.), (30., 30., 0.),
             (500., 0., 0.), (-80., -420., 0.))


def exponential_table(max_h=2500.0):
    refr = workloads.exponential_refractivity(-1000.0, max_h, 1.0)
    return c8mi.Propagation(refractivity=refr, minimum_height_m=-1000.0, maximum_height_m=max_h,
                            step_m=1.0)


@pytest.mark.parametrize("rate,nbins", [(5.0e9, 2048), (1.0e9, 1024), (1.0e10, 4096)])
def test_random_tracks_bit_exact(rate, nbins):
    print(f"random tracks rate {rate} bins {nbins}", flush=True)
    tracks = random_tracks(4096, 5)
    obs = window_observers(tracks, POSITIONS, rate, nbins)
    prop = exponential_table()
    ref = oracle_lib.radio_project(prop, obs, tracks, threads=8)
    assert_same(gpu_project(tracks, obs, prop), ref)


def test_refined_zhs_bit_exact():
    tracks = random_tracks(1024, 9, height=(20.0, 200.0))
    obs = window_observers(tracks, POSITIONS[:3], 5.0e9, 2048)
    prop = exponential_table()
    ref = oracle_lib.radio_project(prop, obs, tracks, coreas=False, refinement=3)
    assert ref["counters"]["zhs_subtracks"] > 3 * 1024
    assert_same(gpu_project(tracks, obs, prop, coreas=False, refinement=3), ref, coreas=False)


@pytest.fixture(scope="module")
def config2_small():
    tracks, obs, refr, (lo, hi) = workloads.config2_setup(count=8192, seed=77)
    prop = c8mi.Propagation(refractivity=refr, minimum_height_m=lo, maximum_height_m=hi, step_m=1.0)
    return tracks, obs, prop


def test_config2_sample_bit_exact(config2_small):
    tracks, obs, prop = config2_small
    print("config2 sample 8192 x 160", flush=True)
    ref = oracle_lib.radio_project(prop, obs, tracks, threads=8)
    assert ref["counters"]["coreas_contributions"] > 0
    assert_same(gpu_project(tracks, obs, prop), ref)


def test_tiling_and_chunking_do_not_change_results(config2_small):
    tracks, obs, prop = config2_small
    base = gpu_project(tracks, obs, prop)
    for tuning in (dict(tile_bins=300), dict(chunk_tracks=256), dict(target_workgroups=64),
                   dict(tile_bins=1000, chunk_tracks=1024), dict(block_threads=256),
                   dict(block_threads=1024), dict(split=False), dict(split=False, block_threads=1024),
                   dict(split=False, tile_bins=500, target_workgroups=2048)):
        print(f"tuning {tuning}", flush=True)
        other = gpu_project(tracks, obs, prop, tuning=tuning)
        assert np.array_equal(other[0], base[0]) and np.array_equal(other[1], base[1])
        assert other[2] == base[2]


def test_single_algorithm_launches_match_fused(config2_small):
    tracks, obs, prop = config2_small
    fused = gpu_project(tracks, obs, prop)
    coreas_only = gpu_project(tracks, obs, prop, zhs=False)
    zhs_only = gpu_project(tracks, obs, prop, coreas=False)
    assert np.array_equal(coreas_only[0], fused[0])
    assert np.array_equal(zhs_only[1], fused[1])


def test_unfused_observer_lists(config2_small):
    tracks, obs, prop = config2_small
    zobs = obs[:40].copy()
    acc = c8mi.RadioAccumulator(obs, prop, zhs_observers=zobs, exact=True)
    assert not acc.fused
    acc.close()
    gpu = gpu_project(tracks, obs, prop, zhs_observers=zobs)
    ref_c = oracle_lib.radio_project(prop, obs, tracks, zhs=False, threads=8)
    ref_z = oracle_lib.radio_project(prop, zobs, tracks, coreas=False, threads=8)
    assert np.array_equal(gpu[0], ref_c["coreas_fixed"])
    assert np.array_equal(gpu[1], ref_z["zhs_fixed"])
    assert gpu[2]["valid_tracks"] == tracks.shape[1]


def test_floating_mode_matches_oracle(config2_small):
    tracks, obs, prop = config2_small
    ref = oracle_lib.radio_project(prop, obs, tracks, threads=8)
    c, z, counters = gpu_project(tracks, obs, prop, deterministic=False)
    for got, want in ((c, ref["coreas_float"]), (z, ref["zhs_float"])):
        peak = np.abs(want).max()
        assert peak > 0
        assert np.abs(got - want).max() <= 1e-10 * peak


def test_lepton_records_paths_match_compact_tracks():
    tracks = random_tracks(10000, 21)
    obs = window_observers(tracks, POSITIONS[:4], 5.0e9, 2048)
    prop = exponential_table()
    records = workloads.tracks_to_records(tracks)
    # photons and muons are not radio tracks (make_radio_track_kinematics)
    records["start"]["pid"][::97] = 22
    records["start"]["pid"][5::113] = 13
    keep = np.ones(tracks.shape[1], dtype=bool)
    keep[::97] = False
    keep[5::113] = False
    ref = oracle_lib.radio_project(prop, obs, tracks[:, keep], threads=8)
    # host enqueue in uneven pieces (batches of 4096 inside)
    acc = c8mi.RadioAccumulator(obs, prop, exact=True)
    for lo, hi in ((0, 1000), (1000, 5000), (5000, 5001), (5001, 10000)):
        acc.enqueue_host_tracks(records[lo:hi])
    assert_same(acc.download_fixed(), ref)
    acc.close()
    # device-resident records
    acc = c8mi.RadioAccumulator(obs, prop, exact=True)
    dev = c8mi.DeviceArray.from_host(records)
    acc.accumulate_records(dev, records.size)
    assert_same(acc.download_fixed(), ref)
    acc.close()


def test_empty_input_and_lifecycle():
    tracks = random_tracks(512, 2)
    obs = window_observers(tracks, POSITIONS[:2], 5.0e9, 512)
    prop = exponential_table()
    acc = c8mi.RadioAccumulator(obs, prop, exact=True)
    acc.accumulate_tracks(0, 0, 0)
    c, z, counters = acc.download()
    assert not c.any() and not z.any() and counters["valid_tracks"] == 0
    with pytest.raises(c8mi.C8miError) as err:
        acc.download()
    assert err.value.status == 2
    acc.reset()
    dev = c8mi.DeviceArray.from_host(tracks)
    acc.accumulate_tracks(dev, 512, 512)
    first = acc.download_fixed()
    acc.reset()
    acc.accumulate_tracks(dev, 512, 512)
    second = acc.download_fixed()
    assert np.array_equal(first[0], second[0]) and np.array_equal(first[1], second[1])
    acc.close()


def test_fixed_point_overflow_is_reported():
    tracks = random_tracks(256, 4)
    obs = window_observers(tracks, POSITIONS[:2], 5.0e9, 1024)
    prop = exponential_table()
    acc = c8mi.RadioAccumulator(obs, prop, fixed_point_limit=1e-40)
    dev = c8mi.DeviceArray.from_host(tracks)
    acc.accumulate_tracks(dev, 256, 256)
    with pytest.raises(c8mi.C8miError) as err:
        acc.download()
    assert err.value.status == 6
    acc.close()


@pytest.fixture(scope="module")
def config2_full():
    tracks, obs, refr, (lo, hi) = workloads.config2_setup()
    prop = c8mi.Propagation(refractivity=refr, minimum_height_m=lo, maximum_height_m=hi, step_m=1.0)
    return tracks, obs, prop


def test_full_size_superposition_and_permutation(config2_full):
    tracks, obs, prop = config2_full
    n = tracks.shape[1]
    print(f"full config2 {n} x {obs.size}", flush=True)
    whole = gpu_project(tracks, obs, prop)
    acc = c8mi.RadioAccumulator(obs, prop, exact=True)
    dev = c8mi.DeviceArray.from_host(tracks)
    half = n // 2
    first = c8mi.DeviceArray.from_host(np.ascontiguousarray(tracks[:, :half]))
    second = c8mi.DeviceArray.from_host(np.ascontiguousarray(tracks[:, half:]))
    acc.accumulate_tracks(first, half, half)
    acc.accumulate_tracks(second, n - half, n - half)
    split = acc.download_fixed()
    acc.close()
    assert np.array_equal(split[0], whole[0]) and np.array_equal(split[1], whole[1])
    perm = np.random.default_rng(1).permutation(n)
    shuffled = gpu_project(np.ascontiguousarray(tracks[:, perm]), obs, prop)
    assert np.array_equal(shuffled[0], whole[0]) and np.array_equal(shuffled[1], whole[1])
    assert whole[2]["valid_tracks"] == n
    del dev


def test_full_size_error_against_fp64_oracle(config2_full):
    tracks, obs, prop = config2_full
    subset = obs[::20].copy()  # 8 antennas, one per arm position class
    print(f"full config2 error check on {subset.size} antennas", flush=True)
    ref = oracle_lib.radio_project(prop, subset, tracks, threads=16, fixed=False)
    for exact in (True, False):
        c, z, counters = gpu_project(tracks, subset, prop, exact=exact)
        cd = oracle_lib.radio_decode(c, subset)
        zd = oracle_lib.radio_decode(z, subset, potential=True)
        for got, want in ((cd, ref["coreas_float"]), (zd, ref["zhs_float"])):
            err = np.linalg.norm(got - want) / np.linalg.norm(want)
            assert err <= 1e-3
            assert err <= 1e-6  # FP64 both ways: fixed-point quantisation dominates


def assert_fast_close(gpu, ref, observers):
    c, z, counters = gpu
    nbins = observers["number_of_bins"].astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(nbins)])
    for fixed, want, potential in ((c, ref["coreas_float"], False), (z, ref["zhs_float"], True)):
        got = oracle_lib.radio_decode(fixed, observers, potential=potential)
        for k in range(observers.size):
            sl = slice(offsets[k], offsets[k + 1])
            w, g = want[:, sl], got[:, sl]
            norm = np.linalg.norm(w)
            if norm == 0:
                assert not np.abs(g).any()
                continue
            assert np.linalg.norm(g - w) <= 1e-6 * norm
            assert np.abs(g - w).max() <= 1e-6 * np.abs(w).max()
    assert counters["valid_tracks"] == ref["counters"]["valid_tracks"]
    assert counters["fixed_point_overflows"] == 0


@pytest.mark.parametrize("name,step", all_reference_cases())
def test_fast_mode_reference_cases(name, step):
    track, obs, propagation = reference_case(name, step)
    ref = oracle_lib.radio_project(propagation, obs, track)
    assert_fast_close(gpu_project(track, obs, propagation, exact=False), ref, obs)


def test_fast_mode_random_and_config2(config2_small):
    tracks = random_tracks(4096, 5)
    obs = window_observers(tracks, POSITIONS, 5.0e9, 2048)
    prop = exponential_table()
    ref = oracle_lib.radio_project(prop, obs, tracks, threads=8)
    assert_fast_close(gpu_project(tracks, obs, prop, exact=False), ref, obs)
    tracks, obs, prop = config2_small
    ref = oracle_lib.radio_project(prop, obs, tracks, threads=8)
    fast = gpu_project(tracks, obs, prop, exact=False)
    assert_fast_close(fast, ref, obs)
    # the fast path is itself deterministic under any launch shape
    other = gpu_project(tracks, obs, prop, exact=False,
                        tuning=dict(block_threads=1024, split=False, tile_bins=700))
    assert np.array_equal(other[0], fast[0]) and np.array_equal(other[1], fast[1])


# ---- in-ice ZHS: homogeneous index n = 1.78, antennas on the Cherenkov cone ----
# The in-ice workload (1 PeV electron in ice, ZHS at 64 antennas on the
# Cherenkov ring 1 km from the shower maximum, 4096 bins @ 0.1 ns) projects
# through the constant-index branch of propagate (RadioProjection.inl:178-181)
# and, on the cone, the near-Cherenkov midpoint form of the endpoint formulae
# (RadioProjectionData.hpp:33). Tracks come from the oracle's in-ice shower.

ICE_INDEX = 1.78
ICE_START, ICE_DOWN = (0.0, 0.0, -100.0), (0.0, 0.0, -1.0)
_ICE_TABLES = {}


def ice_shower_tracks(energy, seed):
    from c8mi import airtables, em

    if "tables" not in _ICE_TABLES:
        _ICE_TABLES["tables"] = airtables.build_ice_tables(cut_mev=0.1)
    ice = em.IceVolume()
    cfg = em.make_config(airtables.IceMedium, seed=seed, cut_mev=0.1, start=ICE_START, direction=ICE_DOWN,
                         atmosphere=ice, thinning=(1.0e-6 * energy, 1.0e4), moliere=True, moliere_table=True,
                         bins=256)
    prim = em.particle(11, energy, ICE_START, ICE_DOWN, history_id=1)
    tracks = oracle_lib.em_run_shower(cfg, _ICE_TABLES["tables"], ice, prim, tracks=True)["tracks"]
    depth = 36.08 / 0.92 / 100.0 * np.log(energy * 1.0e3 / 78.0)
    return np.ascontiguousarray(tracks), depth


def ice_ring(depth, count=64):
    x_max = np.asarray(ICE_START) + depth * np.asarray(ICE_DOWN)
    return workloads.cherenkov_ring_observers(x_max, ICE_DOWN, 1000.0, ICE_INDEX, count=count,
                                              sample_rate_hz=1.0e10, number_of_bins=4096,
                                              emission_time_s=depth / workloads.SPEED_OF_LIGHT)


def near_cone_fraction(tracks, observers, sample=4000):
    """Fraction of (track, antenna) pairs with |1 - n beta.e| < 1e-3 (the
    reference's near-Cherenkov threshold), from the geometry alone."""
    moving = np.flatnonzero((tracks[7] > tracks[6]) & (tracks[8] != 0.0))
    idx = moving[np.linspace(0, moving.size - 1, min(sample, moving.size)).astype(int)]
    t = tracks[:, idx]
    mid = 0.5 * (t[0:3] + t[3:6])
    beta = (t[3:6] - t[0:3]) / (299792458.0 * (t[7] - t[6]))
    near = 0
    for p in observers["position_m"]:
        e = np.asarray(p)[:, None] - mid
        e /= np.linalg.norm(e, axis=0)
        near += np.count_nonzero(np.abs(1.0 - ICE_INDEX * np.einsum("ij,ij->j", beta, e)) < 1.0e-3)
    return near / (idx.size * observers.size)


def assert_zhs_fast_close(got, ref, observers):
    _, z, counters = got
    nbins = observers["number_of_bins"].astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(nbins)])
    decoded = oracle_lib.radio_decode(z, observers, potential=True)
    want = ref["zhs_float"]
    for k in range(observers.size):
        sl = slice(offsets[k], offsets[k + 1])
        norm = np.linalg.norm(want[:, sl])
        assert norm > 0
        assert np.linalg.norm(decoded[:, sl] - want[:, sl]) <= 1e-6 * norm
    assert counters["valid_tracks"] == ref["counters"]["valid_tracks"]
    assert counters["zhs_contributions"] == ref["counters"]["zhs_contributions"]
    assert counters["fixed_point_overflows"] == 0


def test_in_ice_cherenkov_ring_bit_exact_fast_and_tiled():
    tracks, depth = ice_shower_tracks(10.0, 5)
    obs = ice_ring(depth)
    print(f"in-ice ring: {tracks.shape[1]} tracks x {obs.size} antennas x 4096 bins", flush=True)
    assert near_cone_fraction(tracks, obs) > 1.0e-3
    prop = c8mi.Propagation(homogeneous_index=ICE_INDEX)
    ref = oracle_lib.radio_project(prop, obs, tracks, coreas=False, threads=8)
    assert ref["counters"]["zhs_contributions"] > 0
    exact = gpu_project(tracks, obs, prop, coreas=False)
    assert_same(exact, ref, coreas=False)
    # window tiles forced below, at and above the default split of a 4096-bin window
    for tuning in (dict(tile_bins=1000), dict(tile_bins=4096), dict(tile_bins=1365, block_threads=1024),
                   dict(chunk_tracks=512, block_threads=256)):
        print(f"in-ice tuning {tuning}", flush=True)
        other = gpu_project(tracks, obs, prop, coreas=False, tuning=tuning)
        assert np.array_equal(other[1], exact[1]) and other[2] == exact[2]
    ref_float = oracle_lib.radio_project(prop, obs, tracks, coreas=False, threads=8, fixed=False)
    assert_zhs_fast_close(gpu_project(tracks, obs, prop, coreas=False, exact=False), ref_float, obs)
    # CoREAS through the same constant-index branch and near-cone midpoint form
    cref = oracle_lib.radio_project(prop, obs[:16], tracks, zhs=False, threads=8)
    assert_same(gpu_project(tracks, obs[:16], prop, zhs=False), cref, zhs=False)


def test_in_ice_full_size_superposition_permutation_and_error():
    """About 1.3e6 tracks of a 450 GeV in-ice shower on the 64-antenna ring:
    split input and shuffled input give the same int64 sums, and the FP64 error
    of both projection modes on four antennas is the fixed-point quantisation."""
    tracks, depth = ice_shower_tracks(450.0, 11)
    n = tracks.shape[1]
    assert n > 1_000_000
    obs = ice_ring(depth)
    prop = c8mi.Propagation(homogeneous_index=ICE_INDEX)
    print(f"in-ice full size: {n} tracks x {obs.size} antennas", flush=True)
    whole = gpu_project(tracks, obs, prop, coreas=False)
    acc = c8mi.RadioAccumulator(obs, prop, coreas=False, exact=True)
    half = n // 3
    first = c8mi.DeviceArray.from_host(np.ascontiguousarray(tracks[:, :half]))
    second = c8mi.DeviceArray.from_host(np.ascontiguousarray(tracks[:, half:]))
    acc.accumulate_tracks(first, half, half)
    acc.accumulate_tracks(second, n - half, n - half)
    split = acc.download_fixed()
    acc.close()
    assert np.array_equal(split[1], whole[1]) and split[2] == whole[2]
    perm = np.random.default_rng(3).permutation(n)
    shuffled = gpu_project(np.ascontiguousarray(tracks[:, perm]), obs, prop, coreas=False)
    assert np.array_equal(shuffled[1], whole[1])
    subset = obs[::16].copy()
    ref = oracle_lib.radio_project(prop, subset, tracks, coreas=False, threads=16, fixed=False)
    # e+- steps without a track (cut at rest, no movement) carry zero weight and
    # are counted invalid, as the oracle counts them
    assert whole[2]["valid_tracks"] == ref["counters"]["valid_tracks"] > 0.9 * n
    for exact in (True, False):
        _, z, _ = gpu_project(tracks, subset, prop, coreas=False, exact=exact)
        zd = oracle_lib.radio_decode(z, subset, potential=True)
        err = np.linalg.norm(zd - ref["zhs_float"]) / np.linalg.norm(ref["zhs_float"])
        print(f"in-ice FP64 error ({'exact' if exact else 'fast'}): {err:.3e}", flush=True)
        assert err <= 1e-6
