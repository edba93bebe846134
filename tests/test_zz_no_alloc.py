"""The radio projection that allocates nothing (c8mi_radio_project_fixed_noalloc)
against the allocating accumulator path (c8mi_radio_create +
c8mi_radio_accumulate_tracks + c8mi_radio_download_fixed) on the same tracks
and observers: int64 waveforms and counters must be identical bin for bin.

Two calls follow each other on one pair of output buffers and the library's
one scratch array; the second is smaller (fewer tracks, fewer observers,
shorter windows, a different propagation form), so a bin or counter the
second call failed to clear, or a stale observer or table entry, shows up as
a difference. The exact-mode waveforms are also compared with the FP64
oracle."""
import numpy as np
import pytest

import c8mi
import oracle_lib
from radio_cases import random_tracks, uniform_table, window_observers

pytestmark = pytest.mark.gpu


def accumulator_fixed(tracks, obs, prop, exact):
    acc = c8mi.RadioAccumulator(obs, prop, exact=exact)
    n = tracks.shape[1]
    dev = c8mi.DeviceArray.from_host(tracks)
    acc.accumulate_tracks(dev, n, n)
    out = acc.download_fixed()
    acc.close()
    dev.free()
    return out


def noalloc_fixed(tracks, obs, prop, exact, out_coreas, out_zhs):
    n = tracks.shape[1]
    total = int(obs["number_of_bins"].sum())
    dev = c8mi.DeviceArray.from_host(tracks)
    counters = c8mi.radio_project_fixed_noalloc(obs, prop, dev, n, n, device_coreas=out_coreas,
                                                device_zhs=out_zhs, exact=exact)
    dev.free()
    return (out_coreas.to_host(np.int64, (3, total)), out_zhs.to_host(np.int64, (3, total)),
            counters)


def _cases():
    big = random_tracks(24000, 41)
    big_obs = window_observers(big, [(40.0 * k, -25.0 * k, 0.0) for k in range(12)], 5.0e9, 2048)
    small = random_tracks(3000, 42)
    small_obs = window_observers(small, [(-60.0, 80.0, 0.0), (0.0, 0.0, 0.0), (150.0, 10.0, 0.0),
                                         (90.0, -90.0, 0.0), (-5.0, 300.0, 0.0)], 2.0e9, 700)
    return [("24000 tracks x 12 antennas x 2048 bins, tabulated index", big, big_obs,
             uniform_table(1.0003)),
            ("3000 tracks x 5 antennas x 700 bins, constant index", small, small_obs,
             c8mi.Propagation(homogeneous_index=1.0003))]


@pytest.mark.parametrize("exact", [True, False])
def test_noalloc_projection_equals_accumulator_twice_in_a_row(exact):
    cases = _cases()
    largest = max(int(obs["number_of_bins"].sum()) for _, _, obs, _ in cases)
    out_coreas = c8mi.DeviceArray(3 * largest * 8)
    out_zhs = c8mi.DeviceArray(3 * largest * 8)
    for name, tracks, obs, prop in cases:
        print(f"no-alloc projection ({'exact' if exact else 'fast'}): {name}", flush=True)
        want = accumulator_fixed(tracks, obs, prop, exact)
        got = noalloc_fixed(tracks, obs, prop, exact, out_coreas, out_zhs)
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1], want[1])
        assert got[2] == want[2]
        assert want[2]["coreas_contributions"] > 0 and want[2]["zhs_contributions"] > 0
        if exact:
            ref = oracle_lib.radio_project(prop, obs, tracks)
            assert np.array_equal(got[0], ref["coreas_fixed"])
            assert np.array_equal(got[1], ref["zhs_fixed"])
            assert got[2]["valid_tracks"] == ref["counters"]["valid_tracks"]
    out_coreas.free()
    out_zhs.free()


def test_noalloc_projection_rejects_sizes_beyond_its_scratch():
    tracks = random_tracks(64, 43)
    obs = window_observers(tracks, [(float(k), 0.0, 0.0) for k in range(1025)], 1.0e9, 16)
    out = c8mi.DeviceArray(3 * int(obs["number_of_bins"].sum()) * 8)
    with pytest.raises(c8mi.C8miError) as info:
        c8mi.radio_project_fixed_noalloc(obs, uniform_table(1.0003), None, 0, 0,
                                         device_coreas=out, device_zhs=out)
    assert info.value.status == 3
    long_table = c8mi.Propagation(refractivity=np.full(131073, 3.0e-4), minimum_height_m=-1000.0,
                                  maximum_height_m=130072.0, step_m=1.0)
    with pytest.raises(c8mi.C8miError) as info:
        c8mi.radio_project_fixed_noalloc(obs[:4], long_table, None, 0, 0,
                                         device_coreas=out, device_zhs=out)
    assert info.value.status == 3
    out.free()
