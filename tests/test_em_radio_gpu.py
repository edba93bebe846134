"""Radio from the EM lepton fronts on the GPU: every e+- step of the device
shower is projected onto an attached radio accumulator (the reference's
per-front RadioAccumulator::accumulateLeptonTracks call,
corsika/detail/kokkos/em/execution/LeptonCascade.inl:81-82).

The device shower takes the oracle's discrete history (every step, vertex and
secondary), but its positions and times agree with the oracle's only to the
last bits (the device contracts products into FMAs and has its own exp/log). A
track moved by 1e-15 s can put one contribution into the neighbouring time bin,
so whole-shower waveforms are compared by what such shifts leave unchanged: the
contribution counters exactly, the time-integrated field of every observer and
component to 1e-9 of its magnitude sum, and the waveform of every observer to
1e-5 in the L2 norm (measured: below 1e-7; the radio path's stated field
tolerance is 1e-3). The projection itself is bit-exact for identical tracks
(the radio tests)."""
import numpy as np
import pytest

import c8mi
import oracle_lib
from c8mi import airtables, em, workloads
from radio_cases import window_observers

pytestmark = pytest.mark.gpu

START = (0.0, 0.0, 112750.0)
DOWN = (0.0, 0.0, -1.0)


@pytest.fixture(scope="module")
def tables():
    return airtables.build_air_tables()


def propagation():
    refractivity = workloads.exponential_refractivity(-1000.0, 113000.0, 1.0)
    return c8mi.Propagation(refractivity=refractivity, minimum_height_m=-1000.0,
                            maximum_height_m=113000.0, step_m=1.0)


def ground_antennas():
    return [(r * np.cos(a), r * np.sin(a), 0.0) for r in (30.0, 120.0) for a in (0.0, 2.1, 4.2)]


def oracle_tracks(cfg, tables, atm, primaries, shower_id):
    want = oracle_lib.em_run_shower(cfg, tables, atm, primaries, shower_id=shower_id,
                                    first_history_id=2, tracks=True)
    assert want["tracks"].shape == (9, want["stats"]["lepton_steps"])
    return want


def assert_waveforms_close(got, want, observers):
    """got / want: [3][total bins] int64 fixed point of the same observers."""
    got = got.astype(np.float64)
    want = want.astype(np.float64)
    begin = 0
    worst = 0.0
    for n in observers["number_of_bins"]:
        g, w = got[:, begin:begin + n], want[:, begin:begin + n]
        begin += n
        for k in range(3):
            magnitude = np.abs(w[k]).sum()
            assert abs(g[k].sum() - w[k].sum()) <= 1e-9 * magnitude + 8 * n, (k, g[k].sum(), w[k].sum())
        norm = np.linalg.norm(w)
        if norm > 0:
            worst = max(worst, np.linalg.norm(g - w) / norm)
            assert np.linalg.norm(g - w) <= 1e-5 * norm
        else:
            assert not g.any()
    print(f"max relative L2 difference over observers: {worst:.3e}; "
          f"differing samples {np.count_nonzero(got != want)} of {np.count_nonzero(want)} nonzero")


def assert_same_radio(got, ref, observers):
    coreas, zhs, counters = got
    for k in ("valid_tracks", "coreas_contributions", "zhs_contributions", "zhs_subtracks"):
        assert counters[k] == ref["counters"][k], k
    assert counters["fixed_point_overflows"] == 0
    assert_waveforms_close(coreas, ref["coreas_fixed"], observers)
    assert_waveforms_close(zhs, ref["zhs_fixed"], observers)


def device_radio(cfg, tables, atm, observers, prop, showers, exact=True):
    """Runs `showers` [(shower_id, primaries)] one after another in slot 0 with
    the radio attached; returns the fixed-point waveforms and run totals."""
    dev = em.EmBackend(cfg, tables, atm)
    radio = c8mi.RadioAccumulator(observers, prop, exact=exact)
    dev.attach_radio(radio)
    totals = []
    for shower_id, prim in showers:
        dev.begin_shower(shower_id, 2)
        dev.enqueue(prim)
        totals.append(dev.run())
    out = radio.download_fixed()
    dev.attach_radio(None)
    dev.close()
    radio.close()
    return out, totals


@pytest.mark.parametrize("pid,energy,seed,moliere", [(22, 1.0, 1, False), (11, 2.0, 2, True),
                                                     (22, 3.0, 3, True)])
def test_shower_radio_matches_oracle(tables, pid, energy, seed, moliere):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, moliere=moliere)
    prim = em.particle(pid, energy, START, DOWN, history_id=1)
    want = oracle_tracks(cfg, tables, atm, prim, shower_id=1)
    tracks = want["tracks"]
    assert tracks.shape[1] > 100
    observers = window_observers(tracks, ground_antennas(), 1.0e9, 512)
    prop = propagation()
    ref = oracle_lib.radio_project(prop, observers, tracks)
    got, totals = device_radio(cfg, tables, atm, observers, prop, [(1, prim)])
    assert totals[0]["lepton_steps"] == tracks.shape[1]
    assert ref["counters"]["valid_tracks"] > 0 and ref["counters"]["coreas_contributions"] > 0
    assert_same_radio(got, ref, observers)
    assert np.any(got[0] != 0) and np.any(got[1] != 0)


def test_consecutive_showers_accumulate(tables):
    """Two showers, one after the other, into one accumulator: the sum of both
    oracle projections (fixed point adds exactly). Many fronts alternate the
    two track buffers."""
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=7, atmosphere=atm, moliere=True)
    showers = [(11, em.particle(11, 2.0, START, DOWN, history_id=1)),
               (12, em.particle(-11, 1.5, START, DOWN, history_id=1))]
    tracks = np.concatenate([oracle_tracks(cfg, tables, atm, p, sid)["tracks"] for sid, p in showers],
                            axis=1)
    observers = window_observers(tracks, ground_antennas(), 1.0e9, 512)
    prop = propagation()
    ref = oracle_lib.radio_project(prop, observers, tracks)
    got, totals = device_radio(cfg, tables, atm, observers, prop, showers)
    assert sum(t["lepton_steps"] for t in totals) == tracks.shape[1]
    assert sum(t["lepton_fronts"] for t in totals) > 4
    assert_same_radio(got, ref, observers)


def test_pool_slots_share_the_accumulator(tables):
    """Two showers in two pool slots of the same fronts: the attached
    accumulator receives both showers' steps."""
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=9, atmosphere=atm, slots=2, moliere=True)
    cases = [(21, em.particle(22, 2.0, START, DOWN, history_id=1)),
             (22, em.particle(11, 1.0, START, DOWN, history_id=1))]
    tracks = np.concatenate([oracle_tracks(cfg, tables, atm, p, sid)["tracks"] for sid, p in cases],
                            axis=1)
    observers = window_observers(tracks, ground_antennas(), 1.0e9, 512)
    prop = propagation()
    ref = oracle_lib.radio_project(prop, observers, tracks)
    dev = em.EmBackend(cfg, tables, atm)
    radio = c8mi.RadioAccumulator(observers, prop, exact=True)
    dev.attach_radio(radio)
    for slot, (sid, prim) in enumerate(cases):
        dev.slot_begin(slot, sid, 2)
        dev.slot_enqueue(slot, prim)
    totals = dev.run_rounds(1 << 20)
    got = radio.download_fixed()
    dev.close()
    radio.close()
    assert totals["lepton_steps"] == tracks.shape[1]
    assert_same_radio(got, ref, observers)


def test_fast_projection_within_tolerance(tables):
    """The default (non-exact) projection in the coupled path meets the same bar."""
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=5, atmosphere=atm, moliere=True)
    prim = em.particle(11, 2.0, START, DOWN, history_id=1)
    tracks = oracle_tracks(cfg, tables, atm, prim, 1)["tracks"]
    observers = window_observers(tracks, ground_antennas(), 1.0e9, 512)
    prop = propagation()
    ref = oracle_lib.radio_project(prop, observers, tracks)
    got, _ = device_radio(cfg, tables, atm, observers, prop, [(1, prim)], exact=False)
    assert_same_radio(got, ref, observers)


def test_detached_backend_projects_nothing(tables):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=3, atmosphere=atm)
    prim = em.particle(11, 1.0, START, DOWN, history_id=1)
    tracks = oracle_tracks(cfg, tables, atm, prim, 1)["tracks"]
    observers = window_observers(tracks, ground_antennas(), 1.0e9, 256)
    prop = propagation()
    dev = em.EmBackend(cfg, tables, atm)
    radio = c8mi.RadioAccumulator(observers, prop, exact=True)
    dev.attach_radio(radio)
    dev.attach_radio(None)
    dev.begin_shower(1, 2)
    dev.enqueue(prim)
    dev.run()
    coreas, zhs, counters = radio.download_fixed()
    dev.close()
    radio.close()
    assert counters["valid_tracks"] == 0
    assert not coreas.any() and not zhs.any()


# ---- per-shower radio of the pool (every slot its own observers) ----

SLOT_CASES = [  # pid, energy (GeV), shower id, antenna ring radius (m), bins
    (22, 2.0, 31, 40.0, 384), (11, 1.5, 32, 90.0, 512), (-11, 1.0, 33, 25.0, 256), (22, 1.0, 34, 150.0, 320)]


def slot_case(tables, atm, cfg, case):
    pid, energy, shower, radius, nbins = case
    prim = em.particle(pid, energy, START, DOWN, history_id=1)
    tracks = oracle_tracks(cfg, tables, atm, prim, shower)["tracks"]
    antennas = [(radius * np.cos(a), radius * np.sin(a), 0.0) for a in (0.3, 2.0, 4.4)]
    observers = window_observers(tracks, antennas, 1.0e9, nbins)
    return prim, tracks, observers


@pytest.mark.parametrize("exact", [True, False])
def test_pool_slots_have_their_own_radio(tables, exact):
    """Four showers in four slots of the same fronts, each with its own
    antennas and windows: every slot's waveforms are its own shower's tracks
    projected on its own observers (the oracle shower alone), the other slots'
    steps leave no trace."""
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=19, atmosphere=atm, slots=4, moliere=True, moliere_table=True)
    prop = propagation()
    cases = [slot_case(tables, atm, cfg, c) for c in SLOT_CASES]
    dev = em.EmBackend(cfg, tables, atm)
    dev.enable_slot_radio(4, 512, prop, exact=exact)
    for slot, (case, (prim, _, observers)) in enumerate(zip(SLOT_CASES, cases)):
        dev.slot_begin(slot, case[2], 2)
        dev.slot_set_radio(slot, observers)
        dev.slot_enqueue(slot, prim)
    totals = dev.run_rounds(1 << 20)
    got = [dev.slot_download_radio(slot) for slot in range(4)]
    dev.close()
    assert totals["lepton_steps"] == sum(t.shape[1] for _, t, _ in cases)
    for (prim, tracks, observers), g in zip(cases, got):
        ref = oracle_lib.radio_project(prop, observers, tracks)
        assert ref["counters"]["valid_tracks"] > 0
        assert_same_radio(g, ref, observers)


def test_serving_loop_radio_alone_and_in_the_mix(tables):
    """The serving loop with per-shower radio: each request's waveforms (and a
    request without radio in between) equal the oracle's shower alone, whether
    the pool admits them together (budget) or one at a time."""
    atm = em.Atmosphere()
    prop = propagation()
    for slots, budget in ((3, 300.0), (1, 1e9)):
        cfg = em.make_config(airtables, seed=23, atmosphere=atm, slots=slots, moliere=True, moliere_table=True)
        cases = [slot_case(tables, atm, cfg, c) for c in SLOT_CASES]
        requests = []
        for k, (case, (prim, _, observers)) in enumerate(zip(SLOT_CASES, cases)):
            requests.append(dict(primary=prim, shower_id=case[2], arrival_s=0.0, footprint=60.0 * case[1],
                                 radio_observers=None if k == 2 else observers))
        dev = em.EmBackend(cfg, tables, atm)
        dev.enable_slot_radio(3, 512, prop)
        results, _ = dev.serve(requests, budget, realtime=False)
        dev.close()
        for k, ((prim, tracks, observers), res) in enumerate(zip(cases, results)):
            assert res["lepton_steps"] == tracks.shape[1]
            if k == 2:
                assert "radio" not in res and res["radio_counters"]["valid_tracks"] == 0
                continue
            ref = oracle_lib.radio_project(prop, observers, tracks)
            assert_same_radio((res["radio"][0], res["radio"][1], res["radio_counters"]), ref, observers)


def test_slot_radio_contract(tables):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=3, atmosphere=atm, slots=2)
    prop = propagation()
    dev = em.EmBackend(cfg, tables, atm)
    obs = window_observers(np.zeros((9, 1)), [(0.0, 0.0, 0.0)], 1.0e9, 64)
    with pytest.raises(c8mi.C8miError):  # not enabled yet
        dev.slot_set_radio(0, obs)
    dev.enable_slot_radio(2, 64, prop)
    with pytest.raises(c8mi.C8miError):  # more observers than a slot holds
        dev.slot_set_radio(0, np.concatenate([obs, obs, obs]))
    big = window_observers(np.zeros((9, 1)), [(0.0, 0.0, 0.0)], 1.0e9, 65)
    with pytest.raises(c8mi.C8miError):  # more bins than an observer holds
        dev.slot_set_radio(0, big)
    with pytest.raises(c8mi.C8miError):  # slot outside the pool
        dev.slot_set_radio(2, obs)
    dev.slot_set_radio(1, obs)
    c, z, counters = dev.slot_download_radio(1)
    assert c.shape == (3, 64) and not c.any() and not z.any() and counters["valid_tracks"] == 0
    dev.close()


MULTI_TILE_CASES = [  # pid, energy (GeV), shower id, antenna ring radius (m), bins
    (22, 2.0, 41, 40.0, 2500), (11, 1.5, 42, 90.0, 700), (-11, 1.0, 43, 25.0, 3000)]


def test_pool_slot_radio_windows_longer_than_one_tile(tables):
    """Per-shower radio with CoREAS and ZHS on windows of up to 3000 bins: each
    observer's window is split into several LDS tiles (one workgroup per
    chunk, observer and tile), with observers of mixed lengths. Every shower
    equals its oracle projection, in a pool of its own and in the mix."""
    atm = em.Atmosphere()
    prop = propagation()
    for slots in (1, 3):
        cfg = em.make_config(airtables, seed=29, atmosphere=atm, slots=slots, moliere=True, moliere_table=True)
        cases = [slot_case(tables, atm, cfg, c) for c in MULTI_TILE_CASES]
        dev = em.EmBackend(cfg, tables, atm)
        dev.enable_slot_radio(3, 3000, prop, exact=True)
        got = []
        if slots == 1:
            for case, (prim, _, observers) in zip(MULTI_TILE_CASES, cases):
                dev.slot_begin(0, case[2], 2)
                dev.slot_set_radio(0, observers)
                dev.slot_enqueue(0, prim)
                dev.run_rounds(1 << 20)
                got.append(dev.slot_download_radio(0))
                dev.slot_end(0)
        else:
            for slot, (case, (prim, _, observers)) in enumerate(zip(MULTI_TILE_CASES, cases)):
                dev.slot_begin(slot, case[2], 2)
                dev.slot_set_radio(slot, observers)
                dev.slot_enqueue(slot, prim)
            dev.run_rounds(1 << 20)
            got = [dev.slot_download_radio(slot) for slot in range(3)]
        dev.close()
        for (prim, tracks, observers), g in zip(cases, got):
            ref = oracle_lib.radio_project(prop, observers, tracks, threads=8)
            assert ref["counters"]["valid_tracks"] > 0
            assert_same_radio(g, ref, observers)


def test_refused_front_leaves_the_shower_unchanged(tables):
    """A lepton front that the attached accumulator would refuse (downloaded,
    not reset) is refused before any launch: pending counts, statistics and the
    ledger are unchanged, and after a reset the shower runs on to the oracle's
    result."""
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=13, atmosphere=atm, moliere=True)
    prim = em.particle(11, 2.0, START, DOWN, history_id=1)
    want = oracle_lib.em_run_shower(cfg, tables, atm, prim, shower_id=1, first_history_id=2)["stats"]
    observers = window_observers(np.zeros((9, 1)), ground_antennas(), 1.0e9, 64)
    dev = em.EmBackend(cfg, tables, atm)
    radio = c8mi.RadioAccumulator(observers, propagation(), exact=True)
    dev.attach_radio(radio)
    dev.begin_shower(1, 2)
    dev.enqueue(prim)
    first = dev.run(3)
    assert first["pending_leptons"] > 0
    radio.download_fixed()  # commits the accumulator
    before = (dev.pending(), dev.slot_state(0), dev.download_profile()[1])
    with pytest.raises(c8mi.C8miError) as err:
        dev.run_lepton_wavefront([], 1 << 40, (1 << 40) + (1 << 30))  # refused before its first front
    assert err.value.status == 2
    after = (dev.pending(), dev.slot_state(0), dev.download_profile()[1])
    assert after == before
    radio.reset()
    rest = dev.run()
    dev.attach_radio(None)
    dev.close()
    radio.close()
    for k in ("photon_steps", "lepton_steps", "photon_interactions", "lepton_interactions", "secondaries"):
        assert first[k] + rest[k] == want[k], k
    assert rest["next_history_id"] == want["next_history_id"]
