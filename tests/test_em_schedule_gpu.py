"""The reference's call schedule on the device, against the oracle.

The reference's router runs a shower as calls of the backend
(EmRouter.inl:339-416): a photon call of at most 16 fronts over the device
pending photons and the host photons, then a lepton call of at most 1024
fronts, each with a history range it reserves from the stack
(EmRouter.inl:640-679); particles left when a call ends come back to the host
and join the next call of their species after its device pending prefix
(WavefrontQueue.inl:207-260); every front of at least 256 particles is first
stably sorted by the 50-bit bucket key (WavefrontBucketing.inl:72-79,
393-466).

* A shower driven through the IEmBackend-shaped calls by a router written here
  (history reservations as EmRouter's, remaining particles returned and
  re-staged) equals the oracle's restatement of the same schedule, and the
  device-resident driver (c8mi_em_run) equals it too: identical step,
  interaction, secondary, call, bucketing and checkpoint counts, history
  counter, ledger and records (the bar of test_em_gpu.py).
* Showers whose fronts reach thousands of particles are bucketed on almost
  every front, so the parity covers the sort's order and the history IDs the
  children take in it.
* The below-minimum-batch checkpoint (PhotonCascade.inl:296-300) of the
  device-resident driver, and a pool mixing large and small showers.
"""
import numpy as np
import pytest

import c8mi
import oracle_lib
from c8mi import airtables, em
from test_em_gpu import STAT_KEYS, compare

pytestmark = pytest.mark.gpu

START = (0.0, 0.0, 112750.0)
DOWN = (0.0, 0.0, -1.0)


@pytest.fixture(scope="module")
def tables():
    return airtables.build_air_tables()


def photon_reservation(n):
    return n * 2 * 16  # reservePhotonSecondaryRange, EmRouter.inl:640-656


def lepton_reservation(n, multiplier=64):
    minimum = 3 * n  # reserveResidentSecondaryRange, EmRouter.inl:658-679
    return max(min(1 << 20, minimum * multiplier), minimum)


def router_run(dev, primaries, first_history_id=2, multiplier=64):
    """EmRouter::advanceOneWavefrontAndReturn with the device pending queues,
    every staged particle sent to the device: returns summed call results."""
    counter = first_history_id
    staged = np.ascontiguousarray(primaries, dtype=c8mi.PARTICLE_STATE_DTYPE).reshape(-1)
    totals = dict(photon_steps=0, lepton_steps=0, photon_interactions=0, lepton_interactions=0,
                  secondaries=0, photon_fronts=0, lepton_fronts=0, photon_calls=0, lepton_calls=0,
                  bucketed_fronts=0, below=0, exhausted=0, reserved=0)
    while True:
        pending_photons, pending_leptons = dev.pending()
        photons = staged[staged["pid"] == 22]
        leptons = staged[staged["pid"] != 22]
        if not (photons.size or leptons.size or pending_photons or pending_leptons):
            break
        returned = []
        if photons.size or pending_photons:
            n = photons.size + pending_photons
            first, count = counter, photon_reservation(n)
            counter += count
            r = dev.run_photon_wavefront(photons, first, 16, 1)
            assert r["input_particles"] == n
            assert r["wavefronts"] <= 16 and r["next_secondary_history_id"] <= first + count
            totals["photon_steps"] += r["transport_records"]
            totals["photon_interactions"] += r["interactions"]
            totals["photon_fronts"] += r["wavefronts"]
            totals["photon_calls"] += 1
            totals["reserved"] += count
            totals["secondaries"] += r["secondaries"]
            totals["bucketed_fronts"] += r["bucketed_fronts"]
            returned.append(dev.remaining())
        pending_leptons = dev.pending()[1]
        if leptons.size or pending_leptons:
            n = leptons.size + pending_leptons
            first, count = counter, lepton_reservation(n, multiplier)
            counter += count
            r = dev.run_lepton_wavefront(leptons, first, first + count, 1024, 1)
            assert r["input_particles"] == n
            assert r["next_secondary_history_id"] <= first + count
            totals["lepton_steps"] += r["transport_records"]
            totals["lepton_interactions"] += r["interactions"]
            totals["lepton_fronts"] += r["wavefronts"]
            totals["lepton_calls"] += 1
            totals["reserved"] += count
            totals["secondaries"] += r["secondaries"]
            totals["bucketed_fronts"] += r["bucketed_fronts"]
            totals["exhausted"] += r["history_range_exhausted"]
            returned.append(dev.remaining())
        staged = np.concatenate(returned) if returned else staged[:0]
    totals["next_history_id"] = counter
    return totals


def resident_pair(tables, pid, energy, seed, moliere_table=True, multiplier=64, **config):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, moliere=moliere_table is not None,
                         moliere_table=moliere_table, **config)
    prim = em.particle(pid, energy, START, DOWN, history_id=1)
    with oracle_lib.lepton_reservation(multiplier):
        want = oracle_lib.em_run_shower(cfg, tables, atm, prim, shower_id=seed, first_history_id=2)
    return atm, cfg, prim, want


def device_shower(cfg, tables, atm, prim, shower_id, driver, multiplier=64):
    dev = em.EmBackend(cfg, tables, atm)
    dev.begin_shower(shower_id, 2)
    if driver == "resident":
        dev.enqueue(prim)
        stats = dev.run()
    else:
        stats = router_run(dev, prim, multiplier=multiplier)
    hist, ledger = dev.download_profile()
    obs, _ = dev.observations()
    fbs, _ = dev.fallbacks()
    statistics = dev.statistics()
    dev.close()
    return stats, dict(stats=stats, hist=hist, ledger=ledger, observations=obs, fallbacks=fbs), statistics


@pytest.mark.parametrize("pid,energy,seed,multiplier", [(22, 100.0, 31, 64), (11, 30.0, 32, 64),
                                                        (22, 3.0, 33, 64), (22, 30.0, 34, 1), (11, 10.0, 35, 2)])
def test_router_driven_calls_equal_the_oracle_schedule(tables, pid, energy, seed, multiplier):
    """multiplier < 64 reserves tighter lepton ranges (router and oracle alike)
    so that lepton calls end by history-range exhaustion."""
    atm, cfg, prim, want = resident_pair(tables, pid, energy, seed, multiplier=multiplier)
    print(f"router-driven shower: pid {pid}, {energy} GeV, seed {seed}, multiplier {multiplier}", flush=True)
    stats, got, statistics = device_shower(cfg, tables, atm, prim, seed, "router", multiplier)
    w = want["stats"]
    for k in ("photon_steps", "lepton_steps", "photon_interactions", "lepton_interactions", "secondaries",
              "photon_fronts", "lepton_fronts", "photon_calls", "lepton_calls", "bucketed_fronts",
              "next_history_id"):
        assert stats[k] == w[k], (k, stats[k], w[k])
    assert stats["reserved"] == w["reserved_history_ids"]
    assert stats["exhausted"] == w["history_range_checkpoints"]
    got["stats"] = dict(w)  # the integer statistics were compared above
    compare(cfg, want, got, energy)
    assert statistics["photon_calls"] == 0  # the router reserved every range itself
    if energy >= 30.0:
        assert stats["bucketed_fronts"] > 0
    if multiplier < 64:
        assert stats["exhausted"] > 0


@pytest.mark.parametrize("pid,energy,seed,moliere_table", [(22, 100.0, 41, True), (-11, 50.0, 42, None)])
def test_resident_schedule_with_bucketed_fronts_matches_oracle(tables, pid, energy, seed, moliere_table):
    atm, cfg, prim, want = resident_pair(tables, pid, energy, seed, moliere_table=moliere_table)
    print(f"resident shower: pid {pid}, {energy} GeV, seed {seed}", flush=True)
    stats, got, statistics = device_shower(cfg, tables, atm, prim, seed, "resident")
    compare(cfg, want, got, energy)
    assert stats["bucketed_fronts"] > 0 and stats["max_front"] >= 256
    assert statistics["bucketed_fronts"] == stats["bucketed_fronts"]
    assert statistics["reserved_history_ids"] == stats["reserved_history_ids"]


def test_resident_minimum_batch_checkpoint_matches_oracle(tables):
    atm, cfg, prim, want = resident_pair(tables, 22, 30.0, 51, resident_minimum_batch=64)
    assert want["stats"]["below_minimum_batch_checkpoints"] > 0
    stats, got, _ = device_shower(cfg, tables, atm, prim, 51, "resident")
    compare(cfg, want, got, 30.0)


def test_pool_of_large_and_small_showers_equals_solo_runs(tables):
    """Bucketed and small showers in the same fronts, each at its own point of
    its call schedule, parked while its other species' call runs: every shower
    equals its solo oracle run."""
    atm = em.Atmosphere()
    cases = [(22, 60.0, 61), (11, 2.0, 62), (22, 40.0, 63), (-11, 0.8, 64)]
    cfg = em.make_config(airtables, seed=7, atmosphere=atm, slots=6, moliere=True, moliere_table=True)
    dev = em.EmBackend(cfg, tables, atm)
    for slot, (pid, energy, shower) in enumerate(cases):
        dev.slot_begin(slot + 1, shower, 2)
        dev.slot_enqueue(slot + 1, em.particle(pid, energy, START, DOWN, history_id=1))
    totals = dev.run_rounds()
    assert totals["bucketed_fronts"] > 0
    stats = dev.statistics()
    assert stats["parked_fronts"] > 0 and stats["parked_particles"] > 0, stats
    for slot, (pid, energy, shower) in enumerate(cases):
        prim = em.particle(pid, energy, START, DOWN, history_id=1)
        want = oracle_lib.em_run_shower(cfg, tables, atm, prim, shower_id=shower, first_history_id=2)
        st = dev.slot_state(slot + 1)
        for k in ("photon_fronts", "lepton_fronts", "photon_steps", "lepton_steps", "photon_interactions",
                  "lepton_interactions", "secondaries", "observations", "fallbacks", "next_history_id"):
            assert st[k] == want["stats"][k], (shower, k, st[k], want["stats"][k])
        hist, ledger = dev.slot_download_profile(slot + 1)
        obs, _ = dev.slot_observations(slot + 1)
        fbs, _ = dev.slot_fallbacks(slot + 1)
        compare(cfg, want, dict(stats=dict(want["stats"]), hist=hist, ledger=ledger, observations=obs,
                                fallbacks=fbs), energy)
    dev.close()
