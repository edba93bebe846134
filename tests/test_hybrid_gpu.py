"""The host handoff on the GPU: showers that start on the host and reach the
device through the router (c8mi_em_hybrid_run: HybridCascade::run with
EmRouter, HybridCascade.inl:55-175, EmRouter.inl:64-423) against the oracle's
restatement of the same handoff (oracle_em_run_hybrid).

The host side (scalar expansion rounds on the host stepper, and the host steps
of device Molière failures returned by the fallback round trip) and the device
side (IEmBackend wavefront calls with the router's history ranges and minimum
batch; remaining particles staged again, expanded on the host when the device
has nothing pending) are compared separately, with the bar of the device
parity tests (test_em_gpu.py): identical counts, ledger terms to 1e-9 of the
primary energy, histograms, observation records as sets; with a radio
attached to both, the waveforms of the host and device tracks together.
"""
import numpy as np
import pytest

import c8mi
import oracle_lib
from c8mi import airtables, em
from radio_cases import window_observers
from test_em_radio_gpu import assert_same_radio, ground_antennas, propagation

pytestmark = pytest.mark.gpu

START = (0.0, 0.0, 112750.0)
DOWN = (0.0, 0.0, -1.0)


@pytest.fixture(scope="module")
def tables():
    return airtables.build_air_tables()


def compare_side(cfg, got_hist, got_ledger, want, energy):
    es, ws = cfg.struct.profile_energy_scale, cfg.struct.profile_weight_scale
    wl = em.decode_counters(want["counters"], es)
    for k in ("steps", "photon_cuts", "lepton_limits", "fallback_reasons",
              "thinning_hillas_vertices", "thinning_statistical_vertices", "fixed_point_overflows",
              "invalid_records"):
        assert got_ledger[k] == wl[k], (k, got_ledger[k], wl[k])
    # steps with a positive deposit: a continuous loss below one ulp of the
    # energy deposits nothing on one side and one ulp on the other (the device
    # rounds log/exp its own way); measured: one step in 1e5 at most
    assert abs(got_ledger["deposited_steps"] - wl["deposited_steps"]) <= 2, (got_ledger["deposited_steps"],
                                                                             wl["deposited_steps"])
    for k in em.LEDGER_ENERGY_TERMS:
        assert abs(got_ledger[k] - wl[k]) <= 1e-9 * energy, (k, got_ledger[k], wl[k])
    for k in ("photons", "electrons", "positrons"):
        assert np.max(np.abs(got_hist[k] - want["hist"][k] / ws)) <= 1e-6, k
    assert np.max(np.abs(got_hist["energy_loss"] - want["hist"]["energy_loss"] / es)) <= 1e-9 * energy


def same_records(got, want):
    key = lambda a: np.lexsort((a["step_id"], a["history_id"]))
    assert got.size == want.size
    got, want = got[key(got)], want[key(want)]
    assert np.array_equal(got["history_id"], want["history_id"])
    assert np.array_equal(got["pid"], want["pid"])
    np.testing.assert_allclose(got["energy_gev"], want["energy_gev"], rtol=1e-9, atol=0)


def hybrid_shower(cfg, tables, atm, prim, shower_id, hybrid_config, radio=None):
    dev = em.EmBackend(cfg, tables, atm)
    stepper = em.ScalarStepper(cfg, tables, atm)
    cascade = em.HybridCascade(dev, stepper, hybrid_config)
    if radio is not None:
        dev.attach_radio(radio)
        cascade.attach_radio(radio)
    result = cascade.run(prim, shower_id, 2)
    out = dict(result=result, host=stepper.download_profile(), device=dev.download_profile(),
               first=stepper.first_interaction(),
               host_obs=stepper.observations()[0], device_obs=dev.observations()[0],
               pending=dev.pending())
    if radio is not None:
        dev.attach_radio(None)
    cascade.close()
    stepper.close()
    dev.close()
    return out


@pytest.mark.parametrize("pid,energy,seed,minimum,multiplier", [
    (22, 100.0, 1, 16384, 0),  # the benchmark shower with the reference's handoff batch
    (11, 30.0, 2, 64, 0),      # calls of real fronts, checkpoints below 64, expansions in between
    (22, 10.0, 3, 1, 0),       # every staged particle straight to the device
    (-11, 20.0, 4, 32, 2),     # tight lepton ranges: history-range checkpoints
    (22, 100.0, 12, 16384, 0), # Molière failures at a vertex and along a step: host fallback steps
    (22, 100.0, 5, 16384, 0),
])
def test_hybrid_shower_equals_oracle(tables, pid, energy, seed, minimum, multiplier):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, moliere=True, moliere_table=True)
    prim = em.particle(pid, energy, START, DOWN, history_id=1)
    hc = em.hybrid_config(minimum_batch=minimum, multiplier=multiplier)
    print(f"hybrid shower: pid {pid}, {energy} GeV, seed {seed}, minimum batch {minimum}", flush=True)
    want = oracle_lib.em_run_hybrid(cfg, tables, atm, prim, shower_id=seed, hybrid=hc)
    got = hybrid_shower(cfg, tables, atm, prim, seed, hc)
    r, w = got["result"], want["result"]
    for k in em.HYBRID_COUNTS:
        assert r[k] == w[k], (k, r[k], w[k])
    assert got["pending"] == (0, 0)
    assert r["cpu_particles"] > 0 or minimum == 1
    assert r["device_photon_steps"] + r["device_lepton_steps"] > 0
    if minimum == 16384:
        assert r["small_batch_expansions"] >= 8 and r["scalar_expansion_budget_flushes"] >= 1
    if multiplier:
        assert r["history_range_checkpoints"] > 0
    if seed in (5, 12):
        assert r["fallbacks_returned"] > 0  # the round trip ran
    compare_side(cfg, *got["host"], want["host"], energy)
    compare_side(cfg, *got["device"], want["device"], energy)
    same_records(got["host_obs"], want["host"]["observations"])
    same_records(got["device_obs"], want["device"]["observations"])
    assert (got["first"] is None) == (want["host"]["first"] is None)
    if got["first"] is not None:
        assert got["first"]["process_id"] == want["host"]["first"]["process_id"]
    for side in ("host", "device"):
        led = got[side][1]
        if led["fallback_reasons"][21] == 0:
            assert abs(em.ledger_residual(led)) <= 1e-9 * energy, side
    # energy of fallbacks returned to the host: the device's fallback term and
    # the host's injected term at once
    moved = got["device"][1]["fallback_energy"] if r["fallbacks_kept"] == 0 else 0.0
    total = got["host"][1]["injected_energy"] + got["device"][1]["injected_energy"] - moved
    assert abs(total - energy) <= 1e-9 * energy
    # the shower report over both sides (ShowerReport's energy_ledger block):
    # host EM steps make the coverage incomplete by the reference's rule; the
    # closure with every sink booked holds unless a Molière step failed
    host_steps = r["cpu_photon_steps"] + r["cpu_lepton_steps"]
    report = em.energy_ledger_report([got["host"][1], got["device"][1]], energy, scalar_em_steps=host_steps)
    fallbacks = sum(got["host"][1]["fallback_reasons"]) + sum(got["device"][1]["fallback_reasons"])
    assert report["complete_coverage"] == (host_steps == 0 and fallbacks == 0)
    if report["complete_coverage"]:
        assert report["accepted"] and report["relative_closure_error"] <= 1e-9
    if got["host"][1]["fallback_reasons"][21] == 0 and got["device"][1]["fallback_reasons"][21] == 0:
        assert abs(report["extended_residual_gev"] + moved) <= 1e-9 * energy


def test_hybrid_radio_combines_host_and_device_tracks(tables):
    """The host e+- tracks of the prefix reach the radio as host tracks
    (RadioAccumulator::enqueueHostLeptonTracks), the device's through the
    attached accumulator: the waveforms are the oracle's projection of both."""
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=9, atmosphere=atm, moliere=True, moliere_table=True)
    prim = em.particle(11, 3.0, START, DOWN, history_id=1)
    hc = em.hybrid_config(minimum_batch=256)
    want = oracle_lib.em_run_hybrid(cfg, tables, atm, prim, shower_id=9, hybrid=hc)
    tracks = np.concatenate([want["host"]["tracks"], want["device"]["tracks"]], axis=1)
    assert want["host"]["tracks"].shape[1] > 10 and want["device"]["tracks"].shape[1] > 100
    observers = window_observers(tracks, ground_antennas(), 1.0e9, 512)
    prop = propagation()
    ref = oracle_lib.radio_project(prop, observers, tracks)
    radio = c8mi.RadioAccumulator(observers, prop, exact=True)
    got = hybrid_shower(cfg, tables, atm, prim, 9, hc, radio=radio)
    assert got["result"]["cpu_tracks"] == want["host"]["tracks"].shape[1]
    out = radio.download_fixed()
    radio.close()
    assert_same_radio(out, ref, observers)
