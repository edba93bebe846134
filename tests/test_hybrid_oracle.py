"""Host handoff on the CPU: the product's host stepper and prefix against the
oracle, and the oracle's restatement of the router pinned by what its
definition fixes.

* The CPU prefix of a shower (HybridCascade::run with EmRouter's scalar
  expansion, EmRouter.inl:260-293, up to the first device call): the
  product runs the device path's step functions compiled for the host
  (c8mi_em_hybrid_prefix); the oracle runs its own serial restatement of the
  step. The batch handed to the device, the history counter, every host
  e+- track, the host ledger and histograms, observations and fallbacks are
  equal, bit for bit: both sides compute in FP64 without contraction with the
  same libm, and the profile is int64 fixed point.
* The router with a minimum batch of one hands every staged particle to the
  device at once, so no host step happens and the router's calls are the
  device-resident schedule: the oracle's hybrid equals its resident run
  (oracle_em_run_shower with resident_minimum_batch 1) in every count and in
  the ledger.
* Each side closes its own energy ledger: the host books the primaries as
  injected and gives the device what its calls keep.
"""
import numpy as np
import pytest

import oracle_lib
from c8mi import airtables, em

START = (0.0, 0.0, 112750.0)
DOWN = (0.0, 0.0, -1.0)
FIELDS = ("pid", "medium_id", "generation", "energy_gev", "position_m", "direction", "time_s", "weight",
          "history_id", "parent_history_id", "step_id", "environment_node")


@pytest.fixture(scope="module")
def air():
    return airtables.build_air_tables()


def same_particles(a, b):
    assert a.size == b.size
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f


def scene(air, kind, seed):
    if kind == "ice":
        ice = em.IceVolume()
        tables = airtables.build_ice_tables(cut_mev=0.1)
        cfg = em.make_config(airtables.IceMedium, seed=seed, cut_mev=0.1, start=(0.0, 0.0, -100.0),
                             atmosphere=ice, moliere=True, moliere_table=True)
        return cfg, tables, ice, (0.0, 0.0, -100.0)
    field = (0.0, 5.0e-5, 0.0) if kind == "field" else (0.0, 0.0, 0.0)
    atm = em.Atmosphere(magnetic_field_t=field)
    moliere = kind in ("table", "thinned", "field", "analytic")
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, moliere=moliere,
                         moliere_table=True if kind in ("table", "thinned", "field") else None,
                         thinning=(0.5, 50.0) if kind == "thinned" else None)
    return cfg, air, atm, START


CASES = [("table", 22, 100.0, 1), ("table", 11, 10000.0, 2), ("none", -11, 5.0, 3), ("analytic", 22, 30.0, 4),
         ("thinned", 22, 300.0, 5), ("field", 11, 50.0, 6), ("ice", 11, 1000.0, 7)]


@pytest.mark.parametrize("kind,pid,energy,seed", CASES)
def test_host_prefix_equals_oracle(air, kind, pid, energy, seed):
    cfg, tables, env, start = scene(air, kind, seed)
    prim = em.particle(pid, energy, start, DOWN, history_id=1)
    stepper = em.ScalarStepper(cfg, tables, env)
    batch, next_history, tracks, result = stepper.prefix(prim, seed)
    hist, ledger = stepper.download_profile()
    obs, n_obs = stepper.observations()
    fbs, n_fb = stepper.fallbacks()
    first = stepper.first_interaction()
    stepper.close()
    want = oracle_lib.em_run_hybrid(cfg, tables, env, prim, shower_id=seed, prefix_only=True)
    same_particles(batch, want["staged"])
    assert next_history == want["result"]["next_history_id"]
    for k in em.HYBRID_COUNTS:
        assert result[k] == want["result"][k], k
    assert result["small_batch_expansions"] == 8 and result["scalar_expansion_budget_flushes"] == 1
    assert result["cpu_particles"] == result["particles_returned_for_expansion"] > 8
    host = want["host"]
    assert np.array_equal(tracks.T, host["tracks"])
    assert tracks.shape[0] == result["cpu_tracks"] > 0
    es, ws = cfg.struct.profile_energy_scale, cfg.struct.profile_weight_scale
    got = em.decode_counters(host["counters"], es)
    for k in em.LEDGER_ENERGY_TERMS:
        assert ledger[k] == got[k], k
    for k in ("steps", "photon_cuts", "lepton_limits", "deposited_steps", "fallback_reasons",
              "thinning_hillas_vertices", "thinning_statistical_vertices", "thinning_particles_discarded"):
        assert ledger[k] == got[k], k
    for k in ("photons", "electrons", "positrons"):
        assert np.array_equal(hist[k], host["hist"][k] / ws), k
    assert np.array_equal(hist["energy_loss"], host["hist"]["energy_loss"] / es)
    same_particles(obs, host["observations"])
    assert n_fb == host["fallbacks"].size
    same_particles(fbs["particle"], host["fallbacks"]["particle"])
    # the first final state of the primary happened in a host step
    want_first = host["first"]
    assert first is not None and want_first is not None
    assert first["process_id"] == want_first["process_id"]
    assert first["secondary_count"] == want_first["secondary_count"]
    same_particles(first["parent_at_vertex"].reshape(1), want_first["parent_at_vertex"].reshape(1))
    same_particles(first["secondaries"][:first["secondary_count"]],
                   want_first["secondaries"][:want_first["secondary_count"]])
    assert np.all(first["secondaries"]["weight"][:first["secondary_count"]] == prim["weight"])
    # the host side closes: what it kept of the primary is deposited or handed on
    if ledger["fallback_reasons"][21] == 0:
        assert abs(em.ledger_residual(ledger)) <= 1e-9 * energy


def test_host_steps_take_the_next_history_ids(air):
    cfg, tables, env, start = scene(air, "table", 11)
    stepper = em.ScalarStepper(cfg, tables, env)
    stepper.begin_shower(11)
    p = em.particle(22, 50.0, start, DOWN, history_id=1)
    counter, seen = 1000, 0
    for _ in range(200):
        out, children, track = stepper.step(p, counter)
        assert track is None  # photons leave no radio track
        if children:
            assert list(out["history_id"]) == list(range(counter, counter + children))
            assert np.all(out["parent_history_id"] == p["history_id"])
            seen += 1
            break
        assert out.size == 1 and out["history_id"][0] == p["history_id"]
        assert out["step_id"][0] == p["step_id"] + 1
        p = out[0]
    assert seen == 1
    stepper.close()


@pytest.mark.parametrize("pid,energy,seed", [(22, 30.0, 21), (11, 10.0, 22)])
def test_router_with_unit_batch_is_the_resident_schedule(air, pid, energy, seed):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, moliere=True, moliere_table=True,
                         resident_minimum_batch=1)
    prim = em.particle(pid, energy, START, DOWN, history_id=1)
    hybrid = oracle_lib.em_run_hybrid(cfg, air, atm, prim, shower_id=seed, hybrid=em.hybrid_config(minimum_batch=1))
    resident = oracle_lib.em_run_shower(cfg, air, atm, prim, shower_id=seed)
    r, w = hybrid["result"], resident["stats"]
    assert r["cpu_particles"] == 0 and r["small_batch_expansions"] == 0
    assert r["device_photon_steps"] == w["photon_steps"] and r["device_lepton_steps"] == w["lepton_steps"]
    assert r["device_photon_fronts"] == w["photon_fronts"] and r["device_lepton_fronts"] == w["lepton_fronts"]
    assert r["photon_calls"] == w["photon_calls"] and r["lepton_calls"] == w["lepton_calls"]
    assert r["device_secondaries"] == w["secondaries"] and r["next_history_id"] == w["next_history_id"]
    assert r["reserved_history_ids"] == w["reserved_history_ids"]
    assert r["bucketed_fronts"] == w["bucketed_fronts"]
    dev = hybrid["device"]["counters"]
    # every energy term but the injected one (the host books the primary, the
    # device what it keeps) is the resident run's
    assert np.array_equal(dev[1:], resident["counters"][1:])
    assert dev[0] + hybrid["host"]["counters"][0] == resident["counters"][0]
    for k in ("photons", "electrons", "positrons", "energy_loss"):
        assert np.array_equal(hybrid["device"]["hist"][k], resident["hist"][k]), k


@pytest.mark.parametrize("pid,energy,seed,minimum", [(22, 100.0, 31, 16384), (11, 30.0, 32, 64),
                                                     (-11, 20.0, 33, 8)])
def test_each_side_closes_its_ledger(air, pid, energy, seed, minimum):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm)
    prim = em.particle(pid, energy, START, DOWN, history_id=1)
    out = oracle_lib.em_run_hybrid(cfg, air, atm, prim, shower_id=seed,
                                   hybrid=em.hybrid_config(minimum_batch=minimum))
    r = out["result"]
    es = cfg.struct.profile_energy_scale
    host = em.decode_counters(out["host"]["counters"], es)
    dev = em.decode_counters(out["device"]["counters"], es)
    assert abs(host["injected_energy"] + dev["injected_energy"] - energy) <= 1e-9 * energy
    for side in (host, dev):
        assert side["pending_energy"] == 0.0
        assert abs(em.ledger_residual(side)) <= 1e-9 * energy
    assert r["small_batch_expansions"] >= 8 and r["cpu_particles"] > 0 and r["wavefronts"] > 0
    assert r["cpu_tracks"] == out["host"]["tracks"].shape[1]
    assert r["device_lepton_steps"] == out["device"]["tracks"].shape[1]


def test_fallback_round_trip_returns_moliere_failures(air):
    """Device Molière sampling failures go back to the host stack and take a
    host step without Molière scattering (EmRouter::returnFallbacks,
    EmRouter.inl:709-800, with the host's scalar transport): every such record
    is returned, the host steps fail nothing, and the particles' energy moves
    from the device's fallback term to the host's injected term. The shower
    then closes over both sides, except the continuous deposit of a failed
    step that ended at a vertex, which the reference's transport fallback
    does not book (MoliereStep.inl:50-54)."""
    atm = em.Atmosphere()
    returned = closed = 0
    for seed in range(1, 13):
        cfg = em.make_config(airtables, seed=seed, atmosphere=atm, moliere=True, moliere_table=True)
        prim = em.particle(22, 100.0, START, DOWN, history_id=1)
        out = oracle_lib.em_run_hybrid(cfg, air, atm, prim, shower_id=seed,
                                       hybrid=em.hybrid_config(minimum_batch=16384))
        r = out["result"]
        es = cfg.struct.profile_energy_scale
        host = em.decode_counters(out["host"]["counters"], es)
        dev = em.decode_counters(out["device"]["counters"], es)
        assert r["fallbacks_returned"] == dev["fallback_reasons"][21]
        assert r["fallbacks_kept"] == sum(dev["fallback_reasons"]) - dev["fallback_reasons"][21]
        assert host["fallback_reasons"][21] == 0
        # the returned particles' energy is the device's fallback term and the
        # host's injected term at once
        moved = dev["fallback_energy"] if r["fallbacks_kept"] == 0 else 0.0
        assert abs(host["injected_energy"] + dev["injected_energy"] - moved - 100.0) <= 1e-9 * 100.0
        residual = em.ledger_residual(host) + em.ledger_residual(dev)
        assert residual >= -1e-9 * 100.0 and residual <= 1e-2  # an unbooked continuous deposit is >= 0
        if r["fallbacks_returned"]:
            returned += 1
            closed += abs(residual) <= 1e-9 * 100.0
    assert returned >= 4 and closed >= returned // 2


def photoelectric_seeds(cfg_of, tables, atm, prim, count=2):
    """Seeds whose first final state (oracle) is photoelectric."""
    seeds = []
    for seed in range(1, 200):
        want, _ = oracle_lib.em_first_interaction(cfg_of(seed), tables, atm, prim, shower_id=seed)
        if want is not None and want["process_id"] == 1000000016:
            seeds.append(seed)
            if len(seeds) == count:
                break
    return seeds


def test_photoelectric_snapshot_carries_the_assigned_history_id():
    """PhotonFinalStateStep.inl:649-660 gives the photoelectric electron its
    history ID before the snapshot (every other process snapshots its children
    before their IDs): the host stepper's snapshot equals the oracle's."""
    tables = airtables.build_air_tables(cut_mev=0.05)
    atm = em.Atmosphere()
    cfg_of = lambda seed: em.make_config(airtables, seed=seed, atmosphere=atm, cut_mev=0.05)
    prim = em.particle(22, 6e-5, (0.0, 0.0, 20000.0), DOWN, history_id=1)
    seeds = photoelectric_seeds(cfg_of, tables, atm, prim)
    assert len(seeds) == 2
    for seed in seeds:
        cfg = cfg_of(seed)
        want, _ = oracle_lib.em_first_interaction(cfg, tables, atm, prim, shower_id=seed)
        stepper = em.ScalarStepper(cfg, tables, atm)
        stepper.begin_shower(seed)
        counter, stack = 40, [prim]
        while stack:
            out, children, _ = stepper.step(stack.pop(), counter)
            counter += children
            stack.extend(out)
        first = stepper.first_interaction()
        stepper.close()
        assert first["process_id"] == 1000000016 and first["secondary_count"] == 1
        child = first["secondaries"][0]
        assert child["history_id"] == 40 != first["parent_at_vertex"]["history_id"]
        # the oracle's shower gives the photon's first child the first ID of its call
        assert want["secondaries"][0]["history_id"] != want["parent_at_vertex"]["history_id"]
        same_particles(first["parent_at_vertex"].reshape(1), want["parent_at_vertex"].reshape(1))
