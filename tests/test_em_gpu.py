"""GPU parity of the EM transport fronts (c8mi_em.h) against the oracle's
serial restatement (oracle/em_oracle.c) on the same tables, atmosphere,
primaries and Philox keys.

Both run the reference's call schedule (per shower a photon call of at most
16 fronts, then a lepton call of at most 1024, each with its own reserved
history range; fronts of at least 256 particles stably sorted by the bucket
key; call input ordered pending, remaining, host) with the stable source
order of the reference's scans, so history IDs, every Philox draw and
therefore whole showers coincide. The device evaluates log/exp/pow
with the GPU math library, whose results may differ from the host's in the
last ulp; the bar is:
  * identical integer statistics (fronts, steps, interactions, secondaries,
    history IDs, observations, fallbacks, ledger counts);
  * ledger energies and profile bins within 1e-9 of the primary energy;
  * observation-surface particles identical up to 1e-9 relative.
At 100 GeV (the benchmark shower) the checks are size-independent: energy
closure of the device ledger to 1e-9 of E0 and agreement of the shower totals
with the oracle.
"""
import numpy as np
import pytest

import c8mi
import oracle_lib
from c8mi import airtables, em

pytestmark = pytest.mark.gpu

STAT_KEYS = ("photon_fronts", "lepton_fronts", "photon_steps", "lepton_steps", "photon_interactions",
             "lepton_interactions", "secondaries", "observations", "fallbacks", "pending_photons",
             "pending_leptons", "next_history_id", "max_front", "photon_calls", "lepton_calls",
             "bucketed_fronts", "bucketed_particles", "below_minimum_batch_checkpoints",
             "history_range_checkpoints", "reserved_history_ids")


@pytest.fixture(scope="module")
def tables():
    return airtables.build_air_tables()


def run_pair(tables, pid, energy, seed, shower_id=1, thinning=None, start=(0.0, 0.0, 112750.0),
             direction=(0.0, 0.0, -1.0), primaries=None, moliere=False, moliere_table=None, fast_iterations=0,
             field=(0.0, 0.0, 0.0)):
    atm = em.Atmosphere(magnetic_field_t=field)
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, thinning=thinning, start=start,
                         direction=direction, moliere=moliere, moliere_table=moliere_table,
                         moliere_fast_iterations=fast_iterations)
    prim = primaries if primaries is not None else em.particle(pid, energy, start, direction, history_id=1)
    want = oracle_lib.em_run_shower(cfg, tables, atm, prim, shower_id=shower_id, first_history_id=2)
    dev = em.EmBackend(cfg, tables, atm)
    dev.begin_shower(shower_id, 2)
    dev.enqueue(prim)
    got = dev.run()
    hist, ledger = dev.download_profile()
    obs, nobs = dev.observations()
    fbs, nfb = dev.fallbacks()
    dev.close()
    return cfg, want, dict(stats=got, hist=hist, ledger=ledger, observations=obs, fallbacks=fbs)


def closure_bound(cfg, ledger, energy):
    """1e-9 of the primary energy; with thinning also the reference's mass
    convention bookkeeping: a thinned two-child vertex charges the transport /
    native electron mass difference (5.109989461e-4 - 5.109989e-4 GeV) for two
    children at the parent's weight in the thinning term but corrects it with
    the kept children's weights only."""
    th = cfg.struct.thinning
    vertices = ledger["thinning_hillas_vertices"] + ledger["thinning_statistical_vertices"]
    return 1e-9 * energy + (2 * 4.61e-11 * th.maximum_weight * vertices if th.enabled else 0.0)


def compare(cfg, want, got, energy):
    for k in STAT_KEYS:
        assert got["stats"][k] == want["stats"][k], (k, got["stats"][k], want["stats"][k])
    wl = em.decode_counters(want["counters"], cfg.struct.profile_energy_scale)
    gl = got["ledger"]
    for k in ("steps", "photon_cuts", "lepton_limits", "thinning_hillas_vertices",
              "thinning_statistical_vertices", "thinning_particles_discarded", "deposited_steps",
              "fixed_point_overflows", "invalid_records", "fallback_reasons"):
        assert gl[k] == wl[k], (k, gl[k], wl[k])
    for k in em.LEDGER_ENERGY_TERMS:
        assert abs(gl[k] - wl[k]) <= 1e-9 * energy, (k, gl[k], wl[k])
    ws, es = cfg.struct.profile_weight_scale, cfg.struct.profile_energy_scale
    for k in ("photons", "electrons", "positrons"):
        assert np.max(np.abs(got["hist"][k] - want["hist"][k] / ws)) <= 1e-6, k
    assert np.max(np.abs(got["hist"]["energy_loss"] - want["hist"]["energy_loss"] / es)) <= 1e-9 * energy
    # observation-surface particles, as sets
    key = lambda a: np.lexsort((a["step_id"], a["history_id"]))
    go, wo = got["observations"], want["observations"]
    assert go.size == wo.size
    go, wo = go[key(go)], wo[key(wo)]
    assert np.array_equal(go["history_id"], wo["history_id"])
    assert np.array_equal(go["pid"], wo["pid"])
    for f in ("energy_gev", "time_s"):
        np.testing.assert_allclose(go[f], wo[f], rtol=1e-9, atol=0)
    np.testing.assert_allclose(go["position_m"], wo["position_m"], rtol=1e-9, atol=1e-6)
    # A Molière-failed step's transport fallback carries the vertex state, so its
    # continuous deposit is not booked (MoliereStep.inl:60-64): no closure then,
    # but the device books exactly what the oracle does.
    if gl["fallback_reasons"][21] == 0:
        assert abs(em.ledger_residual(gl)) <= closure_bound(cfg, gl, energy)
    assert abs(em.ledger_residual(gl) - em.ledger_residual(wl)) <= 1e-9 * energy


@pytest.mark.parametrize("pid,energy,seed", [(22, 1.0, 1), (22, 1.0, 2), (22, 10.0, 3), (11, 5.0, 4),
                                             (-11, 3.0, 5)])
def test_shower_matches_oracle(tables, pid, energy, seed):
    cfg, want, got = run_pair(tables, pid, energy, seed)
    compare(cfg, want, got, energy)
    assert got["stats"]["photon_steps"] + got["stats"]["lepton_steps"] > 100


def test_thinned_shower_matches_oracle(tables):
    cfg, want, got = run_pair(tables, 22, 10.0, 6, thinning=(1.0, 100.0))
    compare(cfg, want, got, 10.0)
    assert got["ledger"]["thinning_hillas_vertices"] + got["ledger"]["thinning_statistical_vertices"] > 0


def test_inclined_shower_reaches_ground_and_matches(tables):
    # a 60 degree shower started low enough that particles reach the observation level
    d = (np.sin(np.radians(60.0)), 0.0, -np.cos(np.radians(60.0)))
    start = (-2000.0 * d[0] / d[2] * 0, 0.0, 3000.0)
    cfg, want, got = run_pair(tables, 22, 20.0, 7, start=start, direction=d)
    compare(cfg, want, got, 20.0)
    assert got["stats"]["observations"] > 0


@pytest.mark.parametrize("seed", [2, 5, 8])
def test_muon_pair_shower_matches_oracle(seed):
    """Photon showers with the mu-pair columns (at 1e4 times their rate): the
    device's mu-pair final states (PhotonFinalStateStep.inl:326-382, 679-731)
    put both muons on the lepton queue; each is cut at its first step (native
    cut, its signed energy in the ledger) or handed to the host as a fallback
    (no muon tables here), as in the oracle."""
    mu_tables = airtables.build_air_tables(muon_pair=1.0e4)
    cfg, want, got = run_pair(mu_tables, 22, 100.0, seed, moliere=True, moliere_table=True)
    compare(cfg, want, got, 100.0)
    fb = got["fallbacks"]
    muons = fb[np.abs(fb["particle"]["pid"]) == 13]
    assert muons.size >= 2 and np.all(muons["reason"] == 3)
    wf = want["fallbacks"]
    wm = wf[np.abs(wf["particle"]["pid"]) == 13]
    key = lambda a: np.lexsort((a["particle"]["pid"], a["particle"]["history_id"]))
    muons, wm = muons[key(muons)], wm[key(wm)]
    assert np.array_equal(muons["particle"]["history_id"], wm["particle"]["history_id"])
    np.testing.assert_allclose(muons["particle"]["energy_gev"], wm["particle"]["energy_gev"], rtol=1e-9)


@pytest.mark.parametrize("pid,energy,seed,start,fast", [(11, 5.0, 13, 1000.0, 0), (22, 10.0, 14, 112750.0, 0),
                                                        (-11, 3.0, 15, 4000.0, 0), (11, 5.0, 13, 1000.0, 1),
                                                        (22, 30.0, 17, 112750.0, 2)])
def test_moliere_shower_matches_oracle(tables, pid, energy, seed, start, fast):
    """Molière multiple scattering on every finite e+- step (the reference's
    production setting): device and oracle agree shower for shower. With one
    or two Newton iterations in the step kernel nearly every sample continues
    on its own wavefront (the tail kernel), with the same results."""
    cfg, want, got = run_pair(tables, pid, energy, seed, start=(0.0, 0.0, start), moliere=True,
                              fast_iterations=fast)
    compare(cfg, want, got, energy)
    assert got["stats"]["lepton_steps"] > 100


@pytest.mark.parametrize("pid,energy,seed,start,fast", [(11, 5.0, 13, 1000.0, 0), (22, 10.0, 14, 112750.0, 0),
                                                        (-11, 3.0, 16, 4000.0, 0), (22, 30.0, 17, 112750.0, 0),
                                                        (22, 30.0, 17, 112750.0, 1), (11, 5.0, 13, 1000.0, 2)])
def test_moliere_table_shower_matches_oracle(tables, pid, energy, seed, start, fast):
    """Molière scattering through the interpolation table and the tabulated
    Newton initial guess (the reference backend's production path); `fast` 1
    or 2 sends nearly every sample through the tail kernel."""
    cfg, want, got = run_pair(tables, pid, energy, seed, start=(0.0, 0.0, start), moliere=True,
                              moliere_table=True, fast_iterations=fast)
    compare(cfg, want, got, energy)
    assert got["stats"]["lepton_steps"] > 100


@pytest.mark.parametrize("pid,energy,seed,start,field,moliere,magnetic_limits", [
    (22, 10.0, 21, 3000.0, (0.0, 5.0e-5, 0.0), False, False),        # Earth-like field, photon primary
    (11, 5.0, 22, 1000.0, (2.0e-5, -4.0e-5, 3.0e-5), False, False),  # oblique field, electron
    (-11, 3.0, 23, 2000.0, (0.0, 5.0e-2, 0.0), True, True),          # strong field: magnetic limits, Molière
    (22, 100.0, 24, 112750.0, (0.0, 5.0e-5, -2.0e-5), True, True)])  # whole air shower from the top
def test_magnetic_field_shower_matches_oracle(tables, pid, energy, seed, start, field, moliere, magnetic_limits):
    """e+- in a uniform field: leapfrog step limits, curved layer boundaries
    (quartic roots in double-double) and the bent advance on the device equal
    the oracle's, shower for shower."""
    cfg, want, got = run_pair(tables, pid, energy, seed, start=(0.0, 0.0, start), moliere=moliere,
                              moliere_table=True if moliere else None, field=field)
    compare(cfg, want, got, energy)
    if magnetic_limits:
        assert got["ledger"]["lepton_limits"][6] > 0  # steps ended by the magnetic limit


@pytest.fixture(scope="module")
def ice_tables():
    return airtables.build_ice_tables(cut_mev=0.1)


@pytest.mark.parametrize("pid,energy,seed,thinning,moliere", [
    (11, 1.0, 41, None, False),
    (22, 10.0, 42, None, True),
    (11, 100.0, 43, (1.0e-2, 1.0e4), True)])  # thinning below 1e-4 E0
def test_ice_shower_matches_oracle(ice_tables, pid, energy, seed, thinning, moliere):
    """The in-ice scene: homogeneous ice (H2O tables, 0.92 g/cm^3), 100 keV
    cut, e+- process tables from 0.55 MeV, Molière with the ice snapshot's
    table, thinning; device and oracle agree shower for shower."""
    ice = em.IceVolume()
    start = (0.0, 0.0, -100.0)
    cfg = em.make_config(airtables.IceMedium, seed=seed, cut_mev=0.1, start=start, atmosphere=ice,
                         thinning=thinning, moliere=moliere, moliere_table=True if moliere else None, bins=256)
    prim = em.particle(pid, energy, start, (0.0, 0.0, -1.0), history_id=1)
    want = oracle_lib.em_run_shower(cfg, ice_tables, ice, prim, shower_id=1, first_history_id=2)
    dev = em.EmBackend(cfg, ice_tables, ice)
    dev.begin_shower(1, 2)
    dev.enqueue(prim)
    got = dev.run()
    hist, ledger = dev.download_profile()
    obs, _ = dev.observations()
    fbs, _ = dev.fallbacks()
    dev.close()
    compare(cfg, want, dict(stats=got, hist=hist, ledger=ledger, observations=obs, fallbacks=fbs), energy)
    assert got["lepton_steps"] > 1000


def test_mixed_primary_batch(tables):
    prim = np.concatenate([np.atleast_1d(em.particle(22, 2.0, (0, 0, 30000.0), (0, 0, -1), history_id=1)),
                           np.atleast_1d(em.particle(11, 1.5, (10, 0, 30000.0), (0, 0, -1), history_id=2)),
                           np.atleast_1d(em.particle(-11, 0.8, (0, 10, 30000.0), (0, 0, -1), history_id=3))])
    cfg, want, got = run_pair(tables, 0, 0, 8, primaries=prim, start=(0.0, 0.0, 30000.0))
    compare(cfg, want, got, 4.3)


def test_benchmark_shower_closes_energy(tables):
    """100 GeV vertical photon from 112.75 km: device ledger closes to 1e-9 E0
    and the totals agree with the oracle's shower."""
    cfg, want, got = run_pair(tables, 22, 100.0, 1)
    gl = got["ledger"]
    assert abs(em.ledger_residual(gl)) <= 1e-9 * 100.0
    assert gl["invalid_records"] == 0 and gl["fixed_point_overflows"] == 0
    total = got["stats"]["photon_steps"] + got["stats"]["lepton_steps"]
    want_total = want["stats"]["photon_steps"] + want["stats"]["lepton_steps"]
    assert abs(total - want_total) <= 1e-3 * want_total
    wl = em.decode_counters(want["counters"], cfg.struct.profile_energy_scale)
    for k in ("continuous_deposit", "cut_deposit", "observed_energy"):
        assert abs(gl[k] - wl[k]) <= 1e-3 * 100.0, k


def test_backend_contract(tables):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=3, atmosphere=atm)
    dev = em.EmBackend(cfg, tables, atm)
    dev.begin_shower(1, 2)
    assert dev.can_transport(em.particle(22, 1.0, (0, 0, 1e4), (0, 0, -1)))
    assert not dev.can_transport(em.particle(13, 1.0, (0, 0, 1e4), (0, 0, -1)))
    with pytest.raises(c8mi.C8miError) as err:
        dev.enqueue(em.particle(2212, 1.0, (0, 0, 1e4), (0, 0, -1)))
    assert err.value.status == 1
    photon = em.particle(22, 1.0, (0, 0, 1e4), (0, 0, -1), history_id=1)
    # the reference's argument checks (PhotonCascade.inl:211-214, LeptonCascade.inl:127-132)
    for first, fronts, batch in ((0, 16, 1), (2, 0, 1), (2, 16, 0)):
        with pytest.raises(c8mi.C8miError) as err:
            dev.run_photon_wavefront(photon, first, fronts, batch)
        assert err.value.status == 1
    with pytest.raises(c8mi.C8miError) as err:
        dev.run_lepton_wavefront(em.particle(11, 1.0, (0, 0, 1e4), (0, 0, -1)), 10, 10)
    assert err.value.status == 1
    with pytest.raises(c8mi.C8miError) as err:  # species checked at the boundary
        dev.run_photon_wavefront(em.particle(11, 1.0, (0, 0, 1e4), (0, 0, -1)), 2)
    assert err.value.status == 1
    # an empty call completes without checking its limits (PhotonCascade.inl:206-209)
    empty = dev.run_lepton_wavefront(np.zeros(0, dtype=c8mi.PARTICLE_STATE_DTYPE), 0, 0, 0, 0)
    assert empty["completed"] == 1 and empty["wavefronts"] == 0
    r = dev.run_photon_wavefront(photon, 100, 1)
    assert r["wavefronts"] == 1 and r["input_particles"] == 1 and r["transport_records"] == 1
    assert r["next_secondary_history_id"] == 100 + r["secondaries"]
    photons, leptons = dev.pending()
    # the photon left after the one front came back to the host
    assert photons == 0 and r["remaining"] == len(dev.remaining()) == 1 - r["interactions"]
    assert leptons == r["other_species_pending"]
    assert r["completed"] == (r["remaining"] == 0)
    lim = dev.limits()
    assert lim["minimum_batch_size"] == 16384 and lim["maximum_resident_input_batch"] == cfg.struct.queue_capacity // 2
    st = dev.statistics()
    assert st["photon_wavefronts"] == 1 and st["particles_advanced"] == 1 and st["shower_ordinal"] == 1
    # the wavefront calls act on one shower: refused while other pool slots are active
    dev.close()
    pool_cfg = em.make_config(airtables, seed=3, atmosphere=atm, slots=2)
    pool = em.EmBackend(pool_cfg, tables, atm)
    pool.slot_begin(0, 1)
    pool.slot_begin(1, 2)
    with pytest.raises(c8mi.C8miError) as err:
        pool.run_photon_wavefront(photon, 2)
    assert err.value.status == 2
    pool.close()


# ---- shower pool: several showers in the same fronts ----

POOL_CASES = [  # pid, energy, shower id, zenith (deg)
    (22, 1.0, 11, 0.0), (11, 2.0, 12, 0.0), (-11, 1.5, 13, 0.0), (22, 3.0, 14, 30.0), (22, 2.5, 15, 0.0)]


def pool_case(case):
    pid, energy, shower, zenith = case
    d = (np.sin(np.radians(zenith)), 0.0, -np.cos(np.radians(zenith)))
    start = (0.0, 0.0, 30000.0)
    return pid, energy, shower, start, d


def solo_oracle(tables, atm, case, seed=21, moliere_table=None):
    pid, energy, shower, start, d = pool_case(case)
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, start=start, direction=d,
                         moliere=moliere_table is not None, moliere_table=moliere_table)
    prim = em.particle(pid, energy, start, d, history_id=1)
    return cfg, oracle_lib.em_run_shower(cfg, tables, atm, prim, shower_id=shower, first_history_id=2)


def check_slot_against_oracle(dev, slot, cfg, want, energy):
    st = dev.slot_state(slot)
    for k in ("photon_fronts", "lepton_fronts", "photon_steps", "lepton_steps", "photon_interactions",
              "lepton_interactions", "secondaries", "observations", "fallbacks", "next_history_id"):
        assert st[k] == want["stats"][k], (slot, k, st[k], want["stats"][k])
    assert st["pending_photons"] == 0 and st["pending_leptons"] == 0
    hist, ledger = dev.slot_download_profile(slot)
    obs, _ = dev.slot_observations(slot)
    fbs, _ = dev.slot_fallbacks(slot)
    got = dict(stats=dict(want["stats"]), hist=hist, ledger=ledger, observations=obs, fallbacks=fbs)
    compare(cfg, want, got, energy)


def start_slot(dev, atm, slot, case):
    pid, energy, shower, start, d = pool_case(case)
    axis = em.shower_axis(atm, start, d)
    dev.slot_begin(slot, shower, 2, axis)
    dev.slot_enqueue(slot, em.particle(pid, energy, start, d, history_id=1))
    return axis


def test_pool_showers_match_their_solo_runs(tables):
    """Five showers (photons, e-, e+, one inclined) share the device fronts;
    each one's statistics, profile, ledger and records equal the oracle's run
    of that shower alone."""
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=21, atmosphere=atm, slots=8, observation_capacity=1 << 14,
                         fallback_capacity=1 << 10)
    dev = em.EmBackend(cfg, tables, atm)
    slots = [0, 2, 3, 5, 7]
    axes = [start_slot(dev, atm, s, c) for s, c in zip(slots, POOL_CASES)]
    r = dev.run_rounds()
    assert r["pending_photons"] == 0 and r["pending_leptons"] == 0
    for s, case in zip(slots, POOL_CASES):
        ocfg, want = solo_oracle(tables, atm, case)
        check_slot_against_oracle(dev, s, ocfg, want, case[1])
    assert len(axes) == 5
    dev.close()


def test_pool_admits_showers_while_others_run(tables):
    """Continuous batching: two showers start, later ones join free slots and a
    slot is reused after its shower ends; every result equals the solo run."""
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=21, atmosphere=atm, slots=3, observation_capacity=1 << 14,
                         fallback_capacity=1 << 10)
    dev = em.EmBackend(cfg, tables, atm)
    keep = [start_slot(dev, atm, 0, POOL_CASES[0]), start_slot(dev, atm, 1, POOL_CASES[3])]
    dev.run_rounds(3)
    keep.append(start_slot(dev, atm, 2, POOL_CASES[1]))
    done = {}
    waiting = [POOL_CASES[2], POOL_CASES[4]]
    owner = {0: POOL_CASES[0], 1: POOL_CASES[3], 2: POOL_CASES[1]}
    rounds = 0
    while owner:
        dev.run_rounds(2)
        rounds += 1
        for slot in list(owner):
            st = dev.slot_state(slot)
            if st["pending_photons"] or st["pending_leptons"]:
                continue
            case = owner.pop(slot)
            ocfg, want = solo_oracle(tables, atm, case)
            check_slot_against_oracle(dev, slot, ocfg, want, case[1])
            done[case[2]] = True
            dev.slot_end(slot)
            if waiting:
                nxt = waiting.pop(0)
                keep.append(start_slot(dev, atm, slot, nxt))
                owner[slot] = nxt
        assert rounds < 10000
    assert sorted(done) == [11, 12, 13, 14, 15]
    dev.close()


def test_pool_slot_contract(tables):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=21, atmosphere=atm, slots=2, observation_capacity=1 << 10,
                         fallback_capacity=1 << 8)
    dev = em.EmBackend(cfg, tables, atm)
    with pytest.raises(c8mi.C8miError) as err:
        dev.slot_enqueue(1, em.particle(22, 1.0, (0, 0, 1e4), (0, 0, -1)))
    assert err.value.status == 2  # slot 1 not begun
    with pytest.raises(c8mi.C8miError) as err:
        dev.slot_begin(2, 1)
    assert err.value.status == 1  # outside the pool
    dev.slot_begin(1, 7)
    dev.slot_enqueue(1, em.particle(22, 1.0, (0, 0, 1e4), (0, 0, -1), history_id=1))
    with pytest.raises(c8mi.C8miError) as err:
        dev.slot_end(1)
    assert err.value.status == 2  # pending particles
    dev.run_rounds()
    dev.slot_end(1)
    assert dev.slot_state(1)["active"] == 0
    dev.close()


def serve_requests(atm):
    reqs = []
    for k, case in enumerate(POOL_CASES):
        pid, energy, shower, start, d = pool_case(case)
        reqs.append(dict(primary=em.particle(pid, energy, start, d, history_id=1), shower_id=shower,
                         arrival_s=0.0, footprint=60.0 * energy, axis=em.shower_axis(atm, start, d)))
    return reqs


@pytest.mark.parametrize("slots,budget,moliere_table", [(3, 250.0, None), (1, 1e9, None), (4, 250.0, True),
                                                        (1, 1e9, True)])
def test_serving_loop_results_equal_solo_runs(tables, slots, budget, moliere_table):
    """The native serving loop (footprint-limited admission into a pool, or the
    one-at-a-time form with one slot): every request's statistics and ledger
    equal the oracle's run of that shower alone, with and without Molière
    scattering (through the interpolation table)."""
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=21, atmosphere=atm, slots=slots, observation_capacity=1 << 14,
                         fallback_capacity=1 << 10, moliere=moliere_table is not None, moliere_table=moliere_table)
    dev = em.EmBackend(cfg, tables, atm)
    results, totals = dev.serve(serve_requests(atm), budget, realtime=False)
    stats = dev.statistics()
    dev.close()
    if slots > 1:
        # fronts that showers of the other species' call sat out: their entries
        # were partitioned behind the participants and carried over
        assert stats["parked_fronts"] > 0 and stats["parked_particles"] > 0, stats
    else:
        assert stats["parked_fronts"] == 0 and stats["parked_particles"] == 0, stats
    steps = 0
    for case, res in zip(POOL_CASES, results):
        ocfg, want = solo_oracle(tables, atm, case, moliere_table=moliere_table)
        for k in ("photon_steps", "lepton_steps", "photon_interactions", "lepton_interactions", "secondaries",
                  "observations", "fallbacks", "next_history_id"):
            assert res[k] == want["stats"][k], (case, k)
        wl = em.decode_counters(want["counters"], ocfg.struct.profile_energy_scale)
        for k in em.LEDGER_ENERGY_TERMS:
            assert abs(res["ledger"][k] - wl[k]) <= 1e-9 * case[1], (case, k)
        assert res["ledger"]["fallback_reasons"] == wl["fallback_reasons"], case
        assert res["finished_s"] >= res["admitted_s"] >= 0.0
        steps += res["photon_steps"] + res["lepton_steps"]
    assert totals["photon_steps"] + totals["lepton_steps"] == steps
    if slots == 1:
        order = sorted(range(len(results)), key=lambda k: results[k]["admitted_s"])
        for a, b in zip(order, order[1:]):
            assert results[b]["admitted_s"] >= results[a]["finished_s"]
