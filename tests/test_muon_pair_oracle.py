"""Photon mu-pair final states (PhotonFinalStateStep.inl:326-382, 679-731) and
the native cut of their muons (NativeMuonPairCut.inl, NativeMuonPairCutRecord.inl,
ProfileAccumulation.inl:300-345) on the CPU.

The air tables carry photon mu-pair columns at 1e4 times their rate here, so a
100 GeV shower makes several pairs. Pinned by:

* two-body kinematics fixed by the reference's approximation: the muons share
  the photon's energy (native totals E(1-x) and E x), both leave at the polar
  angle m/E around the photon and on opposite azimuths, with the photon's
  weight and no thinning;
* the split x follows the table's CDF: F_E(x) of every sampled x, with F the
  normalised column shape over the mu-pair limits, is uniform (KS);
* a muon made at or below its transport mass is cut at its first step: one
  zero-length step, no track, its signed E - m in the native-cut ledger term,
  the rest mass in the cut rest-mass term, closure exact;
* a muon above it has no table in this set: a selection fallback
  (ParticleTableMissing) that carries its energy;
* the product's host stepper (the device step functions compiled for the host)
  steps photons into mu-pairs, and their muons, bit for bit as the oracle.
"""
import numpy as np
import pytest
from scipy import stats

import oracle_lib
from c8mi import airtables, em
from c8mi import tables as native
from test_hybrid_oracle import same_particles

P = native.PROCESS
START = (0.0, 0.0, 112750.0)
DOWN = (0.0, 0.0, -1.0)
MUON_MASS_GEV = native.CONSTANTS["muon_mass_mev"] / 1000.0
TRANSPORT_MUON_MASS_GEV = 1.056584e-1
FR_PARTICLE_TABLE = 3


@pytest.fixture(scope="module")
def mu_air():
    return airtables.build_air_tables(muon_pair=1.0e4)


@pytest.fixture(scope="module")
def mu_vertices(mu_air):
    atm = em.Atmosphere()
    logs = []
    for seed in range(1, 13):
        cfg = em.make_config(airtables, seed=seed, atmosphere=atm)
        v, _ = oracle_lib.em_vertices(cfg, mu_air, atm, em.particle(22, 100.0, START, DOWN, history_id=1))
        logs.append(v[(v["kind"] == 1) & (v["process"] == P["photon_muon_pair"])])
    return np.concatenate(logs)


def test_muon_pair_kinematics(mu_vertices):
    assert mu_vertices.size >= 20, mu_vertices.size
    for r in mu_vertices:
        assert r["count"] == 2 and r["child_pid"][0] == 13 and r["child_pid"][1] == -13
        assert r["target_mass"] == MUON_MASS_GEV
        e = r["energy"]
        minus, plus = r["child_energy"][0], r["child_energy"][1]
        assert abs(minus + plus - e) <= 1e-12 * e
        # x = E(mu+)/E inside the mu-pair limits of the split
        xmin, xmax = airtables.muon_pair_limits(e * 1000.0)
        assert xmin <= plus / e <= xmax
        d = r["direction"]
        c = np.cos(MUON_MASS_GEV / e)
        for k in range(2):
            assert abs(np.dot(r["child_direction"][k], d) - c) <= 1e-12
        t = [r["child_direction"][k] - np.dot(r["child_direction"][k], d) * d for k in range(2)]
        n = [np.linalg.norm(x) for x in t]
        assert abs(n[0] - n[1]) <= 1e-12
        assert np.dot(t[0], t[1]) / (n[0] * n[1]) <= -1.0 + 1e-9


def test_muon_pair_split_follows_the_table_cdf(mu_vertices):
    e = mu_vertices["energy"] * 1000.0
    x = mu_vertices["child_energy"][:, 1] / mu_vertices["energy"]
    xmin, xmax = airtables.muon_pair_limits(e)
    cdf = airtables.pair_cdf
    u = (cdf(x) - cdf(xmin)) / (cdf(xmax) - cdf(xmin))
    # the column is a bicubic fit of this shape: a wrong draw, limit or
    # transform moves the distribution by far more than the fit error
    assert stats.kstest(u, "uniform").pvalue > 1e-3


def scene(seed):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, moliere=True, moliere_table=True)
    return cfg, atm


def muon(energy, reserved, generation=1, step_id=0, pid=13):
    p = em.particle(pid, energy, (0.0, 0.0, 20000.0), DOWN, history_id=77)
    p["reserved"], p["generation"], p["step_id"] = reserved, generation, step_id
    return p


EVERYTHING_ON_THE_HOST = em.hybrid_config(minimum_batch=1 << 40, expansion_rounds=1 << 40)


def one_step(cfg, tables, atm, p, seed):
    """The muon's step on the host stepper, and the oracle's host step of the
    same particle (a muon never routes to the device)."""
    stepper = em.ScalarStepper(cfg, tables, atm)
    batch, _, tracks, result = stepper.prefix(p, seed, config=EVERYTHING_ON_THE_HOST)
    _, ledger = stepper.download_profile()
    fbs, n_fb = stepper.fallbacks()
    stepper.close()
    want = oracle_lib.em_run_hybrid(cfg, tables, atm, p, shower_id=seed, hybrid=EVERYTHING_ON_THE_HOST,
                                    prefix_only=True)
    assert result["cpu_particles"] == want["result"]["cpu_particles"] == 1
    assert batch.size == want["staged"].size == 0
    return tracks, ledger, fbs[:n_fb], want["host"]


def test_native_cut_muon(mu_air):
    cfg, atm = scene(3)
    w = 2.5
    for energy in (0.1, 0.104, TRANSPORT_MUON_MASS_GEV):
        p = muon(energy, reserved=1)
        p["weight"] = w
        tracks, ledger, fbs, want = one_step(cfg, mu_air, atm, p, 3)
        assert tracks.shape[0] == 0 and fbs.size == 0
        es = cfg.struct.profile_energy_scale
        got = em.decode_counters(want["counters"], es)
        for led in (ledger, got):
            assert led["steps"] == 1 and led["lepton_limits"][2] == 1
            assert abs(led["native_muon_pair_cut_energy"] - (energy - TRANSPORT_MUON_MASS_GEV) * w) <= 1.0 / es
            assert abs(led["cut_rest_mass"] - TRANSPORT_MUON_MASS_GEV * w) <= 1.0 / es
            assert led["cut_deposit"] == 0.0
            assert led["deposited_steps"] == (1 if energy < TRANSPORT_MUON_MASS_GEV else 0)
        for k in em.LEDGER_ENERGY_TERMS:
            assert ledger[k] == got[k], k
        assert want["tracks"].shape[1] == 0
        # the muon's energy is what the cut books: the shower closes
        assert abs(em.ledger_residual(ledger)) <= 2.0 / es


@pytest.mark.parametrize("p", [muon(0.1, reserved=0), muon(0.1, reserved=1, step_id=1),
                               muon(0.1, reserved=1, generation=0), muon(2.0, reserved=1),
                               muon(2.0, reserved=0, pid=-13)])
def test_muons_without_the_native_cut_fall_back(mu_air, p):
    cfg, atm = scene(4)
    tracks, ledger, fbs, want = one_step(cfg, mu_air, atm, p, 4)
    assert tracks.shape[0] == 0
    assert fbs.size == 1 and ledger["fallback_reasons"][FR_PARTICLE_TABLE] == 1
    same_particles(fbs["particle"], want["fallbacks"]["particle"])
    assert fbs["reason"][0] == want["fallbacks"]["reason"][0] == FR_PARTICLE_TABLE
    assert abs(ledger["fallback_energy"] - p["energy_gev"] * p["weight"]) <= 1e-12


def test_host_photon_steps_with_muon_pairs_equal_oracle(mu_air):
    """One host step of each of 400 photons (the router's single expansion
    round, then the host steps of what it made that the device cannot take):
    the product's outputs, ledger and records equal the oracle's bit for bit,
    and many of the steps end in a mu-pair."""
    cfg, atm = scene(6)
    one_round = em.hybrid_config(minimum_batch=1 << 40, expansion_rounds=1)
    rng = np.random.default_rng(11)
    stepper = em.ScalarStepper(cfg, mu_air, atm)
    pairs = 0
    for i in range(400):
        energy = float(np.exp(rng.uniform(np.log(0.5), np.log(2000.0))))
        d = rng.normal(size=3)
        d[2] = -abs(d[2])
        d /= np.linalg.norm(d)
        p = em.particle(22, energy, (0.0, 0.0, float(rng.uniform(2000.0, 40000.0))), tuple(d), history_id=1)
        batch, next_history, tracks, result = stepper.prefix(p, 500 + i, config=one_round)
        _, ledger = stepper.download_profile()
        want = oracle_lib.em_run_hybrid(cfg, mu_air, atm, p, shower_id=500 + i, hybrid=one_round,
                                        prefix_only=True)
        same_particles(batch, want["staged"])
        assert next_history == want["result"]["next_history_id"]
        got = em.decode_counters(want["host"]["counters"], cfg.struct.profile_energy_scale)
        for k in em.LEDGER_ENERGY_TERMS:
            assert ledger[k] == got[k], k
        assert ledger["fallback_reasons"] == got["fallback_reasons"]
        fbs, n_fb = stepper.fallbacks()
        same_particles(fbs["particle"][:n_fb], want["host"]["fallbacks"]["particle"])
        # the muons of a pair: host steps of their own, which hand them on
        muons = fbs["particle"][:n_fb]
        muons = muons[np.abs(muons["pid"]) == 13]
        if muons.size:
            pairs += 1
            assert sorted(muons["pid"].tolist()) == [-13, 13]
            assert np.all(muons["weight"] == p["weight"]) and np.all(muons["generation"] == 1)
            assert np.all(muons["step_id"] == 0) and np.all(muons["parent_history_id"] == 1)
    stepper.close()
    assert pairs >= 20, pairs
