"""Step compositions of the oracle pinned by what the physics fixes.

The reference's step algorithms (InteractionSelection.inl, LeptonVertexSelection.inl,
PhotonFinalStateStep.inl, LeptonFinalStateStep.inl) cannot be compiled here
(they need build-generated particle properties or Kokkos), so the oracle's
restatement of them is checked against laws that a plausible slip breaks:

* interaction depth: X * Sigma(E) is exponential with mean 1, where Sigma is
  the native table's total rate (a table query pinned bit for bit against the
  reference code), so a wrong sign, scale or rate fails;
* process choice: each (process, component) column is chosen with frequency
  sigma_i / Sigma_selection from the native per-column rates;
* vertex reselection (LeptonVertexSelection.inl:41-199): the lepton keeps its
  interaction with probability Sigma(E_vertex) / Sigma(E_start), its uniform
  is U(0, 1);
* final states (two-body kinematics on an electron at rest): Moller / Bhabha
  and in-flight annihilation conserve energy and momentum exactly, Compton
  too; bremsstrahlung conserves energy, emits the photon at the polar angle
  m/E (LeptonFinalStateStep.inl:957-1025) and leaves the lepton along the
  momentum left; pair production conserves energy and puts the pair on
  opposite azimuths. A wrong angle formula, azimuth offset, mass or energy
  fraction fails these.
"""
import numpy as np
import pytest
from scipy import stats

import oracle_lib
from c8mi import airtables, em
from c8mi import tables as native

P = native.PROCESS
START = (0.0, 0.0, 112750.0)
DOWN = (0.0, 0.0, -1.0)


@pytest.fixture(scope="module")
def air():
    return airtables.build_air_tables()


@pytest.fixture(scope="module")
def vertices(air):
    atm = em.Atmosphere()
    logs = []
    for pid, energy, seed in ((22, 10.0, 1), (11, 5.0, 2), (-11, 5.0, 3), (22, 30.0, 4), (-11, 1.0, 5)):
        cfg = em.make_config(airtables, seed=seed, atmosphere=atm)
        v, _ = oracle_lib.em_vertices(cfg, air, atm, em.particle(pid, energy, START, DOWN, history_id=1))
        logs.append(v)
    return np.concatenate(logs)


def _momentum(energy, mass):
    return np.sqrt(np.maximum((energy + mass) * (energy - mass), 0.0))


@pytest.mark.parametrize("process", ["ionization", "annihilation", "compton"])
def test_two_body_final_states_conserve_energy_and_momentum(vertices, process):
    fs = vertices[(vertices["kind"] == 1) & (vertices["process"] == P[process])]
    assert fs.size > 30, fs.size
    worst_e = worst_p = 0.0
    for r in fs:
        e_in = r["energy"] + r["target_mass"]
        p_in = _momentum(r["energy"], r["mass"]) * r["direction"]
        masses = {"ionization": (r["mass"], r["target_mass"]), "annihilation": (0.0, 0.0),
                  "compton": (0.0, r["target_mass"])}[process]
        assert r["count"] == 2
        e_out = r["child_energy"][0] + r["child_energy"][1]
        p_out = sum(_momentum(r["child_energy"][k], masses[k]) * r["child_direction"][k] for k in range(2))
        worst_e = max(worst_e, abs(e_out - e_in) / e_in)
        worst_p = max(worst_p, np.linalg.norm(p_out - p_in) / np.linalg.norm(p_in))
    assert worst_e <= 1e-12, worst_e
    assert worst_p <= 1e-8, worst_p


def test_bremsstrahlung_angles_and_energy(vertices):
    fs = vertices[(vertices["kind"] == 1) & (vertices["process"] == P["brems"])]
    assert fs.size > 100
    for r in fs:
        lepton_e, photon_e = r["child_energy"][0], r["child_energy"][1]
        assert abs(lepton_e + photon_e - r["energy"]) <= 1e-12 * r["energy"]
        assert abs(photon_e - r["v"] * r["energy"]) <= 1e-12 * r["energy"]
        d, dg, dl = r["direction"], r["child_direction"][1], r["child_direction"][0]
        # photon polar angle m / E around the parent
        assert abs(np.dot(dg, d) - np.cos(r["mass"] / r["energy"])) <= 1e-12
        # the lepton moves along the momentum the photon leaves
        left = _momentum(r["energy"], r["mass"]) * d - photon_e * dg
        assert np.linalg.norm(np.cross(dl, left / np.linalg.norm(left))) <= 1e-12
        assert np.dot(dl, left) > 0


def test_pair_production_energy_and_azimuths(vertices):
    fs = vertices[(vertices["kind"] == 1) & (vertices["process"] == P["photon_pair"])]
    assert fs.size > 50
    for r in fs:
        assert r["child_pid"][0] == 11 and r["child_pid"][1] == -11
        assert abs(r["child_energy"][0] + r["child_energy"][1] - r["energy"]) <= 1e-12 * r["energy"]
        assert abs(r["child_energy"][0] - r["v"] * r["energy"]) <= 1e-12 * r["energy"]
        d = r["direction"]
        t = [c - np.dot(c, d) * d for c in r["child_direction"][:2]]
        n = [np.linalg.norm(x) for x in t]
        if min(n) > 1e-9:
            assert np.dot(t[0], t[1]) / (n[0] * n[1]) <= -1.0 + 1e-9


def test_lepton_vertex_reselection_follows_the_rate_ratio(vertices):
    attempts = vertices[vertices["kind"] == 0]
    assert attempts.size > 1000
    u = attempts["process_uniform"]
    assert stats.kstest(u, "uniform").pvalue > 1e-4
    ratio = np.minimum(attempts["rate_vertex"] / attempts["rate_start"], 1.0)
    expected = ratio.mean()
    sigma = np.sqrt(np.sum(ratio * (1.0 - ratio))) / attempts.size
    assert abs(attempts["accepted"].mean() - expected) <= 4.0 * sigma + 1e-12
    # where the rate fell by more than 2%, rejections happen at the right rate
    low = ratio < 0.98
    if low.sum() > 200:
        got = attempts["accepted"][low].mean()
        want = ratio[low].mean()
        sig = np.sqrt(np.sum(ratio[low] * (1.0 - ratio[low]))) / low.sum()
        assert abs(got - want) <= 4.0 * sig


def _native(tables, kind, pid, process, component, energy_mev):
    q = np.zeros(1, dtype=native.QUERY_DTYPE)
    q["kind"], q["pdg_id"], q["process_id"], q["component_hash"], q["energy_mev"] = (
        native.QUERY[kind], pid, process, component, energy_mev)
    r = oracle_lib.native_queries(tables, q)
    assert r["status"][0] == 0
    return float(r["value"][0])


@pytest.mark.parametrize("energy_gev", [0.005, 0.1, 10.0])
def test_photon_interaction_depth_and_column_choice(energy_gev):
    """Photons in homogeneous ice (the in-ice scene's medium, one model)."""
    ice = em.IceVolume()
    tables = airtables.build_ice_tables(cut_mev=0.1)
    cfg = em.make_config(airtables.IceMedium, seed=77, cut_mev=0.1, start=(0.0, 0.0, -100.0), atmosphere=ice)
    n = 6000
    photons = np.concatenate([em.particle(22, energy_gev, (0.0, 0.0, -100.0), DOWN, history_id=h + 1).reshape(1)
                              for h in range(n)])
    sel = oracle_lib.em_select(cfg, tables, ice, photons)
    ok = (sel["fallback"] == 0) & np.isfinite(sel["interaction_grammage"])
    assert ok.mean() > 0.99
    sel = sel[ok]
    e_mev = energy_gev * 1000.0
    total = _native(tables, "total_rate", 22, 0, 0, e_mev)
    assert np.all(sel["total_rate"] == total)
    depth = sel["interaction_grammage"] * total
    assert stats.kstest(depth, "expon").pvalue > 1e-4
    assert abs(depth.mean() - 1.0) <= 5.0 / np.sqrt(depth.size)
    # column frequencies
    columns = tables.dndx[tables.dndx["pdg_id"] == 22]
    selection_total = _native(tables, "selection_total_rate", 22, 0, 0, e_mev)
    rates = np.array([_native(tables, "rate", 22, int(c["process_id"]), int(c["component_hash"]), e_mev)
                      if e_mev >= c["lower_energy_limit_mev"] else 0.0 for c in columns])
    assert abs(rates.sum() - selection_total) <= 1e-9 * selection_total
    observed = np.array([np.count_nonzero((sel["process"] == c["process_id"]) &
                                          (sel["component"] == c["component_hash"])) for c in columns])
    assert observed.sum() == sel.size
    expected = rates / rates.sum() * sel.size
    keep = expected > 5.0
    chi2 = np.sum((observed[keep] - expected[keep]) ** 2 / expected[keep])
    assert stats.chi2.sf(chi2, max(int(keep.sum()) - 1, 1)) > 1e-4, (observed, expected)
    assert observed[~keep].sum() <= max(20.0, 5.0 * expected[~keep].sum())
