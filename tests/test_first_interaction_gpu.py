"""The writers' first-interaction snapshot on the GPU
(IEmBackend::downloadFirstInteractionSnapshot, IEmBackend.hpp:87;
FirstInteractionRecord.hpp:15-27; captured at the first final state of a
generation-zero parent, PhotonFront.inl:95-111, LeptonFinalStateStep.inl:609-615)
against the oracle's capture: the parent at the vertex, the process, and the
children before thinning with the parent's weight. Positions and energies to
1e-12 relative (the device and the oracle round transcendental functions
each in their own way); identities exactly."""
import numpy as np
import pytest

import oracle_lib
from c8mi import airtables, em

pytestmark = pytest.mark.gpu

START = (0.0, 0.0, 112750.0)
DOWN = (0.0, 0.0, -1.0)
INTS = ("pid", "medium_id", "generation", "history_id", "parent_history_id", "step_id", "environment_node")
FLOATS = ("energy_gev", "position_m", "direction", "time_s", "weight")


@pytest.fixture(scope="module")
def tables():
    return airtables.build_air_tables()


def same_states(got, want):
    for f in INTS:
        assert np.array_equal(got[f], want[f]), f
    for f in FLOATS:
        np.testing.assert_allclose(got[f], want[f], rtol=1e-12, atol=1e-9, err_msg=f)


def same_snapshot(got, want):
    assert (got is None) == (want is None)
    if got is None:
        return
    assert got["process_id"] == want["process_id"]
    n = int(got["secondary_count"])
    assert n == want["secondary_count"] and n in (1, 2)
    same_states(got["parent_at_vertex"].reshape(1), want["parent_at_vertex"].reshape(1))
    same_states(got["secondaries"][:n], want["secondaries"][:n])
    assert np.all(got["secondaries"]["weight"][:n] == got["parent_at_vertex"]["weight"])


@pytest.mark.parametrize("pid,energy,seed,thinning", [(22, 100.0, 1, None), (22, 0.005, 2, None),
                                                      (11, 10.0, 3, None), (-11, 0.2, 4, None),
                                                      (22, 50.0, 5, (5.0, 100.0))])
def test_snapshot_equals_oracle(tables, pid, energy, seed, thinning):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, moliere=True, moliere_table=True,
                         thinning=thinning)
    prim = em.particle(pid, energy, START, DOWN, history_id=1)
    want, _ = oracle_lib.em_first_interaction(cfg, tables, atm, prim, shower_id=seed)
    dev = em.EmBackend(cfg, tables, atm)
    dev.begin_shower(seed, 2)
    dev.enqueue(prim)
    dev.run()
    got = dev.first_interaction()
    # a later shower on the same handle starts without a snapshot
    dev.begin_shower(seed + 100, 2)
    assert dev.first_interaction() is None
    dev.close()
    assert want is not None
    same_snapshot(got, want)
    assert got["parent_at_vertex"]["generation"] == 0
    assert np.all(got["secondaries"]["generation"][:got["secondary_count"]] == 1)


def test_pool_slots_keep_their_own_snapshots(tables):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=8, atmosphere=atm, slots=4, moliere=True, moliere_table=True)
    cases = [(22, 30.0, 81), (11, 5.0, 82), (-11, 1.0, 83)]
    dev = em.EmBackend(cfg, tables, atm)
    for slot, (pid, energy, shower) in enumerate(cases):
        dev.slot_begin(slot, shower, 2)
        dev.slot_enqueue(slot, em.particle(pid, energy, START, DOWN, history_id=1))
    dev.run_rounds()
    for slot, (pid, energy, shower) in enumerate(cases):
        want, _ = oracle_lib.em_first_interaction(cfg, tables, atm, em.particle(pid, energy, START, DOWN, history_id=1),
                                                  shower_id=shower)
        same_snapshot(dev.slot_first_interaction(slot), want)
    assert dev.slot_first_interaction(3) is None
    dev.close()


def test_photoelectric_snapshot_child_takes_its_history_id():
    """A photon whose first final state is photoelectric: the snapshot's electron
    carries the history ID the front assigns it (PhotonFinalStateStep.inl:649-660),
    patched in by the scatter once the ID is known; equal to the oracle's."""
    tables = airtables.build_air_tables(cut_mev=0.05)
    atm = em.Atmosphere()
    prim = em.particle(22, 6e-5, (0.0, 0.0, 20000.0), DOWN, history_id=1)
    found = 0
    for seed in range(1, 200):
        cfg = em.make_config(airtables, seed=seed, atmosphere=atm, cut_mev=0.05)
        want, _ = oracle_lib.em_first_interaction(cfg, tables, atm, prim, shower_id=seed)
        if want is None or want["process_id"] != 1000000016:
            continue
        print(f"photoelectric snapshot: seed {seed}", flush=True)
        dev = em.EmBackend(cfg, tables, atm)
        dev.begin_shower(seed, 2)
        dev.enqueue(prim)
        dev.run()
        got = dev.first_interaction()
        dev.close()
        same_snapshot(got, want)
        assert got["secondaries"][0]["history_id"] == 2  # the first ID of the call
        found += 1
        if found == 2:
            break
    assert found == 2
