"""The output sink on the CPU: c8mi_em_output_profile (CorsikaOutputSink::
onGpuProfile, CorsikaOutputSink.inl:117-197) and c8mi_em_energy_ledger_report
(the energy_ledger block of ShowerReport.inl:1221-1287).

* The report's arithmetic is the reference's definition, checked on a ledger
  of known terms (every term a distinct power of two, so a dropped, doubled or
  sign-flipped term changes the result by a term no other combination gives)
  and on oracle showers, where it must agree with the independent closure of
  em.ledger_residual;
* the complete-coverage rule and the 1e-4 acceptance, case by case;
* the sink's checks reject what onGpuProfile rejects (overflows, invalid
  records, negative or non-finite bins and ledger terms) and write nothing;
* the host and device ledgers of a handed-off shower (the oracle's hybrid) sum
  to a closed report whose injected energy is the primary's.
"""
import ctypes
import functools
import os
import subprocess

import numpy as np
import pytest

import c8mi
import oracle_lib
from c8mi import airtables, em

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
START = (0.0, 0.0, 112750.0)
DOWN = (0.0, 0.0, -1.0)


def test_output_structs_match_c_layout(tmp_path):
    names = {"c8mi_em_output_statistics": em.OutputStatistics, "c8mi_em_report_context": em.ReportContext,
             "c8mi_em_energy_ledger": em.EnergyLedger, "c8mi_em_ledger": em.Ledger}
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "c8mi/c8mi_output.h"\nint main(void){\n' +
                   "".join(f'printf("%zu\\n", sizeof({n}));\n' for n in names) + "return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    for (name, cls), size in zip(names.items(), sizes):
        assert ctypes.sizeof(cls) == size, name


def known_ledger():
    terms = dict(injected_energy=64.0, continuous_deposit=2.0 ** 5, cut_deposit=2.0 ** 4, cut_rest_mass=2.0 ** 3,
                 observed_energy=2.0 ** 2, escaped_energy=2.0 ** 1, observation_cut_overlap=2.0 ** -1,
                 mass_convention_correction=2.0 ** -2, medium_rest_mass_input=2.0 ** -3,
                 photoelectric_binding=2.0 ** -4, native_muon_pair_cut_energy=-(2.0 ** -5),
                 fallback_energy=0.0, pending_energy=0.0, thinning_discarded_energy=0.0)
    led = {k: 0 for k, _ in em.Ledger._fields_}
    led.update(terms)
    led["lepton_limits"], led["fallback_reasons"] = [0] * 8, [0] * 32
    return led


def test_report_arithmetic_on_known_terms():
    r = em.energy_ledger_report(known_ledger(), initial_total_gev=100.0)
    assert r["deposited_gev"] == 2.0 ** 5 + 2.0 ** 4 - 2.0 ** -5
    assert r["source_gev"] == 100.0 + 2.0 ** -3 + 2.0 ** -2
    assert r["terminal_gev"] == (2.0 ** 5 + 2.0 ** 4 - 2.0 ** -5) + 2.0 ** 3 + 2.0 ** 2 + 2.0 ** 1 + 2.0 ** -4 - 2.0 ** -1
    assert r["residual_gev"] == r["source_gev"] - r["terminal_gev"]
    assert r["relative_closure_error"] == abs(r["residual_gev"]) / r["source_gev"]
    assert r["acceptance_tolerance"] == 1e-4 and r["complete_coverage"] and not r["accepted"]
    assert r["injected_gev"] == 64.0 and r["definition"].startswith("initial_total")
    # two sides add term by term
    r2 = em.energy_ledger_report([known_ledger(), known_ledger()], initial_total_gev=200.0)
    assert r2["terminal_gev"] == 2 * r["terminal_gev"]
    assert r2["source_gev"] == 200.0 + 2 * (2.0 ** -3 + 2.0 ** -2)


@functools.lru_cache(maxsize=1)
def air():
    return airtables.build_air_tables()


def shower(seed, energy=10.0, pid=22, thinning=None, moliere=False):
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=seed, atmosphere=atm, thinning=thinning, moliere=moliere,
                         moliere_table=True if moliere else None)
    prim = em.particle(pid, energy, START, DOWN, history_id=1)
    want = oracle_lib.em_run_shower(cfg, air(), atm, prim, shower_id=seed)
    hist = {k: v / (cfg.struct.profile_energy_scale if k == "energy_loss" else cfg.struct.profile_weight_scale)
            for k, v in want["hist"].items()}
    return cfg, hist, em.decode_counters(want["counters"], cfg.struct.profile_energy_scale)


@pytest.mark.parametrize("seed,pid,energy", [(1, 22, 10.0), (2, 11, 5.0), (3, -11, 3.0)])
def test_report_of_oracle_showers_closes(seed, pid, energy):
    cfg, hist, led = shower(seed, energy, pid)
    assert sum(led["fallback_reasons"]) == 0
    r = em.energy_ledger_report(led, initial_total_gev=energy)
    assert r["complete_coverage"] and r["accepted"]
    assert r["relative_closure_error"] <= 1e-9
    # the report's residual is the closure em.ledger_residual computes its own way
    assert abs(r["extended_residual_gev"] - em.ledger_residual(led)) <= 1e-12 * energy
    assert abs(r["injected_gev"] - energy) <= 1e-9 * energy


def test_coverage_rules():
    cfg, hist, led = shower(1)
    base = dict(ledgers=led, initial_total_gev=10.0)
    assert em.energy_ledger_report(**base)["accepted"]
    for change in (dict(electromagnetic_primary=False), dict(thinning_fraction=1e-6), dict(scalar_em_steps=1),
                   dict(cpu_fallbacks=1), dict(memory_spills=1), dict(specified_final_states=1)):
        r = em.energy_ledger_report(**base, **change)
        assert not r["complete_coverage"] and not r["accepted"], change
    fb = dict(led)
    fb["fallback_reasons"] = list(led["fallback_reasons"])
    fb["fallback_reasons"][21] = 1
    assert not em.energy_ledger_report(fb, 10.0)["complete_coverage"]
    pending = dict(led, pending_energy=0.5)
    assert not em.energy_ledger_report(pending, 10.0)["complete_coverage"]
    # a closure error above the tolerance is reported, not accepted
    r = em.energy_ledger_report(led, 10.0 * (1 + 2e-4))
    assert r["complete_coverage"] and not r["accepted"] and r["relative_closure_error"] > 1e-4
    assert em.energy_ledger_report(led, 10.0 * (1 + 2e-4), tolerance=1e-3)["accepted"]


def test_thinned_shower_books_its_discarded_energy():
    cfg, hist, led = shower(5, 30.0, thinning=(0.5, 50.0))
    r = em.energy_ledger_report(led, 30.0, thinning_fraction=1e-6)
    assert not r["complete_coverage"] and r["thinning_discarded_energy_gev"] != 0.0
    assert abs(r["extended_residual_gev"] - em.ledger_residual(led)) <= 1e-12 * 30.0


@pytest.mark.parametrize("change", [dict(fixed_point_overflows=1), dict(invalid_records=2),
                                    dict(observed_energy=-1.0), dict(escaped_energy=float("nan")),
                                    dict(native_muon_pair_cut_energy=0.25), dict(cut_rest_mass=-1e-3),
                                    dict(photoelectric_binding=-1e-9), dict(observation_cut_overlap=-1.0),
                                    dict(medium_rest_mass_input=-1.0),
                                    dict(mass_convention_correction=float("inf"))])
def test_invalid_ledgers_are_refused(change):
    cfg, hist, led = shower(1)
    bad = dict(led, **change)
    with pytest.raises(c8mi.C8miError) as err:
        em.energy_ledger_report(bad, 10.0)
    assert err.value.status == 1
    with pytest.raises(c8mi.C8miError):
        em.output_profile(hist, bad)


def test_output_profile_layout():
    cfg, hist, led = shower(2, 5.0, 11)
    out = em.output_profile(hist, led)
    L = out["longitudinal"]
    assert L.shape == (hist["photons"].size, 5)
    assert np.array_equal(L[:, 0], hist["photons"]) and np.array_equal(L[:, 1], hist["electrons"])
    assert np.array_equal(L[:, 2], hist["positrons"]) and not L[:, 3:].any()
    assert np.array_equal(out["energy_loss_em"], hist["energy_loss"]) and not out["energy_loss_muon"].any()
    st = out["statistics"]
    assert st["steps"] == led["steps"] and st["deposited_steps"] == led["deposited_steps"]
    assert st["weighted_deposited_energy_gev"] == led["continuous_deposit"] + led["cut_deposit"] + \
        led["native_muon_pair_cut_energy"]
    for key, value in (("photons", -1.0), ("electrons", float("nan")), ("energy_loss", -1.0)):
        bad = {k: v.copy() for k, v in hist.items()}
        bad[key][3] = value
        with pytest.raises(c8mi.C8miError) as err:
            em.output_profile(bad, led)
        assert err.value.status == 1
    empty = {k: np.zeros(0) for k in hist}
    with pytest.raises(c8mi.C8miError) as err:
        em.output_profile(empty, led)
    assert err.value.status == 3


def test_report_of_a_handed_off_shower():
    """The oracle's host handoff: host and device ledgers together close, and
    their injected energies add up to the primary."""
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=3, atmosphere=atm)
    prim = em.particle(22, 10.0, START, DOWN, history_id=1)
    want = oracle_lib.em_run_hybrid(cfg, air(), atm, prim, shower_id=3, hybrid=em.hybrid_config(minimum_batch=64))
    es = cfg.struct.profile_energy_scale
    sides = [em.decode_counters(want[s]["counters"], es) for s in ("host", "device")]
    steps = want["result"]["cpu_photon_steps"] + want["result"]["cpu_lepton_steps"]
    r = em.energy_ledger_report(sides, 10.0, scalar_em_steps=steps)
    assert steps > 0 and not r["complete_coverage"]  # the reference's rule: host EM steps
    assert abs(r["injected_gev"] - 10.0) <= 1e-9 * 10.0
    assert r["relative_closure_error"] <= 1e-9
