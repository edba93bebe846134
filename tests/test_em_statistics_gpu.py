"""Shower-level observables of the device against the oracle, statistically.

The protocol of the reference's numerical validation
(documentation/kokkos/validation/numerics.rst:1-60): many seeded showers of
the benchmark configuration (100 GeV vertical photons, USStdBK air, 1 MeV cut,
Molière scattering through the interpolation table) from each side, compared
as distributions. Device and oracle showers use different shower IDs, so the
two samples are independent (the same IDs would give identical showers, which
the parity tests check directly); the bar is:
  * Kolmogorov-Smirnov on the depth of the charged-particle maximum (Xmax) and
    on the number of particles reaching the observation level: p > 1e-3;
  * Hotelling's T^2 of the mean charged-particle profile (fine bins summed
    eight at a time; the full sample covariance, since a shower's
    fluctuation moves many bins together): p > 1e-3;
  * the fraction of showers with a particle at the observation level
    (two-proportion z test): p > 1e-3;
  * every device shower's energy ledger closed to 1e-9 of the primary energy
    (showers with a Molière-failed step excepted: that transport fallback
    carries the vertex state, MoliereStep.inl:60-64, and its deposit is the
    host's).
"""
import os

import numpy as np
import pytest
from scipy import stats

import oracle_lib
from c8mi import airtables, em

pytestmark = pytest.mark.gpu

SHOWERS = 1000
ENERGY = 100.0
START = (0.0, 0.0, 112750.0)
DOWN = (0.0, 0.0, -1.0)


def charged_maximum(electrons, positrons, width):
    return (np.argmax(electrons + positrons) + 0.5) * width


def test_shower_observables_agree_with_the_oracle_over_1000_showers():
    tables = airtables.build_air_tables()
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, seed=2026, atmosphere=atm, slots=SHOWERS, queue_capacity=1 << 24,
                         observation_capacity=1 << 8, fallback_capacity=1 << 6, moliere=True,
                         moliere_table=True)
    width = cfg.struct.profile_bin_width_g_per_cm2
    prim = em.particle(22, ENERGY, START, DOWN, history_id=1)
    dev = em.EmBackend(cfg, tables, atm)
    for k in range(SHOWERS):
        dev.slot_begin(k, 1 + k, 2)
        dev.slot_enqueue(k, prim)
    print(f"device: {SHOWERS} showers of {ENERGY} GeV", flush=True)
    totals = dev.run_rounds()
    assert totals["pending_photons"] == 0 and totals["pending_leptons"] == 0
    dev_xmax, dev_ground, dev_profile, residuals = [], [], [], []
    for k in range(SHOWERS):
        hist, ledger = dev.slot_download_profile(k)
        dev_xmax.append(charged_maximum(hist["electrons"], hist["positrons"], width))
        dev_profile.append(hist["electrons"] + hist["positrons"])
        dev_ground.append(dev.slot_state(k)["observations"])
        if ledger["fallback_reasons"][21] == 0:
            residuals.append(abs(em.ledger_residual(ledger)))
    dev.close()
    assert len(residuals) > SHOWERS // 4
    assert max(residuals) <= 1e-9 * ENERGY

    print(f"oracle: {SHOWERS} showers on {min(16, os.cpu_count() or 1)} threads", flush=True)
    primaries = np.concatenate([prim.reshape(1)] * SHOWERS)
    ostats, hist, _ = oracle_lib.em_run_batch_profiles(cfg, tables, atm, primaries, first_shower=SHOWERS + 1,
                                                       threads=min(16, os.cpu_count() or 1))
    ws = cfg.struct.profile_weight_scale
    orc_profile = (hist[:, 1, :] + hist[:, 2, :]) / ws
    orc_xmax = [charged_maximum(hist[k, 1] / ws, hist[k, 2] / ws, width) for k in range(SHOWERS)]
    orc_ground = [s["observations"] for s in ostats]

    ks_xmax = stats.ks_2samp(dev_xmax, orc_xmax)
    ks_ground = stats.ks_2samp(dev_ground, orc_ground)
    print(f"Xmax: device {np.mean(dev_xmax):.1f} +- {np.std(dev_xmax):.1f}, oracle {np.mean(orc_xmax):.1f} "
          f"+- {np.std(orc_xmax):.1f} g/cm^2, KS p = {ks_xmax.pvalue:.3f}; ground particles: device "
          f"{np.mean(dev_ground):.2f}, oracle {np.mean(orc_ground):.2f}, KS p = {ks_ground.pvalue:.3f}", flush=True)
    assert ks_xmax.pvalue > 1e-3
    assert ks_ground.pvalue > 1e-3

    reached = np.array([np.count_nonzero(dev_ground), np.count_nonzero(orc_ground)])
    pooled = reached.sum() / (2 * SHOWERS)
    z = (reached[0] - reached[1]) / SHOWERS / np.sqrt(2 * pooled * (1 - pooled) / SHOWERS)
    print(f"showers reaching the ground: device {reached[0]}, oracle {reached[1]}, z = {z:.2f}", flush=True)
    assert 2 * stats.norm.sf(abs(z)) > 1e-3

    coarse = lambda prof: np.asarray(prof).reshape(SHOWERS, -1, 8).sum(axis=2)
    x, y = coarse(dev_profile), coarse(orc_profile)
    keep = (x.var(axis=0) > 0) & (y.var(axis=0) > 0)
    x, y = x[:, keep], y[:, keep]
    dims = x.shape[1]
    diff = x.mean(axis=0) - y.mean(axis=0)
    pooled_cov = (np.cov(x, rowvar=False) + np.cov(y, rowvar=False)) / 2.0
    t2 = SHOWERS / 2.0 * float(diff @ np.linalg.pinv(pooled_cov) @ diff)
    f_stat = (2 * SHOWERS - dims - 1) / (dims * (2 * SHOWERS - 2)) * t2
    p_profile = stats.f.sf(f_stat, dims, 2 * SHOWERS - dims - 1)
    print(f"mean charged profile: Hotelling T2 {t2:.1f} over {dims} coarse bins, p = {p_profile:.3f}", flush=True)
    assert dims >= 6
    assert p_profile > 1e-3
