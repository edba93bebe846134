"""Format-v7 native tables for air from closed-form cross sections.

The reference exports these tables from PROPOSAL 7.6.2
(src/kokkos/em/proposal/NativeTableExporter.cpp); PROPOSAL is not available
here, so the synthetic workloads use textbook forms of the same processes,
tabulated in exactly the layout the exporter writes:

* photon  pair production   Bethe-Heitler, complete screening, per element:
                            dN/dX dx = (1/X0) (1 - 4/3 x (1 - x)) on [m/E, 1 - m/E],
                            with a threshold factor (1 - 2m/E)^3 / (1 + 10 MeV / E)
          Compton           Klein-Nishina on the air electrons
          photoelectric     Sauter (direct rate model, per element)
* e-+     bremsstrahlung    complete screening, per element:
                            dN/dX dv = (1/X0) (4/3 (1 - v) + v^2) / v, v above the cut
          ionization        Moller (e-) / Bhabha (e+) in the 1/v^2 limit above the cut
          annihilation (e+) Heitler (direct rate model)
          continuous loss   restricted Bethe stopping power below the cut plus
                            bremsstrahlung below the cut; range = integral of 1/(dE/dX)

Each dN/dX column is the bicubic cumulative rate over (E, transformed loss t)
that ProposalNativeQueries.inl:744-1281 integrates and inverts; total rates,
dE/dX and ranges are cubic splines over exponential energy axes.
"""
from __future__ import annotations

import functools
import math

import numpy as np

from . import tables as T

AVOGADRO = 6.02214076e23
ELECTRON_MASS_MEV = 0.5109989461
RE_CM = T.CONSTANTS["classical_electron_radius_cm"]
ALPHA = T.CONSTANTS["fine_structure_constant"]

# Air by mass fraction: (name, Z, A, mass fraction, X0 [g/cm^2], component hash)
AIR = [
    ("N", 7.0, 14.007, 0.7553, 37.99, 0x4E00000000000007),
    ("O", 8.0, 15.999, 0.2318, 34.24, 0x4F00000000000008),
    ("Ar", 18.0, 39.948, 0.0129, 19.55, 0x4172000000000012),
]
AIR_Z_OVER_A = sum(w * z / a for _, z, a, w, _, _ in AIR)
AIR_ELECTRONS_PER_G = AVOGADRO * AIR_Z_OVER_A
AIR_MEAN_EXCITATION_MEV = 85.7e-6
AIR_DENSITY_SEA_LEVEL = 1.205e-3


class Medium:
    """A medium of the closed-form tables: elements by mass fraction, mean
    excitation energy, the density its LPM parameters refer to, molar mass and
    the medium hash written into the table records."""

    def __init__(self, components, mean_excitation_mev, density_g_per_cm3, molar_mass, medium_hash):
        self.components = components
        self.z_over_a = sum(w * z / a for _, z, a, w, _, _ in components)
        self.electrons_per_g = AVOGADRO * self.z_over_a
        self.mean_excitation_mev = mean_excitation_mev
        self.density = density_g_per_cm3
        self.molar_mass = molar_mass
        self.medium_hash = medium_hash

    def x0(self):
        return 1.0 / sum(w / x0 for _, _, _, w, x0, _ in self.components)


AIR_MEDIUM = Medium(AIR, AIR_MEAN_EXCITATION_MEV, AIR_DENSITY_SEA_LEVEL, 28.96, 0xA1)
# Ice (H2O) by mass fraction, I = 75 eV, 0.92 g/cm^3
ICE = [
    ("H", 1.0, 1.008, 0.1119, 63.04, 0x4800000000000001),
    ("O", 8.0, 15.999, 0.8881, 34.24, 0x4F00000000000008),
]
ICE_MEDIUM = Medium(ICE, 75.0e-6, 0.92, 18.015, 0xA2)


def _x0_air():
    return AIR_MEDIUM.x0()


def pair_dndx_scale(e):
    """Threshold factor of the complete-screening pair rate."""
    e = np.asarray(e, dtype=np.float64)
    r = 2.0 * ELECTRON_MASS_MEV / e
    out = np.where(r < 1.0, np.clip(1.0 - r, 0.0, None) ** 3 / (1.0 + 10.0 / e), 0.0)
    return out


def pair_cdf(x):
    """Integral of 1 - 4/3 x (1 - x) from 0 to x."""
    return x - (2.0 / 3.0) * x * x + (4.0 / 9.0) * x * x * x


def brems_cdf(v, vc):
    """Integral of (4/3 (1 - v) + v^2) / v from vc to v."""
    v = np.maximum(v, vc)
    return (4.0 / 3.0) * np.log(v / vc) - (4.0 / 3.0) * (v - vc) + 0.5 * (v * v - vc * vc)


def klein_nishina_dsigma_dv(e, v):
    """d(sigma)/dv per electron, cm^2, v = fraction of E given to the electron."""
    k = e / ELECTRON_MASS_MEV
    eps = 1.0 - v
    eps = np.clip(eps, 1.0 / (1.0 + 2.0 * k), 1.0)
    cos_t = 1.0 - (1.0 - eps) / (k * eps)
    sin2 = np.clip(1.0 - cos_t * cos_t, 0.0, None)
    return math.pi * RE_CM * RE_CM / k * (1.0 / eps + eps) * (1.0 - eps * sin2 / (1.0 + eps * eps))


def restricted_dedx(e, cut_mev, medium=AIR_MEDIUM):
    """Continuous loss (MeV cm^2/g) of an e+- of total energy e below the cut."""
    e = np.asarray(e, dtype=np.float64)
    gamma = e / ELECTRON_MASS_MEV
    beta2 = np.clip(1.0 - 1.0 / (gamma * gamma), 1e-12, None)
    kinetic = np.clip(e - ELECTRON_MASS_MEV, 1e-12, None)
    t_up = np.minimum(cut_mev, 0.5 * kinetic)
    k_half = 0.5 * 0.307075 * medium.z_over_a  # K/2 Z/A, MeV cm^2/g
    arg = 2.0 * ELECTRON_MASS_MEV * beta2 * gamma * gamma * t_up / (medium.mean_excitation_mev ** 2)
    ionization = k_half / beta2 * (np.log(np.clip(arg, 1.0 + 1e-12, None)) - beta2)
    vc = np.minimum(cut_mev / e, 1.0)
    brems = e / medium.x0() * ((4.0 / 3.0) * vc - (2.0 / 3.0) * vc * vc + vc ** 3 / 3.0)
    return np.clip(ionization, 0.05, None) + brems


def _limits_brems(e, cut):
    vmax = np.clip(1.0 - ELECTRON_MASS_MEV / e, 0.0, None)
    vc = np.clip(np.minimum(cut / e, 1.0), 0.0, vmax)
    return vc, vmax


def _limits_moller(e, cut, bhabha):
    vmax = np.clip((1.0 if bhabha else 0.5) * (1.0 - ELECTRON_MASS_MEV / e), 0.0, None)
    vc = np.clip(np.minimum(cut / e, 1.0), 0.0, vmax)
    return vc, vmax


def _rellog(vc, vmax, t):
    with np.errstate(divide="ignore", invalid="ignore"):
        out = vc * np.exp(t * np.log(vmax / vc))
    return np.where(vmax > vc, out, vc)


def _compton_cdf_function(nodes_per_unit=4000, medium=AIR_MEDIUM):
    """Cumulative Klein-Nishina rate in the one-minus-log loss coordinate."""
    tt = np.linspace(0.0, 1.0, nodes_per_unit + 1)

    def f(e, t):
        e = np.asarray(e, dtype=np.float64)
        t = np.asarray(t, dtype=np.float64)
        eb, tb = np.broadcast_arrays(e, t)
        out = np.empty(eb.shape)
        flat_e, flat_t, flat_o = eb.reshape(-1), tb.reshape(-1), out.reshape(-1)
        for ev in np.unique(flat_e):
            vmax = 1.0 - 1.0 / (1.0 + 2.0 * ev / ELECTRON_MASS_MEV)
            xi = math.log(1.0 / (1.0 - vmax))
            v = 1.0 - (1.0 - vmax) * np.exp((1.0 - tt) * xi)
            dvdt = xi * (1.0 - v)
            g = klein_nishina_dsigma_dv(ev, v) * dvdt
            cum = np.concatenate(([0.0], np.cumsum(0.5 * (g[1:] + g[:-1]) * np.diff(tt))))
            m = flat_e == ev
            flat_o[m] = np.interp(np.clip(flat_t[m], 0.0, 1.0), tt, cum) * medium.electrons_per_g
        return out

    return f


def build_air_tables(cut_mev: float = 1.0, e_max_mev: float = 2.0e9, energy_nodes: int = 160,
                     loss_nodes: int = 64, polynomial: bool = True, muon_pair: float = 0.0):
    """Air tables for photons and e+- with an e+- energy-loss cut of `cut_mev`
    (stochastic losses above it, continuous below); `muon_pair` > 0 adds the
    photon mu-pair columns at that multiple of their rate."""
    return build_tables(AIR_MEDIUM, cut_mev, e_max_mev, energy_nodes, loss_nodes, polynomial, muon_pair)


def build_ice_tables(cut_mev: float = 0.1, e_max_mev: float = 2.0e9, energy_nodes: int = 160,
                     loss_nodes: int = 64, polynomial: bool = True):
    """Ice (H2O, 0.92 g/cm^3) tables, by default with the 100 keV cut of the
    in-ice cascade."""
    return build_tables(ICE_MEDIUM, cut_mev, e_max_mev, energy_nodes, loss_nodes, polynomial)


def muon_pair_threshold_mev():
    """Lowest photon energy of the native mu-pair split: 2 m_mu sqrt(e)."""
    return 2.0 * T.CONSTANTS["muon_mass_mev"] * T.CONSTANTS["sqrt_e"]


def muon_pair_limits(e):
    """The split's x = E(mu+)/E range at photon energy e (MeV): the mu-pair limits
    of the native loss-fraction query."""
    width = np.sqrt(np.clip(1.0 - muon_pair_threshold_mev() / np.asarray(e, dtype=np.float64), 0.0, None))
    return 0.5 * (1.0 - width), 0.5 * (1.0 + width)


def muon_pair_scale(e):
    """Threshold factor of the synthetic mu-pair rate: zero at the split's
    threshold, the pair rate's (m_e/m_mu)^2 far above it."""
    e = np.asarray(e, dtype=np.float64)
    r = muon_pair_threshold_mev() / e
    ratio = (ELECTRON_MASS_MEV / T.CONSTANTS["muon_mass_mev"]) ** 2
    return np.where(r < 1.0, ratio * np.clip(1.0 - r, 0.0, None) ** 3, 0.0)


def build_tables(medium, cut_mev, e_max_mev=2.0e9, energy_nodes=160, loss_nodes=64, polynomial=True,
                 muon_pair=0.0):
    """Tables of `medium` for photons and e+-. Below a 0.5 MeV cut the e+-
    process tables start at 0.55 MeV total energy (else 1 MeV), so every
    tracked e+- above the cut lies on them. `muon_pair` > 0: photon mu-pair
    columns per element (the pair rate's shape in x over the mu-pair limits,
    (m_e/m_mu)^2 of its size, times `muon_pair`)."""
    elements = medium.components
    electrons_per_g = medium.electrons_per_g
    lepton_low = 1.0 if cut_mev >= 0.5 else 0.55
    b = T.TableBuilder(constants=dict(T.CONSTANTS, electron_mass_mev=ELECTRON_MASS_MEV))
    me = ELECTRON_MASS_MEV
    lin = T.axis(T.AXIS_LINEAR, 0.0, 1.0, loss_nodes)
    photon_low = 0.05
    ax_pair = T.axis(T.AXIS_EXP, 1.1, e_max_mev, energy_nodes)
    ax_photon = T.axis(T.AXIS_EXP, photon_low, e_max_mev, energy_nodes + 20)
    ax_lepton = T.axis(T.AXIS_EXP, lepton_low, e_max_mev, energy_nodes)
    kin = T.KIN
    proc = T.PROCESS
    ordinal = {22: 0, 11: 0, -11: 0}

    def add(pdg, **fields):
        b.add_dndx(pdg_id=pdg, selection_ordinal=ordinal[pdg], medium_hash=medium.medium_hash,
                   medium_mean_excitation_energy_mev=medium.mean_excitation_mev, **fields)
        ordinal[pdg] += 1

    # photon pair production, per element
    for name, z, a, w, x0, h in elements:
        def f(e, t, w=w, x0=x0):
            xmin = me / e
            xmax = 1.0 - xmin
            x = _rellog(xmin, xmax, t)
            norm = pair_cdf(xmax) - pair_cdf(xmin)
            frac = np.where(norm > 0, (pair_cdf(x) - pair_cdf(xmin)) / np.where(norm > 0, norm, 1), 0)
            return w * (w / x0) * (7.0 / 9.0) * pair_dndx_scale(e) * frac
        add(22, process_id=proc["photon_pair"], component_hash=h, kinematic_model=kin["photon_pair"],
            loss_transform=T.LOSS_RELLOG, rate_model=T.RATE_BICUBIC,
            spline=b.bicubic(ax_pair, lin, f, polynomial=polynomial), component_weight=w,
            particle_mass_mev=0.0, component_nuclear_charge=z, component_atomic_mass=a,
            component_average_nucleon_mass_mev=931.494, lower_energy_limit_mev=1.1)
    mu_low = 1.001 * muon_pair_threshold_mev()
    if muon_pair > 0.0:
        ax_mu = T.axis(T.AXIS_EXP, mu_low, e_max_mev, energy_nodes)
        for name, z, a, w, x0, h in elements:
            def fm(e, t, w=w, x0=x0):
                xmin, xmax = muon_pair_limits(e)
                x = _rellog(xmin, xmax, t)
                norm = pair_cdf(xmax) - pair_cdf(xmin)
                frac = np.where(norm > 0, (pair_cdf(x) - pair_cdf(xmin)) / np.where(norm > 0, norm, 1), 0)
                return w * (w / x0) * (7.0 / 9.0) * muon_pair * muon_pair_scale(e) * frac
            add(22, process_id=proc["photon_muon_pair"], component_hash=h,
                kinematic_model=kin["only_stochastic"], loss_transform=T.LOSS_RELLOG,
                rate_model=T.RATE_BICUBIC, spline=b.bicubic(ax_mu, lin, fm, polynomial=polynomial),
                component_weight=w, particle_mass_mev=0.0, component_nuclear_charge=z,
                component_atomic_mass=a, component_average_nucleon_mass_mev=931.494,
                lower_energy_limit_mev=mu_low)
    # Compton on the medium electrons
    add(22, process_id=proc["compton"], component_hash=0, kinematic_model=kin["compton"],
        loss_transform=T.LOSS_ONEMINUSLOG, rate_model=T.RATE_BICUBIC,
        spline=b.bicubic(ax_photon, lin, _compton_cdf_function(medium=medium), polynomial=polynomial),
        component_weight=1.0, particle_mass_mev=0.0, lower_energy_limit_mev=photon_low)
    # photoelectric, per element (Sauter direct rate)
    for name, z, a, w, x0, h in elements:
        add(22, process_id=proc["photoelectric"], component_hash=h, kinematic_model=kin["only_stochastic"],
            loss_transform=T.LOSS_RELLOG, rate_model=T.RATE_SAUTER, component_weight=w,
            particle_mass_mev=0.0, component_nuclear_charge=z, component_atomic_mass=a,
            direct_rate_factor=AVOGADRO * w / a, lower_energy_limit_mev=1e-3)

    for pdg in (11, -11):
        for name, z, a, w, x0, h in elements:
            def f(e, t, w=w, x0=x0):
                vc, vmax = _limits_brems(e, cut_mev)
                v = _rellog(vc, vmax, t)
                return w * (w / x0) * np.where(vmax > vc, brems_cdf(v, np.where(vc > 0, vc, 1.0)), 0.0)
            add(pdg, process_id=proc["brems"], component_hash=h, kinematic_model=kin["brems_screening"],
                loss_transform=T.LOSS_RELLOG, rate_model=T.RATE_BICUBIC,
                spline=b.bicubic(ax_lepton, lin, f, polynomial=polynomial), component_weight=w,
                particle_mass_mev=me, component_nuclear_charge=z, component_atomic_mass=a,
                absolute_cut_mev=cut_mev, relative_cut=1.0, lower_energy_limit_mev=lepton_low)
        bhabha = pdg == -11

        def g(e, t, bhabha=bhabha):
            vc, vmax = _limits_moller(e, cut_mev, bhabha)
            v = _rellog(vc, vmax, t)
            k = 2.0 * math.pi * RE_CM * RE_CM * me * electrons_per_g
            with np.errstate(divide="ignore", invalid="ignore"):
                val = k / e * (1.0 / vc - 1.0 / v)
            return np.where(vmax > vc, val, 0.0)
        add(pdg, process_id=proc["ionization"], component_hash=0,
            kinematic_model=kin["bhabha" if bhabha else "moller"], loss_transform=T.LOSS_RELLOG,
            rate_model=T.RATE_BICUBIC, spline=b.bicubic(ax_lepton, lin, g, polynomial=polynomial),
            component_weight=1.0, particle_mass_mev=me, absolute_cut_mev=cut_mev, relative_cut=1.0,
            lower_energy_limit_mev=lepton_low)
        if bhabha:
            add(pdg, process_id=proc["annihilation"], component_hash=0,
                kinematic_model=kin["only_stochastic"], loss_transform=T.LOSS_RELLOG,
                rate_model=T.RATE_HEITLER, component_weight=1.0, particle_mass_mev=me,
                direct_rate_factor=math.pi * RE_CM * RE_CM * electrons_per_g,
                lower_energy_limit_mev=me)

    # total rates: sum of the analytic column rates at the nodes
    def photon_total(e):
        e = np.asarray(e, dtype=np.float64)
        pair = sum((w / x0) * (7.0 / 9.0) for _, _, _, w, x0, _ in elements) * pair_dndx_scale(e)
        pair = np.where(e >= 1.1, pair, 0.0)
        compton = _compton_cdf_function(medium=medium)(e, np.ones_like(e))
        photo = sum(_sauter_rate(e, z, AVOGADRO * w / a) for _, z, a, w, _, _ in elements)
        mu = 0.0
        if muon_pair > 0.0:
            mu = sum((w / x0) * (7.0 / 9.0) for _, _, _, w, x0, _ in elements) * muon_pair * muon_pair_scale(e)
            mu = np.where(e >= mu_low, mu, 0.0)
        return pair + compton + photo + mu

    def lepton_total(e, positron):
        e = np.asarray(e, dtype=np.float64)
        out = np.zeros_like(e)
        for _, _, _, w, x0, _ in elements:
            vc, vmax = _limits_brems(e, cut_mev)
            out += (w / x0) * np.where(vmax > vc, brems_cdf(vmax, np.where(vc > 0, vc, 1.0)), 0.0)
        vc, vmax = _limits_moller(e, cut_mev, positron)
        k = 2.0 * math.pi * RE_CM * RE_CM * me * electrons_per_g
        with np.errstate(divide="ignore", invalid="ignore"):
            out += np.where(vmax > vc, k / e * (1.0 / vc - 1.0 / vmax), 0.0)
        if positron:
            out += _heitler_rate(e, math.pi * RE_CM * RE_CM * electrons_per_g)
        out = np.where(e >= lepton_low, out, 0.0)
        return out

    tax_photon = T.axis(T.AXIS_EXP, 0.01, e_max_mev, 400)
    tax_lepton = T.axis(T.AXIS_EXP, 0.6, e_max_mev, 400)
    b.add_total(22, b.cubic(tax_photon, photon_total), lower=0.0)
    b.add_total(11, b.cubic(tax_lepton, lambda e: lepton_total(e, False)), lower=0.0)
    b.add_total(-11, b.cubic(tax_lepton, lambda e: lepton_total(e, True)), lower=0.0)

    # continuous loss and range (utility) of e+-
    ulow = 0.6
    grid = np.exp(np.linspace(math.log(ulow), math.log(e_max_mev), 20001))
    inv = 1.0 / restricted_dedx(grid, cut_mev, medium)
    rng = np.concatenate(([0.0], np.cumsum(0.5 * (inv[1:] + inv[:-1]) * np.diff(grid))))
    lg = np.log(grid)

    def range_of(e):
        e = np.asarray(e, dtype=np.float64)
        return np.interp(np.log(np.clip(e, ulow, e_max_mev)), lg, rng)

    uax = T.axis(T.AXIS_EXP, ulow, e_max_mev, 600)
    for k, pdg in enumerate((11, -11)):
        b.add_dedx(pdg, proc["ionization"], b.cubic(uax, lambda e: restricted_dedx(e, cut_mev, medium)),
                   lower=ulow, ordinal=k)
        b.add_utility(pdg, b.cubic(uax, range_of), mass=me, lower=ulow)
    return b.build()


def _sauter_rate(e, z, factor):
    """Sauter photo-effect rate (cm^2/g), the RATE_SAUTER model of the queries."""
    e = np.asarray(e, dtype=np.float64)
    me = ELECTRON_MASS_MEV
    ion = z * z * ALPHA * ALPHA * me / 2.0
    gamma = 1.0 + (e - ion) / me
    gb = np.sqrt(np.clip(gamma * gamma - 1.0, 1e-300, None))
    sigma_t = 8.0 * math.pi * RE_CM * RE_CM / 3.0
    sigma = 1.5 * sigma_t * z ** 5 * ALPHA ** 4 * (me / e) ** 5 * gb ** 3 * (
        4.0 / 3.0 + gamma * (gamma - 2.0) / (gamma + 1.0) * (1.0 - np.log(gamma + gb) / (gamma * gb)))
    zab = z * ALPHA * gamma / gb
    zpab = math.pi * zab
    corr = np.where(zpab < 10.0,
                    (1 + zab * zab) * zpab / np.sinh(np.minimum(zpab, 700)) * np.exp(zab * (math.pi - 4 * np.arctan(1 / zab))),
                    (1 + zab * zab) * 2 * zpab * np.exp(-4 * zab * np.arctan(1 / zab)))
    lz = math.log(z)
    shell = 1.0 + 0.01481 * lz * lz - 0.000788 * lz ** 3
    out = sigma * corr * shell * factor
    return np.where(e > 1e-3, out, 0.0)


def _heitler_rate(e, factor):
    e = np.asarray(e, dtype=np.float64)
    g = e / ELECTRON_MASS_MEV
    with np.errstate(invalid="ignore", divide="ignore"):
        root = np.sqrt(np.clip(g * g - 1.0, 1e-300, None))
        aux = (g * g + 4 * g + 1) / (g * g - 1) * np.log(g + root) - (g + 3) / root
        out = aux * factor / (g + 1)
    return np.where(g > 1.0, out, 0.0)


def photon_pair_lpm_snapshot(medium=AIR_MEDIUM):
    """Photon-pair LPM parameters (PROPOSAL's PhotoPairLPM inputs) of the medium
    at its reference density (sea-level air by default)."""
    x0_cm = medium.x0() / medium.density
    e_lpm = 7.7e6 * x0_cm  # MeV: 7.7 TeV per cm of radiation length
    mol_mass = medium.molar_mass
    return dict(baseline_mass_density_g_per_cm3=medium.density,
                molecular_density_per_cm3=medium.density * AVOGADRO / mol_mass,
                sum_charge=medium.z_over_a * mol_mass, e_lpm_mev=e_lpm,
                classical_electron_radius_cm=RE_CM, fine_structure_constant=ALPHA,
                components=[(h, z, 183.0) for _, z, _, _, _, h in medium.components])


def brems_lpm_snapshot(medium=AIR_MEDIUM):
    s = photon_pair_lpm_snapshot(medium)
    return dict(baseline_mass_density_g_per_cm3=s["baseline_mass_density_g_per_cm3"],
                molecular_density_per_cm3=s["molecular_density_per_cm3"], sum_charge=s["sum_charge"],
                e_lpm_mev=s["e_lpm_mev"], lepton_mass_mev=0.5109989461,
                electron_mass_mev=0.5109989461, muon_mass_mev=105.6583745,
                classical_electron_radius_cm=RE_CM, fine_structure_constant=ALPHA,
                components=[(h, z, a, 183.0) for _, z, a, _, _, h in medium.components])


# ---- Molière multiple scattering (the MoliereSnapshot of the air medium) ----

HBAR_MEV_S = 6.582119569e-22
SPEED_OF_LIGHT_CM_PER_S = 2.99792458e10
EULER_MASCHERONI = 0.57721566490153286061


@functools.lru_cache(maxsize=None)
def moliere_coefficients():
    """Coefficients of Molière's projected-angle correction functions in
    x = theta^2 / (chi_c^2 B), with f_n(x) = sqrt(pi) g_n(sqrt(x)) and
    g_n(t) = 1/(n! pi) int_0^inf cos(t y) exp(-y^2/4) ((y^2/4) ln(y^2/4))^n dy.

    Small x (power series): c1[k] = (-4)^k Gamma(k+3/2) psi(k+3/2) / (sqrt(pi) (2k)!),
    c2[k] = (-4)^k Gamma(k+5/2) (psi(k+5/2)^2 + psi'(k+5/2)) / (2 sqrt(pi) (2k)!).
    Large x (from the |y|^a ln|y| singularities of the integrand at y = 0):
    f2 ~ sum_i c2l[i] (ln(x)/2 + s2l[i]) x^-(i+1/2), c2l[i] = sqrt(pi) (2i)! / (8 4^(i-2) (i-2)!),
    s2l[i] = ln 2 - psi(2i+1), i = 2..12; the integral of f1 from sqrt(x) to infinity
    gives int_0^sqrt(x) f1 = sum_i c1il[i] x^-i, c1il[i] = -sqrt(pi) (2i-1)! / (4^i (i-1)!).
    Computed at 60 digits and rounded to double."""
    import mpmath as mp

    with mp.workdps(60):
        sq = mp.sqrt(mp.pi)
        c1 = [float((-4) ** k * mp.gamma(k + mp.mpf(3) / 2) * mp.digamma(k + mp.mpf(3) / 2) /
                    (sq * mp.factorial(2 * k))) for k in range(70)]
        c2 = []
        for k in range(70):
            a = k + mp.mpf(5) / 2
            c2.append(float((-4) ** k * mp.gamma(a) * (mp.digamma(a) ** 2 + mp.polygamma(1, a)) /
                            (2 * sq * mp.factorial(2 * k))))
        c2l = [0.0, 0.0] + [float(sq * mp.factorial(2 * i) / (8 * mp.mpf(4) ** (i - 2) * mp.factorial(i - 2)))
                            for i in range(2, 13)]
        s2l = [0.0, 0.0] + [float(mp.log(2) - mp.digamma(2 * i + 1)) for i in range(2, 13)]
        c1il = [0.0] + [float(-sq * mp.factorial(2 * i - 1) / (mp.mpf(4) ** i * mp.factorial(i - 1)))
                        for i in range(1, 15)]
    return tuple(c1), tuple(c2), tuple(c2l), tuple(s2l), tuple(c1il)


def moliere_snapshot(particle_mass_mev=ELECTRON_MASS_MEV, medium=AIR_MEDIUM):
    """Electron Molière snapshot of the medium (air by default). Component fields are derived
    as the reference's auxiliary export derives them from a medium
    (ProposalNativeAux.cpp:99-135): mass weights, weight * Z^2, the mean
    Z(Z+1)/A for electrons, chi_0 = m_e alpha (128 Z / (9 pi^2))^(1/3) and the
    Coulomb correction 3.76 (alpha Z)^2."""
    c1, c2, c2l, s2l, c1il = moliere_coefficients()
    components = []
    mean_a = 0.0
    mean_zz = 0.0
    weight_zz_sum = 0.0
    weights = []
    for _, z, a, w, _, _ in medium.components:
        weights.append(w)
        mean_a += w * a
        wzz = w * z * z
        weight_zz_sum += wzz
        mean_zz += w * z * (z + 1.0) if particle_mass_mev == ELECTRON_MASS_MEV else wzz
        chi0 = ELECTRON_MASS_MEV * ALPHA * (z * 128.0 / (9.0 * math.pi * math.pi)) ** (1.0 / 3.0)
        components.append((float(z), wzz, chi0 * chi0, 3.76 * ALPHA * ALPHA * z * z))
    maximum = 0
    for k in range(len(weights) - 1):
        if weights[k + 1] > weights[k]:
            maximum = k + 1
    return dict(particle_mass_mev=particle_mass_mev, electron_mass_mev=ELECTRON_MASS_MEV,
                fine_structure_constant=ALPHA, avogadro_per_mol=AVOGADRO, hbar_mev_s=HBAR_MEV_S,
                speed_of_light_cm_per_s=SPEED_OF_LIGHT_CM_PER_S, z_squared_over_a_average=mean_zz / mean_a,
                inverse_weight_zz_sum=1.0 / weight_zz_sum, euler_mascheroni=EULER_MASCHERONI,
                maximum_weight_index=maximum, components=components, c1=c1, c2=c2, c2_large=c2l,
                s2_large=s2l, c1_integral_large=c1il)


class IceMedium:
    """The ice medium's snapshots in the form make_config takes (this module's
    functions with medium=ICE_MEDIUM); tables: build_ice_tables()."""

    @staticmethod
    def moliere_snapshot(particle_mass_mev=ELECTRON_MASS_MEV):
        return moliere_snapshot(particle_mass_mev, ICE_MEDIUM)

    @staticmethod
    def photon_pair_lpm_snapshot():
        return photon_pair_lpm_snapshot(ICE_MEDIUM)

    @staticmethod
    def brems_lpm_snapshot():
        return brems_lpm_snapshot(ICE_MEDIUM)
