"""ctypes layer of the EM backend (include/c8mi/c8mi_em.h) and the host-side
builders of its inputs: the layered spherical atmosphere, the profile
projection along the shower axis and the backend configuration.

The device path is libc8mi.so; this module marshals arguments and decodes
results only.
"""
from __future__ import annotations

import ctypes
import math
from ctypes import POINTER, byref, c_double, c_int, c_int32, c_uint32, c_uint64, c_void_p

import numpy as np

from . import (PARTICLE_STATE_DTYPE, RadioCounters as _RadioCounters, _check, library)

EARTH_RADIUS_M = 6371000.0
C_LIGHT = 299792458.0

# USStdBK (corsika/media/CORSIKA7Atmospheres.hpp:182-186): upper altitude [m],
# b [g/cm^2], c [cm]; the last layer is linear in the overburden (homogeneous).
USSTDBK = [(7000.0, 1183.6071, 954248.34), (11400.0, 1143.0425, 800005.34),
           (37000.0, 1322.9748, 629568.93), (100000.0, 655.67307, 737521.77),
           (112800.0, 1.0, 1e9)]

DENSITY_HOMOGENEOUS, DENSITY_SLIDING, DENSITY_FLAT = 1, 2, 3
SHAPE_UNIVERSE, SHAPE_SPHERE = 1, 2


class EnvModel(ctypes.Structure):
    _fields_ = [("density_kind", c_uint32), ("medium_id", c_int32),
                ("reference_density_g_per_cm3", c_double), ("scale_length_m", c_double),
                ("inverse_scale_length_per_m", c_double), ("reference_height_m", c_double),
                ("anchor_m", c_double * 3), ("axis", c_double * 3),
                ("magnetic_field_t", c_double * 3)]


class EnvNode(ctypes.Structure):
    _fields_ = [("shape_kind", c_uint32), ("parent", c_int32), ("model", c_int32),
                ("child_begin", c_uint32), ("child_count", c_uint32), ("pad_", c_uint32),
                ("center_m", c_double * 3), ("radius_m", c_double)]


class Environment(ctypes.Structure):
    _fields_ = [("nodes", POINTER(EnvNode)), ("node_count", c_uint32),
                ("child_index_count", c_uint32), ("children", POINTER(c_int32)),
                ("models", POINTER(EnvModel)), ("model_count", c_uint32),
                ("observation_enabled", c_uint32), ("observation_point_m", c_double * 3),
                ("observation_normal", c_double * 3), ("earth_center_m", c_double * 3),
                ("maximum_magnetic_deflection_rad", c_double)]


class PairLpmComponent(ctypes.Structure):
    _fields_ = [("component_hash", c_uint64), ("nuclear_charge", c_double),
                ("radiation_log_constant", c_double)]


class PairLpm(ctypes.Structure):
    _fields_ = [("baseline_mass_density_g_per_cm3", c_double), ("molecular_density_per_cm3", c_double),
                ("sum_charge", c_double), ("e_lpm_mev", c_double),
                ("classical_electron_radius_cm", c_double), ("fine_structure_constant", c_double),
                ("component_count", c_uint32), ("reserved", c_uint32),
                ("components", PairLpmComponent * 16)]


class BremsLpmComponent(ctypes.Structure):
    _fields_ = [("component_hash", c_uint64), ("nuclear_charge", c_double),
                ("atomic_mass_number", c_double), ("radiation_log_constant", c_double)]


class BremsLpm(ctypes.Structure):
    _fields_ = [("baseline_mass_density_g_per_cm3", c_double), ("molecular_density_per_cm3", c_double),
                ("sum_charge", c_double), ("e_lpm_mev", c_double), ("lepton_mass_mev", c_double),
                ("electron_mass_mev", c_double), ("muon_mass_mev", c_double),
                ("classical_electron_radius_cm", c_double), ("fine_structure_constant", c_double),
                ("component_count", c_uint32), ("reserved", c_uint32),
                ("components", BremsLpmComponent * 16),
                ("muon_baseline_mass_density_g_per_cm3", c_double),
                ("muon_molecular_density_per_cm3", c_double), ("muon_sum_charge", c_double),
                ("muon_e_lpm_mev", c_double)]


class Thinning(ctypes.Structure):
    _fields_ = [("threshold_gev", c_double), ("maximum_weight", c_double), ("enabled", c_uint32),
                ("erase_zero_weight", c_uint32)]


class MoliereComponent(ctypes.Structure):
    _fields_ = [("nuclear_charge", c_double), ("weight_zz", c_double), ("chi0_squared_mev2", c_double),
                ("coulomb_correction", c_double)]


class MoliereSnapshot(ctypes.Structure):
    _fields_ = [(k, c_double) for k in ("particle_mass_mev", "electron_mass_mev", "fine_structure_constant",
                                        "avogadro_per_mol", "hbar_mev_s", "speed_of_light_cm_per_s",
                                        "z_squared_over_a_average", "inverse_weight_zz_sum",
                                        "euler_mascheroni")] + [
        ("component_count", c_uint32), ("maximum_weight_index", c_uint32),
        ("components", MoliereComponent * 16), ("c1", c_double * 70), ("c2", c_double * 70),
        ("c2_large", c_double * 13), ("s2_large", c_double * 13), ("c1_integral_large", c_double * 15)]

    @classmethod
    def from_dict(cls, d):
        m = cls()
        for k in ("particle_mass_mev", "electron_mass_mev", "fine_structure_constant", "avogadro_per_mol",
                  "hbar_mev_s", "speed_of_light_cm_per_s", "z_squared_over_a_average", "inverse_weight_zz_sum",
                  "euler_mascheroni", "maximum_weight_index"):
            setattr(m, k, d[k])
        m.component_count = len(d["components"])
        for i, comp in enumerate(d["components"]):
            m.components[i] = MoliereComponent(*comp)
        for k in ("c1", "c2", "c2_large", "s2_large", "c1_integral_large"):
            arr = getattr(m, k)
            for i, v in enumerate(d[k]):
                arr[i] = v
        return m


class MoliereCubic(ctypes.Structure):
    _fields_ = [("c", c_double * 4)]


MOLIERE_INTERVALS = 99
MOLIERE_GUESS_SHAPE = (64, 8, 128)


class MoliereInterpolation(ctypes.Structure):
    """c8mi_moliere_interpolation (the reference's MoliereInterpolationTable)."""
    _fields_ = [("polynomials", MoliereCubic * (4 * MOLIERE_INTERVALS)),
                ("initial_guess_delta", c_double * (64 * 8 * 128))]

    @classmethod
    def build(cls, snapshot):
        """The table of a MoliereSnapshot (c8mi_em_moliere_interpolation_build)."""
        t = cls()
        _check(library().c8mi_em_moliere_interpolation_build(byref(snapshot), byref(t)))
        return t

    @classmethod
    def read(cls, path, snapshot):
        t = cls()
        _check(library().c8mi_em_moliere_interpolation_read(str(path).encode(), byref(snapshot), byref(t)))
        return t

    def write(self, path, snapshot):
        _check(library().c8mi_em_moliere_interpolation_write(str(path).encode(), byref(snapshot), byref(self)))

    def polynomial_array(self):
        """[4][99][4] coefficients: f1, f2, integral f1, integral f2."""
        return np.ctypeslib.as_array(self.polynomials).view(np.float64).reshape(4, MOLIERE_INTERVALS, 4).copy()

    def guess_array(self):
        return np.ctypeslib.as_array(self.initial_guess_delta).reshape(MOLIERE_GUESS_SHAPE).copy()


_MOLIERE_TABLES = {}


def air_moliere_interpolation(tables_module):
    """The interpolation table of the air electron snapshot, built once."""
    key = tables_module.__name__
    if key not in _MOLIERE_TABLES:
        snap = MoliereSnapshot.from_dict(tables_module.moliere_snapshot())
        _MOLIERE_TABLES[key] = MoliereInterpolation.build(snap)
    return _MOLIERE_TABLES[key]


class EmConfig(ctypes.Structure):
    _fields_ = [("random_seed", c_uint64), ("em_transport_cut_mev", c_double),
                ("muon_transport_cut_mev", c_double), ("queue_capacity", c_uint64),
                ("observation_capacity", c_uint64), ("fallback_capacity", c_uint64),
                ("photon_pair_lpm", PairLpm), ("brems_lpm", BremsLpm), ("thinning", Thinning),
                ("electron_moliere", MoliereSnapshot), ("moliere_enabled", c_uint32),
                ("moliere_fast_iterations", c_uint32),
                ("electron_moliere_interpolation", POINTER(MoliereInterpolation)),
                ("profile_bins", c_uint32), ("profile_crossing_mode", c_uint32),
                ("profile_bin_width_g_per_cm2", c_double),
                ("profile_energy_loss_threshold_g_per_cm2", c_double),
                ("profile_weight_scale", c_double), ("profile_energy_scale", c_double),
                ("axis_start_m", c_double * 3), ("axis_direction", c_double * 3),
                ("axis_step_m", c_double), ("axis_grammage_g_per_cm2", POINTER(c_double)),
                ("axis_support_count", c_uint64), ("shower_slots", c_uint32), ("reserved_", c_uint32),
                ("minimum_batch_size", c_uint64), ("resident_minimum_batch", c_uint64)]


class ShowerAxis(ctypes.Structure):
    _fields_ = [("start_m", c_double * 3), ("direction", c_double * 3), ("step_m", c_double),
                ("grammage_g_per_cm2", POINTER(c_double)), ("support_count", c_uint64)]


class SlotState(ctypes.Structure):
    _fields_ = [(k, c_uint64) for k in (
        "shower_id", "pending_photons", "pending_leptons", "next_history_id", "photon_fronts",
        "lepton_fronts", "photon_steps", "lepton_steps", "photon_interactions", "lepton_interactions",
        "secondaries", "observations", "fallbacks")] + [("active", c_uint32), ("pad_", c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "pad_"}


class _ParticleState(ctypes.Structure):
    # c8mi_particle_state as an opaque 128-byte, 16-aligned record
    _fields_ = [("words", c_uint64 * 16)]


class Request(ctypes.Structure):
    _fields_ = [("primary", _ParticleState), ("shower_id", c_uint64), ("first_history_id", c_uint64),
                ("arrival_s", c_double), ("footprint_particles", c_double), ("axis", ShowerAxis),
                ("radio_observers", c_void_p), ("radio_observer_count", c_uint64),
                ("radio_coreas", c_void_p), ("radio_zhs", c_void_p),
                ("particles", c_void_p), ("particle_count", c_uint64),
                ("align_", c_uint64)]  # the C struct is 16-byte aligned


class ServeOptions(ctypes.Structure):
    _fields_ = [("footprint_budget_particles", c_double), ("realtime", c_uint32), ("pad_", c_uint32)]


class RunResult(ctypes.Structure):
    _fields_ = [(k, c_uint64) for k in (
        "photon_fronts", "lepton_fronts", "photon_steps", "lepton_steps", "photon_interactions",
        "lepton_interactions", "secondaries", "observations", "fallbacks", "pending_photons",
        "pending_leptons", "next_history_id", "max_front", "photon_calls", "lepton_calls",
        "bucketed_fronts", "bucketed_particles", "below_minimum_batch_checkpoints",
        "history_range_checkpoints", "reserved_history_ids")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class WavefrontResult(ctypes.Structure):
    _fields_ = [(k, c_uint64) for k in (
        "input_particles", "wavefronts", "transport_records", "interactions", "secondaries",
        "next_secondary_history_id", "peak_resident", "remaining", "other_species_pending",
        "bucketed_fronts", "bucketed_particles", "small_fronts", "small_particles", "observations",
        "fallbacks")] + [(k, c_uint32) for k in ("completed", "below_minimum_batch_checkpoint",
                                                 "history_range_exhausted", "pad_")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "pad_"}


class Limits(ctypes.Structure):
    _fields_ = [(k, c_uint64) for k in (
        "minimum_batch_size", "maximum_resident_photon_batch", "maximum_resident_lepton_batch",
        "maximum_resident_input_batch", "queue_capacity")] + [("shower_slots", c_uint32), ("pad_", c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "pad_"}


class Statistics(ctypes.Structure):
    _fields_ = [(k, c_uint64) for k in (
        "shower_ordinal", "wavefronts", "photon_wavefronts", "lepton_wavefronts", "particles_enqueued",
        "particles_advanced", "particles_produced", "interactions_selected", "bucketed_fronts",
        "bucketed_particles", "small_fronts", "small_particles", "ordered_fronts", "photon_calls",
        "lepton_calls", "below_minimum_batch_checkpoints", "history_range_checkpoints",
        "reserved_history_ids", "remaining_downloaded", "host_to_device_bytes", "device_to_host_bytes",
        "parked_fronts", "parked_particles")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


# energy terms at counters 0..12, then the native μ-pair cut energy at counter 61
_COUNTER_ENERGY_TERMS = ("injected_energy", "continuous_deposit", "cut_deposit", "cut_rest_mass",
                         "observed_energy", "escaped_energy", "observation_cut_overlap",
                         "mass_convention_correction", "medium_rest_mass_input",
                         "photoelectric_binding", "fallback_energy", "thinning_discarded_energy",
                         "pending_energy")
LEDGER_ENERGY_TERMS = _COUNTER_ENERGY_TERMS + ("native_muon_pair_cut_energy",)


class Ledger(ctypes.Structure):
    _fields_ = ([(k, c_double) for k in _COUNTER_ENERGY_TERMS] +
                [("steps", c_uint64), ("photon_cuts", c_uint64), ("lepton_limits", c_uint64 * 8),
                 ("thinning_hillas_vertices", c_uint64), ("thinning_statistical_vertices", c_uint64),
                 ("thinning_particles_discarded", c_uint64), ("deposited_steps", c_uint64),
                 ("fixed_point_overflows", c_uint64), ("invalid_records", c_uint64),
                 ("fallback_reasons", c_uint64 * 32), ("native_muon_pair_cut_energy", c_double)])

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = list(v) if hasattr(v, "__len__") else v
        return out


class RequestResult(ctypes.Structure):
    _fields_ = [("admitted_s", c_double), ("finished_s", c_double)] + [(k, c_uint64) for k in (
        "photon_steps", "lepton_steps", "photon_interactions", "lepton_interactions", "secondaries",
        "next_history_id", "observations", "fallbacks")] + [("slot", c_uint32), ("pad_", c_uint32),
                                                             ("ledger", Ledger)] + [
        ("radio_counters", _RadioCounters)]

    def as_dict(self):
        out = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("pad_", "ledger", "radio_counters")}
        out["ledger"] = self.ledger.as_dict()
        out["radio_counters"] = self.radio_counters.as_dict()
        return out


FIRST_INTERACTION_DTYPE = np.dtype({"names": ["parent_at_vertex", "secondaries", "process_id", "secondary_count"],
                                    "formats": [PARTICLE_STATE_DTYPE, (PARTICLE_STATE_DTYPE, 3), "<i4", "<u4"],
                                    "offsets": [0, 128, 512, 516], "itemsize": 528})


def _first_interaction(call):
    out = np.zeros(1, dtype=FIRST_INTERACTION_DTYPE)
    captured = c_int(0)
    _check(call(out.ctypes.data_as(c_void_p), byref(captured)))
    return out[0] if captured.value else None


FALLBACK_DTYPE = np.dtype({"names": ["particle", "reason", "process_id", "stage"],
                           "formats": [PARTICLE_STATE_DTYPE, "<i4", "<i4", "<u4"],
                           "offsets": [0, 128, 132, 136], "itemsize": 144})

# Counter layout of the device ledger (and of the oracle's counter array).
COUNTER_INDEX = {name: k for k, name in enumerate(_COUNTER_ENERGY_TERMS)}
COUNTER_INDEX.update(steps=13, photon_cuts=14, lepton_limit0=15, thin_hillas=23,
                     thin_statistical=24, thin_discarded=25, deposited_steps=26, overflows=27,
                     invalid=28, fallback_reason0=29, native_muon_pair_cut_energy=61)
COUNTERS = 64


def ledger_struct(ledger: dict) -> "Ledger":
    """A Ledger struct from a ledger dict (download_profile, decode_counters)."""
    out = Ledger()
    for k, _ in Ledger._fields_:
        v = ledger.get(k, 0)
        if isinstance(v, (list, tuple, np.ndarray)):
            arr = getattr(out, k)
            for i, x in enumerate(v):
                arr[i] = int(x)
        else:
            setattr(out, k, v)
    return out


class OutputStatistics(ctypes.Structure):
    _fields_ = [("steps", c_uint64), ("deposited_steps", c_uint64),
                ("weighted_deposited_energy_gev", c_double),
                ("weighted_native_muon_pair_cut_energy_gev", c_double),
                ("weighted_muon_parent_productions", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


LONGITUDINAL_COLUMNS = ("photon", "electron", "positron", "mu_minus", "mu_plus")


def output_profile(hist: dict, ledger: dict) -> dict:
    """CorsikaOutputSink::onGpuProfile (c8mi_em_output_profile): the checked
    profile in the writers' layout: longitudinal [bins][5] (LONGITUDINAL_COLUMNS),
    EM and muon energy loss per bin (GeV) and the sink's statistics."""
    arrays = [np.ascontiguousarray(hist[k], dtype=np.float64)
              for k in ("photons", "electrons", "positrons", "energy_loss")]
    bins = arrays[0].size
    longitudinal = np.zeros((bins, len(LONGITUDINAL_COLUMNS)))
    em_loss, mu_loss = np.zeros(bins), np.zeros(bins)
    st = OutputStatistics()
    led = ledger_struct(ledger)
    dp = lambda a: a.ctypes.data_as(POINTER(c_double))
    _check(library().c8mi_em_output_profile(*[dp(a) for a in arrays], ctypes.c_uint32(bins), byref(led),
                                            dp(longitudinal), dp(em_loss), dp(mu_loss), byref(st)))
    return dict(longitudinal=longitudinal, energy_loss_em=em_loss, energy_loss_muon=mu_loss,
                statistics=st.as_dict())


class ReportContext(ctypes.Structure):
    _fields_ = [("initial_total_gev", c_double), ("thinning_fraction", c_double),
                ("acceptance_tolerance", c_double), ("electromagnetic_primary", c_uint32), ("pad_", c_uint32),
                ("scalar_em_steps", c_uint64), ("particles_returned_for_cpu_fallback", c_uint64),
                ("particles_returned_for_cpu_memory_spill", c_uint64), ("specified_cpu_final_states", c_uint64)]


class EnergyLedger(ctypes.Structure):
    _fields_ = [(k, c_double) for k in (
        "initial_total_gev", "medium_rest_mass_input_gev", "deposited_gev", "cut_rest_mass_energy_gev",
        "observed_total_energy_gev", "escaped_total_energy_gev", "unwritten_photoelectric_binding_energy_gev",
        "observation_cut_overlap_energy_gev", "mass_convention_correction_gev", "source_gev", "terminal_gev",
        "residual_gev", "relative_closure_error", "acceptance_tolerance", "native_muon_pair_cut_energy_gev",
        "fallback_energy_gev", "pending_energy_gev", "thinning_discarded_energy_gev", "extended_residual_gev",
        "injected_gev")] + [("scalar_em_steps", c_uint64), ("fallbacks", c_uint64),
                            ("complete_coverage", c_uint32), ("accepted", c_uint32)]


ENERGY_LEDGER_DEFINITION = ("initial_total + medium_electron_rest + mass_convention_correction = "
                            "deposit + cut_rest + observed_total + escaped_total + "
                            "unwritten_photoelectric_binding - observation_cut_overlap")


def energy_ledger_report(ledgers, initial_total_gev, electromagnetic_primary=True, thinning_fraction=0.0,
                         scalar_em_steps=0, cpu_fallbacks=0, memory_spills=0, specified_final_states=0,
                         tolerance=0.0) -> dict:
    """The energy_ledger block of the reference's shower report
    (c8mi_em_energy_ledger_report) over the ledgers of a shower's sides, with
    the report's key names (ShowerReport.inl:1254-1276) plus the terms a
    complete-coverage shower never has."""
    ledgers = [ledgers] if isinstance(ledgers, dict) else list(ledgers)
    arr = (Ledger * max(len(ledgers), 1))(*[ledger_struct(l) for l in ledgers])
    ctx = ReportContext(initial_total_gev, thinning_fraction, tolerance, int(bool(electromagnetic_primary)), 0,
                        scalar_em_steps, cpu_fallbacks, memory_spills, specified_final_states)
    out = EnergyLedger()
    _check(library().c8mi_em_energy_ledger_report(arr, c_uint64(len(ledgers)), byref(ctx), byref(out)))
    r = {k: getattr(out, k) for k, _ in out._fields_}
    r["complete_coverage"] = bool(r["complete_coverage"])
    r["accepted"] = bool(r["accepted"])
    r["definition"] = ENERGY_LEDGER_DEFINITION
    return r


def ledger_residual(ledger: dict) -> float:
    """Energy closure of a shower (GeV): injected minus every sink, with the
    native/transport mass convention, atomic-electron rest mass, photoelectric
    binding and thinning terms (ShowerReport.inl:1236-1276)."""
    sinks = (ledger["continuous_deposit"] + ledger["cut_deposit"] + ledger["cut_rest_mass"] +
             ledger["observed_energy"] + ledger["escaped_energy"] - ledger["observation_cut_overlap"] +
             ledger["fallback_energy"] + ledger["pending_energy"] + ledger["thinning_discarded_energy"] +
             ledger["photoelectric_binding"] - ledger["medium_rest_mass_input"] -
             ledger["mass_convention_correction"] + ledger.get("native_muon_pair_cut_energy", 0.0))
    return ledger["injected_energy"] - sinks


class Atmosphere:
    """Layered spherical atmosphere (LayeredSphericalAtmosphereBuilder) as an
    environment tree: universe > outermost layer > ... > innermost layer; each
    layer a sphere around the Earth centre with an exponential (sliding planar)
    or homogeneous medium model. `magnetic_field_t`: a uniform field (tesla) in
    every layer (e+- then follow leapfrog steps)."""

    def __init__(self, layers=USSTDBK, observation_height_m: float = 0.0,
                 earth_radius_m: float = EARTH_RADIUS_M, observation: bool = True,
                 magnetic_field_t=(0.0, 0.0, 0.0)):
        self.earth_center = np.array([0.0, 0.0, -earth_radius_m])
        n = len(layers)
        self.models = (EnvModel * n)()
        self.nodes = (EnvNode * (n + 1))()
        self.children = (c_int32 * n)()
        self.nodes[0].shape_kind = SHAPE_UNIVERSE
        self.nodes[0].parent = -1
        self.nodes[0].model = -1
        # node k (1..n) is layer n-k (outermost first)
        for k in range(1, n + 1):
            li = n - k
            top, b, c = layers[li]
            m = self.models[k - 1]
            rho0 = b / c
            if li == n - 1:
                m.density_kind = DENSITY_HOMOGENEOUS
                m.reference_density_g_per_cm3 = rho0
            else:
                m.density_kind = DENSITY_SLIDING
                m.reference_density_g_per_cm3 = rho0
                scale = -c / 100.0
                m.scale_length_m = scale
                m.inverse_scale_length_per_m = 1.0 / scale
                m.reference_height_m = earth_radius_m
            for a in range(3):
                m.anchor_m[a] = self.earth_center[a]
            m.axis[2] = 1.0
            m.medium_id = 1
            for a in range(3):
                m.magnetic_field_t[a] = magnetic_field_t[a]
            node = self.nodes[k]
            node.shape_kind = SHAPE_SPHERE
            node.parent = k - 1
            node.model = k - 1
            for a in range(3):
                node.center_m[a] = self.earth_center[a]
            node.radius_m = earth_radius_m + top
            self.children[k - 1] = k
            self.nodes[k - 1].child_begin = k - 1
            self.nodes[k - 1].child_count = 1
        self.nodes[n].child_count = 0
        self.struct = Environment()
        self.struct.nodes = self.nodes
        self.struct.node_count = n + 1
        self.struct.child_index_count = n
        self.struct.children = self.children
        self.struct.models = self.models
        self.struct.model_count = n
        self.struct.observation_enabled = 1 if observation else 0
        self.struct.observation_point_m[2] = observation_height_m
        self.struct.observation_normal[2] = 1.0
        for a in range(3):
            self.struct.earth_center_m[a] = self.earth_center[a]
        self.struct.maximum_magnetic_deflection_rad = 0.2
        self.layers = layers
        self.earth_radius = earth_radius_m

    def density(self, p):
        """Density (g/cm^3) at points p [..., 3] (host reference for the projection)."""
        p = np.asarray(p, dtype=np.float64)
        h = np.linalg.norm(p - self.earth_center, axis=-1) - self.earth_radius
        rho = np.zeros(h.shape)
        lower = -np.inf
        for k, (top, b, c) in enumerate(self.layers):
            sel = (h >= lower) & (h < top)
            if k == len(self.layers) - 1:
                rho[sel] = b / c
            else:
                rho[sel] = b / c * np.exp(-h[sel] * 100.0 / c)
            lower = top
        return rho

    def axis_table(self, start, direction, step_m=10.0, length_m=None):
        """Grammage along a straight axis (trapezoidal integration on 1 m cells)."""
        start = np.asarray(start, dtype=np.float64)
        d = np.asarray(direction, dtype=np.float64)
        d = d / np.linalg.norm(d)
        if length_m is None:
            # to the observation plane
            length_m = (start[2] - self.struct.observation_point_m[2]) / max(-d[2], 1e-9)
        n = int(math.ceil(length_m / step_m)) + 2
        fine = 10
        s = np.arange(n * fine + 1) * (step_m / fine)
        rho = self.density(start[None, :] + s[:, None] * d[None, :])
        cum = np.concatenate(([0.0], np.cumsum(0.5 * (rho[1:] + rho[:-1]) * np.diff(s)))) * 100.0
        return np.ascontiguousarray(cum[::fine][:n])


class IceVolume:
    """In-ice scene (applications/c8_ice_cascade.cpp): the universe holding one
    homogeneous ice sphere around the Earth centre whose surface is z = 0
    (0.92 g/cm^3 by default, no magnetic field); the observation plane
    `observation_depth_m` below the surface, or none."""

    def __init__(self, density_g_per_cm3: float = 0.92, observation_depth_m=None,
                 earth_radius_m: float = EARTH_RADIUS_M):
        self.earth_center = np.array([0.0, 0.0, -earth_radius_m])
        self.rho = density_g_per_cm3
        self.models = (EnvModel * 1)()
        self.nodes = (EnvNode * 2)()
        self.children = (c_int32 * 1)(1)
        m = self.models[0]
        m.density_kind = DENSITY_HOMOGENEOUS
        m.reference_density_g_per_cm3 = density_g_per_cm3
        m.medium_id = 2
        m.axis[2] = 1.0
        for a in range(3):
            m.anchor_m[a] = self.earth_center[a]
        n = self.nodes
        n[0].shape_kind, n[0].parent, n[0].model, n[0].child_begin, n[0].child_count = SHAPE_UNIVERSE, -1, -1, 0, 1
        n[1].shape_kind, n[1].parent, n[1].model = SHAPE_SPHERE, 0, 0
        for a in range(3):
            n[1].center_m[a] = self.earth_center[a]
        n[1].radius_m = earth_radius_m
        s = self.struct = Environment()
        s.nodes, s.node_count, s.children, s.child_index_count = n, 2, self.children, 1
        s.models, s.model_count = self.models, 1
        s.observation_enabled = 0 if observation_depth_m is None else 1
        s.observation_point_m[2] = -(observation_depth_m or 0.0)
        s.observation_normal[2] = 1.0
        for a in range(3):
            s.earth_center_m[a] = self.earth_center[a]
        s.maximum_magnetic_deflection_rad = 0.2
        self.earth_radius = earth_radius_m

    def density(self, p):
        p = np.asarray(p, dtype=np.float64)
        inside = np.linalg.norm(p - self.earth_center, axis=-1) < self.earth_radius
        return np.where(inside, self.rho, 0.0)

    def axis_table(self, start, direction, step_m=10.0, length_m=400.0):
        """Grammage along the axis at `step_m` nodes (homogeneous: rho s)."""
        n = int(math.ceil(length_m / step_m)) + 2
        s = np.arange(n) * step_m
        d = np.asarray(direction, dtype=np.float64)
        d = d / np.linalg.norm(d)
        rho = self.density(np.asarray(start, dtype=np.float64)[None, :] + s[:, None] * d[None, :])
        if not np.all(rho == self.rho):
            raise ValueError("the axis leaves the ice within the table")
        return np.ascontiguousarray(s * self.rho * 100.0)


def particle(pid, energy_gev, position, direction, history_id=1, weight=1.0, node=-3):
    p = np.zeros((), dtype=PARTICLE_STATE_DTYPE)
    p["pid"], p["energy_gev"] = pid, energy_gev
    p["position_m"], p["direction"] = position, direction
    p["weight"], p["history_id"], p["environment_node"] = weight, history_id, node
    return p


def make_config(tables_module, seed=1, cut_mev=1.0, start=(0.0, 0.0, 112750.0),
                direction=(0.0, 0.0, -1.0), atmosphere=None, bins=128, bin_width=10.0,
                queue_capacity=1 << 22, observation_capacity=1 << 20, fallback_capacity=1 << 16,
                thinning=None, crossing_mode=0, slots=1, moliere=False, moliere_table=None,
                moliere_fast_iterations=0, resident_minimum_batch=0, minimum_batch_size=0):
    """EmConfig + the arrays it points to (kept alive on the returned object).
    moliere: apply Molière multiple scattering to e+- steps (the reference's
    production setting); off, e+- directions change only at vertices.
    moliere_table: None (analytical correction functions), True (the built
    interpolation table of the snapshot, as the reference backend uses) or a
    MoliereInterpolation. moliere_fast_iterations: Newton iterations per angle
    in the step kernel before a sample moves to its own wavefront (0: 12;
    results do not depend on it). resident_minimum_batch: below-minimum-batch
    checkpoint of the device-resident schedule (0: none); minimum_batch_size:
    the handoff batch a host router stages towards (0: 16384)."""
    cfg = EmConfig()
    cfg.shower_slots = slots
    cfg.resident_minimum_batch = resident_minimum_batch
    cfg.minimum_batch_size = minimum_batch_size
    cfg.moliere_fast_iterations = moliere_fast_iterations
    table_ref = None
    if moliere:
        cfg.electron_moliere = MoliereSnapshot.from_dict(tables_module.moliere_snapshot())
        cfg.moliere_enabled = 1
        if moliere_table is True:
            table_ref = air_moliere_interpolation(tables_module)
        elif moliere_table is not None:
            table_ref = moliere_table
        if table_ref is not None:
            cfg.electron_moliere_interpolation = ctypes.pointer(table_ref)
    cfg.random_seed = seed
    cfg.em_transport_cut_mev = cut_mev
    cfg.muon_transport_cut_mev = cut_mev
    cfg.queue_capacity = queue_capacity
    cfg.observation_capacity = observation_capacity
    cfg.fallback_capacity = fallback_capacity
    s = tables_module.photon_pair_lpm_snapshot()
    for k in ("baseline_mass_density_g_per_cm3", "molecular_density_per_cm3", "sum_charge",
              "e_lpm_mev", "classical_electron_radius_cm", "fine_structure_constant"):
        setattr(cfg.photon_pair_lpm, k, s[k])
    cfg.photon_pair_lpm.component_count = len(s["components"])
    for i, (h, z, r) in enumerate(s["components"]):
        cfg.photon_pair_lpm.components[i].component_hash = h
        cfg.photon_pair_lpm.components[i].nuclear_charge = z
        cfg.photon_pair_lpm.components[i].radiation_log_constant = r
    s = tables_module.brems_lpm_snapshot()
    for k in ("baseline_mass_density_g_per_cm3", "molecular_density_per_cm3", "sum_charge",
              "e_lpm_mev", "lepton_mass_mev", "electron_mass_mev", "muon_mass_mev",
              "classical_electron_radius_cm", "fine_structure_constant"):
        setattr(cfg.brems_lpm, k, s[k])
    cfg.brems_lpm.component_count = len(s["components"])
    for i, (h, z, a, r) in enumerate(s["components"]):
        c = cfg.brems_lpm.components[i]
        c.component_hash, c.nuclear_charge, c.atomic_mass_number, c.radiation_log_constant = h, z, a, r
    if thinning is not None:
        cfg.thinning.threshold_gev, cfg.thinning.maximum_weight = thinning
        cfg.thinning.enabled = 1
    cfg.thinning.erase_zero_weight = 1
    cfg.profile_bins = bins
    cfg.profile_crossing_mode = crossing_mode
    cfg.profile_bin_width_g_per_cm2 = bin_width
    cfg.profile_energy_loss_threshold_g_per_cm2 = 1e-3
    cfg.profile_weight_scale = float(1 << 24)
    cfg.profile_energy_scale = float(1 << 40)
    atmosphere = atmosphere or Atmosphere()
    table = atmosphere.axis_table(start, direction)
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    for a in range(3):
        cfg.axis_start_m[a] = start[a]
        cfg.axis_direction[a] = d[a]
    cfg.axis_step_m = 10.0
    cfg.axis_grammage_g_per_cm2 = table.ctypes.data_as(POINTER(c_double))
    cfg.axis_support_count = table.size

    class Holder:
        pass

    h = Holder()
    h.struct, h.axis_table, h.atmosphere, h.moliere_table = cfg, table, atmosphere, table_ref
    return h


def footprint_estimate(energy_gev):
    """Admission estimate of a shower's peak front (particles of one species in
    flight) at a 1 MeV cut: about 1.5x the largest front of oracle showers
    (52, 305, 2282, 17147 at 1, 10, 100, 1000 GeV)."""
    return 60.0 * float(energy_gev) ** 0.875


def shower_axis(atmosphere, start, direction, step_m=10.0):
    """ShowerAxis of a straight shower axis (the projection table is kept alive
    on the returned object as .table)."""
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    table = atmosphere.axis_table(start, d, step_m=step_m)
    ax = ShowerAxis()
    for a in range(3):
        ax.start_m[a] = start[a]
        ax.direction[a] = d[a]
    ax.step_m = step_m
    ax.grammage_g_per_cm2 = table.ctypes.data_as(POINTER(c_double))
    ax.support_count = table.size
    ax.table = table
    return ax


def decode_counters(counters, energy_scale):
    counters = np.asarray(counters, dtype=np.int64)
    ci = COUNTER_INDEX
    out = {name: float(counters[ci[name]]) / energy_scale for name in LEDGER_ENERGY_TERMS}
    out["steps"] = int(counters[ci["steps"]])
    out["photon_cuts"] = int(counters[ci["photon_cuts"]])
    out["lepton_limits"] = [int(x) for x in counters[ci["lepton_limit0"]:ci["lepton_limit0"] + 8]]
    out["thinning_hillas_vertices"] = int(counters[ci["thin_hillas"]])
    out["thinning_statistical_vertices"] = int(counters[ci["thin_statistical"]])
    out["thinning_particles_discarded"] = int(counters[ci["thin_discarded"]])
    out["deposited_steps"] = int(counters[ci["deposited_steps"]])
    out["fixed_point_overflows"] = int(counters[ci["overflows"]])
    out["invalid_records"] = int(counters[ci["invalid"]])
    out["fallback_reasons"] = [int(x) for x in counters[ci["fallback_reason0"]:ci["fallback_reason0"] + 32]]
    return out


class EmBackend:
    """Binding of the c8mi EM handle (c8mi_em.h; reference IEmBackend)."""

    def __init__(self, config, tables, environment, stream=None):
        self.config = config
        self.tables = tables
        self.environment = environment
        cfg = config.struct if hasattr(config, "struct") else config
        env = environment.struct if hasattr(environment, "struct") else environment
        self.bins = cfg.profile_bins
        self.observation_capacity = cfg.observation_capacity
        self.fallback_capacity = cfg.fallback_capacity
        self.handle = c_void_p()
        _check(library().c8mi_em_create(byref(cfg), byref(tables.struct), byref(env), c_void_p(stream),
                                        byref(self.handle)))

    def close(self):
        if self.handle:
            _check(library().c8mi_em_destroy(self.handle))
            self.handle = c_void_p()
        self._radio = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def attach_radio(self, radio):
        """Project every e+- step of the later lepton fronts onto `radio` (a
        c8mi.RadioAccumulator; None detaches). Keep it open while attached."""
        _check(library().c8mi_em_attach_radio(self.handle, radio.handle if radio is not None else None))
        self._radio = radio

    def begin_shower(self, shower_id=1, first_history_id=2):
        _check(library().c8mi_em_begin_shower(self.handle, c_uint64(shower_id), c_uint64(first_history_id)))

    def can_transport(self, particle_state) -> bool:
        p = np.ascontiguousarray(particle_state, dtype=PARTICLE_STATE_DTYPE).reshape(-1)[:1]
        out = c_int(0)
        _check(library().c8mi_em_can_transport(self.handle, p.ctypes.data_as(c_void_p), byref(out)))
        return bool(out.value)

    def enqueue(self, particles):
        p = np.ascontiguousarray(particles, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
        _check(library().c8mi_em_enqueue(self.handle, p.ctypes.data_as(c_void_p), c_uint64(p.size)))

    def run(self, maximum_rounds=1 << 40) -> dict:
        """Rounds of the reference's call schedule (c8mi_em_run)."""
        r = RunResult()
        _check(library().c8mi_em_run(self.handle, c_uint64(maximum_rounds), byref(r)))
        return r.as_dict()

    def run_photon_wavefront(self, particles, first_history_id, maximum_wavefronts=16,
                             minimum_batch=1) -> dict:
        """IEmBackend::runPhotonWavefront (c8mi_em_run_photon_wavefront)."""
        p = np.ascontiguousarray(particles, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
        r = WavefrontResult()
        _check(library().c8mi_em_run_photon_wavefront(
            self.handle, p.ctypes.data_as(c_void_p), c_uint64(p.size), c_uint64(first_history_id),
            c_uint64(maximum_wavefronts), c_uint64(minimum_batch), byref(r)))
        return r.as_dict()

    def run_lepton_wavefront(self, particles, first_history_id, history_limit, maximum_wavefronts=1024,
                             minimum_batch=1) -> dict:
        """IEmBackend::runLeptonWavefront (c8mi_em_run_lepton_wavefront)."""
        p = np.ascontiguousarray(particles, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
        r = WavefrontResult()
        _check(library().c8mi_em_run_lepton_wavefront(
            self.handle, p.ctypes.data_as(c_void_p), c_uint64(p.size), c_uint64(first_history_id),
            c_uint64(maximum_wavefronts), c_uint64(history_limit), c_uint64(minimum_batch), byref(r)))
        return r.as_dict()

    def remaining(self):
        """remainingPhotons_ / remainingLeptons_ of the last wavefront call."""
        n = c_uint64()
        _check(library().c8mi_em_download_remaining(self.handle, None, c_uint64(0), byref(n)))
        out = np.zeros(n.value, dtype=PARTICLE_STATE_DTYPE)
        if n.value:
            _check(library().c8mi_em_download_remaining(self.handle, out.ctypes.data_as(c_void_p),
                                                        c_uint64(out.size), byref(n)))
        return out

    def limits(self) -> dict:
        r = Limits()
        _check(library().c8mi_em_get_limits(self.handle, byref(r)))
        return r.as_dict()

    def statistics(self) -> dict:
        r = Statistics()
        _check(library().c8mi_em_get_statistics(self.handle, byref(r)))
        return r.as_dict()

    def pending(self):
        a, b = c_uint64(), c_uint64()
        _check(library().c8mi_em_pending_counts(self.handle, byref(a), byref(b)))
        return a.value, b.value

    def download_profile(self):
        hist = {k: np.zeros(self.bins) for k in ("photons", "electrons", "positrons", "energy_loss")}
        led = Ledger()
        dp = lambda a: a.ctypes.data_as(POINTER(c_double))
        _check(library().c8mi_em_download_profile(self.handle, dp(hist["photons"]), dp(hist["electrons"]),
                                                  dp(hist["positrons"]), dp(hist["energy_loss"]), byref(led)))
        return hist, led.as_dict()

    def observations(self):
        out = np.zeros(self.observation_capacity, dtype=PARTICLE_STATE_DTYPE)
        n = c_uint64()
        _check(library().c8mi_em_download_observations(self.handle, out.ctypes.data_as(c_void_p),
                                                       c_uint64(out.size), byref(n)))
        return out[:min(n.value, out.size)], n.value

    def fallbacks(self):
        out = np.zeros(self.fallback_capacity, dtype=FALLBACK_DTYPE)
        n = c_uint64()
        _check(library().c8mi_em_download_fallbacks(self.handle, out.ctypes.data_as(c_void_p),
                                                    c_uint64(out.size), byref(n)))
        return out[:min(n.value, out.size)], n.value

    def first_interaction(self):
        """IEmBackend::downloadFirstInteractionSnapshot: a FIRST_INTERACTION_DTYPE
        record, or None."""
        return _first_interaction(lambda o, c: library().c8mi_em_download_first_interaction(self.handle, o, c))

    def slot_first_interaction(self, slot):
        return _first_interaction(
            lambda o, c: library().c8mi_em_slot_download_first_interaction(self.handle, c_uint32(slot), o, c))

    # ---- shower pool (c8mi_em_slot_*) ----

    def slot_begin(self, slot, shower_id, first_history_id=2, axis=None):
        _check(library().c8mi_em_slot_begin(self.handle, c_uint32(slot), c_uint64(shower_id),
                                            c_uint64(first_history_id), byref(axis) if axis is not None else None))

    def slot_enqueue(self, slot, particles):
        p = np.ascontiguousarray(particles, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
        _check(library().c8mi_em_slot_enqueue(self.handle, c_uint32(slot), p.ctypes.data_as(c_void_p),
                                              c_uint64(p.size)))

    def run_rounds(self, maximum_rounds=1 << 40) -> dict:
        r = RunResult()
        _check(library().c8mi_em_run_rounds(self.handle, c_uint64(maximum_rounds), byref(r)))
        return r.as_dict()

    def slot_state(self, slot) -> dict:
        st = SlotState()
        _check(library().c8mi_em_slot_state_get(self.handle, c_uint32(slot), byref(st)))
        return st.as_dict()

    def slot_download_profile(self, slot):
        hist = {k: np.zeros(self.bins) for k in ("photons", "electrons", "positrons", "energy_loss")}
        led = Ledger()
        dp = lambda a: a.ctypes.data_as(POINTER(c_double))
        _check(library().c8mi_em_slot_download_profile(self.handle, c_uint32(slot), dp(hist["photons"]),
                                                       dp(hist["electrons"]), dp(hist["positrons"]),
                                                       dp(hist["energy_loss"]), byref(led)))
        return hist, led.as_dict()

    def slot_observations(self, slot):
        out = np.zeros(self.observation_capacity, dtype=PARTICLE_STATE_DTYPE)
        n = c_uint64()
        _check(library().c8mi_em_slot_download_observations(self.handle, c_uint32(slot),
                                                            out.ctypes.data_as(c_void_p), c_uint64(out.size),
                                                            byref(n)))
        return out[:min(n.value, out.size)], n.value

    def slot_fallbacks(self, slot):
        out = np.zeros(self.fallback_capacity, dtype=FALLBACK_DTYPE)
        n = c_uint64()
        _check(library().c8mi_em_slot_download_fallbacks(self.handle, c_uint32(slot),
                                                         out.ctypes.data_as(c_void_p), c_uint64(out.size),
                                                         byref(n)))
        return out[:min(n.value, out.size)], n.value

    def slot_pending(self, slot):
        a, b = c_uint64(), c_uint64()
        _check(library().c8mi_em_slot_pending(self.handle, c_uint32(slot), byref(a), byref(b)))
        return a.value, b.value

    def serve(self, requests, footprint_budget, realtime=True):
        """requests: list of dicts with primary (particle), shower_id, arrival_s,
        footprint, axis (ShowerAxis or None), particles (a handed-off batch
        enqueued instead of the primary, or None) and, with per-shower radio,
        radio observers (an OBSERVER_DTYPE array); returns (per-request dicts,
        totals). A request with observers gets its raw fixed-point waveforms
        as result["radio"] = (coreas [3][bins], zhs [3][bins])."""
        from . import observers_array

        n = len(requests)
        arr = (Request * n)()
        keep = []
        radio_out = {}
        for k, r in enumerate(requests):
            p = np.ascontiguousarray(r["primary"], dtype=PARTICLE_STATE_DTYPE).reshape(-1)[:1]
            ctypes.memmove(ctypes.addressof(arr[k].primary), p.ctypes.data, 128)
            arr[k].shower_id = r["shower_id"]
            arr[k].first_history_id = r.get("first_history_id", 2)
            arr[k].arrival_s = r.get("arrival_s", 0.0)
            arr[k].footprint_particles = r.get("footprint", 0.0)
            if r.get("particles") is not None:
                batch = np.ascontiguousarray(r["particles"], dtype=PARTICLE_STATE_DTYPE).reshape(-1)
                keep.append(batch)
                arr[k].particles = batch.ctypes.data
                arr[k].particle_count = batch.size
            if r.get("axis") is not None:
                arr[k].axis = r["axis"]
                keep.append(r["axis"])
            if r.get("radio_observers") is not None:
                obs = observers_array(r["radio_observers"])
                bins = int(obs["number_of_bins"].sum())
                coreas = np.zeros((3, bins), dtype=np.int64)
                zhs = np.zeros((3, bins), dtype=np.int64)
                keep += [obs, coreas, zhs]
                arr[k].radio_observers = obs.ctypes.data
                arr[k].radio_observer_count = obs.size
                arr[k].radio_coreas = coreas.ctypes.data
                arr[k].radio_zhs = zhs.ctypes.data
                radio_out[k] = (coreas, zhs)
        opts = ServeOptions()
        opts.footprint_budget_particles = footprint_budget
        opts.realtime = 1 if realtime else 0
        results = (RequestResult * n)()
        totals = RunResult()
        _check(library().c8mi_em_serve(self.handle, arr, c_uint64(n), byref(opts), results, byref(totals)))
        return _ResultList(results, n, radio_out), totals.as_dict()

    def enable_slot_radio(self, observers_per_slot, bins_per_observer, propagation, coreas=True, zhs=True,
                          exact=False, fixed_point_limit=1.0, zhs_refinement=1):
        """Per-shower radio of the pool (c8mi_em_enable_slot_radio)."""
        from . import RadioConfig

        cfg = RadioConfig()
        cfg.coreas_enabled = 1 if coreas else 0
        cfg.zhs_enabled = 1 if zhs else 0
        cfg.deterministic = 1
        cfg.zhs_subtrack_refinement = zhs_refinement
        cfg.exact_fp64 = 1 if exact else 0
        cfg.fixed_point_field_limit_v_per_m = fixed_point_limit
        cfg.propagation = propagation.struct
        _check(library().c8mi_em_enable_slot_radio(self.handle, byref(cfg), c_uint32(observers_per_slot),
                                                   c_uint32(bins_per_observer)))
        self._slot_radio_propagation = propagation

    def slot_set_radio(self, slot, observers):
        from . import observers_array

        obs = observers_array(observers) if observers is not None else None
        _check(library().c8mi_em_slot_set_radio(self.handle, c_uint32(slot),
                                                obs.ctypes.data_as(c_void_p) if obs is not None else None,
                                                c_uint64(obs.size if obs is not None else 0)))
        if not hasattr(self, "_slot_bins"):
            self._slot_bins = {}
        self._slot_bins[slot] = int(obs["number_of_bins"].sum()) if obs is not None else 0

    def slot_download_radio(self, slot):
        """(coreas [3][bins], zhs [3][bins], counters dict) of a slot."""
        from . import RadioCounters

        bins = self._slot_bins.get(slot, 0)
        coreas = np.zeros((3, bins), dtype=np.int64)
        zhs = np.zeros((3, bins), dtype=np.int64)
        counters = RadioCounters()
        _check(library().c8mi_em_slot_download_radio(self.handle, c_uint32(slot), coreas.ctypes.data_as(c_void_p),
                                                     zhs.ctypes.data_as(c_void_p), byref(counters)))
        return coreas, zhs, counters.as_dict()

    def slot_radio_kernel_time(self):
        ms, n = c_double(), c_uint64()
        _check(library().c8mi_em_slot_radio_kernel_time(self.handle, byref(ms), byref(n)))
        return ms.value, n.value

    def slot_end(self, slot):
        _check(library().c8mi_em_slot_end(self.handle, c_uint32(slot)))

    def enable_kernel_timing(self, enable=True):
        _check(library().c8mi_em_enable_kernel_timing(self.handle, c_int(1 if enable else 0)))

    def kernel_time(self):
        ms, n = (c_double * 6)(), (c_uint64 * 6)()
        _check(library().c8mi_em_kernel_time(self.handle, ms, n))
        return {k: (ms[i], n[i]) for i, k in enumerate(("photon_step", "lepton_transport", "lepton_vertex",
                                                        "count", "scan", "scatter"))}


# ---- host handoff (include/c8mi/c8mi_hybrid.h) ----

HYBRID_COUNTS = ("cpu_particles", "cpu_photon_steps", "cpu_lepton_steps", "cpu_interactions", "cpu_secondaries",
                 "cpu_tracks", "particles_staged", "small_batch_expansions", "particles_returned_for_expansion",
                 "scalar_expansion_budget_flushes", "stalled_boundary_flushes", "wavefronts", "input_batch_splits",
                 "maximum_input_batch", "reserved_history_ids", "below_minimum_batch_checkpoints",
                 "history_range_checkpoints", "photon_calls", "lepton_calls", "device_photon_fronts",
                 "device_lepton_fronts", "device_photon_steps", "device_lepton_steps", "device_interactions",
                 "device_secondaries", "bucketed_fronts", "next_history_id", "fallbacks_returned", "fallbacks_kept")


class HybridConfig(ctypes.Structure):
    """EmRouter's scheduling constants (zero: the reference's value)."""
    _fields_ = [("minimum_batch_size", c_uint64), ("maximum_expansion_rounds", c_uint32),
                ("photon_wavefronts", c_uint32), ("lepton_wavefronts", c_uint32),
                ("lepton_reservation_multiplier", c_uint32), ("lepton_reservation_chunk", c_uint64)]


class HybridResult(ctypes.Structure):
    _fields_ = [(k, c_uint64) for k in HYBRID_COUNTS] + [("host_seconds", c_double), ("device_seconds", c_double)]

    def as_dict(self):
        out = {k: int(getattr(self, k)) for k in HYBRID_COUNTS}
        out["host_seconds"] = self.host_seconds
        out["device_seconds"] = self.device_seconds
        return out


def hybrid_config(minimum_batch=0, expansion_rounds=0, photon_wavefronts=0, lepton_wavefronts=0,
                  multiplier=0, chunk=0) -> HybridConfig:
    return HybridConfig(minimum_batch, expansion_rounds, photon_wavefronts, lepton_wavefronts, multiplier, chunk)


class ScalarStepper:
    """Binding of the host scalar stepper (c8mi_em_scalar_*): one step of one
    particle on the CPU with the device path's step functions."""

    def __init__(self, config, tables, environment):
        cfg = config.struct if hasattr(config, "struct") else config
        env = environment.struct if hasattr(environment, "struct") else environment
        self.config, self.tables, self.environment = config, tables, environment
        self.bins = cfg.profile_bins
        self.observation_capacity = cfg.observation_capacity
        self.fallback_capacity = cfg.fallback_capacity
        self.handle = c_void_p()
        _check(library().c8mi_em_scalar_create(byref(cfg), byref(tables.struct), byref(env), byref(self.handle)))

    def close(self):
        if self.handle:
            _check(library().c8mi_em_scalar_destroy(self.handle))
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def begin_shower(self, shower_id=1):
        _check(library().c8mi_em_scalar_begin_shower(self.handle, c_uint64(shower_id)))

    def step(self, particle_state, first_history_id):
        """One host step: (outputs, children, track or None)."""
        p = np.ascontiguousarray(particle_state, dtype=PARTICLE_STATE_DTYPE).reshape(-1)[:1]
        out = np.zeros(2, dtype=PARTICLE_STATE_DTYPE)
        count, children, has_track = c_uint32(), c_uint32(), c_int()
        track = np.zeros(9)
        _check(library().c8mi_em_scalar_step(self.handle, p.ctypes.data_as(c_void_p), c_uint64(first_history_id),
                                             out.ctypes.data_as(c_void_p), byref(count), byref(children),
                                             track.ctypes.data_as(POINTER(c_double)), byref(has_track)))
        return out[:count.value], children.value, (track if has_track.value else None)

    def download_profile(self):
        hist = {k: np.zeros(self.bins) for k in ("photons", "electrons", "positrons", "energy_loss")}
        led = Ledger()
        dp = lambda a: a.ctypes.data_as(POINTER(c_double))
        _check(library().c8mi_em_scalar_download_profile(self.handle, dp(hist["photons"]), dp(hist["electrons"]),
                                                         dp(hist["positrons"]), dp(hist["energy_loss"]),
                                                         byref(led)))
        return hist, led.as_dict()

    def observations(self):
        out = np.zeros(self.observation_capacity, dtype=PARTICLE_STATE_DTYPE)
        n = c_uint64()
        _check(library().c8mi_em_scalar_download_observations(self.handle, out.ctypes.data_as(c_void_p),
                                                              c_uint64(out.size), byref(n)))
        return out[:min(n.value, out.size)], n.value

    def fallbacks(self):
        out = np.zeros(self.fallback_capacity, dtype=FALLBACK_DTYPE)
        n = c_uint64()
        _check(library().c8mi_em_scalar_download_fallbacks(self.handle, out.ctypes.data_as(c_void_p),
                                                           c_uint64(out.size), byref(n)))
        return out[:min(n.value, out.size)], n.value

    def first_interaction(self):
        return _first_interaction(
            lambda o, c: library().c8mi_em_scalar_download_first_interaction(self.handle, o, c))

    def prefix(self, primaries, shower_id, first_history_id=2, config=None, capacity=1 << 16,
               track_capacity=1 << 16):
        """The CPU prefix of a shower (c8mi_em_hybrid_prefix): (batch handed to
        the device, next history ID, host tracks [n][9], counts)."""
        p = np.ascontiguousarray(primaries, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
        batch = np.zeros(capacity, dtype=PARTICLE_STATE_DTYPE)
        tracks = np.zeros((track_capacity, 9))
        n, nxt, nt = c_uint64(), c_uint64(), c_uint64()
        r = HybridResult()
        _check(library().c8mi_em_hybrid_prefix(
            self.handle, byref(config) if config is not None else None, p.ctypes.data_as(c_void_p), c_uint64(p.size),
            c_uint64(shower_id), c_uint64(first_history_id), batch.ctypes.data_as(c_void_p), c_uint64(capacity),
            byref(n), byref(nxt), tracks.ctypes.data_as(POINTER(c_double)), c_uint64(track_capacity), byref(nt),
            byref(r)))
        if n.value > capacity or nt.value > track_capacity:
            raise RuntimeError("prefix: batch or track capacity exceeded")
        return batch[:n.value], nxt.value, tracks[:nt.value], r.as_dict()


class HybridCascade:
    """Binding of the host handoff driver (c8mi_em_hybrid_*; HybridCascade +
    EmRouter for EM particles) over an EmBackend and a ScalarStepper built
    from the same configuration, tables and environment."""

    def __init__(self, backend, stepper, config=None):
        self.backend, self.stepper = backend, stepper
        self.handle = c_void_p()
        _check(library().c8mi_em_hybrid_create(backend.handle, stepper.handle,
                                               byref(config) if config is not None else None, byref(self.handle)))
        self._radio = None

    def close(self):
        if self.handle:
            _check(library().c8mi_em_hybrid_destroy(self.handle))
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def attach_radio(self, radio):
        """Host e+- tracks go to `radio` (attach it to the backend too)."""
        _check(library().c8mi_em_hybrid_attach_radio(self.handle, radio.handle if radio is not None else None))
        self._radio = radio

    def run(self, primaries, shower_id, first_history_id=2) -> dict:
        p = np.ascontiguousarray(primaries, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
        r = HybridResult()
        _check(library().c8mi_em_hybrid_run(self.handle, p.ctypes.data_as(c_void_p), c_uint64(p.size),
                                            c_uint64(shower_id), c_uint64(first_history_id), byref(r)))
        return r.as_dict()


def hybrid_prefixes(config, tables, environment, primaries, shower_ids, first_history_id=2, threads=16,
                    hybrid=None, capacity=None, histograms=False):
    """CPU prefixes of many showers on host threads (c8mi_em_hybrid_prefixes):
    (list of handed-off batches, next history IDs, host ledgers, totals[, host
    histograms [n][4][bins]])."""
    cfg = config.struct if hasattr(config, "struct") else config
    env = environment.struct if hasattr(environment, "struct") else environment
    prim = np.ascontiguousarray(primaries, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
    ids = np.ascontiguousarray(shower_ids, dtype=np.uint64).reshape(-1)
    if ids.size != prim.size:
        raise ValueError("one shower ID per primary")
    n = prim.size
    capacity = capacity if capacity is not None else 512 * max(n, 1)
    # written only where batches land (the library fills [0, offsets[n]))
    staged = np.empty(capacity, dtype=PARTICLE_STATE_DTYPE)
    offsets = np.zeros(n + 1, dtype=np.uint64)
    nxt = np.zeros(n, dtype=np.uint64)
    ledgers = (Ledger * max(n, 1))()
    hist = np.zeros((n, 4, cfg.profile_bins)) if histograms else None
    totals = HybridResult()
    _check(library().c8mi_em_hybrid_prefixes(
        byref(cfg), byref(tables.struct), byref(env), byref(hybrid) if hybrid is not None else None,
        prim.ctypes.data_as(c_void_p), c_uint64(n), ids.ctypes.data_as(c_void_p), c_uint64(first_history_id),
        c_uint32(threads), staged.ctypes.data_as(c_void_p), c_uint64(capacity), offsets.ctypes.data_as(c_void_p),
        nxt.ctypes.data_as(c_void_p), ledgers, hist.ctypes.data_as(c_void_p) if hist is not None else None,
        byref(totals)))
    batches = [staged[int(offsets[k]):int(offsets[k + 1])] for k in range(n)]
    out = (batches, [int(x) for x in nxt], _LedgerList(ledgers, n), totals.as_dict())
    return out + (hist,) if histograms else out


class _ResultList:
    """The per-request results of serve: each turned into a dict when first
    read (and kept), with its radio waveforms when it had observers."""

    def __init__(self, array, count, radio):
        self._array, self._count, self._radio = array, count, radio
        self._dicts = [None] * count

    def __len__(self):
        return self._count

    def __getitem__(self, k):
        if isinstance(k, slice):
            return [self[i] for i in range(*k.indices(self._count))]
        if k < 0:
            k += self._count
        if not 0 <= k < self._count:
            raise IndexError(k)
        if self._dicts[k] is None:
            d = self._array[k].as_dict()
            if k in self._radio:
                d["radio"] = self._radio[k]
            self._dicts[k] = d
        return self._dicts[k]

    def __iter__(self):
        return (self[k] for k in range(self._count))


class _LedgerList:
    """The host ledgers of hybrid_prefixes, each turned into a dict when read."""

    def __init__(self, array, count):
        self._array, self._count = array, count

    def __len__(self):
        return self._count

    def __getitem__(self, k):
        if isinstance(k, slice):
            return [self[i] for i in range(*k.indices(self._count))]
        if k < 0:
            k += self._count
        if not 0 <= k < self._count:
            raise IndexError(k)
        return self._array[k].as_dict()

    def __iter__(self):
        return (self[k] for k in range(self._count))
