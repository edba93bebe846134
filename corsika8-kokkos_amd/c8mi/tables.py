"""Format-v7 native table sets (include/c8mi/c8mi_tables.h) built in numpy.

The reference exports these tables from live PROPOSAL 7.6.2 calculators
(corsika/kokkos/em/proposal/NativeTableExporter.hpp:30-32), which are not
available offline. This module builds table sets with the same layout from
closed-form inputs: `random_table_set` exercises every axis type, loss
transform, kinematic model and rate model for query parity tests; the spline
data are interpolation inputs only, the queries are what is under test.
"""
from __future__ import annotations

import ctypes
import math
from ctypes import POINTER, c_double, c_int32, c_uint32, c_uint64, c_void_p

import numpy as np

AXIS_DTYPE = np.dtype({"names": ["type", "nodes", "low", "high", "step"],
                       "formats": ["<u4", "<u4", "<f8", "<f8", "<f8"],
                       "offsets": [0, 4, 8, 16, 24], "itemsize": 32})
BICUBIC_DTYPE = np.dtype({"names": ["energy_axis", "loss_axis", "rows", "columns",
                                    "coefficient_offset", "polynomial_coefficient_offset"],
                          "formats": [AXIS_DTYPE, AXIS_DTYPE, "<u4", "<u4", "<u8", "<u8"],
                          "offsets": [0, 32, 64, 68, 72, 80], "itemsize": 88})
CUBIC_DTYPE = np.dtype({"names": ["axis", "value_transform", "coefficient_offset",
                                  "coefficient_count", "has_value_transform"],
                        "formats": [AXIS_DTYPE, AXIS_DTYPE, "<u8", "<u4", "<u4"],
                        "offsets": [0, 32, 64, 72, 76], "itemsize": 80})
DNDX_DTYPE = np.dtype({
    "names": ["pdg_id", "process_id", "medium_hash", "component_hash", "cross_section_hash",
              "table_hash", "loss_transform", "kinematic_model", "rate_model", "spline",
              "component_weight", "particle_mass_mev", "component_nuclear_charge",
              "component_atomic_mass", "component_average_nucleon_mass_mev", "direct_rate_factor",
              "lower_energy_limit_mev", "absolute_cut_mev", "relative_cut",
              "medium_mean_excitation_energy_mev", "selection_ordinal", "reserved"],
    "formats": ["<i4", "<i4", "<u8", "<u8", "<u8", "<u8", "<u4", "<u4", "<u4", BICUBIC_DTYPE,
                "<f8", "<f8", "<f8", "<f8", "<f8", "<f8", "<f8", "<f8", "<f8", "<f8", "<u4", "<u4"],
    "offsets": [0, 4, 8, 16, 24, 32, 40, 44, 48, 56, 144, 152, 160, 168, 176, 184, 192, 200, 208,
                216, 224, 228],
    "itemsize": 232})
TOTAL_DTYPE = np.dtype({"names": ["pdg_id", "reserved", "medium_hash", "table_hash", "spline",
                                  "lower_energy_limit_mev"],
                        "formats": ["<i4", "<u4", "<u8", "<u8", CUBIC_DTYPE, "<f8"],
                        "offsets": [0, 4, 8, 16, 24, 104], "itemsize": 112})
DEDX_DTYPE = np.dtype({"names": ["pdg_id", "process_id", "medium_hash", "cross_section_hash",
                                 "component_index", "table_hash", "spline", "component_weight",
                                 "lower_energy_limit_mev", "accumulation_ordinal", "reserved"],
                       "formats": ["<i4", "<i4", "<u8", "<u8", "<u8", "<u8", CUBIC_DTYPE, "<f8",
                                   "<f8", "<u4", "<u4"],
                       "offsets": [0, 4, 8, 16, 24, 32, 40, 120, 128, 136, 140], "itemsize": 144})
UTILITY_DTYPE = np.dtype({"names": ["pdg_id", "medium_hash", "table_hash", "spline",
                                    "particle_mass_mev", "lower_energy_limit_mev", "reverse"],
                          "formats": ["<i4", "<u8", "<u8", CUBIC_DTYPE, "<f8", "<f8", "<u4"],
                          "offsets": [0, 8, 16, 24, 104, 112, 120], "itemsize": 128})
QUERY_DTYPE = np.dtype({"names": ["kind", "pdg_id", "process_id", "component_hash", "energy_mev",
                                  "argument"],
                        "formats": ["<u4", "<i4", "<i4", "<u8", "<f8", "<f8"],
                        "offsets": [0, 4, 8, 16, 24, 32], "itemsize": 40})
RESULT_DTYPE = np.dtype({"names": ["status", "newton_iterations", "bisection_iterations", "value"],
                         "formats": ["<u4", "<u4", "<u4", "<f8"],
                         "offsets": [0, 4, 8, 16], "itemsize": 24})

NO_POLY = np.uint64(0xFFFFFFFFFFFFFFFF)

AXIS_LINEAR, AXIS_EXP, AXIS_EXPM1 = 0, 1, 2
LOSS_RELLOG, LOSS_RELLOG15, LOSS_ONEMINUSLOG = 0, 1, 2
KIN = dict(unsupported=0, only_stochastic=1, photon_pair=2, compton=3, brems_screening=4,
           electron_pair=5, moller=6, bhabha=7, bethe_bloch_rossi=8, brems_kkp=9, photonuclear=10)
RATE_BICUBIC, RATE_HEITLER, RATE_SAUTER, RATE_HECK = 0, 1, 2, 3
QUERY = dict(total_rate=0, selection_total_rate=1, rate=2, dedx=3, loss_fraction=4,
             photon_split=5, range=6, energy_after_loss=7, energy=8, cumulative_rate=9)
PROCESS = dict(brems=1000000002, ionization=1000000003, electron_pair=1000000004,
               photonuclear=1000000005, compton=1000000010, annihilation=1000000012,
               photon_pair=1000000013, photoproduction=1000000014, photon_muon_pair=1000000015,
               photoelectric=1000000016)

# PROPOSAL-like physical constants (values of the published CODATA set).
CONSTANTS = dict(electron_mass_mev=0.51099895, charged_pion_mass_mev=139.57039,
                 fine_structure_constant=7.2973525693e-3,
                 classical_electron_radius_cm=2.8179403262e-13, pi=math.pi, sqrt_e=math.sqrt(math.e),
                 muon_mass_mev=105.6583755)


class NativeConstants(ctypes.Structure):
    _fields_ = [(k, c_double) for k in ("electron_mass_mev", "charged_pion_mass_mev",
                                         "fine_structure_constant", "classical_electron_radius_cm",
                                         "pi", "sqrt_e", "muon_mass_mev")]


class NativeTableSet(ctypes.Structure):
    _fields_ = [
        ("dndx_columns", c_void_p), ("dndx_column_count", c_uint32),
        ("total_rate_column_count", c_uint32), ("total_rate_columns", c_void_p),
        ("selection_column_indices", c_void_p), ("dedx_columns", c_void_p),
        ("dedx_accumulation_indices", c_void_p), ("dedx_column_count", c_uint32),
        ("utility_column_count", c_uint32), ("utility_columns", c_void_p),
        ("bicubic_values", c_void_p), ("bicubic_derivative_energy", c_void_p),
        ("bicubic_derivative_loss", c_void_p), ("bicubic_mixed_derivative", c_void_p),
        ("bicubic_coefficient_count", c_uint64), ("bicubic_polynomial_coefficients", c_void_p),
        ("bicubic_polynomial_coefficient_count", c_uint64), ("cubic_values", c_void_p),
        ("cubic_node_derivatives", c_void_p), ("cubic_coefficient_count", c_uint64),
        ("constants", NativeConstants),
    ]


def axis(kind, low, high, nodes):
    a = np.zeros((), dtype=AXIS_DTYPE)
    a["type"], a["nodes"], a["low"], a["high"] = kind, nodes, low, high
    if kind == AXIS_LINEAR:
        a["step"] = (high - low) / (nodes - 1)
    elif kind == AXIS_EXP:
        a["step"] = math.log(high / low) / (nodes - 1)
    else:
        a["step"] = (math.log1p(high / low) - math.log(2.0)) / (nodes - 1)
    return a


def back_transform(a, c):
    c = np.asarray(c, dtype=np.float64)
    if a["type"] == AXIS_LINEAR:
        return c * a["step"] + a["low"]
    if a["type"] == AXIS_EXP:
        return a["low"] * np.exp(c * a["step"])
    return a["low"] * np.expm1(c * a["step"] + math.log(2.0))


# Hermite basis in monomial form: h_a(x) = sum_k H[a, k] x^k
_H = np.array([[1., 0., -3., 2.], [0., 0., 3., -2.], [0., 1., -2., 1.], [0., 0., -1., 1.]])


class TableBuilder:
    """Accumulates columns and spline arrays; `build()` returns NativeTables."""

    def __init__(self, constants=None):
        self.constants = dict(CONSTANTS if constants is None else constants)
        self.dndx, self.totals, self.dedx, self.utility = [], [], [], []
        self.bv, self.be, self.bl, self.bm, self.poly = [], [], [], [], []
        self.cv, self.cd = [], []
        self.bicubic_size = 0
        self.poly_size = 0
        self.cubic_size = 0

    # -- splines ------------------------------------------------------------
    def bicubic(self, energy_axis, loss_axis, function, polynomial=False, noise=None):
        """function(E, t) on the node grid (E in MeV, t transformed loss);
        derivatives in node units from central differences of `function`."""
        rows, cols = int(energy_axis["nodes"]), int(loss_axis["nodes"])
        i = np.arange(rows, dtype=np.float64)[:, None]
        j = np.arange(cols, dtype=np.float64)[None, :]
        h = 1e-4

        def f(ci, cj):
            return function(back_transform(energy_axis, ci), back_transform(loss_axis, cj))

        values = f(i, j) * np.ones((rows, cols))
        if noise is not None:
            values = values * (1.0 + noise * np.random.default_rng(rows * 131 + cols).standard_normal(values.shape))
        de = (f(i + h, j) - f(i - h, j)) / (2 * h) * np.ones((rows, cols))
        dl = (f(i, j + h) - f(i, j - h)) / (2 * h) * np.ones((rows, cols))
        dm = (f(i + h, j + h) - f(i + h, j - h) - f(i - h, j + h) + f(i - h, j - h)) / (4 * h * h)
        dm = dm * np.ones((rows, cols))
        d = np.zeros((), dtype=BICUBIC_DTYPE)
        d["energy_axis"], d["loss_axis"], d["rows"], d["columns"] = energy_axis, loss_axis, rows, cols
        d["coefficient_offset"] = self.bicubic_size
        # column-major: element (row, column) at offset + row + rows * column
        for arr, out in ((values, self.bv), (de, self.be), (dl, self.bl), (dm, self.bm)):
            out.append(np.asarray(arr, dtype=np.float64).T.reshape(-1))
        self.bicubic_size += rows * cols
        d["polynomial_coefficient_offset"] = NO_POLY
        if polynomial:
            coeffs = np.zeros(((rows - 1) * (cols - 1), 16))
            for n1 in range(cols - 1):
                for n0 in range(rows - 1):
                    B = np.empty((4, 4))
                    for a in range(4):
                        for b in range(4):
                            src = values if (a < 2 and b < 2) else de if (a >= 2 and b < 2) else \
                                dl if (a < 2 and b >= 2) else dm
                            B[a, b] = src[n0 + (a & 1), n1 + (b & 1)]
                    coeffs[n0 + (rows - 1) * n1] = (_H.T @ B @ _H).reshape(-1)
            d["polynomial_coefficient_offset"] = self.poly_size
            self.poly.append(coeffs.reshape(-1))
            self.poly_size += coeffs.size
        return d

    def cubic(self, ax, function, value_transform=None):
        n = int(ax["nodes"])
        c = np.arange(n, dtype=np.float64)
        h = 1e-4

        def g(cc):
            v = function(back_transform(ax, cc))
            if value_transform is not None:  # store the forward-transformed value
                vt = value_transform
                if vt["type"] == AXIS_LINEAR:
                    v = (v - vt["low"]) / vt["step"]
                elif vt["type"] == AXIS_EXP:
                    v = np.log(v / vt["low"]) / vt["step"]
                else:
                    v = (np.log1p(v / vt["low"]) - math.log(2.0)) / vt["step"]
            return v

        values = g(c)
        deriv = (g(c + h) - g(c - h)) / (2 * h)
        d = np.zeros((), dtype=CUBIC_DTYPE)
        d["axis"] = ax
        d["coefficient_offset"] = self.cubic_size
        d["coefficient_count"] = n
        if value_transform is not None:
            d["value_transform"] = value_transform
            d["has_value_transform"] = 1
        self.cv.append(values)
        self.cd.append(deriv)
        self.cubic_size += n
        return d

    # -- columns ------------------------------------------------------------
    def add_dndx(self, **fields):
        c = np.zeros((), dtype=DNDX_DTYPE)
        c["component_weight"] = 1.0
        for k, v in fields.items():
            c[k] = v
        self.dndx.append(c)

    def add_total(self, pdg_id, spline, lower=0.0):
        c = np.zeros((), dtype=TOTAL_DTYPE)
        c["pdg_id"], c["spline"], c["lower_energy_limit_mev"] = pdg_id, spline, lower
        self.totals.append(c)

    def add_dedx(self, pdg_id, process_id, spline, lower=0.0, weight=1.0, ordinal=0):
        c = np.zeros((), dtype=DEDX_DTYPE)
        c["pdg_id"], c["process_id"], c["spline"] = pdg_id, process_id, spline
        c["lower_energy_limit_mev"], c["component_weight"] = lower, weight
        c["accumulation_ordinal"] = ordinal
        self.dedx.append(c)

    def add_utility(self, pdg_id, spline, mass, lower, reverse=0):
        c = np.zeros((), dtype=UTILITY_DTYPE)
        c["pdg_id"], c["spline"], c["particle_mass_mev"] = pdg_id, spline, mass
        c["lower_energy_limit_mev"], c["reverse"] = lower, reverse
        self.utility.append(c)

    def build(self):
        return NativeTables(self)


class NativeTables:
    """Owns the arrays of one table set and the ctypes struct that points into them."""

    def __init__(self, b: TableBuilder):
        dndx = np.array(b.dndx, dtype=DNDX_DTYPE)
        order = np.lexsort((dndx["component_hash"], dndx["process_id"], dndx["pdg_id"]))
        self.dndx = np.ascontiguousarray(dndx[order])
        # selection order: per particle, ascending selection_ordinal
        self.selection = np.ascontiguousarray(
            np.lexsort((self.dndx["selection_ordinal"], self.dndx["pdg_id"])).astype(np.uint32))
        totals = np.array(b.totals, dtype=TOTAL_DTYPE)
        self.totals = np.ascontiguousarray(totals[np.argsort(totals["pdg_id"], kind="stable")])
        self.dedx = np.ascontiguousarray(np.array(b.dedx, dtype=DEDX_DTYPE))
        self.dedx_order = np.ascontiguousarray(
            np.argsort(self.dedx["accumulation_ordinal"], kind="stable").astype(np.uint32)
            if self.dedx.size else np.zeros(0, dtype=np.uint32))
        self.utility = np.ascontiguousarray(np.array(b.utility, dtype=UTILITY_DTYPE))
        cat = lambda parts: np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(0))
        self.bicubic_values = cat(b.bv)
        self.bicubic_energy = cat(b.be)
        self.bicubic_loss = cat(b.bl)
        self.bicubic_mixed = cat(b.bm)
        self.polynomial = cat(b.poly) if b.poly else None
        self.cubic_values = cat(b.cv)
        self.cubic_derivatives = cat(b.cd)
        self.constants = b.constants
        s = NativeTableSet()
        ptr = lambda a: None if a is None or a.size == 0 else a.ctypes.data
        s.dndx_columns, s.dndx_column_count = ptr(self.dndx), self.dndx.size
        s.total_rate_columns, s.total_rate_column_count = ptr(self.totals), self.totals.size
        s.selection_column_indices = ptr(self.selection)
        s.dedx_columns, s.dedx_column_count = ptr(self.dedx), self.dedx.size
        s.dedx_accumulation_indices = ptr(self.dedx_order)
        s.utility_columns, s.utility_column_count = ptr(self.utility), self.utility.size
        s.bicubic_values = ptr(self.bicubic_values)
        s.bicubic_derivative_energy = ptr(self.bicubic_energy)
        s.bicubic_derivative_loss = ptr(self.bicubic_loss)
        s.bicubic_mixed_derivative = ptr(self.bicubic_mixed)
        s.bicubic_coefficient_count = self.bicubic_values.size
        s.bicubic_polynomial_coefficients = ptr(self.polynomial)
        s.bicubic_polynomial_coefficient_count = 0 if self.polynomial is None else self.polynomial.size
        s.cubic_values, s.cubic_node_derivatives = ptr(self.cubic_values), ptr(self.cubic_derivatives)
        s.cubic_coefficient_count = self.cubic_values.size
        for k, v in self.constants.items():
            setattr(s.constants, k, v)
        self.struct = s

    def column_index(self, pdg_id, process_id, component_hash=0):
        m = (self.dndx["pdg_id"] == pdg_id) & (self.dndx["process_id"] == process_id) & \
            (self.dndx["component_hash"] == component_hash)
        return int(np.flatnonzero(m)[0])


def _smooth_cdf(k):
    def g(t):
        t = np.asarray(t, dtype=np.float64)
        return np.expm1(k * t) / math.expm1(k)
    return g


def random_table_set(seed: int = 1, polynomial_every: int = 3, noise: float = 0.0,
                     wiggle: float = 0.0) -> NativeTables:
    """Table set exercising every axis type, loss transform, kinematic model and
    rate model, with smooth monotone cumulative rates, optional relative noise on
    the node values, and an optional oscillation of relative amplitude `wiggle`
    that makes the cumulative rates non-monotone (reaching the Newton
    safeguards and PROPOSAL's bisection fallback)."""
    rng = np.random.default_rng(seed)
    b = TableBuilder()
    me = CONSTANTS["electron_mass_mev"]
    mmu = CONSTANTS["muon_mass_mev"]
    axes = [AXIS_EXP, AXIS_EXPM1, AXIS_LINEAR]
    specs = [
        # pdg, process, component, kinematic model, loss transform, mass, rate model
        (22, PROCESS["photon_pair"], 0, "photon_pair", LOSS_RELLOG, 0.0, RATE_BICUBIC),
        (22, PROCESS["compton"], 0, "compton", LOSS_RELLOG, 0.0, RATE_BICUBIC),
        (22, PROCESS["photoelectric"], 7, "only_stochastic", LOSS_RELLOG, 0.0, RATE_SAUTER),
        (22, PROCESS["photoproduction"], 14, "only_stochastic", LOSS_RELLOG, 0.0, RATE_HECK),
        (22, PROCESS["photon_muon_pair"], 0, "only_stochastic", LOSS_RELLOG, 0.0, RATE_BICUBIC),
        (11, PROCESS["brems"], 7, "brems_screening", LOSS_RELLOG, me, RATE_BICUBIC),
        (11, PROCESS["brems"], 8, "brems_screening", LOSS_RELLOG15, me, RATE_BICUBIC),
        (11, PROCESS["ionization"], 0, "moller", LOSS_ONEMINUSLOG, me, RATE_BICUBIC),
        (11, PROCESS["electron_pair"], 7, "electron_pair", LOSS_RELLOG, me, RATE_BICUBIC),
        (-11, PROCESS["brems"], 7, "brems_screening", LOSS_RELLOG, me, RATE_BICUBIC),
        (-11, PROCESS["ionization"], 0, "bhabha", LOSS_ONEMINUSLOG, me, RATE_BICUBIC),
        (-11, PROCESS["annihilation"], 7, "only_stochastic", LOSS_RELLOG, me, RATE_HEITLER),
        (13, PROCESS["brems"], 7, "brems_kkp", LOSS_RELLOG, mmu, RATE_BICUBIC),
        (13, PROCESS["ionization"], 0, "bethe_bloch_rossi", LOSS_RELLOG, mmu, RATE_BICUBIC),
        (13, PROCESS["photonuclear"], 8, "photonuclear", LOSS_RELLOG15, mmu, RATE_BICUBIC),
    ]
    for n, (pdg, process, comp, kin, lt, mass, rate) in enumerate(specs):
        e_lo = max(1.05 * mass, 1.0) if mass > 0 else 1.0
        e_hi = 1.0e7
        ea = axis(axes[n % 3] if axes[n % 3] != AXIS_LINEAR else AXIS_EXP, e_lo, e_hi, 60 + 7 * (n % 4))
        la = axis(AXIS_LINEAR, 0.0, 1.0, 30 + 5 * (n % 3))
        scale = 10 ** rng.uniform(-3, -1)
        k = rng.uniform(0.5, 4.0)
        g0 = _smooth_cdf(k)
        g = g0 if wiggle == 0.0 else (lambda t, g0=g0, w=wiggle, m=n % 3 + 2:
                                      g0(t) + w * np.sin(2 * math.pi * m * np.asarray(t) + 0.7))
        spline = b.bicubic(ea, la, lambda e, t, s=scale, g=g: s * (1.0 + 0.2 * np.log(e)) * g(t),
                           polynomial=(polynomial_every > 0 and n % polynomial_every == 0), noise=noise or None)
        b.add_dndx(pdg_id=pdg, process_id=process, component_hash=comp, kinematic_model=KIN[kin],
                   loss_transform=lt, rate_model=rate, spline=spline,
                   component_weight=rng.uniform(0.5, 1.5), particle_mass_mev=mass,
                   component_nuclear_charge=7.3, component_atomic_mass=14.6,
                   component_average_nucleon_mass_mev=938.9, direct_rate_factor=rng.uniform(0.1, 2.0),
                   lower_energy_limit_mev=e_lo * 1.5, absolute_cut_mev=0.5, relative_cut=0.05,
                   medium_mean_excitation_energy_mev=85.7e-6, selection_ordinal=(n * 5) % 7)
    for pdg, mass in ((22, 0.0), (11, me), (-11, me), (13, mmu)):
        e_lo = max(1.05 * mass, 1.0)
        ta = axis(AXIS_EXP, e_lo, 1.0e7, 80)
        vt = axis(AXIS_EXP, 1e-4, 1.0, 2) if pdg == 11 else None
        spline = b.cubic(ta, lambda e: 0.01 + 0.002 * np.log(e), value_transform=vt)
        b.add_total(pdg, spline, lower=e_lo * 1.2)
        for comp in range(2):
            da = axis(axes[comp], e_lo, 1.0e7, 70)
            ds = b.cubic(da, lambda e, c=comp: (2.0 + c) * (1.0 + 0.05 * np.log(e)) + 1e-3 * e)
            b.add_dedx(pdg, PROCESS["ionization"] + comp, ds, lower=e_lo * (1.0 + comp),
                       weight=1.0 + 0.25 * comp, ordinal=(comp + 1) % 2)
        lower = max(mass, 0.5) * 1.1
        ua = axis(AXIS_EXP, lower, 1.0e7, 120)
        # range integral R(E) = int dE / (dE/dX), monotone increasing
        us = b.cubic(ua, lambda e: np.log(e / lower) * 40.0 + 1e-3 * (e - lower))
        b.add_utility(pdg, us, mass=mass if mass > 0 else me, lower=lower, reverse=0)
    return b.build()


def random_queries(tables: NativeTables, count: int, seed: int = 2) -> np.ndarray:
    """Queries of every kind over the table set, inside and outside each domain."""
    rng = np.random.default_rng(seed)
    q = np.zeros(count, dtype=QUERY_DTYPE)
    kinds = rng.integers(0, 10, count)
    cols = rng.integers(0, tables.dndx.size, count)
    pids = np.array([22, 11, -11, 13, 211])
    for i in range(count):
        kind = int(kinds[i])
        q[i]["kind"] = kind
        c = tables.dndx[cols[i]]
        if kind in (QUERY["rate"], QUERY["loss_fraction"], QUERY["photon_split"],
                    QUERY["cumulative_rate"]):
            q[i]["pdg_id"], q[i]["process_id"] = c["pdg_id"], c["process_id"]
            q[i]["component_hash"] = c["component_hash"]
            if rng.random() < 0.03:
                q[i]["process_id"] += 1  # missing column
        else:
            q[i]["pdg_id"] = pids[rng.integers(0, 4 if rng.random() > 0.03 else 5)]
        e = 10 ** rng.uniform(-0.3, 7.3)
        r = rng.random()
        if r < 0.01:
            e = float("nan")
        q[i]["energy_mev"] = e
        if kind in (QUERY["loss_fraction"], QUERY["photon_split"], QUERY["cumulative_rate"]):
            arg = rng.random()
            r = rng.random()
            if r < 0.03:
                arg = 0.0
            elif r < 0.06:
                arg = 1.0
            elif r < 0.08:
                arg = -0.1 if rng.random() < 0.5 else 1.1
        elif kind == QUERY["energy_after_loss"]:
            arg = 10 ** rng.uniform(-4, 2.7)
            if rng.random() < 0.03:
                arg = 0.0
        elif kind == QUERY["energy"]:
            arg = 10 ** rng.uniform(-3, 2.8)
        else:
            arg = 0.0
        q[i]["argument"] = arg
    return q


# ---- table files and content hash (c8mi_native_table_*) ----

class NativeInteractionIdentity(ctypes.Structure):
    _fields_ = [("pdg_id", c_int32), ("reserved", c_uint32), ("medium_hash", c_uint64),
                ("interaction_hash", c_uint64)]


class NativeTableMeta(ctypes.Structure):
    _fields_ = [("format_version", c_uint32), ("pad_", c_uint32), ("proposal_version", ctypes.c_char * 32),
                ("cubic_interpolation_version", ctypes.c_char * 32), ("stochastic_cut_mev", c_double),
                ("interaction_identities", c_void_p), ("interaction_identity_count", c_uint64)]


def table_meta(proposal_version="7.6.2", cubic_interpolation_version="0.1.5", stochastic_cut_mev=0.0,
               identities=()):
    """Metadata of a table set (ProposalNativeTableSet's version strings, cut and
    Interaction identities [(pdg_id, medium_hash, interaction_hash)])."""
    m = NativeTableMeta()
    m.format_version = 7
    m.proposal_version = proposal_version.encode()
    m.cubic_interpolation_version = cubic_interpolation_version.encode()
    m.stochastic_cut_mev = stochastic_cut_mev
    ids = (NativeInteractionIdentity * max(len(identities), 1))()
    for k, (pid, medium, interaction) in enumerate(identities):
        ids[k].pdg_id, ids[k].medium_hash, ids[k].interaction_hash = pid, medium, interaction
    m.interaction_identities = ctypes.cast(ids, c_void_p) if identities else None
    m.interaction_identity_count = len(identities)
    m._keep = ids
    return m


def table_hash(tables, meta) -> bytes:
    """The content hash (calculate_proposal_native_hash) of a table set."""
    from . import _check, library

    digest = (ctypes.c_uint8 * 32)()
    _check(library().c8mi_native_table_hash(ctypes.byref(tables.struct), ctypes.byref(meta), digest))
    return bytes(digest)


def write_table_file(path, tables, meta):
    from . import _check, library

    _check(library().c8mi_native_table_write(str(path).encode(), ctypes.byref(tables.struct), ctypes.byref(meta)))


class TableFile:
    """A table file read back (c8mi_native_table_read): .struct is the set (use
    it where a NativeTables is expected), .meta the metadata, .digest the
    verified content hash. Keep the object alive while the set is in use."""

    def __init__(self, path):
        from . import _check, library

        self.handle = c_void_p()
        _check(library().c8mi_native_table_read(str(path).encode(), ctypes.byref(self.handle)))
        self.struct = NativeTableSet()
        self.meta = NativeTableMeta()
        digest = (ctypes.c_uint8 * 32)()
        _check(library().c8mi_native_table_file_view(self.handle, ctypes.byref(self.struct),
                                                     ctypes.byref(self.meta), digest))
        self.digest = bytes(digest)

    def close(self):
        from . import _check, library

        if self.handle:
            _check(library().c8mi_native_table_file_destroy(self.handle))
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
