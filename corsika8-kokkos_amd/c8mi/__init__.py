"""ctypes binding of the c8mi C ABI (include/c8mi/*.h).

The compute path is libc8mi.so (HIP kernels for gfx950 + C++ host classes).
This module only marshals arguments; it has no numerical fallback. If the
library is missing, every entry point raises instead of computing anything.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

PACKAGE_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBRARY_PATH = os.path.join(PACKAGE_DIR, "libc8mi.so")

STATUS_NAMES = {
    1: "invalid_argument",
    2: "logic_error",
    3: "length_error",
    4: "runtime_error",
    5: "device_error",
    6: "overflow_error",
}


class C8miError(RuntimeError):
    """A c8mi entry point returned a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


_LIB = None


def library() -> ctypes.CDLL:
    """Loads libc8mi.so from the package directory (never a fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBRARY_PATH):
            raise RuntimeError(
                f"{LIBRARY_PATH} is not built; run __graft_entry__.build() (make -C corsika8-kokkos_amd)"
            )
        lib = ctypes.CDLL(LIBRARY_PATH)
        lib.c8mi_last_error.restype = c_char_p
        _LIB = lib
    return _LIB


def _check(status: int) -> None:
    if status != 0:
        raise C8miError(status, library().c8mi_last_error().decode())


# ---- records (layouts of include/c8mi/c8mi_types.h) -------------------------

PARTICLE_STATE_DTYPE = np.dtype(
    {
        "names": ["pid", "medium_id", "generation", "reserved", "energy_gev", "position_m",
                  "direction", "time_s", "weight", "history_id", "parent_history_id", "step_id",
                  "environment_node"],
        "formats": ["<i4", "<i4", "<u4", "<u4", "<f8", ("<f8", 3), ("<f8", 3), "<f8", "<f8",
                    "<u8", "<u8", "<u8", "<i4"],
        "offsets": [0, 4, 8, 12, 16, 24, 48, 72, 80, 88, 96, 104, 112],
        "itemsize": 128,
    }
)

LEPTON_TRANSPORT_RECORD_DTYPE = np.dtype(
    {
        "names": ["start", "end", "input_index", "limit"],
        "formats": [PARTICLE_STATE_DTYPE, PARTICLE_STATE_DTYPE, "<u8", "<i4"],
        "offsets": [320, 448, 576, 584],
        "itemsize": 752,
    }
)


class RadioObserver(ctypes.Structure):
    _fields_ = [
        ("position_m", c_double * 3),
        ("start_time_s", c_double),
        ("duration_s", c_double),
        ("sample_rate_hz", c_double),
        ("number_of_bins", c_uint64),
    ]


class RadioPropagation(ctypes.Structure):
    _fields_ = [
        ("homogeneous_refractive_index", c_double),
        ("minimum_height_m", c_double),
        ("maximum_height_m", c_double),
        ("step_m", c_double),
        ("inverse_step_per_m", c_double),
        ("slope_refractivity_lower", c_double),
        ("slope_integrated_refractivity_lower", c_double),
        ("slope_refractivity_upper", c_double),
        ("slope_integrated_refractivity_upper", c_double),
        ("refractivity", POINTER(c_double)),
        ("integrated_refractivity", POINTER(c_double)),
        ("table_size", c_uint64),
    ]


class RadioConfig(ctypes.Structure):
    _fields_ = [
        ("coreas_enabled", c_int32),
        ("zhs_enabled", c_int32),
        ("deterministic", c_int32),
        ("zhs_subtrack_refinement", c_uint32),
        ("exact_fp64", c_int32),
        ("reserved0", c_int32),
        ("fixed_point_field_limit_v_per_m", c_double),
        ("propagation", RadioPropagation),
        ("coreas_observers", POINTER(RadioObserver)),
        ("coreas_observer_count", c_uint64),
        ("zhs_observers", POINTER(RadioObserver)),
        ("zhs_observer_count", c_uint64),
    ]


class RadioCounters(ctypes.Structure):
    _fields_ = [
        ("coreas_contributions", c_uint64),
        ("zhs_contributions", c_uint64),
        ("zhs_subtracks", c_uint64),
        ("valid_tracks", c_uint64),
        ("fixed_point_overflows", c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


OBSERVER_DTYPE = np.dtype(
    [("position_m", "<f8", 3), ("start_time_s", "<f8"), ("duration_s", "<f8"),
     ("sample_rate_hz", "<f8"), ("number_of_bins", "<u8")]
)


def _dptr(array: np.ndarray):
    return array.ctypes.data_as(POINTER(c_double))


class Propagation:
    """A straight-ray propagation snapshot plus the arrays it points into."""

    def __init__(self, homogeneous_index: float = 0.0, refractivity=None,
                 minimum_height_m: float = 0.0, maximum_height_m: float = 0.0,
                 step_m: float = 1.0):
        self.struct = RadioPropagation()
        self.refractivity = None
        self.integrated = None
        if homogeneous_index > 0.0:
            self.struct.homogeneous_refractive_index = homogeneous_index
            return
        self.refractivity = np.ascontiguousarray(refractivity, dtype=np.float64)
        self.integrated = np.zeros_like(self.refractivity)
        _check(library().c8mi_radio_build_flat_table(
            _dptr(self.refractivity), c_uint64(self.refractivity.size), c_double(minimum_height_m),
            c_double(maximum_height_m), c_double(step_m), _dptr(self.integrated),
            byref(self.struct)))


def observers_array(observers) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(observers, dtype=OBSERVER_DTYPE))


def zhs_electric_field(observers, potential: np.ndarray):
    """The radio writer's ZHS field from a decoded vector potential
    (c8mi_radio_zhs_electric_field): E[i] = -(A[i+1] - A[i]) * sample_rate at
    the bin midpoints, bins - 1 samples per observer. Returns (field [3][n],
    times [n]) with n = total bins - observer count."""
    obs = observers_array(observers)
    potential = np.ascontiguousarray(potential, dtype=np.float64)
    total = int(obs["number_of_bins"].sum())
    if potential.shape != (3, total):
        raise ValueError(f"potential must be [3][{total}], got {potential.shape}")
    n = total - obs.size
    field = np.zeros((3, n))
    times = np.zeros(n)
    _check(library().c8mi_radio_zhs_electric_field(
        obs.ctypes.data_as(c_void_p), c_uint64(obs.size), _dptr(potential), _dptr(field), _dptr(times)))
    return field, times


class DeviceArray:
    """A device allocation made through the C ABI (hipMalloc)."""

    def __init__(self, nbytes: int):
        self.pointer = c_void_p()
        self.nbytes = int(nbytes)
        _check(library().c8mi_device_malloc(byref(self.pointer), c_uint64(self.nbytes)))

    @classmethod
    def from_host(cls, array: np.ndarray) -> "DeviceArray":
        array = np.ascontiguousarray(array)
        device = cls(array.nbytes)
        _check(library().c8mi_memcpy_htod(device.pointer, array.ctypes.data_as(c_void_p),
                                          c_uint64(array.nbytes)))
        return device

    def to_host(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _check(library().c8mi_memcpy_dtoh(out.ctypes.data_as(c_void_p), self.pointer,
                                          c_uint64(out.nbytes)))
        return out

    def free(self) -> None:
        if self.pointer:
            _check(library().c8mi_device_free(self.pointer))
            self.pointer = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count() -> int:
    n = c_int(0)
    _check(library().c8mi_device_count(byref(n)))
    return n.value


def device_synchronize() -> None:
    _check(library().c8mi_device_synchronize())


class RadioAccumulator:
    """Binding of the c8mi radio handle (c8mi_radio.h; reference
    corsika::kokkos::radio::RadioAccumulator)."""

    def __init__(self, observers, propagation: Propagation, coreas: bool = True, zhs: bool = True,
                 deterministic: bool = True, fixed_point_limit: float = 1.0,
                 zhs_refinement: int = 1, zhs_observers=None, stream=None, exact: bool = False):
        self._observers = observers_array(observers)
        self._zhs_observers = (self._observers if zhs_observers is None
                               else observers_array(zhs_observers))
        self.propagation = propagation
        cfg = RadioConfig()
        cfg.coreas_enabled = 1 if coreas else 0
        cfg.zhs_enabled = 1 if zhs else 0
        cfg.deterministic = 1 if deterministic else 0
        cfg.zhs_subtrack_refinement = zhs_refinement
        cfg.exact_fp64 = 1 if exact else 0
        cfg.fixed_point_field_limit_v_per_m = fixed_point_limit
        cfg.propagation = propagation.struct
        cfg.coreas_observers = self._observers.ctypes.data_as(POINTER(RadioObserver))
        cfg.coreas_observer_count = self._observers.size if coreas else 0
        cfg.zhs_observers = self._zhs_observers.ctypes.data_as(POINTER(RadioObserver))
        cfg.zhs_observer_count = self._zhs_observers.size if zhs else 0
        self.handle = c_void_p()
        lib = library()
        _check(lib.c8mi_radio_create(byref(cfg), c_void_p(stream), byref(self.handle)))
        cb, zb, fused = c_uint64(), c_uint64(), c_int()
        _check(lib.c8mi_radio_bins(self.handle, byref(cb), byref(zb), byref(fused)))
        self.coreas_total_bins, self.zhs_total_bins, self.fused = cb.value, zb.value, bool(fused.value)

    def close(self) -> None:
        if self.handle:
            _check(library().c8mi_radio_destroy(self.handle))
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self) -> None:
        _check(library().c8mi_radio_reset(self.handle))

    def enqueue_host_tracks(self, records: np.ndarray) -> None:
        records = np.ascontiguousarray(records, dtype=LEPTON_TRANSPORT_RECORD_DTYPE)
        _check(library().c8mi_radio_enqueue_host_tracks(
            self.handle, records.ctypes.data_as(c_void_p), c_uint64(records.size)))

    def flush_host_tracks(self) -> None:
        _check(library().c8mi_radio_flush_host_tracks(self.handle))

    def accumulate_records(self, device_records: DeviceArray, count: int) -> None:
        _check(library().c8mi_radio_accumulate_records(self.handle, device_records.pointer,
                                                       c_uint64(count)))

    def accumulate_tracks(self, device_tracks, stride: int, count: int) -> None:
        pointer = device_tracks.pointer if isinstance(device_tracks, DeviceArray) else c_void_p(device_tracks)
        _check(library().c8mi_radio_accumulate_tracks(self.handle, pointer, c_uint64(stride),
                                                      c_uint64(count)))

    def download(self):
        coreas = np.zeros((3, self.coreas_total_bins)) if self.coreas_total_bins else None
        zhs = np.zeros((3, self.zhs_total_bins)) if self.zhs_total_bins else None
        counters = RadioCounters()
        _check(library().c8mi_radio_download(
            self.handle, None if coreas is None else _dptr(coreas),
            None if zhs is None else _dptr(zhs), byref(counters)))
        return coreas, zhs, counters.as_dict()

    def download_fixed(self):
        coreas = np.zeros((3, self.coreas_total_bins), dtype=np.int64) if self.coreas_total_bins else None
        zhs = np.zeros((3, self.zhs_total_bins), dtype=np.int64) if self.zhs_total_bins else None
        counters = RadioCounters()
        _check(library().c8mi_radio_download_fixed(
            self.handle, None if coreas is None else coreas.ctypes.data_as(POINTER(c_int64)),
            None if zhs is None else zhs.ctypes.data_as(POINTER(c_int64)), byref(counters)))
        return coreas, zhs, counters.as_dict()

    def set_tuning(self, chunk_tracks: int = 0, tile_bins: int = 0, target_workgroups: int = 0,
                   block_threads: int = 0, split: bool = True) -> None:
        _check(library().c8mi_radio_set_tuning(self.handle, c_uint64(chunk_tracks),
                                               c_uint32(tile_bins), c_uint32(target_workgroups),
                                               c_uint32(block_threads), c_int(1 if split else 0)))

    def enable_kernel_timing(self, enable: bool = True) -> None:
        _check(library().c8mi_radio_enable_kernel_timing(self.handle, c_int(1 if enable else 0)))

    def kernel_time(self):
        """{kind: (total ms, launches)} for kinds 'coreas_zhs', 'coreas', 'zhs'."""
        ms, launches = (c_double * 3)(), (c_uint64 * 3)()
        _check(library().c8mi_radio_kernel_time(self.handle, ms, launches))
        return {kind: (ms[k], launches[k]) for k, kind in enumerate(("coreas_zhs", "coreas", "zhs"))}


def radio_project_fixed_noalloc(observers, propagation: Propagation, device_tracks, stride: int,
                                count: int, device_coreas=None, device_zhs=None, coreas: bool = True,
                                zhs: bool = True, fixed_point_limit: float = 1.0,
                                zhs_refinement: int = 1, zhs_observers=None, exact: bool = False,
                                stream=None):
    """One-shot projection without device allocation
    (c8mi_radio_project_fixed_noalloc): int64 waveforms are written into the
    caller's device buffers (DeviceArray or raw pointers, [3][total bins] each),
    which the call clears. Returns the counters as a dict."""
    obs = observers_array(observers)
    zobs = obs if zhs_observers is None else observers_array(zhs_observers)
    cfg = RadioConfig()
    cfg.coreas_enabled = 1 if coreas else 0
    cfg.zhs_enabled = 1 if zhs else 0
    cfg.deterministic = 1
    cfg.zhs_subtrack_refinement = zhs_refinement
    cfg.exact_fp64 = 1 if exact else 0
    cfg.fixed_point_field_limit_v_per_m = fixed_point_limit
    cfg.propagation = propagation.struct
    cfg.coreas_observers = obs.ctypes.data_as(POINTER(RadioObserver))
    cfg.coreas_observer_count = obs.size if coreas else 0
    cfg.zhs_observers = zobs.ctypes.data_as(POINTER(RadioObserver))
    cfg.zhs_observer_count = zobs.size if zhs else 0

    def pointer(buffer):
        if buffer is None:
            return None
        return buffer.pointer if isinstance(buffer, DeviceArray) else c_void_p(buffer)

    counters = RadioCounters()
    _check(library().c8mi_radio_project_fixed_noalloc(
        byref(cfg), pointer(device_tracks), c_uint64(stride), c_uint64(count),
        pointer(device_coreas), pointer(device_zhs), byref(counters), c_void_p(stream)))
    return counters.as_dict()


# ---- EM: device tables, batched queries, Philox (c8mi_em.h) -----------------

RANDOM_KEY_DTYPE = np.dtype(
    {
        "names": ["seed", "shower_id", "history_id", "step_id", "process_id", "draw_id"],
        "formats": ["<u8", "<u8", "<u8", "<u8", "<u4", "<u8"],
        "offsets": [0, 8, 16, 24, 32, 40],
        "itemsize": 48,
    }
)


class EmTables:
    """Device copy of a format-v7 native table set (c8mi_em_tables)."""

    def __init__(self, tables):
        """tables: a c8mi.tables.NativeTables (host arrays + ctypes struct)."""
        self.host = tables
        self.handle = c_void_p()
        _check(library().c8mi_em_tables_create(byref(tables.struct), byref(self.handle)))

    def close(self) -> None:
        if self.handle:
            _check(library().c8mi_em_tables_destroy(self.handle))
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bytes(self) -> int:
        n = c_uint64()
        _check(library().c8mi_em_tables_bytes(self.handle, byref(n)))
        return n.value

    def query(self, queries: np.ndarray) -> np.ndarray:
        """Runs a batch of c8mi_native_query records on the device."""
        from .tables import QUERY_DTYPE, RESULT_DTYPE

        queries = np.ascontiguousarray(queries, dtype=QUERY_DTYPE)
        count = queries.size
        if count == 0:
            return np.zeros(0, dtype=RESULT_DTYPE)
        dq = DeviceArray.from_host(queries)
        dr = DeviceArray(count * RESULT_DTYPE.itemsize)
        _check(library().c8mi_em_native_queries(self.handle, dq.pointer, c_uint64(count),
                                                dr.pointer, None))
        device_synchronize()
        return dr.to_host(RESULT_DTYPE, count)


def random_draws(keys: np.ndarray):
    """(words [n,4] uint32, uniforms [n,4] float64) for a batch of random keys."""
    keys = np.ascontiguousarray(keys, dtype=RANDOM_KEY_DTYPE)
    n = keys.size
    if n == 0:
        return np.zeros((0, 4), np.uint32), np.zeros((0, 4))
    dk = DeviceArray.from_host(keys)
    dw = DeviceArray(16 * n)
    du = DeviceArray(32 * n)
    _check(library().c8mi_em_random(dk.pointer, c_uint64(n), dw.pointer, du.pointer, None))
    device_synchronize()
    return dw.to_host(np.uint32, (n, 4)), du.to_host(np.float64, (n, 4))
