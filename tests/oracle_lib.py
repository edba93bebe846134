"""ctypes access to the oracle (oracle/liboracle.so) — test infrastructure.

The oracle is the checker: only tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it. It is built from oracle/*.c by oracle/Makefile.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, byref, c_double, c_int, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

_LIB = None


class OrRadioCounters(ctypes.Structure):
    _fields_ = [
        ("coreas_contributions", c_uint64),
        ("zhs_contributions", c_uint64),
        ("zhs_subtracks", c_uint64),
        ("valid_tracks", c_uint64),
        ("fixed_point_overflows", c_uint64),
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_PATH):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        lib = ctypes.CDLL(ORACLE_PATH)
        lib.oracle_radio_nearest_bin.restype = c_double
        lib.oracle_radio_nearest_bin.argtypes = [c_double]
        lib.oracle_radio_doppler.restype = c_double
        lib.oracle_radio_doppler.argtypes = [c_double, POINTER(c_double), POINTER(c_double), c_int]
        lib.oracle_free.argtypes = [c_void_p]
        _LIB = lib
    return _LIB


def _d(a):
    return a.ctypes.data_as(POINTER(c_double))


def nearest_bin(x):
    return load().oracle_radio_nearest_bin(x)


def doppler(index, beta, emit, device_variant=False):
    b = np.ascontiguousarray(beta, dtype=np.float64)
    e = np.ascontiguousarray(emit, dtype=np.float64)
    return load().oracle_radio_doppler(index, _d(b), _d(e), 1 if device_variant else 0)


def propagate(propagation, source, destination):
    """propagation: c8mi.Propagation; returns (time, n_src, n_dst, distance, emit[3])."""
    out = np.zeros(7)
    s = np.ascontiguousarray(source, dtype=np.float64)
    d = np.ascontiguousarray(destination, dtype=np.float64)
    load().oracle_radio_propagate(byref(propagation.struct), _d(s), _d(d), _d(out))
    return out[0], out[1], out[2], out[3], out[4:7].copy()


def separate_same_bin(start, end, grid, signed_order):
    s, e = c_double(start), c_double(end)
    load().oracle_radio_separate_same_bin(byref(s), byref(e), c_double(grid), c_int(1 if signed_order else 0))
    return s.value, e.value


def build_flat_table(refractivity, inverse_step):
    r = np.ascontiguousarray(refractivity, dtype=np.float64)
    integrated = np.zeros_like(r)
    slopes = np.zeros(4)
    load().oracle_radio_build_flat_table(_d(r), c_uint64(r.size), c_double(inverse_step), _d(integrated), _d(slopes))
    return integrated, slopes


def radio_project(propagation, observers, tracks, coreas=True, zhs=True, refinement=1,
                  limit=1.0, fixed=True, floating=True, threads=1):
    """Runs the FP64 restatement. Returns dict with coreas/zhs fixed [3][total],
    floating [3][total] and counters."""
    obs = np.ascontiguousarray(observers)
    tracks = np.ascontiguousarray(tracks, dtype=np.float64)
    total = int(obs["number_of_bins"].sum())
    out = {}
    cf = np.zeros((3, total), dtype=np.int64) if (coreas and fixed) else None
    cd = np.zeros((3, total)) if (coreas and floating) else None
    zf = np.zeros((3, total), dtype=np.int64) if (zhs and fixed) else None
    zd = np.zeros((3, total)) if (zhs and floating) else None
    counters = OrRadioCounters()

    def p64(a):
        return None if a is None else a.ctypes.data_as(POINTER(c_int64))

    def pd(a):
        return None if a is None else _d(a)

    status = load().oracle_radio_project(
        byref(propagation.struct), c_uint32(refinement), c_double(limit),
        obs.ctypes.data_as(c_void_p), c_uint64(obs.size), c_int(1 if coreas else 0),
        c_int(1 if zhs else 0), _d(tracks), c_uint64(tracks.shape[1]), c_uint64(tracks.shape[1]),
        p64(cf), pd(cd), p64(zf), pd(zd), byref(counters), c_int(threads))
    assert status == 0
    out.update(coreas_fixed=cf, coreas_float=cd, zhs_fixed=zf, zhs_float=zd, counters=counters.as_dict())
    return out


def radio_decode(fixed, observers, limit=1.0, potential=False):
    obs = np.ascontiguousarray(observers)
    fixed = np.ascontiguousarray(fixed, dtype=np.int64)
    total = fixed.shape[1]
    out = np.zeros((3, total))
    load().oracle_radio_decode(fixed.ctypes.data_as(POINTER(c_int64)), c_uint64(total),
                               obs.ctypes.data_as(c_void_p), c_uint64(obs.size), c_double(limit),
                               c_int(1 if potential else 0), _d(out))
    return out


# ---- EM: Philox and native tables -------------------------------------------

REF_PATH = os.path.join(ORACLE_DIR, "_ref", "libref.so")
_REF = None


def load_ref():
    """The reference's own header functions compiled in place (oracle/ref.mk);
    None when the library was not built (no reference tree at build time)."""
    global _REF
    if _REF is None and os.path.exists(REF_PATH):
        lib = ctypes.CDLL(REF_PATH)
        lib.ref_uniform_open01.restype = c_double
        lib.ref_uniform_open01.argtypes = [c_uint64, c_uint64, c_uint64, c_uint64, c_uint32,
                                           c_uint64, c_uint32]
        _REF = lib
    return _REF


def philox(counter, key, lib=None):
    lib = lib or load()
    c = (c_uint32 * 4)(*counter)
    k = (c_uint32 * 2)(*key)
    out = (c_uint32 * 4)()
    fn = lib.oracle_philox4x32_10 if lib is _LIB else lib.ref_philox4x32_10
    fn(c, k, out)
    return list(out)


def random_words(seed, shower, history, step, process, draw, lib=None):
    lib = lib or load()
    out = (c_uint32 * 4)()
    fn = lib.oracle_random_words if lib is _LIB else lib.ref_random_words
    fn(c_uint64(seed), c_uint64(shower), c_uint64(history), c_uint64(step), c_uint32(process),
       c_uint64(draw), out)
    return list(out)


def uniform_open01(seed, shower, history, step, process, draw, lane=0):
    lib = load()
    lib.oracle_uniform_open01.restype = c_double
    return lib.oracle_uniform_open01(c_uint64(seed), c_uint64(shower), c_uint64(history),
                                     c_uint64(step), c_uint32(process), c_uint64(draw),
                                     c_uint32(lane))


def native_queries(tables, queries, reference=False):
    """Runs the oracle (or the reference probe) on a c8mi.tables.NativeTables."""
    from c8mi.tables import RESULT_DTYPE

    lib = load_ref() if reference else load()
    out = np.zeros(queries.size, dtype=RESULT_DTYPE)
    fn = lib.ref_native_queries if reference else lib.oracle_native_queries
    fn(byref(tables.struct), queries.ctypes.data_as(c_void_p), c_uint64(queries.size),
       out.ctypes.data_as(c_void_p))
    return out


# ---- EM transport (em_oracle.c) -----------------------------------------------

def em_run_shower(config, tables, environment, primaries, shower_id=1, first_history_id=2,
                  maximum_fronts=1 << 40, tracks=False):
    """Runs the oracle's serial front schedule; returns histograms (int64),
    raw counters, stats and the stored observations / fallbacks; with
    tracks=True also the radio track of every e+- step as a [9][n] array."""
    from c8mi import PARTICLE_STATE_DTYPE
    from c8mi.em import COUNTERS, FALLBACK_DTYPE, RunResult

    lib = load()
    cfg = config.struct if hasattr(config, "struct") else config
    bins = cfg.profile_bins
    hist = {k: np.zeros(bins, dtype=np.int64) for k in ("photons", "electrons", "positrons", "energy_loss")}
    counters = np.zeros(COUNTERS, dtype=np.int64)
    obs = np.zeros(cfg.observation_capacity, dtype=PARTICLE_STATE_DTYPE)
    fbs = np.zeros(cfg.fallback_capacity, dtype=FALLBACK_DTYPE)
    prim = np.ascontiguousarray(primaries, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
    stats = RunResult()
    env = environment.struct if hasattr(environment, "struct") else environment
    p = lambda a: a.ctypes.data_as(c_void_p)
    args = [byref(cfg), byref(tables.struct), byref(env), c_uint64(shower_id),
            c_uint64(first_history_id), p(prim), c_uint64(prim.size),
            c_uint64(maximum_fronts), p(hist["photons"]), p(hist["electrons"]),
            p(hist["positrons"]), p(hist["energy_loss"]), p(counters), p(obs),
            p(fbs), byref(stats)]
    track_array = None
    if tracks:
        buf, n = c_void_p(), c_uint64()
        if lib.oracle_em_run_shower_tracks(*args, byref(buf), byref(n)) != 0:
            raise RuntimeError("oracle_em_run_shower_tracks failed")
        try:
            rows = np.ctypeslib.as_array(ctypes.cast(buf, ctypes.POINTER(ctypes.c_double)),
                                         shape=(n.value * 9,)) if n.value else np.zeros(0)
            track_array = np.ascontiguousarray(rows.reshape(-1, 9).T.copy())
        finally:
            lib.oracle_free(buf)
    else:
        lib.oracle_em_run_shower(*args)
    st = stats.as_dict()
    n_obs = min(st["observations"], obs.size)
    n_fb = min(st["fallbacks"], fbs.size)
    out = dict(hist=hist, counters=counters, stats=st, observations=obs[:n_obs], fallbacks=fbs[:n_fb])
    if tracks:
        out["tracks"] = track_array
    return out


class OrHybridSide(ctypes.Structure):
    _fields_ = [("hist", ctypes.c_void_p), ("counters", ctypes.c_int64 * 64),
                ("observations", ctypes.c_void_p), ("observation_count", c_uint64),
                ("fallbacks", ctypes.c_void_p), ("fallback_count", c_uint64),
                ("tracks", ctypes.c_void_p), ("track_count", c_uint64),
                ("align_", c_uint64),  # the snapshot is 16-byte aligned in C
                ("first", ctypes.c_uint8 * 528), ("first_captured", c_uint64), ("tail_", c_uint64)]


def em_run_hybrid(config, tables, environment, primaries, shower_id=1, first_history_id=2, hybrid=None,
                  limits=None, prefix_only=False):
    """The oracle's restatement of the host handoff (oracle_em_run_hybrid):
    per side ("host" steps, "device" fronts) the int64 histograms, counters,
    records and compact tracks ([9][n]); the staged batch (prefix_only) and
    the counts."""
    from c8mi import PARTICLE_STATE_DTYPE
    from c8mi.em import FALLBACK_DTYPE, HybridResult, Limits

    lib = load()
    cfg = config.struct if hasattr(config, "struct") else config
    env = environment.struct if hasattr(environment, "struct") else environment
    bins = cfg.profile_bins
    sides, keep = {}, []
    for name in ("host", "device"):
        hist = np.zeros(4 * bins, dtype=np.int64)
        obs = np.zeros(cfg.observation_capacity, dtype=PARTICLE_STATE_DTYPE)
        fbs = np.zeros(cfg.fallback_capacity, dtype=FALLBACK_DTYPE)
        side = OrHybridSide()
        side.hist = hist.ctypes.data_as(c_void_p)
        side.observations = obs.ctypes.data_as(c_void_p)
        side.fallbacks = fbs.ctypes.data_as(c_void_p)
        sides[name] = (side, hist, obs, fbs)
    prim = np.ascontiguousarray(primaries, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
    staged, n_staged = c_void_p(), c_uint64()
    result = HybridResult()
    lim = None
    if limits is not None:
        lim = Limits()
        for k, v in limits.items():
            setattr(lim, k, v)
    status = lib.oracle_em_run_hybrid(
        byref(cfg), byref(tables.struct), byref(env), c_uint64(shower_id), c_uint64(first_history_id),
        prim.ctypes.data_as(c_void_p), c_uint64(prim.size), byref(hybrid) if hybrid is not None else None,
        byref(lim) if lim is not None else None, ctypes.c_int(1 if prefix_only else 0), byref(sides["host"][0]),
        byref(sides["device"][0]), byref(staged), byref(n_staged), byref(result))
    out = {"result": result.as_dict()}
    for name, (side, hist, obs, fbs) in sides.items():
        tracks = np.zeros((9, 0))
        if side.tracks:
            if side.track_count:
                rows = np.ctypeslib.as_array(ctypes.cast(side.tracks, ctypes.POINTER(ctypes.c_double)),
                                             shape=(side.track_count * 9,))
                tracks = np.ascontiguousarray(rows.reshape(-1, 9).T.copy())
            lib.oracle_free(side.tracks)
        h = hist.reshape(4, bins)
        from c8mi.em import FIRST_INTERACTION_DTYPE
        first = np.frombuffer(bytes(side.first), dtype=FIRST_INTERACTION_DTYPE)[0].copy() if side.first_captured else None
        out[name] = dict(first=first, hist=dict(photons=h[0].copy(), electrons=h[1].copy(), positrons=h[2].copy(),
                                   energy_loss=h[3].copy()),
                         counters=np.array(side.counters, dtype=np.int64),
                         observations=obs[:min(side.observation_count, obs.size)],
                         fallbacks=fbs[:min(side.fallback_count, fbs.size)], tracks=tracks)
    batch = np.zeros(0, dtype=PARTICLE_STATE_DTYPE)
    if staged:
        if n_staged.value:
            raw = np.ctypeslib.as_array(ctypes.cast(staged, ctypes.POINTER(ctypes.c_uint8)),
                                        shape=(n_staged.value * PARTICLE_STATE_DTYPE.itemsize,))
            batch = raw.view(PARTICLE_STATE_DTYPE).copy()
        lib.oracle_free(staged)
    out["staged"] = batch
    if status != 0:
        raise RuntimeError(f"oracle_em_run_hybrid failed ({status})")
    return out


def em_first_interaction(config, tables, environment, primaries, shower_id=1, first_history_id=2):
    """The oracle's first-interaction snapshot of a shower (or None) and its stats."""
    from c8mi import PARTICLE_STATE_DTYPE
    from c8mi.em import FIRST_INTERACTION_DTYPE, RunResult

    lib = load()
    cfg = config.struct if hasattr(config, "struct") else config
    env = environment.struct if hasattr(environment, "struct") else environment
    prim = np.ascontiguousarray(primaries, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
    out = np.zeros(1, dtype=FIRST_INTERACTION_DTYPE)
    captured = c_int(0)
    stats = RunResult()
    if lib.oracle_em_first_interaction(byref(cfg), byref(tables.struct), byref(env), c_uint64(shower_id),
                                       c_uint64(first_history_id), prim.ctypes.data_as(c_void_p),
                                       c_uint64(prim.size), out.ctypes.data_as(c_void_p), byref(captured),
                                       byref(stats)) != 0:
        raise RuntimeError("oracle_em_first_interaction failed")
    return (out[0] if captured.value else None), stats.as_dict()


VERTEX_DTYPE = np.dtype([
    ("kind", "<i4"), ("pid", "<i4"), ("process", "<i4"), ("accepted", "<u4"), ("count", "<u4"),
    ("pad", "<u4"), ("energy", "<f8"), ("mass", "<f8"), ("target_mass", "<f8"), ("direction", "<f8", 3),
    ("v", "<f8"), ("rate_start", "<f8"), ("rate_vertex", "<f8"), ("process_uniform", "<f8"),
    ("child_pid", "<i4", 3), ("pad2", "<i4"), ("child_energy", "<f8", 3), ("child_direction", "<f8", (3, 3)),
    ("thin_status", "<u4"), ("keep", "<u4"), ("parent_weight", "<f8"), ("child_weight", "<f8", 2)])

SELECTION_DTYPE = np.dtype([
    ("status", "<i4"), ("process", "<i4"), ("component", "<u8"), ("interaction_grammage", "<f8"),
    ("total_rate", "<f8"), ("energy_fraction", "<f8"), ("loss_quantile", "<f8"), ("process_uniform", "<f8"),
    ("fallback", "<i4"), ("fallback_reason", "<i4")])


def em_vertices(config, tables, environment, primaries, shower_id=1, first_history_id=2):
    """Vertex log of an oracle shower (oracle_em_run_shower_vertices): lepton
    vertex attempts (kind 0) and materialised final states (kind 1)."""
    from c8mi import PARTICLE_STATE_DTYPE
    from c8mi.em import RunResult

    lib = load()
    cfg = config.struct if hasattr(config, "struct") else config
    env = environment.struct if hasattr(environment, "struct") else environment
    prim = np.ascontiguousarray(primaries, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
    buf, n, stats = c_void_p(), c_uint64(), RunResult()
    if lib.oracle_em_run_shower_vertices(byref(cfg), byref(tables.struct), byref(env), c_uint64(shower_id),
                                         c_uint64(first_history_id), prim.ctypes.data_as(c_void_p),
                                         c_uint64(prim.size), byref(buf), byref(n), byref(stats)) != 0:
        raise RuntimeError("oracle_em_run_shower_vertices failed")
    try:
        raw = ctypes.string_at(buf, n.value * VERTEX_DTYPE.itemsize) if n.value else b""
    finally:
        lib.oracle_free(buf)
    return np.frombuffer(raw, dtype=VERTEX_DTYPE).copy(), stats.as_dict()


def em_select(config, tables, environment, particles, shower_id=1):
    """select_discrete_interaction of each particle (oracle_em_select)."""
    from c8mi import PARTICLE_STATE_DTYPE

    lib = load()
    cfg = config.struct if hasattr(config, "struct") else config
    env = environment.struct if hasattr(environment, "struct") else environment
    prim = np.ascontiguousarray(particles, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
    out = np.zeros(prim.size, dtype=SELECTION_DTYPE)
    for i in range(prim.size):
        lib.oracle_em_select(byref(cfg), byref(tables.struct), byref(env), c_uint64(shower_id),
                             prim[i:i + 1].ctypes.data_as(c_void_p), out[i:i + 1].ctypes.data_as(c_void_p))
    return out


def em_run_batch_profiles(config, tables, environment, primaries, first_shower=1, threads=1):
    """Independent single-primary showers on `threads` host threads; per
    shower the stats, the four int64 histograms [count][4][bins] and the raw
    counters [count][COUNTERS]."""
    from c8mi import PARTICLE_STATE_DTYPE
    from c8mi.em import COUNTERS, RunResult

    lib = load()
    cfg = config.struct if hasattr(config, "struct") else config
    env = environment.struct if hasattr(environment, "struct") else environment
    prim = np.ascontiguousarray(primaries, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
    bins = max(cfg.profile_bins, 1)
    stats = (RunResult * prim.size)()
    hist = np.zeros((prim.size, 4, bins), dtype=np.int64)
    counters = np.zeros((prim.size, COUNTERS), dtype=np.int64)
    if lib.oracle_em_run_batch_profiles(byref(cfg), byref(tables.struct), byref(env), c_uint64(first_shower),
                                        prim.ctypes.data_as(c_void_p), c_uint64(prim.size), ctypes.c_int(threads),
                                        stats, hist.ctypes.data_as(c_void_p),
                                        counters.ctypes.data_as(c_void_p)) != 0:
        raise RuntimeError("oracle_em_run_batch_profiles failed")
    return [s.as_dict() for s in stats], hist, counters


def em_bucket_key(particle) -> int:
    """oracle_em_bucket_key: the reference's 50-bit wavefront bucket key."""
    from c8mi import PARTICLE_STATE_DTYPE

    lib = load()
    lib.oracle_em_bucket_key.restype = ctypes.c_uint64
    p = np.ascontiguousarray(particle, dtype=PARTICLE_STATE_DTYPE).reshape(-1)[:1]
    return int(lib.oracle_em_bucket_key(p.ctypes.data_as(c_void_p)))


class lepton_reservation:
    """Context manager: the oracle's lepton history reservation
    max(3 n, min(chunk, multiplier * 3 n)) inside the block (the reference's
    64 and 2^20 outside)."""

    def __init__(self, multiplier, chunk=1 << 20):
        self.multiplier, self.chunk = multiplier, chunk

    def __enter__(self):
        load().oracle_em_set_lepton_reservation(c_uint64(self.multiplier), c_uint64(self.chunk))
        return self

    def __exit__(self, *exc):
        load().oracle_em_set_lepton_reservation(c_uint64(0), c_uint64(0))
        return False


def em_run_batch(config, tables, environment, primaries, first_shower=1, threads=1,
                 maximum_fronts=1 << 40):
    """Independent single-primary showers on `threads` host threads; returns
    the list of per-shower statistics dicts."""
    from c8mi import PARTICLE_STATE_DTYPE
    from c8mi.em import RunResult

    lib = load()
    cfg = config.struct if hasattr(config, "struct") else config
    prim = np.ascontiguousarray(primaries, dtype=PARTICLE_STATE_DTYPE).reshape(-1)
    stats = (RunResult * prim.size)()
    env = environment.struct if hasattr(environment, "struct") else environment
    rc = lib.oracle_em_run_batch(byref(cfg), byref(tables.struct), byref(env), c_uint64(first_shower),
                                 prim.ctypes.data_as(c_void_p), c_uint64(prim.size),
                                 c_uint64(maximum_fronts), ctypes.c_int(threads), stats)
    if rc != 0:
        raise RuntimeError("oracle_em_run_batch failed")
    return [s.as_dict() for s in stats]


def _env_ptr(environment):
    return byref(environment.struct if hasattr(environment, "struct") else environment)


def em_env_query(environment, position, direction, node, reference=False):
    lib = load_ref() if reference else load()
    fn = lib.ref_em_environment_query if reference else lib.oracle_em_environment_query
    out_node, dens = ctypes.c_int32(), c_double()
    p = (c_double * 3)(*position)
    d = None if direction is None else (c_double * 3)(*direction)
    st = fn(_env_ptr(environment), p, d, ctypes.c_int32(node), byref(out_node), byref(dens))
    return st, out_node.value, dens.value


def em_boundary(environment, position, direction, node, reference=False):
    lib = load_ref() if reference else load()
    fn = lib.ref_em_boundary if reference else lib.oracle_em_boundary
    dist, nxt = c_double(), ctypes.c_int32()
    st = fn(_env_ptr(environment), (c_double * 3)(*position), (c_double * 3)(*direction),
            ctypes.c_int32(node), byref(dist), byref(nxt))
    return st, dist.value, nxt.value


def em_grammage(environment, node, position, direction, value, arclength=False, reference=False):
    lib = load_ref() if reference else load()
    name = ("ref_em_" if reference else "oracle_em_") + ("arclength" if arclength else "grammage")
    out = c_double()
    st = getattr(lib, name)(_env_ptr(environment), ctypes.c_int32(node), (c_double * 3)(*position),
                            (c_double * 3)(*direction), c_double(value), byref(out))
    return st, out.value


def em_lpm(snapshot, comp, e_mev, x, rho, brems=False, reference=False):
    lib = load_ref() if reference else load()
    name = ("ref_em_" if reference else "oracle_em_") + ("brems_lpm" if brems else "pair_lpm")
    out = c_double()
    st = getattr(lib, name)(byref(snapshot), c_uint64(comp), c_double(e_mev), c_double(x), c_double(rho),
                            byref(out))
    return st, out.value


def em_thinning(cfg, pe, pw, e1, e2, u1, u2, reference=False):
    lib = load_ref() if reference else load()
    fn = lib.ref_em_thinning if reference else lib.oracle_em_thinning
    keep, w1, w2 = c_uint32(), c_double(), c_double()
    st = fn(byref(cfg), c_double(pe), c_double(pw), c_double(e1), c_double(e2), c_double(u1), c_double(u2),
            byref(keep), byref(w1), byref(w2))
    return st, keep.value, w1.value, w2.value


def em_pair_split(e_gev, x, reference=False):
    lib = load_ref() if reference else load()
    fn = lib.ref_em_pair_split if reference else lib.oracle_em_pair_split
    s, d = c_double(), c_double()
    ok = fn(c_double(e_gev), c_double(x), byref(s), byref(d))
    return ok, s.value, d.value


def em_scalar(name, *args):
    """Calls a double-returning or status-returning oracle helper."""
    lib = load()
    fn = getattr(lib, name)
    if name == "oracle_em_sauter_cosine":
        fn.restype = c_double
        return fn(*[c_double(a) for a in args])
    outs = [c_double(), c_double()] if name == "oracle_em_compton_cosines" else [c_double()]
    st = fn(*[c_double(a) for a in args], *[byref(o) for o in outs])
    return (st, *[o.value for o in outs])


# ---- Molière (oracle/em_oracle.c; reference MoliereScattering.hpp via oracle/_ref) ----

def moliere_functions(snapshot, x, reference=False):
    """f1, f2 and their integrals at x: oracle (4 separate functions) or the
    reference header."""
    if reference:
        out = (c_double * 4)()
        load_ref().ref_moliere_functions(byref(snapshot), c_double(x), out)
        return tuple(out)
    lib = load()
    vals = []
    for name in ("oracle_moliere_f1", "oracle_moliere_f2", "oracle_moliere_integral_f1",
                 "oracle_moliere_integral_f2"):
        fn = getattr(lib, name)
        fn.restype = c_double
        vals.append(fn(byref(snapshot), c_double(x)))
    return tuple(vals)


def moliere_table_functions(snapshot, table, x, reference=False):
    """f1, f2 and their integrals at x through an interpolation table: oracle or
    the reference header with a view over the same table."""
    out = (c_double * 4)()
    if reference:
        load_ref().ref_moliere_interpolated_functions(byref(snapshot), byref(table), c_double(x), out)
    else:
        load().oracle_moliere_interpolated_functions(byref(snapshot), byref(table), c_double(x), out)
    return tuple(out)


def moliere_table_angle(snapshot, table, grammage, e_mev, u1, u2, reference=False):
    it, angle = ctypes.c_uint32(), c_double()
    fn = (load_ref().ref_moliere_angle_interpolated if reference
          else load().oracle_moliere_angle_interpolated)
    st = fn(byref(snapshot), byref(table) if table is not None else None, c_double(grammage), c_double(e_mev),
            c_double(u1), c_double(u2), byref(it), byref(angle))
    return st, it.value, angle.value


def moliere_scalar(name, value, reference=False):
    lib = load_ref() if reference else load()
    fn = getattr(lib, ("ref_moliere_" if reference else "oracle_moliere_") + name)
    fn.restype = c_double
    return fn(c_double(value))


def moliere_angle(snapshot, grammage, e_mev, u1, u2, reference=False):
    lib = load_ref() if reference else load()
    fn = lib.ref_moliere_angle if reference else lib.oracle_moliere_angle
    its, angle = ctypes.c_uint32(), c_double()
    st = fn(byref(snapshot), c_double(grammage), c_double(e_mev), c_double(u1), c_double(u2), byref(its),
            byref(angle))
    return st, its.value, angle.value


def moliere_direction(start, tracked, angle, azimuth_uniform, reference=False):
    lib = load_ref() if reference else load()
    fn = lib.ref_moliere_direction if reference else lib.oracle_moliere_direction
    out = (c_double * 3)()
    st = fn((c_double * 3)(*start), (c_double * 3)(*tracked), c_double(angle), c_double(azimuth_uniform), out)
    return st, tuple(out)


def quartic_roots(p, q, r):
    """Real roots of x^4 + p x^2 + q x + r (stable_quartic_roots), ascending."""
    out = (c_double * 4)()
    n = load().oracle_quartic_roots(c_double(p), c_double(q), c_double(r), out)
    return [out[i] for i in range(n)]


def quadratic_roots(a, b, c):
    out = (c_double * 4)()
    n = load().oracle_quadratic_roots(c_double(a), c_double(b), c_double(c), out)
    return [out[i] for i in range(n)]


def _state_ptr(p):
    return np.ascontiguousarray(np.atleast_1d(p)).ctypes.data_as(c_void_p)


def magnetic_step_limit(p, mass_gev, charge, field, max_deflection=0.2):
    """(status, distance_m, gyroradius_m) of maximum_uniform_magnetic_step."""
    out = (c_double * 2)()
    arr = np.ascontiguousarray(np.atleast_1d(p))
    st = load().oracle_magnetic_step_limit(arr.ctypes.data_as(c_void_p), c_double(mass_gev), c_double(charge),
                                           (c_double * 3)(*field), c_double(max_deflection), out)
    return st, out[0], out[1]


def magnetic_advance(p, mass_gev, charge, field, distance):
    """(status, end state) of advance_uniform_magnetic_field."""
    arr = np.ascontiguousarray(np.atleast_1d(p))
    end = np.zeros(1, dtype=arr.dtype)
    st = load().oracle_magnetic_advance(arr.ctypes.data_as(c_void_p), c_double(mass_gev), c_double(charge),
                                        (c_double * 3)(*field), c_double(distance), end.ctypes.data_as(c_void_p))
    return st, end[0]


def lepton_boundary(environment, p, mass_gev, charge, field):
    """(status, boundary distance, next node, observation distance): status is
    the magnetic limit's, -1 for a failed boundary."""
    arr = np.ascontiguousarray(np.atleast_1d(p))
    out = (c_double * 3)()
    st = load().oracle_lepton_boundary(_env_ptr(environment), arr.ctypes.data_as(c_void_p), c_double(mass_gev),
                                       c_double(charge), (c_double * 3)(*field), out)
    return st, out[0], int(out[1]), out[2]
