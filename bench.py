#!/usr/bin/env python3
"""Benchmark of the MI355X radio projection (c8mi) on the standalone radio
workload: 2^20 straight e+- segments along a 60 deg zenith axis
(Gaisser-Hillas depth, Xmax 750 g/cm^2, lambda 60 g/cm^2, lateral sigma 40 m,
0.5-5 m, beta 0.99-1, 20 % negative excess), 160 ground antennas on a star of
8 arms x 20 radii (25-500 m), 2048 bins at 0.2 ns, n(h) = 1 + 3.25e-4
exp(-h / 8 km), CoREAS + ZHS endpoint formulae, int64 fixed-point waveforms.

One step = one pass of the projection over the whole resident track list
(reset of the waveforms + CoREAS+ZHS projection of 2^20 x 160 pairs).
Prints one JSON line (see --help for options).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "corsika8-kokkos_amd"))

import numpy as np  # noqa: E402

METRIC = "EM particle-steps/s; radio track·antenna/s at ≤1e-3 err; p99 shower latency"
UNIT = "track·antenna updates/s"
# Algorithmic FP64 operations per track-antenna update of the fused CoREAS+ZHS
# projection (add/mul/div/sqrt = 1, fma = 2), the per-unit figure of the
# reference's RadioProjection.inl: CoREAS ~130 + ZHS ~155.
FLOPS_PER_UPDATE = {"coreas_zhs": 285.0, "zhs": 155.0, "coreas": 130.0}
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# VALU issue: 256 CUs x 4 SIMDs, each a wave64 FP64/FP32 instruction every 4 clocks, at 2.4 GHz
VALU_ISSUE_PEAK_G = 256 * 2.4
TRACK_BYTES = 72


def parse():
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--tracks", type=int, default=1 << 20)
    p.add_argument("--mode", choices=("coreas_zhs", "zhs", "coreas"), default="coreas_zhs")
    p.add_argument("--cpu-seconds", type=float, default=8.0,
                   help="target wall time of the CPU-baseline sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-error-check", action="store_true")
    p.add_argument("--copy-probe", type=int, default=0,
                   help="only run a device-to-device copy of this many bytes (PMC calibration)")
    p.add_argument("--chunk-tracks", type=int, default=0)
    p.add_argument("--tile-bins", type=int, default=0)
    p.add_argument("--target-workgroups", type=int, default=0)
    p.add_argument("--block-threads", type=int, default=0)
    p.add_argument("--split", action="store_true",
                   help="run CoREAS and ZHS as two launches instead of one fused launch")
    p.add_argument("--exact", action="store_true",
                   help="reference FP64 operation order (bit-identical to the FP64 oracle)")
    p.add_argument("--em-showers", type=int, default=8192,
                   help="timed 100 GeV showers of the EM pool (0 disables the EM section)")
    p.add_argument("--em-plain-showers", type=int, default=20,
                   help="timed 100 GeV showers of the one-at-a-time EM form")
    p.add_argument("--em-slots", type=int, default=8192, help="concurrent showers of the EM pools")
    p.add_argument("--em-moliere-fast-iterations", type=int, default=5,
                   help="Newton iterations of a Moliere angle in the step kernel before the sample moves to "
                        "the tail kernel (0: the library default; results do not depend on it)")
    p.add_argument("--em-pools", type=int, default=2,
                   help="backends the EM pool's slots are split over (own stream and host thread each)")
    p.add_argument("--em-warmup-showers", type=int, default=64, help="untimed showers before the EM pool run")
    p.add_argument("--em-no-moliere", action="store_true",
                   help="EM without Moliere multiple scattering (on by default, as in the reference's production)")
    p.add_argument("--em-analytic-moliere", action="store_true",
                   help="Moliere with the analytical correction functions instead of the interpolation table")
    p.add_argument("--em-only", action="store_true", help="run only the EM section")
    p.add_argument("--em-no-prefix", action="store_true",
                   help="EM showers start on the device from the primary (no CPU prefix / handoff)")
    p.add_argument("--em-handoff-showers", type=int, default=4,
                   help="showers run one at a time through the host handoff driver (0 skips)")
    p.add_argument("--em-radio-showers", type=int, default=2,
                   help="timed showers of the EM + radio section (0 skips it)")
    p.add_argument("--em-radio-energy", type=float, default=1.0e4, help="primary electron energy (GeV)")
    p.add_argument("--em-radio-only", action="store_true", help="run only the EM + radio section")
    p.add_argument("--serve-showers", type=int, default=160,
                   help="requests of the serving-stream section (0 skips it)")
    p.add_argument("--serve-rate", type=float, default=40.0, help="Poisson arrival rate of the stream (1/s)")
    p.add_argument("--serve-slots", type=int, default=256, help="pool slots of the serving section")
    p.add_argument("--serve-plain-showers", type=int, default=24,
                   help="first requests served again alone, pooled and one at a time (latency comparison)")
    p.add_argument("--serve-only", action="store_true", help="run only the serving-stream section")
    p.add_argument("--ice-energy", type=float, default=1.0e6, help="in-ice primary electron energy (GeV)")
    p.add_argument("--ice-only", action="store_true", help="run only the in-ice section")
    return p.parse_args()


def copy_probe(nbytes):
    import c8mi

    src = c8mi.DeviceArray(nbytes)
    dst = c8mi.DeviceArray(nbytes)
    for _ in range(3):
        c8mi._check(c8mi.library().c8mi_memcpy_dtod(dst.pointer, src.pointer, c8mi.c_uint64(nbytes)))
    c8mi.device_synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}))


def cpu_baseline(tracks, observers, propagation, mode, target_seconds):
    """FP64 oracle over a bounded prefix of the same track list, all host
    cores (OpenMP), timed as updates/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    threads = max(1, min(threads, os.cpu_count() or 1))
    coreas = mode in ("coreas_zhs", "coreas")
    zhs = mode in ("coreas_zhs", "zhs")
    n = min(4096, tracks.shape[1])
    while True:
        sample = np.ascontiguousarray(tracks[:, :n])
        t0 = time.perf_counter()
        oracle_lib.radio_project(propagation, observers, sample, coreas=coreas, zhs=zhs,
                                 floating=False, threads=threads)
        dt = time.perf_counter() - t0
        if dt >= target_seconds or n >= tracks.shape[1]:
            break
        n = min(tracks.shape[1], int(n * max(2.0, 1.2 * target_seconds / max(dt, 1e-3))))
    updates = n * observers.size
    return {"value": updates / dt, "unit": UNIT, "cores": threads, "kind": "port",
            "sample": f"first {n} of the {tracks.shape[1]} tracks x {observers.size} antennas "
                      f"({updates:.3e} updates, {dt:.2f} s wall), FP64 oracle, OpenMP"}


def field_error(tracks, observers, propagation, mode, limit, exact):
    """Relative L2 error of the decoded GPU waveforms against the unquantised
    FP64 oracle on 8 of the antennas (every 20th), full track list."""
    import c8mi

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    subset = np.ascontiguousarray(observers[::20])
    coreas = mode in ("coreas_zhs", "coreas")
    zhs = mode in ("coreas_zhs", "zhs")
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0")) or 16, os.cpu_count() or 1))
    ref = oracle_lib.radio_project(propagation, subset, tracks, coreas=coreas, zhs=zhs,
                                   fixed=False, threads=threads)
    acc = c8mi.RadioAccumulator(subset, propagation, coreas=coreas, zhs=zhs,
                                fixed_point_limit=limit, exact=exact)
    dev = c8mi.DeviceArray.from_host(tracks)
    acc.accumulate_tracks(dev, tracks.shape[1], tracks.shape[1])
    c, z, _ = acc.download()
    acc.close()
    dev.free()
    errors = {}
    if coreas:
        errors["coreas"] = float(np.linalg.norm(c - ref["coreas_float"]) / np.linalg.norm(ref["coreas_float"]))
    if zhs:
        errors["zhs"] = float(np.linalg.norm(z - ref["zhs_float"]) / np.linalg.norm(ref["zhs_float"]))
    return errors


EM_BYTES_PER_STEP = 240  # 120 B SoA state (116 B particle + 4 B pool slot) read + written per step


def em_cpu_baseline(tables, atmosphere, energy, target_seconds, moliere=True, moliere_table=True):
    """Oracle EM showers (the serial C restatement, one shower per host thread,
    OpenMP over showers on all host cores) for about `target_seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from c8mi import airtables, em

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    threads = max(1, min(threads, os.cpu_count() or 1))
    cfg = em.make_config(airtables, atmosphere=atmosphere, moliere=moliere,
                         moliere_table=True if moliere_table else None)
    prim = em.particle(22, energy, (0.0, 0.0, 112750.0), (0.0, 0.0, -1.0), history_id=1)
    count = threads
    first = 1000
    while True:
        t0 = time.perf_counter()
        stats = oracle_lib.em_run_batch(cfg, tables, atmosphere, np.stack([prim] * count),
                                        first_shower=first, threads=threads)
        dt = time.perf_counter() - t0
        if dt >= 0.5 * target_seconds:
            break
        first += count
        count = int(count * max(2.0, 1.2 * target_seconds / max(dt, 1e-3)))
    steps = sum(st["photon_steps"] + st["lepton_steps"] for st in stats)
    return {"value": steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{count} x {energy:g} GeV vertical photon showers, one per thread, on {threads} "
                      f"threads (oracle, serial C per shower, OpenMP over showers; {steps} steps, "
                      f"{dt:.1f} s wall)"}


def em_measured_traffic(kernels):
    """HBM bytes per step of the `kernels` together (each one's bytes per step of
    its own kind, summed) from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "em_traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    entries = [data.get("kernels", {}).get(k) for k in kernels]
    if not entries or any(e is None for e in entries):
        return None
    return round(sum(e["traffic_bytes_per_step"] for e in entries), 1)


def radio_valu_issue(kernel, updates, kernel_ms):
    """The projection kernel's share of the VALU issue rate: its counted VALU
    wave-instructions per track-antenna update (committed rocprofv3 counter
    summary) times the launch's updates over the live launch time, or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "radio_valu.json")) as f:
            per = json.load(f)["kernels"][kernel]["valu_per_update"]
    except (OSError, ValueError, KeyError):
        return None
    rate = per * updates / (kernel_ms * 1e-3) / 1e9
    return {"valu_per_update": per, "achieved": rate, "peak": VALU_ISSUE_PEAK_G,
            "unit": "G VALU wave-instructions/s", "frac": rate / VALU_ISSUE_PEAK_G}


def em_valu_per_unit():
    """VALU wave-instructions per unit of work (photon step, e+- step, e+-
    interaction) of each EM kernel from the committed rocprofv3 counter summary
    (SQ_INSTS_VALU over the profiled run's device steps), or {}."""
    try:
        with open(os.path.join(ROOT, "profiles", "em_valu.json")) as f:
            return {k: v["valu_per_unit"] for k, v in json.load(f)["kernels"].items()}
    except (OSError, ValueError, KeyError):
        return {}


def em_showers(energy, count, first_shower, atm=None, pid=22):
    from c8mi import em

    prim = em.particle(pid, energy, (0.0, 0.0, 112750.0), (0.0, 0.0, -1.0), history_id=1)
    return [dict(primary=prim, shower_id=first_shower + k, arrival_s=0.0,
                 footprint=em.footprint_estimate(energy)) for k in range(count)]


def em_handoff_showers(tables, atm, energy, count, moliere, table):
    """`count` showers one after another through the host handoff
    (c8mi_em_hybrid_run, the reference's constants): throughput, latency and
    the host / device split of the wall time."""
    from c8mi import airtables, em

    cfg = em.make_config(airtables, atmosphere=atm, slots=1, queue_capacity=1 << 20, observation_capacity=1 << 14,
                         fallback_capacity=1 << 10, moliere=moliere, moliere_table=table)
    dev = em.EmBackend(cfg, tables, atm)
    stepper = em.ScalarStepper(cfg, tables, atm)
    cascade = em.HybridCascade(dev, stepper)
    prim = em.particle(22, energy, (0.0, 0.0, 112750.0), (0.0, 0.0, -1.0), history_id=1)
    cascade.run(prim, 7000)  # warm-up
    runs = []
    for k in range(count):
        t0 = time.perf_counter()
        r = cascade.run(prim, 7001 + k)
        runs.append((time.perf_counter() - t0, r))
    cascade.close()
    stepper.close()
    dev.close()
    wall = sum(t for t, _ in runs)
    steps = sum(r["cpu_photon_steps"] + r["cpu_lepton_steps"] + r["device_photon_steps"] + r["device_lepton_steps"]
                for _, r in runs)
    return {"value": steps / wall, "showers": count,
            "latency_ms": {"p50": float(np.percentile([t for t, _ in runs], 50) * 1e3),
                           "p99": float(np.percentile([t for t, _ in runs], 99) * 1e3)},
            "host_step_s": sum(r["host_seconds"] for _, r in runs) / count,
            "device_call_s": sum(r["device_seconds"] for _, r in runs) / count,
            "calls_per_shower": sum(r["photon_calls"] + r["lepton_calls"] for _, r in runs) / count,
            "expansion_rounds_per_shower": sum(r["small_batch_expansions"] for _, r in runs) / count,
            "host_steps_per_shower": sum(r["cpu_photon_steps"] + r["cpu_lepton_steps"] for _, r in runs) / count}


def em_section(showers, plain_showers, slots, cpu_seconds, no_cpu_baseline, energy=100.0, warmup_showers=64,
               moliere=True, moliere_table=True, pools=1, fast_iterations=0, prefix=True, handoff_showers=4):
    """EM particle-steps/s on the plain 100 GeV vertical photon shower in the
    USStdBK atmosphere (1 MeV cut, no thinning, FP64). With `prefix` every
    shower starts on the host as the reference's router starts it: scalar
    expansion rounds of the host stepper up to the first device call (the
    16384-particle handoff batch, at most 8 rounds; c8mi_em_hybrid_prefixes on
    host threads, inside the timed region), then the handed-off batch runs on
    the device with the router's calls (a below-16384 checkpoint after every
    first front). Without it the whole cascade runs on the device from the
    primary on. Forms: the pool (`showers` showers served through `slots`
    concurrent slots, continuous batching; the throughput figure), one shower
    at a time on the device (`plain_showers` showers, per-shower latency) and
    `handoff_showers` showers driven one at a time through the host handoff
    (c8mi_em_hybrid_run: every device call returns to the host). With pools >
    1 the slots are split over that many backends, each on its own stream and
    served from its own host thread, so one pool's kernels run while
    another's host turns its front report around."""
    import threading

    from c8mi import airtables, em

    tables = airtables.build_air_tables()
    atm = em.Atmosphere()
    pool_slots = max(1, slots // pools)
    capacity = 1 << 23 if pool_slots <= 1024 else 1 << 24 if pool_slots <= 2048 else 1 << 25
    table = True if moliere_table else None
    handoff = 16384  # BackendConfig minBatchSize: the router's handoff batch and the calls' minimum
    cfg = em.make_config(airtables, atmosphere=atm, slots=pool_slots, queue_capacity=capacity,
                         observation_capacity=1 << 12, fallback_capacity=1 << 8, moliere=moliere,
                         moliere_table=table, moliere_fast_iterations=fast_iterations,
                         resident_minimum_batch=handoff if prefix else 0)
    devs = [em.EmBackend(cfg, tables, atm) for _ in range(pools)]
    budget = 0.4 * capacity
    zero = {"photon_steps": 0, "lepton_steps": 0, "lepton_interactions": 0}
    warm_totals = zero
    if warmup_showers:
        for dev in devs:
            _, warm_totals = dev.serve(em_showers(energy, warmup_showers, 1), budget, realtime=False)
    for dev in devs:
        dev.enable_kernel_timing(True)
    requests = em_showers(energy, showers, 1000)
    parts = [requests[k::pools] for k in range(pools)]
    outcome = [None] * pools
    host_threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0")) or 16, os.cpu_count() or 1))

    def serve_part(k):
        outcome[k] = devs[k].serve(parts[k], budget, realtime=False)

    workers = [threading.Thread(target=serve_part, args=(k,)) for k in range(pools)]
    t0 = time.perf_counter()
    host_ledgers, host_totals, prefix_s = None, None, 0.0
    if prefix:
        prims = np.concatenate([r["primary"].reshape(1) for r in requests])
        batches, next_ids, host_ledgers, host_totals = em.hybrid_prefixes(
            cfg, tables, atm, prims, [r["shower_id"] for r in requests], threads=host_threads,
            hybrid=em.hybrid_config(minimum_batch=handoff))
        for r, b, h in zip(requests, batches, next_ids):
            r["particles"], r["first_history_id"] = b, h
        prefix_s = time.perf_counter() - t0
    for w in workers:
        w.start()
    for w in workers:
        w.join()
    elapsed = time.perf_counter() - t0
    if any(o is None for o in outcome):
        raise RuntimeError("an EM pool failed")
    timings = [dev.kernel_time() for dev in devs]
    timing = {k: (sum(t[k][0] for t in timings), sum(t[k][1] for t in timings)) for k in timings[0]}
    for dev in devs:
        dev.close()
    results = [r for o in outcome for r in o[0]]
    totals = {k: sum(o[1][k] for o in outcome) if k != "max_front" else max(o[1][k] for o in outcome)
              for k in outcome[0][1]}
    host_steps = host_totals["cpu_photon_steps"] + host_totals["cpu_lepton_steps"] if prefix else 0
    steps = totals["photon_steps"] + totals["lepton_steps"] + host_steps
    # per request: its device ledger and, with the prefix, its host ledger
    # (each side closes; the host books the primary as injected less the batch)
    ledgers = [None] * len(requests)
    for k in range(pools):
        for j, r in enumerate(outcome[k][0]):
            ledgers[k + j * pools] = [r["ledger"]] + ([host_ledgers[k + j * pools]] if prefix else [])
    # Closure holds to 1e-9 E0 except in showers with Molière-failed steps: that
    # transport fallback hands over the vertex state, so the failed step's
    # continuous deposit is not booked (MoliereStep.inl:60-64).
    failed = [any(l["fallback_reasons"][21] for l in ls) for ls in ledgers]
    residual = max([abs(sum(em.ledger_residual(l) for l in ls)) for ls, f in zip(ledgers, failed) if not f]
                   or [0.0])
    residual_all = max(abs(sum(em.ledger_residual(l) for l in ls)) for ls in ledgers)
    injected = max(abs(sum(l["injected_energy"] for l in ls) - energy) for ls in ledgers)
    pool_latency = [r["finished_s"] - r["admitted_s"] for r in results]

    plain_warm_totals = plain_totals = zero
    one_at_a_time = None
    if plain_showers:
        plain_cfg = em.make_config(airtables, atmosphere=atm, slots=1, queue_capacity=1 << 20,
                                   observation_capacity=1 << 14, fallback_capacity=1 << 10, moliere=moliere,
                                   moliere_table=table)
        plain = em.EmBackend(plain_cfg, tables, atm)
        _, plain_warm_totals = plain.serve(em_showers(energy, 2, 1), 1e18, realtime=False)  # warm-up
        t1 = time.perf_counter()
        plain_results, plain_totals = plain.serve(em_showers(energy, plain_showers, 5000), 1e18,
                                                  realtime=False)
        plain_elapsed = time.perf_counter() - t1
        plain.close()
        plain_steps = plain_totals["photon_steps"] + plain_totals["lepton_steps"]
        plain_latency = [r["finished_s"] - r["admitted_s"] for r in plain_results]
        one_at_a_time = {"value": plain_steps / plain_elapsed, "showers": plain_showers,
                         "latency_ms": {"p50": float(np.percentile(plain_latency, 50) * 1e3),
                                        "p99": float(np.percentile(plain_latency, 99) * 1e3)}}

    step_kinds = ("photon_step", "lepton_transport", "lepton_vertex")
    kind = max(step_kinds, key=lambda k: timing[k][0])
    step_kernel_ms = sum(timing[k][0] for k in step_kinds)
    total_kernel_ms = sum(v[0] for v in timing.values())
    device_steps = totals["photon_steps"] + totals["lepton_steps"]
    # algorithmic bytes of all device particle-steps over the summed step-kernel time
    hbm_achieved = device_steps * EM_BYTES_PER_STEP / (step_kernel_ms * 1e-3) / 1e9
    # the "lepton_transport" timing covers the fused kernel, or with Molière its
    # move, Molière, tail and finish kernels
    tag = "true" if moliere_table else "false"
    transport_kernels = (["lepton_move_kernel", f"lepton_moliere_kernel<{tag}>", f"lepton_moliere_tail_kernel<{tag}>",
                          "lepton_finish_kernel", "lepton_finish_open_kernel"] if moliere
                         else ["lepton_transport_kernel"])
    kernel_names = {"photon_step": ["photon_step_kernel"], "lepton_transport": transport_kernels,
                    "lepton_vertex": ["lepton_vertex_kernel"]}[kind]
    kernel_name = "+".join(kernel_names)
    # The step kernels are bound by VALU issue (FP64 arithmetic with divergent
    # loops), not HBM: the roofline is the VALU issue rate of the dominant
    # group, its counted instructions per unit of work over the live HIP-event
    # time of its launches.
    valu = em_valu_per_unit()
    units = {"photon_step": totals["photon_steps"], "lepton_transport": totals["lepton_steps"],
             "lepton_vertex": totals["lepton_interactions"]}
    group_valu = {k: (sum(valu[n] for n in names) if all(n in valu for n in names) else None)
                  for k, names in (("photon_step", ["photon_step_kernel"]), ("lepton_transport", transport_kernels),
                                   ("lepton_vertex", ["lepton_vertex_kernel"]))}
    achieved = (group_valu[kind] * units[kind] / (timing[kind][0] * 1e-3) / 1e9
                if group_valu[kind] is not None and timing[kind][0] > 0 else None)
    chip_valu = (sum(group_valu[k] * units[k] for k in step_kinds) / elapsed / 1e9
                 if all(group_valu[k] is not None for k in step_kinds) else None)
    out = {
        "metric": "EM particle-steps/s",
        "value": steps / elapsed,
        "unit": "particle-steps/s",
        "showers": showers,
        "slots": slots,
        "pools": pools,
        "particle_steps": steps,
        "ms_per_shower": elapsed * 1e3 / showers,
        "latency_ms": {"p50": float(np.percentile(pool_latency, 50) * 1e3),
                       "p99": float(np.percentile(pool_latency, 99) * 1e3)},
        "photon_steps": totals["photon_steps"],
        "lepton_steps": totals["lepton_steps"],
        # every device step of the section (warm-up and one-at-a-time runs included),
        # the denominator of per-step counter totals from a profiler run
        "all_device_steps": {k: sum(t[k] for t in (warm_totals, totals, plain_warm_totals, plain_totals))
                             for k in ("photon_steps", "lepton_steps", "lepton_interactions")},
        "fronts": totals["photon_fronts"] + totals["lepton_fronts"],
        "max_front": totals["max_front"],
        "one_at_a_time": one_at_a_time,
        "dtype": "f64",
        "config": {"workload": f"EM plain: {energy:g} GeV vertical photon at 112.75 km, USStdBK 5-layer "
                               "spherical air, observation level 0 m, 1 MeV cut, no thinning, no radio, "
                               f"Moliere scattering {('on (interpolation table)' if moliere_table else 'on (analytical functions)') if moliere else 'off'}, "
                               "closed-form v7 air tables, "
                               + ("CPU prefix (host scalar expansion, at most 8 rounds) to the 16384 handoff, "
                                  "then the device; " if prefix else "whole cascade on the device; ")
                               + f"{showers} showers through {slots} pool slots"
                               + (f" in {pools} pools on their own streams" if pools > 1 else "")},
        "roofline": {
            "bound": "valu",
            "pipe": "fp64 VALU issue: one wave64 instruction per CU per clock, 256 CUs at 2.4 GHz "
                    "(no MFMA instruction is issued)",
            "achieved": achieved,
            "peak": VALU_ISSUE_PEAK_G,
            "unit": "G VALU wave-instructions/s",
            "frac": achieved / VALU_ISSUE_PEAK_G if achieved is not None else None,
            "traffic": em_measured_traffic(kernel_names),
            "kernel": kernel_name,
            "valu_per_unit": group_valu[kind],
            "unit_of_work": {"photon_step": "photon step", "lepton_transport": "e+- step",
                             "lepton_vertex": "e+- interaction"}[kind],
            "chip_valu_frac": chip_valu / VALU_ISSUE_PEAK_G if chip_valu is not None else None,
            "hbm": {"achieved_gbs": hbm_achieved, "frac": hbm_achieved / HBM_PEAK_GBS,
                    "bytes_per_step": EM_BYTES_PER_STEP},
            "kernel_ms": {k: v[0] for k, v in timing.items()},
            "launches": {k: v[1] for k, v in timing.items()},
            "device_busy_fraction": total_kernel_ms * 1e-3 / elapsed,
        },
        "ledger_residual_gev_max": residual,
        # every shower, those with a Molière-failed step included: their residual is
        # the failed step's continuous deposit, which the reference does not book
        "ledger_residual_gev_max_all_showers": residual_all,
        "injected_energy_error_gev_max": injected,
        "showers_with_moliere_fallbacks": int(sum(failed)),
    }
    if prefix:
        out["handoff"] = {
            "minimum_batch": handoff,
            "host_prefix_s": prefix_s,
            "host_threads": host_threads,
            "host_steps": host_steps,
            "device_steps": totals["photon_steps"] + totals["lepton_steps"],
            "mean_batch": float(np.mean([r["particles"].size for r in requests])),
            "expansion_rounds": host_totals["small_batch_expansions"],
            "budget_flushes": host_totals["scalar_expansion_budget_flushes"],
        }
    if handoff_showers:
        out["handoff_one_at_a_time"] = em_handoff_showers(tables, atm, energy, handoff_showers, moliere, table)
    if not no_cpu_baseline:
        base = em_cpu_baseline(tables, atm, energy, cpu_seconds, moliere=moliere, moliere_table=moliere_table)
        out["cpu_baseline"] = base
        out["speedup_vs_cpu_baseline"] = out["value"] / base["value"]
    return out


def em_radio_cpu_baseline(tables, atmosphere, cfg, observers, propagation, target_seconds, primary=None,
                          label="100 GeV vertical electron", coreas=True):
    """Oracle EM showers with their step tracks (one shower of `primary`, by
    default a 100 GeV electron at 112.75 km, per host thread) followed by the
    FP64 radio oracle projecting all their tracks on the same antennas (OpenMP
    over tracks), on all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from concurrent.futures import ThreadPoolExecutor

    import oracle_lib
    from c8mi import em

    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0")) or 16, os.cpu_count() or 1))
    prim = primary if primary is not None else em.particle(11, 100.0, (0.0, 0.0, 112750.0), (0.0, 0.0, -1.0),
                                                          history_id=1)
    count = threads
    while True:
        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as pool:  # ctypes calls release the GIL
            runs = list(pool.map(lambda k: oracle_lib.em_run_shower(cfg, tables, atmosphere, prim,
                                                                    shower_id=500 + k, tracks=True),
                                 range(count)))
        t_em = time.perf_counter() - t0
        tracks = np.concatenate([r["tracks"] for r in runs], axis=1)
        t1 = time.perf_counter()
        oracle_lib.radio_project(propagation, observers, tracks, coreas=coreas, floating=False, threads=threads)
        t_radio = time.perf_counter() - t1
        if t_em + t_radio >= 0.5 * target_seconds or count >= 8 * threads:
            break
        count *= 2
    steps = sum(r["stats"]["photon_steps"] + r["stats"]["lepton_steps"] for r in runs)
    wall = t_em + t_radio
    return {"value": steps / wall, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "updates_per_s": tracks.shape[1] * observers.size / wall,
            "sample": f"{count} x {label} showers with radio tracks, one per thread "
                      f"(oracle, {t_em:.1f} s), then their {tracks.shape[1]} tracks x {observers.size} "
                      f"antennas by the FP64 radio oracle, OpenMP over tracks ({t_radio:.1f} s); "
                      f"{threads} threads"}


def em_radio_section(showers, energy, cpu_seconds, no_cpu_baseline, moliere=True, moliere_table=True, prefix=True):
    """EM transport with radio: a vertical electron of `energy` GeV at 112.75 km
    in USStdBK air (1 MeV cut, no thinning, Molière on), every e+- step
    projected by CoREAS+ZHS onto the 160-antenna star (8 arms x 20 radii,
    25-500 m) at 1024 bins @ 0.5 ns, n(h) = 1 + 3.25e-4 exp(-h / 8 km). With
    `prefix` each shower runs through the host handoff (c8mi_em_hybrid_run:
    host scalar expansion down to the 16384 handoff batch, then the router's
    device calls; the host e+- tracks reach the radio as host tracks), else the
    whole cascade on the device. Showers run one after another, each with its
    own radio reset and fixed-point download; the projection of a front
    overlaps the next fronts on its own stream."""
    import c8mi
    from c8mi import airtables, em, workloads

    tables = airtables.build_air_tables()
    atm = em.Atmosphere()
    cfg = em.make_config(airtables, atmosphere=atm, queue_capacity=1 << 23, observation_capacity=1 << 16,
                         fallback_capacity=1 << 10, moliere=moliere,
                         moliere_table=True if moliere_table else None)
    start, down = (0.0, 0.0, 112750.0), (0.0, 0.0, -1.0)
    observers = workloads.star_observers(arms=8, radii=20, sample_rate_hz=2.0e9, number_of_bins=1024)
    workloads.plane_front_windows(observers, start, down)
    refractivity = workloads.exponential_refractivity(-1000.0, 113000.0, 1.0)
    prop = c8mi.Propagation(refractivity=refractivity, minimum_height_m=-1000.0, maximum_height_m=113000.0,
                            step_m=1.0)
    dev = em.EmBackend(cfg, tables, atm)
    stepper = em.ScalarStepper(cfg, tables, atm) if prefix else None
    cascade = em.HybridCascade(dev, stepper) if prefix else None

    def run(radio, shower_id, e):
        radio.reset()
        prim = em.particle(11, e, start, down, history_id=1)
        t0 = time.perf_counter()
        if prefix:
            r = cascade.run(prim, shower_id, 2)
            totals = {"photon_steps": r["cpu_photon_steps"] + r["device_photon_steps"],
                      "lepton_steps": r["cpu_lepton_steps"] + r["device_lepton_steps"], "handoff": r}
        else:
            dev.begin_shower(shower_id, 2)
            dev.enqueue(prim)
            totals = dev.run()
        fixed = radio.download_fixed()  # waits for the last projections
        return time.perf_counter() - t0, totals, fixed

    def attach(radio):
        dev.attach_radio(radio)
        if prefix:
            cascade.attach_radio(radio)

    radio = c8mi.RadioAccumulator(observers, prop)
    attach(radio)
    run(radio, 999, 100.0)  # warm-up
    dev.enable_kernel_timing(True)
    radio.enable_kernel_timing(True)
    results = [run(radio, 1 + k, energy) for k in range(showers)]
    em_timing = dev.kernel_time()
    radio_timing = radio.kernel_time()
    # field error of this (fast FP64) projection: the same first shower again
    # with the exact projection, whose samples equal the FP64 oracle's
    exact = c8mi.RadioAccumulator(observers, prop, exact=True)
    attach(exact)
    _, _, exact_fixed = run(exact, 1, energy)
    attach(None)
    if prefix:
        cascade.close()
        stepper.close()
    dev.close()
    radio.close()
    exact.close()
    errors = []
    for got, want in zip(results[0][2][:2], exact_fixed[:2]):
        g, w = got.astype(np.float64), want.astype(np.float64)
        errors.append(float(np.linalg.norm(g - w) / max(np.linalg.norm(w), 1.0)))

    elapsed = sum(r[0] for r in results)
    steps = sum(r[1]["photon_steps"] + r[1]["lepton_steps"] for r in results)
    tracks = sum(r[1]["lepton_steps"] for r in results)
    updates = tracks * observers.size
    em_ms = sum(v[0] for v in em_timing.values())
    radio_ms = sum(v[0] for v in radio_timing.values())
    latency = [r[0] for r in results]
    out = {
        "metric": "EM particle-steps/s with radio; radio track·antenna updates/s",
        "value": steps / elapsed,
        "unit": "particle-steps/s",
        "updates_per_s": updates / elapsed,
        "showers": showers,
        "particle_steps": steps,
        "radio_tracks": tracks,
        "valid_tracks": int(results[0][2][2]["valid_tracks"]),
        "latency_ms": {"p50": float(np.percentile(latency, 50) * 1e3),
                       "p99": float(np.percentile(latency, 99) * 1e3)},
        "dtype": "f64",
        "precision": "FP64 transport and ledger; FP64 projection with hardware reciprocal/rsqrt + Newton",
        "field_error_vs_fp64": {"coreas": errors[0], "zhs": errors[1]},
        "config": {"workload": f"EM + radio: {energy:g} GeV vertical electron at 112.75 km, USStdBK, "
                               f"1 MeV cut, Moliere {('on (table)' if moliere_table else 'on (analytical)') if moliere else 'off'}, CoREAS+ZHS at "
                               f"{observers.size} antennas, 1024 bins @ 0.5 ns, "
                               + ("serial CPU prefix down to the 16384 handoff, then the device (host handoff "
                                  "driver, one shower at a time)" if prefix else "whole cascade on the device"),
                   "antennas": int(observers.size), "bins": 1024, "sample_rate_hz": 2.0e9},
        "kernel_ms": {**{k: v[0] for k, v in em_timing.items()},
                      **{"radio_" + k: v[0] for k, v in radio_timing.items() if v[1]}},
        # device time of EM and radio kernels over wall time: above 1 when the
        # projection overlaps the fronts
        "device_busy_fraction": (em_ms + radio_ms) * 1e-3 / elapsed,
    }
    if prefix:
        hs = [r[1]["handoff"] for r in results]
        out["handoff"] = {k: float(np.mean([h[k] for h in hs])) for k in (
            "cpu_photon_steps", "cpu_lepton_steps", "cpu_tracks", "small_batch_expansions", "wavefronts",
            "photon_calls", "lepton_calls", "maximum_input_batch", "below_minimum_batch_checkpoints",
            "host_seconds", "device_seconds")}
    if not no_cpu_baseline:
        base = em_radio_cpu_baseline(tables, atm, cfg, observers, prop, cpu_seconds)
        out["cpu_baseline"] = base
        out["speedup_vs_cpu_baseline"] = out["value"] / base["value"]
        out["updates_speedup_vs_cpu_baseline"] = out["updates_per_s"] / base["updates_per_s"]
    return out


def ice_field_error(fast_zhs, exact_zhs, observers, prop, tables, ice, config, start, down):
    """Relative L2 error of the ZHS vector potential per antenna (the maximum
    over antennas): the timed fast projection of the whole shower against its
    exact FP64 projection, and on a 100 GeV sample shower from the oracle
    (8 of the antennas) both modes against the unquantised FP64 oracle."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import c8mi
    import oracle_lib
    from c8mi import em

    def worst(got, want, obs):
        offsets = np.concatenate([[0], np.cumsum(obs["number_of_bins"].astype(np.int64))])
        out = 0.0
        for k in range(obs.size):
            w = want[:, offsets[k]:offsets[k + 1]]
            norm = np.linalg.norm(w)
            if norm > 0:
                out = max(out, float(np.linalg.norm(got[:, offsets[k]:offsets[k + 1]] - w) / norm))
        return out

    fast = oracle_lib.radio_decode(fast_zhs, observers, potential=True)
    exact = oracle_lib.radio_decode(exact_zhs, observers, potential=True)
    result = {"full_shower_fast_vs_exact": worst(fast, exact, observers)}
    sample_e = 100.0
    tracks = oracle_lib.em_run_shower(config(sample_e), tables, ice, em.particle(11, sample_e, start, down, history_id=1),
                                      tracks=True)["tracks"]
    subset = observers[::8].copy()
    ref = oracle_lib.radio_project(prop, subset, tracks, coreas=False, fixed=False, threads=16)
    for name, exact_mode in (("sample_fast_vs_fp64_oracle", False), ("sample_exact_vs_fp64_oracle", True)):
        acc = c8mi.RadioAccumulator(subset, prop, coreas=False, zhs=True, exact=exact_mode)
        dev_tracks = c8mi.DeviceArray.from_host(np.ascontiguousarray(tracks))
        acc.accumulate_tracks(dev_tracks, tracks.shape[1], tracks.shape[1])
        got = oracle_lib.radio_decode(acc.download_fixed()[1], subset, potential=True)
        acc.close()
        dev_tracks.free()
        result[name] = worst(got, ref["zhs_float"], subset)
    result["sample"] = f"{sample_e:g} GeV in-ice electron shower ({tracks.shape[1]} oracle tracks), " \
                       f"{subset.size} of the {observers.size} ring antennas"
    result["metric"] = "max over antennas of the relative L2 error of the ZHS vector potential"
    return result


def ice_section(energy, cpu_seconds, no_cpu_baseline):
    """In-ice cascade (applications/c8_ice_cascade.cpp scene): a vertical
    electron of `energy` GeV 100 m below the surface of homogeneous ice (0.92
    g/cm^3, H2O), 100 keV cut, thinning below 1e-6 E0 with maximum weight 1e4,
    Molière on (ice table), every e+- step projected by ZHS (homogeneous n =
    1.78) onto 64 antennas on the Cherenkov cone 1 km from the shower maximum,
    4096 bins @ 0.1 ns; the whole cascade on the device."""
    import c8mi
    from c8mi import airtables, em, workloads

    tables = airtables.build_ice_tables(cut_mev=0.1)
    ice = em.IceVolume()
    start, down = (0.0, 0.0, -100.0), (0.0, 0.0, -1.0)
    n_ice = 1.78

    def config(e0, **kw):
        return em.make_config(airtables.IceMedium, cut_mev=0.1, start=start, direction=down, atmosphere=ice,
                              thinning=(1.0e-6 * e0, 1.0e4), moliere=True, moliere_table=True, bins=256, **kw)

    cfg = config(energy, queue_capacity=1 << 24, observation_capacity=1 << 10, fallback_capacity=1 << 16)
    # shower maximum ~ X0 ln(E / E_c) below the start (X0 = 36.08 g/cm^2, E_c ~ 78 MeV)
    depth_max_m = 36.08 / 0.92 / 100.0 * math.log(energy * 1.0e3 / 78.0)
    x_max = np.array(start) + depth_max_m * np.array(down)
    observers = workloads.cherenkov_ring_observers(x_max, down, 1000.0, n_ice, count=64, sample_rate_hz=1.0e10,
                                                   number_of_bins=4096,
                                                   emission_time_s=depth_max_m / workloads.SPEED_OF_LIGHT)
    prop = c8mi.Propagation(homogeneous_index=n_ice)
    dev = em.EmBackend(cfg, tables, ice)
    radio = c8mi.RadioAccumulator(observers, prop, coreas=False, zhs=True)
    dev.attach_radio(radio)

    def run(shower_id, e):
        radio.reset()
        dev.begin_shower(shower_id, 2)
        dev.enqueue(em.particle(11, e, start, down, history_id=1))
        t0 = time.perf_counter()
        totals = dev.run()
        fixed = radio.download_fixed()
        return time.perf_counter() - t0, totals, fixed

    run(999, 100.0)  # warm-up
    dev.enable_kernel_timing(True)
    radio.enable_kernel_timing(True)
    elapsed, totals, fixed = run(1, energy)
    em_timing = dev.kernel_time()
    radio_timing = radio.kernel_time()
    _, ledger = dev.download_profile()
    # Field error of the timed (fast) projection against FP64: the same shower
    # (identical EM history) projected again in the exact mode, which equals the
    # FP64 oracle's int64 sums bit for bit; and, on a 100 GeV sample shower the
    # oracle can run here, both modes against the unquantised FP64 oracle.
    exact_radio = c8mi.RadioAccumulator(observers, prop, coreas=False, zhs=True, exact=True)
    dev.attach_radio(exact_radio)
    exact_radio.reset()
    dev.begin_shower(1, 2)
    dev.enqueue(em.particle(11, energy, start, down, history_id=1))
    dev.run()
    exact_fixed = exact_radio.download_fixed()
    dev.attach_radio(None)
    dev.close()
    radio.close()
    exact_radio.close()
    field_error = ice_field_error(fixed[1], exact_fixed[1], observers, prop, tables, ice, config, start, down)
    steps = totals["photon_steps"] + totals["lepton_steps"]
    updates = totals["lepton_steps"] * observers.size
    zhs = fixed[1].astype(np.float64)
    out = {
        "metric": "EM particle-steps/s and ZHS track·antenna updates/s in ice",
        "value": steps / elapsed,
        "unit": "particle-steps/s",
        "updates_per_s": updates / elapsed,
        "particle_steps": steps,
        "radio_tracks": totals["lepton_steps"],
        "valid_tracks": int(fixed[2]["valid_tracks"]),
        "shower_s": elapsed,
        "thinning_vertices": int(ledger["thinning_hillas_vertices"] + ledger["thinning_statistical_vertices"]),
        "ledger_residual_gev": float(em.ledger_residual(ledger)),
        "zhs_peak_sum": float(np.abs(zhs).max()),
        "field_error_vs_fp64": field_error,
        "dtype": "f64",
        "config": {"workload": f"in-ice cascade: {energy:g} GeV vertical electron in homogeneous ice (0.92 g/cm^3), "
                               "100 keV cut, thinning 1e-6 E0 / max weight 1e4, Moliere on (table), ZHS at 64 "
                               "antennas on the Cherenkov cone 1 km from Xmax (n = 1.78), 4096 bins @ 0.1 ns",
                   "antennas": int(observers.size), "bins": 4096, "sample_rate_hz": 1.0e10},
        "kernel_ms": {**{k: v[0] for k, v in em_timing.items()},
                      **{"radio_" + k: v[0] for k, v in radio_timing.items() if v[1]}},
    }
    if not no_cpu_baseline:
        sample_e = 100.0
        base = em_radio_cpu_baseline(tables, ice, config(sample_e), observers, prop, cpu_seconds,
                                     primary=em.particle(11, sample_e, start, down, history_id=1),
                                     label="100 GeV electron in ice (thinning 1e-6 E0)", coreas=False)
        out["cpu_baseline"] = base
        out["speedup_vs_cpu_baseline"] = out["value"] / base["value"]
        out["updates_speedup_vs_cpu_baseline"] = out["updates_per_s"] / base["updates_per_s"]
    return out


def serving_cpu_baseline(tables, atmosphere, propagation, target_seconds, moliere_table=True):
    """The serving stream's work on the host: each thread serves requests
    alone, one after another (oracle EM shower with its step tracks, then the
    FP64 radio oracle on the request's antennas), on all host cores. Sample:
    requests of the same stream restricted to 100-300 GeV (the per-step and
    per-update costs do not depend on the energy; the TeV tail would take
    minutes per shower here)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from concurrent.futures import ThreadPoolExecutor

    import oracle_lib
    from c8mi import airtables, em, workloads

    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0")) or 16, os.cpu_count() or 1))
    stream = workloads.serving_stream(4 * threads, 10.0, seed=777, e_min_gev=100.0, e_max_gev=300.0)

    def serve_one(k):
        r = stream[k]
        cfg = em.make_config(airtables, atmosphere=atmosphere, start=tuple(r["start_m"]),
                             direction=tuple(r["direction"]), moliere=True,
                             moliere_table=True if moliere_table else None)
        prim = em.particle(r["pid"], r["energy_gev"], tuple(r["start_m"]), tuple(r["direction"]), history_id=1)
        out = oracle_lib.em_run_shower(cfg, tables, atmosphere, prim, shower_id=9000 + k, tracks=True)
        oracle_lib.radio_project(propagation, r["observers"], out["tracks"], floating=False, threads=1)
        return out["stats"]["photon_steps"] + out["stats"]["lepton_steps"], out["tracks"].shape[1] * r["observers"].size

    count = threads
    while True:
        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as pool:
            done = list(pool.map(serve_one, range(count)))
        wall = time.perf_counter() - t0
        if wall >= 0.5 * target_seconds or count >= len(stream):
            break
        count = min(len(stream), 2 * count)
    steps = sum(d[0] for d in done)
    updates = sum(d[1] for d in done)
    return {"value": steps / wall, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "updates_per_s": updates / wall,
            "sample": f"{count} requests of the stream restricted to 100-300 GeV, each served alone by one "
                      f"thread (oracle EM shower, then the FP64 radio oracle on its 40 antennas); {threads} "
                      f"threads, {wall:.1f} s wall, {steps} steps"}


def serving_section(count, rate, slots, plain_count, cpu_seconds, no_cpu_baseline, moliere_table=True):
    """The serving stream: photon and electron primaries 50/50, energy
    log-uniform in 100 GeV - 100 TeV, zenith U(0, 60 deg), azimuth U(0, 360
    deg), Poisson arrivals at `rate` per second (above what the device serves,
    so the pool saturates), 1 MeV cut, Molière on, CoREAS+ZHS radio on 40
    antennas per shower (8 arms x 5 radii, 50-250 m, 1024 bins @ 0.5 ns) kept
    per shower on the device. Requests are admitted as they arrive, into free
    slots and within the footprint budget (continuous batching); per-shower
    latency from arrival to completion. For the latency comparison the first
    `plain_count` requests of the same stream, same arrivals, are served again
    alone: through the pool and through a single slot (one shower at a
    time)."""
    import c8mi
    from c8mi import airtables, em, workloads

    tables = airtables.build_air_tables()
    atm = em.Atmosphere()
    refractivity = workloads.exponential_refractivity(-1000.0, 113000.0, 1.0)
    prop = c8mi.Propagation(refractivity=refractivity, minimum_height_m=-1000.0, maximum_height_m=113000.0,
                            step_m=1.0)
    stream = workloads.serving_stream(count, rate)
    axes = [em.shower_axis(atm, tuple(r["start_m"]), tuple(r["direction"])) for r in stream]

    def requests(n, first_id):
        return [dict(primary=em.particle(r["pid"], r["energy_gev"], tuple(r["start_m"]), tuple(r["direction"]),
                                         history_id=1),
                     shower_id=first_id + k, arrival_s=r["arrival_s"], footprint=em.footprint_estimate(r["energy_gev"]),
                     axis=axes[k], radio_observers=r["observers"]) for k, r in enumerate(stream[:n])]

    table = True if moliere_table else None
    capacity = 1 << 25

    def run(n, pool_slots, timing):
        cfg = em.make_config(airtables, atmosphere=atm, slots=pool_slots, queue_capacity=capacity,
                             observation_capacity=1 << 12, fallback_capacity=1 << 8, moliere=True,
                             moliere_table=table)
        dev = em.EmBackend(cfg, tables, atm)
        dev.enable_slot_radio(40, 1024, prop)
        # warm-up: one small shower through the same pool
        warm = dict(primary=em.particle(22, 100.0, tuple(stream[0]["start_m"]), tuple(stream[0]["direction"]),
                                        history_id=1), shower_id=1, arrival_s=0.0, footprint=1.0, axis=axes[0],
                    radio_observers=stream[0]["observers"])
        dev.serve([warm], 0.4 * capacity, realtime=False)
        dev.enable_kernel_timing(timing)
        t0 = time.perf_counter()
        results, totals = dev.serve(requests(n, 100), 0.4 * capacity, realtime=True)
        wall = time.perf_counter() - t0
        timing_out = (dev.kernel_time(), dev.slot_radio_kernel_time()) if timing else None
        dev.close()
        return results, totals, wall, timing_out

    results, totals, wall, timing = run(count, slots, True)
    steps = totals["photon_steps"] + totals["lepton_steps"]
    updates = sum(r["lepton_steps"] for r in results) * 40
    latency = [r["finished_s"] - s["arrival_s"] for r, s in zip(results, stream)]
    service = [r["finished_s"] - r["admitted_s"] for r in results]
    em_timing, (radio_ms, radio_launches) = timing
    out = {
        "metric": "serving: particle-steps/s, track·antenna updates/s, per-shower latency",
        "value": steps / wall,
        "unit": "particle-steps/s",
        "updates_per_s": updates / wall,
        "showers": count,
        "slots": slots,
        "arrival_rate_hz": rate,
        "offered_s": stream[-1]["arrival_s"],
        "wall_s": wall,
        "particle_steps": steps,
        "latency_ms": {"p50": float(np.percentile(latency, 50) * 1e3), "p99": float(np.percentile(latency, 99) * 1e3)},
        "service_ms": {"p50": float(np.percentile(service, 50) * 1e3), "p99": float(np.percentile(service, 99) * 1e3)},
        "valid_radio_tracks": int(sum(r["radio_counters"]["valid_tracks"] for r in results)),
        "kernel_ms": {**{k: v[0] for k, v in em_timing.items()}, "radio_per_shower": radio_ms},
        "radio_launches": radio_launches,
        "dtype": "f64",
        "config": {"workload": f"serving stream: {count} requests, photon/electron 50/50, E log-uniform "
                               f"100 GeV-100 TeV, zenith U(0,60), azimuth U(0,360), Poisson arrivals at "
                               f"{rate:g}/s, 1 MeV cut, Moliere on, CoREAS+ZHS on 40 antennas per shower "
                               f"(1024 bins @ 0.5 ns), {slots} pool slots"},
    }
    if plain_count:
        def summary(res, tot, w):
            lat = [r["finished_s"] - s["arrival_s"] for r, s in zip(res, stream)]
            srv = [r["finished_s"] - r["admitted_s"] for r in res]
            return {"value": (tot["photon_steps"] + tot["lepton_steps"]) / w, "wall_s": w,
                    "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                                   "p99": float(np.percentile(lat, 99) * 1e3)},
                    "service_ms": {"p50": float(np.percentile(srv, 50) * 1e3),
                                   "p99": float(np.percentile(srv, 99) * 1e3)}}

        out["latency_comparison"] = {
            "showers": plain_count,
            "pool": summary(*run(plain_count, slots, False)[:3]),
            "one_at_a_time": summary(*run(plain_count, 1, False)[:3]),
        }
    if not no_cpu_baseline:
        base = serving_cpu_baseline(tables, atm, prop, cpu_seconds, moliere_table=moliere_table)
        out["cpu_baseline"] = base
        out["speedup_vs_cpu_baseline"] = out["value"] / base["value"]
        out["updates_speedup_vs_cpu_baseline"] = out["updates_per_s"] / base["updates_per_s"]
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "radio_traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    return data["traffic_bytes"] if data.get("kernel") == kernel else None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the radio and EM paths shard by observers/sub-showers without exchange, "
              "which this single-GPU run does not exercise.", file=sys.stderr)
        return 2
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return 0

    import c8mi
    from c8mi import workloads

    if c8mi.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    if args.serve_only:
        out = serving_section(max(args.serve_showers, 1), args.serve_rate, args.serve_slots,
                              args.serve_plain_showers, args.cpu_seconds, args.no_cpu_baseline,
                              moliere_table=not args.em_analytic_moliere)
        out.update({"n_gpus": 1, "steps": out["showers"], "higher_is_better": True, "scaling": "weak",
                    "vs_baseline": None, "data": "synthetic", "ms_per_step": out["latency_ms"]["p50"]})
        print(json.dumps(out), flush=True)
        return 0
    if args.ice_only:
        out = ice_section(args.ice_energy, args.cpu_seconds, args.no_cpu_baseline)
        out.update({"n_gpus": 1, "steps": 1, "higher_is_better": True, "scaling": "weak",
                    "vs_baseline": None, "data": "synthetic", "ms_per_step": out["shower_s"] * 1e3})
        print(json.dumps(out), flush=True)
        return 0
    if args.em_radio_only:
        out = em_radio_section(max(args.em_radio_showers, 1), args.em_radio_energy, args.cpu_seconds,
                               args.no_cpu_baseline, moliere=not args.em_no_moliere,
                               moliere_table=not args.em_analytic_moliere, prefix=not args.em_no_prefix)
        out.update({"n_gpus": 1, "steps": out["showers"], "higher_is_better": True, "scaling": "weak",
                    "vs_baseline": None, "data": "synthetic",
                    "ms_per_step": out["latency_ms"]["p50"]})
        print(json.dumps(out), flush=True)
        return 0
    if args.em_only:
        em_result = em_section(max(args.em_showers, 1), args.em_plain_showers, args.em_slots, args.cpu_seconds,
                               args.no_cpu_baseline, warmup_showers=args.em_warmup_showers,
                               moliere=not args.em_no_moliere,
                               moliere_table=not args.em_analytic_moliere, pools=args.em_pools,
                               fast_iterations=args.em_moliere_fast_iterations, prefix=not args.em_no_prefix,
                               handoff_showers=args.em_handoff_showers)
        em_result.update({"n_gpus": 1, "steps": em_result["showers"], "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "data": "synthetic",
                          "ms_per_step": em_result["ms_per_shower"]})
        print(json.dumps(em_result), flush=True)
        return 0
    tracks, observers, refractivity, (lo, hi) = workloads.config2_setup(count=args.tracks)
    propagation = c8mi.Propagation(refractivity=refractivity, minimum_height_m=lo,
                                   maximum_height_m=hi, step_m=1.0)
    coreas = args.mode in ("coreas_zhs", "coreas")
    zhs = args.mode in ("coreas_zhs", "zhs")
    limit = 1.0
    acc = c8mi.RadioAccumulator(observers, propagation, coreas=coreas, zhs=zhs,
                                fixed_point_limit=limit, exact=args.exact)
    if args.chunk_tracks or args.tile_bins or args.target_workgroups or args.block_threads or args.split:
        acc.set_tuning(args.chunk_tracks, args.tile_bins, args.target_workgroups,
                       args.block_threads, args.split)
    n = tracks.shape[1]
    device_tracks = c8mi.DeviceArray.from_host(tracks)

    def step():
        acc.reset()
        acc.accumulate_tracks(device_tracks, n, n)

    for _ in range(args.warmup):
        step()
    c8mi.device_synchronize()
    acc.enable_kernel_timing(True)
    c8mi.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    c8mi.device_synchronize()
    elapsed = time.perf_counter() - t0
    timing = acc.kernel_time()
    _, _, counters = acc.download_fixed()
    acc.close()

    ms_per_step = elapsed * 1e3 / args.steps
    updates = n * observers.size
    value = updates / (elapsed / args.steps)
    # Dominant kernel: the launch kind with the largest total time.
    kind = max(timing, key=lambda k: timing[k][0])
    kernel_total_ms, launches = timing[kind]
    kernel_avg_ms = kernel_total_ms / max(launches, 1)
    achieved = FLOPS_PER_UPDATE[kind] * updates / (kernel_avg_ms * 1e-3) / 1e12
    labels = {"coreas_zhs": (1, 1), "coreas": (1, 0), "zhs": (0, 1)}
    kernel_name = "radio_project_kernel<coreas=%d,zhs=%d,fixed=1,exact=%d>" % (
        labels[kind] + (int(args.exact),))
    result = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "precision": "exact FP64 order" if args.exact else "FP64, hardware reciprocal/rsqrt + Newton",
        "data": "synthetic",
        "config": {
            "workload": f"radio standalone: {n} tracks x {observers.size} antennas, "
                        f"{int(observers['number_of_bins'][0])} bins @ 0.2 ns, "
                        f"{args.mode.replace('_', '+').upper()}, int64 fixed point",
            "tracks": n,
            "antennas": int(observers.size),
            "bins": int(observers["number_of_bins"][0]),
            "sample_rate_hz": float(observers["sample_rate_hz"][0]),
            "mode": args.mode,
        },
        "roofline": {
            "bound": "valu",
            "pipe": "fp64 VALU (no MFMA instruction is issued; MI355X's FP64 vector and "
                    "matrix peaks are both 78.6 TFLOP/s)",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic(kernel_name),
            "kernel": kernel_name,
            "kernel_avg_ms": kernel_avg_ms,
            "kernel_ms_per_step": {k: v[0] / args.steps for k, v in timing.items() if v[1]},
            "flops_per_update": FLOPS_PER_UPDATE[kind],
            "algorithmic_hbm_gbs": (n * TRACK_BYTES) / (kernel_avg_ms * 1e-3) / 1e9,
            "valu_issue": radio_valu_issue(kernel_name, updates, kernel_avg_ms),
        },
        "counters": counters,
    }
    if not args.no_error_check:
        result["field_error_vs_fp64"] = field_error(tracks, observers, propagation, args.mode, limit,
                                                     args.exact)
    if not args.no_cpu_baseline:
        base = cpu_baseline(tracks, observers, propagation, args.mode, args.cpu_seconds)
        result["cpu_baseline"] = base
        result["speedup_vs_cpu_baseline"] = value / base["value"]
    if args.em_showers > 0:
        result["em"] = em_section(args.em_showers, args.em_plain_showers, args.em_slots, args.cpu_seconds,
                                  args.no_cpu_baseline, warmup_showers=args.em_warmup_showers,
                                  moliere=not args.em_no_moliere,
                                  moliere_table=not args.em_analytic_moliere, pools=args.em_pools,
                                  fast_iterations=args.em_moliere_fast_iterations, prefix=not args.em_no_prefix,
                                  handoff_showers=args.em_handoff_showers)
    if args.serve_showers > 0:
        result["serving"] = serving_section(args.serve_showers, args.serve_rate, args.serve_slots,
                                            args.serve_plain_showers, args.cpu_seconds, args.no_cpu_baseline,
                                            moliere_table=not args.em_analytic_moliere)
    if args.em_radio_showers > 0:
        result["em_radio"] = em_radio_section(args.em_radio_showers, args.em_radio_energy, args.cpu_seconds,
                                              args.no_cpu_baseline, moliere=not args.em_no_moliere,
                                              moliere_table=not args.em_analytic_moliere,
                                              prefix=not args.em_no_prefix)
    if args.ice_energy > 0:
        result["ice"] = ice_section(args.ice_energy, args.cpu_seconds, args.no_cpu_baseline)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
