/*
 * c8mi radio C ABI — MI355X replacement for CORSIKA8-kokkos's device radio
 * accumulator. Every entry point returns a c8mi_status (0 = success); on
 * failure c8mi_last_error() returns the message for the calling thread.
 *
 * Status codes mirror the exceptions of the reference interface:
 *   1 invalid_argument, 2 logic_error, 3 length_error, 4 runtime_error,
 *   5 device (HIP) error, 6 fixed-point overflow (runtime_error in the reference).
 *
 * Threading: one handle is used from one thread at a time; all of a handle's
 * work is ordered on one HIP stream (RadioAccumulator.hpp:28-34).
 */
#ifndef C8MI_RADIO_H
#define C8MI_RADIO_H

#include "c8mi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct c8mi_radio c8mi_radio;

/* Last error message of the calling thread ("" if none). */
const char* c8mi_last_error(void);

/* Replaces RadioAccumulator<Exec>::initialize(RadioConfig, TuningParameters, exec)
 * (corsika/kokkos/radio/RadioAccumulator.hpp:103-105; validation as
 * RadioAccumulator.inl:934-975). Copies the config; the observer lists and
 * propagation tables may be freed after the call. stream: a hipStream_t, or
 * NULL for a stream owned by the handle. */
int c8mi_radio_create(const c8mi_radio_config* config, void* stream, c8mi_radio** out);
int c8mi_radio_destroy(c8mi_radio* radio);

/* RadioAccumulator::reset (RadioAccumulator.hpp:109): zero waveforms/counters. */
int c8mi_radio_reset(c8mi_radio* radio);

/* RadioAccumulator::enqueueHostLeptonTracks (RadioAccumulator.hpp:113-115) and
 * IEmBackend::enqueueRadioTracks (corsika/kokkos/em/execution/IEmBackend.hpp:79-85):
 * host records are copied before return and projected in batches of 4096. */
int c8mi_radio_enqueue_host_tracks(c8mi_radio* radio,
                                   const c8mi_lepton_transport_record* records,
                                   uint64_t count);
/* RadioAccumulator::flushHostLeptonTracks (RadioAccumulator.hpp:117). */
int c8mi_radio_flush_host_tracks(c8mi_radio* radio);

/* RadioAccumulator::accumulateLeptonTracks(device view, count)
 * (RadioAccumulator.hpp:119-121): device-resident records, asynchronous. */
int c8mi_radio_accumulate_records(c8mi_radio* radio,
                                  const c8mi_lepton_transport_record* device_records,
                                  uint64_t count);

/* Native form: device-resident compact tracks, [9][stride] FP64
 * (c8mi_types.h). No counterpart in the reference: its lepton front feeds the
 * accumulator 752-byte records (LeptonCascade.inl:81-82); this build's lepton
 * front emits these 72-byte tracks instead. Asynchronous. */
int c8mi_radio_accumulate_tracks(c8mi_radio* radio, const double* device_tracks,
                                 uint64_t stride, uint64_t count);

/* Total bins of each algorithm (sum over observers; 0 when disabled) and
 * whether CoREAS and ZHS share one fused launch. */
int c8mi_radio_bins(const c8mi_radio* radio, uint64_t* coreas_total_bins,
                    uint64_t* zhs_total_bins, int* fused);

/* RadioAccumulator::download (RadioAccumulator.hpp:129): flushes host tracks,
 * waits, and decodes to V/m (CoREAS) and V s/m (ZHS vector potential; the
 * output layer differentiates, TimeDomainObserver.inl:116-119). Outputs are
 * [3][total_bins] (x block, y block, z block), either may be NULL.
 * Returns 6 (overflow) if the fixed-point range was exceeded. */
int c8mi_radio_download(c8mi_radio* radio, double* coreas, double* zhs,
                        c8mi_radio_counters* counters);

/* RadioAccumulator::downloadFixed (RadioAccumulator.hpp:125-127): the raw int64
 * fixed-point sums (scale 2^62 * (ZHS ? sample rate : 1) / field limit). */
int c8mi_radio_download_fixed(c8mi_radio* radio, int64_t* coreas, int64_t* zhs,
                              c8mi_radio_counters* counters);

/* The ZHS output step of the reference's radio writer
 * (corsika/detail/modules/radio/RadioProcess.inl:212-220; the bin-integrated
 * potential is described at TimeDomainObserver.inl:107-119): from a decoded
 * ZHS vector potential A, [3][total_bins] as c8mi_radio_download returns it
 * (observers' bins in order), the electric field
 *   E[i] = -(A[i+1] - A[i]) * sample_rate   at the bin midpoint
 *   t[i] = (t_axis[i] + t_axis[i+1]) / 2,   t_axis[i] = start_time + i / sample_rate,
 * for i < bins - 1 of every observer. `field` is [3][total_bins - observer_count]
 * (bins - 1 samples per observer, in observer order); `times` (NULL allowed)
 * holds the matching [total_bins - observer_count] midpoint times in seconds.
 * Host only: no device call. */
int c8mi_radio_zhs_electric_field(const c8mi_radio_observer* observers, uint64_t observer_count,
                                  const double* potential, double* field, double* times);

/* TuningParameters (track/observer tiles, team size; TuningParameters.hpp)
 * analogue: tracks per workgroup (multiple of the workgroup size,
 * 0 = automatic), LDS window tile in bins (0 = automatic), target workgroups
 * per launch (0 = default 8192), threads per workgroup (256, 512 or 1024;
 * 0 = default 512) and split: 1 runs fusable CoREAS+ZHS work as two launches
 * (the default), 0 as one fused launch. Results do not depend on any of these. */
int c8mi_radio_set_tuning(c8mi_radio* radio, uint64_t chunk_tracks, uint32_t tile_bins,
                          uint32_t target_workgroups, uint32_t block_threads, int split);

/* Projection-kernel timing with HIP events on the handle's stream. Times and
 * launch counts per kind: [0] fused CoREAS+ZHS, [1] CoREAS only, [2] ZHS only. */
int c8mi_radio_enable_kernel_timing(c8mi_radio* radio, int enable);
int c8mi_radio_kernel_time(c8mi_radio* radio, double ms[3], uint64_t launches[3]);

/* The straight-ray table arithmetic of make_radio_config
 * (corsika/detail/kokkos/radio/atmosphere/FlatAtmosphereConfig.inl:31-73):
 * refractivity[i] = n(minimum_height + i*step) - 1 for i < count (count >= 11).
 * Writes the running integral into `integrated` (count entries) and fills
 * `out` with heights, step and edge slopes; out's table pointers are set to
 * `refractivity` and `integrated`. */
int c8mi_radio_build_flat_table(const double* refractivity, uint64_t count,
                                double minimum_height_m, double maximum_height_m,
                                double step_m, double* integrated,
                                c8mi_radio_propagation* out);

/* One-shot projection that allocates no device memory. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It computes
 * what c8mi_radio_create + c8mi_radio_accumulate_tracks +
 * c8mi_radio_download_fixed compute for `config` (deterministic mode only),
 * but writes the int64 fixed-point waveforms into the caller's DEVICE buffers
 * `device_coreas` / `device_zhs` ([3][total bins] each, cleared by the call;
 * either may be NULL when its algorithm is disabled) and keeps its own device
 * copies of the observers, the propagation table and the counters in a fixed
 * library scratch array (one per GPU, held until the process ends). Limits:
 * at most C8MI_RADIO_NOALLOC_MAX_OBSERVERS observers per algorithm and
 * C8MI_RADIO_NOALLOC_MAX_TABLE_ENTRIES propagation-table entries; larger
 * configurations return 3 (length_error). Work is ordered on `stream` (NULL:
 * the default stream) and the call returns after it completed, filling
 * `counters` (NULL allowed); 6 on fixed-point overflow. Calls that overlap on
 * one GPU (other threads or streams) share the scratch array and race. */
#define C8MI_RADIO_NOALLOC_MAX_OBSERVERS 1024
#define C8MI_RADIO_NOALLOC_MAX_TABLE_ENTRIES 131072
int c8mi_radio_project_fixed_noalloc(const c8mi_radio_config* config, const double* device_tracks,
                                     uint64_t stride, uint64_t count, int64_t* device_coreas,
                                     int64_t* device_zhs, c8mi_radio_counters* counters,
                                     void* stream);

/* Device memory and synchronisation helpers for hosts without their own HIP
 * allocator (tests, ctypes callers). */
int c8mi_device_count(int* count);
int c8mi_device_malloc(void** pointer, uint64_t bytes);
int c8mi_device_free(void* pointer);
int c8mi_memcpy_htod(void* device_destination, const void* host_source, uint64_t bytes);
int c8mi_memcpy_dtoh(void* host_destination, const void* device_source, uint64_t bytes);
/* Device-to-device copy (also the known-size copy used to calibrate HBM
 * traffic counters). */
int c8mi_memcpy_dtod(void* device_destination, const void* device_source, uint64_t bytes);
int c8mi_device_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* C8MI_RADIO_H */
