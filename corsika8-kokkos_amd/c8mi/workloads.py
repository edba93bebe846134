"""Synthetic inputs for the radio and EM workloads (numpy, seeded).

Nothing here computes a result; it only builds inputs of the shapes the
benchmarks and tests feed through the C ABI.
"""
from __future__ import annotations

import math

import numpy as np

from . import LEPTON_TRANSPORT_RECORD_DTYPE, OBSERVER_DTYPE

SPEED_OF_LIGHT = 299792458.0

# US-standard atmosphere after Bernlohr (USStdBK): lower layer boundaries (m),
# b (g/cm^2) and c (cm) per layer; density (b/c) exp(-h/c) in the four
# exponential layers and b/c in the top linear layer
# (corsika/media/CORSIKA7Atmospheres.hpp:182-186,
#  corsika/detail/media/LayeredSphericalAtmosphereBuilder.inl:39-110).
USSTDBK_TOPS_M = (7.0e3, 11.4e3, 37.0e3, 100.0e3, 112.8e3)
USSTDBK_B = (1183.6071, 1143.0425, 1322.9748, 655.67307, 1.0)
USSTDBK_C_CM = (954248.34, 800005.34, 629568.93, 737521.77, 1.0e9)


def usstdbk_density(height_m: np.ndarray) -> np.ndarray:
    """Mass density in g/cm^3 of the five USStdBK layers."""
    h = np.asarray(height_m, dtype=np.float64)
    rho = np.zeros_like(h)
    lower = -np.inf
    for i, top in enumerate(USSTDBK_TOPS_M):
        mask = (h >= lower) & (h < top) if i < 4 else (h >= lower) & (h <= top)
        c_m = USSTDBK_C_CM[i] / 100.0
        if i < 4:
            rho[mask] = USSTDBK_B[i] / USSTDBK_C_CM[i] * np.exp(-h[mask] / c_m)
        else:
            rho[mask] = USSTDBK_B[i] / USSTDBK_C_CM[i]
        lower = top
    return rho


def usstdbk_altitude_of_vertical_depth(depth_g_cm2: np.ndarray) -> np.ndarray:
    """Altitude (m) whose vertical overburden is `depth_g_cm2` (numerical)."""
    grid = np.linspace(0.0, USSTDBK_TOPS_M[-1], 112801)
    rho = usstdbk_density(grid)
    # overburden from the top of the atmosphere, trapezoidal, in g/cm^2
    seg = 0.5 * (rho[1:] + rho[:-1]) * np.diff(grid) * 100.0
    over = np.concatenate([np.cumsum(seg[::-1])[::-1], [0.0]])
    # over is decreasing with altitude; interpolate on the reversed arrays
    return np.interp(np.asarray(depth_g_cm2), over[::-1], grid[::-1])


def exponential_refractivity(minimum_height_m: float, maximum_height_m: float, step_m: float = 1.0,
                             sea_level: float = 3.25e-4, scale_height_m: float = 8000.0) -> np.ndarray:
    """n(h) - 1 = sea_level * exp(-h / scale) sampled like make_radio_config
    (corsika/detail/kokkos/radio/atmosphere/FlatAtmosphereConfig.inl:45-57)."""
    count = int((maximum_height_m - minimum_height_m) * (1.0 / step_m) + 1)
    heights = minimum_height_m + np.arange(count) * step_m
    return sea_level * np.exp(-heights / scale_height_m)


def star_observers(arms: int = 8, radii: int = 20, radius_step_m: float = 25.0,
                   height_m: float = 0.0, sample_rate_hz: float = 5.0e9,
                   number_of_bins: int = 2048, start_times_s=None) -> np.ndarray:
    """Antennas on `arms` azimuths x `radii` rings (radius_step .. radii*radius_step)."""
    obs = np.zeros(arms * radii, dtype=OBSERVER_DTYPE)
    k = 0
    for a in range(arms):
        phi = 2.0 * math.pi * a / arms
        for r in range(1, radii + 1):
            radius = r * radius_step_m
            obs[k]["position_m"] = (radius * math.cos(phi), radius * math.sin(phi), height_m)
            k += 1
    obs["sample_rate_hz"] = sample_rate_hz
    obs["number_of_bins"] = number_of_bins
    obs["duration_s"] = (number_of_bins - 1) / sample_rate_hz
    if start_times_s is not None:
        obs["start_time_s"] = start_times_s
    return obs


def plane_front_windows(observers: np.ndarray, injection_m, direction, pretrigger_s: float = 10.0e-9,
                        start_time_s: float = 0.0) -> None:
    """Sets each observer's window to open `pretrigger_s` before a plane front,
    leaving `injection_m` at `start_time_s` along the unit `direction` at c,
    passes it (the earliest arrival of an air shower's signal, up to the
    refractive delay)."""
    pos = np.stack([observers["position_m"][:, k] for k in range(3)], axis=1)
    along = (pos - np.asarray(injection_m, dtype=np.float64)) @ np.asarray(direction, dtype=np.float64)
    observers["start_time_s"] = start_time_s + along / SPEED_OF_LIGHT - pretrigger_s


def shower_axis(zenith_deg: float, azimuth_deg: float = 0.0):
    th, ph = math.radians(zenith_deg), math.radians(azimuth_deg)
    up = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    e1 = np.array([math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph), -math.sin(th)])
    e2 = np.cross(up, e1)
    return up, e1, e2


def radio_config2_tracks(count: int = 1 << 20, seed: int = 20260, zenith_deg: float = 60.0,
                         xmax: float = 750.0, lam: float = 60.0, lateral_sigma_m: float = 40.0,
                         length_range_m=(0.5, 5.0), beta_range=(0.99, 1.0),
                         electron_fraction: float = 0.6) -> np.ndarray:
    """Config-2 track list as a [9][count] float64 block (c8mi_types.h order).

    Depth ~ Gaisser-Hillas(Xmax, lambda, X0 = 0), i.e. a Gamma(Xmax/lambda + 1,
    lambda) slant depth, mapped to altitude through the USStdBK overburden along
    a `zenith_deg` axis through the origin; Gaussian lateral offset; straight
    segments along the axis with beta ~ U(beta_range); charge -1 with
    probability `electron_fraction` (20 % net negative excess), weight 1.
    Times follow a plane front moving at c that reaches the core at t = 0.
    """
    rng = np.random.default_rng(seed)
    up, e1, e2 = shower_axis(zenith_deg)
    cos_zenith = math.cos(math.radians(zenith_deg))
    ground_slant = 1036.0 / cos_zenith
    slant = rng.gamma(shape=xmax / lam + 1.0, scale=lam, size=count)
    bad = slant >= 0.98 * ground_slant
    while bad.any():
        slant[bad] = rng.gamma(shape=xmax / lam + 1.0, scale=lam, size=int(bad.sum()))
        bad = slant >= 0.98 * ground_slant
    altitude = usstdbk_altitude_of_vertical_depth(slant * cos_zenith)
    along = altitude / cos_zenith
    lateral = rng.normal(0.0, lateral_sigma_m, size=(count, 2))
    start = along[:, None] * up[None, :] + lateral[:, :1] * e1[None, :] + lateral[:, 1:] * e2[None, :]
    length = rng.uniform(length_range_m[0], length_range_m[1], size=count)
    beta = rng.uniform(beta_range[0], beta_range[1], size=count)
    end = start - length[:, None] * up[None, :]
    t0 = -along / SPEED_OF_LIGHT
    t1 = t0 + length / (beta * SPEED_OF_LIGHT)
    electron = rng.random(count) < electron_fraction
    signed_weight = np.where(electron, -1.0, 1.0)
    out = np.empty((9, count), dtype=np.float64)
    out[0:3] = start.T
    out[3:6] = end.T
    out[6] = t0
    out[7] = t1
    out[8] = signed_weight
    return out


def earliest_arrivals(tracks: np.ndarray, observers: np.ndarray, refractive_index: float = 1.0003,
                      sample: int = 8192) -> np.ndarray:
    """Approximate earliest start-point arrival time per observer from a
    subsample of tracks (used only to place each observer's time window)."""
    n = tracks.shape[1]
    idx = np.linspace(0, n - 1, min(sample, n)).astype(np.int64)
    start = tracks[0:3, idx].T
    t0 = tracks[6, idx]
    pos = np.stack([observers["position_m"][:, k] for k in range(3)], axis=1)
    dist = np.linalg.norm(pos[:, None, :] - start[None, :, :], axis=2)
    arrival = t0[None, :] + refractive_index * dist / SPEED_OF_LIGHT
    return arrival.min(axis=1)


def config2_setup(count: int = 1 << 20, seed: int = 20260, number_of_bins: int = 2048,
                  sample_rate_hz: float = 5.0e9, pretrigger_s: float = 10.0e-9, arms: int = 8,
                  radii: int = 20):
    """Tracks, observers and propagation-table samples of the standalone radio
    workload. Returns (tracks [9][count], observers, refractivity, (min_h, max_h))."""
    tracks = radio_config2_tracks(count, seed)
    observers = star_observers(arms=arms, radii=radii, sample_rate_hz=sample_rate_hz,
                               number_of_bins=number_of_bins)
    observers["start_time_s"] = earliest_arrivals(tracks, observers) - pretrigger_s
    max_h = float(np.ceil(max(tracks[2].max(), tracks[5].max()) + 100.0))
    min_h = -1000.0
    refractivity = exponential_refractivity(min_h, max_h, 1.0)
    return tracks, observers, refractivity, (min_h, max_h)


def tracks_to_records(tracks: np.ndarray, energy_gev: float = 1.0) -> np.ndarray:
    """Compact tracks -> LeptonTransportRecord array (pid from the weight sign)."""
    n = tracks.shape[1]
    rec = np.zeros(n, dtype=LEPTON_TRANSPORT_RECORD_DTYPE)
    electron = tracks[8] < 0
    for side, lo in (("start", 0), ("end", 3)):
        rec[side]["pid"] = np.where(electron, 11, -11)
        rec[side]["weight"] = np.abs(tracks[8])
        rec[side]["energy_gev"] = energy_gev
        rec[side]["position_m"] = tracks[lo:lo + 3].T
        rec[side]["environment_node"] = -3
    rec["start"]["time_s"] = tracks[6]
    rec["end"]["time_s"] = tracks[7]
    disp = tracks[3:6] - tracks[0:3]
    norm = np.linalg.norm(disp, axis=0)
    norm[norm == 0] = 1.0
    rec["start"]["direction"] = (disp / norm).T
    rec["end"]["direction"] = (disp / norm).T
    return rec


def serving_stream(count: int, rate_hz: float, seed: int = 4242, e_min_gev: float = 100.0,
                   e_max_gev: float = 1.0e5, max_zenith_deg: float = 60.0, injection_height_m: float = 112000.0,
                   arms: int = 8, radii: int = 5, radius_step_m: float = 50.0, bins: int = 1024,
                   sample_rate_hz: float = 2.0e9, pretrigger_s: float = 10.0e-9,
                   earth_radius_m: float = 6371000.0):
    """A stream of shower requests: photon and electron primaries 50/50, energy
    log-uniform in [e_min, e_max], zenith uniform in [0, max_zenith], azimuth
    uniform, Poisson arrivals at `rate_hz`; each shower's axis hits the core
    (0, 0, 0) from `injection_height_m` above the spherical Earth, and its
    radio antennas (arms x radii star around the core on the ground) open
    their windows `pretrigger_s` before the primary's plane front.

    Returns a list of dicts: pid, energy_gev, zenith_deg, azimuth_deg,
    arrival_s, start_m, direction, observers (OBSERVER_DTYPE)."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, size=count))
    arrivals -= arrivals[0]
    out = []
    for k in range(count):
        pid = 22 if rng.random() < 0.5 else 11
        energy = float(np.exp(rng.uniform(np.log(e_min_gev), np.log(e_max_gev))))
        zenith = float(rng.uniform(0.0, max_zenith_deg))
        azimuth = float(rng.uniform(0.0, 360.0))
        th, ph = math.radians(zenith), math.radians(azimuth)
        d = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), -math.cos(th)])
        # slant distance from the core to the injection height on a sphere
        r, h, c = earth_radius_m, injection_height_m, math.cos(th)
        slant = -r * c + math.sqrt(r * r * c * c + h * h + 2.0 * r * h)
        start = -slant * d
        obs = star_observers(arms=arms, radii=radii, radius_step_m=radius_step_m, sample_rate_hz=sample_rate_hz,
                             number_of_bins=bins)
        plane_front_windows(obs, start, d, pretrigger_s)
        out.append(dict(pid=pid, energy_gev=energy, zenith_deg=zenith, azimuth_deg=azimuth,
                        arrival_s=float(arrivals[k]), start_m=start, direction=d, observers=obs))
    return out


def cherenkov_ring_observers(center_m, axis, distance_m: float, refractive_index: float, count: int = 64,
                             sample_rate_hz: float = 1.0e10, number_of_bins: int = 4096,
                             emission_time_s: float = 0.0, pretrigger_s: float = 100.0e-9) -> np.ndarray:
    """`count` antennas on the Cherenkov cone (cos theta_c = 1/n) around the unit
    `axis` at `distance_m` from `center_m` (the shower maximum), equally spaced in
    azimuth; each window opens `pretrigger_s` before the Cherenkov arrival
    emission_time + n distance / c."""
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    helper = np.array([1.0, 0.0, 0.0]) if abs(axis[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(axis, helper)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(axis, e1)
    cos_c = 1.0 / refractive_index
    sin_c = math.sqrt(1.0 - cos_c * cos_c)
    obs = np.zeros(count, dtype=OBSERVER_DTYPE)
    for k in range(count):
        phi = 2.0 * math.pi * k / count
        radial = math.cos(phi) * e1 + math.sin(phi) * e2
        obs[k]["position_m"] = np.asarray(center_m) + distance_m * (cos_c * axis + sin_c * radial)
    obs["sample_rate_hz"] = sample_rate_hz
    obs["number_of_bins"] = number_of_bins
    obs["duration_s"] = (number_of_bins - 1) / sample_rate_hz
    obs["start_time_s"] = emission_time_s + refractive_index * distance_m / SPEED_OF_LIGHT - pretrigger_s
    return obs
