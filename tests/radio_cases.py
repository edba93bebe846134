"""Radio inputs shared by the oracle and GPU tests.

REFERENCE_CASES restates the single-track geometries of the reference's radio
fixture (tests/kokkos/radio/RadioFixture.cpp:143-196): three observers at
(100, 2, 0), (17, 31, 0), (1, 0, 0) m sampling at 1 GHz from t = 0, a
uniform-index flat table from -1 km to 2 km at 1 m (also 0.5 m and 2 m for the
first case, RadioFixture.cpp:205-207). Those cases exercise the ordinary
endpoint path, weights and charge, Fraunhofer subdivision, the near-Cherenkov
branch, the exact lower table edge, late and window-end tracks, bin-edge and
sub-bin endpoints and vacuum.
"""
import math

import numpy as np

import c8mi

C = 299792458.0
HALF = 0.99 * C * 20.0e-9 / 2.0

# name, start, end, beta, start time (s), window duration (s), n, pid, weight
REFERENCE_CASES = [
    ("ordinary", (3., 4., 100.), (4., 6., 99.), .8, 0., 2.e-6, 1.0003, 11, 1.),
    ("positron_weight", (3., 4., 100.), (4., 6., 99.), .8, 0., 2.e-6, 1.0003, -11, 37.),
    ("fraunhofer", (3., 4., 100.), (4., 6., 10.), .99, 0., 2.e-6, 1.0003, 11, 1.),
    ("near_cone", (0., 0., 1000.), (0., 0., 990.), .999999, -3.e-6, 2.e-6, 1.0003, 11, 1.),
    ("lower_edge", (3., 4., -1000.), (4., 6., -1001.), .8, 0., 8.e-6, 1.0003, 11, 1.),
    ("late_track", (3., 4., 100.), (4., 6., 99.), .8, 1.e-3, 4.e-7, 1.0003, 11, 1.),
    ("window_end", (3., 4., 10.), (8., 4., 10.), .8, 7.181319171927605e-8, 4.e-7, 1.0003, 11, 1.),
    ("edge_first", (100. - HALF, 2., 1000.), (100. + HALF, 2., 1000.), .99,
     -1000. / C - 5.25e-9, 10.e-9, 1., 11, 1.),
    ("edge_last", (100. - HALF, 2., 1000.), (100. + HALF, 2., 1000.), .99,
     -1000. / C - 5.75e-9, 10.e-9, 1., 11, 1.),
    ("subbin_first", (100. - HALF / 100., 2., 1000.), (100. + HALF / 100., 2., 1000.), .99,
     -1000. / C - .35e-9, 10.e-9, 1., 11, 1.),
    ("subbin_last", (100. - HALF / 100., 2., 1000.), (100. + HALF / 100., 2., 1000.), .99,
     -1000. / C + 10.15e-9, 10.e-9, 1., 11, 1.),
    ("vacuum", (3., 4., 100.), (4., 6., 99.), .8, 0., 2.e-6, 1., 11, 1.),
]

FIXTURE_OBSERVERS = ((100., 2., 0.), (17., 31., 0.), (1., 0., 0.))


def case_track(start, end, beta, t0, pid, weight):
    start = np.asarray(start, dtype=np.float64)
    end = np.asarray(end, dtype=np.float64)
    dt = float(np.linalg.norm(end - start)) / (C * beta)
    track = np.zeros((9, 1))
    track[0:3, 0] = start
    track[3:6, 0] = end
    track[6, 0] = t0
    track[7, 0] = t0 + dt
    track[8, 0] = -weight if pid == 11 else weight
    return track


def case_observers(duration, rate=1.0e9, positions=FIXTURE_OBSERVERS, start_time=0.0):
    obs = np.zeros(len(positions), dtype=c8mi.OBSERVER_DTYPE)
    for i, p in enumerate(positions):
        obs[i]["position_m"] = p
    obs["start_time_s"] = start_time
    obs["duration_s"] = duration
    obs["sample_rate_hz"] = rate
    # TimeDomainObserver: num_bins = duration * rate + 1.5, truncated
    # (corsika/detail/modules/radio/observers/TimeDomainObserver.inl:17)
    obs["number_of_bins"] = int(duration * rate + 1.5)
    return obs


def uniform_table(index, step=1.0, lower=-1000.0, upper=2000.0):
    count = int((upper - lower) * (1.0 / step) + 1)
    return c8mi.Propagation(refractivity=np.full(count, index - 1.0), minimum_height_m=lower,
                            maximum_height_m=upper, step_m=step)


def reference_case(name, step=1.0):
    for case in REFERENCE_CASES:
        if case[0] == name:
            _, start, end, beta, t0, duration, n, pid, weight = case
            return (case_track(start, end, beta, t0, pid, weight), case_observers(duration),
                    uniform_table(n, step))
    raise KeyError(name)


def all_reference_cases():
    out = [(c[0], 1.0) for c in REFERENCE_CASES]
    out += [("ordinary", 0.5), ("ordinary", 2.0)]
    return out


def random_tracks(count, seed, box=200.0, height=(50.0, 1500.0), length=(0.2, 6.0),
                  beta=(0.6, 1.0), t_spread=2.0e-8):
    """Tracks with random directions, for parity tests beyond the axis-aligned
    config-2 list."""
    rng = np.random.default_rng(seed)
    start = np.stack([rng.uniform(-box, box, count), rng.uniform(-box, box, count),
                      rng.uniform(height[0], height[1], count)])
    d = rng.normal(size=(3, count))
    d /= np.linalg.norm(d, axis=0)
    d[2] = -np.abs(d[2])
    ell = rng.uniform(length[0], length[1], count)
    b = rng.uniform(beta[0], beta[1], count)
    end = start + d * ell
    t0 = rng.uniform(0.0, t_spread, count) - start[2] / C
    out = np.empty((9, count))
    out[0:3] = start
    out[3:6] = end
    out[6] = t0
    out[7] = t0 + ell / (b * C)
    out[8] = np.where(rng.random(count) < 0.6, -1.0, 1.0) * rng.uniform(0.5, 2.0, count)
    return out


def window_observers(tracks, positions, rate, nbins, index=1.0003, pretrigger=10.0e-9):
    obs = case_observers((nbins - 1) / rate, rate, positions)
    obs["number_of_bins"] = nbins
    start = tracks[0:3].T
    for k, p in enumerate(positions):
        arrival = tracks[6] + index * np.linalg.norm(start - np.asarray(p), axis=1) / C
        obs[k]["start_time_s"] = arrival.min() - pretrigger
    return obs


def em_constant():
    """1 / (4 pi eps0 c) from the SI constants the reference uses
    (corsika/kokkos/ScalarPhysicalConstants.hpp:19-21)."""
    eps0 = 8.8541878128e-12
    return 1.0 / (4.0 * math.pi * eps0 * C)


ELEMENTARY_CHARGE = 1.6021766208e-19
