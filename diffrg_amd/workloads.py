"""Seeded synthetic inputs of the five integrator workloads.

All randomness comes from SplitMix64 (Steele, Lea & Flood 2014) with seed 0xD1FF0000 + c for
workload c; u = (z >> 11) * 2^-53.  The arrays built here are the inputs of both the GPU path
and the CPU oracle, so both see bitwise-identical data.

  0 threshold    256 LPA threshold integrals, m2 ~ U[-0.8, 4], 64-node GL in q^2 over [0, 20]
  1 selfenergy   64 external p on LogarithmicCoordinates1D(64, 1e-3, 40, 6), two-angle form
                 {64, 32, 32}, Z on LogarithmicCoordinates1D(128, 1e-3, 40, 6)
  2 finite_T     512 (p0 = 2 pi n T, |p|) externals, T in {0.01,0.05,0.1,0.2,0.3},
                 explicit Matsubara sum n in [-64, 64], 48-node GL in |q| over [0, 6]
  3 uneven       4096 self-energy integrals, radial order in {16,32,64,128}, angular order in
                 {8,16,32,64}, own 32-256 point Z table and own p (log-uniform in [1e-3, 40])
  4 flows        LPA O(4) flows, n_grid in [64, 1024], T in [0.01, 0.3], mu in [0, 0.4]
"""
from __future__ import annotations

import math

import numpy as np

SEED0 = 0xD1FF0000
GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """First n outputs of SplitMix64 started at `seed` (uint64)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + GAMMA * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


class Stream:
    """Sequential draws from one SplitMix64 stream."""

    def __init__(self, seed: int, size: int = 1 << 16):
        self._u = uniform01(seed, size)
        self._i = 0
        self._seed = seed

    def take(self, n: int) -> np.ndarray:
        if self._i + n > len(self._u):
            self._u = uniform01(self._seed, max(2 * len(self._u), self._i + n))
        out = self._u[self._i:self._i + n]
        self._i += n
        return out


def log_coord_forward(n: int, start: float, stop: float, bias: float) -> np.ndarray:
    """LogarithmicCoordinates1D::forward for i = 0..n-1 (coordinates.hh:170-174)."""
    b = (stop - start) / math.expm1(bias)
    i = np.arange(n)
    return b * np.expm1(bias * i * (1.0 / (n - 1.0))) + start


def dressing_table(n: int, draws: np.ndarray) -> np.ndarray:
    """Z_i = 1 + 0.5 tanh(sum_{m=1..4} A_m sin(m pi i/(n-1) + phi_m)), draws = (A_1, phi_1, ...)."""
    i = np.arange(n)
    xi = np.zeros(n)
    for m in range(1, 5):
        A = (2.0 * draws[2 * (m - 1)] - 1.0) / m
        phi = 2.0 * math.pi * draws[2 * (m - 1) + 1]
        xi += A * np.sin(m * math.pi * i / (n - 1) + phi)
    return 1.0 + 0.5 * np.tanh(xi)


Z_START, Z_STOP, Z_BIAS = 1e-3, 40.0, 6.0
X_EXTENT = 20.0


def threshold(n: int = 256):
    s = Stream(SEED0 + 0)
    m2 = -0.8 + 4.8 * s.take(n)
    return {"m2": m2, "nx": 64, "x_extent": X_EXTENT, "k": 1.0, "d": 4}


def selfenergy(n_p: int = 64, nq: int = 64, na: int = 32, n_tab: int = 128):
    s = Stream(SEED0 + 1)
    Z = dressing_table(n_tab, s.take(8))
    p = log_coord_forward(n_p, Z_START, Z_STOP, Z_BIAS)
    return {"p": p, "nq": nq, "na": na, "Z": Z, "n_tab": n_tab, "coord": (Z_START, Z_STOP, Z_BIAS),
            "x_extent": X_EXTENT, "k": 1.0}


TEMPS = (0.01, 0.05, 0.1, 0.2, 0.3)


def finite_T(n: int = 512, nx: int = 48, n_max: int = 64, q_max: float = 6.0):
    s = Stream(SEED0 + 2)
    u = s.take(5 * n).reshape(n, 5)
    T = np.array(TEMPS)[np.minimum((5 * u[:, 0]).astype(int), 4)]
    n_ext = np.minimum((8 * u[:, 1]).astype(int), 7)
    p0 = 2.0 * math.pi * n_ext * T
    pabs = 2.0 * u[:, 2]
    m1 = -0.5 + 3.5 * u[:, 3]
    m2 = -0.5 + 3.5 * u[:, 4]
    # explicit Matsubara sum n in [-n_max, n_max]: nodes 2 pi n T (n >= 1), weights T, plus T f(0)
    rules = {t: (2.0 * math.pi * t * np.arange(1, n_max + 1), np.full(n_max, t)) for t in TEMPS}
    return {"T": T, "p0": p0, "pabs": pabs, "m1": m1, "m2": m2, "rules": rules, "nx": nx, "q_max": q_max,
            "k": 1.0, "d": 4, "radial_mode": 1}


def uneven(n: int = 4096):
    s = Stream(SEED0 + 3)
    nq = np.empty(n, np.int32)
    na = np.empty(n, np.int32)
    p = np.empty(n)
    tables = []
    for i in range(n):
        u = s.take(4)
        nq[i] = (16, 32, 64, 128)[min(int(4 * u[0]), 3)]
        na[i] = (8, 16, 32, 64)[min(int(4 * u[1]), 3)]
        n_tab = 32 + min(int(225 * u[2]), 224)
        p[i] = math.exp(math.log(1e-3) + u[3] * math.log(4e4))
        tables.append(dressing_table(n_tab, s.take(8)))
    return {"p": p, "nq": nq, "na": na, "tables": tables, "coord": (Z_START, Z_STOP, Z_BIAS), "x_extent": X_EXTENT,
            "k": 1.0}


FLOW_RHO_MAX, FLOW_N, FLOW_M2, FLOW_LAMBDA = 4.0, 4.0, 0.2, 1.0


def flows(n: int = 64, rate: float = 1.0):
    """n flow requests with exponential inter-arrival times of mean 1/rate (seconds)."""
    s = Stream(SEED0 + 4)
    out = []
    t = 0.0
    for i in range(n):
        u = s.take(4)
        t += -math.log(1.0 - u[3]) / rate
        out.append({"n_grid": 64 + min(int(961 * u[0]), 960), "T": 0.01 + 0.29 * u[1], "mu": 0.4 * u[2],
                    "rho_max": FLOW_RHO_MAX, "N": FLOW_N, "m2_0": FLOW_M2, "lambda": FLOW_LAMBDA, "arrival": t})
    return out
