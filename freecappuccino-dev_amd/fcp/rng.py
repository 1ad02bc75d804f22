"""Counter-based SplitMix64 random numbers (accepts seed 0).

``uniform01(seed, stream, i)`` is a pure function of its arguments:

    key = seed*0x9E3779B97F4A7C15 + stream*0xD1B54A32D192ED03 + i   (mod 2^64)
    z   = key + 0x9E3779B97F4A7C15
    z   = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9
    z   = (z ^ (z >> 27)) * 0x94D049BB133111EB
    z   =  z ^ (z >> 31)
    u   = (z >> 11) * 2^-53                                    in [0, 1)

The reference's own generator (Park-Miller ``r8vec_uniform_01``,
``src/linearSolvers/linear_solvers.f90:1557-1631``) rejects seed 0, which the
synthetic workloads use, so inputs are generated here and handed to every
consumer (oracle, reference driver, GPU path) as arrays.
"""
import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_S = np.uint64(0xD1B54A32D192ED03)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform01(seed: int, stream: int, i) -> np.ndarray:
    i = np.asarray(i, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = np.uint64(seed) * _G + np.uint64(stream) * _S + i
        z = _mix(key + _G)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform(seed: int, stream: int, n: int, lo: float, hi: float) -> np.ndarray:
    """n draws U[lo, hi): lo + (hi-lo)*u."""
    u = uniform01(seed, stream, np.arange(n, dtype=np.uint64))
    return lo + (hi - lo) * u
