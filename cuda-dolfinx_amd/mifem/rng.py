"""Counter-based, numbering-invariant random fields.

Every random value is a splitmix64 hash of a tuple of integers that names
the quantity geometrically (seed, lattice indices, sub-cell index, field id),
never of an array position.  Any cell or dof numbering therefore sees the
same field, which is what makes the synthetic workloads reproducible across
mesh orderings.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def hash_u64(*keys) -> np.ndarray:
    """splitmix64 chain over the given integer keys (scalars or arrays)."""
    with np.errstate(over="ignore"):
        h = np.uint64(0x243F6A8885A308D3)
        for k in keys:
            k = np.asarray(k).astype(np.uint64)
            h = _mix(h + _GOLDEN + k)
        return np.asarray(h)


def uniform(*keys) -> np.ndarray:
    """Uniform [0, 1) doubles from a hash chain (53 random bits)."""
    h = hash_u64(*keys)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(*keys) -> np.ndarray:
    """Standard normal via Box-Muller on two decorrelated uniform streams."""
    u1 = uniform(*keys, 0x51)
    u2 = uniform(*keys, 0xA7)
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
