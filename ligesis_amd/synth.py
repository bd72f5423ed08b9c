"""Seeded synthetic inputs: SplitMix64 streams, rejection-sampled into Goldilocks.

The reference draws every input from OS entropy (`rand::rng()`, e.g.
pcs/benches/ligesis.rs:22-25), so no run of it is reproducible. Here every input comes from a
SplitMix64 stream with an explicit seed:

- a base-field element is the next stream value that is < p = 2^64 - 2^32 + 1;
- an extension element is two consecutive base elements (c0, then c1);
- an oracle integer is the raw stream value (the reference's `rng.random::<u64>() as usize`,
  utils/src/oracle.rs:396).

Vectorised with numpy (uint64 arithmetic wraps mod 2^64), identical to the sequential loop.
"""
from __future__ import annotations

import numpy as np

P = 0xFFFFFFFF00000001
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Values start .. start+count-1 of the SplitMix64 stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def field_elems(seed: int, count: int) -> np.ndarray:
    """`count` uniform Goldilocks elements (canonical uint64) by rejection sampling."""
    out = np.empty(0, np.uint64)
    drawn = 0
    while out.size < count:
        need = count - out.size
        chunk = splitmix64(seed, need + 64, drawn)
        drawn += chunk.size
        out = np.concatenate([out, chunk[chunk < np.uint64(P)]])
    return out[:count]


def ext_elems(seed: int, count: int) -> np.ndarray:
    """`count` uniform F_{p^2} elements as a (count, 2) uint64 array."""
    return field_elems(seed, 2 * count).reshape(count, 2)


def oracle_ints(seed: int, count: int) -> np.ndarray:
    return splitmix64(seed, count)


class Inputs:
    """All inputs of one LigeSIS instance, derived from one base seed."""

    def __init__(self, log_m: int, log_n: int, seed: int = 1, n_fields: int = 512,
                 n_ints: int = 1024, with_a: bool = True):
        self.log_m, self.log_n = log_m, log_n
        mu = log_m + log_n
        self.f = field_elems(seed * 8 + 0, 1 << mu)
        self.a = field_elems(seed * 8 + 1, 32 * (64 << log_m)) if with_a else None
        self.z = ext_elems(seed * 8 + 2, mu)
        self.fields = ext_elems(seed * 8 + 3, n_fields)
        self.ints = oracle_ints(seed * 8 + 4, n_ints)
