"""Seeded synthetic inputs: SplitMix64 words with 255-bit rejection sampling.

The sampling scheme (four SplitMix64 words per candidate, top word masked to 63
bits, accept if below the Fr modulus) is the one the reference's own edge
tests use (backend/src/ascend910b3/native/tests/sumcheck_tensor_edge.cpp:62-80);
it is vectorised here with numpy and yields exactly the sequential stream.
"""
from __future__ import annotations

import numpy as np

from .fields import R_MODULUS

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_R_LIMBS = [np.uint64((R_MODULUS >> (64 * i)) & 0xFFFFFFFFFFFFFFFF) for i in range(4)]


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class SplitMix64Stream:
    """Vectorised SplitMix64; `words(k)` returns the next k outputs."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def words(self, k: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            steps = np.arange(1, k + 1, dtype=np.uint64) * _GAMMA
            out = _mix(self.state + steps)
            self.state = self.state + np.uint64(k) * _GAMMA
        return out

    def fr_limbs(self, n: int) -> np.ndarray:
        """n canonical Fr values as (n, 4) uint64 limbs (little-endian)."""
        got = []
        have = 0
        while have < n:
            want = max(16, int((n - have) * 2.3) + 16)
            w = self.words(4 * want).reshape(want, 4)
            w[:, 3] &= np.uint64(0x7FFFFFFFFFFFFFFF)
            lt = np.zeros(want, dtype=bool)
            eq = np.ones(want, dtype=bool)
            for limb in (3, 2, 1, 0):
                lt |= eq & (w[:, limb] < _R_LIMBS[limb])
                eq &= w[:, limb] == _R_LIMBS[limb]
            acc = w[lt]
            # the stream must not skip accepted candidates beyond what we keep:
            # rewind to just after the last candidate actually used
            need = n - have
            if acc.shape[0] > need:
                idx = np.flatnonzero(lt)[need - 1]
                unused = want - (idx + 1)
                with np.errstate(over="ignore"):
                    self.state = self.state - np.uint64(4 * unused) * _GAMMA
                acc = acc[:need]
            got.append(acc)
            have += acc.shape[0]
        return np.concatenate(got, axis=0) if got else np.zeros((0, 4), dtype=np.uint64)

    def fr_bytes(self, n: int) -> np.ndarray:
        """n canonical Fr values as an (n, 32) uint8 array."""
        return self.fr_limbs(n).astype("<u8").view(np.uint8).reshape(n, 32)


def random_fr_table(n: int, seed: int) -> np.ndarray:
    return SplitMix64Stream(seed).fr_bytes(n)


def limbs_to_ints(limbs: np.ndarray) -> list:
    return [int(a) | (int(b) << 64) | (int(c) << 128) | (int(d) << 192) for a, b, c, d in limbs.tolist()]
