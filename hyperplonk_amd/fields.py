"""BLS12-381 constants and the canonical host encodings of the C ABI.

Fr values cross the ABI as 32-byte little-endian canonical integers, G1 points
as 97-byte packets (tag 0 finite / 1 identity, x LE48, y LE48) — the formats of
backend/src/ascend910b3/encoding.rs:12-108.  Tables are handled as numpy uint8
arrays of shape (n, 32) so large inputs never pass through Python integers.
"""
from __future__ import annotations

from typing import Iterable, Optional, Sequence, Tuple, Union

import numpy as np

R_MODULUS = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
Q_MODULUS = int(
    "1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f6241eabfffeb153ffffb9feffffffffaaab",
    16,
)
FR_BYTES = 32
G1_BYTES = 97

G1Point = Optional[Tuple[int, int]]  # None is the identity
FrTable = Union[np.ndarray, bytes, Sequence[int]]


def fr_to_bytes(values: Iterable[int]) -> bytes:
    return b"".join((int(v) % R_MODULUS).to_bytes(32, "little") for v in values)


def fr_from_bytes(data: bytes) -> list:
    return [int.from_bytes(data[i : i + 32], "little") for i in range(0, len(data), 32)]


def table_bytes(values: FrTable) -> bytes:
    """Canonical LE32 bytes of a table given as ints, bytes or an (n, 32) uint8 array."""
    if isinstance(values, np.ndarray):
        arr = np.ascontiguousarray(values, dtype=np.uint8)
        return arr.tobytes()
    if isinstance(values, (bytes, bytearray, memoryview)):
        return bytes(values)
    return fr_to_bytes(values)


def g1_to_packet(point: G1Point) -> bytes:
    if point is None:
        return b"\x01" + bytes(96)
    x, y = point
    return b"\x00" + int(x).to_bytes(48, "little") + int(y).to_bytes(48, "little")


def g1_from_packet(packet: bytes) -> G1Point:
    if packet[0] == 1:
        return None
    return (int.from_bytes(packet[1:49], "little"), int.from_bytes(packet[49:97], "little"))
