"""Host implementation of the library's counter-based random stream (Philox4x32-10,
Salmon et al., SC'11), vectorised with numpy. Used for host-side field initialisation
(``Initial_s``) and for replaying the exact draws the device consumed.

Stream layout (identical on the device):
    key = (seed_lo, seed_hi)
    proposal draw: ctr = (elem, lt | dir << 31, sweep, stream << 24 | chain)
        choice = floor(3 x0 / 2^32)  (index into the other three field values)
        u      = ((x1 >> 5) 2^26 + (x2 >> 6)) / 2^53
    initial field: ctr = (elem, lt - 1, 0xFFFFFFFF, stream << 24 | chain),
        value = 1 + floor(4 x0 / 2^32)
"""
from __future__ import annotations

import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint32) for x in (c0, c1, c2, c3))
    k0 = np.asarray(k0, dtype=np.uint32)
    k1 = np.asarray(k1, dtype=np.uint32)
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = _M0 * c0.astype(np.uint64)
            p1 = _M1 * c2.astype(np.uint64)
            n0 = (p1 >> np.uint64(32)).astype(np.uint32) ^ c1 ^ k0
            n1 = (p1 & _MASK).astype(np.uint32)
            n2 = (p0 >> np.uint64(32)).astype(np.uint32) ^ c3 ^ k1
            n3 = (p0 & _MASK).astype(np.uint32)
            c0, c1, c2, c3 = n0, n1, n2, n3
            k0 = k0 + _W0
            k1 = k1 + _W1
    return c0, c1, c2, c3


def _key(seed: int):
    return np.uint32(seed & 0xFFFFFFFF), np.uint32((seed >> 32) & 0xFFFFFFFF)


def proposal_draws(seed: int, chain: int, stream: int, sweep: int, lt: int, direction: int, n: int):
    """(choice[n], u[n]) for field elements 0..n-1 of slice lt."""
    k0, k1 = _key(seed)
    elem = np.arange(n, dtype=np.uint32)
    tag = np.uint32((lt | (direction << 31)) & 0xFFFFFFFF)
    x0, x1, x2, _ = philox4x32_10(elem, np.full(n, tag, np.uint32), np.full(n, sweep, np.uint32),
                                  np.full(n, (stream << 24) | (chain & 0xFFFFFF), np.uint32), k0, k1)
    choice = ((x0.astype(np.uint64) * np.uint64(3)) >> np.uint64(32)).astype(np.int64)
    u = ((x1 >> np.uint32(5)).astype(np.float64) * 67108864.0 + (x2 >> np.uint32(6)).astype(np.float64)) / 9007199254740992.0
    return choice, u


def initial_fields(seed: int, chain: int, ns: int, nt: int, stream: int = 0) -> np.ndarray:
    """Initial_s replacement: UInt8 [ns, nt] i.i.d. uniform over {1,2,3,4}."""
    k0, k1 = _key(seed)
    i = np.tile(np.arange(ns, dtype=np.uint32), nt)
    lt = np.repeat(np.arange(nt, dtype=np.uint32), ns)
    x0, _, _, _ = philox4x32_10(i, lt, np.full(ns * nt, 0xFFFFFFFF, np.uint32),
                                np.full(ns * nt, (stream << 24) | (chain & 0xFFFFFF), np.uint32), k0, k1)
    v = 1 + ((x0.astype(np.uint64) * np.uint64(4)) >> np.uint64(32)).astype(np.int64)
    return v.reshape(nt, ns).T.astype(np.uint8).copy()
