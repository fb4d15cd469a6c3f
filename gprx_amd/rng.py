"""gprx's SeededRng (src/rng.rs:15-87): SplitMix64-seeded Xoshiro256++.

Used to draw the synthetic benchmark inputs exactly as the reference's
seeded configurations define them.  The first outputs for seeds 0 and 42 are
pinned by src/rng.rs:105-116 and tested in tests/test_rng.py.  Bulk draws
(units, normals) go through the same generator in libgprx_rng.so when it is
built, which yields the identical stream at native speed.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

_M64 = (1 << 64) - 1
_RNG_LIB_PATH = Path(__file__).resolve().parent / "libgprx_rng.so"
_rng_lib = None


def _native():
    global _rng_lib
    if _rng_lib is None and _RNG_LIB_PATH.exists():
        lib = C.CDLL(str(_RNG_LIB_PATH))
        u64p = C.POINTER(C.c_uint64)
        dp = C.POINTER(C.c_double)
        lib.gprx_rng_units.argtypes = [u64p, dp, C.c_size_t]
        lib.gprx_rng_units.restype = None
        lib.gprx_rng_normals.argtypes = [u64p, dp, C.c_size_t]
        lib.gprx_rng_normals.restype = None
        _rng_lib = lib
    return _rng_lib


def _rotl(x: int, k: int) -> int:
    return ((x << k) | (x >> (64 - k))) & _M64


class SeededRng:
    def __init__(self, seed: int):
        x = seed & _M64
        s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & _M64
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            s.append(z ^ (z >> 31))
        self.s = s

    def next_u64(self) -> int:
        s0, s1, s2, s3 = self.s
        result = (_rotl((s0 + s3) & _M64, 23) + s0) & _M64
        t = (s1 << 17) & _M64
        s2 ^= s0
        s3 ^= s1
        s1 ^= s2
        s0 ^= s3
        s2 ^= t
        s3 = _rotl(s3, 45)
        self.s = [s0, s1, s2, s3]
        return result

    def unit(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def _bulk(self, fn_name: str, n: int):
        lib = _native()
        if lib is None:
            return None
        st = np.array(self.s, dtype=np.uint64)
        out = np.empty(n, dtype=np.float64)
        getattr(lib, fn_name)(st.ctypes.data_as(C.POINTER(C.c_uint64)),
                              out.ctypes.data_as(C.POINTER(C.c_double)), n)
        self.s = [int(v) for v in st]
        return out

    def units(self, n: int, native: bool = True) -> np.ndarray:
        if native and n > 64:
            out = self._bulk("gprx_rng_units", n)
            if out is not None:
                return out
        return np.fromiter((self.unit() for _ in range(n)), dtype=np.float64, count=n)

    def open_unit(self) -> float:
        u = self.unit()
        eps = 1.0 / 9007199254740992.0
        return eps if u <= eps else (1.0 - eps if u >= 1.0 - eps else u)

    def unit_normal(self) -> float:
        u1 = max(self.open_unit(), 2.2250738585072014e-308)
        u2 = self.open_unit()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def normals(self, n: int, native: bool = True) -> np.ndarray:
        if native and n > 64:
            out = self._bulk("gprx_rng_normals", n)
            if out is not None:
                return out
        return np.fromiter((self.unit_normal() for _ in range(n)), dtype=np.float64, count=n)

    def up_to(self, mx: int) -> int:
        rng = mx + 1
        if rng > _M64:
            return self.next_u64()
        threshold = ((-rng) & _M64) % rng
        while True:
            wide = self.next_u64() * rng
            if (wide & _M64) >= threshold:
                return wide >> 64

    def shuffle(self, idx: np.ndarray) -> None:
        """Fisher-Yates of src/svgp/factor/adam.rs:85-90."""
        for i in range(len(idx) - 1, 0, -1):
            j = self.up_to(i)
            idx[i], idx[j] = idx[j], idx[i]
