"""The C++ standard library generators the reference's radar and CRPA draw their receiver noise
from (src/Radar.cpp:581-582, src/Crpa.cpp:436-437): std::mt19937 and libstdc++'s
std::normal_distribution<double>, restated so that the same seeds give the same noise.

- std::mt19937 (ISO C++ [rand.predef]: the 10000th output of a default-constructed engine is
  4123659995) is numpy's legacy MT19937 with the same 32-bit seeding (init_genrand).
- generate_canonical<double, 53> over a 32-bit engine takes two draws:
  u = (g1 + g2 * 2^32) / 2^64, held below 1.
- normal_distribution (libstdc++, Marsaglia's polar method): x, y = 2u - 1 until
  0 < r2 = x^2 + y^2 <= 1; returns y sqrt(-2 ln r2 / r2) and keeps x sqrt(-2 ln r2 / r2) for the
  next call; value = mean + stddev * z.
"""
from __future__ import annotations

import math

import numpy as np

_TWO32 = 4294967296.0
_TWO64 = 18446744073709551616.0


class Mt19937:
    """std::mt19937 seeded with one 32-bit value; raw draws in blocks."""

    def __init__(self, seed: int = 5489):
        self._bg = np.random.RandomState(int(seed) & 0xFFFFFFFF)._bit_generator
        self._buf = np.empty(0, np.uint64)
        self._pos = 0

    def __call__(self) -> int:
        if self._pos >= self._buf.size:
            self._buf = self._bg.random_raw(4096)
            self._pos = 0
        v = int(self._buf[self._pos])
        self._pos += 1
        return v

    def canonical(self) -> float:
        """generate_canonical<double, 53>(engine)."""
        s = float(self()) + float(self()) * _TWO32
        r = s / _TWO64
        return r if r < 1.0 else math.nextafter(1.0, 0.0)


class NormalDistribution:
    """libstdc++ std::normal_distribution<double>(mean, stddev)."""

    def __init__(self, mean: float = 0.0, stddev: float = 1.0):
        self.mean = float(mean)
        self.stddev = float(stddev)
        self._saved = None

    def __call__(self, eng: Mt19937) -> float:
        if self._saved is not None:
            z, self._saved = self._saved, None
        else:
            while True:
                x = 2.0 * eng.canonical() - 1.0
                y = 2.0 * eng.canonical() - 1.0
                r2 = x * x + y * y
                if 0.0 < r2 <= 1.0:
                    break
            mult = math.sqrt(-2.0 * math.log(r2) / r2)
            self._saved = x * mult
            z = y * mult
        return z * self.stddev + self.mean
