"""java.util.Random, restated from its published specification, for regenerating seeded scenes.

The reference builds every test scene from ``new java.util.Random(seed)`` (e.g.
``ctest/sphere/UntilDoneTest.java:58-67``, ``ctest/sphere/GridTest.java:55-66``), so the same seeds must give the
same numbers here. The generator is the 48-bit linear congruential one the Java SE API documentation specifies:
``seed = (seed * 0x5DEECE66D + 0xB) mod 2**48``, ``next(bits) = (int) (seed >>> (48 - bits))``, with
``nextDouble() = ((long) next(26) << 27 + next(27)) * 2**-53`` and ``nextGaussian()`` the polar method.

``nextGaussian`` uses ``math.log``/``math.sqrt``; the JDK's ``StrictMath`` may differ from them in the last bit for
some arguments, which moves a Gaussian draw by one ulp at most.
"""

import math

import numpy as np

_MULT = 0x5DEECE66D
_ADD = 0xB
_MASK = (1 << 48) - 1


class JavaRandom:
    def __init__(self, seed: int):
        self._seed = (seed ^ _MULT) & _MASK
        self._next_gaussian = None

    def next(self, bits: int) -> int:
        self._seed = (self._seed * _MULT + _ADD) & _MASK
        r = self._seed >> (48 - bits)
        if r >= 1 << 31:
            r -= 1 << 32
        return r

    def nextInt(self, bound: int = None) -> int:
        if bound is None:
            return self.next(32)
        if bound <= 0:
            raise ValueError("bound must be positive")
        r = self.next(31)
        m = bound - 1
        if bound & m == 0:
            return (bound * r) >> 31
        u = r
        while True:
            r = u % bound
            if u - r + m < (1 << 31):
                return r
            u = self.next(31)

    def nextLong(self) -> int:
        v = (self.next(32) << 32) + self.next(32)
        v &= (1 << 64) - 1
        return v - (1 << 64) if v >= 1 << 63 else v

    def nextFloat(self) -> float:
        return self.next(24) / float(1 << 24)

    def nextDouble(self) -> float:
        return ((self.next(26) << 27) + self.next(27)) * (1.0 / (1 << 53))

    def nextGaussian(self) -> float:
        if self._next_gaussian is not None:
            g = self._next_gaussian
            self._next_gaussian = None
            return g
        while True:
            v1 = 2 * self.nextDouble() - 1
            v2 = 2 * self.nextDouble() - 1
            s = v1 * v1 + v2 * v2
            if 0 < s < 1:
                break
        multiplier = math.sqrt(-2 * math.log(s) / s)
        self._next_gaussian = v2 * multiplier
        return v1 * multiplier

    def doubles(self, count: int) -> np.ndarray:
        """``count`` successive ``nextDouble()`` values, as float64."""
        return np.array([self.nextDouble() for _ in range(count)], dtype=np.float64)
