"""SplitMix64 (Steele, Lea & Flood 2014): the seeded generator the uneven
batch of Newton searches and the scan sample points are drawn from, so that
the draws do not depend on numpy's generator versions."""
from __future__ import annotations

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int = 0):
        self.state = seed & MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def below(self, n: int) -> int:
        """Integer in [0, n) (multiply-shift; bias < n / 2^64)."""
        return (self.next_u64() * n) >> 64
