"""Random number generator handles with the reference's factory signature.

``create(name, bufsize)`` mirrors steps.rng.create (src/steps/rng/create.cpp:44-55,
pysteps/cysteps_rng.pyx). The engine does not consume a sequential stream: every
(tet, period, channel) draws from its own counter-based Philox4x32-8 sequence,
and the seed given to ``initialize`` is the Philox key. All reference generator
names are accepted and give the same keyed streams.
"""
from __future__ import annotations

_NAMES = ("r123", "mt19937", "std_mt19937")


class RNG:
    def __init__(self, name: str, bufsize: int):
        if name not in _NAMES:
            raise ValueError(f"unknown RNG {name!r}; expected one of {_NAMES}")
        if bufsize <= 0:
            raise ValueError("bufsize must be positive")
        self.name = name
        self.bufsize = int(bufsize)
        self.seed = 0
        self.initialized = False

    def initialize(self, seed: int) -> None:
        seed = int(seed)
        if seed < 0 or seed >= 1 << 64:
            raise ValueError("seed must fit in 64 unsigned bits")
        self.seed = seed
        self.initialized = True


def create(name: str, bufsize: int) -> RNG:
    return RNG(name, bufsize)


def create_mt19937(bufsize: int) -> RNG:
    return RNG("mt19937", bufsize)
