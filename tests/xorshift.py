"""The reference tests' xorshift32 generator (src/fluid_core.rs:11232-11239,
:11276-11283): three shifts, then the top 24 bits as an f32 in [0, 1)."""
import numpy as np


class XorShift32:
    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFF

    def __call__(self) -> np.float32:
        s = self.s
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        self.s = s
        return np.float32(s >> 8) / np.float32(1 << 24)
