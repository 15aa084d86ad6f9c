"""A Python std::mt19937_64 and the two libstdc++ draws the validation inputs need.

Test helper: produces, for a seed, the same values the reference's host code draws before and
around its sweeps — State::random / SQAState::random spins (uniform_int_distribution<int>(0,1):
one engine call per spin, spin = +1 iff the call's top bit is set), raw engine calls for the
per-replica seeds, and uniform_real_distribution(0,1) values (generate_canonical: one call x,
x * 2^-64 clamped below 1). The engine is the published 64-bit Mersenne Twister (Matsumoto &
Nishimura; parameters of std::mt19937_64). It is pinned by the standard's 10000th-output value
and by agreement with the oracle's libstdc++ runs in tests/test_validation_inputs.py.
"""
import numpy as np

_NN, _MM = 312, 156
_MATRIX_A = 0xB5026F5AA96619E9
_UM, _LM = 0xFFFFFFFF80000000, 0x7FFFFFFF
_MASK = (1 << 64) - 1


class MT64:
    def __init__(self, seed: int = 5489):
        mt = [0] * _NN
        mt[0] = seed & _MASK
        for i in range(1, _NN):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & _MASK
        self.mt, self.i = mt, _NN

    def _twist(self):
        mt = self.mt
        for i in range(_NN):
            x = (mt[i] & _UM) | (mt[(i + 1) % _NN] & _LM)
            xa = x >> 1
            if x & 1:
                xa ^= _MATRIX_A
            mt[i] = mt[(i + _MM) % _NN] ^ xa
        self.i = 0

    def __call__(self) -> int:
        if self.i >= _NN:
            self._twist()
        x = self.mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & _MASK

    def raw(self, k: int) -> np.ndarray:
        return np.array([self() for _ in range(k)], dtype=np.uint64)

    def spins(self, k: int) -> np.ndarray:
        return np.array([1 if (self() >> 63) else -1 for _ in range(k)], dtype=np.int8)

    def uniforms(self, k: int) -> np.ndarray:
        out = np.empty(k)
        for j in range(k):
            u = float(self()) * 2.0 ** -64
            out[j] = u if u < 1.0 else np.nextafter(1.0, 0.0)
        return out
