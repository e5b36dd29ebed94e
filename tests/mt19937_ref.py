"""Pure-Python MT19937 and imfit's bootstrap draw rule, for checking the library's
bootstrap samples (test helper; the published generator of Matsumoto & Nishimura,
as core/mersenne_twister.cpp:60-172 uses it)."""
import numpy as np


class MT19937:
    """Matsumoto & Nishimura's generator (init_genrand seeding, genrand_int32)."""

    def __init__(self, seed):
        self.mt = [seed & 0xffffffff]
        for i in range(1, 624):
            prev = self.mt[-1]
            self.mt.append((1812433253 * (prev ^ (prev >> 30)) + i) & 0xffffffff)
        self.i = 624

    def int32(self):
        mt = self.mt
        if self.i >= 624:
            for k in range(624):
                y = (mt[k] & 0x80000000) | (mt[(k + 1) % 624] & 0x7fffffff)
                mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908b0df if y & 1 else 0)
            self.i = 0
        y = mt[self.i]
        self.i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9d2c5680
        y ^= (y << 15) & 0xefc60000
        y ^= y >> 18
        return y

    def real2(self):
        return self.int32() * (1.0 / 4294967296.0)


def draw_sample(rng, n_data, valid):
    """MakeBootstrapSample's draws (core/model_object.cpp:2073-2084)."""
    out = []
    for _ in range(int(valid.sum())):
        while True:
            n = int(np.floor(rng.real2() * n_data))
            if valid[n]:
                break
        out.append(n)
    return np.array(out)
