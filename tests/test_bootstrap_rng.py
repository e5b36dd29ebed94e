"""The pure-Python MT19937 used to check bootstrap samples, against the C++ standard."""
from mt19937_ref import MT19937


def test_python_mt19937_matches_the_standard():
    # C++ [rand.predef]: the 10000th output of a default-constructed mt19937 (seed 5489)
    g = MT19937(5489)
    v = [g.int32() for _ in range(10000)]
    assert v[0] == 3499211612 and v[-1] == 4123659995
