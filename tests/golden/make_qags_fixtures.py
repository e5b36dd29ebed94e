"""Write tests/golden/qags.npz: QUADPACK QAGS (scipy.integrate.quad, which wraps
QUADPACK's dqagse -- the algorithm GSL's gsl_integration_qags reimplements) on the
integrands the oracle's QAGS restatement is checked with.

Run in this container (scipy is installed here; nothing of the reference is run or
imported):  python tests/golden/make_qags_fixtures.py

Integrands
  * four test functions (kink, smooth, end-point singularities), the same ones the
    oracle exposes as orc_qags_test;
  * the line-of-sight luminosity density of ExponentialDisk3D as defined by
    function_objects/func_expdisk3d.cpp:177-268 (sky-plane rotation, tilt by the
    inclination, exponential radial profile, sech^(2/n) vertical profile switching to
    2^(2/n) exp(-|z|/z_0) beyond |z| / (2 z_0 / n) > 100), integrated over
    [-20 h, 20 h] with epsabs 0, epsrel 1e-6, limit 1000
    (function_objects/integrator.cpp:42-55).
"""
import math
import os
import sys

import numpy as np
from scipy import integrate

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import cases   # noqa: E402  (tests/cases.py: the integrand restated from the reference)


def test_function(which):
    if which == 0:
        return lambda s: math.exp(-abs(s))
    if which == 1:
        return lambda s: 1.0 / (1.0 + s * s)
    if which == 2:
        return lambda s: 1.0 / math.sqrt(s) if s > 0 else 0.0
    return lambda s: math.log(s) * math.sqrt(s) if s > 0 else 0.0


TEST_CASES = [(0, -3.0, 5.0), (0, -1.0, 1.0), (1, -10.0, 10.0), (1, 0.0, 3.0), (2, 0.0, 1.0), (3, 0.0, 1.0)]


def los_density(params6, x0, y0, x, y):
    return cases.disk_los_density(params6, x0, y0, x, y)[0]


# (PA, inc, J_0, h, n, z_0), centre, pixels
DISK_CASES = [
    ((0.0, 85.0, 1.0, 50.0, 100.0, 5.0), (256.5, 256.5),
     [(256.5, 256.5), (257.0, 256.0), (300.0, 256.5), (256.5, 300.0), (400.0, 270.0), (120.0, 200.0),
      (256.5, 262.0), (500.0, 500.0), (10.0, 256.0)]),
    ((30.0, 60.0, 2.0, 20.0, 1.0, 3.0), (64.0, 64.0),
     [(64.0, 64.0), (70.0, 60.0), (90.0, 64.0), (64.0, 100.0), (20.0, 30.0)]),
    ((-20.0, 89.5, 1.5, 30.0, 4.0, 2.0), (100.0, 100.0),
     [(100.0, 100.0), (100.0, 103.0), (140.0, 101.0), (60.0, 95.0)]),
    ((45.0, 30.0, 0.7, 15.0, 2.0, 4.0), (50.0, 50.0),
     [(50.0, 50.0), (55.0, 50.0), (50.0, 70.0), (80.0, 20.0)]),
]


def main():
    tv, tinfo = [], []
    for which, a, b in TEST_CASES:
        val, err, info = integrate.quad(test_function(which), a, b, epsabs=0.0, epsrel=1e-6, limit=1000,
                                        full_output=1)[:3]
        tv.append([val, err, info["neval"]])
        tinfo.append([which, a, b])
    dcase, dval = [], []
    for params, (x0, y0), pix in DISK_CASES:
        h = params[3]
        for (x, y) in pix:
            f = los_density(params, x0, y0, x, y)
            val, err, info = integrate.quad(f, -20.0 * h, 20.0 * h, epsabs=0.0, epsrel=1e-6, limit=1000,
                                            full_output=1)[:3]
            dcase.append(list(params) + [x0, y0, x, y])
            dval.append([val, err, info["neval"]])
    np.savez(os.path.join(HERE, "qags.npz"), test_cases=np.array(tinfo), test_results=np.array(tv),
             disk_cases=np.array(dcase), disk_results=np.array(dval))
    print("wrote qags.npz: %d test integrals, %d disk integrals" % (len(tv), len(dval)))


if __name__ == "__main__":
    main()
