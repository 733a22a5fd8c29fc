"""6j / 9j symbols of the oracle and of the engine's table builder, pinned
against sympy's exact Wigner symbols and the 6j orthogonality relation.

The reference evaluates 6j through anglib's log-factorial Racah sum
(anglib.f:285-377) and looks them up through sixj (basic_IMSRG.f90:2509-2555);
9j only appear in its brute-force test (coef9, anglib.f:546-566).
"""
import itertools

import numpy as np
import pytest
from sympy import Rational
from sympy.physics.wigner import wigner_6j, wigner_9j

from imsrg_mi355x import Basis
from oracle import oracle as O


def h(x):  # doubled -> sympy value
    return Rational(x, 2)


@pytest.fixture(scope="module")
def ob():
    return O.OBasis(Basis.ho(4, 3))  # jj up to 9/2


def test_sixj_table_vs_sympy(ob):
    rng = np.random.Generator(np.random.PCG64(3))
    checked = 0
    halves = [1, 3, 5, 7, 9]
    for _ in range(400):
        j1, j2, j4, j5 = (int(x) for x in rng.choice(halves, 4))
        # mostly triangle-allowed arguments, plus some that are not (must give 0)
        lo3, hi3 = max(abs(j1 - j2), abs(j4 - j5)), min(j1 + j2, j4 + j5)
        lo6, hi6 = max(abs(j1 - j5), abs(j4 - j2)), min(j1 + j5, j4 + j2)
        if lo3 > hi3 or lo6 > hi6 or rng.random() < 0.1:
            J3, J6 = (int(x) * 2 for x in rng.integers(0, 10, 2))
        else:
            J3 = int(rng.choice(np.arange(lo3, hi3 + 1, 2)))
            J6 = int(rng.choice(np.arange(lo6, hi6 + 1, 2)))
        want = float(wigner_6j(h(j1), h(j2), h(J3), h(j4), h(j5), h(J6)))
        got = ob.sixj(j1, j2, J3, j4, j5, J6)
        assert abs(got - want) <= 1e-14 * max(1.0, abs(want)), (j1, j2, J3, j4, j5, J6, got, want)
        checked += want != 0
    assert checked > 100


def test_sixj_orthogonality(ob):
    # sum_x (2x+1)(2f+1) {a b x; c d f}{a b x; c d g} = delta_fg
    for a, b, c, d in itertools.product([1, 3, 5], repeat=4):
        fs = [f for f in range(0, 12, 2)]
        for f in fs:
            for g in fs:
                s = sum((x + 1) * (f + 1) * ob.sixj(a, b, x, c, d, f) * ob.sixj(a, b, x, c, d, g)
                        for x in range(0, 12, 2))
                tri_f = abs(a - d) <= f <= a + d and abs(c - b) <= f <= c + b
                want = 1.0 if (f == g and tri_f) else 0.0
                assert abs(s - want) < 1e-12, (a, b, c, d, f, g, s)


def test_ninej_vs_sympy():
    rng = np.random.Generator(np.random.PCG64(5))
    n = 0
    while n < 60:
        # 9j as used by the brute-force sums: {jj J1 ja; J2 ji jb; jc jd J}
        jj, ja, ji, jb, jc, jd = rng.choice([1, 3, 5], 6)
        J1, J2, J = (int(x) * 2 for x in rng.integers(0, 4, 3))
        args = (jj, J1, ja, J2, ji, jb, jc, jd, J)
        want = float(wigner_9j(*[h(x) for x in args], prec=30))
        got = O.lib().or_ninej(*[int(x) for x in args])
        assert abs(got - want) <= 1e-13 * max(1.0, abs(want)), (args, got, want)
        n += want != 0
