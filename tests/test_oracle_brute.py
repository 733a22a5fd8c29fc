"""The oracle's blocked commutator against the unblocked brute-force J-coupled
sums (restated from brute_force_testing.f90:1765-2076), the reference's only
test of this engine.  The reference checks 0-body and 1-body to 1e-10 and
2-body to 1e-6 absolute (brute_force_testing.f90:517,546,591); the same
bounds are used here, on every 1-body element and a seeded set of 2-body
elements, at emax=2.
"""
import numpy as np
import pytest

from imsrg_mi355x import Basis, random_operator
from oracle import oracle as O


@pytest.fixture(scope="module", params=[(-1, 1), (-1, -1), (1, 1)], ids=str)
def case(request):
    hA, hB = request.param
    B = Basis.ho(2, 2)
    ob = O.OBasis(B)
    A = ob.op(random_operator(B, hA, 21), hA)
    Bo = ob.op(random_operator(B, hB, 22), hB)
    return B, ob, A, Bo, O.commutator(A, Bo)


def test_zero_body(case):
    B, ob, A, Bo, C = case
    assert abs(C.e0 - O.lib().or_brute_0body(A.h, Bo.h)) <= 1e-10


def test_one_body_all_elements(case):
    B, ob, A, Bo, C = case
    for a in range(B.nsp):
        for b in range(B.nsp):
            want = O.lib().or_brute_1body(A.h, Bo.h, a, b)
            assert abs(C.f(a, b) - want) <= 1e-10, (a, b)


def test_two_body_sampled(case):
    B, ob, A, Bo, C = case
    rng = np.random.Generator(np.random.PCG64(9))
    n = 0
    while n < 25:
        a, b, c, d = (int(x) for x in rng.integers(0, B.nsp, 4))
        if (B.l[a] + B.l[b]) % 2 != (B.l[c] + B.l[d]) % 2 or B.tz[a] + B.tz[b] != B.tz[c] + B.tz[d]:
            continue
        jmin = max(abs(B.jj[a] - B.jj[b]), abs(B.jj[c] - B.jj[d]))
        jmax = min(B.jj[a] + B.jj[b], B.jj[c] + B.jj[d])
        for J in range(jmin, jmax + 1, 2):
            want = O.lib().or_brute_2body(A.h, Bo.h, a, b, c, d, J)
            assert abs(C.v(a, b, c, d, J) - want) <= 1e-6, (a, b, c, d, J)
        n += 1
