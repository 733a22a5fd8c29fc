"""Closed-form and structural pins of the oracle that a plausible mistake (a
dropped term, a wrong sign or index, a transposed operand) would break.
"""
import numpy as np
import pytest

from helpers import assert_close
from imsrg_mi355x import Basis, random_operator
from oracle import oracle as O


@pytest.fixture(scope="module")
def env():
    B = Basis.ho(4, 3)
    return B, O.OBasis(B)


def one_body_only(B, herm, seed):
    v = random_operator(B, herm, seed)
    e0, f, mats = B.full_from_vec(v, herm)
    return B.vec_from_full(0.0, f, [m * 0 for m in mats]), f


def test_antisymmetry(env):
    # [A,B] = -[B,A]
    B, ob = env
    A = ob.op(random_operator(B, -1, 1), -1)
    H = ob.op(random_operator(B, 1, 2), 1)
    assert_close(O.commutator(A, H).vec(), -O.commutator(H, A).vec(), tol=1e-12, what="antisymmetry")


def test_one_body_closed_form(env):
    # for one-body operators the commutator is the matrix commutator (111) with
    # E0 = sum_ij (n_i - n_j)(2j_i+1) A_ij B_ji = sum_{p,h} (2j_h+1)(A_hp B_ph - A_ph B_hp);
    # the two-body part vanishes
    B, ob = env
    va, fa = one_body_only(B, -1, 3)
    vb, fb = one_body_only(B, 1, 4)
    C = O.commutator(ob.op(va, -1), ob.op(vb, 1))
    e0, f, mats = B.full_from_vec(C.vec(), 1)
    assert_close(f, fa @ fb - fb @ fa, tol=1e-12, what="111 matrix commutator")
    assert max(np.abs(m).max() if m.size else 0 for m in mats) == 0.0
    P, H = B.parts, B.holes
    want = sum((B.jj[h] + 1) * (fa[h, p] * fb[p, h] - fa[p, h] * fb[h, p]) for p in P for h in H)
    assert abs(e0 - want) <= 1e-10 * max(1.0, abs(want))


def test_bch_zero_generator_is_identity(env):
    B, ob = env
    H = ob.op(random_operator(B, 1, 5), 1)
    G = ob.op(np.zeros(B.neq), -1)
    hs, n = O.bch(G, H)
    assert n == 1
    assert_close(hs.vec(), H.vec(), tol=0, what="BCH with Omega=0")


def test_bch_one_body_rotation(env):
    # a purely one-body Omega: IMSRG(2) BCH is exact, so the one-body part of
    # e^G f e^-G equals U f U^T with U = expm(G_1b) (orthogonal, (l,j,tz)-blocked)
    from scipy.linalg import expm
    B, ob = env
    vg, g = one_body_only(B, -1, 6)
    g *= 0.05
    vg = B.vec_from_full(0.0, g, [np.zeros((b.dim, b.dim)) for b in B.blocks])
    vh, f = one_body_only(B, 1, 7)
    hs, n = O.bch(ob.op(vg, -1), ob.op(vh, 1))
    _, fs, _ = B.full_from_vec(hs.vec(), 1)
    U = expm(g)
    assert_close(fs, U @ f @ U.T, tol=1e-8, what="one-body BCH rotation")


def test_norm_is_frobenius(env):
    B, ob = env
    for herm in (1, -1):
        v = random_operator(B, herm, 8)
        e0, f, mats = B.full_from_vec(v, herm)
        want = np.sqrt((f ** 2).sum() + sum((m ** 2).sum() for m in mats))
        assert abs(O.frob_norm(ob.op(v, herm)) - want) <= 1e-12 * want


def test_mbpt2_closed_form(env):
    B, ob = env
    v = random_operator(B, 1, 9)
    e0, f, mats = B.full_from_vec(v, 1)
    f = f + np.diag(np.arange(B.nsp) * 10.0 + 50 * (1 - B.con))  # separate p and h energies
    v = B.vec_from_full(e0, f, mats)
    want = 0.0
    for blk, M in zip(B.blocks, mats):
        for ii, (i, j) in enumerate(blk.pairs[0]):
            for jj_, (k, l) in enumerate(blk.pairs[2]):
                x = M[ii, blk.npp + blk.nph + jj_]
                want += (blk.J + 1) * x * x / (f[k, k] + f[l, l] - f[i, i] - f[j, j])
    assert abs(O.mbpt2(ob.op(v, 1)) - want) <= 1e-12 * abs(want)
