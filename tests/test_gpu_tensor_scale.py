"""The scalar-tensor commutator at the workload size (emax=8, 6 holes per
isospin, a rank-2 tensor of 2.0e7 stored elements), through properties that
do not depend on the size:

* the whole commutator equals the sum of its terms, each run alone on a
  cleared RES (1e-12);
* linearity in the tensor: every term but 222_ph to 1e-12, 222_ph to 1e-7
  (its |W| > float32(1e-6) test, TS_commutators.f90:1562, 1714, is not
  linear: of the 2e8 product entries, the ~1e2 that fall within 1e-6 of zero
  for one of the three tensors but not the others each drop a ~1e-5
  contribution from up to a few hundred outputs; measured 1.05e-8 relative to
  the 222_ph term's norm);
* a hermitian tensor under an anti-hermitian G stays hermitian in its
  J1 = J2 blocks (the reference's own outputs keep this to 1e-15 at emax 2
  and 4; a transposed operand or a wrong phase in one half of any term
  breaks it);
* one-body operands: [G1, T1] is the matrix commutator F_G F_T - F_T F_G of
  the full one-body matrices (F_T[h][p] = herm (-1)^((jh-jp)/2) F_T[p][h]),
  with no two-body part.
The tensor is random with the symmetries of construct_random_rankX
(brute_force_testing.f90:123-265).
"""
import numpy as np
import pytest

from helpers import rel
from imsrg_mi355x import Basis, random_operator

pytestmark = pytest.mark.gpu
RANK, DPAR = 4, 0


def _sgn(e):
    return -1.0 if int(e) % 2 else 1.0


def random_tensor(B, rank, dpar, herm, seed):
    """tensor stream (E0, fhh, fpp, fph, tgam 1..9 per block) of a random
    rank/dpar tensor: selection rules on the one-body part, Pauli-forbidden
    rows and columns zero, J1 = J2 blocks with M = herm M^T"""
    rng = np.random.Generator(np.random.PCG64(seed))
    con = np.asarray(B.con)
    hol, par = np.flatnonzero(con == 1), np.flatnonzero(con == 0)
    jj, tz, l = np.asarray(B.jj), np.asarray(B.tz), np.asarray(B.l)

    def allowed(a, b):
        return (abs(jj[a] - jj[b]) <= rank <= jj[a] + jj[b] and tz[a] == tz[b]
                and l[a] % 2 == (l[b] + dpar // 2) % 2)

    def sq(lst):
        n = len(lst)
        M = np.zeros((n, n))
        for i in range(n):
            for j in range(i, n):
                if allowed(lst[i], lst[j]):
                    x = rng.uniform(-5, 5)
                    M[i, j] = x
                    M[j, i] = _sgn((jj[lst[i]] - jj[lst[j]]) // 2) * x * herm
            M[i, i] = (herm + 1) * M[i, i] / 2
        return M

    fhh, fpp = sq(hol), sq(par)
    fph = np.array([[rng.uniform(-5, 5) if allowed(p, h) else 0.0 for h in hol] for p in par]).reshape(len(par), len(hol))
    out = [np.zeros(1), fhh.ravel(order="F"), fpp.ravel(order="F"), fph.ravel(order="F")]
    dims = {(b.J, b.Tz, b.P): b for b in B.blocks}
    jm = max(b.J for b in B.blocks)
    sea1, sea2 = [0, 0, 0, 1, 2, 1, 2, 1, 2], [0, 1, 2, 1, 2, 2, 0, 0, 1]
    for J1 in range(0, jm + 1, 2):
        for J2 in range(max(abs(J1 - rank), J1), J1 + rank + 1, 2):
            for Tz in (-1, 0, 1):
                for P1 in (0, 1):
                    if J2 > jm:
                        continue
                    b1, b2 = dims[(J1, Tz, P1)], dims[(J2, Tz, P1 ^ ((dpar // 2) & 1))]
                    if b1.dim * b2.dim == 0:
                        continue
                    M = rng.uniform(-5, 5, (b1.dim, b2.dim))
                    p1 = [tuple(x) for c in range(3) for x in b1.pairs[c]]
                    p2 = [tuple(x) for c in range(3) for x in b2.pairs[c]]
                    for i, (a, b) in enumerate(p1):
                        if a == b and (J1 // 2) % 2 == 1:
                            M[i, :] = 0.0
                    for i, (a, b) in enumerate(p2):
                        if a == b and (J2 // 2) % 2 == 1:
                            M[:, i] = 0.0
                    if J1 == J2 and dpar == 0:
                        M = 0.5 * (M + herm * M.T)
                    s1 = [0, b1.npp, b1.npp + b1.nph, b1.dim]
                    s2 = [0, b2.npp, b2.npp + b2.nph, b2.dim]
                    for g in range(9):
                        out.append(M[s1[sea1[g]]:s1[sea1[g] + 1], s2[sea2[g]]:s2[sea2[g] + 1]].ravel(order="F"))
    return np.concatenate(out)


def square_blocks(B, rank, dpar, v):
    """the J1 = J2 blocks of a tensor stream as full matrices"""
    A = int(np.sum(B.con))
    P = B.nsp - A
    k = 1 + A * A + P * P + P * A
    dims = {(b.J, b.Tz, b.P): b for b in B.blocks}
    jm = max(b.J for b in B.blocks)
    sea1, sea2 = [0, 0, 0, 1, 2, 1, 2, 1, 2], [0, 1, 2, 1, 2, 2, 0, 0, 1]
    for J1 in range(0, jm + 1, 2):
        for J2 in range(max(abs(J1 - rank), J1), J1 + rank + 1, 2):
            for Tz in (-1, 0, 1):
                for P1 in (0, 1):
                    if J2 > jm:
                        continue
                    b1, b2 = dims[(J1, Tz, P1)], dims[(J2, Tz, P1 ^ ((dpar // 2) & 1))]
                    if b1.dim * b2.dim == 0:
                        continue
                    s1 = [0, b1.npp, b1.npp + b1.nph, b1.dim]
                    s2 = [0, b2.npp, b2.npp + b2.nph, b2.dim]
                    M = np.zeros((b1.dim, b2.dim))
                    for g in range(9):
                        r0, r1, c0, c1 = s1[sea1[g]], s1[sea1[g] + 1], s2[sea2[g]], s2[sea2[g] + 1]
                        m = (r1 - r0) * (c1 - c0)
                        M[r0:r1, c0:c1] = v[k:k + m].reshape((c1 - c0, r1 - r0)).T
                        k += m
                    if J1 == J2:
                        yield M
    assert k == len(v)


@pytest.fixture(scope="module")
def env():
    from imsrg_mi355x import engine as E
    B = Basis.ho(8, 6)
    ctx = E.Context(B)
    G = random_operator(B, -1, 31) * 0.05
    T1 = random_tensor(B, RANK, DPAR, 1, 41)
    T2 = random_tensor(B, RANK, DPAR, 1, 42)
    print(f"[tensor scale] emax=8 tensor stream {len(T1)}", flush=True)
    return E, B, ctx, G, T1, T2


def _comm(E, ctx, G, T, mask=None):
    dG = ctx.op(-1, G)
    dT = ctx.tensor(RANK, DPAR, 1, vec=T)
    return ctx.tensor_commutator(dG, dT, mask=E.TSALL if mask is None else mask).download()


def test_whole_equals_sum_of_terms(env):
    E, B, ctx, G, T1, _ = env
    full = _comm(E, ctx, G, T1)
    parts = sum(_comm(E, ctx, G, T1, m) for m in E.TS_TERMS.values())
    assert rel(parts, full) <= 1e-12, rel(parts, full)


def test_linear_in_the_tensor(env):
    E, B, ctx, G, T1, T2 = env
    exact = E.TSALL & ~E.TS_TERMS["222ph"]
    got = _comm(E, ctx, G, T1 + 2.0 * T2, exact)
    want = _comm(E, ctx, G, T1, exact) + 2.0 * _comm(E, ctx, G, T2, exact)
    assert rel(got, want) <= 1e-12, rel(got, want)
    ph = E.TS_TERMS["222ph"]
    got = _comm(E, ctx, G, T1 + 2.0 * T2, ph)
    want = _comm(E, ctx, G, T1, ph) + 2.0 * _comm(E, ctx, G, T2, ph)
    assert rel(got, want) <= 1e-7, rel(got, want)


def test_square_blocks_stay_hermitian(env):
    E, B, ctx, G, T1, _ = env
    num = den = 0.0
    for M in square_blocks(B, RANK, DPAR, T1):
        assert np.abs(M - M.T).max() <= 1e-12 * max(1.0, np.abs(M).max())
    for M in square_blocks(B, RANK, DPAR, _comm(E, ctx, G, T1)):  # RES herm = -(-1)(+1) = +1
        num += np.sum((M - M.T) ** 2)
        den += np.sum(M ** 2)
    assert np.sqrt(num / den) <= 1e-12, np.sqrt(num / den)


def test_one_body_operands_closed_form(env):
    E, B, ctx, G, T1, _ = env
    A = int(np.sum(B.con))
    P = B.nsp - A
    n1 = 1 + A * A + P * P + P * A
    _, fG, _ = B.full_from_vec(G, -1)  # G's full one-body matrix; the stream below drops its two-body part
    G1 = B.vec_from_full(0.0, fG, [np.zeros((b.dim, b.dim)) for b in B.blocks])
    T = T1.copy()
    T[n1:] = 0.0
    got = _comm(E, ctx, G1, T)
    assert np.abs(got[n1:]).max() == 0.0
    hol, par = np.flatnonzero(np.asarray(B.con) == 1), np.flatnonzero(np.asarray(B.con) == 0)
    jj = np.asarray(B.jj)
    FT = np.zeros((B.nsp, B.nsp))
    fhh = T[1:1 + A * A].reshape(A, A, order="F")
    fpp = T[1 + A * A:1 + A * A + P * P].reshape(P, P, order="F")
    fph = T[1 + A * A + P * P:n1].reshape(P, A, order="F")
    FT[np.ix_(hol, hol)] = fhh
    FT[np.ix_(par, par)] = fpp
    FT[np.ix_(par, hol)] = fph
    for i, p in enumerate(par):
        for j, h in enumerate(hol):
            FT[h, p] = fph[i, j] * _sgn((jj[h] - jj[p]) // 2)
    C = fG @ FT - FT @ fG
    want = np.concatenate([[0.0], C[np.ix_(hol, hol)].ravel(order="F"), C[np.ix_(par, par)].ravel(order="F"),
                           C[np.ix_(par, hol)].ravel(order="F")])
    assert rel(got[:n1], want) <= 1e-12, rel(got[:n1], want)
