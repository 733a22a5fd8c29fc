"""Rank > 0 operators: the scalar-tensor commutator terms (TS_commutators.f90)
and BCH_TENSOR (IMSRG_MAGNUS.f90:389-465) on the device against the
reference's own routines, run by oracle/_ref/ref_harness on the same inputs.

The tensor T comes from the reference's generator (construct_random_rankX,
brute_force_testing.f90:123-265) through the harness; G is a random scalar
operator.  Every term alone (on a cleared RES with herm -herm(G) herm(T), as
test_scalar_tensor_commutator sets OUT%herm), the whole commutator in
BCH_TENSOR's call order, and BCH_TENSOR with a small anti-hermitian G must
agree to 1e-10 relative.  Cases: E2-like (rank 2, even), M1-like (rank 1,
even, anti-hermitian T) and E1-like (rank 1, odd parity).
"""
import numpy as np
import pytest

from helpers import assert_close
from imsrg_mi355x import Basis, random_operator
from oracle import oracle as O

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.ref_available(), reason="the reference harness is built where the reference is")]

CASES = [(4, 0, 1, -1), (2, 0, -1, -1), (2, 2, 1, 1)]  # (rank, dpar, herm T, herm G), rank and dpar doubled
CASE_IDS = ["E2", "M1anti", "E1hermG"]


def tensor_blocks(B, rank, dpar):
    """(J1, J2, n1, n2) of every tensor block in allocate_tensor's order"""
    dims = {(b.J, b.Tz, b.P): b.dim for b in B.blocks}
    jm = max(b.J for b in B.blocks)
    out = []
    for J1 in range(0, jm + 1, 2):
        for J2 in range(max(abs(J1 - rank), J1), J1 + rank + 1, 2):
            for Tz in (-1, 0, 1):
                for P1 in (0, 1):
                    P2 = P1 ^ ((dpar // 2) & 1)
                    n1 = dims.get((J1, Tz, P1), 0) if J2 <= jm else 0
                    n2 = dims.get((J2, Tz, P2), 0) if J2 <= jm else 0
                    out.append((J1, J2, n1, n2))
    return out


def tensor_norm(B, rank, dpar, v):
    """mat_frob_norm of a tensor stream (basic_IMSRG.f90:3921-3952)"""
    A = int(np.sum(B.con))
    P = B.nsp - A
    k = 1
    s = np.sum(v[k:k + A * A] ** 2); k += A * A
    s += np.sum(v[k:k + P * P] ** 2); k += P * P
    s += 2 * np.sum(v[k:k + P * A] ** 2); k += P * A
    for J1, J2, n1, n2 in tensor_blocks(B, rank, dpar):
        s += (1 if J1 == J2 else 2) * np.sum(v[k:k + n1 * n2] ** 2)
        k += n1 * n2
    assert k == len(v)
    return float(np.sqrt(s))


@pytest.fixture(scope="module", params=[(2, 2), (4, 3)], ids=["emax2", "emax4"])
def env(request):
    from imsrg_mi355x import engine as E
    emax, k = request.param
    B = Basis.ho(emax, k)
    print(f"[tensor] emax={emax} nsp={B.nsp}", flush=True)
    return B, E.Context(B), O.RefRunner(B, threads=8), {}


def _case(env, case):
    B, ctx, ref, cache = env
    if case not in cache:
        rank, dpar, hT, hG = case
        G = random_operator(B, hG, 11 + rank + dpar)
        T, terms = ref.tensor(hG, G, rank, dpar, hT, 5 + rank)
        cache[case] = (G, T, terms)
    return cache[case]


@pytest.mark.parametrize("case", CASES, ids=CASE_IDS)
def test_tensor_stream_and_norm(env, case):
    B, ctx, ref, _ = env
    rank, dpar, hT, hG = case
    G, T, terms = _case(env, case)
    t = ctx.tensor(rank, dpar, hT, vec=T)
    A = int(np.sum(B.con))
    P = B.nsp - A
    assert t.neq == len(T) == 1 + A * A + P * P + P * A + sum(n1 * n2 for *_, n1, n2 in tensor_blocks(B, rank, dpar))
    assert np.array_equal(t.download(), T)
    assert abs(t.norm() - tensor_norm(B, rank, dpar, T)) <= 1e-12 * tensor_norm(B, rank, dpar, T)


@pytest.mark.parametrize("term", ["111", "121", "211", "122", "212", "222pphh", "221", "222ph", "full"])
@pytest.mark.parametrize("case", CASES, ids=CASE_IDS)
def test_ts_terms_vs_reference(env, case, term):
    from imsrg_mi355x import engine as E
    B, ctx, ref, _ = env
    rank, dpar, hT, hG = case
    G, T, terms = _case(env, case)
    dG = ctx.op(hG, G)
    dT = ctx.tensor(rank, dpar, hT, vec=T)
    mask = E.TSALL if term == "full" else E.TS_TERMS[term]
    got = ctx.tensor_commutator(dG, dT, mask=mask).download()
    want = terms[term]
    assert np.linalg.norm(want) > 0, f"{term}: reference output is zero"
    assert_close(got, want, tol=1e-10, what=f"TS_commutator {term}")


@pytest.mark.parametrize("case", CASES[:2], ids=CASE_IDS[:2])
def test_bch_tensor_vs_reference(env, case):
    B, ctx, ref, _ = env
    rank, dpar, hT, _ = case
    G, T, _ = _case(env, case)
    Gs = random_operator(B, -1, 23) * 0.02  # a small Omega: the series ends well before 30 terms
    want = ref.bch_tensor(-1, Gs, rank, dpar, hT, T)
    dT = ctx.tensor(rank, dpar, hT, vec=T)
    n = ctx.bch_tensor(ctx.op(-1, Gs), dT)
    print(f"[tensor] BCH_TENSOR {n} commutators", flush=True)
    assert 2 <= n < 29
    assert_close(dT.download(), want, tol=1e-10, what="BCH_TENSOR")
