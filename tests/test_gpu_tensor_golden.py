"""The tensor path against the reference's own outputs committed as fixtures
(tests/golden/tensor_emax2_k2.npz, written by tests/golden/make_golden.py
tensor from oracle/_ref/ref_harness): every TS_commutator term of [G, T]
alone on a cleared RES (herm -herm(G) herm(T)), the whole commutator in
BCH_TENSOR's call order, and BCH_TENSOR with a small anti-hermitian G, at
emax=2 for an E2 tensor (rank 2, even) under an anti-hermitian G and an E1
tensor (rank 1, odd parity) under a hermitian G; 1e-10 relative.  Needs no
reference build at run time."""
import os

import numpy as np
import pytest

from helpers import assert_close
from imsrg_mi355x import Basis

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tensor_emax2_k2.npz")
CASES = {"E2": (4, 0, 1, -1), "E1hermG": (2, 2, 1, 1)}
TERMS = ["111", "121", "211", "122", "212", "222pphh", "221", "222ph", "full"]


@pytest.fixture(scope="module")
def env():
    from imsrg_mi355x import engine as E
    B = Basis.ho(2, 2)
    return E, E.Context(B), np.load(FIX)


@pytest.mark.parametrize("term", TERMS)
@pytest.mark.parametrize("case", list(CASES))
def test_tensor_terms_vs_fixture(env, case, term):
    E, ctx, d = env
    rank, dpar, hT, hG = CASES[case]
    G, T = ctx.op(hG, d[f"{case}_G"]), ctx.tensor(rank, dpar, hT, vec=d[f"{case}_T"])
    mask = E.TSALL if term == "full" else E.TS_TERMS[term]
    got = ctx.tensor_commutator(G, T, mask=mask).download()
    assert_close(got, d[f"{case}_{term}"], tol=1e-10, what=f"{case} {term}")


@pytest.mark.parametrize("case", list(CASES))
def test_bch_tensor_vs_fixture(env, case):
    E, ctx, d = env
    rank, dpar, hT, _ = CASES[case]
    T = ctx.tensor(rank, dpar, hT, vec=d[f"{case}_T"])
    n = ctx.bch_tensor(ctx.op(-1, d[f"{case}_bchG"]), T)
    assert 2 <= n < 29
    assert_close(T.download(), d[f"{case}_bch"], tol=1e-10, what=f"{case} BCH_TENSOR")
