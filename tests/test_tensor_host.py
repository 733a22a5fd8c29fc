"""Host-side checks of the tensor path that need no GPU: the reference
harness's tensor stream has the length and block order the device layout
assumes, the whole scalar-tensor commutator equals the sum of its terms
(the harness calls each TS_commutator term the way BCH_TENSOR does), and the
library exports the tensor entry points."""
import numpy as np
import pytest

from imsrg_mi355x import Basis, random_operator
from oracle import oracle as O
from test_gpu_tensor import CASES, CASE_IDS, tensor_blocks

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="the reference harness is built where the reference is")


@needs_ref
@pytest.mark.parametrize("case", CASES, ids=CASE_IDS)
def test_harness_tensor_stream_layout(case):
    B = Basis.ho(2, 2)
    rank, dpar, hT, hG = case
    ref = O.RefRunner(B, threads=2)
    T, terms = ref.tensor(hG, random_operator(B, hG, 3), rank, dpar, hT, 7)
    A = int(np.sum(B.con))
    P = B.nsp - A
    blocks = tensor_blocks(B, rank, dpar)
    assert len(T) == 1 + A * A + P * P + P * A + sum(n1 * n2 for *_, n1, n2 in blocks)
    assert T[0] == 0.0
    # construct_random_rankX: the one-body part obeys the tensor's selection rules
    fhh = T[1:1 + A * A].reshape(A, A, order="F")
    hol = np.flatnonzero(B.con)
    for i in range(A):
        for j in range(A):
            a, b = hol[i], hol[j]
            allowed = (abs(B.jj[a] - B.jj[b]) <= rank <= B.jj[a] + B.jj[b] and B.tz[a] == B.tz[b]
                       and (B.l[a] % 2) == ((B.l[b] + dpar // 2) % 2))
            if not allowed:
                assert fhh[i, j] == 0.0
    s = sum(terms[k] for k in ("111", "121", "211", "122", "212", "222pphh", "221", "222ph"))
    full = terms["full"]
    assert np.linalg.norm(s - full) <= 1e-12 * np.linalg.norm(full)


@needs_ref
def test_tensor_fixture_matches_live_reference():
    """the committed tensor fixture is what the reference build writes today"""
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tensor_emax2_k2.npz"))
    B = Basis.ho(2, 2)
    ref = O.RefRunner(B, threads=2)
    rank, dpar, hT, hG = 4, 0, 1, -1
    T, terms = ref.tensor(hG, d["E2_G"], rank, dpar, hT, 5 + rank)
    assert np.array_equal(T, d["E2_T"])
    for k, v in terms.items():
        assert np.linalg.norm(v - d[f"E2_{k}"]) <= 1e-13 * np.linalg.norm(v), k


def test_tensor_symbols_exported():
    from imsrg_mi355x import engine as E
    names = [n for n in E.header_symbols() if "tensor" in n]
    assert len(names) >= 12
    for n in names:
        assert hasattr(E._lib, n), n
