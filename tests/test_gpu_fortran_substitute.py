"""The link-time substitutes for the reference's `module commutators` and
`module TS_commutators` (eom-imsrg_full_amd/fortran/commutators_gpu.f90,
ts_commutators_gpu.f90) against the reference itself.

oracle/ref.mk builds the same harness program twice: ref_harness against the
reference's own module commutators, and sub_harness against the substitute,
with the reference's IMSRG_MAGNUS.f90, generators.f90, IMSRG_ODE.f90 and
brute_force_testing.f90 compiled unmodified against it (their calls to
commutator_111 ... commutator_222_ph then run on the GPU).  Both run on the
same inputs; every output must agree to 1e-10 relative:

* each routine in the reference's call contract (overwrite / accumulate),
  including commutator_221 served from the preceding commutator_222_pp_hh's
  product, and the full call sequence of IMSRG_MAGNUS.f90:207-219;
* the reference's build_gs_w2 (its inner commutator through the substitute),
  BCH_EXPAND, MAGNUS_EXPAND, IMSRG_ODE's dHds;
* magnus_decouple itself (its step lines, H(s), Omega(s)), with the
  reference's generators and with the device generators passed as its
  external build_generator;
* every TS_commutator routine in its contract (TS_commutator_221 on the w1,
  w2 that TS_commutator_222_pp_hh left, 211 and 222_ph on the caller's
  Pandya matrices), BCH_TENSOR's call sequence and BCH_TENSOR itself
  (IMSRG_MAGNUS.f90:389-465, compiled unmodified against the substitute).
"""
import os

import numpy as np
import pytest

from helpers import assert_close, rel
from imsrg_mi355x import Basis, random_operator, synthetic_hamiltonian
from oracle import oracle as O

HAVE = os.path.exists(O.SUB_HARNESS) and os.path.exists(O.REF_HARNESS)
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not HAVE, reason="the harnesses are built where the reference is")]


@pytest.fixture(scope="module", params=[(2, 2), (4, 3)], ids=["emax2", "emax4"])
def pair(request):
    emax, k = request.param
    B = Basis.ho(emax, k)
    ref = O.RefRunner(B, threads=8)
    sub = O.RefRunner(B, threads=8, binary=O.SUB_HARNESS)
    Hn = ref.normal(synthetic_hamiltonian(B, seed=0))
    print(f"[substitute] emax={emax} neq={B.neq}", flush=True)
    return B, ref, sub, Hn


@pytest.mark.parametrize("hL,hR", [(-1, 1), (-1, -1)])
def test_each_routine_and_full_sequence(pair, hL, hR):
    B, ref, sub, Hn = pair
    L, R = random_operator(B, hL, 1), random_operator(B, hR, 2)
    want, got = ref.terms(hL, L, hR, R), sub.terms(hL, L, hR, R)
    for k in ("e110", "e220"):
        assert abs(got[k] - want[k]) <= 1e-10 * max(1.0, abs(want[k])), k
    for k in ("111", "121", "122", "222pphh", "221", "222ph"):
        print(f"[substitute] {k}: rel {rel(got[k], want[k]):.2e}", flush=True)
        assert_close(got[k], want[k], what=f"commutator_{k}")
    assert_close(sub.comm(hL, L, hR, R), ref.comm(hL, L, hR, R), what="IMSRG_MAGNUS.f90:207-219 sequence")


def test_generators_bch_magnus_dhds(pair):
    B, ref, sub, Hn = pair
    for kind in ("white", "atan", "w2"):
        assert_close(sub.gen(kind, Hn), ref.gen(kind, Hn), what=f"reference build_gs_{kind}")
    for kind in ("white", "atan", "w2"):
        assert_close(sub.gen("gpu_" + kind, Hn), ref.gen(kind, Hn), what=f"gpu_build_gs_{kind}")
    eta = ref.gen("white", Hn)
    G = 0.5 * eta
    assert_close(sub.bch(G, Hn), ref.bch(G, Hn), what="BCH_EXPAND")
    assert_close(sub.magexp(eta, G), ref.magexp(eta, G), what="MAGNUS_EXPAND")
    for kind in ("white", "w2"):
        assert_close(sub.dhds(kind, Hn), ref.dhds(kind, Hn), what=f"dHds {kind}")
    assert_close(sub.dhds("gpu_white", Hn), ref.dhds("white", Hn), what="dHds with gpu_build_gs_white")


@pytest.mark.parametrize("kind,ref_kind", [("white", "white"), ("w2", "w2"), ("gpu_white", "white"),
                                           ("gpu_w2", "w2")])
def test_magnus_decouple(pair, kind, ref_kind):
    B, ref, sub, Hn = pair
    hs, g, rows = sub.decouple(kind, Hn)
    hs0, g0, rows0 = ref.decouple(ref_kind, Hn)
    print(f"[substitute] magnus_decouple {kind}: {len(rows)} lines, final E0 {rows[-1][2]}", flush=True)
    assert rows.shape == rows0.shape
    # the step lines carry 7 significant digits (e15.7); the first line's
    # E0 + MBPT2 column prints a variable magnus_decouple has not set yet
    # (IMSRG_MAGNUS.f90:72-74), so that entry is left out
    cols = [0, 1, 2, 4]
    assert np.allclose(rows[:, cols], rows0[:, cols], rtol=1e-6, atol=0)
    assert np.allclose(rows[1:, 3], rows0[1:, 3], rtol=1e-6, atol=0)
    assert_close(hs, hs0, what=f"magnus_decouple H(s) {kind}")
    assert_close(g, g0, what=f"magnus_decouple Omega(s) {kind}")


@pytest.mark.parametrize("case", [(4, 0, 1, -1), (2, 2, -1, 1)], ids=["E2", "E1anti"])
def test_ts_routines_and_bch_tensor(pair, case):
    B, ref, sub, Hn = pair
    rank, dpar, hT, hG = case
    G = random_operator(B, hG, 11 + rank)
    T, want = ref.tensor(hG, G, rank, dpar, hT, 5 + rank)
    T2, got = sub.tensor(hG, G, rank, dpar, hT, 5 + rank)
    assert np.array_equal(T, T2)  # the reference's generator, same seed, in both programs
    for k in ref.TENSOR_TERMS:
        print(f"[substitute] TS {k}: rel {rel(got[k], want[k]):.2e}", flush=True)
        assert_close(got[k], want[k], what=f"TS_commutator {k}")
    Gs = random_operator(B, -1, 23) * 0.02
    assert_close(sub.bch_tensor(-1, Gs, rank, dpar, hT, T), ref.bch_tensor(-1, Gs, rank, dpar, hT, T),
                 what="BCH_TENSOR through the substitute")
