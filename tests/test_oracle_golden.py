"""The oracle (C restatement) against golden vectors written by the
reference's own compiled FP64 routines (tests/golden/make_golden.py drives
oracle/_ref/ref_harness, built from the reference sources).

emax=2 (12 orbits, 2 holes/isospin): every commutator term for two
hermiticity pairs, the full commutator, Pandya matrices, normal ordering,
norm/MBPT2, the white/atan/w2 generators, BCH_EXPAND, MAGNUS_EXPAND, one
commutator [eta,H] plus one Magnus step, and complete white and w2 flows.
emax=4 (30 orbits, 3 holes/isospin): a complete white-generator flow.
"""
import os

import numpy as np
import pytest

from helpers import assert_close, rel
from imsrg_mi355x import Basis
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def g2():
    d = np.load(os.path.join(GOLD, "emax2_k2.npz"), allow_pickle=False)
    B = Basis.ho(2, 2)
    return B, O.OBasis(B), d


def test_fixture_inputs_match_generator(g2):
    # the stored inputs are the seeded synthetic/random operators
    from imsrg_mi355x import random_operator, synthetic_hamiltonian
    B, ob, d = g2
    assert d["H0"].size == B.neq
    assert np.allclose(d["L1"], random_operator(B, -1, 1))
    assert np.allclose(d["H0"], synthetic_hamiltonian(B, seed=0))


@pytest.mark.parametrize("tag,hL,hR,L,R", [("a", -1, 1, "L1", "R1"), ("b", -1, -1, "L2", "R2")])
def test_terms(g2, tag, hL, hR, L, R):
    B, ob, d = g2
    oL, oR = ob.op(d[L], hL), ob.op(d[R], hR)
    t = O.terms(oL, oR)
    for k in ("111", "121", "122", "222pphh", "221", "222ph"):
        assert_close(t[k], d[f"{tag}_{k}"], what=f"term {k}")
    for k in ("e110", "e220"):
        assert abs(t[k] - float(d[f"{tag}_{k}"][0])) <= 1e-10 * max(1.0, abs(t[k]))
    assert_close(O.commutator(oL, oR).vec(), d[f"{tag}_full"], what="full commutator")


def test_pandya(g2):
    B, ob, d = g2
    assert_close(O.cross_coupled(ob.op(d["L1"], -1)), d["cc_L1"], what="Pandya L1")
    assert_close(O.cross_coupled(ob.op(d["R1"], 1)), d["cc_R1"], what="Pandya R1")


def test_normal_order_and_scalars(g2):
    B, ob, d = g2
    H = O.normal_order(ob.op(d["H0"], 1))
    assert_close(H.vec(), d["Hn"], what="normal_order")
    nrm, mp2 = d["scal_Hn"]
    assert abs(O.frob_norm(H) - nrm) <= 1e-12 * nrm
    assert abs(O.mbpt2(H) - mp2) <= 1e-12 * abs(mp2)


@pytest.mark.parametrize("kind", ["white", "atan", "w2"])
def test_generators(g2, kind):
    B, ob, d = g2
    assert_close(O.generator(kind, ob.op(d["Hn"], 1)).vec(), d[f"gen_{kind}"], what=kind)


def test_bch_magnus_expand(g2):
    B, ob, d = g2
    hs, _ = O.bch(ob.op(d["G"], -1), ob.op(d["Hn"], 1))
    assert_close(hs.vec(), d["bch"], what="BCH_EXPAND")
    dg, _ = O.magnus_expand(ob.op(d["gen_white"], -1), ob.op(d["G"], -1))
    assert_close(dg.vec(), d["magexp"], what="MAGNUS_EXPAND")


def test_emax2_workload_one_commutator_one_step(g2):
    # the emax=2 CPU workload: one commutator [eta,H] and one Magnus step
    B, ob, d = g2
    eta = ob.op(d["gen_white"], -1)
    assert_close(O.commutator(eta, ob.op(d["Hn"], 1)).vec(), d["comm_eta_H"], what="[eta,H]")
    r = O.magnus_flow(ob.op(d["Hn"], 1), "white", 0.5, 1)
    rows = d["step1_rows"]
    assert r["steps"] == 1 and abs(r["e0"][0] - rows[0, 2]) <= 1e-12 * abs(rows[0, 2])
    assert_close(r["HS"].vec(), d["step1_HS"], what="H(s=0.5)")
    assert_close(r["G"].vec(), d["step1_G"], what="Omega(s=0.5)")


@pytest.mark.parametrize("kind", ["white", "w2"])
def test_full_flow_emax2(g2, kind):
    B, ob, d = g2
    rows = d[f"flow_{kind}_rows"]
    r = O.magnus_flow(ob.op(d["Hn"], 1), kind, 0.5, 200)
    assert r["steps"] == len(rows)
    assert rel(r["e0"], rows[:, 2]) <= 1e-12
    assert rel(r["e0_mbpt2"], rows[:, 3]) <= 1e-12
    assert_close(r["HS"].vec(), d[f"flow_{kind}_HS"], what="final H(s)")
    assert_close(r["G"].vec(), d[f"flow_{kind}_G"], what="final Omega(s)")


def test_full_flow_emax4():
    d = np.load(os.path.join(GOLD, "emax4_k3.npz"), allow_pickle=False)
    B = Basis.ho(4, 3)
    ob = O.OBasis(B)
    H = O.normal_order(ob.op(d["H0"], 1))
    rows = d["flow_white_rows"]
    r = O.magnus_flow(H, "white", 0.5, 200)
    assert r["steps"] == len(rows)
    assert rel(r["e0"], rows[:, 2]) <= 1e-12
    assert_close(r["HS"].vec(), d["flow_white_HS"], what="final H(s) emax=4")


def test_brute_force_golden(g2):
    # the reference's brute-force J-coupled sums at sampled points, evaluated by
    # the oracle's restatement of them
    B, ob, d = g2
    pts = d["brute_pts"]
    for tag, (hA, A, hB, Bn) in {"a": (-1, "L1", 1, "R1"), "b": (-1, "L2", -1, "R2")}.items():
        oA, oB = ob.op(d[A], hA), ob.op(d[Bn], hB)
        want = d[f"brute_{tag}"]
        for p, w in zip(pts, want):
            kind, a, b, c, dd, J = (int(x) for x in p)
            if kind == 0:
                got = O.lib().or_brute_0body(oA.h, oB.h)
            elif kind == 1:
                got = O.lib().or_brute_1body(oA.h, oB.h, a, b)
            else:
                got = O.lib().or_brute_2body(oA.h, oB.h, a, b, c, dd, J)
            assert abs(got - w) <= 1e-10 * max(1.0, abs(w)), (tag, p, got, w)


@pytest.mark.parametrize("emax,k", [(6, 3), (8, 6)], ids=["emax6", "emax8"])
def test_terms_at_scale_vs_reference_sketch(emax, k):
    """every term and the full commutator of [L, R] at the batching (emax=6)
    and precision (emax=8) sizes against the reference build's sketches
    (norm, projections, sampled entries; tests/helpers.py)"""
    from helpers import sketch_check
    from imsrg_mi355x import random_operator
    d = np.load(os.path.join(GOLD, f"emax{emax}_k{k}_sketch.npz"), allow_pickle=False)
    B = Basis.ho(emax, k)
    O.lib().or_set_threads(os.cpu_count() or 1)
    ob = O.OBasis(B)
    L, R = random_operator(B, -1, 1), random_operator(B, 1, 2)
    sketch_check(d, "input_L", L, exact=True)
    sketch_check(d, "input_R", R, exact=True)
    oL, oR = ob.op(L, -1), ob.op(R, 1)
    t = O.terms(oL, oR)
    assert abs(t["e110"] - float(d["e110"])) <= 1e-10 * max(1.0, abs(float(d["e110"])))
    assert abs(t["e220"] - float(d["e220"])) <= 1e-10 * max(1.0, abs(float(d["e220"])))
    for name in ("111", "121", "122", "222pphh", "221", "222ph"):
        sketch_check(d, f"term_{name}", t[name])
    sketch_check(d, "full", O.commutator(oL, oR).vec())
