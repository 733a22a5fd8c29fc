"""The emax=8 workload (90 orbits, 6 holes per isospin, 108 channel blocks,
Pandya blocks of up to 1296 rows by 116 ph columns) against the oracle, at the
size the benchmark times.

* every commutator term and the full commutator, for all four hermiticity
  pairs, against the oracle's restatement of commutators.f90 at 1e-10
  relative;
* the first two steps of magnus_decouple (IMSRG_MAGNUS.f90:98-126) from the
  benchmark's normal-ordered Hamiltonian: E0 per step to 1e-10 relative, the
  same MAGNUS_EXPAND and BCH_EXPAND term counts, and H(s), Omega(s) after the
  second step to 1e-10;
* the reduced-precision 222 products (int8x5, int8x4, split-FP32, split-BF16) on the
  same two steps: |dE0| per step against the oracle's FP64 steps below
  1e-6 MeV and the FP64 term counts;
* the sketches (norm, four Gaussian projections, 256 sampled entries) of the
  reference build's own outputs for the same inputs (tests/golden).

The oracle runs one emax=8 commutator in about a second on 8 host threads and
the two flow steps (32 commutators) in about 40 s.
"""
import os

import numpy as np
import pytest

from helpers import assert_close, rel, sketch_check
from imsrg_mi355x import Basis, random_operator, synthetic_hamiltonian
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
HERMS = [(-1, 1), (-1, -1), (1, 1), (1, -1)]


@pytest.fixture(scope="module")
def w8():
    from imsrg_mi355x import engine as E
    B = Basis.ho(8, 6)
    print(f"[emax8 oracle] neq={B.neq}", flush=True)
    O.lib().or_set_threads(min(16, os.cpu_count() or 1))
    ctx = E.Context(B)
    yield B, ctx, O.OBasis(B), E
    ctx.close()


@pytest.mark.parametrize("hL,hR", HERMS)
def test_emax8_each_term_and_full(w8, hL, hR):
    B, ctx, ob, E = w8
    L = random_operator(B, hL, 1)
    R = random_operator(B, hR, 2)
    oL, oR = ob.op(L, hL), ob.op(R, hR)
    want = O.terms(oL, oR)
    dL, dR = ctx.op(hL, L), ctx.op(hR, R)
    assert abs(E.commutator_110(dL, dR) - want["e110"]) <= 1e-10 * max(1.0, abs(want["e110"]))
    assert abs(E.commutator_220(dL, dR) - want["e220"]) <= 1e-10 * max(1.0, abs(want["e220"]))
    masks = {"111": E.T111, "121": E.T121, "122": E.T122, "222pphh": E.T222PPHH, "221": E.T221,
             "222ph": E.T222PH}
    for name, m in masks.items():
        got = ctx.commutator(dL, dR, mask=m).download()
        print(f"[emax8 oracle] herm {hL},{hR} term {name}: rel {rel(got, want[name]):.2e}", flush=True)
        assert_close(got, want[name], what=f"emax8 term {name} herm {hL},{hR}")
    full = ctx.commutator(dL, dR).download()
    assert_close(full, O.commutator(oL, oR).vec(), what=f"emax8 full commutator herm {hL},{hR}")


def test_emax8_pandya_vs_oracle(w8):
    B, ctx, ob, E = w8
    for herm in (1, -1):
        A = random_operator(B, herm, 5)
        assert_close(ctx.cross_coupled(ctx.op(herm, A)), O.cross_coupled(ob.op(A, herm)),
                     what=f"emax8 Pandya herm {herm}")


@pytest.fixture(scope="module")
def oracle_steps(w8):
    """the benchmark's Hamiltonian (seed 0, hw = 20 MeV) normal-ordered on the
    device, and two oracle Magnus steps from it"""
    B, ctx, ob, E = w8
    H = ctx.op(1, synthetic_hamiltonian(B, seed=0, hw=20.0, sigma1=0.5, sigma2=1.0))
    H.normal_order()
    h0 = H.download()
    oH = ob.op(h0, 1)
    print("[emax8 oracle] two oracle Magnus steps (about 40 s on 8 threads)", flush=True)
    r = O.magnus_flow(oH, "white", 0.5, 2)
    assert r["steps"] == 2
    return h0, r


def _gpu_steps(ctx, h0, prec, n=2):
    ctx.set_precision(prec)
    HS = ctx.op(1, h0)
    f = ctx.flow(HS, "white", 0.5)
    recs = [f.step() for _ in range(n)]
    G = f.G.download()
    f.close()
    ctx.set_precision(0)
    return recs, HS.download(), G


def test_emax8_two_magnus_steps_fp64(w8, oracle_steps):
    B, ctx, ob, E = w8
    h0, r = oracle_steps
    recs, hs, g = _gpu_steps(ctx, h0, 0)
    for k, rec in enumerate(recs):
        print(f"[emax8 oracle] step {k + 1}: E0 gpu {rec['e0']:.12f} oracle {r['e0'][k]:.12f} "
              f"terms {rec['n_magnus']}/{rec['n_bch']} vs {r['n_magnus'][k]}/{r['n_bch'][k]}", flush=True)
        assert rec["n_magnus"] == r["n_magnus"][k] and rec["n_bch"] == r["n_bch"][k]
        assert abs(rec["e0"] - r["e0"][k]) <= 1e-10 * abs(r["e0"][k])
        assert abs(rec["crit"] - r["crit"][k]) <= 1e-10 * abs(r["crit"][k])
    assert_close(hs, r["HS"].vec(), what="emax8 H(s) after two steps")
    assert_close(g, r["G"].vec(), what="emax8 Omega(s) after two steps")


@pytest.mark.parametrize("prec,name", [(3, "int8x5"), (4, "int8x4"), (1, "split32"), (2, "bf16")])
def test_emax8_reduced_precision_steps(w8, oracle_steps, prec, name):
    B, ctx, ob, E = w8
    h0, r = oracle_steps
    recs, hs, g = _gpu_steps(ctx, h0, prec)
    for k, rec in enumerate(recs):
        de = abs(rec["e0"] - r["e0"][k])
        print(f"[emax8 oracle] {name} step {k + 1}: |dE0| {de:.3e} MeV", flush=True)
        assert rec["n_magnus"] == r["n_magnus"][k] and rec["n_bch"] == r["n_bch"][k]
        assert de < 1e-6
    dg = float(np.linalg.norm(g - r["G"].vec()))
    print(f"[emax8 oracle] {name} ||dOmega||_F after two steps {dg:.3e}", flush=True)
    assert dg < 1e-6 * float(np.linalg.norm(r["G"].vec()))


def test_emax8_reference_sketches(w8, oracle_steps):
    """the GPU results against the reference build's own outputs for the same
    inputs (tests/golden/emax8_k6_sketch.npz, written by make_golden.py)"""
    path = os.path.join(GOLD, "emax8_k6_sketch.npz")
    d = np.load(path, allow_pickle=False)
    B, ctx, ob, E = w8
    L, R = random_operator(B, -1, 1), random_operator(B, 1, 2)
    sketch_check(d, "input_L", L, exact=True)
    sketch_check(d, "input_R", R, exact=True)
    dL, dR = ctx.op(-1, L), ctx.op(1, R)
    masks = {"111": E.T111, "121": E.T121, "122": E.T122, "222pphh": E.T222PPHH, "221": E.T221,
             "222ph": E.T222PH}
    for name, m in masks.items():
        sketch_check(d, f"term_{name}", ctx.commutator(dL, dR, mask=m).download())
    sketch_check(d, "full", ctx.commutator(dL, dR).download())
    h0, r = oracle_steps
    sketch_check(d, "input_Hn", h0)
    recs, hs, g = _gpu_steps(ctx, h0, 0)
    rows = d["flow_rows"]
    assert len(rows) == 2
    for k, rec in enumerate(recs):
        assert abs(rec["e0"] - rows[k, 2]) <= 1e-10 * abs(rows[k, 2])
    sketch_check(d, "flow_HS", hs)
    sketch_check(d, "flow_G", g)
