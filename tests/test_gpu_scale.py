"""GPU checks at the larger workloads: emax=6 against the oracle, emax=8 through
size-independent properties, and the GPU flow against the reference's golden
flows directly."""
import os

import numpy as np
import pytest

from helpers import assert_close, rel
from imsrg_mi355x import Basis, random_operator, synthetic_hamiltonian
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_emax6_commutator_vs_oracle():
    from imsrg_mi355x import engine as E
    B = Basis.ho(6, 3)
    print(f"[gpu scale] emax=6 neq={B.neq}", flush=True)
    ctx = E.Context(B)
    O.lib().or_set_threads(min(16, os.cpu_count() or 1))
    ob = O.OBasis(B)
    L, R = random_operator(B, -1, 1), random_operator(B, 1, 2)
    got = ctx.commutator(ctx.op(-1, L), ctx.op(1, R)).download()
    want = O.commutator(ob.op(L, -1), ob.op(R, 1)).vec()
    assert_close(got, want, what="emax6 commutator")


@pytest.fixture(scope="module")
def e8():
    from imsrg_mi355x import engine as E
    B = Basis.ho(8, 6)
    print(f"[gpu scale] emax=8 neq={B.neq}", flush=True)
    return B, E.Context(B), E


def test_emax8_antisymmetry(e8):
    B, ctx, E = e8
    A, H = ctx.op(-1, random_operator(B, -1, 1)), ctx.op(1, random_operator(B, 1, 2))
    ab = ctx.commutator(A, H).download()
    ba = ctx.commutator(H, A).download()
    assert_close(ab, -ba, tol=1e-12, what="[A,B] = -[B,A] at emax=8")


def test_emax8_one_body_closed_form_and_rotation(e8):
    from scipy.linalg import expm
    B, ctx, E = e8
    zero = [np.zeros((b.dim, b.dim)) for b in B.blocks]
    _, fa, _ = B.full_from_vec(random_operator(B, -1, 3), -1)
    _, fb, _ = B.full_from_vec(random_operator(B, 1, 4), 1)
    va, vb = B.vec_from_full(0.0, fa, zero), B.vec_from_full(0.0, fb, zero)
    C = ctx.commutator(ctx.op(-1, va), ctx.op(1, vb)).download()
    e0, f, mats = B.full_from_vec(C, 1)
    assert_close(f, fa @ fb - fb @ fa, tol=1e-12, what="111 closed form")
    assert max(np.abs(m).max() for m in mats if m.size) == 0.0
    g = 0.02 * fa
    hs, n = ctx.bch(ctx.op(-1, B.vec_from_full(0.0, g, zero)), ctx.op(1, vb))
    _, fs, _ = B.full_from_vec(hs.download(), 1)
    U = expm(g)
    assert_close(fs, U @ fb @ U.T, tol=1e-8, what="one-body BCH rotation")


def test_emax8_bch_identity_and_norms(e8):
    B, ctx, E = e8
    H = synthetic_hamiltonian(B, 0)
    dH = ctx.op(1, H)
    hs, n = ctx.bch(ctx.op(-1, np.zeros(B.neq)), dH)
    assert n == 1
    assert np.array_equal(hs.download(), H)
    e0, f, mats = B.full_from_vec(H, 1)
    want = np.sqrt((f ** 2).sum() + sum((m ** 2).sum() for m in mats))
    assert abs(dH.norm() - want) <= 1e-12 * want


@pytest.mark.parametrize("name,emax,k,kinds", [("emax2_k2", 2, 2, ("white", "w2")), ("emax4_k3", 4, 3, ("white",))])
def test_gpu_flow_vs_reference_golden(name, emax, k, kinds):
    from imsrg_mi355x import engine as E
    d = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    B = Basis.ho(emax, k)
    ctx = E.Context(B)
    for kind in kinds:
        H = ctx.op(1, d["H0"])
        H.normal_order()
        rows = d[f"flow_{kind}_rows"]
        r = ctx.magnus_flow(H, kind, 0.5, 200)
        assert r["steps"] == len(rows)
        assert rel(r["e0"], rows[:, 2]) <= 1e-10
        assert_close(r["HS"].download(), d[f"flow_{kind}_HS"], what=f"{name} {kind} final H(s)")
