"""HIP engine vs the oracle (the C restatement of the reference), FP64.

Every check runs the MI355X path through the C ABI and the oracle on the same
seeded inputs and requires agreement to 1e-10 relative (max-abs over the
operator stream).  Sizes are the emax=2 (k=2) and emax=4 (k=3) workloads,
which the oracle finishes in seconds.
"""
import numpy as np
import pytest

from helpers import assert_close, rel
from imsrg_mi355x import Basis, random_operator, synthetic_hamiltonian
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [(2, 2), (4, 3)]
HERMS = [(-1, 1), (-1, -1), (1, 1), (1, -1)]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"emax{c[0]}")
def env(request):
    from imsrg_mi355x import engine as E
    emax, k = request.param
    B = Basis.ho(emax, k)
    ctx = E.Context(B)
    ob = O.OBasis(B)
    print(f"[gpu parity] emax={emax} k={k} neq={B.neq}", flush=True)
    yield B, ctx, ob, E
    ctx.close()


def test_upload_download_roundtrip(env):
    B, ctx, ob, E = env
    for herm in (1, -1):
        v = random_operator(B, herm, 11)
        back = ctx.op(herm, v).download()
        # the oracle's repackage/vectorize pair is the reference behaviour
        ref = ob.op(v, herm).vec()
        assert np.array_equal(back, ref)


@pytest.mark.parametrize("hL,hR", HERMS)
def test_each_term_and_full(env, hL, hR):
    B, ctx, ob, E = env
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
        assert_close(got, want[name], what=f"term {name} herm {hL},{hR}")
    full = ctx.commutator(dL, dR).download()
    assert_close(full, O.commutator(oL, oR).vec(), what=f"full commutator herm {hL},{hR}")


def test_cross_coupled(env):
    B, ctx, ob, E = env
    for herm in (1, -1):
        A = random_operator(B, herm, 5)
        assert_close(ctx.cross_coupled(ctx.op(herm, A)), O.cross_coupled(ob.op(A, herm)),
                     what="Pandya")


def test_scalars_normal_order(env):
    B, ctx, ob, E = env
    H = synthetic_hamiltonian(B, seed=3)
    dH = ctx.op(1, H)
    oH = ob.op(H, 1)
    assert abs(dH.norm() - O.frob_norm(oH)) <= 1e-12 * O.frob_norm(oH)
    dH.normal_order()
    O.normal_order(oH)
    assert_close(dH.download(), oH.vec(), what="normal_order")
    assert abs(dH.mbpt2() - O.mbpt2(oH)) <= 1e-10 * abs(O.mbpt2(oH))


@pytest.mark.parametrize("kind", ["white", "atan", "w2"])
def test_generators(env, kind):
    B, ctx, ob, E = env
    oH = O.normal_order(ob.op(synthetic_hamiltonian(B, seed=0), 1))
    Hn = oH.vec()
    got = ctx.generator(kind, ctx.op(1, Hn)).download()
    assert_close(got, O.generator(kind, oH).vec(), what=f"generator {kind}")


def test_bch_and_magnus_expand(env):
    B, ctx, ob, E = env
    oH = O.normal_order(ob.op(synthetic_hamiltonian(B, seed=0), 1))
    Hn = oH.vec()
    eta = O.generator("white", oH).vec()
    G = 0.5 * eta
    want_hs, want_n = O.bch(ob.op(G, -1), ob.op(Hn, 1))
    got_hs, got_n = ctx.bch(ctx.op(-1, G), ctx.op(1, Hn))
    assert got_n == want_n
    assert_close(got_hs.download(), want_hs.vec(), what="BCH")
    want_dg, want_m = O.magnus_expand(ob.op(eta, -1), ob.op(G, -1))
    got_dg, got_m = ctx.magnus_expand(ctx.op(-1, eta), ctx.op(-1, G))
    assert got_m == want_m
    assert_close(got_dg.download(), want_dg.vec(), what="MAGNUS_EXPAND")


def test_transform_bch_and_dhds(env):
    """transform_sq_op_BCH (rank 0) of a random hermitian observable, in place,
    against the oracle's BCH_EXPAND of it; dHds = [eta(H), H] against the
    oracle's generator and commutator"""
    B, ctx, ob, E = env
    oH = O.normal_order(ob.op(synthetic_hamiltonian(B, seed=0), 1))
    Hn = oH.vec()
    eta = O.generator("white", oH)
    G = 0.5 * eta.vec()
    obs = random_operator(B, 1, 17)
    want, want_n = O.bch(ob.op(G, -1), ob.op(obs, 1))
    Op = ctx.op(1, obs)
    n = ctx.transform_bch(Op, ctx.op(-1, G))
    assert n == want_n
    assert_close(Op.download(), want.vec(), what="transform_sq_op_BCH")
    got = ctx.dhds(ctx.op(1, Hn), "white").download()
    assert_close(got, O.commutator(eta, oH).vec(), what="dHds")


def test_magnus_flow(env):
    B, ctx, ob, E = env
    oH = O.normal_order(ob.op(synthetic_hamiltonian(B, seed=0), 1))
    Hn = oH.vec()
    want = O.magnus_flow(ob.op(Hn, 1), "white", 0.5, 100)
    got = ctx.magnus_flow(ctx.op(1, Hn), "white", 0.5, 100)
    print(f"[gpu parity] flow steps gpu={got['steps']} oracle={want['steps']}", flush=True)
    assert got["steps"] == want["steps"]
    assert np.array_equal(got["n_magnus"], want["n_magnus"])
    assert np.array_equal(got["n_bch"], want["n_bch"])
    assert rel(got["e0"], want["e0"]) <= 1e-10
    assert_close(got["HS"].download(), want["HS"].vec(), what="final H(s)")
    assert_close(got["G"].download(), want["G"].vec(), what="final Omega(s)")


@pytest.mark.parametrize("shapes", [(0, 0), (1, 1), (0, 1), (1, 0)], ids=["64/64", "128x64/128x64", "64/128x64", "128x64/64"])
def test_fp64_tile_shapes(env, shapes):
    """both FP64 GEMM tile shapes the autotuner chooses between give the oracle's commutator"""
    B, ctx, ob, E = env
    L = random_operator(B, -1, 3)
    R = random_operator(B, 1, 4)
    want = O.commutator(ob.op(L, -1), ob.op(R, 1)).vec()
    ctx.set_tile_shapes(*shapes)
    got = ctx.commutator(ctx.op(-1, L), ctx.op(1, R)).download()
    ctx.set_tile_shapes(-1, -1)
    assert_close(got, want, what=f"tile shapes {shapes}")
