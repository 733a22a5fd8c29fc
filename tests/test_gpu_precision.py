"""Split-precision GEMM variants of the 222 products against the FP64 path.
Everything outside the two GEMMs stays FP64, so the only error sources are
the FP32 accumulation inside a K slab (split32, bf16: about 2^-24 relative)
and, for the int8 slices, the dropped slice products (about 2^-36 of the row
maxima with five slices, 2^-29 with four)."""
import numpy as np
import pytest

from helpers import rel
from imsrg_mi355x import Basis, random_operator, synthetic_hamiltonian

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def e4():
    from imsrg_mi355x import engine as E
    B = Basis.ho(4, 3)
    return B, E.Context(B), E


@pytest.mark.parametrize("prec", [1, 2, 3, 4], ids=["split32", "bf16", "int8x5", "int8x4"])
def test_split_commutator_close_to_fp64(e4, prec):
    B, ctx, E = e4
    L, R = ctx.op(-1, random_operator(B, -1, 1)), ctx.op(1, random_operator(B, 1, 2))
    ctx.set_precision(0)
    ref = ctx.commutator(L, R).download()
    for mask in (E.T222PPHH, E.T222PH, E.T221, E.TALL):
        ctx.set_precision(0)
        want = ctx.commutator(L, R, mask=mask).download()
        ctx.set_precision(prec)
        got = ctx.commutator(L, R, mask=mask).download()
        err = rel(got, want)
        print(f"[precision {prec}] mask {mask}: rel err {err:.2e}", flush=True)
        assert err < (1e-7 if prec == 3 else 2e-6)  # int8x4 (prec 4): 2^-29 per product
        assert err > 0 or mask == E.T221  # the split path really ran
    ctx.set_precision(0)
    assert np.array_equal(ctx.commutator(L, R).download(), ref)  # FP64 path restored


@pytest.mark.parametrize("prec", [1, 2, 3, 4], ids=["split32", "bf16", "int8x5", "int8x4"])
def test_split_flow_energy(e4, prec):
    B, ctx, E = e4
    H0 = ctx.op(1, synthetic_hamiltonian(B, 0))
    H0.normal_order()
    v = H0.download()
    ctx.set_precision(0)
    a = ctx.magnus_flow(ctx.op(1, v), "white", 0.5, 200)
    ctx.set_precision(prec)
    b = ctx.magnus_flow(ctx.op(1, v), "white", 0.5, 200)
    ctx.set_precision(0)
    n = min(a["steps"], b["steps"])
    dE = np.abs(a["e0"][:n] - b["e0"][:n]).max()
    print(f"[precision {prec}] emax=4 flow: {a['steps']} vs {b['steps']} steps, max |dE0| {dE:.2e} MeV", flush=True)
    assert dE < 1e-6


@pytest.mark.parametrize("prec,name,tol", [(3, "int8x5", 1e-7), (4, "int8x4", 2e-6)])
def test_int8_slices_every_hermiticity_and_batched(prec, name, tol):
    """int8 slices at emax=6 for every hermiticity pair (the B rows of the
    pp/hh product carry the operand's hermiticity sign) against FP64, and a
    batched context gives each instance the single-instance result bit for bit"""
    from imsrg_mi355x import engine as E
    B = Basis.ho(6, 3)
    ctx = E.Context(B)
    errs = {}
    for hl, hr in ((-1, 1), (1, 1), (-1, -1)):
        L, R = ctx.op(hl, random_operator(B, hl, 91)), ctx.op(hr, random_operator(B, hr, 92))
        ctx.set_precision(0)
        want = ctx.commutator(L, R).download()
        ctx.set_precision(prec)
        got = ctx.commutator(L, R).download()
        err = rel(got, want)
        errs[(hl, hr)] = err
        print(f"[{name}] emax=6 herm ({hl},{hr}): rel err {err:.2e}", flush=True)
        assert 0 < err < tol
    if prec == 4:  # four slices drop the 2^-35 level: larger than five slices' error
        ctx.set_precision(0)
        want = ctx.commutator(L, R).download()
        ctx.set_precision(3)
        assert rel(ctx.commutator(L, R).download(), want) < errs[(-1, -1)]
        ctx.set_precision(prec)
    nb = 3
    Ls = np.stack([random_operator(B, -1, 100 + k) for k in range(nb)])
    Rs = np.stack([random_operator(B, 1, 200 + k) for k in range(nb)])
    bctx = E.Context(B, nbatch=nb, precision=prec)
    got = bctx.commutator(bctx.op(-1, Ls), bctx.op(1, Rs)).download()
    for k in range(nb):
        one = ctx.commutator(ctx.op(-1, Ls[k]), ctx.op(1, Rs[k])).download()
        assert np.array_equal(got[k], one)
