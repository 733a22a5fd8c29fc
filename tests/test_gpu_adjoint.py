"""Adjoint commutator: for random P, Y and a cotangent C of the output's
hermiticity, <C, [P, Y]> = <Ybar, Y> with Ybar = commutator_adjoint(P, C),
where <,> is the Frobenius product over E0, the full one-body matrix and the
full channel matrices.  This pins every transposed kernel against its forward
kernel, term by term and for every hermiticity combination; the X-side
gradient follows from antisymmetry, <C, [X, Y]> = <-adjoint(Y, C), X>."""
import numpy as np
import pytest

from imsrg_mi355x import Basis, random_operator

pytestmark = pytest.mark.gpu


def frob(B, x, hx, y, hy):
    e0x, fx, Mx = B.full_from_vec(x, hx)
    e0y, fy, My = B.full_from_vec(y, hy)
    return e0x * e0y + float((fx * fy).sum()) + sum(float((a * b).sum()) for a, b in zip(Mx, My))


@pytest.fixture(scope="module", params=[(2, 2), (4, 3)], ids=["emax2", "emax4"])
def env(request):
    from imsrg_mi355x import engine as E
    emax, k = request.param
    B = Basis.ho(emax, k)
    return B, E.Context(B), E


TERMS = ["T110", "T111", "T121", "T122", "T220", "T221", "T222PPHH", "T222PH", "TALL"]


@pytest.mark.parametrize("hP,hY", [(-1, 1), (1, 1), (-1, -1), (1, -1)])
def test_adjoint_identity_each_term(env, hP, hY):
    B, ctx, E = env
    hC = -hP * hY
    P = random_operator(B, hP, 11)
    Y = random_operator(B, hY, 12)
    Cb = random_operator(B, hC, 13)
    Cb[0] = 1.3  # E0 cotangent, so that the 110 and 220 terms are exercised
    dP, dY, dC = ctx.op(hP, P), ctx.op(hY, Y), ctx.op(hC, Cb)
    for name in TERMS:
        mask = getattr(E, name)
        out = ctx.op(hC)
        out.e0 = 0.0
        ctx.commutator(dP, dY, out, mask=mask)
        lhs = frob(B, Cb, hC, out.download(), hC)
        yb = ctx.commutator_adjoint(dP, dC, hY, mask=mask).download()
        rhs = frob(B, yb, hY, Y, hY)
        xb = ctx.commutator_adjoint(dY, dC, hP, mask=mask).download()
        rhs_x = -frob(B, xb, hP, P, hP)
        scale = max(abs(lhs), 1e-300)
        print(f"[adjoint] herm ({hP},{hY}) {name}: <C,[P,Y]> {lhs:.12e} <Ybar,Y> {rhs:.12e} "
              f"-<Xbar(Y),P> {rhs_x:.12e}", flush=True)
        assert abs(lhs - rhs) <= 1e-11 * scale + 1e-12
        assert abs(lhs - rhs_x) <= 1e-11 * scale + 1e-12
