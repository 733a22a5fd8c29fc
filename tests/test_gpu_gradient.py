"""Gradient of E0(s_N) after N Magnus steps with respect to the bare
Hamiltonian (the device adjoint flow, imsrg_flow_gradient), with N the step
at which the unperturbed flow converges (|MBPT2| < 1e-6), so the loss is the
decoupled E0(s_final):

* the transposes it is built from (normal ordering, repackage) satisfy the
  adjoint identity <C, A x> = <A^T C, x>;
* central finite differences (h = 1e-5 MeV) of the GPU FP64 flow along 64
  seeded Gaussian directions (all 128 perturbed flows in one batched context)
  agree with <grad, d> to 1e-6 relative, at emax=4 with 8 checkpoints;
* the oracle (the C restatement of the reference, FP64 on the CPU) gives the
  reference gradient at emax=2 by finite differences along Gaussian
  directions, every single-particle energy and single two-body elements;
* the torch.autograd.Function returns the same gradient.
"""
import numpy as np
import pytest

from helpers import rel
from imsrg_mi355x import Basis, gaussian_direction, random_operator, synthetic_hamiltonian
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EPS = 1e-5


def frob(B, x, hx, y, hy):
    e0x, fx, Mx = B.full_from_vec(x, hx)
    e0y, fy, My = B.full_from_vec(y, hy)
    return e0x * e0y + float((fx * fy).sum()) + sum(float((a * b).sum()) for a, b in zip(Mx, My))


@pytest.fixture(scope="module")
def e4():
    from imsrg_mi355x import engine as E
    B = Basis.ho(4, 3)
    return B, E.Context(B), E


def test_normal_order_and_repackage_transposes(e4):
    B, ctx, E = e4
    h = random_operator(B, 1, 31)
    cb = random_operator(B, 1, 32)
    H = ctx.op(1, h).normal_order()
    lhs = frob(B, cb, 1, H.download(), 1)
    Hb = ctx.op(1, cb).normal_order_adjoint()
    rhs = float(Hb.download_cotangent() @ h)
    print(f"[grad] <C, NO(H)> {lhs:.12e}  <NO^T C, H> {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
    x = random_operator(B, 1, 33)
    v = random_operator(B, 1, 34)
    assert abs(float(ctx.op(1, x).download_cotangent() @ v) - frob(B, x, 1, v, 1)) <= 1e-11 * abs(frob(B, x, 1, v, 1))


def test_gradient_vs_central_differences_64_directions(e4):
    from imsrg_mi355x import autograd as AG
    from imsrg_mi355x import engine as E
    B, ctx, _ = e4
    h0 = synthetic_hamiltonian(B, seed=0)
    nsteps = ctx.magnus_flow(ctx.op(1, h0).normal_order(), "white", 0.5, 200)["steps"]
    ncheck = 8
    print(f"[grad] emax=4 flow converges after {nsteps} steps", flush=True)
    e0, g = AG.energy_and_gradient(ctx, h0, "white", 0.5, nsteps, ncheck)
    D = np.stack([gaussian_direction(B, k) for k in range(64)])
    hs = np.concatenate([h0 + EPS * D, h0 - EPS * D])
    bctx = E.Context(B, nbatch=len(hs))
    bH = bctx.op(1, hs).normal_order()
    f = bctx.flow(bH, "white", 0.5)
    for _ in range(nsteps):
        assert all(x["stepped"] for x in f.step())
    f.close()
    L = np.array(bH.e0)
    fd = (L[:64] - L[64:]) / (2 * EPS)
    an = D @ g
    err = np.abs(fd - an) / np.maximum(np.abs(an), 1e-3)
    print(f"[grad] E0(s={nsteps * 0.5}) = {e0:.10f}; max rel |FD - <g,d>| over 64 directions {err.max():.2e}; "
          f"|<g,d>| in [{np.abs(an).min():.3e}, {np.abs(an).max():.3e}]", flush=True)
    assert abs(0.5 * (L[0] + L[64]) - e0) < 1e-8  # the unperturbed energy, to O(eps^2)
    assert err.max() < 1e-6


def test_gradient_vs_oracle_reference_emax2():
    from imsrg_mi355x import autograd as AG
    from imsrg_mi355x import engine as E
    B = Basis.ho(2, 2)
    ctx = E.Context(B)
    h0 = synthetic_hamiltonian(B, seed=0)
    ob = O.OBasis(B)
    nsteps = abs(O.magnus_flow(O.normal_order(ob.op(h0, 1)), "white", 0.5, 200)["steps"])
    ncheck = 4
    print(f"[grad] emax=2 flow converges after {nsteps} steps", flush=True)
    e0, g = AG.energy_and_gradient(ctx, h0, "white", 0.5, nsteps, ncheck)

    def loss(h):
        H = O.normal_order(ob.op(h, 1))
        r = O.magnus_flow(H, "white", 0.5, max_steps=nsteps)
        assert abs(r["steps"]) == nsteps  # the perturbed flow converges at the same step
        return r["e0"][-1]

    assert abs(loss(h0) - e0) <= 1e-10 * abs(e0)
    dirs = [gaussian_direction(B, 100 + k) for k in range(16)]
    # single parameters that are Hermitian by themselves: every f_ii and 24 two-body elements
    spe = AG.spe_indices(B)
    first2b = 1 + len(B.holes) ** 2 + len(B.parts) ** 2 + len(B.holes) * len(B.parts)
    rng = np.random.default_rng(7)
    nz = np.nonzero(h0[first2b:])[0] + first2b
    for idx in list(spe) + list(rng.choice(nz, 24, replace=False)):
        d = np.zeros_like(h0)
        d[idx] = 1.0
        dirs.append(d)
    # Central differences of E0 ~ 300 MeV carry round-off of a few ulp(E0)/(2h) ~ 1e-8
    # absolute; single elements with tiny derivatives are held to that floor, the
    # Gaussian directions (|<g,d>| ~ 1-10) to 1e-6 relative.
    floor = 5e-8
    worst = 0.0
    for k, d in enumerate(dirs):
        fd = (loss(h0 + EPS * d) - loss(h0 - EPS * d)) / (2 * EPS)
        an = float(g @ d)
        print(f"[grad] dir {k}: <g,d> {an:+.10e} FD {fd:+.10e} |diff| {abs(fd - an):.2e}", flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an) + floor
        if k < 16:
            worst = max(worst, abs(fd - an) / abs(an))
    print(f"[grad] emax2 oracle reference gradient: {len(dirs)} directions; Gaussian directions max rel "
          f"diff {worst:.2e}", flush=True)
    assert worst < 1e-6


def test_torch_autograd_function(e4):
    import torch
    from imsrg_mi355x import autograd as AG
    B, ctx, _ = e4
    h0 = synthetic_hamiltonian(B, seed=1)
    h = torch.tensor(h0, dtype=torch.float64, requires_grad=True)
    L = AG.flow_e0(h, ctx, "white", 0.5, 4, 2)
    L.backward()
    e0, g = AG.energy_and_gradient(ctx, h0, "white", 0.5, 4, 2)
    assert abs(L.item() - e0) == 0.0
    assert rel(h.grad.numpy(), g) < 1e-12
