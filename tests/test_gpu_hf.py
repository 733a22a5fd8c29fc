"""Hartree-Fock on the MI355X (imsrg_hartree_fock, imsrg_observable_to_hf)
against the reference's outputs (golden fixtures written by the reference's
calc_HF and sqop_to_HF) and against the oracle at a size beyond the fixtures.
Tolerance 1e-10 relative (max-abs over the vectorize stream); the iteration
count must equal the oracle's."""
import os

import numpy as np
import pytest

from imsrg_mi355x import Basis, random_operator, synthetic_hamiltonian
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("emax,k", [(2, 2), (4, 3)])
def test_hf_matches_reference_fixture(emax, k):
    from imsrg_mi355x import engine as E
    g = np.load(os.path.join(GOLD, f"hf_emax{emax}_k{k}.npz"))
    B = Basis.ho(emax, k)
    print(f"[gpu hf] emax={emax} k={k}", flush=True)
    ctx = E.Context(B)
    H = ctx.op(1, g["H0"])
    D, e, it = H.hartree_fock()
    _, _, _, it_o = O.hartree_fock(B, g["H0"])
    assert it == it_o
    assert abs(e - g["E_hf"][0]) < 1e-10 * abs(e)
    assert np.abs(D - g["D"]).max() < 1e-10
    assert _rel(H.download(), g["H_hf"]) < 1e-10
    Ob = ctx.op(1, g["O"]).to_hf(D)
    assert _rel(Ob.download(), g["O_hf"]) < 1e-10
    # the Hartree-Fock Hamiltonian drives a Magnus step like any other
    rec = ctx.flow(H, "white", 0.5).step()
    ref = O.magnus_flow(O.OOp(O.OBasis(B), g["H_hf"], 1), "white", 0.5, 2)
    k = int(np.argmin(np.abs(ref["s"] - 0.5)))
    assert ref["s"][k] == 0.5 and abs(rec["e0"] - ref["e0"][k]) < 1e-9 * abs(ref["e0"][k])
    ctx.close()


@pytest.mark.parametrize("emax,k", [(6, 3), (8, 6)])
def test_hf_matches_oracle(emax, k):
    """emax=6 and the emax=8 workload basis: the device loop against the oracle"""
    from imsrg_mi355x import engine as E
    B = Basis.ho(emax, k)
    print(f"[gpu hf] emax={emax} k={k}", flush=True)
    H0 = synthetic_hamiltonian(B, seed=4)
    h_o, D_o, e_o, it_o = O.hartree_fock(B, H0)
    ctx = E.Context(B)
    H = ctx.op(1, H0)
    D, e, it = H.hartree_fock()
    assert it == it_o
    assert abs(e - e_o) < 1e-10 * abs(e_o)
    assert np.abs(D - D_o).max() < 1e-10
    assert _rel(H.download(), h_o) < 1e-10
    Ob = random_operator(B, 1, 21)
    assert _rel(ctx.op(1, Ob).to_hf(D).download(), O.observable_to_hf(B, Ob, D_o)) < 1e-10
    ctx.close()


def test_hf_rejects_batched_and_antihermitian():
    from imsrg_mi355x import engine as E
    B = Basis.ho(2, 2)
    ctx = E.Context(B)
    with pytest.raises(E.ImsrgError):
        ctx.op(-1, random_operator(B, -1, 1)).hartree_fock()
    ctx.close()
    ctxb = E.Context(B, nbatch=2)
    H0 = synthetic_hamiltonian(B, seed=0)
    with pytest.raises(E.ImsrgError):
        ctxb.op(1, np.concatenate([H0, H0])).hartree_fock()
    ctxb.close()
