"""The oracle's Hartree-Fock (calc_HF / sqop_to_HF restated, oracle/oracle.py)
against the reference's own outputs and against what the mathematics fixes:
the Brillouin condition, orthogonality of the coefficients, a closed form
without two-body force, and the orthogonality of the pair-basis transform."""
import os

import numpy as np
import pytest

from imsrg_mi355x import Basis, random_operator, synthetic_hamiltonian
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("emax,k", [(2, 2), (4, 3)])
def test_oracle_hf_matches_reference_fixture(emax, k):
    g = np.load(os.path.join(GOLD, f"hf_emax{emax}_k{k}.npz"))
    B = Basis.ho(emax, k)
    h, D, e, it = O.hartree_fock(B, g["H0"])
    assert it > 1
    assert abs(e - g["E_hf"][0]) < 1e-10 * abs(e)
    assert np.abs(D - g["D"]).max() < 1e-10
    assert np.abs(h - g["H_hf"]).max() < 1e-10 * np.abs(g["H_hf"]).max()
    o = O.observable_to_hf(B, g["O"], D)
    assert np.abs(o - g["O_hf"]).max() < 1e-10 * np.abs(g["O_hf"]).max()


@pytest.mark.skipif(not O.ref_available(), reason="needs the reference-built harness")
def test_oracle_hf_matches_reference_other_seed():
    B = Basis.ho(4, 3)
    H0 = synthetic_hamiltonian(B, seed=3)
    h_ref, D_ref, e_ref = O.RefRunner(B).hf(H0)
    h, D, e, _ = O.hartree_fock(B, H0)
    assert abs(e - e_ref) < 1e-10 * abs(e)
    assert np.abs(D - D_ref).max() < 1e-10
    assert np.abs(h - h_ref).max() < 1e-10 * np.abs(h_ref).max()


def test_brillouin_and_orthogonality():
    """at self-consistency the particle-hole Fock elements vanish; D is
    orthogonal and does not mix (l, j, tz) groups"""
    B = Basis.ho(4, 3)
    h, D, e, _ = O.hartree_fock(B, synthetic_hamiltonian(B, seed=1))
    e0, f, _ = B.full_from_vec(h, 1)
    assert e0 == e
    assert np.abs(f[np.ix_(B.parts, B.holes)]).max() < 1e-6
    assert np.abs(D.T @ D - np.eye(B.nsp)).max() < 1e-13
    assert np.abs(D[~B.same_ljt()]).max() == 0.0
    # the Fock matrix is diagonal in the Hartree-Fock basis within each group
    off = f * B.same_ljt() - np.diag(np.diag(f))
    assert np.abs(off).max() < 1e-6


def test_no_two_body_force_closed_form():
    """V = 0: the loop diagonalises the one-body part once; E_HF is the sum of
    the lowest (number of holes) eigenvalues of each group times 2j+1"""
    B = Basis.ho(4, 3)
    rng = np.random.default_rng(5)
    _, f, mats = B.full_from_vec(synthetic_hamiltonian(B, seed=2), 1)
    S = B.same_ljt()
    X = rng.normal(0, 1.0, f.shape) * S
    f = np.diag(np.diag(f)) + (X + X.T)
    vec = B.vec_from_full(0.0, f, [np.zeros_like(m) for m in mats])
    _, D, e, it = O.hartree_fock(B, vec)
    want = 0.0
    for g in O.sp_blocks(B):
        w = np.linalg.eigvalsh(f[np.ix_(g, g)])
        nh = int(B.con[g].sum())
        want += (B.jj[g[0]] + 1) * w[:nh].sum()
    assert it == 2  # the second diagonalisation sees the same Fock matrix
    assert abs(e - want) < 1e-12 * abs(want)


def test_pair_transform_is_orthogonal():
    """with orthogonal block-diagonal D the pair-basis transform is orthogonal
    on the physical pairs: every channel block keeps its Frobenius norm, and
    D = 1 changes nothing"""
    B = Basis.ho(4, 3)
    rng = np.random.default_rng(8)
    _, _, mats = B.full_from_vec(random_operator(B, 1, 9), 1)
    ident = O.transform_2b(B, mats, np.eye(B.nsp))
    assert all(np.array_equal(a, b) for a, b in zip(ident, mats))
    D = np.zeros((B.nsp, B.nsp))
    for g in O.sp_blocks(B):
        q, _ = np.linalg.qr(rng.normal(size=(len(g), len(g))))
        D[np.ix_(g, g)] = q
    out = O.transform_2b(B, mats, D)
    for blk, a, b in zip(B.blocks, mats, out):
        ok = B.pauli_mask(blk)
        assert abs(np.linalg.norm(b) - np.linalg.norm(a)) < 1e-12 * max(1.0, np.linalg.norm(a))
        assert np.abs(b[~ok]).max(initial=0.0) < 1e-13
