"""Seeded synthetic operators in the reference's vectorize order.

The workload definitions (harmonic-oscillator basis, closed-shell reference,
Gaussian one- and two-body matrix elements) are the synthetic Hamiltonians the
Magnus-flow benchmarks are quoted on.  Everything is generated with numpy's
PCG64 from a fixed seed and serialised as the operator stream itself, so no
result depends on a random-number generator being portable.

* one-body: f_ii = hw (2n + l + 3/2); symmetric Gaussian off-diagonal
  (sigma1) only between orbits of equal (l, 2j, 2tz);
* two-body: normalised J-coupled antisymmetrised matrix elements, Gaussian
  (sigma2), symmetric per (J,P,Tz) channel; identical-orbit pairs with odd J
  are Pauli-forbidden and zero (brute_force_testing.f90:102-109).

``random_operator`` builds test operators of either hermiticity the way
construct_random_rank0 does (brute_force_testing.f90:15-120): uniform in
[-5, 5], one-body only within (l, j, tz), Pauli zeros.
"""
from __future__ import annotations

import numpy as np

from .basis import Basis


def _sym_gauss(rng, n, sigma):
    A = rng.normal(0.0, sigma, (n, n))
    U = np.triu(A)
    return U + np.triu(A, 1).T


def synthetic_hamiltonian(basis: Basis, seed: int = 0, hw: float = 20.0,
                          sigma1: float = 0.5, sigma2: float = 1.0):
    """Bare (not yet normal-ordered) Hermitian H as a vectorize stream."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N = basis.nsp
    f = np.zeros((N, N))
    same = basis.same_ljt()
    off = _sym_gauss(rng, N, sigma1)
    f[same] = off[same]
    np.fill_diagonal(f, hw * (2 * basis.n + basis.l + 1.5))
    mats = []
    for blk in basis.blocks:
        M = _sym_gauss(rng, blk.dim, sigma2)
        ok = basis.pauli_mask(blk)
        M[~ok, :] = 0.0
        M[:, ~ok] = 0.0
        mats.append(M)
    return basis.vec_from_full(0.0, f, mats)


def batch_parameters(seed: int):
    """(hw, sigma2) of batch member `seed`: hw uniform in [16, 24] MeV and the
    two-body width uniform in [0.5, 1.5] MeV, drawn from PCG64(1000 + seed)."""
    rng = np.random.Generator(np.random.PCG64(1000 + seed))
    return float(rng.uniform(16.0, 24.0)), float(rng.uniform(0.5, 1.5))


def batch_hamiltonians(basis: Basis, seeds):
    """Bare Hamiltonians of a batch of independent instances, shape
    (len(seeds), neq): member k has seed seeds[k] and batch_parameters(seed)."""
    out = []
    for sd in seeds:
        hw, s2 = batch_parameters(sd)
        out.append(synthetic_hamiltonian(basis, seed=sd, hw=hw, sigma1=0.5, sigma2=s2))
    return np.stack(out)


def gaussian_direction(basis: Basis, seed: int, sigma: float = 1.0):
    """A random Hermitian perturbation as a vectorize stream: symmetric
    Gaussian one-body part within (l, j, tz) (diagonal included, i.e. the
    single-particle energies), symmetric Gaussian two-body blocks with the
    Pauli zeros, E0 = 0 (normal ordering overwrites it)."""
    rng = np.random.Generator(np.random.PCG64(5000 + seed))
    N = basis.nsp
    f = np.zeros((N, N))
    same = basis.same_ljt()
    off = _sym_gauss(rng, N, sigma)
    f[same] = off[same]
    mats = []
    for blk in basis.blocks:
        M = _sym_gauss(rng, blk.dim, sigma)
        ok = basis.pauli_mask(blk)
        M[~ok, :] = 0.0
        M[:, ~ok] = 0.0
        mats.append(M)
    return basis.vec_from_full(0.0, f, mats)


def random_operator(basis: Basis, herm: int, seed: int):
    """Random operator with hermiticity ``herm`` (+1/-1) as a vectorize stream."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N = basis.nsp
    same = basis.same_ljt()
    U = rng.uniform(-5.0, 5.0, (N, N))
    f = np.triu(U, 1) + herm * np.triu(U, 1).T
    if herm == 1:
        f += np.diag(np.diag(U))
    f[~same] = 0.0
    mats = []
    for blk in basis.blocks:
        n = blk.dim
        U = rng.uniform(-5.0, 5.0, (n, n))
        M = np.triu(U, 1) + herm * np.triu(U, 1).T
        if herm == 1:
            M += np.diag(np.diag(U))
        ok = basis.pauli_mask(blk)
        M[~ok, :] = 0.0
        M[:, ~ok] = 0.0
        mats.append(M)
    return basis.vec_from_full(0.0, f, mats)


def config_basis(name: str) -> Basis:
    """Basis of a named workload: emax2 (k=2), emax4 (k=3), emax6 (k=3), emax8 (k=6)."""
    table = {"emax2": (2, 2), "emax4": (4, 3), "emax6": (6, 3), "emax8": (8, 6)}
    emax, k = table[name]
    return Basis.ho(emax, k)
