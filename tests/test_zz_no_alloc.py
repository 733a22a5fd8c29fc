"""imsrg_hartree_fock_noalloc, the entry point that allocates no device memory
(its buffers live in one static per-GPU array), against imsrg_hartree_fock
through the C ABI: two no-alloc calls in a row, the emax=8 workload
Hamiltonian (its two-body transform runs in several windows of channel
blocks) and then a smaller emax=4 one on another context, each equal bit for
bit to the allocating entry point's Hamiltonian, coefficients, E_HF and
iteration count."""
import numpy as np
import pytest

from imsrg_mi355x import Basis, synthetic_hamiltonian

pytestmark = pytest.mark.gpu


def _hf(E, ctx, bare, fn):
    n = ctx.basis.nsp
    H = ctx.op(1, bare)
    D = np.zeros((n, n))
    it = np.zeros(1, np.int32)
    e = np.zeros(1)
    rc = fn(H.h, E._dptr(D), E._iptr(it), E._dptr(e))
    assert rc == 0, E._lib.imsrg_last_error().decode()
    return H.download(), D, float(e[0]), int(it[0])


def test_noalloc_hartree_fock_equals_allocating_entry_point():
    from imsrg_mi355x import engine as E
    cases = [(8, 6), (4, 3)]  # the second call has the smaller input
    ctxs = [E.Context(Basis.ho(emax, k)) for emax, k in cases]
    bares = [synthetic_hamiltonian(c.basis, seed=0) for c in ctxs]
    got = []
    for (emax, k), ctx, bare in zip(cases, ctxs, bares):
        print(f"[no-alloc] imsrg_hartree_fock_noalloc emax={emax} nsp={ctx.basis.nsp}", flush=True)
        got.append(_hf(E, ctx, bare, E._lib.imsrg_hartree_fock_noalloc))
    for (emax, k), ctx, bare, g in zip(cases, ctxs, bares, got):
        print(f"[no-alloc] imsrg_hartree_fock emax={emax}", flush=True)
        w = _hf(E, ctx, bare, E._lib.imsrg_hartree_fock)
        assert g[3] == w[3] and g[2] == w[2], f"emax={emax}: E_HF/iterations {g[2:]} vs {w[2:]}"
        assert np.array_equal(g[1], w[1]), f"emax={emax}: coefficients differ"
        assert np.array_equal(g[0], w[0]), f"emax={emax}: HF Hamiltonian differs"
    for c in ctxs:
        c.close()
