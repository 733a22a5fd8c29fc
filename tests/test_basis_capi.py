"""Host-side logic: basis layout, vectorize-stream conversions, synthetic
workloads, and the C ABI library (loads, exports every declared symbol; no
compute calls -- there is no GPU here)."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from imsrg_mi355x import Basis, gemm_flops_222, hk_orbits, pandya_blocks, random_operator, synthetic_hamiltonian
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_hk_orbits_and_sizes():
    orb = hk_orbits(2)
    assert orb[:6] == [(0, 0, 1, -1), (0, 0, 1, 1), (0, 1, 1, -1), (0, 1, 1, 1), (0, 1, 3, -1), (0, 1, 3, 1)]
    for emax, n in ((2, 12), (4, 30), (6, 56), (8, 90)):
        assert len(hk_orbits(emax)) == n


@pytest.mark.parametrize("emax,k,nblocks", [(2, 2, 36), (4, 3, 60), (6, 3, 84)])
def test_layout_matches_oracle(emax, k, nblocks):
    B = Basis.ho(emax, k)
    ob = O.OBasis(B)
    assert len(B.blocks) == nblocks
    assert B.neq == ob.neq
    info = np.zeros(6, np.int32)
    for q, blk in enumerate(B.blocks):
        O.lib().or_basis_block_info.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        O.lib().or_basis_block_info(ob.h, q, info.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        assert tuple(info) == (blk.J, blk.P, blk.Tz, blk.npp, blk.nph, blk.nhh)
    assert [x for x in pandya_blocks(B)] == ob.cc_blocks()


@pytest.mark.parametrize("herm", [1, -1])
def test_vectorize_roundtrip_matches_repackage(herm):
    B = Basis.ho(4, 3)
    ob = O.OBasis(B)
    v = random_operator(B, herm, 3)
    assert np.array_equal(ob.op(v, herm).vec(), v)
    e0, f, mats = B.full_from_vec(v, herm)
    assert np.array_equal(B.vec_from_full(e0, f, mats), v)
    # one-body: f_ph stored, f_hp = herm * f_ph^T
    assert np.array_equal(f[np.ix_(B.holes, B.parts)], herm * f[np.ix_(B.parts, B.holes)].T)


def test_synthetic_hamiltonian_properties():
    B = Basis.ho(4, 3)
    v = synthetic_hamiltonian(B, seed=0)
    e0, f, mats = B.full_from_vec(v, 1)
    assert np.allclose(f, f.T)
    assert np.allclose(np.diag(f), 20.0 * (2 * B.n + B.l + 1.5))
    assert np.all(f[~B.same_ljt()] == 0)
    for blk, M in zip(B.blocks, mats):
        assert np.allclose(M, M.T)
        ok = B.pauli_mask(blk)
        assert np.all(M[~ok] == 0) and np.all(M[:, ~ok] == 0)


def test_flop_counts_emax8():
    # algorithmic 222 GEMM work per commutator at emax=8 (90 orbits, 6 holes/iso)
    pphh, ph = gemm_flops_222(Basis.ho(8, 6))
    assert abs(pphh - 5.369e9) / 5.369e9 < 1e-3
    assert abs(ph - 6.486e9) / 6.486e9 < 1e-3


def test_library_loads_and_exports_header_symbols():
    from imsrg_mi355x import engine as E
    syms = E.header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(E._lib, s), s


def test_library_missing_fails_loudly(tmp_path):
    # the product path has no fallback: without the .so the engine import raises
    pkg = os.path.join(ROOT, "eom-imsrg_full_amd")
    code = ("import sys, os; sys.path.insert(0, %r); import imsrg_mi355x.engine as E" % pkg)
    env = dict(os.environ)
    lib = os.path.join(pkg, "imsrg_mi355x", "lib", "libimsrg_mi355x.so")
    env["IMSRG_MI355X_LIB"] = str(tmp_path / "missing.so")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True)
    assert os.path.exists(lib)
    assert r.returncode != 0 and "no CPU fallback" in r.stderr


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
