"""Bases beyond the benchmark's: emax=12 (182 orbits, 156 channel blocks of up
to 2.2e3 pair states, 1.5e8 ph recoupling J' terms), where the Pandya kernel
stages more than 64 KiB of tables per workgroup and the ph apply's term
offsets exceed the row-independent 27-bit range, so both the raised LDS limit
and the row-relative descriptors are exercised.  One FP64 commutator [L, R]
against the oracle at 1e-10 relative, and [A, B] = -[B, A].
"""
import os
import time

import numpy as np
import pytest

from helpers import assert_close
from imsrg_mi355x import Basis, random_operator
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_emax12_commutator_vs_oracle():
    from imsrg_mi355x import engine as E
    B = Basis.ho(12, 6)
    t0 = time.time()
    ctx = E.Context(B)
    print(f"[large basis] emax=12 nsp={B.nsp} neq={B.neq} context {time.time() - t0:.1f} s", flush=True)
    L, R = random_operator(B, -1, 1), random_operator(B, 1, 2)
    dL, dR = ctx.op(-1, L), ctx.op(1, R)
    got = ctx.commutator(dL, dR).download()
    assert_close(got, -ctx.commutator(dR, dL).download(), tol=1e-12, what="[A,B] = -[B,A] at emax=12")
    O.lib().or_set_threads(min(16, os.cpu_count() or 1))
    ob = O.OBasis(B)
    t0 = time.time()
    want = O.commutator(ob.op(L, -1), ob.op(R, 1)).vec()
    print(f"[large basis] oracle commutator {time.time() - t0:.1f} s", flush=True)
    assert_close(got, want, what="emax=12 commutator")
    ctx.close()
