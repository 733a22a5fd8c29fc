"""The Fortran drop-in module (eom-imsrg_full_amd/fortran/imsrg_gpu.f90)
against the reference's own routines on the same sq_op inputs: a program built
from the reference objects (oracle/_ref/drop_in_test) calls, e.g.,
commutator_111 and gpu_commutator_111, BCH_EXPAND and gpu_BCH_EXPAND,
transform_sq_op_BCH and gpu_transform_sq_op_BCH, dHds and gpu_dHds,
magnus_decouple and gpu_magnus_decouple (also resumed from the reference's own
Omega checkpoint files), and reports relative differences."""
import os
import subprocess

import numpy as np
import pytest

from imsrg_mi355x import Basis, random_operator, synthetic_hamiltonian
from oracle import oracle as O

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(O.__file__), "_ref", "drop_in_test")


@pytest.mark.skipif(not os.path.exists(BIN), reason="drop_in_test is built only where the reference is mounted")
@pytest.mark.parametrize("emax,k", [(2, 2), (4, 3)])
def test_fortran_drop_in(tmp_path, emax, k):
    B = Basis.ho(emax, k)
    B.write_basis_file(tmp_path / "basis.txt")
    ob = O.OBasis(B)
    random_operator(B, -1, 1).tofile(tmp_path / "L.bin")
    random_operator(B, 1, 2).tofile(tmp_path / "R.bin")
    O.normal_order(ob.op(synthetic_hamiltonian(B, 0), 1)).vec().tofile(tmp_path / "H.bin")
    synthetic_hamiltonian(B, 0).tofile(tmp_path / "H0.bin")
    print(f"[drop-in] emax={emax}", flush=True)
    r = subprocess.run([BIN, "basis.txt", "L.bin", "R.bin", "H.bin", "H0.bin"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=600)
    checks = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("CHECK")]
    print("\n".join(" ".join(c) for c in checks))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    names = {c[1] for c in checks}
    assert {"full_commutator", "per_routine_sequence", "cross_coupled", "build_gs_w2", "BCH_EXPAND",
            "MAGNUS_EXPAND", "magnus_decouple_H", "magnus_decouple_Omega", "transform_sq_op_BCH",
            "dHds", "checkpoint_restart_H", "checkpoint_restart_Omega", "calc_HF", "calc_HF_coefficients",
            "sqop_to_HF"} <= names
    assert all(float(c[2]) <= 1e-10 for c in checks)
