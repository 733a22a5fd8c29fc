"""imsrg_mi355x -- MI355X-native J-coupled IMSRG(2) Magnus commutator engine.

Host-side Python view of the C-ABI library ``libimsrg_mi355x.so``:

* ``basis`` / ``synthetic``: basis layout and seeded synthetic operators in
  the reference's vectorize order (pure numpy, no device needed);
* ``engine``: ctypes binding of the HIP library (loads the in-tree .so and
  raises if it is missing -- there is no CPU fallback).
"""
from .basis import Basis, hk_orbits, closed_shell_con, gemm_flops_222, pandya_blocks  # noqa: F401
from .synthetic import synthetic_hamiltonian, random_operator, config_basis, batch_hamiltonians, batch_parameters, gaussian_direction  # noqa: F401
