"""The precision gate over the whole emax=8 flow (the benchmark's workload:
90 orbits, 6 holes per isospin, synthetic Gaussian H with seed 0 and hw = 20,
White generator, ds = 0.5): each int8 variant runs in lockstep with the FP64
flow from s = 0 until both have converged (IMSRG_MAGNUS.f90:98-126, the
reference's MBPT2 stop test), and |E0 - E0_FP64| stays below 1e-6 MeV at
every step with the FP64 flow's Magnus and BCH term counts.  The FP64 flow is
the one tests/test_gpu_emax8_oracle.py pins against the oracle."""
import numpy as np
import pytest

from imsrg_mi355x import Basis, synthetic_hamiltonian

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,name", [(4, "int8x4"), (3, "int8x5")])
def test_emax8_int8_gate_over_the_whole_flow(prec, name):
    from imsrg_mi355x import engine as E
    B = Basis.ho(8, 6)
    ctx = E.Context(B)
    H = ctx.op(1, synthetic_hamiltonian(B, seed=0, hw=20.0, sigma1=0.5, sigma2=1.0))
    H.normal_order()
    h0 = H.download()
    ha, hb = ctx.op(1, h0), ctx.op(1, h0)
    ctx.set_precision(0)
    fa = ctx.flow(ha, "white", 0.5)
    ctx.set_precision(prec)
    fb = ctx.flow(hb, "white", 0.5)
    worst, steps = 0.0, 0
    ra = rb = None
    for _ in range(200):
        if fa.converged and fb.converged:
            break
        if not fa.converged:
            ctx.set_precision(0)
            ra = fa.step()
        if not fb.converged:
            ctx.set_precision(prec)
            rb = fb.step()
        steps += 1
        assert ra["n_magnus"] == rb["n_magnus"] and ra["n_bch"] == rb["n_bch"]
        worst = max(worst, abs(ra["e0"] - rb["e0"]))
    both = fa.converged and fb.converged
    fa.close()
    fb.close()
    ctx.set_precision(0)
    print(f"[emax8 gate] {name}: {steps} lockstep steps, max |dE0| {worst:.3e} MeV", flush=True)
    assert both
    assert worst < 1e-6
