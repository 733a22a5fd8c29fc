#!/usr/bin/env python3
"""Magnus-flow throughput of the MI355X engine (steps/s at emax=8).

A "step" is one pass of magnus_decouple's loop (IMSRG_MAGNUS.f90:98-126):
MAGNUS_EXPAND + Euler step + BCH_EXPAND + generator + MBPT2, on the synthetic
emax=8 workload (90 HO orbits, 40Ca-like closed shell: lowest 6 orbits per
isospin filled, hw = 20 MeV, Gaussian one-/two-body matrix elements, seed 0,
normal-ordered on the device).  W untimed warmup steps advance the flow from
s = 0; then exactly K steps of the same flow are timed between two stream
synchronizations.  Inputs are resident in HBM when the timer starts.

Prints ONE JSON line on stdout (progress goes to stderr).
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "eom-imsrg_full_amd"))

PEAK = {"fp64_mfma_tflops": 78.6, "fp32_mfma_tflops": 157.3, "hbm_gbs": 8000.0}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=4)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--emax", type=int, default=8)
    p.add_argument("--holes", type=int, default=6, help="filled orbits per isospin")
    p.add_argument("--generator", default="white", choices=["white", "atan", "w2"])
    p.add_argument("--precision", default="fp64", choices=["fp64", "split32"])
    p.add_argument("--cpu-baseline", default="auto", choices=["auto", "reference", "port", "none"])
    p.add_argument("--cpu-reps", type=int, default=3, help="commutators in the CPU sample")
    return p.parse_args()


def kernel_work(info, basis):
    """Algorithmic flop (GEMMs) or bytes (gathers) per launch of each kernel class."""
    dims = np.array([b.dim for b in basis.blocks], dtype=np.float64)
    upper = float((dims * (dims + 1) / 2).sum()) * 8.0  # bytes of the upper triangles
    full = float((dims * dims).sum()) * 8.0
    cc = info["ccsize"] * 8.0
    w = info["cwsize"] * 8.0
    return {
        "gemm_pphh": ("mfma", info["flop_pphh"]),
        "gemm_ph": ("mfma", info["flop_ph"]),
        # Pandya: read every channel block once, write the cross-coupled matrix
        "pandya": ("hbm", full + cc),
        # 122: read L and R channel blocks once, write the upper triangle of RES
        "k122": ("hbm", 2 * full + upper),
        # ph recoupling: read W once, read-modify-write the upper triangle of RES
        "ph_gather": ("hbm", w + 2 * upper),
        "pphh_combine": ("hbm", 2 * full + 2 * upper),
        "mirror": ("hbm", 2 * upper),
    }


def cpu_baseline(mode, basis, Hn_vec, eta_vec, comm_per_step, reps):
    """Time the reference Fortran (oracle/_ref) or the C restatement on the
    host cores: one sample of `reps` emax=8 commutators [eta, H], converted to
    steps/s with the measured commutators per step."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    kind = None
    if mode in ("auto", "reference") and O.ref_available():
        kind = "reference"
        ref = O.RefRunner(basis, threads=threads)
        sec = ref.time_comm(-1, eta_vec, 1, Hn_vec, reps)
    elif mode in ("auto", "port"):
        kind = "port"
        O.lib().or_set_threads(threads)
        ob = O.OBasis(basis)
        L, R = ob.op(eta_vec, -1), ob.op(Hn_vec, 1)
        t0 = time.perf_counter()
        for _ in range(reps):
            O.commutator(L, R)
        sec = (time.perf_counter() - t0) / reps
    else:
        return None
    return {"value": 1.0 / (sec * comm_per_step), "unit": "steps/s", "cores": threads, "kind": kind,
            "sample": f"{reps} emax={basis_tag(basis)} commutators [eta,H] on {threads} host threads, "
                      f"{sec:.3f} s each, x {comm_per_step:.2f} commutators per Magnus step"}


def basis_tag(basis):
    return int(max(2 * basis.n + basis.l))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); multi-GPU sharding of "
              "this path is design-only", file=sys.stderr)
        sys.exit(2)
    from imsrg_mi355x import Basis, synthetic_hamiltonian
    from imsrg_mi355x import engine as E

    basis = Basis.ho(a.emax, a.holes)
    log(f"emax={a.emax} holes/iso={a.holes} orbits={basis.nsp} blocks={len(basis.blocks)} neq={basis.neq}")
    ctx = E.Context(basis, precision=E.PREC_SPLIT_FP32 if a.precision == "split32" else E.PREC_FP64)
    info = ctx.info()
    H0 = synthetic_hamiltonian(basis, seed=0, hw=20.0, sigma1=0.5, sigma2=1.0)
    HS = ctx.op(1, H0)
    HS.normal_order()
    Hn_vec = HS.download() if a.cpu_baseline != "none" else None
    eta_vec = ctx.generator(a.generator, HS).download() if a.cpu_baseline != "none" else None
    flow = ctx.flow(HS, a.generator, 0.5)
    log(f"flow set up; device bytes/operator {info['op_bytes'] / 2**20:.1f} MiB; warmup {a.warmup} steps")
    for k in range(a.warmup):
        r = flow.step()
        log(f"warmup step {r['step']} s={r['s']} E0={r['e0']:.10f} magnus={r['n_magnus']} bch={r['n_bch']}")
    ctx.sync()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    recs = []
    t0 = time.perf_counter()
    for k in range(a.steps):
        recs.append(flow.step())
    ctx.sync()
    t1 = time.perf_counter()
    prof = ctx.profile_read(reset=True)
    ctx.profile(False)
    for r in recs:
        log(f"timed step {r['step']} s={r['s']} E0={r['e0']:.10f} crit={r['crit']:.3e} "
            f"magnus={r['n_magnus']} bch={r['n_bch']}")
    elapsed = t1 - t0
    steps_per_s = a.steps / elapsed
    gen_comm = 1 if a.generator == "w2" else 0
    ncomm = sum(r["n_magnus"] + r["n_bch"] + gen_comm for r in recs)
    comm_per_step = ncomm / a.steps
    # roofline of the dominant kernel class
    work = kernel_work(info, basis)
    dom = max(work, key=lambda k: prof[k][0])
    ms, n = prof[dom]
    bound, per_launch = work[dom]
    avg_s = ms / 1e3 / n
    if bound == "mfma":
        achieved = per_launch / avg_s / 1e12
        peak = PEAK["fp64_mfma_tflops"] if a.precision == "fp64" else PEAK["fp32_mfma_tflops"]
        unit = "TFLOP/s"
    else:
        achieved = per_launch / avg_s / 1e9
        peak = PEAK["hbm_gbs"]
        unit = "GB/s"
    kernels = {k: {"ms_total": round(v[0], 3), "launches": v[1]} for k, v in prof.items() if v[1]}
    log("kernel classes (ms, launches):", json.dumps(kernels))
    cpu = None
    if a.cpu_baseline != "none":
        log("timing the CPU baseline sample")
        cpu = cpu_baseline(a.cpu_baseline, basis, Hn_vec, eta_vec, comm_per_step, a.cpu_reps)
    out = {
        "metric": "Magnus BCH steps/s at emax=8: split-FP32 vs FP64 GPU at |dE0|<1e-6 MeV",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "split-f32",
        "data": "synthetic",
        "config": {
            "workload": f"Magnus flow emax={a.emax}, {a.holes} holes/isospin, {a.generator} generator, "
                        "ds=0.5, synthetic Gaussian H (seed 0, hw=20)",
            "orbits": basis.nsp, "channel_blocks": len(basis.blocks), "neq": basis.neq,
            "commutators_per_step": comm_per_step, "commutators_per_s": ncomm / elapsed,
            "precision": a.precision, "first_timed_s": recs[0]["s"], "E0_last": recs[-1]["e0"],
        },
        "roofline": {"kernel": dom, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
                     "frac": achieved / peak, "traffic": None,
                     "per_launch": per_launch, "avg_launch_ms": avg_s * 1e3, "launches": n},
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
