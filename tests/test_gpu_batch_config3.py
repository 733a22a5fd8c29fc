"""The batching workload at its real size: 16 independent synthetic
Hamiltonians at emax=6 (56 orbits, 3 holes per isospin, channel blocks of 1 to
265 pair states; seeds 0-15, hw uniform in [16, 24] MeV, two-body width
uniform in [0.5, 1.5] MeV) flowed together on one batched context with every
commutator captured in a hipGraph and replayed.

* every instance's trajectory (E0, |MBPT2|, term counts per step, final H(s)
  and Omega(s)) equals its single-instance eager flow bit for bit;
* the graphs are really used (a few builds, many more replays);
* two instances' first two Magnus steps against the oracle (restated
  IMSRG_MAGNUS.f90:98-126) at 1e-10 relative with equal term counts.
"""
import os

import numpy as np
import pytest

from imsrg_mi355x import Basis, batch_hamiltonians
from oracle import oracle as O

pytestmark = pytest.mark.gpu
NB = 16


@pytest.fixture(scope="module")
def batch_run():
    from imsrg_mi355x import engine as E
    B = Basis.ho(6, 3)
    H = batch_hamiltonians(B, range(NB))
    bctx = E.Context(B, nbatch=NB)
    bctx.set_tile_shapes(0, 0)
    bctx.set_graphs(True)
    bH = bctx.op(1, H).normal_order()
    h0 = bH.download()
    f = bctx.flow(bH, "white", 0.5)
    rows = [[] for _ in range(NB)]
    nsteps = 0
    while not f.converged and nsteps < 200:
        for k, r in enumerate(f.step()):
            if r["stepped"]:
                rows[k].append((r["e0"], r["crit"], r["n_magnus"], r["n_bch"]))
        nsteps += 1
    stats = bctx.graph_stats()
    print(f"[config3] batched flow: {nsteps} batched steps, per-instance steps "
          f"{[len(r) for r in rows]}, graph builds/launches {stats}", flush=True)
    out = dict(B=B, E=E, H=H, h0=h0, rows=rows, HS=bH.download(), G=f.G.download(), stats=stats)
    f.close()
    bctx.close()
    return out


def test_config3_batched_graph_flow_equals_single_flows(batch_run):
    d = batch_run
    B, E = d["B"], d["E"]
    assert all(len(r) > 0 for r in d["rows"])
    sctx = E.Context(B)
    sctx.set_tile_shapes(0, 0)
    for k in range(NB):
        sH = sctx.op(1, d["H"][k]).normal_order()
        assert np.array_equal(sH.download(), d["h0"][k])
        r = sctx.magnus_flow(sH, "white", 0.5, 200)
        got = d["rows"][k]
        print(f"[config3] instance {k}: {r['steps']} steps, E0 {r['e0'][-1]:.10f}", flush=True)
        assert r["steps"] == len(got)
        want = list(zip(r["e0"], r["crit"], r["n_magnus"], r["n_bch"]))
        assert [tuple(x) for x in want] == [tuple(x) for x in got]
        assert np.array_equal(r["HS"].download(), d["HS"][k])
        assert np.array_equal(r["G"].download(), d["G"][k])
    sctx.close()


def test_config3_graphs_replayed(batch_run):
    builds, launches = batch_run["stats"]
    assert builds > 0 and launches > 10 * builds


@pytest.mark.parametrize("k", [0, 11])
def test_config3_instances_vs_oracle(batch_run, k):
    d = batch_run
    O.lib().or_set_threads(min(16, os.cpu_count() or 1))
    ob = O.OBasis(d["B"])
    r = O.magnus_flow(ob.op(d["h0"][k], 1), "white", 0.5, 2)
    for s in range(2):
        e0, crit, nm, nb = d["rows"][k][s]
        assert nm == r["n_magnus"][s] and nb == r["n_bch"][s]
        assert abs(e0 - r["e0"][s]) <= 1e-10 * abs(r["e0"][s])
        assert abs(crit - r["crit"][s]) <= 1e-10 * abs(r["crit"][s])
