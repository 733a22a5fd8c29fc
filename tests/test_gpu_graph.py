"""Graph mode: commutators captured into hipGraphs and replayed.  Every replay
is checked against the eager result bit for bit, with eager work and stream
synchronizations between the replays; whole flows (single and batched) give
the eager trajectory exactly while reusing a handful of graphs."""
import numpy as np
import pytest

from imsrg_mi355x import Basis, batch_hamiltonians, random_operator, synthetic_hamiltonian

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def e4():
    from imsrg_mi355x import engine as E
    return Basis.ho(4, 3), E


def test_graph_replays_match_eager(e4):
    B, E = e4
    ctx = E.Context(B)
    pairs = [(ctx.op(-1, random_operator(B, -1, 60 + k)), ctx.op(1, random_operator(B, 1, 70 + k))) for k in range(3)]
    outs = [ctx.op(1) for _ in pairs]
    want = []
    for (L, R), o in zip(pairs, outs):
        ctx.commutator(L, R, o)
        want.append(o.download())
    ctx.set_graphs(True)
    X = ctx.op(1, synthetic_hamiltonian(B, seed=3))
    for rep in range(4):
        for k, ((L, R), o) in enumerate(zip(pairs, outs)):
            o.clear()
            ctx.commutator(L, R, o)  # graph build on the first pass, replay afterwards
            X.add(1.0, o, 0.5)        # eager work between replays
            X.normal_order()
            ctx.sync()
            got = o.download()
            print(f"[graph] replay {rep} pair {k}: max |diff| {np.abs(got - want[k]).max():.1e}", flush=True)
            assert np.array_equal(got, want[k])
    builds, launches = ctx.graph_stats()
    assert builds == 3 and launches == 12


def test_graph_flow_matches_eager(e4):
    B, E = e4
    H = synthetic_hamiltonian(B, seed=0)

    def run(graphs, ctx, stats):
        ctx.set_graphs(graphs)
        HS = ctx.op(1, H).normal_order()
        f = ctx.flow(HS, "white", 0.5)
        recs = []
        for _ in range(6):
            recs.append(f.step())
            stats.append(ctx.graph_stats())
        f.close()
        return recs, HS.download()

    ctx = E.Context(B)
    r0, h0 = run(False, ctx, [])
    stats = []
    r1, h1 = run(True, ctx, stats)
    assert [r["e0"] for r in r0] == [r["e0"] for r in r1]
    assert np.array_equal(h0, h1)
    builds, launches = stats[-1]
    print(f"[graph] flow: {builds} graphs built, {launches} replays; per step {stats}", flush=True)
    # one graph per nested commutator of a series (the series coefficient is
    # part of the key): once the series lengths have been seen, steps only
    # replay graphs
    assert stats[-1][0] == stats[-3][0] and launches > 2 * builds


def test_graph_batched_flow_matches_eager(e4):
    B, E = e4
    Hb = batch_hamiltonians(B, range(4))
    out = []
    for graphs in (False, True):
        ctx = E.Context(B, nbatch=4)
        ctx.set_graphs(graphs)
        HS = ctx.op(1, Hb).normal_order()
        f = ctx.flow(HS, "white", 0.5)
        recs = [f.step() for _ in range(5)]
        f.close()
        out.append(([[x["e0"] for x in r] for r in recs], HS.download()))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


def test_two_stream_commutators_match_one_stream():
    """the pp/hh and ph chains on two streams (the default) give the
    single-stream results bit for bit: single commutators of every hermiticity
    pair at emax=6, a BCH series and a batched one at emax=4"""
    from imsrg_mi355x import engine as E
    B = Basis.ho(6, 3)
    ctx = E.Context(B)
    for hl, hr in ((-1, 1), (1, 1), (-1, -1)):
        L, R = ctx.op(hl, random_operator(B, hl, 81)), ctx.op(hr, random_operator(B, hr, 82))
        outs = []
        for dual in (False, True, False, True):
            ctx.set_concurrency(dual)
            o = ctx.commutator(L, R)
            outs.append(o.download())
        print(f"[streams] emax=6 herm ({hl},{hr}): max |diff| {np.abs(outs[1] - outs[0]).max():.1e}", flush=True)
        for o in outs[1:]:
            assert np.array_equal(o, outs[0])
    B4 = Basis.ho(4, 3)
    for nb in (1, 3):
        ctxb = E.Context(B4, nbatch=nb)
        Hv = batch_hamiltonians(B4, range(nb)) if nb > 1 else synthetic_hamiltonian(B4, seed=5)
        res = []
        for dual in (False, True):
            ctxb.set_concurrency(dual)
            H = ctxb.op(1, Hv).normal_order()
            G = ctxb.generator("white", H)
            G.scale(0.3)
            HS, nt = ctxb.bch(G, H)
            res.append((HS.download(), np.atleast_1d(nt)))
        print(f"[streams] emax=4 nbatch={nb} BCH terms {res[0][1]}", flush=True)
        assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_checkpoint_restart_continues_the_flow(tmp_path):
    """magnus_decouple's checkpoint restart: Omega(s) written as the
    reference's .bin stream after 3 steps, read back into a fresh context, and
    the flow resumed there (HS = BCH(Omega, H(0)), generator rebuilt) takes the
    same later steps bit for bit as the uninterrupted flow"""
    from imsrg_mi355x import engine as E
    B = Basis.ho(4, 3)
    H = synthetic_hamiltonian(B, seed=0)
    ctx = E.Context(B)
    HS = ctx.op(1, H).normal_order()
    H0 = HS.download()
    f = ctx.flow(HS, "white", 0.5)
    recs = [f.step() for _ in range(6)]
    f.close()
    path = tmp_path / "omega_s1.500.bin"
    # replay the first 3 steps and checkpoint Omega(1.5)
    HS2 = ctx.op(1, H0)
    f2 = ctx.flow(HS2, "white", 0.5)
    for _ in range(3):
        f2.step()
    f2.G.write_stream(path)
    assert np.array_equal(np.fromfile(path, dtype=np.float64), f2.G.download())
    f2.close()
    ctx2 = E.Context(B)
    G = ctx2.op(-1).read_stream(path)
    HS3 = ctx2.op(1, H0)
    f3 = ctx2.flow(HS3, "white", 0.5, G=G, restart_s=1.5)
    later = [f3.step() for _ in range(3)]
    f3.close()
    print(f"[restart] E0 uninterrupted {[r['e0'] for r in recs[3:]]}, resumed {[r['e0'] for r in later]}", flush=True)
    for a, b in zip(recs[3:], later):
        assert a["e0"] == b["e0"] and a["s"] == b["s"] and a["step"] == b["step"]
        assert a["n_magnus"] == b["n_magnus"] and a["n_bch"] == b["n_bch"]
    with pytest.raises(E.ImsrgError):
        ctx2.op(1).read_stream(tmp_path / "missing.bin")
    (tmp_path / "short.bin").write_bytes(b"\0" * 16)
    with pytest.raises(E.ImsrgError):
        ctx2.op(1).read_stream(tmp_path / "short.bin")
