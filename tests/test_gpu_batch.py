"""Batched contexts: nbatch independent operators per handle, one launch per
kernel for all of them.  Every instance must come out exactly as the same
computation on a context of its own (same kernels, same arithmetic), and the
batched Magnus/BCH drivers must truncate every instance on its own."""
import numpy as np
import pytest

from imsrg_mi355x import Basis, batch_hamiltonians, random_operator

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def e4():
    from imsrg_mi355x import engine as E
    return Basis.ho(4, 3), E


def test_batched_commutator_each_instance_exact(e4):
    B, E = e4
    nb = 3
    Ls = np.stack([random_operator(B, -1, 10 + k) for k in range(nb)])
    Rs = np.stack([random_operator(B, 1, 20 + k) for k in range(nb)])
    bctx = E.Context(B, nbatch=nb)
    got = bctx.commutator(bctx.op(-1, Ls), bctx.op(1, Rs)).download()
    assert got.shape == (nb, B.neq)
    single = E.Context(B)
    for k in range(nb):
        want = single.commutator(single.op(-1, Ls[k]), single.op(1, Rs[k])).download()
        print(f"[batch] instance {k}: max |diff| {np.abs(got[k] - want).max():.1e}", flush=True)
        assert np.array_equal(got[k], want)


def test_batched_scalars_and_generator(e4):
    B, E = e4
    nb = 4
    H = batch_hamiltonians(B, range(nb))
    bctx = E.Context(B, nbatch=nb)
    bH = bctx.op(1, H).normal_order()
    single = E.Context(B)
    eta_b = bctx.generator("white", bH).download()
    for k in range(nb):
        sH = single.op(1, H[k]).normal_order()
        assert bH.e0[k] == sH.e0
        assert bH.norm()[k] == sH.norm()
        assert bH.mbpt2()[k] == sH.mbpt2()
        assert np.array_equal(eta_b[k], single.generator("white", sH).download())


def test_batched_bch_per_instance_truncation(e4):
    B, E = e4
    nb = 3
    H = batch_hamiltonians(B, range(nb))
    # Omega of very different sizes so the instances need different term counts
    G = np.stack([random_operator(B, -1, 40 + k) * s for k, s in enumerate((1e-4, 1e-3, 3e-3))])
    bctx = E.Context(B, nbatch=nb)
    bH = bctx.op(1, H).normal_order()
    HS, nt = bctx.bch(bctx.op(-1, G), bH)
    print(f"[batch] BCH terms per instance {list(nt)}", flush=True)
    assert len(set(nt)) > 1
    out = HS.download()
    single = E.Context(B)
    for k in range(nb):
        sH = single.op(1, H[k]).normal_order()
        sHS, n1 = single.bch(single.op(-1, G[k]), sH)
        assert n1 == nt[k]
        assert np.array_equal(out[k], sHS.download())


def test_batched_magnus_expand_per_instance_exits(e4):
    """MAGNUS_EXPAND's early exits are tested per instance on the device: an
    Omega of zero norm (no term), a tiny one (exits after the first test) and
    ordinary ones give each instance the terms and the result of its own
    single-instance run, and a second call on the same context agrees"""
    B, E = e4
    scales = (0.0, 1e-9, 1e-3, 0.3)
    nb = len(scales)
    G = np.stack([random_operator(B, -1, 50 + k) * s for k, s in enumerate(scales)])
    eta = np.stack([random_operator(B, -1, 60 + k) for k in range(nb)])
    bctx = E.Context(B, nbatch=nb)
    res = []
    for _ in range(2):
        DG, nt = bctx.magnus_expand(bctx.op(-1, eta), bctx.op(-1, G))
        res.append((DG.download(), np.array(nt)))
    print(f"[batch] Magnus terms per instance {list(res[0][1])}", flush=True)
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    assert res[0][1][0] == 0 and len(set(res[0][1])) > 1
    single = E.Context(B)
    for k in range(nb):
        sDG, n1 = single.magnus_expand(single.op(-1, eta[k]), single.op(-1, G[k]))
        assert n1 == res[0][1][k]
        assert np.array_equal(res[0][0][k], sDG.download())


def test_batched_flow_matches_single_flows(e4):
    B, E = e4
    nb = 4
    H = batch_hamiltonians(B, range(nb))
    bctx = E.Context(B, nbatch=nb)
    bH = bctx.op(1, H).normal_order()
    f = bctx.flow(bH, "white", 0.5)
    recs = []
    while not f.converged and len(recs) < 60:
        recs.append(f.step())
    f.close()
    single = E.Context(B)
    nsteps = []
    for k in range(nb):
        sH = single.op(1, H[k]).normal_order()
        sf = single.flow(sH, "white", 0.5)
        srec = []
        while not sf.converged and len(srec) < 60:
            srec.append(sf.step())
        sf.close()
        mine = [r[k] for r in recs if r[k]["stepped"]]
        nsteps.append(len(srec))
        assert len(mine) == len(srec)
        for a, b in zip(mine, srec):
            assert a["e0"] == b["e0"] and a["n_magnus"] == b["n_magnus"] and a["n_bch"] == b["n_bch"]
        print(f"[batch] instance {k}: {len(srec)} steps, E0 {srec[-1]['e0']:.10f}", flush=True)
    assert len(set(nsteps)) > 1  # the instances really stop at different steps
