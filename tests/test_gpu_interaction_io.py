"""The interaction-file readers (imsrg_read_interaction, interaction_IO.cpp)
against the reference's own readers (interaction_IO.f90, run by
oracle/_ref/ref_harness readint) on the same files, written in each format by
tests/interaction_files.py: the operator each reader fills from zero must
agree to 1e-12 relative.

* ascii (read_interaction) and gz (read_gz): random elements of random
  channel blocks, some pairs in swapped order (exchange phase), one element
  repeated (last value kept), with and without the Lawson / centre-of-mass
  terms (r1.r2, p1.p2);
* me2j (read_me2j_interaction): an isospin-coupled file over a larger basis
  than the context's (eMax + 2, lMax + 1: the reader's cuts), with and
  without the centre-of-mass terms;
* me2b (read_me2b_interaction): a two-body file and its one-body file (E0 and
  the one-body part).
"""
import os

import pytest

import interaction_files as F
from helpers import assert_close
from imsrg_mi355x import Basis
from oracle import oracle as O

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.ref_available(), reason="the reference harness is built where the reference is")]
COM = dict(hw=20.0, com_hw=18.0, beta=0.4, A=16, htype=1)
PLAIN = dict(hw=20.0, com_hw=20.0, beta=0.0, A=16, htype=0)


@pytest.fixture(scope="module", params=[(2, 2), (4, 3)], ids=["emax2", "emax4"])
def env(request, tmp_path_factory):
    from imsrg_mi355x import engine as E
    emax, k = request.param
    B = Basis.ho(emax, k)
    d = tmp_path_factory.mktemp(f"int{emax}")
    rows = F.morten_elements(B, 3)
    F.write_ascii(d / "v.int", rows)
    F.write_gz(d / "v.gz", rows)
    F.write_me2j(d / "v.me2j.gz", emax + 2, emax + 1, 5)
    F.write_me2b(d / "v.me2b.gz", d / "v.me1b.gz", B, 7)
    print(f"[interaction io] emax={emax}: files in {d}", flush=True)
    return B, E.Context(B), O.RefRunner(B, threads=4), d, emax


def _ours(ctx, fmt, path, kw, **extra):
    H = ctx.op(1)
    H.read_interaction(fmt, path, hw=kw["hw"], com_hw=kw["com_hw"], lawson_beta=kw["beta"], A=kw["A"],
                       htype=kw["htype"], **extra)
    return H.download()


@pytest.mark.parametrize("fmt,name", [("ascii", "v.int"), ("gz", "v.gz")])
@pytest.mark.parametrize("kw", [PLAIN, COM], ids=["plain", "com"])
def test_morten_formats(env, fmt, name, kw):
    B, ctx, ref, d, emax = env
    got = _ours(ctx, fmt, d / name, kw)
    want = ref.readint(fmt, str(d / name), **kw)
    assert_close(got, want, tol=1e-12, what=f"{fmt} reader")


@pytest.mark.parametrize("kw", [PLAIN, COM], ids=["plain", "com"])
def test_me2j(env, kw):
    B, ctx, ref, d, emax = env
    ef, lf = emax + 2, emax + 1
    got = _ours(ctx, "me2j", d / "v.me2j.gz", kw, emax_file=ef, lmax_file=lf)
    want = ref.readint("me2j", str(d / "v.me2j.gz"), emax_file=ef, lmax_file=lf, **kw)
    assert_close(got, want, tol=1e-12, what="me2j reader")


def test_me2b(env):
    B, ctx, ref, d, emax = env
    H = ctx.op(1)
    H.read_interaction("me2b", d / "v.me2b.gz", path_1b=d / "v.me1b.gz")
    want = ref.readint("me2b", str(d / "v.me2b.gz"), **PLAIN)
    assert_close(H.download(), want, tol=1e-12, what="me2b reader")


def test_errors(env):
    from imsrg_mi355x import engine as E
    B, ctx, ref, d, emax = env
    with pytest.raises(E.ImsrgError):
        ctx.op(1).read_interaction("ascii", os.path.join(str(d), "missing.int"), hw=20.0, A=16)
    with pytest.raises(E.ImsrgError):
        ctx.op(-1).read_interaction("ascii", d / "v.int", hw=20.0, A=16)
