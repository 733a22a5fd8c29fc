"""The interaction-file readers on the host (imsrg_read_interaction_vec, no
device) against the reference's own readers (interaction_IO.f90 through
oracle/_ref/ref_harness readint) on the same files, written in each format
by tests/interaction_files.py.  Without the centre-of-mass terms the streams
must agree bit for bit; with them (r1.r2 and p1.p2 built from 6j symbols,
basic_IMSRG.f90:4310-4530) to 1e-13 relative.  The device entry point
(imsrg_read_interaction) is compared with this one in
tests/test_gpu_interaction_io.py."""
import numpy as np
import pytest

import interaction_files as F
from helpers import assert_close
from imsrg_mi355x import Basis
from oracle import oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="the reference harness is built where the reference is")
COM = dict(hw=20.0, com_hw=18.0, beta=0.4, A=16, htype=1)
PLAIN = dict(hw=20.0, com_hw=20.0, beta=0.0, A=16, htype=0)


@pytest.fixture(scope="module", params=[(2, 2), (4, 3)], ids=["emax2", "emax4"])
def env(request, tmp_path_factory):
    emax, k = request.param
    B = Basis.ho(emax, k)
    d = tmp_path_factory.mktemp(f"inth{emax}")
    rows = F.morten_elements(B, 3)
    F.write_ascii(d / "v.int", rows)
    F.write_gz(d / "v.gz", rows)
    F.write_me2j(d / "v.me2j.gz", emax + 2, emax + 1, 5)
    F.write_me2b(d / "v.me2b.gz", d / "v.me1b.gz", B, 7)
    return B, O.RefRunner(B, threads=4), d, emax


def _ours(B, fmt, path, kw, **extra):
    from imsrg_mi355x import engine as E
    return E.read_interaction_vec(B, fmt, path, hw=kw["hw"], com_hw=kw["com_hw"], lawson_beta=kw["beta"],
                                  A=kw["A"], htype=kw["htype"], **extra)


@pytest.mark.parametrize("fmt,name", [("ascii", "v.int"), ("gz", "v.gz")])
def test_morten_formats(env, fmt, name):
    B, ref, d, emax = env
    assert np.array_equal(_ours(B, fmt, d / name, PLAIN), ref.readint(fmt, str(d / name), **PLAIN))
    assert_close(_ours(B, fmt, d / name, COM), ref.readint(fmt, str(d / name), **COM), tol=1e-13,
                 what=f"{fmt} with centre-of-mass terms")


def test_me2j(env):
    B, ref, d, emax = env
    ef, lf = emax + 2, emax + 1
    got = _ours(B, "me2j", d / "v.me2j.gz", PLAIN, emax_file=ef, lmax_file=lf)
    assert np.array_equal(got, ref.readint("me2j", str(d / "v.me2j.gz"), emax_file=ef, lmax_file=lf, **PLAIN))
    got = _ours(B, "me2j", d / "v.me2j.gz", COM, emax_file=ef, lmax_file=lf)
    assert_close(got, ref.readint("me2j", str(d / "v.me2j.gz"), emax_file=ef, lmax_file=lf, **COM), tol=1e-13,
                 what="me2j with centre-of-mass terms")


def test_me2b(env):
    from imsrg_mi355x import engine as E
    B, ref, d, emax = env
    got = E.read_interaction_vec(B, "me2b", d / "v.me2b.gz", path_1b=d / "v.me1b.gz")
    assert np.array_equal(got, ref.readint("me2b", str(d / "v.me2b.gz"), **PLAIN))


def test_keeps_other_elements_and_errors(env):
    """elements the file does not list keep their values; a missing file or a
    label outside the basis is an error"""
    from imsrg_mi355x import engine as E
    from imsrg_mi355x import random_operator
    B, ref, d, emax = env
    base = random_operator(B, 1, 9)
    fresh = _ours(B, "ascii", d / "v.int", PLAIN)
    got = E.read_interaction_vec(B, "ascii", d / "v.int", vec=base, hw=20.0, A=16)
    listed = fresh != 0
    assert np.array_equal(got[listed], fresh[listed])
    assert np.array_equal(got[~listed], base[~listed])
    with pytest.raises(E.ImsrgError):
        E.read_interaction_vec(B, "ascii", d / "missing.int", hw=20.0, A=16)
    bad = d / "bad.int"
    bad.write_text("\n" * 8 + f"0 0 0 1 {B.nsp + 5} 1 1 1.0\n")
    with pytest.raises(E.ImsrgError):
        E.read_interaction_vec(B, "ascii", bad, hw=20.0, A=16)
