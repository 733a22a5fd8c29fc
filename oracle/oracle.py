"""ctypes view of the oracle (the C restatement) and a runner for the
reference-built harness (oracle/_ref/ref_harness).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle_imsrg.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")
# the same harness linked against the product's substitute module commutators
SUB_HARNESS = os.path.join(HERE, "_ref", "sub_harness")

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        sig = {
            "or_basis_create": (vp, [C.c_int, ip, ip, ip, ip, ip]),
            "or_basis_free": (None, [vp]),
            "or_basis_neq": (C.c_int64, [vp]),
            "or_sixj": (C.c_double, [vp] + [C.c_int] * 6),
            "or_ninej": (C.c_double, [C.c_int] * 9),
            "or_op_create": (vp, [vp, C.c_int]),
            "or_op_free": (None, [vp]),
            "or_op_set_herm": (None, [vp, C.c_int]),
            "or_op_get_herm": (C.c_int, [vp]),
            "or_op_from_vec": (None, [vp, dp]),
            "or_op_to_vec": (None, [vp, dp]),
            "or_op_get_e0": (C.c_double, [vp]),
            "or_op_set_e0": (None, [vp, C.c_double]),
            "or_f_elem": (C.c_double, [C.c_int, C.c_int, vp]),
            "or_v_elem": (C.c_double, [C.c_int] * 5 + [vp]),
            "or_comm110": (C.c_double, [vp, vp]),
            "or_comm220": (C.c_double, [vp, vp]),
            "or_comm111": (None, [vp, vp, vp]),
            "or_comm121": (None, [vp, vp, vp]),
            "or_comm122": (None, [vp, vp, vp]),
            "or_comm222_pp_hh": (None, [vp, vp, vp, vp, vp]),
            "or_comm221": (None, [vp, vp, vp, vp, vp]),
            "or_cc_create": (vp, [vp]),
            "or_cc_free": (None, [vp]),
            "or_calc_cross_coupled": (None, [vp, vp]),
            "or_cc_block": (dp, [vp, C.c_int]),
            "or_basis_ncc": (C.c_int, [vp]),
            "or_basis_cc_info": (None, [vp, C.c_int, ip]),
            "or_comm222_ph": (None, [vp, vp, vp, vp]),
            "or_commutator": (None, [vp, vp, vp, C.c_int]),
            "or_add": (None, [vp, C.c_double, vp, C.c_double, vp]),
            "or_copy": (None, [vp, vp]),
            "or_clear": (None, [vp]),
            "or_frob_norm": (C.c_double, [vp]),
            "or_mbpt2": (C.c_double, [vp]),
            "or_normal_order": (None, [vp]),
            "or_build_generator": (None, [C.c_int, vp, vp]),
            "or_bch_expand": (C.c_int, [vp, vp, vp]),
            "or_magnus_expand": (C.c_int, [vp, vp]),
            "or_magnus_flow": (C.c_int, [vp, vp, C.c_int, C.c_double, C.c_int, dp, dp, dp, dp, ip, ip]),
            "or_brute_0body": (C.c_double, [vp, vp]),
            "or_brute_1body": (C.c_double, [vp, vp, C.c_int, C.c_int]),
            "or_brute_2body": (C.c_double, [vp, vp] + [C.c_int] * 5),
            "or_set_threads": (None, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


GEN_KIND = {"white": 0, "atan": 1, "w2": 2}


class OBasis:
    """Oracle basis built from an imsrg_mi355x.Basis orbit table."""

    def __init__(self, basis):
        t = np.ascontiguousarray(basis.orbit_table())
        cols = [np.ascontiguousarray(t[:, k]) for k in range(5)]
        self._keep = cols
        self.h = lib().or_basis_create(len(t), *[_iptr(c) for c in cols])
        self.neq = int(lib().or_basis_neq(self.h))
        self.basis = basis

    def __del__(self):
        try:
            lib().or_basis_free(self.h)
        except Exception:
            pass

    def sixj(self, *a):
        return lib().or_sixj(self.h, *[int(x) for x in a])

    def op(self, vec=None, herm=1):
        return OOp(self, vec, herm)

    def cc_blocks(self):
        out = []
        info = np.zeros(3, np.int32)
        for q in range(lib().or_basis_ncc(self.h)):
            lib().or_basis_cc_info(self.h, q, _iptr(info))
            out.append(tuple(int(x) for x in info))
        return out


class OOp:
    def __init__(self, ob: OBasis, vec=None, herm=1):
        self.ob = ob
        self.h = lib().or_op_create(ob.h, herm)
        if vec is not None:
            v = np.ascontiguousarray(vec, dtype=np.float64)
            assert v.size == ob.neq
            lib().or_op_from_vec(self.h, _dptr(v))

    def __del__(self):
        try:
            lib().or_op_free(self.h)
        except Exception:
            pass

    @property
    def herm(self):
        return lib().or_op_get_herm(self.h)

    @herm.setter
    def herm(self, h):
        lib().or_op_set_herm(self.h, int(h))

    @property
    def e0(self):
        return lib().or_op_get_e0(self.h)

    def vec(self):
        v = np.zeros(self.ob.neq)
        lib().or_op_to_vec(self.h, _dptr(v))
        return v

    def f(self, a, b):
        return lib().or_f_elem(a, b, self.h)

    def v(self, a, b, c, d, J):
        return lib().or_v_elem(a, b, c, d, J, self.h)


def commutator(L: OOp, R: OOp, with_e0=True):
    ob = L.ob
    RES = OOp(ob, None, -L.herm * R.herm)
    lib().or_commutator(L.h, R.h, RES.h, 1 if with_e0 else 0)
    return RES


def terms(L: OOp, R: OOp):
    """Per-term outputs of [L,R] into zeroed results (reference call contract)."""
    ob = L.ob
    hres = -L.herm * R.herm
    out = {"e110": lib().or_comm110(L.h, R.h), "e220": lib().or_comm220(L.h, R.h)}
    for name, fn in (("111", lib().or_comm111), ("121", lib().or_comm121), ("122", lib().or_comm122)):
        RES = OOp(ob, None, hres)
        fn(L.h, R.h, RES.h)
        out[name] = RES.vec()
    RES = OOp(ob, None, hres)
    w1, w2 = OOp(ob, None, 1), OOp(ob, None, 1)
    lib().or_comm222_pp_hh(L.h, R.h, RES.h, w1.h, w2.h)
    out["222pphh"] = RES.vec()
    RES = OOp(ob, None, hres)
    lib().or_comm221(L.h, R.h, RES.h, w1.h, w2.h)
    out["221"] = RES.vec()
    LCC, RCC, WCC = (lib().or_cc_create(ob.h) for _ in range(3))
    lib().or_calc_cross_coupled(L.h, LCC)
    lib().or_calc_cross_coupled(R.h, RCC)
    RES = OOp(ob, None, hres)
    lib().or_comm222_ph(LCC, RCC, RES.h, WCC)
    out["222ph"] = RES.vec()
    for c in (LCC, RCC, WCC):
        lib().or_cc_free(c)
    return out


def cross_coupled(A: OOp):
    """Pandya matrices of A, concatenated column-major over non-empty blocks."""
    ob = A.ob
    cc = lib().or_cc_create(ob.h)
    lib().or_calc_cross_coupled(A.h, cc)
    parts = []
    for q, (_, r, nb) in enumerate(ob.cc_blocks()):
        if r * nb == 0:
            continue
        p = lib().or_cc_block(cc, q)
        parts.append(np.ctypeslib.as_array(p, shape=(r * nb,)).copy())
    lib().or_cc_free(cc)
    return np.concatenate(parts) if parts else np.zeros(0)


def generator(kind, H: OOp):
    ETA = OOp(H.ob, None, -1)
    lib().or_build_generator(GEN_KIND[kind], H.h, ETA.h)
    return ETA


def normal_order(H: OOp):
    lib().or_normal_order(H.h)
    return H


def bch(G: OOp, H: OOp):
    HS = OOp(H.ob, None, 1)
    n = lib().or_bch_expand(HS.h, G.h, H.h)
    return HS, n


def magnus_expand(DG: OOp, G: OOp):
    n = lib().or_magnus_expand(DG.h, G.h)
    return DG, n


def magnus_flow(HS: OOp, kind="white", ds=0.5, max_steps=200):
    ob = HS.ob
    G = OOp(ob, None, -1)
    s = np.zeros(max_steps); e0 = np.zeros(max_steps); emp2 = np.zeros(max_steps)
    crit = np.zeros(max_steps)
    nm = np.zeros(max_steps, np.int32); nb = np.zeros(max_steps, np.int32)
    steps = lib().or_magnus_flow(HS.h, G.h, GEN_KIND[kind], ds, max_steps, _dptr(s), _dptr(e0),
                                 _dptr(emp2), _dptr(crit), _iptr(nm), _iptr(nb))
    n = abs(steps)
    return dict(steps=steps, s=s[:n], e0=e0[:n], e0_mbpt2=emp2[:n], crit=crit[:n],
                n_magnus=nm[:n], n_bch=nb[:n], HS=HS, G=G)


def frob_norm(A: OOp):
    return lib().or_frob_norm(A.h)


def mbpt2(A: OOp):
    return lib().or_mbpt2(A.h)


# ----------------------------------------------------------------------
# Hartree-Fock without three-body forces (HF_mod.f90), restated in numpy on the
# C oracle's f_elem/v_elem.  The block eigenproblems go to numpy.linalg.eigh
# where the reference calls LAPACK dsyev (basic_IMSRG.f90:2163-2180).
HF_CONV = float(np.float32(1e-6))  # the reference's `crit > 1e-6` (a default-real literal)


def sp_blocks(B):
    """single-particle blocks of allocate_sp_mat (basic_IMSRG.f90:2600-2660):
    2tz outer, then l, then 2j; members in orbit order; empty blocks dropped"""
    out = []
    for tz in (-1, 1):
        for l in range(int(B.l.max()) + 1):
            for j in range(1, B.jtot_max + 1, 2):
                m = [i for i in range(B.nsp) if B.l[i] == l and B.jj[i] == j and B.tz[i] == tz]
                if m:
                    out.append(np.array(m))
    return out


def _write_1b(B, f, groups, Fb):
    """transform_1b_to_HF's store (HF_mod.f90:461-488): for a <= b in a
    block, pp and hh take F(a,b) in both orders; a particle-hole entry takes
    F(particle, hole) of the block matrix"""
    for g, M in zip(groups, Fb):
        for x in range(len(g)):
            for y in range(x, len(g)):
                a, b = g[x], g[y]
                if B.con[a] == B.con[b]:
                    f[a, b] = f[b, a] = M[x, y]
                elif B.con[a] == 0:
                    f[a, b] = M[x, y]
                else:
                    f[b, a] = M[y, x]


def transform_2b(B, mats, D):
    """transform_2b_to_HF (HF_mod.f90:494-619): per (J,P,Tz) block
    V' = C^T V C - (-1)^J (C^T V Crev + Crev^T V C) + Crev^T V Crev with
    C(ab,ij) = D(a,i) D(b,j) sqrt(1+d_ab)/sqrt(1+d_ij) and
    Crev(ab,ij) = D(b,i) D(a,j) (1-d_ab) (-1)^((ja+jb)/2) sqrt(1+d_ab)/sqrt(1+d_ij)"""
    out = []
    for blk, M in zip(B.blocks, mats):
        pr = np.concatenate(blk.pairs).reshape(-1, 2)
        if len(pr) == 0:
            out.append(M)
            continue
        a, b = pr[:, 0], pr[:, 1]
        dab = (a == b).astype(float)
        nrm = np.sqrt(1.0 + dab)[:, None] / np.sqrt(1.0 + dab)[None, :]
        C = D[a[:, None], a[None, :]] * D[b[:, None], b[None, :]] * nrm
        ph = (-1.0) ** ((B.jj[a] + B.jj[b]) // 2)
        Cr = D[b[:, None], a[None, :]] * D[a[:, None], b[None, :]] * ((1.0 - dab) * ph)[:, None] * nrm
        s = (-1.0) ** (blk.J // 2)
        out.append(C.T @ M @ C - s * (C.T @ M @ Cr + Cr.T @ M @ C) + Cr.T @ M @ Cr)
    return out


def hartree_fock(B, Hvec, max_iter=100000):
    """calc_HF (HF_mod.f90:14-106) on the bare Hamiltonian stream Hvec:
    returns (stream of H in the normal-ordered HF basis, D[a, i] (old orbit a,
    new orbit i), E_HF, iterations)"""
    ob = OBasis(B)
    H = OOp(ob, Hvec, 1)
    groups = sp_blocks(B)
    con = B.con.astype(float)
    T = [np.array([[H.f(a, b) for b in g] for a in g]) for g in groups]   # write_kin_matrix
    # gamma_matrix (HF_mod.f90:169-232) is linear in rho:
    #   gam(a,c) = 1/(ja+1) sum_{blocks, b, d} rho(d,b) sum_J (J+1) v(a,b,c,d;J)
    # (the reference scales rho by 2j+1 and divides by it again); tabulated once
    K = {}
    for p, g in enumerate(groups):
        ja = int(B.jj[g[0]])
        for q, h in enumerate(groups):
            jr = int(B.jj[h[0]])
            k = np.zeros((len(g), len(g), len(h), len(h)))
            for x, a in enumerate(g):
                for y in range(x, len(g)):
                    c = g[y]
                    for u, b in enumerate(h):
                        for w, d in enumerate(h):
                            k[x, y, u, w] = sum((J + 1.0) * H.v(int(a), int(b), int(c), int(d), J)
                                                for J in range(abs(jr - ja), jr + ja + 1, 2))
            K[p, q] = k / (ja + 1.0)
    D = [np.eye(len(g)) for g in groups]
    eig = [np.zeros(len(g)) for g in groups]
    crit, it = 10.0, 0
    while crit > HF_CONV:
        rho = []
        for Dq, g in zip(D, groups):       # density_matrix (HF_mod.f90:346-373)
            Dx = Dq * con[g][None, :]
            rho.append(Dx @ Dx.T)
        Vg = []
        for p, g in enumerate(groups):
            m = np.zeros((len(g), len(g)))
            for q in range(len(groups)):
                m += np.einsum("xyuw,wu->xy", K[p, q], rho[q])
            m = np.triu(m) + np.triu(m, 1).T
            Vg.append(m)
        F = [t + v for t, v in zip(T, Vg)]
        crit = 0.0
        for q in range(len(groups)):
            w, v = np.linalg.eigh(F[q])
            crit += np.sqrt(np.sum((eig[q] - w) ** 2))
            D[q], eig[q] = v, w
        it += 1
        if it > max_iter:
            raise RuntimeError("Hartree-Fock did not converge")
    for q in range(len(groups)):          # columns with a negative diagonal flip sign
        D[q] = D[q] * np.where(np.diag(D[q]) < 0.0, -1.0, 1.0)[None, :]
    Ft = [Dq.T @ Fq @ Dq for Dq, Fq in zip(D, F)]
    Tt = [Dq.T @ Tq @ Dq for Dq, Tq in zip(D, T)]
    Vt = [Dq.T @ Vq @ Dq for Dq, Vq in zip(D, Vg)]
    ehf = 0.0                              # e_HF (HF_mod.f90:376-405)
    for g, t, v in zip(groups, Tt, Vt):
        for i in range(len(g)):
            ehf += (t[i, i] + 0.5 * v[i, i]) * con[g[i]] * (B.jj[g[i]] + 1.0)
    Dfull = np.zeros((B.nsp, B.nsp))
    for g, Dq in zip(groups, D):
        Dfull[np.ix_(g, g)] = Dq
    _, f, mats = B.full_from_vec(Hvec, 1)
    _write_1b(B, f, groups, Ft)
    return B.vec_from_full(ehf, f, transform_2b(B, mats, Dfull)), Dfull, ehf, it


def observable_to_hf(B, Ovec, Dfull, herm=1):
    """sqop_to_HF (HF_mod.f90:108-136) for a scalar observable: one-body
    D^T f D per block, the two-body transform, then normal_order"""
    groups = sp_blocks(B)
    e0, f, mats = B.full_from_vec(Ovec, herm)
    _write_1b(B, f, groups, [Dfull[np.ix_(g, g)].T @ f[np.ix_(g, g)] @ Dfull[np.ix_(g, g)] for g in groups])
    v = B.vec_from_full(e0, f, transform_2b(B, mats, Dfull))
    return normal_order(OOp(OBasis(B), v, herm)).vec()


# ----------------------------------------------------------------------
# the reference-built harness (exists only where oracle/ref.mk was run)
def ref_available():
    return os.path.exists(REF_HARNESS)


class RefRunner:
    """Runs oracle/_ref/ref_harness (or another build of it, `binary`) on
    operator streams in a temp directory."""

    def __init__(self, basis, threads=None, binary=None):
        self.basis = basis
        self.binary = binary or REF_HARNESS
        self.tmp = tempfile.mkdtemp(prefix="refh_")
        self.bfile = os.path.join(self.tmp, "basis.txt")
        basis.write_basis_file(self.bfile)
        self.env = dict(os.environ)
        if threads:
            self.env["OMP_NUM_THREADS"] = str(threads)
            self.env["MKL_NUM_THREADS"] = str(threads)

    def _w(self, name, v):
        p = os.path.join(self.tmp, name)
        np.asarray(v, dtype=np.float64).tofile(p)
        return p

    def _r(self, name):
        return np.fromfile(os.path.join(self.tmp, name), dtype=np.float64)

    def run(self, *args):
        r = subprocess.run([self.binary, self.bfile] + [str(a) for a in args], cwd=self.tmp,
                           env=self.env, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"ref_harness {args[0]} failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
        return r.stdout

    def terms(self, hL, L, hR, R):
        self.run("terms", hL, self._w("L.bin", L), hR, self._w("R.bin", R), os.path.join(self.tmp, "t"))
        out = {}
        e = open(os.path.join(self.tmp, "t_e0.txt")).read().split()
        out["e110"], out["e220"] = float(e[0]), float(e[1])
        for k in ("111", "121", "122", "222pphh", "221", "222ph"):
            out[k] = self._r(f"t_{k}.bin")
        return out

    TENSOR_TERMS = ("111", "121", "211", "122", "212", "222pphh", "221", "222ph", "full")

    def tensor(self, hG, G, rank, dpar, hT, seed):
        """a random tensor T of the reference's generator (rank, dpar doubled)
        and the TS_commutator terms of [G, T]: (T, {term: stream})"""
        self.run("tensor", hG, self._w("G.bin", G), rank, dpar, hT, seed, os.path.join(self.tmp, "ts"))
        return self._r("ts_T.bin"), {k: self._r(f"ts_{k}.bin") for k in self.TENSOR_TERMS}

    def bch_tensor(self, hG, G, rank, dpar, hT, T):
        self.run("tbch", hG, self._w("G.bin", G), rank, dpar, hT, self._w("T.bin", T), "tb.bin")
        return self._r("tb.bin")

    def time_tensor(self, hG, G, rank, dpar, hT, T, reps):
        out = self.run("ttime", hG, self._w("G.bin", G), rank, dpar, hT, self._w("T.bin", T), reps)
        return float(out.split("SECONDS_PER_TENSOR_COMMUTATOR")[1].split()[0])

    def comm(self, hL, L, hR, R):
        self.run("comm", hL, self._w("L.bin", L), hR, self._w("R.bin", R), "c.bin")
        return self._r("c.bin")

    def cc(self, h, A):
        self.run("cc", h, self._w("A.bin", A), "cc.bin")
        return self._r("cc.bin")

    def bch(self, G, H):
        self.run("bch", self._w("G.bin", G), self._w("H.bin", H), "b.bin")
        return self._r("b.bin")

    def magexp(self, DG, G):
        self.run("magexp", self._w("DG.bin", DG), self._w("G.bin", G), "m.bin")
        return self._r("m.bin")

    def gen(self, kind, H):
        self.run("gen", kind, self._w("H.bin", H), "g.bin")
        return self._r("g.bin")

    def hf(self, H, O=None):
        """calc_HF on the bare H (and sqop_to_HF on O): (H_HF, D, E_HF[, O_HF])"""
        args = ["hf", self._w("H.bin", H), "hf.bin", "D.bin"]
        if O is not None:
            args += [self._w("O.bin", O), "hfo.bin"]
        out = self.run(*args)
        e = float([ln for ln in out.splitlines() if ln.startswith("EHF ")][-1].split()[1])
        n = self.basis.nsp
        D = self._r("D.bin").reshape(n, n).T  # written column-major: D[a, i]
        res = (self._r("hf.bin"), D, e)
        return res + (self._r("hfo.bin"),) if O is not None else res

    def normal(self, H):
        self.run("normal", self._w("H.bin", H), "n.bin")
        return self._r("n.bin")

    def scal(self, h, A):
        out = self.run("scal", h, self._w("A.bin", A)).strip().splitlines()[-1].split()
        return float(out[0]), float(out[1])

    def flow(self, kind, HS, maxsteps):
        self.run("flow", kind, self._w("HS0.bin", HS), maxsteps, os.path.join(self.tmp, "f"))
        rows = [ln.split() for ln in open(os.path.join(self.tmp, "f_steps.txt")) if ln.strip()]
        steps = np.array([[float(x) for x in r] for r in rows]).reshape(-1, 5)
        return steps, self._r("f_HS.bin"), self._r("f_G.bin")

    def brute(self, hA, A, hB, B, points):
        pf = os.path.join(self.tmp, "pts.txt")
        with open(pf, "w") as f:
            for p in points:
                f.write(" ".join(str(int(x)) for x in p) + "\n")
        self.run("brute", hA, self._w("A.bin", A), hB, self._w("B.bin", B), pf, "bo.txt")
        return [float(ln.split()[-1]) for ln in open(os.path.join(self.tmp, "bo.txt")) if ln.strip()]

    def decouple(self, kind, HS):
        """magnus_decouple itself: (H(s), Omega(s), its printed step lines)"""
        self.run("decouple", kind, self._w("HSd.bin", HS), "d")
        rows = [ln.split() for ln in open(os.path.join(self.tmp, "d_0b_magnus_flow.dat")) if ln.strip()]
        return self._r("d_HS.bin"), self._r("d_G.bin"), np.array([[float(x) for x in r] for r in rows])

    def dhds(self, kind, HS):
        self.run("dhds", kind, self._w("HSh.bin", HS), "dh.bin")
        return self._r("dh.bin")

    def readint(self, fmt, path, hw=0.0, com_hw=0.0, beta=0.0, A=0, htype=0, emax_file=0, lmax_file=0):
        """the reference's reader `fmt` on path (copied beside the harness; an me2b
        file brings its .me1b file) into a zero operator"""
        import shutil
        name = os.path.basename(path)
        shutil.copy(path, os.path.join(self.tmp, name))
        if fmt == "me2b":
            shutil.copy(path[:-5] + "1b.gz", os.path.join(self.tmp, name[:-5] + "1b.gz"))
        self.run("readint", fmt, name, "ri.bin", repr(float(hw)), repr(float(com_hw)), repr(float(beta)), A, htype,
                 emax_file, lmax_file)
        return self._r("ri.bin")

    def time_comm(self, hL, L, hR, R, reps=1):
        out = self.run("time", hL, self._w("L.bin", L), hR, self._w("R.bin", R), reps)
        for ln in out.splitlines():
            if ln.startswith("SECONDS_PER_COMMUTATOR"):
                return float(ln.split()[1])
        raise RuntimeError(out)
