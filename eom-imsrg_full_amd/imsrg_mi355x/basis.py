"""Single-particle basis and two-body channel layout (host side, numpy).

Mirrors the reference's basis conventions so that operator vectors produced
here are byte-compatible with its ``vectorize``/``repackage`` stream:

* orbit order "hk": e = 0..emax, l = e mod 2 .. e step 2, 2j = |2l-1|, 2l+1,
  2tz = -1 (proton) then +1 (interaction_IO.f90:257-274);
* holes = the lowest ``k`` orbits per isospin in that order
  (a fixed closed-shell reference, basic_IMSRG.f90:345-400);
* two-body (J,P,Tz) blocks with J (doubled) slowest, Tz = -1..1, P = 0..1
  fastest; pair lists i<=j split into pp / ph / hh classes
  (allocate_blocks, basic_IMSRG.f90:477-641);
* the six stored sub-blocks gam1..gam6 = Vpppp, Vppph, Vpphh, Vphph, Vhhhh,
  Vphhh and the vectorize order of basic_IMSRG.f90:3794-3854 (square
  sub-blocks store only rows j >= i, column-major).

All angular momenta are doubled; orbit indices are 0-based.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SEA1 = (1, 1, 1, 2, 3, 2)  # row pair class of gam1..gam6 (1=pp, 2=ph, 3=hh)
SEA2 = (1, 2, 3, 2, 3, 3)  # column pair class
SQS = (True, False, False, True, True, False)  # square sub-blocks


def hk_orbits(emax: int):
    """(n, l, 2j, 2tz) of every HO orbit with 2n+l <= emax, hk order."""
    out = []
    for e in range(emax + 1):
        for l in range(e % 2, e + 1, 2):
            n = (e - l) // 2
            for jj in range(abs(2 * l - 1), 2 * l + 2, 2):
                for tz in (-1, 1):
                    out.append((n, l, jj, tz))
    return out


def closed_shell_con(orbits, holes_per_isospin: int):
    """con[i] = 1 for the lowest ``holes_per_isospin`` orbits of each isospin."""
    con = [0] * len(orbits)
    seen = {-1: 0, 1: 0}
    for i, (_, _, _, tz) in enumerate(orbits):
        if seen[tz] < holes_per_isospin:
            con[i] = 1
            seen[tz] += 1
    return con


def _tri(j1, j2, J):
    return abs(j1 - j2) <= J <= j1 + j2


@dataclass
class Block:
    J: int
    P: int
    Tz: int
    pairs: tuple  # (pp, ph, hh) each an (n,2) int array of 0-based orbits

    @property
    def npp(self):
        return len(self.pairs[0])

    @property
    def nph(self):
        return len(self.pairs[1])

    @property
    def nhh(self):
        return len(self.pairs[2])

    @property
    def dim(self):
        return self.npp + self.nph + self.nhh

    def nclass(self, c):
        return len(self.pairs[c - 1])


@dataclass
class Basis:
    n: np.ndarray
    l: np.ndarray
    jj: np.ndarray
    tz: np.ndarray
    con: np.ndarray
    blocks: list = field(default_factory=list)

    @classmethod
    def from_orbits(cls, orbits, con):
        orbits = list(orbits)
        b = cls(
            n=np.array([o[0] for o in orbits], dtype=np.int32),
            l=np.array([o[1] for o in orbits], dtype=np.int32),
            jj=np.array([o[2] for o in orbits], dtype=np.int32),
            tz=np.array([o[3] for o in orbits], dtype=np.int32),
            con=np.array(con, dtype=np.int32),
        )
        b._build()
        return b

    @classmethod
    def ho(cls, emax: int, holes_per_isospin: int):
        orb = hk_orbits(emax)
        return cls.from_orbits(orb, closed_shell_con(orb, holes_per_isospin))

    # ------------------------------------------------------------------
    @property
    def nsp(self):
        return len(self.n)

    @property
    def holes(self):
        return np.nonzero(self.con == 1)[0]

    @property
    def parts(self):
        return np.nonzero(self.con == 0)[0]

    @property
    def jtot_max(self):
        return int(self.jj.max())

    def _build(self):
        N = self.nsp
        jt = self.jtot_max
        self.blocks = []
        for J in range(0, 2 * jt + 1, 2):
            for Tz in (-1, 0, 1):
                for P in (0, 1):
                    lists = ([], [], [])
                    for i in range(N):
                        for j in range(i, N):
                            if Tz != (self.tz[i] + self.tz[j]) // 2:
                                continue
                            if P != (self.l[i] + self.l[j]) % 2:
                                continue
                            if not _tri(self.jj[i], self.jj[j], J):
                                continue
                            lists[self.con[i] + self.con[j]].append((i, j))
                    pairs = tuple(np.array(x, dtype=np.int32).reshape(-1, 2) for x in lists)
                    self.blocks.append(Block(J, P, Tz, pairs))
        A = int(self.con.sum())
        Pn = N - A
        neq = 1 + A * A + A * Pn + Pn * Pn
        for blk in self.blocks:
            np_, nb, nh = blk.npp, blk.nph, blk.nhh
            neq += (nh * nh + nh + nb * nb + nb + np_ * np_ + np_) // 2 + nb * np_ + nh * np_ + nh * nb
        self.neq = neq

    def orbit_table(self):
        """(nsp, 5) int array n, l, 2j, 2tz, con (the C-ABI basis input)."""
        return np.stack([self.n, self.l, self.jj, self.tz, self.con], axis=1).astype(np.int32)

    def write_basis_file(self, path):
        with open(path, "w") as f:
            f.write(f"{self.nsp}\n")
            for row in self.orbit_table():
                f.write(" ".join(str(int(x)) for x in row) + "\n")

    def same_ljt(self):
        """boolean (N,N): orbits share (l, 2j, 2tz)."""
        return (
            (self.l[:, None] == self.l[None, :])
            & (self.jj[:, None] == self.jj[None, :])
            & (self.tz[:, None] == self.tz[None, :])
        )

    def pauli_mask(self, blk: Block):
        """per full-matrix index: False where the pair is (a,a) with J/2 odd."""
        ok = []
        for c in range(3):
            pr = blk.pairs[c]
            bad = (pr[:, 0] == pr[:, 1]) & ((blk.J // 2) % 2 == 1)
            ok.append(~bad)
        return np.concatenate(ok) if ok else np.zeros(0, bool)

    # ------------------------------------------------------------------
    # conversions between a full (herm-symmetric) representation and the
    # reference vectorize stream
    def vec_from_full(self, e0, f_full, v_blocks):
        """Build a vectorize-order stream from E0, the full one-body matrix
        f_full (N,N) and per-block full matrices v_blocks[q] (dim,dim) with
        pair order [pp | ph | hh]."""
        A = self.holes
        P = self.parts
        out = [np.array([e0], dtype=np.float64)]
        out.append(np.asarray(f_full)[np.ix_(A, A)].reshape(-1, order="F"))
        out.append(np.asarray(f_full)[np.ix_(P, P)].reshape(-1, order="F"))
        out.append(np.asarray(f_full)[np.ix_(P, A)].reshape(-1, order="F"))
        for blk, M in zip(self.blocks, v_blocks):
            off = (0, blk.npp, blk.npp + blk.nph, blk.dim)
            for g in range(6):
                r0, r1 = off[SEA1[g] - 1], off[SEA1[g]]
                c0, c1 = off[SEA2[g] - 1], off[SEA2[g]]
                X = M[r0:r1, c0:c1]
                if X.size == 0:
                    continue
                if SQS[g]:
                    # column-major, rows j >= i: X(j,i) for i, j>=i
                    ii, jj = np.tril_indices(X.shape[0])
                    # order: for col i ascending, row j ascending -> sort by (i, j)
                    order = np.lexsort((ii, jj))
                    out.append(X[ii[order], jj[order]])
                else:
                    out.append(X.reshape(-1, order="F"))
        v = np.concatenate(out)
        assert v.size == self.neq, (v.size, self.neq)
        return v

    def full_from_vec(self, v, herm):
        """Inverse of vec_from_full (repackage semantics, basic_IMSRG.f90:3857-3918)."""
        v = np.asarray(v, dtype=np.float64)
        A = self.holes
        P = self.parts
        N = self.nsp
        nA, nP = len(A), len(P)
        k = 0
        e0 = v[k]
        k += 1
        f = np.zeros((N, N))
        fhh = v[k:k + nA * nA].reshape(nA, nA, order="F"); k += nA * nA
        fpp = v[k:k + nP * nP].reshape(nP, nP, order="F"); k += nP * nP
        fph = v[k:k + nP * nA].reshape(nP, nA, order="F"); k += nP * nA
        f[np.ix_(A, A)] = fhh
        f[np.ix_(P, P)] = fpp
        f[np.ix_(P, A)] = fph
        f[np.ix_(A, P)] = herm * fph.T
        mats = []
        for blk in self.blocks:
            off = (0, blk.npp, blk.npp + blk.nph, blk.dim)
            M = np.zeros((blk.dim, blk.dim))
            for g in range(6):
                r0, r1 = off[SEA1[g] - 1], off[SEA1[g]]
                c0, c1 = off[SEA2[g] - 1], off[SEA2[g]]
                nr, nc = r1 - r0, c1 - c0
                if nr * nc == 0:
                    continue
                if SQS[g]:
                    n = nr
                    cnt = n * (n + 1) // 2
                    ii, jj = np.tril_indices(n)
                    order = np.lexsort((ii, jj))
                    X = np.zeros((n, n))
                    X[ii[order], jj[order]] = v[k:k + cnt]
                    # repackage also writes X(i,j) = v*herm (diagonal included, last)
                    X[jj[order], ii[order]] = herm * v[k:k + cnt]
                    k += cnt
                else:
                    X = v[k:k + nr * nc].reshape(nr, nc, order="F")
                    k += nr * nc
                M[r0:r1, c0:c1] = X
                if not SQS[g]:
                    M[c0:c1, r0:r1] = herm * X.T
            mats.append(M)
        assert k == self.neq
        return e0, f, mats


def gemm_flops_222(basis: Basis):
    """Algorithmic GEMM flop of one commutator's 222 terms.

    pp/hh ladder: X = L D R over the full channel matrix with the intermediate
    index running over pp and hh pairs: 2 * dim^2 * (npp + nhh) per block.
    ph: W = LCC RCC^T per Pandya block: 2 * rlen^2 * nph.
    """
    pphh = sum(2.0 * b.dim * b.dim * (b.npp + b.nhh) for b in basis.blocks)
    ph = sum(2.0 * r * r * nb for (_, r, nb) in pandya_blocks(basis))
    return pphh, ph


def pandya_blocks(basis: Basis):
    """(J, rlen, nph) of every Pandya block in reference order
    (cross_coupled.f90:61-177: JC fastest, |Tz|, parity slowest)."""
    N = basis.nsp
    JTM = basis.jtot_max
    out = []
    for q1 in range((JTM + 1) * 4):
        JC = (q1 % (JTM + 1)) * 2
        PAR = q1 // (2 * JTM + 2)
        TZ = (q1 % (2 * JTM + 2)) // (JTM + 1)
        r = nb = 0
        for i in range(N):
            for j in range(N):
                if not _tri(basis.jj[i], basis.jj[j], JC):
                    continue
                if (basis.l[i] + basis.l[j]) % 2 != PAR:
                    continue
                if abs(basis.tz[i] - basis.tz[j]) // 2 != TZ:
                    continue
                r += 1
                if basis.con[i] == 0 and basis.con[j] == 1:
                    nb += 1
        out.append((JC, r, nb))
    return out
