"""Writes small interaction files in the four formats interaction_IO.f90 reads,
with seeded random matrix elements, for the reader parity tests.

The layouts are the ones the reference's readers parse (interaction_IO.f90):
* ascii  8 header lines, then free-format "Tz Par J a b c d V" (orbits 1-based
         in the basis order, J doubled);
* gz     gzip: "<count> XXXX", then per element seven I4 fields, an E19.6
         value, a blank and the marker XXXX (52 characters: read_morten_gz
         keeps 52, and read_gz cuts each line two characters before XXXX);
* me2j   gzip: a header line, then the isospin-coupled values (four per
         (nlj1, nlj2, nnlj1, nnlj2, J) of the file's basis) ten per line as
         F12.7 and nine F13.7;
* me2b   gzip two-body file (three header lines, one size line per block,
         then per block a blank and a label line and "A  B  V" lines ending
         with the block's last diagonal element) and its .me1b file (version,
         count, E0, then " t lj  n1 n2  V" lines).
Orbit labels, block orders and counts follow the loops of those readers.
"""
import gzip

import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _block_pairs(B):
    """(J, Tz, P, list of (a, b) with a <= b) of every channel block"""
    out = []
    for blk in B.blocks:
        prs = [tuple(int(x) for x in p) for c in range(3) for p in blk.pairs[c]]
        out.append((blk.J, blk.Tz, blk.P, prs))
    return out


def morten_elements(B, seed, nmax=400):
    """(Tz, Par, J, a, b, c, d, V) with 1-based orbits: random upper-triangle
    elements of random blocks, some pairs written in swapped order, one
    element repeated with a new value (the reader keeps the last)"""
    rng = _rng(seed)
    blocks = [b for b in _block_pairs(B) if b[3]]
    rows = []
    while len(rows) < nmax:
        J, Tz, P, prs = blocks[rng.integers(len(blocks))]
        i, k = sorted(rng.integers(len(prs), size=2))
        (a, b), (c, d) = prs[i], prs[k]
        if rng.random() < 0.3 and a != b:
            a, b = b, a
        if rng.random() < 0.3 and c != d:
            c, d = d, c
        rows.append((Tz, P, J, a + 1, b + 1, c + 1, d + 1, float(np.round(rng.uniform(-5, 5), 6))))
    rows.append(rows[7][:7] + (float(np.round(rng.uniform(-5, 5), 6)),))
    return rows


def write_ascii(path, rows):
    with open(path, "w") as f:
        for k in range(8):
            f.write(f"# header line {k + 1}\n")
        for r in rows:
            f.write(" ".join(str(x) for x in r[:7]) + f" {r[7]:.6f}\n")


def write_gz(path, rows):
    with gzip.open(path, "wt") as f:
        f.write(f"{len(rows):10d} XXXX\n")
        for r in rows:
            line = "".join(f"{x:4d}" for x in r[:7]) + f"{r[7]:19.6E}" + " XXXX"
            assert len(line) == 52
            f.write(line + "\n")


def me2j_basis(emax_file, lmax_file):
    """the file's orbits (n, l, 2j, 2tz) in hk order, proton then neutron"""
    out = []
    for e in range(emax_file + 1):
        for l in range(e % 2, min(e, lmax_file) + 1, 2):
            for jj in range(abs(2 * l - 1), 2 * l + 2, 2):
                for tz in (-1, 1):
                    out.append(((e - l) // 2, l, jj, tz))
    return out


def me2j_count(emax_file, lmax_file):
    fb = me2j_basis(emax_file, lmax_file)
    nf = len(fb) // 2
    n = 0
    for k1 in range(nf):
        for k2 in range(k1 + 1):
            for k3 in range(k1 + 1):
                for k4 in range((k2 if k1 == k3 else k3) + 1):
                    o1, o2, o3, o4 = fb[2 * k1], fb[2 * k2], fb[2 * k3], fb[2 * k4]
                    if (o1[1] + o2[1]) % 2 != (o3[1] + o4[1]) % 2:
                        continue
                    jmin = max(abs(o1[2] - o2[2]), abs(o3[2] - o4[2]))
                    jmax = min(o1[2] + o2[2], o3[2] + o4[2])
                    if jmin <= jmax:
                        n += 4 * ((jmax - jmin) // 2 + 1)
    return n


def write_me2j(path, emax_file, lmax_file, seed):
    vals = np.round(_rng(seed).uniform(-5, 5, me2j_count(emax_file, lmax_file)), 7)
    with gzip.open(path, "wt") as f:
        f.write("    me2j synthetic file\n")
        for i in range(0, len(vals), 10):
            chunk = vals[i:i + 10]
            f.write(f"{chunk[0]:12.7f}" + "".join(f"{v:13.7f}" for v in chunk[1:]) + "\n")
    return vals


def _me2b_blocks(B):
    """per block of the file, in its order (Tz = 1, 0, -1; parity; J), the
    pair count the reader's enumeration gives"""
    lmax = int(max(B.l))
    emax = int(max(2 * B.n + B.l))
    ljs = [(twol // 2, twoj) for twol in range(0, 2 * lmax + 1, 2) for twoj in range(abs(twol - 1), twol + 2, 2)]
    nmax = [(emax - l) // 2 for (l, _) in ljs]
    jtm = int(max(B.jj))
    out = []
    for Tz in (1, 0, -1):
        for Pi in (0, 1):
            for JT in range(0, 2 * jtm + 1, 2):
                if lmax == emax and JT == 2 * jtm and abs(Tz) == 1:
                    continue
                if JT == 2 * jtm and Pi == 1:
                    continue
                n = 0
                for k1, (l1, j1) in enumerate(ljs):
                    for k2, (l2, j2) in enumerate(ljs):
                        if JT < abs(j1 - j2) or JT > j1 + j2 or (l1 + l2) % 2 != Pi:
                            continue
                        for n1 in range(nmax[k1] + 1):
                            idx = k1 * (nmax[0] + 1) + n1
                            for n2 in range(nmax[k2] + 1):
                                idxx = k2 * (nmax[0] + 1) + n2
                                if Tz != 0 and idx > idxx:
                                    continue
                                if (JT // 2) % 2 == 1 and k1 == k2 and n1 == n2 and Tz != 0:
                                    continue
                                n += 1
                out.append((Tz, Pi, JT, n))
    return out, ljs, nmax


def write_me2b(path2b, path1b, B, seed, per_block=40):
    """two-body file with up to per_block elements per block (upper triangle,
    ending on the last diagonal element) and the one-body file with every
    same-(l, j, tz) orbit pair of the basis and E0"""
    rng = _rng(seed)
    blocks, ljs, nmax = _me2b_blocks(B)
    with gzip.open(path2b, "wt") as f:
        f.write(f"  me2b synthetic file\n  version 0\nbMax {len(blocks):4d}\n")
        for Tz, Pi, JT, n in blocks:  # the block size in columns 10-15
            f.write(f"{Tz:3d}{Pi:3d}{JT:3d}{n:6d}\n")
        for Tz, Pi, JT, n in blocks:
            f.write("\n")
            f.write(f"  block Tz={Tz} P={Pi} J={JT} n={n}\n")
            if n == 0:
                continue
            picks = set()
            for _ in range(per_block):
                a, b = sorted(rng.integers(n, size=2))
                picks.add((int(a), int(b)))
            picks.discard((n - 1, n - 1))
            for a, b in sorted(picks) + [(n - 1, n - 1)]:
                v = float(np.round(rng.uniform(-0.99, 0.99), 8))
                f.write(f"{a}  {b}  {v:.8f}\n")
    orbs = set(zip(B.n.tolist(), B.l.tolist(), B.jj.tolist(), B.tz.tolist()))
    lines = []
    for k, (l, jj) in enumerate(ljs):
        for t in (0, 1):
            tz = -2 * t + 1
            for n1 in range(nmax[k] + 1):
                for n2 in range(nmax[k] + 1):
                    if (n1, l, jj, tz) in orbs and (n2, l, jj, tz) in orbs:
                        v = float(np.round(rng.uniform(-9, 9), 10))
                        lines.append(f" {t} {k:2d}  {n1:2d} {n2:2d}  {v:14.10f}")
    e0 = float(np.round(rng.uniform(-99, 99), 6))
    with gzip.open(path1b, "wt") as f:
        f.write(" me1b version\n")
        f.write(f"{len(lines):4d}\n")
        f.write(f"{e0:10.6f}\n")
        for ln in lines:
            f.write(ln + "\n")
