"""Shared helpers for the test suite."""
import numpy as np

RTOL = 1e-10  # FP64 parity bar: max |a-b| / max |b|


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(float(np.abs(b).max()) if b.size else 0.0, 1e-300)
    return float(np.abs(a - b).max()) / scale if b.size else 0.0


def assert_close(a, b, tol=RTOL, what=""):
    r = rel(a, b)
    assert r <= tol, f"{what}: relative error {r:.3e} > {tol:.1e}"


# ---------------------------------------------------------------- sketches
# A sketch stands for an operator stream too large to commit (emax=6 and 8):
# its 2-norm, its projections on four seeded Gaussian directions, its 32
# largest entries and 224 entries sampled among the non-zero ones, and the
# SHA-256 of its bytes.  A plausible mistake (a dropped term, a flipped sign in
# one block, a transposed operand) moves the norm or a projection by far more
# than rounding does.
SKETCH_SEED = 20260501


def _direction(n, k):
    return np.random.Generator(np.random.PCG64(SKETCH_SEED + k)).standard_normal(n)


def sketch(v, nproj=4):
    import hashlib
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    nz = np.flatnonzero(v)
    top = nz[np.argsort(-np.abs(v[nz]), kind="stable")[:32]] if nz.size else nz
    rng = np.random.Generator(np.random.PCG64(SKETCH_SEED + v.size))
    rest = np.setdiff1d(nz, top)
    samp = rng.choice(rest, size=min(224, rest.size), replace=False) if rest.size else rest
    idx = np.sort(np.concatenate([top, samp])).astype(np.int64)
    return {"n": np.int64(v.size), "norm": np.float64(np.linalg.norm(v)), "maxabs": np.float64(np.abs(v).max()),
            "proj": np.array([float(_direction(v.size, k) @ v) for k in range(nproj)]),
            "idx": idx, "val": v[idx], "sha256": np.array(hashlib.sha256(v.tobytes()).hexdigest())}


def sketch_store(d, key, v):
    for k, x in sketch(v).items():
        d[f"{key}__{k}"] = x


def sketch_check(d, key, v, tol=1e-10, exact=False):
    """v against the stored sketch `key`: bytes identical with exact, else the
    norm and sampled entries to tol relative (of the norm, of the largest
    entry) and each projection to 10*tol of the norm"""
    import hashlib
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    assert v.size == int(d[f"{key}__n"]), f"{key}: length {v.size} != {int(d[f'{key}__n'])}"
    if exact:
        assert hashlib.sha256(v.tobytes()).hexdigest() == str(d[f"{key}__sha256"]), f"{key}: input differs"
        return
    norm, mx = float(d[f"{key}__norm"]), float(d[f"{key}__maxabs"])
    assert abs(np.linalg.norm(v) - norm) <= tol * norm, f"{key}: norm {np.linalg.norm(v)} vs {norm}"
    idx = d[f"{key}__idx"]
    err = float(np.abs(v[idx] - d[f"{key}__val"]).max()) if idx.size else 0.0
    assert err <= tol * mx, f"{key}: sampled entries differ by {err:.3e} (max entry {mx:.3e})"
    for k, p in enumerate(d[f"{key}__proj"]):
        q = float(_direction(v.size, k) @ v)
        assert abs(q - p) <= 10 * tol * norm, f"{key}: projection {k} {q} vs {p}"
