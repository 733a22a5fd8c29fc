"""ctypes binding of libimsrg_mi355x.so (the C ABI in include/imsrg_mi355x.h).

The shared library is loaded from this package's ``lib/`` directory, where
``__graft_entry__.build()`` (or ``make -C eom-imsrg_full_amd/csrc``) puts it.
There is no CPU fallback: if the library is missing this module raises on
import, and every call that fails on the device raises ``ImsrgError``.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from .basis import Basis

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("IMSRG_MI355X_LIB", os.path.join(HERE, "lib", "libimsrg_mi355x.so"))
HEADER = os.path.abspath(os.path.join(HERE, "..", "..", "include", "imsrg_mi355x.h"))

T110, T111, T121, T122, T220, T221, T222PPHH, T222PH = (1 << k for k in range(8))
TALL = 0xFF
# scalar-tensor commutator terms (TS_commutators.f90), IMSRG_TS* in the header
TS_TERMS = {"111": 1 << 0, "121": 1 << 1, "211": 1 << 2, "122": 1 << 3, "212": 1 << 4,
            "222pphh": 1 << 5, "221": 1 << 6, "222ph": 1 << 7}
TSALL = 0xFF
PREC_FP64, PREC_SPLIT_FP32, PREC_SPLIT_BF16, PREC_INT8X5, PREC_INT8X4 = 0, 1, 2, 3, 4
GEN = {"white": 0, "atan": 1, "w2": 2}


class ImsrgError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C eom-imsrg_full_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    return C.CDLL(LIB_PATH)


_lib = _load()
_vp, _ip, _dp, _i64 = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int64


class InteractionParams(C.Structure):
    """imsrg_interaction_params"""
    _fields_ = [("hw", C.c_double), ("com_hw", C.c_double), ("lawson_beta", C.c_double), ("A", C.c_int),
                ("htype", C.c_int), ("emax_file", C.c_int), ("lmax_file", C.c_int)]


INTERACTION_FORMATS = {"ascii": 0, "gz": 1, "me2j": 2, "me2b": 3}
_SIG = {
    "imsrg_last_error": (C.c_char_p, []),
    "imsrg_ctx_create": (C.c_int, [C.c_int, _ip, _ip, _ip, _ip, _ip, C.c_int, C.POINTER(_vp)]),
    "imsrg_ctx_create_batch": (C.c_int, [C.c_int, _ip, _ip, _ip, _ip, _ip, C.c_int, C.c_int, C.POINTER(_vp)]),
    "imsrg_ctx_nbatch": (C.c_int, [_vp]),
    "imsrg_ctx_destroy": (C.c_int, [_vp]),
    "imsrg_ctx_neq": (_i64, [_vp]),
    "imsrg_ctx_info": (C.c_int, [_vp, _dp]),
    "imsrg_ctx_set_precision": (C.c_int, [_vp, C.c_int]),
    "imsrg_ctx_synchronize": (C.c_int, [_vp]),
    "imsrg_ctx_set_graphs": (C.c_int, [_vp, C.c_int]),
    "imsrg_ctx_tuning": (C.c_int, [_vp, _ip, _ip, C.POINTER(C.c_float)]),
    "imsrg_ctx_set_tile_shapes": (C.c_int, [_vp, C.c_int, C.c_int]),
    "imsrg_ctx_graph_stats": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "imsrg_ctx_stream": (_vp, [_vp]),
    "imsrg_op_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "imsrg_op_destroy": (C.c_int, [_vp]),
    "imsrg_op_set_herm": (C.c_int, [_vp, C.c_int]),
    "imsrg_op_get_herm": (C.c_int, [_vp, _ip]),
    "imsrg_op_upload": (C.c_int, [_vp, _dp, _i64]),
    "imsrg_op_download": (C.c_int, [_vp, _dp, _i64]),
    "imsrg_op_copy": (C.c_int, [_vp, _vp]),
    "imsrg_op_clear": (C.c_int, [_vp]),
    "imsrg_op_scale": (C.c_int, [_vp, C.c_double]),
    "imsrg_op_add": (C.c_int, [_vp, C.c_double, _vp, C.c_double, _vp]),
    "imsrg_op_get_e0": (C.c_int, [_vp, _dp]),
    "imsrg_op_set_e0": (C.c_int, [_vp, C.c_double]),
    "imsrg_op_norm": (C.c_int, [_vp, _dp]),
    "imsrg_mbpt2": (C.c_int, [_vp, _dp]),
    "imsrg_op_e0_batch": (C.c_int, [_vp, _dp]),
    "imsrg_op_norm_batch": (C.c_int, [_vp, _dp]),
    "imsrg_mbpt2_batch": (C.c_int, [_vp, _dp]),
    "imsrg_normal_order": (C.c_int, [_vp]),
    "imsrg_commutator": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint]),
    "imsrg_commutator_adjoint": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_uint]),
    "imsrg_flow_gradient": (C.c_int, [_vp, _vp, C.c_int, C.c_double, C.c_int, C.c_int, _dp, _vp]),
    "imsrg_normal_order_adjoint": (C.c_int, [_vp, _vp]),
    "imsrg_op_download_cotangent": (C.c_int, [_vp, _dp, _i64]),
    "imsrg_commutator_110": (C.c_int, [_vp, _vp, _dp]),
    "imsrg_commutator_220": (C.c_int, [_vp, _vp, _dp]),
    "imsrg_cross_coupled": (C.c_int, [_vp, _dp, _i64]),
    "imsrg_cross_coupled_size": (_i64, [_vp]),
    "imsrg_commutator_222_ph_cc": (C.c_int, [_vp, _dp, C.c_int, _dp, C.c_int, _i64, _vp]),
    "imsrg_generator": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "imsrg_generator_atanxx": (C.c_int, [_vp, _vp, _vp, _vp]),
    "imsrg_bch_expand": (C.c_int, [_vp, _vp, _vp, _vp, _ip]),
    "imsrg_magnus_expand": (C.c_int, [_vp, _vp, _vp, _ip]),
    "imsrg_magnus_flow": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, _dp,
                                    _ip, _ip, _ip]),
    "imsrg_flow_create": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double, C.POINTER(_vp)]),
    "imsrg_flow_restart": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double, C.c_double, C.POINTER(_vp)]),
    "imsrg_op_write_stream": (C.c_int, [_vp, C.c_char_p]),
    "imsrg_op_read_stream": (C.c_int, [_vp, C.c_char_p]),
    "imsrg_flow_step": (C.c_int, [_vp, _dp]),
    "imsrg_flow_converged": (C.c_int, [_vp, _ip]),
    "imsrg_flow_destroy": (C.c_int, [_vp]),
    "imsrg_ctx_profile": (C.c_int, [_vp, C.c_int]),
    "imsrg_ctx_profile_classes": (C.c_int, [_vp, C.c_uint]),
    "imsrg_ctx_set_concurrency": (C.c_int, [_vp, C.c_int]),
    "imsrg_transform_bch": (C.c_int, [_vp, _vp, C.POINTER(C.c_int)]),
    "imsrg_hartree_fock": (C.c_int, [_vp, _dp, C.POINTER(C.c_int), _dp]),
    "imsrg_hartree_fock_noalloc": (C.c_int, [_vp, _dp, C.POINTER(C.c_int), _dp]),
    "imsrg_read_interaction": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_char_p, C.POINTER(InteractionParams)]),
    "imsrg_read_interaction_vec": (C.c_int, [C.c_int, _ip, _ip, _ip, _ip, _ip, C.c_int, C.c_char_p, C.c_char_p,
                                             C.POINTER(InteractionParams), _dp, _i64]),
    "imsrg_observable_to_hf": (C.c_int, [_vp, _dp]),
    "imsrg_dhds": (C.c_int, [_vp, C.c_int, _vp]),
    "imsrg_ctx_profile_read": (C.c_int, [_vp, _dp, C.POINTER(C.c_longlong), C.c_int, C.c_int]),
    "imsrg_tensor_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "imsrg_tensor_destroy": (C.c_int, [_vp]),
    "imsrg_tensor_neq": (_i64, [_vp]),
    "imsrg_tensor_set_herm": (C.c_int, [_vp, C.c_int]),
    "imsrg_tensor_get_herm": (C.c_int, [_vp, _ip]),
    "imsrg_tensor_upload": (C.c_int, [_vp, _dp, _i64]),
    "imsrg_tensor_download": (C.c_int, [_vp, _dp, _i64]),
    "imsrg_tensor_copy": (C.c_int, [_vp, _vp]),
    "imsrg_tensor_clear": (C.c_int, [_vp]),
    "imsrg_tensor_add": (C.c_int, [_vp, C.c_double, _vp, C.c_double, _vp]),
    "imsrg_tensor_norm": (C.c_int, [_vp, _dp]),
    "imsrg_tensor_commutator": (C.c_int, [_vp, _vp, _vp, C.c_uint]),
    "imsrg_bch_tensor": (C.c_int, [_vp, _vp, _ip]),
}
PROF_CLASSES = ["gemm_pphh", "gemm_ph", "pandya", "k122", "ph_apply", "ph_terms", "sym_mirror",
                "onebody", "e0", "norm", "axpby", "generator", "slice"]
for _name, (_res, _args) in _SIG.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def header_symbols(path=HEADER):
    """Every function name the C ABI header declares."""
    txt = open(path).read()
    return sorted(set(re.findall(r"\b(imsrg_\w+)\s*\(", txt)))


def _chk(rc):
    if rc != 0:
        raise ImsrgError(f"imsrg error {rc}: {_lib.imsrg_last_error().decode()}")


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


class Context:
    """A device context bound to one basis (imsrg_ctx_create / _batch).

    With nbatch > 1 every operator holds nbatch independent instances; vectors
    are then arrays of shape (nbatch, neq) and scalars arrays of nbatch."""

    def __init__(self, basis: Basis, device: int = 0, precision: int = PREC_FP64, nbatch: int = 1):
        self.basis = basis
        t = np.ascontiguousarray(basis.orbit_table())
        cols = [np.ascontiguousarray(t[:, k]) for k in range(5)]
        h = _vp()
        _chk(_lib.imsrg_ctx_create_batch(len(t), *[_iptr(c) for c in cols], device, int(nbatch), C.byref(h)))
        self.h = h
        self.nbatch = int(_lib.imsrg_ctx_nbatch(h))
        self.neq = int(_lib.imsrg_ctx_neq(h))
        if self.neq != basis.neq:
            raise ImsrgError(f"neq mismatch {self.neq} vs {basis.neq}")
        self.set_precision(precision)

    def close(self):
        if getattr(self, "h", None):
            _lib.imsrg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, prec):
        _chk(_lib.imsrg_ctx_set_precision(self.h, int(prec)))
        self.precision = prec

    def info(self):
        v = np.zeros(16)
        _chk(_lib.imsrg_ctx_info(self.h, _dptr(v)))
        keys = ["nsp", "nhole", "nblocks", "ncc", "npairs", "max_dim", "max_cc_rows", "op_bytes",
                "vsize", "ccsize", "cwsize", "flop_pphh", "flop_ph", "nphterm", "nupper"]
        return dict(zip(keys, v[:len(keys)]))

    def sync(self):
        _chk(_lib.imsrg_ctx_synchronize(self.h))

    def set_graphs(self, enable=True):
        """replay every commutator from a cached hipGraph"""
        _chk(_lib.imsrg_ctx_set_graphs(self.h, 1 if enable else 0))

    def set_concurrency(self, enable=True):
        """run the pp/hh and ph chains of each commutator on two streams (default on)"""
        _chk(_lib.imsrg_ctx_set_concurrency(self.h, 1 if enable else 0))

    def tuning(self):
        """FP64 tile shapes chosen by the autotuner and its four timings (ms)"""
        a, b = C.c_int(0), C.c_int(0)
        ms = (C.c_float * 4)()
        _chk(_lib.imsrg_ctx_tuning(self.h, C.byref(a), C.byref(b), ms))
        names = {-1: "untuned", 0: "64x64", 1: "128x64"}
        return dict(pphh=names[a.value], ph=names[b.value], ms_pphh_64=ms[0], ms_pphh_128x64=ms[1],
                    ms_ph_64=ms[2], ms_ph_128x64=ms[3])

    def set_tile_shapes(self, pphh, ph):
        _chk(_lib.imsrg_ctx_set_tile_shapes(self.h, int(pphh), int(ph)))

    def graph_stats(self):
        b, n = C.c_longlong(0), C.c_longlong(0)
        _chk(_lib.imsrg_ctx_graph_stats(self.h, C.byref(b), C.byref(n)))
        return b.value, n.value

    @property
    def stream(self):
        return _lib.imsrg_ctx_stream(self.h)

    def op(self, herm=1, vec=None):
        o = Operator(self, herm)
        if vec is not None:
            o.upload(vec)
        return o

    def tensor(self, rank, dpar=0, herm=1, vec=None):
        """a rank > 0 operator; rank and dpar doubled as the reference stores them"""
        t = TensorOp(self, rank, dpar, herm)
        if vec is not None:
            t.upload(vec)
        return t

    def tensor_commutator(self, G, T, RES=None, mask=TSALL, herm=None):
        """[G, T] terms of TS_commutators.f90 into RES (herm -herm(G) herm(T) unless given)"""
        if RES is None:
            RES = TensorOp(self, T.rank, T.dpar, herm if herm is not None else -G.herm * T.herm)
        _chk(_lib.imsrg_tensor_commutator(G.h, T.h, RES.h, int(mask)))
        return RES

    def bch_tensor(self, G, HS):
        """BCH_TENSOR(G, HS) in place; returns the number of commutators evaluated"""
        n = C.c_int(0)
        _chk(_lib.imsrg_bch_tensor(G.h, HS.h, C.byref(n)))
        return n.value

    # -- engine
    def commutator(self, L, R, RES=None, mask=TALL):
        if RES is None:
            RES = Operator(self, -L.herm * R.herm)
        _chk(_lib.imsrg_commutator(self.h, L.h, R.h, RES.h, mask))
        return RES

    def commutator_adjoint(self, P, Cbar, hY, mask=TALL, Ybar=None):
        """Ybar = d<Cbar, [P, Y]>/dY for Y of hermiticity hY (Frobenius product
        over the full layout), projected onto hY."""
        Ybar = Ybar or Operator(self, hY)
        _chk(_lib.imsrg_commutator_adjoint(self.h, P.h, Cbar.h, Ybar.h, int(hY), mask))
        return Ybar

    def flow_gradient(self, H0, kind="white", ds=0.5, nsteps=8, ncheck=8):
        """E0 after nsteps Magnus steps from the normal-ordered H0, and dE0/dH0
        as an operator (full-layout Frobenius gradient, hermitian projection)."""
        Hb = Operator(self, 1)
        e0 = np.zeros(1)
        _chk(_lib.imsrg_flow_gradient(self.h, H0.h, GEN[kind], float(ds), int(nsteps), int(ncheck), _dptr(e0),
                                      Hb.h))
        return float(e0[0]), Hb

    def generator(self, kind, H, ETA=None):
        ETA = ETA or Operator(self, -1)
        _chk(_lib.imsrg_generator(self.h, GEN[kind], H.h, ETA.h))
        return ETA

    def generator_atanxx(self, Hold, H, ETA=None):
        ETA = ETA or Operator(self, -1)
        _chk(_lib.imsrg_generator_atanxx(self.h, Hold.h, H.h, ETA.h))
        return ETA

    def _per_inst(self, a):
        return int(a[0]) if self.nbatch == 1 else a.copy()

    def bch(self, G, H, HS=None):
        """HS = e^G H e^-G; also the nested-commutator count (array per instance if batched)"""
        HS = HS or Operator(self, 1)
        n = np.zeros(self.nbatch, np.int32)
        _chk(_lib.imsrg_bch_expand(self.h, HS.h, G.h, H.h, _iptr(n)))
        return HS, self._per_inst(n)

    def transform_bch(self, Op, G):
        """Op <- e^G Op e^-G in place (transform_sq_op_BCH); the nested-commutator count"""
        n = np.zeros(self.nbatch, np.int32)
        _chk(_lib.imsrg_transform_bch(Op.h, G.h, _iptr(n)))
        return self._per_inst(n)

    def dhds(self, HS, kind="white", DH=None):
        """DH = [eta(HS), HS] with the generator `kind` built from HS (dHds)"""
        DH = DH or Operator(self, HS.herm)
        _chk(_lib.imsrg_dhds(HS.h, GEN[kind], DH.h))
        return DH

    def magnus_expand(self, DG, G):
        n = np.zeros(self.nbatch, np.int32)
        _chk(_lib.imsrg_magnus_expand(self.h, DG.h, G.h, _iptr(n)))
        return DG, self._per_inst(n)

    def magnus_flow(self, HS, kind="white", ds=0.5, max_steps=200, G=None):
        G = G or Operator(self, -1)
        s = np.zeros(max_steps); e0 = np.zeros(max_steps); emp2 = np.zeros(max_steps)
        crit = np.zeros(max_steps)
        nm = np.zeros(max_steps, np.int32); nb = np.zeros(max_steps, np.int32)
        steps = C.c_int(0)
        _chk(_lib.imsrg_magnus_flow(self.h, HS.h, G.h, GEN[kind], ds, max_steps, _dptr(s), _dptr(e0),
                                    _dptr(emp2), _dptr(crit), _iptr(nm), _iptr(nb), C.byref(steps)))
        n = steps.value
        return dict(steps=n, s=s[:n], e0=e0[:n], e0_mbpt2=emp2[:n], crit=crit[:n], n_magnus=nm[:n],
                    n_bch=nb[:n], HS=HS, G=G)

    def flow(self, HS, kind="white", ds=0.5, G=None, restart_s=None):
        return Flow(self, HS, kind, ds, G, restart_s)

    def profile(self, enable=True, classes=None):
        """time every kernel class with HIP events, or only those named in classes"""
        if enable and classes is not None:
            mask = 0
            for name in classes:
                mask |= 1 << PROF_CLASSES.index(name)
            _chk(_lib.imsrg_ctx_profile_classes(self.h, mask))
        else:
            _chk(_lib.imsrg_ctx_profile(self.h, 1 if enable else 0))

    def profile_read(self, reset=True):
        """{kernel class: (total ms, launches)} since the last reset."""
        n = 16
        ms = np.zeros(n)
        cnt = (C.c_longlong * n)()
        _chk(_lib.imsrg_ctx_profile_read(self.h, _dptr(ms), cnt, n, 1 if reset else 0))
        return {name: (float(ms[k]), int(cnt[k])) for k, name in enumerate(PROF_CLASSES)}

    def cross_coupled(self, A):
        n = int(_lib.imsrg_cross_coupled_size(self.h))
        out = np.zeros(n)
        _chk(_lib.imsrg_cross_coupled(A.h, _dptr(out), n))
        return out


class TensorOp:
    """A device-resident rank > 0 sq_op (imsrg_tensor)."""

    def __init__(self, ctx: Context, rank, dpar=0, herm=1):
        self.ctx, self.rank, self.dpar = ctx, int(rank), int(dpar)
        h = _vp()
        _chk(_lib.imsrg_tensor_create(ctx.h, self.rank, self.dpar, int(herm), C.byref(h)))
        self.h = h
        self.neq = int(_lib.imsrg_tensor_neq(h))

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                _lib.imsrg_tensor_destroy(self.h)
        except Exception:
            pass

    @property
    def herm(self):
        v = C.c_int(0)
        _chk(_lib.imsrg_tensor_get_herm(self.h, C.byref(v)))
        return v.value

    @herm.setter
    def herm(self, h):
        _chk(_lib.imsrg_tensor_set_herm(self.h, int(h)))

    def upload(self, vec):
        v = np.ascontiguousarray(vec, dtype=np.float64)
        _chk(_lib.imsrg_tensor_upload(self.h, _dptr(v), v.size))
        return self

    def download(self):
        v = np.zeros(self.neq)
        _chk(_lib.imsrg_tensor_download(self.h, _dptr(v), v.size))
        return v

    def copy_from(self, other):
        _chk(_lib.imsrg_tensor_copy(other.h, self.h))
        return self

    def clear(self):
        _chk(_lib.imsrg_tensor_clear(self.h))
        return self

    def add(self, ax, B, bx, out=None):
        out = out or self
        _chk(_lib.imsrg_tensor_add(self.h, float(ax), B.h, float(bx), out.h))
        return out

    def norm(self):
        v = C.c_double(0.0)
        _chk(_lib.imsrg_tensor_norm(self.h, C.byref(v)))
        return v.value


class Operator:
    """A device-resident sq_op (imsrg_op)."""

    def __init__(self, ctx: Context, herm=1):
        self.ctx = ctx
        h = _vp()
        _chk(_lib.imsrg_op_create(ctx.h, int(herm), C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                _lib.imsrg_op_destroy(self.h)
        except Exception:
            pass

    @property
    def herm(self):
        v = C.c_int(0)
        _chk(_lib.imsrg_op_get_herm(self.h, C.byref(v)))
        return v.value

    @herm.setter
    def herm(self, h):
        _chk(_lib.imsrg_op_set_herm(self.h, int(h)))

    def upload(self, vec):
        """vectorize-order vector, or (nbatch, neq) array for a batched context"""
        v = np.ascontiguousarray(vec, dtype=np.float64)
        _chk(_lib.imsrg_op_upload(self.h, _dptr(v), v.size))
        return self

    def write_stream(self, path):
        """the reference's .bin operator file: the vectorize stream, no header"""
        _chk(_lib.imsrg_op_write_stream(self.h, os.fsencode(path)))

    def read_stream(self, path):
        _chk(_lib.imsrg_op_read_stream(self.h, os.fsencode(path)))
        return self

    def download(self):
        B = self.ctx.nbatch
        v = np.zeros(self.ctx.neq * B)
        _chk(_lib.imsrg_op_download(self.h, _dptr(v), v.size))
        return v if B == 1 else v.reshape(B, self.ctx.neq)

    def _scalars(self, fn):
        v = np.zeros(self.ctx.nbatch)
        _chk(fn(self.h, _dptr(v)))
        return float(v[0]) if self.ctx.nbatch == 1 else v

    def copy_from(self, other):
        _chk(_lib.imsrg_op_copy(other.h, self.h))
        return self

    def clear(self):
        _chk(_lib.imsrg_op_clear(self.h))

    def scale(self, a):
        _chk(_lib.imsrg_op_scale(self.h, float(a)))

    def add(self, ax, B, bx, out=None):
        out = out or self
        _chk(_lib.imsrg_op_add(self.h, float(ax), B.h, float(bx), out.h))
        return out

    @property
    def e0(self):
        return self._scalars(_lib.imsrg_op_e0_batch)

    @e0.setter
    def e0(self, x):
        _chk(_lib.imsrg_op_set_e0(self.h, float(x)))

    def norm(self):
        return self._scalars(_lib.imsrg_op_norm_batch)

    def mbpt2(self):
        return self._scalars(_lib.imsrg_mbpt2_batch)

    def normal_order(self):
        _chk(_lib.imsrg_normal_order(self.h))
        return self

    def hartree_fock(self):
        """calc_HF on this bare Hamiltonian, in place: it becomes the normal-ordered
        Hamiltonian in the Hartree-Fock basis.  Returns (D, E_HF, iterations),
        D[a, i] = component of old orbit a in new orbit i."""
        n = self.ctx.basis.nsp
        D = np.zeros((n, n))
        it = np.zeros(1, np.int32)
        e = np.zeros(1)
        _chk(_lib.imsrg_hartree_fock(self.h, _dptr(D), _iptr(it), _dptr(e)))
        return D, float(e[0]), int(it[0])

    def to_hf(self, D):
        """sqop_to_HF on this bare scalar observable, in place, with coefficients D"""
        D = np.ascontiguousarray(D, dtype=np.float64)
        n = self.ctx.basis.nsp
        if D.shape != (n, n):
            raise ValueError("D must be nsp x nsp")
        _chk(_lib.imsrg_observable_to_hf(self.h, _dptr(D)))
        return self

    def read_interaction(self, fmt, path, path_1b=None, hw=0.0, com_hw=0.0, lawson_beta=0.0, A=0, htype=0,
                         emax_file=0, lmax_file=0):
        """set the two-body matrix elements of an interaction file (interaction_IO.f90
        formats: ascii, gz, me2j, me2b) into this hermitian operator, in place"""
        p = InteractionParams(hw, com_hw, lawson_beta, A, htype, emax_file, lmax_file)
        _chk(_lib.imsrg_read_interaction(self.h, INTERACTION_FORMATS[fmt], str(path).encode(),
                                         str(path_1b).encode() if path_1b else None, C.byref(p)))
        return self

    def normal_order_adjoint(self):
        """transpose of normal ordering applied to this cotangent (a new operator)"""
        out = Operator(self.ctx, self.herm)
        _chk(_lib.imsrg_normal_order_adjoint(self.h, out.h))
        return out

    def download_cotangent(self):
        """this full-layout cotangent as a gradient with respect to the vectorize stream"""
        B = self.ctx.nbatch
        v = np.zeros(self.ctx.neq * B)
        _chk(_lib.imsrg_op_download_cotangent(self.h, _dptr(v), v.size))
        return v if B == 1 else v.reshape(B, self.ctx.neq)


def read_interaction_vec(basis, fmt, path, path_1b=None, vec=None, hw=0.0, com_hw=0.0, lawson_beta=0.0, A=0,
                         htype=0, emax_file=0, lmax_file=0):
    """an interaction file (interaction_IO.f90 formats: ascii, gz, me2j, me2b) read
    on the host into a vectorize stream of `basis` (zeros, or a copy of vec)"""
    n, l, jj, tz, con = (np.ascontiguousarray(x, dtype=np.int32) for x in basis.orbit_table().T)
    v = np.zeros(basis.neq) if vec is None else np.array(vec, dtype=np.float64)
    p = InteractionParams(hw, com_hw, lawson_beta, A, htype, emax_file, lmax_file)
    _chk(_lib.imsrg_read_interaction_vec(basis.nsp, _iptr(n), _iptr(l), _iptr(jj), _iptr(tz), _iptr(con),
                                         INTERACTION_FORMATS[fmt], str(path).encode(),
                                         str(path_1b).encode() if path_1b else None, C.byref(p), _dptr(v), v.size))
    return v


def commutator_110(L, R):
    v = C.c_double(0)
    _chk(_lib.imsrg_commutator_110(L.h, R.h, C.byref(v)))
    return v.value


def commutator_220(L, R):
    v = C.c_double(0)
    _chk(_lib.imsrg_commutator_220(L.h, R.h, C.byref(v)))
    return v.value


class Flow:
    """magnus_decouple one step at a time (imsrg_flow_*)."""

    def __init__(self, ctx, HS, kind="white", ds=0.5, G=None, restart_s=None):
        """HS = normal-ordered H(0).  With restart_s, G holds Omega(restart_s) (a
        checkpoint) and the flow resumes there (imsrg_flow_restart)."""
        self.ctx = ctx
        self.HS = HS
        self.G = G or Operator(ctx, -1)
        h = _vp()
        if restart_s is None:
            _chk(_lib.imsrg_flow_create(ctx.h, HS.h, self.G.h, GEN[kind], float(ds), C.byref(h)))
        else:
            _chk(_lib.imsrg_flow_restart(ctx.h, HS.h, self.G.h, GEN[kind], float(ds), float(restart_s),
                                         C.byref(h)))
        self.h = h

    def step(self):
        """one Magnus step; a record dict (a list of them, one per instance, if batched)"""
        B = self.ctx.nbatch
        r = np.zeros(8 * B)
        _chk(_lib.imsrg_flow_step(self.h, _dptr(r)))
        recs = [dict(s=x[0], e0=x[1], e0_mbpt2=x[2], crit=x[3], n_magnus=int(x[4]), n_bch=int(x[5]),
                     step=int(x[6]), stepped=bool(x[7])) for x in r.reshape(B, 8)]
        return recs[0] if B == 1 else recs

    @property
    def converged(self):
        v = C.c_int(0)
        _chk(_lib.imsrg_flow_converged(self.h, C.byref(v)))
        return bool(v.value)

    def close(self):
        if getattr(self, "h", None):
            _lib.imsrg_flow_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
