"""ORACLE - test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker of the HIP apply.  It wraps
oracle/mw_oracle.cpp, a plain FP64 OpenMP restatement of mrcpp::apply (see the
header of that file for the reference call sites it follows and the parity
status of each part), and oracle/closed_forms.py (Appendix-B closed forms of
Gaussian Coulomb and Yukawa interactions).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _cpu_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = line.split()
                    return all(x in fl for x in ("avx512f", "avx512bw", "avx512vl", "avx512dq"))
    except OSError:
        pass
    return False


def _load():
    tags = ["v4", "v3"] if _cpu_has_avx512() else ["v3"]
    for t in tags:
        p = os.path.join(HERE, "libmworacle_%s.so" % t)
        if os.path.exists(p):
            return ctypes.CDLL(p), p
    raise ImportError("oracle library not built; run __graft_entry__.build()")


_i, _d, _p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_llp = ctypes.POINTER(ctypes.c_longlong)

_LIB = None
LIB_PATH = None


def _lib():
    """Load the oracle library on first use (so that building it needs no import)."""
    global _LIB, LIB_PATH
    if _LIB is not None:
        return _LIB
    lib, LIB_PATH = _load()
    _declare(lib)
    _LIB = lib
    return lib


def _declare(lib):
    lib.oracle_multiply.argtypes = [_i, _i, _ip, _ip, _dp, _dp, _i, _ip, _llp, _dp, _i, _ip, _llp, _dp, _i, _ip, _llp,
                                    _dp, _d, _dp]
    lib.oracle_multiply.restype = _i
    lib.oracle_set_threads.argtypes = [_i]
    lib.oracle_set_threads.restype = None
    lib.oracle_apply.argtypes = [_i, _i, _ip, _ip, _dp, _i, _i, _ip, _llp, _llp, _dp, _dp, _i, _ip, _llp, _dp, _i, _ip,
                                 _llp, _dp, _i, _i, _dp, _dp, _dp]
    lib.oracle_apply.restype = _i
    lib.oracle_apply_adaptive.argtypes = [_i, _i, _ip, _ip, _dp, _i, _i, _ip, _llp, _llp, _dp, _dp, _i, _ip, _llp,
                                          _dp, _d, _i, _i, _i, _dp]
    lib.oracle_apply_adaptive.restype = _p
    lib.oracle_result_size.argtypes = [_p]
    lib.oracle_result_size.restype = _i
    lib.oracle_result_get.argtypes = [_p, _ip, _llp, _dp, _dp]
    lib.oracle_result_get.restype = None
    lib.oracle_result_free.argtypes = [_p]
    lib.oracle_result_free.restype = None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def set_threads(n):
    _lib().oracle_set_threads(int(n))


def _common(mra, tables, in_tree):
    corner = _c(mra.corner, np.int32)
    boxes = _c(mra.boxes, np.int32)
    F = _c(mra.filter(), np.float64)
    dep, l, coefs = in_tree
    return (corner, boxes, F, _c(tables["L"], np.int32), _c(tables["boff"], np.int64), _c(tables["noff"], np.int64),
            _c(tables["blocks"], np.float64), _c(tables["norms"], np.float64), _c(dep, np.int32), _c(l, np.int64),
            _c(coefs, np.float64))


def apply_fixed(mra, tables, in_tree, out_depth, out_l, thr, band_w=-1, want_ns=False, post=True):
    """Frozen-grid FP64 apply.  in_tree = (depth, l, coefs) of a complete
    compressed input tree; returns (out_blocks, stats[, ns_blocks]).  With
    post=False the output nodes may be any sample (no post transforms; the
    returned blocks are the raw non-standard contributions)."""
    corner, boxes, F, L, boff, noff, blocks, norms, dep, l, coefs = _common(mra, tables, in_tree)
    od = _c(out_depth, np.int32)
    ol = _c(out_l, np.int64)
    th = _c(thr, np.float64)
    n_out = len(od)
    out = np.zeros((n_out, mra.block))
    ns = np.zeros((n_out, mra.block)) if want_ns else None
    stats = np.zeros(5)
    lib = _lib()
    rc = lib.oracle_apply(mra.kp1, mra.min_scale, _ptr(corner, _i), _ptr(boxes, _i), _ptr(F, _d), tables["M"],
                          tables["ndepth"], _ptr(L, _i), _ptr(boff, ctypes.c_longlong),
                          _ptr(noff, ctypes.c_longlong), _ptr(blocks, _d), _ptr(norms, _d), len(dep), _ptr(dep, _i),
                          _ptr(l, ctypes.c_longlong), _ptr(coefs, _d), n_out, _ptr(od, _i),
                          _ptr(ol, ctypes.c_longlong), _ptr(th, _d), band_w, 0 if post else 1,
                          _ptr(ns, _d) if want_ns else None, _ptr(out, _d), _ptr(stats, _d))
    if rc != 0:
        raise RuntimeError("oracle_apply failed with %d" % rc)
    st = {"term_apps": stats[0], "candidates": stats[1], "slots": stats[2], "generated": stats[3],
          "band_w": int(stats[4])}
    return (out, st, ns) if want_ns else (out, st)


def multiply(mra, a_tree, b_tree, out_depth, out_l, coef=1.0):
    """coef * a * b at the given output nodes (interpolating basis): each
    node's compressed block from the product at its 8 children's points.
    a_tree, b_tree = (depth, l, coefs) of complete compressed trees."""
    corner = _c(mra.corner, np.int32)
    boxes = _c(mra.boxes, np.int32)
    F = _c(mra.filter(), np.float64)
    _, w = np.polynomial.legendre.leggauss(mra.kp1)
    wq = _c(w / 2.0, np.float64)
    ad, al, ac = (_c(a_tree[0], np.int32), _c(a_tree[1], np.int64), _c(a_tree[2], np.float64))
    bd, bl, bc = (_c(b_tree[0], np.int32), _c(b_tree[1], np.int64), _c(b_tree[2], np.float64))
    od = _c(out_depth, np.int32)
    ol = _c(out_l, np.int64)
    hfac = _c([mra.box_size(int(d) + 1) ** -1.5 for d in od], np.float64)
    out = np.zeros((len(od), mra.block))
    bad = _lib().oracle_multiply(mra.kp1, mra.min_scale, _ptr(corner, _i), _ptr(boxes, _i), _ptr(F, _d), _ptr(wq, _d),
                                 len(ad), _ptr(ad, _i), _ptr(al, ctypes.c_longlong), _ptr(ac, _d), len(bd),
                                 _ptr(bd, _i), _ptr(bl, ctypes.c_longlong), _ptr(bc, _d), len(od), _ptr(od, _i),
                                 _ptr(ol, ctypes.c_longlong), _ptr(hfac, _d), float(coef), _ptr(out, _d))
    if bad:
        raise RuntimeError("oracle_multiply: %d output nodes outside the inputs' trees" % bad)
    return out


def apply_adaptive(mra, tables, in_tree, prec, abs_prec=False, max_iter=-1, max_depth=-1):
    """Adaptive FP64 apply from the roots.  Returns (depth, l, thr, blocks, stats)."""
    corner, boxes, F, L, boff, noff, blocks, norms, dep, l, coefs = _common(mra, tables, in_tree)
    stats = np.zeros(5)
    lib = _lib()
    h = lib.oracle_apply_adaptive(mra.kp1, mra.min_scale, _ptr(corner, _i), _ptr(boxes, _i), _ptr(F, _d),
                                  tables["M"], tables["ndepth"], _ptr(L, _i), _ptr(boff, ctypes.c_longlong),
                                  _ptr(noff, ctypes.c_longlong), _ptr(blocks, _d), _ptr(norms, _d), len(dep),
                                  _ptr(dep, _i), _ptr(l, ctypes.c_longlong), _ptr(coefs, _d), prec, int(abs_prec),
                                  max_iter, max_depth, _ptr(stats, _d))
    n = lib.oracle_result_size(h)
    od = np.zeros(n, dtype=np.int32)
    ol = np.zeros((n, 3), dtype=np.int64)
    th = np.zeros(n)
    oc = np.zeros((n, mra.block))
    lib.oracle_result_get(h, _ptr(od, _i), _ptr(ol, ctypes.c_longlong), _ptr(th, _d), _ptr(oc, _d))
    lib.oracle_result_free(h)
    st = {"term_apps": stats[0], "candidates": stats[1], "slots": stats[2], "generated": stats[3],
          "band_w": int(stats[4])}
    return od, ol, th, oc, st
