"""mrchem_amd: MI355X-native multiwavelet Green's-function apply.

Host-side mirror of the MRCPP interface MRChem uses for its Green's-function
applications, over the C ABI in include/mwgreen.h (libmwgreen.so):

    MultiResolutionAnalysis   mrcpp::MultiResolutionAnalysis<3>   (src/mrenv.cpp:99-127)
    PoissonOperator           mrcpp::PoissonOperator(MRA, prec)    (src/driver.cpp:1131)
    HelmholtzOperator         mrcpp::HelmholtzOperator(MRA, mu, prec)
                                                   (src/scf_solver/HelmholtzVector.cpp:144)
    FunctionTree              mrcpp::FunctionTree<3>
    apply(prec, out, oper, inp, max_iter=-1, abs_prec=False, prec_trees=None)
                              mrcpp::apply            (CoulombPotential.cpp:131,
                                                       ExchangePotential.cpp:200,
                                                       HelmholtzVector.cpp:149)

Every compute call runs the HIP kernels; there is no CPU fallback.  Importing
this module fails when libmwgreen.so has not been built.
"""
import ctypes
import os

import numpy as np

__all__ = ["MultiResolutionAnalysis", "PoissonOperator", "HelmholtzOperator", "GaussianOperator",
           "FunctionTree", "Context", "ApplyPlan", "DeviceBuffer", "apply", "apply_fixed", "MWGError", "lib", "LIB_PATH"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmwgreen.so")
if not os.path.exists(LIB_PATH):
    raise ImportError("libmwgreen.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
lib = ctypes.CDLL(LIB_PATH)

_i = ctypes.c_int
_d = ctypes.c_double
_ll = ctypes.c_longlong
_p = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_llp = ctypes.POINTER(ctypes.c_longlong)


def _sig(name, *args):
    f = getattr(lib, name)
    f.argtypes = list(args)
    f.restype = ctypes.c_int
    return f


lib.mwg_last_error.restype = ctypes.c_char_p
lib.mwg_last_error.argtypes = []
_sig("mwg_context_create", _i, ctypes.POINTER(_p))
_sig("mwg_context_destroy", _p)
_sig("mwg_mra_create", _i, _i, _i, _ip, _ip, _d, _i, ctypes.POINTER(_p))
_sig("mwg_mra_destroy", _p)
_sig("mwg_mra_filter", _p, _dp)
_sig("mwg_poisson_create", _p, _d, ctypes.POINTER(_p))
_sig("mwg_helmholtz_create", _p, _d, _d, ctypes.POINTER(_p))
_sig("mwg_gaussian_operator_create", _p, _i, _dp, _dp, _i, ctypes.POINTER(_p))
_sig("mwg_operator_destroy", _p)
_sig("mwg_operators_create_device", _p, _p, _i, _dp, _d, ctypes.POINTER(_p))
_sig("mwg_multiply", _p, _d, _p, _d, _p, _p, _i, _i)
_sig("mwg_multiply_noalloc", _p, _d, _p, _d, _p, _p, _i, _i)
_sig("mwg_multiply_ex", _p, _d, _p, _d, _p, _p, _i, _i, _i)
_sig("mwg_add", _p, _d, _p, _i, _dp, ctypes.POINTER(_p), _i, _i)
_sig("mwg_tree_add", _p, _p, _d, _p)
_sig("mwg_tree_crop", _p, _d, _d, _i)
_sig("mwg_tree_pack_size", _p, _llp)
_sig("mwg_tree_pack", _p, ctypes.c_void_p, ctypes.c_longlong)
_sig("mwg_tree_unpack", _p, ctypes.c_void_p, ctypes.c_longlong)
_sig("mwg_tree_pack_device", _p, _p, ctypes.c_void_p, ctypes.c_longlong)
_sig("mwg_tree_unpack_device", _p, _p, ctypes.c_void_p, ctypes.c_longlong)
_sig("mwg_device_buffer_create", _p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_void_p))
_sig("mwg_device_buffer_destroy", _p, ctypes.c_void_p)
_sig("mwg_device_buffer_read", _p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong)
_sig("mwg_tree_save", _p, ctypes.c_char_p)
_sig("mwg_tree_load", _p, ctypes.c_char_p)
_sig("mwg_operator_info", _p, _ip, _ip, _ip, _dp, _dp, _dp)
_sig("mwg_operator_terms", _p, _dp, _dp)
_sig("mwg_operator_table_sizes", _p, _llp, _llp)
_sig("mwg_operator_tables", _p, _ip, _llp, _llp, _dp, _dp)
_sig("mwg_tree_create", _p, ctypes.POINTER(_p))
_sig("mwg_tree_destroy", _p)
_sig("mwg_tree_set_resident", _p, _p, _i)
_sig("mwg_tree_is_resident", _p, _ip, _ip)
_sig("mwg_context_transfer_bytes", _p, _llp, _llp, _i)
_sig("mwg_tree_project_gaussians", _p, _d, _i, _i, _i, _dp, _dp, _dp)
_sig("mwg_tree_size", _p, _ip)
_sig("mwg_tree_get", _p, _ip, _llp, _dp)
_sig("mwg_tree_set", _p, _i, _ip, _llp, _dp)
_sig("mwg_tree_square_norm", _p, _dp)
_sig("mwg_tree_dot", _p, _p, _dp)
_sig("mwg_tree_eval", _p, _i, _dp, _dp)
_sig("mwg_tree_rescale", _p, _d)
_sig("mwg_tree_thresholds", _p, _dp)
_sig("mwg_apply", _p, _d, _p, _p, _p, _i, _i)
_sig("mwg_apply_prec_trees", _p, _d, _p, _p, _p, ctypes.POINTER(_p), _i, _i, _i)
_sig("mwg_apply_fixed", _p, _p, _p, _p, _dp, _i)
_sig("mwg_last_apply_stats", _p, _dp, _i)
_sig("mwg_plan_create", _p, _p, _p, _p, _dp, _i, ctypes.POINTER(_p))
_sig("mwg_plan_create_batch", _p, _i, ctypes.POINTER(_p), ctypes.POINTER(_p), ctypes.POINTER(_p),
     ctypes.POINTER(_dp), _i, ctypes.POINTER(_p))


class _ApplyDesc(ctypes.Structure):
    """mwg_apply_desc of include/mwgreen.h."""
    _fields_ = [("out", _p), ("op", _p), ("inp", _p), ("prec", _d), ("max_iter", _i), ("abs_prec", _i),
                ("prec_trees", ctypes.POINTER(_p)), ("n_prec", _i)]


_sig("mwg_apply_batch", _p, _i, ctypes.POINTER(_ApplyDesc))
_sig("mwg_plan_execute", _p, _dp, _i)
_sig("mwg_plan_apply", _p, _dp, _i)
_sig("mwg_plan_use_graph", _p, _i)
_sig("mwg_plan_tune", _p, _d, _i, _dp, _i)
_sig("mwg_context_set_precision", _p, _i, _d)
_sig("mwg_plan_finish", _p)
_sig("mwg_plan_info", _p, _dp, _i)
_sig("mwg_plan_download_ns", _p, _dp, _ll)
lib.mwg_context_last_error.restype = ctypes.c_char_p
lib.mwg_context_last_error.argtypes = [_p]
_sig("mwg_plan_destroy", _p)

MODES = {"fp64": 0, "mixed": 1, "split16": 2}


class MWGError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise MWGError("mwgreen error %d: %s" % (rc, lib.mwg_last_error().decode()))


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class Context:
    """One HIP device and stream (the OpenMP environment of the reference apply)."""

    _default = None

    def __init__(self, device=0):
        h = _p()
        _check(lib.mwg_context_create(device, ctypes.byref(h)))
        self._h = h
        self.mode, self.tau = "fp64", 65536.0

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    def set_precision(self, mode="fp64", tau=65536.0):
        """Arithmetic of mw.apply on this context: 'fp64', 'mixed' (FP64 for
        terms whose bound exceeds tau x threshold, FP32 MFMA for the rest) or
        'split16' (the rest as FP16 hi/lo split products on FP16 MFMA)."""
        _check(lib.mwg_context_set_precision(self._h, MODES[mode], tau))
        self.mode, self.tau = mode, float(tau)

    def last_error(self):
        """Message of the last failed call on this context ("" if none)."""
        return lib.mwg_context_last_error(self._h).decode()

    def transfer_bytes(self, reset=False):
        """(host-to-device, device-to-host) bytes of tree blocks moved by calls
        on this context since it was made or last reset."""
        a, b = ctypes.c_longlong(0), ctypes.c_longlong(0)
        _check(lib.mwg_context_transfer_bytes(self._h, ctypes.byref(a), ctypes.byref(b), int(reset)))
        return a.value, b.value

    def last_stats(self):
        s = np.zeros(16)
        _check(lib.mwg_last_apply_stats(self._h, _ptr(s, _d), 16))
        return {"kernel_ms": s[0], "out_nodes": int(s[1]), "in_slots": int(s[2]), "generated": int(s[3]),
                "plan_ms": s[4], "launches": int(s[5]), "term_apps": int(s[6]), "device_ms": s[7],
                "split_ms": s[8], "post_ms": s[9], "wall_ms": s[10], "device_create_ms": s[11],
                "device_run_ms": s[12], "plan_grid_ms": s[13], "plan_pool_ms": s[14], "plan_counts_ms": s[15]}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.mwg_context_destroy(h)
            self._h = None


class DeviceBuffer:
    """Device memory on a context's GPU (a packed-tree send or receive
    buffer); ptr is its device address."""

    def __init__(self, nbytes, ctx=None):
        self.ctx = ctx or Context.default()
        p = ctypes.c_void_p()
        _check(lib.mwg_device_buffer_create(self.ctx._h, nbytes, ctypes.byref(p)))
        self.ptr, self.nbytes = p.value, nbytes

    def read(self):
        """The buffer's bytes, copied to the host."""
        out = ctypes.create_string_buffer(self.nbytes)
        _check(lib.mwg_device_buffer_read(self.ctx._h, ctypes.c_void_p(self.ptr), out, self.nbytes))
        return out.raw

    def __del__(self):
        if getattr(self, "ptr", None):
            lib.mwg_device_buffer_destroy(self.ctx._h, ctypes.c_void_p(self.ptr))
            self.ptr = None


class MultiResolutionAnalysis:
    """MRA: BoundingBox(min_scale, corner, boxes) x scaling factor, basis order/type, max depth."""

    def __init__(self, order, basis="interpolating", min_scale=-4, corner=(-1, -1, -1), boxes=(2, 2, 2),
                 scaling_factor=1.0, max_depth=20):
        btype = {"interpolating": 0, "legendre": 1}[basis]
        c = np.ascontiguousarray(corner, dtype=np.int32)
        b = np.ascontiguousarray(boxes, dtype=np.int32)
        h = _p()
        _check(lib.mwg_mra_create(order, btype, min_scale, _ptr(c, _i), _ptr(b, _i), scaling_factor, max_depth,
                                  ctypes.byref(h)))
        self._h = h
        self.order, self.basis, self.min_scale = order, basis, min_scale
        self.corner, self.boxes = tuple(corner), tuple(boxes)
        self.scaling_factor, self.max_depth = scaling_factor, max_depth
        self.kp1 = order + 1

    @property
    def block(self):
        return 8 * self.kp1 ** 3

    def filter(self):
        F = np.zeros((2 * self.kp1, 2 * self.kp1))
        _check(lib.mwg_mra_filter(self._h, _ptr(F, _d)))
        return F

    def box_size(self, depth):
        return self.scaling_factor * 2.0 ** (-(self.min_scale + depth))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.mwg_mra_destroy(h)
            self._h = None


class _Operator:
    def __init__(self, mra, h):
        self.mra = mra
        self._h = h

    def info(self):
        n, nd, kp1 = _i(), _i(), _i()
        rmin, rmax, err = _d(), _d(), _d()
        _check(lib.mwg_operator_info(self._h, ctypes.byref(n), ctypes.byref(nd), ctypes.byref(kp1),
                                     ctypes.byref(rmin), ctypes.byref(rmax), ctypes.byref(err)))
        return {"n_terms": n.value, "n_depth": nd.value, "kp1": kp1.value, "r_min": rmin.value,
                "r_max": rmax.value, "fit_err": err.value}

    def size(self):
        return self.info()["n_terms"]

    def terms(self):
        n = self.size()
        a, b = np.zeros(n), np.zeros(n)
        _check(lib.mwg_operator_terms(self._h, _ptr(a, _d), _ptr(b, _d)))
        return a, b

    def tables(self):
        info = self.info()
        nb, nn = _ll(), _ll()
        _check(lib.mwg_operator_table_sizes(self._h, ctypes.byref(nb), ctypes.byref(nn)))
        k = info["n_terms"] * info["n_depth"]
        L = np.zeros(k, dtype=np.int32)
        boff = np.zeros(k, dtype=np.int64)
        noff = np.zeros(k, dtype=np.int64)
        blocks = np.zeros(max(nb.value, 1))
        norms = np.zeros(max(nn.value, 1))
        _check(lib.mwg_operator_tables(self._h, _ptr(L, _i), _ptr(boff, _ll), _ptr(noff, _ll), _ptr(blocks, _d),
                                       _ptr(norms, _d)))
        return {"M": info["n_terms"], "ndepth": info["n_depth"], "L": L, "boff": boff, "noff": noff,
                "blocks": blocks, "norms": norms}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.mwg_operator_destroy(h)
            self._h = None


class PoissonOperator(_Operator):
    """Separated-Gaussian NS operator of the kernel 1/r."""

    def __init__(self, mra, prec):
        h = _p()
        _check(lib.mwg_poisson_create(mra._h, prec, ctypes.byref(h)))
        super().__init__(mra, h)
        self.prec = prec


class HelmholtzOperator(_Operator):
    """Separated-Gaussian NS operator of the bound-state kernel exp(-mu r)/r."""

    def __init__(self, mra, mu, prec):
        h = _p()
        _check(lib.mwg_helmholtz_create(mra._h, mu, prec, ctypes.byref(h)))
        super().__init__(mra, h)
        self.mu, self.prec = mu, prec


def build_operators(mra, mus, prec, ctx=None):
    """PoissonOperator (mu = 0) or HelmholtzOperator (mu > 0) for every mu,
    built together on the device of ctx (HelmholtzVector.cpp:144 builds one
    per orbital in every SCF iteration).  Same kernel fit and tables as the
    host constructors; the tables stay on the device and the operators work
    only with ctx."""
    ctx = ctx or Context.default()
    mus = np.ascontiguousarray(np.atleast_1d(mus), dtype=np.float64)
    hs = (_p * len(mus))()
    _check(lib.mwg_operators_create_device(ctx._h, mra._h, len(mus), _ptr(mus, _d), float(prec), hs))
    ops = []
    for i, mu in enumerate(mus):
        op = _Operator(mra, _p(hs[i]))
        op.mu, op.prec, op.ctx = float(mu), float(prec), ctx
        ops.append(op)
    return ops


class GaussianOperator(_Operator):
    """NS operator of an explicit kernel sum_m beta_m exp(-alpha_m r^2)."""

    def __init__(self, mra, alpha, beta, max_depth=-1):
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        b = np.ascontiguousarray(beta, dtype=np.float64)
        h = _p()
        _check(lib.mwg_gaussian_operator_create(mra._h, len(a), _ptr(a, _d), _ptr(b, _d), max_depth,
                                                ctypes.byref(h)))
        super().__init__(mra, h)


class FunctionTree:
    """Adaptive 2^3-tree of compressed 8(k+1)^3 blocks (mrcpp::FunctionTree<3>)."""

    def __init__(self, mra):
        h = _p()
        _check(lib.mwg_tree_create(mra._h, ctypes.byref(h)))
        self._h = h
        self.mra = mra

    def project_gaussians(self, prec, coef, exponent, center, abs_prec=False, max_depth=-1):
        c = np.ascontiguousarray(coef, dtype=np.float64).ravel()
        e = np.ascontiguousarray(exponent, dtype=np.float64).ravel()
        r = np.ascontiguousarray(center, dtype=np.float64).reshape(-1, 3)
        _check(lib.mwg_tree_project_gaussians(self._h, prec, int(abs_prec), max_depth, len(c), _ptr(c, _d),
                                              _ptr(e, _d), _ptr(r, _d)))
        return self

    def size(self):
        n = _i()
        _check(lib.mwg_tree_size(self._h, ctypes.byref(n)))
        return n.value

    def nodes(self):
        n = self.size()
        depth = np.zeros(n, dtype=np.int32)
        l = np.zeros((n, 3), dtype=np.int64)
        coefs = np.zeros((n, self.mra.block))
        _check(lib.mwg_tree_get(self._h, _ptr(depth, _i), _ptr(l, _ll), _ptr(coefs, _d)))
        return depth, l, coefs

    def set_nodes(self, depth, l, coefs=None):
        depth = np.ascontiguousarray(depth, dtype=np.int32)
        l = np.ascontiguousarray(l, dtype=np.int64).reshape(-1, 3)
        n = len(depth)
        if coefs is None:
            coefs = np.zeros((n, self.mra.block))
        coefs = np.ascontiguousarray(coefs, dtype=np.float64).reshape(n, self.mra.block)
        _check(lib.mwg_tree_set(self._h, n, _ptr(depth, _i), _ptr(l, _ll), _ptr(coefs, _d)))
        return self

    def square_norm(self):
        s = _d()
        _check(lib.mwg_tree_square_norm(self._h, ctypes.byref(s)))
        return s.value

    def norm(self):
        return float(np.sqrt(max(self.square_norm(), 0.0)))

    def dot(self, other):
        s = _d()
        _check(lib.mwg_tree_dot(self._h, other._h, ctypes.byref(s)))
        return s.value

    def evaluate(self, points):
        r = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        v = np.zeros(len(r))
        _check(lib.mwg_tree_eval(self._h, len(r), _ptr(r, _d), _ptr(v, _d)))
        return v

    def rescale(self, f):
        _check(lib.mwg_tree_rescale(self._h, f))
        return self

    def set_resident(self, ctx=None, on=True):
        """Keep the tree's blocks in HBM of ctx's GPU across calls (on=False:
        back to host memory).  Device calls on ctx read and write them there;
        host reads download them lazily."""
        ctx = ctx or Context.default()
        _check(lib.mwg_tree_set_resident(self._h, ctx._h, int(on)))
        # the context must outlive the device copy
        self._ctx = ctx if on else None
        return self

    def residency(self):
        """(resident, host copy current, device copy current)."""
        a, b = _i(), _i()
        r = lib.mwg_tree_is_resident(self._h, ctypes.byref(a), ctypes.byref(b))
        return bool(r), bool(a.value), bool(b.value)

    def add(self, c, inp, ctx=None):
        """self += c * inp on the union of the grids (FunctionTree::add; the
        node sums run on the device)."""
        ctx = ctx or Context.default()
        _check(lib.mwg_tree_add(ctx._h, self._h, c, inp._h))
        return self

    def pack(self):
        """The tree's packed form (bytes): the wire and file format."""
        n = ctypes.c_longlong(0)
        _check(lib.mwg_tree_pack_size(self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib.mwg_tree_pack(self._h, buf, n.value))
        return buf.raw

    def pack_size(self):
        n = ctypes.c_longlong(0)
        _check(lib.mwg_tree_pack_size(self._h, ctypes.byref(n)))
        return n.value

    def pack_device(self, ptr, nbytes, ctx=None):
        """Write the packed form into device memory at ptr (an address on
        ctx's GPU of at least pack_size() bytes)."""
        ctx = ctx or Context.default()
        _check(lib.mwg_tree_pack_device(self._h, ctx._h, ctypes.c_void_p(ptr), nbytes))
        return self

    def unpack_device(self, ptr, nbytes, ctx=None):
        """Replace the tree by the packed tree at device address ptr; the tree
        becomes resident on ctx."""
        ctx = ctx or Context.default()
        _check(lib.mwg_tree_unpack_device(self._h, ctx._h, ctypes.c_void_p(ptr), nbytes))
        self._ctx = ctx
        return self

    def unpack(self, data):
        """Replace the tree by a packed one written on the same MRA."""
        b = bytes(data)
        _check(lib.mwg_tree_unpack(self._h, b, len(b)))
        return self

    def save(self, path):
        _check(lib.mwg_tree_save(self._h, os.fsencode(path)))
        return self

    def load(self, path):
        _check(lib.mwg_tree_load(self._h, os.fsencode(path)))
        return self

    def crop(self, prec, split_fac=1.0, abs_prec=True):
        """FunctionTree::crop(prec, splitFac, absPrec): drop the children of
        branch nodes with small wavelet norms, bottom-up."""
        _check(lib.mwg_tree_crop(self._h, prec, split_fac, 1 if abs_prec else 0))
        return self

    def thresholds(self):
        thr = np.zeros(self.size())
        _check(lib.mwg_tree_thresholds(self._h, _ptr(thr, _d)))
        return thr

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.mwg_tree_destroy(h)
            self._h = None


def apply(prec, out, oper, inp, max_iter=-1, abs_prec=False, prec_trees=None, ctx=None):
    """mrcpp::apply(prec, out, oper, inp[, precTrees], maxIter, absPrec) on the GPU."""
    ctx = ctx or Context.default()
    if prec_trees:
        arr = (_p * len(prec_trees))(*[t._h for t in prec_trees])
        _check(lib.mwg_apply_prec_trees(ctx._h, prec, out._h, oper._h, inp._h, arr, len(prec_trees), max_iter,
                                        int(abs_prec)))
    else:
        _check(lib.mwg_apply(ctx._h, prec, out._h, oper._h, inp._h, max_iter, int(abs_prec)))
    return out


def multiply(prec, out, a, b, coef=1.0, max_iter=-1, abs_prec=False, ctx=None, noalloc=False, use_max_norms=False):
    """out = coef * a * b to prec on the device (mrcpp::multiply; interpolating
    basis).  noalloc: the variant that allocates no device memory (its
    buffers come from the library's fixed device array; inputs that do not
    fit are an error).  use_max_norms: the max-norm refinement from the
    roots (useMaxNorms)."""
    ctx = ctx or Context.default()
    if use_max_norms:
        if noalloc:
            raise ValueError("use_max_norms has no allocation-free form")
        _check(lib.mwg_multiply_ex(ctx._h, prec, out._h, coef, a._h, b._h, max_iter, 1 if abs_prec else 0, 1))
        return out
    f = lib.mwg_multiply_noalloc if noalloc else lib.mwg_multiply
    _check(f(ctx._h, prec, out._h, coef, a._h, b._h, max_iter, 1 if abs_prec else 0))
    return out


def add(prec, out, coefs, trees, max_iter=-1, abs_prec=False, ctx=None):
    """out = sum_i coefs[i] trees[i] (mrcpp::add / cplxfunc::linear_combination);
    prec < 0: on the union grid, else built adaptively to prec."""
    ctx = ctx or Context.default()
    c = np.ascontiguousarray(coefs, dtype=np.float64)
    if len(c) != len(trees):
        raise ValueError("one coefficient per tree")
    arr = (_p * len(trees))(*[t._h for t in trees])
    _check(lib.mwg_add(ctx._h, prec, out._h, len(trees), _ptr(c, _d), arr, max_iter, 1 if abs_prec else 0))
    return out


def apply_fixed(out, oper, inp, thr, mode="fp64", ctx=None):
    """Frozen-grid apply: out keeps its grid; thr[g] is node g's screening threshold."""
    ctx = ctx or Context.default()
    t = np.ascontiguousarray(thr, dtype=np.float64)
    if t.size != out.size():
        raise ValueError("one threshold per output node is required")
    _check(lib.mwg_apply_fixed(ctx._h, out._h, oper._h, inp._h, _ptr(t, _d), MODES[mode]))
    return out


def apply_batch(entries, ctx=None):
    """Many applies in shared launches (mwg_apply_batch).  Each entry is a dict
    with keys prec, out, oper, inp and optionally max_iter, abs_prec,
    prec_trees, with the meaning of apply()."""
    ctx = ctx or Context.default()
    n = len(entries)
    descs = (_ApplyDesc * max(n, 1))()
    keep = []
    for i, e in enumerate(entries):
        pt = e.get("prec_trees") or []
        arr = (_p * max(len(pt), 1))(*[t._h for t in pt])
        keep.append(arr)
        descs[i] = _ApplyDesc(e["out"]._h, e["oper"]._h, e["inp"]._h, float(e["prec"]), int(e.get("max_iter", -1)),
                              int(bool(e.get("abs_prec", False))), ctypes.cast(arr, ctypes.POINTER(_p)), len(pt))
    _check(lib.mwg_apply_batch(ctx._h, n, descs))
    return [e["out"] for e in entries]


class ApplyPlan:
    """Frozen-grid apply resident on the device; execute() reruns the kernel.

    out, oper, inp and thr may be lists (one entry per apply) for a batched
    plan whose applies run in the same launches."""

    def __init__(self, out, oper, inp, thr, mode="fp64", ctx=None):
        self.ctx = ctx or Context.default()
        batched = isinstance(out, (list, tuple))
        outs = list(out) if batched else [out]
        n = len(outs)
        opers = list(oper) if isinstance(oper, (list, tuple)) else [oper] * n
        inps = list(inp) if isinstance(inp, (list, tuple)) else [inp] * n
        thrs = list(thr) if batched else [thr]
        if not (len(opers) == len(inps) == len(thrs) == n) or n == 0:
            raise ValueError("one operator, input tree and threshold array per output tree")
        ts = []
        for o, t in zip(outs, thrs):
            a = np.ascontiguousarray(t, dtype=np.float64)
            if a.size != o.size():
                raise ValueError("one threshold per output node is required")
            ts.append(a)
        h = _p()
        arr_o = (_p * n)(*[o._h for o in outs])
        arr_p = (_p * n)(*[o._h for o in opers])
        arr_i = (_p * n)(*[o._h for o in inps])
        arr_t = (_dp * n)(*[_ptr(a, _d) for a in ts])
        _check(lib.mwg_plan_create_batch(self.ctx._h, n, arr_o, arr_p, arr_i, arr_t, MODES[mode], ctypes.byref(h)))
        self._h = h
        self._outs = outs
        self._keep = (outs, opers, inps, ts)

    def execute(self):
        """One execution; returns a dict: ms (device), term_apps, lp_term_apps
        (on the low-precision path: FP32 MFMA or FP16 split products), ms_fp64
        and ms_lp (device time of the two launches)."""
        v = np.zeros(5)
        _check(lib.mwg_plan_execute(self._h, _ptr(v, _d), 5))
        return {"ms": v[0], "term_apps": int(v[1]), "lp_term_apps": int(v[2]), "ms_fp64": v[3], "ms_lp": v[4]}

    def tune(self, max_rel_err, reps=3):
        """Autotune the plan's knobs (FP64-class cut tau, single-rounding cut,
        queue order) against an FP64 run of the same plan; the choice stays
        in the plan and is cached in the context for plans of the same
        (k + 1, terms, band width, mode).  Returns the picks and times."""
        v = np.zeros(9)
        _check(lib.mwg_plan_tune(self._h, float(max_rel_err), int(reps), _ptr(v, _d), 9))
        return {"tau": v[0], "single_ratio": v[1], "order": int(v[2]), "ms": v[3], "start_ms": v[4],
                "rel_err": v[5], "tried": int(v[6]), "speedup": v[7], "start_rel_err": v[8]}

    def use_graph(self, on=True):
        """Later execute()/apply() calls replay one captured hipGraph of the
        whole step (after one eager run); a replay reports the whole device
        time only."""
        _check(lib.mwg_plan_use_graph(self._h, 1 if on else 0))
        return self

    def apply(self):
        """One whole apply on resident data: the calculator, then the TopDown
        and BottomUp transforms and the norms on the device, in place (what
        mrcpp::apply does after its calculator).  Returns the execute() dict
        with ms covering all of it and post_ms the post transforms' share;
        finish() then only downloads."""
        v = np.zeros(6)
        _check(lib.mwg_plan_apply(self._h, _ptr(v, _d), 6))
        return {"ms": v[0], "term_apps": int(v[1]), "lp_term_apps": int(v[2]), "ms_fp64": v[3], "ms_lp": v[4],
                "post_ms": v[5]}

    def finish(self):
        """Result into the output tree(s); returns the tree (or the list)."""
        _check(lib.mwg_plan_finish(self._h))
        return self._outs[0] if len(self._outs) == 1 else self._outs

    def download_ns(self):
        """Raw non-standard blocks, all applies one after the other."""
        n = self.info()["out_nodes"]
        ns = np.empty((n, self._outs[0].mra.block))
        _check(lib.mwg_plan_download_ns(self._h, _ptr(ns, _d), ns.size))
        return ns

    def info(self):
        v = np.zeros(11)
        _check(lib.mwg_plan_info(self._h, _ptr(v, _d), 11))
        return {"out_nodes": int(v[0]), "in_slots": int(v[1]), "generated": int(v[2]), "plan_ms": v[3],
                "host_term_apps": int(v[4]), "segments": int(v[5]), "applies": int(v[6]), "band_w": int(v[7]),
                "tau": v[8], "single_ratio": v[9], "order": int(v[10])}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.mwg_plan_destroy(h)
            self._h = None
