"""ctypes binding of libcrux_amd.so (the C ABI declared in include/crux_lik.h).

The library is built in-tree by ``python -m crux_amd.build``.  There is no
CPU fallback: importing this module without the built library raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcrux_amd.so")

CRUX_OK, CRUX_EINVAL, CRUX_ENOMEM, CRUX_ENEGLEN, CRUX_EDECOMP, CRUX_EDEVICE, \
    CRUX_ETAXON, CRUX_ESTATE = range(8)
CRUX_TUNE_STACK_DEPTH = 1
CRUX_TUNE_PRECISION = 2
CRUX_TUNE_CHUNK = 3
CRUX_TUNE_WAVES = 4
CRUX_TUNE_SHORT_PLANS = 5
CRUX_TUNE_SPLIT = 6
CRUX_PREC_F32, CRUX_PREC_BF16, CRUX_PREC_F16, CRUX_PREC_F16X3 = range(4)

# Mixture components (over all models, +I counted) a likelihood may have:
# the largest component class of the traversal kernels per state count.
MAX_COMPONENTS = {4: 16, 20: 8, 22: 8}

#: Every entry point the header declares, with (restype, argtypes).
_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "crux_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "crux_last_error": (ctypes.c_char_p, []),
    "crux_q_decomp": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "crux_pt": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, _dp]),
    "crux_gamma_rates": (ctypes.c_int, [ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp]),
    "crux_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "crux_ctx_destroy": (ctypes.c_int, [_vp]),
    "crux_ctx_bytes": (ctypes.c_uint64, [_vp]),
    "crux_tree_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_uint8), ctypes.c_int, _dp,
                                        ctypes.POINTER(ctypes.c_uint32),
                                        ctypes.POINTER(_vp)]),
    "crux_tree_destroy": (ctypes.c_int, [_vp]),
    "crux_tree_set_topology": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                              _i32p, _i32p, _dp]),
    "crux_tree_set_lengths": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _dp]),
    "crux_tree_update_rings": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _i32p,
                                              ctypes.c_int, _i32p, _i32p]),
    "crux_lik_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "crux_lik_destroy": (ctypes.c_int, [_vp]),
    "crux_lik_set_model": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int,
                                          _i32p, _dp, _dp, ctypes.c_int]),
    "crux_lik_lnl": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "crux_lik_lnl_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "crux_lik_site_lnl": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int]),
    "crux_lik_flush": (ctypes.c_int, [_vp]),
    "crux_batch_lnl": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), _i32p, _dp]),
    "crux_batch_submit": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), _i32p,
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "crux_batch_wait": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_int, _dp]),
    "crux_batch_poll": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "crux_server_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "crux_server_destroy": (ctypes.c_int, [_vp]),
    "crux_server_add_lik": (ctypes.c_int, [_vp, _vp]),
    "crux_server_remove_lik": (ctypes.c_int, [_vp, _vp]),
    "crux_server_submit": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _i32p, _dp,
                                          ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]),
    "crux_server_poll": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "crux_server_wait": (ctypes.c_int, [_vp, ctypes.c_uint64, _dp, ctypes.POINTER(ctypes.c_int64)]),
    "crux_server_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "crux_ctx_set_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "crux_ctx_set_tuning": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "crux_ctx_get_tuning": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "crux_ctx_set_class_tuning": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int]),
    "crux_ctx_get_class_tuning": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_int)]),
    "crux_ctx_launch_shape": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int)]),
    "crux_ctx_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "crux_ctx_last_timing": (ctypes.c_int, [_vp, _dp]),
    "crux_ctx_last_host_timing": (ctypes.c_int, [_vp, _dp]),
    "crux_ctx_last_traffic": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64),
                                             ctypes.POINTER(ctypes.c_uint64),
                                             ctypes.POINTER(ctypes.c_int)]),
    "crux_ctx_last_launches": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_int)]),
}


class CruxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


_lib = None
_lock = threading.Lock()


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        "libcrux_amd.so is not built (run `python -m crux_amd.build`); "
                        "there is no CPU fallback")
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in SIGNATURES.items():
                    f = getattr(L, name)
                    f.restype = res
                    f.argtypes = args
                _lib = L
    return _lib


def check(rc):
    """Raise the Python exception the reference raises for this status."""
    if rc == CRUX_OK:
        return
    msg = lib().crux_last_error().decode(errors="replace")
    if rc in (CRUX_EINVAL, CRUX_ENEGLEN, CRUX_EDECOMP, CRUX_ETAXON):
        raise ValueError(msg)
    if rc == CRUX_ENOMEM:
        raise MemoryError(msg)
    raise CruxError(rc, "%s: %s" % (lib().crux_strerror(rc).decode(), msg))


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_i32p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---- host FP64 model math ---------------------------------------------------

def q_decomp(rtri, pi):
    """(V, Vinv, evals, pi_norm, q_norm) of the Q matrix (crux_q_decomp)."""
    rtri = f64(rtri)
    pi = f64(pi)
    n = len(pi)
    V = np.empty((n, n))
    Vi = np.empty((n, n))
    w = np.empty(n)
    pin = np.empty(n)
    qn = ctypes.c_double()
    check(lib().crux_q_decomp(n, dptr(rtri), dptr(pi), dptr(pin), ctypes.byref(qn),
                              dptr(V), dptr(Vi), dptr(w)))
    return V, Vi, w, pin, qn.value


def pt(V, Vi, w, t):
    n = len(w)
    P = np.empty((n, n))
    check(lib().crux_pt(n, dptr(f64(V)), dptr(f64(Vi)), dptr(f64(w)), float(t), dptr(P)))
    return P


def gamma_rates(alpha, ncat, median=False):
    out = np.empty(ncat)
    rc = lib().crux_gamma_rates(float(alpha), int(ncat), 1 if median else 0, dptr(out))
    if rc == CRUX_EINVAL and "discretizing" in lib().crux_last_error().decode():
        raise OverflowError(lib().crux_last_error().decode())
    check(rc)
    return out


# ---- device context ------------------------------------------------------------

class Context:
    """One device context (stream, staging buffers, HBM budget)."""

    def __init__(self, device=0, hbm_budget=0):
        h = _vp()
        check(lib().crux_ctx_create(int(device), int(hbm_budget), ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().crux_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def bytes(self):
        return lib().crux_ctx_bytes(self.handle)

    def set_mode(self, mode):
        check(lib().crux_ctx_set_mode(self.handle, int(mode)))

    def set_tuning(self, key, value):
        check(lib().crux_ctx_set_tuning(self.handle, int(key), int(value)))

    def get_tuning(self, key):
        v = ctypes.c_int()
        check(lib().crux_ctx_get_tuning(self.handle, int(key), ctypes.byref(v)))
        return v.value

    def set_class_tuning(self, K, C, key, value):
        """A shape key for the traversal class of K states and C components
        (-1 clears it: the context's value applies)."""
        check(lib().crux_ctx_set_class_tuning(self.handle, int(K), int(C), int(key), int(value)))

    def get_class_tuning(self, K, C, key):
        v = ctypes.c_int()
        check(lib().crux_ctx_get_class_tuning(self.handle, int(K), int(C), int(key),
                                              ctypes.byref(v)))
        return v.value

    def launch_shape(self, K, C, S):
        """(stack depth, updates per chunk, waves per SIMD, CM) a batch of one
        evaluation of S patterns would launch the class of K states and C
        components with now."""
        v = (ctypes.c_int * 4)()
        check(lib().crux_ctx_launch_shape(self.handle, int(K), int(C), int(S), v))
        return tuple(int(x) for x in v)

    def set_timing(self, on):
        check(lib().crux_ctx_set_timing(self.handle, 1 if on else 0))

    def last_timing(self):
        out = np.zeros(4)
        check(lib().crux_ctx_last_timing(self.handle, dptr(out)))
        return out

    def last_host_timing(self):
        """Host microseconds of the last batch: planning, building, launch
        calls, completion wait + final sums."""
        v = np.zeros(4)
        check(lib().crux_ctx_last_host_timing(self.handle, dptr(v)))
        return v

    def last_traffic(self):
        """(bytes moved, bytes under the all-children-from-HBM model, node
        updates) of the last evaluation."""
        b = ctypes.c_uint64()
        f = ctypes.c_uint64()
        n = ctypes.c_int()
        check(lib().crux_ctx_last_traffic(self.handle, ctypes.byref(b), ctypes.byref(f),
                                          ctypes.byref(n)))
        return b.value, f.value, n.value

    def last_launches(self):
        """(traversal launches, of which short-plan kernel launches) of the
        last batch."""
        n = ctypes.c_int()
        k = ctypes.c_int()
        check(lib().crux_ctx_last_launches(self.handle, ctypes.byref(n), ctypes.byref(k)))
        return n.value, k.value


_contexts = {}


def get_context(device=0):
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        _contexts[device] = ctx
    return ctx


def batch_lnl(ctx, liks, roots):
    """lnL of several device likelihoods in one fused pass."""
    n = len(liks)
    arr = (_vp * max(1, n))(*[l.handle for l in liks])
    r = i32(roots)
    out = np.empty(n)
    check(lib().crux_batch_lnl(ctx.handle, n, arr, iptr(r), dptr(out)))
    return out


def batch_submit(ctx, liks, roots):
    """Plan and enqueue a fused pass; returns (ticket, n) for batch_wait."""
    n = len(liks)
    arr = (_vp * max(1, n))(*[l.handle for l in liks])
    r = i32(roots)
    t = ctypes.c_uint64()
    check(lib().crux_batch_submit(ctx.handle, n, arr, iptr(r), ctypes.byref(t)))
    return (t.value, n)


def batch_wait(ctx, ticket):
    """Results of a submitted pass (blocks until it is done)."""
    t, n = ticket
    out = np.empty(n)
    check(lib().crux_batch_wait(ctx.handle, t, n, dptr(out)))
    return out


def batch_poll(ctx, ticket):
    """True once a submitted pass has finished (never blocks)."""
    t, _ = ticket
    done = ctypes.c_int()
    check(lib().crux_batch_poll(ctx.handle, t, ctypes.byref(done)))
    return bool(done.value)
