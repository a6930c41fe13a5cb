"""FP64 CPU oracle for the likelihood call surface.  TEST INFRASTRUCTURE ONLY:
imported by tests/, by __graft_entry__.smoke() as the checker and by
bench.py's cpu_baseline leg; never by the product package.

``OracleLik`` restates the reference's ``Lik`` (pkg/Crux/Tree/Lik.pyx) in
Python over the C core in cxlik_oracle.c:

* model slots, defaults and setters: Lik.pyx:388-468, 900-1393;
* ``prep``: weight scaling and wNorm with the +I exclusion, Lik.pyx:1603-1665;
* Q construction and its eigensystem: CxLik.c:59-250.  The eigensystem uses
  numpy.linalg.eig and numpy.linalg.inv, which call LAPACK dgeev and an LU
  factorisation -- the same LAPACK routines the reference calls
  (CxLik.c:194-226, CxLapack.h:5-15);
* Gamma categories: oracle/gamma.py (Lik.pyx:1196-1278);
* the plan: a full post-order plan built exactly as Lik._plan/_planRecurse
  build it when every cache is invalid (Lik.pyx:1427-1601), so every call
  recomputes every CLV (the reference's LikDebug mode compares against such
  a fresh plan, Lik.pyx:1714-1735);
* execution: orc_execute, a restatement of CxLikExecute (CxLik.c:294-968).
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

from oracle.gamma import gamma_cmult

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

ORC_COMPUTE_L, ORC_COMPUTE_I, ORC_MERGE_L, ORC_MERGE_I = 0, 1, 2, 3


class _Step(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("ntrail", ctypes.c_int),
                ("parentMat", ctypes.c_void_p), ("parentLnScale", ctypes.c_void_p),
                ("childMat", ctypes.c_void_p), ("childLnScale", ctypes.c_void_p),
                ("edgeLen", ctypes.c_double)]


class _OrcLik(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("ncomp", ctypes.c_int),
                ("nchars", ctypes.c_uint), ("stripeWidth", ctypes.c_uint),
                ("nstripes", ctypes.c_uint), ("charFreqs", ctypes.c_void_p),
                ("wNorm", ctypes.c_double), ("compWeight", ctypes.c_void_p),
                ("compCmult", ctypes.c_void_p), ("compRmult", ctypes.c_void_p),
                ("compModel", ctypes.c_void_p), ("cube", ctypes.c_void_p),
                ("vals", ctypes.c_void_p), ("piNorm", ctypes.c_void_p),
                ("nsteps", ctypes.c_int), ("steps", ctypes.c_void_p),
                ("rootMat", ctypes.c_void_p), ("rootLnScale", ctypes.c_void_p),
                ("siteLnL", ctypes.c_void_p), ("stripeLnL", ctypes.c_void_p)]


def build(force=False):
    """Compile the C oracle with the reference's optimisation flags
    (configure.ac:57,257)."""
    src = os.path.join(_HERE, "cxlik_oracle.c")
    if (not force and os.path.exists(_LIB_PATH)
            and os.path.getmtime(_LIB_PATH) >= os.path.getmtime(src)):
        return _LIB_PATH
    tmp = _LIB_PATH + ".tmp%d" % os.getpid()
    subprocess.check_call(["gcc", "-std=gnu99", "-O3", "-funroll-loops",
                           "-fomit-frame-pointer", "-fPIC", "-shared",
                           "-pthread", "-o", tmp, src, "-lm"])
    os.replace(tmp, _LIB_PATH)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_execute.argtypes = [ctypes.POINTER(_OrcLik), ctypes.c_int]
        L.orc_execute.restype = None
        L.orc_stripe_width.argtypes = [ctypes.c_uint, ctypes.c_uint]
        L.orc_stripe_width.restype = ctypes.c_uint
        L.orc_pt.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p, ctypes.c_double]
        L.orc_pt.restype = None
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _tri_index(n, x, y):
    """Upper-triangle index of (x<y), CxLik.c:36-53 (order AC,AG,AT,CG,CT,GT
    for DNA)."""
    return n * x + y - (((x + 3) * x) >> 1) - 1


def q_matrix(rtri, pi):
    """Q (row i = from-state), normalised pi and qNorm, CxLik.c:59-141."""
    n = len(pi)
    pi = np.asarray(pi, dtype=np.float64)
    pin = pi / pi.sum()
    Q = np.zeros((n, n))
    qscale = 0.0
    for i in range(n):
        for j in range(i):
            elm = rtri[_tri_index(n, j, i)] * pin[j]
            qscale += elm * pin[i] * 2.0
            Q[i, j] = elm
        for j in range(i + 1, n):
            Q[i, j] = rtri[_tri_index(n, i, j)] * pin[j]
    for i in range(n):
        Q[i, i] = -(Q[i, :i].sum() + Q[i, i + 1:].sum())
    qnorm = 1.0 / qscale if qscale != 0.0 else math.inf
    return Q, pin, qnorm


def q_decomp_symmetric(Q, pin):
    """Eigensystem of a reversible Q through the symmetric Pi^1/2 Q Pi^-1/2:
    its orthonormal eigenvectors U give V = Pi^-1/2 U and V^-1 = U^T Pi^1/2.
    Returns (eigenvalues, V, V^-1)."""
    sq = np.sqrt(pin)
    w, U = np.linalg.eigh(Q * sq[:, None] / sq[None, :])
    return w, U / sq[:, None], U.T * sq[None, :]


def q_decomp(rtri, pi):
    """(cube[i][j][k] = V[i,k]*Vinv[k,j], eigenvalues, piNorm, qNorm),
    CxLik.c:177-250."""
    Q, pin, qnorm = q_matrix(rtri, pi)
    w, V = np.linalg.eig(Q)
    w = np.real(w)
    V = np.real(V)
    try:
        Vi = np.linalg.inv(V)
    except np.linalg.LinAlgError:
        # The general solver can return linearly dependent eigenvectors for
        # a repeated eigenvalue of this non-symmetric Q (the reference then
        # reports a decomposition failure).
        w, V, Vi = q_decomp_symmetric(Q, pin)
    cube = np.einsum("ik,kj->ijk", V, Vi)
    return np.ascontiguousarray(cube), np.ascontiguousarray(w), pin, qnorm


def p_matrix(cube, vals, t):
    """P(t), CxLik.c:258-292."""
    n = len(vals)
    P = np.empty((n, n))
    lib().orc_pt(n, _ptr(P), _ptr(np.ascontiguousarray(cube)),
                 _ptr(np.ascontiguousarray(vals)), float(t))
    return P


class _Model:
    def __init__(self, dim, weight, ncat, catMedian, invar):
        self.dim = dim
        self.weight = float(weight)
        self.rmult = 1.0
        rlen = dim * (dim - 1) // 2
        self.rclass = [0] * rlen
        self.rtri = [1.0] * rlen
        self.pi = [1.0 / dim] * dim
        self.alpha = math.inf
        self.catMedian = bool(catMedian)
        self.invar = bool(invar)
        self.clen = ncat + (1 if invar else 0)
        self.decomp = True
        self.cube = self.vals = self.pin = None
        self.qnorm = None


class _Comp:
    def __init__(self, model, cweight, cmult):
        self.model = model
        self.cweight = cweight
        self.cmult = cmult
        self.weightScaled = 0.0


class OracleLik:
    """FP64 likelihood with the reference's call surface (subset used by
    the tests and the CPU baseline)."""

    def __init__(self, tree, alignment, nmodels=1, ncat=1, catMedian=False,
                 invar=False, nthreads=1):
        if tree.rooted:
            raise ValueError("Tree must be unrooted")
        self.tree = tree
        self.alignment = alignment
        self.char_ = alignment.charType.get()
        self.dim = self.char_.nstates
        self.models = []
        self.comps = []
        self.wNorm = 0.0
        self.nthreads = max(1, int(nthreads))
        self._rows = None
        # CL buffers per ring (and the root's), kept between evaluations as
        # the reference keeps its CL objects on Ring.aux (Lik.pyx:32-108);
        # like its malloc'd cLMat they are not cleared: every entry an
        # evaluation reads it has written first.
        self._cls = {}
        for _ in range(nmodels):
            self.addModel(1.0, ncat, catMedian, invar)

    # -- mixture bookkeeping (Lik.pyx:388-468,900-949) ------------------------
    def nmodels(self):
        return len(self.models)

    def addModel(self, weight, ncat=1, catMedian=False, invar=False):
        assert ncat > 0
        m = _Model(self.dim, weight, ncat, catMedian, invar)
        comps = [_Comp(m, 1.0, 1.0)] + [_Comp(m, 0.0, 0.0)
                                         for _ in range(m.clen - 1)]
        m.comps = comps
        self.models.append(m)
        self._rebuild_comps()
        return len(self.models) - 1

    def delModel(self, i):
        del self.models[i]
        self._rebuild_comps()

    def _rebuild_comps(self):
        self.comps = [c for m in self.models for c in m.comps]

    def setWeight(self, m, w):
        assert w >= 0.0
        self.models[m].weight = float(w)

    def getWeight(self, m):
        return self.models[m].weight

    def setRmult(self, m, r):
        assert r >= 0.0
        self.models[m].rmult = float(r)

    def setRclass(self, m, rclass, rates=None):
        mod = self.models[m]
        rmax = 0
        for r in rclass:
            if r == rmax + 1:
                rmax += 1
            elif r > rmax + 1:
                raise ValueError("Invalid rclass specification")
        mod.rclass = list(rclass)
        if rates is None:
            mod.rtri = [1.0] * len(rclass)
        else:
            if len(rates) != rmax + 1:
                raise ValueError("Incorrect rates list length")
            mod.rtri = [float(rates[c]) for c in rclass]
        mod.decomp = True

    def setRate(self, m, i, rate):
        mod = self.models[m]
        ok = False
        for j, c in enumerate(mod.rclass):
            if c == i:
                mod.rtri[j] = float(rate)
                ok = True
        if not ok:
            raise ValueError("Invalid rate class index")
        mod.decomp = True

    def setFreq(self, m, i, f):
        assert f > 0.0
        self.models[m].pi[i] = float(f)
        self.models[m].decomp = True

    def setAlpha(self, m, alpha):
        mod = self.models[m]
        ncat = mod.clen - (1 if mod.invar else 0)
        assert ncat > 1
        if mod.alpha == alpha:
            return
        mod.alpha = alpha
        cs = mod.comps[:ncat]
        if alpha != math.inf:
            cm = gamma_cmult(alpha, ncat, mod.catMedian)
            for c, v in zip(cs, cm):
                c.cmult = v
            w = sum(c.cweight for c in cs) / ncat
            for c in cs:
                c.cweight = w
        else:
            w = sum(c.cweight for c in cs)
            cs[0].cweight = w
            cs[0].cmult = 1.0
            for c in cs[1:]:
                c.cweight = 0.0
                c.cmult = 0.0

    def setWVar(self, m, w):
        mod = self.models[m]
        ncat = mod.clen - (1 if mod.invar else 0)
        if mod.alpha == math.inf:
            mod.comps[0].cweight = w
        else:
            for c in mod.comps[:ncat]:
                c.cweight = w / ncat

    def setWInvar(self, m, w):
        mod = self.models[m]
        if not mod.invar and w != 0.0:
            raise ValueError("Model does not support invariable sites.")
        mod.comps[mod.clen - 1].cweight = w

    # -- prep (Lik.pyx:1603-1665) -------------------------------------------
    def prep(self):
        if not self.models:
            raise ValueError("Empty model mixture")
        wsum = sum(m.weight for m in self.models)
        if wsum == 0.0:
            raise ValueError("At least one model must have non-zero weight")
        for m in self.models:
            wc = sum(c.cweight for c in m.comps)
            assert wc > 0.0
            for c in m.comps:
                c.weightScaled = (m.weight / wsum) * (c.cweight / wc)
        for m in self.models:
            if m.decomp:
                m.cube, m.vals, m.pin, m.qnorm = q_decomp(m.rtri, m.pi)
                m.decomp = False
        qmean = 0.0
        for c in self.comps:
            if c.weightScaled != 0.0 and c.cmult != 0.0:
                qmean += c.weightScaled * (c.model.rmult / c.model.qnorm)
        self.wNorm = 1.0 / qmean

    # -- leaf CLVs -----------------------------------------------------------
    def _leaf(self, taxon):
        if self._rows is None:
            self._rows = {}
        cl = self._rows.get(taxon)
        if cl is None:
            ind = self.alignment.taxaMap.get(taxon, -1)
            if ind == -1:
                raise ValueError("Taxon %r missing from alignment's taxa map"
                                 % taxon)
            vals = self.char_.table[self.alignment.getRow(ind)]
            vals = np.where(vals == 0, self.char_.any, vals)
            mat = ((vals[:, None] >> np.arange(self.dim)[None, :]) & 1)
            cl = (np.ascontiguousarray(mat, dtype=np.float64),
                  np.zeros(len(vals)))
            self._rows[taxon] = cl
        return cl

    # -- plan (Lik.pyx:1427-1601, all caches invalid) ------------------------
    def _plan(self, root):
        n = self.alignment.nchars
        ncomp = len(self.comps)
        dn = ncomp * self.dim
        steps = []
        keep = []

        def new_cl(key):
            cl = self._cls.get(key)
            if cl is None or cl[0].shape != (n, dn):
                cl = (np.empty((n, dn)), np.empty(n))
                self._cls[key] = cl
            keep.append(cl)
            return cl

        def cl_of(ring):
            node = ring.node
            if node.degree == 1:
                if node.taxon is None:
                    raise ValueError("Leaf node missing taxon")
                return self._leaf(node.taxon), True
            return ring_cls[ring.id], False

        ring_cls = {}

        def append(variant, ntrail, parent, child, elen):
            if elen < 0.0:
                raise ValueError("Negative branch length")
            steps.append((variant, ntrail, parent, child, elen))

        def recurse(ring):
            node = ring.node
            deg = node.degree
            if deg == 1:
                cl_of(ring)
                return
            cl = new_cl(ring.id)
            ring_cls[ring.id] = cl
            for r in ring.siblings():
                recurse(r.other)
            r = ring.next
            for j in range(1, deg):
                ccl, leaf = cl_of(r.other)
                if leaf:
                    var = ORC_COMPUTE_L if j == 1 else ORC_MERGE_L
                else:
                    var = ORC_COMPUTE_I if j == 1 else ORC_MERGE_I
                append(var, deg - j - 1, cl, ccl, r.edge.length)
                r = r.next

        if root is None:
            root = self.tree.base
        root_cl = new_cl("root")
        ring = root.ring
        deg = root.degree
        if deg == 1:
            recurse(ring.other)
            ccl, leaf = cl_of(ring.other)
            append(ORC_COMPUTE_L if leaf else ORC_COMPUTE_I, 1, root_cl, ccl,
                   ring.edge.length)
            lcl, _ = cl_of(ring)
            append(ORC_MERGE_L, 0, root_cl, lcl, 0.0)
        else:
            for r in ring:
                recurse(r.other)
            ccl, leaf = cl_of(ring.other)
            ntrail = deg - 1
            append(ORC_COMPUTE_L if leaf else ORC_COMPUTE_I, ntrail, root_cl,
                   ccl, ring.edge.length)
            for r in ring.siblings():
                ccl, leaf = cl_of(r.other)
                ntrail -= 1
                append(ORC_MERGE_L if leaf else ORC_MERGE_I, ntrail, root_cl,
                       ccl, r.edge.length)
        return steps, root_cl, keep

    # -- execution --------------------------------------------------------
    def _execute(self, root, stripe_width=None):
        self.prep()
        steps, root_cl, keep = self._plan(root)
        site, stripe = self._run(steps, root_cl, stripe_width)
        self._last = (steps, keep)
        return site, stripe

    def _run(self, steps, root_cl, stripe_width=None):
        n = self.alignment.nchars
        ncomp = len(self.comps)
        if stripe_width is None:
            stripe_width = (lib().orc_stripe_width(n, self.nthreads)
                            if self.nthreads > 1 else n)
        nstripes = (n + stripe_width - 1) // stripe_width
        st = (_Step * max(1, len(steps)))()
        for i, (var, ntr, par, ch, el) in enumerate(steps):
            st[i].variant = var
            st[i].ntrail = ntr
            st[i].parentMat = _ptr(par[0])
            st[i].parentLnScale = _ptr(par[1])
            st[i].childMat = _ptr(ch[0])
            st[i].childLnScale = _ptr(ch[1])
            st[i].edgeLen = el
        mods = self.models
        midx = {id(m): i for i, m in enumerate(mods)}
        cw = np.array([c.weightScaled for c in self.comps])
        cc = np.array([c.cmult for c in self.comps])
        cr = np.array([c.model.rmult for c in self.comps])
        cm = np.array([midx[id(c.model)] for c in self.comps], dtype=np.int32)
        cube = np.ascontiguousarray(np.stack([m.cube for m in mods]))
        vals = np.ascontiguousarray(np.stack([m.vals for m in mods]))
        pin = np.ascontiguousarray(np.stack([m.pin for m in mods]))
        freqs = np.ascontiguousarray(self.alignment.freqs, dtype=np.uint32)
        site = np.zeros(n)
        stripe = np.zeros(nstripes)
        L = _OrcLik(self.dim, ncomp, n, stripe_width, nstripes, _ptr(freqs),
                    self.wNorm, _ptr(cw), _ptr(cc), _ptr(cr), _ptr(cm),
                    _ptr(cube), _ptr(vals), _ptr(pin), len(steps),
                    ctypes.cast(st, ctypes.c_void_p), _ptr(root_cl[0]),
                    _ptr(root_cl[1]), _ptr(site), _ptr(stripe))
        lib().orc_execute(ctypes.byref(L), self.nthreads)
        return site, stripe

    def time_full(self, reps=20, warmup=3, root=None):
        """Median CPU seconds of a full evaluation (plan and execution, CL
        buffers reused), after ``warmup`` untimed ones."""
        import time
        ts = []
        for i in range(warmup + reps):
            t = time.perf_counter()
            self.lnL(root)
            if i >= warmup:
                ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    def time_partial(self, nupdates, reps=20, warmup=3, root=None):
        """Median CPU seconds of a partial re-evaluation's work: the plan's
        last ``nupdates`` CLV updates and the root's steps, executed on the
        buffers a full execution has just filled.  The reference's planner
        recomputes only invalidated CLVs (Lik.pyx:1427-1601), and each of its
        updates costs the same (patterns x components x states^2) work, so
        this times the CPU work of a device partial re-evaluation with as
        many node updates."""
        import time
        self.prep()
        steps, root_cl, keep = self._plan(root)
        self._run(steps, root_cl)
        # Steps are grouped by parent CLV in execution order; the root's
        # group is last.
        groups = []
        for st in steps:
            if not groups or groups[-1][0] is not st[2]:
                groups.append((st[2], []))
            groups[-1][1].append(st)
        tail = [st for _, g in groups[-(nupdates + 1):] for st in g]
        ts = []
        for i in range(warmup + reps):
            t = time.perf_counter()
            self._run(tail, root_cl)
            if i >= warmup:
                ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    def lnL(self, root=None):
        site, stripe = self._execute(root)
        ret = 0.0
        for v in stripe:
            ret += v
        return ret

    def siteLnLs(self, root=None):
        site, _ = self._execute(root)
        return list(site[:self.alignment.nchars - self.alignment.npad])

    def getWNorm(self):
        return self.wNorm
