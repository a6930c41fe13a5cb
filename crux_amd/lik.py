"""``Lik``: the reference's likelihood call surface over the MI355X engine.

Mirrors ``Crux.Tree.Lik.Lik`` (pkg/Crux/Tree/Lik.pyx, declared in Lik.pxd:
23-98): construction from an unrooted tree and an alignment with
``nmodels``/``ncat``/``catMedian``/``invar``; ``lnL(root)`` and
``siteLnLs(root)``; the mixture setters and getters; ``prep``, ``flush``,
``clone`` (polarity mate on the same tree), ``dup``, ``simulate`` and
pickling.  Errors are the reference's Python exceptions.

What lives where:

* model parameters, mixture weights and the invalidation flags are kept here,
  with the reference's semantics (Lik.pyx:388-468, 900-1393, 1603-1665);
* Q's eigensystem and the Gamma categories come from the native FP64 host
  code (crux_q_decomp, crux_gamma_rates);
* the tree topology and branch lengths are mirrored into a device tree; the
  native planner re-derives which CLVs are stale with the reference's cache
  keys and runs them on the GPU (crux_lik_lnl).

Columns are not padded: the reference pads only in threaded mode, to a
multiple of its CPU stripe width (Lik.pyx:206-250); sites map to GPU lanes
here, so there are no stripes to fill.
"""

from __future__ import annotations

import math
import random

import numpy as np

from . import _native as nat
from .alignment import Alignment
from .tree import Tree

__all__ = ["Lik"]


class _Model:
    __slots__ = ("weight", "rmult", "rclass", "rtri", "pi", "alpha", "catMedian",
                 "invar", "comps", "decomp", "V", "Vi", "w", "pin", "qnorm")

    def __init__(self, dim, weight, ncat, catMedian, invar):
        rlen = dim * (dim - 1) // 2
        self.weight = float(weight)
        self.rmult = 1.0
        self.rclass = [0] * rlen
        self.rtri = [1.0] * rlen
        self.pi = [1.0 / dim] * dim
        self.alpha = math.inf
        self.catMedian = bool(catMedian)
        self.invar = bool(invar)
        n = ncat + (1 if invar else 0)
        # Only the first component carries weight until alpha or the +I
        # weight is set (Lik.pyx:459-468).
        self.comps = [_Comp(self, 1.0, 1.0)] + [_Comp(self, 0.0, 0.0) for _ in range(n - 1)]
        self.decomp = True
        self.V = self.Vi = self.w = self.pin = None
        self.qnorm = None

    @property
    def ncat(self):
        return len(self.comps) - (1 if self.invar else 0)


class _Comp:
    __slots__ = ("model", "cweight", "cmult", "weightScaled")

    def __init__(self, model, cweight, cmult):
        self.model = model
        self.cweight = float(cweight)
        self.cmult = float(cmult)
        self.weightScaled = 0.0


class _DeviceTree:
    """Device mirror of one (tree, alignment) pair, shared by a Lik and its
    clone."""

    def __init__(self, ctx, tree, alignment):
        self.ctx = ctx
        self.tree = tree
        self.alignment = alignment
        codes, masks = alignment.encode()
        vecs = alignment.mask_vectors(masks)
        self.dim = vecs.shape[1]
        h = nat._vp()
        nat.check(nat.lib().crux_tree_create(
            ctx.handle, self.dim, alignment.nchars, alignment.ntaxa,
            codes.ctypes.data_as(nat.ctypes.POINTER(nat.ctypes.c_uint8)),
            len(masks), nat.dptr(nat.f64(vecs)),
            np.ascontiguousarray(alignment.freqs, dtype=np.uint32).ctypes.data_as(
                nat.ctypes.POINTER(nat.ctypes.c_uint32)),
            nat.ctypes.byref(h)))
        self.handle = h
        self.sn = None
        self.ids = None
        self.sink = tree.length_sink()
        self.tsink = tree.topology_sink()
        # Length uploads of a few edges (an MC3 move) reuse these arrays: a
        # new ctypes array type per call costs more than the call itself.
        self._nbuf = 16
        self._idbuf = (nat.ctypes.c_int32 * self._nbuf)()
        self._lenbuf = (nat.ctypes.c_double * self._nbuf)()
        self._set_lengths = nat.lib().crux_tree_set_lengths
        self._tbuf_n = 64
        self._tbuf = [(nat.ctypes.c_int32 * self._tbuf_n)() for _ in range(5)]
        self._update_rings = nat.lib().crux_tree_update_rings

    def sync(self):
        t = self.tree
        # Only attach/detach change what the device mirrors (moving the base
        # does not), so key the topology upload on the topology serial; when
        # no node or edge was created since the last upload, only the rings
        # and nodes that attach/detach touched are sent.
        if self.sn != t.tsn:
            rings, nodes = self.tsink
            if self.ids == (t.max_node_id, t.max_edge_id):
                # A few rings and nodes (an MC3 move), in any order, through
                # arrays kept for the purpose (see _nbuf).
                nr, nn = len(rings), len(nodes)
                if nr <= self._tbuf_n and nn <= self._tbuf_n:
                    rid, rnode, rnext, nid, nring = self._tbuf
                else:
                    rid, rnode, rnext = [(nat.ctypes.c_int32 * nr)() for _ in range(3)]
                    nid, nring = [(nat.ctypes.c_int32 * nn)() for _ in range(2)]
                for i, r in enumerate(rings):
                    rid[i] = r.id
                    if r.node is not None:
                        rnode[i] = r.node.id
                        rnext[i] = r.next.id
                    else:
                        rnode[i] = -1
                        rnext[i] = -1
                for i, n in enumerate(nodes):
                    nid[i] = n.id
                    nring[i] = n.ring.id if n.ring is not None else -1
                nat.check(self._update_rings(self.handle, nr, rid, rnode, rnext, nn, nid, nring))
            else:
                rn, rx, nr, nt, el = t.topology_arrays(self.alignment.taxaMap)
                nat.check(nat.lib().crux_tree_set_topology(
                    self.handle, len(nr), len(el), nat.iptr(rn), nat.iptr(rx), nat.iptr(nr),
                    nat.iptr(nt), nat.dptr(el)))
                self.ids = (t.max_node_id, t.max_edge_id)
                self.sink.clear()
            rings.clear()
            nodes.clear()
            self.sn = t.tsn
        if self.sink:
            # Usually one or two edges (an MC3 move).
            n = len(self.sink)
            if n <= self._nbuf:
                ids, lens = self._idbuf, self._lenbuf
                for i, e in enumerate(self.sink):
                    ids[i] = e.id
                    lens[i] = e._length
            else:
                edges = list(self.sink)
                ids = (nat.ctypes.c_int32 * n)(*[e.id for e in edges])
                lens = (nat.ctypes.c_double * n)(*[e.length for e in edges])
            self.sink.clear()
            rc = self._set_lengths(self.handle, n, ids, lens)
            if rc:
                nat.check(rc)

    def close(self):
        if self.handle:
            nat.lib().crux_tree_destroy(self.handle)
            self.handle = None
            # Stop collecting edits for a mirror that no longer exists.
            t = self.tree
            t._sinks = [x for x in t._sinks if x is not self.sink]
            t._topo_sinks = [x for x in t._topo_sinks if x is not self.tsink]

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Lik:
    """Log-likelihood of a tree and an alignment under a weighted mixture of
    reversible CTMC models with optional +G rate categories and +I."""

    def __init__(self, tree=None, alignment=None, nmodels=1, ncat=1, catMedian=False,
                 invar=False, context=None):
        self._dev = None
        self._handle = None
        if alignment is None:
            return
        self._init0(tree)
        self._init1(tree, alignment, 0, context, None)
        for _ in range(nmodels):
            self.addModel(1.0, ncat, catMedian, invar)

    # -- construction -----------------------------------------------------
    def _init0(self, tree):
        if tree.rooted:
            raise ValueError("Tree must be unrooted")
        tree.clearAux()

    def _init1(self, tree, alignment, polarity, context, dev):
        self.tree = tree
        self.alignment = alignment
        self.char_ = alignment.charType.get()
        self.dim = self.char_.nstates
        self.polarity = polarity
        self.mate = None
        self.models = []
        self.comps = []
        self.wNorm = 0.0
        self.invalidate = True
        self.reweight = True
        self._model_dirty = True
        ctx = context if context is not None else nat.get_context(0)
        self._ctx = ctx
        self._dev = dev if dev is not None else _DeviceTree(ctx, tree, alignment)
        h = nat._vp()
        nat.check(nat.lib().crux_lik_create(self._dev.handle, polarity, nat.ctypes.byref(h)))
        self._handle = h

    @property
    def handle(self):
        return self._handle

    def __del__(self):
        try:
            if self._handle:
                nat.lib().crux_lik_destroy(self._handle)
                self._handle = None
        except Exception:
            pass

    # -- mixture bookkeeping (Lik.pyx:900-949) --------------------------------
    def nmodels(self):
        return len(self.models)

    def addModel(self, weight, ncat=1, catMedian=False, invar=False):
        assert ncat > 0
        self.models.append(_Model(self.dim, weight, ncat, catMedian, invar))
        self._comps_changed()
        return len(self.models) - 1

    def delModel(self, model):
        assert model < len(self.models)
        del self.models[model]
        self._comps_changed()

    def _comps_changed(self):
        self.comps = [c for m in self.models for c in m.comps]
        self.reweight = True
        self.invalidate = True

    def _m(self, model):
        assert model < len(self.models)
        return self.models[model]

    # -- parameters (Lik.pyx:951-1393) ------------------------------------------
    def getWeight(self, model):
        return self._m(model).weight

    def setWeight(self, model, weight):
        assert weight >= 0.0
        self._m(model).weight = float(weight)
        self.reweight = True

    def getRmult(self, model):
        return self._m(model).rmult

    def setRmult(self, model, rmult):
        assert rmult >= 0.0
        self._m(model).rmult = float(rmult)
        self.reweight = True

    def getRclass(self, model):
        return list(self._m(model).rclass)

    def setRclass(self, model, rclass, rates=None):
        m = self._m(model)
        rclass = list(rclass)
        if len(rclass) != len(m.rclass):
            raise ValueError("Invalid rclass specification")
        rmax = 0
        for r in rclass:
            if r == rmax + 1:
                rmax += 1
            elif r > rmax + 1:
                raise ValueError("Invalid rclass specification")
        if rates is None:
            rtri = [1.0] * len(rclass)
        else:
            if len(rates) != rmax + 1:
                raise ValueError("Incorrect rates list length")
            for c in rclass:
                if not isinstance(rates[c], float):
                    raise ValueError("Incorrect type for rate %d (%r)" % (c, rates[c]))
            rtri = [rates[c] for c in rclass]
        m.rclass = rclass
        m.rtri = rtri
        m.decomp = True

    def getNrates(self, model):
        return max(self._m(model).rclass) + 1

    def getRate(self, model, i):
        m = self._m(model)
        for c, r in zip(m.rclass, m.rtri):
            if c == i:
                return r
        raise ValueError("Invalid rate class index")

    def setRate(self, model, i, rate):
        m = self._m(model)
        assert rate >= 0.0
        ok = False
        for j, c in enumerate(m.rclass):
            if c == i:
                if m.rtri[j] != rate:
                    m.decomp = True
                m.rtri[j] = float(rate)
                ok = True
        if not ok:
            raise ValueError("Invalid rate class index")

    def getFreq(self, model, i):
        return self._m(model).pi[i]

    def setFreq(self, model, i, freq):
        m = self._m(model)
        assert freq > 0.0
        if m.pi[i] != freq:
            m.pi[i] = float(freq)
            m.decomp = True

    def getAlpha(self, model):
        return self._m(model).alpha

    def setAlpha(self, model, alpha):
        """Gamma shape of the model's rate categories (Lik.pyx:1196-1278)."""
        m = self._m(model)
        ncat = m.ncat
        assert ncat > 1
        if m.alpha == alpha:
            return
        cs = m.comps[:ncat]
        if alpha != math.inf:
            cm = nat.gamma_rates(alpha, ncat, m.catMedian)
            m.alpha = alpha
            for c, v in zip(cs, cm):
                c.cmult = float(v)
            w = sum(c.cweight for c in cs) / ncat
            for c in cs:
                c.cweight = w
        else:
            m.alpha = alpha
            w = sum(c.cweight for c in cs)
            cs[0].cweight = w
            cs[0].cmult = 1.0
            for c in cs[1:]:
                c.cweight = 0.0
                c.cmult = 0.0
        self.invalidate = True
        self._model_dirty = True

    def getNcat(self, model):
        return self._m(model).ncat

    def getCatMedian(self, model):
        return self._m(model).catMedian

    def getInvar(self, model):
        return self._m(model).invar

    def getWVar(self, model):
        m = self._m(model)
        if m.alpha == math.inf:
            return m.comps[0].cweight
        return sum(c.cweight for c in m.comps[:m.ncat])

    def setWVar(self, model, wVar):
        m = self._m(model)
        assert wVar >= 0.0
        if m.alpha == math.inf:
            m.comps[0].cweight = float(wVar)
        else:
            for c in m.comps[:m.ncat]:
                c.cweight = wVar / m.ncat
        self.reweight = True

    def getWInvar(self, model):
        m = self._m(model)
        return m.comps[-1].cweight if m.invar else 0.0

    def setWInvar(self, model, wInvar):
        m = self._m(model)
        if not m.invar and wInvar != 0.0:
            raise ValueError("Model does not support invariable sites.")
        assert wInvar >= 0.0
        m.comps[-1].cweight = float(wInvar)
        self.reweight = True

    def getWNorm(self):
        return self.wNorm

    # -- prep (Lik.pyx:1603-1665) ----------------------------------------
    def prep(self):
        if not self.models:
            raise ValueError("Empty model mixture")
        if self.reweight:
            wsum = sum(m.weight for m in self.models)
            if wsum == 0.0:
                raise ValueError("At least one model must have non-zero weight")
            for m in self.models:
                wc = sum(c.cweight for c in m.comps)
                assert wc > 0.0
                for c in m.comps:
                    c.weightScaled = (m.weight / wsum) * (c.cweight / wc)
            self.invalidate = True
            self.reweight = False
            self._model_dirty = True
        for m in self.models:
            if m.decomp:
                m.V, m.Vi, m.w, m.pin, m.qnorm = nat.q_decomp(m.rtri, m.pi)
                m.decomp = False
                # A new Q invalidates every CLV computed under the old one.
                self.invalidate = True
                self._model_dirty = True
        qmean = 0.0
        for c in self.comps:
            if c.weightScaled != 0.0 and c.cmult != 0.0:
                qmean += c.weightScaled * (c.model.rmult / c.model.qnorm)
        wnorm = 1.0 / qmean if qmean != 0.0 else math.inf
        if wnorm != self.wNorm:
            self.invalidate = True
            self.wNorm = wnorm
            self._model_dirty = True

    def _push_model(self):
        if not self._model_dirty and not self.invalidate:
            return
        active = [c for c in self.comps if c.weightScaled != 0.0]
        mods = self.models
        midx = {id(m): i for i, m in enumerate(mods)}
        V = nat.f64(np.stack([m.V for m in mods]))
        Vi = nat.f64(np.stack([m.Vi for m in mods]))
        w = nat.f64(np.stack([m.w for m in mods]))
        pin = nat.f64(np.stack([m.pin for m in mods]))
        cm = nat.i32([midx[id(c.model)] for c in active])
        cw = nat.f64([c.weightScaled for c in active])
        cr = nat.f64([c.cmult * c.model.rmult * self.wNorm for c in active])
        nat.check(nat.lib().crux_lik_set_model(
            self._handle, len(mods), nat.dptr(V), nat.dptr(Vi), nat.dptr(w), nat.dptr(pin),
            len(active), nat.iptr(cm), nat.dptr(cw), nat.dptr(cr),
            1 if self.invalidate else 0))
        self.invalidate = False
        self._model_dirty = False

    def _ready(self, root):
        self.prep()
        self._push_model()
        self._dev.sync()
        if root is None:
            root = self.tree.base
        return root.id

    # -- evaluation (Lik.pyx:1687-1763) -----------------------------------
    def lnL(self, root=None):
        rid = self._ready(root)
        out = nat.ctypes.c_double()
        nat.check(nat.lib().crux_lik_lnl(self._handle, rid, nat.ctypes.byref(out)))
        return out.value

    def siteLnLs(self, root=None):
        rid = self._ready(root)
        n = self.alignment.nchars - self.alignment.npad
        out = np.empty(self.alignment.nchars)
        nat.check(nat.lib().crux_lik_site_lnl(self._handle, rid, nat.dptr(out),
                                              self.alignment.nchars))
        return list(out[:n])

    def flush(self):
        nat.check(nat.lib().crux_lik_flush(self._handle))

    # -- copies (Lik.pyx:587-660, 835-883) -----------------------------------
    def _copy_models_to(self, other):
        other.models = []
        for m in self.models:
            nm = _Model(self.dim, m.weight, m.ncat, m.catMedian, m.invar)
            nm.rmult = m.rmult
            nm.rclass = list(m.rclass)
            nm.rtri = list(m.rtri)
            nm.pi = list(m.pi)
            nm.alpha = m.alpha
            for c, d in zip(m.comps, nm.comps):
                d.cweight = c.cweight
                d.cmult = c.cmult
            other.models.append(nm)
        other.comps = [c for m in other.models for c in m.comps]
        other.reweight = True
        other.invalidate = True
        other._model_dirty = True

    def clone(self):
        """Lik with the same parameters sharing this tree and its device CLV
        pool through the opposite polarity (Lik.pyx:835-883)."""
        if self.mate is not None:
            ret = self.mate
        else:
            assert self.polarity == 0
            ret = Lik.__new__(Lik)
            ret._dev = None
            ret._handle = None
            ret._init1(self.tree, self.alignment, 1, self._ctx, self._dev)
            ret.mate = self
            self.mate = ret
        self._copy_models_to(ret)
        return ret

    def dup(self):
        """Independent Lik on a copy of the tree, same alignment
        (Lik.pyx:642-660)."""
        ret = Lik.__new__(Lik)
        ret._dev = None
        ret._handle = None
        tree = self.tree.dup()
        ret._init0(tree)
        ret._init1(tree, self.alignment, 0, self._ctx, None)
        self._copy_models_to(ret)
        return ret

    # -- simulation (Lik.pyx:662-833) -----------------------------------------
    def simulate(self, nchars=0, rng=None):
        """Lik on a copy of the tree with an alignment simulated under the
        current parameters.  Sites are split among mixture components in
        proportion to their scaled weights, as the reference does; the random
        stream is numpy's (``rng``) rather than Python's."""
        rng = rng if rng is not None else np.random.default_rng(random.getrandbits(64))
        self.prep()
        if nchars == 0:
            nchars = self.alignment.nchars - self.alignment.npad
        brks = []
        wsum = 0.0
        for c in self.comps[:-1]:
            wsum = min(1.0, wsum + c.weightScaled)
            brks.append(int(math.ceil(wsum * nchars)))
        brks.append(nchars)
        comp_of_site = np.zeros(nchars, dtype=np.int64)
        lo = 0
        for i, hi in enumerate(brks):
            comp_of_site[lo:hi] = i
            lo = max(lo, hi)
        tree = self.tree.dup()
        taxa = list(self.alignment.taxa)
        rows = np.zeros((len(taxa), nchars), dtype=np.uint8)
        states = np.frombuffer("".join(self.char_.pcodes()).encode(), dtype=np.uint8)
        root_seq = np.empty(nchars, dtype=np.int64)
        for i, c in enumerate(self.comps):
            sel = comp_of_site == i
            if sel.any():
                root_seq[sel] = rng.choice(self.dim, size=int(sel.sum()), p=c.model.pin)
        base = tree.base

        def put(node, seq):
            if node.taxon is not None:
                rows[self.alignment.taxaMap[node.taxon]] = states[seq]

        put(base, root_seq)
        stack = []
        if base.ring is not None:
            stack = [(r.other, root_seq) for r in base.ring]
        while stack:
            ring, pseq = stack.pop()
            seq = np.empty(nchars, dtype=np.int64)
            for i, c in enumerate(self.comps):
                sel = np.nonzero(comp_of_site == i)[0]
                if len(sel) == 0:
                    continue
                m = c.model
                P = nat.pt(m.V, m.Vi, m.w, ring.edge.length * c.cmult * m.rmult * self.wNorm)
                cdf = np.cumsum(P, axis=1)
                cdf[:, -1] = 1.0
                u = rng.random(len(sel))
                seq[sel] = (u[:, None] > cdf[pseq[sel]]).sum(axis=1)
            put(ring.node, seq)
            stack.extend((r.other, seq) for r in ring.siblings())
        aln = Alignment(charType=self.alignment.charType, taxa=taxa, rows=rows)
        ret = Lik.__new__(Lik)
        ret._dev = None
        ret._handle = None
        ret._init0(tree)
        ret._init1(tree, aln, 0, self._ctx, None)
        self._copy_models_to(ret)
        return ret

    # -- pickling (Lik.pyx:326-386) -------------------------------------------
    def __getstate__(self):
        models = []
        for i, m in enumerate(self.models):
            models.append((m.weight, m.rmult, list(m.rclass),
                           [self.getRate(i, j) for j in range(self.getNrates(i))],
                           list(m.pi), m.alpha, m.ncat, m.catMedian, m.invar,
                           self.getWVar(i), self.getWInvar(i)))
        return ((self.tree, self.alignment), models)

    def __setstate__(self, data):
        (tree, alignment), models = data
        self._dev = None
        self._handle = None
        self._init0(tree)
        self._init1(tree, alignment, 0, None, None)
        for i, (weight, rmult, rclass, rates, freqs, alpha, ncat, catMedian, invar, wVar,
                wInvar) in enumerate(models):
            self.addModel(weight, ncat, catMedian, invar)
            self.setRmult(i, rmult)
            self.setRclass(i, rclass, [float(r) for r in rates])
            for j, f in enumerate(freqs):
                self.setFreq(i, j, f)
            if alpha != math.inf:
                self.setAlpha(i, alpha)
            if invar:
                self.setWVar(i, wVar)
                self.setWInvar(i, wInvar)
