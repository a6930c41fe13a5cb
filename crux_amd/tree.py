"""Phylogenetic tree model: nodes, edges and the directed half-edge rings that
the likelihood planner walks.

This mirrors the data model the reference's likelihood code is written
against (pkg/Crux/Tree/__init__.pyx:32-1076):

* an Edge owns two Rings, one per end (``ring`` and ``ring.other``);
* the Rings of one Node form a circular list through ``next``/``prev``, whose
  order is the order children are visited in by the planner
  (pkg/Crux/Tree/Lik.pyx:1491,1514-1529);
* ``Tree.base`` is the default root; ``Tree.deroot`` merges the two edges of a
  degree-2 root through a float32 sum (pkg/Crux/Tree/__init__.pyx:397-437).

Two additions serve the device mirror of the tree: every Node and Edge gets a
stable integer id (rings are ``2*edge.id + side``), and the tree keeps a
topology serial number ``sn`` and a branch-length serial number ``lsn`` plus
per-consumer sets of edges whose length changed, so a likelihood object can
push only what changed to the device.
"""

from __future__ import annotations

import re
import numpy as np

__all__ = ["Tree", "Node", "Edge", "Ring", "NewickError"]


class NewickError(SyntaxError):
    pass


class Ring:
    """One end of an Edge, attached to ``node``."""

    __slots__ = ("node", "edge", "other", "next", "prev", "side", "aux")

    def __init__(self, edge, side):
        self.node = None
        self.edge = edge
        self.other = None
        self.next = self
        self.prev = self
        self.side = side
        self.aux = None

    @property
    def id(self):
        return 2 * self.edge.id + self.side

    def __iter__(self):
        """All rings of this node, starting with self."""
        r = self
        while True:
            yield r
            r = r.next
            if r is self:
                return

    def siblings(self):
        """All rings of this node except self, in ring order."""
        r = self.next
        while r is not self:
            yield r
            r = r.next

    def __repr__(self):
        return "<Ring %d>" % self.id


class Node:
    __slots__ = ("tree", "ring", "taxon", "id", "aux")

    def __init__(self, tree):
        self.tree = tree
        self.ring = None
        self.taxon = None
        self.id = tree._new_node_id()
        self.aux = None

    @property
    def degree(self):
        if self.ring is None:
            return 0
        n = 0
        for _ in self.ring:
            n += 1
        return n

    def getDegree(self):
        return self.degree

    def getTaxon(self):
        return self.taxon

    def neighbors(self):
        if self.ring is None:
            return []
        return [r.other.node for r in self.ring]

    def __repr__(self):
        return "<Node %d %s>" % (self.id, self.taxon or "")


class Edge:
    __slots__ = ("tree", "ring", "_length", "id", "aux")

    def __init__(self, tree, length=0.0):
        self.tree = tree
        self.id = tree._new_edge_id()
        self.ring = Ring(self, 0)
        other = Ring(self, 1)
        self.ring.other = other
        other.other = self.ring
        self._length = float(length)
        self.aux = None

    @property
    def length(self):
        return self._length

    @length.setter
    def length(self, v):
        v = float(v)
        if v != self._length:
            self._length = v
            self.tree._length_changed(self)

    def attach(self, a, b):
        """Attach this edge between nodes a and b.  The new ring becomes the
        head of each node's ring list, inserted before the previous head, as
        in the reference (pkg/Crux/Tree/__init__.pyx Edge.attach)."""
        assert self.ring.node is None and self.ring.other.node is None
        assert a is not b
        for ring, node in ((self.ring, a), (self.ring.other, b)):
            ring.node = node
            head = node.ring
            if head is not None:
                prev = head.prev
                ring.next = head
                ring.prev = prev
                head.prev = ring
                prev.next = ring
            node.ring = ring
            self.tree._topo_touch(ring, ring.prev, node)
        self.tree.sn += 1
        self.tree.tsn += 1

    def detach(self):
        for ring in (self.ring, self.ring.other):
            node = ring.node
            nxt = ring.next
            if nxt is ring:
                node.ring = None
                self.tree._topo_touch(ring, None, node)
            else:
                if node.ring is ring:
                    node.ring = nxt
                prev = ring.prev
                nxt.prev = prev
                prev.next = nxt
                ring.next = ring
                ring.prev = ring
                self.tree._topo_touch(ring, prev, node)
            ring.node = None
        self.tree.sn += 1
        self.tree.tsn += 1

    def __repr__(self):
        return "<Edge %d %g>" % (self.id, self._length)


_TOKEN = re.compile(
    r"\s*(?:(\()|(\))|(,)|(:)|(;)|'((?:[^']|'')*)'|([^\s():;,\[\]']+)|(\[[^\]]*\]))")


class _NewickParser:
    """Recursive-descent Newick reader.  Children keep their Newick order in
    the parent's ring (pkg/Crux/Newick.pyx:176-192: each attach is followed by
    advancing the ring head)."""

    def __init__(self, tree, text):
        self.tree = tree
        self.toks = []
        pos = 0
        text = text.strip()
        while pos < len(text):
            m = _TOKEN.match(text, pos)
            if m is None or m.end() == pos:
                raise NewickError("Bad Newick input near %r" % text[pos:pos + 20])
            pos = m.end()
            if m.group(8) is not None:  # comment
                continue
            if m.group(6) is not None:
                self.toks.append(("label", m.group(6).replace("''", "'")))
            elif m.group(7) is not None:
                self.toks.append(("label", m.group(7).replace("_", " ")))
            else:
                for g, kind in ((1, "("), (2, ")"), (3, ","), (4, ":"), (5, ";")):
                    if m.group(g) is not None:
                        self.toks.append((kind, None))
                        break
        self.i = 0

    def peek(self):
        return self.toks[self.i][0] if self.i < len(self.toks) else None

    def take(self, kind=None):
        if self.i >= len(self.toks):
            raise NewickError("Unexpected end of Newick input")
        t = self.toks[self.i]
        if kind is not None and t[0] != kind:
            raise NewickError("Expected %r, got %r" % (kind, t))
        self.i += 1
        return t

    def label(self):
        if self.peek() == "label":
            return self.take()[1]
        return None

    def brlen(self):
        if self.peek() == ":":
            self.take()
            t = self.take("label")
            try:
                return float(t[1])
            except ValueError:
                raise NewickError("Bad branch length %r" % t[1])
        return None

    def subtree(self):
        if self.peek() == "(":
            node = self.descendants()
        else:
            node = None
        lab = self.label()
        bl = self.brlen()
        if node is None:
            node = Node(self.tree)
        if lab is not None:
            self.tree._label(node, lab)
        return node, (0.0 if bl is None else bl)

    def descendants(self):
        self.take("(")
        ret = Node(self.tree)
        while True:
            child, bl = self.subtree()
            e = Edge(self.tree, bl)
            e.attach(ret, child)
            ret.ring = ret.ring.next
            if self.peek() == ",":
                self.take()
                continue
            break
        self.take(")")
        return ret

    def parse(self):
        node = self.descendants() if self.peek() == "(" else None
        lab = self.label()
        bl = self.brlen()
        self.take(";")
        root = Node(self.tree)
        self.tree.base = root
        if node is None:
            if lab is None and bl is None:
                return
            node = Node(self.tree)
        if lab is not None:
            self.tree._label(node, lab)
        e = Edge(self.tree, 0.0 if bl is None else bl)
        e.attach(root, node)


class Tree:
    """Rooted or unrooted tree.  ``Tree(newick)`` parses a Newick string; like
    the reference, a tree built with ``rooted=False`` is derooted after
    parsing."""

    def __init__(self, with_=None, rooted=True):
        self._next_node_id = 0
        self._next_edge_id = 0
        self._base = None
        self.rooted = rooted
        self.sn = 0          # serial number of topology or base changes
        self.tsn = 0         # serial number of attach/detach only
        self.lsn = 0         # branch-length serial number
        self._sinks = []     # sets that collect edges whose length changed
        self._topo_sinks = []   # (rings, nodes) touched by attach/detach
        self.aux = None
        self._cache_sn = -1
        self._nodes = []
        self._edges = []
        self._taxa = []
        if isinstance(with_, str):
            self.rooted = True
            _NewickParser(self, with_).parse()
            if not rooted:
                self.deroot()
        elif with_ is not None:
            raise ValueError("Unexpected with_ argument")

    def __getstate__(self):
        # Edit sinks belong to the device mirrors of this object, not to a
        # copy.
        d = self.__dict__.copy()
        d["_sinks"] = []
        d["_topo_sinks"] = []
        return d

    # -- ids and serials ---------------------------------------------------
    def _new_node_id(self):
        i = self._next_node_id
        self._next_node_id += 1
        return i

    def _new_edge_id(self):
        i = self._next_edge_id
        self._next_edge_id += 1
        return i

    @property
    def max_node_id(self):
        return self._next_node_id

    @property
    def max_edge_id(self):
        return self._next_edge_id

    def _length_changed(self, edge):
        self.lsn += 1
        for sink in self._sinks:
            sink.add(edge)

    def length_sink(self):
        """A set that from now on collects every edge whose length changes
        (device mirrors drain it before each evaluation)."""
        s = set()
        self._sinks.append(s)
        return s

    def topology_sink(self):
        """A pair of sets (rings, nodes) that from now on collects every ring
        and node whose links attach/detach change (device mirrors send just
        those)."""
        s = (set(), set())
        self._topo_sinks.append(s)
        return s

    def _topo_touch(self, ring, prev, node):
        for rings, nodes in self._topo_sinks:
            rings.add(ring)
            if prev is not None:
                rings.add(prev)
            nodes.add(node)

    def _label(self, node, label):
        node.taxon = label

    # -- base ----------------------------------------------------------------
    @property
    def base(self):
        return self._base

    @base.setter
    def base(self, node):
        self._base = node
        self.sn += 1

    def getBase(self):
        return self._base

    def setBase(self, node):
        self.base = node

    # -- cached traversals (same order as the reference's _recache) ----------
    def _recache(self):
        nodes, edges, taxa = [], [], []
        b = self._base
        if b is not None:
            nodes.append(b)
            if b.taxon is not None:
                taxa.append(b.taxon)
            if b.ring is not None:
                stack = [r for r in reversed(list(b.ring))]
                while stack:
                    r = stack.pop()
                    edges.append(r.edge)
                    o = r.other
                    n = o.node
                    nodes.append(n)
                    if n.taxon is not None:
                        taxa.append(n.taxon)
                    stack.extend(reversed(list(o.siblings())))
        taxa.sort()
        self._nodes, self._edges, self._taxa = nodes, edges, taxa
        self._cache_sn = self.sn

    @property
    def nodes(self):
        if self._cache_sn != self.sn:
            self._recache()
        return self._nodes

    @property
    def edges(self):
        if self._cache_sn != self.sn:
            self._recache()
        return self._edges

    @property
    def taxa(self):
        if self._cache_sn != self.sn:
            self._recache()
        return self._taxa

    @property
    def ntaxa(self):
        return len(self.taxa)

    def getNodes(self):
        return self.nodes

    def getEdges(self):
        return self.edges

    def getTaxa(self):
        return self.taxa

    def clearAux(self):
        self.aux = None
        for n in self.nodes:
            n.aux = None
        for e in self.edges:
            e.aux = None
            e.ring.aux = None
            e.ring.other.aux = None

    # -- structure edits --------------------------------------------------
    def deroot(self):
        """Remove the root.  When that leaves a degree-2 node, splice it out
        and merge its two edges; the removed length is held in float32 before
        the sum, as the reference does (pkg/Crux/Tree/__init__.pyx:397-437)."""
        if not self.rooted:
            return
        self.rooted = False
        ring = self._base.ring if self._base is not None else None
        if ring is None:
            self.base = None
            return
        ring = ring.other
        node = ring.node
        ring.edge.detach()
        if node.degree == 2:
            ring = node.ring
            self.base = ring.other.node
            edge = ring.edge
            removed = float(np.float32(edge.length))
            edge.detach()
            ring = node.ring
            edge = ring.edge
            other = ring.other.node
            edge.detach()
            edge.attach(self._base, other)
            edge.length = edge.length + removed
            self._base.ring = self._base.ring.next
        else:
            self.base = node

    def canonize(self, taxa_order=None):
        """Re-order rings so that subtrees are visited in order of their
        minimum taxon (pkg/Crux/Tree/__init__.pyx:439-473).  ``taxa_order``
        maps taxon label -> rank; the default ranks labels alphabetically."""
        if taxa_order is None:
            taxa_order = {t: i for i, t in enumerate(sorted(self.taxa))}
        elif not isinstance(taxa_order, dict):
            taxa_order = {t: i for i, t in enumerate(taxa_order)}

        def rank(node):
            return taxa_order[node.taxon]

        def min_taxon(ring):
            ret = ring.node
            for r in ring.siblings():
                m = min_taxon(r.other)
                if ret.taxon is None or rank(m) < rank(ret):
                    ret = m
            return ret

        def canon(ring):
            node = ring.node
            ret = node
            if node.degree > 1:
                rings = []
                for r in list(ring.siblings()):
                    m = canon(r.other)
                    if ret.taxon is None or rank(m) < rank(ret):
                        ret = m
                    rings.append((rank(m), r))
                rings.sort(key=lambda x: x[0])
                for _, r in reversed(rings):
                    other = r.other.node
                    e = r.edge
                    e.detach()
                    e.attach(node, other)
                node.ring = ring
            return ret

        if self.rooted:
            base = self._base
            if base is None:
                return
        else:
            node = self._base
            if node is None or node.ring is None:
                return
            base = min_taxon(node.ring)
            other = min_taxon(node.ring.other)
            if base.taxon is None or rank(other) < rank(base):
                base = other
            self.base = base
        if base.ring is None:
            return
        canon(base.ring.other)

    def collapse(self):
        """Collapse internal edges of length <= 0 and clamp leaf edges to 0
        (pkg/Crux/Tree/__init__.pyx:475-517).  Returns the number removed."""
        if self._base is None:
            return 0
        if self._base.degree > 1:
            r = self._base.ring.other
            while r.next is not r:
                r = r.next.other
            self.base = r.node
        ring = self._base.ring
        if ring is None:
            return 0
        coll, clamp = [], []

        def walk(ring):
            for r in ring.siblings():
                walk(r.other)
            e = ring.edge
            if e.length <= 0.0:
                if ring.node.degree > 1 and ring.other.node.degree > 1:
                    coll.append(ring)
                else:
                    clamp.append(e)

        for r in ring:
            walk(r.other)
        for r in coll:
            node = r.node
            ro = r.other
            t = ro.next
            while t is not ro:
                e = t.edge
                n2 = t.other.node
                e.detach()
                e.attach(node, n2)
                t = ro.next
            r.edge.detach()
        for e in clamp:
            e.length = 0.0
        return len(coll)

    def dup(self):
        """Independent copy with the same topology, ring order and lengths."""
        t = Tree()
        t.rooted = self.rooted

        def copy(node, prev_ring):
            nn = Node(t)
            nn.taxon = node.taxon
            if node.ring is not None:
                for r in node.ring:
                    if r is prev_ring:
                        continue
                    o = r.other
                    child = copy(o.node, o)
                    e = Edge(t, r.edge.length)
                    e.attach(nn, child)
            return nn

        if self._base is not None:
            t._base = copy(self._base, None)
        t.sn += 1
        return t

    # -- output --------------------------------------------------------------
    def render(self, lengths=False, lengthFormat="%.7e", outFile=None):
        out = []

        def rr(node, via, zero, nolen):
            opened = False
            if node.ring is not None:
                for r in node.ring:
                    if r.edge is via:
                        continue
                    out.append("," if opened else "(")
                    opened = True
                    rr(r.other.node, r.edge, False, False)
                if opened:
                    out.append(")")
            if node.taxon is not None:
                lab = node.taxon
                if re.search(r"[_()[\]':;,]", lab):
                    out.append("'%s'" % lab.replace("'", "''"))
                else:
                    out.append(lab.replace(" ", "_"))
            if lengths and node.degree > 0:
                if zero:
                    out.append((":" + lengthFormat) % 0.0)
                elif not nolen:
                    out.append((":" + lengthFormat) % via.length)

        n = self._base
        if n is not None:
            if self.rooted:
                if n.degree == 1:
                    r = n.ring
                    rr(r.other.node, r.edge, False, False)
            else:
                d = n.degree
                if d == 0:
                    rr(n, None, False, False)
                elif d == 1:
                    nb = n.ring.other.node
                    if nb.degree > 1:
                        rr(nb, None, False, True)
                    else:
                        out.append("(")
                        rr(n, n.ring.edge, False, False)
                        out.append(",")
                        rr(nb, n.ring.edge, True, False)
                        out.append(")")
                else:
                    rr(n, None, False, True)
        out.append(";")
        s = "".join(out)
        if outFile is not None:
            outFile.write(s + "\n")
            return None
        return s

    # -- helpers for the device mirror -------------------------------------
    def topology_arrays(self, taxon_index):
        """Flatten the topology into integer arrays indexed by node id and
        ring id: ring_node, ring_next, node_ring, node_taxon (row index or
        -1), edge_len.  ``taxon_index`` maps labels to alignment rows."""
        nn = self.max_node_id
        ne = self.max_edge_id
        ring_node = np.full(2 * ne, -1, dtype=np.int32)
        ring_next = np.full(2 * ne, -1, dtype=np.int32)
        node_ring = np.full(nn, -1, dtype=np.int32)
        node_taxon = np.full(nn, -1, dtype=np.int32)
        edge_len = np.zeros(ne, dtype=np.float64)
        for n in self.nodes:
            if n.ring is not None:
                node_ring[n.id] = n.ring.id
                for r in n.ring:
                    ring_node[r.id] = n.id
                    ring_next[r.id] = r.next.id
            if n.taxon is not None:
                node_taxon[n.id] = taxon_index.get(n.taxon, -2)
        for e in self.edges:
            edge_len[e.id] = e.length
        return ring_node, ring_next, node_ring, node_taxon, edge_len
