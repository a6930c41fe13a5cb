"""Helpers that replay the reference's golden scripts against any object with
the reference Lik call surface (the oracle or the device Lik)."""

from crux_amd.alignment import Alignment
from crux_amd.tree import Tree


def build_tree(newick, taxa=None):
    """Parse, deroot and canonize by the alignment's taxa order, as every
    reference Lik test does."""
    t = Tree(newick)
    t.deroot()
    t.canonize(taxa)
    return t


def apply_ops(lik, ops):
    for op in ops:
        name, args = op[0], op[1:]
        getattr(lik, name)(*args)


def run_case(case, make_lik):
    """Yield (eval_spec, lnL, siteLnLs or None) for each evaluation."""
    aln = Alignment(case["fasta"])
    tree = build_tree(case["tree"], aln.taxa)
    lik = make_lik(tree, aln, **case.get("ctor", {}))
    apply_ops(lik, case["ops"])
    nodes = list(tree.nodes)
    for ev in case["evals"]:
        apply_ops(lik, ev.get("pre", []))
        root = None if ev["root"] is None else nodes[ev["root"]]
        if ev.get("dup"):
            # Lik.dup(): an independent tree copy on the same alignment
            # (Lik.pyx:642-660); objects without it evaluate a fresh
            # likelihood on a copy of the tree.
            if hasattr(lik, "dup"):
                d = lik.dup()
            else:
                d = make_lik(tree.dup(), aln, **case.get("ctor", {}))
                apply_ops(d, case["ops"])
            yield ev, d.lnL(), None
            continue
        lnl = lik.lnL(root)
        sites = lik.siteLnLs(root) if "sites" in ev else None
        yield ev, lnl, sites


def tol(ev, key="dec", approx=False):
    """Half a unit in the last printed decimal (plus slack for trees known
    only to 4 significant digits)."""
    t = 0.5 * 10.0 ** (-ev[key]) * 1.0001
    return max(t, 2e-3) if approx else t
