"""MI355X-native tree-likelihood engine with the reference's Lik call surface.

    from crux_amd import Tree, Alignment, Lik
    t = Tree("((A:0.1,B:0.2):0.05,C:0.3,D:0.4);", rooted=False)
    lik = Lik(t, Alignment(fasta_text), ncat=4)
    lik.setAlpha(0, 0.5)
    lik.lnL()
"""

from .alignment import Alignment, AminoAcid20, Character, Dna, Protein, parse_fasta
from .tree import Edge, Node, Ring, Tree

__all__ = ["Alignment", "AminoAcid20", "Character", "Dna", "Protein", "parse_fasta",
           "Tree", "Node", "Edge", "Ring", "Lik"]


def __getattr__(name):
    # Lik needs the native library; import it lazily so the data model is
    # usable without a build.
    if name == "Lik":
        from .lik import Lik
        return Lik
    raise AttributeError(name)
