"""Character state codes and alignments: the leaf data of the likelihood.

* ``Character`` maps a state code to a bitmask over the primary states, with
  ambiguity codes as unions and the gap as 0 (pkg/Crux/Character.pyx:12-123).
  The likelihood treats a 0 mask as "any state" (pkg/Crux/Tree/Lik.pyx:1468).
* ``Dna`` and ``Protein`` carry the reference's alphabets
  (pkg/Crux/Character.pyx:130-254; Protein has 22 states including U and *).
  ``AminoAcid20`` is the standard 20-state alphabet used by 20-state models.
* ``Alignment`` holds a taxa x sites matrix of codes plus an integer weight
  per column (``freqs``), with ``pad`` and ``compact`` as in
  pkg/Crux/CTMatrix.pyx:563-662,856-1032.

The alignment is stored as a numpy uint8 matrix of ASCII codes, and
``encode()`` turns it into the dense per-taxon state-mask indices the device
consumes: one byte per (taxon, site) indexing a table of distinct masks.
"""

from __future__ import annotations

import numpy as np

__all__ = ["Character", "Dna", "Protein", "AminoAcid20", "Alignment",
           "parse_fasta"]


class Character:
    _instance = None

    def __init__(self, states, ambiguities, aliases):
        self.states = list(states)
        self.nstates = len(self.states)
        self.any = (1 << self.nstates) - 1
        self._val = {}
        self._code = {}
        for i, s in enumerate(self.states):
            self._val[s] = 1 << i
            self._code.setdefault(1 << i, s)
        for code, members in ambiguities.items():
            v = 0
            for m in members:
                v |= self._val[m]
            self._val[code] = v
            self._code.setdefault(v, code)
        for code, target in aliases.items():
            self._val[code] = self._val[target]
        # ASCII lookup table (-1 = undefined code).
        self.table = np.full(256, -1, dtype=np.int64)
        for c, v in self._val.items():
            self.table[ord(c)] = v

    @classmethod
    def get(cls):
        if cls.__dict__.get("_instance") is None:
            cls._instance = cls()
        return cls._instance

    def code2val(self, code):
        if code not in self._val:
            raise ValueError("State not defined: %r" % code)
        return self._val[code]

    def val2code(self, val):
        if val not in self._code:
            raise ValueError("Value not defined: %r" % val)
        return self._code[val]

    def pcodes(self):
        return list(self.states)

    def codes(self):
        return list(self._val.keys())


class Dna(Character):
    def __init__(self):
        super().__init__(
            ["A", "C", "G", "T"],
            {"K": "GT", "Y": "CT", "S": "CG", "B": "CGT", "W": "AT", "R": "AG",
             "D": "AGT", "M": "AC", "H": "ACT", "V": "ACG", "N": "ACGT",
             "-": ""},
            {"U": "T", "u": "T", "t": "T", "g": "G", "k": "K", "c": "C",
             "y": "Y", "s": "S", "b": "B", "a": "A", "w": "W", "r": "R",
             "d": "D", "m": "M", "h": "H", "v": "V", "n": "N", "?": "N",
             ".": "-"})


class Protein(Character):
    def __init__(self):
        st = list("ACDEFGHIKLMNPQRSTUVWY*")
        amb = {"B": "DN", "Z": "EQ", "X": st, "-": ""}
        ali = {c.lower(): c for c in "ABCDEFGHIKLMNPQRSTUVWXYZ"}
        ali.update({"?": "X", ".": "-"})
        super().__init__(st, amb, ali)


class AminoAcid20(Character):
    """The 20 standard amino acids (no U, no *)."""

    def __init__(self):
        st = list("ACDEFGHIKLMNPQRSTVWY")
        amb = {"B": "DN", "Z": "EQ", "X": st, "-": ""}
        ali = {c.lower(): c for c in "ABCDEFGHIKLMNPQRSTVWXYZ"}
        ali.update({"?": "X", ".": "-"})
        super().__init__(st, amb, ali)


def parse_fasta(text):
    """Return [(label, sequence)] from FASTA text (whitespace ignored)."""
    out = []
    label, seq = None, []
    for line in text.splitlines():
        line = line.strip()
        if not line:
            continue
        if line.startswith(">"):
            if label is not None:
                out.append((label, "".join(seq)))
            label = line[1:].strip()
            seq = []
        else:
            seq.append("".join(line.split()))
    if label is not None:
        out.append((label, "".join(seq)))
    return out


class Alignment:
    """Taxa x sites matrix with integer column weights.

    ``Alignment(fasta_text)`` or ``Alignment(rows=..., taxa=...)``;
    ``taxaMap`` maps label -> row index in input order.
    """

    def __init__(self, fasta=None, charType=Dna, taxa=None, rows=None,
                 freqs=None):
        self.charType = charType
        if fasta is not None:
            recs = parse_fasta(fasta)
            if not recs:
                raise ValueError("Invalid alignment size: 0x0")
            taxa = [r[0] for r in recs]
            n = len(recs[0][1])
            if n == 0:
                raise ValueError("Invalid alignment size: %dx0" % len(recs))
            for lab, s in recs:
                if len(s) != n:
                    raise ValueError("Inconsistent sequence lengths (%d vs. %d)"
                                     % (n, len(s)))
            rows = np.frombuffer("".join(r[1] for r in recs).encode("ascii"),
                                 dtype=np.uint8).reshape(len(recs), n).copy()
        if rows is None or taxa is None:
            raise ValueError("Specify fasta, or rows and taxa")
        rows = np.ascontiguousarray(rows, dtype=np.uint8)
        self.rows = rows
        self.taxa = list(taxa)
        self.taxaMap = {t: i for i, t in enumerate(self.taxa)}
        self.ntaxa, self.nchars = rows.shape
        self.npad = 0
        if freqs is None:
            freqs = np.ones(self.nchars, dtype=np.uint32)
        self.freqs = np.ascontiguousarray(freqs, dtype=np.uint32)
        ch = charType.get()
        bad = ch.table[self.rows] < 0
        if bad.any():
            i, j = np.argwhere(bad)[0]
            raise ValueError("State not defined: %r" % chr(self.rows[i, j]))

    def getRow(self, i):
        return self.rows[i]

    def render(self):
        w = max(len(t) for t in self.taxa)
        return "\n".join("%-*s %s" % (w, t, self.rows[i].tobytes().decode())
                         for i, t in enumerate(self.taxa))

    def pad(self, code, npad):
        """Append npad columns of ``code`` with weight 1
        (pkg/Crux/CTMatrix.pyx:610-662)."""
        assert npad > 0
        col = np.full((self.ntaxa, npad), ord(code), dtype=np.uint8)
        self.rows = np.ascontiguousarray(np.concatenate([self.rows, col], 1))
        self.freqs = np.concatenate(
            [self.freqs, np.ones(npad, dtype=np.uint32)])
        self.nchars += npad
        self.npad += npad

    def canonize(self):
        """Replace alias codes by their primary code."""
        ch = self.charType.get()
        trans = np.arange(256, dtype=np.uint8)
        for c in ch.codes():
            trans[ord(c)] = ord(ch.val2code(ch.code2val(c)))
        self.rows = trans[self.rows]

    def compact(self):
        """Merge identical columns into one pattern each, summing weights
        (pkg/Crux/CTMatrix.pyx:949-1032).  Patterns come out sorted by column
        contents, as the reference's sort of column strings leaves them."""
        self.canonize()
        cols = np.ascontiguousarray(self.rows.T)
        v = cols.view(np.dtype((np.void, cols.shape[1])))[:, 0]
        uniq, inv = np.unique(v, return_inverse=True)
        w = np.bincount(inv.ravel(), weights=self.freqs,
                        minlength=len(uniq)).astype(np.uint32)
        ucols = np.frombuffer(uniq.tobytes(), dtype=np.uint8).reshape(
            len(uniq), self.ntaxa)
        self.rows = np.ascontiguousarray(ucols.T)
        self.freqs = w
        self.nchars = len(uniq)
        self.npad = 0
        return 0

    def encode(self):
        """Per-taxon state masks as (codes[ntaxa, nchars] uint8 indices,
        masks[ncodes] int64).  A 0 (gap) mask becomes ``any``
        (pkg/Crux/Tree/Lik.pyx:1466-1475)."""
        ch = self.charType.get()
        vals = ch.table[self.rows]
        vals = np.where(vals == 0, ch.any, vals)
        masks, inv = np.unique(vals, return_inverse=True)
        if len(masks) > 256:
            raise ValueError("More than 256 distinct state sets")
        codes = inv.reshape(vals.shape).astype(np.uint8)
        return np.ascontiguousarray(codes), masks.astype(np.int64)

    def mask_vectors(self, masks):
        """0/1 leaf vectors (ncodes x nstates) for the given masks."""
        k = self.charType.get().nstates
        bits = (masks[:, None] >> np.arange(k)[None, :]) & 1
        return bits.astype(np.float64)
