"""Extract the reference's likelihood golden vectors into a JSON fixture.

Reads test/Lik_*.cx (inputs: FASTA strings, Newick trees, model settings) and
test/Lik_*.cx.exp (expected printed outputs) from the reference checkout and
writes tests/golden/lik_golden.json.  Only data is written: alignments, trees,
the model operations each script performs, and the printed numbers with the
number of decimals they were printed with.

Run in the build container (the reference is not present on the GPU box):
    python tests/golden/make_lik_golden.py [/root/reference/crux/test]
"""

import json
import os
import re
import sys

TEST = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/crux/test"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lik_golden.json")


def read(name):
    with open(os.path.join(TEST, name)) as f:
        return f.read()


def fasta_strs(src):
    return [m.replace("\\\n", "") for m in re.findall(r'"""\\\n(.*?)"""', src, re.S)]


def trees_list(src):
    m = re.search(r"trees = \[(.*?)\]\n", src, re.S)
    return re.findall(r'"([^"]+)"', m.group(1)) if m else []


def nums(line):
    return [float(x) for x in re.findall(r"-?\d+\.\d+(?:e[-+]?\d+)?|-?inf", line)]


def decimals(tok):
    return len(tok.split(".")[1]) if "." in tok else 0


def lnl_lines(exp, prefix=r"lnL"):
    out = []
    for line in exp.splitlines():
        m = re.match(r"\s*" + prefix + r"[^:]*:\s*(-?\d+\.\d+)\s*$", line)
        if m:
            out.append((float(m.group(1)), decimals(m.group(1))))
    return out


def site_lines(exp, prefix="Site lnL's:"):
    out = []
    for line in exp.splitlines():
        if prefix in line:
            toks = line.split(":", 1)[1].split()
            out.append(([float(t) for t in toks], decimals(toks[0])))
    return out


cases = []

# Lik_b: JC lnL and site lnLs on 4 explicit trees.
src, exp = read("Lik_b.cx"), read("Lik_b.cx.exp")
fa = fasta_strs(src)[0]
L = lnl_lines(exp)
S = site_lines(exp)
for i, t in enumerate(trees_list(src)):
    cases.append({"name": "Lik_b[%d]" % i, "fasta": fa, "tree": t, "ops": [],
                  "evals": [{"root": None, "lnL": L[i][0], "dec": L[i][1],
                             "sites": S[i][0], "site_dec": S[i][1]}]})

# Lik_c: lnL at every node as root.
src, exp = read("Lik_c.cx"), read("Lik_c.cx.exp")
fa = fasta_strs(src)[0]
L = lnl_lines(exp)
S = site_lines(exp)
t = trees_list(src)[0]
cases.append({"name": "Lik_c", "fasta": fa, "tree": t, "ops": [],
              "evals": [{"root": k, "lnL": L[k][0], "dec": L[k][1],
                         "sites": S[k][0], "site_dec": S[k][1]}
                        for k in range(len(L))]})

# Lik_e: Gamma(4) means and medians, alpha in 0.5, 1, 5, 300.
src, exp = read("Lik_e.cx"), read("Lik_e.cx.exp")
fa = fasta_strs(src)[0]
L = lnl_lines(exp, r"lnL\(alpha")
k = 0
for i, t in enumerate(trees_list(src)):
    for med in (False, True):
        ev = []
        for a in (0.5, 1.0, 5.0, 300.0):
            ev.append({"root": None, "pre": [["setAlpha", 0, a]],
                       "lnL": L[k][0], "dec": L[k][1]})
            k += 1
        cases.append({"name": "Lik_e[%d,median=%s]" % (i, med), "fasta": fa,
                      "tree": t, "ctor": {"ncat": 4, "catMedian": med},
                      "ops": [], "evals": ev})

# Lik_g: zero branch lengths with conflicting data -> lnL < -50 at every root.
src = read("Lik_g.cx")
cases.append({"name": "Lik_g", "fasta": fasta_strs(src)[0],
              "tree": trees_list(src)[0], "ops": [],
              "evals": [{"root": k, "lnL_below": -50.0} for k in range(4)]})

# Lik_h: gaps are equivalent to N.
src, exp = read("Lik_h.cx"), read("Lik_h.cx.exp")
L = lnl_lines(exp)
k = 0
for j, fa in enumerate(fasta_strs(src)):
    for i, t in enumerate(trees_list(src)):
        cases.append({"name": "Lik_h[%d,%d]" % (j, i), "fasta": fa, "tree": t,
                      "ops": [], "evals": [{"root": None, "lnL": L[k][0],
                                            "dec": L[k][1]}]})
        k += 1

# Lik_j: F81 frequencies (0.4, 0.3, 0.2, 0.1), every root.
src, exp = read("Lik_j.cx"), read("Lik_j.cx.exp")
fa = fasta_strs(src)[0]
L = lnl_lines(exp)
S = site_lines(exp)
k = 0
for i, t in enumerate(trees_list(src)):
    ev = []
    for r in range(4):
        ev.append({"root": r, "lnL": L[k][0], "dec": L[k][1],
                   "sites": S[k][0], "site_dec": S[k][1]})
        k += 1
    cases.append({"name": "Lik_j[%d]" % i, "fasta": fa, "tree": t,
                  "ops": [["setFreq", 0, q, f] for q, f in
                          enumerate((0.4, 0.3, 0.2, 0.1))], "evals": ev})

# Lik_k: 2-category Gamma vs an equivalent 2-model mixture (3 decimals).  The
# trees come from RNJ and are known only as printed (4 significant digits).
src, exp = read("Lik_k.cx"), read("Lik_k.cx.exp")
fas = fasta_strs(src)
rtrees = re.findall(r"^(\(.*;)$", exp, re.M)
L2c = lnl_lines(exp, "lik2cat lnL")
L2m = lnl_lines(exp, "lik2mod lnL")
S2c = site_lines(exp, "lik2cat site lnL's:")
S2m = site_lines(exp, "lik2mod site lnL's:")
freqops = lambda m: [["setFreq", m, q, f] for q, f in
                     enumerate((0.4, 0.3, 0.2, 0.1))] + \
    [["setRclass", m, [0, 1, 2, 3, 4, 5], [0.1, 0.2, 0.3, 0.4, 0.5, 0.6]]]
for i, (fa, t) in enumerate(zip(fas, rtrees)):
    cases.append({"name": "Lik_k[%d,2cat]" % i, "fasta": fa, "tree": t,
                  "approx_tree": True, "ctor": {"nmodels": 1, "ncat": 2},
                  "ops": freqops(0) + [["setAlpha", 0, 0.179062]],
                  "evals": [{"root": None, "lnL": L2c[i][0], "dec": L2c[i][1],
                             "sites": S2c[i][0], "site_dec": S2c[i][1]}]})
    cases.append({"name": "Lik_k[%d,2mod]" % i, "fasta": fa, "tree": t,
                  "approx_tree": True, "ctor": {"nmodels": 2, "ncat": 1},
                  "ops": freqops(0) + freqops(1) +
                  [["setRmult", 0, 0.011437], ["setRmult", 1, 1.988563]],
                  "evals": [{"root": None, "lnL": L2m[i][0], "dec": L2m[i][1],
                             "sites": S2m[i][0], "site_dec": S2m[i][1]}]})

# Lik_a / Lik_d: RNJ trees (printed at 4 significant digits); Lik_d checks
# invariance under rate/frequency normalisation and mixtures of identical
# models.
src, exp = read("Lik_d.cx"), read("Lik_d.cx.exp")
fas = fasta_strs(src)
rtrees = re.findall(r"^(\(.*;)$", exp, re.M)
L = lnl_lines(exp)
seq = [[], [["setRate", 0, 0, 0.75]] + [["setFreq", 0, q, 0.2] for q in range(4)],
       [["setRate", 0, 0, 1.0], ["addModel", 0.0]],
       [["setWeight", 0, 0.5], ["setWeight", 1, 0.5]],
       [["setWeight", 0, 0.0], ["setWeight", 1, 0.3], ["addModel", 0.7]],
       [["setWeight", 0, 1.0], ["setWeight", 1, 1.0], ["setWeight", 2, 1.0]],
       [["delModel", 1]], [["delModel", 1]]]
k = 0
for i, (fa, t) in enumerate(zip(fas, rtrees)):
    ev = []
    for ops in seq:
        ev.append({"root": None, "pre": ops, "lnL": L[k][0], "dec": L[k][1]})
        k += 1
    cases.append({"name": "Lik_d[%d]" % i, "fasta": fa, "tree": t,
                  "approx_tree": True, "ops": [], "evals": ev})

# Lik_l: direct +I vs a rate-0 mixture model.
src, exp = read("Lik_l.cx"), read("Lik_l.cx.exp")
fa = fasta_strs(src)[0]
t = trees_list(src)[0]
L = lnl_lines(exp, r"\s*lnL")
S = site_lines(exp)
ev = []
for j, p in enumerate((0.0, 0.25, 0.5, 0.75)):
    ev.append({"root": None, "pre": [["setWVar", 0, 1.0 - p], ["setWInvar", 0, p]],
               "lnL": L[j][0], "dec": L[j][1], "sites": S[j][0],
               "site_dec": S[j][1]})
cases.append({"name": "Lik_l[direct]", "fasta": fa, "tree": t,
              "ctor": {"invar": True}, "ops": [], "evals": ev})
ev = []
for j, p in enumerate((0.0, 0.25, 0.5, 0.75)):
    ev.append({"root": None, "pre": [["setWeight", 0, 1.0 - p],
                                     ["setWeight", 1, p]],
               "lnL": L[4 + j][0], "dec": L[4 + j][1], "sites": S[4 + j][0],
               "site_dec": S[4 + j][1]})
cases.append({"name": "Lik_l[emulated]", "fasta": fa, "tree": t,
              "ctor": {"invar": True},
              "ops": [["setWInvar", 0, 0.0], ["addModel", 1.0],
                      ["setRate", 1, 0, 0.0]], "evals": ev})

# Lik_m: JC lnL of the script's alignment and of the simulated alignments the
# .exp prints (12 significant digits).  The last value depends on parameters
# drawn by the MC3 sampler that are never printed, so it is not used.
src, exp = read("Lik_m.cx"), read("Lik_m.cx.exp")
t = trees_list(src)[0]
fa0 = fasta_strs(src)[0]
blocks, vals, cur = [], [], {}
for line in exp.splitlines()[1:]:
    if re.match(r"^-?\d+\.\d+$", line.strip()):
        vals.append(line.strip())
        blocks.append(cur)
        cur = {}
        continue
    m = re.match(r"^([A-Z]\d?)\s+([A-Z ]+)$", line)
    if m and not line.startswith("("):
        cur[m.group(1)] = cur.get(m.group(1), "") + m.group(2).replace(" ", "")
order = ["A1", "A2", "B", "C", "D"]
for j in range(4):
    fa = fa0 if j == 0 else "".join(">%s\n%s\n" % (n, blocks[j][n]) for n in order)
    v = vals[j]
    sig = len(v.replace("-", "").replace(".", "").lstrip("0"))
    cases.append({"name": "Lik_m[%d]" % j, "fasta": fa, "tree": t, "ops": [],
                  "evals": [{"root": None, "lnL": float(v), "dec": decimals(v),
                             "sig": sig}]})

# Lik_i: lnL of each tree and of lik.dup() (an independent copy of the tree
# on the same alignment); trees canonized by the alignment's taxa order (the
# distance matrix's taxa map in the script).
src, exp = read("Lik_i.cx"), read("Lik_i.cx.exp")
fa = fasta_strs(src)[0]
L = lnl_lines(exp)
for i, t in enumerate(trees_list(src)):
    cases.append({"name": "Lik_i[%d]" % i, "fasta": fa, "tree": t, "ops": [],
                  "evals": [{"root": None, "lnL": L[2 * i][0], "dec": L[2 * i][1]},
                            {"root": None, "lnL": L[2 * i + 1][0], "dec": L[2 * i + 1][1],
                             "dup": True}]})

# Rendered trees after deroot() + canonize() (every script prints them with
# %.7e lengths before evaluating).
renders = []
for name in ("Lik_b", "Lik_c", "Lik_e", "Lik_g", "Lik_j", "Lik_l", "Lik_m"):
    src, exp = read(name + ".cx"), read(name + ".cx.exp")
    outs = re.findall(r"^(\(.*;)$", exp, re.M)
    ins = trees_list(src)
    # Lik_e prints each tree once; Lik_h prints every tree per alignment.
    taxa = re.findall(r"^>(\S+)", fasta_strs(src)[0], re.M)
    for tin, tout in zip(ins, outs):
        renders.append({"case": name, "newick": tin, "render": tout,
                        "taxa": taxa})

with open(OUT, "w") as f:
    json.dump({"source": "reference test/Lik_*.cx and test/Lik_*.cx.exp",
               "cases": cases, "renders": renders}, f, indent=1)
print("wrote %d cases to %s" % (len(cases), OUT))
