"""Writes element_matrices.json: the exact element matrices and assembled
stencils that the reference publishes as text in its TikZ figures
(docs/figures/fem_stencils_q1.tex, docs/figures/fem_stencils_q2_elements.tex).

Run in the build container, where the reference tree is readable:
    python tests/golden/make_element_fixtures.py /root/reference
The JSON holds only the numbers (fractions as strings) and the node orderings
the figures use; the tests read the JSON, never the reference.
"""
import json
import os
import re
import sys
from fractions import Fraction

FRAC = re.compile(r"\\frac\{(\d+)\}\{(\d+)\}")


def parse_value(tok):
    tok = tok.replace(" ", "")
    tok = FRAC.sub(r"\1/\2", tok)
    return str(Fraction(tok))


def parse_pmatrix(body):
    rows = [r for r in body.split(r"\\") if r.strip()]
    return [[parse_value(t) for t in r.split("&")] for r in rows]


def matrices(text):
    """label -> (scale, matrix) for every '\\node[lab] ... {$X = s$}' followed by a pmatrix."""
    out = {}
    lab = re.compile(r"\\node\[lab\] at \(([-\d.]+),([-\d.]+)\) \{\$(.*?)\$\};")
    mat = re.compile(r"\\begin\{pmatrix\}(.*?)\\end\{pmatrix\}", re.S)
    for m in lab.finditer(text):
        label = m.group(3)
        name, _, scale = label.partition("=")
        name = name.strip().replace("\\,", "").replace(" ", "")
        scale = scale.replace("\\,", "").replace("\\times", "").strip()
        mm = mat.search(text, m.end())
        out[name] = {"scale": scale, "matrix": parse_pmatrix(mm.group(1))}
    return out


def q1_stencils(text):
    """label -> 3x3 assembled stencil [dy+1][dx+1] from the scopes shifted to x = 11.6."""
    out = {}
    names = [n for n in matrices(text)]
    scopes = re.findall(r"\\begin\{scope\}\[shift=\{\(11\.6,([-\d.]+)\)\}\](.*?)\\end\{scope\}", text, re.S)
    node = re.compile(r"\\node\[stencil(?:,central)?\] at \(([-\d.]+),([-\d.]+)\) \{\$(.*?)\$\};")
    for name, (_, body) in zip(names, scopes):
        st = [["0"] * 3 for _ in range(3)]
        for x, y, v in node.findall(body):
            i = int(round(float(x) / 0.75)) + 1
            j = int(round(float(y) / 0.75)) + 1
            st[j][i] = parse_value(v)
        out[name] = st
    return out


def main(ref):
    fig = os.path.join(ref, "docs", "figures")
    q1 = open(os.path.join(fig, "fem_stencils_q1.tex")).read()
    q2 = open(os.path.join(fig, "fem_stencils_q2_elements.tex")).read()
    data = {
        "source": "docs/figures/fem_stencils_q1.tex, docs/figures/fem_stencils_q2_elements.tex",
        "q1": {
            "node_order": [[0, 0], [1, 0], [1, 1], [0, 1]],
            "elements": matrices(q1),
            "stencils": q1_stencils(q1),
        },
        "q2": {
            "node_order": [[0, 0], [2, 0], [2, 2], [0, 2], [1, 0], [2, 1], [1, 2], [0, 1], [1, 1]],
            "elements": matrices(q2),
        },
    }
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "element_matrices.json"), "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", os.path.join(here, "element_matrices.json"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
