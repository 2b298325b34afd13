"""Write tests/golden/small_strato_doc.json from KPP's user manual.

The fixtures are the reference-held known answers for the small_strato
mechanism in docs/source/output/code_f90.rst: parameter sizes (:194-211),
species order (:234-244), the generated Fun code (:388-409), the Jacobian
in closed form (:423-444) and the LU sparsity arrays (:501-506).  Usage:
    python tests/golden/make_small_strato_doc.py [KPP_HOME]
"""
import json
import re
import sys
from pathlib import Path

home = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
lines = (home / "docs/source/output/code_f90.rst").read_text().splitlines()
out = {}

# sizes table: "* - ``NAME``" ... "- value"
sizes = {}
for i, l in enumerate(lines):
    m = re.match(r"\s*\* - ``(\w+)``", l)
    if m and i + 2 < len(lines):
        v = lines[i + 2].strip()
        if v.startswith("- ") and v[2:].strip().isdigit():
            sizes[m.group(1)] = int(v[2:])
out["sizes"] = sizes

# species order
for l in lines:
    if "ind_O1D=1" in l:
        out["var_order"] = [s.split("=")[0].replace("ind_", "").strip()
                            for s in l.strip().split(",")]
    if "ind_M = 6" in l:
        out["fix_order"] = [s.split("=")[0].replace("ind_", "").strip()
                            for s in l.strip().rstrip(".").split(",")]

# Fun code: A(..) and Vdot(..) assignments (continuation lines joined)
code = []
in_fun = False
for l in lines:
    if "SUBROUTINE Fun (" in l:
        in_fun = True
    elif in_fun and "END SUBROUTINE Fun" in l:
        break
    elif in_fun:
        s = l.strip()
        if code and code[-1].endswith("&"):
            code[-1] = code[-1][:-1].rstrip() + s
        elif re.match(r"(A|Vdot)\(\d+\) =", s):
            code.append(s)
out["fun_f90"] = code

# Jacobian closed forms "\mathbf{J}(i,j) & = & terms\\"
jac = {}
cur = None
for l in lines:
    m = re.search(r"\\mathbf\{J\}\((\d),(\d)\) & = & (.*)", l)
    if m:
        cur = (int(m.group(1)), int(m.group(2)))
        jac[cur] = m.group(3)
    elif cur and l.strip() and not l.strip().startswith("\\end") and "&" not in l and "\\\\" in l:
        jac[cur] += " " + l.strip()
    elif l.strip().startswith("\\end"):
        cur = None
terms = {}
for (i, j), t in jac.items():
    t = t.replace("\\\\", "").replace("\\,", " ")
    lst = []
    for sgn, k, spc in re.findall(r"([+-]?)\s*k_\{(\d+)\}\s*(?:\[\{?([A-Za-z0-9_(^){}\\]+?)\}?\])?", t):
        spc = spc.replace("{", "").replace("}", "")
        spc = {"O(^1D)": "O1D", "O_3": "O3", "NO_2": "NO2", "O_2": "O2"}.get(spc, spc)
        lst.append([-1 if sgn == "-" else 1, int(k), [spc] if spc else []])
    terms[f"{i},{j}"] = lst
out["jacobian"] = terms

# LU arrays (1-based)
txt = "\n".join(lines)
for name in ("LU_ICOL", "LU_IROW", "LU_CROW", "LU_DIAG"):
    m = re.search(name + r" = \(/(.*?)/\)", txt, re.S)
    out[name] = [int(x) for x in re.findall(r"\d+", m.group(1))]

Path(__file__).with_name("small_strato_doc.json").write_text(json.dumps(out, indent=1))
print(json.dumps(out)[:600])
