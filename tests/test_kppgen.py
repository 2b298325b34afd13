"""The mechanism compiler against the reference's known answers.

small_strato: every number and code line that KPP's manual prints for this
mechanism (docs/source/output/code_f90.rst:194-506; fixture written by
tests/golden/make_small_strato_doc.py).  saprc99: the generated model is
compared with an independent dense evaluation of the mass-action ODE built
here from a separate parse of the equation file, and the sparsity patterns
with their defining invariants.
"""
import ctypes
import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import ref

GOLD = json.loads((Path(__file__).parent / "golden" / "small_strato_doc.json").read_text())


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_small_strato_sizes_and_order():
    s = ref.sizes("small_strato")
    for k in ("NVAR", "NFIX", "NSPEC", "NREACT", "NONZERO", "LU_NONZERO"):
        assert s[k] == GOLD["sizes"][k], k
    L = ref.lib("small_strato")
    names = [L.ros_ref_species_name(k).decode() for k in range(s["NSPEC"])]
    assert names == GOLD["var_order"] + GOLD["fix_order"]


def test_small_strato_lu_pattern():
    s = ref.sizes("small_strato")
    icol = np.zeros(s["LU_NONZERO"], np.int32)
    crow = np.zeros(s["NVAR"] + 1, np.int32)
    diag = np.zeros(s["NVAR"] + 1, np.int32)
    ref.lib("small_strato").ros_ref_lu_pattern(_p(icol), _p(crow), _p(diag))
    assert (icol + 1).tolist() == GOLD["LU_ICOL"]
    assert (crow + 1).tolist() == GOLD["LU_CROW"]
    assert (diag + 1).tolist() == GOLD["LU_DIAG"]
    irow = np.repeat(np.arange(s["NVAR"]), np.diff(crow))
    assert (irow + 1).tolist() == GOLD["LU_IROW"]


def _f90_to_c(line):
    # A(1) = RCT(1)*F(2) -> A[0] = RCT[0]*F[1]
    return re.sub(r"(\w+)\((\d+)\)", lambda m: f"{m.group(1)}[{int(m.group(2)) - 1}]", line)


def test_small_strato_fun_code_matches_manual():
    from kpp_amd.build import GEN
    text = (GEN / "small_strato_model.h").read_text()
    body = text[text.index("static void kpp_Fun("):]
    body = body[: body.index("\n}\n")]
    got = [l.strip().rstrip(";") for l in body.splitlines() if re.match(r"\s*(A|Vdot)\[", l)]
    got = [l.replace("2.0*", "2*") for l in got]
    assert got == [_f90_to_c(l) for l in GOLD["fun_f90"]]


def _species_index():
    L = ref.lib("small_strato")
    s = ref.sizes("small_strato")
    return {L.ros_ref_species_name(k).decode(): k for k in range(s["NSPEC"])}


def test_small_strato_jacobian_closed_forms():
    s = ref.sizes("small_strato")
    nv, nf = s["NVAR"], s["NFIX"]
    idx = _species_index()
    icol = np.zeros(s["LU_NONZERO"], np.int32)
    crow = np.zeros(nv + 1, np.int32)
    diag = np.zeros(nv + 1, np.int32)
    L = ref.lib("small_strato")
    L.ros_ref_lu_pattern(_p(icol), _p(crow), _p(diag))
    rng = np.random.default_rng(5)
    for _ in range(5):
        V = rng.uniform(0.5, 2.0, nv)
        F = rng.uniform(0.5, 2.0, nf)
        R = rng.uniform(0.5, 2.0, s["NREACT"])
        C = np.concatenate([V, F])
        J = np.zeros(s["LU_NONZERO"])
        L.ros_ref_jac_sp(_p(V), _p(F), _p(R), _p(J))
        seen = set()
        for key, terms in GOLD["jacobian"].items():
            i, j = (int(x) - 1 for x in key.split(","))
            expect = sum(sg * R[k - 1] * np.prod([C[idx[sp]] for sp in spc]) for sg, k, spc in terms)
            pos = [p for p in range(crow[i], crow[i + 1]) if icol[p] == j]
            assert len(pos) == 1
            assert J[pos[0]] == pytest.approx(expect, rel=1e-14), key
            seen.add(pos[0])
        # every other entry is an LU fill-in and must be zero
        for p in set(range(s["LU_NONZERO"])) - seen:
            assert J[p] == 0.0
        assert len(seen) == s["NONZERO"]


# ---------------------------------------------------------------------------
# saprc99: independent dense model from a separate parse of the KPP files
# ---------------------------------------------------------------------------
def _strip_comments(t):
    t = re.sub(r"\{[^}]*\}", " ", t)
    return re.sub(r"//[^\n]*", " ", t)


def _kpp_const(x):
    """A constant as KPP's C output carries it: the float value printed with
    "%g" (src/code_c.c:66) and read back by the C compiler as a double."""
    return float(f"{float(np.float32(x)):g}")


def _number_rate(rate):
    """Value of a rate that is one floating literal (src/scanner.c:604-628),
    else None."""
    txt = "".join(rate.split())
    try:
        return float(txt)
    except ValueError:
        return None


def _parse_eqn(path):
    """Equations as (reactants{name:coef}, products{name:coef}, rate text)."""
    t = _strip_comments(Path(path).read_text())
    t = t[t.upper().index("#EQUATIONS") + len("#EQUATIONS"):]
    eqs = []
    for item in t.split(";"):
        item = item.strip()
        if not item:
            continue
        item = re.sub(r"^<[^>]*>", "", item)
        lhs, rest = item.split("=", 1)
        rhs, rate = rest.split(":", 1)

        def terms(e):
            d = {}
            for sign, coef, name in re.findall(r"([+-]?)\s*([0-9.]+(?:[eE][+-]?\d+)?)?\s*([A-Za-z_]\w*)", e):
                c = float(coef) if coef else 1.0
                c = -c if sign == "-" else c
                d[name.upper()] = d.get(name.upper(), 0.0) + c
            return d
        eqs.append((terms(lhs), terms(rhs), rate))
    return eqs


def _dense(kpp_home, mech):
    eqs = _parse_eqn(kpp_home / "models" / f"{mech}.eqn")
    L = ref.lib(mech)
    s = ref.sizes(mech)
    names = [L.ros_ref_species_name(k).decode().upper() for k in range(s["NSPEC"])]
    return eqs, names, s


@pytest.fixture(scope="module")
def saprc99_dense(kpp_home):
    return _dense(kpp_home, "saprc99")


def _dense_fun_jac(eqs, names, s, C, R):
    """Vdot = S (k o prod c^nu) and its Jacobian, with stoichiometric and
    NUMBER-rate constants as KPP's generated code carries them."""
    nv = s["NVAR"]
    pos = {n: i for i, n in enumerate(names)}
    vdot = np.zeros(nv)
    J = np.zeros((nv, nv))
    for j, (lhs, rhs, rtext) in enumerate(eqs):
        react = {k: v for k, v in lhs.items() if k in pos}
        num = _number_rate(rtext) if "HV" not in lhs else None
        k_j = R[j] if num is None else _kpp_const(num)
        rate = k_j * np.prod([C[pos[k]] ** int(v) for k, v in react.items()])
        net = {}
        for k, v in lhs.items():
            net[k] = net.get(k, 0.0) - v
        for k, v in rhs.items():
            net[k] = net.get(k, 0.0) + v
        for k, v in net.items():
            if k in pos and pos[k] < nv:
                vdot[pos[k]] += _kpp_const(v) * rate
        for k, v in react.items():
            q = pos[k]
            if q >= nv:
                continue
            others = np.prod([C[pos[kk]] ** int(vv) for kk, vv in react.items() if kk != k])
            d = k_j * v * C[q] ** (int(v) - 1) * others
            for kk, vv in net.items():
                if kk in pos and pos[kk] < nv:
                    J[pos[kk], q] += _kpp_const(vv) * d
    return vdot, J


def test_saprc99_sizes(saprc99_dense):
    eqs, names, s = saprc99_dense
    assert s["NVAR"] == 74 and s["NFIX"] == 5 and s["NREACT"] == 211 == len(eqs)
    assert set(names[s["NVAR"]:]) == {"AIR", "O2", "H2O", "H2", "CH4"}


def test_saprcnov_sizes(kpp_home):
    eqs, names, s = _dense(kpp_home, "saprcnov")
    # models/saprcnov.spc declares N2 fixed, but no equation uses it
    assert s["NVAR"] == 88 and s["NFIX"] == 5 and s["NREACT"] == 235 == len(eqs)
    assert set(names[s["NVAR"]:]) == {"AIR", "O2", "H2O", "H2", "CH4"}


@pytest.mark.parametrize("mech", ["saprc99", "saprcnov"])
def test_saprc_fun_and_jac_match_dense_model(kpp_home, mech):
    eqs, names, s = _dense(kpp_home, mech)
    L = ref.lib(mech)
    nv, nf, nr = s["NVAR"], s["NFIX"], s["NREACT"]
    icol = np.zeros(s["LU_NONZERO"], np.int32)
    crow = np.zeros(nv + 1, np.int32)
    diag = np.zeros(nv + 1, np.int32)
    L.ros_ref_lu_pattern(_p(icol), _p(crow), _p(diag))
    rng = np.random.default_rng(11)
    for _ in range(3):
        C = rng.uniform(0.5, 2.0, nv + nf)
        R = rng.uniform(0.5, 2.0, nr)
        V, F = C[:nv].copy(), C[nv:].copy()
        vd = np.zeros(nv)
        L.ros_ref_fun(_p(V), _p(F), _p(R), _p(vd))
        jv = np.zeros(s["LU_NONZERO"])
        L.ros_ref_jac_sp(_p(V), _p(F), _p(R), _p(jv))
        vdot, J = _dense_fun_jac(eqs, names, s, C, R)
        scale = np.abs(vdot).max()
        assert np.allclose(vd, vdot, rtol=1e-12, atol=1e-12 * scale)
        Jg = np.zeros((nv, nv))
        for i in range(nv):
            for p in range(crow[i], crow[i + 1]):
                Jg[i, icol[p]] = jv[p]
        assert np.allclose(Jg, J, rtol=1e-12, atol=1e-12 * np.abs(J).max())


@pytest.mark.parametrize("mech", ["saprc99", "saprcnov"])
def test_saprc_pattern_invariants(mech):
    s = ref.sizes(mech)
    nv = s["NVAR"]
    icol = np.zeros(s["LU_NONZERO"], np.int32)
    crow = np.zeros(nv + 1, np.int32)
    diag = np.zeros(nv + 1, np.int32)
    ref.lib(mech).ros_ref_lu_pattern(_p(icol), _p(crow), _p(diag))
    P = np.zeros((nv, nv), bool)
    for i in range(nv):
        cols = icol[crow[i]:crow[i + 1]]
        assert (np.diff(cols) > 0).all()          # row-major, columns ascending
        assert icol[diag[i]] == i                  # diagonal present and indexed
        P[i, cols] = True
    # closed under symbolic elimination: no fill-in left to create
    F = P.copy()
    for k in range(nv):
        rows = np.nonzero(F[k + 1:, k])[0] + k + 1
        for r in rows:
            F[r, k:] |= F[k, k:]
    assert (F == P).all()


# ------------------------------------------------------ commands it refuses
KPPGEN = Path(__file__).resolve().parents[1] / "kpp_amd" / "_obj" / "kppgen"

_BASE = """#DEFVAR
X = IGNORE;
Y = IGNORE;
#DEFFIX
P = IGNORE;
#EQUATIONS
<R1> X + P = Y : TEMP;
<R2> Y = X : 2.0E-3;
"""


def _kppgen_on(tmp_path, text, *opts):
    if not KPPGEN.exists():
        pytest.skip("kppgen not built")
    src = tmp_path / "m.kpp"
    src.write_text(text)
    out = tmp_path / "out"
    out.mkdir(exist_ok=True)
    return subprocess.run([str(KPPGEN), "--root", "m", "-o", str(out), *opts, str(src)],
                          capture_output=True, text=True)


def test_kppgen_compiles_plain_mechanism(tmp_path):
    r = _kppgen_on(tmp_path, _BASE + "#LOOKATALL\n#MONITOR X;\n#INTEGRATOR rosenbrock\n")
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("extra,why", [
    ("#FAMILIES\nFX : X + Q;\n", "FX"),  # a member that is no species
    ("#INLINE C_RCONST\n  RCONST[0] = 1.0;\n#ENDINLINE\n", "C_RCONST"),
    ("#INLINE C_RATES\ndouble k(double t) { return t; }\n#ENDINLINE\n", "C_RATES"),
    ("#FUNCTION SPARSE\n", "FUNCTION"),
    ("#JACOBIAN FULL\n", "JACOBIAN"),
    ("#DOUBLE OFF\n", "DOUBLE"),
    ("#FLUX ON\n", "FLUX"),
    ("#NOSUCHCOMMAND\n", "NOSUCHCOMMAND"),
])
def test_kppgen_refuses_what_changes_the_model(tmp_path, extra, why):
    """Commands that would change the model functions and that kppgen does not
    implement are errors, not silently skipped."""
    r = _kppgen_on(tmp_path, _BASE + extra)
    assert r.returncode != 0
    assert why in r.stderr + r.stdout


def test_kppgen_reorder_off(tmp_path):
    r = _kppgen_on(tmp_path, _BASE + "#REORDER OFF\n")
    assert r.returncode == 0, r.stderr
