"""#FAMILIES (src/scan.y:244-268, src/scanner.c:921-1097) and #FUNCTION SPLIT
(src/gen.c:713-812) in kppgen, on the strato_fam test mechanism
(kpp_amd/mechs/strato_fam.kpp: small_strato with production and loss
families of odd oxygen Ox = O + O1D + O3, compiled with #FUNCTION SPLIT).

What pins it: per reaction the family species gets max(P - L, 0) (POx) or
max(L - P, 0) (LOx) of the members' net production, so for any state
POx' - LOx' = O' + O1D' + O3' (checked on the generated Fun at random states),
and the oracle's integration keeps POx - LOx - Ox at its initial value (a
linear invariant, which Rosenbrock methods preserve up to rounding); the
family species are not reactive and read by no rate, so the other species'
ODEs are small_strato's (same Fun values at the same state); the reference's
quirks: member coefficients truncated to integers, families declared before
the equations see none, a name without a P/L prefix refused.
Fun_SPLIT: P - D*V equals Fun's value up to rounding at random states, the
loss frequencies D are those of the reactions consuming each species, and
the F90 flavour (which calls Fun_SPLIT) and the C integrator (Fun) reach the
same state within the tolerances; a mechanism without SPLIT has none.
"""
import subprocess

import numpy as np
import pytest

from kpp_amd import build, workloads
from oracle import ref


@pytest.fixture(scope="module", autouse=True)
def _have_mech():
    try:
        ref.sizes("strato_fam")
    except (FileNotFoundError, OSError):
        pytest.skip("oracle for strato_fam not built")


def _names(mech):
    n = ref.sizes(mech)["NSPEC"]
    return [ref.species_name(mech, k) for k in range(n)]


def _fun(mech, V, F, R):
    L = ref.lib(mech)
    out = np.zeros(ref.sizes(mech)["NVAR"])
    L.ros_ref_fun(ref._p(np.ascontiguousarray(V)), ref._p(np.ascontiguousarray(F)),
                  ref._p(np.ascontiguousarray(R)), ref._p(out))
    return out


def test_family_species_and_fun():
    nf, ns = _names("strato_fam"), _names("small_strato")
    szf, szs = ref.sizes("strato_fam"), ref.sizes("small_strato")
    assert szf["NVAR"] == szs["NVAR"] + 2 and szf["NREACT"] == szs["NREACT"]
    assert "POx" in nf and "LOx" in nf
    rng = np.random.default_rng(1)
    for _ in range(20):
        st = {n: 10.0 ** rng.uniform(0, 12) for n in ns}
        R = 10.0 ** rng.uniform(-16, -2, szs["NREACT"])
        vf = np.array([st.get(n, 0.0) for n in nf[:szf["NVAR"]]])
        ff = np.array([st[n] for n in nf[szf["NVAR"]:]])
        vs = np.array([st[n] for n in ns[:szs["NVAR"]]])
        fs = np.array([st[n] for n in ns[szs["NVAR"]:]])
        df, ds = _fun("strato_fam", vf, ff, R), _fun("small_strato", vs, fs, R)
        i = {n: k for k, n in enumerate(nf)}
        j = {n: k for k, n in enumerate(ns)}
        for n in ns[:szs["NVAR"]]:
            # the same terms; the species order changes the order of each
            # product's factors, so a few ULPs apart
            assert df[i[n]] == pytest.approx(ds[j[n]], rel=1e-14, abs=0.0), n
        ox = df[i["O"]] + df[i["O1D"]] + df[i["O3"]]
        # the tendencies nearly cancel in both sums: the bar is relative to them
        gross = (df[i["POx"]] + df[i["LOx"]] + abs(df[i["O"]]) + abs(df[i["O1D"]]) +
                 abs(df[i["O3"]]))
        assert df[i["POx"]] - df[i["LOx"]] == pytest.approx(ox, rel=0.0, abs=1e-12 * gross)
        assert df[i["POx"]] >= 0 and df[i["LOx"]] >= 0


def test_integration_keeps_the_family_balance():
    b = workloads.strato_fam_tod(ncells=16)
    nf = _names("strato_fam")
    i = {n: k for k, n in enumerate(nf)}
    o = ref.integrate(b)
    assert (o["ierr"] == 1).all()
    C0, C1 = b.C, o["C"]
    for c in range(b.ncells):
        inv0 = C0[i["POx"], c] - C0[i["LOx"], c] - (C0[i["O"], c] + C0[i["O1D"], c] + C0[i["O3"], c])
        inv1 = C1[i["POx"], c] - C1[i["LOx"], c] - (C1[i["O"], c] + C1[i["O1D"], c] + C1[i["O3"], c])
        scale = C1[i["O3"], c] + C1[i["POx"], c] + C1[i["LOx"], c]
        # rounding of ~100 stiff solves; a family term missed drifts by O(1)
        assert abs(inv1 - inv0) <= 1e-9 * scale, c
        assert C1[i["POx"], c] >= 0 and C1[i["LOx"], c] >= 0  # zero at night
    assert (C1[i["POx"]] > 0).any() and (C1[i["LOx"]] > 0).any()


def _kppgen(tmp_path, text):
    src = tmp_path / "fam.kpp"
    src.write_text(text)
    exe = build.OBJ / "kppgen"
    return subprocess.run([str(exe), "--kpp-home", str(build.kpp_home()), "--root", "fam",
                           "-o", str(tmp_path), str(src)], capture_output=True, text=True)


def test_reference_quirks(tmp_path, kpp_home):
    base = "#INCLUDE small_strato.spc\n#INCLUDE small_strato.eqn\n"
    # a coefficient below 1 truncates to 0: no family species
    r = _kppgen(tmp_path, base + "#FAMILIES\n  POx : 0.5 O3;\n")
    assert r.returncode == 0, r.stderr
    par = (tmp_path / "fam_Parameters.h").read_text()
    assert "ind_POx" not in par and "NVAR                 5" in par and "NFAM                 1" in par
    # families before the equations see none
    r = _kppgen(tmp_path, "#INCLUDE small_strato.spc\n#FAMILIES\n  POx : O3;\n"
                          "#INCLUDE small_strato.eqn\n")
    assert r.returncode == 0 and "ind_POx" not in (tmp_path / "fam_Parameters.h").read_text()
    # "2 O" counts twice: R1 (O2 + hv = 2O) makes 4 PO2x
    r = _kppgen(tmp_path, base + "#FAMILIES\n  PO2x : 2 O;\n")
    assert r.returncode == 0
    model = (tmp_path / "fam_model.h").read_text()
    assert "4.0*A[0]" in model
    # the prefix decides the family type
    r = _kppgen(tmp_path, base + "#FAMILIES\n  Ox : O3;\n")
    assert r.returncode != 0 and "P/L" in r.stderr


def test_fun_split_against_fun():
    L = ref.lib("strato_fam")
    sz = ref.sizes("strato_fam")
    nv, nf = sz["NVAR"], sz["NFIX"]
    rng = np.random.default_rng(3)
    for _ in range(20):
        V = 10.0 ** rng.uniform(0, 12, nv)
        F = 10.0 ** rng.uniform(14, 17, nf)
        R = 10.0 ** rng.uniform(-16, -2, sz["NREACT"])
        agg = _fun("strato_fam", V, F, R)
        spl = np.zeros(nv)
        assert L.ros_ref_fun_split(ref._p(V), ref._p(F), ref._p(R), ref._p(spl)) == 1
        # P - D*V against the chain of signed terms: the same value up to the
        # rounding of the larger of production and loss
        A = np.abs(agg) + np.abs(spl)
        assert np.all(np.abs(spl - agg) <= 1e-9 * np.maximum(A, 1e-9 * A.max())), (spl, agg)
    small = np.zeros(ref.sizes("small_strato")["NVAR"])
    assert ref.lib("small_strato").ros_ref_fun_split(ref._p(small), ref._p(small), ref._p(small),
                                                     ref._p(small)) == 0


def test_f90_flavour_with_fun_split():
    b = workloads.strato_fam_tod(ncells=8)
    b.rtol = np.array([1e-6])
    ic = np.zeros(20, np.int32)
    ic[1], ic[2] = 1, 2  # scalar tolerances, Ros3
    o90 = ref.integrate_f90(b, ic)
    oc = ref.integrate(b)
    assert (o90["ierr"] == 1).all() and (oc["ierr"] == 1).all()
    nv = ref.sizes("strato_fam")["NVAR"]
    big = np.abs(oc["C"][:nv]) > 1e3
    assert np.allclose(o90["C"][:nv][big], oc["C"][:nv][big], rtol=1e-3)
