"""KPP's on-disk formats written by this project, pinned on the CPU:

* <root>_Parameters.h and <root>_JacobianSP.c (kppgen; the layout of KPP's C
  output, src/gen.c:528-580 and :2627-2700) compile into a C program whose
  sizes, species indices and LU arrays are the ones the KPP manual prints for
  small_strato (docs/source/output/code_f90.rst:186-244, 476-506, fixture
  tests/golden/small_strato_doc.json; the C indices are the manual's Fortran
  indices minus one, src/gen.c:4116-4121) and, for saprc99, the pattern the
  library itself uses;
* ROOT.dat lines (kpp_amd.datfile.SaveData) are byte for byte what a C
  program printing with util/util.c:31-36's formats writes.
"""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from kpp_amd import datfile, solver

REPO = Path(__file__).resolve().parents[1]
KPPGEN = REPO / "kpp_amd" / "_obj" / "kppgen"


def _gen(tmp_path, kpp_home, root, rel):
    if not KPPGEN.exists():
        pytest.skip("kppgen not built")
    subprocess.run([str(KPPGEN), "--kpp-home", str(kpp_home), "--root", root, "-o",
                    str(tmp_path), str(kpp_home / rel)], check=True, capture_output=True)
    return tmp_path / f"{root}_Parameters.h", tmp_path / f"{root}_JacobianSP.c"


def _run_c(tmp_path, hdr, jsp, body):
    src = tmp_path / "main.c"
    src.write_text(f'#include <stdio.h>\n#include "{hdr.name}"\n'
                   "extern int LU_IROW[], LU_ICOL[], LU_CROW[], LU_DIAG[];\n"
                   f"int main(void) {{\n{body}\nreturn 0;\n}}\n")
    exe = tmp_path / "main"
    subprocess.run(["gcc", "-Wall", "-Werror", "-I", str(tmp_path), str(src), str(jsp), "-o",
                    str(exe)], check=True, capture_output=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


ARRAYS = ('for (int i = 0; i < LU_NONZERO; i++) printf("%d %d ", LU_IROW[i], LU_ICOL[i]);'
          'printf("\\n");'
          'for (int i = 0; i <= NVAR; i++) printf("%d %d ", LU_CROW[i], LU_DIAG[i]);'
          'printf("\\n");')


def test_small_strato_files_match_the_manual(tmp_path, kpp_home):
    hdr, jsp = _gen(tmp_path, kpp_home, "small_strato", "models/small_strato.def")
    doc = json.loads((REPO / "tests" / "golden" / "small_strato_doc.json").read_text())
    body = ('printf("%d %d %d %d %d %d %d %d\\n", NSPEC, NVAR, NFIX, NREACT, NONZERO, '
            'LU_NONZERO, NVARST, NFIXST);'
            'printf("%d %d %d %d %d %d %d %d %d\\n", ind_O1D, ind_O, ind_O3, ind_NO, ind_NO2, '
            'ind_M, ind_O2, indf_M, indf_O2);' + ARRAYS)
    lines = _run_c(tmp_path, hdr, jsp, body).split("\n")
    sz = doc["sizes"]
    assert list(map(int, lines[0].split())) == [sz["NSPEC"], sz["NVAR"], sz["NFIX"],
                                                 sz["NREACT"], sz["NONZERO"],
                                                 sz["LU_NONZERO"], 0, sz["NVAR"]]
    order = doc["var_order"] + doc["fix_order"]
    assert order == ["O1D", "O", "O3", "NO", "NO2", "M", "O2"]
    assert list(map(int, lines[1].split())) == [0, 1, 2, 3, 4, 5, 6, 0, 1]
    a = list(map(int, lines[2].split()))
    assert a[0::2] == [v - 1 for v in doc["LU_IROW"]]
    assert a[1::2] == [v - 1 for v in doc["LU_ICOL"]]
    b = list(map(int, lines[3].split()))
    assert b[0::2] == [v - 1 for v in doc["LU_CROW"]]
    assert b[1::2] == [v - 1 for v in doc["LU_DIAG"]]


def test_saprc99_files_match_the_library(tmp_path, kpp_home):
    hdr, jsp = _gen(tmp_path, kpp_home, "saprc99", "models/saprc99.def")
    body = 'printf("%d %d %d %d %d\\n", NSPEC, NVAR, NFIX, NREACT, LU_NONZERO);' + ARRAYS
    lines = _run_c(tmp_path, hdr, jsp, body).split("\n")
    st = json.loads((solver.GEN / "saprc99_stats.json").read_text())
    assert list(map(int, lines[0].split())) == [st["NSPEC"], st["NVAR"], st["NFIX"],
                                                 st["NREACT"], st["LU_NONZERO"]]
    a = list(map(int, lines[1].split()))
    b = list(map(int, lines[2].split()))
    assert a[1::2] == st["lu_icol"] and b[0::2] == st["lu_crow"] and b[1::2] == st["lu_diag"]
    # LU_IROW is the row of each entry of the row-compressed pattern
    rows = np.repeat(np.arange(st["NVAR"]), np.diff(st["lu_crow"]))
    assert a[0::2] == list(rows)


def test_root_dat_lines_match_util_c(tmp_path):
    species = ["O1D", "O", "O3", "NO", "NO2", "M", "O2"]
    lookat = ["O3", "NO", "NO2", "M"]
    rng = np.random.default_rng(0)
    C = 10.0 ** rng.uniform(-5, 18, (7, 3))
    C[2, 1] = 0.0
    times = [0.0, 5400.0, 86399.0 + 1.0 / 3.0]
    cf = np.array([1.0, 2.5e19, 3.1e-7])
    with datfile.SaveData(str(tmp_path / "box{cell}.dat"), species, lookat, [0, 1, 2]) as sd:
        for t in times:
            sd.save(t, C, cf)
    # the same lines from C's printf with util/util.c's formats
    vals = ",".join("%.17g" % v for v in C.ravel())
    idx = ",".join(str(species.index(s)) for s in lookat)
    src = tmp_path / "dat.c"
    src.write_text(
        "#include <stdio.h>\n"
        f"static const double C[] = {{{vals}}};\n"
        f"static const int L[] = {{{idx}}};\n"
        f"static const double T[] = {{{','.join('%.17g' % t for t in times)}}};\n"
        f"static const double CF[] = {{{','.join('%.17g' % c for c in cf)}}};\n"
        "int main(void) { for (int c = 0; c < 3; c++) for (int t = 0; t < 3; t++) {\n"
        '  printf("%6.1f ", T[t]/3600.0);\n'
        f'  for (int i = 0; i < {len(lookat)}; i++) printf("%24.16e ", C[L[i]*3 + c]/CF[c]);\n'
        '  printf("\\n"); } return 0; }\n')
    exe = tmp_path / "dat"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True, capture_output=True)
    want = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = "".join((tmp_path / f"box{c}.dat").read_text() for c in range(3))
    assert got == want


def test_save_data_rejects_unknown_species(tmp_path):
    with pytest.raises(KeyError):
        datfile.SaveData(str(tmp_path / "x.dat"), ["A", "B"], ["C"], [0])
