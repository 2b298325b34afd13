"""The mechanism compiler's Fortran 90 rate code (MCM-style mechanisms).

KPP's F90 back end pastes rate expressions and #INLINE F90_RCONST code into
Update_RCONST, and util/Makefile_f90:38-39 compiles it with gfortran without
-fdefault-real-8, so a real literal without a D exponent is single precision
and an operation on two such literals is done in single precision.  These
tests pin the translation against an independent evaluation: Python scalars
(C doubles, glibc's exp/pow/log10/cos/acos through the math module) with
numpy.float32 standing in for default-REAL literals and their products.  A
translator that kept literals in double, or promoted a single-precision
product early, fails them bit for bit.
"""
import ctypes
import json
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
KPPGEN = REPO / "kpp_amd" / "_obj" / "kppgen"

MECH = """
#LANGUAGE FORTRAN90
#DEFVAR
A = IGNORE ;
B = IGNORE ;
R1 = IGNORE ;
#INLINE F90_RCONST
  ! a sum of concentrations, continued over two lines
  RSUM = C(ind_R1) + &
      C(ind_A)
  CALL define_consts
#ENDINLINE
#EQUATIONS
<1> A = B : 2.40E-12*0.8*RSUM ;
<2> B = A : 5.6E-34*N2*(TEMP/300.)**(-2.6)*O2 ;
<3> A + hv = B : J(J_ONE)+J(J_TWO) ;
<4> R1 = A : KMT ;
<5> B = R1 : 3.0E7*EXP(-5300./TEMP) ;
<6> R1 = B : 1.0D-12*TEMP**2 ;
<7> A = R1 : 7.5E-12 ;
"""

MODULE = """
MODULE consts
  IMPLICIT NONE
  INTEGER, PARAMETER :: J_ONE = 1 ! first photolysis
  INTEGER, PARAMETER :: J_TWO = 2
  REAL(dp) :: KMT, FC, K0, KI
  REAL(dp), DIMENSION(2) :: J
CONTAINS
  SUBROUTINE define_consts()
    FC = 0.35
    K0 = 3.6E-30*M*(TEMP/300.)**(-4.1)
    KI = 1.9E-12*(TEMP/300.)**(0.2)
    KMT = (K0*KI)*10.**(LOG10(FC)/(1.+(LOG10(K0/KI)/(0.75-1.27*LOG10(FC)))**(2.)))/(K0+KI)
    J(J_ONE) = 6.073E-05*(cos(zenith)**1.743)*exp(-0.474*(1./cos(zenith)))
    J(J_TWO) = 4.775E-04*(cos(zenith)**0.298)*exp(-0.08*(1./cos(zenith)))
  END SUBROUTINE define_consts
END MODULE consts
"""

ENV = "M = CFACTOR*1.0D6; O2 = 0.21*M; N2 = 0.78*M; zenith = ACOS(SUN)"

WRAP = """
#include "f90t_model.h"
void rates(double TEMP, double SUN, double CFACTOR, const double* C, double* out) {
  kpp_Update_RCONST(TEMP, SUN, CFACTOR, 0.0, C, C + KPP_NVAR, out);
}
"""


def f4(x):
    """A default-REAL literal: the float nearest x, as a double."""
    return float(np.float32(x))


def mul4(a, b):
    """Single-precision product of two default-REAL literals."""
    return float(np.float32(np.float32(a) * np.float32(b)))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    if not KPPGEN.exists():
        pytest.skip("kppgen not built")
    d = tmp_path_factory.mktemp("f90")
    (d / "f90t.kpp").write_text(MECH)
    (d / "consts.f90").write_text(MODULE)
    subprocess.run([str(KPPGEN), "--root", "f90t", "-o", str(d), "--rates-f90", "--f90-module",
                    str(d / "consts.f90"), "--env", ENV, str(d / "f90t.kpp")], check=True,
                   capture_output=True)
    (d / "w.c").write_text(WRAP)
    so = d / "libf90t.so"
    subprocess.run(["gcc", "-O", "-ffp-contract=off", "-fPIC", "-shared", "-I", str(d), "-o",
                    str(so), str(d / "w.c"), "-lm"], check=True, capture_output=True)
    L = ctypes.CDLL(str(so))
    L.rates.argtypes = [ctypes.c_double] * 3 + [ctypes.c_void_p, ctypes.c_void_p]
    stats = json.loads((d / "f90t_stats.json").read_text())
    hip = (d / "f90t_hip.h").read_text()
    return L, stats, hip


def expected(T, SUN, CF, conc):
    M = CF * 1.0e6
    O2 = f4(0.21) * M
    N2 = f4(0.78) * M
    zen = math.acos(SUN)
    rsum = conc["R1"] + conc["A"]
    FC = f4(0.35)
    K0 = (f4(3.6e-30) * M) * (T / f4(300.0)) ** f4(-4.1)
    KI = f4(1.9e-12) * (T / f4(300.0)) ** f4(0.2)
    ncf = f4(0.75) - f4(1.27) * math.log10(FC)
    KMT = (K0 * KI) * f4(10.0) ** (math.log10(FC) / (f4(1.0) + (math.log10(K0 / KI) / ncf)
                                                    ** f4(2.0))) / (K0 + KI)
    c = math.cos(zen)
    J1 = (f4(6.073e-5) * (c ** f4(1.743))) * math.exp(-(f4(0.474) * (f4(1.0) / c)))
    J2 = (f4(4.775e-4) * (c ** f4(0.298))) * math.exp(-(f4(0.08) * (f4(1.0) / c)))
    return [mul4(2.40e-12, 0.8) * rsum,
            ((f4(5.6e-34) * N2) * (T / f4(300.0)) ** f4(-2.6)) * O2,
            J1 + J2,
            KMT,
            f4(3.0e7) * math.exp(-(f4(5300.0) / T)),
            1.0e-12 * (T * T),
            7.5e-12]


@pytest.mark.parametrize("T,SUN", [(298.0, 1.0), (260.5, 0.37), (309.9, 0.02)])
def test_f90_rates_bitwise(lib, T, SUN):
    L, stats, _ = lib
    names = stats["species"]
    conc = {"A": 3.0e11, "B": 7.0e9, "R1": 1.25e8}
    C = np.array([conc[n] for n in names])
    out = np.zeros(stats["NREACT"])
    L.rates(T, SUN, 2.5e13, C.ctypes.data, out.ctypes.data)
    exp = expected(T, SUN, 2.5e13, conc)
    for j, (a, b) in enumerate(zip(out, exp)):
        assert a == b, (j + 1, a, b)


def test_single_precision_matters(lib):
    """The literal product differs from its double-precision value, so the test bites."""
    assert mul4(2.40e-12, 0.8) != 2.40e-12 * 0.8
    assert f4(-2.6) != -2.6


def test_rate_classes_and_state_rates(lib):
    _, stats, hip = lib
    assert "RATES_USE_STATE = true" in hip
    assert "N_STATE_RATES = 1;" in hip   # reaction 1 (RSUM)
    assert "N_SUN_RATES = 1;" in hip     # reaction 3 (J through zenith = ACOS(SUN))
    # the state-dependent rate reads the concentrations through the accessors
    body = hip[hip.index("void rates_state"):]
    body = body[:body.index("\n}\n")]
    assert "V[" in body and "RCT[0]" in body


def test_errors_name_the_problem(tmp_path):
    if not KPPGEN.exists():
        pytest.skip("kppgen not built")
    (tmp_path / "bad.kpp").write_text(MECH.replace("2.40E-12*0.8*RSUM", "2.40E-12*UNDEF"))
    (tmp_path / "consts.f90").write_text(MODULE)
    r = subprocess.run([str(KPPGEN), "--root", "bad", "-o", str(tmp_path), "--rates-f90",
                        "--f90-module", str(tmp_path / "consts.f90"), "--env", ENV,
                        str(tmp_path / "bad.kpp")], capture_output=True, text=True)
    assert r.returncode != 0 and "UNDEF" in r.stderr
    r = subprocess.run([str(KPPGEN), "--root", "bad2", "-o", str(tmp_path),
                        str(tmp_path / "bad.kpp")], capture_output=True, text=True)
    assert r.returncode != 0 and "--rates-f90" in r.stderr


def test_mcm_rates_from_the_oracle():
    """The real MCM export through the oracle's C translation, at states where
    each checked rate reduces to a closed form of its rate expression."""
    from oracle import ref
    try:
        L = ref.lib("mcm")
    except OSError:
        pytest.skip("MCM oracle not built")
    s = ref.sizes("mcm")
    assert (s["NVAR"], s["NFIX"], s["NREACT"]) == (610, 0, 1944)
    names = [L.ros_ref_species_name(k).decode() for k in range(s["NSPEC"])]
    T, SUN, CF = 298.0, 1.0, 2.5e13
    M = CF * 1.0e6
    O2, N2 = f4(0.21) * M, f4(0.78) * M
    out = np.zeros(s["NREACT"])
    C = np.zeros(s["NSPEC"])
    L.ros_ref_rconst_c(T, SUN, CF, 0.0, C.ctypes.data, out.ctypes.data)
    # <1> O = O3 : 5.6E-34*N2*(TEMP/300.)**(-2.6)*O2+6.0E-34*O2*(TEMP/300.)**(-2.6)*O2
    p = (T / f4(300.0)) ** f4(-2.6)
    assert out[0] == ((f4(5.6e-34) * N2) * p) * O2 + ((f4(6.0e-34) * O2) * p) * O2
    # <2> O + O3 = PROD : 8.0E-12*EXP(-2060./TEMP)
    assert out[1] == f4(8.0e-12) * math.exp(-(f4(2060.0) / T))
    # <1929> DHPMEK + hv = ... : J(J_CH3OOH); J = l*cos(z)**m*exp(-n/cos(z)), cos(z) = SUN
    c = math.cos(math.acos(SUN))
    J = (f4(7.649e-06) * (c ** f4(0.682))) * math.exp(-(f4(0.279) * (f4(1.0) / c)))
    assert out[1928] == J
    # <1932> CISOPAO2 = CISOPAO : 2.40E-12*0.8*RO2 -- zero without peroxy radicals,
    # then proportional to CH3O2 (one of the RO2 species)
    assert out[1931] == 0.0
    C[names.index("CH3O2")] = 1.0e9
    L.ros_ref_rconst_c(T, SUN, CF, 0.0, C.ctypes.data, out.ctypes.data)
    assert out[1931] == mul4(2.40e-12, 0.8) * 1.0e9
    # a species outside the RO2 list changes nothing
    C[names.index("NO2")] = 1.0e12
    before = out.copy()
    L.ros_ref_rconst_c(T, SUN, CF, 0.0, C.ctypes.data, out.ctypes.data)
    assert np.array_equal(out, before)
