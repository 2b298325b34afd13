"""KPP's GEOS-Chem carbon-gases example (models/carbon.*) through kppgen's F90
rate translator (`2.45d-12*EXP(-1775.0d0/TEMP)`: D exponents, EXP) and the
oracle.  Every reaction is first order in one variable species, so the
system has a closed form: CH4 decays with lam = k1 OH + k2 Cl, the loss
trackers integrate k1 OH CH4 and k2 Cl CH4, CO relaxes to S/mu with mu = k4 OH
and S = k3 DummyCH4 + k5 DummyNMVOC, and the production trackers grow
linearly.  Pinned: the rate constants against the formulas of
models/carbon.eqn, and an RTOL 1e-8 integration against the closed form.
"""
import ctypes

import numpy as np
import pytest

from kpp_amd import workloads
from oracle import ref


@pytest.fixture(scope="module", autouse=True)
def _have_mech():
    try:
        ref.sizes("carbon")
    except (FileNotFoundError, OSError):
        pytest.skip("oracle for carbon not built")


def _idx():
    n = ref.sizes("carbon")["NSPEC"]
    return {ref.species_name("carbon", k): k for k in range(n)}


def _k(T):
    return np.array([2.45e-12 * np.exp(-1775.0 / T), 9.60e-12 * np.exp(-1360.0 / T),
                     4.2566446e-15, 7.3679649e-14, 3.8199012e+04])


def test_rate_constants():
    L = ref.lib("carbon")
    R = np.zeros(5)
    for T in (260.0, 283.7, 310.0):
        L.ros_ref_rconst(T, 0.0, 1.0, 0.0, R.ctypes.data_as(ctypes.c_void_p))
        assert np.allclose(R, _k(T), rtol=2e-16, atol=0.0), T


def test_closed_form():
    b = workloads.carbon_box(ncells=8)
    b.rtol = np.array([1e-9])
    b.atol = np.array([1e-6])
    o = ref.integrate(b)
    assert (o["ierr"] == 1).all()
    i = _idx()
    for c in range(b.ncells):
        C0 = b.C[:, c]
        k1, k2, k3, k4, k5 = _k(b.temp[c])
        t = b.tend[c] - b.tstart[c]
        OH, Cl = C0[i["FixedOH"]], C0[i["FixedCl"]]
        lam = k1 * OH + k2 * Cl
        ch4 = C0[i["CH4"]] * np.exp(-lam * t)
        lost = C0[i["CH4"]] * -np.expm1(-lam * t) / lam
        mu = k4 * OH
        S = k3 * C0[i["DummyCH4"]] + k5 * C0[i["DummyNMVOC"]]
        co = S / mu + (C0[i["CO"]] - S / mu) * np.exp(-mu * t)
        co_lost = (S * t - (co - C0[i["CO"]])) / 1.0  # int of k4 OH CO = S t - (CO(t) - CO0)
        want = {"CH4": ch4, "LCH4byOH": C0[i["LCH4byOH"]] + k1 * OH * lost,
                "LCH4byCl": C0[i["LCH4byCl"]] + k2 * Cl * lost, "CO": co,
                "LCObyOH": C0[i["LCObyOH"]] + co_lost,
                "PCOfromCH4": C0[i["PCOfromCH4"]] + k3 * C0[i["DummyCH4"]] * t,
                "PCOfromNMVOC": C0[i["PCOfromNMVOC"]] + k5 * C0[i["DummyNMVOC"]] * t}
        for name, w in want.items():
            assert o["C"][i[name], c] == pytest.approx(w, rel=1e-6, abs=1e-3), (name, c)
