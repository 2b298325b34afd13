"""Synthetic cell batches for the solver's workloads.

Every random number is a counter-based draw keyed by (seed, field, cell), so
any subset of cells of a batch can be regenerated exactly (the CPU oracle
and the GPU see identical inputs).  Workloads:

  small_strato_tod  small_strato, Ros3, cells at uniform time of day; SUN
                    from Update_SUN(TIME); TEMP 270 K (models/small_strato.def
                    C_INIT); 1 h; RTOL 1e-3, ATOL 1 (scalar)
  saprc99_day       saprc99, Rodas3, T ~ U(260,310) K, p ~ U(500,1013) hPa,
                    SUN = 1 held fixed; 1 h; RTOL 1e-3, ATOL 1
  saprc99_uneven    as saprc99_day but 50 % night (SUN = 0) and 50 % day
                    with SUN ~ U(0.2, 1)

Initial concentrations are the mechanism's #INITVALUES (x CFACTOR) with
each variable species scaled by 10**U(-1, 1); fixed species keep their
#INITVALUES value x CFACTOR, so AIR equals M for saprc99.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import solver

K_BOLTZMANN = 1.380649e-23  # J/K

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, fld: int, cells: np.ndarray) -> np.ndarray:
    """U[0,1) draws for (seed, field, cell), cell-wise independent."""
    cells = np.asarray(cells, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(fld))
        u = _splitmix64(cells ^ key)
    return (u >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# field numbers
F_TEMP, F_PRES, F_SUN, F_DAY, F_TOD = 1, 2, 3, 4, 5
F_SPC = 100  # + species index


@dataclass
class Batch:
    mech: str
    method: int
    C: np.ndarray          # [NSPEC, n]
    temp: np.ndarray
    cfactor: np.ndarray
    sun: np.ndarray
    sun_mode: int
    tstart: np.ndarray
    tend: np.ndarray
    atol: np.ndarray
    rtol: np.ndarray
    IPAR: np.ndarray
    RPAR: np.ndarray
    cells: np.ndarray
    info: dict = field(default_factory=dict)

    @property
    def ncells(self):
        return self.C.shape[1]

    def to_device(self, device="cuda"):
        import torch
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        return {
            "C": t(self.C), "temp": t(self.temp), "cfactor": t(self.cfactor),
            "sun": t(self.sun), "tstart": t(self.tstart), "tend": t(self.tend),
        }


def mech_init(mech: str):
    """#INITVALUES values (in the mechanism's concentration units, before
    CFACTOR) and the #INITVALUES CFACTOR, from the compiler's stats file."""
    st = json.loads((solver.GEN / f"{mech}_stats.json").read_text())
    vals = np.array([float(eval(v, {"__builtins__": {}}, {})) for v in st["init"]], np.float64)
    return vals, float(st["cfactor"]), st


def _scale_var(vals, nvar, seed, cells):
    C = np.repeat(vals[:, None], len(cells), axis=1)
    for s in range(nvar):
        C[s] *= 10.0 ** (2.0 * uniform(seed, F_SPC + s, cells) - 1.0)
    return C


def small_strato_tod(ncells=4096, seed=1, cells=None, rtol=1e-3, atol=1.0, method=solver.ROS3,
                     mech="small_strato"):
    cells = np.arange(ncells, dtype=np.int64) if cells is None else np.asarray(cells)
    vals, cf, st = mech_init(mech)
    nvar = st["NVAR"]
    C = _scale_var(vals * cf, nvar, seed, cells)
    tod = 24.0 * uniform(seed, F_TOD, cells)
    tstart = tod * 3600.0
    n = len(cells)
    return Batch(mech, method, C, np.full(n, 270.0), np.full(n, cf), np.zeros(n),
                 solver.SUN_TOD, tstart, tstart + 3600.0, np.array([atol]), np.array([rtol]),
                 solver.default_ipar(method), np.zeros(20), cells,
                 {"workload": f"{mech}_tod"})


def strato_fam_tod(ncells=4096, seed=1, cells=None, rtol=1e-3, atol=1.0, method=solver.ROS3):
    """small_strato_tod's recipe on the #FAMILIES test mechanism
    (kpp_amd/mechs/strato_fam.kpp); the family species POx, LOx start at 0."""
    return small_strato_tod(ncells, seed, cells, rtol, atol, method, mech="strato_fam")


def _saprc99(ncells, seed, cells, uneven, rtol, atol, method, mech="saprc99"):
    cells = np.arange(ncells, dtype=np.int64) if cells is None else np.asarray(cells)
    vals, _, st = mech_init(mech)
    nvar = st["NVAR"]
    T = 260.0 + 50.0 * uniform(seed, F_TEMP, cells)
    p_hpa = 500.0 + 513.0 * uniform(seed, F_PRES, cells)
    M = p_hpa * 100.0 / (K_BOLTZMANN * T) * 1e-6  # molecules / cm3
    cf = M / 1e6                                   # ppm -> molecules / cm3
    C = _scale_var(vals, nvar, seed, cells) * cf[None, :]
    if uneven:
        day = uniform(seed, F_DAY, cells) < 0.5
        sun = np.where(day, 0.2 + 0.8 * uniform(seed, F_SUN, cells), 0.0)
    else:
        sun = np.ones(len(cells))
    tstart = np.full(len(cells), 12.0 * 3600.0)
    return Batch(mech, method, C, T, cf, sun, solver.SUN_FIXED, tstart, tstart + 3600.0,
                 np.array([atol]), np.array([rtol]), solver.default_ipar(method), np.zeros(20),
                 cells, {"workload": f"{mech}_uneven" if uneven else f"{mech}_day"})


def saprc99_day(ncells=65536, seed=2, cells=None, rtol=1e-3, atol=1.0, method=solver.RODAS3):
    return _saprc99(ncells, seed, cells, False, rtol, atol, method)


def saprc99_uneven(ncells=1048576, seed=3, cells=None, rtol=1e-3, atol=1.0,
                   method=solver.RODAS3):
    return _saprc99(ncells, seed, cells, True, rtol, atol, method)


def saprcnov_uneven(ncells=65536, seed=9, cells=None, rtol=1e-3, atol=1.0,
                    method=solver.RODAS3):
    """models/saprcnov (KPP's other SAPRC-99 variant: 88 variable species, N2 fixed):
    the saprc99_uneven recipe -- .def ppm values x 10^U(-1,1) x CFACTOR per cell,
    T ~ U(260, 310) K, p ~ U(500, 1013) hPa, half the cells at night."""
    return _saprc99(ncells, seed, cells, True, rtol, atol, method, mech="saprcnov")


# MCM isoprene example (examples/mcm/driver_mcm.f90): air density M = 2.5E19
# molecules/cm3, initial mixing ratios of O3, NO2, CH4 and C5H8 times M, all
# other species from Initialize (zero).  The mechanism's rate code takes M as
# CFACTOR*1e6 and O2, N2, H2O as the driver's fractions of M.
MCM_M = 2.5e19
MCM_INIT = {"O3": 3.0e-8, "NO2": 1.0e-10, "CH4": 1.8e-6, "C5H8": 1.0e-9}


def mcm_mixed(ncells=16384, seed=5, cells=None, rtol=1e-3, atol=1.0, method=solver.RODAS3):
    """MCM isoprene subset: T ~ U(260, 310) K, SUN ~ U(0, 1) taken as cos(zenith),
    initial concentrations = the driver's values x 10^U(-1, 1), 1 h, Rodas3."""
    cells = np.arange(ncells, dtype=np.int64) if cells is None else np.asarray(cells)
    vals, cf0, st = mech_init("mcm")
    nvar = st["NVAR"]
    base = vals * cf0
    for name, mix in MCM_INIT.items():
        base[st["species"].index(name)] = mix * MCM_M
    C = _scale_var(base, nvar, seed, cells)
    n = len(cells)
    T = 260.0 + 50.0 * uniform(seed, F_TEMP, cells)
    sun = uniform(seed, F_SUN, cells)
    tstart = np.zeros(n)
    return Batch("mcm", method, C, T, np.full(n, MCM_M / 1e6), sun, solver.SUN_FIXED, tstart,
                 tstart + 3600.0, np.array([atol]), np.array([rtol]), solver.default_ipar(method),
                 np.zeros(20), cells, {"workload": "mcm_mixed"})


def linear6(ncells=1, method=solver.RODAS3, rtol=1e-3, atol=1.0, temp=1.0e-3):
    """The linear6 test mechanism (kpp_amd/mechs/linear6.kpp): every species
    at its #INITVALUES value (1), P = 1, every rate constant TEMP (growth and
    decay cancel: dX/dt = 0), SUN fixed at 0, [0, 3600 s]; callers set C,
    TEMP, the times or the rate constants per test."""
    cells = np.arange(ncells, dtype=np.int64)
    vals, cf, _ = mech_init("linear6")
    C = np.repeat((vals * cf)[:, None], ncells, axis=1)
    z = np.zeros(ncells)
    return Batch("linear6", method, C, np.full(ncells, temp), np.full(ncells, cf), z.copy(),
                 solver.SUN_FIXED, z.copy(), z + 3600.0, np.array([atol]), np.array([rtol]),
                 solver.default_ipar(method), np.zeros(20), cells, {"workload": "linear6"})


def carbon_box(ncells=4096, seed=11, cells=None, rtol=1e-3, atol=1.0, method=solver.RODAS3):
    """KPP's carbon example (models/carbon.*): its #INITVALUES (CH4, CO, the
    fixed OH and the dummy precursors) x 10^U(-1,1) per species and cell,
    T ~ U(260, 310) K, three days from t = 0 as the example's driver runs."""
    cells = np.arange(ncells, dtype=np.int64) if cells is None else np.asarray(cells)
    vals, cf, st = mech_init("carbon")
    nvar, nspec = st["NVAR"], len(vals)
    C = np.repeat((vals * cf)[:, None], len(cells), axis=1)
    for s in range(nspec):
        C[s] *= 10.0 ** (2.0 * uniform(seed, F_SPC + s, cells) - 1.0)
    n = len(cells)
    T = 260.0 + 50.0 * uniform(seed, F_TEMP, cells)
    tstart = np.zeros(n)
    return Batch("carbon", method, C, T, np.full(n, cf), np.zeros(n), solver.SUN_FIXED, tstart,
                 tstart + 3.0 * 24.0 * 3600.0, np.array([atol]), np.array([rtol]),
                 solver.default_ipar(method), np.zeros(20), cells, {"workload": "carbon_box"})


WORKLOADS = {
    "small_strato_tod": small_strato_tod,
    "saprc99_day": saprc99_day,
    "saprc99_uneven": saprc99_uneven,
    "saprcnov_uneven": saprcnov_uneven,
    "mcm_mixed": mcm_mixed,
}
