"""Build the solver library and its generated mechanism code.

Steps (all outputs stay in-tree and are git-ignored build products):
  1. compile the mechanism compiler ``kppgen`` (g++, C++17);
  2. run it on each mechanism's KPP files -> ``kpp_amd/_gen/<root>_{hip,model}.h``
     and ``<root>_stats.json``;
  3. compile one HIP translation unit per mechanism plus the C ABI for gfx950
     and link ``kpp_amd/lib/libkpphip.so``.
``build_oracle()`` compiles the CPU oracle (test infrastructure) from the
generated C model headers into ``oracle/build/``.

The KPP mechanism sources are read from ``$KPP_HOME`` (default
``/root/reference``), the KPP distribution whose ``models/`` directory holds
``small_strato.*`` and ``saprc99.*``.
"""
from __future__ import annotations

import concurrent.futures as cf
import json
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
GEN = PKG / "_gen"
OBJ = PKG / "_obj"
LIB = PKG / "lib"
ORACLE = REPO / "oracle"
ORACLE_BUILD = ORACLE / "build"

# mechanism root -> KPP input file relative to KPP_HOME, and kppgen options
# (paths in them are relative to KPP_HOME too)
MECHANISMS = {
    "small_strato": ("models/small_strato.def", []),
    "saprc99": ("models/saprc99.def", ["--lds-split", "24"]),
    # KPP's second SAPRC-99 model (models/saprcnov.*): same rate-law syntax, 88
    # variable species, N2 among the fixed species
    "saprcnov": ("models/saprcnov.def", ["--lds-split", "24"]),
    # MCM isoprene export: rate code in Fortran 90 (examples/mcm/constants_mcm.f90,
    # inlined through #INLINE F90_RCONST), with the air density and its O2, N2,
    # H2O fractions and the solar zenith angle from examples/mcm/driver_mcm.f90
    # (cos(zenith) is the cell's SUN)
    # this project's test mechanism (kpp_amd/mechs/linear6.kpp; an absolute
    # path, so not under KPP_HOME): six independent linear species whose
    # steps and singular iteration matrices have closed forms
    "linear6": (str(PKG / "mechs" / "linear6.kpp"), []),
    # this project's #FAMILIES test mechanism (kpp_amd/mechs/strato_fam.kpp):
    # small_strato with production/loss families of odd oxygen
    "strato_fam": (str(PKG / "mechs" / "strato_fam.kpp"), []),
    # KPP's GEOS-Chem carbon-gases example (models/carbon.*): F90 rate
    # expressions (EXP, D exponents), fixed OH/Cl and diagnostic species
    "carbon": ("models/carbon.def", ["--rates-f90"]),
    "mcm": ("examples/mcm/mcm.kpp",
            ["--no-pipelined", "--part-ops", "3000", "--rates-f90", "--f90-module", "examples/mcm/constants_mcm.f90",
             "--env", "M = CFACTOR*1.0D6; O2 = 0.21*M; N2 = 0.78*M; H2O = 1.0E-2*M; "
                      "zenith = ACOS(SUN)"]),
}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # parity build: no FMA contraction, so device arithmetic is the oracle's
    "-ffp-contract=off",
    # packed FP32 instructions stay off in every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-function", "-Wno-unused-variable",
]

# Register budget of each mechanism's integrator kernels, as the number of
# wavefronts per SIMD the allocator must leave room for (512/N VGPRs per
# lane, shared by the out-of-line model routines).  Measured on MI355X
# (persistent kernel, index order): saprc99 on 1,048,576 uneven cells takes
# 1.85 s at 4 waves/SIMD (128 VGPRs, 2.6 KB of spills per lane), 1.71 s at 2
# (256 VGPRs) and 2.07 s at 1; MCM on 16,384 cells 2.35 s at 4 (its 16,384
# cells fill one wavefront per SIMD at most), 2.03 s at 2 and 1.88 s at 1
# (VGPRs plus AGPRs for spills).  small_strato: 2, so that its register
# kernel (the whole working set in VGPRs, 71-86 doubles per lane) fits.
WAVES_PER_EU = {"small_strato": 2, "saprc99": 2, "saprcnov": 2, "linear6": 2, "strato_fam": 2,
                "carbon": 2, "mcm": 1}


def kpp_home() -> Path:
    return Path(os.environ.get("KPP_HOME", "/root/reference"))


def run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(map(str, cmd))}\n"
                           f"{r.stdout[-4000:]}\n{r.stderr[-8000:]}")
    return r


def build_kppgen() -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    exe = OBJ / "kppgen"
    run(["g++", "-std=c++17", "-O2", "-Wall", "-Wno-unused-function", "-o", str(exe),
         str(CSRC / "kppgen" / "kppgen.cpp")])
    return exe


def generate(mechs=None) -> list[str]:
    """Run kppgen on the mechanisms whose sources are available."""
    home = kpp_home()
    exe = build_kppgen()
    GEN.mkdir(parents=True, exist_ok=True)
    done = []
    for root, (rel, opts) in (mechs or MECHANISMS).items():
        src = home / rel
        if not src.exists():
            raise FileNotFoundError(f"KPP mechanism source {src} not found (set KPP_HOME)")
        args = []
        for k, a in enumerate(opts):
            args.append(str(home / a) if k > 0 and opts[k - 1] == "--f90-module" else a)
        run([str(exe), "--kpp-home", str(home), "--root", root, "-o", str(GEN), "--coop", *args,
             str(src)])
        (GEN / f"kpphip_{root}.hip").write_text(
            f'#include "{root}_hip.h"\n#include "mech_launch.h"\nKPPHIP_DEFINE_MECH({root})\n')
        # the mixed-precision integrator and the batched model functions in
        # translation units of their own (they compile in parallel)
        (GEN / f"kpphip_{root}_mixed.hip").write_text(
            f'#include "{root}_hip.h"\n#include "mech_launch.h"\n'
            f'KPPHIP_DEFINE_MECH_MIXED({root})\n')
        (GEN / f"kpphip_{root}_comp.hip").write_text(
            f'#include "{root}_hip.h"\n#include "mech_launch.h"\n'
            f'KPPHIP_DEFINE_MECH_COMPONENTS({root})\n')
        (GEN / f"kpphip_{root}_sdirk.hip").write_text(
            f'#include "{root}_hip.h"\n#include "implicit_launch.h"\n'
            f'KPPHIP_DEFINE_MECH_SDIRK({root})\n')
        (GEN / f"kpphip_{root}_rk.hip").write_text(
            f'#include "{root}_hip.h"\n#include "implicit_launch.h"\n'
            f'KPPHIP_DEFINE_MECH_RK({root})\n')
        (GEN / f"kpphip_{root}_adj.hip").write_text(
            f'#include "{root}_hip.h"\n#include "adj_launch.h"\n'
            f'KPPHIP_DEFINE_MECH_ADJ({root})\n')
        done.append(root)
    (GEN / "kpphip_registry.inc").write_text("".join(f"KPPHIP_MECH({r})\n" for r in done))
    return done


def _mech_of(src: Path) -> str:
    stem = src.stem.replace("kpphip_", "", 1)
    for suffix in ("_mixed", "_comp", "_sdirk", "_rk", "_adj"):
        if stem.endswith(suffix):
            return stem[:-len(suffix)]
    return stem


def _includes(src: Path) -> list[Path]:
    """The project headers a source includes, directly or not (quoted
    #include lines resolved in its own folder, csrc, _gen and include)."""
    import re
    seen, todo = [], [src]
    dirs = [CSRC, GEN, REPO / "include"]
    while todo:
        f = todo.pop()
        for name in re.findall(r'^\s*#\s*include\s+"([^"]+)"', f.read_text(), re.M):
            for d in (f.parent, *dirs):
                p = d / name
                if p.exists():
                    if p not in seen:
                        seen.append(p)
                        todo.append(p)
                    break
    return sorted(seen)


def _stamp(src: Path, cmd: list) -> str:
    """Digest of a translation unit's inputs: the command, the source and
    every project header it includes."""
    import hashlib
    h = hashlib.sha256(" ".join(cmd).encode())
    for d in [src, *_includes(src)]:
        h.update(str(d).encode())
        h.update(d.read_bytes())
    return h.hexdigest()


def _hipcc_obj(src: Path, out: Path):
    waves = WAVES_PER_EU.get(_mech_of(src), 4)
    cmd = [HIPCC, *HIP_FLAGS, f"-DKPPHIP_WAVES_PER_EU={waves}", "-I", str(CSRC), "-I", str(GEN),
           "-I", str(REPO / "include"), "-c", "-x", "hip", str(src), "-o", str(out)]
    # KPPHIP_INCREMENTAL=1 (development): skip a translation unit whose
    # inputs are unchanged since its object was built
    stamp_file = out.with_suffix(".stamp")
    stamp = _stamp(src, cmd) if os.environ.get("KPPHIP_INCREMENTAL") == "1" else None
    if stamp and out.exists() and stamp_file.exists() and stamp_file.read_text() == stamp:
        return out
    run(cmd)
    if stamp:
        stamp_file.write_text(stamp)
    return out


def build_lib(mechs: list[str], jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(parents=True, exist_ok=True)
    srcs = ([GEN / f"kpphip_{m}.hip" for m in mechs] +
            [GEN / f"kpphip_{m}_mixed.hip" for m in mechs] +
            [GEN / f"kpphip_{m}_comp.hip" for m in mechs] +
            [GEN / f"kpphip_{m}_sdirk.hip" for m in mechs] +
            [GEN / f"kpphip_{m}_rk.hip" for m in mechs] +
            [GEN / f"kpphip_{m}_adj.hip" for m in mechs] +
            [CSRC / "kpphip_api.cpp", CSRC / "mlp.hip", CSRC / "sched.cpp"])
    # largest mechanism first, so the slowest translation units start at once
    def weight(src):
        h = GEN / (_mech_of(src) + "_hip.h")
        return -(h.stat().st_size if h.exists() else 0)
    srcs.sort(key=weight)
    objs = [OBJ / (s.stem + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_hipcc_obj, srcs, objs))
    out = LIB / "libkpphip.so"
    tmp = LIB / "libkpphip.so.tmp"
    run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_driver(lib: Path) -> Path:
    """The C box-model driver (kpp_amd/drv/box_driver.c) linked to the library."""
    exe = LIB / "box_driver"
    run(["gcc", "-O2", "-Wall", "-I", str(REPO / "include"), "-o", str(exe),
         str(PKG / "drv" / "box_driver.c"), "-L", str(LIB), "-lkpphip",
         "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-lm"])
    return exe


def build_oracle(mechs: list[str]) -> list[Path]:
    """CPU oracle per mechanism: parity build (-O -ffp-contract=off, the flags
    of util/Makefile_c:3 plus no contraction) and a timing build."""
    ORACLE_BUILD.mkdir(parents=True, exist_ok=True)
    outs = []
    for m in mechs:
        hdr = f'-DKPP_MODEL_HEADER="{m}_model.h"'
        common = ["-fPIC", "-shared", "-fopenmp", "-Wall", "-Wno-unused-function", hdr,
                  "-I", str(GEN), str(ORACLE / "ros_ref.c"), "-lm"]
        parity = ORACLE_BUILD / f"libros_ref_{m}.so"
        run(["gcc", "-O", "-ffp-contract=off", *common[:-2], "-o", str(parity), *common[-2:]])
        fast = ORACLE_BUILD / f"libros_ref_{m}_fast.so"
        run(["gcc", "-O2", "-march=x86-64-v3", *common[:-2], "-o", str(fast), *common[-2:]])
        outs += [parity, fast]
    return outs


def export_weights(mechs: list[str]) -> list[Path]:
    """The shipped scheduler weights (kpp_amd/weights/<mech>_step_mlp.npz) as
    the flat images C hosts load (kpp_sched_load), next to the library."""
    from . import learned
    outs = []
    for m in mechs:
        src = PKG / "weights" / f"{m}_step_mlp.npz"
        if src.exists():
            out = LIB / f"{m}_step_mlp.kppw"
            learned.export_flat(src, out)
            outs.append(out)
    return outs


def build_all(with_oracle: bool = True) -> dict:
    mechs = generate()
    lib = build_lib(mechs)
    info = {"lib": str(lib), "mechanisms": mechs, "driver": str(build_driver(lib)),
            "weights": [str(p) for p in export_weights(mechs)]}
    if with_oracle:
        info["oracle"] = [str(p) for p in build_oracle(mechs)]
    return info


if __name__ == "__main__":
    print(json.dumps(build_all("--no-oracle" not in sys.argv), indent=1))
