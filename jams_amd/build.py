"""Build the native parts of jams_amd in-tree.

- ``lib/libjamship.so``: the gfx950 HIP engine behind ``include/jams_hip.h``
  (hipcc, ``--offload-arch=gfx950``);
- ``lib/libjamshost.so`` and ``lib/jams``: the C++ host layer (config parser,
  lattice, neighbour list, CSR builder, Solver/Hamiltonian/Thermostat plugins)
  and the command-line driver, linked against the engine.

Every call recompiles everything (the driver's "does it build" check); the
translation units are compiled in parallel.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
LIB = PKG / "lib"
OBJ = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# -ffp-contract=off: the FP64 path reproduces the CPU solver's rounding bit
# for bit.  Packed-FP32 instructions are disabled on this platform.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", str(ROOT / "include"),
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wextra",
             "-Wno-unused-parameter", "-I", str(ROOT / "include"), "-I", str(HOST)]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def _compile(src: Path, obj: Path, hip: bool) -> None:
    if hip:
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    else:
        _run([CXX, *CXX_FLAGS, "-c", str(src), "-o", str(obj)])


def build(verbose: bool = False) -> dict:
    if OBJ.exists():
        shutil.rmtree(OBJ)
    OBJ.mkdir(parents=True)
    LIB.mkdir(exist_ok=True)
    hip_srcs = sorted(CSRC.glob("*.hip"))
    host_srcs = sorted(p for p in HOST.glob("*.cc") if p.name != "jams_main.cc")
    jobs = [(s, OBJ / (s.stem + ".hip.o"), True) for s in hip_srcs]
    jobs += [(s, OBJ / (s.stem + ".o"), False) for s in host_srcs]
    main_src = HOST / "jams_main.cc"
    if main_src.exists():
        jobs.append((main_src, OBJ / "jams_main.o", False))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(lambda j: _compile(*j), jobs))
    engine = LIB / "libjamship.so"
    tmp = engine.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", str(tmp),
          *[str(OBJ / (s.stem + ".hip.o")) for s in hip_srcs], "-lpthread"])
    tmp.replace(engine)
    out = {"engine": str(engine)}
    if host_srcs:
        host = LIB / "libjamshost.so"
        tmp = host.with_suffix(".so.tmp")
        _run([CXX, "-shared", "-fopenmp", "-o", str(tmp),
              *[str(OBJ / (s.stem + ".o")) for s in host_srcs],
              "-L", str(LIB), "-ljamship", "-ldl", "-Wl,-rpath,$ORIGIN"])
        tmp.replace(host)
        out["host"] = str(host)
        if main_src.exists():
            exe = LIB / "jams"
            tmp = LIB / "jams.tmp"
            _run([CXX, "-fopenmp", "-o", str(tmp), str(OBJ / "jams_main.o"),
                  "-L", str(LIB), "-ljamshost", "-ljamship", "-Wl,-rpath,$ORIGIN"])
            tmp.replace(exe)
            out["driver"] = str(exe)
    shutil.rmtree(OBJ)
    return out


if __name__ == "__main__":
    print(build())
