"""Build recipe of the native parts: libe3d_hip.so (HIP kernels for gfx950 + host C++), the
E3D_PRE-, E3D_Solver- and E3D_POST-compatible executables, and the CPU oracle used by the tests.

Run as ``python -m euler3d_amd.build`` or through ``__graft_entry__.build()``. Outputs go next to
the sources (in-tree) so that they travel with the repository snapshot:

    euler3d_amd/lib/libe3d_hip.so
    euler3d_amd/bin/E3D_Solver
    euler3d_amd/bin/E3D_POST
    euler3d_amd/bin/E3D_PRE
    oracle/lib/libe3d_oracle.so          (test infrastructure)

Host C++ is compiled with g++ -O2 -ffp-contract=off (the arithmetic of the mesh metrics must
round as the reference's g++ build does); device code with hipcc --offload-arch=gfx950
-ffp-contract=off (every FP64 add and multiply rounds separately, as in the reference).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "euler3d_amd"
CSRC = PKG / "csrc"
OBJ = ROOT / "_build" / "obj"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("E3D_OFFLOAD_ARCH", "gfx950")

HOST_SRCS = ["mesh_io.cpp", "config.cpp", "preprocess.cpp", "partitions.cpp", "generators.cpp", "plan.cpp", "post.cpp",
             "capi.cpp", "batch.cpp"]
MAINS = {"E3D_Solver": "solver_main.cpp", "E3D_POST": "post_main.cpp", "E3D_PRE": "pre_main.cpp"}
DEVICE_SRCS = ["kernels.hip", "kernels_mixed.hip"]

CXXFLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
            "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{ROOT}/include"]
HIPFLAGS = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            f"-I{ROOT}/include"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def _hipcc() -> str:
    p = shutil.which("hipcc") or str(ROCM / "bin" / "hipcc")
    return p


def build_native(jobs: int = 8) -> dict[str, Path]:
    OBJ.mkdir(parents=True, exist_ok=True)
    (PKG / "lib").mkdir(exist_ok=True)
    (PKG / "bin").mkdir(exist_ok=True)
    hipcc = _hipcc()
    steps = []
    objs = []
    for s in HOST_SRCS:
        o = OBJ / (s + ".o")
        steps.append(["g++", *CXXFLAGS, "-c", str(CSRC / s), "-o", str(o)])
        objs.append(o)
    for s in DEVICE_SRCS:
        o = OBJ / (s + ".o")
        steps.append([hipcc, *HIPFLAGS, "-c", str(CSRC / s), "-o", str(o)])
        objs.append(o)
    for src in MAINS.values():
        steps.append(["g++", *CXXFLAGS, "-c", str(CSRC / src), "-o", str(OBJ / (src + ".o"))])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, steps))
    lib = PKG / "lib" / "libe3d_hip.so"
    tmp = lib.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
          f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, lib)
    out = {"lib": lib}
    for name, src in MAINS.items():
        exe = PKG / "bin" / name
        _run(["g++", str(OBJ / (src + ".o")), "-o", str(exe), f"-L{PKG}/lib", "-le3d_hip",
              "-Wl,-rpath,$ORIGIN/../lib", f"-Wl,-rpath,{ROCM}/lib"])
        out[name] = exe
    return out


def build_oracle() -> Path:
    """CPU oracle (test infrastructure): plain C restatement of the reference's arithmetic."""
    out = ROOT / "oracle" / "lib"
    out.mkdir(exist_ok=True)
    lib = out / "libe3d_oracle.so"
    tmp = lib.with_suffix(".so.tmp")
    _run(["gcc", "-std=c99", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-Wall",
          str(ROOT / "oracle" / "e3d_oracle.c"), "-o", str(tmp), "-lm"])
    os.replace(tmp, lib)
    return lib


def build_reference() -> bool:
    """Reference solver + golden harness into oracle/_ref (only where the reference tree exists)."""
    if not Path("/root/reference/src").is_dir():
        return False
    _run(["make", "-s", "-f", str(ROOT / "oracle" / "ref.mk"), "-C", str(ROOT), "-j8"])
    return True


def build_all(jobs: int = 8) -> None:
    build_native(jobs)
    build_oracle()
    build_reference()


def build_sanitized(jobs: int = 8) -> Path:
    """The library with its host C++ under AddressSanitizer and UBSan (device code as built by
    build_native, whose objects it links), into _build/asan/ for the CPU tests:

        python -m euler3d_amd.build --sanitize
        E3D_LIB=_build/asan/libe3d_hip.so ASAN_OPTIONS=detect_leaks=0 \\
          LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libstdc++.so.6)" \\
          python -m pytest tests -m "not gpu"

    (libstdc++ is preloaded with the sanitizer runtime so that it can intercept C++ throws in a
    Python process.)"""
    out = ROOT / "_build" / "asan"
    out.mkdir(parents=True, exist_ok=True)
    flags = [f for f in CXXFLAGS if f != "-O2"] + ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined"]
    steps, objs = [], []
    for s in HOST_SRCS:
        o = out / (s + ".o")
        steps.append(["g++", *flags, "-c", str(CSRC / s), "-o", str(o)])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, steps))
    dev = [OBJ / (s + ".o") for s in DEVICE_SRCS]
    if not all(o.exists() for o in dev):
        build_native(jobs)
    lib = out / "libe3d_hip.so"
    _run(["g++", "-shared", "-fsanitize=address,undefined", *map(str, objs), *map(str, dev), "-o", str(lib),
          f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"])
    return lib


if __name__ == "__main__":
    if "--sanitize" in sys.argv[1:]:
        print("built", build_sanitized())
    else:
        build_all()
        print("built", PKG / "lib" / "libe3d_hip.so")
