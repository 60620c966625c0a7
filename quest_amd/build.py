"""Builds libquest_amd.so in-tree for gfx950 with hipcc.

    python -m quest_amd.build [--jobs N]

Every HIP source is compiled with packed FP32 instructions disabled
(-target-feature -packed-fp32-ops): the kernels use scalar FP32 FMAs.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libquest_amd.so")
ROOT = os.path.dirname(HERE)

SOURCES = ["quest_kernels.hip", "quest_capi.cpp", "pc_weights.cpp"]
ARCH = os.environ.get("QUEST_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build quest_amd)")


def common_flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]


def compile_one(src: str, obj: str, extra=()) -> None:
    cmd = [hipcc()] + common_flags() + list(extra) + ["-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile of {src} failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    errs = [ln for ln in res.stderr.splitlines()
            if "warning" in ln and "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
    if errs:
        sys.stderr.write("\n".join(errs[:40]) + "\n")


def build(jobs: int = 4, out: str = LIB, defines=()) -> str:
    libdir = os.path.dirname(out)
    os.makedirs(libdir, exist_ok=True)
    objdir = os.path.join(libdir, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    extra = [f"-D{d}" for d in defines]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda so: compile_one(*so, extra), zip(SOURCES, objs)))
    tmp = out + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


BINDIR = os.path.join(HERE, "bin")
PROGRAMS = {  # output name -> source (host C++ over the C ABI, linked to the library)
    "quest_hip": os.path.join(HERE, "host", "quest_main.cpp"),
    "mirror_test": os.path.join(ROOT, "tests", "cpp", "mirror_test.cpp"),
}


def build_programs(lib: str = LIB) -> list:
    """The C++ driver (`quest_hip`, main.cpp's replacement) and the mirror's test."""
    os.makedirs(BINDIR, exist_ok=True)
    out = []
    libdir = os.path.dirname(lib)
    for name, src in PROGRAMS.items():
        exe = os.path.join(BINDIR, name)
        cmd = [hipcc(), "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src,
               "-o", exe + ".tmp", f"-L{libdir}", "-lquest_amd", "-Wl,-rpath,$ORIGIN/../lib"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"build of {name} failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--out", default=LIB, help="library path (default: in-tree quest_amd/lib)")
    ap.add_argument("-D", dest="defines", action="append", default=[], help="extra macro definition")
    args = ap.parse_args(argv)
    print(build(args.jobs, args.out, args.defines))
    if args.out == LIB:
        print("\n".join(build_programs()))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
