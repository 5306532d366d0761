"""Build libzeus_amd.so (HIP, gfx950) in-tree.

    python -m zeus_amd.build [--force] [--jobs N]

Every .hip file under zeus_amd/csrc is compiled with hipcc for gfx950 only
into zeus_amd/_build/, then linked into zeus_amd/lib/libzeus_amd.so.
Without --force, objects newer than every source and header are reused.
The example drivers under zeus_amd/drivers/*.hip (zeus_amd/bin/zeus_optimize,
with optimize_<function> links to it) and the C++ front-end test programs
under tests/cpp/*.hip (tests/cpp/bin/) are then compiled against the library
(with an rpath to zeus_amd/lib).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libzeus_amd.so")
DRIVERS = os.path.join(PKG, "drivers")
BINDIR = os.path.join(PKG, "bin")
# names the example-driver program answers to (examples/optimize_*.cu)
DRIVER_FUNCTIONS = ("rosenbrock", "rastrigin", "ackley", "goldstein_price", "himmelblau")
CPP_TESTS = os.path.join(ROOT, "tests", "cpp")
CPP_TEST_BIN = os.path.join(CPP_TESTS, "bin")

ARCH = "gfx950"
# -ffp-contract=off: the fused multiply-adds of the method are written out as
# fma() calls so that the per-thread form matches the CPU oracle bit for bit.
# -packed-fp32-ops: no packed FP32 instructions (none are wanted here).
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
    "-Wno-cuda-compat", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _headers():
    hs = glob.glob(os.path.join(CSRC, "**", "*.hpp"), recursive=True)
    hs += glob.glob(os.path.join(ROOT, "include", "*.h"))
    return hs


def _compile(src: str, obj: str) -> tuple[str, int, str]:
    cmd = [hipcc(), *HIPFLAGS, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l)
    return src, p.returncode, err


def build(force: bool = False, jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    newest_hdr = max((os.path.getmtime(h) for h in _headers()), default=0.0)
    todo, objs = [], []
    for s in srcs:
        o = os.path.join(BUILD, os.path.basename(s)[:-4] + ".o")
        objs.append(o)
        stale = (force or not os.path.exists(o) or os.path.getmtime(o) < os.path.getmtime(s)
                 or os.path.getmtime(o) < newest_hdr)
        if stale:
            todo.append((s, o))
    jobs = jobs or min(8, os.cpu_count() or 1, max(1, len(todo)))
    failed = []
    if todo:
        with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
            for src, rc, err in ex.map(lambda t: _compile(*t), todo):
                name = os.path.relpath(src, ROOT)
                if rc != 0:
                    failed.append(name)
                    print(f"[zeus_amd.build] FAILED {name}\n{err}", file=sys.stderr)
                elif verbose:
                    print(f"[zeus_amd.build] compiled {name}", flush=True)
    if failed:
        raise RuntimeError(f"compilation failed: {failed}")
    if todo or not os.path.exists(LIB):
        tmp = LIB + ".tmp"
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError(f"link failed:\n{p.stderr}")
        os.replace(tmp, LIB)
        if verbose:
            print(f"[zeus_amd.build] linked {os.path.relpath(LIB, ROOT)}", flush=True)
    build_programs(DRIVERS, BINDIR, force=force or bool(todo), verbose=verbose)
    for fn in DRIVER_FUNCTIONS:
        link = os.path.join(BINDIR, "optimize_" + fn)
        if not os.path.islink(link):
            if os.path.exists(link):
                os.remove(link)
            os.symlink("zeus_optimize", link)
    build_cpp_tests(force=force or bool(todo), verbose=verbose)
    return LIB


def build_programs(src_dir: str, bin_dir: str, force: bool = False,
                   verbose: bool = True) -> list[str]:
    """Compile each src_dir/*.hip into bin_dir/<name> linked to the library
    (the example drivers: optimize.hip -> zeus_optimize)."""
    srcs = sorted(glob.glob(os.path.join(src_dir, "*.hip")))
    if not srcs:
        return []
    os.makedirs(bin_dir, exist_ok=True)
    hdrs = _headers() + glob.glob(os.path.join(src_dir, "*.hpp"))
    newest = max([os.path.getmtime(h) for h in hdrs] + [os.path.getmtime(LIB)])
    rpath = "-Wl,-rpath,$ORIGIN/" + os.path.relpath(LIBDIR, bin_dir)
    out = []
    for s in srcs:
        name = os.path.basename(s)[:-4]
        exe = os.path.join(bin_dir, "zeus_" + name if src_dir == DRIVERS else name)
        out.append(exe)
        if (not force and os.path.exists(exe) and os.path.getmtime(exe) >= os.path.getmtime(s)
                and os.path.getmtime(exe) >= newest):
            continue
        cmd = [hipcc(), *HIPFLAGS, f"-I{src_dir}", s, "-o", exe, f"-L{LIBDIR}", "-lzeus_amd",
               rpath]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l)
            raise RuntimeError(f"compilation of {os.path.relpath(s, ROOT)} failed:\n{err}")
        if verbose:
            print(f"[zeus_amd.build] compiled {os.path.relpath(exe, ROOT)}", flush=True)
    return out


def build_cpp_tests(force: bool = False, verbose: bool = True) -> list[str]:
    """Compile each tests/cpp/*.hip into an executable linked to the library."""
    srcs = sorted(glob.glob(os.path.join(CPP_TESTS, "*.hip")))
    if not srcs:
        return []
    os.makedirs(CPP_TEST_BIN, exist_ok=True)
    newest = max([os.path.getmtime(h) for h in _headers()] + [os.path.getmtime(LIB)])
    rpath = "-Wl,-rpath,$ORIGIN/" + os.path.relpath(LIBDIR, CPP_TEST_BIN)
    out = []
    for s in srcs:
        exe = os.path.join(CPP_TEST_BIN, os.path.basename(s)[:-4])
        out.append(exe)
        if (not force and os.path.exists(exe) and os.path.getmtime(exe) >= os.path.getmtime(s)
                and os.path.getmtime(exe) >= newest):
            continue
        cmd = [hipcc(), *HIPFLAGS, s, "-o", exe, f"-L{LIBDIR}", "-lzeus_amd", rpath]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l)
            raise RuntimeError(f"compilation of {os.path.relpath(s, ROOT)} failed:\n{err}")
        if verbose:
            print(f"[zeus_amd.build] compiled {os.path.relpath(exe, ROOT)}", flush=True)
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(force=a.force, jobs=a.jobs)


if __name__ == "__main__":
    main()
