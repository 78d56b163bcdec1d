"""Builds libfiphip.so (gfx950 HIP kernels + C ABI) and the fip CLI in-tree.

Every HIP translation unit is compiled with packed-FP32 instructions disabled. The exact
stencil and the small kernels use -ffp-contract=off (bit-exact with the IEEE host
oracle); the fast stencil uses -ffp-contract=fast.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfiphip.so")
CLI = os.path.join(PKG, "bin", "fip")
ARCH = os.environ.get("FIP_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
NO_PK = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-command-line-argument"] + NO_PK

EMBED = ["fip_ctl.h", "fip_detmath.h", "fip_sched.h", "fip_step_common.h", "fip_bc_common.h"]

UNITS = [
    ("stencil_exact.hip", ["-ffp-contract=off"]),
    ("stencil_fast.hip", ["-ffp-contract=fast", "-ffinite-math-only"]),
    ("kernels_misc.hip", ["-ffp-contract=off"]),
    ("ensemble.hip", ["-ffp-contract=off"]),
    ("closure.hip", ["-ffp-contract=off"]),
    ("fip_lib.cpp", ["-ffp-contract=off", "-x", "hip"]),
    ("fip_ens.cpp", ["-ffp-contract=off", "-x", "hip"]),
    ("closure_host.cpp", ["-ffp-contract=off", "-x", "hip"]),
]


def gen_embed() -> str:
    out = os.path.join(CSRC, "fip_embed.h")
    parts = ["// Generated by fip_amd/build.py: device headers handed to hipRTC.",
             "#pragma once", f"static const int kEmbedCount = {len(EMBED)};",
             "static const char* const kEmbedNames[] = {" + ", ".join(f'"{n}"' for n in EMBED) + "};",
             "static const char* const kEmbedSrc[] = {"]
    for n in EMBED:
        with open(os.path.join(CSRC, n)) as f:
            txt = f.read()
        assert ")FIPX\"" not in txt
        parts.append('R"FIPX(' + txt + ')FIPX",')
    parts.append("};")
    text = "\n".join(parts) + "\n"
    old = open(out).read() if os.path.exists(out) else None
    if old != text:
        with open(out, "w") as f:
            f.write(text)
    return out


def _run_all(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
             for c in cmds]
    errs = []
    for c, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            errs.append(" ".join(c) + "\n" + out)
    if errs:
        raise RuntimeError("build failed:\n" + "\n".join(errs))


def build_lib(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    gen_embed()
    cmds, objs = [], []
    for src, flags in UNITS:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC] + COMMON + flags + ["-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
    _run_all(cmds)
    tmp = LIB + ".tmp"
    _run_all([[HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs +
              ["-L/opt/rocm/lib", "-lhiprtc", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_cli(verbose: bool = False) -> str | None:
    src = os.path.join(PKG, "cli", "fip_main.cpp")
    if not os.path.exists(src):
        return None
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    srcs = [os.path.join(PKG, "cli", f) for f in sorted(os.listdir(os.path.join(PKG, "cli")))
            if f.endswith(".cpp")]
    tmp = CLI + ".tmp"
    _run_all([["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", tmp]
              + srcs + ["-L", PKG, "-lfiphip", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]])
    os.replace(tmp, CLI)
    if verbose:
        print("built", CLI)
    return CLI


PROBE = os.path.join(PKG, "bin", "stream_probe")


def build_probe(verbose: bool = False) -> str:
    """The memory-ceiling probe of the dense stencil's traffic (csrc/stream_probe.hip)."""
    os.makedirs(os.path.dirname(PROBE), exist_ok=True)
    tmp = PROBE + ".tmp"
    _run_all([[HIPCC] + COMMON + ["-o", tmp, os.path.join(CSRC, "stream_probe.hip")]])
    os.replace(tmp, PROBE)
    if verbose:
        print("built", PROBE)
    return PROBE


def build_oracle(verbose: bool = False) -> str:
    d = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", d, "-B"], check=True)
    # the reference's own host sources, compiled where they lie, when that tree is present
    if os.path.isfile("/root/reference/src/timeseries.cpp"):
        subprocess.run(["make", "-s", "-C", d, "-B", "ref"], check=True)
    if verbose:
        print("built", os.path.join(d, "liboracle.so"))
    return os.path.join(d, "liboracle.so")


def build_all(verbose: bool = False) -> None:
    build_lib(verbose)
    build_cli(verbose)
    build_probe(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
