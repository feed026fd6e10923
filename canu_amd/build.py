"""Compile the native libraries in-tree (hipcc for gfx950, gcc for host-only code)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             # double-precision branch-point scores must round like the reference: no FMA contraction
             "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)


def build_hip():
    out = os.path.join(HERE, "liboic_hip.so")
    _run([HIPCC] + HIP_FLAGS + ["-o", out, os.path.join(CSRC, "oic.hip")])
    return out


def build_readsim():
    out = os.path.join(HERE, "libreadsim.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-Wall", "-o", out, os.path.join(CSRC, "readsim.c"), "-lm"])
    return out


def build_cli():
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    out = os.path.join(bindir, "overlapInCore")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", out, os.path.join(CSRC, "oic_cli.cpp"),
          "-I" + os.path.join(ROOT, "include"), "-L" + HERE, "-loic_hip", "-Wl,-rpath,$ORIGIN/.."])
    return out


def build_all():
    return [build_readsim(), build_hip()] + ([build_cli()] if os.path.exists(os.path.join(CSRC, "oic_cli.cpp")) else [])
