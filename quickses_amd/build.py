"""Build the HIP library (libquickses.so, gfx950) and the CPU oracle library in-tree."""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libquickses.so")
CLI = os.path.join(PKG, "quickses")
OBJ_DIR = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # every floating-point operation rounds once: parity with the single-precision formulas
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
SOURCES = ["qs_kernels.hip", "qs_engine.hip", "qs_serve.hip", "qs_io.cpp", "qs_synth.cpp"]


def _run(cmd, cwd=None):
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return p.stdout


def build_library():
    objs = []
    procs = []
    os.makedirs(OBJ_DIR, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(OBJ_DIR, src.rsplit(".", 1)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            raise RuntimeError("build step failed: " + " ".join(cmd))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_cli():
    src = os.path.join(CSRC, "qs_cli.cpp")
    if not os.path.exists(src):
        return None
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-o", tmp, src, "-L" + PKG, "-lquickses", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath-link,/opt/rocm/lib"])
    os.replace(tmp, CLI)
    return CLI


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_library()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
