"""Build of the gfx950 shared library libsponge_amd.so (in-tree).

hipcc compiles every kernel file for gfx950 only and links them with the C
ABI into sponge_amd/libsponge_amd.so.  Packed-FP32 instruction selection is
disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libsponge_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def build(verbose=True, extra=None):
    """Compile the library (always recompiles)."""
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    procs = []
    for src in sources():
        obj = os.path.join(bdir, os.path.basename(src).replace(".hip", ".o"))
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-c", src, "-o", obj] + (extra or [])
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    # host-only C++ (module facades over the C ABI)
    for src in sorted(glob.glob(os.path.join(HERE, "host", "*.cpp"))):
        obj = os.path.join(bdir, "host_" + os.path.basename(src).replace(".cpp", ".o"))
        cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-c", src, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    ok = True
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            ok = False
            sys.stderr.write(" ".join(cmd) + "\n" + out.decode(errors="replace"))
        elif verbose and out:
            txt = out.decode(errors="replace")
            # hipcc's host pass does not know the packed-fp32 feature; that is expected
            lines = [l for l in txt.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if lines:
                sys.stderr.write("\n".join(lines) + "\n")
    if not ok:
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
