"""Build the in-tree native libraries.

    python -m drnum_amd.build          # libdrnhip.so (hipcc, gfx950) + libdrnum.so (g++)
                                       # + the applications and the VALU rate probe

Every HIP source is compiled for gfx950 only.  Packed FP32 instructions are
disabled (-packed-fp32-ops): hipcc would otherwise emit v_pk_* forms that are
not trusted on the target machines.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    src = os.path.join(HERE, "csrc", "drnhip.hip")
    out = os.path.join(LIB, "libdrnhip.so")
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-o", tmp, src])
    os.replace(tmp, out)
    return out


def build_probe():
    """The FP32 VALU issue-rate probe (drnum_amd/csrc/valu_probe.hip -> lib/valu_probe)."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "valu_probe")
    tmp = out + ".tmp"
    flags = [f for f in HIP_FLAGS if f not in ("-fPIC", "-shared")]
    _run([HIPCC] + flags + ["-o", tmp, os.path.join(HERE, "csrc", "valu_probe.hip")])
    os.replace(tmp, out)
    return out


def build_host():
    src_dir = os.path.join(HERE, "host")
    srcs = sorted(os.path.join(src_dir, f) for f in os.listdir(src_dir) if f.endswith(".cpp"))
    if not srcs:
        return None
    out = os.path.join(LIB, "libdrnum.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fopenmp",
          "-I" + os.path.join(ROOT, "include"), "-o", tmp] + srcs +
         ["-L" + LIB, "-ldrnhip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_apps():
    """C++ applications on the drop-in classes (drnum_amd/apps/*.cpp -> lib/<name>)."""
    src_dir = os.path.join(HERE, "apps")
    outs = []
    for f in sorted(os.listdir(src_dir)):
        if not f.endswith(".cpp"):
            continue
        out = os.path.join(LIB, f[:-4])
        tmp = out + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"), "-o", tmp,
              os.path.join(src_dir, f), "-L" + LIB, "-ldrnum", "-ldrnhip", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, out)
        outs.append(out)
    return outs


def build_all():
    build_hip()
    build_host()
    build_apps()
    build_probe()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
