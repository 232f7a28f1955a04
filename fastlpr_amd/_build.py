"""Build libfastlpr_amd.so in-tree with hipcc for gfx950."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/flpr_points.hip", "csrc/flpr_conv.hip", "csrc/flpr_select.hip",
           "csrc/flpr_resample.hip", "csrc/flpr_mlp.hip"]
OUT = os.path.join(HERE, "libfastlpr_amd.so")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 math stays off (v_pk_*_f32 are not generated)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every source to an object in parallel (one hipcc process per
    file), then link the shared library."""
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in FLAGS if f != "-shared"] + ["-c"]
    procs = []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        procs.append((src, obj, subprocess.Popen([hipcc()] + cflags + ["-o", obj, src],
                                                 stdout=subprocess.PIPE,
                                                 stderr=subprocess.PIPE, text=True)))
    errs = []
    for src, obj, p in procs:
        _out, err = p.communicate()
        if p.returncode != 0:
            errs.append(f"{os.path.basename(src)} ({p.returncode}):\n{err[-4000:]}")
        elif verbose and err:
            print(err)
    if errs:
        raise RuntimeError("hipcc failed: " + "\n".join(errs))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for _s, o, _p in procs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed ({res.returncode}):\n{res.stderr[-4000:]}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
