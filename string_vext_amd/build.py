"""Build libsv_vext.so in-tree for gfx950 with hipcc.

    python -m string_vext_amd.build [--arch gfx950]
"""
from __future__ import annotations

import argparse
import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
SOURCES = ["sv_kernels.hip", "sv_sites.hip", "sv_runtime.hip", "sv_string.hip"]
LIB = HERE / "libsv_vext.so"

# Packed FP32 (v_pk_*_f32) is disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(arch: str = "gfx950", verbose: bool = False) -> Path:
    out = HERE / "_build"
    out.mkdir(exist_ok=True)
    flags = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={arch}", "-mcode-object-version=5",
             *DEVICE_FLAGS, f"-I{ROOT / 'include'}", f"-I{HERE / 'csrc'}"]
    procs, objs = [], []
    for src in SOURCES:
        obj = out / (src + ".o")
        cmd = [hipcc(), *flags, "-c", str(HERE / "csrc" / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(str(obj))
    for cmd, proc in procs:
        text = proc.communicate()[0].decode(errors="replace")
        if proc.returncode:
            raise RuntimeError(f"hipcc failed ({proc.returncode}): {' '.join(cmd)}\n{text}")
    tmp = LIB.with_name(LIB.name + ".tmp")
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={arch}", *objs, "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


def main():
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--arch", default="gfx950")
    parser.add_argument("-v", "--verbose", action="store_true")
    args = parser.parse_args()
    print(build(args.arch, args.verbose))


if __name__ == "__main__":
    main()
