"""Build ``libnsvf_amd.so`` (gfx950) in-tree with hipcc.

Usage: ``python -m nsvf_amd.build`` (or ``nsvf_amd.build.build()``).

Every HIP source is compiled with:
  -ffp-contract=off         no implicit FMA: the kernels' float arithmetic is
                            written operation by operation to match the
                            reference's rounding; the few fused multiply-adds
                            are explicit __fmaf_rn calls;
  -fhip-fp32-correctly-rounded-divide-sqrt
                            IEEE division and square root;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                            no packed FP32 (v_pk_*_f32) instructions;
  -mllvm -amdgpu-mfma-vgpr-form=1
                            MFMA accumulators in VGPRs, so the layer
                            epilogues read them without v_accvgpr_read.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_PATH = os.path.join(HERE, "libnsvf_amd.so")
ARCH = os.environ.get("NSVF_AMD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-Wall",
    "-Wno-unused-function",
]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, jobs: int = 8) -> str:
    """Compile every source to an object in a temp dir, link the shared
    library next to this file.  Always recompiles."""
    hipcc = _hipcc()
    srcs = sources()
    with tempfile.TemporaryDirectory(prefix="nsvf_amd_build_") as tmp:
        procs = []
        objs = []
        for src in srcs:
            obj = os.path.join(tmp, os.path.basename(src) + ".o")
            objs.append(obj)
            lang = ["-x", "hip"] if src.endswith(".hip") else []
            cmd = [hipcc, *HIP_FLAGS, *lang, "-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
            if len(procs) >= jobs:
                _wait(procs.pop(0))
        for p in procs:
            _wait(p)
        tmp_lib = LIB_PATH + ".tmp"
        cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp_lib]
        if verbose:
            print(" ".join(cmd), flush=True)
        out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if out.returncode != 0:
            raise RuntimeError("link failed:\n" + out.stdout.decode(errors="replace"))
        os.replace(tmp_lib, LIB_PATH)
    return LIB_PATH


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n" + out.decode(errors="replace"))
    text = out.decode(errors="replace")
    # hipcc's host pass does not know the packed-fp32 feature flag; that
    # warning is expected and dropped from the log.
    lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
