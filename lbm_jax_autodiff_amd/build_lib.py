"""Build liblbm_hip.so (gfx950) in-tree with hipcc.

Run ``python -m lbm_jax_autodiff_amd.build_lib`` (or ``__graft_entry__.build()``).
The library is written next to this file so it travels with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblbm_hip.so")
SOURCES = ["lbm_ops.hip", "lbm_ops_vjp.hip", "lbm_fused.hip", "lbm_adjoint.hip"]
ARCH = os.environ.get("LBM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm at /opt/rocm)")


def flags() -> list:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-mcode-object-version=5",
        # packed FP32 instructions are disabled for all device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-ffp-contract=fast",
        # lets x + 0 fold to x in sums that start from zero; changes no
        # result except the sign of a zero
        "-fno-signed-zeros",
        f"-I{os.path.join(ROOT, 'include')}",
        "-Wno-unused-result",
    ]


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    cc = hipcc()
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with tempfile.TemporaryDirectory(prefix="lbm_build_") as tmp:
        objs = []
        procs = []
        for src in srcs:
            obj = os.path.join(tmp, os.path.basename(src) + ".o")
            cmd = [cc, *flags(), "-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), src))
            objs.append(obj)
        for p, src in procs:
            out, _ = p.communicate()
            if p.returncode != 0:
                raise RuntimeError(f"hipcc failed on {src}:\n{out.decode(errors='replace')}")
            if verbose and out:
                print(out.decode(errors="replace"))
        tmp_lib = LIB + ".tmp"
        cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp_lib]
        res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if res.returncode != 0:
            raise RuntimeError("link failed:\n" + res.stdout.decode(errors="replace"))
        os.replace(tmp_lib, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
