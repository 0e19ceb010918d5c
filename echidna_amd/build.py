"""Build libechidna_hip.so (HIP for gfx950) in-tree with hipcc.

Each source compiles to its own object (in parallel, in a temporary
directory), then hipcc links the shared library next to this file.
"""

from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libechidna_hip.so")

SOURCES = ["ehip_api.cpp", "ehip_program.cpp", "ehip_kernels.hip", "ehip_stde.hip", "ehip_cv.hip",
           "ehip_cv_api.cpp", "ehip_sparse.cpp", "ehip_serde.cpp", "ehip_spec.cpp", "ehip_noalloc.hip", "ehip_jetgemm.hip"]
# Headers whose text the library carries for hipRTC (tape-specialised kernels):
# (file, C symbol).
EMBED = [("ehip_math.h", "ehip_math_h_text"), ("philox.h", "philox_h_text")]
# The interpreters are built once per program-fetch mode, each in its own
# namespace (object name, defines).
VARIANTS = {"ehip_kernels.hip": [("ehip_kernels_lds", ["-DEHIP_SFETCH=0", "-DEHIP_NS=lds_fetch"]),
                                 ("ehip_kernels_sfetch", ["-DEHIP_SFETCH=1", "-DEHIP_NS=scalar_fetch"])]}
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # Reference arithmetic never contracts a*b+c into an FMA.
    "-ffp-contract=off",
    # Keep the compiler from forming packed FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# The interpreters are one loop around every opcode's code.  Machine LICM
# hoists each math routine's polynomial constants out of that loop into
# VGPRs held for the whole kernel (f64 VOP3 instructions take no 64-bit
# literals), which took the f64 interpreter to 256 VGPRs plus AGPR spills
# and one wave per SIMD; without it each constant is materialised where it
# is used and the f64 interpreter needs about 100 VGPRs.
NO_LICM = ["-mllvm", "-disable-machine-licm"]
EXTRA = {"ehip_kernels.hip": NO_LICM, "ehip_stde.hip": NO_LICM}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    return r.returncode


def _embed_source(path: str) -> str:
    """A C++ source defining each EMBED header's text as a char array."""
    out = []
    for name, sym in EMBED:
        text = open(os.path.join(CSRC, name), "rb").read()
        body = ",".join(str(b) for b in text)
        out.append(f'extern "C" const char {sym}[] = {{{body},0}};\n')
    with open(path, "w") as f:
        f.write("".join(out))
    return path


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8"))))
    with tempfile.TemporaryDirectory(prefix="ehip_build_") as tmpdir:
        objs, cmds = [], []
        for s in SOURCES:
            for name, defs in VARIANTS.get(s, [(s, [])]):
                obj = os.path.join(tmpdir, name + ".o")
                objs.append(obj)
                cmds.append([hipcc, *FLAGS, *EXTRA.get(s, []), *defs, "-c", "-o", obj, os.path.join(CSRC, s)])
        emb = _embed_source(os.path.join(tmpdir, "ehip_embed.cpp"))
        objs.append(os.path.join(tmpdir, "ehip_embed.o"))
        cmds.append([hipcc, "-O1", "-fPIC", "-c", "-o", objs[-1], emb])
        with ThreadPoolExecutor(jobs) as ex:
            rcs = list(ex.map(lambda c: _run(c, verbose), cmds))
        if any(rcs):
            raise RuntimeError("hipcc failed building libechidna_hip.so")
        tmp = LIB + ".tmp"
        if _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp, *objs,
                 "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"],
                verbose):
            raise RuntimeError("hipcc failed linking libechidna_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
