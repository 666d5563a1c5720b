"""Compiles libissampler.so for gfx950 with hipcc (in-tree).

The host engine (csrc/engine.hip), the batched AMIS host side (csrc/batch.hip),
the serving loop (csrc/serve.hip),
four objects per dimension instantiation (csrc/dim.hip with -DIS_DIM=D, D in the
IS_DIMS list of csrc/launch.h, and -DIS_DIM_PART=0..3: new-sample kernels and the
sweep / round / target kernels, each in FP32 and FP64) and one per batched-kernel shape (csrc/batch_k.hip
with -DIS_BATCH_FP and -DIS_BATCH_W) are compiled in parallel, then linked into
one shared library."""

import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libissampler.so")
OBJDIR = os.path.join(HERE, "_build")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # keep the compiler from emitting packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]
NOISE = "packed-fp32-ops"
# MFMA accumulators in VGPRs (one register file on gfx950): no copies between the
# accumulation registers and the VGPRs that rescale or read the tiles.  Not for the
# FP64 extended-range unit (dim64 part 5), which this compiler fails to build with it.
VGPR_MFMA = ["-mllvm", "-amdgpu-mfma-vgpr-form"]
BATCH_WAVES = (2, 4, 8)  # waves per workgroup of the batched AMIS kernel (batch_k.hip)


def dims():
    with open(os.path.join(CSRC, "launch.h")) as f:
        text = f.read()
    body = text[text.index("#define IS_DIMS(X)"):]
    body = body[: body.index("\n\n")]
    return [int(v) for v in re.findall(r"X\((\d+)\)", body)]


def _unit_cost(extra):
    for flag in extra:
        if flag.startswith("-DIS_DIM="):
            return int(flag.split("=")[1])
    return 20 if any(f.startswith("-DIS_BATCH") for f in extra) else 10  # batch kernels, then the host-side units


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(line for line in res.stderr.splitlines() if NOISE not in line)
    return res.returncode, err


def units():
    """(source, object, extra flags) of every compilation unit of the library."""
    units = [(os.path.join(CSRC, "engine.hip"), os.path.join(OBJDIR, "engine.o"), list(VGPR_MFMA)),
             (os.path.join(CSRC, "batch.hip"), os.path.join(OBJDIR, "batch.o"), list(VGPR_MFMA)),
             (os.path.join(CSRC, "serve.hip"), os.path.join(OBJDIR, "serve.o"), list(VGPR_MFMA))]
    for fp in (32, 64):
        for w in BATCH_WAVES:
            units.append((os.path.join(CSRC, "batch_k.hip"), os.path.join(OBJDIR, f"batch_k{fp}_w{w}.o"),
                          [f"-DIS_BATCH_FP={fp}", f"-DIS_BATCH_W={w}", *VGPR_MFMA]))
    for d in dims():
        # parts 4, 5: the extended-range kernels; 6, 7: the grouped fused rounds (d <= 15)
        parts = list(range(4)) + ([4, 5] if d == 64 else []) + ([6, 7] if d <= 15 else [])
        for part in parts:
            vgpr = [] if (d == 64 and part == 5) else VGPR_MFMA
            units.append((os.path.join(CSRC, "dim.hip"), os.path.join(OBJDIR, f"dim{d}_p{part}.o"),
                          [f"-DIS_DIM={d}", f"-DIS_DIM_PART={part}", *vgpr]))
    return units


def build(verbose=False, jobs=None, skip=None):
    """Compiles every unit (or those `skip(unit)` does not exclude) and links the library."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJDIR, exist_ok=True)
    all_units = units()
    todo = [u for u in all_units if not (skip and skip(u))]
    jobs = jobs or min(max(1, len(todo)), max(1, (os.cpu_count() or 4)))
    # the wide dimension units take the longest (their density loops unroll over up to
    # 64 coordinates): start them first so the pool stays full to the end
    order = sorted(range(len(todo)), key=lambda i: -_unit_cost(todo[i][2]))
    cmds = [[hipcc, *FLAGS, *todo[i][2], "-c", todo[i][0], "-o", todo[i][1]] for i in order]
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(_run, cmds))
    failed = [(cmd, err) for cmd, (rc, err) in zip(cmds, results) if rc != 0]
    for cmd, err in failed:
        sys.stderr.write(" ".join(cmd[-4:]) + "\n" + err + "\n")
    if failed:
        raise RuntimeError("hipcc failed for libissampler.so")
    if verbose:
        for (rc, err) in results:
            if err.strip():
                sys.stderr.write(err + "\n")
    rc, err = _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp", *[o for _, o, _ in all_units]])
    if rc != 0:
        sys.stderr.write(err + "\n")
        raise RuntimeError("link failed for libissampler.so")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
