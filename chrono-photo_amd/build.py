"""Builds the HIP libraries of this package in-tree for gfx950.

libchrono_hip.so   -- the compositing library behind include/chrono_hip.h
libchrono_synth.so -- seeded synthetic stacks (tests and benchmarks)
chrono-photo-hip   -- the command-line tool (cli/), linked against libchrono_hip.so

Every kernel translation unit is compiled with the flags that keep the reference's f32
semantics: no FMA contraction, correctly rounded division and square root, no denormal
flushing, and no packed-f32 instruction selection.
"""
import os
import re
import subprocess
import sys
import concurrent.futures as cf

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
CLI = os.path.join(HERE, "cli")
OUT = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function",
]

LIBS = {
    "libchrono_hip.so": ["abi.hip", "kernels.hip", "kernels_fast_b1_c3.hip", "kernels_fast_b1_c4.hip",
                         "kernels_fast_b2_c3.hip", "kernels_fast_b2_c4.hip", "kernels_u8_c1.hip",
                         "kernels_u8_c2.hip", "kernels_u8_c3.hip", "kernels_u8_c4.hip", "kernels_batch_u8_c1.hip",
                         "kernels_batch_u8_c2.hip", "kernels_batch_u8_c3.hip", "kernels_batch_u8_c4.hip",
                         "kernels_batch_fast_b1_c3.hip", "kernels_batch_fast_b1_c4.hip",
                         "kernels_batch_fast_b2_c3.hip", "kernels_batch_fast_b2_c4.hip",
                         "shake.hip", "batch.hip", "server.hip", "outofcore.hip"],
    "libchrono_synth.so": ["synth.hip"],
}


def _deps(src, seen=None):
    """The source and every local header it includes, transitively."""
    seen = set() if seen is None else seen
    path = os.path.join(CSRC, src)
    if path in seen or not os.path.exists(path):
        return seen
    seen.add(path)
    with open(path) as f:
        for line in f:
            m = re.match(r'\s*#include\s+"([^"]+)"', line)
            if m:
                _deps(os.path.normpath(os.path.join(os.path.dirname(src), m.group(1))), seen)
    return seen


def _compile(src, changed_only=False):
    obj = os.path.join(OUT, "obj", os.path.basename(src) + ".o")
    if changed_only and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if all(os.path.getmtime(d) <= t for d in _deps(src)):
            return obj
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    lines = [l for l in (r.stdout + r.stderr).splitlines()
             if "packed-fp32-ops" not in l or "error" in l]
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n" + "\n".join(lines))
    return obj


def build(verbose=False, changed_only=False):
    """Compiles every translation unit (changed_only: only those whose source or local
    headers are newer than their object, for development) and links the libraries."""
    os.makedirs(os.path.join(OUT, "obj"), exist_ok=True)
    # the kernel translation units take minutes each: start the biggest first (the packed-byte
    # ones, then the 16-bit-halves ones), so that the small ones fill in at the end
    def cost(src):
        return (0 if "_u8_" in src else 1 if "_fast_" in src else 2, src)
    srcs = sorted({s for v in LIBS.values() for s in v}, key=cost)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = dict(zip(srcs, ex.map(lambda s: _compile(s, changed_only), srcs)))
    for lib, parts in LIBS.items():
        out = os.path.join(OUT, lib)
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp",
               *[objs[p] for p in parts]]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}{r.stderr}")
        os.replace(out + ".tmp", out)
        if verbose:
            print("built", out)
    build_cli(verbose)
    return OUT


def build_cli(verbose=False):
    """Host-only C++ (g++): the CLI reaches the GPU through the C ABI alone."""
    out = os.path.join(OUT, "chrono-photo-hip")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", out + ".tmp",
           os.path.join(CLI, "main.cpp"), os.path.join(CLI, "imageio.cpp"), os.path.join(CLI, "jpeg.cpp"),
           "-L" + OUT, "-lchrono_hip", "-lz", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"CLI build failed: {' '.join(cmd)}\n{r.stdout}{r.stderr}")
    os.replace(out + ".tmp", out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    if sys.argv[1:] == ["cli"]:
        build_cli(verbose=True)
    else:
        build(verbose=True, changed_only="--changed" in sys.argv[1:])
    sys.exit(0)
