"""Build the native parts of tweakseq_amd in-tree (no JIT cache).

  libtweakalign.so  HIP kernels (gfx950) + host runtime + C ABI
  tweakalign        CLI honouring the external-aligner "-in FILE -out FILE" contract
  tweakalign-serve-bench   open-loop serving benchmark
  tweakalign-pair-bench    call -> CIGAR latency of one pair (single-pair path and batch path)
  tweakalign-valu-peak     measured VALU throughput of the fill's instruction kinds
  tweakalign-gpualign-check  the tweakseq integration class (qt/GpuAlign) against stand-in Qt headers

Usage: python -m tweakseq_amd.build   (or build() from __graft_entry__)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtweakalign.so")
CLI = os.path.join(HERE, "tweakalign")
ARCH = os.environ.get("TWEAKALIGN_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # packed-FP32 instruction forms stay off (none of this code is FP32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def run(cmd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode:
        raise RuntimeError("build step failed: " + cmd[0])


def build_lib():
    srcs = [os.path.join(CSRC, f) for f in ("runtime.hip", "serve.hip", "consensus.hip", "msa.hip", "pair.hip", "devq.hip")]
    objs = []
    for s in srcs:
        o = os.path.join(CSRC, os.path.basename(s) + ".o")
        run(["hipcc", "-c"] + HIP_FLAGS + [s, "-o", o])
        objs.append(o)
    wo = os.path.join(CSRC, "workload.o")
    run(["g++", "-c", "-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", os.path.join(CSRC, "workload.cpp"), "-o", wo])
    objs.append(wo)
    tmp = LIB + ".tmp"
    run(["hipcc", "-shared", "--offload-arch=" + ARCH, "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)


def build_tools():
    """Executables linked against libtweakalign.so (rpath = package dir)."""
    tools = [("cli.cpp", CLI), ("serve_bench.cpp", os.path.join(HERE, "tweakalign-serve-bench")),
             ("pair_bench.cpp", os.path.join(HERE, "tweakalign-pair-bench"))]
    for src, out in tools:
        src = os.path.join(CSRC, src)
        if not os.path.exists(src):
            continue
        run(["g++", "-O2", "-std=c++17", "-Wall", "-pthread", src, "-o", out, "-L" + HERE, "-ltweakalign",
             "-Wl,-rpath,$ORIGIN"])
    run(["hipcc"] + HIP_FLAGS + [os.path.join(CSRC, "valu_peak.hip"), "-o", os.path.join(HERE, "tweakalign-valu-peak")])
    # The tweakseq integration class (qt/GpuAlign.*) built against stand-in Qt
    # headers with its driver: the check that the maintainer-facing code
    # compiles, renders CIGARs and calls the library.
    native = os.path.join(ROOT, "tests", "native")
    run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(native, "fakeqt"), "-I" + os.path.join(ROOT, "include"),
         "-I" + os.path.join(HERE, "qt"), os.path.join(HERE, "qt", "GpuAlign.cpp"),
         os.path.join(native, "gpualign_check.cpp"), "-o", os.path.join(HERE, "tweakalign-gpualign-check"),
         "-L" + HERE, "-ltweakalign", "-Wl,-rpath,$ORIGIN"])


def build():
    build_lib()
    build_tools()


if __name__ == "__main__":
    build()
    sys.exit(0)
