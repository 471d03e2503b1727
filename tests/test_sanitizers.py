"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

tests/native/sanitize_host.cpp drives the CLI's FASTA reader / writer and
CIGAR helpers, the seeded workload generator and the CPU oracle on random
and edge-case inputs; built with g++/gcc -fsanitize=address,undefined and
-fno-sanitize-recover, any report fails the run.  GPU code is not built
with sanitizers (the GPU sanitizers are not available on the GPU hosts)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no host compiler")
def test_host_code_under_asan_ubsan(tmp_path):
    objs = []
    for c in ("gotoh.c", "consensus.c", "msa.c", "timing.c"):
        o = str(tmp_path / (c + ".o"))
        subprocess.run(["gcc", "-std=c11", "-fopenmp", "-c", os.path.join(ROOT, "oracle", c), "-o", o] + SAN,
                       check=True)
        objs.append(o)
    exe = str(tmp_path / "sanitize_host")
    subprocess.run(["g++", "-std=c++17", "-pthread", "-fopenmp",
                    os.path.join(ROOT, "tests", "native", "sanitize_host.cpp"),
                    os.path.join(ROOT, "tweakseq_amd", "csrc", "workload.cpp"), "-o", exe] + objs + SAN,
                   check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and "sanitize ok" in r.stdout, r.stdout + r.stderr
