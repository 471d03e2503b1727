"""The tweakseq integration class, GpuAlign (tweakseq_amd/qt), compiled against
stand-in Qt headers (tests/native/fakeqt; the image has no Qt) into
tweakalign-gpualign-check by the build:

* its process path hands the CLI MUSCLE's argument list, "-in FIN -out FOUT"
  with the configured executable (tweakseq/Core/Muscle.cpp:49-53), which the
  tweakalign CLI accepts (tests/test_cli.py);
* its CIGAR rendering turns the oracle's CIGARs into two gapped rows whose
  letters are the inputs and whose columns rescore to the oracle's score,
  and rejects runs that do not consume both sequences;
* without a GPU, alignPair fails with the library's message (CPU); on the
  GPU it returns the oracle's alignment (gpu test)."""
import os
import random
import subprocess

import pytest

from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tweakseq_amd", "tweakalign-gpualign-check")
AA = "ARNDCQEGHILKMFPSTWYV"

needs_check = pytest.mark.skipif(not os.path.exists(CHECK), reason="tweakalign-gpualign-check not built")


def run(*args, stdin=None, env=None):
    r = subprocess.run([CHECK] + list(args), input=stdin, capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 0, r.stderr
    return r.stdout


def rescore(ga, gb, go=-10, ge=-1):
    """Score of two gapped rows: BLOSUM62 per aligned column, open + (k-1) * ext per gap run."""
    smat = po.protein_matrix()[0]
    score, run_a, run_b = 0, 0, 0
    for x, y in zip(ga, gb):
        if x != "-" and y != "-":
            score += int(smat[po.encode(x)[0], po.encode(y)[0]])
            run_a = run_b = 0
        elif x == "-":
            score += ge if run_a else go
            run_a, run_b = 1, 0
        else:
            score += ge if run_b else go
            run_a, run_b = 0, 1
    return score


@needs_check
def test_process_path_is_the_muscle_argument_list():
    out = run("command", "/tmp/in.fasta", "/tmp/out.fasta", "/opt/tweakalign/tweakalign").splitlines()
    assert out == ["name tweakalign", "exec /opt/tweakalign/tweakalign", "arg -in", "arg /tmp/in.fasta",
                   "arg -out", "arg /tmp/out.fasta", "stdout 0"]


@needs_check
def test_render_oracle_cigars():
    rng = random.Random(5)
    p = po.params(mode=0, gap_open=-10, gap_extend=-1)
    lines, want = [], []
    for _ in range(60):
        a = "".join(rng.choice(AA) for _ in range(rng.randint(1, 80)))
        b = a[: len(a) // 2] + "".join(rng.choice(AA) for _ in range(rng.randint(1, 40))) \
            if rng.random() < 0.6 else "".join(rng.choice(AA) for _ in range(rng.randint(1, 80)))
        res = po.align(p, po.encode(a), po.encode(b))
        lines.append("%s %s %s" % (a, b, res[5]))
        want.append((a, b, res[0]))
    got = run("render", stdin="\n".join(lines) + "\n").splitlines()
    assert len(got) == len(want)
    for line, (a, b, score) in zip(got, want):
        ga, gb = line.split()
        assert len(ga) == len(gb)
        assert ga.replace("-", "") == a and gb.replace("-", "") == b
        assert not any(x == "-" and y == "-" for x, y in zip(ga, gb))
        assert rescore(ga, gb) == score


@needs_check
def test_render_rejects_malformed_runs():
    bad = ["AC ACD 3M", "AC ACD 2M", "AC AC 2M1", "AC AC 0M2M", "AC AC 2X", "AC AC 1M1M1I", "- A 1I"]
    assert run("render", stdin="\n".join(bad) + "\n").splitlines() == ["FAIL"] * len(bad)
    assert run("render", stdin="- A 1D\nA - 1I\n").splitlines() == ["- A", "A -"]


@needs_check
def test_align_pair_without_gpu_reports_the_library_error():
    env = {**os.environ, "HIP_VISIBLE_DEVICES": "-1", "ROCR_VISIBLE_DEVICES": "-1"}
    assert run("pair", "ACDE", "ACE", env=env).strip() == "ERR no usable GPU"


@needs_check
@pytest.mark.gpu
def test_align_pair_on_gpu_equals_oracle():
    rng = random.Random(9)
    p = po.params(mode=0, gap_open=-10, gap_extend=-1)
    for _ in range(6):
        a = "".join(rng.choice(AA) for _ in range(rng.randint(20, 300)))
        b = a[5:] + "".join(rng.choice(AA) for _ in range(rng.randint(1, 30)))
        out = run("pair", a, b).split()
        assert out[0] == "OK", out
        ga, gb = out[1], out[2]
        res = po.align(p, po.encode(a), po.encode(b))
        assert ga.replace("-", "") == a and gb.replace("-", "") == b
        assert rescore(ga, gb) == res[0]
        want = run("render", stdin="%s %s %s\n" % (a, b, res[5])).split()
        assert [ga, gb] == want
