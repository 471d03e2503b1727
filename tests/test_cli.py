"""The tweakalign CLI: the drop-in for tweakseq's external aligner slot.

CPU tests cover the FASTA contract (tweakseq/Core/FASTAFile.cpp:71-187) and
argument handling; GPU tests cover the alignments it writes."""
import os
import random
import subprocess

import pytest

import tweakseq_amd as ta
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "tweakseq_amd", "tweakalign")
AA = "ARNDCQEGHILKMFPSTWYV"


def run(args, **kw):
    return subprocess.run([CLI] + args, capture_output=True, text=True, **kw)


def read_fasta(path):
    recs, cur = [], None
    for line in open(path):
        line = line.strip()
        if not line:
            continue
        if line[0] in ">;":
            cur = [line, ""]
            recs.append(cur)
        else:
            cur[1] += line
    return [(h, s) for h, s in recs]


def write_fasta(path, recs, width=60):
    with open(path, "w") as f:
        for h, s in recs:
            f.write(h + "\n")
            for p in range(0, len(s), width):
                f.write(s[p:p + width] + "\n")


def test_version_probe_second_token():
    out = run(["-version"])
    assert out.returncode == 0 and len(out.stdout.split()) >= 2  # Muscle.cpp:100-113 takes token 1


def test_fasta_roundtrip_rules(tmp_path):
    src = tmp_path / "in.fa"
    src.write_text(">a first  \nAC-DE\n  FG..H\n\n>b\nWW\nWW\n>c\n" + "K" * 170 + "\n")
    out = tmp_path / "out.fa"
    r = run(["-in", str(src), "-out", str(out), "--fasta-roundtrip"])
    assert r.returncode == 0, r.stderr
    lines = out.read_text().splitlines()
    # header verbatim (trimmed), gaps removed, 80-column wrap
    assert lines[0] == ">a first" and lines[1] == "ACDEFGH"
    assert lines[2] == ">b" and lines[3] == "WWWW"
    assert lines[4] == ">c" and lines[5] == "K" * 80 and lines[6] == "K" * 80 and lines[7] == "K" * 10


def test_argument_styles_and_errors(tmp_path):
    src = tmp_path / "in.fa"
    src.write_text(">x\nACD\n")
    # ClustalO-style flags are accepted
    r = run(["--force", "-v", "--outfmt=fa", "--output-order=tree-order", "-i", str(src), "-o",
             str(tmp_path / "o.fa"), "--fasta-roundtrip"])
    assert r.returncode == 0, r.stderr
    # MAFFT style: output on stdout
    r = run(["--auto", "--thread", "-1", str(src), "--fasta-roundtrip"])
    assert r.returncode == 0 and r.stdout.splitlines() == [">x", "ACD"]
    assert run(["-in", str(tmp_path / "missing.fa"), "-out", str(tmp_path / "o.fa")]).returncode != 0
    assert run(["-bogus"]).returncode != 0


def rand_seq(rng, n):
    return "".join(rng.choice(AA) for _ in range(n))


def mutate(rng, s):
    out = []
    for c in s:
        r = rng.random()
        if r < 0.05:
            continue
        if r < 0.1:
            out.append(rng.choice(AA))
        out.append(c if rng.random() > 0.2 else rng.choice(AA))
    return "".join(out)


@pytest.mark.gpu
def test_two_sequences_is_the_global_alignment(tmp_path):
    rng = random.Random(1)
    a = rand_seq(rng, 230)
    b = mutate(rng, a)
    src, out = tmp_path / "in.fa", tmp_path / "out.fa"
    write_fasta(str(src), [(">q query", a[:50] + "-" + a[50:]), (">r", b)])
    r = run(["-in", str(src), "-out", str(out)])
    assert r.returncode == 0, r.stderr
    recs = read_fasta(str(out))
    assert [h for h, _ in recs] == [">q query", ">r"]
    ra, rb = recs[0][1], recs[1][1]
    assert len(ra) == len(rb) and ra.replace("-", "") == a and rb.replace("-", "") == b
    want = po.align(po.params(mode=0), po.encode(a), po.encode(b))
    ops = "".join("M" if x != "-" and y != "-" else ("I" if y == "-" else "D") for x, y in zip(ra, rb))
    cig, k = "", 0
    while k < len(ops):
        e = k
        while e < len(ops) and ops[e] == ops[k]:
            e += 1
        cig += "%d%s" % (e - k, ops[k])
        k = e
    assert cig == want[5]


@pytest.mark.gpu
def test_progressive_msa(tmp_path):
    rng = random.Random(2)
    base = rand_seq(rng, 180)
    seqs = [mutate(rng, base) for _ in range(7)]
    src, out = tmp_path / "in.fa", tmp_path / "out.fa"
    write_fasta(str(src), [(">s%d" % k, s) for k, s in enumerate(seqs)])
    r = run(["-in", str(src), "-out", str(out)])
    assert r.returncode == 0, r.stderr
    recs = read_fasta(str(out))
    rows = [s for _, s in recs]
    assert [h for h, _ in recs] == [">s%d" % k for k in range(7)]
    assert [x.replace("-", "") for x in rows] == seqs
    # The oracle's progressive alignment of the same sequences, row for row.
    alen, cols, _ = po.msa([po.encode(x) for x in seqs])
    assert rows == po.msa_rows(seqs, alen, cols)

