"""The restated selection splice (tests/splice.py) on hand-worked cases:
the host behaviour the realign-selection path feeds (CPU only)."""
import splice as sp


PROJECT = [(">s0 first", "AC-DE"), (">s1", "WWW"), (">s2 x y", "K-LM"), (">s3", "NNPP"), (">s4", "Q")]


def test_label_and_fasta_rules():
    assert sp.label_of(">s2 x y") == "s2" and sp.label_of(">abc") == "abc"
    text = sp.write_fasta([(">a", "A" * 170)])
    assert text.splitlines() == [">a", "A" * 80, "A" * 80, "A" * 10]
    labels, seqs, headers = sp.read_fasta(">a d\n;c\nAC\nDE\n\n>b\nW\n")
    assert labels == ["a", "b"] and seqs == ["ACDE", "W"] and headers == [">a d", ">b"]


def test_export_keeps_gaps_and_headers():
    text = sp.export_selection(PROJECT, [2, 3])
    assert text == ">s2 x y\nK-LM\n>s3\nNNPP\n"


def test_splice_replaces_in_place_and_keeps_the_rest():
    out = ">s2 x y\nK-LM--\n>s3\nN-NP-P\n"
    new = sp.splice(PROJECT, [2, 3], out)
    assert new == [(">s0 first", "AC-DE"), (">s1", "WWW"), (">s2 x y", "K-LM--"), (">s3", "N-NP-P"), (">s4", "Q")]


def test_splice_follows_output_order():
    # the aligner may reorder the selection: the output order wins
    out = ">s3\nNNPP\n>s2\nKLM-\n"
    new = sp.splice(PROJECT, [2, 3], out)
    assert [h for h, _ in new] == [">s0 first", ">s1", ">s3", ">s2 x y", ">s4"]
    assert new[3][1] == "KLM-"


def test_splice_skips_unknown_labels():
    # a record whose label is not in the project is skipped, but it still
    # counts in the target index first + l (Project.cpp:1008-1024)
    new = sp.splice(PROJECT, [1, 2], ">zz\nAAA\n>s1\nW-WW\n")
    assert [h for h, _ in new] == [">s0 first", ">s2 x y", ">s1", ">s3", ">s4"]
    assert new[2] == (">s1", "W-WW")
