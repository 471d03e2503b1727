"""The host side of "realign selection", restated for tests (the Qt host
cannot be built here): what tweakseq exports for a selection and how it
splices the aligner's output back.  Test infrastructure only.

  * export: tweakseq/Core/Project.cpp:883-894 (exportSelectionFASTA) writes
    each selected sequence's header line verbatim and its residues as
    Sequence::filter returns them (gaps kept), through FASTAFile::write
    (tweakseq/Core/FASTAFile.cpp:149-171: 80-column lines);
  * read: FASTAFile::read (FASTAFile.cpp:71-147) with labels from
    parseComment (:177-187, the header up to the first space);
  * splice: Project::readNewAlignment, selection branch
    (tweakseq/Core/Project.cpp:964-1026): copy the old sequences, find the
    first selected index, then for output record l set that label's
    residues to the aligned row and move it to first + l (positions taken in
    the list as it is after the previous moves).
"""


def label_of(header):
    i = header.find(" ", 1)
    return header[1:] if i == -1 else header[1:i]


def write_fasta(records):
    out = []
    for header, seq in records:
        out.append(header)
        for k in range(0, len(seq), 80):
            out.append(seq[k:k + 80])
    return "\n".join(out) + "\n"


def read_fasta(text):
    """(labels, seqs, headers) as FASTAFile::read returns them."""
    labels, seqs, headers = [], [], []
    state = "seek"
    for line in text.splitlines():
        s = line.strip()
        if not s:
            continue
        c = s[0]
        if state == "seek" or state == "seq":
            if c in ";>":
                headers.append(s)
                labels.append(label_of(s))
                seqs.append("")
                state = "comment"
            elif state == "seq":
                seqs[-1] += s
        elif state == "comment":
            if c == ";":
                continue
            seqs[-1] = s
            state = "seq"
    return labels, seqs, headers


def export_selection(project, selected):
    """project: list of (header, residues); selected: indices in selection order."""
    return write_fasta([project[k] for k in selected])


def splice(project, selected, out_text):
    """The project after readNewAlignment(fname, isFullAlignment=false)."""
    labels, seqs, _ = read_fasta(out_text)
    new = [list(x) for x in project]  # copies of the old sequences
    first = min(selected)             # index of the first selected sequence in the old order
    for l, (lab, row) in enumerate(zip(labels, seqs)):
        idx = next((k for k, (h, _) in enumerate(new) if label_of(h) == lab), None)
        if idx is None:
            continue  # "missed aligned sequence"
        new[idx][1] = row
        item = new.pop(idx)  # QList::move(from, to)
        new.insert(first + l, item)
    return [tuple(x) for x in new]
