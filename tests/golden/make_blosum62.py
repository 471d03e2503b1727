"""Extract the reference's substitution table as a data fixture.

Reads tweakseq/Core/Annotations/Consensus.cpp as *text* (nothing is compiled or
imported) and writes blosum62.json beside this script:

  order  -- the row/column alphabet, taken from the row comments (":34-59")
  matrix -- the 23x23 integer table ("static int BLOSUM62[23][23]", :34-59)
  map    -- the 26-entry letter map 'A'..'Z' -> row ("BLOSUM62map", :61-69)

Run once in a container that has the reference mounted:
    python tests/golden/make_blosum62.py /root/reference
"""
import json
import os
import re
import sys


def extract(ref_root):
    path = os.path.join(ref_root, "tweakseq", "Core", "Annotations", "Consensus.cpp")
    text = open(path, encoding="latin-1").read()
    m = re.search(r"BLOSUM62\s*\[23\]\s*\[23\]\s*=\s*\{(.*?)\};", text, re.S)
    body = m.group(1)
    rows, order = [], []
    for line in body.splitlines():
        if "{" not in line:
            continue
        nums = re.search(r"\{([^}]*)\}", line).group(1)
        rows.append([int(x) for x in nums.replace(" ", "").split(",")])
        order.append(re.search(r"//\s*\d+\s+([A-Z])", line).group(1))
    m2 = re.search(r"BLOSUM62map\s*\[26\]\s*=\s*\{(.*?)\};", text, re.S)
    cmap = [int(x) for x in re.findall(r"-?\d+", m2.group(1))]
    assert len(rows) == 23 and all(len(r) == 23 for r in rows), "table shape"
    assert len(cmap) == 26, "map shape"
    return {"order": "".join(order), "matrix": rows, "map": cmap,
            "source": "tweakseq/Core/Annotations/Consensus.cpp:34-69"}


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = extract(root)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "blosum62.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", out, data["order"])
