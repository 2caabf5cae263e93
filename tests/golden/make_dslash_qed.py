"""Writes dslash_qed.json from the reference's examples-py/dslash-qed.log.json:
every entry with a value, as [name, [[re, im], ...], eps] in the order the
script appends them, plus the geometry lines as notes."""

import json
import os
import sys

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/examples-py/dslash-qed.log.json"
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dslash_qed.json")
rows = []
for e in json.load(open(src)):
    if len(e) == 3:
        vals = [[z["real"], z["imag"]] for z in e[1]["value"]]
        rows.append([e[0], vals, e[2]])
    else:
        rows.append([e[0]])
json.dump({"source": "examples-py/dslash-qed.log.json of Qlattice v1.11 (examples-py/dslash-qed.py)",
           "entries": rows}, open(out, "w"), indent=1)
print(out, len(rows))
