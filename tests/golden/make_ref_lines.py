"""Writes ref_lines.json: the line count of every source file of the LA3
reference tree (path relative to the tree's root -> lines). Run in a
container where the reference is present; the table is data the citation
test checks `file:line` references against, so the test itself never reads
the reference."""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
out = {}
for d, _, files in os.walk(REF):
    if "/.git" in d:
        continue
    for f in files:
        if not f.endswith((".h", ".hpp", ".cpp", ".c", ".md")) and f != "Makefile":
            continue
        p = os.path.join(d, f)
        with open(p, "rb") as fh:
            out[os.path.relpath(p, REF)] = sum(1 for _ in fh)
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_lines.json"), "w") as fh:
    json.dump(dict(sorted(out.items())), fh, indent=0)
    fh.write("\n")
print(len(out), "files")
