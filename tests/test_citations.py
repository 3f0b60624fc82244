"""Every `file:line` citation of a reference file in the sources, headers and
documents of this repository points inside that file. The reference's file
lengths come from tests/golden/ref_lines.json (written by
tests/golden/make_ref_lines.py), so the test does not need the reference."""
import json
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCAN = ["oracle", "include", "la3_amd", "tests", "DESIGN.md", "INTEGRATION.md", "README.md", "bench.py",
        "__graft_entry__.py"]
CITE = re.compile(r"([A-Za-z0-9_./-]+\.(?:hpp|h|cpp))[:]([0-9]+)(?:-([0-9]+))?((?:,\s?[0-9]+(?:-[0-9]+)?)*)")


def _files():
    for r in SCAN:
        p = os.path.join(ROOT, r)
        if os.path.isfile(p):
            yield p
            continue
        for d, dirs, fs in os.walk(p):
            dirs[:] = [x for x in dirs if x not in ("__pycache__", "_build", "bin")]
            for f in fs:
                if f.endswith((".py", ".h", ".hpp", ".cpp", ".hip", ".md")):
                    yield os.path.join(d, f)


def test_reference_citations_lie_inside_their_files():
    with open(os.path.join(ROOT, "tests", "golden", "ref_lines.json")) as f:
        lines = json.load(f)
    by_base = {}
    for path, n in lines.items():
        by_base.setdefault(os.path.basename(path), []).append((path, n))
    bad, checked = [], 0
    for p in _files():
        with open(p, errors="replace") as fh:
            for ln, text in enumerate(fh, 1):
                for m in CITE.finditer(text):
                    name = m.group(1)
                    cands = by_base.get(os.path.basename(name))
                    if not cands:
                        continue  # not a reference file (e.g. this build's own sources)
                    exact = [c for c in cands if c[0].endswith(name.lstrip("./"))]
                    nmax = max(n for _, n in (exact or cands))
                    nums = [int(m.group(2))] + ([int(m.group(3))] if m.group(3) else [])
                    for part in re.findall(r"[0-9]+", m.group(4) or ""):
                        nums.append(int(part))
                    checked += 1
                    lo = int(m.group(2))
                    if max(nums) > nmax or (m.group(3) and int(m.group(3)) < lo):
                        bad.append(f"{os.path.relpath(p, ROOT)}:{ln}: {m.group(0)} (file has {nmax} lines)")
    assert checked > 100
    assert not bad, "\n".join(bad)
