"""Every reference citation in this repository's sources and documents
(src/*.f90:N or :N-M, also tc/, doc/, tools/ paths) names a file of the
reference tree and lines inside it.  Runs where the reference tree is
present (it is read as text, nothing is executed)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("DGBGK_REFERENCE", "/root/reference")
# files this project does not write (they quote the reference freely)
NOT_OURS = {"VERDICT.md", "ADVICE.md", "SURVEY.md", "BASELINE.md", "BASELINE.json"}
CITE = re.compile(r"((?:src|tc/\w+|doc|tools)/[\w./-]+?\.(?:f90|txt|md|con|OUT|plt|py|gnu|inp))"
                  r"((?::\d+(?:-\d+)?)(?:,\d+(?:-\d+)?)*)")


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="no reference tree")
def test_cited_lines_exist():
    files = subprocess.run(["git", "ls-files"], cwd=ROOT, capture_output=True,
                           text=True).stdout.split()
    nlines, bad, count = {}, [], 0
    for f in files:
        if f in NOT_OURS or f.startswith("tests/golden/") or f.endswith((".npz", ".so")):
            continue
        try:
            text = open(os.path.join(ROOT, f), encoding="utf-8").read()
        except (UnicodeDecodeError, FileNotFoundError):
            continue
        for ln, line in enumerate(text.splitlines(), 1):
            for m in CITE.finditer(line):
                path, ranges = m.group(1), m.group(2)
                full = os.path.join(REF, path)
                if not os.path.isfile(full):
                    bad.append(f"{f}:{ln}: {path} does not exist")
                    continue
                if path not in nlines:
                    with open(full, errors="ignore") as fh:
                        nlines[path] = sum(1 for _ in fh)
                for part in ranges[1:].split(","):
                    a = [int(x) for x in part.split("-")]
                    count += 1
                    if min(a) < 1 or max(a) > nlines[path] or a != sorted(a):
                        bad.append(f"{f}:{ln}: {path}:{part} (file has {nlines[path]} lines)")
        # a citation broken over two lines escapes the check above
        for ln, line in enumerate(text.splitlines(), 1):
            if re.search(r"\.f90:\s*$", line):
                bad.append(f"{f}:{ln}: citation broken over a line end")
    assert count > 100
    assert not bad, "\n".join(bad)
