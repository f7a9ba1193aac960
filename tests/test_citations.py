"""Every ``file.jl:a-b`` citation of the reference in this repository's sources and
documents points inside the cited file.

The reference's file lengths come from ``tests/golden/reference_files.json`` (written by
``tests/golden/make_reference_files.py`` from the reference checkout), so this test runs
without the checkout.  Where the checkout is present the fixture is compared with it.
"""
from __future__ import annotations

import json
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_files.json")
REF = "/root/reference"

# a file name, ':' and a line or range, then optionally more ranges after commas
# (", 143-171"); a continuation must end at a non-word boundary so that
# "3d_wave.jl:83, 3d_maxwell.jl:92" stays two citations
CITE = re.compile(r"([\w./-]*[\w-]+\.(?:jl|toml|md)):(\d+)(?:-(\d+))?((?:,\s*\d+(?:-\d+)?(?![\w.]))*)")
SKIP_DIRS = {".git", "_work", "gpurun_out", "__pycache__", "_build", ".pytest_cache", "oracle/_ref"}
# files written outside this build (the survey of the reference, reviews, the baseline)
SKIP_FILES = {"SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md"}
SUFFIXES = (".py", ".c", ".h", ".hip", ".hpp", ".md", ".jl")


def _counts() -> dict:
    with open(GOLDEN) as fh:
        return json.load(fh)


def _resolve(name: str, counts: dict) -> str | None:
    name = name.lstrip("./")
    if name.startswith("root/reference/"):
        name = name[len("root/reference/"):]
    if name in counts:
        return name
    cands = [p for p in counts if p.endswith("/" + name)]
    if not cands:
        cands = [p for p in counts if os.path.basename(p) == os.path.basename(name)]
    return cands[0] if len(cands) == 1 else None


def _citations():
    counts = _counts()
    for dp, dn, fn in os.walk(ROOT):
        rel_dir = os.path.relpath(dp, ROOT)
        dn[:] = [d for d in dn if d not in SKIP_DIRS and os.path.join(rel_dir, d) not in SKIP_DIRS]
        for f in fn:
            if f in SKIP_FILES or not f.endswith(SUFFIXES):
                continue
            path = os.path.join(dp, f)
            with open(path, errors="replace") as fh:
                for ln, line in enumerate(fh, 1):
                    for m in CITE.finditer(line):
                        target = _resolve(m.group(1), counts)
                        if target is None:
                            continue
                        ranges = [(int(m.group(2)), int(m.group(3) or m.group(2)))]
                        for a, b in re.findall(r"(\d+)(?:-(\d+))?", m.group(4) or ""):
                            ranges.append((int(a), int(b or a)))
                        yield os.path.relpath(path, ROOT), ln, m.group(0), target, ranges


def test_citations_inside_cited_files():
    counts = _counts()
    bad, n = [], 0
    for path, ln, text, target, ranges in _citations():
        n += 1
        for a, b in ranges:
            if not (1 <= a <= b <= counts[target]):
                bad.append(f"{path}:{ln}: {text} ({target} has {counts[target]} lines)")
    assert n > 100, f"only {n} citations found: the scan is broken"
    assert not bad, "\n".join(bad)


def test_checker_catches_a_range_past_the_end(tmp_path):
    """the scan itself: a planted citation past the end of a file is found"""
    counts = _counts()
    # assembled at run time so that the scan of this file does not report it
    line = "see {}:180-237 and {}:83, {}:236".format("examples/3d_maxwell.jl", "3d_wave.jl", "3d_maxwell.jl")
    found = []
    for m in CITE.finditer(line):
        t = _resolve(m.group(1), counts)
        found.append((t, int(m.group(2)), int(m.group(3) or m.group(2))))
    assert found == [("examples/3d_maxwell.jl", 180, 237), ("examples/3d_wave.jl", 83, 83),
                     ("examples/3d_maxwell.jl", 236, 236)]
    assert all(b > counts[t] for t, a, b in (found[0], found[2]))


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not present")
def test_fixture_matches_reference_checkout():
    for rel, n in _counts().items():
        with open(os.path.join(REF, rel), "rb") as fh:
            b = fh.read()
        assert b.count(b"\n") + (1 if b and not b.endswith(b"\n") else 0) == n, rel
