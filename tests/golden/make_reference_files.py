"""Writes reference_files.json: the line count of every text file of the reference
checkout (BernsteinBasis.jl), so that citations of the form ``file.jl:a-b`` in this
repository can be checked without the checkout being present.

Run in a container that holds the reference at /root/reference:
    python tests/golden/make_reference_files.py
Only line counts are recorded; no text of the reference is copied.
"""
from __future__ import annotations

import json
import os

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_files.json")


def line_count(path: str) -> int:
    with open(path, "rb") as fh:
        b = fh.read()
    return b.count(b"\n") + (1 if b and not b.endswith(b"\n") else 0)


def main() -> None:
    counts = {}
    for dirpath, dirnames, filenames in os.walk(REF):
        dirnames[:] = [d for d in dirnames if d != ".git"]
        for name in filenames:
            if name.endswith((".jl", ".md", ".toml")):
                full = os.path.join(dirpath, name)
                counts[os.path.relpath(full, REF)] = line_count(full)
    with open(OUT, "w") as fh:
        json.dump(dict(sorted(counts.items())), fh, indent=1)
        fh.write("\n")
    print(f"wrote {len(counts)} entries to {OUT}")


if __name__ == "__main__":
    main()
