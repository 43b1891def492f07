"""Writes beat_tables.json: the Fourier/polynomial coefficients of every fitted cilium beat
as the reference's filament.cpp assigns them (filament.cpp:684-960), read mechanically
from the assignments A?(i,j) = value / B?(i,j) = value under each #if / #elif branch, so
that the hand-typed oracle tables can be checked against them.

Run in the build container (the reference tree is not on the GPU box):
    python tests/golden/extract_beat_tables.py /root/reference/src/cilia/filament.cpp
"""
import json
import os
import re
import sys

BRANCHES = {  # branch condition -> beat names of the oracle using that table
    "FULFORD_AND_BLAKE_BEAT || FULFORD_AND_BLAKE_BEAT_NO_WALL": ["fulford_blake"],
    "VOLVOX_BEAT": ["volvox"],
    "FULFORD_AND_BLAKE_BEAT_ORIGINAL": ["fulford_blake_original"],
    "CORAL_LARVAE_BEAT": ["coral_larvae"],
    "BICILIA or BICILIA_LONGT": ["bicilia"],
}
ASSIGN = re.compile(r"\b([AB][xy])\((\d+),\s*(\d+)\)\s*=\s*([-+0-9.eE]+)\s*;")
DIRECTIVE = re.compile(r"^\s*#\s*(if|elif)\s+(.*?)\s*$")


def extract(path, first=680, last=965):
    tables, label = {}, None
    with open(path) as fh:
        for no, line in enumerate(fh, 1):
            if no < first or no > last:
                continue
            m = DIRECTIVE.match(line)
            if m:
                label = m.group(2) if m.group(2) in BRANCHES else None
                continue
            if label is None:
                continue
            for name, i, j, v in ASSIGN.findall(line):
                t = tables.setdefault(label, {}).setdefault(name, {})
                t[(int(i), int(j))] = float(v)
    out = {}
    for label, mats in tables.items():
        dense = {}
        for name, entries in mats.items():
            rows = 1 + max(i for i, _ in entries)
            cols = 1 + max(j for _, j in entries)
            dense[name] = [[entries[(i, j)] for j in range(cols)] for i in range(rows)]
        for beat in BRANCHES[label]:
            out[beat] = dense
    return out


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/cilia/filament.cpp"
    tables = extract(src)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "beat_tables.json")
    with open(dst, "w") as fh:
        json.dump(tables, fh, indent=1, sort_keys=True)
    print(f"{dst}: {sorted(tables)}")
