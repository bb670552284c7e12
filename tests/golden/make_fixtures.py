#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the oracle.

  example/                          the reference's example inputs (data only)
  example_rc_pt1e-4.occurrences.txt `scan -rc -pt 0.0001` of the example, sorted
  example_rc_pt1e-4.PWMthresholds.txt

Run from the repository root: python tests/golden/make_fixtures.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import oracle  # noqa: E402


def main():
    oracle.build()
    with tempfile.TemporaryDirectory() as d:
        shutil.copytree(os.path.join(HERE, "example"), d, dirs_exist_ok=True)
        for args in (["dict", "sequences.mf"], ["hist", "motifs.jaspar", "sequences.mf"],
                     ["scan", "-rc", "-pt", "0.0001", "motifs.jaspar", "sequences.mf"]):
            oracle.run_cli(args, d)
        lines = sorted(open(os.path.join(d, "occurrences.txt")).readlines())
        with open(os.path.join(HERE, "example_rc_pt1e-4.occurrences.txt"), "w") as f:
            f.writelines(lines)
        shutil.copy(os.path.join(d, "PWMthresholds.txt"), os.path.join(HERE, "example_rc_pt1e-4.PWMthresholds.txt"))
    print("fixtures written")


if __name__ == "__main__":
    main()
