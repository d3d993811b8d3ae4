"""Copies the reference's own test inputs into tests/golden/ as fixtures.

Run in a container where the reference is mounted (default /root/reference).
Copies, as data:
  * tests/astronomy_parameters-{11,12,79,86}-small[-aux].txt (parameter files)
  * the first N_STARS rows of tests/stars-{11,12,79,86}.txt (star catalogues;
    the full files are 2-4 MB each)
  * the CTest parameter vectors of tests/CMakeLists.txt:24-41 as JSON.
The GPU tests read only these fixtures, never the reference tree.
"""
import json
import re
import sys
from pathlib import Path

N_STARS = 2000
REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent
IDS = ["11", "12", "79", "86"]


def main():
    tests = REF / "tests"
    for i in IDS:
        for suffix in ("-small", "-small-aux"):
            src = tests / f"astronomy_parameters-{i}{suffix}.txt"
            if src.exists():
                (OUT / src.name).write_text(src.read_text())
        lines = (tests / f"stars-{i}.txt").read_text().splitlines()
        n = min(N_STARS, int(lines[0].split()[0]))
        rows = [ln for ln in lines[1:] if ln.strip()][:n]
        (OUT / f"stars-{i}-first{N_STARS}.txt").write_text("\n".join([str(len(rows))] + rows) + "\n")
    # CTest parameter vectors (tests/CMakeLists.txt:24-41)
    cm = (tests / "CMakeLists.txt").read_text()
    sets = {}
    for name, body in re.findall(r"set\((\w+)\s+([^)]*)\)", cm):
        toks = body.split()
        if name in ("param_set1", "test_79_params", "test_82_params", "test_86_params"):
            sets[name] = [float(t) for t in toks]
    vectors = {
        "11": sets["param_set1"], "12": sets["param_set1"],
        "79": sets["test_79_params"], "86": sets["test_86_params"],
    }
    (OUT / "ctest_parameter_vectors.json").write_text(json.dumps(vectors, indent=1) + "\n")


if __name__ == "__main__":
    main()
