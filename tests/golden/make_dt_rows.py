"""Extract Doyle-Turner rows from the reference's Muslib Curves.cpp bank
(Musli console 1.3 (STEM) 2/Muslib/Curves.cpp:83-179) into dt_rows.json:
a fixture of data values, parsed independently of the table generator.

Usage: python make_dt_rows.py /path/to/reference
"""
import json
import os
import sys

ZS = [1, 6, 8, 14, 22, 38, 79, 92]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    path = os.path.join(ref, "Musli console 1.3 (STEM) 2", "Muslib", "Curves.cpp")
    rows = {}
    for line in open(path, "rb").read().decode("latin-1").splitlines():
        parts = [p.strip() for p in line.strip().rstrip(",").split(",")]
        if len(parts) == 10 and parts[0].isdigit() and parts[1].startswith('"'):
            z = int(parts[0])
            if z in ZS:
                rows[str(z)] = [float(v) for v in parts[2:]]
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "dt_rows.json"), "w") as f:
        json.dump({"source": "Muslib/Curves.cpp bank[] (a1 b1 a2 b2 a3 b3 a4 b4)", "rows": rows}, f, indent=1)


if __name__ == "__main__":
    main()
