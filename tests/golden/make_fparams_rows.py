"""Write tests/golden/fparams_rows.json: six rows of each fparams table, read
from the reference's data (FPARAMS.DAT) and from slicelib.cpp's hard-coded
table, to pin the tables shipped in multislice_amd/csrc/fparams_tables.inc.

    python tests/golden/make_fparams_rows.py [/path/to/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ZS = (1, 8, 14, 38, 79, 103)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    rows = {"lmultislice": {}, "kirkland": {}}
    txt = open(os.path.join(ref, "ImageSimulation/parserFPARAMS/FPARAMS.DAT"), "rb").read().decode("latin-1")
    for z in ZS:
        m = re.search(r"Z=\s*%d,\s*chisq=[^\n]*\n((?:\s*[-+0-9.eE]+){12})" % z, txt)
        rows["lmultislice"][z] = [float(v) for v in m.group(1).split()]
    src = open(os.path.join(ref, "ImageSimulation/multislice/slicelib.cpp"), "rb").read().decode("latin-1")
    for z in ZS:
        rows["kirkland"][z] = [float(re.search(r"fparams\[%d\]\[%d\]\s*=\s*([-+0-9.eE]+)" % (z, i), src).group(1))
                               for i in range(12)]
    with open(os.path.join(HERE, "fparams_rows.json"), "w") as f:
        json.dump({"source": "ImageSimulation/parserFPARAMS/FPARAMS.DAT; "
                             "ImageSimulation/multislice/slicelib.cpp:903-2138", "rows": rows}, f, indent=0)
    print("wrote fparams_rows.json")


if __name__ == "__main__":
    main()
