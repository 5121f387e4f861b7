"""Writes tests/golden/blocks.json from the reference's blocked-pocket data
files (read as data, nothing is executed):

  structures/zeolites/block/DDR.block
  structures/zeolites/block/MFI_Olson.block

Each file holds the number of pockets, then one line per pocket with its
fractional centre (a, b, c) and radius [A] (format read by
src/framework.c:6200-6212).  Run where the reference is at /root/reference:
    python tests/golden/make_blocks.py
"""
import json
import os

REF = "/root/reference/structures/zeolites/block"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_block(path):
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    n = int(lines[0][0])
    rows = [[float(v) for v in ln[:4]] for ln in lines[1:1 + n]]
    assert len(rows) == n
    return rows


def main():
    out = {}
    for name in ("DDR", "MFI_Olson"):
        rows = read_block(os.path.join(REF, name + ".block"))
        out[name] = {"source": f"structures/zeolites/block/{name}.block",
                     "fractional_centres": [r[:3] for r in rows],
                     "radii": [r[3] for r in rows]}
    with open(os.path.join(HERE, "blocks.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
