"""Writes mc_tables_digest.json: SHA-256 digests of the reference's marching-cubes tables (rx_mc_tables.h:19-580).

The library's tables (fltk_sph_ice_amd/csrc/mc_tables.h) are written independently; the test compares their digests
with these, so the fixture holds three hashes and no table content.  Canonical form, per table:
  edgeTable      "m0,m1,...,m255"                  (decimal)
  triTable       rows joined by ";", each row its edge numbers up to the terminator joined by ","
  numVertsTable  "n0,n1,...,n255"
Run in a checkout that has the reference at /root/reference:  python tests/golden/mc_tables_digest.py
"""
import hashlib
import json
import re
from pathlib import Path

SRC = Path("/root/reference/rx_mc_tables.h")
OUT = Path(__file__).with_name("mc_tables_digest.json")


def _block(text, name):
    i = text.index(name)
    j = text.index("{", i)
    depth, k = 0, j
    while True:
        if text[k] == "{":
            depth += 1
        elif text[k] == "}":
            depth -= 1
            if depth == 0:
                return text[j + 1:k]
        k += 1


def _ints(s):
    return [int(t, 0) for t in re.findall(r"0x[0-9a-fA-F]+|\d+", s)]


def canonical():
    text = re.sub(r"//[^\n]*", "", SRC.read_text(encoding="latin-1"))
    edge = _ints(_block(text, "edgeTable[256]"))
    rows = re.findall(r"\{([^{}]*)\}", _block(text, "triTable[256][16]"))
    tri = []
    for r in rows:
        vals = []
        for tok in r.split(","):
            tok = tok.strip()
            if not tok.isdigit():
                break
            vals.append(int(tok))
        tri.append(vals)
    nv = _ints(_block(text, "numVertsTable[256]"))
    assert len(edge) == 256 and len(tri) == 256 and len(nv) == 256
    return {
        "edge": ",".join(map(str, edge)),
        "tri": ";".join(",".join(map(str, r)) for r in tri),
        "nverts": ",".join(map(str, nv)),
    }


def digests(c):
    return {k: hashlib.sha256(v.encode()).hexdigest() for k, v in c.items()}


if __name__ == "__main__":
    OUT.write_text(json.dumps(digests(canonical()), indent=1) + "\n")
    print(f"wrote {OUT}")
