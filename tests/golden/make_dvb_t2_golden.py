"""Writes dvb_t2_tables.json: digests of the reference's twelve DVB-T2 address
tables (crates/gf2-coding/src/ldpc/dvb_t2/dvb_t2_matrices.rs) and code
parameters (dvb_t2/params.rs:59-73), the fixture that pins the library's
compiled-in tables.  Reads the reference here only; the fixture is committed.

    python tests/golden/make_dvb_t2_golden.py [/root/reference]

Per table: n, k, q, row count, entry count, a histogram of row lengths, the
first and last rows, and the SHA-256 of (row lengths, then addresses) as
little-endian u32.  The normal rate-1/2 table is also read from the
reference's plain-text copy (dvb_t2/table_a_1_R1_2_N64800.txt) and must agree.
"""
import hashlib
import json
import os
import re
import struct
import sys

RATES = {"1_2": "Rate1_2", "3_5": "Rate3_5", "2_3": "Rate2_3", "3_4": "Rate3_4", "4_5": "Rate4_5",
         "5_6": "Rate5_6"}


def tables_from_module(text):
    out, cur = {}, None
    for line in text.splitlines():
        m = re.match(r"\s*pub const (SHORT|NORMAL)_RATE_(\d_\d)_TABLE", line)
        if m:
            cur = (m.group(1).lower(), m.group(2))
            out[cur] = []
            continue
        if cur and line.strip() == "];":
            cur = None
            continue
        if cur:
            m = re.match(r"\s*&\[([\d,\s]*)\],?\s*$", line)
            if m:
                out[cur].append([int(v) for v in m.group(1).replace(",", " ").split()])
    return out


def params(text):
    out = {}
    for frame, rate, n, k, q in re.findall(r"\(FrameSize::(Normal|Short), CodeRate::(Rate\d_\d)\) => \((\d+), (\d+), (\d+)\)",
                                           text):
        out[(frame.lower(), rate)] = (int(n), int(k), int(q))
    return out


def digest(rows):
    b = b"".join(struct.pack("<I", len(r)) for r in rows) + b"".join(struct.pack("<I", a) for r in rows for a in r)
    return hashlib.sha256(b).hexdigest()


def main(argv):
    ref = argv[1] if len(argv) > 1 else "/root/reference"
    d = os.path.join(ref, "crates/gf2-coding/src/ldpc/dvb_t2")
    tables = tables_from_module(open(os.path.join(d, "dvb_t2_matrices.rs")).read())
    prm = params(open(os.path.join(d, "params.rs")).read())
    txt = [[int(v) for v in l.split()] for l in open(os.path.join(d, "table_a_1_R1_2_N64800.txt")) if l.strip()]
    assert txt == tables[("normal", "1_2")], "text copy of table A.1 disagrees with the module"
    assert len(tables) == 12 and len(prm) == 12
    fixture = {}
    for (frame, r), rows in sorted(tables.items()):
        n, k, q = prm[(frame, RATES[r])]
        hist = {}
        for row in rows:
            hist[str(len(row))] = hist.get(str(len(row)), 0) + 1
        fixture[f"{frame}_{r}"] = {"frame": 0 if frame == "short" else 1, "rate": list(RATES).index(r), "n": n,
                                   "k": k, "q": q, "rows": len(rows), "entries": sum(map(len, rows)),
                                   "row_lengths": hist, "first_row": rows[0], "last_row": rows[-1],
                                   "sha256": digest(rows)}
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "dvb_t2_tables.json"), "w") as f:
        json.dump({"source": "crates/gf2-coding/src/ldpc/dvb_t2/dvb_t2_matrices.rs, params.rs",
                   "digest": "sha256 of LE u32 row lengths then LE u32 addresses", "tables": fixture}, f, indent=1)
    print("wrote", len(fixture), "table digests")


if __name__ == "__main__":
    main(sys.argv)
