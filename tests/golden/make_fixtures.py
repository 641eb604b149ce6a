"""Writes the golden fixtures of tests/golden/ from data held in the reference.

Run where the reference is checked out (read-only) at /root/reference; the
tests only read its outputs. Outputs are data only:

* horner_sigma2.json / horner_sigma125.json: the Horner coefficient tables of
  include/cufinufft/contrib/ker_horner_allw_loop.inc and
  ker_lowupsampfac_horner_allw_loop.inc for widths 2, 5, 7, 8, 12 and 16
  (rows c0..cd, one value per interval).
* next235_even.json: the next235(n, 2) table for n = 0..99 of
  test/testutils.cpp:39-44.
"""

import json
import re
from pathlib import Path

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent


def horner_table(inc, widths):
    txt = (REF / "include/cufinufft/contrib" / inc).read_text()
    out = {}
    for w in widths:
        blk = txt.split(f"if constexpr (w=={w})")[1].split("for (int i")[0]
        rows = re.findall(r"c\d+\[\] = \{([^}]*)\}", blk)
        out[str(w)] = [[float(v) for v in r.split(",")] for r in rows]
    return out


def main():
    widths = [2, 5, 7, 8, 12, 16]
    (OUT / "horner_sigma2.json").write_text(
        json.dumps(horner_table("ker_horner_allw_loop.inc", widths), indent=1))
    (OUT / "horner_sigma125.json").write_text(
        json.dumps(horner_table("ker_lowupsampfac_horner_allw_loop.inc", widths), indent=1))
    src = (REF / "test/testutils.cpp").read_text()
    table = src.split("next235even_true[100] = {")[1].split("}")[0]
    vals = [int(v) for v in re.findall(r"\d+", table)]
    assert len(vals) == 100
    (OUT / "next235_even.json").write_text(json.dumps(vals))


if __name__ == "__main__":
    main()
