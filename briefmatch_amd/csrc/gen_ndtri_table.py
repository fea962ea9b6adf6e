"""Writes bm_ndtri_q16.inc: the 4096-entry fixed-point inverse normal CDF table.

Entry k is round(ndtri((k + 0.5) / 4096) * 65536) as a signed 32-bit integer.
Both the random-search offsets and the BRIEF test pattern draw standard normals
by indexing this table with the top 12 bits of a Philox4x32-10 output word, so
the GPU and the CPU oracle see the same integers.

Run:  python briefmatch_amd/csrc/gen_ndtri_table.py
"""
import os

import numpy as np
from scipy.special import ndtri

TABLE_BITS = 12


def table():
    n = 1 << TABLE_BITS
    k = np.arange(n, dtype=np.float64)
    q = np.rint(ndtri((k + 0.5) / n) * 65536.0).astype(np.int64)
    return q


def main():
    q = table()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bm_ndtri_q16.inc")
    lines = ["// Generated by gen_ndtri_table.py: round(ndtri((k + 0.5) / 4096) * 65536)."]
    for i in range(0, len(q), 8):
        lines.append("  " + ", ".join("%d" % v for v in q[i:i + 8]) + ",")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
