"""Golden vectors: the Kelvin-function table of the LV-ELRA Green's function.

The reference tabulates kei_0(x) for x = 0.00, 0.01, ..., 10.00
(src/kelvin_function.f90:167-1190, "originally obtained from the IMSL special
math library function AKEIO") with five significant digits.  This script
evaluates kei_0 with mpmath (an independent arbitrary-precision library) at
those points and rounds it to five significant digits; the result is written
as decimal strings, the form in which the reference holds them.

When the reference tree is present (in the build container only), the script
also checks that every entry agrees with the reference's table, read as text.

    python tests/golden/make_kei_table.py   # writes kei_table.json
"""
import json
import os
import re

import mpmath

mpmath.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/src/kelvin_function.f90"


def table():
    xs, ks = [], []
    for k in range(1001):
        x = mpmath.mpf(k) / 100
        v = -mpmath.pi / 4 if k == 0 else mpmath.kei(0, x)
        xs.append("%.4f" % (k / 100.0))
        ks.append("%.4e" % float(v))
    return xs, ks


def check_against_reference(xs, ks):
    if not os.path.exists(REF):
        return None
    src = open(REF).read()
    body = src[src.index("tbl = ["):]
    body = body[:body.index("]")]
    nums = [float(v) for v in re.findall(r"-?\d+\.\d+(?:E[+-]\d+)?", body)]
    ref_x, ref_k = nums[0::2], nums[1::2]
    assert len(ref_x) == 1001
    bad = [i for i in range(1001) if float(xs[i]) != ref_x[i] or float(ks[i]) != ref_k[i]]
    assert not bad, f"entries differing from the reference table: {bad[:10]}"
    return len(ref_x)


if __name__ == "__main__":
    xs, ks = table()
    n = check_against_reference(xs, ks)
    out = {"source": "kei_0(x) by mpmath, rounded to 5 significant digits",
           "checked_against_reference_entries": n, "x": xs, "kei": ks}
    with open(os.path.join(HERE, "kei_table.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote kei_table.json", "(matches the reference table)" if n else "")
