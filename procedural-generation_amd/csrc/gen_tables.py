"""Regenerate the double-double constant tables of pg_common.h with mpmath.

Each table entry is (hi, lo): hi = the double nearest the value, lo = the
double nearest the remainder, computed at 300 bits.  Run from any directory;
rewrites the tables in place and checks them.
"""
import ast
import os
import re

import mpmath as mp

mp.mp.prec = 300
HEADER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pg_common.h")


def dd(x):
    hi = float(x)
    return hi, float(x - mp.mpf(hi))


def tables():
    return {
        "atan_tab[65][2]": [dd(mp.atan(mp.mpf(k) / 64)) for k in range(65)],
        "tan_tab[41][2]": [dd(mp.tan(mp.mpf(k) / 32)) for k in range(41)],
        "inv_odd[12][2]": [dd(mp.mpf(1) / (2 * n + 1)) for n in range(12)],
        "inv_fact[30][2]": [dd(mp.mpf(1) / mp.factorial(n)) for n in range(30)],
    }


def main():
    s = open(HEADER).read()
    for name, vals in tables().items():
        pat = re.compile(r"(static constexpr double " + re.escape(name) + r" = \{\n)(.*?)(\n    \};)", re.S)
        m = pat.search(s)
        body = "\n".join(f"        {{{h!r}, {l!r}}}," for h, l in vals)
        s = s[:m.start(2)] + body + s[m.end(2):]
    h, l = dd(mp.pi / 2)
    s = re.sub(r"static constexpr double half_pi\[2\] = \{[^}]*\};", f"static constexpr double half_pi[2] = {{{h!r}, {l!r}}};", s)
    open(HEADER, "w").write(s)
    for name, vals in tables().items():
        m = re.search(r"static constexpr double " + re.escape(name) + r" = \{\n(.*?)\n    \};", s, re.S)
        rows = [ast.literal_eval(r.strip().rstrip(",").replace("{", "(").replace("}", ")")) for r in m.group(1).splitlines()]
        assert rows == [tuple(v) for v in vals], name


if __name__ == "__main__":
    main()
