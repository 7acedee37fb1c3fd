"""Writes tests/golden/cbrt_tight_cases.json: doubles t in [1, 2) whose cube
root lies within about 2^-20 units in the last place of a rounding midpoint
-- the inputs on which a fast cube root must hand the last bit to an exact
method.  Search: for random K in [2^52, 2^53), the midpoint between K 2^-52
and (K+1) 2^-52 is m = (2K+1) 2^-53, and m^3 = (2K+1)^3 2^-159; when that
is within 2^-18 of a multiple of 2^-52 (fraction of (2K+1)^3 / 2^107 near 0
or 1), the two doubles X 2^-52 around it qualify (K is kept below
2^(1/3) 2^52 so that X 2^-52 < 2 is a double).  Exact integer arithmetic
only.  Run from the repo root:  python tests/golden/make_cbrt_cases.py"""
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))


def main(n_keep=64, seed=11):
    rng = random.Random(seed)
    one107 = 1 << 107
    tol = 1 << 89  # 2^-18 of 2^107
    out = []
    while len(out) < n_keep:
        K = (1 << 52) + rng.randrange(1169017446160875)  # below (2^(1/3) - 1) 2^52
        c = (2 * K + 1) ** 3
        frac = c & (one107 - 1)
        if frac < tol or frac > one107 - tol:
            X = c >> 107
            for Xd in (X, X + 1):
                out.append(float.hex(Xd * 2.0 ** -52))
    with open(os.path.join(HERE, "cbrt_tight_cases.json"), "w") as f:
        json.dump({"inputs_hex": out}, f, indent=0)
    print(len(out), "inputs")


if __name__ == "__main__":
    main()
