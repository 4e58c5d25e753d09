#!/usr/bin/env python3
"""Writes ziggurat_tables.h: the layer edges of the Marsaglia-Tsang ziggurats that
boost::random::normal_distribution<double> (128 layers of the half normal) and
exponential_distribution<double> (256 layers) sample with (the reference draws them in
src/normal.rs:144-214).

The tables are derived here from their definition, in 80-digit decimal arithmetic, and rounded
once to double:
  f(x) = exp(-x^2/2) (normal) or exp(-x) (exponential); x_1 = R, the tail start;
  v = R f(R) + integral_R^inf f, the common area of every layer; x_0 = v / f(R);
  x_{i+1} = f^-1(f(x_i) + v / x_i) for i >= 1, and R is the root of x_L = 0 (L the layer count);
  y_0 = 0, y_i = f(x_i), y_L = 1.
Run from the repository root: python3 gpl_amd/csrc/gen_ziggurat.py"""
import os
from decimal import Decimal as D, getcontext

getcontext().prec = 80
PI = D("3.14159265358979323846264338327950288419716939937510582097494459230781640628620899862803")


def erfc(z):
    # erf(z) = 2/sqrt(pi) exp(-z^2) sum_n 2^n z^(2n+1) / (1*3*...*(2n+1)): positive terms only
    t = z
    s = z
    n = 0
    eps = D(10) ** -(getcontext().prec + 5)
    while t > eps * s:
        n += 1
        t = t * 2 * z * z / (2 * n + 1)
        s += t
    return 1 - 2 / PI.sqrt() * (-z * z).exp() * s


NORMAL = dict(layers=128, f=lambda x: (-x * x / 2).exp(), finv=lambda y: (-2 * y.ln()).sqrt(),
              tail=lambda r: (PI / 2).sqrt() * erfc(r / D(2).sqrt()), lo=D(3), hi=D(4))
EXPO = dict(layers=256, f=lambda x: (-x).exp(), finv=lambda y: -y.ln(),
            tail=lambda r: (-r).exp(), lo=D(7), hi=D(8))


def edges(z, r):
    """x_0..x_{L-1} for tail start r, or None when the layers reach the mode before the last one
    (the layer area v too large)."""
    f, finv, L = z["f"], z["finv"], z["layers"]
    v = r * f(r) + z["tail"](r)
    x = [v / f(r), r]
    for i in range(1, L - 1):
        y = f(x[i]) + v / x[i]
        if y >= 1:
            return None, i
        x.append(finv(y))
    return x, L


def solve(z):
    lo, hi = z["lo"], z["hi"]
    for _ in range(280):  # 2^-280: far past 80 digits
        mid = (lo + hi) / 2
        x, _ = edges(z, mid)
        if x is None:  # the layers reach the mode early: v too large, R too small
            lo = mid
        else:
            # the top of the last layer, above x_{L-1}, must be f = 1 exactly
            yL = z["f"](x[-1]) + (mid * z["f"](mid) + z["tail"](mid)) / x[-1]
            if yL > 1:
                lo = mid
            else:
                hi = mid
    x, _ = edges(z, lo)
    x.append(D(0))
    y = [D(0)] + [z["f"](t) for t in x[1:-1]] + [D(1)]
    return x, y


def bits(d):
    import struct
    return struct.unpack("<Q", struct.pack("<d", float(d)))[0]


def table(name, vals):
    out = [f"static const uint64_t {name}[{len(vals)}] = {{"]
    row = []
    for k, v in enumerate(vals):
        row.append(f"0x{bits(v):016x}ull")
        if len(row) == 4 or k == len(vals) - 1:
            out.append("    " + ", ".join(row) + ",")
            row = []
    out.append("};")
    return "\n".join(out)


def main():
    nx, ny = solve(NORMAL)
    ex, ey = solve(EXPO)
    text = "\n".join([
        "// Generated by gen_ziggurat.py (the layer edges derived from the ziggurat definition in",
        "// 80-digit arithmetic, each rounded once to double, stored as bit patterns). Do not edit.",
        "#pragma once",
        "#include <cstdint>",
        "",
        "namespace gplz {",
        "// normal: 128 layers of exp(-x^2/2); x[1] is the tail start R = " + str(nx[1])[:22],
        table("kNormalX", nx), table("kNormalY", ny),
        "// exponential: 256 layers of exp(-x); x[1] = " + str(ex[1])[:22],
        table("kExpX", ex), table("kExpY", ey),
        "}  // namespace gplz",
        "",
    ])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ziggurat_tables.h")
    with open(path, "w") as f:
        f.write(text)
    print(f"wrote {path}: R_normal {str(nx[1])[:22]}, R_exp {str(ex[1])[:22]}")


if __name__ == "__main__":
    main()
