"""Writes the constants of elem_cr.h: atan(k/32) for k = 0..32, the arctangent
series coefficients (-1)^n / (2n + 1) for n = 1..8, pi/2 and pi, each as a
pair of doubles (the value rounded to nearest, then the rounded remainder),
evaluated in 300-bit arithmetic.

    python3 glisser_amd/csrc/gen_elem_cr_table.py > /tmp/table.txt
"""
import mpmath as mp

mp.mp.prec = 300


def pair(v):
    hi = float(v)
    lo = float(v - mp.mpf(hi))
    return f"{{{hi.hex()}, {lo.hex()}}}"


def main():
    print("constexpr Pair kAtanTable[33] = {")
    for k in range(33):
        print(f"    {pair(mp.atan(mp.mpf(k) / 32))},  // atan({k}/32)")
    print("};")
    print("constexpr Pair kAtanSeries[9] = {")
    print("    {0x1p+0, 0x0p+0},")
    for n in range(1, 9):
        print(f"    {pair(mp.mpf((-1) ** n) / (2 * n + 1))},  // (-1)^{n} / {2 * n + 1}")
    print("};")
    print(f"constexpr Pair kHalfPi = {pair(mp.pi / 2)};")
    print(f"constexpr Pair kPiPair = {pair(mp.pi)};")


if __name__ == "__main__":
    main()
