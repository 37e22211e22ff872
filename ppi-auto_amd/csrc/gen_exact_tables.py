#!/usr/bin/env python3
"""Writes exact_tables.h: the constants of the table-driven transcendentals in exact_math.h.

Every value is computed with mpmath at 200 bits and rounded once to the nearest
double (mpmath's float conversion rounds to nearest).  Run from this directory:

    python3 gen_exact_tables.py > exact_tables.h
"""
import mpmath as mp

mp.mp.prec = 200


def hexd(v) -> str:
    return float(v).hex()


def split(v):
    hi = mp.mpf(float(v))
    lo = mp.mpf(float(v - hi))
    return hi, lo


def main() -> None:
    out = []
    w = out.append
    w("// exact_tables.h — generated by gen_exact_tables.py (mpmath, 200 bits, rounded to nearest double).")
    w("// Constants of the table-driven transcendentals in exact_math.h.  Do not edit.")
    w("#ifndef PPI_EXACT_TABLES_H")
    w("#define PPI_EXACT_TABLES_H")
    w("")
    w("namespace ppi {")
    w("namespace dev {")
    w("")

    # exp: x = k*ln2/128 + r, exp(x) = 2^(k>>7) * 2^((k&127)/128) * exp(r).
    ln2 = mp.log(2)
    l1, l2 = split(ln2 / 128)
    w("// exp(x) = 2^(k >> 7) * 2^((k & 127)/128) * exp(r),  k = rint(x * 128/ln2),  r = x - k*(L1 + L2)")
    w(f"constexpr double kExpInvL = {hexd(128 / ln2)};  // 128/ln2")
    w(f"constexpr double kExpL1 = {hexd(l1)};  // ln2/128 rounded")
    w(f"constexpr double kExpL2 = {hexd(l2)};  // ln2/128 - kExpL1 rounded")
    tables = []  # (comment, member declaration, rows of initialiser text); the sine table goes first (below)
    tables.append(("2^(j/128)", "double exp2j[128];",
                   [", ".join(hexd(mp.power(2, mp.mpf(i) / 128)) for i in range(j, j + 4)) for j in range(0, 128, 4)]))

    # log: x = 2^k * z, z in [OFF, 2 OFF) with OFF = 0x3f330000 (0.69921875); bucket i = top 7 bits
    # of z's offset from OFF; log(z) = log1p(z*invc - 1) - log(invc).
    OFF = 0x3F330000
    w("// log(x) = k*ln2 - log(invc_i) + log1p(z*invc_i - 1), x = 2^k z, z in [0.69921875, 1.3984375);")
    w("// bucket i = bits 16..22 of (bits(x) - 0x3f330000); invc_i ~ 1 / (bucket centre), 1 exactly for the")
    w("// two buckets around z = 1 (no cancellation between the table term and the polynomial there).")
    import struct

    def f_of_bits(b):
        return mp.mpf(struct.unpack("<f", struct.pack("<I", b & 0xFFFFFFFF))[0])

    invs, logs = [], []
    for i in range(128):
        lo = f_of_bits(OFF + (i << 16))
        hi = f_of_bits(OFF + ((i + 1) << 16))
        if lo <= 1 < hi or hi == 1:
            inv = mp.mpf(1)
        else:
            inv = mp.mpf(float(2 / (lo + hi)))
        invs.append(inv)
        logs.append(mp.mpf(0) if inv == 1 else -mp.log(inv))
    lh, ll = split(ln2)
    w(f"constexpr double kLn2Hi = {hexd(lh)};")
    w(f"constexpr double kLn2Lo = {hexd(ll)};")
    tables.append(("{invc_i, -log(invc_i)}", "double log[128][2];",
                   ["{" + hexd(invs[i]) + ", " + hexd(logs[i]) + "}, {" + hexd(invs[i + 1]) + ", " + hexd(logs[i + 1])
                    + "}" for i in range(0, 128, 2)]))

    # sin/cos: |x| = k*pi/128 + r, sin(|x|) = S_k cos r + C_k sin r with S_k = sin(k pi/128), C_k = S_{k+64}.
    p1, p2 = split(mp.pi / 128)

    def sin_pi128(i):  # sin(i pi/128); exactly 0 at multiples of pi (mpmath leaves ~2^-200 there)
        return mp.mpf(0) if i % 128 == 0 else mp.sin(mp.mpf(i) * mp.pi / 128)

    w("// sin(a) = S_j cos(r) + S_{j+64} sin(r),  a = |x| = k*pi/128 + r,  j = k mod 256,  S_j = sin(j pi/128)")
    w(f"constexpr double kTrigInvP = {hexd(128 / mp.pi)};  // 128/pi")
    w(f"constexpr double kTrigP1 = {hexd(p1)};  // pi/128 rounded")
    w(f"constexpr double kTrigP2 = {hexd(p2)};  // pi/128 - kTrigP1 rounded")
    # First in the struct (LDS offset 0): both of its reads then fit the 8-bit offsets of one ds_read2_b64.
    tables.insert(0, ("sin(j pi/128), j < 384: sin and cos read entries j and j + 64 (cos: j + 64 and j + 128) "
                      "of one index j < 256", "double sinpi128[384];",
                      [", ".join(hexd(sin_pi128(i)) for i in range(j, j + 4)) for j in range(0, 384, 4)]))

    # One struct, so that kernels can stage all tables into LDS with one copy.
    w("")
    w("struct ExactTables {")
    for comment, decl, _ in tables:
        w(f"    {decl}  // {comment}")
    w("};")
    w("__device__ const ExactTables kExactTables = {")
    for _, _, rows in tables:
        w("    {")
        for row in rows:
            w("        " + row + ",")
        w("    },")
    w("};")
    w("")

    # Taylor coefficients.
    fac = mp.factorial
    w("// Taylor coefficients: exp(r) - 1 to r^4; log1p(r) to r^7; sin(r) to r^5; cos(r) - 1 to r^6.")
    for n in range(2, 5):
        w(f"constexpr double kE{n} = {hexd(1 / fac(n))};")
    for n in range(2, 8):
        w(f"constexpr double kL{n} = {hexd(mp.mpf((-1) ** (n + 1)) / n)};")
    for n in (3, 5):
        w(f"constexpr double kS{n} = {hexd(mp.mpf((-1) ** ((n - 1) // 2)) / fac(n))};")
    for n in (2, 4, 6):
        w(f"constexpr double kC{n} = {hexd(mp.mpf((-1) ** (n // 2)) / fac(n))};")
    w("")
    w("}  // namespace dev")
    w("}  // namespace ppi")
    w("")
    w("#endif")
    print("\n".join(out))


if __name__ == "__main__":
    main()
