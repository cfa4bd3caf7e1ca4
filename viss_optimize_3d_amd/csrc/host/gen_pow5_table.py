"""Writes pow5_table.h: the 128-bit powers of five of the Eisel-Lemire
decimal-to-binary conversion (D. Lemire, "Number Parsing at a Gigabyte per
Second", Software: Practice and Experience 51(8), 2021), for decimal exponents
q in [-342, 308].

Entry q holds 5^q scaled by a power of two into [2^127, 2^128): truncated for
q >= 0; for q < 0 the reciprocal floor(2^b / 5^-q) + 1 (rounded up), with b
chosen so that it has 128 bits (q >= -27) or then truncated to 128 bits.  The
two 64-bit halves are stored high word first.

    python gen_pow5_table.py > pow5_table.h
"""

Q_MIN, Q_MAX = -342, 308


def entry(q):
    if q >= 0:
        p = 5 ** q
        while p < (1 << 127):
            p <<= 1
        while p >= (1 << 128):
            p >>= 1
        return p
    p = 5 ** -q
    z = p.bit_length() if p & (p - 1) else p.bit_length() - 1  # smallest z with 2^z >= p
    b = z + 127 if q >= -27 else 2 * z + 128
    c = (1 << b) // p + 1
    while c >= (1 << 128):
        c >>= 1
    return c


def main():
    print("// Generated by gen_pow5_table.py (128-bit powers of five for the Eisel-Lemire")
    print("// decimal-to-binary conversion); do not edit.")
    print("#pragma once")
    print("#include <cstdint>")
    print("namespace vissio {")
    print(f"constexpr int kPow5QMin = {Q_MIN}, kPow5QMax = {Q_MAX};")
    print(f"alignas(16) inline constexpr uint64_t kPow5[{Q_MAX - Q_MIN + 1}][2] = {{")
    for q in range(Q_MIN, Q_MAX + 1):
        c = entry(q)
        assert (1 << 127) <= c < (1 << 128)
        print(f"    {{0x{c >> 64:016x}ull, 0x{c & ((1 << 64) - 1):016x}ull}},  // 5^{q}")
    print("};")
    print("}  // namespace vissio")


if __name__ == "__main__":
    main()
