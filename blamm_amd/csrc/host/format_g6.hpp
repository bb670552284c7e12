// A float as ostream's default formatting prints it: printf's %g with
// precision 6 of the value converted to double (the reference's occurrence
// writer, pwmscan.cpp:88-95, streams the score with default precision).
//
// %g with precision P = 6: X = the decimal exponent of the value rounded to
// 6 significant digits (round half to even on the exact binary value); fixed
// notation with 5 - X decimals when -4 <= X < 6, exponent notation otherwise;
// trailing zeros and a bare decimal point removed.
//
// A finite nonzero float is m * 2^e with m < 2^24.  For 1e-4 <= |v| < 1e6 the
// six digits are round(m * 10^k / 2^s) with k = 5 - X <= 10 and s = -e <= 40,
// an exact 64-bit integer computation (m * 10^10 < 2^58).  Every other value
// (zero, exponent notation, subnormal, inf, nan) goes to std::to_chars.
#pragma once

#include <charconv>
#include <cstdint>
#include <cstring>

namespace blamm {

// Writes the text at o (o must have room for 17 bytes) and returns its end.
inline char* put_g6(char* o, float f)
{
    static constexpr uint64_t kPow10[11] = {1ull,         10ull,         100ull,         1000ull,
                                            10000ull,     100000ull,     1000000ull,     10000000ull,
                                            100000000ull, 1000000000ull, 10000000000ull};
    static constexpr char kDigits2[201] =
        "00010203040506070809101112131415161718192021222324252627282930313233343536373839"
        "40414243444546474849505152535455565758596061626364656667686970717273747576777879"
        "8081828384858687888990919293949596979899";
    uint32_t u;
    std::memcpy(&u, &f, sizeof u);
    const uint32_t be = (u >> 23) & 0xFFu;
    // be in [110, 150): v = m / 2^s with 1 <= s <= 40, i.e. 2^-17 <= |v| < 2^23,
    // which covers 1e-4 <= |v| < 1e6
    if (be >= 110 && be < 150) {
        const uint64_t m = (u & 0x7FFFFFu) | 0x800000u;
        const int s = 150 - (int)be;
        // X0 = floor((be - 127) log10 2) <= floor(log10 |v|) <= X0 + 1, since
        // 2^(be-127) <= |v| < 2^(be-126); 78913 / 2^18 approximates log10 2
        // closely enough for |be - 127| <= 40
        int X = (int)((((int)be - 127) * 78913) >> 18);
        int k = 5 - X;
        if (k >= 1 && k <= 10) {
            uint64_t N = m * kPow10[k];   // < 2^24 * 10^10 < 2^58
            if ((N >> s) >= 1000000u) {    // floor(log10 |v|) = X0 + 1: one digit less
                N /= 10u;                  // exact: k >= 1
                ++X;
            }
            uint64_t I = N >> s;
            const uint64_t rem = N & ((1ull << s) - 1), half = 1ull << (s - 1);
            if (I >= 100000u && I < 1000000u) {
                if (rem > half || (rem == half && (I & 1u))) ++I;   // round half to even
                if (I == 1000000u) {
                    I = 100000u;
                    ++X;
                }
                if (X >= -4 && X < 6) {
                    // six digits from three two-digit pairs; the text is laid
                    // out in t with fixed-size copies, then trimmed by length
                    const uint32_t J = (uint32_t)I;
                    char d[16] = {};
                    std::memcpy(d, kDigits2 + 2 * (J / 10000u), 2);
                    std::memcpy(d + 2, kDigits2 + 2 * ((J / 100u) % 100u), 2);
                    std::memcpy(d + 4, kDigits2 + 2 * (J % 100u), 2);
                    int last = 5;   // last digit kept: trailing zeros of the fraction dropped
                    while (last > X && last > 0 && d[last] == '0') --last;
                    char t[24] = {};
                    int len;
                    if (X >= 0) {
                        std::memcpy(t, d, 8);                 // integer digits d[0..X]
                        t[X + 1] = '.';
                        std::memcpy(t + X + 2, d + X + 1, 8);  // fraction d[X+1..5]
                        len = last > X ? last + 2 : X + 1;
                    } else {
                        std::memcpy(t, "0.000000", 8);        // "0." and -X-1 zeros
                        std::memcpy(t + 1 - X, d, 8);
                        len = 1 - X + last + 1;
                    }
                    if (u >> 31) *o++ = '-';
                    std::memcpy(o, t, 16);
                    return o + len;
                }
            }
        }
    }
    return std::to_chars(o, o + 16, (double)f, std::chars_format::general, 6).ptr;
}

}  // namespace blamm
