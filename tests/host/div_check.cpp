/*
 * div_check.cpp — test-side checker of div_rcp (code_mpm_amd/csrc/mpm_device.h),
 * the remap's division through a reciprocal: the header's own function,
 * compiled for the host, against the IEEE quotient a / b on cases that stress
 * the rounding of a quotient:
 *   0  random a, b over div_rcp's whole range;
 *   1  divisors whose significand is all ones or nearly 1 (their reciprocals
 *      carry the largest relative rounding error);
 *   2  quotients within a few ulp of a midpoint between two doubles (a is the
 *      double nearest to mid * b, in 128-bit arithmetic);
 *   3  the same with divisors near a power of two and quotients whose
 *      significand is near 2, where RN(a RN(1/b)) is most often 2 ulp off;
 *   4  exact quotients (a = k b) and signed zeros;
 *   5  quotients at the workloads' scales: x / h, (i L) / cells,
 *      (s_j s_l) / h and the remap's N (m v) / m_i.
 * Prints one line per class and the total number of mismatches.
 *
 *   div_check N_PER_CLASS
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "mpm_device.h"

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static uint64_t rnd() {
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return g_state;
}
static double from_bits(uint64_t u) {
    double d;
    std::memcpy(&d, &u, 8);
    return d;
}
static uint64_t to_bits(double d) {
    uint64_t u;
    std::memcpy(&u, &d, 8);
    return u;
}
/* a double with a random significand and a binary exponent in [lo, hi] */
static double rand_exp(int lo, int hi, bool sign) {
    const uint64_t m = rnd() & ((1ull << 52) - 1);
    const int e = lo + (int)(rnd() % (uint64_t)(hi - lo + 1));
    uint64_t u = ((uint64_t)(e + 1023) << 52) | m;
    if (sign && (rnd() & 1)) u |= 1ull << 63;
    return from_bits(u);
}
static double uni() { return (double)(rnd() >> 11) * 0x1p-53; }

int main(int argc, char **argv) {
    const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
    long total = 0;
    for (int cls = 0; cls < 6; cls++) {
        long bad = 0;
        for (long i = 0; i < n; i++) {
            double a = 0.0, b = 1.0;
            switch (cls) {
            case 0:
                a = rand_exp(-860, 859, true);
                b = rand_exp(-100, 99, false);
                break;
            case 1: {
                const int k = (int)(rnd() % 52);
                const uint64_t low = rnd() & ((1ull << k) - 1);
                const uint64_t m = (rnd() & 1) ? ((1ull << 52) - 1 - low) : low;
                b = from_bits(((uint64_t)(1023 + (int)(rnd() % 200) - 100) << 52) | m);
                a = rand_exp(-700, 700, true);
                break;
            }
            case 2:
            case 3: {
                if (cls == 2) {
                    b = rand_exp(-90, 90, false);
                } else {
                    const int k = (int)(rnd() % 52);
                    const uint64_t low = rnd() & ((1ull << k) - 1);
                    b = from_bits(((uint64_t)(1023 + (int)(rnd() % 40) - 20) << 52) |
                                  ((rnd() & 1) ? low : ((1ull << 52) - 1 - low)));
                }
                double q = rand_exp(-300, 300, true);
                if (cls == 3) {
                    const int k = (int)(rnd() % 52);
                    const uint64_t low = rnd() & ((1ull << k) - 1);
                    q = from_bits((to_bits(q) & ~((1ull << 52) - 1)) | ((1ull << 52) - 1 - low));
                }
                const __float128 mid = ((__float128)q + (__float128)std::nextafter(q, 2.0 * q)) / 2;
                a = (double)(mid * (__float128)b);
                const uint64_t r = rnd() % 3;
                if (r == 1) a = std::nextafter(a, INFINITY);
                if (r == 2) a = std::nextafter(a, -INFINITY);
                break;
            }
            case 4: {
                b = from_bits(((uint64_t)(1023 + (int)(rnd() % 120) - 60) << 52) | ((rnd() & 0xffffull) << 36));
                const double k = (double)(int64_t)(rnd() % 2000001) - 1000000.0;
                a = k * b;
                if (rnd() % 16 == 0) a = (rnd() & 1) ? 0.0 : -0.0;
                break;
            }
            default: {
                /* grids of 8 to 2^20 cells over 1e-3 to 10 m, particle
                 * positions over the grid, node coordinates, products of
                 * shape factors, node masses of 1e-15 to 1e-3 kg */
                const int cells = 8 + (int)(rnd() % (1u << 20));
                const double L = 1.0e-3 * std::pow(1.0e4, uni());
                const double h = L / (double)cells;
                switch (rnd() % 4) {
                case 0:
                    b = h;
                    a = (uni() * 1.2 - 0.1) * L;
                    break;
                case 1:
                    b = (double)cells;
                    a = (double)((int)(rnd() % (unsigned)(cells + 9)) - 4) * L;
                    break;
                case 2:
                    b = h;
                    a = (1.0 - uni()) * (1.0 - uni());
                    break;
                default:
                    b = 1.0e-15 * std::pow(1.0e12, uni());
                    a = uni() * uni() * uni() * (b * (uni() * 8.0)) * ((uni() - 0.5) * 40.0);
                    break;
                }
                break;
            }
            }
            const double y = 1.0 / b;
            const double got = div_rcp(a, b, y), want = a / b;
            if (to_bits(got) != to_bits(want)) {
                if (bad < 5)
                    std::printf("class %d: a = %a, b = %a: div_rcp %a, a / b %a\n", cls, a, b, got, want);
                bad++;
            }
        }
        std::printf("class %d: %ld cases, %ld mismatches\n", cls, n, bad);
        total += bad;
    }
    std::printf("mismatches %ld\n", total);
    return total == 0 ? 0 : 1;
}
