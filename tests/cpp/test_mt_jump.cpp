// CPU check of the jump-ahead polynomials of pagmo2_amd/csrc/mt_jump_host.hpp
// against std::mt19937::discard (the standard library's own generator, the
// reference's detail::random_engine_type, include/pagmo/rng.hpp:43).
//
// The jump is applied the way the device applies it: from the word sequence x
// of the recurrence (x[0..623] the seeded state), the state after C words
// handed out is x[C .. C+623]; after a jump of J words it is
//   y[k] = XOR over the set coefficients i of z^J mod phi of x[C+i+k],
// and the generator continues from y.  Every output after the jump must equal
// std::mt19937's after discard(C + J).  Also checked: z^a z^b = z^(a+b) and
// z^J for J below the degree.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "mt_jump_host.hpp"

using pgh::mtj::poly;

static uint32_t temper(uint32_t y)
{
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

static uint32_t next_word(uint32_t a, uint32_t b, uint32_t c)
{
    const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return c ^ (y >> 1) ^ ((b & 1u) ? 0x9908b0dfu : 0u);
}

// x[0 .. n) of the recurrence from a 624-word start
static std::vector<uint32_t> extend(std::vector<uint32_t> x, size_t n)
{
    x.reserve(n);
    while (x.size() < n) {
        const size_t k = x.size() - 624u;
        x.push_back(next_word(x[k], x[k + 1u], x[k + 397u]));
    }
    return x;
}

static int failures = 0;

static void check_jump(uint32_t seed, uint64_t C, uint64_t J)
{
    std::vector<uint32_t> x(624);
    x[0] = seed;
    for (uint32_t i = 1; i < 624u; ++i) x[i] = 1812433253u * (x[i - 1u] ^ (x[i - 1u] >> 30)) + i;
    x = extend(x, C + 19937u + 624u);
    const poly &p = pgh::mtj::jump_poly(J);
    std::vector<uint32_t> y(624, 0u);
    for (uint64_t i = 0; i < 19937u; ++i) {
        if (!pgh::mtj::bit(p, i)) continue;
        for (uint32_t k = 0; k < 624u; ++k) y[k] ^= x[C + i + k];
    }
    // outputs after the jump: temper(y-sequence word 624 + j)
    const std::vector<uint32_t> z = extend(y, 624u + 2000u);
    std::mt19937 e(seed);
    e.discard(C + J);
    for (uint32_t j = 0; j < 2000u; ++j) {
        const uint32_t want = e();
        if (temper(z[624u + j]) != want) {
            std::printf("FAIL seed %u C %llu J %llu output %u\n", seed, (unsigned long long)C, (unsigned long long)J,
                        j);
            ++failures;
            return;
        }
    }
}

int main()
{
    const auto &f = pgh::mtj::the_field();
    if (!f.ok) {
        std::printf("FAIL characteristic polynomial not of degree 19937\n");
        return 1;
    }
    // below the degree the polynomial is z^J itself
    for (uint64_t J : {0ull, 1ull, 63ull, 64ull, 19936ull}) {
        const poly &p = pgh::mtj::jump_poly(J);
        for (uint64_t i = 0; i < 19937u; ++i) {
            if (pgh::mtj::bit(p, i) != (i == J)) {
                std::printf("FAIL z^%llu\n", (unsigned long long)J);
                ++failures;
                break;
            }
        }
    }
    // z^a z^b = z^(a+b)
    const uint64_t a = 1234567u, b = 987654321u;
    if (f.mul(pgh::mtj::jump_poly(a), pgh::mtj::jump_poly(b)) != pgh::mtj::jump_poly(a + b)) {
        std::printf("FAIL z^a z^b != z^(a+b)\n");
        ++failures;
    }
    // distances std::mt19937::discard reaches in reasonable time (it steps
    // one word at a time); larger ones follow from the products above
    const uint64_t Js[] = {1u, 2u, 623u, 624u, 625u, 19936u, 19937u, 19938u, 100000u, 1u << 20, 255u << 17,
                           (1u << 26) + 12345u};
    for (uint64_t J : Js) {
        check_jump(5489u, 0u, J);
        check_jump(42u, 1000u, J);
    }
    check_jump(7u, 227u * 3u + 5u, 131072u);
    // z^(2^40) = (z^(2^39))^2 and z^(2^40 + 2^39) = z^(2^40) z^(2^39)
    const poly &p39 = pgh::mtj::jump_poly(1ull << 39), &p40 = pgh::mtj::jump_poly(1ull << 40);
    if (f.sqr(p39) != p40 || f.mul(p40, p39) != pgh::mtj::jump_poly((1ull << 40) + (1ull << 39))) {
        std::printf("FAIL large powers\n");
        ++failures;
    }
    // distances computed on several threads equal those computed one by one
    const std::vector<uint64_t> Jp = {(1ull << 33) + 7u, 3ull << 30, 123456789012ull, 3ull << 30, 98765432109ull};
    pgh::mtj::prefetch_jump_polys(Jp, 4u);
    for (uint64_t J : Jp) {
        if (pgh::mtj::jump_poly(J) != f.zpow(J)) {
            std::printf("FAIL prefetched z^%llu\n", (unsigned long long)J);
            ++failures;
        }
    }
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
