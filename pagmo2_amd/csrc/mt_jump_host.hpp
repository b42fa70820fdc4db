// mt_jump_host.hpp -- jump-ahead of std::mt19937 (host part): the polynomials
// that move a generator's state forward by J words.
//
// Method (Haramoto, Matsumoto, Nishimura, Panneton, L'Ecuyer, "Efficient jump
// ahead for F2-linear random number generators", INFORMS J. Computing 20,
// 2008).  With T the one-word transition of the state S_j = (x[j], ...,
// x[j+623]) of the recurrence in mt_dev.hpp and phi the characteristic
// polynomial of T, T^J = p(T) for p = z^J mod phi, so that
//   S_J = p(T) S_0,   word k of S_J = XOR over the set coefficients i of p of x[i+k].
// T never reads the 31 low bits of x[j], so the identity holds on every bit
// that a later word depends on.  phi has degree 19937; it is obtained here by
// the Berlekamp-Massey algorithm from 2 x 19937 + 64 output bits of the
// generator itself and checked to have that degree.
//
// Polynomials over GF(2) are bit vectors (bit i = coefficient of z^i), 312
// 64-bit words below degree 19937.  Products use carry-less multiplication
// (PCLMULQDQ when the CPU has it, a shift-and-add loop otherwise).  A product
// of degree < 2 x 19968 is reduced through a table of z^(64 w) mod phi for the
// words w >= 312 it has: each such word c contributes c * (z^(64 w) mod phi),
// so the reduction is 312 x 312 independent word products; the bits 19937 ..
// 19967 left over are reduced one by one.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

namespace pgh
{
namespace mtj
{

constexpr uint32_t kDeg = 19937;
constexpr uint32_t kWords = (kDeg + 63u) / 64u; // 312

using poly = std::vector<uint64_t>;

inline bool bit(const poly &p, uint64_t i)
{
    return i / 64u < p.size() && ((p[i >> 6] >> (i & 63u)) & 1u);
}

// 64 bits of p from bit `at` (zeros past the end).
inline uint64_t bits64(const poly &p, uint64_t at)
{
    const uint64_t w = at >> 6, s = at & 63u;
    const uint64_t lo = w < p.size() ? p[w] : 0u;
    if (s == 0u) return lo;
    const uint64_t hi = w + 1u < p.size() ? p[w + 1u] : 0u;
    return (lo >> s) | (hi << (64u - s));
}

// ---- carry-less 64 x 64 -> 128 -------------------------------------------------
inline void clmul_soft(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi)
{
    lo = hi = 0u;
    for (uint32_t i = 0; i < 64u; ++i) {
        if ((b >> i) & 1u) {
            lo ^= a << i;
            if (i) hi ^= a >> (64u - i);
        }
    }
}

#if defined(__x86_64__)
__attribute__((target("pclmul,sse2"))) inline void clmul_hw(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi)
{
    const __m128i r = _mm_clmulepi64_si128(_mm_cvtsi64_si128(static_cast<long long>(a)),
                                           _mm_cvtsi64_si128(static_cast<long long>(b)), 0x00);
    lo = static_cast<uint64_t>(_mm_cvtsi128_si64(r));
    hi = static_cast<uint64_t>(_mm_cvtsi128_si64(_mm_unpackhi_epi64(r, r)));
}

// acc[0 .. n] ^= a[0 .. n) * c (a word vector times one word)
__attribute__((target("pclmul,sse2"))) inline void mul_word_acc_hw(const uint64_t *a, uint64_t n, uint64_t c,
                                                                   uint64_t *acc)
{
    const __m128i cv = _mm_cvtsi64_si128(static_cast<long long>(c));
    __m128i carry = _mm_setzero_si128();
    for (uint64_t i = 0; i < n; ++i) {
        const __m128i r = _mm_clmulepi64_si128(_mm_cvtsi64_si128(static_cast<long long>(a[i])), cv, 0x00);
        // r = lo | hi << 64: lo goes to word i, hi to word i + 1
        const __m128i cur = _mm_xor_si128(r, carry);
        acc[i] ^= static_cast<uint64_t>(_mm_cvtsi128_si64(cur));
        carry = _mm_srli_si128(r, 8);
    }
    acc[n] ^= static_cast<uint64_t>(_mm_cvtsi128_si64(carry));
}

inline bool have_pclmul()
{
    static const bool yes = __builtin_cpu_supports("pclmul");
    return yes;
}
#endif

inline void mul_word_acc(const uint64_t *a, uint64_t n, uint64_t c, uint64_t *acc)
{
#if defined(__x86_64__)
    if (have_pclmul()) {
        mul_word_acc_hw(a, n, c, acc);
        return;
    }
#endif
    uint64_t carry = 0u;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t lo, hi;
        clmul_soft(a[i], c, lo, hi);
        acc[i] ^= lo ^ carry;
        carry = hi;
    }
    acc[n] ^= carry;
}

// ---- Berlekamp-Massey over GF(2) ------------------------------------------------
// Connection polynomial C (c_0 = 1, degree L) of the shortest linear
// recurrence s[n] = sum_{i=1..L} c_i s[n-i] of the bit sequence s.
inline void berlekamp_massey(const std::vector<uint8_t> &s, poly &C, uint32_t &L)
{
    const uint64_t N = s.size();
    const uint64_t words = (N + 64u) / 64u + 1u;
    // rs: s reversed, so that the window s[n-1], s[n-2], ... is ascending
    poly rs(words, 0u);
    for (uint64_t k = 0; k < N; ++k) {
        if (s[N - 1u - k]) rs[k >> 6] |= uint64_t(1) << (k & 63u);
    }
    C.assign(words, 0u);
    poly B(words, 0u), Tm;
    C[0] = B[0] = 1u;
    L = 0u;
    uint64_t m = 1u;
    for (uint64_t n = 0; n < N; ++n) {
        // d = s[n] + sum_{i=1..L} c_i s[n-i], with s[n-i] = rs[N-1-n+i]
        uint64_t acc = 0u;
        const uint64_t o = N - 1u - n;
        for (uint64_t w = 0; w * 64u <= L; ++w) {
            uint64_t c = C[w];
            if (w == 0u) c &= ~uint64_t(1); // i >= 1
            const uint64_t top = L - w * 64u; // keep i <= L
            if (top < 63u) c &= (uint64_t(2) << top) - 1u;
            if (c) acc ^= c & bits64(rs, o + w * 64u);
        }
        const unsigned d = (static_cast<unsigned>(__builtin_popcountll(acc)) & 1u) ^ s[n];
        if (!d) {
            ++m;
            continue;
        }
        Tm = C;
        // C ^= B z^m
        const uint64_t ws = m >> 6, bs = m & 63u;
        for (uint64_t w = 0; w + ws < words; ++w) {
            uint64_t v = B[w] << bs;
            if (bs && w > 0u) v |= B[w - 1u] >> (64u - bs);
            C[w + ws] ^= v;
        }
        if (2u * L <= n) {
            L = static_cast<uint32_t>(n + 1u - L);
            B = Tm;
            m = 1u;
        } else {
            ++m;
        }
    }
}

// ---- arithmetic modulo phi ------------------------------------------------------
class field
{
  public:
    bool ok = false;
    poly phi; // kWords + 1 words, degree kDeg

    field()
    {
        std::mt19937 e(5489u);
        std::vector<uint8_t> s(2u * kDeg + 64u);
        for (auto &b : s) b = static_cast<uint8_t>(e() & 1u);
        poly C;
        uint32_t L = 0;
        berlekamp_massey(s, C, L);
        if (L != kDeg) return;
        // phi(z) = z^L C(1/z): coefficient i of phi is c_{L-i}
        phi.assign(kWords + 1u, 0u);
        for (uint32_t i = 0; i <= L; ++i) {
            if (bit(C, L - i)) phi[i >> 6] |= uint64_t(1) << (i & 63u);
        }
        // tab[w - kWords] = z^(64 w) mod phi, w = kWords .. 2 kWords - 1
        poly t = shift_reduce_bits(unit_at(64u * kWords));
        tab.push_back(t);
        for (uint32_t w = kWords + 1u; w < 2u * kWords; ++w) {
            t = times_z64(t);
            tab.push_back(t);
        }
        ok = true;
    }

    // a * b mod phi (a, b reduced)
    poly mul(const poly &a, const poly &b) const
    {
        poly prod(2u * kWords + 1u, 0u);
        for (uint32_t j = 0; j < kWords; ++j) {
            if (b[j]) mul_word_acc(a.data(), kWords, b[j], prod.data() + j);
        }
        return reduce(prod);
    }

    poly sqr(const poly &a) const
    {
        return mul(a, a);
    }

    // z^J mod phi: left to right over J's bits, squaring and shifting by z
    poly zpow(uint64_t J) const
    {
        poly r(kWords, 0u);
        if (J < kDeg) {
            r[J >> 6] = uint64_t(1) << (J & 63u);
            return r;
        }
        r[0] = 1u;
        for (int b = 63; b >= 0; --b) {
            r = sqr(r);
            if ((J >> b) & 1u) r = times_z(r);
        }
        return r;
    }

  private:
    std::vector<poly> tab;

    static poly unit_at(uint64_t i)
    {
        poly r(i / 64u + 1u, 0u);
        r[i / 64u] = uint64_t(1) << (i & 63u);
        return r;
    }

    // XOR of phi << s into r at bit position `at` (phi's leading bit lands on at + kDeg)
    void add_phi_at(poly &r, uint64_t at) const
    {
        const uint64_t ws = at >> 6, bs = at & 63u;
        for (uint64_t w = 0; w <= kWords; ++w) {
            const uint64_t lo = phi[w] << bs;
            const uint64_t hi = bs ? (phi[w] >> (64u - bs)) : 0u;
            if (ws + w < r.size()) r[ws + w] ^= lo;
            if (bs && ws + w + 1u < r.size()) r[ws + w + 1u] ^= hi;
        }
    }

    // bit-by-bit reduction (used for the table and the last 31 bits)
    poly shift_reduce_bits(poly r) const
    {
        for (uint64_t k = static_cast<uint64_t>(r.size()) * 64u; k-- > kDeg;) {
            if (bit(r, k)) add_phi_at(r, k - kDeg);
        }
        r.resize(kWords);
        return r;
    }

    poly times_z64(const poly &a) const
    {
        poly r(kWords + 1u, 0u);
        for (uint32_t w = 0; w < kWords; ++w) r[w + 1u] = a[w];
        return shift_reduce_bits(r);
    }

    poly times_z(const poly &a) const
    {
        poly r(kWords + 1u, 0u);
        uint64_t carry = 0u;
        for (uint32_t w = 0; w < kWords; ++w) {
            r[w] = (a[w] << 1) | carry;
            carry = a[w] >> 63;
        }
        r[kWords] = carry;
        return shift_reduce_bits(r);
    }

    // prod has up to 2 kWords + 1 words: words w >= kWords fold through the
    // table (word 2 kWords is zero for products of reduced operands: their
    // degree is below 2 x 19937 - 1 < 64 x 2 kWords)
    poly reduce(const poly &prod) const
    {
        poly r(kWords + 1u, 0u);
        for (uint32_t w = 0; w < kWords; ++w) r[w] = prod[w];
        for (uint32_t w = kWords; w < 2u * kWords; ++w) {
            if (prod[w]) mul_word_acc(tab[w - kWords].data(), kWords, prod[w], r.data());
        }
        return shift_reduce_bits(r);
    }
};

inline const field &the_field()
{
    static const field f;
    return f;
}

struct poly_cache {
    std::mutex mu;
    std::map<uint64_t, poly> map; // references stay valid as entries are added
};

inline poly_cache &the_cache()
{
    static poly_cache c;
    return c;
}

// z^J mod phi, memoised (the same few distances recur).
inline const poly &jump_poly(uint64_t J)
{
    poly_cache &c = the_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.map.find(J);
    if (it != c.map.end()) return it->second;
    return c.map.emplace(J, the_field().zpow(J)).first->second;
}

// Computes the polynomials of the distances not yet memoised on up to
// `threads` host threads (each z^J is independent), so that the jump_poly
// calls that follow find them.
inline void prefetch_jump_polys(const std::vector<uint64_t> &Js, unsigned threads)
{
    const field &f = the_field();
    poly_cache &c = the_cache();
    std::vector<uint64_t> todo;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        for (uint64_t J : Js) {
            if (!c.map.count(J) && std::find(todo.begin(), todo.end(), J) == todo.end()) todo.push_back(J);
        }
    }
    if (todo.size() < 2u || threads < 2u) return; // jump_poly computes them as they come
    std::vector<poly> res(todo.size());
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t i = next++; i < todo.size(); i = next++) res[i] = f.zpow(todo[i]);
    };
    std::vector<std::thread> pool;
    const size_t n = std::min<size_t>(threads, todo.size());
    try {
        pool.reserve(n);
        for (size_t t = 1; t < n; ++t) pool.emplace_back(work);
    } catch (...) {
        // fewer helper threads: the work queue is shared, so this thread and
        // those started finish it
    }
    work();
    for (auto &th : pool) th.join();
    std::lock_guard<std::mutex> lk(c.mu);
    for (size_t i = 0; i < todo.size(); ++i) c.map.emplace(todo[i], std::move(res[i]));
}

// z^(k S) mod phi for k = 1 .. count, memoised: one multiplication by z^S
// each rather than one exponentiation each.
inline void prefetch_multiples(uint64_t S, uint64_t count)
{
    if (count == 0u || S == 0u) return;
    {
        // made by an earlier call (the multiples come in order, the last one
        // is entered with the rest)
        poly_cache &c = the_cache();
        std::lock_guard<std::mutex> lk(c.mu);
        if (count == 1u ? c.map.count(S) != 0u : c.map.count(count * S) != 0u) return;
    }
    const field &f = the_field();
    const poly step = jump_poly(S);
    poly cur = step;
    std::vector<std::pair<uint64_t, poly>> res;
    for (uint64_t k = 2; k <= count; ++k) {
        cur = f.mul(cur, step);
        res.emplace_back(k * S, cur);
    }
    poly_cache &c = the_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    for (auto &kv : res) c.map.emplace(kv.first, std::move(kv.second));
}

} // namespace mtj
} // namespace pgh
