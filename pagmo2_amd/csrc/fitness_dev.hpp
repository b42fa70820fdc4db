// fitness_dev.hpp -- per-individual fitness/gradient device routines.
//
// Shared by the batch evaluator kernels (fitness.hip) and the archipelago
// kernels (pso.hip), so that an individual is reduced with exactly the same
// operation order whichever launch evaluates it.  Every mapping parameter
// (lanes per individual) is a function of the problem shape only.
//
// Build with -ffp-contract=off: products and sums below must round separately,
// as in the reference's x86-64 build.  Fused multiply-adds appear only where
// written explicitly.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "trig_dev.hpp"

namespace pgh
{
namespace dev
{

// include/pagmo/detail/constants.hpp:37-39
constexpr double kPi = 3.1415926535897932384626433832795028841971693993;
// rastrigin.cpp:65, ackley.cpp:65: omega = 2. * pi(), an exact doubling.
constexpr double kOmega = 2. * kPi;

// constants.hpp:41-44
constexpr double kPiHalf = 3.1415926535897932384626433832795028841971693993 / 2.;

enum : int {
    ROSENBROCK = 1,
    RASTRIGIN = 2,
    ACKLEY = 3,
    SCHWEFEL = 4,
    GRIEWANK = 5,
    LENNARD_JONES = 6,
    MO_SPHERE = 7,
    ZDT = 8,
    DTLZ = 9
};

// Box bounds of each problem (the get_bounds() of every UDP: rosenbrock.cpp:72-75,
// rastrigin.cpp:80-85, ackley.cpp:85-90, schwefel.cpp:77-82, griewank.cpp:84-89,
// lennard_jones.cpp:99-111); MO_SPHERE draws from [0, 1].
__host__ __device__ __forceinline__ void bounds_of(int kind, uint32_t j, double &lb, double &ub)
{
    switch (kind) {
        case ROSENBROCK:
            lb = -5.;
            ub = 10.;
            return;
        case RASTRIGIN:
            lb = -5.12;
            ub = 5.12;
            return;
        case ACKLEY:
            lb = -15;
            ub = 30;
            return;
        case SCHWEFEL:
            lb = -500;
            ub = 500;
            return;
        case GRIEWANK:
            lb = -600;
            ub = 600;
            return;
        case LENNARD_JONES:
            if (j != 0u && j % 3u == 0u) {
                lb = 0.0;
                ub = 6.0;
            } else {
                lb = -3.0;
                ub = 3.0;
            }
            return;
        default:
            lb = 0.;
            ub = 1.;
    }
}

// ---------------------------------------------------------------------------
// Lanes per individual
// ---------------------------------------------------------------------------

// Separable problems: an L-lane group per individual.  Wide rows use 32 or 64
// lanes so that a row streams as 512 B / 1 KiB coalesced loads; ackley takes
// 32 (two individuals per wavefront), which halves the share of its
// per-individual epilogue (two group sums, two exp, a sqrt).
__host__ __device__ constexpr int sep_lanes(int kind, uint32_t nx)
{
    return nx >= 256u ? (kind == ACKLEY ? 32 : 64) : (nx >= 32u ? 16 : 4);
}

// Lennard-Jones gradients of 22 to 128 atoms visit every pair once, in
// 16-lane groups up to 64 atoms and 32-lane groups beyond
// (lj_energy_grad_blk); the others visit each pair from both atoms (below 22
// atoms most of the blocks' slots would be padding, above 128 the blocks no
// longer fit the registers).
__host__ __device__ constexpr bool lj_grad_blocks(uint32_t atoms)
{
    return atoms >= 22u && atoms <= 128u;
}

// The energy takes the same blocks (lj_energy_blk) only where at most 1/16 of
// the block slots are padding; otherwise its cyclic schedule wastes less.
__host__ __device__ constexpr bool lj_energy_blocks(uint32_t atoms)
{
    const uint32_t R = atoms <= 64u ? 16u : 32u;
    const uint32_t slots = R * ((atoms + R - 1u) / R);
    return lj_grad_blocks(atoms) && 16u * (slots - atoms) <= slots;
}

// Lennard-Jones: lanes per individual.  At least ~8 pair evaluations per
// lane, at most 16 lanes owning atoms; from 32 atoms on, 32 lanes: two halves
// of 16 that own the same atoms and split the schedule's steps between them
// (twice the wavefronts for the same work, to hide latency when a batch fills
// the device only once) (lj_lanes_cyclic).  The pair-once blocks of 22 to 64
// atoms take 16 lanes, of 65 to 128 atoms 32 (lj_grad_blocks,
// lj_energy_blocks).
__host__ __device__ constexpr int lj_lanes_cyclic(uint32_t atoms)
{
    const uint32_t pairs = atoms * (atoms - 1u) / 2u;
    const uint32_t want = pairs / 8u;
    return atoms >= 32u ? 32 : (want >= 16u ? 16 : (want >= 8u ? 8 : (want >= 4u ? 4 : (want >= 2u ? 2 : 1))));
}

__host__ __device__ constexpr int lj_lanes(uint32_t atoms, bool grad)
{
    if (grad ? lj_grad_blocks(atoms) : lj_energy_blocks(atoms)) return atoms <= 64u ? 16 : 32;
    return lj_lanes_cyclic(atoms);
}

// Lanes of a group that own atoms (the rest repeat them for other steps).
__host__ __device__ constexpr int lj_atom_lanes(int L)
{
    return L >= 32 ? 16 : L;
}

// Lennard-Jones individuals are staged in LDS with the atoms' (x, y, z) and a
// second copy of the first lj_copy_atoms atoms after them (see lj_stage): the
// energy visits partners a + k with k <= A/2, the gradient k <= A - 1.
__host__ __device__ constexpr uint32_t lj_copy_atoms(uint32_t atoms, bool grad)
{
    return grad ? atoms - 1u : atoms / 2u;
}

// Doubles of LDS per staged individual.
__host__ __device__ constexpr uint32_t lj_stride(uint32_t atoms, bool grad)
{
    return 3u * (atoms + lj_copy_atoms(atoms, grad));
}

// Individuals per workgroup for Lennard-Jones: a workgroup uses at most 64 KiB
// of LDS.
__host__ __device__ constexpr uint32_t lj_group_count(uint32_t atoms, bool grad)
{
    const uint32_t L = static_cast<uint32_t>(lj_lanes(atoms, grad));
    const uint32_t bytes = 8u * lj_stride(atoms, grad);
    uint32_t g = 256u / L;
    while (g > 1u && g * bytes > 65536u) g >>= 1;
    return (g * bytes > 65536u) ? 0u : g;
}

// ---------------------------------------------------------------------------
// Group reductions (fixed butterfly: the order depends on L only)
// ---------------------------------------------------------------------------

template <int L>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, L);
    return v;
}

template <int L>
__device__ __forceinline__ double group_prod(double v)
{
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v *= __shfl_xor(v, o, L);
    return v;
}

template <int L>
__device__ __forceinline__ int group_max(int v)
{
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, L));
    return v;
}

// ---------------------------------------------------------------------------
// Separable problems (rastrigin, ackley, schwefel, griewank)
// ---------------------------------------------------------------------------

// Lane r of an L-group owns element pairs c = r, r+L, r+2L, ... (elements 2c
// and 2c+1) and accumulates them in increasing order; the L partial results
// are then combined by group_sum/group_prod.  ALIGNED selects 16-byte loads
// (even nx and a 16-byte aligned row) without changing that order.
// FAST (rastrigin and ackley): the cosines come from cos_2pi_small, valid for
// |x| < 2^10; the caller checks at the end, from the lane's sum of squares,
// that every element was in that range (else the individual is evaluated
// again with FAST = false, the reference's expression with an accurate
// cosine).
template <int KIND, bool FAST = false>
struct sep_acc {
    double a = 0., b = 0.;
    __device__ __forceinline__ void init()
    {
        a = 0.;
        b = (KIND == GRIEWANK) ? 1. : 0.;
    }
    __device__ __forceinline__ void add(double xi, uint32_t i)
    {
        if constexpr (KIND == RASTRIGIN) {
            if constexpr (FAST) {
                a = fma(xi, xi, a);
                a -= cos_2pi_small<true>(xi);
            } else {
                // rastrigin.cpp:67
                a += xi * xi - 10. * cos_any(kOmega * xi);
            }
        } else if constexpr (KIND == ACKLEY) {
            if constexpr (FAST) {
                a = fma(xi, xi, a);
                b += cos_2pi_small<false>(xi);
            } else {
                // ackley.cpp:69-70
                a += xi * xi;
                b += cos_any(kOmega * xi);
            }
        } else if constexpr (KIND == SCHWEFEL) {
            // schwefel.cpp:64
            a += xi * sin(sqrt(fabs(xi)));
        } else if constexpr (KIND == GRIEWANK) {
            // griewank.cpp:67-72 (two passes in the reference; each is a
            // reduction over the same elements, so they fuse)
            a += xi * xi;
            b *= cos(xi / sqrt(static_cast<double>(i) + 1.0));
        }
        (void)i;
    }
    // FAST: true when every element added had |x| < 2^10, known from the sum
    // of squares (< 2^20); rastrigin's a holds sum(x^2 - 10 cos), at least
    // sum(x^2) - 10 * elements.  NaN and inf fail the test.
    __device__ __forceinline__ bool in_fast_range(uint32_t elements) const
    {
        if constexpr (KIND == RASTRIGIN) {
            return a + 10. * static_cast<double>(elements) < 0x1p20;
        } else {
            return a < 0x1p20;
        }
    }
};

template <int KIND, int L, bool ALIGNED, bool FAST>
__device__ __forceinline__ void separable_pair(sep_acc<KIND, FAST> &acc, const double *__restrict__ x, uint32_t nx,
                                               uint32_t c)
{
    const uint32_t i = 2u * c;
    if constexpr (ALIGNED) {
        const double2 v = *reinterpret_cast<const double2 *>(x + i);
        acc.add(v.x, i);
        acc.add(v.y, i + 1u);
    } else {
        acc.add(x[i], i);
        if (i + 1u < nx) acc.add(x[i + 1u], i + 1u);
    }
}

// Two gradient elements to a 16-byte aligned slot, non-temporal (written
// once, never read back by the kernel).
__device__ __forceinline__ void store_pair_nt(double2 *p, double a, double b)
{
    typedef double v2 __attribute__((ext_vector_type(2)));
    const v2 v = {a, b};
    __builtin_nontemporal_store(v, reinterpret_cast<v2 *>(p));
}

// rastrigin.cpp:103, g = 2 x + 10 omega sin(omega x) with sin_2pi_small:
// valid for |x| < 2^10.  A row holding a larger element (or inf, NaN) is
// flagged by the fitness pass, and its gradient is computed again with the
// general routine.
__device__ __forceinline__ double rastrigin_grad_fast(double xi)
{
    return 2 * xi + 10.0 * kOmega * sin_2pi_small(xi);
}

// GRAD (rastrigin, 16-byte aligned rows): the same pass also writes the
// gradient of every element it adds to grad (the row's gradient).
template <int KIND, int L, bool ALIGNED, bool FAST, bool GRAD = false>
__device__ __forceinline__ double separable_eval_impl(const double *__restrict__ x, uint32_t nx, int r,
                                                      bool &out_of_range, double *__restrict__ grad = nullptr)
{
    static_assert(!GRAD || (KIND == RASTRIGIN && ALIGNED), "the fused gradient is rastrigin's, on aligned rows");
    sep_acc<KIND, FAST> acc;
    acc.init();
    const uint32_t npairs = (nx + 1u) >> 1;
    uint32_t c = static_cast<uint32_t>(r);
    if constexpr (ALIGNED) {
        // `full` rounds in which every lane of the group has a pair (a trip
        // count uniform over the wavefront, so the loop is scalar-controlled
        // and the loads take a scalar offset), then at most one more pair per
        // lane.  The next pair's load is issued before the current pair is
        // used.  Pairs are added in increasing c either way.
        const double2 *row = reinterpret_cast<const double2 *>(x) + c;
        double2 *grow = GRAD ? reinterpret_cast<double2 *>(grad) + c : nullptr;
        const uint32_t full = npairs / L;
        const auto add_round = [&](const double2 &v, uint32_t j) {
            acc.add(v.x, 2u * (c + j * L));
            acc.add(v.y, 2u * (c + j * L) + 1u);
            if constexpr (GRAD) store_pair_nt(grow + j * L, rastrigin_grad_fast(v.x), rastrigin_grad_fast(v.y));
        };
        if (full > 0u) {
            // two rounds per iteration with the loads alternating between a
            // and b, so that no register copy moves a loaded pair: b's load
            // is in flight while a is used, then a's next one while b is.
            // The barriers keep each load above the use that follows it (the
            // scheduler would otherwise sink it to reuse registers).
            double2 a = row[0];
            uint32_t j = 1u;
            for (; j + 1u < full; j += 2u) {
                const double2 b = row[j * L];
                __builtin_amdgcn_sched_barrier(0);
                add_round(a, j - 1u);
                a = row[(j + 1u) * L];
                __builtin_amdgcn_sched_barrier(0);
                add_round(b, j);
            }
            if (j < full) {
                const double2 b = row[j * L];
                __builtin_amdgcn_sched_barrier(0);
                add_round(a, j - 1u);
                add_round(b, j);
            } else {
                add_round(a, j - 1u);
            }
        }
        if (c + full * L < npairs) {
            const double2 v = row[full * L];
            acc.add(v.x, 2u * (c + full * L));
            acc.add(v.y, 2u * (c + full * L) + 1u);
            if constexpr (GRAD) store_pair_nt(grow + full * L, rastrigin_grad_fast(v.x), rastrigin_grad_fast(v.y));
        }
    } else {
        for (; c < npairs; c += L) separable_pair<KIND, L, ALIGNED, FAST>(acc, x, nx, c);
    }
    if constexpr (FAST) {
        out_of_range = !acc.in_fast_range(2u * ((npairs + L - 1u) / L));
    } else {
        out_of_range = false;
    }
    const double n = static_cast<double>(nx);
    if constexpr (KIND == RASTRIGIN) {
        // rastrigin.cpp:69
        return group_sum<L>(acc.a) + 10. * n;
    } else if constexpr (KIND == ACKLEY) {
        // ackley.cpp:72-73, on the group's first lane only (the exp / sqrt
        // sequence then runs with one active lane) and broadcast
        const double s1 = group_sum<L>(acc.a);
        const double s2 = group_sum<L>(acc.b);
        double f = 0.;
        if (r == 0) {
            const double nepero = exp(1.0);
            f = -20 * exp(-0.2 * sqrt(1.0 / n * s1)) - exp(1.0 / n * s2) + 20 + nepero;
        }
        return L > 1 ? __shfl(f, 0, L) : f;
    } else if constexpr (KIND == SCHWEFEL) {
        // schwefel.cpp:66
        return 418.9828872724338 * n - group_sum<L>(acc.a);
    } else {
        // griewank.cpp:74
        const double s = group_sum<L>(acc.a);
        const double p = group_prod<L>(acc.b);
        return s / 4000. - p + 1.;
    }
}

// A signalling NaN: arithmetic never produces one, so it can mark a result
// that must be recomputed.
constexpr unsigned long long kRecomputeBits = 0x7ff44f5253464958ull;

__device__ __forceinline__ bool is_recompute_mark(double f)
{
    return static_cast<unsigned long long>(__double_as_longlong(f)) == kRecomputeBits;
}

// Fitness of one individual by the L lanes of its group (lane r); the same
// value is returned to every lane.  The per-element operations and the
// reduction tree depend on nx and L only.  For rastrigin and ackley the
// cosines come from cos_2pi_small; an individual with an element outside its
// range (|x| >= 2^10, inf, NaN) gets the recompute mark instead, and
// separable_eval_exact produces its value (kept out of this routine so that
// its registers do not weigh on the streaming kernel).
template <int KIND, int L, bool ALIGNED, bool GRAD = false>
__device__ __forceinline__ double separable_eval(const double *__restrict__ x, uint32_t nx, int r,
                                                 double *__restrict__ grad = nullptr)
{
    constexpr bool fast = (KIND == RASTRIGIN || KIND == ACKLEY);
    bool bad = false;
    const double f = separable_eval_impl<KIND, L, ALIGNED, fast, GRAD>(x, nx, r, bad, grad);
    if constexpr (fast) {
        if (group_max<L>(bad ? 1 : 0)) return __longlong_as_double(static_cast<long long>(kRecomputeBits));
    }
    return f;
}

// The same individual with the math library's cosine for every element.
template <int KIND, int L, bool ALIGNED>
__device__ __forceinline__ double separable_eval_exact(const double *__restrict__ x, uint32_t nx, int r)
{
    bool unused;
    return separable_eval_impl<KIND, L, ALIGNED, false>(x, nx, r, unused);
}

// ---------------------------------------------------------------------------
// Rosenbrock: one lane per individual, sequential in i as rosenbrock.cpp:62-64
// (bit-identical to the reference: only + - * in the same order).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rosenbrock_eval(const double *__restrict__ x, uint32_t m_dim)
{
    double retval = 0.;
    double xi = x[0];
    for (uint32_t i = 0u; i + 1u < m_dim; ++i) {
        const double xn = x[i + 1u];
        const double t = xi * xi - xn;
        retval += 100. * t * t + (xi - 1) * (xi - 1);
        xi = xn;
    }
    return retval;
}

// Rosenbrock fitness of `rows` (1..8) consecutive rows of nx elements (at
// `first`, row stride nx) by one wavefront: the lanes take column chunks of
// 64, each lane computing one element's term for all rows (contiguous loads),
// the terms go to the wavefront's LDS tile (8 x kRosenWaveStride doubles) and
// lane r (r < rows) adds row r's terms in the reference's order
// (rosenbrock.cpp:62-64), so the sum is rosenbrock_eval's bit for bit.  The
// value is valid in lanes r < rows.  Clamped indices keep every load in the
// rows given; a clamped term is never summed.
constexpr uint32_t kRosenWaveRows = 8, kRosenWaveStride = 65; // odd stride: the summing lanes hit distinct banks

__device__ __forceinline__ double rosenbrock_wave_rows(const double *__restrict__ first, uint32_t rows, uint32_t nx,
                                                       double *__restrict__ tile, uint32_t lane)
{
    constexpr uint32_t W = 64;
    double acc = 0.;
    for (uint32_t c0 = 0; c0 < nx; c0 += W) {
        const uint32_t ii = min(c0 + lane, nx - 1u), in = min(ii + 1u, nx - 1u);
        double xi[kRosenWaveRows], xn[kRosenWaveRows];
#pragma unroll
        for (uint32_t r = 0; r < kRosenWaveRows; ++r) {
            const double *x = first + static_cast<uint64_t>(min(r, rows - 1u)) * nx;
            xi[r] = x[ii];
            xn[r] = x[in];
        }
#pragma unroll
        for (uint32_t r = 0; r < kRosenWaveRows; ++r) {
            const double d = xi[r] * xi[r] - xn[r];
            tile[r * kRosenWaveStride + lane] = 100. * d * d + (xi[r] - 1) * (xi[r] - 1);
        }
        __builtin_amdgcn_wave_barrier();
        const uint32_t cnt = c0 + 1u < nx ? min(W, nx - 1u - c0) : 0u;
        if (lane < rows) {
            const double *t = tile + lane * kRosenWaveStride;
            for (uint32_t j = 0; j < cnt; ++j) acc += t[j];
        }
        __builtin_amdgcn_wave_barrier();
    }
    return acc;
}

// rosenbrock.cpp:83-92, element i of m_dim from its neighbours xm = x[i-1]
// (unused for i = 0), x0 = x[i] and xp = x[i+1] (unused for i = m_dim - 1).
// Both parts are always evaluated (no branch, so that callers' loads of xm
// and xp are not sunk into one); the middle elements add them in the
// reference's association, (a - b) + c.
__device__ __forceinline__ double rosenbrock_grad3(double xm, double x0, double xp, uint32_t m_dim, uint32_t i)
{
    const double right = -400. * x0 * (xp - x0 * x0) - 2. * (1 - x0);
    const double left = 200. * (x0 - xm * xm);
    return i == 0u ? right : (i == m_dim - 1u ? left : right + left);
}

// rastrigin.cpp:103.
__device__ __forceinline__ double rastrigin_grad(double xi)
{
    return 2 * xi + 10.0 * kOmega * sin_any(kOmega * xi);
}

// A row's gradient elements r, r + L, ... with the general routine.
__device__ __noinline__ void rastrigin_grad_row(const double *__restrict__ x, uint32_t nx, uint32_t r, uint32_t L,
                                                double *__restrict__ g)
{
    for (uint32_t i = r; i < nx; i += L) g[i] = rastrigin_grad(x[i]);
}

// ---------------------------------------------------------------------------
// ZDT (zdt.cpp:233-356): one lane per individual, each expression in the
// reference's order and association.  Only + - * / sqrt and round for ZDT1,
// ZDT2 and ZDT5 (bit-exact); ZDT3, ZDT4 and ZDT6 call the device sin, cos,
// exp and pow.
// ---------------------------------------------------------------------------
// 10 cos(4 pi x) (zdt.cpp:281-296) as cos_2pi_small at the exact 2x for
// |2x| < 2^10 (the error against the reference's cos(4 pi x) is the rounding
// of 4 pi x there, about |4 pi x| 1e-16 absolute; ZDT4's g >= 1 sums such
// terms); the math library's cosine of the reference's argument beyond.
__device__ __forceinline__ double zdt4_cos10(double x)
{
    const double y = 2. * x;
    if (fabs(y) < 0x1p10) return cos_2pi_small<true>(y);
    return 10. * cos(4. * kPi * x);
}

__device__ __forceinline__ void zdt_eval(uint32_t id, const double *__restrict__ x, uint32_t N, double &f0,
                                         double &f1)
{
    if (id == 5u) {
        // zdt.cpp:298-339: 30 bits, then blocks of 5; x rounded half away
        // from zero and compared with 1
        const uint32_t nvec = (N - 30u) / 5u + 1u;
        uint32_t u0 = 0u;
        for (uint32_t j = 0u; j < 30u; ++j) u0 += round(x[j]) == 1. ? 1u : 0u;
        double g = 0.;
        for (uint32_t i = 1u; i < nvec; ++i) {
            uint32_t ui = 0u;
            for (uint32_t j = 0u; j < 5u; ++j) ui += round(x[30u + 5u * (i - 1u) + j]) == 1. ? 1u : 0u;
            const uint32_t vi = ui < 5u ? 2u + ui : 1u;
            g += static_cast<double>(vi);
        }
        f0 = 1.0 + static_cast<double>(u0);
        f1 = g * (1. / f0);
        return;
    }
    const double x0 = x[0];
    if (id == 4u) {
        // zdt.cpp:281-296
        double g = 1 + 10 * static_cast<double>(N - 1u);
        for (uint32_t i = 1u; i < N; ++i) g += x[i] * x[i] - zdt4_cos10(x[i]);
        f0 = x0;
        f1 = g * (1. - sqrt(x0 / g));
        return;
    }
    double g = 0.;
    for (uint32_t i = 1u; i < N; ++i) g += x[i];
    if (id == 6u) {
        // zdt.cpp:341-356
        f0 = 1 - exp(-4 * x0) * pow(sin(6 * kPi * x0), 6);
        g = 1 + 9 * pow((g / static_cast<double>(N - 1u)), 0.25);
        f1 = g * (1 - (f0 / g) * (f0 / g));
        return;
    }
    g = 1. + (9. * g) / static_cast<double>(N - 1u);
    f0 = x0;
    if (id == 1u) {
        f1 = g * (1. - sqrt(x0 / g)); // zdt.cpp:233-247
    } else if (id == 2u) {
        f1 = g * (1. - (x0 / g) * (x0 / g)); // zdt.cpp:249-263
    } else {
        f1 = g * (1. - sqrt(x0 / g) - x0 / g * sin(10. * kPi * x0)); // zdt.cpp:265-279
    }
}

// ---------------------------------------------------------------------------
// DTLZ (dtlz.cpp:181-407): one lane per individual; M = fdim objectives, the
// distance variables x_M = x[M-1 .. nx-1].  pow(t, 2) is written t * t, the
// correctly rounded square (glibc's pow returns it too, to within 0.52 ulp).
// ---------------------------------------------------------------------------
// cos(20 pi t) (dtlz.cpp:207-214) as cos(2 pi (10 t)) through cos_2pi_small
// for |10 t| < 2^10 (about |20 pi t| 1e-16 absolute from the reference's
// rounded argument), the math library's cosine beyond.
__device__ __forceinline__ double dtlz_cos20pi(double t)
{
    const double y = 10. * t;
    if (fabs(y) < 0x1p10) return cos_2pi_small<false>(y);
    return cos(20. * kPi * t);
}

__device__ __forceinline__ double dtlz_g(uint32_t id, const double *__restrict__ xm, uint32_t K)
{
    double y = 0.;
    switch (id) {
        case 1u:
        case 3u: // dtlz.cpp:207-214
            for (uint32_t i = 0u; i < K; ++i) {
                const double t = xm[i] - 0.5;
                y += t * t - dtlz_cos20pi(t);
            }
            return 100. * (y + static_cast<double>(K));
        case 6u: // dtlz.cpp:225-232
            for (uint32_t i = 0u; i < K; ++i) y += pow(xm[i], 0.1);
            return y;
        case 7u: // dtlz.cpp:234-244
            for (uint32_t i = 0u; i < K; ++i) y += xm[i];
            return (9. / static_cast<double>(K)) * y;
        default: // 2, 4, 5: dtlz.cpp:216-223
            for (uint32_t i = 0u; i < K; ++i) {
                const double t = xm[i] - 0.5;
                y += t * t;
            }
            return y;
    }
}

// Position variable j of the shape functions: x_j (DTLZ1-3), x_j^alpha
// (DTLZ4, dtlz.cpp:320-341), theta_j (DTLZ5-6, dtlz.cpp:343-383).
__device__ __forceinline__ double dtlz_pos(uint32_t id, const double *__restrict__ x, uint32_t j, double g,
                                           double alpha)
{
    if (id == 4u) return pow(x[j], alpha);
    if (id == 5u || id == 6u) {
        if (j == 0u) return x[0];
        const double t = 1. / (2. * (1. + g));
        return t + ((g * x[j]) / (1.0 + g));
    }
    return x[j];
}

__device__ __forceinline__ void dtlz_eval(uint32_t id, const double *__restrict__ x, uint32_t nx, uint32_t M,
                                          uint32_t alpha, double *__restrict__ f)
{
    const double gd = dtlz_g(id, x + (M - 1u), nx - M + 1u);
    if (id == 7u) {
        // dtlz.cpp:385-407 with h7 (dtlz.cpp:246-256)
        const double g = 1. + gd;
        double y = 0.;
        for (uint32_t i = 0u; i + 1u < M; ++i) {
            f[i] = x[i];
            y += (x[i] / (1.0 + g)) * (1.0 + sin(3 * kPi * x[i]));
        }
        f[M - 1u] = (1. + g) * (static_cast<double>(M) - y);
        return;
    }
    const double g = gd;
    if (id == 1u) {
        // dtlz.cpp:258-283
        double f0 = 0.5 * (1. + g);
        for (uint32_t i = 0u; i + 1u < M; ++i) f0 *= x[i];
        f[0] = f0;
        for (uint32_t i = 1u; i + 1u < M; ++i) {
            double fi = 0.5 * (1.0 + g);
            for (uint32_t j = 0u; j < M - (i + 1u); ++j) fi *= x[j];
            fi *= 1. - x[M - (i + 1u)];
            f[i] = fi;
        }
        f[M - 1u] = 0.5 * (1. - x[0]) * (1. + g);
        return;
    }
    // DTLZ2-6 (dtlz.cpp:285-383): f[M-1-k] = (1+g) c_0 ... c_{k-1} s_k and
    // f[0] = (1+g) c_0 ... c_{M-2}, with c_j, s_j = cos, sin(v_j pi/2).  The
    // reference multiplies each objective's product from (1+g) left to right;
    // the running product P_k = (1+g) c_0 ... c_{k-1} is the same sequence of
    // roundings, so each position variable's sine and cosine are taken once
    // (M-1 sincos instead of M(M-1)/2 + M-1 separate calls).
    const double a = static_cast<double>(alpha);
    double P = 1. + g;
    for (uint32_t k = 0u; k + 1u < M; ++k) {
        double s, c;
        sincos(dtlz_pos(id, x, k, g, a) * kPiHalf, &s, &c);
        f[M - 1u - k] = P * s;
        P *= c;
    }
    f[0] = P;
}

// ---------------------------------------------------------------------------
// Multi-objective sphere: f_k = sum_i (x_i - c_{k,i})^2, sequential in i.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double mo_sphere_eval(const double *__restrict__ x, uint32_t nx,
                                                 const double *__restrict__ ck)
{
    double acc = 0.;
    for (uint32_t i = 0u; i < nx; ++i) {
        const double d = x[i] - ck[i];
        acc += d * d;
    }
    return acc;
}

// ---------------------------------------------------------------------------
// Lennard-Jones (lennard_jones.cpp:72-91 with the encoding of :132-151)
// ---------------------------------------------------------------------------

// Decision-vector element e -> (atom, coordinate).
__device__ __forceinline__ void lj_slot(uint32_t e, uint32_t &atom, uint32_t &coord)
{
    if (e == 0u) {
        atom = 1u;
        coord = 2u;
    } else if (e < 3u) {
        atom = 2u;
        coord = e;
    } else {
        atom = 3u + (e - 3u) / 3u;
        coord = (e - 3u) % 3u;
    }
}

// Stages `count` individuals (rows of nx = 3*atoms-6) into LDS, lj_stride
// doubles each: atom a's (x, y, z) at [3a, 3a+3), and for a < lj_copy_atoms
// again at [3(a+A), 3(a+A)+3), so that a partner a + k is read without a wrap
// and at a constant offset from atom a.  The rows are contiguous in global
// memory and read as one flat range, kStageLoads loads in flight per thread
// before any LDS write.  Called by all threads of the workgroup; a
// __syncthreads() must follow.
constexpr uint32_t kStageLoads = 4;

__device__ __forceinline__ void lj_stage(const double *__restrict__ rows, uint32_t count, uint32_t atoms, bool grad,
                                         double *__restrict__ lds)
{
    const uint32_t nx = 3u * atoms - 6u;
    const uint32_t stride = lj_stride(atoms, grad);
    const uint32_t copy = 3u * atoms;
    const uint32_t copied = 3u * lj_copy_atoms(atoms, grad); // flat positions below this are copied
    const uint32_t nt = blockDim.x, tid = threadIdx.x;
    // the six fixed coordinates of each individual: atom 0 (x,y,z), atom 1
    // (x,y), atom 2 (x) at flat positions 0..4 and 6
    for (uint32_t z = tid; z < 6u * count; z += nt) {
        const uint32_t g = z / 6u, zz = z - 6u * g;
        const uint32_t f = zz < 5u ? zz : 6u;
        double *ind = lds + g * stride;
        ind[f] = 0.0;
        if (f < copied) ind[copy + f] = 0.0;
    }
    // element e of the flat range is element el = e mod nx of row g = e / nx,
    // at flat position el + 6, except x[0] (atom 1's z) at 5 (lj_slot).  A
    // thread's e advances by nt: (g, el) by (nt / nx, nt mod nx) and a carry.
    const uint32_t total = count * nx;
    const uint32_t dq = nt / nx, dr = nt - dq * nx;
    uint32_t g = tid / nx, el = tid - g * nx;
    for (uint32_t base = 0; base < total; base += kStageLoads * nt) {
        double v[kStageLoads];
#pragma unroll
        for (uint32_t u = 0; u < kStageLoads; ++u) {
            const uint32_t e = base + u * nt + tid;
            v[u] = e < total ? rows[e] : 0.0;
        }
#pragma unroll
        for (uint32_t u = 0; u < kStageLoads; ++u) {
            if (base + u * nt + tid < total) {
                const uint32_t f = el == 0u ? 5u : el + 6u;
                double *ind = lds + g * stride;
                ind[f] = v[u];
                if (f < copied) ind[copy + f] = v[u];
            }
            g += dq;
            el += dr;
            if (el >= nx) {
                el -= nx;
                ++g;
            }
        }
    }
}

// First index (i-major, lennard_jones.cpp:74-75) of the pairs with first atom i.
__device__ __forceinline__ uint32_t lj_row_start(uint32_t i, uint32_t atoms)
{
    return i * atoms - (i * (i + 1u)) / 2u;
}

// Index, in the reference's loop order, of the pair {a, b}.
__device__ __forceinline__ int lj_pair_index(uint32_t a, uint32_t b, uint32_t atoms)
{
    const uint32_t i = a < b ? a : b, j = a < b ? b : a;
    return static_cast<int>(lj_row_start(i, atoms) + (j - i - 1u));
}

// The careful paths below are out of line, so the compiler no longer sees
// that their staged individual lies in LDS; they address it through an LDS
// pointer (LDS loads instead of generic ones).  Their callers always pass a
// pointer into the workgroup's LDS block.
using lds_double = const __attribute__((address_space(3))) double;

__device__ __forceinline__ lds_double *lds_ptr(const double *p)
{
    return (lds_double *)p;
}

// 1 / d by v_rcp_f64 refined with two Newton steps: accurate to about an ulp
// for d in [2^-1000, 2^1000]; outside that range (0, subnormal, huge, inf,
// NaN) the result is not 1 / d, and callers that can meet such d use lj_recip.
__device__ __forceinline__ double lj_recip_fast(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
}

// 1 / d for any d: lj_recip_fast where it is valid, IEEE division otherwise.
__device__ __forceinline__ double lj_recip(double d)
{
    if (d > 0x1p-1000 && d < 0x1p1000) return lj_recip_fast(d);
    return 1.0 / d;
}

// Cyclic pair schedule of an individual's atoms over the L lanes of its group:
// lane r owns atoms a = r, r+L, ...; step k (1 <= k <= A/2) pairs each owned
// atom a with b = a + k (mod A, read from the second copy), and when A is even
// the last step keeps only a < A/2, so every unordered pair is visited exactly
// once.  Within a step the lanes of a group read consecutive atoms.
struct lj_cycle {
    uint32_t atoms, half;
    __device__ __forceinline__ explicit lj_cycle(uint32_t a) : atoms(a), half(a / 2u) {}
    __device__ __forceinline__ bool counts(uint32_t a, uint32_t k) const
    {
        return 2u * k < atoms || a < half; // k == A/2 with A even: one of the two visits
    }
    // steps in which every visit counts
    __device__ __forceinline__ uint32_t full_steps() const
    {
        return (atoms & 1u) ? half : half - 1u;
    }
    __device__ __forceinline__ uint32_t partner(uint32_t a, uint32_t k) const
    {
        const uint32_t b = a + k;
        return b >= atoms ? b - atoms : b;
    }
};

// Steps of one lane: the schedule's T steps are dealt round-robin to the KS
// halves of a group, half h taking k = h + 1 + i KS.  All halves run the same
// n = ceil(T / KS) steps: n - 1 steps that are always full, then a last one
// that is full, the partial step A/2 (A even: only a < A/2 counts) or, past
// T, a dummy of weight 0.
struct lj_steps {
    uint32_t first, stride, n, last, T;
    bool dummy;
    __device__ __forceinline__ lj_steps(uint32_t T_, uint32_t h, uint32_t KS)
    {
        T = T_;
        first = h + 1u;
        stride = KS;
        n = (T + KS - 1u) / KS;
        const uint32_t kt = first + (n - 1u) * KS;
        dummy = kt > T;
        last = dummy ? T : kt;
    }
};

// ---- careful paths: any input, the reference's special cases -----------------

// The reference's dist == 0 rule (lennard_jones.cpp:81-83,89): the result is
// +inf unless a NaN term comes after the last zero-distance pair in the
// reference's pair order, then NaN.  Only reached when some pair of the
// individual has zero distance; recomputes the group's terms to find the NaNs.
template <int L>
__device__ __noinline__ double lj_zero_rule(const double *__restrict__ P, uint32_t atoms, int r, int zlast)
{
    const lj_cycle cyc(atoms);
    int nlast = -1;
    for (uint32_t k = 1u; k <= cyc.half; ++k) {
        for (uint32_t a = static_cast<uint32_t>(r); a < atoms; a += L) {
            if (!cyc.counts(a, k)) continue;
            lds_double *pa = lds_ptr(P) + 3u * a, *pb = pa + 3u * k;
            const double dx = pa[0] - pb[0], dy = pa[1] - pb[1], dz = pa[2] - pb[2];
            const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
            if (r2 == 0.0) continue;
            const double s = lj_recip(r2 * r2 * r2);
            const double t = fma(s, s, -s);
            if (t != t) nlast = max(nlast, lj_pair_index(a, cyc.partner(a, k), atoms));
        }
    }
    nlast = group_max<L>(nlast);
    return nlast > zlast ? __builtin_nan("") : __builtin_inf();
}

// Energy with every special case: lj_recip's full range, zero distances and
// the reference's rule for them (see lj_zero_rule), in one pass over the
// streaming kernel's own schedule (lj_steps over the group's KS halves,
// kLjCarefulChains owned atoms per pass, each its own accumulation chain), so
// the group's lanes share the pairs as in lj_energy.  Two chains, not more:
// the kernels that call this function reserve its registers as well.
constexpr uint32_t kLjCarefulChains = 2;
// TAG gives a caller its own copy, compiled for that caller's register budget.
template <int L, uint32_t TAG = 0>
__device__ __noinline__ double lj_energy_careful(const double *__restrict__ P, uint32_t atoms, int r)
{
    constexpr int LA = lj_atom_lanes(L);
    constexpr uint32_t KS = static_cast<uint32_t>(L / LA);
    const uint32_t ra = static_cast<uint32_t>(r) % LA, h = static_cast<uint32_t>(r) / LA;
    const lj_cycle cyc(atoms);
    const lj_steps st(cyc.half, h, KS);
    lds_double *Q = lds_ptr(P);
    double acc[kLjCarefulChains] = {};
    int zlast = -1, nlast = -1;
    for (uint32_t a0 = ra; a0 < atoms; a0 += kLjCarefulChains * LA) {
        uint32_t a[kLjCarefulChains], aa[kLjCarefulChains];
        double x[kLjCarefulChains], y[kLjCarefulChains], z[kLjCarefulChains];
#pragma unroll
        for (uint32_t q = 0; q < kLjCarefulChains; ++q) {
            a[q] = a0 + q * LA;
            aa[q] = a[q] < atoms ? a[q] : ra; // a slot past the last atom repeats atom ra, uncounted
            x[q] = Q[3u * aa[q]];
            y[q] = Q[3u * aa[q] + 1u];
            z[q] = Q[3u * aa[q] + 2u];
        }
        for (uint32_t i = 0u; i < st.n; ++i) {
            const bool last = i + 1u == st.n;
            const uint32_t k = last ? st.last : st.first + i * st.stride;
            double r2[kLjCarefulChains], d[kLjCarefulChains], s[kLjCarefulChains];
            bool cnt[kLjCarefulChains], odd = false;
#pragma unroll
            for (uint32_t q = 0; q < kLjCarefulChains; ++q) {
                lds_double *pb = Q + 3u * (aa[q] + k);
                const double dx = x[q] - pb[0], dy = y[q] - pb[1], dz = z[q] - pb[2];
                r2[q] = fma(dz, dz, fma(dy, dy, dx * dx));
                d[q] = r2[q] * r2[q] * r2[q];
                cnt[q] = a[q] < atoms && !(last && st.dummy) && cyc.counts(a[q], k);
                s[q] = lj_recip_fast(d[q]);
                odd |= cnt[q] && !(d[q] > 0x1p-1000 && d[q] < 0x1p1000);
            }
            // rare: a counted visit outside lj_recip_fast's range (zero,
            // subnormal, huge, inf or NaN r^6) takes the IEEE reciprocal;
            // zero distances and NaN terms are recorded by pair index
            if (odd) {
#pragma unroll
                for (uint32_t q = 0; q < kLjCarefulChains; ++q) {
                    if (!cnt[q] || (d[q] > 0x1p-1000 && d[q] < 0x1p1000)) continue;
                    const int pi = lj_pair_index(a[q], cyc.partner(a[q], k), atoms);
                    if (r2[q] == 0.0) {
                        zlast = max(zlast, pi);
                        cnt[q] = false;
                        continue;
                    }
                    s[q] = 1.0 / d[q];
                    const double t = fma(s[q], s[q], -s[q]);
                    if (t != t) nlast = max(nlast, pi);
                }
            }
#pragma unroll
            for (uint32_t q = 0; q < kLjCarefulChains; ++q) {
                const double t = fma(s[q], s[q], -s[q]);
                acc[q] += cnt[q] ? t : 0.0;
            }
        }
    }
    const double sum = group_sum<L>(acc[0] + acc[1]);
    zlast = group_max<L>(zlast);
    nlast = group_max<L>(nlast);
    // lennard_jones.cpp:81-83,89: after a zero distance f is DBL_MAX plus the
    // later terms (each >= -1/4), so 4 f overflows: +inf, or NaN if a NaN term
    // follows the last zero-distance pair in the reference's pair order
    if (zlast >= 0) return nlast > zlast ? __builtin_nan("") : __builtin_inf();
    return 4 * sum;
}

// Writes the gradient components of atom a (its force f) into the gradient row
// g (nx doubles) through the encoding of lennard_jones.cpp:132-151: atom 0 is
// fixed, atom 1 moves along z only, atom 2 in y and z.
__device__ __forceinline__ void lj_write_grad(uint32_t a, double fx, double fy, double fz, double *__restrict__ g)
{
    if (a == 1u) {
        g[0] = fz;
    } else if (a == 2u) {
        g[1] = fy;
        g[2] = fz;
    } else if (a >= 3u) {
        const uint32_t o = 3u * (a - 2u);
        g[o] = fx;
        g[o + 1u] = fy;
        g[o + 2u] = fz;
    }
}

// Energy and gradient with every special case (see lj_energy_careful).
template <int L>
__device__ __noinline__ double lj_energy_grad_careful(const double *__restrict__ P, uint32_t atoms, int r,
                                                      double *__restrict__ g)
{
    const lj_cycle cyc(atoms);
    double acc = 0.;
    int zlast = -1;
    for (uint32_t a = static_cast<uint32_t>(r); a < atoms; a += L) {
        lds_double *pa = lds_ptr(P) + 3u * a;
        double fx = 0., fy = 0., fz = 0.;
        for (uint32_t k = 1u; k < atoms; ++k) {
            lds_double *pb = pa + 3u * k;
            const double dx = pa[0] - pb[0], dy = pa[1] - pb[1], dz = pa[2] - pb[2];
            const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
            const double s = lj_recip(r2 * r2 * r2);
            const double inv_r2 = (s * r2) * r2;
            const double coef = (fma(-48., s, 24.) * s) * inv_r2;
            fx = fma(coef, dx, fx);
            fy = fma(coef, dy, fy);
            fz = fma(coef, dz, fz);
            if (k <= cyc.half && cyc.counts(a, k)) {
                if (r2 == 0.0) {
                    zlast = max(zlast, lj_pair_index(a, cyc.partner(a, k), atoms));
                } else {
                    acc += fma(s, s, -s);
                }
            }
        }
        lj_write_grad(a, fx, fy, fz, g);
    }
    acc = group_sum<L>(acc);
    zlast = group_max<L>(zlast);
    if (zlast >= 0) return lj_zero_rule<L>(P, atoms, r, zlast);
    return 4 * acc;
}

// ---- streaming paths ------------------------------------------------------------

// Lane r of a group owns atoms a_q = r + q L.  APL > 0: exactly APL owned
// slots (the ceiling of atoms / L), all held in registers; APL == 0: any
// count, taken four slots at a time.  A slot past the last atom repeats atom r
// with weight 0, so that no visit needs a mask.
__host__ __device__ constexpr uint32_t lj_apl(uint32_t atoms, int L)
{
    const uint32_t la = static_cast<uint32_t>(lj_atom_lanes(L));
    const uint32_t apl = (atoms + la - 1u) / la;
    // up to 4 slots for every lane count; 32-lane groups (from 32 atoms on)
    // up to 8, i.e. 128 atoms
    return apl <= (L >= 32 ? 8u : 4u) ? apl : 0u;
}


template <int L, uint32_t NS>
struct lj_slots {
    const double *p[NS];
    double x[NS], y[NS], z[NS], w[NS];
    uint32_t a[NS];
    // RANGE: a coordinate beyond 2^100 (or NaN) makes the slot's weight NaN,
    // so that the individual takes the careful path (see lj_recip_group)
    template <bool RANGE>
    __device__ __forceinline__ void load(const double *__restrict__ P, uint32_t a0, uint32_t r, uint32_t atoms)
    {
#pragma unroll
        for (uint32_t q = 0; q < NS; ++q) {
            a[q] = a0 + q * L;
            const bool own = a[q] < atoms;
            p[q] = P + 3u * (own ? a[q] : r);
            x[q] = p[q][0];
            y[q] = p[q][1];
            z[q] = p[q][2];
            w[q] = own ? 1.0 : 0.0;
            if constexpr (RANGE) {
                const bool in_range = fabs(x[q]) <= 0x1p100 && fabs(y[q]) <= 0x1p100 && fabs(z[q]) <= 0x1p100;
                if (!in_range) w[q] = __builtin_nan("");
            }
        }
    }
};

// One visit: the pair (owned slot, partner at pb).  inv = 1 / r^2 from
// v_rcp_f64 (24 bits) and one Newton step (measured: 48.7 bits, a relative
// error of 2.2e-15), and s = inv^3 = r^-6; an energy term is then within about
// 1e-14 of its own size.  For r^2 outside [2^-1000, 2^1000] these are not the
// reference's values, but then the energy term (or, for r^2 = 0 or inf, every
// result) is not finite unless the true term underflows to 0 as well, and the
// callers' finiteness test sends such an individual to the careful path.
struct lj_pair {
    double dx, dy, dz, inv, s;
    __device__ __forceinline__ lj_pair(double x, double y, double z, const double *__restrict__ pb)
    {
        dx = x - pb[0];
        dy = y - pb[1];
        dz = z - pb[2];
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double r = __builtin_amdgcn_rcp(r2);
        inv = fma(r, fma(-r2, r, 1.0), r);
        s = (inv * inv) * inv;
    }
};

// 1 / d[q] for N <= 4 values by one reciprocal of their product (Montgomery's
// batch inversion): one v_rcp_f64 refined by one Newton step as in lj_pair,
// then 3 (N - 1) products, each adding a rounding (a relative error of a few
// 1e-16 beyond lj_pair's).  For squared distances in [2^-250, 2^204] every
// product stays inside [2^-1000, 2^1000], where the reciprocal is accurate.
// Larger ones need a coordinate beyond 2^100, which the callers test for and
// send to a careful path; a smaller one makes its own r^-12 term overflow,
// so the sum is not finite and the individual takes the careful path too.
template <uint32_t N>
__device__ __forceinline__ void lj_recip_group(const double *d, double *inv)
{
    double p[N];
    p[0] = d[0];
#pragma unroll
    for (uint32_t q = 1; q < N; ++q) p[q] = p[q - 1] * d[q];
    const double r0 = __builtin_amdgcn_rcp(p[N - 1]);
    double t = fma(r0, fma(-p[N - 1], r0, 1.0), r0);
#pragma unroll
    for (uint32_t q = N - 1; q > 0; --q) {
        inv[q] = t * p[q - 1];
        t = t * d[q];
    }
    inv[0] = t;
}

// 1 / d[q] for NS values, in groups of at most four.
template <uint32_t NS>
__device__ __forceinline__ void lj_recip_slots(const double (&d)[NS], double (&inv)[NS])
{
#pragma unroll
    for (uint32_t q0 = 0; q0 < NS; q0 += 4u) {
        if (NS - q0 >= 4u) lj_recip_group<4>(d + q0, inv + q0);
        else if (NS - q0 == 3u) lj_recip_group<3>(d + q0, inv + q0);
        else if (NS - q0 == 2u) lj_recip_group<2>(d + q0, inv + q0);
        else lj_recip_group<1>(d + q0, inv + q0);
    }
}

// One step of the cyclic schedule for every owned slot: the pairs (slot q,
// partner at sl.p[q] + 3 k), their reciprocals taken together
// (lj_recip_slots), s = r^-6.
template <int L, uint32_t NS>
struct lj_visit {
    double s[NS];
    __device__ __forceinline__ lj_visit(const lj_slots<L, NS> &sl, uint32_t k)
    {
        double r2[NS], inv[NS];
#pragma unroll
        for (uint32_t q = 0; q < NS; ++q) {
            const double *pb = sl.p[q] + 3u * k;
            const double dx = sl.x[q] - pb[0], dy = sl.y[q] - pb[1], dz = sl.z[q] - pb[2];
            r2[q] = fma(dz, dz, fma(dy, dy, dx * dx));
        }
        lj_recip_slots<NS>(r2, inv);
#pragma unroll
        for (uint32_t q = 0; q < NS; ++q) s[q] = (inv[q] * inv[q]) * inv[q];
    }
};

// The energy terms of the owned slots over the lane's steps.  BATCH: each
// step's reciprocals taken together (lj_visit); else one per pair (lj_pair),
// which holds fewer registers (for callers that are register-bound).
template <int L, uint32_t NS, bool BATCH>
__device__ __forceinline__ void lj_energy_slots(const double *__restrict__ P, uint32_t atoms, uint32_t a0, uint32_t r,
                                                const lj_steps &st, double &acc)
{
    const lj_cycle cyc(atoms);
    lj_slots<L, NS> sl;
    sl.template load<BATCH>(P, a0, r, atoms);
    const bool partial = !st.dummy && st.last == cyc.half && !(atoms & 1u);
    if constexpr (BATCH) {
        for (uint32_t i = 0u; i + 1u < st.n; ++i) {
            const uint32_t k = st.first + i * st.stride;
            const lj_visit<L, NS> v(sl, k);
#pragma unroll
            for (uint32_t q = 0; q < NS; ++q) acc = fma(sl.w[q], fma(v.s[q], v.s[q], -v.s[q]), acc);
        }
        const lj_visit<L, NS> v(sl, st.last);
#pragma unroll
        for (uint32_t q = 0; q < NS; ++q) {
            const double w = st.dummy || (partial && sl.a[q] >= cyc.half) ? 0.0 : sl.w[q];
            acc = fma(w, fma(v.s[q], v.s[q], -v.s[q]), acc);
        }
    } else {
        for (uint32_t i = 0u; i + 1u < st.n; ++i) {
            const uint32_t k = st.first + i * st.stride;
#pragma unroll
            for (uint32_t q = 0; q < NS; ++q) {
                const lj_pair v(sl.x[q], sl.y[q], sl.z[q], sl.p[q] + 3u * k);
                acc = fma(sl.w[q], fma(v.s, v.s, -v.s), acc);
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < NS; ++q) {
            const lj_pair v(sl.x[q], sl.y[q], sl.z[q], sl.p[q] + 3u * st.last);
            const double w = st.dummy || (partial && sl.a[q] >= cyc.half) ? 0.0 : sl.w[q];
            acc = fma(w, fma(v.s, v.s, -v.s), acc);
        }
    }
}

// Energy of the individual staged at P by the L lanes of a group (lane r):
// sum of 4 (s^2 - s), s = r^-6, over the cyclic schedule, then the group's
// butterfly sum.  The streaming loop has no masks and no range checks: a
// zero, tiny or NaN distance makes the sum NaN or infinite (see lj_pair), and
// such an individual goes to lj_energy_careful, which handles the reference's
// special cases.
template <uint32_t R, uint32_t S, bool PAD>
__device__ __forceinline__ double lj_energy_blk_sum(const double *__restrict__ P, uint32_t atoms, int r);

template <int L, uint32_t APL, bool BATCH = true>
__device__ __forceinline__ double lj_energy(const double *__restrict__ P, uint32_t atoms, int r)
{
    if constexpr (BATCH && ((L == 16 && APL >= 2u && APL <= 4u) || (L == 32 && APL >= 5u && APL <= 8u))) {
        // the pair-once blocks (lj_lanes): 16 lanes and blocks of APL atoms, or
        // 32 lanes and blocks of ceil(APL / 2)
        if (lj_energy_blocks(atoms)) {
            constexpr uint32_t R = static_cast<uint32_t>(L), S = L == 16 ? APL : (APL + 1u) / 2u;
            const double acc = atoms % R == 0u ? lj_energy_blk_sum<R, S, false>(P, atoms, r)
                                               : lj_energy_blk_sum<R, S, true>(P, atoms, r);
            if (!(acc - acc == 0.0)) return lj_energy_careful<L, APL>(P, atoms, r);
            return 4 * acc;
        }
    }
    constexpr int LA = lj_atom_lanes(L);
    constexpr uint32_t KS = static_cast<uint32_t>(L / LA);
    const uint32_t ra = static_cast<uint32_t>(r) % LA, h = static_cast<uint32_t>(r) / LA;
    const lj_steps st(atoms / 2u, h, KS);
    double acc = 0.;
    if constexpr (APL > 0) {
        lj_energy_slots<LA, APL, BATCH>(P, atoms, ra, ra, st, acc);
    } else {
        for (uint32_t a0 = ra; a0 < atoms; a0 += 4u * LA) {
            lj_energy_slots<LA, 4, BATCH>(P, atoms, a0, ra, st, acc);
        }
    }
    acc = group_sum<L>(acc);
    if (!(acc - acc == 0.0)) return lj_energy_careful<L, BATCH ? APL : 0u>(P, atoms, r);
    return 4 * acc;
}

template <int L, int LA, uint32_t NS>
__device__ __forceinline__ void lj_grad_slots(const double *__restrict__ P, uint32_t atoms, uint32_t a0, uint32_t r,
                                              const lj_steps &st, bool owner, double &acc, double &fsum,
                                              double *__restrict__ g)
{
    lj_slots<LA, NS> sl;
    sl.template load<false>(P, a0, r, atoms);
    double fx[NS], fy[NS], fz[NS];
#pragma unroll
    for (uint32_t q = 0; q < NS; ++q) fx[q] = fy[q] = fz[q] = 0.;
    for (uint32_t i = 0u; i + 1u < st.n; ++i) {
        const uint32_t k = st.first + i * st.stride;
#pragma unroll
        for (uint32_t q = 0; q < NS; ++q) {
            const lj_pair v(sl.x[q], sl.y[q], sl.z[q], sl.p[q] + 3u * k);
            const double coef = (fma(-2., v.s, 1.) * v.s) * v.inv;
            fx[q] = fma(coef, v.dx, fx[q]);
            fy[q] = fma(coef, v.dy, fy[q]);
            fz[q] = fma(coef, v.dz, fz[q]);
            // every pair is visited from both of its atoms: the sum is halved
            acc = fma(sl.w[q], fma(v.s, v.s, -v.s), acc);
        }
    }
    {
        // the last step, possibly a dummy: weight 0 for the force and the energy
        const double wk = st.dummy ? 0.0 : 1.0;
#pragma unroll
        for (uint32_t q = 0; q < NS; ++q) {
            const lj_pair v(sl.x[q], sl.y[q], sl.z[q], sl.p[q] + 3u * st.last);
            const double coef = ((fma(-2., v.s, 1.) * v.s) * v.inv) * wk;
            fx[q] = fma(coef, v.dx, fx[q]);
            fy[q] = fma(coef, v.dy, fy[q]);
            fz[q] = fma(coef, v.dz, fz[q]);
            acc = fma(sl.w[q] * wk, fma(v.s, v.s, -v.s), acc);
        }
    }
    if constexpr (L > LA) {
        // the two halves' partial forces of the same atoms (lanes r, r + LA)
#pragma unroll
        for (uint32_t q = 0; q < NS; ++q) {
            fx[q] += __shfl_xor(fx[q], LA, L);
            fy[q] += __shfl_xor(fy[q], LA, L);
            fz[q] += __shfl_xor(fz[q], LA, L);
        }
    }
    if (!owner) return;
#pragma unroll
    for (uint32_t q = 0; q < NS; ++q) {
        fsum += fabs(fx[q]) + fabs(fy[q]) + fabs(fz[q]); // inf or NaN if a force is
        if (sl.w[q] != 0.0) lj_write_grad(sl.a[q], 24. * fx[q], 24. * fy[q], 24. * fz[q], g);
    }
}

// ---- energy and gradient with every pair visited once (22 to 128 atoms) ------
// A group of R lanes (R = 16 up to 64 atoms, 32 up to 128).  The A atoms are
// dealt to R blocks, block I holding atoms I + R j (j < S = ceil(A / R) <= 4);
// lane I keeps block I's positions and force accumulators in registers.  Every
// unordered block pair {I, I + d mod R} is visited once, by lane I at distance
// d = 1..R/2 (at d = R/2 only I < R/2 counts), plus the pairs inside block I.
// The partner block's positions come from LDS; its force accumulator travels:
// at distance d lane I holds the accumulator of block I + d, passes it to lane
// I - 1 before the next distance (which then holds block I + d + 1's), and
// after d = R/2 it goes home from lane I - R/2 to lane I.  Each sum is taken in
// an order fixed by the schedule alone.
// A slot past the last atom (only j = S - 1 can be one) holds weight 0 and a
// point of its own, 2^12 apart from the other such points and far outside the
// problem's bounds (PAD); a real atom placed exactly on one makes the sum NaN,
// which sends the individual to the careful path.
template <uint32_t R>
__device__ __forceinline__ double lj_shfl_row(double v, uint32_t src)
{
    return __shfl(v, static_cast<int>(src), static_cast<int>(R));
}

template <uint32_t R, uint32_t S, bool PAD>
struct lj_block {
    double x[S], y[S], z[S];
    double w = 1.0; // weight of slot S - 1
    // block B's atoms from the individual staged at P
    __device__ __forceinline__ void load(const double *__restrict__ P, uint32_t B, uint32_t atoms)
    {
#pragma unroll
        for (uint32_t j = 0; j < S; ++j) {
            const uint32_t a = B + R * j;
            if (!PAD || j + 1u < S || a < atoms) {
                x[j] = P[3u * a];
                y[j] = P[3u * a + 1u];
                z[j] = P[3u * a + 2u];
            } else {
                x[j] = 4096.0 * static_cast<double>(1u + a);
                y[j] = -4096.0;
                z[j] = 4096.0;
                w = 0.0;
            }
        }
    }
};

// Force factor c (dE/dp_a = 24 c (p_a - p_b)) and energy term s^2 - s of one
// pair, 1/r^2 given.
__device__ __forceinline__ void lj_blk_terms(double inv, double &c, double &e)
{
    const double sv = (inv * inv) * inv;
    c = (fma(-2., sv, 1.) * sv) * inv;
    e = fma(sv, sv, -sv);
}

// The S x S pairs of the own block o (forces f) with the partner block t
// (forces g); energy terms into acc.
template <uint32_t R, uint32_t S, bool PAD>
__device__ __forceinline__ void lj_blk_pairs(const lj_block<R, S, PAD> &o, const lj_block<R, S, PAD> &t, double (&fx)[S],
                                             double (&fy)[S], double (&fz)[S], double (&gx)[S], double (&gy)[S],
                                             double (&gz)[S], double &acc)
{
#pragma unroll
    for (uint32_t j1 = 0; j1 < S; ++j1) {
        double dx[S], dy[S], dz[S], r2[S], inv[S];
#pragma unroll
        for (uint32_t j2 = 0; j2 < S; ++j2) {
            dx[j2] = o.x[j1] - t.x[j2];
            dy[j2] = o.y[j1] - t.y[j2];
            dz[j2] = o.z[j1] - t.z[j2];
            r2[j2] = fma(dz[j2], dz[j2], fma(dy[j2], dy[j2], dx[j2] * dx[j2]));
        }
        lj_recip_group<S>(r2, inv);
#pragma unroll
        for (uint32_t j2 = 0; j2 < S; ++j2) {
            double c, e;
            lj_blk_terms(inv[j2], c, e);
            if constexpr (PAD) {
                if (j1 + 1u == S || j2 + 1u == S) {
                    const double w = j1 + 1u < S ? t.w : (j2 + 1u < S ? o.w : o.w * t.w);
                    c *= w;
                    e *= w;
                }
            }
            fx[j1] = fma(c, dx[j2], fx[j1]);
            fy[j1] = fma(c, dy[j2], fy[j1]);
            fz[j1] = fma(c, dz[j2], fz[j1]);
            gx[j2] = fma(-c, dx[j2], gx[j2]);
            gy[j2] = fma(-c, dy[j2], gy[j2]);
            gz[j2] = fma(-c, dz[j2], gz[j2]);
            acc += e;
        }
    }
}

// Energy and gradient of the individual staged at P (lj_stage with the
// gradient's layout) by the R lanes of a group (lane r); the gradient row is
// written to g.  A coordinate beyond 2^100, or a non-finite energy or force,
// sends the individual to lj_energy_grad_careful.
template <uint32_t R, uint32_t S, bool PAD>
__device__ __forceinline__ double lj_energy_grad_blk(const double *__restrict__ P, uint32_t atoms, int r,
                                                     double *__restrict__ g)
{
    const uint32_t I = static_cast<uint32_t>(r);
    lj_block<R, S, PAD> o;
    o.load(P, I, atoms);
    bool in_range = true;
#pragma unroll
    for (uint32_t j = 0; j < S; ++j)
        in_range = in_range && fabs(o.x[j]) <= 0x1p100 && fabs(o.y[j]) <= 0x1p100 && fabs(o.z[j]) <= 0x1p100;
    double fx[S], fy[S], fz[S], gx[S], gy[S], gz[S];
#pragma unroll
    for (uint32_t j = 0; j < S; ++j) fx[j] = fy[j] = fz[j] = gx[j] = gy[j] = gz[j] = 0.0;
    double acc = 0.0;
    // the pairs inside block I: both forces in the own accumulators
#pragma unroll
    for (uint32_t j1 = 0; j1 < S; ++j1) {
#pragma unroll
        for (uint32_t j2 = j1 + 1u; j2 < S; ++j2) {
            const double dx = o.x[j1] - o.x[j2], dy = o.y[j1] - o.y[j2], dz = o.z[j1] - o.z[j2];
            double r2 = fma(dz, dz, fma(dy, dy, dx * dx)), inv, c, e;
            lj_recip_group<1>(&r2, &inv);
            lj_blk_terms(inv, c, e);
            if constexpr (PAD) {
                if (j2 + 1u == S) {
                    c *= o.w;
                    e *= o.w;
                }
            }
            fx[j1] = fma(c, dx, fx[j1]);
            fy[j1] = fma(c, dy, fy[j1]);
            fz[j1] = fma(c, dz, fz[j1]);
            fx[j2] = fma(-c, dx, fx[j2]);
            fy[j2] = fma(-c, dy, fy[j2]);
            fz[j2] = fma(-c, dz, fz[j2]);
            acc += e;
        }
    }
    const uint32_t next = (I + 1u) % R;
#pragma unroll 1
    for (uint32_t d = 1u; d <= R / 2u; ++d) {
        if (d > 1u) {
#pragma unroll
            for (uint32_t j = 0; j < S; ++j) {
                gx[j] = lj_shfl_row<R>(gx[j], next);
                gy[j] = lj_shfl_row<R>(gy[j], next);
                gz[j] = lj_shfl_row<R>(gz[j], next);
            }
        }
        if (d < R / 2u || I < R / 2u) {
            lj_block<R, S, PAD> t;
            t.load(P, (I + d) % R, atoms);
            lj_blk_pairs<R, S, PAD>(o, t, fx, fy, fz, gx, gy, gz, acc);
        }
    }
    // home: block I's travelling accumulator is at lane I - R/2
    const uint32_t from = (I + R / 2u) % R;
    double fsum = 0.0;
#pragma unroll
    for (uint32_t j = 0; j < S; ++j) {
        const double Fx = fx[j] + lj_shfl_row<R>(gx[j], from);
        const double Fy = fy[j] + lj_shfl_row<R>(gy[j], from);
        const double Fz = fz[j] + lj_shfl_row<R>(gz[j], from);
        fsum += fabs(Fx) + fabs(Fy) + fabs(Fz); // inf or NaN if a force is
        const uint32_t a = I + R * j;
        if (!PAD || a < atoms) lj_write_grad(a, 24. * Fx, 24. * Fy, 24. * Fz, g);
    }
    if (!in_range) acc = __builtin_nan("");
    acc = group_sum<static_cast<int>(R)>(acc);
    fsum = group_sum<static_cast<int>(R)>(fsum);
    const bool fast_ok = acc - acc == 0.0 && fsum - fsum == 0.0;
    if (!fast_ok) return lj_energy_grad_careful<static_cast<int>(R)>(P, atoms, r, g);
    return 4 * acc;
}

// Energy alone on the blocks of lj_energy_grad_blk (no forces, nothing
// travels): the group's sum of s^2 - s over every pair, NaN if a coordinate
// lies beyond 2^100.  The caller tests it and takes the careful path.
template <uint32_t R, uint32_t S, bool PAD>
__device__ __forceinline__ double lj_energy_blk_sum(const double *__restrict__ P, uint32_t atoms, int r)
{
    const uint32_t I = static_cast<uint32_t>(r);
    lj_block<R, S, PAD> o;
    o.load(P, I, atoms);
    bool in_range = true;
#pragma unroll
    for (uint32_t j = 0; j < S; ++j)
        in_range = in_range && fabs(o.x[j]) <= 0x1p100 && fabs(o.y[j]) <= 0x1p100 && fabs(o.z[j]) <= 0x1p100;
    double acc = 0.0;
#pragma unroll
    for (uint32_t j1 = 0; j1 < S; ++j1) {
#pragma unroll
        for (uint32_t j2 = j1 + 1u; j2 < S; ++j2) {
            const double dx = o.x[j1] - o.x[j2], dy = o.y[j1] - o.y[j2], dz = o.z[j1] - o.z[j2];
            double r2 = fma(dz, dz, fma(dy, dy, dx * dx)), inv, c, e;
            lj_recip_group<1>(&r2, &inv);
            lj_blk_terms(inv, c, e);
            if constexpr (PAD) {
                if (j2 + 1u == S) e *= o.w;
            }
            acc += e;
        }
    }
#pragma unroll 1
    for (uint32_t d = 1u; d <= R / 2u; ++d) {
        if (d < R / 2u || I < R / 2u) {
            lj_block<R, S, PAD> t;
            t.load(P, (I + d) % R, atoms);
#pragma unroll
            for (uint32_t j1 = 0; j1 < S; ++j1) {
                double r2[S], inv[S];
#pragma unroll
                for (uint32_t j2 = 0; j2 < S; ++j2) {
                    const double dx = o.x[j1] - t.x[j2], dy = o.y[j1] - t.y[j2], dz = o.z[j1] - t.z[j2];
                    r2[j2] = fma(dz, dz, fma(dy, dy, dx * dx));
                }
                lj_recip_group<S>(r2, inv);
#pragma unroll
                for (uint32_t j2 = 0; j2 < S; ++j2) {
                    const double sv = (inv[j2] * inv[j2]) * inv[j2];
                    double e = fma(sv, sv, -sv);
                    if constexpr (PAD) {
                        if (j1 + 1u == S || j2 + 1u == S) e *= j1 + 1u < S ? t.w : (j2 + 1u < S ? o.w : o.w * t.w);
                    }
                    acc += e;
                }
            }
        }
    }
    if (!in_range) acc = __builtin_nan("");
    return group_sum<static_cast<int>(R)>(acc);
}

// Energy and gradient: lane r owns atoms a = r, r+LA, ... (LA =
// lj_atom_lanes(L); with L = 2 LA the two halves take the steps k = 1..A-1 in
// two ranges and add their forces) and visits every
// other atom a + k, k = 1..A-1, accumulating the force
// dE/dp_a = 24 sum_b (p_a - p_b) (s - 2 s^2) / r^2 (s = r^-6; the factor 24
// is applied to the sum, so that the FP64 constants of the loop are inline
// ones) and the energy terms of all its visits (each pair twice, so the sum
// is halved).  The gradient row (nx doubles) is written to g; returns the
// energy.  A non-finite energy or force (a zero, tiny or NaN distance, see
// lj_pair) sends the individual to lj_energy_grad_careful.
template <int L, uint32_t APL>
__device__ __forceinline__ double lj_energy_grad(const double *__restrict__ P, uint32_t atoms, int r,
                                                 double *__restrict__ g)
{
    if constexpr (L == 16 && APL >= 2u && APL <= 4u) {
        // 16-lane groups, 17 to 64 atoms (lj_lanes): blocks of APL atoms
        if (lj_grad_blocks(atoms)) {
            if (atoms % 16u == 0u) return lj_energy_grad_blk<16u, APL, false>(P, atoms, r, g);
            return lj_energy_grad_blk<16u, APL, true>(P, atoms, r, g);
        }
    }
    if constexpr (L == 32 && APL >= 5u && APL <= 8u) {
        // 32-lane groups, 65 to 128 atoms: blocks of ceil(A / 32) atoms
        constexpr uint32_t S = (APL + 1u) / 2u;
        if (lj_grad_blocks(atoms)) {
            if (atoms % 32u == 0u) return lj_energy_grad_blk<32u, S, false>(P, atoms, r, g);
            return lj_energy_grad_blk<32u, S, true>(P, atoms, r, g);
        }
    }
    constexpr int LA = lj_atom_lanes(L);
    constexpr uint32_t KS = static_cast<uint32_t>(L / LA);
    const uint32_t ra = static_cast<uint32_t>(r) % LA, h = static_cast<uint32_t>(r) / LA;
    const lj_steps st(atoms - 1u, h, KS);
    double acc = 0., fsum = 0.;
    if constexpr (APL > 0) {
        lj_grad_slots<L, LA, APL>(P, atoms, ra, ra, st, h == 0u, acc, fsum, g);
    } else {
        for (uint32_t a0 = ra; a0 < atoms; a0 += 4u * LA) {
            lj_grad_slots<L, LA, 4>(P, atoms, a0, ra, st, h == 0u, acc, fsum, g);
        }
    }
    acc = group_sum<L>(acc);
    fsum = group_sum<L>(fsum);
    const bool fast_ok = acc - acc == 0.0 && fsum - fsum == 0.0;
    if (!fast_ok) return lj_energy_grad_careful<L>(P, atoms, r, g);
    return 2 * acc;
}

} // namespace dev
} // namespace pgh
