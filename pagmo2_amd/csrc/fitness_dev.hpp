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

enum : int { ROSENBROCK = 1, RASTRIGIN = 2, ACKLEY = 3, SCHWEFEL = 4, GRIEWANK = 5, LENNARD_JONES = 6, MO_SPHERE = 7 };

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

// Separable problems: an L-lane group per individual.  Wide rows use a whole
// wavefront so that a row streams as 1 KiB coalesced loads.
__host__ __device__ constexpr int sep_lanes(uint32_t nx)
{
    return nx >= 256u ? 64 : (nx >= 32u ? 16 : 4);
}

// Lennard-Jones: at least ~8 pair evaluations per lane, at most 16 lanes.
__host__ __device__ constexpr int lj_lanes(uint32_t atoms)
{
    const uint32_t pairs = atoms * (atoms - 1u) / 2u;
    const uint32_t want = pairs / 8u;
    return want >= 16u ? 16 : (want >= 8u ? 8 : (want >= 4u ? 4 : (want >= 2u ? 2 : 1)));
}

// Individuals per workgroup for Lennard-Jones: the LDS holds 3*atoms doubles
// per individual and a workgroup uses at most 64 KiB.
__host__ __device__ constexpr uint32_t lj_group_count(uint32_t atoms)
{
    const uint32_t L = static_cast<uint32_t>(lj_lanes(atoms));
    uint32_t g = 256u / L;
    while (g > 1u && g * 24u * atoms > 65536u) g >>= 1;
    return (g * 24u * atoms > 65536u) ? 0u : g;
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
// FAST: rastrigin and ackley take the cosine from cos_pi_reduced and only
// record whether an argument fell outside its range (|y| >= 2^20, inf, NaN);
// the caller then evaluates the whole individual again with FAST = false.
template <int KIND, bool FAST = false>
struct sep_acc {
    double a = 0., b = 0.;
    bool out_of_range = false;
    __device__ __forceinline__ void init()
    {
        a = 0.;
        b = (KIND == GRIEWANK) ? 1. : 0.;
    }
    __device__ __forceinline__ double cosine(double y)
    {
        if constexpr (FAST) {
            out_of_range |= !(fabs(y) < 0x1p20);
            return cos_pi_reduced(y);
        } else {
            return cos_any(y);
        }
    }
    __device__ __forceinline__ void add(double xi, uint32_t i)
    {
        if constexpr (KIND == RASTRIGIN) {
            // rastrigin.cpp:67
            a += xi * xi - 10. * cosine(kOmega * xi);
        } else if constexpr (KIND == ACKLEY) {
            // ackley.cpp:69-70
            a += xi * xi;
            b += cosine(kOmega * xi);
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

template <int KIND, int L, bool ALIGNED, bool FAST, int UNROLL = 1>
__device__ __forceinline__ double separable_eval_impl(const double *__restrict__ x, uint32_t nx, int r,
                                                      bool &out_of_range)
{
    sep_acc<KIND, FAST> acc;
    acc.init();
    const uint32_t npairs = (nx + 1u) >> 1;
    uint32_t c = static_cast<uint32_t>(r);
    if constexpr (ALIGNED && UNROLL == 2) {
        for (; c + L < npairs; c += 2u * L) {
            const double2 v0 = *reinterpret_cast<const double2 *>(x + 2u * c);
            const double2 v1 = *reinterpret_cast<const double2 *>(x + 2u * (c + L));
            acc.add(v0.x, 2u * c);
            acc.add(v0.y, 2u * c + 1u);
            acc.add(v1.x, 2u * (c + L));
            acc.add(v1.y, 2u * (c + L) + 1u);
        }
    }
    if constexpr (ALIGNED && UNROLL == 4) {
        // four 16-byte loads in flight per lane; the pairs are still added in
        // increasing c
        for (; c + 3u * L < npairs; c += 4u * L) {
            const double2 v0 = *reinterpret_cast<const double2 *>(x + 2u * c);
            const double2 v1 = *reinterpret_cast<const double2 *>(x + 2u * (c + L));
            const double2 v2 = *reinterpret_cast<const double2 *>(x + 2u * (c + 2u * L));
            const double2 v3 = *reinterpret_cast<const double2 *>(x + 2u * (c + 3u * L));
            acc.add(v0.x, 2u * c);
            acc.add(v0.y, 2u * c + 1u);
            acc.add(v1.x, 2u * (c + L));
            acc.add(v1.y, 2u * (c + L) + 1u);
            acc.add(v2.x, 2u * (c + 2u * L));
            acc.add(v2.y, 2u * (c + 2u * L) + 1u);
            acc.add(v3.x, 2u * (c + 3u * L));
            acc.add(v3.y, 2u * (c + 3u * L) + 1u);
        }
    }
    for (; c < npairs; c += L) separable_pair<KIND, L, ALIGNED, FAST>(acc, x, nx, c);
    out_of_range = acc.out_of_range;
    const double n = static_cast<double>(nx);
    if constexpr (KIND == RASTRIGIN) {
        // rastrigin.cpp:69
        return group_sum<L>(acc.a) + 10. * n;
    } else if constexpr (KIND == ACKLEY) {
        // ackley.cpp:72-73
        const double s1 = group_sum<L>(acc.a);
        const double s2 = group_sum<L>(acc.b);
        const double nepero = exp(1.0);
        return -20 * exp(-0.2 * sqrt(1.0 / n * s1)) - exp(1.0 / n * s2) + 20 + nepero;
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
// cosines come from cos_pi_reduced; an individual with an argument outside its
// range (|y| >= 2^20, inf, NaN) gets the recompute mark instead, and
// separable_eval_exact produces its value (kept out of this routine so that
// its registers do not weigh on the streaming kernel).
template <int KIND, int L, bool ALIGNED>
__device__ __forceinline__ double separable_eval(const double *__restrict__ x, uint32_t nx, int r)
{
    constexpr bool fast = (KIND == RASTRIGIN || KIND == ACKLEY);
    bool bad = false;
    const double f = separable_eval_impl<KIND, L, ALIGNED, fast>(x, nx, r, bad);
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

// rosenbrock.cpp:83-92, element i of the gradient (m_dim >= 2).
__device__ __forceinline__ double rosenbrock_grad(const double *__restrict__ x, uint32_t m_dim, uint32_t i)
{
    if (i == 0u) {
        return -400. * x[0] * (x[1] - x[0] * x[0]) - 2. * (1 - x[0]);
    }
    if (i == m_dim - 1u) {
        return 200. * (x[m_dim - 1u] - x[m_dim - 2u] * x[m_dim - 2u]);
    }
    return -400. * x[i] * (x[i + 1u] - x[i] * x[i]) - 2. * (1 - x[i]) + 200. * (x[i] - x[i - 1u] * x[i - 1u]);
}

// rastrigin.cpp:103.
__device__ __forceinline__ double rastrigin_grad(double xi)
{
    return 2 * xi + 10.0 * kOmega * sin_any(kOmega * xi);
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

// Stages `count` individuals (rows of nx = 3*atoms-6) into LDS as per-
// individual structure-of-arrays: [x0..x_{A-1}, y0.., z0..].  Called by all
// threads of the workgroup; a __syncthreads() must follow.
__device__ __forceinline__ void lj_stage(const double *__restrict__ rows, uint32_t count, uint32_t atoms,
                                         double *__restrict__ lds)
{
    const uint32_t nx = 3u * atoms - 6u;
    const uint32_t stride = 3u * atoms;
    for (uint32_t t = threadIdx.x; t < count * 6u; t += blockDim.x) {
        // the six fixed coordinates: atom 0 (x,y,z), atom 1 (x,y), atom 2 (x)
        const uint32_t g = t / 6u, f = t % 6u;
        const uint32_t atom = f < 3u ? 0u : (f < 5u ? 1u : 2u);
        const uint32_t coord = f < 3u ? f : (f < 5u ? f - 3u : 0u);
        lds[g * stride + coord * atoms + atom] = 0.0;
    }
    for (uint32_t t = threadIdx.x; t < count * nx; t += blockDim.x) {
        const uint32_t g = t / nx, e = t - g * nx;
        uint32_t atom, coord;
        lj_slot(e, atom, coord);
        lds[g * stride + coord * atoms + atom] = rows[static_cast<uint64_t>(g) * nx + e];
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

// 1 / d for the sixth power of a distance: v_rcp_f64 refined by two Newton
// steps, and the IEEE division where the refinement is not valid (d zero,
// subnormal, infinite or NaN).
__device__ __forceinline__ double lj_recip(double d)
{
    if (d > 0x1p-1000 && d < 0x1p1000) {
        double r = __builtin_amdgcn_rcp(d);
        r = fma(r, fma(-d, r, 1.0), r);
        return fma(r, fma(-d, r, 1.0), r);
    }
    return 1.0 / d;
}

// Cyclic pair schedule of an individual's atoms over the L lanes of its group:
// lane r owns atoms a = r, r+L, ...; step k (1 <= k <= A/2) pairs each owned
// atom a with b = (a + k) mod A, and when A is even the last step keeps only
// a < A/2, so every unordered pair is visited exactly once.  Within a step the
// lanes of a group read consecutive atoms.
struct lj_cycle {
    uint32_t atoms, half;
    __device__ __forceinline__ explicit lj_cycle(uint32_t a) : atoms(a), half(a / 2u) {}
    __device__ __forceinline__ bool counts(uint32_t a, uint32_t k) const
    {
        return 2u * k < atoms || a < half; // k == A/2 with A even: one of the two visits
    }
    __device__ __forceinline__ uint32_t partner(uint32_t a, uint32_t k) const
    {
        const uint32_t b = a + k;
        return b >= atoms ? b - atoms : b;
    }
};

// The reference's dist == 0 rule (lennard_jones.cpp:81-83,89): the result is
// +inf unless a NaN term comes after the last zero-distance pair in the
// reference's pair order, then NaN.  Only reached when some pair of the
// individual has zero distance; recomputes the group's terms to find the NaNs.
template <int L>
__device__ __noinline__ double lj_zero_rule(const double *__restrict__ X, uint32_t atoms, int r, int zlast)
{
    const double *__restrict__ Y = X + atoms;
    const double *__restrict__ Z = X + 2u * atoms;
    const lj_cycle cyc(atoms);
    int nlast = -1;
    for (uint32_t k = 1u; k <= cyc.half; ++k) {
        for (uint32_t a = static_cast<uint32_t>(r); a < atoms; a += L) {
            if (!cyc.counts(a, k)) continue;
            const uint32_t b = cyc.partner(a, k);
            const double dx = X[a] - X[b], dy = Y[a] - Y[b], dz = Z[a] - Z[b];
            const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
            if (r2 == 0.0) continue;
            const double s = lj_recip(r2 * r2 * r2);
            const double t = fma(s, s, -s);
            if (t != t) nlast = max(nlast, lj_pair_index(a, b, atoms));
        }
    }
    nlast = group_max<L>(nlast);
    return nlast > zlast ? __builtin_nan("") : __builtin_inf();
}

// Lane r of a group owns atoms a_q = r + q L.  APL > 0: exactly APL owned
// slots (the ceiling of atoms / L), all held in registers; APL == 0: any
// count, taken four slots at a time.  Owned slots beyond the last atom are
// masked.  Processing a lane's slots together gives each step of the
// schedule several independent chains of arithmetic.
__host__ __device__ constexpr uint32_t lj_apl(uint32_t atoms, int L)
{
    const uint32_t apl = (atoms + static_cast<uint32_t>(L) - 1u) / static_cast<uint32_t>(L);
    return apl <= 4u ? apl : 0u;
}

template <int L, uint32_t NS>
struct lj_slots {
    uint32_t a[NS];
    bool own[NS];
    double x[NS], y[NS], z[NS];
    __device__ __forceinline__ void load(const double *__restrict__ X, const double *__restrict__ Y,
                                         const double *__restrict__ Z, uint32_t a0, uint32_t atoms)
    {
#pragma unroll
        for (uint32_t q = 0; q < NS; ++q) {
            a[q] = a0 + q * L;
            own[q] = a[q] < atoms;
            if (!own[q]) a[q] = a0;
            x[q] = X[a[q]];
            y[q] = Y[a[q]];
            z[q] = Z[a[q]];
        }
    }
};

template <int L, uint32_t NS>
__device__ __forceinline__ void lj_energy_slots(const double *__restrict__ X, const double *__restrict__ Y,
                                                const double *__restrict__ Z, uint32_t atoms, uint32_t a0,
                                                double &acc, int &zlast)
{
    const lj_cycle cyc(atoms);
    lj_slots<L, NS> sl;
    sl.load(X, Y, Z, a0, atoms);
    for (uint32_t k = 1u; k <= cyc.half; ++k) {
#pragma unroll
        for (uint32_t q = 0; q < NS; ++q) {
            const uint32_t b = cyc.partner(sl.a[q], k);
            const double dx = sl.x[q] - X[b], dy = sl.y[q] - Y[b], dz = sl.z[q] - Z[b];
            const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
            const bool valid = sl.own[q] && cyc.counts(sl.a[q], k);
            const double s = lj_recip(r2 * r2 * r2);
            const double t = fma(s, s, -s);
            acc += (valid && r2 != 0.0) ? t : 0.0;
            if (valid && r2 == 0.0) zlast = max(zlast, lj_pair_index(sl.a[q], b, atoms));
        }
    }
}

// Energy of the individual staged at X (Y = X + A, Z = X + 2A) by the L lanes
// of a group (lane r): sum of 4 (s^2 - s), s = r^-6, over the cyclic
// schedule, then the group's butterfly sum.
template <int L, uint32_t APL>
__device__ __forceinline__ double lj_energy(const double *__restrict__ X, uint32_t atoms, int r)
{
    const double *__restrict__ Y = X + atoms;
    const double *__restrict__ Z = X + 2u * atoms;
    double acc = 0.;
    int zlast = -1;
    if constexpr (APL > 0) {
        lj_energy_slots<L, APL>(X, Y, Z, atoms, static_cast<uint32_t>(r), acc, zlast);
    } else {
        for (uint32_t a0 = static_cast<uint32_t>(r); a0 < atoms; a0 += 4u * L) {
            lj_energy_slots<L, 4>(X, Y, Z, atoms, a0, acc, zlast);
        }
    }
    acc = group_sum<L>(acc);
    zlast = group_max<L>(zlast);
    if (zlast >= 0) return lj_zero_rule<L>(X, atoms, r, zlast);
    return 4 * acc;
}

template <int L, uint32_t NS>
__device__ __forceinline__ void lj_grad_slots(const double *__restrict__ X, const double *__restrict__ Y,
                                              const double *__restrict__ Z, uint32_t atoms, uint32_t a0,
                                              double &acc, int &zlast, double *__restrict__ g)
{
    const lj_cycle cyc(atoms);
    lj_slots<L, NS> sl;
    sl.load(X, Y, Z, a0, atoms);
    double gx[NS], gy[NS], gz[NS];
#pragma unroll
    for (uint32_t q = 0; q < NS; ++q) gx[q] = gy[q] = gz[q] = 0.;
    for (uint32_t k = 1u; k < atoms; ++k) {
#pragma unroll
        for (uint32_t q = 0; q < NS; ++q) {
            const uint32_t b = cyc.partner(sl.a[q], k);
            const double dx = sl.x[q] - X[b], dy = sl.y[q] - Y[b], dz = sl.z[q] - Z[b];
            const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
            const double s = lj_recip(r2 * r2 * r2);
            const double inv_r2 = (s * r2) * r2;
            const double coef = (fma(-48., s, 24.) * s) * inv_r2;
            gx[q] = fma(coef, dx, gx[q]);
            gy[q] = fma(coef, dy, gy[q]);
            gz[q] = fma(coef, dz, gz[q]);
            const bool valid = sl.own[q] && k <= cyc.half && cyc.counts(sl.a[q], k);
            const double t = fma(s, s, -s);
            acc += (valid && r2 != 0.0) ? t : 0.0;
            if (valid && r2 == 0.0) zlast = max(zlast, lj_pair_index(sl.a[q], b, atoms));
        }
    }
#pragma unroll
    for (uint32_t q = 0; q < NS; ++q) {
        if (!sl.own[q]) continue;
        const uint32_t a = sl.a[q];
        if (a == 1u) {
            g[0] = gz[q];
        } else if (a == 2u) {
            g[1] = gy[q];
            g[2] = gz[q];
        } else if (a >= 3u) {
            const uint32_t o = 3u * (a - 2u);
            g[o] = gx[q];
            g[o + 1u] = gy[q];
            g[o + 2u] = gz[q];
        }
    }
}

// Energy and gradient: lane r owns atoms a = r, r+L, ... and visits every
// other atom b = (a + k) mod A, k = 1..A-1, accumulating the force
// dE/dp_a = sum_b (p_a - p_b) (24 s - 48 s^2) / r^2 (s = r^-6; 1/r^2 = s r^4)
// and, on the steps the cyclic schedule counts, the energy terms.  The
// gradient row (nx doubles) is written to g; returns the energy.
template <int L, uint32_t APL>
__device__ __forceinline__ double lj_energy_grad(const double *__restrict__ X, uint32_t atoms, int r,
                                                 double *__restrict__ g)
{
    const double *__restrict__ Y = X + atoms;
    const double *__restrict__ Z = X + 2u * atoms;
    double acc = 0.;
    int zlast = -1;
    if constexpr (APL > 0) {
        lj_grad_slots<L, APL>(X, Y, Z, atoms, static_cast<uint32_t>(r), acc, zlast, g);
    } else {
        for (uint32_t a0 = static_cast<uint32_t>(r); a0 < atoms; a0 += 4u * L) {
            lj_grad_slots<L, 4>(X, Y, Z, atoms, a0, acc, zlast, g);
        }
    }
    acc = group_sum<L>(acc);
    zlast = group_max<L>(zlast);
    if (zlast >= 0) return lj_zero_rule<L>(X, atoms, r, zlast);
    return 4 * acc;
}

} // namespace dev
} // namespace pgh
