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
template <int KIND>
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
            // rastrigin.cpp:67
            a += xi * xi - 10. * cos_any(kOmega * xi);
        } else if constexpr (KIND == ACKLEY) {
            // ackley.cpp:69-70
            a += xi * xi;
            b += cos_any(kOmega * xi);
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

template <int KIND, int L, bool ALIGNED>
__device__ __forceinline__ void separable_pair(sep_acc<KIND> &acc, const double *__restrict__ x, uint32_t nx,
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

template <int KIND, int L, bool ALIGNED>
__device__ __forceinline__ double separable_eval(const double *__restrict__ x, uint32_t nx, int r)
{
    sep_acc<KIND> acc;
    acc.init();
    const uint32_t npairs = (nx + 1u) >> 1;
    uint32_t c = static_cast<uint32_t>(r);
    if constexpr (ALIGNED) {
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
    for (; c < npairs; c += L) separable_pair<KIND, L, ALIGNED>(acc, x, nx, c);
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

// Combines per-lane partial sums with the reference's dist == 0 semantics:
// a zero distance overwrites the accumulator with DBL_MAX and later finite
// terms cannot bring it back, so the result is +inf after the final *4,
// unless a NaN term follows the last zero pair (then NaN).
template <int L>
__device__ __forceinline__ double lj_finish(double acc, int zlast, int nlast)
{
    acc = group_sum<L>(acc);
    zlast = group_max<L>(zlast);
    nlast = group_max<L>(nlast);
    if (zlast >= 0) {
        return nlast > zlast ? __builtin_nan("") : __builtin_inf();
    }
    return 4 * acc;
}

// Energy of the individual staged at X/Y/Z (= X + atoms, X + 2*atoms); lane r
// of L takes a contiguous slice of the pair sequence.
template <int L>
__device__ __forceinline__ double lj_energy(const double *__restrict__ X, uint32_t atoms, int r)
{
    const double *__restrict__ Y = X + atoms;
    const double *__restrict__ Z = X + 2u * atoms;
    const uint32_t P = atoms * (atoms - 1u) / 2u;
    const uint32_t k0 = static_cast<uint32_t>((static_cast<uint64_t>(P) * r) / L);
    const uint32_t k1 = static_cast<uint32_t>((static_cast<uint64_t>(P) * (r + 1)) / L);
    uint32_t i = 0u, rem = k0;
    while (i + 1u < atoms && rem >= atoms - 1u - i) {
        rem -= atoms - 1u - i;
        ++i;
    }
    uint32_t j = i + 1u + rem;
    double acc = 0.;
    int zlast = -1, nlast = -1;
    double xi = X[i], yi = Y[i], zi = Z[i];
    for (uint32_t k = k0; k < k1; ++k) {
        const double dx = xi - X[j], dy = yi - Y[j], dz = zi - Z[j];
        const double dist = dx * dx + dy * dy + dz * dz;
        if (dist == 0.0) {
            zlast = static_cast<int>(k);
        } else {
            const double inv = 1.0 / dist;
            const double sixth = inv * inv * inv;
            const double t = sixth * sixth - sixth;
            if (t != t) nlast = static_cast<int>(k);
            acc += t;
        }
        if (++j == atoms) {
            ++i;
            j = i + 1u;
            if (i < atoms) {
                xi = X[i];
                yi = Y[i];
                zi = Z[i];
            }
        }
    }
    return lj_finish<L>(acc, zlast, nlast);
}

// Energy and gradient: lane r owns atoms a = r, r+L, ...; each owned atom
// sums the force of every other atom (dE/dp_a = sum_j (p_a - p_j) *
// (24 s - 48 s^2) / r^2, s = r^-6) and the pairs (a, j>a) of the energy.
// The gradient row (nx doubles) is written to g; returns the energy.
template <int L>
__device__ __forceinline__ double lj_energy_grad(const double *__restrict__ X, uint32_t atoms, int r,
                                                 double *__restrict__ g)
{
    const double *__restrict__ Y = X + atoms;
    const double *__restrict__ Z = X + 2u * atoms;
    double acc = 0.;
    int zlast = -1, nlast = -1;
    for (uint32_t a = static_cast<uint32_t>(r); a < atoms; a += L) {
        const double xa = X[a], ya = Y[a], za = Z[a];
        double gx = 0., gy = 0., gz = 0.;
        const uint32_t base = lj_row_start(a, atoms) - a - 1u;
        for (uint32_t j = 0u; j < atoms; ++j) {
            if (j == a) continue;
            const double dx = xa - X[j], dy = ya - Y[j], dz = za - Z[j];
            const double dist = dx * dx + dy * dy + dz * dz;
            const double inv = 1.0 / dist;
            const double sixth = inv * inv * inv;
            const double coef = (24. * sixth - 48. * sixth * sixth) * inv;
            gx += coef * dx;
            gy += coef * dy;
            gz += coef * dz;
            if (j > a) {
                const int k = static_cast<int>(base + j);
                if (dist == 0.0) {
                    zlast = max(zlast, k);
                } else {
                    const double t = sixth * sixth - sixth;
                    if (t != t) nlast = max(nlast, k);
                    acc += t;
                }
            }
        }
        if (a == 1u) {
            g[0] = gz;
        } else if (a == 2u) {
            g[1] = gy;
            g[2] = gz;
        } else if (a >= 3u) {
            const uint32_t o = 3u * (a - 2u);
            g[o] = gx;
            g[o + 1u] = gy;
            g[o + 2u] = gz;
        }
    }
    return lj_finish<L>(acc, zlast, nlast);
}

} // namespace dev
} // namespace pgh
