// oracle.hpp -- CPU restatement of pagmo 2.19.1's batch-fitness-evaluation path.
//
// TEST INFRASTRUCTURE ONLY.  This header is the checker that the HIP library
// (pagmo2_amd/libpagmo_hip.so) is compared against.  Only tests/, the smoke
// check in __graft_entry__.py and the cpu_baseline leg of bench.py may load it;
// the product path never links, calls or falls back to anything in oracle/.
//
// Every function restates the reference algorithm it cites (file:line under the
// pagmo source tree) with the same operation order, so that on the same host
// libm it reproduces the reference bit for bit.  Compile with
// -ffp-contract=off and without -ffast-math (the reference is built that way:
// generic x86-64, no FMA contraction).
//
// How each part is pinned (see tests/test_oracle.py):
//   * UDP fitness/gradients: the reference's inline known answers
//     (tests/{rosenbrock,rastrigin,ackley,schwefel,griewank,lennard_jones}.cpp)
//     plus closed forms at best_known() and brute-force identities.
//   * estimate_gradient(_h): tests/gradients_and_hessians.cpp:144-152 and
//     analytic gradients of polynomials.
//   * pareto_dominance / fnds: tests/multi_objective.cpp:44-144 plus a
//     brute-force longest-chain check on random inputs.
//   * MT19937: the C++ standard's 10000th-output value (4123659995) and
//     libstdc++'s generate_canonical source (bits/random.tcc).
//   * LJ gradient (no reference implementation): central finite differences
//     and estimate_gradient_h.
//   * crowding_distance / select_best_N_mo / sort_population_mo:
//     tests/multi_objective.cpp:147-262 plus brute-force crowding on ties.
//   * local_descent (the memetic step, no reference implementation):
//     hand-derived traces (rastrigin from 0.1, rosenbrock from the origin)
//     and invariants (f never increases, f is the fitness of x, bounds hold).
//   * pso_gen trajectories: PARITY UNPINNED by the reference (its tests only
//     check self-consistency, tests/pso_gen.cpp:154-240,302-336); pinned here by
//     the same self-consistency checks plus hand-checked single steps.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <utility>
#include <numeric>
#include <vector>

namespace oracle
{

using vector_double = std::vector<double>;
using pop_size_t = std::vector<double>::size_type;

// include/pagmo/detail/constants.hpp:37-39
constexpr double pi()
{
    return 3.1415926535897932384626433832795028841971693993;
}

// Problem kinds; numbering shared with the C-ABI (pgh_udp_kind).
enum kind_t : int {
    ROSENBROCK = 1,
    RASTRIGIN = 2,
    ACKLEY = 3,
    SCHWEFEL = 4,
    GRIEWANK = 5,
    LENNARD_JONES = 6,
    MO_SPHERE = 7,
    ZDT = 8,
    DTLZ = 9,
};

// ---------------------------------------------------------------------------
// UDP fitness functions
// ---------------------------------------------------------------------------

// src/problems/rosenbrock.cpp:59-66 (loop bound is m_dim, not x.size()).
inline vector_double rosenbrock_fitness(std::size_t m_dim, const vector_double &x)
{
    double retval = 0.;
    for (std::size_t i = 0u; i < m_dim - 1u; ++i) {
        retval += 100. * (x[i] * x[i] - x[i + 1]) * (x[i] * x[i] - x[i + 1]) + (x[i] - 1) * (x[i] - 1);
    }
    return {retval};
}

// src/problems/rastrigin.cpp:62-72 (n = x.size()).
inline vector_double rastrigin_fitness(const vector_double &x)
{
    vector_double f(1, 0.);
    const auto omega = 2. * pi();
    auto n = x.size();
    for (decltype(n) i = 0u; i < n; ++i) {
        f[0] += x[i] * x[i] - 10. * std::cos(omega * x[i]);
    }
    f[0] += 10. * static_cast<double>(n);
    return f;
}

// src/problems/ackley.cpp:61-76.
inline vector_double ackley_fitness(const vector_double &x)
{
    vector_double f(1, 0.);
    auto n = x.size();
    double omega = 2. * pi();
    double s1 = 0., s2 = 0.;
    double nepero = std::exp(1.0);
    for (decltype(n) i = 0u; i < n; i++) {
        s1 += x[i] * x[i];
        s2 += std::cos(omega * x[i]);
    }
    f[0] = -20 * std::exp(-0.2 * std::sqrt(1.0 / static_cast<double>(n) * s1))
           - std::exp(1.0 / static_cast<double>(n) * s2) + 20 + nepero;
    return f;
}

// src/problems/schwefel.cpp:60-69.
inline vector_double schwefel_fitness(const vector_double &x)
{
    vector_double f(1, 0.);
    auto n = x.size();
    for (decltype(n) i = 0u; i < n; i++) {
        f[0] += x[i] * std::sin(std::sqrt(std::abs(x[i])));
    }
    f[0] = 418.9828872724338 * static_cast<double>(n) - f[0];
    return f;
}

// src/problems/griewank.cpp:60-76 (sum pass, then product pass).
inline vector_double griewank_fitness(const vector_double &x)
{
    vector_double f(1, 0.);
    auto n = x.size();
    double fr = 4000.;
    double retval = 0.;
    double p = 1.;
    for (decltype(n) i = 0u; i < n; i++) {
        retval += x[i] * x[i];
    }
    for (decltype(n) i = 0u; i < n; i++) {
        p *= std::cos(x[i] / std::sqrt(static_cast<double>(i) + 1.0));
    }
    f[0] = (retval / fr - p + 1.);
    return f;
}

// src/problems/lennard_jones.cpp:132-151: decision vector -> atom coordinate.
inline double lj_r(unsigned atom, unsigned coord, const vector_double &x)
{
    if (atom == 0u) {
        return 0.0;
    } else if (atom == 1u) {
        return coord < 2u ? 0.0 : x[0];
    } else if (atom == 2u) {
        return coord == 0u ? 0.0 : x[coord];
    }
    return x[3u * (atom - 2u) + coord];
}

// src/problems/lennard_jones.cpp:72-91.  dist == 0 overwrites the accumulator
// with DBL_MAX and the loop continues; the final *4 overflows to +inf.
inline vector_double lennard_jones_fitness(unsigned m_atoms, const vector_double &x)
{
    vector_double f(1, 0.);
    for (unsigned i = 0u; i < (m_atoms - 1u); ++i) {
        for (unsigned j = (i + 1u); j < m_atoms; ++j) {
            double sixth, dist;
            dist = std::pow(lj_r(i, 0u, x) - lj_r(j, 0u, x), 2) + std::pow(lj_r(i, 1u, x) - lj_r(j, 1u, x), 2)
                   + std::pow(lj_r(i, 2u, x) - lj_r(j, 2u, x), 2);
            if (dist == 0.0) {
                f[0] = std::numeric_limits<double>::max();
            } else {
                sixth = std::pow(dist, -3);
                f[0] += (std::pow(sixth, 2) - sixth);
            }
        }
    }
    f[0] = 4 * f[0];
    return f;
}

// ZDT1-6 (src/problems/zdt.cpp:233-356), restated in the reference's order.
inline vector_double zdt_fitness(unsigned id, const vector_double &x)
{
    const double pi = 3.1415926535897932384626433832795028841971693993;
    vector_double f(2, 0.);
    const auto N = x.size();
    if (id == 5u) {
        // zdt.cpp:298-339
        const auto n_vectors = ((N - 30u) / 5u) + 1u;
        std::vector<std::size_t> u(n_vectors, 0u), v(n_vectors);
        vector_double r;
        for (double item : x) r.push_back(std::round(item));
        u[0] = static_cast<std::size_t>(std::count(r.begin(), r.begin() + 30, 1.));
        std::size_t k = 30u;
        for (std::size_t i = 1u; i < n_vectors; ++i) {
            for (int j = 0; j < 5; ++j) {
                if (r[k] == 1.) ++u[i];
                ++k;
            }
        }
        f[0] = 1.0 + static_cast<double>(u[0]);
        for (std::size_t i = 1u; i < n_vectors; ++i) v[i] = u[i] < 5u ? 2u + u[i] : 1u;
        double g = 0.;
        for (std::size_t i = 1u; i < n_vectors; ++i) g += static_cast<double>(v[i]);
        f[1] = g * (1. / f[0]);
        return f;
    }
    double g = 0.;
    f[0] = x[0];
    if (id == 4u) {
        // zdt.cpp:281-296
        g = 1 + 10 * static_cast<double>(N - 1u);
        for (std::size_t i = 1u; i < N; ++i) g += x[i] * x[i] - 10. * std::cos(4. * pi * x[i]);
        f[1] = g * (1. - std::sqrt(x[0] / g));
        return f;
    }
    for (std::size_t i = 1u; i < N; ++i) g += x[i];
    if (id == 6u) {
        // zdt.cpp:341-356
        f[0] = 1 - std::exp(-4 * x[0]) * std::pow(std::sin(6 * pi * x[0]), 6);
        g = 1 + 9 * std::pow((g / static_cast<double>(N - 1u)), 0.25);
        f[1] = g * (1 - (f[0] / g) * (f[0] / g));
        return f;
    }
    g = 1. + (9. * g) / static_cast<double>(N - 1u);
    if (id == 1u) {
        f[1] = g * (1. - std::sqrt(x[0] / g));
    } else if (id == 2u) {
        f[1] = g * (1. - (x[0] / g) * (x[0] / g));
    } else {
        f[1] = g * (1. - std::sqrt(x[0] / g) - x[0] / g * std::sin(10. * pi * x[0]));
    }
    return f;
}

// DTLZ1-7 (src/problems/dtlz.cpp:181-407), restated in the reference's order
// with its std::pow calls.
inline double dtlz_g(unsigned id, const vector_double &xm)
{
    const double pi = 3.1415926535897932384626433832795028841971693993;
    double y = 0.;
    switch (id) {
        case 1u:
        case 3u:
            for (std::size_t i = 0u; i < xm.size(); ++i) {
                y += std::pow(xm[i] - 0.5, 2) - std::cos(20. * pi * (xm[i] - 0.5));
            }
            return 100. * (y + static_cast<double>(xm.size()));
        case 6u:
            for (std::size_t i = 0u; i < xm.size(); ++i) y += std::pow(xm[i], 0.1);
            return y;
        case 7u:
            for (std::size_t i = 0u; i < xm.size(); ++i) y += xm[i];
            return (9. / static_cast<double>(xm.size())) * y;
        default:
            for (std::size_t i = 0u; i < xm.size(); ++i) y += std::pow(xm[i] - 0.5, 2);
            return y;
    }
}

inline vector_double dtlz_fitness(unsigned id, std::size_t M, unsigned alpha, const vector_double &x)
{
    const double pi = 3.1415926535897932384626433832795028841971693993;
    const double pi_half = 3.1415926535897932384626433832795028841971693993 / 2.;
    vector_double f(M);
    const vector_double xm(x.begin() + static_cast<std::ptrdiff_t>(M - 1u), x.end());
    if (id == 7u) {
        const double g = 1. + dtlz_g(id, xm);
        for (std::size_t i = 0u; i < M - 1u; ++i) f[i] = x[i];
        double y = 0.;
        for (std::size_t i = 0u; i < M - 1u; ++i) y += (f[i] / (1.0 + g)) * (1.0 + std::sin(3 * pi * f[i]));
        f[M - 1u] = (1. + g) * (static_cast<double>(M) - y);
        return f;
    }
    const double g = dtlz_g(id, xm);
    if (id == 1u) {
        f[0] = 0.5 * (1. + g);
        for (std::size_t i = 0u; i < M - 1u; ++i) f[0] *= x[i];
        for (std::size_t i = 1u; i < M - 1u; ++i) {
            f[i] = 0.5 * (1.0 + g);
            for (std::size_t j = 0u; j < M - (i + 1); ++j) f[i] *= x[j];
            f[i] *= 1. - x[M - (i + 1u)];
        }
        f[M - 1u] = 0.5 * (1. - x[0]) * (1. + g);
        return f;
    }
    vector_double v(M);
    if (id == 5u || id == 6u) {
        v[0] = x[0];
        const double t = 1. / (2. * (1. + g));
        for (std::size_t i = 1u; i < M; ++i) v[i] = t + ((g * x[i]) / (1.0 + g));
    } else {
        for (std::size_t i = 0u; i < M; ++i) v[i] = id == 4u ? std::pow(x[i], alpha) : x[i];
    }
    f[0] = (1. + g);
    for (std::size_t i = 0u; i < M - 1u; ++i) f[0] *= std::cos(v[i] * pi_half);
    for (std::size_t i = 1u; i < M - 1u; ++i) {
        f[i] = (1. + g);
        for (std::size_t j = 0u; j < M - (i + 1u); ++j) f[i] *= std::cos(v[j] * pi_half);
        f[i] *= std::sin(v[M - (i + 1u)] * pi_half);
    }
    f[M - 1u] = (1. + g) * std::sin(v[0] * pi_half);
    return f;
}

// Synthetic multi-objective problem of the ranking workload:
// f_k(x) = sum_i (x_i - c_{k,i})^2, accumulated left to right from 0.
// Not a pagmo UDP (a user-defined problem of the benchmark).
inline vector_double mo_sphere_fitness(const vector_double &centers, unsigned nobj, const vector_double &x)
{
    const auto n = x.size();
    vector_double f(nobj, 0.);
    for (unsigned k = 0; k < nobj; ++k) {
        double acc = 0.;
        for (std::size_t i = 0; i < n; ++i) {
            const double d = x[i] - centers[k * n + i];
            acc += d * d;
        }
        f[k] = acc;
    }
    return f;
}

// ---------------------------------------------------------------------------
// Analytic gradients
// ---------------------------------------------------------------------------

// src/problems/rosenbrock.cpp:83-92.
inline vector_double rosenbrock_gradient(std::size_t m_dim, const vector_double &x)
{
    vector_double retval(m_dim);
    retval[0] = -400. * x[0] * (x[1] - x[0] * x[0]) - 2. * (1 - x[0]);
    for (unsigned i = 1; i < m_dim - 1u; ++i) {
        retval[i] = -400. * x[i] * (x[i + 1u] - x[i] * x[i]) - 2. * (1 - x[i]) + 200. * (x[i] - x[i - 1u] * x[i - 1u]);
    }
    retval[m_dim - 1u] = 200. * (x[m_dim - 1u] - x[m_dim - 2u] * x[m_dim - 2u]);
    return retval;
}

// src/problems/rastrigin.cpp:98-107.
inline vector_double rastrigin_gradient(const vector_double &x)
{
    auto n = x.size();
    vector_double g(n);
    const auto omega = 2. * pi();
    for (decltype(n) i = 0u; i < n; ++i) {
        g[i] = 2 * x[i] + 10.0 * omega * std::sin(omega * x[i]);
    }
    return g;
}

// Lennard-Jones gradient.  The reference has none (include/pagmo/problems/
// lennard_jones.hpp:58-88); this is the derivative of lennard_jones_fitness:
//   dE/dp_i = sum_{j != i} (p_i - p_j) * (24 s - 48 s^2) / r^2,  s = r^-6,
// mapped back through the encoding of lj_r.  Pinned only by finite
// differences and estimate_gradient_h.
inline vector_double lennard_jones_gradient(unsigned m_atoms, const vector_double &x)
{
    std::vector<double> p(3u * m_atoms), g(3u * m_atoms, 0.);
    for (unsigned a = 0; a < m_atoms; ++a) {
        for (unsigned c = 0; c < 3u; ++c) {
            p[3u * a + c] = lj_r(a, c, x);
        }
    }
    for (unsigned i = 0u; i < m_atoms - 1u; ++i) {
        for (unsigned j = i + 1u; j < m_atoms; ++j) {
            const double dx = p[3u * i] - p[3u * j];
            const double dy = p[3u * i + 1u] - p[3u * j + 1u];
            const double dz = p[3u * i + 2u] - p[3u * j + 2u];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double inv = 1. / r2;
            const double s = inv * inv * inv;
            const double coef = (24. * s - 48. * s * s) * inv;
            g[3u * i] += coef * dx;
            g[3u * i + 1u] += coef * dy;
            g[3u * i + 2u] += coef * dz;
            g[3u * j] -= coef * dx;
            g[3u * j + 1u] -= coef * dy;
            g[3u * j + 2u] -= coef * dz;
        }
    }
    vector_double out(x.size(), 0.);
    out[0] = g[3u * 1u + 2u];
    out[1] = g[3u * 2u + 1u];
    out[2] = g[3u * 2u + 2u];
    for (unsigned a = 3u; a < m_atoms; ++a) {
        for (unsigned c = 0; c < 3u; ++c) {
            out[3u * (a - 2u) + c] = g[3u * a + c];
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// Problem descriptor (the subset of pagmo::problem this path needs)
// ---------------------------------------------------------------------------

struct udp {
    int kind = ROSENBROCK;
    unsigned dim = 2;      // rosenbrock m_dim / rastrigin..griewank m_dim / mo_sphere nx
    unsigned atoms = 3;    // lennard_jones m_atoms
    unsigned nobj = 1;     // mo_sphere / dtlz number of objectives
    vector_double centers; // mo_sphere centres, nobj x dim row-major
    unsigned prob_id = 0;  // zdt / dtlz problem number
    unsigned alpha = 100;  // dtlz alpha (dtlz.hpp:105-106)

    std::size_t get_nx() const
    {
        if (kind == LENNARD_JONES) return 3u * atoms - 6u;
        if (kind == ZDT && prob_id == 5u) return 30u + 5u * (dim - 1u); // zdt.cpp:126-129
        return dim;
    }
    std::size_t get_nf() const
    {
        return kind == MO_SPHERE || kind == DTLZ ? nobj : (kind == ZDT ? 2u : 1u);
    }
    // Box bounds: rosenbrock.cpp:72-75, rastrigin.cpp:80-85, ackley.cpp:85-90,
    // schwefel.cpp:77-82, griewank.cpp:84-89, lennard_jones.cpp:99-111.
    std::pair<vector_double, vector_double> get_bounds() const
    {
        const auto nx = get_nx();
        switch (kind) {
            case ROSENBROCK:
                return {vector_double(nx, -5.), vector_double(nx, 10.)};
            case RASTRIGIN:
                return {vector_double(nx, -5.12), vector_double(nx, 5.12)};
            case ACKLEY:
                return {vector_double(nx, -15), vector_double(nx, 30)};
            case SCHWEFEL:
                return {vector_double(nx, -500), vector_double(nx, 500)};
            case GRIEWANK:
                return {vector_double(nx, -600), vector_double(nx, 600)};
            case LENNARD_JONES: {
                vector_double lb(nx, -3.0), ub(nx, 3.0);
                for (unsigned i = 0u; i < nx; ++i) {
                    if ((i != 0) && (i % 3) == 0) {
                        lb[i] = 0.0;
                        ub[i] = 6.0;
                    }
                }
                return {lb, ub};
            }
            case ZDT:
                if (prob_id == 4u) { // zdt.cpp:118-125
                    vector_double lb(nx, -5.), ub(nx, 5.);
                    lb[0] = 0.0;
                    ub[0] = 1.0;
                    return {lb, ub};
                }
                return {vector_double(nx, 0.), vector_double(nx, 1.)};
            default:
                return {vector_double(nx, 0.), vector_double(nx, 1.)};
        }
    }
    vector_double fitness(const vector_double &x) const
    {
        switch (kind) {
            case ROSENBROCK:
                return rosenbrock_fitness(dim, x);
            case RASTRIGIN:
                return rastrigin_fitness(x);
            case ACKLEY:
                return ackley_fitness(x);
            case SCHWEFEL:
                return schwefel_fitness(x);
            case GRIEWANK:
                return griewank_fitness(x);
            case LENNARD_JONES:
                return lennard_jones_fitness(atoms, x);
            case MO_SPHERE:
                return mo_sphere_fitness(centers, nobj, x);
            case ZDT:
                return zdt_fitness(prob_id, x);
            case DTLZ:
                return dtlz_fitness(prob_id, nobj, alpha, x);
        }
        throw std::invalid_argument("unknown problem kind");
    }
    bool has_gradient() const
    {
        return kind == ROSENBROCK || kind == RASTRIGIN || kind == LENNARD_JONES;
    }
    vector_double gradient(const vector_double &x) const
    {
        switch (kind) {
            case ROSENBROCK:
                return rosenbrock_gradient(dim, x);
            case RASTRIGIN:
                return rastrigin_gradient(x);
            case LENNARD_JONES:
                return lennard_jones_gradient(atoms, x);
        }
        throw std::invalid_argument("the problem has no gradient");
    }
};

// ---------------------------------------------------------------------------
// thread_bfe restatement (the CPU baseline)
// ---------------------------------------------------------------------------

// src/detail/bfe_impl.cpp:63-88 (the per-dv check only compares lengths,
// src/problem.cpp:937-950).
inline void bfe_check_input_dvs(const udp &p, const vector_double &dvs)
{
    const auto n_dim = p.get_nx();
    if (dvs.size() % n_dim) {
        throw std::invalid_argument("Invalid argument for a batch fitness evaluation: the length of the vector "
                                    "representing the decision vectors, "
                                    + std::to_string(dvs.size())
                                    + ", is not an exact multiple of the dimension of the problem, "
                                    + std::to_string(n_dim));
    }
}

// src/detail/bfe_impl.cpp:92-129.
inline void bfe_check_output_fvs(const udp &p, const vector_double &dvs, const vector_double &fvs)
{
    const auto n_dvs = dvs.size() / p.get_nx();
    const auto f_dim = p.get_nf();
    const auto n_fvs = fvs.size() / f_dim;
    if (fvs.size() % f_dim) {
        throw std::invalid_argument("An invalid result was produced by a batch fitness evaluation: the length of "
                                    "the vector representing the fitness vectors, "
                                    + std::to_string(fvs.size())
                                    + ", is not an exact multiple of the fitness dimension of the problem, "
                                    + std::to_string(f_dim));
    }
    if (n_fvs != n_dvs) {
        throw std::invalid_argument(
            "An invalid result was produced by a batch fitness evaluation: the number of produced fitness vectors, "
            + std::to_string(n_fvs) + ", differs from the number of input decision vectors, " + std::to_string(n_dvs));
    }
}

// src/batch_evaluators/thread_bfe.cpp:64-144 with TBB's parallel_for replaced
// by a static split over n_threads std::threads (per-individual results do not
// depend on the split).  As in the reference's `basic` branch each worker
// evaluates on its own copy of the problem, copies each dv into a temporary and
// receives a freshly allocated fitness vector per call.
inline vector_double thread_bfe(const udp &p, const vector_double &dvs, unsigned n_threads)
{
    const auto n_dim = p.get_nx();
    const auto f_dim = p.get_nf();
    const auto n_dvs = dvs.size() / n_dim;
    vector_double retval(n_dvs * f_dim);
    auto range_evaluator = [&dvs, &retval, n_dim, f_dim](const udp &prob, std::size_t begin, std::size_t end) {
        vector_double tmp_dv(n_dim);
        for (; begin != end; ++begin) {
            auto in_ptr = dvs.data() + begin * n_dim;
            auto out_ptr = retval.data() + begin * f_dim;
            std::copy(in_ptr, in_ptr + n_dim, tmp_dv.begin());
            const auto fv = prob.fitness(tmp_dv);
            std::copy(fv.begin(), fv.end(), out_ptr);
        }
    };
    if (n_threads <= 1u || n_dvs < 2u) {
        range_evaluator(udp(p), 0, n_dvs);
        return retval;
    }
    n_threads = static_cast<unsigned>(std::min<std::size_t>(n_threads, n_dvs));
    std::vector<std::thread> pool;
    pool.reserve(n_threads);
    for (unsigned t = 0; t < n_threads; ++t) {
        const std::size_t b = n_dvs * t / n_threads, e = n_dvs * (t + 1u) / n_threads;
        pool.emplace_back([p, b, e, &range_evaluator]() { range_evaluator(p, b, e); });
    }
    for (auto &th : pool) {
        th.join();
    }
    return retval;
}

// src/bfe.cpp:91-110: validate, evaluate, validate (fevals are the caller's).
inline vector_double bfe_call(const udp &p, const vector_double &dvs, unsigned n_threads)
{
    bfe_check_input_dvs(p, dvs);
    auto retval = thread_bfe(p, dvs, n_threads);
    bfe_check_output_fvs(p, dvs, retval);
    return retval;
}

// ---------------------------------------------------------------------------
// Memetic local step on the analytic gradient.  The reference has no such
// routine; this states the rule of pgh_local_descent (pagmo_hip/pgh.h) one
// individual at a time: per round, m = max_j |g_j|, s = alpha / m,
// y_j = x_j - s * g_j clamped to [lb_j, ub_j] (when bounds are given), and y
// replaces x when f(y) < f(x) (alpha doubled, kept finite), otherwise alpha is
// halved.  A non-finite gradient element or m == 0 leaves y = x.
// ---------------------------------------------------------------------------
inline void local_descent(const udp &p, double *X, std::size_t n, double *F, const double *lb, const double *ub,
                          unsigned iters, double step0)
{
    const std::size_t nx = p.get_nx();
    for (std::size_t i = 0; i < n; ++i) {
        vector_double x(X + i * nx, X + (i + 1u) * nx);
        double f = p.fitness(x)[0];
        double alpha = step0;
        for (unsigned it = 0; it < iters; ++it) {
            const auto g = p.gradient(x);
            double m = 0.;
            bool bad = false;
            for (std::size_t j = 0; j < nx; ++j) {
                const double a = std::abs(g[j]);
                if (!(a <= std::numeric_limits<double>::max())) bad = true;
                if (a > m) m = a;
            }
            const bool move = !bad && m > 0.;
            const double s = move ? alpha / m : 0.;
            vector_double y(x);
            if (move) {
                for (std::size_t j = 0; j < nx; ++j) {
                    double v = x[j] - s * g[j];
                    if (lb) {
                        if (v < lb[j]) v = lb[j];
                        if (v > ub[j]) v = ub[j];
                    }
                    y[j] = v;
                }
            }
            const double fy = p.fitness(y)[0];
            if (fy < f) {
                x = y;
                f = fy;
                const double a2 = 2. * alpha;
                if (a2 <= std::numeric_limits<double>::max()) alpha = a2;
            } else {
                alpha = 0.5 * alpha;
            }
        }
        std::copy(x.begin(), x.end(), X + i * nx);
        F[i] = f;
    }
}

// ---------------------------------------------------------------------------
// Gradient estimation (include/pagmo/utils/gradients_and_hessians.hpp)
// ---------------------------------------------------------------------------

// gradients_and_hessians.hpp:131-155.
template <typename Func>
inline vector_double estimate_gradient(Func f, const vector_double &x, double dx = 1e-8)
{
    vector_double f0 = f(x);
    vector_double gradient(f0.size() * x.size(), 0.);
    vector_double x_r = x, x_l = x;
    for (decltype(x.size()) j = 0u; j < x.size(); ++j) {
        double h = std::max(std::abs(x[j]), 1.0) * dx;
        x_r[j] = x[j] + h;
        x_l[j] = x[j] - h;
        vector_double f_r = f(x_r);
        vector_double f_l = f(x_l);
        if (f_r.size() != f0.size() || f_l.size() != f0.size()) {
            throw std::invalid_argument("Change in the size of the returned vector detected around the "
                                        "reference point. Cannot compute a gradient");
        }
        for (decltype(f_r.size()) i = 0u; i < f_r.size(); ++i) {
            gradient[j + i * x.size()] = (f_r[i] - f_l[i]) / 2. / h;
        }
        x_r[j] = x[j];
        x_l[j] = x[j];
    }
    return gradient;
}

// gradients_and_hessians.hpp:199-244.
template <typename Func>
inline vector_double estimate_gradient_h(Func f, const vector_double &x, double dx = 1e-2)
{
    vector_double f0 = f(x);
    vector_double gradient(f0.size() * x.size(), 0.);
    vector_double x_r1 = x, x_l1 = x;
    vector_double x_r2 = x, x_l2 = x;
    vector_double x_r3 = x, x_l3 = x;
    for (decltype(x.size()) j = 0u; j < x.size(); ++j) {
        double h = std::max(std::abs(x[j]), 1.0) * dx;
        x_r1[j] = x[j] + h;
        x_l1[j] = x[j] - h;
        x_r2[j] = x[j] + 2. * h;
        x_l2[j] = x[j] - 2. * h;
        x_r3[j] = x[j] + 3. * h;
        x_l3[j] = x[j] - 3. * h;
        vector_double f_r1 = f(x_r1);
        vector_double f_l1 = f(x_l1);
        vector_double f_r2 = f(x_r2);
        vector_double f_l2 = f(x_l2);
        vector_double f_r3 = f(x_r3);
        vector_double f_l3 = f(x_l3);
        if (f_r1.size() != f0.size() || f_l1.size() != f0.size() || f_r2.size() != f0.size() || f_l2.size() != f0.size()
            || f_r3.size() != f0.size() || f_l3.size() != f0.size()) {
            throw std::invalid_argument("Change in the size of the returned vector detected around the "
                                        "reference point. Cannot compute a gradient");
        }
        for (decltype(f_r1.size()) i = 0u; i < f_r1.size(); ++i) {
            double m1 = (f_r1[i] - f_l1[i]) / 2.;
            double m2 = (f_r2[i] - f_l2[i]) / 4.;
            double m3 = (f_r3[i] - f_l3[i]) / 6.;
            double fifteen_m1 = 15. * m1;
            double six_m2 = 6. * m2;
            double ten_h = 10. * h;
            gradient[j + i * x.size()] = ((fifteen_m1 - six_m2) + m3) / ten_h;
        }
        x_r1[j] = x[j];
        x_l1[j] = x[j];
        x_r2[j] = x[j];
        x_l2[j] = x[j];
        x_r3[j] = x[j];
        x_l3[j] = x[j];
    }
    return gradient;
}

// ---------------------------------------------------------------------------
// Multi-objective utilities
// ---------------------------------------------------------------------------

// include/pagmo/detail/custom_comparisons.hpp:54-100 (NaN placed after +inf).
inline bool less_than_f(double a, double b)
{
    if (!std::isnan(a)) {
        return !std::isnan(b) ? a < b : true;
    }
    return false;
}
inline bool greater_than_f(double a, double b)
{
    if (!std::isnan(a)) {
        return !std::isnan(b) ? a > b : false;
    }
    return !std::isnan(b);
}
inline bool equal_to_f(double a, double b)
{
    if (!std::isnan(a) && !std::isnan(b)) {
        return a == b;
    }
    return std::isnan(a) && std::isnan(b);
}

// src/utils/multi_objective.cpp:97-113.
inline bool pareto_dominance(const double *o1, const double *o2, std::size_t m)
{
    bool found_strictly_dominating_dimension = false;
    for (std::size_t i = 0u; i < m; ++i) {
        if (greater_than_f(o1[i], o2[i])) {
            return false;
        } else if (less_than_f(o1[i], o2[i])) {
            found_strictly_dominating_dimension = true;
        }
    }
    return found_strictly_dominating_dimension;
}

struct fnds_result {
    std::vector<std::vector<pop_size_t>> fronts;
    std::vector<std::vector<pop_size_t>> dom_list;
    std::vector<pop_size_t> dom_count;
    std::vector<pop_size_t> rank;
};

// src/utils/multi_objective.cpp:200-257 over a row-major n x m matrix.
inline fnds_result fast_non_dominated_sorting(const double *points, std::size_t N, std::size_t m)
{
    if (N < 2u) {
        throw std::invalid_argument("At least two points are needed for fast_non_dominated_sorting: "
                                    + std::to_string(N) + " detected.");
    }
    fnds_result r;
    r.fronts.resize(1u);
    r.dom_list.resize(N);
    r.dom_count.assign(N, 0u);
    r.rank.assign(N, 0u);
    for (std::size_t i = 0u; i < N; ++i) {
        for (std::size_t j = 0u; j < i; ++j) {
            if (pareto_dominance(points + i * m, points + j * m, m)) {
                r.dom_list[i].push_back(j);
                ++r.dom_count[j];
            } else if (pareto_dominance(points + j * m, points + i * m, m)) {
                r.dom_list[j].push_back(i);
                ++r.dom_count[i];
            }
        }
    }
    for (std::size_t i = 0u; i < N; ++i) {
        if (r.dom_count[i] == 0u) {
            r.rank[i] = 0u;
            r.fronts[0].push_back(i);
        }
    }
    auto dom_count_copy(r.dom_count);
    auto current_front = r.fronts[0];
    std::size_t front_counter = 0u;
    while (!current_front.empty()) {
        std::vector<pop_size_t> next_front;
        for (std::size_t p = 0u; p < current_front.size(); ++p) {
            const auto &dl = r.dom_list[current_front[p]];
            for (std::size_t q = 0u; q < dl.size(); ++q) {
                --dom_count_copy[dl[q]];
                if (dom_count_copy[dl[q]] == 0u) {
                    r.rank[dl[q]] = front_counter + 1u;
                    next_front.push_back(dl[q]);
                }
            }
        }
        ++front_counter;
        current_front = next_front;
        if (!current_front.empty()) {
            r.fronts.push_back(current_front);
        }
    }
    return r;
}

// The reference sorts with std::sort, whose order among equal keys the
// standard leaves open.  STABLE = false restates the reference (std::sort,
// libstdc++'s introsort, so the same order as the reference for a given input
// on this standard library); STABLE = true breaks ties by input position
// (std::stable_sort), the rule the device implements.
template <typename It, typename Cmp>
inline void mo_sort(It first, It last, Cmp cmp, bool stable)
{
    if (stable) {
        std::stable_sort(first, last, cmp);
    } else {
        std::sort(first, last, cmp);
    }
}

// src/utils/multi_objective.cpp:280-318 over the N points (m objectives each)
// of a front given row-major.
inline vector_double crowding_distance(const double *pts, std::size_t N, std::size_t M, bool stable = false)
{
    if (N < 2u) {
        throw std::invalid_argument("A non dominated front must contain at least two points: " + std::to_string(N)
                                    + " detected.");
    }
    if (M < 2u) {
        throw std::invalid_argument("Points in the non dominated front must contain at least two objectives: "
                                    + std::to_string(M) + " detected.");
    }
    std::vector<pop_size_t> indexes(N);
    std::iota(indexes.begin(), indexes.end(), pop_size_t(0u));
    vector_double retval(N, 0.);
    for (std::size_t i = 0u; i < M; ++i) {
        mo_sort(indexes.begin(), indexes.end(),
                [i, pts, M](pop_size_t a, pop_size_t b) { return less_than_f(pts[a * M + i], pts[b * M + i]); },
                stable);
        retval[indexes[0]] = std::numeric_limits<double>::infinity();
        retval[indexes[N - 1u]] = std::numeric_limits<double>::infinity();
        const double df = pts[indexes[N - 1u] * M + i] - pts[indexes[0] * M + i];
        for (std::size_t j = 1u; j < N - 1u; ++j) {
            retval[indexes[j]] += (pts[indexes[j + 1u] * M + i] - pts[indexes[j - 1u] * M + i]) / df;
        }
    }
    return retval;
}

inline vector_double front_points(const double *pts, std::size_t M, const std::vector<pop_size_t> &front)
{
    vector_double out(front.size() * M);
    for (std::size_t k = 0; k < front.size(); ++k) {
        for (std::size_t i = 0; i < M; ++i) out[k * M + i] = pts[front[k] * M + i];
    }
    return out;
}

// src/utils/multi_objective.cpp:344-396.
inline std::vector<pop_size_t> select_best_N_mo(const double *pts, std::size_t n, std::size_t M, std::size_t N,
                                                bool stable = false)
{
    if (N == 0u || n == 0u) return {};
    if (n == 1u) return {0u};
    if (N >= n) {
        std::vector<pop_size_t> r(n);
        std::iota(r.begin(), r.end(), pop_size_t(0u));
        return r;
    }
    std::vector<pop_size_t> retval;
    std::size_t front_id = 0u;
    const auto t = fast_non_dominated_sorting(pts, n, M);
    for (const auto &front : t.fronts) {
        if (retval.size() + front.size() <= N) {
            for (auto i : front) retval.push_back(i);
            if (retval.size() == N) return retval;
            ++front_id;
        } else {
            break;
        }
    }
    const auto &front = t.fronts[front_id];
    const auto fp = front_points(pts, M, front);
    const vector_double cds = crowding_distance(fp.data(), front.size(), M, stable);
    std::vector<pop_size_t> idxs(front.size());
    std::iota(idxs.begin(), idxs.end(), pop_size_t(0u));
    mo_sort(idxs.begin(), idxs.end(), [&cds](pop_size_t a, pop_size_t b) { return greater_than_f(cds[a], cds[b]); },
            stable);
    const auto remaining = N - retval.size();
    for (std::size_t i = 0u; i < remaining; ++i) retval.push_back(front[idxs[i]]);
    return retval;
}

// src/utils/multi_objective.cpp:425-465.
inline std::vector<pop_size_t> sort_population_mo(const double *pts, std::size_t n, std::size_t M, bool stable = false)
{
    if (n < 2u) {
        if (n == 0u) return {};
        return {0u};
    }
    std::vector<pop_size_t> retval(n);
    std::iota(retval.begin(), retval.end(), pop_size_t(0u));
    const auto t = fast_non_dominated_sorting(pts, n, M);
    vector_double crowding(n);
    for (const auto &front : t.fronts) {
        if (front.size() == 1u) {
            crowding[front[0]] = 0.;
        } else {
            const auto fp = front_points(pts, M, front);
            const vector_double tmp = crowding_distance(fp.data(), front.size(), M, stable);
            for (std::size_t i = 0u; i < front.size(); ++i) crowding[front[i]] = tmp[i];
        }
    }
    mo_sort(retval.begin(), retval.end(),
            [&t, &crowding](pop_size_t a, pop_size_t b) {
                if (t.rank[a] == t.rank[b]) return greater_than_f(crowding[a], crowding[b]);
                return t.rank[a] < t.rank[b];
            },
            stable);
    return retval;
}

// ---------------------------------------------------------------------------
// Random decision vectors (include/pagmo/utils/generic.hpp, libstdc++ <random>)
// ---------------------------------------------------------------------------

// include/pagmo/rng.hpp:43: detail::random_engine_type is std::mt19937.
using random_engine_type = std::mt19937;

// generic.hpp:97-104: equal bounds consume no draw.
inline double uniform_real_from_range(double lb, double ub, random_engine_type &e)
{
    return (lb == ub) ? lb : std::uniform_real_distribution<double>(lb, ub)(e);
}

// generic.hpp:326-388 (continuous part; the six problems have no integer part).
inline vector_double batch_random_decision_vector(const udp &p, std::size_t n, random_engine_type &e)
{
    const auto nx = p.get_nx();
    const auto b = p.get_bounds();
    vector_double out(nx * n);
    std::uniform_real_distribution<double> rdist;
    for (std::size_t i = 0; i < out.size(); i += nx) {
        for (std::size_t j = 0; j < nx; ++j) {
            out[i + j] = (b.first[j] == b.second[j])
                             ? b.first[j]
                             : rdist(e, std::uniform_real_distribution<double>::param_type(b.first[j], b.second[j]));
        }
    }
    return out;
}

// src/population.cpp:62-78 (serial ctor): random_decision_vector per
// individual (generic.hpp:271-297), then one fitness call each.
inline void population_init(const udp &p, std::size_t pop_size, unsigned seed, vector_double &X, vector_double &F)
{
    random_engine_type e(seed);
    const auto nx = p.get_nx(), nf = p.get_nf();
    const auto b = p.get_bounds();
    X.assign(pop_size * nx, 0.);
    F.assign(pop_size * nf, 0.);
    for (std::size_t i = 0; i < pop_size; ++i) {
        vector_double x(nx);
        for (std::size_t j = 0; j < nx; ++j) {
            x[j] = uniform_real_from_range(b.first[j], b.second[j], e);
        }
        const auto f = p.fitness(x);
        std::copy(x.begin(), x.end(), X.begin() + static_cast<std::ptrdiff_t>(i * nx));
        std::copy(f.begin(), f.end(), F.begin() + static_cast<std::ptrdiff_t>(i * nf));
    }
}

// ---------------------------------------------------------------------------
// pso_gen (src/algorithms/pso_gen.cpp, include/pagmo/algorithms/pso_gen.hpp)
// ---------------------------------------------------------------------------

struct pso_params {
    unsigned gen = 1u;       // pso_gen.hpp:109-111 defaults
    double omega = 0.7298;
    double eta1 = 2.05;
    double eta2 = 2.05;
    double max_vel = 0.5;
    unsigned variant = 5u;
    unsigned neighb_type = 2u;
    unsigned neighb_param = 4u;
    bool memory = false;
    unsigned seed = 0u;
};

// Batch evaluator used inside the generation loop (bfe::operator() semantics).
using batch_eval_fn = std::function<vector_double(const udp &, const vector_double &)>;

class pso_gen
{
public:
    explicit pso_gen(const pso_params &prm) : m(prm), m_e(prm.seed) {}

    // pso_gen.cpp:120-530 for a non-stochastic, unconstrained, single-objective
    // problem.  X/F hold the population (pop x nx, pop x 1) and receive
    // pop.set_xf(i, lbX[i], lbfit[i]) at the end.  If `bfe` is set, every
    // generation's packed positions go through it (pso_gen.cpp:418-432);
    // otherwise through per-individual fitness calls.  `trace_X`/`trace_F`,
    // when non-null, receive the evaluated batch of every generation.
    void evolve(const udp &prob, vector_double &Xpop, vector_double &Fpop, const batch_eval_fn *bfe,
                vector_double *trace_X = nullptr, vector_double *trace_F = nullptr)
    {
        const auto dim = prob.get_nx();
        const auto bounds = prob.get_bounds();
        const auto &lb = bounds.first;
        const auto &ub = bounds.second;
        if (prob.get_nf() != 1u) {
            throw std::invalid_argument("Multiple objectives detected: pso_gen cannot deal with them");
        }
        const auto swarm_size = Xpop.size() / dim;
        if (!swarm_size) {
            throw std::invalid_argument("pso_gen does not work on an empty population");
        }
        vector_double dummy(dim, 0.);
        std::vector<vector_double> X(swarm_size, dummy), fit(swarm_size), lbX(swarm_size, dummy), lbfit(swarm_size);
        std::vector<std::vector<std::size_t>> neighb(swarm_size);
        vector_double best_neighb(dim, 0.);
        vector_double best_fit;
        bool best_fit_improved;
        vector_double minv(dim), maxv(dim);
        std::uniform_real_distribution<double> drng(0., 1.);
        for (std::size_t i = 0u; i < dim; ++i) {
            const double vwidth = (ub[i] - lb[i]) * m.max_vel;
            minv[i] = -1. * vwidth;
            maxv[i] = vwidth;
        }
        for (std::size_t i = 0u; i < swarm_size; ++i) {
            X[i].assign(Xpop.begin() + static_cast<std::ptrdiff_t>(i * dim),
                        Xpop.begin() + static_cast<std::ptrdiff_t>((i + 1) * dim));
            lbX[i] = X[i];
            fit[i] = {Fpop[i]};
            lbfit[i] = fit[i];
        }
        // pso_gen.cpp:193-200
        if ((m_V.size() != swarm_size) || (!m.memory)) {
            m_V = std::vector<vector_double>(swarm_size, dummy);
            for (std::size_t i = 0u; i < swarm_size; ++i) {
                for (std::size_t j = 0u; j < dim; ++j) {
                    m_V[i][j] = uniform_real_from_range(minv[j], maxv[j], m_e);
                }
            }
        }
        // pso_gen.cpp:203-218
        switch (m.neighb_type) {
            case 1:
                init_gbest(X, fit, best_neighb, best_fit, neighb);
                break;
            case 3:
                init_von(neighb);
                break;
            case 4:
                init_adaptive_random(neighb);
                best_fit = fit[best_idx(fit)];
                break;
            case 2:
            default:
                init_lbest(neighb);
        }
        const double acceleration_coefficient = m.eta1 + m.eta2;
        double r1 = 0., r2 = 0.;
        for (unsigned gen = 1u; gen <= m.gen; ++gen) {
            // velocity update, pso_gen.cpp:232-327
            for (std::size_t p = 0u; p < swarm_size; ++p) {
                if (m.neighb_type != 1u && m.variant != 6u) {
                    best_neighb = get_best_neighbor(p, neighb, lbX, lbfit);
                }
                auto &V = m_V[p];
                if (m.variant == 1u) {
                    for (std::size_t d = 0u; d < dim; ++d) {
                        r1 = drng(m_e);
                        r2 = drng(m_e);
                        V[d] = m.omega * V[d] + m.eta1 * r1 * (lbX[p][d] - X[p][d])
                               + m.eta2 * r2 * (best_neighb[d] - X[p][d]);
                    }
                } else if (m.variant == 2u) {
                    for (std::size_t d = 0u; d < dim; ++d) {
                        r1 = drng(m_e);
                        V[d] = m.omega * V[d] + m.eta1 * r1 * (lbX[p][d] - X[p][d])
                               + m.eta2 * r1 * (best_neighb[d] - X[p][d]);
                    }
                } else if (m.variant == 3u) {
                    r1 = drng(m_e);
                    r2 = drng(m_e);
                    for (std::size_t d = 0u; d < dim; ++d) {
                        V[d] = m.omega * V[d] + m.eta1 * r1 * (lbX[p][d] - X[p][d])
                               + m.eta2 * r2 * (best_neighb[d] - X[p][d]);
                    }
                } else if (m.variant == 4u) {
                    r1 = drng(m_e);
                    for (std::size_t d = 0u; d < dim; ++d) {
                        V[d] = m.omega * V[d] + m.eta1 * r1 * (lbX[p][d] - X[p][d])
                               + m.eta2 * r1 * (best_neighb[d] - X[p][d]);
                    }
                } else if (m.variant == 5u) {
                    for (std::size_t d = 0u; d < dim; ++d) {
                        r1 = drng(m_e);
                        r2 = drng(m_e);
                        V[d] = m.omega
                               * (V[d] + m.eta1 * r1 * (lbX[p][d] - X[p][d]) + m.eta2 * r2 * (best_neighb[d] - X[p][d]));
                    }
                } else if (m.variant == 6u) {
                    for (std::size_t d = 0u; d < dim; ++d) {
                        double sum_forces = 0.;
                        for (std::size_t n = 0u; n < neighb[p].size(); ++n) {
                            sum_forces += drng(m_e) * acceleration_coefficient * (lbX[neighb[p][n]][d] - X[p][d]);
                        }
                        V[d] = m.omega * (V[d] + sum_forces / static_cast<double>(neighb[p].size()));
                    }
                }
            }
            // position update, pso_gen.cpp:330-363
            for (std::size_t p = 0u; p < swarm_size; ++p) {
                for (std::size_t d = 0u; d < dim; ++d) {
                    auto &v = m_V[p][d];
                    if (v > maxv[d]) {
                        v = maxv[d];
                    } else if (v < minv[d]) {
                        v = minv[d];
                    }
                    double new_x = X[p][d] + v;
                    if (new_x < lb[d]) {
                        new_x = lb[d];
                        v = 0.;
                    } else if (new_x > ub[d]) {
                        new_x = ub[d];
                        v = 0.;
                    }
                    X[p][d] = new_x;
                }
            }
            // evaluation, pso_gen.cpp:408-440 (non-stochastic branch)
            vector_double decision_vectors(swarm_size * dim);
            for (std::size_t i = 0; i < swarm_size; ++i) {
                std::copy(X[i].begin(), X[i].end(), decision_vectors.data() + i * dim);
            }
            vector_double fitnesses;
            if (bfe) {
                fitnesses = (*bfe)(prob, decision_vectors);
            } else {
                fitnesses.resize(swarm_size);
                for (std::size_t p = 0; p < swarm_size; ++p) {
                    fitnesses[p] = prob.fitness(X[p])[0];
                }
            }
            for (std::size_t i = 0; i < swarm_size; ++i) {
                fit[i] = {fitnesses[i]};
            }
            if (trace_X) {
                trace_X->insert(trace_X->end(), decision_vectors.begin(), decision_vectors.end());
            }
            if (trace_F) {
                trace_F->insert(trace_F->end(), fitnesses.begin(), fitnesses.end());
            }
            // memory update, pso_gen.cpp:443-459
            best_fit_improved = false;
            for (std::size_t p = 0; p < swarm_size; p++) {
                if (less_than_f(fit[p][0], lbfit[p][0]) || equal_to_f(fit[p][0], lbfit[p][0])) {
                    lbfit[p] = fit[p];
                    lbX[p] = X[p];
                    if ((m.neighb_type == 1u || m.neighb_type == 4u)
                        && ((less_than_f(fit[p][0], best_fit[0]) || equal_to_f(fit[p][0], best_fit[0])))) {
                        best_neighb = X[p];
                        best_fit = fit[p];
                        best_fit_improved = true;
                    }
                }
            }
            if (m.neighb_type == 4u && !best_fit_improved) {
                init_adaptive_random(neighb);
            }
        }
        // pso_gen.cpp:526-528
        for (std::size_t i = 0u; i < swarm_size; ++i) {
            std::copy(lbX[i].begin(), lbX[i].end(), Xpop.begin() + static_cast<std::ptrdiff_t>(i * dim));
            Fpop[i] = lbfit[i][0];
        }
    }

    random_engine_type &engine()
    {
        return m_e;
    }

private:
    // src/population.cpp:285-311 for a single-objective unconstrained problem:
    // std::min_element over the fitness vectors (plain operator<), i.e. the
    // first index of the minimum.
    static std::size_t best_idx(const std::vector<vector_double> &fit)
    {
        std::size_t b = 0;
        for (std::size_t i = 1; i < fit.size(); ++i) {
            if (fit[i][0] < fit[b][0]) {
                b = i;
            }
        }
        return b;
    }
    // pso_gen.cpp:593-622: later neighbours win ties.
    vector_double get_best_neighbor(std::size_t pidx, const std::vector<std::vector<std::size_t>> &neighb,
                                    const std::vector<vector_double> &lbX,
                                    const std::vector<vector_double> &lbfit) const
    {
        auto bnidx = neighb[pidx][0];
        for (std::size_t nidx = 1u; nidx < neighb[pidx].size(); ++nidx) {
            if (less_than_f(lbfit[neighb[pidx][nidx]][0], lbfit[bnidx][0])
                || equal_to_f(lbfit[neighb[pidx][nidx]][0], lbfit[bnidx][0])) {
                bnidx = neighb[pidx][nidx];
            }
        }
        return lbX[bnidx];
    }
    // pso_gen.cpp:644-662
    void init_gbest(const std::vector<vector_double> &X, const std::vector<vector_double> &fit, vector_double &gbX,
                    vector_double &gbfit, std::vector<std::vector<std::size_t>> &neighb) const
    {
        const auto b = best_idx(fit);
        gbX = X[b];
        gbfit = fit[b];
        if (m.variant == 6u) {
            for (std::size_t i = 0u; i < neighb.size(); i++) {
                neighb[0].push_back(i);
            }
            for (std::size_t i = 0u; i < neighb.size(); i++) {
                neighb[i] = neighb[0];
            }
        }
    }
    // pso_gen.cpp:679-699: left neighbours radius..1, then right 1..radius.
    void init_lbest(std::vector<std::vector<std::size_t>> &neighb) const
    {
        const auto swarm_size = neighb.size();
        const unsigned radius = m.neighb_param / 2u;
        for (std::size_t pidx = 0u; pidx < swarm_size; ++pidx) {
            for (unsigned j = radius; j > 0u; j--) {
                neighb[pidx].push_back(pidx < j ? pidx - j + swarm_size : pidx - j);
            }
            for (unsigned j = 1u; j <= radius; j++) {
                neighb[pidx].push_back(pidx + j >= swarm_size ? pidx + j - swarm_size : pidx + j);
            }
        }
    }
    // pso_gen.cpp:719-746 (offsets {-1,0},{1,0},{0,-1},{0,1}, pso_gen.cpp:72).
    void init_von(std::vector<std::vector<std::size_t>> &neighb) const
    {
        static const int diff[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
        const int swarm_size = static_cast<int>(neighb.size());
        int rows = static_cast<int>(std::sqrt(swarm_size));
        while (swarm_size % rows != 0) {
            rows -= 1;
        }
        const int cols = swarm_size / rows;
        for (int pidx = 0; pidx < swarm_size; pidx++) {
            const int p_x = pidx % cols, p_y = pidx / cols;
            for (unsigned nidx = 0u; nidx < 4u; nidx++) {
                int n_x = (p_x + diff[nidx][0]) % cols;
                if (n_x < 0) n_x = cols + n_x;
                int n_y = (p_y + diff[nidx][1]) % rows;
                if (n_y < 0) n_y = rows + n_y;
                neighb[static_cast<unsigned>(pidx)].push_back(static_cast<unsigned>(n_y * cols + n_x));
            }
        }
    }
    // pso_gen.cpp:769-792
    void init_adaptive_random(std::vector<std::vector<std::size_t>> &neighb)
    {
        const auto swarm_size = neighb.size();
        std::uniform_int_distribution<std::size_t> dis(0u, swarm_size - 1u);
        for (std::size_t pidx = 0u; pidx < swarm_size; ++pidx) {
            neighb[pidx].clear();
        }
        for (std::size_t pidx = 0u; pidx < swarm_size; ++pidx) {
            neighb[pidx].push_back(pidx);
            for (unsigned j = 1u; j < m.neighb_param; ++j) {
                neighb[dis(m_e)].push_back(pidx);
            }
        }
    }

    pso_params m;
    random_engine_type m_e;
    std::vector<vector_double> m_V;
};

// ---------------------------------------------------------------------------
// Migration: select_best, fair_replace, topologies and the archipelago's
// migration step (src/island.cpp:452-642)
// ---------------------------------------------------------------------------

// A policy's migration rate: an absolute count or a fraction of the input
// individuals (detail/base_sr_policy.hpp, boost::variant<pop_size_t, double>).
struct migr_rate {
    bool fractional = false;
    std::uint64_t count = 1u;
    double fraction = 0.;
};

// src/detail/base_sr_policy.cpp:54-66
inline void check_migr_rate(const migr_rate &r)
{
    if (r.fractional && (!std::isfinite(r.fraction) || r.fraction < 0. || r.fraction > 1.)) {
        throw std::invalid_argument("Invalid fractional migration rate specified in the constructor of a "
                                    "replacement/selection policy: the rate must be in the [0., 1.] range, but it is "
                                    + std::to_string(r.fraction) + " instead");
    }
}

// The number of migrants of a rate over inds_size individuals
// (select_best.cpp:76-97, fair_replace.cpp:79-103 before its min with the
// migrants' count): numeric_cast truncates the fractional product.
// `policy` is "'Select best' selection policy" or "'fair_replace'
// replacement policy".
inline std::uint64_t migr_count(const migr_rate &r, std::uint64_t inds_size, const std::string &policy)
{
    if (r.fractional) {
        return std::min(static_cast<std::uint64_t>(r.fraction * static_cast<double>(inds_size)), inds_size);
    }
    if (r.count > inds_size) {
        throw std::invalid_argument("The absolute migration rate (" + std::to_string(r.count) + ") in a " + policy
                                    + " is larger than the number of input individuals ("
                                    + std::to_string(inds_size) + ")");
    }
    return r.count;
}

// individuals_group_t (IDs, decision vectors, fitness vectors).
struct ind_group {
    std::vector<std::uint64_t> id;
    std::vector<vector_double> x, f;
};

// Indices of g ordered by less_than_f on the first objective (std::sort as
// in the reference, or std::stable_sort: see mo_sort).
inline std::vector<pop_size_t> order_by_f(const ind_group &g, bool stable)
{
    std::vector<pop_size_t> idx(g.f.size());
    std::iota(idx.begin(), idx.end(), pop_size_t(0));
    mo_sort(idx.begin(), idx.end(), [&g](pop_size_t a, pop_size_t b) { return less_than_f(g.f[a][0], g.f[b][0]); },
            stable);
    return idx;
}

inline void push_ind(ind_group &to, const ind_group &from, pop_size_t i)
{
    to.id.push_back(from.id[i]);
    to.x.push_back(from.x[i]);
    to.f.push_back(from.f[i]);
}

// src/s_policies/select_best.cpp:60-116, single-objective unconstrained: the
// n_migr best individuals, best first.
inline ind_group select_best(const ind_group &inds, const migr_rate &r, bool stable = false)
{
    const auto n_migr = migr_count(r, inds.f.size(), "'Select best' selection policy");
    const auto order = order_by_f(inds, stable);
    ind_group out;
    for (pop_size_t i = 0; i < n_migr; ++i) {
        push_ind(out, inds, order[i]);
    }
    return out;
}

// src/r_policies/fair_replace.cpp:62-129, single-objective unconstrained: the
// best n_migr migrants join the population, and the best inds_size of the
// merged group, best first, are the new population.
inline ind_group fair_replace(const ind_group &inds, const migr_rate &r, const ind_group &mig, bool stable = false)
{
    const auto inds_size = inds.f.size();
    const auto n_migr = std::min<std::uint64_t>(migr_count(r, inds_size, "'fair_replace' replacement policy"),
                                                mig.f.size());
    const auto mig_order = order_by_f(mig, stable);
    ind_group merged(inds);
    for (pop_size_t i = 0; i < n_migr; ++i) {
        push_ind(merged, mig, mig_order[i]);
    }
    const auto order = order_by_f(merged, stable);
    ind_group out;
    for (pop_size_t i = 0; i < inds_size; ++i) {
        push_ind(out, merged, order[i]);
    }
    return out;
}

// Islands with an edge towards each island, in the order that
// topology::get_connections returns them, and the edge weights.
struct connections {
    std::vector<std::vector<std::size_t>> src;
    std::vector<vector_double> w;
};

// src/topology.cpp:58-68
inline void check_edge_weight(double w)
{
    if (!std::isfinite(w)) {
        throw std::invalid_argument("invalid weight for the edge of a topology: the value " + std::to_string(w)
                                    + " is not finite");
    }
    if (w < 0. || w > 1.) {
        throw std::invalid_argument("invalid weight for the edge of a topology: the value " + std::to_string(w)
                                    + " is not in the [0., 1.] range");
    }
}

// ring(n, w): the edges of ring::push_back (src/topologies/ring.cpp:61-100)
// added and removed vertex by vertex.  base_bgl_topology::get_connections
// (base_bgl_topology.cpp:214-230) walks inv_adjacent_vertices of a
// bidirectional vecS adjacency_list (topology.hpp:109-115): each vertex's
// in-edges in the order they were added, a removal erasing in place.
inline connections ring_topology(std::size_t n, double w)
{
    check_edge_weight(w);
    connections c;
    auto add = [&c, w](std::size_t a, std::size_t b) {
        c.src[b].push_back(a);
        c.w[b].push_back(w);
    };
    auto remove = [&c](std::size_t a, std::size_t b) {
        for (std::size_t j = 0; j < c.src[b].size(); ++j) {
            if (c.src[b][j] == a) {
                c.src[b].erase(c.src[b].begin() + static_cast<std::ptrdiff_t>(j));
                c.w[b].erase(c.w[b].begin() + static_cast<std::ptrdiff_t>(j));
                return;
            }
        }
    };
    for (std::size_t size = 1; size <= n; ++size) {
        c.src.emplace_back();
        c.w.emplace_back();
        if (size == 2u) {
            add(0, 1);
            add(1, 0);
        } else if (size == 3u) {
            add(1, 2);
            add(2, 1);
            add(2, 0);
            add(0, 2);
        } else if (size > 3u) {
            remove(size - 2u, 0);
            remove(0, size - 2u);
            add(size - 2u, size - 1u);
            add(size - 1u, size - 2u);
            add(0, size - 1u);
            add(size - 1u, 0);
        }
    }
    return c;
}

// fully_connected(n, w): every other vertex in ascending order
// (src/topologies/fully_connected.cpp:86-116).
inline connections fully_connected_topology(std::size_t n, double w)
{
    check_edge_weight(w);
    connections c;
    c.src.resize(n);
    c.w.resize(n);
    for (std::size_t i = 0; i < n; ++i) {
        for (std::size_t j = 0; j < n; ++j) {
            if (j != i) {
                c.src[i].push_back(j);
                c.w[i].push_back(w);
            }
        }
    }
    return c;
}

// One island of a migrating archipelago: a single-objective problem, its
// pso_gen and population, the policies' rates and the migration engine.
struct migr_island {
    udp prob;
    pso_gen algo;
    ind_group pop;
    migr_rate sel, rep;
    std::mt19937 eng;
};

// One archipelago::evolve(1) round in which every island migrates in before
// any island evolves, and emigrates after every island has evolved -- one of
// the interleavings the reference's island threads can take; with it every
// island reads the migrants database as the previous round left it.
//   immigration (island.cpp:452-597), island by island in index order:
//     p2p: one source drawn with uniform_int_distribution<size_t>(0, n-1)
//     and kept if uniform_real_distribution<>{} < its weight; fair_replace
//     only then.  broadcast: one draw per connection in order, the kept
//     sources' migrants concatenated; fair_replace always.  The source's
//     entry is copied (preserve) or taken and emptied (evict,
//     archipelago.cpp:658-678).
//   evolution: algo.evolve on the population (pso_gen, per-individual
//     fitness).
//   emigration (island.cpp:603-620): db[k] = select_best(population).
// type: 0 p2p, 1 broadcast; handling: 0 preserve, 1 evict.
inline void archi_migrate_round(std::vector<migr_island> &isl, const connections &c, int type, int handling,
                                std::vector<ind_group> &db, bool stable)
{
    for (std::size_t k = 0; k < isl.size(); ++k) {
        const auto &src = c.src[k];
        if (src.empty()) {
            continue;
        }
        auto take = [&db, handling](std::size_t s) {
            ind_group g = db[s];
            if (handling == 1) {
                db[s] = ind_group{};
            }
            return g;
        };
        if (type == 0) {
            const auto j = std::uniform_int_distribution<std::size_t>(0, src.size() - 1u)(isl[k].eng);
            if (std::uniform_real_distribution<>{}(isl[k].eng) < c.w[k][j]) {
                const auto mig = take(src[j]);
                isl[k].pop = fair_replace(isl[k].pop, isl[k].rep, mig, stable);
            }
        } else {
            ind_group mig;
            for (std::size_t j = 0; j < src.size(); ++j) {
                if (std::uniform_real_distribution<>{}(isl[k].eng) < c.w[k][j]) {
                    const auto g = take(src[j]);
                    for (pop_size_t i = 0; i < g.f.size(); ++i) {
                        push_ind(mig, g, i);
                    }
                }
            }
            isl[k].pop = fair_replace(isl[k].pop, isl[k].rep, mig, stable);
        }
    }
    for (auto &is : isl) {
        const auto nx = is.prob.get_nx();
        vector_double X, F;
        for (pop_size_t i = 0; i < is.pop.x.size(); ++i) {
            X.insert(X.end(), is.pop.x[i].begin(), is.pop.x[i].end());
            F.push_back(is.pop.f[i][0]);
        }
        is.algo.evolve(is.prob, X, F, nullptr);
        for (pop_size_t i = 0; i < is.pop.x.size(); ++i) {
            is.pop.x[i].assign(X.begin() + static_cast<std::ptrdiff_t>(i * nx),
                               X.begin() + static_cast<std::ptrdiff_t>((i + 1u) * nx));
            is.pop.f[i] = {F[i]};
        }
    }
    for (std::size_t k = 0; k < isl.size(); ++k) {
        db[k] = select_best(isl[k].pop, isl[k].sel, stable);
    }
}

} // namespace oracle
