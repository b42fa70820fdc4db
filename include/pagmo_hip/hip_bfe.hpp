// hip_bfe.hpp -- the MI355X batch fitness evaluator as a pagmo UDBFE.
//
//     pagmo::pso_gen algo{100u};
//     algo.set_bfe(pagmo::bfe{pagmo_hip::hip_bfe{}});
//
// hip_bfe is a user-defined batch fitness evaluator in the sense of
// include/pagmo/bfe.hpp:68-108: default/copy/move constructible and callable as
// vector_double(const problem &, const vector_double &).  It is stateless
// (like thread_bfe, src/batch_evaluators/thread_bfe.cpp:147-150) and never
// touches the problem's counters: the pagmo::bfe wrapper increments fevals
// (src/bfe.cpp:102-107).
//
// Problems with a device kernel (rosenbrock, rastrigin, ackley, schwefel,
// griewank, lennard_jones, zdt, dtlz, found with problem::extract<T>(),
// include/pagmo/problem.hpp:1104-1147) are evaluated by libpagmo_hip.so.  Any
// other problem goes through default_bfe's dispatch
// (src/batch_evaluators/default_bfe.cpp:53-68): member_bfe if it has
// batch_fitness(), else thread_bfe if its thread safety allows it, else the
// same std::invalid_argument.  A supported problem with no usable device is an
// error, never a silent CPU evaluation.
//
// Header-only; link with -lpagmo_hip.  Besides the UDBFE the header offers the
// batched gradient (problem::gradient, src/problem.cpp:456-467, for a whole
// population), fast_non_dominated_sorting, crowding_distance,
// sort_population_mo and select_best_N_mo (src/utils/multi_objective.cpp:
// 200-465) on the device, and local_descent, a memetic local step on the
// batched gradient for a population between two evolve() calls.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include <pagmo/batch_evaluators/member_bfe.hpp>
#include <pagmo/batch_evaluators/thread_bfe.hpp>
#include <pagmo/exceptions.hpp>
#include <pagmo/population.hpp>
#include <pagmo/problem.hpp>
#include <pagmo/problems/ackley.hpp>
#include <pagmo/problems/dtlz.hpp>
#include <pagmo/problems/griewank.hpp>
#include <pagmo/problems/lennard_jones.hpp>
#include <pagmo/problems/rastrigin.hpp>
#include <pagmo/problems/rosenbrock.hpp>
#include <pagmo/problems/schwefel.hpp>
#include <pagmo/problems/translate.hpp>
#include <pagmo/problems/zdt.hpp>
#include <pagmo/threading.hpp>
#include <pagmo/types.hpp>
#include <pagmo/utils/multi_objective.hpp>

#include "pgh.h"

namespace pagmo_hip
{

namespace detail
{

// Throws the exception type of the reference for a failed status.
[[noreturn]] inline void raise(pgh_status st)
{
    const std::string msg = pgh_last_error();
    switch (st) {
        case PGH_INVALID_ARGUMENT:
            pagmo_throw(std::invalid_argument, msg);
        case PGH_OVERFLOW:
            pagmo_throw(std::overflow_error, msg);
        default:
            pagmo_throw(std::runtime_error, "libpagmo_hip: " + msg);
    }
}

inline void check(pgh_status st)
{
    if (st != PGH_OK) raise(st);
}

} // namespace detail

namespace detail
{
// The number at the end of a UDP name "<prefix><id>" (zdt.cpp:164-167,
// dtlz.cpp:168-171), or 0.
inline unsigned name_id(const std::string &name, const std::string &prefix)
{
    if (name.size() <= prefix.size() || name.compare(0, prefix.size(), prefix) != 0) return 0u;
    unsigned id = 0u;
    for (auto c = name.begin() + static_cast<std::ptrdiff_t>(prefix.size()); c != name.end(); ++c) {
        if (*c < '0' || *c > '9' || id > 1000u) return 0u;
        id = 10u * id + static_cast<unsigned>(*c - '0');
    }
    return id;
}

// dtlz keeps alpha private (dtlz.hpp:163); it enters DTLZ4 only through
// x_j^alpha (dtlz.cpp:320-341).  At x = (0.5, ..., 0.5) the distance function
// is 0 and the last objective is sin(0.5^alpha * pi / 2): alpha is read back
// from it and confirmed by recomputing that objective (same libm as the UDP).
// Values it cannot pin (0.5^alpha underflowing, above 1000) report false.
inline bool dtlz4_alpha(const pagmo::dtlz &q, std::size_t nx, uint64_t &alpha)
{
    const double pi_half = 3.1415926535897932384626433832795028841971693993 / 2.;
    const auto f = q.fitness(pagmo::vector_double(nx, 0.5));
    const double last = f.back();
    if (!(last > 0.) || !(last <= 1.)) return false;
    const double a = std::round(-std::log2(std::asin(last) / pi_half));
    if (!(a >= 0.) || a > 1000.) return false;
    const auto cand = static_cast<unsigned>(a);
    if (std::sin(std::pow(0.5, cand) * pi_half) != last) return false;
    alpha = cand;
    return true;
}
// The device description of p itself (no meta-problem), if a kernel exists.
inline bool describe_udp(const pagmo::problem &p, pgh_udp &u)
{
    u = pgh_udp{};
    if (const auto *r = p.extract<pagmo::rosenbrock>()) {
        u.kind = PGH_ROSENBROCK;
        u.dim = r->m_dim;
    } else if (p.extract<pagmo::rastrigin>()) {
        u.kind = PGH_RASTRIGIN;
        u.dim = p.get_nx();
    } else if (p.extract<pagmo::ackley>()) {
        u.kind = PGH_ACKLEY;
        u.dim = p.get_nx();
    } else if (p.extract<pagmo::schwefel>()) {
        u.kind = PGH_SCHWEFEL;
        u.dim = p.get_nx();
    } else if (p.extract<pagmo::griewank>()) {
        u.kind = PGH_GRIEWANK;
        u.dim = p.get_nx();
    } else if (p.extract<pagmo::lennard_jones>()) {
        // m_atoms is private (lennard_jones.hpp:87); nx = 3*atoms - 6
        u.kind = PGH_LENNARD_JONES;
        u.dim = (p.get_nx() + 6u) / 3u;
    } else if (const auto *z = p.extract<pagmo::zdt>()) {
        // m_prob_id and m_param are private (zdt.hpp:205-206): the id is in
        // the name, the param follows from nx (zdt.cpp:108-132)
        u.kind = PGH_ZDT;
        u.prob_id = detail::name_id(z->get_name(), "ZDT");
        const auto nx = p.get_nx();
        if (u.prob_id == 5u && (nx < 30u || (nx - 30u) % 5u != 0u)) return false;
        u.dim = u.prob_id == 5u ? (nx - 30u) / 5u + 1u : nx;
        u.nobj = 2;
        return pgh_supported(&u) == 1;
    } else if (const auto *d = p.extract<pagmo::dtlz>()) {
        u.kind = PGH_DTLZ;
        u.prob_id = detail::name_id(d->get_name(), "DTLZ");
        u.dim = p.get_nx();
        u.nobj = d->get_nobj();
        if (u.prob_id == 4u && !detail::dtlz4_alpha(*d, p.get_nx(), u.alpha)) return false;
        return pgh_supported(&u) == 1;
    } else {
        return false;
    }
    u.nobj = 1;
    return pgh_supported(&u) == 1;
}
} // namespace detail

// The device description of p, if a kernel exists for it: one of the UDPs
// above, or a translate (translate.hpp) of one, whose vector the host-buffer
// entry points subtract on the device (x - t, translate.cpp:414-423).  A
// translate of a translate rounds twice and is left to the CPU evaluator;
// decompose keeps its weights and method private (decompose.hpp:219-224,
// printed rounded by get_extra_info) and is left there too.
inline bool describe(const pagmo::problem &p, pgh_udp &u)
{
    if (const auto *t = p.extract<pagmo::translate>()) {
        const pagmo::problem &inner = t->get_inner_problem();
        if (inner.extract<pagmo::translate>() || !detail::describe_udp(inner, u)) return false;
        u.shift = t->get_translation().data();
        return pgh_supported(&u) == 1;
    }
    return detail::describe_udp(p, u);
}

class hip_bfe
{
public:
    pagmo::vector_double operator()(const pagmo::problem &p, const pagmo::vector_double &dvs) const
    {
        pgh_udp u;
        if (!describe(p, u)) {
            return fallback(p, dvs);
        }
        const auto nx = p.get_nx(), nf = p.get_nf();
        pagmo::vector_double out(nx ? dvs.size() / nx * nf : 0u);
        detail::check(pgh_bfe_eval_host(&u, dvs.data(), dvs.size(), out.data(), out.size()));
        return out;
    }
    std::string get_name() const
    {
        return "HIP batch fitness evaluator (MI355X, gfx950)";
    }
    std::string get_extra_info() const
    {
        return std::string("\tDevice available: ") + (pgh_device_available() ? "yes" : "no")
               + "\n\tC ABI version: " + std::to_string(pgh_abi_version());
    }
    pagmo::thread_safety get_thread_safety() const
    {
        // every host thread gets its own stream and staging buffers
        return pagmo::thread_safety::basic;
    }
    template <typename Archive>
    void serialize(Archive &, unsigned)
    {
    }

private:
    // default_bfe's heuristic (src/batch_evaluators/default_bfe.cpp:53-68).
    static pagmo::vector_double fallback(const pagmo::problem &p, const pagmo::vector_double &dvs)
    {
        if (p.has_batch_fitness()) {
            return pagmo::member_bfe{}(p, dvs);
        }
        if (p.get_thread_safety() >= pagmo::thread_safety::basic) {
            return pagmo::thread_bfe{}(p, dvs);
        }
        pagmo_throw(std::invalid_argument,
                    "Cannot execute fitness evaluations in batch mode for a problem of type '" + p.get_name()
                        + "': the problem does not implement the batch_fitness() member function, and its thread "
                          "safety level is not sufficient to run a thread-based batch fitness evaluation "
                          "implementation");
    }
};

// Gradients of a whole population: one problem::gradient vector per
// decision vector of dvs (row-major; gradient_sparsity() order, which is the
// dense 1 x nx pattern of problem.cpp:132-140 for the device kernels).
//   - rosenbrock and rastrigin (their UDP's gradient()) and lennard_jones
//     (the library's adjoint kernel; pagmo's UDP has none), also inside a
//     translate, run on the device.  The problem's gradient counter is not
//     advanced there: pagmo's problem offers no public way to do so.
//   - every other problem goes through problem::gradient row by row, with
//     its semantics (src/problem.cpp:456-467): the length check, one gevals
//     per row, and not_implemented_error for a UDP without a gradient.
inline pagmo::vector_double batch_gradient(const pagmo::problem &p, const pagmo::vector_double &dvs)
{
    const auto nx = p.get_nx();
    if (nx == 0u || dvs.size() % nx) {
        pagmo_throw(std::invalid_argument, "Invalid argument for a batch gradient evaluation: the length of the "
                                           "vector representing the decision vectors, "
                                               + std::to_string(dvs.size())
                                               + ", is not an exact multiple of the dimension of the problem, "
                                               + std::to_string(nx));
    }
    pgh_udp u;
    if (describe(p, u)
        && (u.kind == PGH_ROSENBROCK || u.kind == PGH_RASTRIGIN || u.kind == PGH_LENNARD_JONES)) {
        pagmo::vector_double g(dvs.size());
        detail::check(pgh_batch_gradient_host(&u, dvs.data(), dvs.size(), g.data(), g.size(), nullptr));
        return g;
    }
    pagmo::vector_double out, row(nx);
    for (std::size_t i = 0; i < dvs.size() / nx; ++i) {
        std::copy(dvs.begin() + static_cast<std::ptrdiff_t>(i * nx),
                  dvs.begin() + static_cast<std::ptrdiff_t>((i + 1u) * nx), row.begin());
        const auto g = p.gradient(row);
        out.insert(out.end(), g.begin(), g.end());
    }
    return out;
}

// Memetic local step: `iters` rounds of the projected gradient descent of
// pgh_local_descent (pagmo_hip/pgh.h) on every individual of `pop`, inside the
// problem's bounds, starting from the step length `step0`.  Individuals are
// written back with population::set_xf (src/population.cpp:490-519), which
// keeps the champion current; an individual's fitness never increases.  No
// algorithm's selection logic is involved: a caller interleaves it with
// evolve(), e.g.
//     for (...) { pop = algo.evolve(pop); pagmo_hip::local_descent(pop, 5u, 0.1); }
// The problem's fevals grow by (iters + 1) * pop.size(), the fitness
// evaluations made on the device.
inline void local_descent(pagmo::population &pop, unsigned iters, double step0)
{
    const pagmo::problem &p = pop.get_problem();
    pgh_udp u;
    if (!describe(p, u)) {
        pagmo_throw(std::invalid_argument, "local_descent: no device gradient for the problem '" + p.get_name() + "'");
    }
    const auto n = pop.size();
    if (n == 0u) return;
    const auto nx = p.get_nx();
    std::vector<double> x;
    x.reserve(n * nx);
    for (const auto &xi : pop.get_x()) x.insert(x.end(), xi.begin(), xi.end());
    std::vector<double> f(n);
    const auto bounds = p.get_bounds();
    detail::check(pgh_local_descent_host(&u, x.data(), x.size(), f.data(), f.size(), bounds.first.data(),
                                         bounds.second.data(), iters, step0));
    for (decltype(pop.size()) i = 0; i < n; ++i) {
        pop.set_xf(i, pagmo::vector_double(x.begin() + static_cast<std::ptrdiff_t>(i * nx),
                                           x.begin() + static_cast<std::ptrdiff_t>((i + 1u) * nx)),
                   pagmo::vector_double{f[i]});
    }
    p.increment_fevals(static_cast<unsigned long long>(iters + 1u) * n);
}

namespace detail
{
// Row-major copy of equally sized objective vectors; a size mismatch throws
// pareto_dominance's message (multi_objective.cpp:99-104).
inline std::vector<double> flatten_objectives(const std::vector<pagmo::vector_double> &points, std::size_t m)
{
    std::vector<double> flat;
    flat.reserve(points.size() * m);
    for (const auto &v : points) {
        if (v.size() != m) {
            pagmo_throw(std::invalid_argument, "Different number of objectives found in input fitnesses: "
                                                   + std::to_string(m) + " and " + std::to_string(v.size())
                                                   + ". I cannot define dominance");
        }
        flat.insert(flat.end(), v.begin(), v.end());
    }
    return flat;
}
} // namespace detail

// fast_non_dominated_sorting on the device, same return value and exceptions.
inline pagmo::fnds_return_type fast_non_dominated_sorting(const std::vector<pagmo::vector_double> &points)
{
    const auto n = points.size();
    const auto m = n ? points[0].size() : 0u;
    const auto flat = detail::flatten_objectives(points, m);
    // the calling thread's workspace: kept between calls, so a generation
    // loop (nsga2.cpp:183) allocates device memory and a stream once
    pgh_fnds_ws *ws = nullptr;
    detail::check(pgh_fnds_thread_workspace(&ws));
    detail::check(pgh_fnds_run_host(ws, flat.empty() ? nullptr : flat.data(), n, m));
    uint64_t n_fronts = 0, n_dom = 0;
    detail::check(pgh_fnds_sizes(ws, &n_fronts, &n_dom));
    std::vector<uint64_t> dom_count(n), rank(n), fo(n_fronts + 1u), fi(n), doff(n + 1u), di(n_dom);
    detail::check(pgh_fnds_copy(ws, dom_count.data(), rank.data(), fo.data(), fi.data(), doff.data(),
                                di.empty() ? nullptr : di.data(), 1));
    std::vector<std::vector<pagmo::pop_size_t>> fronts(n_fronts), dom_list(n);
    for (uint64_t f = 0; f < n_fronts; ++f) {
        fronts[f].assign(fi.begin() + static_cast<std::ptrdiff_t>(fo[f]), fi.begin() + static_cast<std::ptrdiff_t>(fo[f + 1u]));
    }
    for (uint64_t i = 0; i < n; ++i) {
        dom_list[i].assign(di.begin() + static_cast<std::ptrdiff_t>(doff[i]),
                           di.begin() + static_cast<std::ptrdiff_t>(doff[i + 1u]));
    }
    return std::make_tuple(std::move(fronts), std::move(dom_list),
                           std::vector<pagmo::pop_size_t>(dom_count.begin(), dom_count.end()),
                           std::vector<pagmo::pop_size_t>(rank.begin(), rank.end()));
}


// crowding_distance (src/utils/multi_objective.cpp:280-318) on the device, same
// checks and messages.  Equal objective values keep their order from the
// previous objective's sort (std::sort leaves it open in the reference).
inline pagmo::vector_double crowding_distance(const std::vector<pagmo::vector_double> &non_dom_front)
{
    const auto N = non_dom_front.size();
    if (N < 2u) {
        pagmo_throw(std::invalid_argument,
                    "A non dominated front must contain at least two points: " + std::to_string(N) + " detected.");
    }
    const auto M = non_dom_front[0].size();
    if (M < 2u) {
        pagmo_throw(std::invalid_argument, "Points in the non dominated front must contain at least two objectives: "
                                               + std::to_string(M) + " detected.");
    }
    for (const auto &v : non_dom_front) {
        if (v.size() != M) {
            pagmo_throw(std::invalid_argument, "A non dominated front must contain points of uniform dimensionality. "
                                               "Some different sizes were instead detected.");
        }
    }
    const auto flat = detail::flatten_objectives(non_dom_front, M);
    pagmo::vector_double out(N);
    detail::check(pgh_crowding_distance_host(flat.data(), N, M, out.data()));
    return out;
}

// sort_population_mo (multi_objective.cpp:425-465) on the device; equal
// (rank, crowding) pairs in index order.
inline std::vector<pagmo::pop_size_t> sort_population_mo(const std::vector<pagmo::vector_double> &input_f)
{
    const auto n = input_f.size();
    const auto m = n ? input_f[0].size() : 0u;
    const auto flat = detail::flatten_objectives(input_f, m);
    std::vector<uint64_t> out(n);
    detail::check(pgh_sort_population_mo_host(flat.empty() ? nullptr : flat.data(), n, m, out.data()));
    return std::vector<pagmo::pop_size_t>(out.begin(), out.end());
}

// select_best_N_mo (multi_objective.cpp:344-396) on the device.
inline std::vector<pagmo::pop_size_t> select_best_N_mo(const std::vector<pagmo::vector_double> &input_f,
                                                       pagmo::pop_size_t N)
{
    const auto n = input_f.size();
    if (N == 0u || n == 0u) return {};
    const auto m = input_f[0].size();
    const auto flat = detail::flatten_objectives(input_f, m);
    std::vector<uint64_t> out(n);
    uint64_t len = 0;
    detail::check(pgh_select_best_N_mo_host(flat.empty() ? nullptr : flat.data(), n, m, N, out.data(), &len));
    return std::vector<pagmo::pop_size_t>(out.begin(), out.begin() + static_cast<std::ptrdiff_t>(len));
}

} // namespace pagmo_hip
