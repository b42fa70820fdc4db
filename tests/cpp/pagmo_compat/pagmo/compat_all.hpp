// compat_all.hpp -- TEST SCAFFOLDING: a minimal stand-in for the part of
// pagmo 2.19.1's API that include/pagmo_hip/hip_bfe.hpp and its drop-in test
// use (pagmo itself cannot be built here: it needs Boost and TBB).  The
// headers next to this one keep pagmo's include paths so that hip_bfe.hpp
// compiles unchanged against either.  UDP arithmetic comes from the oracle.
//
// Mirrored semantics (citations into the pagmo sources):
//   problem: fitness() checks lengths and counts fevals (src/problem.cpp:353-379),
//            gradient() checks the length and counts gevals (src/problem.cpp:456-467),
//            not_implemented_error for a UDP without one (problem.hpp:638-643),
//            extract<T>() (include/pagmo/problem.hpp:1104-1147), thread safety
//            defaults to basic (problem.hpp:847-851)
//   bfe:     validate -> call -> validate -> increment_fevals (src/bfe.cpp:91-110)
//   thread_bfe / member_bfe (src/batch_evaluators/*.cpp); member_bfe does not
//            count fevals, the bfe wrapper does (member_bfe.cpp:40-45)
//   population(prob, size, seed) (src/population.cpp:62-78)
//   pso_gen with set_bfe (src/algorithms/pso_gen.cpp)
#pragma once

#include <algorithm>
#include <atomic>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "oracle.hpp"

#define pagmo_throw(exception_type, ...) throw exception_type(__VA_ARGS__)

namespace pagmo
{

// include/pagmo/exceptions.hpp
struct not_implemented_error final : std::runtime_error {
    using std::runtime_error::runtime_error;
};

using vector_double = std::vector<double>;
typedef std::vector<vector_double>::size_type pop_size_t;
enum class thread_safety { none, basic, constant };
using fnds_return_type = std::tuple<std::vector<std::vector<pop_size_t>>, std::vector<std::vector<pop_size_t>>,
                                    std::vector<pop_size_t>, std::vector<pop_size_t>>;

// ---- the six UDPs (fitness and bounds from the oracle's restatement) ----------
#define PAGMO_COMPAT_UDP(NAME, KIND, TITLE)                                                                            \
    class NAME                                                                                                         \
    {                                                                                                                  \
    public:                                                                                                            \
        explicit NAME(unsigned dim = 1u) : m_dim(dim)                                                                  \
        {                                                                                                              \
            if (dim < 1u) throw std::invalid_argument(TITLE " Function must have minimum 1 dimension");                \
        }                                                                                                              \
        vector_double fitness(const vector_double &x) const                                                            \
        {                                                                                                              \
            return as_udp().fitness(x);                                                                                \
        }                                                                                                              \
        std::pair<vector_double, vector_double> get_bounds() const                                                     \
        {                                                                                                              \
            return as_udp().get_bounds();                                                                              \
        }                                                                                                              \
        std::string get_name() const                                                                                   \
        {                                                                                                              \
            return TITLE " Function";                                                                                  \
        }                                                                                                              \
        oracle::udp as_udp() const                                                                                     \
        {                                                                                                              \
            oracle::udp u;                                                                                             \
            u.kind = KIND;                                                                                             \
            u.dim = m_dim;                                                                                             \
            return u;                                                                                                  \
        }                                                                                                              \
                                                                                                                       \
    private:                                                                                                           \
        unsigned m_dim;                                                                                                \
    };

PAGMO_COMPAT_UDP(rastrigin, oracle::RASTRIGIN, "Rastrigin")
PAGMO_COMPAT_UDP(ackley, oracle::ACKLEY, "Ackley")
PAGMO_COMPAT_UDP(schwefel, oracle::SCHWEFEL, "Schwefel")
PAGMO_COMPAT_UDP(griewank, oracle::GRIEWANK, "Griewank")
#undef PAGMO_COMPAT_UDP

class rosenbrock
{
public:
    explicit rosenbrock(vector_double::size_type dim = 2u) : m_dim(dim)
    {
        if (dim < 2u) throw std::invalid_argument("Rosenbrock Function must have minimum 2 dimensions");
    }
    vector_double fitness(const vector_double &x) const
    {
        return oracle::rosenbrock_fitness(m_dim, x);
    }
    std::pair<vector_double, vector_double> get_bounds() const
    {
        return as_udp().get_bounds();
    }
    vector_double gradient(const vector_double &x) const
    {
        return oracle::rosenbrock_gradient(m_dim, x);
    }
    thread_safety get_thread_safety() const
    {
        return thread_safety::constant; // rosenbrock.hpp:86-89
    }
    std::string get_name() const
    {
        return "Multidimensional Rosenbrock Function";
    }
    oracle::udp as_udp() const
    {
        oracle::udp u;
        u.kind = oracle::ROSENBROCK;
        u.dim = static_cast<unsigned>(m_dim);
        return u;
    }
    vector_double::size_type m_dim; // public, as in rosenbrock.hpp:91
};

class lennard_jones
{
public:
    explicit lennard_jones(unsigned atoms = 3u) : m_atoms(atoms)
    {
        if (atoms < 3u) throw std::invalid_argument("The number of atoms must be greater than 2");
    }
    vector_double fitness(const vector_double &x) const
    {
        return oracle::lennard_jones_fitness(m_atoms, x);
    }
    std::pair<vector_double, vector_double> get_bounds() const
    {
        return as_udp().get_bounds();
    }
    std::string get_name() const
    {
        return "Lennard Jones Cluster (" + std::to_string(m_atoms) + " atoms)";
    }
    oracle::udp as_udp() const
    {
        oracle::udp u;
        u.kind = oracle::LENNARD_JONES;
        u.atoms = m_atoms;
        return u;
    }

private:
    unsigned m_atoms;
};

// ---- ZDT and DTLZ (zdt.hpp, dtlz.hpp: parameters private, as there) --------
class zdt
{
public:
    zdt(unsigned prob_id = 1u, unsigned param = 30u) : m_prob_id(prob_id), m_param(param)
    {
        if (param < 2u) {
            throw std::invalid_argument("ZDT test problems must have a minimum value of 2 for the constructing "
                                        "parameter (representing the dimension except for ZDT5), "
                                        + std::to_string(param) + " requested");
        }
        if (prob_id == 0u || prob_id > 6u) {
            throw std::invalid_argument("ZDT test suite contains six (prob_id=[1 ... 6]) problems, prob_id="
                                        + std::to_string(prob_id) + " was detected");
        }
    }
    vector_double fitness(const vector_double &x) const
    {
        return oracle::zdt_fitness(m_prob_id, x);
    }
    std::pair<vector_double, vector_double> get_bounds() const
    {
        return as_udp().get_bounds();
    }
    vector_double::size_type get_nobj() const
    {
        return 2u;
    }
    std::string get_name() const
    {
        return "ZDT" + std::to_string(m_prob_id);
    }
    oracle::udp as_udp() const
    {
        oracle::udp u;
        u.kind = oracle::ZDT;
        u.prob_id = m_prob_id;
        u.dim = m_param;
        return u;
    }

private:
    unsigned m_prob_id;
    unsigned m_param;
};

class dtlz
{
public:
    dtlz(unsigned prob_id = 1u, vector_double::size_type dim = 5u, vector_double::size_type fdim = 3u,
         unsigned alpha = 100u)
        : m_prob_id(prob_id), m_alpha(alpha), m_dim(dim), m_fdim(fdim)
    {
        if (prob_id == 0u || prob_id > 7u) {
            throw std::invalid_argument("DTLZ test suite contains seven (prob_id = [1 ... 7]) problems, prob_id="
                                        + std::to_string(prob_id) + " was detected");
        }
        if (fdim < 2u) {
            throw std::invalid_argument("DTLZ test problem have a minimum of 2 objectives: fdim="
                                        + std::to_string(fdim) + " was detected");
        }
        if (dim <= fdim) {
            throw std::invalid_argument("The problem dimension has to be larger than the number of objectives.");
        }
    }
    vector_double fitness(const vector_double &x) const
    {
        return oracle::dtlz_fitness(m_prob_id, m_fdim, m_alpha, x);
    }
    std::pair<vector_double, vector_double> get_bounds() const
    {
        return {vector_double(m_dim, 0.), vector_double(m_dim, 1.)};
    }
    vector_double::size_type get_nobj() const
    {
        return m_fdim;
    }
    std::string get_name() const
    {
        return "DTLZ" + std::to_string(m_prob_id);
    }

private:
    unsigned m_prob_id;
    unsigned m_alpha;
    vector_double::size_type m_dim;
    vector_double::size_type m_fdim;
};

// ---- problem ---------------------------------------------------------------
namespace detail
{
template <typename T, typename = void>
struct has_bf : std::false_type {
};
template <typename T>
struct has_bf<T, std::void_t<decltype(std::declval<const T &>().batch_fitness(std::declval<const vector_double &>()))>>
    : std::true_type {
};
template <typename T, typename = void>
struct has_grad : std::false_type {
};
template <typename T>
struct has_grad<T, std::void_t<decltype(std::declval<const T &>().gradient(std::declval<const vector_double &>()))>>
    : std::true_type {
};
template <typename T, typename = void>
struct has_ts : std::false_type {
};
template <typename T>
struct has_ts<T, std::void_t<decltype(std::declval<const T &>().get_thread_safety())>> : std::true_type {
};
template <typename T, typename = void>
struct has_name : std::false_type {
};
template <typename T>
struct has_name<T, std::void_t<decltype(std::declval<const T &>().get_name())>> : std::true_type {
};
} // namespace detail

class problem
{
    struct base {
        virtual ~base() = default;
        virtual std::unique_ptr<base> clone() const = 0;
        virtual vector_double fitness(const vector_double &) const = 0;
        virtual vector_double batch_fitness(const vector_double &) const = 0;
        virtual bool has_batch_fitness() const = 0;
        virtual vector_double gradient(const vector_double &) const = 0;
        virtual bool has_gradient() const = 0;
        virtual std::pair<vector_double, vector_double> get_bounds() const = 0;
        virtual thread_safety get_thread_safety() const = 0;
        virtual std::string get_name() const = 0;
    };
    template <typename T>
    struct inner final : base {
        T m_value;
        explicit inner(T v) : m_value(std::move(v)) {}
        std::unique_ptr<base> clone() const override
        {
            return std::make_unique<inner>(m_value);
        }
        vector_double fitness(const vector_double &x) const override
        {
            return m_value.fitness(x);
        }
        vector_double batch_fitness(const vector_double &x) const override
        {
            if constexpr (detail::has_bf<T>::value) return m_value.batch_fitness(x);
            throw std::logic_error("no batch_fitness");
        }
        bool has_batch_fitness() const override
        {
            return detail::has_bf<T>::value;
        }
        vector_double gradient(const vector_double &x) const override
        {
            if constexpr (detail::has_grad<T>::value) {
                return m_value.gradient(x);
            } else {
                (void)x;
                throw not_implemented_error("The gradient has been requested, but it is not implemented in a UDP "
                                            "of type '"
                                            + get_name() + "'");
            }
        }
        bool has_gradient() const override
        {
            return detail::has_grad<T>::value;
        }
        std::pair<vector_double, vector_double> get_bounds() const override
        {
            return m_value.get_bounds();
        }
        thread_safety get_thread_safety() const override
        {
            if constexpr (detail::has_ts<T>::value) return m_value.get_thread_safety();
            return thread_safety::basic;
        }
        std::string get_name() const override
        {
            if constexpr (detail::has_name<T>::value) return m_value.get_name();
            return typeid(T).name();
        }
    };

public:
    template <typename T, std::enable_if_t<!std::is_same<std::decay_t<T>, problem>::value, int> = 0>
    explicit problem(T udp) : m_ptr(std::make_unique<inner<T>>(std::move(udp)))
    {
        const auto b = m_ptr->get_bounds();
        m_nx = b.first.size();
        m_nf = m_ptr->fitness(b.first).size();
    }
    problem(const problem &o)
        : m_ptr(o.m_ptr->clone()), m_nx(o.m_nx), m_nf(o.m_nf), m_fevals(o.m_fevals.load()),
          m_gevals(o.m_gevals.load())
    {
    }
    problem &operator=(const problem &o)
    {
        m_ptr = o.m_ptr->clone();
        m_nx = o.m_nx;
        m_nf = o.m_nf;
        m_fevals = o.m_fevals.load();
        m_gevals = o.m_gevals.load();
        return *this;
    }
    vector_double fitness(const vector_double &dv) const
    {
        if (dv.size() != m_nx) throw std::invalid_argument("A decision vector is incompatible with a problem");
        auto f = m_ptr->fitness(dv);
        ++m_fevals;
        return f;
    }
    bool has_batch_fitness() const
    {
        return m_ptr->has_batch_fitness();
    }
    vector_double batch_fitness(const vector_double &dvs) const
    {
        auto r = m_ptr->batch_fitness(dvs);
        m_fevals += dvs.size() / m_nx;
        return r;
    }
    bool has_gradient() const
    {
        return m_ptr->has_gradient();
    }
    vector_double gradient(const vector_double &dv) const
    {
        if (dv.size() != m_nx) throw std::invalid_argument("A decision vector is incompatible with a problem");
        auto g = m_ptr->gradient(dv);
        ++m_gevals;
        return g;
    }
    unsigned long long get_gevals() const
    {
        return m_gevals.load();
    }
    std::pair<vector_double, vector_double> get_bounds() const
    {
        return m_ptr->get_bounds();
    }
    vector_double::size_type get_nx() const
    {
        return m_nx;
    }
    vector_double::size_type get_nf() const
    {
        return m_nf;
    }
    thread_safety get_thread_safety() const
    {
        return m_ptr->get_thread_safety();
    }
    std::string get_name() const
    {
        return m_ptr->get_name();
    }
    template <typename T>
    const T *extract() const noexcept
    {
        auto p = dynamic_cast<const inner<T> *>(m_ptr.get());
        return p ? &p->m_value : nullptr;
    }
    void increment_fevals(unsigned long long n) const
    {
        m_fevals += n;
    }
    unsigned long long get_fevals() const
    {
        return m_fevals.load();
    }

private:
    std::unique_ptr<base> m_ptr;
    vector_double::size_type m_nx = 0, m_nf = 0;
    mutable std::atomic<unsigned long long> m_fevals{0};
    mutable std::atomic<unsigned long long> m_gevals{0};

    friend vector_double member_batch_fitness(const problem &, const vector_double &);
};

// the UDP's batch_fitness without counting (prob_invoke_mem_batch_fitness(p,
// dvs, false), src/problem.cpp:972-984)
inline vector_double member_batch_fitness(const problem &p, const vector_double &dvs)
{
    if (!p.has_batch_fitness()) {
        throw not_implemented_error("The batch_fitness() method has been invoked, but it is not implemented in a UDP "
                                    "of type '"
                                    + p.get_name() + "'");
    }
    return p.m_ptr->batch_fitness(dvs);
}

// ---- batch evaluators --------------------------------------------------------
class thread_bfe
{
public:
    // src/batch_evaluators/thread_bfe.cpp:64-144 (std::thread static split)
    vector_double operator()(problem p, const vector_double &dvs) const
    {
        const auto nx = p.get_nx(), nf = p.get_nf(), n = dvs.size() / nx;
        if (p.get_thread_safety() < thread_safety::basic) {
            throw std::invalid_argument("Cannot use a thread_bfe on the problem '" + p.get_name()
                                        + "', which does not provide the required level of thread safety");
        }
        vector_double out(n * nf);
        const unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < nt; ++t) {
            pool.emplace_back([&, t, p]() {
                vector_double tmp(nx);
                for (auto i = n * t / nt; i < n * (t + 1) / nt; ++i) {
                    std::copy(dvs.begin() + static_cast<std::ptrdiff_t>(i * nx),
                              dvs.begin() + static_cast<std::ptrdiff_t>((i + 1) * nx), tmp.begin());
                    const auto f = p.fitness(tmp);
                    std::copy(f.begin(), f.end(), out.begin() + static_cast<std::ptrdiff_t>(i * nf));
                }
            });
        }
        for (auto &th : pool) th.join();
        return out;
    }
    std::string get_name() const
    {
        return "Multi-threaded batch fitness evaluator";
    }
};

class member_bfe
{
public:
    // src/batch_evaluators/member_bfe.cpp:40-45: fevals are left to pagmo::bfe
    vector_double operator()(const problem &p, const vector_double &dvs) const
    {
        return member_batch_fitness(p, dvs);
    }
};

class bfe
{
public:
    template <typename T, std::enable_if_t<!std::is_same<std::decay_t<T>, bfe>::value, int> = 0>
    explicit bfe(T udbfe)
        : m_f([u = std::move(udbfe)](const problem &p, const vector_double &dvs) { return u(p, dvs); })
    {
    }
    // src/bfe.cpp:91-110 with the checks of src/detail/bfe_impl.cpp:63-129
    vector_double operator()(const problem &p, const vector_double &dvs) const
    {
        oracle::udp shape;
        const auto nx = p.get_nx(), nf = p.get_nf();
        if (dvs.size() % nx) {
            throw std::invalid_argument("Invalid argument for a batch fitness evaluation: the length of the vector "
                                        "representing the decision vectors, "
                                        + std::to_string(dvs.size())
                                        + ", is not an exact multiple of the dimension of the problem, "
                                        + std::to_string(nx));
        }
        auto r = m_f(p, dvs);
        if (r.size() % nf || r.size() / nf != dvs.size() / nx) {
            throw std::invalid_argument("An invalid result was produced by a batch fitness evaluation");
        }
        p.increment_fevals(dvs.size() / nx);
        return r;
    }

private:
    std::function<vector_double(const problem &, const vector_double &)> m_f;
};

// ---- multi-objective utilities -------------------------------------------------
inline fnds_return_type fast_non_dominated_sorting(const std::vector<vector_double> &points)
{
    const auto n = points.size(), m = n ? points[0].size() : 0u;
    std::vector<double> flat;
    for (const auto &v : points) flat.insert(flat.end(), v.begin(), v.end());
    auto r = oracle::fast_non_dominated_sorting(flat.data(), n, m);
    return std::make_tuple(r.fronts, r.dom_list, r.dom_count, r.rank);
}

inline vector_double crowding_distance(const std::vector<vector_double> &front)
{
    const auto n = front.size(), m = n ? front[0].size() : 0u;
    std::vector<double> flat;
    for (const auto &v : front) flat.insert(flat.end(), v.begin(), v.end());
    return oracle::crowding_distance(flat.data(), n, m);
}

inline std::vector<pop_size_t> sort_population_mo(const std::vector<vector_double> &f)
{
    const auto n = f.size(), m = n ? f[0].size() : 0u;
    std::vector<double> flat;
    for (const auto &v : f) flat.insert(flat.end(), v.begin(), v.end());
    return oracle::sort_population_mo(flat.data(), n, m);
}

inline std::vector<pop_size_t> select_best_N_mo(const std::vector<vector_double> &f, pop_size_t N)
{
    const auto n = f.size(), m = n ? f[0].size() : 0u;
    std::vector<double> flat;
    for (const auto &v : f) flat.insert(flat.end(), v.begin(), v.end());
    return oracle::select_best_N_mo(flat.data(), n, m, N);
}

// ---- population and pso_gen ------------------------------------------------------
// ---- translate (translate.hpp / translate.cpp) -----------------------------
class translate
{
public:
    template <typename T>
    translate(T udp, const vector_double &translation) : m_problem(std::move(udp)), m_translation(translation)
    {
        if (translation.size() != m_problem.get_nx()) {
            throw std::invalid_argument("Length of shift vector is: " + std::to_string(translation.size())
                                        + " while the problem dimension is: " + std::to_string(m_problem.get_nx()));
        }
    }
    vector_double fitness(const vector_double &x) const
    {
        return m_problem.fitness(translate_back(x));
    }
    vector_double gradient(const vector_double &x) const
    {
        return m_problem.gradient(translate_back(x));
    }
    std::pair<vector_double, vector_double> get_bounds() const
    {
        auto b = m_problem.get_bounds();
        for (std::size_t i = 0; i < b.first.size(); ++i) {
            b.first[i] += m_translation[i];
            b.second[i] += m_translation[i];
        }
        return b;
    }
    std::string get_name() const
    {
        return m_problem.get_name() + " [translated]";
    }
    thread_safety get_thread_safety() const
    {
        return m_problem.get_thread_safety();
    }
    const problem &get_inner_problem() const
    {
        return m_problem;
    }
    const vector_double &get_translation() const
    {
        return m_translation;
    }

private:
    vector_double translate_back(const vector_double &x) const
    {
        vector_double xs(x.size());
        std::transform(x.begin(), x.end(), m_translation.begin(), xs.begin(), std::minus<double>());
        return xs;
    }
    problem m_problem;
    vector_double m_translation;
};

inline oracle::udp to_oracle(const problem &p)
{
    if (auto q = p.extract<rosenbrock>()) return q->as_udp();
    if (auto q = p.extract<rastrigin>()) return q->as_udp();
    if (auto q = p.extract<ackley>()) return q->as_udp();
    if (auto q = p.extract<schwefel>()) return q->as_udp();
    if (auto q = p.extract<griewank>()) return q->as_udp();
    if (auto q = p.extract<lennard_jones>()) return q->as_udp();
    throw std::invalid_argument("the stand-in pso_gen supports the built-in problems only");
}

class population
{
public:
    population(problem p, pop_size_t size, unsigned seed) : m_prob(std::move(p))
    {
        vector_double X, F;
        oracle::population_init(to_oracle(m_prob), size, seed, X, F);
        const auto nx = m_prob.get_nx(), nf = m_prob.get_nf();
        for (pop_size_t i = 0; i < size; ++i) {
            m_x.emplace_back(X.begin() + static_cast<std::ptrdiff_t>(i * nx), X.begin() + static_cast<std::ptrdiff_t>((i + 1) * nx));
            m_f.emplace_back(F.begin() + static_cast<std::ptrdiff_t>(i * nf), F.begin() + static_cast<std::ptrdiff_t>((i + 1) * nf));
        }
        m_prob.increment_fevals(size);
    }
    const std::vector<vector_double> &get_x() const
    {
        return m_x;
    }
    const std::vector<vector_double> &get_f() const
    {
        return m_f;
    }
    const problem &get_problem() const
    {
        return m_prob;
    }
    void set_xf(pop_size_t i, const vector_double &x, const vector_double &f)
    {
        m_x[i] = x;
        m_f[i] = f;
    }
    pop_size_t size() const
    {
        return m_x.size();
    }

private:
    problem m_prob;
    std::vector<vector_double> m_x, m_f;
};

class pso_gen
{
public:
    explicit pso_gen(unsigned gen = 1u, double omega = 0.7298, double eta1 = 2.05, double eta2 = 2.05,
                     double max_vel = 0.5, unsigned variant = 5u, unsigned neighb_type = 2u,
                     unsigned neighb_param = 4u, bool memory = false, unsigned seed = 0u)
    {
        oracle::pso_params p;
        p.gen = gen;
        p.omega = omega;
        p.eta1 = eta1;
        p.eta2 = eta2;
        p.max_vel = max_vel;
        p.variant = variant;
        p.neighb_type = neighb_type;
        p.neighb_param = neighb_param;
        p.memory = memory;
        p.seed = seed;
        m_algo = std::make_shared<oracle::pso_gen>(p);
    }
    void set_bfe(const bfe &b)
    {
        m_bfe = std::make_shared<bfe>(b);
    }
    population evolve(population pop) const
    {
        const problem &prob = pop.get_problem();
        const auto nx = prob.get_nx();
        vector_double X, F;
        for (pop_size_t i = 0; i < pop.size(); ++i) {
            X.insert(X.end(), pop.get_x()[i].begin(), pop.get_x()[i].end());
            F.push_back(pop.get_f()[i][0]);
        }
        oracle::batch_eval_fn fn = [this, &prob](const oracle::udp &, const vector_double &dvs) {
            return (*m_bfe)(prob, dvs);
        };
        m_algo->evolve(to_oracle(prob), X, F, m_bfe ? &fn : nullptr);
        for (pop_size_t i = 0; i < pop.size(); ++i) {
            pop.set_xf(i, vector_double(X.begin() + static_cast<std::ptrdiff_t>(i * nx), X.begin() + static_cast<std::ptrdiff_t>((i + 1) * nx)),
                       vector_double{F[i]});
        }
        return pop;
    }

private:
    std::shared_ptr<oracle::pso_gen> m_algo;
    std::shared_ptr<bfe> m_bfe;
};

} // namespace pagmo
