// test_hip_bfe.cpp -- hip_bfe as a drop-in pagmo UDBFE.
//
//   test_hip_bfe cpu   contract checks that need no GPU (type requirements,
//                      fallback to the CPU evaluators, validation messages,
//                      the loud failure of a supported problem without device)
//   test_hip_bfe gpu   parity with thread_bfe through pagmo::bfe, fevals
//                      accounting, pso_gen with set_bfe (tests/pso_gen.cpp:
//                      302-318 style), batch_gradient, fast_non_dominated_sorting
//
// Built against pagmo's headers when they exist, and otherwise against the
// stand-in in tests/cpp/pagmo_compat.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include <pagmo/algorithms/pso_gen.hpp>
#include <pagmo/population.hpp>

#include "pagmo_hip/hip_bfe.hpp"

using namespace pagmo;
using pagmo_hip::hip_bfe;

static int g_fail = 0;
#define CHECK(cond)                                                                                                  \
    do {                                                                                                             \
        if (!(cond)) {                                                                                               \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);                                           \
            ++g_fail;                                                                                                \
        }                                                                                                            \
    } while (0)

template <typename E, typename F>
static std::string throws(F &&f)
{
    try {
        f();
    } catch (const E &e) {
        return e.what();
    } catch (...) {
        return "<other exception>";
    }
    return "<no exception>";
}

static vector_double random_dvs(const problem &p, std::size_t n, unsigned seed)
{
    std::mt19937 e(seed);
    const auto b = p.get_bounds();
    vector_double out(n * p.get_nx());
    for (std::size_t i = 0; i < out.size(); ++i) {
        const auto j = i % p.get_nx();
        out[i] = std::uniform_real_distribution<double>(b.first[j], b.second[j])(e);
    }
    return out;
}

// A user problem with no device kernel (thread safety basic by default).
struct user_udp {
    vector_double fitness(const vector_double &x) const
    {
        return {x[0] * x[0] + 3. * x[1]};
    }
    std::pair<vector_double, vector_double> get_bounds() const
    {
        return {{-1., -1.}, {1., 1.}};
    }
};

// A UDP with batch_fitness (tests/member_bfe.cpp:64-77): default_bfe's
// first choice, member_bfe (default_bfe.cpp:56-58).
struct batch_udp {
    vector_double fitness(const vector_double &) const
    {
        return {0};
    }
    std::pair<vector_double, vector_double> get_bounds() const
    {
        return {{0}, {1}};
    }
    vector_double batch_fitness(const vector_double &dvs) const
    {
        return vector_double(dvs.size(), 1.);
    }
};

// A user problem with a gradient and no device kernel.
struct grad_udp {
    vector_double fitness(const vector_double &x) const
    {
        return {x[0] * x[0] + 3. * x[1] * x[2]};
    }
    vector_double gradient(const vector_double &x) const
    {
        return {2. * x[0], 3. * x[2], 3. * x[1]};
    }
    std::pair<vector_double, vector_double> get_bounds() const
    {
        return {{-1., -1., -1.}, {1., 1., 1.}};
    }
    std::string get_name() const
    {
        return "grad_udp";
    }
};

// A problem that forbids concurrent use (tests/thread_bfe.cpp:99-116).
struct unsafe_udp {
    vector_double fitness(const vector_double &) const
    {
        return {1.};
    }
    std::pair<vector_double, vector_double> get_bounds() const
    {
        return {{0.}, {1.}};
    }
    thread_safety get_thread_safety() const
    {
        return thread_safety::none;
    }
    std::string get_name() const
    {
        return "unsafe_prob";
    }
};

static void cpu_tests()
{
    // UDBFE requirements (bfe.hpp:97-108)
    static_assert(std::is_default_constructible<hip_bfe>::value, "");
    static_assert(std::is_copy_constructible<hip_bfe>::value, "");
    static_assert(std::is_move_constructible<hip_bfe>::value, "");
    static_assert(std::is_destructible<hip_bfe>::value, "");
    CHECK(hip_bfe{}.get_thread_safety() == thread_safety::basic);
    CHECK(hip_bfe{}.get_name().find("HIP") != std::string::npos);

    // problems without a kernel go through default_bfe's dispatch
    problem up{user_udp{}};
    pgh_udp u;
    CHECK(!pagmo_hip::describe(up, u));
    const auto dvs = random_dvs(up, 100, 1);
    bfe b{hip_bfe{}};
    const auto f = b(up, dvs);
    CHECK(f.size() == 100u);
    bool same = true;
    for (std::size_t i = 0; i < 100; ++i) {
        same = same && f[i] == user_udp{}.fitness({dvs[2 * i], dvs[2 * i + 1]})[0];
    }
    CHECK(same);
    // fevals: the problem's per-call counter through thread_bfe's copies is not
    // shared; the wrapper adds n_dvs once (bfe.cpp:102-107)
    CHECK(up.get_fevals() == 100u);

    // a UDP with batch_fitness goes to member_bfe, and pagmo::bfe counts its
    // fevals once (tests/member_bfe.cpp:84-87)
    {
        problem bp{batch_udp{}};
        bfe bb{hip_bfe{}};
        CHECK(bp.get_fevals() == 0u);
        CHECK(bb(bp, {1., 2., 3.}) == vector_double(3, 1.));
        CHECK(bp.get_fevals() == 3u);
        CHECK(hip_bfe{}(bp, {0.5, 0.25}) == vector_double(2, 1.));
        CHECK(bp.get_fevals() == 3u); // the UDBFE itself never counts
    }

    // batch_gradient of a problem with a gradient but no kernel: problem::
    // gradient row by row (src/problem.cpp:456-467), one gevals per row
    {
        problem gp{grad_udp{}};
        const auto dvs = random_dvs(gp, 37, 4);
        const auto g = pagmo_hip::batch_gradient(gp, dvs);
        bool same = g.size() == dvs.size();
        for (std::size_t i = 0; i < 37u && same; ++i) {
            const auto gi = grad_udp{}.gradient({dvs[3 * i], dvs[3 * i + 1], dvs[3 * i + 2]});
            for (std::size_t j = 0; j < 3u; ++j) same = same && gi[j] == g[3 * i + j];
        }
        CHECK(same);
        CHECK(gp.get_gevals() == 37u);
        CHECK(pagmo_hip::batch_gradient(gp, vector_double{}).empty());
        const auto mlen = throws<std::invalid_argument>([&] { pagmo_hip::batch_gradient(gp, vector_double(4, 0.)); });
        CHECK(mlen.find("is not an exact multiple of the dimension of the problem, 3") != std::string::npos);
        // a UDP without a gradient, with or without a fitness kernel, throws
        // problem::gradient's not_implemented_error
        const auto m4 = throws<not_implemented_error>(
            [&] { pagmo_hip::batch_gradient(problem{user_udp{}}, vector_double(4, 0.)); });
        CHECK(m4.find("The gradient has been requested, but it is not implemented in a UDP of type") == 0u);
        const auto m5 = throws<not_implemented_error>(
            [&] { pagmo_hip::batch_gradient(problem{ackley{3u}}, vector_double(6, 0.)); });
        CHECK(m5.find("not implemented in a UDP of type 'Ackley Function'") != std::string::npos);
    }

    problem unsafe{unsafe_udp{}};
    const auto msg = throws<std::invalid_argument>([&] { hip_bfe{}(unsafe, vector_double(4, 0.5)); });
    CHECK(msg.find("Cannot execute fitness evaluations in batch mode for a problem of type 'unsafe_prob'")
          != std::string::npos);

    // supported problems are recognised from the UDP type
    CHECK(pagmo_hip::describe(problem{rosenbrock{10u}}, u) && u.kind == PGH_ROSENBROCK && u.dim == 10u);
    CHECK(pagmo_hip::describe(problem{lennard_jones{48u}}, u) && u.kind == PGH_LENNARD_JONES && u.dim == 48u);
    CHECK(pagmo_hip::describe(problem{ackley{7u}}, u) && u.kind == PGH_ACKLEY && u.dim == 7u);

    // the library validates the batch like bfe_check_input_dvs (bfe_impl.cpp:63-88)
    const auto m2 = throws<std::invalid_argument>([&] { hip_bfe{}(problem{rosenbrock{3u}}, vector_double(7, 0.)); });
    CHECK(m2 == "Invalid argument for a batch fitness evaluation: the length of the vector representing the decision "
                "vectors, 7, is not an exact multiple of the dimension of the problem, 3");

    if (!pgh_device_available()) {
        // a supported problem with no device fails loudly
        const auto m3 = throws<std::runtime_error>([&] { hip_bfe{}(problem{rosenbrock{3u}}, vector_double(6, 0.)); });
        CHECK(m3.find("libpagmo_hip") != std::string::npos);
    }
}

static double rel(double a, double b)
{
    return std::abs(a - b) / std::max(std::abs(b), 1e-300);
}

// 4 sum over pairs of |r^-12| + |r^-6| of one decision vector of a
// Lennard-Jones cluster (lennard_jones.cpp:132-151 encoding): the scale of
// the cancelling sum that the energy is
static double lj_abs_scale(const double *x, std::size_t nx)
{
    const std::size_t atoms = (nx + 6u) / 3u;
    std::vector<double> p(3u * atoms, 0.);
    p[5] = x[0];
    p[7] = x[1];
    p[8] = x[2];
    for (std::size_t i = 3; i < nx; ++i) p[6u + (i - 3u)] = x[i];
    double s = 0.;
    for (std::size_t i = 0; i < atoms; ++i) {
        for (std::size_t j = i + 1u; j < atoms; ++j) {
            double r2 = 0.;
            for (int c = 0; c < 3; ++c) r2 += (p[3u * i + c] - p[3u * j + c]) * (p[3u * i + c] - p[3u * j + c]);
            const double s6 = 1. / (r2 * r2 * r2);
            s += std::abs(s6 * s6) + std::abs(s6);
        }
    }
    return 4. * s;
}

// exact: bit-identical; otherwise within tol relative, or for Lennard-Jones
// within tol of the sum of the absolute pair terms
template <typename UDP>
static void parity(UDP udp, std::size_t n, bool exact, double tol)
{
    problem p{udp};
    const auto dvs = random_dvs(p, n, 7);
    bfe hb{hip_bfe{}}, tb{thread_bfe{}};
    const auto a = hb(p, dvs);
    const auto b = tb(p, dvs);
    CHECK(a.size() == b.size());
    const bool lj = std::is_same<UDP, lennard_jones>::value;
    const std::size_t nx = p.get_nx();
    double worst = 0.;
    for (std::size_t i = 0; i < a.size(); ++i) {
        if (exact) {
            CHECK(a[i] == b[i]);
        } else if (lj) {
            worst = std::max(worst, a[i] == b[i] ? 0. : std::abs(a[i] - b[i]) / lj_abs_scale(&dvs[i * nx], nx));
        } else {
            worst = std::max(worst, rel(a[i], b[i]));
        }
    }
    if (!exact && !(worst <= tol)) {
        std::printf("FAILED parity %s: worst rel %.3e\n", p.get_name().c_str(), worst);
        ++g_fail;
    }
    // the wrapper counted both calls, each once (tests/thread_bfe.cpp:66-96)
    CHECK(p.get_fevals() == 2u * n);
}

// Islands call their bfe from their own threads (island.cpp:438): 8 host
// threads, each with its own problems and batches (some large enough for the
// pinned two-chunk upload), call hip_bfe, batch_gradient and
// fast_non_dominated_sorting at once; every result equals the serial call's.
static void concurrent_tests()
{
    struct job {
        problem p;
        vector_double dvs;
        vector_double f, g;
        pagmo::fnds_return_type nd;
    };
    std::vector<job> jobs;
    for (unsigned t = 0; t < 8u; ++t) {
        problem p = t % 3u == 0u   ? problem{rosenbrock{10u + 60u * t}}
                    : t % 3u == 1u ? problem{rastrigin{500u + 100u * t}}
                                   : problem{lennard_jones{20u + t}};
        const std::size_t n = t == 1u || t == 4u ? 6000u : 700u + 50u * t; // 6000 x 600+ doubles > 16 MiB
        jobs.push_back(job{p, random_dvs(p, n, 100u + t), {}, {}, {}});
    }
    std::vector<job> serial = jobs;
    auto run = [](job &j, unsigned t) {
        j.f = hip_bfe{}(j.p, j.dvs);
        if (t % 3u != 1u) j.g = pagmo_hip::batch_gradient(j.p, j.dvs);
        std::vector<vector_double> pts;
        for (std::size_t i = 0; i + 1u < j.f.size(); i += 2u) pts.push_back({j.f[i], j.f[i + 1u], double(i % 5u)});
        j.nd = pagmo_hip::fast_non_dominated_sorting(pts);
    };
    for (unsigned t = 0; t < 8u; ++t) run(serial[t], t);
    for (int round = 0; round < 3; ++round) {
        std::vector<job> par = jobs;
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < 8u; ++t) pool.emplace_back([&, t] { run(par[t], t); });
        for (auto &th : pool) th.join();
        for (unsigned t = 0; t < 8u; ++t) {
            CHECK(par[t].f == serial[t].f);
            CHECK(par[t].g == serial[t].g);
            CHECK(par[t].nd == serial[t].nd);
        }
    }
}

static void gpu_tests()
{
    if (!pgh_device_available()) {
        std::printf("FAILED: no gfx950 device\n");
        ++g_fail;
        return;
    }
    parity(rosenbrock{10u}, 1000, true, 0.);
    parity(rosenbrock{501u}, 300, true, 0.);
    parity(rastrigin{1000u}, 512, false, 1e-12);
    parity(ackley{1000u}, 512, false, 1e-12);
    parity(schwefel{33u}, 700, false, 1e-12);
    parity(griewank{100u}, 700, false, 1e-12);
    parity(lennard_jones{5u}, 300, false, 1e-12);
    parity(lennard_jones{48u}, 400, false, 1e-12);

    // pso_gen with bfe equals pso_gen without (tests/pso_gen.cpp:302-318), and
    // hip_bfe reproduces thread_bfe's trajectory bit for bit on rosenbrock
    {
        population pop{problem{rosenbrock{10u}}, 64u, 42u};
        pso_gen a1{100u, 0.7298, 2.05, 2.05, 0.5, 5u, 2u, 4u, false, 42u};
        pso_gen a2{100u, 0.7298, 2.05, 2.05, 0.5, 5u, 2u, 4u, false, 42u};
        pso_gen a3{100u, 0.7298, 2.05, 2.05, 0.5, 5u, 2u, 4u, false, 42u};
        a1.set_bfe(bfe{hip_bfe{}});
        a2.set_bfe(bfe{thread_bfe{}});
        const auto p1 = a1.evolve(pop), p2 = a2.evolve(pop), p3 = a3.evolve(pop);
        CHECK(p1.get_f() == p2.get_f());
        CHECK(p1.get_x() == p2.get_x());
        CHECK(p1.get_f() == p3.get_f());
    }

    // batched gradient: rosenbrock bit-exact with its UDP gradient
    {
        problem p{rosenbrock{7u}};
        const auto dvs = random_dvs(p, 50, 3);
        const auto g = pagmo_hip::batch_gradient(p, dvs);
        bool same = true;
        for (std::size_t i = 0; i < 50; ++i) {
            const auto gi = p.gradient(vector_double(dvs.begin() + static_cast<std::ptrdiff_t>(7 * i),
                                                     dvs.begin() + static_cast<std::ptrdiff_t>(7 * i + 7)));
            for (std::size_t j = 0; j < 7; ++j) same = same && gi[j] == g[7 * i + j];
        }
        CHECK(same);
    }

    // fast_non_dominated_sorting: the reference's known answer and a tied random set
    {
        std::vector<vector_double> ex = {{0, 7}, {1, 5}, {2, 3}, {4, 2}, {7, 1}, {10, 0},
                                         {2, 6}, {4, 4}, {10, 2}, {6, 6}, {9, 5}};
        CHECK(pagmo_hip::fast_non_dominated_sorting(ex) == fast_non_dominated_sorting(ex));
        std::mt19937 e(5);
        std::vector<vector_double> pts(2000, vector_double(3));
        for (auto &v : pts)
            for (auto &x : v) x = static_cast<double>(e() % 7u);
        CHECK(pagmo_hip::fast_non_dominated_sorting(pts) == fast_non_dominated_sorting(pts));
        const auto m = throws<std::invalid_argument>(
            [&] { pagmo_hip::fast_non_dominated_sorting(std::vector<vector_double>{{1., 2.}}); });
        CHECK(m == "At least two points are needed for fast_non_dominated_sorting: 1 detected.");
    }

    // crowding_distance / sort_population_mo / select_best_N_mo: the reference's
    // known answers (tests/multi_objective.cpp:147-262) and agreement with the
    // CPU restatement on a larger tie-free set
    {
        const double inf = std::numeric_limits<double>::infinity();
        CHECK((pagmo_hip::crowding_distance({{0, 0}, {-1, 1}, {2, -2}}) == vector_double{2, inf, inf}));
        CHECK((pagmo_hip::crowding_distance({{0, 0}, {1, -1}, {2, -2}, {4, -4}}) == vector_double{inf, 1., 1.5, inf}));
        CHECK((pagmo_hip::crowding_distance({{0, 0}, {0, 0}}) == vector_double{inf, inf}));
        CHECK(throws<std::invalid_argument>([] { pagmo_hip::crowding_distance({{1, 2}}); })
              == "A non dominated front must contain at least two points: 1 detected.");
        CHECK(throws<std::invalid_argument>([] { pagmo_hip::crowding_distance({{1}, {2}}); })
              == "Points in the non dominated front must contain at least two objectives: 1 detected.");
        CHECK(throws<std::invalid_argument>([] { pagmo_hip::crowding_distance({{2, 3}, {3, 4}, {2, 4, 5}}); })
              == "A non dominated front must contain points of uniform dimensionality. Some different sizes were "
                 "instead detected.");
        const std::vector<vector_double> ex = {{0, 7}, {1, 5}, {2, 3}, {4, 2}, {7, 1}, {10, 0},
                                               {2, 6}, {4, 4}, {10, 2}, {6, 6}, {9, 5}};
        CHECK((pagmo_hip::sort_population_mo(ex) == std::vector<pop_size_t>{0, 5, 4, 3, 1, 2, 6, 8, 7, 9, 10}));
        CHECK((pagmo_hip::select_best_N_mo({{1, 2}, {2, 4}, {-3, 2}, {-3, -3}}, 4u) == std::vector<pop_size_t>{0, 1, 2, 3}));
        CHECK(pagmo_hip::select_best_N_mo({{1, 2}, {3, 4}}, 0u).empty());
        std::mt19937 e(11);
        std::uniform_real_distribution<double> u(0., 1.);
        std::vector<vector_double> pts(600, vector_double(3));
        for (auto &v : pts)
            for (auto &x : v) x = u(e);
        CHECK(pagmo_hip::crowding_distance(std::vector<vector_double>(pts.begin(), pts.begin() + 40))
              == crowding_distance(std::vector<vector_double>(pts.begin(), pts.begin() + 40)));
        // equal (rank, crowding) pairs (the fronts' infinite ends) are ordered by
        // index on the device: the restatement's std::stable_sort form
        std::vector<double> flat;
        for (const auto &v : pts) flat.insert(flat.end(), v.begin(), v.end());
        CHECK(pagmo_hip::sort_population_mo(pts) == oracle::sort_population_mo(flat.data(), 600u, 3u, true));
        CHECK(pagmo_hip::select_best_N_mo(pts, 250u) == oracle::select_best_N_mo(flat.data(), 600u, 3u, 250u, true));
    }

    // ZDT and DTLZ through hip_bfe: bit-exact where only + - * / sqrt round
    // (ZDT1, ZDT2, ZDT5), else within 1e-12 relative of the UDP
    {
        auto check_udp = [](const problem &p, std::size_t n, bool exact) {
            const auto dvs = random_dvs(p, n, 21);
            const auto f = hip_bfe{}(p, dvs);
            const auto nx = p.get_nx(), nf = p.get_nf();
            bool ok = f.size() == n * nf;
            for (std::size_t i = 0; i < n && ok; ++i) {
                const auto fi = p.fitness(vector_double(dvs.begin() + static_cast<std::ptrdiff_t>(nx * i),
                                                        dvs.begin() + static_cast<std::ptrdiff_t>(nx * (i + 1))));
                for (std::size_t k = 0; k < nf; ++k) {
                    ok = ok && (exact ? fi[k] == f[nf * i + k]
                                      : std::abs(fi[k] - f[nf * i + k]) <= 1e-12 * std::abs(fi[k]) + 1e-300);
                }
            }
            return ok;
        };
        for (unsigned id = 1u; id <= 6u; ++id) {
            CHECK(check_udp(problem{zdt{id, 30u}}, 500, id == 1u || id == 2u || id == 5u));
        }
        for (unsigned id = 1u; id <= 7u; ++id) {
            CHECK(check_udp(problem{dtlz{id, 12u, 3u}}, 500, false));
        }
        CHECK(check_udp(problem{dtlz{4u, 10u, 4u, 9u}}, 300, false));
    }

    // translate around a kernel problem: the vector is subtracted on the
    // device (x - t rounds exactly as translate_back does), so rosenbrock and
    // ZDT1 stay bit-exact; the gradient is the inner one at x - t
    {
        const auto check_t = [](const problem &p, std::size_t n, bool exact) {
            const auto dvs = random_dvs(p, n, 5);
            const auto f = hip_bfe{}(p, dvs);
            const auto nx = p.get_nx(), nf = p.get_nf();
            bool ok = f.size() == n * nf;
            for (std::size_t i = 0; i < n && ok; ++i) {
                const auto fi = p.fitness(vector_double(dvs.begin() + static_cast<std::ptrdiff_t>(nx * i),
                                                        dvs.begin() + static_cast<std::ptrdiff_t>(nx * (i + 1))));
                for (std::size_t k = 0; k < nf; ++k) {
                    ok = ok && (exact ? fi[k] == f[nf * i + k] : std::abs(fi[k] - f[nf * i + k]) <= 1e-12 * std::abs(fi[k]));
                }
            }
            return ok;
        };
        std::mt19937 e(3);
        std::uniform_real_distribution<double> d(-1., 1.);
        vector_double t10(10), t50(50), t30(30);
        for (auto &v : t10) v = d(e);
        for (auto &v : t50) v = d(e);
        for (auto &v : t30) v = 0.1 * d(e);
        CHECK(check_t(problem{translate{rosenbrock{10u}, t10}}, 700, true));
        CHECK(check_t(problem{translate{rastrigin{50u}, t50}}, 700, false));
        CHECK(check_t(problem{translate{zdt{1u, 30u}, t30}}, 700, true));
        problem tr{translate{rosenbrock{10u}, t10}};
        const auto dvs = random_dvs(tr, 40, 6);
        const auto g = pagmo_hip::batch_gradient(tr, dvs);
        bool same = g.size() == dvs.size();
        for (std::size_t i = 0; i < 40u && same; ++i) {
            const auto gi = tr.gradient(vector_double(dvs.begin() + static_cast<std::ptrdiff_t>(10 * i),
                                                      dvs.begin() + static_cast<std::ptrdiff_t>(10 * i + 10)));
            for (std::size_t j = 0; j < 10u; ++j) same = same && gi[j] == g[10 * i + j];
        }
        CHECK(same);
    }

    // memetic local step: a hand-derived trace (rosenbrock at the origin:
    // g = (-2, 0); alpha 0.5 gives (0.5, 0), f = 6.5, rejected; alpha 0.25
    // gives (0.25, 0), f = 0.953125, accepted), then bit-exact parity with the
    // oracle's statement of the rule on a seeded population
    {
        population pop{problem{rosenbrock{2u}}, 1u, 1u};
        pop.set_xf(0u, {0., 0.}, {1.});
        const auto fe = pop.get_problem().get_fevals();
        pagmo_hip::local_descent(pop, 2u, 0.5);
        CHECK((pop.get_x()[0] == vector_double{0.25, 0.}));
        CHECK((pop.get_f()[0] == vector_double{0.953125}));
        CHECK(pop.get_problem().get_fevals() == fe + 3u);

        population p10{problem{rosenbrock{10u}}, 64u, 42u};
        std::vector<double> X;
        for (const auto &x : p10.get_x()) X.insert(X.end(), x.begin(), x.end());
        std::vector<double> F(64);
        oracle::udp ou;
        ou.kind = oracle::ROSENBROCK;
        ou.dim = 10u;
        const auto b = ou.get_bounds();
        oracle::local_descent(ou, X.data(), 64u, F.data(), b.first.data(), b.second.data(), 30u, 0.5);
        pagmo_hip::local_descent(p10, 30u, 0.5);
        bool same = true;
        for (std::size_t i = 0; i < 64u; ++i) {
            same = same && p10.get_f()[i][0] == F[i];
            for (std::size_t j = 0; j < 10u; ++j) same = same && p10.get_x()[i][j] == X[10u * i + j];
        }
        CHECK(same);
    }

    // the memetic loop: pso_gen generations (hip_bfe) interleaved with local
    // steps on lennard_jones; the local step never raises an individual's
    // fitness and leaves x inside the bounds, and pso_gen runs unchanged
    {
        problem prob{lennard_jones{8u}};
        population pop{prob, 32u, 7u};
        pso_gen algo{3u, 0.7298, 2.05, 2.05, 0.5, 5u, 2u, 4u, false, 7u};
        algo.set_bfe(bfe{hip_bfe{}});
        const auto b = prob.get_bounds();
        bool monotone = true, inside = true, consistent = true;
        for (int round = 0; round < 4; ++round) {
            pop = algo.evolve(pop);
            const auto before = pop.get_f();
            pagmo_hip::local_descent(pop, 10u, 0.05);
            const auto fx = hip_bfe{}(pop.get_problem(), [&] {
                vector_double all;
                for (const auto &x : pop.get_x()) all.insert(all.end(), x.begin(), x.end());
                return all;
            }());
            for (std::size_t i = 0; i < pop.size(); ++i) {
                monotone = monotone && pop.get_f()[i][0] <= before[i][0];
                consistent = consistent && pop.get_f()[i][0] == fx[i];
                for (std::size_t j = 0; j < b.first.size(); ++j) {
                    inside = inside && pop.get_x()[i][j] >= b.first[j] && pop.get_x()[i][j] <= b.second[j];
                }
            }
        }
        CHECK(monotone);
        CHECK(inside);
        CHECK(consistent);
    }
}

// ZDT / DTLZ descriptions (no device needed): ids from the names, ZDT5's
// param from nx, DTLZ4's alpha read back from one fitness probe
static void mo_problem_cpu_tests()
{
    pgh_udp u;
    CHECK(pagmo_hip::describe(problem{zdt{5u, 11u}}, u) && u.kind == PGH_ZDT && u.prob_id == 5u && u.dim == 11u);
    CHECK(pagmo_hip::describe(problem{zdt{3u, 17u}}, u) && u.prob_id == 3u && u.dim == 17u && u.nobj == 2u);
    CHECK(pagmo_hip::describe(problem{dtlz{2u, 9u, 4u}}, u) && u.kind == PGH_DTLZ && u.prob_id == 2u && u.dim == 9u
          && u.nobj == 4u);
    CHECK(pagmo_hip::describe(problem{dtlz{4u, 7u, 3u, 100u}}, u) && u.alpha == 100u);
    CHECK(pagmo_hip::describe(problem{dtlz{4u, 7u, 3u, 7u}}, u) && u.alpha == 7u);
    CHECK(pagmo_hip::describe(problem{dtlz{4u, 7u, 3u, 0u}}, u) && u.alpha == 0u);
    // an alpha the probe cannot pin falls back to the CPU evaluator, exactly
    problem big{dtlz{4u, 6u, 3u, 2000u}};
    CHECK(!pagmo_hip::describe(big, u));
    const auto dvs = random_dvs(big, 20, 9);
    const auto f = hip_bfe{}(big, dvs);
    bool same = f.size() == 60u;
    for (std::size_t i = 0; i < 20u && same; ++i) {
        const auto fi = big.fitness(vector_double(dvs.begin() + static_cast<std::ptrdiff_t>(6 * i),
                                                  dvs.begin() + static_cast<std::ptrdiff_t>(6 * i + 6)));
        for (std::size_t k = 0; k < 3u; ++k) same = same && fi[k] == f[3 * i + k];
    }
    CHECK(same);
}

// translate: forwarded with its vector when it wraps a kernel problem; a
// translate of a translate goes to the CPU evaluator, exactly
static void translate_cpu_tests()
{
    pgh_udp u;
    const vector_double t{0.5, -0.25, 1.0, 2.0};
    problem tp{translate{rosenbrock{4u}, t}};
    CHECK(pagmo_hip::describe(tp, u) && u.kind == PGH_ROSENBROCK && u.dim == 4u && u.shift != nullptr
          && u.shift[3] == 2.0);
    problem twice{translate{translate{rosenbrock{4u}, t}, vector_double{-0.5, 0.25, -1.0, -2.0}}};
    CHECK(!pagmo_hip::describe(twice, u));
    const auto dvs = random_dvs(twice, 30, 4);
    const auto f = hip_bfe{}(twice, dvs);
    bool same = f.size() == 30u;
    for (std::size_t i = 0; i < 30u && same; ++i) {
        same = twice.fitness(vector_double(dvs.begin() + static_cast<std::ptrdiff_t>(4 * i),
                                           dvs.begin() + static_cast<std::ptrdiff_t>(4 * i + 4)))[0]
               == f[i];
    }
    CHECK(same);
    CHECK(!pagmo_hip::describe(problem{translate{user_udp{}, vector_double{1., 2.}}}, u));
}

// local_descent's argument checks happen before any device work
static void local_descent_cpu_tests()
{
    population pa{problem{ackley{3u}}, 2u, 1u};
    CHECK(throws<std::runtime_error>([&] { pagmo_hip::local_descent(pa, 1u, 0.1); }).find("gradient has not been")
          != std::string::npos);
    population pr{problem{rosenbrock{3u}}, 2u, 1u};
    CHECK(throws<std::invalid_argument>([&] { pagmo_hip::local_descent(pr, 1u, -1.); }).find("initial step")
          != std::string::npos);
}

int main(int argc, char **argv)
{
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    cpu_tests();
    mo_problem_cpu_tests();
    translate_cpu_tests();
    local_descent_cpu_tests();
    if (gpu) {
        gpu_tests();
        concurrent_tests();
    }
    std::printf("%s: %d failure(s)\n", gpu ? "gpu" : "cpu", g_fail);
    return g_fail ? 1 : 0;
}
