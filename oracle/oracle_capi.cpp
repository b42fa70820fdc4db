// oracle_capi.cpp -- extern "C" face of oracle.hpp for the Python test suite
// and the cpu_baseline leg of bench.py.  TEST INFRASTRUCTURE ONLY: the product
// library never links this file.
//
// Conventions: matrices are row-major, sizes are uint64, every function returns
// 0 on success and a negative value after an exception (message available from
// orc_last_error()).
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>

#include "oracle.hpp"

namespace
{
thread_local std::string g_err;

oracle::udp make_udp(int kind, uint64_t dim, const double *centers, uint64_t nobj)
{
    oracle::udp p;
    p.kind = kind;
    if (kind == oracle::LENNARD_JONES) {
        p.atoms = static_cast<unsigned>(dim);
    } else {
        p.dim = static_cast<unsigned>(dim);
    }
    if (kind == oracle::MO_SPHERE) {
        p.nobj = static_cast<unsigned>(nobj);
        p.centers.assign(centers, centers + nobj * dim);
    }
    return p;
}

template <typename F>
int guarded(F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}
oracle::udp make_mo_udp(int kind, uint32_t prob_id, uint64_t dim, uint64_t nobj, uint32_t alpha)
{
    oracle::udp p;
    p.kind = kind;
    p.prob_id = prob_id;
    p.dim = static_cast<unsigned>(dim);
    p.nobj = static_cast<unsigned>(nobj);
    p.alpha = alpha;
    return p;
}

oracle::migr_rate make_rate(int fractional, uint64_t count, double fraction)
{
    oracle::migr_rate r;
    r.fractional = fractional != 0;
    r.count = count;
    r.fraction = fraction;
    oracle::check_migr_rate(r);
    return r;
}

oracle::ind_group make_group(uint64_t n, uint64_t nx, const uint64_t *ids, const double *X, const double *F)
{
    oracle::ind_group g;
    for (uint64_t i = 0; i < n; ++i) {
        g.id.push_back(ids ? ids[i] : i);
        g.x.emplace_back(X + i * nx, X + (i + 1u) * nx);
        g.f.push_back({F[i]});
    }
    return g;
}

void put_group(const oracle::ind_group &g, uint64_t nx, uint64_t *n, uint64_t *ids, double *X, double *F)
{
    *n = g.f.size();
    for (uint64_t i = 0; i < g.f.size(); ++i) {
        if (ids) ids[i] = g.id[i];
        std::memcpy(X + i * nx, g.x[i].data(), nx * sizeof(double));
        F[i] = g.f[i][0];
    }
}

} // namespace

extern "C" {

const char *orc_last_error(void)
{
    return g_err.c_str();
}

// n outputs of std::mt19937(seed) after discarding `skip` outputs.
int orc_mt19937_u32(uint32_t seed, uint64_t skip, uint64_t n, uint32_t *out)
{
    return guarded([&] {
        std::mt19937 e(seed);
        e.discard(skip);
        for (uint64_t i = 0; i < n; ++i) {
            out[i] = static_cast<uint32_t>(e());
        }
    });
}

// n draws of std::uniform_real_distribution<double>(a, b) on std::mt19937(seed).
int orc_uniform_real(uint32_t seed, double a, double b, uint64_t n, double *out)
{
    return guarded([&] {
        std::mt19937 e(seed);
        std::uniform_real_distribution<double> d(a, b);
        for (uint64_t i = 0; i < n; ++i) {
            out[i] = d(e);
        }
    });
}

// Problem metadata: nx, nf, bounds (lb/ub may be null).
int orc_problem_info(int kind, uint64_t dim, uint64_t nobj, uint64_t *nx, uint64_t *nf, double *lb, double *ub)
{
    return guarded([&] {
        std::vector<double> c(nobj * (kind == oracle::MO_SPHERE ? dim : 0), 0.);
        auto p = make_udp(kind, dim, c.data(), nobj);
        *nx = p.get_nx();
        *nf = p.get_nf();
        const auto b = p.get_bounds();
        if (lb) std::memcpy(lb, b.first.data(), b.first.size() * sizeof(double));
        if (ub) std::memcpy(ub, b.second.data(), b.second.size() * sizeof(double));
    });
}

// ZDT / DTLZ: problem shape and batch fitness (single-threaded).
int orc_mo_info(int kind, uint32_t prob_id, uint64_t dim, uint64_t nobj, uint32_t alpha, uint64_t *nx, uint64_t *nf,
                double *lb, double *ub)
{
    return guarded([&] {
        auto p = make_mo_udp(kind, prob_id, dim, nobj, alpha);
        *nx = p.get_nx();
        *nf = p.get_nf();
        const auto b = p.get_bounds();
        if (lb) std::memcpy(lb, b.first.data(), b.first.size() * sizeof(double));
        if (ub) std::memcpy(ub, b.second.data(), b.second.size() * sizeof(double));
    });
}

int orc_mo_fitness(int kind, uint32_t prob_id, uint64_t dim, uint64_t nobj, uint32_t alpha, const double *dvs,
                   uint64_t n_dvs, uint64_t nx, double *fvs)
{
    return guarded([&] {
        auto p = make_mo_udp(kind, prob_id, dim, nobj, alpha);
        const auto nf = p.get_nf();
        for (uint64_t i = 0; i < n_dvs; ++i) {
            const auto f = p.fitness(std::vector<double>(dvs + i * nx, dvs + (i + 1u) * nx));
            std::memcpy(fvs + i * nf, f.data(), nf * sizeof(double));
        }
    });
}

// Single fitness call on an arbitrary-length x (the reference's known answers
// evaluate e.g. a 3-vector on a 1-dimensional ackley).
int orc_fitness_one(int kind, uint64_t dim, const double *centers, uint64_t nobj, const double *x, uint64_t n,
                    double *f)
{
    return guarded([&] {
        auto p = make_udp(kind, dim, centers, nobj);
        const auto r = p.fitness(std::vector<double>(x, x + n));
        std::memcpy(f, r.data(), r.size() * sizeof(double));
    });
}

// bfe::operator() with thread_bfe on n_threads host threads.
int orc_bfe(int kind, uint64_t dim, const double *centers, uint64_t nobj, const double *dvs, uint64_t n_dvs,
            uint64_t nx, double *fvs, uint32_t n_threads)
{
    return guarded([&] {
        auto p = make_udp(kind, dim, centers, nobj);
        std::vector<double> in(dvs, dvs + n_dvs * nx);
        const auto r = oracle::bfe_call(p, in, n_threads);
        std::memcpy(fvs, r.data(), r.size() * sizeof(double));
    });
}

// Timed bfe: runs `reps` calls and returns the median wall time per call in *sec.
int orc_bfe_timed(int kind, uint64_t dim, const double *centers, uint64_t nobj, const double *dvs, uint64_t n_dvs,
                  uint64_t nx, double *fvs, uint32_t n_threads, uint32_t reps, double *sec)
{
    return guarded([&] {
        auto p = make_udp(kind, dim, centers, nobj);
        std::vector<double> in(dvs, dvs + n_dvs * nx);
        std::vector<double> t;
        std::vector<double> r;
        for (uint32_t k = 0; k < reps; ++k) {
            const auto t0 = std::chrono::steady_clock::now();
            r = oracle::bfe_call(p, in, n_threads);
            const auto t1 = std::chrono::steady_clock::now();
            t.push_back(std::chrono::duration<double>(t1 - t0).count());
        }
        std::sort(t.begin(), t.end());
        *sec = t[t.size() / 2];
        std::memcpy(fvs, r.data(), r.size() * sizeof(double));
    });
}

// Analytic gradient of every dv (n_dvs x nx).
int orc_gradient(int kind, uint64_t dim, const double *dvs, uint64_t n_dvs, uint64_t nx, double *grad)
{
    return guarded([&] {
        auto p = make_udp(kind, dim, nullptr, 1);
        for (uint64_t i = 0; i < n_dvs; ++i) {
            const auto g = p.gradient(std::vector<double>(dvs + i * nx, dvs + (i + 1) * nx));
            std::memcpy(grad + i * nx, g.data(), nx * sizeof(double));
        }
    });
}

// Memetic local step on n_dvs x nx decision vectors, in place; F gets f(x).
// lb and ub are both NULL (no clamping) or nx values each.
int orc_local_descent(int kind, uint64_t dim, double *dvs, uint64_t n_dvs, double *fvs, const double *lb,
                      const double *ub, uint32_t iters, double step0)
{
    return guarded([&] {
        auto p = make_udp(kind, dim, nullptr, 1);
        if (!p.has_gradient()) throw std::invalid_argument("the problem has no gradient");
        oracle::local_descent(p, dvs, n_dvs, fvs, lb, ub, iters, step0);
    });
}

// estimate_gradient (high = 0) or estimate_gradient_h (high = 1) of the
// problem's fitness at x; out has nf * nx entries (layout grad[j + i*nx]).
int orc_estimate_gradient(int kind, uint64_t dim, const double *centers, uint64_t nobj, const double *x, uint64_t nx,
                          double dx, int high, double *out)
{
    return guarded([&] {
        auto p = make_udp(kind, dim, centers, nobj);
        auto f = [&p](const std::vector<double> &v) { return p.fitness(v); };
        const std::vector<double> xv(x, x + nx);
        const auto g = high ? oracle::estimate_gradient_h(f, xv, dx) : oracle::estimate_gradient(f, xv, dx);
        std::memcpy(out, g.data(), g.size() * sizeof(double));
    });
}

// The known-answer problem of tests/gradients_and_hessians.cpp:120-142
// (dummy_problem_easy_grad), for pinning estimate_gradient(_h).
int orc_estimate_gradient_kat(const double *x, double dx, int high, double *out)
{
    return guarded([&] {
        auto f = [](const std::vector<double> &dv) {
            std::vector<double> r(3, 0.);
            r[0] = dv[0] + dv[1] * dv[1] + dv[2] * dv[2] + dv[3] * dv[3] * dv[3];
            r[1] = dv[1] * dv[2] * dv[3] * dv[0];
            r[2] = std::sin(dv[0]);
            return r;
        };
        const std::vector<double> xv(x, x + 4);
        const auto g = high ? oracle::estimate_gradient_h(f, xv, dx) : oracle::estimate_gradient(f, xv, dx);
        std::memcpy(out, g.data(), g.size() * sizeof(double));
    });
}

int orc_pareto_dominance(const double *a, const double *b, uint64_t m, int *out)
{
    return guarded([&] { *out = oracle::pareto_dominance(a, b, m) ? 1 : 0; });
}

// fast_non_dominated_sorting: handle-based because output sizes are data
// dependent.  orc_fnds_run -> orc_fnds_sizes -> orc_fnds_copy -> orc_fnds_free.
void *orc_fnds_run(const double *objs, uint64_t n, uint64_t m)
{
    try {
        return new oracle::fnds_result(oracle::fast_non_dominated_sorting(objs, n, m));
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_fnds_sizes(void *h, uint64_t *n_fronts, uint64_t *n_dom_items)
{
    auto r = static_cast<oracle::fnds_result *>(h);
    *n_fronts = r->fronts.size();
    uint64_t t = 0;
    for (const auto &d : r->dom_list) t += d.size();
    *n_dom_items = t;
}

// front_offsets: n_fronts+1, front_items: n; dom_offsets: n+1, dom_items: total.
void orc_fnds_copy(void *h, uint64_t *dom_count, uint64_t *rank, uint64_t *front_offsets, uint64_t *front_items,
                   uint64_t *dom_offsets, uint64_t *dom_items)
{
    auto r = static_cast<oracle::fnds_result *>(h);
    const auto n = r->dom_count.size();
    for (std::size_t i = 0; i < n; ++i) {
        dom_count[i] = r->dom_count[i];
        rank[i] = r->rank[i];
    }
    uint64_t o = 0;
    front_offsets[0] = 0;
    for (std::size_t f = 0; f < r->fronts.size(); ++f) {
        for (auto v : r->fronts[f]) front_items[o++] = v;
        front_offsets[f + 1] = o;
    }
    o = 0;
    dom_offsets[0] = 0;
    for (std::size_t i = 0; i < n; ++i) {
        for (auto v : r->dom_list[i]) dom_items[o++] = v;
        dom_offsets[i + 1] = o;
    }
}

void orc_fnds_free(void *h)
{
    delete static_cast<oracle::fnds_result *>(h);
}

// crowding_distance / select_best_N_mo / sort_population_mo; stable != 0 breaks
// ties by input position (the device's rule), else std::sort as the reference.
int orc_crowding_distance(const double *pts, uint64_t n, uint64_t m, int stable, double *out)
{
    return guarded([&] {
        const auto r = oracle::crowding_distance(pts, n, m, stable != 0);
        std::copy(r.begin(), r.end(), out);
    });
}

int orc_select_best_N_mo(const double *pts, uint64_t n, uint64_t m, uint64_t N, int stable, uint64_t *out,
                         uint64_t *out_len)
{
    return guarded([&] {
        const auto r = oracle::select_best_N_mo(pts, n, m, N, stable != 0);
        std::copy(r.begin(), r.end(), out);
        *out_len = r.size();
    });
}

int orc_sort_population_mo(const double *pts, uint64_t n, uint64_t m, int stable, uint64_t *out)
{
    return guarded([&] {
        const auto r = oracle::sort_population_mo(pts, n, m, stable != 0);
        std::copy(r.begin(), r.end(), out);
    });
}

// Timed sort_population_mo (single-threaded, as in the reference).
int orc_sort_population_mo_timed(const double *pts, uint64_t n, uint64_t m, uint32_t reps, double *sec)
{
    return guarded([&] {
        std::vector<double> t;
        for (uint32_t k = 0; k < reps; ++k) {
            const auto t0 = std::chrono::steady_clock::now();
            auto r = oracle::sort_population_mo(pts, n, m, false);
            const auto t1 = std::chrono::steady_clock::now();
            t.push_back(std::chrono::duration<double>(t1 - t0).count());
        }
        std::sort(t.begin(), t.end());
        *sec = t[t.size() / 2];
    });
}

// Timed fnds (single-threaded, as in the reference): median seconds per call.
int orc_fnds_timed(const double *objs, uint64_t n, uint64_t m, uint32_t reps, double *sec)
{
    return guarded([&] {
        std::vector<double> t;
        for (uint32_t k = 0; k < reps; ++k) {
            const auto t0 = std::chrono::steady_clock::now();
            auto r = oracle::fast_non_dominated_sorting(objs, n, m);
            const auto t1 = std::chrono::steady_clock::now();
            t.push_back(std::chrono::duration<double>(t1 - t0).count());
        }
        std::sort(t.begin(), t.end());
        *sec = t[t.size() / 2];
    });
}

// population(problem, pop_size, seed): X (pop x nx) and F (pop x nf).
int orc_population_init(int kind, uint64_t dim, const double *centers, uint64_t nobj, uint64_t pop_size,
                        uint32_t seed, double *X, double *F)
{
    return guarded([&] {
        auto p = make_udp(kind, dim, centers, nobj);
        std::vector<double> x, f;
        oracle::population_init(p, pop_size, seed, x, f);
        std::memcpy(X, x.data(), x.size() * sizeof(double));
        std::memcpy(F, f.data(), f.size() * sizeof(double));
    });
}

// batch_random_decision_vector(prob, n, mt19937(seed)).
int orc_batch_random_dv(int kind, uint64_t dim, uint64_t n, uint32_t seed, double *out)
{
    return guarded([&] {
        auto p = make_udp(kind, dim, nullptr, 1);
        std::mt19937 e(seed);
        const auto r = oracle::batch_random_decision_vector(p, n, e);
        std::memcpy(out, r.data(), r.size() * sizeof(double));
    });
}

// Uniform draws over explicit bounds in batch_random_decision_vector order.
int orc_batch_uniform(const double *lb, const double *ub, uint64_t nx, uint64_t n, uint32_t seed, double *out)
{
    return guarded([&] {
        std::mt19937 e(seed);
        std::uniform_real_distribution<double> rdist;
        for (uint64_t i = 0; i < n; ++i) {
            for (uint64_t j = 0; j < nx; ++j) {
                out[i * nx + j] = (lb[j] == ub[j]) ? lb[j]
                                                   : rdist(e, std::uniform_real_distribution<double>::param_type(
                                                                  lb[j], ub[j]));
            }
        }
    });
}

// pso_gen(gen, omega, eta1, eta2, max_vel, variant, neighb_type, neighb_param,
// memory, seed).evolve(pop) with X/F in and out.  use_bfe selects the bfe
// branch (thread_bfe on n_threads) versus per-individual fitness calls.
// trace_X (calls x gen x pop x nx) and trace_F (calls x gen x pop) may be
// null.  `calls` successive evolve calls share the algorithm (and its engine).
int orc_pso_gen_evolve(int kind, uint64_t dim, uint64_t pop, double *X, double *F, uint32_t gen, double omega,
                       double eta1, double eta2, double max_vel, uint32_t variant, uint32_t neighb_type,
                       uint32_t neighb_param, int memory, uint32_t seed, int use_bfe, uint32_t n_threads,
                       uint32_t calls, double *trace_X, double *trace_F)
{
    return guarded([&] {
        auto p = make_udp(kind, dim, nullptr, 1);
        const auto nx = p.get_nx();
        oracle::pso_params prm;
        prm.gen = gen;
        prm.omega = omega;
        prm.eta1 = eta1;
        prm.eta2 = eta2;
        prm.max_vel = max_vel;
        prm.variant = variant;
        prm.neighb_type = neighb_type;
        prm.neighb_param = neighb_param;
        prm.memory = memory != 0;
        prm.seed = seed;
        oracle::pso_gen algo(prm);
        std::vector<double> x(X, X + pop * nx), f(F, F + pop);
        oracle::batch_eval_fn bfe = [n_threads](const oracle::udp &q, const std::vector<double> &d) {
            return oracle::bfe_call(q, d, n_threads);
        };
        std::vector<double> tx, tf;
        for (uint32_t c = 0; c < calls; ++c) {
            algo.evolve(p, x, f, use_bfe ? &bfe : nullptr, trace_X ? &tx : nullptr, trace_F ? &tf : nullptr);
        }
        std::memcpy(X, x.data(), x.size() * sizeof(double));
        std::memcpy(F, f.data(), f.size() * sizeof(double));
        if (trace_X) std::memcpy(trace_X, tx.data(), tx.size() * sizeof(double));
        if (trace_F) std::memcpy(trace_F, tf.data(), tf.size() * sizeof(double));
    });
}

// Timed pso_gen evolve (median seconds over reps, each from the same input).
int orc_pso_gen_timed(int kind, uint64_t dim, uint64_t pop, const double *X, const double *F, uint32_t gen,
                      uint32_t seed, uint32_t n_threads, uint32_t reps, double *sec)
{
    return guarded([&] {
        auto p = make_udp(kind, dim, nullptr, 1);
        const auto nx = p.get_nx();
        std::vector<double> t;
        for (uint32_t k = 0; k < reps; ++k) {
            oracle::pso_params prm;
            prm.gen = gen;
            prm.seed = seed;
            oracle::pso_gen algo(prm);
            std::vector<double> x(X, X + pop * nx), f(F, F + pop);
            oracle::batch_eval_fn bfe = [n_threads](const oracle::udp &q, const std::vector<double> &d) {
                return oracle::bfe_call(q, d, n_threads);
            };
            const auto t0 = std::chrono::steady_clock::now();
            algo.evolve(p, x, f, &bfe);
            const auto t1 = std::chrono::steady_clock::now();
            t.push_back(std::chrono::duration<double>(t1 - t0).count());
        }
        std::sort(t.begin(), t.end());
        *sec = t[t.size() / 2];
    });
}

// select_best (src/s_policies/select_best.cpp:60-116); the outputs hold up to n
// individuals.
int orc_select_best(uint64_t n, uint64_t nx, const uint64_t *ids, const double *X, const double *F, int frac,
                    uint64_t count, double fraction, int stable, uint64_t *n_out, uint64_t *ids_out, double *X_out,
                    double *F_out)
{
    return guarded([&] {
        const auto g = oracle::select_best(make_group(n, nx, ids, X, F), make_rate(frac, count, fraction), stable != 0);
        put_group(g, nx, n_out, ids_out, X_out, F_out);
    });
}

// fair_replace (src/r_policies/fair_replace.cpp:62-129); the outputs hold n
// individuals.
int orc_fair_replace(uint64_t n, uint64_t nx, const uint64_t *ids, const double *X, const double *F, uint64_t m,
                     const uint64_t *mids, const double *mX, const double *mF, int frac, uint64_t count,
                     double fraction, int stable, uint64_t *n_out, uint64_t *ids_out, double *X_out, double *F_out)
{
    return guarded([&] {
        const auto g = oracle::fair_replace(make_group(n, nx, ids, X, F), make_rate(frac, count, fraction),
                                            make_group(m, nx, mids, mX, mF), stable != 0);
        put_group(g, nx, n_out, ids_out, X_out, F_out);
    });
}

// Inbound connections of ring(n, w) (which = 0) or fully_connected(n, w)
// (which = 1) as CSR: offsets (n + 1), sources and weights (n (n - 1) each at
// most).
int orc_topology(int which, uint64_t n, double w, uint64_t *offsets, uint64_t *src, double *wout)
{
    return guarded([&] {
        const auto c = which == 0 ? oracle::ring_topology(n, w) : oracle::fully_connected_topology(n, w);
        uint64_t o = 0;
        for (uint64_t i = 0; i < n; ++i) {
            offsets[i] = o;
            for (std::size_t j = 0; j < c.src[i].size(); ++j, ++o) {
                src[o] = c.src[i][j];
                wout[o] = c.w[i][j];
            }
        }
        offsets[n] = o;
    });
}

// `rounds` archipelago::evolve(1) rounds with migration (oracle::
// archi_migrate_round) over n_isl pso_gen islands.  Populations are
// concatenated row-major (X0/X_out: pops[k] x nx_k after island k-1's rows,
// F0/F_out: pops[k]); the final migrants database goes to db_count (per
// island) and dbX/dbF at the same offsets as X_out/F_out.
int orc_archi_migrate(uint64_t n_isl, const int *kinds, const uint64_t *dims, const uint64_t *pops, const double *X0,
                      const double *F0, const uint32_t *pso_seeds, const uint32_t *gens, int type, int handling,
                      const uint64_t *conn_off, const uint64_t *conn_src, const double *conn_w, const int *sel_frac,
                      const uint64_t *sel_count, const double *sel_fraction, const int *rep_frac,
                      const uint64_t *rep_count, const double *rep_fraction, const uint32_t *migr_seeds,
                      uint32_t rounds, int stable, double *X_out, double *F_out, uint64_t *db_count, double *dbX,
                      double *dbF)
{
    return guarded([&] {
        std::vector<oracle::migr_island> isl;
        oracle::connections c;
        uint64_t xo = 0, fo = 0;
        for (uint64_t k = 0; k < n_isl; ++k) {
            auto p = make_udp(kinds[k], dims[k], nullptr, 1);
            const auto nx = p.get_nx();
            oracle::pso_params prm;
            prm.gen = gens[k];
            prm.seed = pso_seeds[k];
            oracle::migr_island is{p, oracle::pso_gen(prm),
                                   make_group(pops[k], nx, nullptr, X0 + xo, F0 + fo),
                                   make_rate(sel_frac[k], sel_count[k], sel_fraction[k]),
                                   make_rate(rep_frac[k], rep_count[k], rep_fraction[k]),
                                   std::mt19937(migr_seeds[k])};
            isl.push_back(std::move(is));
            xo += pops[k] * nx;
            fo += pops[k];
            c.src.emplace_back(conn_src + conn_off[k], conn_src + conn_off[k + 1u]);
            c.w.emplace_back(conn_w + conn_off[k], conn_w + conn_off[k + 1u]);
        }
        std::vector<oracle::ind_group> db(n_isl);
        for (uint32_t r = 0; r < rounds; ++r) {
            oracle::archi_migrate_round(isl, c, type, handling, db, stable != 0);
        }
        xo = fo = 0;
        for (uint64_t k = 0; k < n_isl; ++k) {
            const auto nx = isl[k].prob.get_nx();
            uint64_t n = 0;
            put_group(isl[k].pop, nx, &n, nullptr, X_out + xo, F_out + fo);
            put_group(db[k], nx, &db_count[k], nullptr, dbX + xo, dbF + fo);
            xo += pops[k] * nx;
            fo += pops[k];
        }
    });
}

} // extern "C"
