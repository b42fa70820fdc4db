/* pgh.h -- C ABI of libpagmo_hip.so, the MI355X (gfx950) batch fitness
 * evaluator for pagmo 2.x.
 *
 * Plain pointers and sizes only.  Matrices are row-major and 64-bit sizes
 * count elements.  Every entry point returns a pgh_status; on failure a
 * thread-local message is available from pgh_last_error().  Entry points whose
 * name ends in _host take host buffers and are synchronous; the others take
 * device pointers and enqueue on `stream` (a hipStream_t, NULL = the default
 * stream) without synchronising.
 *
 * Citations are to the pagmo 2.19.1 sources each entry point replaces or
 * backs.
 */
#ifndef PAGMO_HIP_PGH_H
#define PAGMO_HIP_PGH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGH_ABI_VERSION 3

/* Status codes.  The C++ layer (pagmo_hip/hip_bfe.hpp) maps them onto the
 * exception types of the reference: INVALID_ARGUMENT -> std::invalid_argument
 * (src/detail/bfe_impl.cpp:71-77,100-118), OVERFLOW -> std::overflow_error
 * (src/batch_evaluators/thread_bfe.cpp:84-87), anything else ->
 * std::runtime_error.  UNSUPPORTED means "no kernel for this problem": the
 * caller falls back to the CPU evaluator. */
typedef enum pgh_status {
    PGH_OK = 0,
    PGH_INVALID_ARGUMENT = 1,
    PGH_OVERFLOW = 2,
    PGH_UNSUPPORTED = 3,
    PGH_DEVICE_ERROR = 4,
    PGH_NO_DEVICE = 5,
    PGH_OUT_OF_MEMORY = 6
} pgh_status;

/* Problems with a device kernel.  pagmo's UDPs
 * src/problems/{rosenbrock,rastrigin,ackley,schwefel,griewank,lennard_jones,
 * zdt,dtlz}.cpp; MO_SPHERE is the user-defined problem
 * f_k(x) = sum_i (x_i - c_{k,i})^2. */
typedef enum pgh_udp_kind {
    PGH_ROSENBROCK = 1,
    PGH_RASTRIGIN = 2,
    PGH_ACKLEY = 3,
    PGH_SCHWEFEL = 4,
    PGH_GRIEWANK = 5,
    PGH_LENNARD_JONES = 6,
    PGH_MO_SPHERE = 7,
    PGH_ZDT = 8,
    PGH_DTLZ = 9
} pgh_udp_kind;

/* What hip_bfe extracts from a pagmo::problem (problem.hpp:1104-1147).
 *  dim: rosenbrock m_dim (rosenbrock.hpp:91); rastrigin/ackley/schwefel/
 *       griewank get_nx(); lennard_jones atom count (nx = 3*atoms-6,
 *       lennard_jones.cpp:99-111); MO_SPHERE nx; ZDT the constructor's
 *       `param` (zdt.hpp:145: nx = param, or 30 + 5 (param - 1) for ZDT5);
 *       DTLZ the constructor's `dim` (= nx, dtlz.hpp:105).
 *  prob_id: ZDT 1..6, DTLZ 1..7 (zdt.cpp:55-66, dtlz.cpp:53-74); 0 otherwise.
 *  nobj: MO_SPHERE objectives; DTLZ `fdim`; ignored otherwise.
 *  centers: MO_SPHERE only (nobj x dim, row-major; a device pointer for
 *       device entry points, a host pointer for _host ones).
 *  alpha: DTLZ's `alpha` (used by DTLZ4); ignored otherwise.
 *  shift: NULL, or the translation vector t (nx values, host memory) of a
 *       pagmo::translate around the problem: the fitness of x is the problem's
 *       fitness of x - t, element by element (translate.cpp:101-105,
 *       414-423).  Only the host-buffer entry points of fitness and gradient
 *       accept it (they subtract t from their staged copy of dvs); the device
 *       entry points, which may not modify their input, return UNSUPPORTED. */
typedef struct pgh_udp {
    int32_t kind;
    uint32_t prob_id;
    uint64_t dim;
    uint64_t nobj;
    const double *centers;
    uint64_t alpha;
    const double *shift;
} pgh_udp;

/* ---- library ------------------------------------------------------------ */

/* Thread-local description of the last failure on this thread. */
const char *pgh_last_error(void);
int pgh_abi_version(void);
/* 1 if a gfx950 device is usable by this process, 0 otherwise (never fails). */
int pgh_device_available(void);

/* nx / nf of a problem description (problem::get_nx/get_nf,
 * problem.hpp:1360-1400).  INVALID_ARGUMENT for an impossible description
 * (rosenbrock dim < 2: rosenbrock.cpp:43-49; lennard_jones atoms < 3:
 * lennard_jones.cpp:51-62; others dim < 1). */
pgh_status pgh_problem_dims(const pgh_udp *udp, uint64_t *nx, uint64_t *nf);

/* 1 if a kernel exists for this problem, 0 otherwise (no side effects). */
int pgh_supported(const pgh_udp *udp);

/* ---- batch fitness: bfe::operator() (include/pagmo/bfe.hpp:320,
 *      src/bfe.cpp:91-110), backing the UDBFE call operator of hip_bfe.
 *      dvs: n_dvs x nx in, fvs: n_dvs x nf out.  Counters are not touched
 *      (the pagmo::bfe wrapper increments fevals, bfe.cpp:102-107). ------- */
pgh_status pgh_bfe_eval(const pgh_udp *udp, const double *dvs, uint64_t n_dvs, uint64_t nx, double *fvs,
                        uint64_t nf, void *stream);

/* Host-buffer form with the reference's validation: dvs_len must be a
 * multiple of nx (bfe_impl.cpp:63-88, same message); fvs_len must equal
 * n_dvs*nf.  Synchronous. */
pgh_status pgh_bfe_eval_host(const pgh_udp *udp, const double *dvs, uint64_t dvs_len, double *fvs,
                             uint64_t fvs_len);

/* ---- batched gradient: problem::gradient (src/problem.cpp:456-467) for a
 *      whole population, dense layout n_dvs x (nf*nx) (problem.cpp:488-507).
 *      rosenbrock.cpp:83-92, rastrigin.cpp:98-107, and an adjoint kernel for
 *      lennard_jones (which has no gradient in the reference).  fvs may be
 *      NULL; when given it receives the fitness computed in the same pass. */
pgh_status pgh_batch_gradient(const pgh_udp *udp, const double *dvs, uint64_t n_dvs, uint64_t nx, double *grad,
                              double *fvs, void *stream);
pgh_status pgh_batch_gradient_host(const pgh_udp *udp, const double *dvs, uint64_t dvs_len, double *grad,
                                   uint64_t grad_len, double *fvs);

/* ---- memetic local step on the batched gradient.  No counterpart in the
 *      reference: it feeds the gradient of pgh_batch_gradient into a local
 *      refinement that a caller applies to a population between two calls of
 *      an algorithm's evolve() (pagmo's population::set_xf,
 *      src/population.cpp:490-519), so no algorithm's selection logic changes.
 *      For each individual, `iters` rounds of
 *          m = max_j |g_j|,  s = alpha / m,  y_j = x_j - s * g_j,
 *          y_j = lb_j if y_j < lb_j,  y_j = ub_j if y_j > ub_j,
 *          accept y (x = y, f = f(y), alpha = 2 alpha) if f(y) < f(x),
 *          otherwise alpha = alpha / 2,
 *      starting from alpha = step0.  A gradient with a non-finite element or
 *      m == 0 gives y = x, which is rejected.  Problems with a device
 *      gradient only (rosenbrock, rastrigin, lennard_jones; nf = 1).
 *      dvs (n_dvs x nx) is updated in place; fvs (n_dvs) receives f(x) at the
 *      end.  lb and ub (nx each) are both NULL (no clamping) or both given.
 *      work holds at least pgh_local_descent_work_len(n_dvs, nx) doubles of
 *      device memory; the device form allocates nothing. */
uint64_t pgh_local_descent_work_len(uint64_t n_dvs, uint64_t nx);
pgh_status pgh_local_descent(const pgh_udp *udp, double *dvs, uint64_t n_dvs, uint64_t nx, double *fvs,
                             const double *lb, const double *ub, uint32_t iters, double step0, double *work,
                             uint64_t work_len, void *stream);
/* Host-buffer form: dvs (dvs_len = n_dvs * nx) in/out, fvs (n_dvs) out;
 * lb/ub host arrays of nx (or both NULL).  Synchronous. */
pgh_status pgh_local_descent_host(const pgh_udp *udp, double *dvs, uint64_t dvs_len, double *fvs, uint64_t fvs_len,
                                  const double *lb, const double *ub, uint32_t iters, double step0);
/* No counterpart in the reference.  pgh_local_descent for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same arguments and results without the
 * work buffer.  The work lives in a library-owned device array of
 * PGH_LOCAL_DESCENT_NOALLOC_MAX_WORK doubles (64 MiB) per GPU, made when the
 * library's code is loaded on that GPU and held until the process ends; a
 * batch needing more (pgh_local_descent_work_len(n_dvs, nx) beyond it) gets
 * INVALID_ARGUMENT.  Allocates nothing.  `stream` must belong to the current
 * device.  Every call on a GPU uses that GPU's one array: calls whose work
 * overlaps in time on one GPU (several streams or host threads) race and give
 * wrong results; serialise them. */
#define PGH_LOCAL_DESCENT_NOALLOC_MAX_WORK (UINT64_C(1) << 23)
pgh_status pgh_local_descent_noalloc(const pgh_udp *udp, double *dvs, uint64_t n_dvs, uint64_t nx, double *fvs,
                                     const double *lb, const double *ub, uint32_t iters, double step0,
                                     void *stream);

/* ---- fast_non_dominated_sorting (src/utils/multi_objective.cpp:200-257,
 *      return type include/pagmo/utils/multi_objective.hpp:68-72) over a
 *      device-resident n x m objective matrix.  Outputs are CSR:
 *        fronts:   front_offsets[n_fronts+1], front_items[n]
 *        dom_list: dom_offsets[n+1], dom_items[n_dom_items] (ascending rows)
 *        dom_count[n], rank[n]
 *      all uint64 (pop_size_t).  A workspace owns the device results:
 *      run (async) -> sizes (synchronises) -> copy (to host or device). ---- */
typedef struct pgh_fnds_ws pgh_fnds_ws;
pgh_status pgh_fnds_create(pgh_fnds_ws **ws);
/* The calling host thread's own workspace, made on its first call and kept
 * until the thread ends (never at process exit), so that a caller sorting
 * every generation (nsga2.cpp:183, nspso.cpp:162,377, maco.cpp:176) allocates
 * once per thread.  pgh_fnds_destroy refuses it. */
pgh_status pgh_fnds_thread_workspace(pgh_fnds_ws **ws);
pgh_status pgh_fnds_destroy(pgh_fnds_ws *ws);
/* INVALID_ARGUMENT when n < 2, with the reference's message
 * (multi_objective.cpp:203-207). */
pgh_status pgh_fnds_run(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m, void *stream);
/* Host-buffer form: objs (n x m, host memory) is staged into the workspace and
 * sorted on a stream the workspace owns. */
pgh_status pgh_fnds_run_host(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m);
pgh_status pgh_fnds_sizes(pgh_fnds_ws *ws, uint64_t *n_fronts, uint64_t *n_dom_items);
/* Any output pointer may be NULL.  to_host != 0: destinations are host
 * memory (synchronous); 0: device memory (async on the run's stream). */
pgh_status pgh_fnds_copy(pgh_fnds_ws *ws, uint64_t *dom_count, uint64_t *rank, uint64_t *front_offsets,
                         uint64_t *front_items, uint64_t *dom_offsets, uint64_t *dom_items, int to_host);

/* ---- autotuning of launch mappings.  No counterpart in the reference (its
 *      evaluators have no launch shape).  Only parameters that cannot change
 *      any result are tuned: the workgroup count of the front peel (for
 *      n <= 32768) and the number of Lennard-Jones individuals per workgroup;
 *      lanes per individual, reduction trees and operation order never depend
 *      on them.  A tuner times each candidate `reps` times with HIP events on
 *      `stream` over the caller's own input, keeps the fastest, and finishes
 *      with one more run so that the outputs hold that run's results.
 *      Synchronises `stream`.
 *
 *      Peel workgroups of a workspace: 0 = the default rule (one per 256
 *      points); at most 4096 may be set, and a run clamps the count to the
 *      peel's range: at least one per 1024 points, at most one per 64 points,
 *      the CU count and 256.  pgh_fnds_tune tries one per 64, 128, 256, 512
 *      and 1024 points within that range, and sets the fastest on the
 *      workspace (*chosen = 0 when n > 32768: nothing to tune). */
pgh_status pgh_fnds_set_peel_blocks(pgh_fnds_ws *ws, uint32_t blocks);
pgh_status pgh_fnds_tune(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m, uint32_t reps, void *stream,
                         uint32_t *chosen);
/* Lennard-Jones individuals per workgroup for the energy (grad = 0) or the
 * energy-and-gradient kernel (grad != 0), process-wide, also used by
 * archipelagos created afterwards: 0 = the default (the most whose LDS fits
 * 64 KiB); at most that default.  pgh_lj_tune tries the default and its
 * halvings (up to four values) on dvs; grad == NULL tunes the energy kernel
 * (fvs required), otherwise the gradient kernel (fvs may be NULL). */
pgh_status pgh_lj_set_groups(uint64_t atoms, int grad, uint32_t groups);
pgh_status pgh_lj_tune(const pgh_udp *udp, const double *dvs, uint64_t n_dvs, double *fvs, double *grad,
                       uint32_t reps, void *stream, uint32_t *chosen);

/* ---- crowding distance and the multi-objective orderings built on fnds
 *      (src/utils/multi_objective.cpp:280-465).  The reference sorts with
 *      std::sort, which leaves the order of equal keys open; here equal keys
 *      keep their input order (std::stable_sort over the reference's own index
 *      vectors): equal values of objective i within a front in their order
 *      after objective i-1 (the first: position in the front), equal
 *      (rank, crowding) pairs by index, equal crowding values of the partially
 *      selected front by position.  With distinct keys the results are the
 *      reference's bit for bit.  These calls use the workspace's fnds
 *      buffers (the fnds results are replaced) and synchronise the stream. */
/* crowding_distance (multi_objective.cpp:280-318) of the n points (row-major
 * n x m) of one non-dominated front; out: n doubles.  INVALID_ARGUMENT with the
 * reference's messages when n < 2 or m < 2. */
pgh_status pgh_crowding_distance(pgh_fnds_ws *ws, const double *points, uint64_t n, uint64_t m, double *out,
                                 void *stream);
pgh_status pgh_crowding_distance_host(const double *points, uint64_t n, uint64_t m, double *out);
/* sort_population_mo (multi_objective.cpp:425-465): out receives the n indices
 * ordered by (non-domination rank, crowding distance descending). */
pgh_status pgh_sort_population_mo(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m, uint64_t *out,
                                  void *stream);
pgh_status pgh_sort_population_mo_host(const double *objs, uint64_t n, uint64_t m, uint64_t *out);
/* select_best_N_mo (multi_objective.cpp:344-396): out (room for min(N, n)
 * indices) receives the selection, *out_len its length. */
pgh_status pgh_select_best_N_mo(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m, uint64_t N,
                                uint64_t *out, uint64_t *out_len, void *stream);
pgh_status pgh_select_best_N_mo_host(const double *objs, uint64_t n, uint64_t m, uint64_t N, uint64_t *out,
                                     uint64_t *out_len);

/* ---- seeded inputs: std::mt19937 (include/pagmo/rng.hpp:43) and
 *      batch_random_decision_vector (include/pagmo/utils/generic.hpp:326-388)
 *      generated on the device, bit-identical to libstdc++. ------------- */
pgh_status pgh_mt19937_generate(uint32_t seed, uint64_t skip, uint64_t n, uint32_t *out, void *stream);
pgh_status pgh_batch_random_dv(const pgh_udp *udp, uint64_t n, uint32_t seed, double *out, void *stream);

/* ---- pso_gen (src/algorithms/pso_gen.cpp:120-530) on the device.
 *      Parameters and defaults of pso_gen.hpp:109-111.  Supported on the
 *      device: variants 1-6 (pso_gen.cpp:241-327) with the gbest (neighb_type
 *      1), lbest (2) and von Neumann (3) topologies, FIPS (variant 6) with
 *      lbest and von Neumann only; the adaptive random topology (4), whose
 *      draws depend on each generation's results, and FIPS with gbest return
 *      UNSUPPORTED. ------------------------------------------------------- */
typedef struct pgh_pso_params {
    uint32_t gen;
    uint32_t variant;
    double omega;
    double eta1;
    double eta2;
    double max_vel;
    uint32_t neighb_type;
    uint32_t neighb_param;
    uint32_t memory;
    uint32_t seed;
} pgh_pso_params;

/* ---- archipelago of independent pso_gen islands (archipelago::evolve,
 *      src/archipelago.cpp:325-330, one island per task_queue thread in the
 *      reference, src/detail/task_queue.cpp:40-107).  Islands may differ in
 *      problem, dimension and population size.  X0/F0: per-island host
 *      arrays (pop x nx, pop x 1) of the initial populations. ------------- */
typedef struct pgh_archi pgh_archi;
typedef struct pgh_island_spec {
    pgh_udp udp;
    uint64_t pop;
    pgh_pso_params algo; /* per-island seed in algo.seed */
} pgh_island_spec;

enum pgh_archi_mode {
    PGH_ARCHI_BATCHED = 0,   /* a generation of all islands is one launch (or one per problem class), graph-captured */
    PGH_ARCHI_PER_ISLAND = 1 /* each island launches its own kernels, one island after another */
};

pgh_status pgh_archi_create(pgh_archi **a, const pgh_island_spec *islands, uint64_t n_islands,
                            const double *const *X0, const double *const *F0);
/* The same with every island's initial population made on the device, as
 * population(problem, bfe, pop, pop_seeds[k]) does (population.cpp:82-103):
 * batch_random_decision_vector from std::mt19937(pop_seeds[k])
 * (generic.hpp:326-388), bit for bit, then one batch evaluation by the device
 * kernels.  No host population arrays are involved.  Synchronous. */
pgh_status pgh_archi_create_seeded(pgh_archi **a, const pgh_island_spec *islands, uint64_t n_islands,
                                   const uint32_t *pop_seeds);
pgh_status pgh_archi_destroy(pgh_archi *a);
/* One pso_gen::evolve call of algo.gen generations on every island
 * (velocities re-drawn first unless memory is set, pso_gen.cpp:193-200). */
pgh_status pgh_archi_evolve(pgh_archi *a, int mode, void *stream);
/* Population of island k after the last evolve (pop.set_xf(i, lbX, lbfit),
 * pso_gen.cpp:526-528), to host memory.  Synchronous. */
pgh_status pgh_archi_get(pgh_archi *a, uint64_t k, double *X, double *F);
/* Every generation's evaluated fitness of island k (gen x pop), recorded
 * when the archipelago was created with tracing on; see pgh_archi_set_trace. */
pgh_status pgh_archi_set_trace(pgh_archi *a, int on);
pgh_status pgh_archi_get_trace(pgh_archi *a, uint64_t k, double *trace_F);

/* Launch mapping of a generation (no counterpart in the reference; no
 * result depends on it): a generation moves, evaluates and updates the memory
 * of every island's individuals in items of rows, handed out by a queue to a
 * persistent grid.
 *   form:          PGH_ARCHI_STEP_ONE (the default), one launch per
 *                  generation over every item; PGH_ARCHI_STEP_SPLIT, one
 *                  launch per problem class (rosenbrock, separable,
 *                  lennard_jones), each with its own LDS and register
 *                  footprint (batched mode only)
 *   lj_lds:        LDS bytes for a Lennard-Jones item's staged individuals
 *                  (as many individuals per item as fit, at least one; 0 =
 *                  64 KiB, at most 64 KiB)
 *   blocks_per_cu: persistent workgroups per CU and launch (0 = as many as
 *                  can be resident)
 * pgh_archi_tune times candidate mappings (form x lj_lds in {64, 32, 16}
 * KiB x blocks_per_cu in {0, 2}) with one batched evolve call each, `reps`
 * times, from the archipelago's current state, which it restores afterwards;
 * every candidate must leave a population bit-identical to the first's
 * (else DEVICE_ERROR).  It keeps the fastest and returns it in *chosen.
 * Synchronous; refused while migration is on. */
enum pgh_archi_step_form {
    PGH_ARCHI_STEP_ONE = 0,
    PGH_ARCHI_STEP_SPLIT = 1
};
typedef struct pgh_archi_mapping {
    uint32_t form;
    uint32_t lj_lds;
    uint32_t blocks_per_cu;
    uint32_t pad;
} pgh_archi_mapping;
pgh_status pgh_archi_set_mapping(pgh_archi *a, const pgh_archi_mapping *m);
pgh_status pgh_archi_get_mapping(pgh_archi *a, pgh_archi_mapping *m);
pgh_status pgh_archi_tune(pgh_archi *a, uint32_t reps, pgh_archi_mapping *chosen);

/* ---- migration between the islands (island::evolve, src/island.cpp:452-642;
 *      the migrants database, src/archipelago.cpp:658-713).  Once set, every
 *      pgh_archi_evolve call is one archipelago::evolve(1) round: each island
 *      takes migrants in (island index order), then every island evolves,
 *      then each island's select_best emigrants replace its database entry.
 *      The migrants stay on the device.  Single-objective, unconstrained
 *      problems with the select_best and fair_replace policies. ------------ */
enum pgh_migration_type {
    PGH_MIGRATION_P2P = 0,      /* migration_type::p2p (archipelago.hpp) */
    PGH_MIGRATION_BROADCAST = 1 /* migration_type::broadcast */
};
enum pgh_migrant_handling {
    PGH_MIGRANTS_PRESERVE = 0, /* migrant_handling::preserve: sources keep their migrants */
    PGH_MIGRANTS_EVICT = 1     /* migrant_handling::evict: taking migrants empties the source's entry */
};
/* A policy's migration rate (detail/base_sr_policy.hpp): `count`
 * individuals, or with `fractional` set, `fraction` of the input individuals
 * (truncated). */
typedef struct pgh_migr_rate {
    int32_t fractional;
    uint32_t pad;
    uint64_t count;
    double fraction;
} pgh_migr_rate;
/* Turns migration on for the archipelago (conn_offsets null turns it off).
 * Island k's inbound connections are conn_src / conn_weight over
 * [conn_offsets[k], conn_offsets[k+1]), in topology::get_connections order
 * (topology.hpp:173, e.g. ring.cpp, fully_connected.cpp); connected islands
 * must have the same nx.  select_rate / replace_rate: each island's
 * select_best and fair_replace rates (select_best.cpp:60-116,
 * fair_replace.cpp:62-129); migr_seeds: the seed of each island's migration
 * engine (std::mt19937; the reference seeds it from random_device, once per
 * island::evolve call, island.cpp:466-468).  The engines persist across
 * pgh_archi_evolve calls until the next pgh_archi_set_migration, so that a
 * set followed by n evolve calls draws as one island::evolve(n) does.
 * The database starts empty (archipelago.cpp:511).  Messages of the
 * reference's checks: edge weights (topology.cpp:58-68), fractional rates
 * (base_sr_policy.cpp:54-66) and absolute rates above the population size
 * (select_best.cpp:90-96, fair_replace.cpp:93-99, raised here at set-up).
 * Islands with pso_gen memory set are UNSUPPORTED. */
pgh_status pgh_archi_set_migration(pgh_archi *a, int type, int handling, const uint64_t *conn_offsets,
                                   const uint64_t *conn_src, const double *conn_weight,
                                   const pgh_migr_rate *select_rate, const pgh_migr_rate *replace_rate,
                                   const uint32_t *migr_seeds);
/* Island k's entry of the migrants database (archipelago::get_migrants,
 * archipelago.cpp:682-694): *count individuals, best first, to host X
 * (count x nx) and F (count) when non-null.  Synchronous. */
pgh_status pgh_archi_get_migrants(pgh_archi *a, uint64_t k, uint64_t *count, double *X, double *F);

/* Single-island convenience: pso_gen(prm).evolve(pop) with host X/F in/out
 * (pso_gen.cpp:120-530).  Results equal an archipelago of that one island
 * evolved once; an island that is one evaluation tile runs a batch of 32
 * generations per launch in one workgroup.  Synchronous. */
pgh_status pgh_pso_gen_evolve_host(const pgh_udp *udp, const pgh_pso_params *prm, double *X, double *F,
                                   uint64_t pop, uint64_t nx);

#ifdef __cplusplus
}
#endif

#endif /* PAGMO_HIP_PGH_H */
