// pgh_internal.hpp -- shared declarations of libpagmo_hip.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "pagmo_hip/pgh.h"

namespace pgh
{

// ---- error state (capi.hip) ------------------------------------------------
void set_error(const std::string &msg);
pgh_status fail(pgh_status st, const std::string &msg);
pgh_status hip_fail(hipError_t e, const char *what);

// PGH_OK, or the status and message of a failed HIP call.
inline pgh_status hip_status(hipError_t e, const char *what)
{
    return e == hipSuccess ? PGH_OK : hip_fail(e, what);
}

#define PGH_HIP(call)                                                                                                \
    do {                                                                                                             \
        hipError_t pgh_e_ = (call);                                                                                  \
        if (pgh_e_ != hipSuccess) return ::pgh::hip_fail(pgh_e_, #call);                                             \
    } while (0)

// Propagates a non-OK status.
#define PGH_TRY(expr)                                                                                                \
    do {                                                                                                             \
        const pgh_status pgh_st_ = (expr);                                                                           \
        if (pgh_st_ != PGH_OK) return pgh_st_;                                                                       \
    } while (0)

// ---- problem description after validation ---------------------------------
struct prob {
    int kind = 0;
    uint32_t dim = 0;   // rosenbrock m_dim, lj atoms, otherwise nx
    uint32_t nx = 0;
    uint32_t nf = 0;
    const double *centers = nullptr; // MO_SPHERE, device pointer
    uint32_t prob_id = 0;            // ZDT / DTLZ problem number
    uint32_t alpha = 0;              // DTLZ alpha
    const double *shift = nullptr;   // translate's vector, host pointer
};

// Validates a pgh_udp; fills `out`.  Does not touch centers' memory.
pgh_status resolve(const pgh_udp *udp, prob &out);

// ---- launchers (fitness.hip) -----------------------------------------------
// x[e] -= t[e % nx] for the count elements of a row-major matrix (translate).
pgh_status launch_shift(double *x, uint64_t count, uint32_t nx, const double *t, hipStream_t s);
pgh_status launch_fitness(const prob &p, const double *dvs, uint64_t n_dvs, double *fvs, hipStream_t s);
pgh_status launch_gradient(const prob &p, const double *dvs, uint64_t n_dvs, double *grad, double *fvs,
                           hipStream_t s);

// Individuals per Lennard-Jones workgroup in use for (atoms, gradient kernel):
// a value set or tuned through the C ABI, else lj_group_count (fitness.hip).
uint32_t lj_groups_for(uint32_t atoms, bool grad);

// Refuses grids beyond what one launch addresses (fitness.hip).
pgh_status check_launch_size(uint64_t n, uint32_t nx);

// ---- argument checks shared by entry points (capi.hip) ------------------------
// The reference's messages for a decision / fitness vector of the wrong size.
pgh_status check_shape(const prob &p, uint64_t nx, uint64_t nf);
// pgh_local_descent's arguments beyond the problem and the buffers.
pgh_status check_local_descent(const prob &p, const double *lb, const double *ub, double step0);

// ---- memetic local step (local.hip) -------------------------------------------
uint64_t local_descent_work_len(uint64_t n, uint64_t nx);
pgh_status launch_local_descent(const prob &p, double *x, uint64_t n, double *f, const double *lb, const double *ub,
                                uint32_t iters, double step0, double *work, hipStream_t s);

// ---- rng (pso.hip) ------------------------------------------------------------
pgh_status launch_mt19937(uint32_t seed, uint64_t skip, uint64_t n, uint32_t *out, hipStream_t s);
pgh_status launch_batch_random_dv(const prob &p, uint64_t n, uint32_t seed, double *out, hipStream_t s);

// ---- std::mt19937 jump-ahead (mt_jump.hip) ------------------------------------
// Device table of the coefficient lists of z^d mod phi for each distance d
// (layout: mt_dev.hpp, kMtIdxStride); made on first use per device and kept.
// Synchronous the first time.
pgh_status jump_table_device(const std::vector<uint64_t> &dists, const void **out);
// The same table as bytes, for a caller that keeps its own device copy.
pgh_status jump_table_bytes(const std::vector<uint64_t> &dists, std::vector<uint8_t> &out);
// Memoises z^(k S) mod phi for k = 1 .. count (host), so that a table of those
// distances costs one multiplication per entry.
pgh_status jump_polys_multiples(uint64_t S, uint64_t count);
// Bytes of the table's header (one 32-bit length per distance, 16-byte padded).
constexpr size_t jump_table_head(size_t n_dists)
{
    return (n_dists * 4u + 15u) / 16u * 16u;
}

// ---- migration between archipelago islands (migration.hip) ------------------
struct migr_island {
    uint32_t nx, pop;
    uint64_t xoff, foff, eoff; // the island's rows, fitness and first element in the archipelago's arrays
};
struct migr_ctx;
pgh_status migr_create(const std::vector<migr_island> &isl, int type, int handling, const uint64_t *conn_off,
                       const uint64_t *conn_src, const double *conn_w, const pgh_migr_rate *sel,
                       const pgh_migr_rate *rep, const uint32_t *seeds, migr_ctx **out);
void migr_destroy(migr_ctx *m);
// Immigration of one round from the populations in (sX, sF) into (dX, dF);
// *moved is false (and nothing is launched) when no island replaces.
pgh_status migr_immigrate(migr_ctx *m, const double *sX, const double *sF, double *dX, double *dF, bool *moved,
                          hipStream_t s);
// Emigration of one round from the populations in (X, F).
pgh_status migr_emigrate(migr_ctx *m, const double *X, const double *F, hipStream_t s);
// Island k's database entry to host memory (the caller synchronises first).
pgh_status migr_get(migr_ctx *m, uint64_t k, uint64_t *count, double *X, double *F);

// ---- device scratch owned by the calling host thread ---------------------
struct dev_buf {
    void *ptr = nullptr;
    size_t cap = 0;
    pgh_status reserve(size_t bytes);
    ~dev_buf();
};

// Grow-only pinned host memory (hipHostMalloc) with two events, one per half.
struct pinned_buf {
    void *ptr = nullptr;
    size_t cap = 0;
    hipEvent_t done[2] = {nullptr, nullptr};
    pgh_status reserve(size_t bytes);
    ~pinned_buf();
};

// Per-host-thread stream, staging buffers and fnds workspace for the _host
// entry points.  Made on a thread's first call and released when that thread
// ends (a pthread key destructor); the process's main thread keeps its
// context until the process ends, so exit() makes no HIP call.
struct thread_ctx {
    hipStream_t stream = nullptr;
    dev_buf a, b, c, d, e;
    pinned_buf pin;
    pgh_fnds_ws *fnds = nullptr; // pgh_fnds_thread_workspace, made on first use
    ~thread_ctx();
};

// Host-to-device copy of a host buffer on the context's stream.  Large
// buffers go through two pinned chunks filled by host threads, each sent by
// DMA while the other fills; small ones are one hipMemcpyAsync.
pgh_status upload_host(thread_ctx *ctx, void *dst, const void *src, size_t bytes);
pgh_status get_thread_ctx(thread_ctx **ctx);

// The calling thread's fnds workspace (fnds.hip): owned by its thread_ctx.
pgh_status thread_fnds_ws(pgh_fnds_ws **ws);
// Releases a workspace whatever its owner (fnds.hip).
void fnds_free(pgh_fnds_ws *ws);

// Status and message of an exception caught at the C ABI boundary: every
// extern "C" entry point is a function-try-block ending in
//     catch (...) { return pgh::api_exception(); }
// so that no C++ exception unwinds into a C or ctypes caller.
pgh_status api_exception() noexcept;

} // namespace pgh
