// capi.hip -- the extern "C" entry points of libpagmo_hip.so: validation,
// error reporting, host staging and dispatch to the kernel launchers.
#include <pthread.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <exception>
#include <limits>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "pgh_internal.hpp"

namespace pgh
{

namespace
{
thread_local std::string g_last_error;
}

void set_error(const std::string &msg)
{
    g_last_error = msg;
}

pgh_status fail(pgh_status st, const std::string &msg)
{
    g_last_error = msg;
    return st;
}

pgh_status api_exception() noexcept
{
    try {
        throw;
    } catch (const std::bad_alloc &) {
        g_last_error = "host memory exhausted";
        return PGH_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        try {
            g_last_error = std::string("host error: ") + e.what();
        } catch (...) {
        }
        return PGH_DEVICE_ERROR;
    } catch (...) {
        return PGH_DEVICE_ERROR;
    }
}

pgh_status hip_fail(hipError_t e, const char *what)
{
    const pgh_status st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PGH_OUT_OF_MEMORY
                          : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? PGH_NO_DEVICE
                                                                                   : PGH_DEVICE_ERROR;
    g_last_error = std::string("HIP error '") + hipGetErrorString(e) + "' in " + what;
    return st;
}

pgh_status dev_buf::reserve(size_t bytes)
{
    if (bytes <= cap) return PGH_OK;
    if (ptr) {
        (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    // round up to 2 MiB so that slowly growing batches do not reallocate
    const size_t want = (bytes + (size_t(2) << 20) - 1u) & ~((size_t(2) << 20) - 1u);
    PGH_HIP(hipMalloc(&ptr, want));
    cap = want;
    return PGH_OK;
}

dev_buf::~dev_buf()
{
    // Runs when the owner is released explicitly (an archipelago, a
    // workspace) or when a thread_ctx's thread ends, never from exit().
    if (ptr) (void)hipFree(ptr);
}

pgh_status pinned_buf::reserve(size_t bytes)
{
    if (!done[0]) {
        PGH_HIP(hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
        PGH_HIP(hipEventCreateWithFlags(&done[1], hipEventDisableTiming));
    }
    if (bytes <= cap) return PGH_OK;
    if (ptr) {
        (void)hipHostFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    PGH_HIP(hipHostMalloc(&ptr, bytes, hipHostMallocDefault));
    cap = bytes;
    return PGH_OK;
}

pinned_buf::~pinned_buf()
{
    // as for dev_buf
    if (ptr) (void)hipHostFree(ptr);
    if (done[0]) (void)hipEventDestroy(done[0]);
    if (done[1]) (void)hipEventDestroy(done[1]);
}

namespace
{
// chunk of a pinned upload and the size from which uploads use them
constexpr size_t kPinChunk = size_t(16) << 20, kPinMin = size_t(16) << 20;

// memcpy split over up to `threads` host threads
void parallel_copy(void *dst, const void *src, size_t bytes, unsigned threads)
{
    if (threads <= 1u || bytes < (size_t(1) << 20)) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = (bytes / threads + 63u) & ~size_t(63);
    const auto copy_part = [=](size_t off) {
        if (off < bytes) {
            std::memcpy(static_cast<char *>(dst) + off, static_cast<const char *>(src) + off,
                        std::min(part, bytes - off));
        }
    };
    std::vector<std::thread> pool;
    unsigned t = 1;
    try {
        pool.reserve(threads);
        for (; t < threads && t * part < bytes; ++t) pool.emplace_back(copy_part, t * part);
    } catch (...) {
        // no more threads (or no memory for them): this thread copies the
        // parts that have none
    }
    for (unsigned u = t; u < threads; ++u) copy_part(u * part);
    copy_part(0);
    for (auto &th : pool) th.join();
}
} // namespace

pgh_status upload_host(thread_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (bytes < kPinMin) {
        PGH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return PGH_OK;
    }
    PGH_TRY(ctx->pin.reserve(2u * kPinChunk));
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned threads = std::max(1u, std::min(8u, hw / 2u));
    auto *stage = static_cast<char *>(ctx->pin.ptr);
    size_t i = 0;
    for (size_t off = 0; off < bytes; off += kPinChunk, ++i) {
        const size_t len = std::min(kPinChunk, bytes - off);
        char *half = stage + (i % 2u) * kPinChunk;
        // the DMA that last read this half (in this call or an earlier one)
        // has finished; an event never recorded is complete
        PGH_HIP(hipEventSynchronize(ctx->pin.done[i % 2u]));
        parallel_copy(half, static_cast<const char *>(src) + off, len, threads);
        PGH_HIP(hipMemcpyAsync(static_cast<char *>(dst) + off, half, len, hipMemcpyHostToDevice, ctx->stream));
        PGH_HIP(hipEventRecord(ctx->pin.done[i % 2u], ctx->stream));
    }
    return PGH_OK;
}

thread_ctx::~thread_ctx()
{
    // the owning thread has ended (pthread key destructor): the runtime is
    // alive, and nothing of this thread is in flight after the synchronise
    if (fnds) fnds_free(fnds);
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
}

namespace
{
// The contexts hang off a pthread key rather than a C++ thread_local: key
// destructors run when a thread ends, but not for the thread that calls
// exit(), whose context the process end reclaims.  A thread_local's
// destructor would run inside exit()'s handlers, where HIP calls are not safe
// (the runtime and a profiler's tool library tear down there too).
pthread_key_t g_ctx_key;
int g_ctx_key_status = -1;
pthread_once_t g_ctx_once = PTHREAD_ONCE_INIT;

void release_ctx(void *p)
{
    delete static_cast<thread_ctx *>(p);
}

void make_ctx_key()
{
    g_ctx_key_status = pthread_key_create(&g_ctx_key, release_ctx);
}
} // namespace

pgh_status get_thread_ctx(thread_ctx **out)
{
    (void)pthread_once(&g_ctx_once, make_ctx_key);
    if (g_ctx_key_status != 0) return fail(PGH_DEVICE_ERROR, "pthread_key_create failed");
    auto *ctx = static_cast<thread_ctx *>(pthread_getspecific(g_ctx_key));
    if (!ctx) {
        ctx = new thread_ctx();
        if (pthread_setspecific(g_ctx_key, ctx) != 0) {
            delete ctx;
            return fail(PGH_OUT_OF_MEMORY, "pthread_setspecific failed");
        }
    }
    if (!ctx->stream) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            return fail(PGH_NO_DEVICE, "no HIP device is available");
        }
        PGH_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    }
    *out = ctx;
    return PGH_OK;
}

pgh_status resolve(const pgh_udp *udp, prob &out)
{
    if (!udp) return fail(PGH_INVALID_ARGUMENT, "null problem description");
    out = prob{};
    out.kind = udp->kind;
    const uint64_t dim = udp->dim;
    const uint64_t lim = std::numeric_limits<uint32_t>::max() / 4u;
    switch (udp->kind) {
        case PGH_ROSENBROCK:
            // rosenbrock.cpp:43-49
            if (dim < 2u) {
                return fail(PGH_INVALID_ARGUMENT, "Rosenbrock Function must have minimum 2 dimensions, "
                                                      + std::to_string(dim) + " requested");
            }
            break;
        case PGH_RASTRIGIN:
        case PGH_ACKLEY:
        case PGH_SCHWEFEL:
        case PGH_GRIEWANK:
            if (dim < 1u) {
                static const char *names[] = {"", "", "Rastrigin", "Ackley", "Schwefel", "Griewank"};
                return fail(PGH_INVALID_ARGUMENT, std::string(names[udp->kind]) + " Function must have minimum 1 dimension, "
                                                      + std::to_string(dim) + " requested");
            }
            break;
        case PGH_LENNARD_JONES:
            // lennard_jones.cpp:51-62
            if (dim < 3u) {
                return fail(PGH_INVALID_ARGUMENT, "The number of atoms in a Lennard Jones Clusters problem must be "
                                                  "positive and greater than 2, while a number of "
                                                      + std::to_string(dim) + " was detected.");
            }
            break;
        case PGH_MO_SPHERE:
            if (dim < 1u || udp->nobj < 1u) {
                return fail(PGH_INVALID_ARGUMENT, "MO_SPHERE needs dim >= 1 and nobj >= 1");
            }
            if (!udp->centers) return fail(PGH_INVALID_ARGUMENT, "MO_SPHERE needs centers");
            if (udp->nobj > 1024u) return fail(PGH_UNSUPPORTED, "MO_SPHERE supports at most 1024 objectives");
            break;
        case PGH_ZDT:
            // zdt.cpp:55-66, in its order
            if (dim < 2u) {
                return fail(PGH_INVALID_ARGUMENT, "ZDT test problems must have a minimum value of 2 for the "
                                                  "constructing parameter (representing the dimension except for "
                                                  "ZDT5), "
                                                      + std::to_string(dim) + " requested");
            }
            if (udp->prob_id == 0u || udp->prob_id > 6u) {
                return fail(PGH_INVALID_ARGUMENT, "ZDT test suite contains six (prob_id=[1 ... 6]) problems, prob_id="
                                                      + std::to_string(udp->prob_id) + " was detected");
            }
            if (udp->prob_id == 5u && dim > lim / 5u) {
                return fail(PGH_OVERFLOW, "problem dimension too large: " + std::to_string(dim));
            }
            break;
        case PGH_DTLZ:
            // dtlz.cpp:53-74, in its order (the overflow limits are reached
            // only beyond this library's dimension limit, checked below)
            if (udp->prob_id == 0u || udp->prob_id > 7u) {
                return fail(PGH_INVALID_ARGUMENT, "DTLZ test suite contains seven (prob_id = [1 ... 7]) problems, "
                                                  "prob_id="
                                                      + std::to_string(udp->prob_id) + " was detected");
            }
            if (udp->nobj < 2u) {
                return fail(PGH_INVALID_ARGUMENT, "DTLZ test problem have a minimum of 2 objectives: fdim="
                                                      + std::to_string(udp->nobj) + " was detected");
            }
            if (dim <= udp->nobj) {
                return fail(PGH_INVALID_ARGUMENT,
                            "The problem dimension has to be larger than the number of objectives.");
            }
            if (udp->alpha > std::numeric_limits<uint32_t>::max()) {
                return fail(PGH_INVALID_ARGUMENT, "DTLZ alpha out of range: " + std::to_string(udp->alpha));
            }
            if (udp->nobj > 1024u) return fail(PGH_UNSUPPORTED, "DTLZ supports at most 1024 objectives");
            break;
        default:
            return fail(PGH_UNSUPPORTED, "no kernel for problem kind " + std::to_string(udp->kind));
    }
    if (dim > lim) return fail(PGH_OVERFLOW, "problem dimension too large: " + std::to_string(dim));
    out.dim = static_cast<uint32_t>(dim);
    switch (udp->kind) {
        case PGH_LENNARD_JONES:
            out.nx = 3u * out.dim - 6u;
            break;
        case PGH_ZDT:
            // ZDT5: 30 bits, then 5 per further block (zdt.cpp:126-129)
            out.nx = udp->prob_id == 5u ? 30u + 5u * (out.dim - 1u) : out.dim;
            break;
        default:
            out.nx = out.dim;
    }
    out.nf = udp->kind == PGH_MO_SPHERE || udp->kind == PGH_DTLZ ? static_cast<uint32_t>(udp->nobj)
             : udp->kind == PGH_ZDT                              ? 2u
                                                                 : 1u;
    out.centers = udp->centers;
    out.prob_id = udp->kind == PGH_ZDT || udp->kind == PGH_DTLZ ? udp->prob_id : 0u;
    out.alpha = udp->kind == PGH_DTLZ ? static_cast<uint32_t>(udp->alpha) : 0u;
    out.shift = udp->shift;
    return PGH_OK;
}

pgh_status check_shape(const prob &p, uint64_t nx, uint64_t nf)
{
    if (nx != p.nx) {
        return fail(PGH_INVALID_ARGUMENT, "A decision vector is incompatible with the problem: the number of "
                                          "dimensions of the problem is "
                                              + std::to_string(p.nx) + ", while the decision vector has a size of "
                                              + std::to_string(nx) + " (the two values should be equal)");
    }
    if (nf != p.nf) {
        return fail(PGH_INVALID_ARGUMENT, "A fitness vector is incompatible with the problem: the dimension of the "
                                          "fitness of the problem is "
                                              + std::to_string(p.nf) + ", while the fitness vector has a size of "
                                              + std::to_string(nf) + " (the two values should be equal)");
    }
    return PGH_OK;
}

namespace
{

// problem::gradient throws not_implemented_error for UDPs without a gradient
// (problem.hpp:633-645); lennard_jones gets the new adjoint kernel.
pgh_status check_has_gradient(const prob &p)
{
    if (p.kind == PGH_ROSENBROCK || p.kind == PGH_RASTRIGIN || p.kind == PGH_LENNARD_JONES) return PGH_OK;
    return fail(PGH_UNSUPPORTED, "The gradient has not been implemented in the user-defined problem");
}

// src/detail/bfe_impl.cpp:63-88
pgh_status check_input_len(const prob &p, uint64_t len, uint64_t &n_dvs)
{
    if (len % p.nx) {
        return fail(PGH_INVALID_ARGUMENT,
                    "Invalid argument for a batch fitness evaluation: the length of the vector representing the "
                    "decision vectors, "
                        + std::to_string(len) + ", is not an exact multiple of the dimension of the problem, "
                        + std::to_string(p.nx));
    }
    n_dvs = len / p.nx;
    return PGH_OK;
}

// A translation is applied only by the host-buffer fitness and gradient entry
// points, on their staged copy of the decision vectors.
pgh_status reject_shift(const prob &p, const char *what)
{
    if (!p.shift) return PGH_OK;
    return fail(PGH_UNSUPPORTED, std::string(what) + " does not take a translated problem");
}

// Subtracts the translation from the staged decision vectors x (n_elems).
pgh_status apply_shift(const prob &p, double *x, uint64_t n_elems, thread_ctx *ctx)
{
    if (!p.shift) return PGH_OK;
    PGH_TRY(ctx->e.reserve(p.nx * sizeof(double)));
    PGH_HIP(hipMemcpyAsync(ctx->e.ptr, p.shift, p.nx * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    return launch_shift(x, n_elems, p.nx, static_cast<const double *>(ctx->e.ptr), ctx->stream);
}

} // namespace

// pgh_local_descent's arguments beyond the problem and the buffers
pgh_status check_local_descent(const prob &p, const double *lb, const double *ub, double step0)
{
    PGH_TRY(reject_shift(p, "the local descent"));
    PGH_TRY(check_has_gradient(p));
    if ((lb == nullptr) != (ub == nullptr)) {
        return fail(PGH_INVALID_ARGUMENT, "the bounds of a local descent are both given or both absent");
    }
    if (!(step0 > 0.) || !(step0 <= DBL_MAX)) {
        return fail(PGH_INVALID_ARGUMENT, "the initial step of a local descent must be positive and finite, got "
                                              + std::to_string(step0));
    }
    return PGH_OK;
}

namespace
{

// Copies host MO_SPHERE centres to the device buffer `buf` and repoints p.
pgh_status stage_centers(prob &p, const pgh_udp *udp, dev_buf &buf, hipStream_t s)
{
    if (p.kind != PGH_MO_SPHERE) return PGH_OK;
    const size_t bytes = static_cast<size_t>(p.nf) * p.nx * sizeof(double);
    pgh_status st = buf.reserve(bytes);
    if (st != PGH_OK) return st;
    PGH_HIP(hipMemcpyAsync(buf.ptr, udp->centers, bytes, hipMemcpyHostToDevice, s));
    p.centers = static_cast<const double *>(buf.ptr);
    return PGH_OK;
}

} // namespace
} // namespace pgh

using namespace pgh;

extern "C" {

const char *pgh_last_error(void)
{
    return g_last_error.c_str();
}

int pgh_abi_version(void)
{
    return PGH_ABI_VERSION;
}

int pgh_device_available(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

pgh_status pgh_problem_dims(const pgh_udp *udp, uint64_t *nx, uint64_t *nf)
try {
    prob p;
    PGH_TRY(resolve(udp, p));
    if (nx) *nx = p.nx;
    if (nf) *nf = p.nf;
    return PGH_OK;
} catch (...) {
    return api_exception();
}

int pgh_supported(const pgh_udp *udp)
try {
    const std::string saved = g_last_error;
    prob p;
    const bool ok = resolve(udp, p) == PGH_OK;
    g_last_error = saved;
    return ok ? 1 : 0;
} catch (...) {
    return 0;
}

pgh_status pgh_bfe_eval(const pgh_udp *udp, const double *dvs, uint64_t n_dvs, uint64_t nx, double *fvs, uint64_t nf,
                        void *stream)
try {
    prob p;
    PGH_TRY(resolve(udp, p));
    PGH_TRY(reject_shift(p, "pgh_bfe_eval"));
    PGH_TRY(check_shape(p, nx, nf));
    if (n_dvs && (!dvs || !fvs)) return fail(PGH_INVALID_ARGUMENT, "null buffer");
    return launch_fitness(p, dvs, n_dvs, fvs, static_cast<hipStream_t>(stream));
} catch (...) {
    return api_exception();
}

pgh_status pgh_bfe_eval_host(const pgh_udp *udp, const double *dvs, uint64_t dvs_len, double *fvs, uint64_t fvs_len)
try {
    prob p;
    PGH_TRY(resolve(udp, p));
    uint64_t n_dvs = 0;
    PGH_TRY(check_input_len(p, dvs_len, n_dvs));
    if (fvs_len != n_dvs * p.nf) {
        return fail(PGH_INVALID_ARGUMENT, "the fitness buffer holds " + std::to_string(fvs_len)
                                              + " values, while " + std::to_string(n_dvs * p.nf) + " are produced");
    }
    if (n_dvs == 0u) return PGH_OK;
    thread_ctx *ctx = nullptr;
    PGH_TRY(get_thread_ctx(&ctx));
    const size_t in_bytes = dvs_len * sizeof(double), out_bytes = fvs_len * sizeof(double);
    PGH_TRY(ctx->a.reserve(in_bytes));
    PGH_TRY(ctx->b.reserve(out_bytes));
    PGH_TRY(stage_centers(p, udp, ctx->c, ctx->stream));
    PGH_TRY(upload_host(ctx, ctx->a.ptr, dvs, in_bytes));
    PGH_TRY(apply_shift(p, static_cast<double *>(ctx->a.ptr), dvs_len, ctx));
    PGH_TRY(launch_fitness(p, static_cast<const double *>(ctx->a.ptr), n_dvs, static_cast<double *>(ctx->b.ptr),
                           ctx->stream));
    PGH_HIP(hipMemcpyAsync(fvs, ctx->b.ptr, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
    PGH_HIP(hipStreamSynchronize(ctx->stream));
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_batch_gradient(const pgh_udp *udp, const double *dvs, uint64_t n_dvs, uint64_t nx, double *grad,
                              double *fvs, void *stream)
try {
    prob p;
    PGH_TRY(resolve(udp, p));
    PGH_TRY(reject_shift(p, "pgh_batch_gradient"));
    PGH_TRY(check_shape(p, nx, p.nf));
    PGH_TRY(check_has_gradient(p));
    if (n_dvs && (!dvs || !grad)) return fail(PGH_INVALID_ARGUMENT, "null buffer");
    return launch_gradient(p, dvs, n_dvs, grad, fvs, static_cast<hipStream_t>(stream));
} catch (...) {
    return api_exception();
}

pgh_status pgh_batch_gradient_host(const pgh_udp *udp, const double *dvs, uint64_t dvs_len, double *grad,
                                   uint64_t grad_len, double *fvs)
try {
    prob p;
    PGH_TRY(resolve(udp, p));
    PGH_TRY(check_has_gradient(p));
    uint64_t n_dvs = 0;
    PGH_TRY(check_input_len(p, dvs_len, n_dvs));
    if (grad_len != dvs_len) {
        return fail(PGH_INVALID_ARGUMENT, "the gradient buffer holds " + std::to_string(grad_len)
                                              + " values, while " + std::to_string(dvs_len) + " are produced");
    }
    if (n_dvs == 0u) return PGH_OK;
    thread_ctx *ctx = nullptr;
    PGH_TRY(get_thread_ctx(&ctx));
    const size_t bytes = dvs_len * sizeof(double);
    PGH_TRY(ctx->a.reserve(bytes));
    PGH_TRY(ctx->b.reserve(bytes));
    PGH_TRY(ctx->d.reserve(n_dvs * sizeof(double)));
    PGH_TRY(upload_host(ctx, ctx->a.ptr, dvs, bytes));
    // translate forwards the gradient at x - t (translate.cpp:245-249)
    PGH_TRY(apply_shift(p, static_cast<double *>(ctx->a.ptr), dvs_len, ctx));
    PGH_TRY(launch_gradient(p, static_cast<const double *>(ctx->a.ptr), n_dvs, static_cast<double *>(ctx->b.ptr),
                            fvs ? static_cast<double *>(ctx->d.ptr) : nullptr, ctx->stream));
    PGH_HIP(hipMemcpyAsync(grad, ctx->b.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    if (fvs) {
        PGH_HIP(hipMemcpyAsync(fvs, ctx->d.ptr, n_dvs * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    PGH_HIP(hipStreamSynchronize(ctx->stream));
    return PGH_OK;
} catch (...) {
    return api_exception();
}

uint64_t pgh_local_descent_work_len(uint64_t n_dvs, uint64_t nx)
{
    return local_descent_work_len(n_dvs, nx);
}

pgh_status pgh_local_descent(const pgh_udp *udp, double *dvs, uint64_t n_dvs, uint64_t nx, double *fvs,
                             const double *lb, const double *ub, uint32_t iters, double step0, double *work,
                             uint64_t work_len, void *stream)
try {
    prob p;
    PGH_TRY(resolve(udp, p));
    PGH_TRY(check_shape(p, nx, p.nf));
    PGH_TRY(check_local_descent(p, lb, ub, step0));
    if (n_dvs == 0u) return PGH_OK;
    if (!dvs || !fvs || !work) return fail(PGH_INVALID_ARGUMENT, "null buffer");
    const uint64_t need = local_descent_work_len(n_dvs, nx);
    if (work_len < need) {
        return fail(PGH_INVALID_ARGUMENT, "the work buffer holds " + std::to_string(work_len) + " doubles, while "
                                              + std::to_string(need) + " are needed");
    }
    return launch_local_descent(p, dvs, n_dvs, fvs, lb, ub, iters, step0, work, static_cast<hipStream_t>(stream));
} catch (...) {
    return api_exception();
}

pgh_status pgh_local_descent_host(const pgh_udp *udp, double *dvs, uint64_t dvs_len, double *fvs, uint64_t fvs_len,
                                  const double *lb, const double *ub, uint32_t iters, double step0)
try {
    prob p;
    PGH_TRY(resolve(udp, p));
    PGH_TRY(check_local_descent(p, lb, ub, step0));
    uint64_t n_dvs = 0;
    PGH_TRY(check_input_len(p, dvs_len, n_dvs));
    if (fvs_len != n_dvs * p.nf) {
        return fail(PGH_INVALID_ARGUMENT, "the fitness buffer holds " + std::to_string(fvs_len)
                                              + " values, while " + std::to_string(n_dvs * p.nf) + " are produced");
    }
    if (n_dvs == 0u) return PGH_OK;
    thread_ctx *ctx = nullptr;
    PGH_TRY(get_thread_ctx(&ctx));
    const size_t x_bytes = dvs_len * sizeof(double);
    PGH_TRY(ctx->a.reserve(x_bytes));
    PGH_TRY(ctx->b.reserve(local_descent_work_len(n_dvs, p.nx) * sizeof(double)));
    PGH_TRY(ctx->d.reserve(n_dvs * sizeof(double)));
    const double *dlb = nullptr, *dub = nullptr;
    if (lb) {
        PGH_TRY(ctx->c.reserve(2u * p.nx * sizeof(double)));
        double *c = static_cast<double *>(ctx->c.ptr);
        PGH_HIP(hipMemcpyAsync(c, lb, p.nx * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        PGH_HIP(hipMemcpyAsync(c + p.nx, ub, p.nx * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        dlb = c;
        dub = c + p.nx;
    }
    double *x = static_cast<double *>(ctx->a.ptr);
    double *f = static_cast<double *>(ctx->d.ptr);
    PGH_TRY(upload_host(ctx, x, dvs, x_bytes));
    PGH_TRY(launch_local_descent(p, x, n_dvs, f, dlb, dub, iters, step0, static_cast<double *>(ctx->b.ptr),
                                 ctx->stream));
    PGH_HIP(hipMemcpyAsync(dvs, x, x_bytes, hipMemcpyDeviceToHost, ctx->stream));
    PGH_HIP(hipMemcpyAsync(fvs, f, n_dvs * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PGH_HIP(hipStreamSynchronize(ctx->stream));
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_mt19937_generate(uint32_t seed, uint64_t skip, uint64_t n, uint32_t *out, void *stream)
try {
    if (n && !out) return fail(PGH_INVALID_ARGUMENT, "null buffer");
    return launch_mt19937(seed, skip, n, out, static_cast<hipStream_t>(stream));
} catch (...) {
    return api_exception();
}

pgh_status pgh_batch_random_dv(const pgh_udp *udp, uint64_t n, uint32_t seed, double *out, void *stream)
try {
    prob p;
    PGH_TRY(resolve(udp, p));
    PGH_TRY(reject_shift(p, "pgh_batch_random_dv"));
    if (p.kind == PGH_ZDT || p.kind == PGH_DTLZ) {
        // ZDT4's first bound and ZDT5's integer part are not drawn here
        return fail(PGH_UNSUPPORTED, "batch_random_decision_vector on the device covers the continuous problems "
                                     "with a kernel except ZDT and DTLZ");
    }
    if (n && !out) return fail(PGH_INVALID_ARGUMENT, "null buffer");
    return launch_batch_random_dv(p, n, seed, out, static_cast<hipStream_t>(stream));
} catch (...) {
    return api_exception();
}

} // extern "C"
