// local.hip -- memetic local step: batched projected gradient descent with a
// step length per individual, on the gradient and fitness kernels of
// fitness.hip.  The reference has no such routine; the update rule is the one
// documented for pgh_local_descent in pagmo_hip/pgh.h.  Every individual is
// handled by one wavefront and its arithmetic does not depend on the launch,
// so results are the same for any batch composition.
#include <cfloat>
#include <mutex>
#include <string>

#include "pgh_internal.hpp"

namespace pgh
{
namespace
{

constexpr uint32_t kBlock = 256;
constexpr uint32_t kWaves = kBlock / 64;

__global__ __launch_bounds__(kBlock) void k_fill_alpha(double *__restrict__ alpha, uint64_t n, double v)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) alpha[i] = v;
}

// y_i = clamp(x_i - (alpha_i / max_j |g_ij|) * g_i, lb, ub); y_i = x_i when the
// gradient has a non-finite element or is zero.
__global__ __launch_bounds__(kBlock) void k_trial(const double *__restrict__ x, const double *__restrict__ g,
                                                  const double *__restrict__ alpha, const double *__restrict__ lb,
                                                  const double *__restrict__ ub, uint64_t n, uint32_t nx,
                                                  double *__restrict__ y)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kWaves + threadIdx.x / 64u;
    const uint32_t lane = threadIdx.x % 64u;
    if (i >= n) return; // i is uniform over the wavefront
    const double *xi = x + i * nx;
    const double *gi = g + i * nx;
    double *yi = y + i * nx;
    double m = 0.;
    int bad = 0;
    for (uint32_t j = lane; j < nx; j += 64u) {
        const double a = fabs(gi[j]);
        bad |= !(a <= DBL_MAX); // inf or NaN
        m = a > m ? a : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double mo = __shfl_xor(m, o, 64);
        m = mo > m ? mo : m;
        bad |= __shfl_xor(bad, o, 64);
    }
    const bool move = !bad && m > 0.;
    const double s = move ? alpha[i] / m : 0.;
    for (uint32_t j = lane; j < nx; j += 64u) {
        double v = xi[j];
        if (move) {
            v = v - s * gi[j];
            if (lb) {
                if (v < lb[j]) v = lb[j];
                if (v > ub[j]) v = ub[j];
            }
        }
        yi[j] = v;
    }
}

// Accept y_i where f(y_i) < f(x_i): x_i = y_i, f_i = f(y_i), alpha_i doubled
// (kept finite); otherwise alpha_i halved.  f_i and fy_i are read by every
// lane before lane 0 writes f_i.
__global__ __launch_bounds__(kBlock) void k_accept(const double *__restrict__ y, const double *__restrict__ fy,
                                                   uint64_t n, uint32_t nx, double *__restrict__ x,
                                                   double *__restrict__ f, double *__restrict__ alpha)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kWaves + threadIdx.x / 64u;
    const uint32_t lane = threadIdx.x % 64u;
    if (i >= n) return;
    const double fi = f[i];
    const double fyi = fy[i];
    if (fyi < fi) {
        for (uint32_t j = lane; j < nx; j += 64u) x[i * nx + j] = y[i * nx + j];
        if (lane == 0u) {
            f[i] = fyi;
            const double a2 = 2. * alpha[i];
            if (a2 <= DBL_MAX) alpha[i] = a2;
        }
    } else if (lane == 0u) {
        alpha[i] = 0.5 * alpha[i];
    }
}

} // namespace

uint64_t local_descent_work_len(uint64_t n, uint64_t nx)
{
    return 2u * n * nx + 2u * n;
}

pgh_status launch_local_descent(const prob &p, double *x, uint64_t n, double *f, const double *lb, const double *ub,
                                uint32_t iters, double step0, double *work, hipStream_t s)
{
    if (n == 0u) return PGH_OK;
    PGH_TRY(check_launch_size(n, p.nx));
    const uint64_t nx = p.nx;
    double *y = work;
    double *g = y + n * nx;
    double *fy = g + n * nx;
    double *alpha = fy + n;
    const dim3 rows(static_cast<uint32_t>((n + kWaves - 1u) / kWaves));
    PGH_TRY(launch_fitness(p, x, n, f, s));
    k_fill_alpha<<<dim3(static_cast<uint32_t>((n + kBlock - 1u) / kBlock)), dim3(kBlock), 0, s>>>(alpha, n, step0);
    PGH_HIP(hipGetLastError());
    for (uint32_t it = 0u; it < iters; ++it) {
        PGH_TRY(launch_gradient(p, x, n, g, nullptr, s));
        k_trial<<<rows, dim3(kBlock), 0, s>>>(x, g, alpha, lb, ub, n, p.nx, y);
        PGH_HIP(hipGetLastError());
        PGH_TRY(launch_fitness(p, y, n, fy, s));
        k_accept<<<rows, dim3(kBlock), 0, s>>>(y, fy, n, p.nx, x, f, alpha);
        PGH_HIP(hipGetLastError());
    }
    return PGH_OK;
}

namespace
{

// The work buffer of pgh_local_descent_noalloc: one copy per device, made
// when the code object is loaded and kept until the process ends.  In host
// code `scratch` only names the variable; its device address on the current
// device comes from hipGetSymbolAddress (scratch_on_current_device).
static __device__ __align__(256) double scratch[PGH_LOCAL_DESCENT_NOALLOC_MAX_WORK];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call made with that device current.
pgh_status scratch_on_current_device(double **out)
{
    int dev = 0;
    PGH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(PGH_DEVICE_ERROR, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(PGH_DEVICE_ERROR, "hipGetSymbolAddress(scratch) gave a null address");
        g_scratch_ptr[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return PGH_OK;
}

} // namespace
} // namespace pgh

using namespace pgh;

extern "C" {

pgh_status pgh_local_descent_noalloc(const pgh_udp *udp, double *dvs, uint64_t n_dvs, uint64_t nx, double *fvs,
                                     const double *lb, const double *ub, uint32_t iters, double step0, void *stream)
try {
    prob p;
    PGH_TRY(resolve(udp, p));
    PGH_TRY(check_shape(p, nx, p.nf));
    PGH_TRY(check_local_descent(p, lb, ub, step0));
    if (n_dvs == 0u) return PGH_OK;
    if (!dvs || !fvs) return fail(PGH_INVALID_ARGUMENT, "null buffer");
    // local_descent_work_len without overflow: 2 n (nx + 1) doubles
    const uint64_t cap = PGH_LOCAL_DESCENT_NOALLOC_MAX_WORK;
    if (n_dvs > cap / 2u / (nx + 1u)) {
        return fail(PGH_INVALID_ARGUMENT, "pgh_local_descent_noalloc: " + std::to_string(n_dvs) + " individuals of "
                                              + std::to_string(nx) + " dimensions need a work buffer beyond its "
                                              + std::to_string(cap) + " doubles; use pgh_local_descent");
    }
    double *work = nullptr;
    PGH_TRY(scratch_on_current_device(&work));
    return launch_local_descent(p, dvs, n_dvs, fvs, lb, ub, iters, step0, work, static_cast<hipStream_t>(stream));
} catch (...) {
    return api_exception();
}

} // extern "C"
