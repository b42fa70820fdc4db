// rng.hip -- seeded inputs on the device: std::mt19937 words and
// batch_random_decision_vector (include/pagmo/utils/generic.hpp:326-388).
#include "fitness_dev.hpp"
#include "mt_dev.hpp"
#include "pgh_internal.hpp"

namespace pgh
{
using namespace dev;

namespace
{

constexpr uint32_t kMtBlock = 256;

// One workgroup: seed, discard `skip` words, hand out n words.
__global__ __launch_bounds__(kMtBlock) void k_mt19937(uint32_t seed, uint64_t skip, uint64_t n,
                                                      uint32_t *__restrict__ out)
{
    __shared__ uint32_t ring[kMtRing];
    mt_seed_lds(seed, ring);
    uint64_t K = 0;
    if (skip) mt_run(ring, K, 0, skip, nullptr);
    mt_run(ring, K, skip, n, out);
}

// In place: slot m holds the words 2m, 2m+1 and becomes U(lb_j, ub_j).
__global__ __launch_bounds__(256) void k_uniform_inplace(double *__restrict__ io, uint64_t count, uint32_t nx,
                                                         int kind)
{
    const uint64_t m = static_cast<uint64_t>(blockIdx.x) * 256u + threadIdx.x;
    if (m >= count) return;
    const uint2 w = reinterpret_cast<const uint2 *>(io)[m];
    double lb, ub;
    bounds_of(kind, static_cast<uint32_t>(m % nx), lb, ub);
    io[m] = mt_uniform(w.x, w.y, lb, ub);
}

} // namespace

pgh_status launch_mt19937(uint32_t seed, uint64_t skip, uint64_t n, uint32_t *out, hipStream_t s)
{
    if (n == 0u) return PGH_OK;
    k_mt19937<<<dim3(1), dim3(kMtBlock), 0, s>>>(seed, skip, n, out);
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

pgh_status launch_batch_random_dv(const prob &p, uint64_t n, uint32_t seed, double *out, hipStream_t s)
{
    if (p.kind == ZDT || p.kind == DTLZ) return fail(PGH_UNSUPPORTED, "no device draws for ZDT and DTLZ");
    const uint64_t count = n * p.nx;
    if (count == 0u) return PGH_OK;
    // every bound pair of the supported problems has lb < ub, so each element
    // consumes exactly two words (generic.hpp:376-380)
    k_mt19937<<<dim3(1), dim3(kMtBlock), 0, s>>>(seed, 0, 2u * count, reinterpret_cast<uint32_t *>(out));
    PGH_HIP(hipGetLastError());
    k_uniform_inplace<<<dim3(static_cast<uint32_t>((count + 255u) / 256u)), dim3(256), 0, s>>>(out, count, p.nx,
                                                                                                p.kind);
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

} // namespace pgh
