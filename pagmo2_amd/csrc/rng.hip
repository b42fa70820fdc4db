// rng.hip -- seeded inputs on the device: std::mt19937 words and
// batch_random_decision_vector (include/pagmo/utils/generic.hpp:326-388).
#include "fitness_dev.hpp"
#include "mt_dev.hpp"
#include "pgh_internal.hpp"

#include <vector>

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

// Segment blockIdx.x = s of a stream split in segments of B words: words
// [s B, min((s + 1) B, n)) of std::mt19937(seed), written to out at their
// stream positions.  The segment's generator reaches s B by one jump,
// z^(s B) (entry s - 1 of the jump table).
__global__ __launch_bounds__(kMtJumpBlock) void k_mt19937_seg(uint32_t seed, uint64_t n, uint64_t B, uint32_t nd,
                                                               const void *__restrict__ table,
                                                               uint32_t *__restrict__ out)
{
    __shared__ uint32_t ring[kMtRing];
    __shared__ uint4 lds_idx[kMtIdxStride / 8u];
    mt_seed_lds(seed, ring);
    uint64_t K = 0, C = 0;
    const uint32_t s = blockIdx.x;
    const uint32_t *lens = static_cast<const uint32_t *>(table);
    const uint16_t *lists = reinterpret_cast<const uint16_t *>(static_cast<const char *>(table) + jump_table_head(nd));
    if (s > 0u) {
        mt_jump(ring, K, C, static_cast<uint64_t>(s) * B, lists + static_cast<uint64_t>(s - 1u) * kMtIdxStride,
                lens[s - 1u], lds_idx);
    }
    const uint64_t begin = static_cast<uint64_t>(s) * B;
    mt_run(ring, K, C, min(B, n - begin), out + begin);
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

// The first n words of std::mt19937(seed) into out: one workgroup up to
// kMtSplitMin words, else up to kMtMaxSegs workgroups, each starting its
// segment by one jump.  The jump polynomials z^(s B) depend on B alone: the
// host makes them once per process, as consecutive multiples
// (jump_polys_multiples), and one device table per B holds all kMtMaxSegs - 1
// of them (kept for the process: at most one per power of two B, 12 MB each),
// whatever the number of segments of a call.
constexpr uint64_t kMtSplitMin = uint64_t(1) << 21;
constexpr uint32_t kMtMaxSegs = 256;

pgh_status launch_mt_words(uint32_t seed, uint64_t n, uint32_t *out, hipStream_t s)
{
    if (n <= kMtSplitMin) {
        k_mt19937<<<dim3(1), dim3(kMtBlock), 0, s>>>(seed, 0, n, out);
        PGH_HIP(hipGetLastError());
        return PGH_OK;
    }
    // B: the power of two that splits n into at most kMtMaxSegs segments
    uint64_t B = 1;
    while (B * kMtMaxSegs < n) B <<= 1;
    const uint32_t segs = static_cast<uint32_t>((n + B - 1u) / B);
    std::vector<uint64_t> dists;
    for (uint32_t k = 1; k < kMtMaxSegs; ++k) dists.push_back(static_cast<uint64_t>(k) * B);
    PGH_TRY(jump_polys_multiples(B, kMtMaxSegs - 1u));
    const void *table = nullptr;
    PGH_TRY(jump_table_device(dists, &table));
    k_mt19937_seg<<<dim3(segs), dim3(kMtJumpBlock), 0, s>>>(seed, n, B, static_cast<uint32_t>(dists.size()), table,
                                                             out);
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

pgh_status launch_mt19937(uint32_t seed, uint64_t skip, uint64_t n, uint32_t *out, hipStream_t s)
{
    if (n == 0u) return PGH_OK;
    if (skip == 0u) return launch_mt_words(seed, n, out, s);
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
    PGH_TRY(launch_mt_words(seed, 2u * count, reinterpret_cast<uint32_t *>(out), s));
    k_uniform_inplace<<<dim3(static_cast<uint32_t>((count + 255u) / 256u)), dim3(256), 0, s>>>(out, count, p.nx,
                                                                                                p.kind);
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

} // namespace pgh
