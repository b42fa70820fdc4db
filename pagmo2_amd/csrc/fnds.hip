// fnds.hip -- fast_non_dominated_sorting (src/utils/multi_objective.cpp:200-257)
// on the device, bit-exact with the reference including the order of every
// front and of every domination list.
//
// Pipeline (all on one stream):
//   k_keys      objectives -> order-preserving int64 keys in which NaN sorts
//               after +inf and equals every NaN, and -0 equals +0
//               (detail::less_than_f/greater_than_f/equal_to_f,
//               include/pagmo/detail/custom_comparisons.hpp:54-100)
//   k_pairs     one lane per row i against a chunk of columns j: builds the
//               row-major dominance bit matrix D (bit j of row i set iff i
//               dominates j, multi_objective.cpp:97-113), the number of rows
//               dominating i (dom_count) and |dom_list[i]| (popcount of row i)
//   k_scan      exclusive prefix of |dom_list[i]| -> dom_offsets
//   k_emit      dom_list[i] = set bits of row i in increasing j: exactly the
//               reference's push_back order (multi_objective.cpp:215-227 appends
//               j < i while scanning row i, then larger indices in scan order)
//   k_peel      the front peel (multi_objective.cpp:228-254) in one workgroup:
//               F0 ascending; then per level, decrement the counters of the
//               rows dominated by the current front, mark the elements whose
//               counter reached zero with the position (in the current front)
//               of their last dominator, and emit them ordered by (that
//               position, index) -- the order in which the reference's nested
//               loop appends them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "pgh_internal.hpp"

struct pgh_fnds_ws {
    hipStream_t stream = nullptr;
    hipStream_t own = nullptr; // created by pgh_fnds_run_host
    pgh::dev_buf staged;
    uint64_t n = 0, m = 0, words = 0;
    pgh::dev_buf keys, D, cnt32, ndom32, key32, dom_count, rank, front_items, front_offsets, dom_offsets, dom_items,
        scalars;
    uint64_t h_sizes[2] = {0, 0}; // n_fronts, n_dom_items
    bool sized = false;
};

namespace pgh
{
namespace
{

constexpr uint32_t kBlock = 256;
constexpr uint32_t kPeelBlock = 1024;

__device__ __forceinline__ int64_t order_key(double x)
{
    if (x != x) return INT64_MAX;
    if (x == 0.0) return 0; // -0 == +0
    const int64_t b = __double_as_longlong(x);
    return b >= 0 ? b : (b ^ INT64_MAX);
}

__global__ __launch_bounds__(kBlock) void k_keys(const double *__restrict__ objs, uint64_t count,
                                                 int64_t *__restrict__ keys)
{
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (t < count) keys[t] = order_key(objs[t]);
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *__restrict__ p, uint64_t count, uint32_t v)
{
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (t < count) p[t] = v;
}

// Row i (one lane) against columns [j0, j0 + JC).  MT > 0: the row's keys are
// held in registers; MT == 0: any m, keys re-read from memory.
template <int MT>
__global__ __launch_bounds__(kBlock) void k_pairs(const int64_t *__restrict__ keys, uint32_t n, uint32_t m,
                                                  uint32_t jc, uint64_t *__restrict__ D, uint32_t words,
                                                  uint32_t *__restrict__ dom_count, uint32_t *__restrict__ ndom)
{
    extern __shared__ int64_t kj[]; // jc x m
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t j0 = blockIdx.y * jc;
    const uint32_t jn = min(jc, n - j0);
    for (uint32_t t = threadIdx.x; t < jn * m; t += kBlock) kj[t] = keys[static_cast<uint64_t>(j0) * m + t];
    __syncthreads();
    if (i >= n) return;
    int64_t ki[MT > 0 ? MT : 1];
    if constexpr (MT > 0) {
#pragma unroll
        for (int q = 0; q < MT; ++q) ki[q] = keys[static_cast<uint64_t>(i) * MT + q];
    }
    uint32_t dominated_by = 0, dominates = 0;
    for (uint32_t jb = 0; jb < jn; jb += 64u) {
        uint64_t word = 0;
        const uint32_t je = min(jb + 64u, jn);
        for (uint32_t j = jb; j < je; ++j) {
            bool lt = false, gt = false;
            if constexpr (MT > 0) {
#pragma unroll
                for (int q = 0; q < MT; ++q) {
                    const int64_t a = ki[q], b = kj[j * MT + q];
                    lt |= a < b;
                    gt |= a > b;
                }
            } else {
                for (uint32_t q = 0; q < m; ++q) {
                    const int64_t a = keys[static_cast<uint64_t>(i) * m + q], b = kj[j * m + q];
                    lt |= a < b;
                    gt |= a > b;
                }
            }
            // pareto_dominance(i, j): no worse anywhere, better somewhere
            const bool idomj = lt && !gt;
            dominated_by += (gt && !lt) ? 1u : 0u;
            word |= idomj ? (uint64_t(1) << (j - jb)) : uint64_t(0);
        }
        D[static_cast<uint64_t>(i) * words + (j0 + jb) / 64u] = word;
        dominates += static_cast<uint32_t>(__popcll(word));
    }
    if (dominated_by) atomicAdd(&dom_count[i], dominated_by);
    if (dominates) atomicAdd(&ndom[i], dominates);
}

__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t v, uint32_t d)
{
    const uint32_t lo = __shfl_up(static_cast<uint32_t>(v), d), hi = __shfl_up(static_cast<uint32_t>(v >> 32), d);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Block-wide exclusive scan of one value per thread: wavefront scans, then a
// scan of the wavefront totals.  sh holds at least blockDim.x / 64 values.
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *sh, uint64_t &total)
{
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        const uint64_t y = shfl_up_u64(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63u) sh[wid] = x;
    __syncthreads();
    if (wid == 0) {
        uint64_t t = lane < nw ? sh[lane] : 0u;
#pragma unroll
        for (uint32_t d = 1; d < 64u; d <<= 1) {
            const uint64_t y = shfl_up_u64(t, d);
            if (lane >= d) t += y;
        }
        if (lane < nw) sh[lane] = t;
    }
    __syncthreads();
    total = sh[nw - 1u];
    const uint64_t excl = (wid ? sh[wid - 1u] : 0u) + x - v;
    __syncthreads();
    return excl;
}

// dom_offsets (n+1) = exclusive prefix of ndom; dom_count -> u64; scalars[1] = total.
__global__ __launch_bounds__(kPeelBlock) void k_scan(const uint32_t *__restrict__ ndom,
                                                     const uint32_t *__restrict__ cnt32, uint32_t n,
                                                     uint64_t *__restrict__ offsets, uint64_t *__restrict__ dom_count,
                                                     uint64_t *__restrict__ scalars)
{
    __shared__ uint64_t sh[kPeelBlock];
    const uint32_t per = (n + kPeelBlock - 1u) / kPeelBlock;
    const uint32_t b = threadIdx.x * per, e = min(b + per, n);
    uint64_t local = 0;
    for (uint32_t i = b; i < e; ++i) local += ndom[i];
    uint64_t total;
    uint64_t off = block_exclusive_scan(local, sh, total);
    for (uint32_t i = b; i < e; ++i) {
        offsets[i] = off;
        off += ndom[i];
        dom_count[i] = cnt32[i];
    }
    if (threadIdx.x == 0) {
        offsets[n] = total;
        scalars[1] = total;
    }
}

// dom_items: one wavefront per row; lane l takes words l, l+64, ... of the row
// (coalesced), and a wavefront scan of the popcounts places every set bit,
// so each row's items come out in increasing column order.
__global__ __launch_bounds__(kBlock) void k_emit(const uint64_t *__restrict__ D, uint32_t n, uint32_t words,
                                                 const uint64_t *__restrict__ offsets, uint64_t *__restrict__ items)
{
    const uint32_t row = blockIdx.x * (kBlock / 64u) + threadIdx.x / 64u;
    const uint32_t lane = threadIdx.x & 63u;
    if (row >= n) return; // whole wavefronts retire together
    uint64_t o = offsets[row];
    const uint64_t *r = D + static_cast<uint64_t>(row) * words;
    for (uint32_t w0 = 0; w0 < words; w0 += 64u) {
        const uint32_t w = w0 + lane;
        uint64_t word = w < words ? r[w] : 0u;
        const uint32_t c = static_cast<uint32_t>(__popcll(word));
        uint32_t x = c;
#pragma unroll
        for (uint32_t d = 1; d < 64u; d <<= 1) {
            const uint32_t y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        const uint32_t total = __shfl(x, 63);
        uint64_t pos = o + (x - c);
        while (word) {
            items[pos++] = static_cast<uint64_t>(w) * 64u +
                           static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(word)) - 1);
            word &= word - 1u;
        }
        o += total;
    }
}

// A thread's next (member, word) items of a level, loaded together so that
// several global loads are in flight before any bit is processed.
constexpr uint32_t kPeelBatch = 8u;
struct peel_batch {
    uint64_t word[kPeelBatch];
    uint32_t p[kPeelBatch], w[kPeelBatch];
    __device__ __forceinline__ void load(const uint64_t *__restrict__ D, const uint64_t *__restrict__ fr,
                                         uint32_t words, uint64_t it0, uint64_t ie)
    {
#pragma unroll
        for (uint32_t q = 0; q < kPeelBatch; ++q) {
            const uint64_t it = it0 + q;
            const bool ok = it < ie;
            const uint64_t itc = ok ? it : it0;
            p[q] = static_cast<uint32_t>(itc / words);
            w[q] = static_cast<uint32_t>(itc - static_cast<uint64_t>(p[q]) * words);
            const uint64_t v = D[fr[p[q]] * words + w[q]];
            word[q] = ok ? v : 0u;
        }
    }
};

// The front peel with the counters in LDS (n <= 65535 and 8n bytes of LDS).
// Level r spreads the (member, word) pairs of front r over the workgroup in
// order; key[j] = ((r+1) << 16) | p is the largest position p of a member of
// front r dominating j, tagged with the level so that values left by earlier
// levels never win.
__global__ __launch_bounds__(kPeelBlock) void k_peel_lds(const uint64_t *__restrict__ D, uint32_t n, uint32_t words,
                                                         const uint32_t *__restrict__ dom_count,
                                                         uint64_t *__restrict__ front_items,
                                                         uint64_t *__restrict__ front_offsets,
                                                         uint64_t *__restrict__ rank, uint64_t *__restrict__ scalars)
{
    extern __shared__ uint32_t lds_u32[];
    uint32_t *cnt = lds_u32;
    uint32_t *key = lds_u32 + n;
    __shared__ uint64_t sh[kPeelBlock / 64u];
    const uint32_t t = threadIdx.x;
    for (uint32_t j = t; j < n; j += kPeelBlock) {
        cnt[j] = dom_count[j];
        key[j] = 0u;
    }
    __syncthreads();
    uint64_t total;
    {
        // F0: ascending indices with dom_count == 0 (multi_objective.cpp:228-233)
        const uint32_t per = (n + kPeelBlock - 1u) / kPeelBlock;
        const uint32_t b = min(t * per, n), e = min(b + per, n);
        uint64_t c = 0;
        for (uint32_t i = b; i < e; ++i) c += cnt[i] == 0u;
        uint64_t o = block_exclusive_scan(c, sh, total);
        for (uint32_t i = b; i < e; ++i) {
            if (cnt[i] == 0u) {
                front_items[o++] = i;
                rank[i] = 0;
            }
        }
        if (t == 0) {
            front_offsets[0] = 0;
            front_offsets[1] = total;
        }
    }
    uint64_t s = 0, e = total;
    uint32_t level = 0;
    __syncthreads();
    while (e > s) {
        const uint64_t items = (e - s) * words;
        const uint64_t per = (items + kPeelBlock - 1u) / kPeelBlock;
        const uint64_t ib = min(static_cast<uint64_t>(t) * per, items), ie = min(ib + per, items);
        const uint32_t tag = (level + 1u) << 16;
        const uint64_t *fr = front_items + s;
        // pass 1: decrement the counters of everything front r dominates
        for (uint64_t it0 = ib; it0 < ie; it0 += kPeelBatch) {
            peel_batch b;
            b.load(D, fr, words, it0, ie);
#pragma unroll
            for (uint32_t q = 0; q < kPeelBatch; ++q) {
                uint64_t word = b.word[q];
                while (word) {
                    const uint32_t j =
                        b.w[q] * 64u + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(word)) - 1);
                    atomicSub(&cnt[j], 1u);
                    atomicMax(&key[j], tag | b.p[q]);
                    word &= word - 1u;
                }
            }
        }
        __syncthreads();
        // pass 2: count, in (member, word, bit) order, the elements whose last
        // dominator is this member; pass 3 writes them at their positions
        uint64_t c = 0;
        for (uint64_t it0 = ib; it0 < ie; it0 += kPeelBatch) {
            peel_batch b;
            b.load(D, fr, words, it0, ie);
#pragma unroll
            for (uint32_t q = 0; q < kPeelBatch; ++q) {
                uint64_t word = b.word[q];
                while (word) {
                    const uint32_t j =
                        b.w[q] * 64u + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(word)) - 1);
                    c += (cnt[j] == 0u && key[j] == (tag | b.p[q])) ? 1u : 0u;
                    word &= word - 1u;
                }
            }
        }
        uint64_t o = e + block_exclusive_scan(c, sh, total);
        for (uint64_t it0 = ib; it0 < ie; it0 += kPeelBatch) {
            peel_batch b;
            b.load(D, fr, words, it0, ie);
#pragma unroll
            for (uint32_t q = 0; q < kPeelBatch; ++q) {
                uint64_t word = b.word[q];
                while (word) {
                    const uint32_t j =
                        b.w[q] * 64u + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(word)) - 1);
                    if (cnt[j] == 0u && key[j] == (tag | b.p[q])) {
                        front_items[o++] = j;
                        rank[j] = level + 1u;
                    }
                    word &= word - 1u;
                }
            }
        }
        ++level;
        if (t == 0 && total) front_offsets[level + 1u] = e + total;
        __syncthreads();
        s = e;
        e = e + total;
    }
    if (t == 0) scalars[0] = level;
}

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The front peel for large n: counters in global memory, touched only with
// atomics (which bypass the non-coherent L1).
__global__ __launch_bounds__(kPeelBlock) void k_peel(const uint64_t *__restrict__ D, uint32_t n, uint32_t words,
                                                     uint32_t *__restrict__ cnt, uint32_t *__restrict__ key,
                                                     uint64_t *__restrict__ front_items,
                                                     uint64_t *__restrict__ front_offsets, uint64_t *__restrict__ rank,
                                                     uint64_t *__restrict__ scalars)
{
    __shared__ uint64_t sh[kPeelBlock];
    const uint32_t t = threadIdx.x;
    // F0: ascending indices with dom_count == 0
    {
        const uint32_t per = (n + kPeelBlock - 1u) / kPeelBlock;
        const uint32_t b = t * per, e = min(b + per, n);
        uint64_t c = 0;
        for (uint32_t i = b; i < e; ++i) c += ld_relaxed(&cnt[i]) == 0u;
        uint64_t total;
        uint64_t o = block_exclusive_scan(c, sh, total);
        for (uint32_t i = b; i < e; ++i) {
            if (ld_relaxed(&cnt[i]) == 0u) {
                front_items[o++] = i;
                rank[i] = 0;
            }
        }
        if (t == 0) {
            front_offsets[0] = 0;
            front_offsets[1] = total;
        }
        __syncthreads();
        sh[0] = total;
        __syncthreads();
    }
    uint64_t s = 0, e = sh[0];
    uint64_t level = 0;
    __syncthreads();
    while (e > s) {
        const uint64_t len = e - s;
        const uint64_t per = (len + kPeelBlock - 1u) / kPeelBlock;
        const uint64_t pb = s + min(static_cast<uint64_t>(t) * per, len), pe = s + min((t + 1u) * per, len);
        // pass 1: decrement the counters of everything the front dominates
        for (uint64_t p = pb; p < pe; ++p) {
            const uint64_t *row = D + front_items[p] * words;
            for (uint32_t w = 0; w < words; ++w) {
                uint64_t word = row[w];
                while (word) {
                    const uint32_t j = w * 64u + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(word)) - 1);
                    atomicSub(&cnt[j], 1u);
                    word &= word - 1u;
                }
            }
        }
        __threadfence();
        __syncthreads();
        // pass 2: elements completed at this level record their last dominator
        for (uint64_t p = pb; p < pe; ++p) {
            const uint64_t *row = D + front_items[p] * words;
            const uint32_t tag = static_cast<uint32_t>(p - s) + 1u;
            for (uint32_t w = 0; w < words; ++w) {
                uint64_t word = row[w];
                while (word) {
                    const uint32_t j = w * 64u + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(word)) - 1);
                    if (ld_relaxed(&cnt[j]) == 0u) atomicMax(&key[j], tag);
                    word &= word - 1u;
                }
            }
        }
        __threadfence();
        __syncthreads();
        // pass 3: emit in (last dominator position, index) order
        uint64_t c = 0;
        for (uint64_t p = pb; p < pe; ++p) {
            const uint64_t *row = D + front_items[p] * words;
            const uint32_t tag = static_cast<uint32_t>(p - s) + 1u;
            for (uint32_t w = 0; w < words; ++w) {
                uint64_t word = row[w];
                while (word) {
                    const uint32_t j = w * 64u + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(word)) - 1);
                    c += (ld_relaxed(&cnt[j]) == 0u && ld_relaxed(&key[j]) == tag) ? 1u : 0u;
                    word &= word - 1u;
                }
            }
        }
        uint64_t total;
        uint64_t o = e + block_exclusive_scan(c, sh, total);
        for (uint64_t p = pb; p < pe; ++p) {
            const uint64_t *row = D + front_items[p] * words;
            const uint32_t tag = static_cast<uint32_t>(p - s) + 1u;
            for (uint32_t w = 0; w < words; ++w) {
                uint64_t word = row[w];
                while (word) {
                    const uint32_t j = w * 64u + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(word)) - 1);
                    if (ld_relaxed(&cnt[j]) == 0u && ld_relaxed(&key[j]) == tag) {
                        front_items[o++] = j;
                        rank[j] = level + 1u;
                    }
                    word &= word - 1u;
                }
            }
        }
        ++level;
        if (t == 0 && total) front_offsets[level + 1u] = e + total;
        __threadfence();
        __syncthreads();
        s = e;
        e = e + total;
    }
    if (t == 0) scalars[0] = level;
}

uint32_t grid_for(uint64_t count)
{
    return static_cast<uint32_t>((count + kBlock - 1u) / kBlock);
}

// Dynamic LDS a workgroup of k_peel_lds may use on the current device (the
// static scan scratch aside), with the kernel's limit raised to it once.
size_t max_dynamic_lds()
{
    static thread_local int dev_cached = -1;
    static thread_local size_t bytes = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (dev != dev_cached) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
        bytes = v > 1024 ? static_cast<size_t>(v) - 1024u : 0u;
        if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_peel_lds), hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(bytes)) != hipSuccess) {
            (void)hipGetLastError();
            bytes = 65536u - 1024u;
        }
        dev_cached = dev;
    }
    return bytes;
}

#define PGH_TRY(expr)                                                                                                \
    do {                                                                                                             \
        const pgh_status pgh_st_ = (expr);                                                                           \
        if (pgh_st_ != PGH_OK) return pgh_st_;                                                                       \
    } while (0)

} // namespace
} // namespace pgh

using namespace pgh;

extern "C" {

pgh_status pgh_fnds_create(pgh_fnds_ws **ws)
{
    if (!ws) return fail(PGH_INVALID_ARGUMENT, "null workspace pointer");
    *ws = new pgh_fnds_ws();
    return PGH_OK;
}

pgh_status pgh_fnds_destroy(pgh_fnds_ws *ws)
{
    if (ws && ws->stream) (void)hipStreamSynchronize(ws->stream);
    if (ws && ws->own) (void)hipStreamDestroy(ws->own);
    delete ws;
    return PGH_OK;
}

pgh_status pgh_fnds_run_host(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m)
{
    if (!ws) return fail(PGH_INVALID_ARGUMENT, "null workspace");
    if (n < 2u) {
        ws->sized = false;
        return fail(PGH_INVALID_ARGUMENT, "At least two points are needed for fast_non_dominated_sorting: "
                                              + std::to_string(n) + " detected.");
    }
    if (!ws->own) {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(PGH_NO_DEVICE, "no HIP device is available");
        PGH_HIP(hipStreamCreateWithFlags(&ws->own, hipStreamNonBlocking));
    }
    const uint64_t bytes = n * m * sizeof(double);
    PGH_TRY(ws->staged.reserve(std::max<uint64_t>(bytes, 8u)));
    if (bytes) PGH_HIP(hipMemcpyAsync(ws->staged.ptr, objs, bytes, hipMemcpyHostToDevice, ws->own));
    return pgh_fnds_run(ws, static_cast<const double *>(ws->staged.ptr), n, m, ws->own);
}

pgh_status pgh_fnds_run(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m, void *stream)
{
    if (!ws) return fail(PGH_INVALID_ARGUMENT, "null workspace");
    ws->sized = false;
    if (n < 2u) {
        return fail(PGH_INVALID_ARGUMENT, "At least two points are needed for fast_non_dominated_sorting: "
                                              + std::to_string(n) + " detected.");
    }
    if (n >= (uint64_t(1) << 31)) return fail(PGH_OVERFLOW, "too many points for fast_non_dominated_sorting");
    if (m && !objs) return fail(PGH_INVALID_ARGUMENT, "null objective matrix");
    if (m > 512u) return fail(PGH_UNSUPPORTED, "at most 512 objectives are supported");
    hipStream_t s = static_cast<hipStream_t>(stream);
    ws->stream = s;
    ws->n = n;
    ws->m = m;
    const uint64_t words = (n + 63u) / 64u;
    ws->words = words;
    PGH_TRY(ws->keys.reserve(std::max<uint64_t>(n * m, 1u) * 8u));
    PGH_TRY(ws->D.reserve(n * words * 8u));
    PGH_TRY(ws->cnt32.reserve(n * 4u));
    PGH_TRY(ws->ndom32.reserve(n * 4u));
    PGH_TRY(ws->key32.reserve(n * 4u));
    PGH_TRY(ws->dom_count.reserve(n * 8u));
    PGH_TRY(ws->rank.reserve(n * 8u));
    PGH_TRY(ws->front_items.reserve(n * 8u));
    PGH_TRY(ws->front_offsets.reserve((n + 1u) * 8u));
    PGH_TRY(ws->dom_offsets.reserve((n + 1u) * 8u));
    PGH_TRY(ws->scalars.reserve(16u));
    auto *keys = static_cast<int64_t *>(ws->keys.ptr);
    auto *D = static_cast<uint64_t *>(ws->D.ptr);
    auto *cnt = static_cast<uint32_t *>(ws->cnt32.ptr);
    auto *ndom = static_cast<uint32_t *>(ws->ndom32.ptr);
    auto *key = static_cast<uint32_t *>(ws->key32.ptr);
    const uint32_t n32 = static_cast<uint32_t>(n), m32 = static_cast<uint32_t>(m);
    if (m) {
        k_keys<<<grid_for(n * m), kBlock, 0, s>>>(objs, n * m, keys);
        PGH_HIP(hipGetLastError());
    }
    k_fill_u32<<<grid_for(n), kBlock, 0, s>>>(cnt, n, 0u);
    k_fill_u32<<<grid_for(n), kBlock, 0, s>>>(ndom, n, 0u);
    k_fill_u32<<<grid_for(n), kBlock, 0, s>>>(key, n, 0u);
    PGH_HIP(hipGetLastError());
    // columns per workgroup: a multiple of 64 whose keys fit 32 KiB of LDS
    uint32_t jc = 1024u;
    while (jc > 64u && static_cast<uint64_t>(jc) * std::max<uint64_t>(m, 1u) * 8u > 32768u) jc >>= 1;
    const size_t lds = static_cast<size_t>(jc) * std::max<uint64_t>(m, 1u) * 8u;
    const dim3 grid(grid_for(n), static_cast<uint32_t>((n + jc - 1u) / jc));
    const uint32_t w32 = static_cast<uint32_t>(words);
    switch (m) {
        case 1:
            k_pairs<1><<<grid, kBlock, lds, s>>>(keys, n32, m32, jc, D, w32, cnt, ndom);
            break;
        case 2:
            k_pairs<2><<<grid, kBlock, lds, s>>>(keys, n32, m32, jc, D, w32, cnt, ndom);
            break;
        case 3:
            k_pairs<3><<<grid, kBlock, lds, s>>>(keys, n32, m32, jc, D, w32, cnt, ndom);
            break;
        case 4:
            k_pairs<4><<<grid, kBlock, lds, s>>>(keys, n32, m32, jc, D, w32, cnt, ndom);
            break;
        default:
            k_pairs<0><<<grid, kBlock, lds, s>>>(keys, n32, m32, jc, D, w32, cnt, ndom);
    }
    PGH_HIP(hipGetLastError());
    auto *dom_offsets = static_cast<uint64_t *>(ws->dom_offsets.ptr);
    auto *scalars = static_cast<uint64_t *>(ws->scalars.ptr);
    k_scan<<<1, kPeelBlock, 0, s>>>(ndom, cnt, n32, dom_offsets, static_cast<uint64_t *>(ws->dom_count.ptr), scalars);
    PGH_HIP(hipGetLastError());
    const size_t peel_lds = 8u * static_cast<size_t>(n);
    if (n < 65536u && peel_lds <= max_dynamic_lds()) {
        k_peel_lds<<<1, kPeelBlock, peel_lds, s>>>(D, n32, w32, cnt, static_cast<uint64_t *>(ws->front_items.ptr),
                                                   static_cast<uint64_t *>(ws->front_offsets.ptr),
                                                   static_cast<uint64_t *>(ws->rank.ptr), scalars);
    } else {
        k_peel<<<1, kPeelBlock, 0, s>>>(D, n32, w32, cnt, key, static_cast<uint64_t *>(ws->front_items.ptr),
                                        static_cast<uint64_t *>(ws->front_offsets.ptr),
                                        static_cast<uint64_t *>(ws->rank.ptr), scalars);
    }
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

pgh_status pgh_fnds_sizes(pgh_fnds_ws *ws, uint64_t *n_fronts, uint64_t *n_dom_items)
{
    if (!ws || !ws->n) return fail(PGH_INVALID_ARGUMENT, "fnds workspace has no result");
    if (!ws->sized) {
        PGH_HIP(hipMemcpyAsync(ws->h_sizes, ws->scalars.ptr, 16u, hipMemcpyDeviceToHost, ws->stream));
        PGH_HIP(hipStreamSynchronize(ws->stream));
        // dom_items is written once its size is known
        const uint64_t total = ws->h_sizes[1];
        PGH_TRY(ws->dom_items.reserve(std::max<uint64_t>(total, 1u) * 8u));
        k_emit<<<static_cast<uint32_t>((ws->n + 3u) / 4u), kBlock, 0, ws->stream>>>(
            static_cast<const uint64_t *>(ws->D.ptr), static_cast<uint32_t>(ws->n), static_cast<uint32_t>(ws->words),
            static_cast<const uint64_t *>(ws->dom_offsets.ptr), static_cast<uint64_t *>(ws->dom_items.ptr));
        PGH_HIP(hipGetLastError());
        ws->sized = true;
    }
    if (n_fronts) *n_fronts = ws->h_sizes[0];
    if (n_dom_items) *n_dom_items = ws->h_sizes[1];
    return PGH_OK;
}

pgh_status pgh_fnds_copy(pgh_fnds_ws *ws, uint64_t *dom_count, uint64_t *rank, uint64_t *front_offsets,
                         uint64_t *front_items, uint64_t *dom_offsets, uint64_t *dom_items, int to_host)
{
    uint64_t nfr = 0, ndi = 0;
    PGH_TRY(pgh_fnds_sizes(ws, &nfr, &ndi));
    const hipMemcpyKind kind = to_host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    const uint64_t n = ws->n;
    struct item {
        void *dst;
        const void *src;
        uint64_t count;
    } items[] = {
        {dom_count, ws->dom_count.ptr, n},       {rank, ws->rank.ptr, n},
        {front_offsets, ws->front_offsets.ptr, nfr + 1u}, {front_items, ws->front_items.ptr, n},
        {dom_offsets, ws->dom_offsets.ptr, n + 1u},       {dom_items, ws->dom_items.ptr, ndi},
    };
    for (const auto &it : items) {
        if (it.dst && it.count) {
            PGH_HIP(hipMemcpyAsync(it.dst, it.src, it.count * 8u, kind, ws->stream));
        }
    }
    if (to_host) PGH_HIP(hipStreamSynchronize(ws->stream));
    return PGH_OK;
}

} // extern "C"
