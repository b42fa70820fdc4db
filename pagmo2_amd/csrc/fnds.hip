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
//   k_front0 +  the front peel (multi_objective.cpp:228-254), n <= 32768:
//   k_peel_cols F0 ascending; then per level, each workgroup decrements the
//               counters of its own columns (in LDS) for the rows of the
//               current front and records for each element the position of
//               its last dominator; the elements whose counter reached zero are
//               exchanged between workgroups and ordered by (that position,
//               index) -- the order in which the reference's nested loop
//               appends them -- in every workgroup.
//   k_peel      the same peel in one workgroup for larger n.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "pgh_internal.hpp"

#include "fnds_ws.hpp"

namespace pgh
{
namespace
{

constexpr uint32_t kBlock = 256;
constexpr uint32_t kPeelBlock = 1024;


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


// ---- column-partitioned peel (n <= 32768) -----------------------------------
// Workgroup b owns the columns (dominated elements) of D's words
// [b wpw, (b + 1) wpw): their counters and keys live in its LDS and only it
// changes them.  Every workgroup holds the current front in LDS (slot = the
// member's position in the front, low 16 bits = its index).  Level L, in
// every workgroup:
//   1. for each (member at position i, own word w), every set bit c of
//      D[member][w] raises key[c] to ((L+1) << 16 | i) and decrements cnt[c];
//      the decrement that reaches zero appends c to the workgroup's list;
//   2. wavefront 0 writes the list as sort keys (position of the element's
//      last dominator in front L << 16 | index) into the level's half of comp
//      at the workgroup's first column, writes the elements' ranks, and
//      publishes its flag ((L+1) << 16 | list length);
//   3. thread g < G waits for workgroup g's flag of level L; the lists are
//      gathered into LDS and sorted by key -- the order in which the
//      reference's nested loop appends them (multi_objective.cpp:236-251) --
//      which is front L+1.  Workgroup 0 writes it out.
// So a level costs one pass of D loads and one exchange of flags and lists.
// Every workgroup waits for every other one, so all must be resident at once:
// the kernel is launched cooperatively.  Each half of comp and flag is written
// again two levels later, only after every workgroup has published the level
// in between, which each does after gathering this one.  Every wait is
// bounded: a spin limit sets a flag that the host reports, and workgroups
// that see it stop waiting.
constexpr uint32_t kGridBlock = 1024;
constexpr uint32_t kGridMaxN = 32768;
// Columns per workgroup: at most 64 kColWords (counters, keys and list in LDS
// beside the front).  Workgroups: at most kColMaxG (one flag poller each).
constexpr uint32_t kColWords = 16;
constexpr uint32_t kColMaxG = 256;
// The largest count pgh_fnds_set_peel_blocks accepts; the kernel clamps any
// count to the range it supports (the fronts do not depend on it).
constexpr uint32_t kMaxPeelBlocks = 4096;
constexpr uint32_t kSpinLimit = 1u << 22; // polls of about 1 us each: a few seconds
// Fronts of up to kColSortMax elements are ordered by counting, larger ones by
// a bitonic sort.
constexpr uint32_t kColSortMax = 1024;
constexpr size_t kColMaxLds = (static_cast<size_t>(kGridMaxN) + 3u * 64u * kColWords) * 4u;

__device__ __forceinline__ uint64_t ld_relaxed64(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// F0 in ascending order (multi_objective.cpp:228-233).
__global__ __launch_bounds__(kPeelBlock) void k_front0(const uint32_t *__restrict__ dom_count, uint32_t n,
                                                       uint64_t *__restrict__ front_items,
                                                       uint64_t *__restrict__ front_offsets,
                                                       uint64_t *__restrict__ rank)
{
    __shared__ uint64_t sh[kPeelBlock / 64u];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + kPeelBlock - 1u) / kPeelBlock;
    const uint32_t b = min(t * per, n), e = min(b + per, n);
    uint64_t c = 0;
    for (uint32_t i = b; i < e; ++i) c += dom_count[i] == 0u;
    uint64_t total;
    uint64_t o = block_exclusive_scan(c, sh, total);
    for (uint32_t i = b; i < e; ++i) {
        if (dom_count[i] == 0u) {
            front_items[o++] = i;
            rank[i] = 0;
        }
    }
    if (t == 0) {
        front_offsets[0] = 0;
        front_offsets[1] = total;
    }
}

// Bitonic sort of sh[0..P) (P a power of two) by one workgroup.
__device__ __forceinline__ void block_bitonic_sort(uint32_t *sh, uint32_t P)
{
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint32_t a = sh[i], b = sh[l];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        sh[i] = b;
                        sh[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Dynamic LDS: F (P words: the front, then the next front's keys), then cnt,
// key and list (64 wpw words each).  The grid is ceil(words / wpw)
// workgroups, at most kColMaxG, all resident; P >= n is a power of two.
__global__ __launch_bounds__(kGridBlock) void k_peel_cols(const uint64_t *__restrict__ D, uint32_t n, uint32_t words,
                                                          uint32_t wpw, uint32_t P, const uint32_t *__restrict__ cnt0,
                                                          uint32_t *__restrict__ comp, uint32_t *__restrict__ flag,
                                                          uint64_t *__restrict__ front_items,
                                                          uint64_t *__restrict__ front_offsets,
                                                          uint64_t *__restrict__ rank, uint64_t *__restrict__ scalars)
{
    extern __shared__ __align__(16) uint32_t lds[];
    uint32_t *F = lds;
    uint32_t *cnt = F + P;
    uint32_t *key = cnt + 64u * wpw;
    uint32_t *list = key + 64u * wpw;
    __shared__ uint32_t s_off[kColMaxG + 1u];
    __shared__ uint32_t s_rank[kColSortMax]; // the counting sort's ranks
    __shared__ uint32_t s_len, s_abort;
    const uint32_t G = gridDim.x, b = blockIdx.x, t = threadIdx.x;
    const uint32_t w0 = b * wpw;
    const uint32_t wn = min(wpw, words - w0); // >= 1 by the grid size
    const uint32_t c0 = 64u * w0;
    const uint32_t cn = min(64u * wn, n - c0);
    for (uint32_t c = t; c < 64u * wpw; c += kGridBlock) {
        cnt[c] = c < cn ? cnt0[c0 + c] : 0u;
        key[c] = 0u;
    }
    uint32_t s = min(static_cast<uint32_t>(front_offsets[1]), n); // front 0 (k_front0)
    for (uint32_t i = t; i < s; i += kGridBlock) F[i] = static_cast<uint32_t>(front_items[i]);
    uint32_t e = s; // end of the current front in front_items
    if (t == 0) s_abort = 0u;
    for (uint32_t L = 0;; ++L) {
        const uint32_t q = L & 1u;
        const uint32_t tag = (L + 1u) << 16; // L + 1 <= n <= 2^15
        if (t == 0) s_len = 0u;
        __syncthreads();
        // 1. the decrements by front L, in 64-, 32- or 16-bit pieces of the
        // own words so that small fronts spread over the workgroup
        const uint32_t items = s * wn;
        const uint32_t ps = items * 4u <= kGridBlock ? 2u : (items * 2u <= kGridBlock ? 1u : 0u);
        const uint32_t pbits = 64u >> ps;
        const uint64_t pmask = ps ? (uint64_t(1) << pbits) - 1u : ~uint64_t(0);
        for (uint32_t it = t; it < (items << ps); it += kGridBlock) {
            const uint32_t wi = it >> ps, piece = it & ((1u << ps) - 1u);
            const uint32_t i = wi / wn, w = wi - i * wn;
            const uint64_t word = D[static_cast<uint64_t>(F[i] & 0xffffu) * words + w0 + w];
            uint64_t bits = (word >> (piece * pbits)) & pmask;
            const uint32_t base = w * 64u + piece * pbits;
            while (bits) {
                const uint32_t c = base + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(bits)) - 1);
                bits &= bits - 1u;
                atomicMax(&key[c], tag | i);
                if (atomicSub(&cnt[c], 1u) == 1u) list[atomicAdd(&s_len, 1u)] = c;
            }
        }
        __syncthreads();
        // 2. publish (wavefront 0; its release covers its own stores) and the
        // ranks (wavefront 1, read only after the kernel)
        if (t < 64u) {
            const uint32_t len = s_len;
            uint32_t *dst = comp + static_cast<uint64_t>(q) * n + c0;
            for (uint32_t k = t; k < len; k += 64u) {
                const uint32_t c = list[k];
                st_relaxed(&dst[k], ((key[c] & 0xffffu) << 16) | (c0 + c));
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (t == 0) st_relaxed(&flag[q * G + b], tag | len);
            // 3. every workgroup's flag of level L: lane t waits for those of
            // workgroups 4t .. 4t + 3 (G <= 256), then the exclusive offsets
            // of the lists in workgroup order
            uint32_t c4[4] = {0u, 0u, 0u, 0u}, pending = 0u, spins = 0u;
#pragma unroll
            for (uint32_t r = 0; r < 4u; ++r) pending |= (4u * t + r < G ? 1u : 0u) << r;
            while (pending) {
                uint32_t v[4];
#pragma unroll
                for (uint32_t r = 0; r < 4u; ++r) v[r] = (pending >> r) & 1u ? ld_relaxed(&flag[q * G + 4u * t + r]) : 0u;
#pragma unroll
                for (uint32_t r = 0; r < 4u; ++r) {
                    if (((pending >> r) & 1u) && (v[r] & 0xffff0000u) == tag) {
                        c4[r] = v[r] & 0xffffu;
                        pending &= ~(1u << r);
                    }
                }
                if (!pending) break;
                // another workgroup's time-out, checked every 256 polls
                if ((++spins & 255u) == 0u && (ld_relaxed64(&scalars[2]) != 0u || spins >= kSpinLimit)) {
                    atomicOr(reinterpret_cast<unsigned long long *>(&scalars[2]), 1ull);
                    s_abort = 1u;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const uint32_t sum = c4[0] + c4[1] + c4[2] + c4[3];
            uint32_t x = sum;
#pragma unroll
            for (uint32_t d = 1; d < 64u; d <<= 1) {
                const uint32_t y = __shfl_up(x, d);
                if (t >= d) x += y;
            }
            uint32_t o = x - sum;
#pragma unroll
            for (uint32_t r = 0; r < 4u; ++r) {
                const uint32_t g = 4u * t + r;
                if (g < G) s_off[g] = o;
                o += c4[r];
            }
            if (t == 63u) s_off[G] = x;
        } else if (t < 128u) {
            const uint32_t len = s_len;
            for (uint32_t k = t - 64u; k < len; k += 64u) rank[c0 + list[k]] = L + 1u;
        }
        __syncthreads();
        if (s_abort) return;
        const uint32_t m = min(s_off[G], n - e);
        if (m == 0u) {
            if (b == 0u && t == 0u) scalars[0] = L + 1u; // number of fronts
            return;
        }
        // gather (front L is no longer read): entry k is in the list of the
        // workgroup g with s_off[g] <= k < s_off[g + 1]
        for (uint32_t k = t; k < m; k += kGridBlock) {
            uint32_t lo = 0u, hi = G;
            while (hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_off[mid] <= k) {
                    lo = mid;
                } else {
                    hi = mid;
                }
            }
            F[k] = ld_relaxed(comp + static_cast<uint64_t>(q) * n + 64u * lo * wpw + (k - s_off[lo]));
            if (k < kColSortMax) s_rank[k] = 0u;
        }
        __syncthreads();
        // sort by key (distinct).  Up to kColSortMax keys: the rank of each
        // by counting, wavefront (chunk, part) counting the keys of part
        // `part` of F below each key of chunk `chunk` (a broadcast LDS read
        // per key), the parts' counts added in s_rank; beyond, bitonic.
        if (m <= kColSortMax) {
            const uint32_t kw = (m + 63u) / 64u; // chunks of 64 keys, <= 16
            const uint32_t parts = (kGridBlock / 64u) / kw;
            const uint32_t wave = t / 64u;
            if (wave < kw * parts) {
                const uint32_t chunk = wave % kw, part = wave / kw;
                const uint32_t k = 64u * chunk + (t % 64u);
                const uint32_t x = k < m ? F[k] : 0xffffffffu;
                const uint32_t i0 = part * m / parts, i1 = (part + 1u) * m / parts;
                uint32_t r = 0u;
#pragma unroll 8
                for (uint32_t i = i0; i < i1; ++i) r += F[i] < x ? 1u : 0u;
                if (k < m && r) atomicAdd(&s_rank[k], r);
            }
            __syncthreads();
            const uint32_t x = t < m ? F[t] : 0u, r = t < m ? s_rank[t] : 0u;
            __syncthreads();
            if (t < m) F[r] = x;
            __syncthreads();
        } else {
            uint32_t Pm = 1u;
            while (Pm < m) Pm <<= 1;
            for (uint32_t i = m + t; i < Pm; i += kGridBlock) F[i] = 0xffffffffu;
            __syncthreads();
            block_bitonic_sort(F, Pm);
        }
        if (b == 0u) {
            for (uint32_t k = t; k < m; k += kGridBlock) front_items[e + k] = F[k] & 0xffffu;
            if (t == 0u) front_offsets[L + 2u] = e + m;
        }
        e += m;
        s = m;
    }
}

uint32_t grid_for(uint64_t count)
{
    return static_cast<uint32_t>((count + kBlock - 1u) / kBlock);
}

uint32_t device_cus()
{
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256u;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 256u;
    return static_cast<uint32_t>(std::max(v, 1));
}

// The peel's workgroup range for n points of w32 words: at least enough that
// a workgroup owns at most kColWords words, at most one per word, per CU and
// kColMaxG.
void peel_range(uint32_t w32, uint32_t cus, uint32_t &gmin, uint32_t &gmax)
{
    gmin = (w32 + kColWords - 1u) / kColWords;
    gmax = std::min({w32, cus, kColMaxG});
}

// Default: one workgroup per four words (256 points), the fastest of the
// tuner's candidates at 8192 x 2 and 16384 x 3.
uint32_t default_peel_blocks(uint32_t w32)
{
    return (w32 + 3u) / 4u;
}

pgh_status launch_grid_peel(pgh_fnds_ws *ws, hipStream_t s, uint32_t n32, uint32_t w32)
{
    static thread_local int dev_cached = -1;
    static thread_local uint32_t cus = 0;
    int dev = 0;
    PGH_HIP(hipGetDevice(&dev));
    if (dev != dev_cached) {
        int v = 0;
        PGH_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
        PGH_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_peel_cols),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kColMaxLds)));
        cus = static_cast<uint32_t>(std::max(v, 1));
        dev_cached = dev;
    }
    uint32_t gmin = 0, gmax = 0;
    peel_range(w32, cus, gmin, gmax);
    if (gmin > gmax) return fail(PGH_UNSUPPORTED, "fast_non_dominated_sorting: too few CUs for the grid peel");
    const uint32_t want = ws->peel_blocks ? ws->peel_blocks : default_peel_blocks(w32);
    const uint32_t g = std::min(std::max(want, gmin), gmax);
    const uint32_t wpw = (w32 + g - 1u) / g; // <= kColWords, since g >= gmin
    const uint32_t G = (w32 + wpw - 1u) / wpw; // <= g, and no workgroup without a word
    uint32_t P = 1;
    while (P < n32) P <<= 1;
    const size_t lds = (static_cast<size_t>(P) + 3u * 64u * wpw) * 4u;
    int per_cu = 0;
    PGH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&k_peel_cols),
                                                         static_cast<int>(kGridBlock), lds));
    if (per_cu < 1 || static_cast<uint64_t>(per_cu) * cus < G) {
        return fail(PGH_DEVICE_ERROR, "fast_non_dominated_sorting: the grid peel does not fit the device");
    }
    const uint64_t *D = static_cast<const uint64_t *>(ws->D.ptr);
    const uint32_t *cnt = static_cast<const uint32_t *>(ws->cnt32.ptr);
    auto *comp = static_cast<uint32_t *>(ws->comp.ptr);
    auto *flag = static_cast<uint32_t *>(ws->ctl.ptr);
    auto *fi = static_cast<uint64_t *>(ws->front_items.ptr);
    auto *fo = static_cast<uint64_t *>(ws->front_offsets.ptr);
    auto *rk = static_cast<uint64_t *>(ws->rank.ptr);
    auto *sc = static_cast<uint64_t *>(ws->scalars.ptr);
    k_fill_u32<<<grid_for(2u * G), kBlock, 0, s>>>(flag, 2u * G, 0u);
    k_front0<<<1, kPeelBlock, 0, s>>>(cnt, n32, fi, fo, rk);
    PGH_HIP(hipGetLastError());
    uint32_t n_a = n32, w_a = w32, wpw_a = wpw, P_a = P;
    void *args[] = {&D, &n_a, &w_a, &wpw_a, &P_a, &cnt, &comp, &flag, &fi, &fo, &rk, &sc};
    const hipError_t le = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_peel_cols), dim3(G),
                                                     dim3(kGridBlock), args, static_cast<unsigned int>(lds), s);
    if (le != hipSuccess) {
        // the runtime refused a co-resident grid: the single-workgroup peel
        // gives the same fronts (it restarts from the dominator counts)
        (void)hipGetLastError();
        k_peel<<<1, kPeelBlock, 0, s>>>(D, n32, w32, static_cast<uint32_t *>(ws->cnt32.ptr),
                                        static_cast<uint32_t *>(ws->key32.ptr), fi, fo, rk, sc);
        PGH_HIP(hipGetLastError());
    }
    return PGH_OK;
}

} // namespace

void fnds_free(pgh_fnds_ws *ws)
{
    if (!ws) return;
    if (ws->stream) (void)hipStreamSynchronize(ws->stream);
    if (ws->own) (void)hipStreamDestroy(ws->own);
    delete ws;
}

pgh_status thread_fnds_ws(pgh_fnds_ws **ws)
{
    thread_ctx *ctx = nullptr;
    PGH_TRY(get_thread_ctx(&ctx));
    if (!ctx->fnds) {
        ctx->fnds = new pgh_fnds_ws();
        ctx->fnds->thread_owned = true;
    }
    *ws = ctx->fnds;
    return PGH_OK;
}

} // namespace pgh

using namespace pgh;

extern "C" {

pgh_status pgh_fnds_create(pgh_fnds_ws **ws)
try {
    if (!ws) return fail(PGH_INVALID_ARGUMENT, "null workspace pointer");
    *ws = new pgh_fnds_ws();
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_fnds_thread_workspace(pgh_fnds_ws **ws)
try {
    if (!ws) return fail(PGH_INVALID_ARGUMENT, "null workspace pointer");
    return thread_fnds_ws(ws);
} catch (...) {
    return api_exception();
}

pgh_status pgh_fnds_destroy(pgh_fnds_ws *ws)
try {
    if (ws && ws->thread_owned) {
        return fail(PGH_INVALID_ARGUMENT, "a thread's own fnds workspace is released with its thread");
    }
    fnds_free(ws);
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_fnds_run_host(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m)
try {
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
} catch (...) {
    return api_exception();
}

pgh_status pgh_fnds_run(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m, void *stream)
try {
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
    PGH_TRY(ws->scalars.reserve(32u));
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
    k_fill_u32<<<1, kBlock, 0, s>>>(static_cast<uint32_t *>(ws->scalars.ptr) + 4, 2u, 0u); // timeout flag
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
    if (n <= kGridMaxN) {
        PGH_TRY(ws->comp.reserve(n * 8u));
        PGH_TRY(ws->ctl.reserve(2u * kColMaxG * 4u));
        PGH_TRY(launch_grid_peel(ws, s, n32, w32));
    } else {
        k_peel<<<1, kPeelBlock, 0, s>>>(D, n32, w32, cnt, key, static_cast<uint64_t *>(ws->front_items.ptr),
                                        static_cast<uint64_t *>(ws->front_offsets.ptr),
                                        static_cast<uint64_t *>(ws->rank.ptr), scalars);
    }
    PGH_HIP(hipGetLastError());
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_fnds_set_peel_blocks(pgh_fnds_ws *ws, uint32_t blocks)
try {
    if (!ws) return fail(PGH_INVALID_ARGUMENT, "null workspace");
    if (blocks > kMaxPeelBlocks) {
        return fail(PGH_INVALID_ARGUMENT, "at most " + std::to_string(kMaxPeelBlocks) + " peel workgroups, got "
                                              + std::to_string(blocks));
    }
    ws->peel_blocks = blocks;
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_fnds_tune(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m, uint32_t reps, void *stream,
                         uint32_t *chosen)
try {
    if (!ws) return fail(PGH_INVALID_ARGUMENT, "null workspace");
    if (reps == 0u) return fail(PGH_INVALID_ARGUMENT, "the tuner needs at least one timed repetition");
    if (chosen) *chosen = 0u;
    if (n > kGridMaxN) { // the single-workgroup peel: nothing to tune
        ws->peel_blocks = 0u;
        return pgh_fnds_run(ws, objs, n, m, stream);
    }
    const uint32_t n32 = static_cast<uint32_t>(n);
    const uint32_t w32 = (n32 + 63u) / 64u;
    uint32_t gmin = 0, gmax = 0;
    peel_range(w32, device_cus(), gmin, gmax);
    std::vector<uint32_t> cands;
    for (uint32_t per : {64u, 128u, 256u, 512u, 1024u}) {
        const uint32_t b = std::min(gmax, std::max(gmin, (n32 + per - 1u) / per));
        if (std::find(cands.begin(), cands.end(), b) == cands.end()) cands.push_back(b);
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PGH_HIP(hipEventCreate(&e0));
    const hipError_t ce = hipEventCreate(&e1);
    if (ce != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(ce, "hipEventCreate");
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    float best = 0.f;
    uint32_t best_b = 0u;
    pgh_status st = PGH_OK;
    for (uint32_t b : cands) {
        ws->peel_blocks = b;
        st = pgh_fnds_run(ws, objs, n, m, stream); // warm-up (and argument checks)
        if (st != PGH_OK) break;
        if ((st = hip_status(hipEventRecord(e0, s), "hipEventRecord")) != PGH_OK) break;
        for (uint32_t r = 0; r < reps && st == PGH_OK; ++r) st = pgh_fnds_run(ws, objs, n, m, stream);
        if (st != PGH_OK) break;
        if ((st = hip_status(hipEventRecord(e1, s), "hipEventRecord")) != PGH_OK) break;
        if ((st = hip_status(hipEventSynchronize(e1), "hipEventSynchronize")) != PGH_OK) break;
        // a peel that ran out of its spin limit is an error, not a timing
        uint64_t flag[3] = {0, 0, 0};
        if ((st = hip_status(hipMemcpy(flag, ws->scalars.ptr, 24u, hipMemcpyDeviceToHost), "hipMemcpy")) != PGH_OK) {
            break;
        }
        if (flag[2]) {
            st = fail(PGH_DEVICE_ERROR, "fast_non_dominated_sorting: a grid barrier timed out");
            break;
        }
        float ms = 0.f;
        if ((st = hip_status(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) != PGH_OK) break;
        if (best_b == 0u || ms < best) {
            best = ms;
            best_b = b;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != PGH_OK) {
        ws->peel_blocks = 0u;
        return st;
    }
    ws->peel_blocks = best_b;
    if (chosen) *chosen = best_b;
    // leave the workspace holding a result computed with the chosen grid
    return pgh_fnds_run(ws, objs, n, m, stream);
} catch (...) {
    return api_exception();
}

pgh_status pgh_fnds_sizes(pgh_fnds_ws *ws, uint64_t *n_fronts, uint64_t *n_dom_items)
try {
    if (!ws || !ws->n) return fail(PGH_INVALID_ARGUMENT, "fnds workspace has no result");
    if (!ws->sized) {
        PGH_HIP(hipMemcpyAsync(ws->h_sizes, ws->scalars.ptr, 24u, hipMemcpyDeviceToHost, ws->stream));
        PGH_HIP(hipStreamSynchronize(ws->stream));
        if (ws->h_sizes[2]) return fail(PGH_DEVICE_ERROR, "fast_non_dominated_sorting: a grid barrier timed out");
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
} catch (...) {
    return api_exception();
}

pgh_status pgh_fnds_copy(pgh_fnds_ws *ws, uint64_t *dom_count, uint64_t *rank, uint64_t *front_offsets,
                         uint64_t *front_items, uint64_t *dom_offsets, uint64_t *dom_items, int to_host)
try {
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
} catch (...) {
    return api_exception();
}

} // extern "C"
