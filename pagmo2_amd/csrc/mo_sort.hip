// mo_sort.hip -- crowding_distance, sort_population_mo and select_best_N_mo
// (src/utils/multi_objective.cpp:280-465) on the device, after the device
// fast_non_dominated_sorting of fnds.hip.
//
// Crowding distance of every front at once, objective by objective in the
// reference's order i = 0..M-1 (the per-element sums are then the reference's
// bit for bit):
//   1. (key_i, element) pairs in the current order (fronts contiguous: front
//      order for the first objective, the previous objective's order after) are
//      sorted by key_i with a stable radix sort, then by rank with a stable
//      radix sort: each front is contiguous again, ordered by objective i,
//      equal values in their previous order (std::stable_sort of the
//      reference's one index vector);
//   2. one lane per position: the front's ends become +inf, interior elements
//      add (f[next] - f[prev]) / (f[last] - f[first]) (multi_objective.cpp:
//      305-315).
// sort_population_mo then sorts the indices by crowding (descending, pagmo's
// greater_than_f order: NaN first) and by rank with the same stable sorts.
// The radix sorts are rocPRIM's (header-only, stable).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "fnds_ws.hpp"
#include "pgh_internal.hpp"

namespace pgh
{
namespace
{

constexpr uint32_t kBlock = 256;

uint32_t grid_for(uint64_t count)
{
    return static_cast<uint32_t>((count + kBlock - 1u) / kBlock);
}

// Unsigned radix key: ascending in pagmo's less_than_f order (asc) or in its
// greater_than_f order (desc).
__device__ __forceinline__ uint64_t radix_key(double x, bool desc)
{
    const uint64_t k = static_cast<uint64_t>(order_key(x)) ^ (uint64_t(1) << 63);
    return desc ? ~k : k;
}

// Keys of objective i in the current order: front_items (items, the first
// objective), the previous objective's order (prev), or 0..n-1 (neither).
// prev may alias vout (each lane reads and writes its own position).
__global__ __launch_bounds__(kBlock) void k_objective_keys(const double *__restrict__ objs, uint32_t n, uint32_t m,
                                                           uint32_t i, const uint64_t *__restrict__ items,
                                                           const uint32_t *prev, uint64_t *__restrict__ kout,
                                                           uint32_t *vout)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const uint32_t e = prev ? prev[p] : (items ? static_cast<uint32_t>(items[p]) : p);
    kout[p] = radix_key(objs[static_cast<uint64_t>(e) * m + i], false);
    vout[p] = e;
}

__global__ __launch_bounds__(kBlock) void k_rank_keys(const uint32_t *__restrict__ v, const uint64_t *__restrict__ rank,
                                                      uint32_t n, uint32_t *__restrict__ rk)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p < n) rk[p] = static_cast<uint32_t>(rank[v[p]]);
}

// One objective's contribution (multi_objective.cpp:305-315).  order holds the
// elements front by front, each front sorted by objective i; rank and
// front_offsets locate an element's front (NULL: one front of n).
__global__ __launch_bounds__(kBlock) void k_crowd(const double *__restrict__ objs, uint32_t n, uint32_t m, uint32_t i,
                                                  const uint32_t *__restrict__ order,
                                                  const uint64_t *__restrict__ rank,
                                                  const uint64_t *__restrict__ front_offsets,
                                                  double *__restrict__ crowd)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const uint32_t e = order[p];
    uint32_t fb = 0u, fe = n;
    if (rank) {
        const uint64_t r = rank[e];
        fb = static_cast<uint32_t>(front_offsets[r]);
        fe = static_cast<uint32_t>(front_offsets[r + 1u]);
    }
    if (fe - fb < 2u) return; // a front of one keeps 0 (multi_objective.cpp:443-446)
    if (p == fb || p == fe - 1u) {
        crowd[e] = __builtin_inf();
        return;
    }
    const double df = objs[static_cast<uint64_t>(order[fe - 1u]) * m + i] - objs[static_cast<uint64_t>(order[fb]) * m + i];
    crowd[e] += (objs[static_cast<uint64_t>(order[p + 1u]) * m + i] - objs[static_cast<uint64_t>(order[p - 1u]) * m + i])
                / df;
}

__global__ __launch_bounds__(kBlock) void k_fill_f64(double *__restrict__ p, uint32_t n, double v)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n) p[t] = v;
}

// keys: crowding descending (greater_than_f order); values: index, or the
// position j of front_items[base + j] when items is given
__global__ __launch_bounds__(kBlock) void k_crowd_keys(const double *__restrict__ crowd, uint32_t n,
                                                       const uint64_t *__restrict__ items, uint64_t base,
                                                       uint64_t *__restrict__ kout, uint32_t *__restrict__ vout)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const uint32_t e = items ? static_cast<uint32_t>(items[base + p]) : p;
    kout[p] = radix_key(crowd[e], true);
    vout[p] = p;
}

__global__ __launch_bounds__(kBlock) void k_widen(const uint32_t *__restrict__ v, uint32_t n,
                                                  const uint64_t *__restrict__ items, uint64_t base,
                                                  uint64_t *__restrict__ out)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p < n) out[p] = items ? items[base + v[p]] : v[p];
}

uint32_t bits_for(uint64_t count)
{
    uint32_t b = 1u;
    while (b < 32u && (uint64_t(1) << b) < count) ++b;
    return b;
}

struct sorter {
    pgh_fnds_ws *ws;
    hipStream_t s;
    uint32_t n;
    // stable sort of (k0, v0) by the low `bits` of 64-bit keys -> (k1, v1)
    pgh_status by_u64(uint32_t count, uint32_t bits)
    {
        size_t bytes = 0;
        auto *k0 = static_cast<uint64_t *>(ws->sk0.ptr), *k1 = static_cast<uint64_t *>(ws->sk1.ptr);
        auto *v0 = static_cast<uint32_t *>(ws->sv0.ptr), *v1 = static_cast<uint32_t *>(ws->sv1.ptr);
        PGH_HIP(rocprim::radix_sort_pairs(nullptr, bytes, k0, k1, v0, v1, count, 0u, bits, s));
        PGH_TRY(ws->sort_tmp.reserve(std::max<size_t>(bytes, 8u)));
        PGH_HIP(rocprim::radix_sort_pairs(ws->sort_tmp.ptr, bytes, k0, k1, v0, v1, count, 0u, bits, s));
        return PGH_OK;
    }
    // stable sort of (v1) by rank -> v0 (rank keys staged in sk0 as uint32)
    pgh_status by_rank(uint32_t count, uint64_t n_fronts)
    {
        auto *r0 = static_cast<uint32_t *>(ws->sk0.ptr), *r1 = static_cast<uint32_t *>(ws->sk1.ptr);
        auto *v0 = static_cast<uint32_t *>(ws->sv0.ptr), *v1 = static_cast<uint32_t *>(ws->sv1.ptr);
        k_rank_keys<<<grid_for(count), kBlock, 0, s>>>(v1, static_cast<const uint64_t *>(ws->rank.ptr), count, r0);
        PGH_HIP(hipGetLastError());
        size_t bytes = 0;
        const uint32_t bits = bits_for(n_fronts);
        PGH_HIP(rocprim::radix_sort_pairs(nullptr, bytes, r0, r1, v1, v0, count, 0u, bits, s));
        PGH_TRY(ws->sort_tmp.reserve(std::max<size_t>(bytes, 8u)));
        PGH_HIP(rocprim::radix_sort_pairs(ws->sort_tmp.ptr, bytes, r0, r1, v1, v0, count, 0u, bits, s));
        return PGH_OK;
    }
};

pgh_status reserve_sort(pgh_fnds_ws *ws, uint64_t n)
{
    PGH_TRY(ws->sk0.reserve(std::max<uint64_t>(n, 1u) * 8u));
    PGH_TRY(ws->sk1.reserve(std::max<uint64_t>(n, 1u) * 8u));
    PGH_TRY(ws->sv0.reserve(std::max<uint64_t>(n, 1u) * 4u));
    PGH_TRY(ws->sv1.reserve(std::max<uint64_t>(n, 1u) * 4u));
    PGH_TRY(ws->crowd.reserve(std::max<uint64_t>(n, 1u) * 8u));
    return PGH_OK;
}

// Crowding of every front of the fnds result held by ws (with_fronts) or of
// the n points as one front; leaves it in ws->crowd.
pgh_status crowding_all(pgh_fnds_ws *ws, const double *objs, uint32_t n, uint32_t m, bool with_fronts,
                        uint64_t n_fronts, hipStream_t s)
{
    sorter so{ws, s, n};
    auto *crowd = static_cast<double *>(ws->crowd.ptr);
    k_fill_f64<<<grid_for(n), kBlock, 0, s>>>(crowd, n, 0.0);
    PGH_HIP(hipGetLastError());
    const uint64_t *items = with_fronts ? static_cast<const uint64_t *>(ws->front_items.ptr) : nullptr;
    const uint64_t *rank = with_fronts ? static_cast<const uint64_t *>(ws->rank.ptr) : nullptr;
    const uint64_t *fo = with_fronts ? static_cast<const uint64_t *>(ws->front_offsets.ptr) : nullptr;
    // As in the reference, which sorts one index vector objective after
    // objective, each objective's sort starts from the previous objective's
    // order, so equal values keep that order.
    const uint32_t *order = nullptr;
    for (uint32_t i = 0; i < m; ++i) {
        k_objective_keys<<<grid_for(n), kBlock, 0, s>>>(objs, n, m, i, items, order,
                                                        static_cast<uint64_t *>(ws->sk0.ptr),
                                                        static_cast<uint32_t *>(ws->sv0.ptr));
        PGH_HIP(hipGetLastError());
        PGH_TRY(so.by_u64(n, 64u)); // -> sv1
        order = static_cast<const uint32_t *>(ws->sv1.ptr);
        if (with_fronts && n_fronts > 1u) {
            PGH_TRY(so.by_rank(n, n_fronts)); // -> sv0
            order = static_cast<const uint32_t *>(ws->sv0.ptr);
        }
        k_crowd<<<grid_for(n), kBlock, 0, s>>>(objs, n, m, i, order, rank, fo, crowd);
        PGH_HIP(hipGetLastError());
    }
    return PGH_OK;
}

std::string crowding_points_message(uint64_t n)
{
    return "A non dominated front must contain at least two points: " + std::to_string(n) + " detected.";
}
std::string crowding_objectives_message(uint64_t m)
{
    return "Points in the non dominated front must contain at least two objectives: " + std::to_string(m)
           + " detected.";
}

// Runs fnds on ws and reads the front offsets back (synchronises).
pgh_status fnds_fronts(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m, hipStream_t s,
                       std::vector<uint64_t> &offsets)
{
    PGH_TRY(pgh_fnds_run(ws, objs, n, m, s));
    uint64_t nf = 0;
    PGH_TRY(pgh_fnds_sizes(ws, &nf, nullptr));
    offsets.resize(nf + 1u);
    PGH_HIP(hipMemcpyAsync(offsets.data(), ws->front_offsets.ptr, (nf + 1u) * 8u, hipMemcpyDeviceToHost, s));
    PGH_HIP(hipStreamSynchronize(s));
    return PGH_OK;
}

// Host-buffer forms: the calling thread's own workspace (kept between calls,
// so a generation loop allocates once), its stream and a staged copy.
struct host_call {
    pgh_fnds_ws *ws = nullptr;
    pgh_status begin(const double *objs, uint64_t n, uint64_t m, const double **dev)
    {
        PGH_TRY(thread_fnds_ws(&ws));
        if (!ws->own) PGH_HIP(hipStreamCreateWithFlags(&ws->own, hipStreamNonBlocking));
        const uint64_t bytes = n * m * sizeof(double);
        PGH_TRY(ws->staged.reserve(std::max<uint64_t>(bytes, 8u)));
        if (bytes) PGH_HIP(hipMemcpyAsync(ws->staged.ptr, objs, bytes, hipMemcpyHostToDevice, ws->own));
        *dev = static_cast<const double *>(ws->staged.ptr);
        return PGH_OK;
    }
};

} // namespace
} // namespace pgh

using namespace pgh;

extern "C" {

pgh_status pgh_crowding_distance(pgh_fnds_ws *ws, const double *points, uint64_t n, uint64_t m, double *out,
                                 void *stream)
try {
    if (!ws) return fail(PGH_INVALID_ARGUMENT, "null workspace");
    if (n < 2u) return fail(PGH_INVALID_ARGUMENT, crowding_points_message(n));
    if (m < 2u) return fail(PGH_INVALID_ARGUMENT, crowding_objectives_message(m));
    if (!points || !out) return fail(PGH_INVALID_ARGUMENT, "null buffer");
    if (n >= (uint64_t(1) << 31) || m > 512u) return fail(PGH_UNSUPPORTED, "crowding_distance: front too large");
    hipStream_t s = static_cast<hipStream_t>(stream);
    PGH_TRY(reserve_sort(ws, n));
    PGH_TRY(crowding_all(ws, points, static_cast<uint32_t>(n), static_cast<uint32_t>(m), false, 1u, s));
    PGH_HIP(hipMemcpyAsync(out, ws->crowd.ptr, n * 8u, hipMemcpyDeviceToDevice, s));
    PGH_HIP(hipStreamSynchronize(s));
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_crowding_distance_host(const double *points, uint64_t n, uint64_t m, double *out)
try {
    if (n < 2u) return fail(PGH_INVALID_ARGUMENT, crowding_points_message(n));
    if (m < 2u) return fail(PGH_INVALID_ARGUMENT, crowding_objectives_message(m));
    host_call h;
    const double *dev = nullptr;
    PGH_TRY(h.begin(points, n, m, &dev));
    PGH_TRY(reserve_sort(h.ws, n));
    PGH_TRY(crowding_all(h.ws, dev, static_cast<uint32_t>(n), static_cast<uint32_t>(m), false, 1u, h.ws->own));
    PGH_HIP(hipMemcpyAsync(out, h.ws->crowd.ptr, n * 8u, hipMemcpyDeviceToHost, h.ws->own));
    PGH_HIP(hipStreamSynchronize(h.ws->own));
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_sort_population_mo(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m, uint64_t *out,
                                  void *stream)
try {
    if (!ws) return fail(PGH_INVALID_ARGUMENT, "null workspace");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n < 2u) { // corner cases (multi_objective.cpp:427-434)
        if (n == 1u) {
            const uint64_t zero = 0;
            PGH_HIP(hipMemcpyAsync(out, &zero, 8u, hipMemcpyHostToDevice, s));
            PGH_HIP(hipStreamSynchronize(s));
        }
        return PGH_OK;
    }
    if (!objs || !out) return fail(PGH_INVALID_ARGUMENT, "null buffer");
    std::vector<uint64_t> fo;
    PGH_TRY(fnds_fronts(ws, objs, n, m, s, fo));
    uint64_t largest = 0;
    for (size_t f = 0; f + 1u < fo.size(); ++f) largest = std::max(largest, fo[f + 1u] - fo[f]);
    if (largest >= 2u && m < 2u) return fail(PGH_INVALID_ARGUMENT, crowding_objectives_message(m));
    const uint32_t n32 = static_cast<uint32_t>(n), m32 = static_cast<uint32_t>(m);
    PGH_TRY(reserve_sort(ws, n));
    PGH_TRY(crowding_all(ws, objs, n32, m32, true, fo.size() - 1u, s));
    // (rank ascending, crowding descending), equal pairs by index
    sorter so{ws, s, n32};
    k_crowd_keys<<<grid_for(n), kBlock, 0, s>>>(static_cast<const double *>(ws->crowd.ptr), n32, nullptr, 0u,
                                                static_cast<uint64_t *>(ws->sk0.ptr),
                                                static_cast<uint32_t *>(ws->sv0.ptr));
    PGH_HIP(hipGetLastError());
    PGH_TRY(so.by_u64(n32, 64u)); // -> sv1
    const uint32_t *order = static_cast<const uint32_t *>(ws->sv1.ptr);
    if (fo.size() > 2u) {
        PGH_TRY(so.by_rank(n32, fo.size() - 1u)); // -> sv0
        order = static_cast<const uint32_t *>(ws->sv0.ptr);
    }
    k_widen<<<grid_for(n), kBlock, 0, s>>>(order, n32, nullptr, 0u, out);
    PGH_HIP(hipGetLastError());
    PGH_HIP(hipStreamSynchronize(s));
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_sort_population_mo_host(const double *objs, uint64_t n, uint64_t m, uint64_t *out)
try {
    if (n < 2u) {
        if (n == 1u) out[0] = 0u;
        return PGH_OK;
    }
    host_call h;
    const double *dev = nullptr;
    PGH_TRY(h.begin(objs, n, m, &dev));
    PGH_TRY(h.ws->out_staged.reserve(n * 8u));
    PGH_TRY(pgh_sort_population_mo(h.ws, dev, n, m, static_cast<uint64_t *>(h.ws->out_staged.ptr), h.ws->own));
    PGH_HIP(hipMemcpyAsync(out, h.ws->out_staged.ptr, n * 8u, hipMemcpyDeviceToHost, h.ws->own));
    PGH_HIP(hipStreamSynchronize(h.ws->own));
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_select_best_N_mo(pgh_fnds_ws *ws, const double *objs, uint64_t n, uint64_t m, uint64_t N,
                                uint64_t *out, uint64_t *out_len, void *stream)
try {
    if (!ws || !out_len) return fail(PGH_INVALID_ARGUMENT, "null workspace or length");
    hipStream_t s = static_cast<hipStream_t>(stream);
    *out_len = 0;
    // corner cases (multi_objective.cpp:346-360)
    if (N == 0u || n == 0u) return PGH_OK;
    if (!out) return fail(PGH_INVALID_ARGUMENT, "null buffer");
    if (n == 1u || N >= n) {
        const uint64_t count = n == 1u ? 1u : n;
        std::vector<uint64_t> iota(count);
        for (uint64_t k = 0; k < count; ++k) iota[k] = k;
        PGH_HIP(hipMemcpyAsync(out, iota.data(), count * 8u, hipMemcpyHostToDevice, s));
        PGH_HIP(hipStreamSynchronize(s));
        *out_len = count;
        return PGH_OK;
    }
    if (!objs) return fail(PGH_INVALID_ARGUMENT, "null buffer");
    std::vector<uint64_t> fo;
    PGH_TRY(fnds_fronts(ws, objs, n, m, s, fo));
    // whole fronts while they fit (multi_objective.cpp:365-377)
    size_t f = 0;
    while (f + 1u < fo.size() && fo[f + 1u] <= N) ++f;
    const uint64_t whole = fo[f];
    if (whole) PGH_HIP(hipMemcpyAsync(out, ws->front_items.ptr, whole * 8u, hipMemcpyDeviceToDevice, s));
    if (whole == N) {
        PGH_HIP(hipStreamSynchronize(s));
        *out_len = N;
        return PGH_OK;
    }
    // the next front by crowding distance, descending (:378-394); its size is >= 2
    if (m < 2u) return fail(PGH_INVALID_ARGUMENT, crowding_objectives_message(m));
    const uint64_t fb = fo[f], size = fo[f + 1u] - fo[f];
    const uint32_t n32 = static_cast<uint32_t>(n), m32 = static_cast<uint32_t>(m);
    PGH_TRY(reserve_sort(ws, n));
    PGH_TRY(crowding_all(ws, objs, n32, m32, true, fo.size() - 1u, s));
    sorter so{ws, s, static_cast<uint32_t>(size)};
    k_crowd_keys<<<grid_for(size), kBlock, 0, s>>>(static_cast<const double *>(ws->crowd.ptr),
                                                   static_cast<uint32_t>(size),
                                                   static_cast<const uint64_t *>(ws->front_items.ptr), fb,
                                                   static_cast<uint64_t *>(ws->sk0.ptr),
                                                   static_cast<uint32_t *>(ws->sv0.ptr));
    PGH_HIP(hipGetLastError());
    PGH_TRY(so.by_u64(static_cast<uint32_t>(size), 64u)); // -> sv1 (positions)
    const uint64_t remaining = N - whole;
    k_widen<<<grid_for(remaining), kBlock, 0, s>>>(static_cast<const uint32_t *>(ws->sv1.ptr),
                                                   static_cast<uint32_t>(remaining),
                                                   static_cast<const uint64_t *>(ws->front_items.ptr), fb, out + whole);
    PGH_HIP(hipGetLastError());
    PGH_HIP(hipStreamSynchronize(s));
    *out_len = N;
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_select_best_N_mo_host(const double *objs, uint64_t n, uint64_t m, uint64_t N, uint64_t *out,
                                     uint64_t *out_len)
try {
    if (!out_len) return fail(PGH_INVALID_ARGUMENT, "null length");
    *out_len = 0;
    if (N == 0u || n == 0u) return PGH_OK;
    if (n == 1u || N >= n) {
        const uint64_t count = n == 1u ? 1u : n;
        for (uint64_t k = 0; k < count; ++k) out[k] = k;
        *out_len = count;
        return PGH_OK;
    }
    host_call h;
    const double *dev = nullptr;
    PGH_TRY(h.begin(objs, n, m, &dev));
    PGH_TRY(h.ws->out_staged.reserve(n * 8u));
    PGH_TRY(pgh_select_best_N_mo(h.ws, dev, n, m, N, static_cast<uint64_t *>(h.ws->out_staged.ptr), out_len,
                                 h.ws->own));
    if (*out_len) PGH_HIP(hipMemcpyAsync(out, h.ws->out_staged.ptr, *out_len * 8u, hipMemcpyDeviceToHost, h.ws->own));
    PGH_HIP(hipStreamSynchronize(h.ws->own));
    return PGH_OK;
} catch (...) {
    return api_exception();
}

} // extern "C"
