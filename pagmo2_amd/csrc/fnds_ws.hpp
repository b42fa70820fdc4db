// fnds_ws.hpp -- the fast_non_dominated_sorting workspace (fnds.hip) and the
// multi-objective orderings that reuse it (mo_sort.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pgh_internal.hpp"

namespace pgh
{
// Order-preserving int64 key of a double under pagmo's comparisons
// (include/pagmo/detail/custom_comparisons.hpp:54-100): NaN sorts after +inf
// and equals every NaN, -0 equals +0.
__device__ __forceinline__ int64_t order_key(double x)
{
    if (x != x) return INT64_MAX;
    if (x == 0.0) return 0; // -0 == +0
    const int64_t b = __double_as_longlong(x);
    return b >= 0 ? b : (b ^ INT64_MAX);
}
} // namespace pgh

struct pgh_fnds_ws {
    hipStream_t stream = nullptr;
    hipStream_t own = nullptr; // created by the host-buffer entry points
    pgh::dev_buf staged;
    uint64_t n = 0, m = 0, words = 0;
    pgh::dev_buf keys, D, cnt32, ndom32, key32, dom_count, rank, front_items, front_offsets, dom_offsets, dom_items,
        scalars;
    pgh::dev_buf comp, ctl; // the grid peel's sort-key lists and flags (two halves each)
    uint64_t h_sizes[3] = {0, 0, 0}; // n_fronts, n_dom_items, peel barrier timeout
    bool sized = false;
    // workgroups of the grid-wide front peel: 0 = the default rule, else the
    // value set or found by pgh_fnds_tune (any count gives the same result)
    uint32_t peel_blocks = 0;
    // mo_sort.hip: sort keys and values (two of each), crowding, radix-sort
    // scratch, and a staged output
    pgh::dev_buf sk0, sk1, sv0, sv1, crowd, sort_tmp, out_staged;
    // a host thread's own workspace (pgh_fnds_thread_workspace): released
    // with its thread, never by pgh_fnds_destroy
    bool thread_owned = false;
};
