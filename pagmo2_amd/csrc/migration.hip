// migration.hip -- migration between the islands of a device archipelago
// (island::evolve, src/island.cpp:452-642) with the select_best selection
// policy (src/s_policies/select_best.cpp:60-116) and the fair_replace
// replacement policy (src/r_policies/fair_replace.cpp:62-129), for
// single-objective unconstrained problems.
//
// Who migrates where, and how many, never depends on the individuals: the
// migration engines' draws (uniform_int / uniform_real over std::mt19937),
// the database sizes and the policies' counts are fixed by the topology, the
// rates and the population sizes.  The host therefore plans each round with
// the same standard-library calls as the reference, and the device moves the
// individuals:
//   immigration  the candidates of each replacing island (its sources'
//                database entries, concatenated in connection order) are
//                sorted by fitness and the first n_keep kept; the population
//                and the kept migrants are sorted by fitness; the first pop
//                become the new population, best first.
//   emigration   each island's population sorted by fitness; the first
//                n_emig rows become its database entry.
// Both sorts are rocPRIM's segmented radix sort (one segment per island) on
// an order-preserving key of pagmo's less_than_f (NaN last, -0 == +0), which
// is stable: equal fitness values keep their input order.  The reference's
// std::sort leaves that order open; on distinct values the results agree.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "fnds_ws.hpp"
#include "pgh_internal.hpp"

namespace pgh
{

namespace
{

constexpr uint32_t kBlock = 256;
constexpr uint64_t kFromDb = uint64_t(1) << 63; // reference bit: a database row, else a population row

struct dev_isl {
    uint32_t nx, pop;
    uint64_t xoff, foff, eoff; // population rows, fitness, first element (flat over all islands)
    uint64_t dbx, dbf;         // database entry: rows, fitness
};

// A replacing island of the round: its candidates start at cbeg (sorted
// order) and n_keep of them join its pop individuals.
struct rep_seg {
    uint32_t island, pop, n_keep, cbeg;
};

uint32_t grid_for(uint64_t count)
{
    return static_cast<uint32_t>((count + kBlock - 1u) / kBlock);
}

__device__ __forceinline__ uint64_t fit_key(double f)
{
    return static_cast<uint64_t>(order_key(f)) ^ (uint64_t(1) << 63);
}

// Last s in [0, n) with off[s] <= i (off ascending, off[0] = 0).
template <typename T>
__device__ __forceinline__ uint32_t seg_of(const T *__restrict__ off, uint32_t n, uint64_t i)
{
    uint32_t lo = 0, hi = n - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (static_cast<uint64_t>(off[mid]) <= i) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

__device__ __forceinline__ uint32_t island_of(const dev_isl *__restrict__ isl, uint32_t n, uint64_t e)
{
    uint32_t lo = 0, hi = n - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (isl[mid].eoff <= e) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

// Candidate i: the fitness of the database row it names.
__global__ __launch_bounds__(kBlock) void k_cand_keys(const uint64_t *__restrict__ cref, uint32_t C,
                                                      const dev_isl *__restrict__ isl, const double *__restrict__ dbF,
                                                      uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= C) return;
    const uint64_t ref = cref[i];
    const uint32_t src = static_cast<uint32_t>(ref >> 32) & 0x7fffffffu;
    key[i] = fit_key(dbF[isl[src].dbf + static_cast<uint32_t>(ref)]);
    val[i] = i;
}

// Merged item m of a replacing island: population row `local` (local < pop),
// else its (local - pop)-th best candidate.
__global__ __launch_bounds__(kBlock) void k_merge_fill(const rep_seg *__restrict__ seg, uint32_t nseg,
                                                       const uint32_t *__restrict__ moff,
                                                       const dev_isl *__restrict__ isl, const double *__restrict__ sF,
                                                       const uint64_t *__restrict__ cref,
                                                       const uint64_t *__restrict__ ckey,
                                                       const uint32_t *__restrict__ cval, uint32_t M,
                                                       uint64_t *__restrict__ mref, uint64_t *__restrict__ mkey,
                                                       uint32_t *__restrict__ mval)
{
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= M) return;
    const uint32_t s = seg_of(moff, nseg, m);
    const rep_seg sg = seg[s];
    const uint32_t local = m - moff[s];
    if (local < sg.pop) {
        mref[m] = (static_cast<uint64_t>(sg.island) << 32) | local;
        mkey[m] = fit_key(sF[isl[sg.island].foff + local]);
    } else {
        const uint32_t c = sg.cbeg + (local - sg.pop);
        mref[m] = cref[cval[c]];
        mkey[m] = ckey[c];
    }
    mval[m] = m;
}

// One lane per element of every island: replacing islands take row p of
// their new population from the merged item sorted to position p; the others
// copy their population unchanged.
__global__ __launch_bounds__(kBlock) void k_mig_copy(const dev_isl *__restrict__ isl, uint32_t n_isl,
                                                     const int32_t *__restrict__ seg_of_isl,
                                                     const uint32_t *__restrict__ moff,
                                                     const uint32_t *__restrict__ mval,
                                                     const uint64_t *__restrict__ mref, const double *__restrict__ sX,
                                                     const double *__restrict__ sF, const double *__restrict__ dbX,
                                                     const double *__restrict__ dbF, double *__restrict__ dX,
                                                     double *__restrict__ dF, uint64_t total)
{
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= total) return;
    const uint32_t k = island_of(isl, n_isl, t);
    const dev_isl di = isl[k];
    const uint64_t m = t - di.eoff;
    const uint32_t p = static_cast<uint32_t>(m / di.nx);
    const uint32_t d = static_cast<uint32_t>(m - static_cast<uint64_t>(p) * di.nx);
    const int32_t s = seg_of_isl[k];
    if (s < 0) {
        dX[di.xoff + m] = sX[di.xoff + m];
        if (d == 0u) dF[di.foff + p] = sF[di.foff + p];
        return;
    }
    const uint64_t ref = mref[mval[moff[s] + p]];
    const double *row;
    double f;
    if (ref & kFromDb) {
        const dev_isl ds = isl[static_cast<uint32_t>(ref >> 32) & 0x7fffffffu];
        const uint32_t j = static_cast<uint32_t>(ref);
        row = dbX + ds.dbx + static_cast<uint64_t>(j) * di.nx;
        f = dbF[ds.dbf + j];
    } else {
        const uint32_t r = static_cast<uint32_t>(ref);
        row = sX + di.xoff + static_cast<uint64_t>(r) * di.nx;
        f = sF[di.foff + r];
    }
    dX[di.xoff + static_cast<uint64_t>(p) * di.nx + d] = row[d];
    if (d == 0u) dF[di.foff + p] = f;
}

// Emigration keys: row r of island k at poff[k] + r.
__global__ __launch_bounds__(kBlock) void k_pop_keys(const dev_isl *__restrict__ isl, uint32_t n_isl,
                                                     const uint32_t *__restrict__ poff, const double *__restrict__ F,
                                                     uint32_t E, uint64_t *__restrict__ key,
                                                     uint32_t *__restrict__ val)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= E) return;
    const uint32_t k = seg_of(poff, n_isl, i);
    const uint32_t r = i - poff[k];
    key[i] = fit_key(F[isl[k].foff + r]);
    val[i] = r;
}

// Database entry of island k: its rows in sorted order, the first n_emig
// (goff: flat element offsets of the entries, n_isl + 1 values).
__global__ __launch_bounds__(kBlock) void k_emig_copy(const dev_isl *__restrict__ isl, uint32_t n_isl,
                                                      const uint32_t *__restrict__ poff,
                                                      const uint64_t *__restrict__ goff,
                                                      const uint32_t *__restrict__ val, const double *__restrict__ X,
                                                      const double *__restrict__ F, double *__restrict__ dbX,
                                                      double *__restrict__ dbF, uint64_t G)
{
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= G) return;
    const uint32_t k = seg_of(goff, n_isl, t);
    const dev_isl di = isl[k];
    const uint64_t m = t - goff[k];
    const uint32_t j = static_cast<uint32_t>(m / di.nx);
    const uint32_t d = static_cast<uint32_t>(m - static_cast<uint64_t>(j) * di.nx);
    const uint32_t r = val[poff[k] + j];
    dbX[di.dbx + m] = X[di.xoff + static_cast<uint64_t>(r) * di.nx + d];
    if (d == 0u) dbF[di.dbf + j] = F[di.foff + r];
}

// Stable segmented sort of (key, val) pairs, segments [off[s], off[s+1]).
pgh_status seg_sort(dev_buf &tmp, const uint64_t *k0, uint64_t *k1, const uint32_t *v0, uint32_t *v1, uint32_t size,
                    uint32_t nseg, const uint32_t *off, hipStream_t s)
{
    size_t bytes = 0;
    PGH_HIP(rocprim::segmented_radix_sort_pairs(nullptr, bytes, k0, k1, v0, v1, size, nseg, off, off + 1, 0u, 64u,
                                                s));
    PGH_TRY(tmp.reserve(std::max<size_t>(bytes, 8u)));
    PGH_HIP(rocprim::segmented_radix_sort_pairs(tmp.ptr, bytes, k0, k1, v0, v1, size, nseg, off, off + 1, 0u, 64u,
                                                s));
    return PGH_OK;
}

std::string rate_str(double r)
{
    return std::to_string(r);
}

} // namespace

struct migr_ctx {
    int type = 0, handling = 0;
    uint32_t n = 0;
    std::vector<migr_island> isl;
    std::vector<uint64_t> conn_off, conn_src;
    std::vector<double> conn_w;
    std::vector<uint64_t> n_emig, n_rep; // the policies' counts over each island's population
    std::vector<std::mt19937> eng;
    std::vector<uint64_t> db_count; // host view of the database sizes
    std::vector<uint64_t> dbx_off;
    uint64_t total_elems = 0, c_max = 0, m_max = 0, e_total = 0, g_total = 0;
    dev_buf d_isl, dbX, dbF, plan, cref, ckey0, ckey1, cval0, cval1, mref, mkey0, mkey1, mval0, mval1, emig, sort_tmp;
    // the round's plan, staged in pinned memory: segs | coff | moff | seg_of_isl | cref
    void *staging = nullptr;
    size_t staging_bytes = 0;
    size_t o_coff = 0, o_moff = 0, o_sof = 0, o_cref = 0;
    hipEvent_t staged = nullptr;
    bool staged_pending = false;

    ~migr_ctx()
    {
        if (staged) {
            if (staged_pending) (void)hipEventSynchronize(staged);
            (void)hipEventDestroy(staged);
        }
        if (staging) (void)hipHostFree(staging);
    }
};

static pgh_status check_rate(const pgh_migr_rate &r)
{
    // base_sr_policy.cpp:54-66
    if (r.fractional && (!std::isfinite(r.fraction) || r.fraction < 0. || r.fraction > 1.)) {
        return fail(PGH_INVALID_ARGUMENT, "Invalid fractional migration rate specified in the constructor of a "
                                          "replacement/selection policy: the rate must be in the [0., 1.] range, "
                                          "but it is "
                                              + rate_str(r.fraction) + " instead");
    }
    return PGH_OK;
}

// select_best.cpp:76-97 / fair_replace.cpp:79-101 (before the min with the
// migrants' count): numeric_cast truncates the fractional product.
static pgh_status rate_count(const pgh_migr_rate &r, uint64_t inds, const char *policy, uint64_t &out)
{
    if (r.fractional) {
        out = std::min(static_cast<uint64_t>(r.fraction * static_cast<double>(inds)), inds);
        return PGH_OK;
    }
    if (r.count > inds) {
        return fail(PGH_INVALID_ARGUMENT, "The absolute migration rate (" + std::to_string(r.count) + ") in a "
                                              + policy + " is larger than the number of input individuals ("
                                              + std::to_string(inds) + ")");
    }
    out = r.count;
    return PGH_OK;
}

void migr_destroy(migr_ctx *m)
{
    delete m;
}

pgh_status migr_create(const std::vector<migr_island> &isl, int type, int handling, const uint64_t *off,
                       const uint64_t *src, const double *w, const pgh_migr_rate *sel, const pgh_migr_rate *rep,
                       const uint32_t *seeds, migr_ctx **out)
{
    if (!off || !sel || !rep || !seeds) return fail(PGH_INVALID_ARGUMENT, "null argument");
    if (type != PGH_MIGRATION_P2P && type != PGH_MIGRATION_BROADCAST) {
        return fail(PGH_INVALID_ARGUMENT, "unknown migration type " + std::to_string(type));
    }
    if (handling != PGH_MIGRANTS_PRESERVE && handling != PGH_MIGRANTS_EVICT) {
        return fail(PGH_INVALID_ARGUMENT, "unknown migrant handling " + std::to_string(handling));
    }
    const uint32_t n = static_cast<uint32_t>(isl.size());
    if (off[0] != 0u) return fail(PGH_INVALID_ARGUMENT, "conn_offsets must start at 0");
    for (uint32_t k = 0; k < n; ++k) {
        if (off[k + 1u] < off[k]) return fail(PGH_INVALID_ARGUMENT, "conn_offsets must not decrease");
    }
    if (off[n] != 0u && (!src || !w)) return fail(PGH_INVALID_ARGUMENT, "null argument");
    auto m = std::make_unique<migr_ctx>();
    m->type = type;
    m->handling = handling;
    m->n = n;
    m->isl = isl;
    m->conn_off.assign(off, off + n + 1u);
    m->conn_src.assign(src, src + off[n]);
    m->conn_w.assign(w, w + off[n]);
    for (uint64_t j = 0; j < off[n]; ++j) {
        // topology.cpp:58-68
        const double x = w[j];
        if (!std::isfinite(x)) {
            return fail(PGH_INVALID_ARGUMENT,
                        "invalid weight for the edge of a topology: the value " + rate_str(x) + " is not finite");
        }
        if (x < 0. || x > 1.) {
            return fail(PGH_INVALID_ARGUMENT, "invalid weight for the edge of a topology: the value " + rate_str(x)
                                                  + " is not in the [0., 1.] range");
        }
    }
    std::vector<dev_isl> h(n);
    uint64_t dbx = 0, dbf = 0;
    for (uint32_t k = 0; k < n; ++k) {
        PGH_TRY(check_rate(sel[k]));
        PGH_TRY(check_rate(rep[k]));
        uint64_t ne = 0, nr = 0;
        PGH_TRY(rate_count(sel[k], isl[k].pop, "'Select best' selection policy", ne));
        PGH_TRY(rate_count(rep[k], isl[k].pop, "'fair_replace' replacement policy", nr));
        m->n_emig.push_back(ne);
        m->n_rep.push_back(nr);
        h[k] = dev_isl{isl[k].nx, isl[k].pop, isl[k].xoff, isl[k].foff, isl[k].eoff, dbx, dbf};
        m->dbx_off.push_back(dbx);
        dbx += ne * isl[k].nx;
        dbf += ne;
        m->total_elems = isl[k].eoff + static_cast<uint64_t>(isl[k].pop) * isl[k].nx;
        m->e_total += isl[k].pop;
        m->g_total += ne * isl[k].nx;
    }
    for (uint32_t k = 0; k < n; ++k) {
        uint64_t ck = 0;
        for (uint64_t j = off[k]; j < off[k + 1u]; ++j) {
            if (src[j] >= n) return fail(PGH_INVALID_ARGUMENT, "connection from island " + std::to_string(src[j])
                                                                   + ", which does not exist");
            if (isl[src[j]].nx != isl[k].nx) {
                return fail(PGH_INVALID_ARGUMENT, "islands " + std::to_string(src[j]) + " and " + std::to_string(k)
                                                      + " are connected but their problems differ in dimension");
            }
            ck += m->n_emig[src[j]];
        }
        m->c_max += ck;
        if (off[k + 1u] > off[k]) m->m_max += isl[k].pop + std::min(m->n_rep[k], ck);
    }
    if (m->c_max >= (uint64_t(1) << 31) || m->m_max >= (uint64_t(1) << 31) || m->e_total >= (uint64_t(1) << 31)) {
        return fail(PGH_UNSUPPORTED, "too many individuals for one migration sort");
    }
    for (uint32_t k = 0; k < n; ++k) m->eng.emplace_back(seeds[k]);
    m->db_count.assign(n, 0u); // archipelago.cpp:511: empty entries

    PGH_TRY(m->d_isl.reserve(n * sizeof(dev_isl)));
    PGH_HIP(hipMemcpy(m->d_isl.ptr, h.data(), n * sizeof(dev_isl), hipMemcpyHostToDevice));
    PGH_TRY(m->dbX.reserve(std::max<uint64_t>(dbx, 1u) * 8u));
    PGH_TRY(m->dbF.reserve(std::max<uint64_t>(dbf, 1u) * 8u));
    const uint64_t cm = std::max<uint64_t>(m->c_max, 1u), mm = std::max<uint64_t>(std::max(m->m_max, m->e_total), 1u);
    PGH_TRY(m->cref.reserve(cm * 8u));
    PGH_TRY(m->ckey0.reserve(cm * 8u));
    PGH_TRY(m->ckey1.reserve(cm * 8u));
    PGH_TRY(m->cval0.reserve(cm * 4u));
    PGH_TRY(m->cval1.reserve(cm * 4u));
    PGH_TRY(m->mref.reserve(mm * 8u));
    PGH_TRY(m->mkey0.reserve(mm * 8u));
    PGH_TRY(m->mkey1.reserve(mm * 8u));
    PGH_TRY(m->mval0.reserve(mm * 4u));
    PGH_TRY(m->mval1.reserve(mm * 4u));
    // emigration offsets, fixed: poff (n + 1 x u32) | goff (n + 1 x u64)
    std::vector<uint32_t> poff(n + 1u, 0u);
    std::vector<uint64_t> goff(n + 1u, 0u);
    for (uint32_t k = 0; k < n; ++k) {
        poff[k + 1u] = poff[k] + isl[k].pop;
        goff[k + 1u] = goff[k] + m->n_emig[k] * isl[k].nx;
    }
    const size_t pbytes = (n + 1u) * 4u, gpos = (pbytes + 7u) & ~size_t(7);
    PGH_TRY(m->emig.reserve(gpos + (n + 1u) * 8u));
    PGH_HIP(hipMemcpy(m->emig.ptr, poff.data(), pbytes, hipMemcpyHostToDevice));
    PGH_HIP(hipMemcpy(static_cast<char *>(m->emig.ptr) + gpos, goff.data(), (n + 1u) * 8u, hipMemcpyHostToDevice));
    // the round's plan
    m->o_coff = ((n * sizeof(rep_seg)) + 7u) & ~size_t(7);
    m->o_moff = m->o_coff + (((n + 1u) * 4u + 7u) & ~size_t(7));
    m->o_sof = m->o_moff + (((n + 1u) * 4u + 7u) & ~size_t(7));
    m->o_cref = m->o_sof + ((n * 4u + 7u) & ~size_t(7));
    m->staging_bytes = m->o_cref + cm * 8u;
    PGH_TRY(m->plan.reserve(m->staging_bytes));
    PGH_HIP(hipHostMalloc(&m->staging, m->staging_bytes, hipHostMallocDefault));
    PGH_HIP(hipEventCreateWithFlags(&m->staged, hipEventDisableTiming));
    *out = m.release();
    return PGH_OK;
}

pgh_status migr_immigrate(migr_ctx *m, const double *sX, const double *sF, double *dX, double *dF, bool *moved,
                          hipStream_t s)
{
    *moved = false;
    // the pinned plan of the previous round must have reached the device
    if (m->staged_pending) {
        PGH_HIP(hipEventSynchronize(m->staged));
        m->staged_pending = false;
    }
    char *st = static_cast<char *>(m->staging);
    auto *segs = reinterpret_cast<rep_seg *>(st);
    auto *coff = reinterpret_cast<uint32_t *>(st + m->o_coff);
    auto *moff = reinterpret_cast<uint32_t *>(st + m->o_moff);
    auto *sof = reinterpret_cast<int32_t *>(st + m->o_sof);
    auto *cref = reinterpret_cast<uint64_t *>(st + m->o_cref);
    uint32_t nseg = 0, C = 0;
    coff[0] = 0u;
    moff[0] = 0u;
    std::vector<uint32_t> taken;
    for (uint32_t k = 0; k < m->n; ++k) {
        sof[k] = -1;
        const uint64_t b = m->conn_off[k], e = m->conn_off[k + 1u];
        if (b == e) continue;
        // island.cpp:458-597: the draws of one migration step
        taken.clear();
        if (m->type == PGH_MIGRATION_P2P) {
            const auto j = std::uniform_int_distribution<std::size_t>(0, e - b - 1u)(m->eng[k]);
            if (!(std::uniform_real_distribution<>{}(m->eng[k]) < m->conn_w[b + j])) continue;
            taken.push_back(static_cast<uint32_t>(m->conn_src[b + j]));
        } else {
            for (uint64_t j = b; j < e; ++j) {
                if (std::uniform_real_distribution<>{}(m->eng[k]) < m->conn_w[j]) {
                    taken.push_back(static_cast<uint32_t>(m->conn_src[j]));
                }
            }
        }
        const uint32_t cbeg = C;
        for (const uint32_t src : taken) {
            for (uint64_t i = 0; i < m->db_count[src]; ++i) {
                cref[C++] = kFromDb | (static_cast<uint64_t>(src) << 32) | i;
            }
            // archipelago.cpp:658-678: extract_migrants leaves the entry empty
            if (m->handling == PGH_MIGRANTS_EVICT) m->db_count[src] = 0u;
        }
        const uint32_t n_keep = static_cast<uint32_t>(std::min<uint64_t>(m->n_rep[k], C - cbeg));
        sof[k] = static_cast<int32_t>(nseg);
        segs[nseg] = rep_seg{k, m->isl[k].pop, n_keep, cbeg};
        coff[nseg + 1u] = C;
        moff[nseg + 1u] = moff[nseg] + m->isl[k].pop + n_keep;
        ++nseg;
    }
    if (nseg == 0u) return PGH_OK;
    const uint32_t M = moff[nseg];
    char *pl = static_cast<char *>(m->plan.ptr);
    PGH_HIP(hipMemcpyAsync(pl, st, m->o_cref + static_cast<size_t>(C) * 8u, hipMemcpyHostToDevice, s));
    PGH_HIP(hipEventRecord(m->staged, s));
    m->staged_pending = true;
    const auto *d_segs = reinterpret_cast<const rep_seg *>(pl);
    const auto *d_coff = reinterpret_cast<const uint32_t *>(pl + m->o_coff);
    const auto *d_moff = reinterpret_cast<const uint32_t *>(pl + m->o_moff);
    const auto *d_sof = reinterpret_cast<const int32_t *>(pl + m->o_sof);
    const auto *d_cref = reinterpret_cast<const uint64_t *>(pl + m->o_cref);
    const auto *isl = static_cast<const dev_isl *>(m->d_isl.ptr);
    auto *ck0 = static_cast<uint64_t *>(m->ckey0.ptr);
    auto *ck1 = static_cast<uint64_t *>(m->ckey1.ptr);
    auto *cv0 = static_cast<uint32_t *>(m->cval0.ptr);
    auto *cv1 = static_cast<uint32_t *>(m->cval1.ptr);
    if (C > 0u) {
        k_cand_keys<<<grid_for(C), kBlock, 0, s>>>(d_cref, C, isl, static_cast<const double *>(m->dbF.ptr), ck0, cv0);
        PGH_HIP(hipGetLastError());
        PGH_TRY(seg_sort(m->sort_tmp, ck0, ck1, cv0, cv1, C, nseg, d_coff, s));
    }
    auto *mr = static_cast<uint64_t *>(m->mref.ptr);
    auto *mk0 = static_cast<uint64_t *>(m->mkey0.ptr);
    auto *mk1 = static_cast<uint64_t *>(m->mkey1.ptr);
    auto *mv0 = static_cast<uint32_t *>(m->mval0.ptr);
    auto *mv1 = static_cast<uint32_t *>(m->mval1.ptr);
    k_merge_fill<<<grid_for(M), kBlock, 0, s>>>(d_segs, nseg, d_moff, isl, sF, d_cref, ck1, cv1, M, mr, mk0, mv0);
    PGH_HIP(hipGetLastError());
    PGH_TRY(seg_sort(m->sort_tmp, mk0, mk1, mv0, mv1, M, nseg, d_moff, s));
    k_mig_copy<<<grid_for(m->total_elems), kBlock, 0, s>>>(isl, m->n, d_sof, d_moff, mv1, mr, sX, sF,
                                                           static_cast<const double *>(m->dbX.ptr),
                                                           static_cast<const double *>(m->dbF.ptr), dX, dF,
                                                           m->total_elems);
    PGH_HIP(hipGetLastError());
    *moved = true;
    return PGH_OK;
}

pgh_status migr_emigrate(migr_ctx *m, const double *X, const double *F, hipStream_t s)
{
    for (uint32_t k = 0; k < m->n; ++k) m->db_count[k] = m->n_emig[k];
    if (m->g_total == 0u) return PGH_OK;
    const auto *isl = static_cast<const dev_isl *>(m->d_isl.ptr);
    const auto *poff = static_cast<const uint32_t *>(m->emig.ptr);
    const size_t gpos = ((m->n + 1u) * 4u + 7u) & ~size_t(7);
    const auto *goff = reinterpret_cast<const uint64_t *>(static_cast<const char *>(m->emig.ptr) + gpos);
    const uint32_t E = static_cast<uint32_t>(m->e_total);
    auto *k0 = static_cast<uint64_t *>(m->mkey0.ptr);
    auto *k1 = static_cast<uint64_t *>(m->mkey1.ptr);
    auto *v0 = static_cast<uint32_t *>(m->mval0.ptr);
    auto *v1 = static_cast<uint32_t *>(m->mval1.ptr);
    k_pop_keys<<<grid_for(E), kBlock, 0, s>>>(isl, m->n, poff, F, E, k0, v0);
    PGH_HIP(hipGetLastError());
    PGH_TRY(seg_sort(m->sort_tmp, k0, k1, v0, v1, E, m->n, poff, s));
    k_emig_copy<<<grid_for(m->g_total), kBlock, 0, s>>>(isl, m->n, poff, goff, v1, X, F,
                                                        static_cast<double *>(m->dbX.ptr),
                                                        static_cast<double *>(m->dbF.ptr), m->g_total);
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

pgh_status migr_get(migr_ctx *m, uint64_t k, uint64_t *count, double *X, double *F)
{
    if (k >= m->n) return fail(PGH_INVALID_ARGUMENT, "no such island");
    const uint64_t c = m->db_count[k];
    if (count) *count = c;
    const auto &is = m->isl[k];
    if (X && c) {
        PGH_HIP(hipMemcpy(X, static_cast<const double *>(m->dbX.ptr) + m->dbx_off[k], c * is.nx * 8u,
                          hipMemcpyDeviceToHost));
    }
    if (F && c) {
        uint64_t dbf = 0;
        for (uint64_t i = 0; i < k; ++i) dbf += m->n_emig[i];
        PGH_HIP(hipMemcpy(F, static_cast<const double *>(m->dbF.ptr) + dbf, c * 8u, hipMemcpyDeviceToHost));
    }
    return PGH_OK;
}

} // namespace pgh
