// pso.hip -- pso_gen (src/algorithms/pso_gen.cpp:120-530) for an archipelago
// of independent islands of uneven size, on the device.
//
// Per island the device holds, contiguously: X, V, lbX (pop x nx, the memory
// double-buffered with a per-row selector), fit, lbfit (pop), the island's
// std::mt19937 state and a word buffer for a batch of kGenBatch generations
// (in draw order: r1 and r2 of pso_gen.cpp:300-301 for variants 1 and 5, the
// other variants' words as they draw them).  A generation is:
//   k_step      one persistent launch (or one per problem class): workgroups
//               take (island, rows) items from a queue and, per item, find
//               each row's best neighbour (pso_gen.cpp:593-622), move the
//               rows (:302-361), evaluate them with the batch evaluator's
//               device routines and update the memory (:443-449)
//   k_rng,      on a second stream, during the current batch: the next
//   k_rng_gen   batch's words, the island's generator split in segments that
//               reach their own positions by jump-ahead (mt_dev.hpp,
//               mt_jump): exactly the draws of the reference's sequential loop
// plus k_gbest_update for islands of the gbest topology.  Batched mode
// captures a whole evolve call, every island at once, in a hipGraph;
// per-island mode launches the same kernels island by island.  Both run
// identical arithmetic per island.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fitness_dev.hpp"
#include "mt_dev.hpp"
#include "pgh_internal.hpp"

namespace pgh
{
using namespace dev;

struct island_desc {
    int32_t kind;
    uint32_t dim, nx, pop;
    uint32_t gens, radius, lanes;
    uint32_t memory; // pso_gen's memory flag: velocities kept between evolve calls
    double omega, eta1, eta2, max_vel;
    uint64_t xoff; // X, V, lbX offset (doubles)
    uint64_t foff; // fit, lbfit, offset (doubles)
    uint64_t woff; // word buffer offset (u32)
    uint64_t whalf; // offset of the second half of a double-buffered word buffer (batches alternate), 0: one
    uint64_t toff; // trace offset (doubles), gens x pop
    uint64_t eoff; // cumulative pop*nx of the islands before this one
    // the generator split in `segs` segments: a batch of gps generations
    // (kGenBatch, or all of a call's when fewer) is bwords = gps W words, and
    // segment s hands out its words s swords .. (s + 1) swords - 1 (the last
    // segment the rest); seg0 is the first segment's index; j_* index the
    // archipelago's jump table: z^n0 (the velocity draws), z^(s swords) at
    // j_pos + s - 1 (to segment s), z^(bwords - swords) and
    // z^((segs - 1) swords) (from the end of a full segment, and of the last
    // one, to the same segment of the next batch)
    uint64_t bwords, swords;
    uint32_t segs, seg0, gps;
    uint32_t j_n0, j_pos, j_main, j_last;
    // pso_gen's variant (1..6) and topology (1 gbest, 2 lbest, 3 von Neumann),
    // the neighbours per particle (lbest 2 radius, von Neumann 4) and the von
    // Neumann lattice's columns; the random words of a generation (wgen) and
    // of one unit of it (wpu words per element, or per particle for variants
    // 3 and 4, pso_gen.cpp:241-327); gbest: the swarm's best position (nx
    // doubles at goff of the gbest buffer, its fitness after them)
    uint32_t variant, topo, nbk, vcols;
    uint64_t wgen, goff;
    uint32_t wpu, pad_;
};

struct seg_desc {
    uint32_t island, s;
};

struct tile {
    uint32_t island, first, count, pad;
};

namespace
{

constexpr uint32_t kBlock = 256;
// The generators hand out the words of kGenBatch generations at a time; an
// island's generator is split in segments of a fixed number of words (a power
// of two chosen per archipelago between kSegMin and kSegMax words) so that
// the batch's words are made by many workgroups at once
constexpr uint32_t kGenBatch = 32;
constexpr uint64_t kSegMin = uint64_t(1) << 16, kSegMax = uint64_t(1) << 22, kSoloWords = uint64_t(1) << 20;
// k_rng_gen's workgroup: the generator steps 227 words at a time (mt_run)
constexpr uint32_t kMtGenBlock = 256;

// "a <= b" in the sense of less_than_f(a, b) || equal_to_f(a, b)
// (custom_comparisons.hpp:54-100): NaN is larger than everything but itself.
__device__ __forceinline__ bool le_f(double a, double b)
{
    const bool an = a != a, bn = b != b;
    if (an) return bn;
    if (bn) return true;
    return a <= b;
}

__device__ __forceinline__ uint32_t find_island(const island_desc *__restrict__ d, uint32_t i0, uint32_t i1,
                                                uint64_t e)
{
    // last island whose eoff <= e (islands [i0, i1) are contiguous in eoff)
    uint32_t lo = i0, hi = i1 - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (d[mid].eoff <= e) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

struct jump_tab {
    const uint32_t *lens;
    const uint16_t *lists; // kMtIdxStride values per distance
};

__device__ __forceinline__ void load_state(const mt_state *__restrict__ src, uint32_t *ring, uint64_t &K, uint64_t &C)
{
    for (uint32_t i = threadIdx.x; i < kMtRing; i += blockDim.x) ring[i] = src->ring[i];
    K = src->K;
    C = src->C;
    __syncthreads();
}

__device__ __forceinline__ void store_state(mt_state *__restrict__ dst, const uint32_t *ring, uint64_t K, uint64_t C)
{
    for (uint32_t i = threadIdx.x; i < kMtRing; i += blockDim.x) dst->ring[i] = ring[i];
    if (threadIdx.x == 0) {
        dst->K = K;
        dst->C = C;
    }
}

// The island's std::mt19937 draws, by its segments (one workgroup each).
// Every word of an evolve call has a position in the island's one sequential
// stream: first the velocity draws (n0 = 2 words per element,
// pso_gen.cpp:193-200, unless the island keeps its velocities: memory, once
// they exist), then generation g's 4 words per element (r1 and r2 of
// pso_gen.cpp:300-301) at n0 + g W.  Batch b holds generations
// b gps .. (b + 1) gps - 1, and segment s of the batch hands out its words
// s swords .. (s + 1) swords - 1 (the last segment the rest) into their place
// in the word buffer.
//
// k_rng phase 0, the start of an evolve call: every segment takes the
// island's canonical generator (everything handed out so far); segment 0
// hands out the velocity draws and every segment moves to its share of the
// first batch, n0 + s swords (jumps z^n0 and z^(s swords)).  k_rng_gen, batch b: each segment hands out its share; the segment
// that hands out the island's last word keeps the generator there as the
// canonical one.  k_rng phase 2, after batch b when another follows: each
// segment with more to hand out jumps from the end of its share to its share
// of batch b + 1.  The words are those of one sequential generator, in order.
__device__ __forceinline__ void seg_span(const island_desc &di, uint32_t s, uint32_t b, uint64_t n0, uint64_t &start,
                                         uint64_t &count, uint64_t &end)
{
    const uint64_t W = di.wgen;
    end = n0 + static_cast<uint64_t>(di.gens) * W; // one past the island's last word of the call
    start = n0 + static_cast<uint64_t>(b) * di.bwords + static_cast<uint64_t>(s) * di.swords;
    const uint64_t share = s + 1u < di.segs ? di.swords : di.bwords - static_cast<uint64_t>(s) * di.swords;
    // the words left from start on, 0 past the end: end - min(start, end)
    // cannot wrap.  (Written as start < end ? min(share, end - start) : 0,
    // the compiled code lost the comparison and took a full share past the
    // end, overwriting the live batch of an island whose last batch was
    // already in hand.)
    const uint64_t left = end - min(start, end);
    count = min(share, left);
}

__device__ __forceinline__ uint64_t velocity_words(const island_desc &di, uint32_t v_valid)
{
    return (di.memory && v_valid) ? 0u : 2ull * di.pop * di.nx;
}

__global__ __launch_bounds__(kMtJumpBlock) void k_rng(const island_desc *__restrict__ d,
                                                      const seg_desc *__restrict__ sd, mt_state *__restrict__ st,
                                                      mt_state *__restrict__ segst, jump_tab jt,
                                                      uint32_t *__restrict__ words, uint32_t sg0, uint32_t phase,
                                                      uint32_t b, uint32_t v_valid)
{
    __shared__ uint32_t ring[kMtRing];
    __shared__ uint4 lds_idx[kMtIdxStride / 8u];
    const seg_desc sg = sd[sg0 + blockIdx.x];
    const island_desc di = d[sg.island];
    const auto jump = [&](uint64_t &K, uint64_t &C, uint64_t J, uint32_t j) {
        mt_jump(ring, K, C, J, jt.lists + static_cast<uint64_t>(j) * kMtIdxStride, jt.lens[j], lds_idx);
    };
    const uint64_t n0 = velocity_words(di, v_valid);
    uint64_t K, C;
    mt_state *mine = &segst[di.seg0 + sg.s];
    if (phase == 0u) {
        load_state(&st[sg.island], ring, K, C);
        if (sg.s == 0u) {
            if (n0) {
                mt_run(ring, K, C, n0, words + di.woff);
                C += n0;
            }
            if (di.gens == 0u) store_state(&st[sg.island], ring, K, C);
        } else {
            if (n0) jump(K, C, n0, di.j_n0);
            jump(K, C, static_cast<uint64_t>(sg.s) * di.swords, di.j_pos + sg.s - 1u);
        }
        store_state(mine, ring, K, C);
        return;
    }
    // phase 2: after batch b, to batch b + 1 (a one-segment generator ends
    // its share where the next begins)
    if (di.segs == 1u) return;
    uint64_t start, count, end;
    seg_span(di, sg.s, b, n0, start, count, end);
    if (count == 0u || start + count == end || start + di.bwords >= end) return;
    load_state(mine, ring, K, C);
    if (sg.s + 1u < di.segs) {
        jump(K, C, di.bwords - di.swords, di.j_main);
    } else {
        jump(K, C, static_cast<uint64_t>(di.segs - 1u) * di.swords, di.j_last);
    }
    store_state(mine, ring, K, C);
}

// Batch b's words of one segment (see k_rng).
__device__ __forceinline__ void rng_gen_segment(const island_desc &di, uint32_t s, uint32_t b, uint64_t n0,
                                                mt_state *__restrict__ st, mt_state *__restrict__ mine,
                                                uint32_t *__restrict__ batch_words, uint32_t *ring)
{
    uint64_t start, count, end;
    seg_span(di, s, b, n0, start, count, end);
    if (count == 0u) return;
    uint64_t K, C;
    load_state(mine, ring, K, C);
    mt_run(ring, K, C, count, batch_words + static_cast<uint64_t>(s) * di.swords);
    C += count;
    store_state(start + count == end ? st : mine, ring, K, C);
}

__global__ __launch_bounds__(kMtGenBlock) void k_rng_gen(const island_desc *__restrict__ d,
                                                         const seg_desc *__restrict__ sd, mt_state *__restrict__ st,
                                                         mt_state *__restrict__ segst, uint32_t *__restrict__ words,
                                                         uint32_t sg0, uint32_t b, uint32_t v_valid)
{
    __shared__ uint32_t ring[kMtRing];
    const seg_desc sg = sd[sg0 + blockIdx.x];
    const island_desc di = d[sg.island];
    rng_gen_segment(di, sg.s, b, velocity_words(di, v_valid), &st[sg.island], &segst[di.seg0 + sg.s],
                    words + di.woff + ((b & 1u) ? di.whalf : 0u), ring);
}

// Start of an evolve call (pso_gen.cpp:183-200): V = U(minv, maxv) per
// element, and X = lbX = the population.  The population is X/fit before the
// first call and lbX/lbfit after one (pop.set_xf(i, lbX, lbfit)).
__global__ __launch_bounds__(kBlock) void k_vinit(const island_desc *__restrict__ d, uint32_t i0, uint32_t i1,
                                                  uint64_t e0, uint64_t total, const uint32_t *__restrict__ words,
                                                  double *__restrict__ X, double *__restrict__ V,
                                                  double *__restrict__ lbX, double *__restrict__ fit,
                                                  double *__restrict__ lbfit, uint32_t *__restrict__ selA,
                                                  uint32_t v_valid, uint32_t from_lb)
{
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= total) return;
    const uint64_t e = e0 + t;
    const uint32_t k = find_island(d, i0, i1, e);
    const island_desc di = d[k];
    const uint64_t m = e - di.eoff;
    const uint32_t j = static_cast<uint32_t>(m % di.nx);
    if (!(di.memory && v_valid)) {
        double lb, ub;
        bounds_of(di.kind, j, lb, ub);
        const double vwidth = (ub - lb) * di.max_vel; // pso_gen.cpp:177-181
        const double minv = -1. * vwidth, maxv = vwidth;
        const uint32_t *w = words + di.woff + 2u * m;
        V[di.xoff + m] = mt_uniform(w[0], w[1], minv, maxv);
    }
    const uint64_t p = m / di.nx;
    if (j == 0u) selA[di.foff + p] = 0u; // every row's memory starts in lbX0 (k_step)
    if (from_lb) {
        X[di.xoff + m] = lbX[di.xoff + m];
        if (j == 0u) fit[di.foff + p] = lbfit[di.foff + p];
    } else {
        lbX[di.xoff + m] = X[di.xoff + m];
        if (j == 0u) lbfit[di.foff + p] = fit[di.foff + p];
    }
}

// Best neighbour of particle p on the lbest ring (pso_gen.cpp:679-699):
// radius..1 to the left, then 1..radius to the right; the later neighbour
// wins ties (:615-617).  lbf: the island's memory fitness.
// Neighbour n of particle p: lbest (pso_gen.cpp:679-699) left radius..1, then
// right 1..radius; von Neumann (pso_gen.cpp:719-746) west, east, north,
// south on a (pop / cols) x cols lattice with wrapping edges.
__device__ __forceinline__ uint32_t neighbour(const island_desc &di, uint32_t p, uint32_t n)
{
    const uint32_t S = di.pop;
    if (di.topo == 3u) {
        const uint32_t cols = di.vcols, rows = S / cols;
        const uint32_t px = p % cols, py = p / cols;
        uint32_t qx = px, qy = py;
        if (n == 0u) qx = (px + cols - 1u) % cols;
        else if (n == 1u) qx = (px + 1u) % cols;
        else if (n == 2u) qy = (py + rows - 1u) % rows;
        else qy = (py + 1u) % rows;
        return qy * cols + qx;
    }
    const uint32_t R = di.radius;
    if (n < R) {
        const uint32_t j = R - n;
        return p < j ? p - j + S : p - j;
    }
    const uint32_t j = n - R + 1u;
    return p + j >= S ? p + j - S : p + j;
}

// The best of the neighbours of p (pso_gen.cpp:593-622: a later neighbour wins
// ties) for the lbest and von Neumann topologies.
__device__ __noinline__ uint32_t best_neighbour_any(const island_desc &di, uint32_t p, const double *__restrict__ lbf)
{
    uint32_t bn = neighbour(di, p, 0u);
    for (uint32_t n = 1u; n < di.nbk; ++n) {
        const uint32_t q = neighbour(di, p, n);
        if (le_f(lbf[q], lbf[bn])) bn = q;
    }
    return bn;
}

__device__ __forceinline__ uint32_t best_neighbour(uint32_t p, uint32_t S, uint32_t R, const double *__restrict__ lbf)
{
    uint32_t bn = p < R ? p - R + S : p - R;
    for (uint32_t j = R - 1u; j >= 1u; --j) {
        const uint32_t q = p < j ? p - j + S : p - j;
        if (le_f(lbf[q], lbf[bn])) bn = q;
    }
    for (uint32_t j = 1u; j <= R; ++j) {
        const uint32_t q = p + j >= S ? p + j - S : p + j;
        if (le_f(lbf[q], lbf[bn])) bn = q;
    }
    return bn;
}

// The velocity of variants 1-5 (pso_gen.cpp:241-306), each expression in the
// reference's order: 1 and 3 with r1 and r2 (drawn per element or per
// particle), 2 and 4 with r1 for both terms, 5 with the constriction factor
// over the whole sum.
__device__ __forceinline__ double pso_velocity(const island_desc &di, double v_old, double x, double lbx, double best,
                                               double r1, double r2)
{
    switch (di.variant) {
        case 1u:
        case 3u:
            return di.omega * v_old + di.eta1 * r1 * (lbx - x) + di.eta2 * r2 * (best - x);
        case 2u:
        case 4u:
            return di.omega * v_old + di.eta1 * r1 * (lbx - x) + di.eta2 * r1 * (best - x);
        default:
            return di.omega * (v_old + di.eta1 * r1 * (lbx - x) + di.eta2 * r2 * (best - x));
    }
}

// The velocity clamp (pso_gen.cpp:336-342) and the bounded move (:345-361) of
// dimension dd: returns the new position, v the new velocity.
__device__ __forceinline__ double pso_clamp_move(const island_desc &di, uint32_t dd, double x, double &v)
{
    double lb, ub;
    bounds_of(di.kind, dd, lb, ub);
    const double vwidth = (ub - lb) * di.max_vel;
    const double minv = -1. * vwidth, maxv = vwidth;
    if (v > maxv) {
        v = maxv;
    } else if (v < minv) {
        v = minv;
    }
    double nxv = x + v;
    if (nxv < lb) {
        nxv = lb;
        v = 0.;
    } else if (nxv > ub) {
        nxv = ub;
        v = 0.;
    }
    return nxv;
}

// Variant 5 velocity (pso_gen.cpp:302-305), its clamp (:336-342) and the
// bounded move (:345-361) of dimension dd: returns the new position, v the new
// velocity.
__device__ __forceinline__ double pso_move(const island_desc &di, uint32_t dd, double x, double v_old, double lbx,
                                           double best, double r1, double r2, double &v)
{
    v = di.omega * (v_old + di.eta1 * r1 * (lbx - x) + di.eta2 * r2 * (best - x));
    return pso_clamp_move(di, dd, x, v);
}

// The words of unit u (an element, or a particle for variants 3 and 4) in
// generation gen, in draw order: wpu words per unit, wgen per generation.
__device__ __forceinline__ const uint32_t *gen_words(const island_desc &di, const uint32_t *__restrict__ words,
                                                     uint32_t gen, uint64_t u)
{
    const uint64_t half = ((gen / kGenBatch) & 1u) ? di.whalf : 0u;
    return words + di.woff + half + static_cast<uint64_t>(gen % kGenBatch) * di.wgen + di.wpu * u;
}

// gen_words for the variants of 4 words per element (1 and 5): the element's
// r1, r2 words as one aligned 16-byte unit.
__device__ __forceinline__ const uint4 *gen_words4(const island_desc &di, const uint32_t *__restrict__ words,
                                                   uint32_t gen, uint64_t m)
{
    const uint64_t half = ((gen / kGenBatch) & 1u) ? di.whalf : 0u;
    return reinterpret_cast<const uint4 *>(words + di.woff + half + static_cast<uint64_t>(gen % kGenBatch) * di.wgen) +
           m;
}

// Velocity and position of element m of island di in generation gen.
__device__ __forceinline__ void update_elem(const island_desc &di, uint64_t m, const uint32_t *__restrict__ words,
                                            double *__restrict__ X, double *__restrict__ V,
                                            const double *__restrict__ lbX, const double *__restrict__ lbfit,
                                            uint32_t gen)
{
    const uint32_t p = static_cast<uint32_t>(m / di.nx);
    const uint32_t dd = static_cast<uint32_t>(m - static_cast<uint64_t>(p) * di.nx);
    const uint32_t bn = best_neighbour(p, di.pop, di.radius, lbfit + di.foff);
    const uint4 w = *gen_words4(di, words, gen, m);
    const double r1 = mt_canonical(w.x, w.y);
    const double r2 = mt_canonical(w.z, w.w);
    const uint64_t xi = di.xoff + m;
    const double best = lbX[di.xoff + static_cast<uint64_t>(bn) * di.nx + dd];
    double v;
    X[xi] = pso_move(di, dd, X[xi], V[xi], lbX[xi], best, r1, r2, v);
    V[xi] = v;
}

struct eval_ctx {
    double *X, *lbX, *fit, *lbfit, *trace;
    uint32_t gen;
};

// Memory update of individual p by the L lanes of its group (lane r copies
// elements r, r+L, ...).
__device__ __forceinline__ void remember(const island_desc &di, const eval_ctx &c, uint32_t p, double f, uint32_t r,
                                         uint32_t L)
{
    const uint64_t row = di.xoff + static_cast<uint64_t>(p) * di.nx;
    if (le_f(f, c.lbfit[di.foff + p])) {
        for (uint32_t i = r; i < di.nx; i += L) c.lbX[row + i] = c.X[row + i];
        if (r == 0u) c.lbfit[di.foff + p] = f;
    }
    if (r == 0u) {
        c.fit[di.foff + p] = f;
        if (c.trace) c.trace[di.toff + static_cast<uint64_t>(c.gen) * di.pop + p] = f;
    }
}

template <int KIND, int L>
__device__ __forceinline__ void eval_separable(const island_desc &di, const tile &tl, const eval_ctx &c)
{
    const uint32_t g = threadIdx.x / L, r = threadIdx.x % L;
    if (g * L >= kBlock) return;
    const uint32_t gg = g < tl.count ? g : tl.count - 1u;
    const uint32_t p = tl.first + gg;
    const double *x = c.X + di.xoff + static_cast<uint64_t>(p) * di.nx;
    double f = (di.nx % 2u == 0u) ? separable_eval<KIND, L, true>(x, di.nx, static_cast<int>(r))
                                  : separable_eval<KIND, L, false>(x, di.nx, static_cast<int>(r));
    if (is_recompute_mark(f)) {
        // group-uniform: an element beyond the streaming cosine's range.  One
        // lane recomputes the individual as the batch evaluator does (the
        // reference's sequential order) and the group takes its value.
        double e = 0.;
        if (r == 0u) {
            e = (di.nx % 2u == 0u) ? separable_eval_exact<KIND, 1, true>(x, di.nx, 0)
                                   : separable_eval_exact<KIND, 1, false>(x, di.nx, 0);
        }
        f = __shfl(e, 0, L);
    }
    if (g < tl.count) remember(di, c, p, f, r, L);
}

template <int KIND>
__device__ __forceinline__ void eval_sep_dispatch(const island_desc &di, const tile &tl, const eval_ctx &c)
{
    switch (di.lanes) {
        case 64:
            eval_separable<KIND, 64>(di, tl, c);
            return;
        case 32:
            eval_separable<KIND, 32>(di, tl, c);
            return;
        case 16:
            eval_separable<KIND, 16>(di, tl, c);
            return;
        default:
            eval_separable<KIND, 4>(di, tl, c);
    }
}

template <int L, uint32_t APL>
__device__ __forceinline__ void eval_lj_a(const island_desc &di, const tile &tl, const eval_ctx &c, double *lds)
{
    const uint32_t atoms = di.dim;
    lj_stage(c.X + di.xoff + static_cast<uint64_t>(tl.first) * di.nx, tl.count, atoms, false, lds);
    __syncthreads();
    const uint32_t g = threadIdx.x / L, r = threadIdx.x % L;
    if (g >= lj_group_count(atoms, false)) return;
    const uint32_t gg = g < tl.count ? g : tl.count - 1u;
    const double f = lj_energy<L, APL, false>(lds + gg * lj_stride(atoms, false), atoms, static_cast<int>(r));
    if (g < tl.count) remember(di, c, tl.first + gg, f, r, L);
}

template <int L>
__device__ __forceinline__ void eval_lj(const island_desc &di, const tile &tl, const eval_ctx &c, double *lds)
{
    switch (lj_apl(di.dim, L)) {
        case 1:
            eval_lj_a<L, 1>(di, tl, c, lds);
            return;
        case 2:
            eval_lj_a<L, 2>(di, tl, c, lds);
            return;
        case 3:
            eval_lj_a<L, 3>(di, tl, c, lds);
            return;
        case 4:
            eval_lj_a<L, 4>(di, tl, c, lds);
            return;
        default:
            break;
    }
    if constexpr (L >= 32) {
        switch (lj_apl(di.dim, L)) {
            case 5:
                eval_lj_a<L, 5>(di, tl, c, lds);
                return;
            case 6:
                eval_lj_a<L, 6>(di, tl, c, lds);
                return;
            case 7:
                eval_lj_a<L, 7>(di, tl, c, lds);
                return;
            case 8:
                eval_lj_a<L, 8>(di, tl, c, lds);
                return;
            default:
                break;
        }
    }
    eval_lj_a<L, 0>(di, tl, c, lds);
}

// Fitness and memory update of the individuals of tile tl (all threads of
// the workgroup call it).
__device__ __forceinline__ void eval_tile(const island_desc &di, const tile &tl, const eval_ctx &c, double *lds)
{
    switch (di.kind) {
        case ROSENBROCK:
            if (threadIdx.x < tl.count) {
                const uint32_t p = tl.first + threadIdx.x;
                const double f = rosenbrock_eval(c.X + di.xoff + static_cast<uint64_t>(p) * di.nx, di.nx);
                remember(di, c, p, f, 0u, 1u);
            }
            return;
        case RASTRIGIN:
            eval_sep_dispatch<RASTRIGIN>(di, tl, c);
            return;
        case ACKLEY:
            eval_sep_dispatch<ACKLEY>(di, tl, c);
            return;
        case SCHWEFEL:
            eval_sep_dispatch<SCHWEFEL>(di, tl, c);
            return;
        case GRIEWANK:
            eval_sep_dispatch<GRIEWANK>(di, tl, c);
            return;
        case LENNARD_JONES:
            switch (di.lanes) {
                case 64:
                    eval_lj<64>(di, tl, c, lds);
                    return;
                case 32:
                    eval_lj<32>(di, tl, c, lds);
                    return;
                case 16:
                    eval_lj<16>(di, tl, c, lds);
                    return;
                case 8:
                    eval_lj<8>(di, tl, c, lds);
                    return;
                case 4:
                    eval_lj<4>(di, tl, c, lds);
                    return;
                case 2:
                    eval_lj<2>(di, tl, c, lds);
                    return;
                default:
                    eval_lj<1>(di, tl, c, lds);
                    return;
            }
    }
}

// ---- one generation of every island in one launch (k_step) ------------------
// A generation's work is split in items: rows [first, first + count) of one
// island.  A workgroup takes an item, moves its rows (best neighbour, velocity,
// position), evaluates them and updates their memory; a generation is one
// launch of workgroups that take items from a queue until it is empty.
//
// Items of different islands, and of the same island, run concurrently.  A
// particle's move reads the memory (lbX, lbfit) of its neighbours, which
// other items of the same generation update; the reference moves every
// particle before it evaluates any (pso_gen.cpp:232-449).  So the memory is
// kept twice: lbfit in two buffers, read from one and written to the other
// in alternate generations, and each row's lbX in one of two buffers, named
// by a per-row selector (itself kept in two buffers like lbfit).  A row that
// improves is written to the buffer its selector does not name, and the next
// selector names that buffer; a row that does not improve keeps its buffer.
// So no generation writes anything that the same generation reads from
// another item, and the arithmetic per element is that of the launch-per-phase
// evolve (update_elem, eval_tile), bit for bit.

enum : uint32_t { kStepRosen = 0u, kStepSep = 1u, kStepLj = 2u, kStepAll = 3u };

// Lanes per Lennard-Jones individual in an archipelago: the batch evaluator's
// for its cyclic schedule (lj_lanes_cyclic), or a whole wavefront from 80
// atoms on (16 lanes own the atoms, four quarters of them split the
// schedule's steps), so that one large cluster's evaluation, which bounds a
// generation, takes half as many steps.
__host__ __device__ constexpr uint32_t archi_lj_lanes(uint32_t atoms)
{
    return atoms >= 80u ? 64u : static_cast<uint32_t>(lj_lanes_cyclic(atoms));
}

__host__ __device__ constexpr uint32_t step_class(int kind)
{
    return kind == ROSENBROCK ? kStepRosen : (kind == LENNARD_JONES ? kStepLj : kStepSep);
}

// LDS of a rosenbrock item: each wavefront's term tile (rosenbrock_wave_rows)
constexpr uint32_t kStepRosenRows = (kBlock / 64u) * kRosenWaveRows;
constexpr size_t kStepRosenLds = (kBlock / 64u) * kRosenWaveRows * kRosenWaveStride * sizeof(double);

struct step_args {
    const island_desc *d;
    const tile *items;
    const uint32_t *order; // item indices in the order they are handed out; null: base + i
    uint32_t base, n_items;
    uint32_t *queue; // this launch's counter, zero before the launch
    const uint32_t *words;
    double *X, *V, *lbX0, *lbX1, *fit, *trace;
    const double *lbf_cur;
    double *lbf_nxt;
    const uint32_t *sel_cur;
    uint32_t *sel_nxt;
    const double *gbest; // the gbest islands' best positions (island_desc::goff)
    uint32_t gen;
};

// Per row of an item (at most kBlock rows): its best neighbour and the
// memory rows of the particle and of that neighbour, found once per row
// rather than once per element.
struct step_rows {
    uint32_t bn[kBlock];
    uint32_t sel[kBlock]; // bit 0: the particle's memory is in lbX1, bit 1: the neighbour's
};

// step_move for every variant and topology but variant 5 with lbest
// (pso_gen.cpp:232-327): the best neighbour of the island's topology, or the
// swarm's best position (gbest); words per particle for variants 3 and 4; for
// variant 6 (FIPS) every neighbour's memory with one draw each.
__device__ __noinline__ void step_move_general(const island_desc &di, const tile &tl, const step_args &a,
                                               step_rows &rw)
{
    const double *lbf = a.lbf_cur + di.foff;
    const uint32_t *sel = a.sel_cur + di.foff;
    const bool fips = di.variant == 6u, gbest = di.topo == 1u;
    for (uint32_t t = threadIdx.x; t < tl.count; t += kBlock) {
        const uint32_t p = tl.first + t;
        const uint32_t bn = (fips || gbest) ? p : best_neighbour_any(di, p, lbf);
        rw.bn[t] = bn;
        rw.sel[t] = (sel[p] ? 1u : 0u) | (sel[bn] ? 2u : 0u);
    }
    __syncthreads();
    const uint64_t e0 = static_cast<uint64_t>(tl.first) * di.nx;
    const uint32_t span = tl.count * di.nx;
    const double *gb = a.gbest + di.goff;
    const bool per_particle = di.variant == 3u || di.variant == 4u;
    const bool one_draw = di.variant == 2u || di.variant == 4u;
    const double acc_coef = di.eta1 + di.eta2; // pso_gen.cpp:221
    for (uint32_t e = threadIdx.x; e < span; e += kBlock) {
        const uint32_t t = e / di.nx, dd = e - t * di.nx;
        const uint64_t m = e0 + e;
        const uint32_t p = tl.first + t;
        const uint32_t sb = rw.sel[t];
        const uint64_t xi = di.xoff + m;
        const double x = a.X[xi], v_old = a.V[xi];
        double v;
        if (fips) {
            const uint32_t *w = gen_words(di, a.words, a.gen, m);
            double sum = 0.;
            for (uint32_t n = 0; n < di.nbk; ++n) {
                const uint32_t q = neighbour(di, p, n);
                const double lq = (sel[q] ? a.lbX1 : a.lbX0)[di.xoff + static_cast<uint64_t>(q) * di.nx + dd];
                sum += mt_canonical(w[2u * n], w[2u * n + 1u]) * acc_coef * (lq - x);
            }
            v = di.omega * (v_old + sum / static_cast<double>(di.nbk));
        } else {
            const uint32_t *w = gen_words(di, a.words, a.gen, per_particle ? p : m);
            const double r1 = mt_canonical(w[0], w[1]);
            const double r2 = one_draw ? 0. : mt_canonical(w[2], w[3]);
            const double lbx = ((sb & 1u) ? a.lbX1 : a.lbX0)[xi];
            const double best = gbest ? gb[dd]
                                      : ((sb & 2u) ? a.lbX1 : a.lbX0)[di.xoff + static_cast<uint64_t>(rw.bn[t]) * di.nx + dd];
            v = pso_velocity(di, v_old, x, lbx, best, r1, r2);
        }
        a.X[xi] = pso_clamp_move(di, dd, x, v);
        a.V[xi] = v;
    }
}

// Moves the rows of item tl (update_elem's arithmetic, the memory read through
// the selectors).  All threads of the workgroup; a __syncthreads() must
// follow.
// GEN: the kernel serves an archipelago with islands of other variants or
// topologies (a separate instantiation, so that the call of the out-of-line
// path does not weigh on the register allocation of the default kernel).
template <bool GEN>
__device__ __forceinline__ void step_move(const island_desc &di, const tile &tl, const step_args &a, step_rows &rw)
{
    if constexpr (GEN) {
        if (di.variant != 5u || di.topo != 2u) { // island-uniform
            step_move_general(di, tl, a, rw);
            return;
        }
    }
    const double *lbf = a.lbf_cur + di.foff;
    const uint32_t *sel = a.sel_cur + di.foff;
    for (uint32_t t = threadIdx.x; t < tl.count; t += kBlock) {
        const uint32_t p = tl.first + t;
        const uint32_t bn = best_neighbour(p, di.pop, di.radius, lbf);
        rw.bn[t] = bn;
        rw.sel[t] = (sel[p] ? 1u : 0u) | (sel[bn] ? 2u : 0u);
    }
    __syncthreads();
    // item-local element e (count * nx < 2^32, build_items): row t, column dd
    const uint64_t e0 = static_cast<uint64_t>(tl.first) * di.nx;
    const uint32_t span = tl.count * di.nx;
    for (uint32_t e = threadIdx.x; e < span; e += kBlock) {
        const uint32_t t = e / di.nx, dd = e - t * di.nx;
        const uint64_t m = e0 + e;
        const uint32_t bn = rw.bn[t], sb = rw.sel[t];
        const uint4 w = *gen_words4(di, a.words, a.gen, m);
        const double r1 = mt_canonical(w.x, w.y);
        const double r2 = mt_canonical(w.z, w.w);
        const uint64_t xi = di.xoff + m;
        const double best = ((sb & 2u) ? a.lbX1 : a.lbX0)[di.xoff + static_cast<uint64_t>(bn) * di.nx + dd];
        const double lbx = ((sb & 1u) ? a.lbX1 : a.lbX0)[xi];
        double v;
        a.X[xi] = pso_move(di, dd, a.X[xi], a.V[xi], lbx, best, r1, r2, v);
        a.V[xi] = v;
    }
}

// Memory update of particle p with fitness f (pso_gen.cpp:443-449) by the L
// lanes of its group (lane r copies elements r, r + L, ...), and its fitness
// and trace.
__device__ __forceinline__ void step_remember(const island_desc &di, const step_args &a, uint32_t p, double f,
                                              uint32_t r, uint32_t L)
{
    const uint64_t fi = di.foff + p;
    const double lbf = a.lbf_cur[fi];
    const uint32_t s = a.sel_cur[fi];
    if (le_f(f, lbf)) {
        const uint64_t row = di.xoff + static_cast<uint64_t>(p) * di.nx;
        double *dst = (s ? a.lbX0 : a.lbX1) + row;
        for (uint32_t i = r; i < di.nx; i += L) dst[i] = a.X[row + i];
        if (r == 0u) {
            a.lbf_nxt[fi] = f;
            a.sel_nxt[fi] = s ^ 1u;
        }
    } else if (r == 0u) {
        a.lbf_nxt[fi] = lbf;
        a.sel_nxt[fi] = s;
    }
    if (r == 0u) {
        a.fit[fi] = f;
        if (a.trace) a.trace[di.toff + static_cast<uint64_t>(a.gen) * di.pop + p] = f;
    }
}

// Rosenbrock: 8 rows per wavefront (rosenbrock_wave_rows), then the memory
// update by the wavefront's 64 lanes, row by row.
__device__ __forceinline__ void step_eval_rosen(const island_desc &di, const tile &tl, const step_args &a,
                                                double *lds)
{
    const uint32_t wave = threadIdx.x / 64u, lane = threadIdx.x % 64u;
    const uint32_t r0 = wave * kRosenWaveRows;
    if (r0 >= tl.count) return; // wavefront-uniform
    const uint32_t rows = min(kRosenWaveRows, tl.count - r0);
    const uint32_t p0 = tl.first + r0;
    const double f = rosenbrock_wave_rows(a.X + di.xoff + static_cast<uint64_t>(p0) * di.nx, rows, di.nx,
                                          lds + wave * kRosenWaveRows * kRosenWaveStride, lane);
    for (uint32_t r = 0; r < rows; ++r) step_remember(di, a, p0 + r, __shfl(f, static_cast<int>(r), 64), lane, 64u);
}

// An individual with an element beyond the streaming cosine's range, again
// with the reference's expression (rare; out of line so that its registers do
// not weigh on the streaming code).
template <int KIND>
__device__ __noinline__ double step_sep_cold(const double *__restrict__ x, uint32_t nx)
{
    return (nx % 2u == 0u) ? separable_eval_exact<KIND, 1, true>(x, nx, 0)
                           : separable_eval_exact<KIND, 1, false>(x, nx, 0);
}

template <int KIND, int L>
__device__ __forceinline__ void step_eval_sep_l(const island_desc &di, const tile &tl, const step_args &a)
{
    const uint32_t g = threadIdx.x / L, r = threadIdx.x % L;
    const uint32_t gg = g < tl.count ? g : tl.count - 1u;
    const uint32_t p = tl.first + gg;
    const double *x = a.X + di.xoff + static_cast<uint64_t>(p) * di.nx;
    double f = (di.nx % 2u == 0u) ? separable_eval<KIND, L, true>(x, di.nx, static_cast<int>(r))
                                  : separable_eval<KIND, L, false>(x, di.nx, static_cast<int>(r));
    if (is_recompute_mark(f)) {
        double e = 0.;
        if (r == 0u) e = step_sep_cold<KIND>(x, di.nx);
        f = __shfl(e, 0, L);
    }
    if (g < tl.count) step_remember(di, a, p, f, r, L);
}

template <int KIND>
__device__ __forceinline__ void step_eval_sep(const island_desc &di, const tile &tl, const step_args &a)
{
    switch (di.lanes) {
        case 64:
            step_eval_sep_l<KIND, 64>(di, tl, a);
            return;
        case 32:
            step_eval_sep_l<KIND, 32>(di, tl, a);
            return;
        case 16:
            step_eval_sep_l<KIND, 16>(di, tl, a);
            return;
        default:
            step_eval_sep_l<KIND, 4>(di, tl, a);
    }
}

template <int L, uint32_t APL>
__device__ __forceinline__ void step_eval_lj_a(const island_desc &di, const tile &tl, const step_args &a, double *lds)
{
    const uint32_t atoms = di.dim;
    lj_stage(a.X + di.xoff + static_cast<uint64_t>(tl.first) * di.nx, tl.count, atoms, false, lds);
    __syncthreads();
    const uint32_t g = threadIdx.x / L, r = threadIdx.x % L;
    if (g * L >= kBlock || g >= tl.count) return; // the item's groups
    const double f = lj_energy<L, APL, false>(lds + g * lj_stride(atoms, false), atoms, static_cast<int>(r));
    step_remember(di, a, tl.first + g, f, r, L);
}

template <int L>
__device__ __forceinline__ void step_eval_lj_l(const island_desc &di, const tile &tl, const step_args &a, double *lds)
{
    switch (lj_apl(di.dim, L)) {
        case 1:
            step_eval_lj_a<L, 1>(di, tl, a, lds);
            return;
        case 2:
            step_eval_lj_a<L, 2>(di, tl, a, lds);
            return;
        case 3:
            step_eval_lj_a<L, 3>(di, tl, a, lds);
            return;
        case 4:
            step_eval_lj_a<L, 4>(di, tl, a, lds);
            return;
        default:
            break;
    }
    if constexpr (L >= 32) {
        switch (lj_apl(di.dim, L)) {
            case 5:
                step_eval_lj_a<L, 5>(di, tl, a, lds);
                return;
            case 6:
                step_eval_lj_a<L, 6>(di, tl, a, lds);
                return;
            case 7:
                step_eval_lj_a<L, 7>(di, tl, a, lds);
                return;
            case 8:
                step_eval_lj_a<L, 8>(di, tl, a, lds);
                return;
            default:
                break;
        }
    }
    step_eval_lj_a<L, 0>(di, tl, a, lds);
}

__device__ __forceinline__ void step_eval_lj(const island_desc &di, const tile &tl, const step_args &a, double *lds)
{
    switch (di.lanes) {
        case 64:
            step_eval_lj_l<64>(di, tl, a, lds);
            return;
        case 32:
            step_eval_lj_l<32>(di, tl, a, lds);
            return;
        case 16:
            step_eval_lj_l<16>(di, tl, a, lds);
            return;
        case 8:
            step_eval_lj_l<8>(di, tl, a, lds);
            return;
        case 4:
            step_eval_lj_l<4>(di, tl, a, lds);
            return;
        case 2:
            step_eval_lj_l<2>(di, tl, a, lds);
            return;
        default:
            step_eval_lj_l<1>(di, tl, a, lds);
            return;
    }
}

template <uint32_t CLASS, bool GEN>
__device__ __forceinline__ void step_item(const island_desc &di, const tile &tl, const step_args &a, double *lds,
                                          step_rows &rw)
{
    step_move<GEN>(di, tl, a, rw);
    __syncthreads(); // the item's new rows are visible to the whole workgroup
    const uint32_t cls = CLASS == kStepAll ? step_class(di.kind) : CLASS;
    if (cls == kStepRosen) {
        if constexpr (CLASS == kStepRosen || CLASS == kStepAll) step_eval_rosen(di, tl, a, lds);
    } else if (cls == kStepLj) {
        if constexpr (CLASS == kStepLj || CLASS == kStepAll) step_eval_lj(di, tl, a, lds);
    } else {
        if constexpr (CLASS == kStepSep || CLASS == kStepAll) {
            switch (di.kind) {
                case RASTRIGIN:
                    step_eval_sep<RASTRIGIN>(di, tl, a);
                    break;
                case ACKLEY:
                    step_eval_sep<ACKLEY>(di, tl, a);
                    break;
                case SCHWEFEL:
                    step_eval_sep<SCHWEFEL>(di, tl, a);
                    break;
                default:
                    step_eval_sep<GRIEWANK>(di, tl, a);
                    break;
            }
        }
    }
}

// One generation of the items order[0 .. n_items) (or base + i): a persistent
// grid whose workgroups take items from the queue until it is empty (every
// wavefront reads the same index, so all leave the loop together).
// Wavefronts per SIMD each class's kernel is built for (its register budget):
// the move phase streams from HBM and wants occupancy; the Lennard-Jones
// evaluation wants registers for its atom slots.
#ifndef PGH_STEP_BOUNDED
#define PGH_STEP_BOUNDED 1
#endif
constexpr int step_min_waves(uint32_t cls)
{
    return !PGH_STEP_BOUNDED ? 1 : (cls == kStepRosen || cls == kStepSep ? 8 : (cls == kStepLj ? 4 : 3));
}

template <uint32_t CLASS, bool GEN>
__global__ __launch_bounds__(kBlock, step_min_waves(CLASS)) void k_step(step_args a)
{
    extern __shared__ double lds[];
    __shared__ uint32_t s_next;
    __shared__ step_rows rw;
    for (;;) {
        if (threadIdx.x == 0) s_next = atomicAdd(a.queue, 1u);
        __syncthreads();
        const uint32_t i = s_next;
        if (i >= a.n_items) return;
        const uint32_t it = a.order ? a.order[i] : a.base + i;
        const tile tl = a.items[it];
        const island_desc di = a.d[tl.island];
        if (a.gen < di.gens) step_item<CLASS, GEN>(di, tl, a, lds, rw);
        __syncthreads(); // LDS and s_next are reused by the next item
    }
}

// gbest (pso_gen.cpp:644-662): the swarm's best position at the start of an
// evolve call is pop.best_idx()'s (population.cpp:285-311 for one objective):
// the first index of the minimum by operator<, so a NaN at index 0 stays and
// no other NaN is taken.  One workgroup per island of gb_list; each thread
// takes a contiguous chunk, and the chunks are combined in index order.
__global__ __launch_bounds__(kBlock) void k_gbest_init(const island_desc *__restrict__ d,
                                                       const uint32_t *__restrict__ gb_list,
                                                       const double *__restrict__ X, const double *__restrict__ fit,
                                                       double *__restrict__ gbest)
{
    __shared__ uint32_t s_idx[kBlock];
    __shared__ double s_f[kBlock];
    const island_desc di = d[gb_list[blockIdx.x]];
    const double *f = fit + di.foff;
    const uint32_t t = threadIdx.x;
    const uint32_t per = (di.pop + kBlock - 1u) / kBlock;
    const uint32_t c0 = min(t * per, di.pop), c1 = min(c0 + per, di.pop);
    uint32_t bi = UINT32_MAX;
    double bf = 0.;
    for (uint32_t i = c0; i < c1; ++i) {
        const double v = f[i];
        if (v != v) continue;
        if (bi == UINT32_MAX || v < bf) {
            bi = i;
            bf = v;
        }
    }
    s_idx[t] = bi;
    s_f[t] = bf;
    __syncthreads();
    if (t == 0u) {
        uint32_t w = UINT32_MAX;
        double wf = 0.;
        for (uint32_t u = 0; u < kBlock; ++u) {
            if (s_idx[u] != UINT32_MAX && (w == UINT32_MAX || s_f[u] < wf)) {
                w = s_idx[u];
                wf = s_f[u];
            }
        }
        if (f[0] != f[0] || w == UINT32_MAX) w = 0u;
        s_idx[0] = w;
    }
    __syncthreads();
    const uint32_t w = s_idx[0];
    double *g = gbest + di.goff;
    for (uint32_t dd = t; dd < di.nx; dd += kBlock) g[dd] = X[di.xoff + static_cast<uint64_t>(w) * di.nx + dd];
    if (t == 0u) g[di.nx] = f[w];
}

// gbest after a generation's memory update (pso_gen.cpp:443-459): in index
// order, every particle that improved (its selector flipped) whose fitness is
// <= the swarm's best (less_than_f or equal_to_f) becomes the best.  That is
// the last improved particle of the least fitness, if it is <= the best.
__global__ __launch_bounds__(kBlock) void k_gbest_update(const island_desc *__restrict__ d,
                                                         const uint32_t *__restrict__ gb_list,
                                                         const double *__restrict__ X, const double *__restrict__ fit,
                                                         const uint32_t *__restrict__ sel_cur,
                                                         const uint32_t *__restrict__ sel_nxt, double *__restrict__ gbest,
                                                         uint32_t gen)
{
    __shared__ uint32_t s_idx[kBlock];
    __shared__ double s_f[kBlock];
    const island_desc di = d[gb_list[blockIdx.x]];
    if (gen >= di.gens) return; // workgroup-uniform
    const double *f = fit + di.foff;
    const uint32_t *sc = sel_cur + di.foff, *sn = sel_nxt + di.foff;
    const uint32_t t = threadIdx.x;
    const uint32_t per = (di.pop + kBlock - 1u) / kBlock;
    const uint32_t c0 = min(t * per, di.pop), c1 = min(c0 + per, di.pop);
    uint32_t bi = UINT32_MAX;
    double bf = 0.;
    for (uint32_t i = c0; i < c1; ++i) {
        if (sc[i] == sn[i]) continue;
        const double v = f[i];
        if (bi == UINT32_MAX || le_f(v, bf)) {
            bi = i;
            bf = v;
        }
    }
    s_idx[t] = bi;
    s_f[t] = bf;
    __syncthreads();
    double *g = gbest + di.goff;
    if (t == 0u) {
        uint32_t w = UINT32_MAX;
        double wf = 0.;
        for (uint32_t u = 0; u < kBlock; ++u) {
            if (s_idx[u] != UINT32_MAX && (w == UINT32_MAX || le_f(s_f[u], wf))) {
                w = s_idx[u];
                wf = s_f[u];
            }
        }
        if (w != UINT32_MAX && !le_f(wf, g[di.nx])) w = UINT32_MAX;
        s_idx[0] = w;
    }
    __syncthreads();
    const uint32_t w = s_idx[0];
    if (w == UINT32_MAX) return; // workgroup-uniform
    for (uint32_t dd = t; dd < di.nx; dd += kBlock) g[dd] = X[di.xoff + static_cast<uint64_t>(w) * di.nx + dd];
    if (t == 0u) g[di.nx] = f[w];
}

__global__ __launch_bounds__(kBlock) void k_zero_u32(uint32_t *__restrict__ p, uint64_t n)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// *differs = 1 when the n words of x and y are not all equal (the tuner's
// check that two mappings leave identical states).
__global__ __launch_bounds__(kBlock) void k_differs(const uint32_t *__restrict__ x, const uint32_t *__restrict__ y,
                                                    uint64_t n, uint32_t *__restrict__ differs)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n && x[i] != y[i]) *differs = 1u;
}

// End of an evolve call: island k's memory back into lbX0 and lbfit (A), the
// buffers the rest of the library reads.  After the island's gens
// generations its memory fitness is in A (gens even) or B, and row p's memory
// in the buffer named by its final selector (in selA or selB likewise).  The
// selectors are only read here (k_vinit clears selA at the next call's start).
__global__ __launch_bounds__(kBlock) void k_settle(const island_desc *__restrict__ d, uint32_t i0, uint32_t i1,
                                                   uint64_t e0, uint64_t total, double *__restrict__ lbX0,
                                                   const double *__restrict__ lbX1, double *__restrict__ lbfA,
                                                   const double *__restrict__ lbfB, const uint32_t *__restrict__ selA,
                                                   const uint32_t *__restrict__ selB)
{
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= total) return;
    const uint64_t e = e0 + t;
    const uint32_t k = find_island(d, i0, i1, e);
    const island_desc di = d[k];
    const uint64_t m = e - di.eoff;
    const uint32_t p = static_cast<uint32_t>(m / di.nx);
    const bool odd = (di.gens & 1u) != 0u;
    const uint64_t fi = di.foff + p;
    if ((odd ? selB[fi] : selA[fi]) != 0u) lbX0[di.xoff + m] = lbX1[di.xoff + m];
    if (odd && m == static_cast<uint64_t>(p) * di.nx) lbfA[fi] = lbfB[fi];
}

// Generations g0 .. g1 - 1 (within one batch of random words) of one island
// that is a single evaluation tile, in one workgroup: k_update's and k_eval's
// work in turn, with workgroup barriers where their launches would end.
// Only this workgroup touches the island's X, V, lbX, fit and lbfit, so the
// barriers order every write before the reads of the next phase.  STAGE: the
// island is held in LDS for the batch, at byte offset stage_off of the
// dynamic LDS (after the evaluation's own staging area), read at the start
// and written back at the end; the arithmetic is the same.
//
// With a second workgroup (gridDim.x == 2, an island whose generator is one
// segment), that workgroup hands out the next batch's words (g1 ..) into
// words_next, as k_rng's phase 1 would after this launch: the two workgroups
// share no data, and the next launch reads those words.
template <bool STAGE>
__global__ __launch_bounds__(kBlock) void k_gens_one_tile(const island_desc *__restrict__ d,
                                                          const tile *__restrict__ tiles, uint32_t island,
                                                          uint32_t tl_index, const uint32_t *__restrict__ words,
                                                          double *V, eval_ctx c, uint32_t g0, uint32_t g1,
                                                          uint32_t stage_off, uint32_t *__restrict__ words_next,
                                                          mt_state *__restrict__ st, mt_state *__restrict__ segst)
{
    extern __shared__ double lds[];
    const tile tl = tiles[tl_index];
    const island_desc di = d[island];
    const uint64_t elems = static_cast<uint64_t>(di.pop) * di.nx;
    if (blockIdx.x == 1u) {
        // batch g1 / gps of a one-segment generator (k_rng_gen's work; the
        // velocity draws shift the whole call's words alike)
        __shared__ uint32_t ring[kMtRing];
        rng_gen_segment(di, 0u, g1 / di.gps, 0u, &st[island], &segst[di.seg0], words_next + di.woff, ring);
        return;
    }
    const uint32_t end = min(g1, di.gens);
    if (g0 >= end) return;
    island_desc dl = di;
    eval_ctx cl = c;
    double *Vl = V;
    double *sX = nullptr, *sV = nullptr, *sL = nullptr, *sF = nullptr, *sLf = nullptr;
    if constexpr (STAGE) {
        sX = lds + stage_off / 8u;
        sV = sX + elems;
        sL = sV + elems;
        sF = sL + elems;
        sLf = sF + di.pop;
        for (uint64_t m = threadIdx.x; m < elems; m += kBlock) {
            sX[m] = c.X[di.xoff + m];
            sV[m] = V[di.xoff + m];
            sL[m] = c.lbX[di.xoff + m];
        }
        for (uint32_t p = threadIdx.x; p < di.pop; p += kBlock) {
            sF[p] = c.fit[di.foff + p];
            sLf[p] = c.lbfit[di.foff + p];
        }
        __syncthreads();
        dl.xoff = 0u;
        dl.foff = 0u;
        dl.toff = 0u;
        cl.X = sX;
        cl.lbX = sL;
        cl.fit = sF;
        cl.lbfit = sLf;
        cl.trace = c.trace ? c.trace + di.toff : nullptr;
        Vl = sV;
    }
    for (uint32_t g = g0; g < end; ++g) {
        for (uint64_t m = threadIdx.x; m < elems; m += kBlock) update_elem(dl, m, words, cl.X, Vl, cl.lbX, cl.lbfit, g);
        __syncthreads();
        eval_ctx cg = cl;
        cg.gen = g;
        eval_tile(dl, tl, cg, lds);
        __syncthreads();
    }
    if constexpr (STAGE) {
        for (uint64_t m = threadIdx.x; m < elems; m += kBlock) {
            c.X[di.xoff + m] = sX[m];
            V[di.xoff + m] = sV[m];
            c.lbX[di.xoff + m] = sL[m];
        }
        for (uint32_t p = threadIdx.x; p < di.pop; p += kBlock) {
            c.fit[di.foff + p] = sF[p];
            c.lbfit[di.foff + p] = sLf[p];
        }
    }
}


uint64_t align_up(uint64_t v, uint64_t a)
{
    return (v + a - 1u) / a * a;
}

// A buffer inside an archipelago's arena: its byte offset while the layout
// is planned, its device address afterwards.
struct arena_view {
    void *ptr = nullptr;
    uint64_t off = 0, bytes = 0;
};

} // namespace
} // namespace pgh

using namespace pgh;

struct pgh_archi {
    std::vector<island_desc> h_desc;
    std::vector<tile> h_tiles;
    std::vector<uint32_t> tile_begin; // per island, into h_tiles (+ end sentinel)
    uint64_t total_elems = 0, total_words = 0, total_trace = 0;
    uint32_t max_gens = 0;
    size_t lds_bytes = 0;
    bool trace = false;
    bool v_valid = false;
    bool evolved = false; // the population lives in lbX/lbfit once evolved
    std::vector<uint32_t> memory;
    // one device allocation (arena) holds every buffer below but trace_buf;
    // the small tables (desc, tiles, segs, jump_table, states) come first
    // and go up in one copy
    dev_buf arena, trace_buf;
    arena_view desc, tiles, segs, jump_table, states, X, V, lbX, fit, lbfit, words, seg_states;
    // the gbest islands (their indices, in order) and their best positions
    arena_view gblist, gbest;
    std::vector<uint32_t> h_gb;
    bool general = false; // an island other than variant 5 with lbest: the step kernels with that path
    // k_step: the second memory buffers, the selectors and one queue counter
    // per launch (in the arena); its items (per island, contiguous) and their
    // dispatch orders in a table of their own, rebuilt with the mapping
    arena_view lbX1, lbfitB, selA, selB, queues;
    uint64_t queue_stride = 0; // counters per generation: one per island, then one per step class
    dev_buf step_tab;
    const tile *d_items = nullptr;
    const uint32_t *d_order_all = nullptr, *d_order_cls = nullptr;
    std::vector<tile> h_items;
    std::vector<uint32_t> item_begin;       // per island, into h_items (+ end sentinel)
    uint32_t class_begin[4] = {0, 0, 0, 0}; // per step class, into the class order (+ end)
    size_t lj_lds = 0;                      // LDS of the Lennard-Jones items
    uint32_t n_cus = 0;
    pgh_archi_mapping map{PGH_ARCHI_STEP_ONE, 0u, 0u, 0u}; // the 64-island workload's fastest form
    uint32_t step_per_cu[4] = {1u, 1u, 1u, 1u}; // k_step workgroups per CU, per class
    uint32_t n_segs = 0;
    jump_tab jt{nullptr, nullptr};
    hipStream_t own = nullptr;
    hipStream_t side = nullptr; // a batched evolve's random words for the next batch
    hipEvent_t ev_main = nullptr, ev_side = nullptr;
    hipStream_t last = nullptr;
    hipEvent_t order = nullptr; // orders a call after the previous call's stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_v_valid = false, graph_from_lb = false;
    migr_ctx *migr = nullptr; // set by pgh_archi_set_migration
    bool one_shot = false;    // pgh_pso_gen_evolve_host's own archipelago
    uint64_t words_half = 0;  // offset of a lone island's second batch of words (0: none)

    ~pgh_archi()
    {
        if (last) (void)hipStreamSynchronize(last);
        if (migr) migr_destroy(migr);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (order) (void)hipEventDestroy(order);
        if (own) (void)hipStreamDestroy(own);
        if (side) (void)hipStreamDestroy(side);
        if (ev_main) (void)hipEventDestroy(ev_main);
        if (ev_side) (void)hipEventDestroy(ev_side);
    }
};

namespace
{

eval_ctx make_ctx(pgh_archi *a, uint32_t gen)
{
    eval_ctx c;
    c.X = static_cast<double *>(a->X.ptr);
    c.lbX = static_cast<double *>(a->lbX.ptr);
    c.fit = static_cast<double *>(a->fit.ptr);
    c.lbfit = static_cast<double *>(a->lbfit.ptr);
    c.trace = a->trace ? static_cast<double *>(a->trace_buf.ptr) : nullptr;
    c.gen = gen;
    return c;
}

uint32_t blocks_for(uint64_t count)
{
    return static_cast<uint32_t>((count + kBlock - 1u) / kBlock);
}

// The items of every island for the mapping a->map: 32 rows for rosenbrock
// (8 per wavefront), one L-lane group per row for the separable problems
// (256 / L rows), and for lennard_jones as many individuals as fit the
// mapping's LDS (at most one group of L lanes each per workgroup, at least
// one individual).  Batched launches hand them out by decreasing estimated
// cost, over all items (one launch) or within each problem class (a launch
// per class).  The cost only orders the items; no result depends on it.
void step_occupancy(pgh_archi *a);

pgh_status build_items(pgh_archi *a)
{
    const uint64_t budget = a->map.lj_lds ? a->map.lj_lds : 65536u;
    a->h_items.clear();
    a->item_begin.assign(1u, 0u);
    a->lj_lds = 0u;
    std::vector<double> cost;
    std::vector<uint32_t> cls_of;
    for (size_t k = 0; k < a->h_desc.size(); ++k) {
        const island_desc &di = a->h_desc[k];
        const uint32_t cls = step_class(di.kind);
        uint32_t per = 0u;
        double unit = 0.; // per row, in units of one element's move
        if (cls == kStepRosen) {
            // 32 rows (one per lane of the evaluation's 4 x 8), fewer for long
            // rows so that an item's move stays near 4096 elements: the
            // largest item bounds the generation
            per = di.nx > 256u ? kStepRosenRows / 4u : (di.nx > 128u ? kStepRosenRows / 2u : kStepRosenRows);
            unit = 1.2 * di.nx;
        } else if (cls == kStepLj) {
            const uint64_t bytes = static_cast<uint64_t>(lj_stride(di.dim, false)) * sizeof(double);
            per = static_cast<uint32_t>(std::min<uint64_t>(kBlock / di.lanes, budget / bytes));
            per = std::max(per, 1u);
            a->lj_lds = std::max<size_t>(a->lj_lds, per * bytes);
            unit = di.nx + 0.25 * di.dim * (di.dim - 1u);
        } else {
            per = kBlock / di.lanes;
            unit = 1.3 * di.nx;
        }
        // an item's elements are counted in 32 bits (step_move)
        per = std::max<uint32_t>(1u, std::min<uint64_t>(per, UINT32_MAX / std::max(di.nx, 1u)));
        for (uint32_t f = 0; f < di.pop; f += per) {
            const uint32_t c = std::min(per, di.pop - f);
            a->h_items.push_back(tile{static_cast<uint32_t>(k), f, c, 0u});
            cost.push_back(unit * c);
            cls_of.push_back(cls);
        }
        a->item_begin.push_back(static_cast<uint32_t>(a->h_items.size()));
    }
    const size_t n = a->h_items.size();
    std::vector<uint32_t> all(n), by_cls;
    for (size_t i = 0; i < n; ++i) all[i] = static_cast<uint32_t>(i);
    std::stable_sort(all.begin(), all.end(), [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
    for (uint32_t c = 0; c < 3u; ++c) {
        a->class_begin[c] = static_cast<uint32_t>(by_cls.size());
        for (uint32_t i : all) {
            if (cls_of[i] == c) by_cls.push_back(i);
        }
    }
    a->class_begin[3] = static_cast<uint32_t>(by_cls.size());
    // the table: items, then the two orders
    const size_t ib = n * sizeof(tile), ob = n * sizeof(uint32_t);
    std::vector<uint8_t> tab(ib + 2u * ob + 16u, 0u);
    std::memcpy(tab.data(), a->h_items.data(), ib);
    std::memcpy(tab.data() + ib, all.data(), ob);
    std::memcpy(tab.data() + ib + ob, by_cls.data(), ob);
    PGH_TRY(a->step_tab.reserve(tab.size()));
    PGH_HIP(hipMemcpy(a->step_tab.ptr, tab.data(), tab.size(), hipMemcpyHostToDevice));
    auto *base = static_cast<uint8_t *>(a->step_tab.ptr);
    a->d_items = reinterpret_cast<const tile *>(base);
    a->d_order_all = reinterpret_cast<const uint32_t *>(base + ib);
    a->d_order_cls = reinterpret_cast<const uint32_t *>(base + ib + ob);
    step_occupancy(a);
    return PGH_OK;
}

size_t step_lds(const pgh_archi *a, uint32_t cls);

// k_step workgroups per CU for each class at the current items' LDS: as many
// as can be resident, or the mapping's count.
void step_occupancy(pgh_archi *a)
{
    const void *fn[4] = {a->general ? reinterpret_cast<const void *>(&k_step<kStepRosen, true>)
                                    : reinterpret_cast<const void *>(&k_step<kStepRosen, false>),
                         a->general ? reinterpret_cast<const void *>(&k_step<kStepSep, true>)
                                    : reinterpret_cast<const void *>(&k_step<kStepSep, false>),
                         a->general ? reinterpret_cast<const void *>(&k_step<kStepLj, true>)
                                    : reinterpret_cast<const void *>(&k_step<kStepLj, false>),
                         a->general ? reinterpret_cast<const void *>(&k_step<kStepAll, true>)
                                    : reinterpret_cast<const void *>(&k_step<kStepAll, false>)};
    for (uint32_t c = 0; c < 4u; ++c) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn[c], static_cast<int>(kBlock), step_lds(a, c)) !=
                hipSuccess ||
            per_cu < 1) {
            (void)hipGetLastError();
            per_cu = 1;
        }
        if (a->map.blocks_per_cu) per_cu = std::min<int>(per_cu, static_cast<int>(a->map.blocks_per_cu));
        a->step_per_cu[c] = static_cast<uint32_t>(per_cu);
    }
}

size_t step_lds(const pgh_archi *a, uint32_t cls)
{
    switch (cls) {
        case kStepRosen:
            return kStepRosenLds;
        case kStepSep:
            return 0u;
        case kStepLj:
            return a->lj_lds;
        default:
            return std::max(kStepRosenLds, a->lj_lds);
    }
}

// Launches k_step<cls> over n_items items as a persistent grid: as many
// workgroups as can be resident (or the mapping's count per CU), at most one
// per item.
pgh_status launch_step(const pgh_archi *a, uint32_t cls, const step_args &args, hipStream_t s)
{
    const size_t lds = step_lds(a, cls);
    const uint32_t grid = std::min<uint32_t>(args.n_items, a->step_per_cu[cls] * a->n_cus);
    if (grid == 0u) return PGH_OK;
    if (a->general) {
        switch (cls) {
            case kStepRosen:
                k_step<kStepRosen, true><<<grid, kBlock, lds, s>>>(args);
                break;
            case kStepSep:
                k_step<kStepSep, true><<<grid, kBlock, lds, s>>>(args);
                break;
            case kStepLj:
                k_step<kStepLj, true><<<grid, kBlock, lds, s>>>(args);
                break;
            default:
                k_step<kStepAll, true><<<grid, kBlock, lds, s>>>(args);
                break;
        }
    } else {
        switch (cls) {
            case kStepRosen:
                k_step<kStepRosen, false><<<grid, kBlock, lds, s>>>(args);
                break;
            case kStepSep:
                k_step<kStepSep, false><<<grid, kBlock, lds, s>>>(args);
                break;
            case kStepLj:
                k_step<kStepLj, false><<<grid, kBlock, lds, s>>>(args);
                break;
            default:
                k_step<kStepAll, false><<<grid, kBlock, lds, s>>>(args);
                break;
        }
    }
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

// Launches the evolve call for islands [i0, i1) on stream s.  v_valid: the
// islands have velocities from an earlier call (kept by those with memory).
pgh_status launch_range(pgh_archi *a, uint32_t i0, uint32_t i1, bool v_valid, bool from_lb, hipStream_t s)
{
    const auto *d = static_cast<const island_desc *>(a->desc.ptr);
    auto *st = static_cast<mt_state *>(a->states.ptr);
    auto *w = static_cast<uint32_t *>(a->words.ptr);
    auto *X = static_cast<double *>(a->X.ptr);
    auto *V = static_cast<double *>(a->V.ptr);
    auto *lbX = static_cast<double *>(a->lbX.ptr);
    auto *fit = static_cast<double *>(a->fit.ptr);
    auto *lbfit = static_cast<double *>(a->lbfit.ptr);
    const uint64_t e0 = a->h_desc[i0].eoff;
    const uint64_t e1 = i1 < a->h_desc.size() ? a->h_desc[i1].eoff : a->total_elems;
    const uint64_t total = e1 - e0;
    const uint32_t t0 = a->tile_begin[i0], nt = a->tile_begin[i1] - t0;
    uint32_t gens = 0;
    for (uint32_t k = i0; k < i1; ++k) gens = std::max(gens, a->h_desc[k].gens);
    const uint32_t sg0 = a->h_desc[i0].seg0;
    const uint32_t ng = (i1 < a->h_desc.size() ? a->h_desc[i1].seg0 : a->n_segs) - sg0;
    const auto *sd = static_cast<const seg_desc *>(a->segs.ptr);
    auto *segst = static_cast<mt_state *>(a->seg_states.ptr);
    const jump_tab jt = a->jt;
    const uint32_t vv = v_valid ? 1u : 0u;
    k_rng<<<ng, kMtJumpBlock, 0, s>>>(d, sd, st, segst, jt, w, sg0, 0u, 0u, vv);
    PGH_HIP(hipGetLastError());
    auto *selA = static_cast<uint32_t *>(a->selA.ptr);
    k_vinit<<<blocks_for(total), kBlock, 0, s>>>(d, i0, i1, e0, total, w, X, V, lbX, fit, lbfit, selA,
                                                 v_valid ? 1u : 0u, from_lb ? 1u : 0u);
    PGH_HIP(hipGetLastError());
    // a single island of one tile, evolved by the one-shot entry point: a
    // batch of generations per launch (k_gens_one_tile)
    const bool fused = a->one_shot && i1 == i0 + 1u && nt == 1u && a->h_desc[i0].variant == 5u &&
                       a->h_desc[i0].topo == 2u;
    // the gbest islands of the range: the swarm's best at the call's start
    const auto gb_lo = std::lower_bound(a->h_gb.begin(), a->h_gb.end(), i0);
    const auto gb_hi = std::lower_bound(a->h_gb.begin(), a->h_gb.end(), i1);
    const uint32_t n_gb = static_cast<uint32_t>(gb_hi - gb_lo);
    const uint32_t *gbl = static_cast<const uint32_t *>(a->gblist.ptr) + (gb_lo - a->h_gb.begin());
    auto *gbest = static_cast<double *>(a->gbest.ptr);
    if (n_gb) {
        k_gbest_init<<<n_gb, kBlock, 0, s>>>(d, gbl, X, fit, gbest);
        PGH_HIP(hipGetLastError());
    }
    // the island in LDS when it fits beside the evaluation's staging area
    // within 64 KiB: X, V, lbX (pop x nx) and fit, lbfit (pop)
    const uint64_t stage_off = align_up(a->lds_bytes, 256u);
    uint64_t stage_bytes = 0u;
    if (fused) {
        const island_desc &di = a->h_desc[i0];
        const uint64_t need = (3u * static_cast<uint64_t>(di.pop) * di.nx + 2u * di.pop) * 8u;
        if (stage_off + need <= 65536u) stage_bytes = need;
    }
    const auto *tl = static_cast<const tile *>(a->tiles.ptr);
    auto *lbX1 = static_cast<double *>(a->lbX1.ptr);
    auto *lbfitB = static_cast<double *>(a->lbfitB.ptr);
    auto *selB = static_cast<uint32_t *>(a->selB.ptr);
    auto *queues = static_cast<uint32_t *>(a->queues.ptr);
    const uint32_t n_isl = static_cast<uint32_t>(a->h_desc.size());
    // ... and the next batch's words handed out beside it, into the other
    // half of the word buffer, when the generator is one segment
    const bool overlap = fused && a->words_half != 0u && a->h_desc[i0].segs == 1u;
    // several islands at once: the next batch's words are made on the side
    // stream while the current batch's generations run (each island with more
    // than one batch alternates between the two halves of its word buffer)
    const bool side_rng = !fused && i1 > i0 + 1u;
    for (uint32_t g = 0; g < gens; ++g) {
        if (g % kGenBatch == 0u) {
            const uint32_t b = g / kGenBatch;
            if (side_rng) {
                if (b == 0u) {
                    k_rng_gen<<<ng, kMtGenBlock, 0, s>>>(d, sd, st, segst, w, sg0, 0u, vv);
                    PGH_HIP(hipGetLastError());
                } else {
                    PGH_HIP(hipStreamWaitEvent(s, a->ev_side, 0)); // batch b's words
                }
                if (g + kGenBatch < gens) {
                    // batch b - 1's generations are done: its half may take
                    // batch b + 1's words
                    PGH_HIP(hipEventRecord(a->ev_main, s));
                    PGH_HIP(hipStreamWaitEvent(a->side, a->ev_main, 0));
                    k_rng<<<ng, kMtJumpBlock, 0, a->side>>>(d, sd, st, segst, jt, w, sg0, 2u, b, vv);
                    k_rng_gen<<<ng, kMtGenBlock, 0, a->side>>>(d, sd, st, segst, w, sg0, b + 1u, vv);
                    PGH_HIP(hipGetLastError());
                    PGH_HIP(hipEventRecord(a->ev_side, a->side));
                }
            } else if (!overlap || g == 0u) {
                if (b > 0u) k_rng<<<ng, kMtJumpBlock, 0, s>>>(d, sd, st, segst, jt, w, sg0, 2u, b - 1u, vv);
                k_rng_gen<<<ng, kMtGenBlock, 0, s>>>(d, sd, st, segst, w, sg0, b, vv);
                PGH_HIP(hipGetLastError());
            }
            if (fused) {
                const uint32_t *wc = w + (overlap ? (b % 2u) * a->words_half : 0u);
                uint32_t *wn = w + (overlap ? ((b + 1u) % 2u) * a->words_half : 0u);
                const uint32_t grid = overlap && g + kGenBatch < gens ? 2u : 1u;
                if (stage_bytes) {
                    k_gens_one_tile<true><<<grid, kBlock, stage_off + stage_bytes, s>>>(
                        d, tl, i0, t0, wc, V, make_ctx(a, g), g, g + kGenBatch, stage_off, wn, st, segst);
                } else {
                    k_gens_one_tile<false><<<grid, kBlock, a->lds_bytes, s>>>(d, tl, i0, t0, wc, V, make_ctx(a, g), g,
                                                                              g + kGenBatch, 0u, wn, st, segst);
                }
                PGH_HIP(hipGetLastError());
            }
        }
        if (fused) continue;
        if (g == 0u) {
            // this call's queue counters
            const uint64_t nq = static_cast<uint64_t>(a->max_gens) * a->queue_stride;
            k_zero_u32<<<blocks_for(nq), kBlock, 0, s>>>(queues, nq);
            PGH_HIP(hipGetLastError());
        }
        step_args args{};
        args.d = d;
        args.items = a->d_items;
        args.words = w;
        args.X = X;
        args.V = V;
        args.lbX0 = lbX;
        args.lbX1 = lbX1;
        args.fit = fit;
        args.trace = a->trace ? static_cast<double *>(a->trace_buf.ptr) : nullptr;
        // the memory fitness and the selectors alternate between their buffers
        args.lbf_cur = (g % 2u) ? lbfitB : lbfit;
        args.lbf_nxt = (g % 2u) ? lbfit : lbfitB;
        args.sel_cur = (g % 2u) ? selB : selA;
        args.sel_nxt = (g % 2u) ? selA : selB;
        args.gbest = gbest;
        args.gen = g;
        uint32_t *qg = queues + static_cast<uint64_t>(g) * a->queue_stride;
        if (i1 == i0 + 1u) {
            // one island: its own items, in order
            args.base = a->item_begin[i0];
            args.n_items = a->item_begin[i1] - a->item_begin[i0];
            args.queue = qg + i0;
            PGH_TRY(launch_step(a, step_class(a->h_desc[i0].kind), args, s));
        } else if (a->map.form == PGH_ARCHI_STEP_ONE) {
            args.order = a->d_order_all + a->item_begin[i0];
            args.n_items = a->item_begin[i1] - a->item_begin[i0];
            args.queue = qg + n_isl + kStepAll;
            PGH_TRY(launch_step(a, kStepAll, args, s));
        } else {
            for (uint32_t c = 0; c < 3u; ++c) {
                args.order = a->d_order_cls + a->class_begin[c];
                args.n_items = a->class_begin[c + 1u] - a->class_begin[c];
                args.queue = qg + n_isl + c;
                PGH_TRY(launch_step(a, c, args, s));
            }
        }
        if (n_gb) {
            // gbest after the generation's memory update
            k_gbest_update<<<n_gb, kBlock, 0, s>>>(d, gbl, X, fit, args.sel_cur, args.sel_nxt, gbest, g);
            PGH_HIP(hipGetLastError());
        }
    }
    if (!fused) {
        k_settle<<<blocks_for(total), kBlock, 0, s>>>(d, i0, i1, e0, total, lbX, lbX1, lbfit, lbfitB, selA, selB);
        PGH_HIP(hipGetLastError());
    }
    return PGH_OK;
}

// Random words of one generation (two per draw, pso_gen.cpp:241-327): per
// element two draws (variants 1 and 5), one (2) or one per neighbour (6, K
// neighbours); per particle two (3) or one (4).  wpu: words per unit.
uint64_t words_per_gen(uint32_t variant, uint32_t K, uint64_t pop, uint64_t nx, uint32_t &wpu)
{
    switch (variant) {
        case 2u:
            wpu = 2u;
            return 2u * pop * nx;
        case 3u:
            wpu = 4u;
            return 4u * pop;
        case 4u:
            wpu = 2u;
            return 2u * pop;
        case 6u:
            wpu = 2u * K;
            return 2u * static_cast<uint64_t>(K) * pop * nx;
        default:
            wpu = 4u;
            return 4u * pop * nx;
    }
}

// Columns of the von Neumann lattice (pso_gen.cpp:719-746): rows is the
// largest divisor of the swarm size not above its square root.
uint32_t von_cols(uint32_t pop)
{
    uint32_t rows = static_cast<uint32_t>(std::sqrt(static_cast<double>(pop)));
    while (rows > 1u && pop % rows != 0u) --rows;
    rows = std::max(rows, 1u);
    return pop / rows;
}

// Initial populations from host arrays (X0, F0) or, with `seeds`, drawn and
// evaluated on the device.
pgh_status archi_create(pgh_archi **out, const pgh_island_spec *islands, uint64_t n_islands,
                        const double *const *X0, const double *const *F0, const uint32_t *seeds)
{
    if (!out || !islands || (!seeds && (!X0 || !F0))) return fail(PGH_INVALID_ARGUMENT, "null argument");
    if (n_islands == 0u) return fail(PGH_INVALID_ARGUMENT, "an archipelago needs at least one island");
    auto a = std::make_unique<pgh_archi>();
    std::vector<prob> probs;
    uint64_t xoff = 0, foff = 0, woff = 0, toff = 0, eoff = 0, gboff = 0;
    std::vector<seg_desc> seg_list;
    std::vector<uint64_t> dists;
    {
        int dev = 0, cus = 0;
        PGH_HIP(hipGetDevice(&dev));
        PGH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        a->n_cus = static_cast<uint32_t>(std::max(cus, 1));
    }
    // The generators' share per segment: about two shares per CU over all
    // islands' batches (a power of two in [kSegMin, kSegMax]), so that the
    // batch's words are made by workgroups spread over the device while the
    // jumps that position the segments (each costs about as much as making
    // a few hundred thousand words) stay few.  The jumps z^(s seg_words) that
    // move segment s to its first share are the same for every island and
    // come first in the jump table (one multiplication each on the host);
    // other distances are entered once each.
    uint64_t seg_words = kSegMin, max_segs = 1u;
    {
        uint64_t total = 0;
        std::vector<uint64_t> bw(n_islands, 0u);
        for (uint64_t k = 0; k < n_islands; ++k) {
            prob p;
            PGH_TRY(resolve(&islands[k].udp, p));
            const uint64_t gps = std::min<uint64_t>(kGenBatch, std::max<uint64_t>(islands[k].algo.gen, 1u));
            const pgh_pso_params &q = islands[k].algo;
            const uint32_t K = q.neighb_type == 3u ? 4u : 2u * (q.neighb_param / 2u);
            uint32_t wpu = 0u;
            bw[k] = gps * words_per_gen(q.variant, K, islands[k].pop, p.nx, wpu);
            total += bw[k];
        }
        const uint64_t want = total / (2u * a->n_cus);
        while (seg_words * 2u <= want && seg_words < kSegMax) seg_words *= 2u;
        // a lone island of a small batch stays one generator (the one-shot
        // evolve hands out its next batch beside the current one)
        if (n_islands == 1u && bw[0] <= kSoloWords) seg_words = std::max(seg_words, bw[0]);
        for (uint64_t k = 0; k < n_islands; ++k) {
            if (bw[k] > seg_words) max_segs = std::max<uint64_t>(max_segs, (bw[k] + seg_words - 1u) / seg_words);
        }
    }
    for (uint64_t k = 1; k < max_segs; ++k) dists.push_back(k * seg_words);
    PGH_TRY(jump_polys_multiples(seg_words, max_segs - 1u));
    const auto dist_index = [&dists](uint64_t J) {
        for (size_t i = 0; i < dists.size(); ++i) {
            if (dists[i] == J) return static_cast<uint32_t>(i);
        }
        dists.push_back(J);
        return static_cast<uint32_t>(dists.size() - 1u);
    };
    a->tile_begin.push_back(0);
    for (uint64_t k = 0; k < n_islands; ++k) {
        const pgh_island_spec &sp = islands[k];
        prob p;
        PGH_TRY(resolve(&sp.udp, p));
        if (p.nf != 1u) return fail(PGH_UNSUPPORTED, "pso_gen needs a single-objective problem");
        if (p.shift) return fail(PGH_UNSUPPORTED, "the archipelago does not take translated problems");
        if (p.kind != PGH_ROSENBROCK && p.kind != PGH_RASTRIGIN && p.kind != PGH_ACKLEY && p.kind != PGH_SCHWEFEL &&
            p.kind != PGH_GRIEWANK && p.kind != PGH_LENNARD_JONES) {
            return fail(PGH_UNSUPPORTED, "the device archipelago evolves rosenbrock, rastrigin, ackley, schwefel, "
                                         "griewank and lennard_jones islands");
        }
        probs.push_back(p);
        const pgh_pso_params &q = sp.algo;
        // pso_gen.cpp:83-112 parameter checks
        if (q.omega < 0. || q.omega > 1.) {
            return fail(PGH_INVALID_ARGUMENT, "The particles' inertia (or the constriction factor) must be in the "
                                              "[0,1] range");
        }
        if (q.eta1 < 0. || q.eta2 < 0. || q.eta1 > 4. || q.eta2 > 4.) {
            return fail(PGH_INVALID_ARGUMENT, "The eta parameters must be in the [0,4] range");
        }
        if (q.max_vel <= 0. || q.max_vel > 1.) {
            return fail(PGH_INVALID_ARGUMENT, "The maximum particle velocity (as a fraction of the bounds) should be "
                                              "in the (0,1] range");
        }
        if (q.variant < 1u || q.variant > 6u) {
            return fail(PGH_INVALID_ARGUMENT, "The PSO variant must be in [1,6], while a value of "
                                                  + std::to_string(q.variant) + " was detected");
        }
        if (q.neighb_type < 1u || q.neighb_type > 4u) {
            return fail(PGH_INVALID_ARGUMENT, "The swarm topology variant must be in [1,4], while a value of "
                                                  + std::to_string(q.neighb_type) + " was detected");
        }
        if (q.neighb_param < 1u) {
            return fail(PGH_INVALID_ARGUMENT, "The neighborhood parameter must be in (0, inf), while a value of "
                                                  + std::to_string(q.neighb_param) + " was detected");
        }
        if (q.neighb_type == 4u) {
            return fail(PGH_UNSUPPORTED, "the device pso_gen does not implement the adaptive random topology (its "
                                         "draws depend on each generation's results)");
        }
        if (q.variant == 6u && q.neighb_type == 1u) {
            return fail(PGH_UNSUPPORTED, "the device pso_gen does not implement FIPS (variant 6) with the gbest "
                                         "topology");
        }
        if (sp.pop == 0u) return fail(PGH_INVALID_ARGUMENT, "pso_gen does not work on an empty population");
        const uint32_t radius = q.neighb_param / 2u;
        if (sp.pop > (uint64_t(1) << 24) || (q.neighb_type == 2u && (radius == 0u || sp.pop < radius))) {
            return fail(PGH_UNSUPPORTED, "population size incompatible with the lbest radius");
        }
        island_desc di{};
        di.kind = p.kind;
        di.dim = p.dim;
        di.nx = p.nx;
        di.pop = static_cast<uint32_t>(sp.pop);
        di.gens = q.gen;
        di.radius = radius;
        di.omega = q.omega;
        di.eta1 = q.eta1;
        di.eta2 = q.eta2;
        di.max_vel = q.max_vel;
        di.memory = q.memory ? 1u : 0u;
        di.variant = q.variant;
        di.topo = q.neighb_type;
        if (q.variant != 5u || q.neighb_type != 2u) a->general = true;
        di.nbk = q.neighb_type == 3u ? 4u : 2u * radius;
        di.vcols = q.neighb_type == 3u ? von_cols(static_cast<uint32_t>(sp.pop)) : 1u;
        di.wgen = words_per_gen(di.variant, di.nbk, sp.pop, p.nx, di.wpu);
        if (di.topo == 1u) {
            di.goff = gboff;
            gboff += align_up(p.nx + 1u, 32u);
            a->h_gb.push_back(static_cast<uint32_t>(k));
        }
        const uint64_t elems = sp.pop * p.nx;
        di.xoff = xoff;
        di.foff = foff;
        di.woff = woff;
        di.toff = toff;
        di.eoff = eoff;
        xoff += align_up(elems, 32u); // 256-byte aligned rows blocks
        foff += align_up(sp.pop, 32u);
        // the words of a batch of generations.  Islands of more than one batch
        // in an archipelago of several get room for two, used in turn: a
        // batched evolve makes the next batch's words while the current batch's
        // generations run.  A lone island of up to 2^22 words per batch gets
        // room for a second batch too, which the one-shot evolve fills while it
        // evolves the first (k_gens_one_tile, its own alternation).
        const uint64_t wbatch = align_up(di.wgen * kGenBatch, 64u);
        const bool second = n_islands == 1u && wbatch <= (uint64_t(1) << 22);
        const bool alternate = n_islands > 1u && q.gen > kGenBatch;
        di.whalf = alternate ? wbatch : 0u;
        woff += (second || alternate) ? 2u * wbatch : wbatch;
        if (second) a->words_half = wbatch;
        toff += static_cast<uint64_t>(q.gen) * sp.pop;
        eoff += elems;
        a->max_gens = std::max(a->max_gens, q.gen);
        a->memory.push_back(q.memory ? 1u : 0u);
        // evaluation tiles
        uint32_t per = 256u;
        if (p.kind == PGH_LENNARD_JONES) {
            di.lanes = archi_lj_lanes(p.dim);
            per = std::min(lj_groups_for(p.dim, false), 256u / di.lanes);
            if (per == 0u) return fail(PGH_UNSUPPORTED, "lennard_jones cluster too large for one workgroup");
            a->lds_bytes =
                std::max<size_t>(a->lds_bytes, static_cast<size_t>(per) * lj_stride(p.dim, false) * sizeof(double));
        } else if (p.kind != PGH_ROSENBROCK) {
            di.lanes = static_cast<uint32_t>(sep_lanes(p.kind, p.nx));
            per = 256u / di.lanes;
        } else {
            di.lanes = 1u;
        }
        // the generator's segments (k_rng): batches of gps generations split in
        // shares of seg_words words (one share when the batch is smaller)
        {
            const uint64_t W = di.wgen;
            di.gps = std::min<uint32_t>(kGenBatch, std::max<uint32_t>(q.gen, 1u));
            di.bwords = static_cast<uint64_t>(di.gps) * W;
            di.seg0 = static_cast<uint32_t>(seg_list.size());
            di.swords = std::min(di.bwords, seg_words);
            di.segs = static_cast<uint32_t>((di.bwords + di.swords - 1u) / di.swords);
            di.j_n0 = di.j_pos = di.j_main = di.j_last = 0u;
            if (di.segs > 1u) {
                di.j_n0 = dist_index(2u * elems);
                di.j_pos = 0u; // the shared z^(s seg_words) block
                di.j_main = dist_index(di.bwords - di.swords);
                di.j_last = dist_index(static_cast<uint64_t>(di.segs - 1u) * di.swords);
            }
            for (uint32_t sgi = 0; sgi < di.segs; ++sgi) seg_list.push_back(seg_desc{static_cast<uint32_t>(k), sgi});
        }
        for (uint64_t f = 0; f < sp.pop; f += per) {
            a->h_tiles.push_back(tile{static_cast<uint32_t>(k), static_cast<uint32_t>(f),
                                      static_cast<uint32_t>(std::min<uint64_t>(per, sp.pop - f)), 0u});
        }
        a->tile_begin.push_back(static_cast<uint32_t>(a->h_tiles.size()));
        a->h_desc.push_back(di);
    }
    a->total_elems = eoff;
    a->total_words = woff;
    a->total_trace = toff;
    PGH_HIP(hipStreamCreateWithFlags(&a->own, hipStreamNonBlocking));
    PGH_HIP(hipStreamCreateWithFlags(&a->side, hipStreamNonBlocking));
    PGH_HIP(hipEventCreateWithFlags(&a->ev_main, hipEventDisableTiming));
    PGH_HIP(hipEventCreateWithFlags(&a->ev_side, hipEventDisableTiming));
    PGH_HIP(hipEventCreateWithFlags(&a->order, hipEventDisableTiming));
    a->n_segs = static_cast<uint32_t>(seg_list.size());
    std::vector<uint8_t> tab;
    PGH_TRY(jump_table_bytes(dists, tab));
    std::vector<mt_state> states(n_islands);
    for (uint64_t k = 0; k < n_islands; ++k) mt_seed_host(islands[k].algo.seed, states[k]);
    // arena layout, every buffer on a 256-byte boundary
    uint64_t arena_bytes = 0;
    auto place = [&](arena_view &v, uint64_t bytes) {
        v.off = arena_bytes;
        v.bytes = bytes;
        arena_bytes = align_up(arena_bytes + std::max<uint64_t>(bytes, 1u), 256u);
    };
    place(a->desc, a->h_desc.size() * sizeof(island_desc));
    place(a->tiles, a->h_tiles.size() * sizeof(tile));
    place(a->segs, seg_list.size() * sizeof(seg_desc));
    place(a->jump_table, std::max<size_t>(tab.size(), 16u));
    place(a->states, n_islands * sizeof(mt_state));
    place(a->gblist, a->h_gb.size() * sizeof(uint32_t));
    const uint64_t head_bytes = arena_bytes;
    place(a->X, xoff * 8u);
    place(a->V, xoff * 8u);
    place(a->lbX, xoff * 8u);
    place(a->fit, foff * 8u);
    place(a->lbfit, foff * 8u);
    place(a->gbest, gboff * 8u);
    place(a->words, std::max<uint64_t>(woff, 1u) * 4u);
    place(a->seg_states, seg_list.size() * sizeof(mt_state));
    place(a->lbX1, xoff * 8u);
    place(a->lbfitB, foff * 8u);
    place(a->selA, foff * 4u);
    place(a->selB, foff * 4u);
    a->queue_stride = n_islands + 4u;
    place(a->queues, std::max<uint64_t>(a->max_gens, 1u) * a->queue_stride * 4u);
    PGH_TRY(a->arena.reserve(arena_bytes));
    auto *base = static_cast<uint8_t *>(a->arena.ptr);
    for (arena_view *v : {&a->desc, &a->tiles, &a->segs, &a->jump_table, &a->states, &a->gblist, &a->X, &a->V,
                          &a->lbX, &a->fit, &a->lbfit, &a->gbest, &a->words, &a->seg_states, &a->lbX1, &a->lbfitB,
                          &a->selA, &a->selB, &a->queues}) {
        v->ptr = base + v->off;
    }
    PGH_TRY(build_items(a.get()));
    {
        std::vector<uint8_t> head(head_bytes, 0u);
        auto put = [&](const arena_view &v, const void *src, uint64_t bytes) {
            if (bytes) std::memcpy(head.data() + v.off, src, bytes);
        };
        put(a->desc, a->h_desc.data(), a->h_desc.size() * sizeof(island_desc));
        put(a->tiles, a->h_tiles.data(), a->h_tiles.size() * sizeof(tile));
        put(a->segs, seg_list.data(), seg_list.size() * sizeof(seg_desc));
        put(a->jump_table, tab.data(), tab.size());
        put(a->states, states.data(), n_islands * sizeof(mt_state));
        put(a->gblist, a->h_gb.data(), a->h_gb.size() * sizeof(uint32_t));
        PGH_HIP(hipMemcpy(base, head.data(), head_bytes, hipMemcpyHostToDevice));
    }
    a->jt.lens = static_cast<const uint32_t *>(a->jump_table.ptr);
    a->jt.lists = reinterpret_cast<const uint16_t *>(static_cast<const char *>(a->jump_table.ptr) +
                                                     jump_table_head(dists.size()));
    for (uint64_t k = 0; k < n_islands; ++k) {
        const island_desc &di = a->h_desc[k];
        double *Xk = static_cast<double *>(a->X.ptr) + di.xoff;
        double *Fk = static_cast<double *>(a->fit.ptr) + di.foff;
        if (seeds) {
            // population(problem, bfe, pop, seed), population.cpp:82-103:
            // batch_random_decision_vector from std::mt19937(seed), then one
            // batch evaluation, both on the device
            PGH_TRY(launch_batch_random_dv(probs[k], di.pop, seeds[k], Xk, a->own));
            PGH_TRY(launch_fitness(probs[k], Xk, di.pop, Fk, a->own));
        } else {
            PGH_HIP(hipMemcpy(Xk, X0[k], static_cast<uint64_t>(di.pop) * di.nx * 8u, hipMemcpyHostToDevice));
            PGH_HIP(hipMemcpy(Fk, F0[k], static_cast<uint64_t>(di.pop) * 8u, hipMemcpyHostToDevice));
        }
    }
    if (seeds) PGH_HIP(hipStreamSynchronize(a->own));
    *out = a.release();
    return PGH_OK;
}
} // namespace

extern "C" {

pgh_status pgh_archi_create(pgh_archi **out, const pgh_island_spec *islands, uint64_t n_islands,
                            const double *const *X0, const double *const *F0)
try {
    return archi_create(out, islands, n_islands, X0, F0, nullptr);
} catch (...) {
    return api_exception();
}

pgh_status pgh_archi_create_seeded(pgh_archi **out, const pgh_island_spec *islands, uint64_t n_islands,
                                   const uint32_t *pop_seeds)
try {
    if (!pop_seeds) return fail(PGH_INVALID_ARGUMENT, "null argument");
    return archi_create(out, islands, n_islands, nullptr, nullptr, pop_seeds);
} catch (...) {
    return api_exception();
}

pgh_status pgh_archi_destroy(pgh_archi *a)
try {
    delete a;
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_archi_set_trace(pgh_archi *a, int on)
try {
    if (!a) return fail(PGH_INVALID_ARGUMENT, "null archipelago");
    if (on && !a->trace_buf.ptr) PGH_TRY(a->trace_buf.reserve(std::max<uint64_t>(a->total_trace, 1u) * 8u));
    if (a->trace != (on != 0)) {
        // the captured graph bakes the trace pointer in
        if (a->exec) (void)hipGraphExecDestroy(a->exec);
        if (a->graph) (void)hipGraphDestroy(a->graph);
        a->exec = nullptr;
        a->graph = nullptr;
    }
    a->trace = on != 0;
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_archi_evolve(pgh_archi *a, int mode, void *stream)
try {
    if (!a) return fail(PGH_INVALID_ARGUMENT, "null archipelago");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : a->own;
    // a call on another stream than the previous one starts after it: the
    // archipelago's buffers are read and written by every call
    if (a->last && a->last != s) {
        PGH_HIP(hipEventRecord(a->order, a->last));
        PGH_HIP(hipStreamWaitEvent(s, a->order, 0));
    }
    // velocities are re-drawn unless an island's memory flag is set and they
    // exist (pso_gen.cpp:193-200); the kernels decide per island
    const bool v_valid = a->v_valid;
    bool from_lb = a->evolved;
    const uint32_t n = static_cast<uint32_t>(a->h_desc.size());
    auto *X = static_cast<double *>(a->X.ptr);
    auto *fit = static_cast<double *>(a->fit.ptr);
    auto *lbX = static_cast<double *>(a->lbX.ptr);
    auto *lbfit = static_cast<double *>(a->lbfit.ptr);
    if (a->migr) {
        // immigration (island.cpp:452-597) out of the buffers that hold the
        // population into the other pair, from which the evolve call starts
        bool moved = false;
        if (a->evolved) {
            PGH_TRY(migr_immigrate(a->migr, lbX, lbfit, X, fit, &moved, s));
        } else {
            PGH_TRY(migr_immigrate(a->migr, X, fit, lbX, lbfit, &moved, s));
        }
        if (moved) from_lb = !a->evolved;
    }
    if (mode == PGH_ARCHI_PER_ISLAND) {
        for (uint32_t k = 0; k < n; ++k) PGH_TRY(launch_range(a, k, k + 1u, v_valid, from_lb, s));
    } else if (mode == PGH_ARCHI_BATCHED) {
        if (!a->exec || a->graph_v_valid != v_valid || a->graph_from_lb != from_lb) {
            if (a->exec) (void)hipGraphExecDestroy(a->exec);
            if (a->graph) (void)hipGraphDestroy(a->graph);
            a->exec = nullptr;
            a->graph = nullptr;
            PGH_HIP(hipStreamBeginCapture(a->own, hipStreamCaptureModeThreadLocal));
            const pgh_status st = launch_range(a, 0u, n, v_valid, from_lb, a->own);
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(a->own, &g);
            if (st != PGH_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
            a->graph = g;
            PGH_HIP(hipGraphInstantiate(&a->exec, a->graph, nullptr, nullptr, 0));
            a->graph_v_valid = v_valid;
            a->graph_from_lb = from_lb;
        }
        PGH_HIP(hipGraphLaunch(a->exec, s));
    } else {
        return fail(PGH_INVALID_ARGUMENT, "unknown archipelago mode " + std::to_string(mode));
    }
    // emigration (island.cpp:599-620) from the evolved populations
    if (a->migr) PGH_TRY(migr_emigrate(a->migr, lbX, lbfit, s));
    a->v_valid = true;
    a->evolved = true;
    a->last = s;
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_archi_set_mapping(pgh_archi *a, const pgh_archi_mapping *m)
try {
    if (!a || !m) return fail(PGH_INVALID_ARGUMENT, "null argument");
    if (m->form != PGH_ARCHI_STEP_ONE && m->form != PGH_ARCHI_STEP_SPLIT) {
        return fail(PGH_INVALID_ARGUMENT, "unknown generation launch form " + std::to_string(m->form));
    }
    if (m->lj_lds > 65536u) return fail(PGH_INVALID_ARGUMENT, "lj_lds is at most 65536 bytes");
    if (m->blocks_per_cu > 64u) return fail(PGH_INVALID_ARGUMENT, "blocks_per_cu is at most 64");
    if (a->last) PGH_HIP(hipStreamSynchronize(a->last));
    a->map = *m;
    a->map.pad = 0u;
    // the captured graph bakes the items and launch shapes in
    if (a->exec) (void)hipGraphExecDestroy(a->exec);
    if (a->graph) (void)hipGraphDestroy(a->graph);
    a->exec = nullptr;
    a->graph = nullptr;
    return build_items(a);
} catch (...) {
    return api_exception();
}

pgh_status pgh_archi_get_mapping(pgh_archi *a, pgh_archi_mapping *m)
try {
    if (!a || !m) return fail(PGH_INVALID_ARGUMENT, "null argument");
    *m = a->map;
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_archi_tune(pgh_archi *a, uint32_t reps, pgh_archi_mapping *chosen)
try {
    if (!a) return fail(PGH_INVALID_ARGUMENT, "null archipelago");
    if (a->migr) return fail(PGH_UNSUPPORTED, "pgh_archi_tune: turn migration off first");
    reps = std::max(reps, 1u);
    if (a->last) PGH_HIP(hipStreamSynchronize(a->last));
    // the state an evolve call reads and writes: the generators, X, V, lbX,
    // fit and lbfit (contiguous in the arena), and the host flags
    auto *arena = static_cast<uint8_t *>(a->arena.ptr);
    const uint64_t off = a->states.off, bytes = a->words.off - a->states.off;
    const bool v_valid = a->v_valid, evolved = a->evolved, trace = a->trace;
    const pgh_archi_mapping before = a->map;
    dev_buf snap, ref, flag;
    PGH_TRY(snap.reserve(bytes));
    PGH_TRY(ref.reserve(bytes));
    PGH_TRY(flag.reserve(4u));
    PGH_HIP(hipMemcpyAsync(snap.ptr, arena + off, bytes, hipMemcpyDeviceToDevice, a->own));
    hipEvent_t t0 = nullptr, t1 = nullptr;
    PGH_HIP(hipEventCreate(&t0));
    PGH_HIP(hipEventCreate(&t1));
    const auto release = [&]() {
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
    };
    const auto restore = [&]() -> pgh_status {
        PGH_HIP(hipMemcpyAsync(arena + off, snap.ptr, bytes, hipMemcpyDeviceToDevice, a->own));
        a->v_valid = v_valid;
        a->evolved = evolved;
        return PGH_OK;
    };
    a->trace = false;
    std::vector<pgh_archi_mapping> cand;
    for (uint32_t form : {static_cast<uint32_t>(PGH_ARCHI_STEP_SPLIT), static_cast<uint32_t>(PGH_ARCHI_STEP_ONE)}) {
        for (uint32_t lds : {65536u, 32768u, 16384u}) {
            for (uint32_t bpc : {0u, 2u}) cand.push_back(pgh_archi_mapping{form, lds, bpc, 0u});
        }
    }
    pgh_status st = PGH_OK;
    float best_ms = 0.f;
    pgh_archi_mapping best = cand[0];
    for (size_t c = 0; c < cand.size() && st == PGH_OK; ++c) {
        st = pgh_archi_set_mapping(a, &cand[c]);
        float ms = 0.f;
        for (uint32_t r = 0; r <= reps && st == PGH_OK; ++r) {
            // the first run captures the graph; the others are timed
            st = restore();
            if (st != PGH_OK) break;
            if (r > 0u && (st = hip_status(hipEventRecord(t0, a->own), "hipEventRecord")) != PGH_OK) break;
            st = pgh_archi_evolve(a, PGH_ARCHI_BATCHED, nullptr);
            if (st != PGH_OK) break;
            if (r > 0u) {
                if ((st = hip_status(hipEventRecord(t1, a->own), "hipEventRecord")) != PGH_OK) break;
                if ((st = hip_status(hipEventSynchronize(t1), "hipEventSynchronize")) != PGH_OK) break;
                float e = 0.f;
                if ((st = hip_status(hipEventElapsedTime(&e, t0, t1), "hipEventElapsedTime")) != PGH_OK) break;
                ms = r == 1u ? e : std::min(ms, e);
            }
        }
        if (st != PGH_OK) break;
        if (c == 0u) {
            st = hip_status(hipMemcpyAsync(ref.ptr, arena + off, bytes, hipMemcpyDeviceToDevice, a->own), "copy");
        } else {
            uint32_t differs = 0u;
            k_zero_u32<<<1, kBlock, 0, a->own>>>(static_cast<uint32_t *>(flag.ptr), 1u);
            k_differs<<<blocks_for(bytes / 4u), kBlock, 0, a->own>>>(static_cast<const uint32_t *>(ref.ptr),
                                                                      reinterpret_cast<const uint32_t *>(arena + off),
                                                                      bytes / 4u, static_cast<uint32_t *>(flag.ptr));
            if ((st = hip_status(hipMemcpyAsync(&differs, flag.ptr, 4u, hipMemcpyDeviceToHost, a->own), "copy")) !=
                    PGH_OK ||
                (st = hip_status(hipStreamSynchronize(a->own), "hipStreamSynchronize")) != PGH_OK) {
                break;
            }
            if (differs) {
                st = fail(PGH_DEVICE_ERROR, "pgh_archi_tune: the mapping {form " + std::to_string(cand[c].form) +
                                                ", lj_lds " + std::to_string(cand[c].lj_lds) + ", blocks_per_cu " +
                                                std::to_string(cand[c].blocks_per_cu) +
                                                "} left a state different from the first mapping's");
                break;
            }
        }
        if (c == 0u || ms < best_ms) {
            best_ms = ms;
            best = cand[c];
        }
    }
    // back to the state before tuning, with the fastest mapping (or the old
    // one after a failure)
    const pgh_status rs = restore();
    a->trace = trace;
    const pgh_status mst = pgh_archi_set_mapping(a, st == PGH_OK ? &best : &before);
    const hipError_t ss = hipStreamSynchronize(a->own);
    release();
    if (st != PGH_OK) return st;
    PGH_TRY(rs);
    PGH_TRY(mst);
    PGH_HIP(ss);
    if (chosen) *chosen = best;
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_archi_set_migration(pgh_archi *a, int type, int handling, const uint64_t *conn_offsets,
                                   const uint64_t *conn_src, const double *conn_weight,
                                   const pgh_migr_rate *select_rate, const pgh_migr_rate *replace_rate,
                                   const uint32_t *migr_seeds)
try {
    if (!a) return fail(PGH_INVALID_ARGUMENT, "null archipelago");
    if (a->last) PGH_HIP(hipStreamSynchronize(a->last));
    if (!conn_offsets) {
        if (a->migr) migr_destroy(a->migr);
        a->migr = nullptr;
        return PGH_OK;
    }
    std::vector<migr_island> isl;
    for (size_t k = 0; k < a->h_desc.size(); ++k) {
        const island_desc &di = a->h_desc[k];
        if (a->memory[k]) return fail(PGH_UNSUPPORTED, "migration with pso_gen memory is not supported");
        isl.push_back(migr_island{di.nx, di.pop, di.xoff, di.foff, di.eoff});
    }
    migr_ctx *m = nullptr;
    PGH_TRY(migr_create(isl, type, handling, conn_offsets, conn_src, conn_weight, select_rate, replace_rate,
                        migr_seeds, &m));
    if (a->migr) migr_destroy(a->migr);
    a->migr = m;
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_archi_get_migrants(pgh_archi *a, uint64_t k, uint64_t *count, double *X, double *F)
try {
    if (!a || k >= a->h_desc.size()) return fail(PGH_INVALID_ARGUMENT, "no such island");
    if (!a->migr) {
        if (count) *count = 0u;
        return PGH_OK;
    }
    if (a->last) PGH_HIP(hipStreamSynchronize(a->last));
    return migr_get(a->migr, k, count, X, F);
} catch (...) {
    return api_exception();
}

pgh_status pgh_archi_get(pgh_archi *a, uint64_t k, double *X, double *F)
try {
    if (!a || k >= a->h_desc.size()) return fail(PGH_INVALID_ARGUMENT, "no such island");
    if (a->last) PGH_HIP(hipStreamSynchronize(a->last));
    const island_desc &di = a->h_desc[k];
    // pop.set_xf(i, lbX[i], lbfit[i]), pso_gen.cpp:526-528
    const double *xs = static_cast<const double *>(a->evolved ? a->lbX.ptr : a->X.ptr);
    const double *fs = static_cast<const double *>(a->evolved ? a->lbfit.ptr : a->fit.ptr);
    if (X) {
        PGH_HIP(hipMemcpy(X, xs + di.xoff, static_cast<uint64_t>(di.pop) * di.nx * 8u, hipMemcpyDeviceToHost));
    }
    if (F) {
        PGH_HIP(hipMemcpy(F, fs + di.foff, static_cast<uint64_t>(di.pop) * 8u, hipMemcpyDeviceToHost));
    }
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_archi_get_trace(pgh_archi *a, uint64_t k, double *trace_F)
try {
    if (!a || k >= a->h_desc.size()) return fail(PGH_INVALID_ARGUMENT, "no such island");
    if (!a->trace) return fail(PGH_INVALID_ARGUMENT, "tracing is off");
    if (a->last) PGH_HIP(hipStreamSynchronize(a->last));
    const island_desc &di = a->h_desc[k];
    PGH_HIP(hipMemcpy(trace_F, static_cast<double *>(a->trace_buf.ptr) + di.toff,
                      static_cast<uint64_t>(di.gens) * di.pop * 8u, hipMemcpyDeviceToHost));
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_pso_gen_evolve_host(const pgh_udp *udp, const pgh_pso_params *prm, double *X, double *F, uint64_t pop,
                                   uint64_t nx)
try {
    if (!udp || !prm || !X || !F) return fail(PGH_INVALID_ARGUMENT, "null argument");
    prob p;
    PGH_TRY(resolve(udp, p));
    if (nx != p.nx) return fail(PGH_INVALID_ARGUMENT, "population dimension does not match the problem");
    pgh_island_spec sp{};
    sp.udp = *udp;
    sp.pop = pop;
    sp.algo = *prm;
    pgh_archi *a = nullptr;
    const double *xs[1] = {X};
    const double *fs[1] = {F};
    PGH_TRY(pgh_archi_create(&a, &sp, 1u, xs, fs));
    std::unique_ptr<pgh_archi> guard(a);
    // one island, evolved once: direct launches (capturing and instantiating
    // a graph that runs once costs more than it saves), a batch of
    // generations per launch when the island is one evaluation tile
    a->one_shot = true;
    PGH_TRY(pgh_archi_evolve(a, PGH_ARCHI_PER_ISLAND, nullptr));
    return pgh_archi_get(a, 0u, X, F);
} catch (...) {
    return api_exception();
}

} // extern "C"
