// pso.hip -- pso_gen (src/algorithms/pso_gen.cpp:120-530) for an archipelago
// of independent islands of uneven size, on the device.
//
// Per island the device holds, contiguously: X, V, lbX (pop x nx), fit, lbfit
// (pop), the island's std::mt19937 state and a word buffer for a batch of
// kGenBatch generations (pop x nx x 4 words each: r1 and r2 of
// pso_gen.cpp:300-301 in draw order).  A generation is up to three launches:
//   k_rng     (every kGenBatch generations) the island's generator split in
//             segments, one workgroup each, each handing out whole
//             generations of the batch and reaching its own by jump-ahead
//             (mt_dev.hpp, mt_jump): exactly the draws of the reference's
//             sequential loop
//   k_update  one lane per (island, particle, dimension): best neighbour
//             (pso_gen.cpp:593-622), velocity (variant 5, :302-305), clamp and
//             position (:336-361) -- arithmetic in the reference's order
//   k_eval    one workgroup per tile of individuals: fitness with the same
//             device routine as hip_bfe, then the memory update
//             (fit <= lbfit, pso_gen.cpp:443-449) and the optional trace
// Batched mode launches each phase once for all islands and captures the
// whole generation loop in a hipGraph; per-island mode launches the same
// kernels island by island.  Both run identical arithmetic per island.
#include <hip/hip_runtime.h>

#include <algorithm>
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
    uint64_t toff; // trace offset (doubles), gens x pop
    uint64_t eoff; // cumulative pop*nx of the islands before this one
    // the generator split in `segs` segments (k_rng): of every batch of
    // kGenBatch generations, segment s hands out the words of generations
    // s gps .. (s + 1) gps - 1; seg0 is the first segment's index; j_* index
    // the archipelago's jump table (z^n0, z^(gps W 2^b) for b = 0..,
    // z^((segs - 1) gps W))
    uint32_t segs, seg0, gps;
    uint32_t j_n0, j_tree, j_main;
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
// k_rng hands out the words of kGenBatch generations at a time; an island's
// generator is split in segments (k_rng) so that each hands out at most
// kSegWords words of a batch, or a single generation's
constexpr uint32_t kGenBatch = 32;
constexpr uint64_t kSegWords = uint64_t(1) << 20;

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
// phase 0, the start of an evolve call: every segment takes the island's
// canonical generator (everything handed out so far); segment 0 hands out
// the velocity draws (2 words per element, pso_gen.cpp:193-200) unless the
// island keeps its velocities (memory, once they exist: v_valid), and every
// segment moves to the start of its first generation, s gps
// (jumps z^n0 and z^(gps W 2^b) over the bits of s).  phase 1, the batch of
// generations g0 .. g0 + kGenBatch - 1 (4 words per element and generation,
// r1 and r2 of pso_gen.cpp:300-301): segment s hands out the words of its
// gps generations (those before the island's last) into their slots of the
// word buffer, then jumps over the other segments' generations to its share
// of the next batch.  The segment that hands out the island's last
// generation stops there and becomes the canonical generator.  The words are
// those of one sequential generator, in order.
__global__ __launch_bounds__(kMtJumpBlock) void k_rng(const island_desc *__restrict__ d,
                                                      const seg_desc *__restrict__ sd, mt_state *__restrict__ st,
                                                      mt_state *__restrict__ segst, jump_tab jt,
                                                      uint32_t *__restrict__ words, uint32_t sg0, uint32_t phase,
                                                      uint32_t g0, uint32_t v_valid)
{
    __shared__ uint32_t ring[kMtRing];
    __shared__ uint4 lds_idx[kMtIdxStride / 8u];
    const seg_desc sg = sd[sg0 + blockIdx.x];
    const island_desc di = d[sg.island];
    const uint64_t elems = static_cast<uint64_t>(di.pop) * di.nx;
    const uint64_t W = 4u * elems;
    const auto jump = [&](uint64_t &K, uint64_t &C, uint64_t J, uint32_t j) {
        mt_jump(ring, K, C, J, jt.lists + static_cast<uint64_t>(j) * kMtIdxStride, jt.lens[j], lds_idx);
    };
    uint64_t K, C;
    mt_state *mine = &segst[di.seg0 + sg.s];
    if (phase == 0u) {
        load_state(&st[sg.island], ring, K, C);
        // velocities are drawn unless the island keeps them (pso_gen.cpp:193-200)
        const bool draw_v = !(di.memory && v_valid);
        const uint64_t n0 = draw_v ? 2u * elems : 0u;
        if (sg.s == 0u) {
            if (n0) {
                mt_run(ring, K, C, n0, words + di.woff);
                C += n0;
            }
            if (di.gens == 0u) store_state(&st[sg.island], ring, K, C);
        } else {
            if (n0) jump(K, C, n0, di.j_n0);
            for (uint32_t b = 0; (sg.s >> b) != 0u; ++b) {
                if ((sg.s >> b) & 1u) jump(K, C, (di.gps * W) << b, di.j_tree + b);
            }
        }
        store_state(mine, ring, K, C);
        return;
    }
    const uint32_t first = g0 + sg.s * di.gps;
    if (first >= di.gens) return; // no generation of this batch, nor of a later one
    const uint32_t last = min(first + di.gps, di.gens); // one past
    load_state(mine, ring, K, C);
    const uint64_t n = (last - first) * W;
    mt_run(ring, K, C, n, words + di.woff + static_cast<uint64_t>(first - g0) * W);
    C += n;
    if (last == di.gens) {
        store_state(&st[sg.island], ring, K, C);
    } else if (first + kGenBatch < di.gens) {
        if (di.segs > 1u) jump(K, C, static_cast<uint64_t>(di.segs - 1u) * di.gps * W, di.j_main);
        store_state(mine, ring, K, C);
    }
}

// Start of an evolve call (pso_gen.cpp:183-200): V = U(minv, maxv) per
// element, and X = lbX = the population.  The population is X/fit before the
// first call and lbX/lbfit after one (pop.set_xf(i, lbX, lbfit)).
__global__ __launch_bounds__(kBlock) void k_vinit(const island_desc *__restrict__ d, uint32_t i0, uint32_t i1,
                                                  uint64_t e0, uint64_t total, const uint32_t *__restrict__ words,
                                                  double *__restrict__ X, double *__restrict__ V,
                                                  double *__restrict__ lbX, double *__restrict__ fit,
                                                  double *__restrict__ lbfit, uint32_t v_valid, uint32_t from_lb)
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
    if (from_lb) {
        X[di.xoff + m] = lbX[di.xoff + m];
        if (j == 0u) fit[di.foff + p] = lbfit[di.foff + p];
    } else {
        lbX[di.xoff + m] = X[di.xoff + m];
        if (j == 0u) lbfit[di.foff + p] = fit[di.foff + p];
    }
}

// Velocity and position of element m of island di in generation gen.
__device__ __forceinline__ void update_elem(const island_desc &di, uint64_t m, const uint32_t *__restrict__ words,
                                            double *__restrict__ X, double *__restrict__ V,
                                            const double *__restrict__ lbX, const double *__restrict__ lbfit,
                                            uint32_t gen)
{
    const uint32_t p = static_cast<uint32_t>(m / di.nx);
    const uint32_t dd = static_cast<uint32_t>(m - static_cast<uint64_t>(p) * di.nx);
    const uint32_t S = di.pop, R = di.radius;
    const double *lbf = lbfit + di.foff;
    // lbest ring (pso_gen.cpp:679-699): radius..1 to the left, then 1..radius
    // to the right; the later neighbour wins ties (:615-617)
    uint32_t bn = p < R ? p - R + S : p - R;
    for (uint32_t j = R - 1u; j >= 1u; --j) {
        const uint32_t q = p < j ? p - j + S : p - j;
        if (le_f(lbf[q], lbf[bn])) bn = q;
    }
    for (uint32_t j = 1u; j <= R; ++j) {
        const uint32_t q = p + j >= S ? p + j - S : p + j;
        if (le_f(lbf[q], lbf[bn])) bn = q;
    }
    const uint32_t *w = words + di.woff + static_cast<uint64_t>(gen % kGenBatch) * (4ull * di.pop * di.nx) + 4u * m;
    const double r1 = mt_canonical(w[0], w[1]);
    const double r2 = mt_canonical(w[2], w[3]);
    const uint64_t xi = di.xoff + m;
    const double x = X[xi];
    const double best = lbX[di.xoff + static_cast<uint64_t>(bn) * di.nx + dd];
    double v = di.omega * (V[xi] + di.eta1 * r1 * (lbX[xi] - x) + di.eta2 * r2 * (best - x));
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
    X[xi] = nxv;
    V[xi] = v;
}

__global__ __launch_bounds__(kBlock) void k_update(const island_desc *__restrict__ d, uint32_t i0, uint32_t i1,
                                                   uint64_t e0, uint64_t total, const uint32_t *__restrict__ words,
                                                   double *__restrict__ X, double *__restrict__ V,
                                                   const double *__restrict__ lbX, const double *__restrict__ lbfit,
                                                   uint32_t gen)
{
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= total) return;
    const uint64_t e = e0 + t;
    const uint32_t k = find_island(d, i0, i1, e);
    const island_desc di = d[k];
    if (gen >= di.gens) return;
    update_elem(di, e - di.eoff, words, X, V, lbX, lbfit, gen);
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
    const double f = lj_energy<L, APL>(lds + gg * lj_stride(atoms, false), atoms, static_cast<int>(r));
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

__global__ __launch_bounds__(kBlock) void k_eval(const island_desc *__restrict__ d, const tile *__restrict__ tiles,
                                                 uint32_t t0, eval_ctx c)
{
    extern __shared__ double lds[];
    const tile tl = tiles[t0 + blockIdx.x];
    const island_desc di = d[tl.island];
    if (c.gen >= di.gens) return;
    eval_tile(di, tl, c, lds);
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
        __shared__ uint32_t ring[kMtRing];
        const uint32_t first = g1;
        if (first >= di.gens) return;
        const uint32_t last = min(first + di.gps, di.gens);
        const uint64_t W = 4u * elems;
        mt_state *mine = &segst[di.seg0];
        uint64_t K, C;
        load_state(mine, ring, K, C);
        const uint64_t n = (last - first) * W;
        mt_run(ring, K, C, n, words_next + di.woff);
        C += n;
        if (last == di.gens) {
            store_state(&st[island], ring, K, C);
        } else if (first + kGenBatch < di.gens) {
            store_state(mine, ring, K, C);
        }
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
    uint32_t n_segs = 0;
    jump_tab jt{nullptr, nullptr};
    hipStream_t own = nullptr;
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
    k_rng<<<ng, kMtJumpBlock, 0, s>>>(d, sd, st, segst, jt, w, sg0, 0u, 0u, v_valid ? 1u : 0u);
    PGH_HIP(hipGetLastError());
    k_vinit<<<blocks_for(total), kBlock, 0, s>>>(d, i0, i1, e0, total, w, X, V, lbX, fit, lbfit, v_valid ? 1u : 0u,
                                                 from_lb ? 1u : 0u);
    PGH_HIP(hipGetLastError());
    // a single island of one tile, evolved by the one-shot entry point: a
    // batch of generations per launch (k_gens_one_tile)
    const bool fused = a->one_shot && i1 == i0 + 1u && nt == 1u;
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
    // ... and the next batch's words handed out beside it, into the other
    // half of the word buffer, when the generator is one segment
    const bool overlap = fused && a->words_half != 0u && a->h_desc[i0].segs == 1u;
    for (uint32_t g = 0; g < gens; ++g) {
        if (g % kGenBatch == 0u) {
            const uint32_t b = g / kGenBatch;
            if (!overlap || g == 0u) k_rng<<<ng, kMtJumpBlock, 0, s>>>(d, sd, st, segst, jt, w, sg0, 1u, g, 0u);
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
        k_update<<<blocks_for(total), kBlock, 0, s>>>(d, i0, i1, e0, total, w, X, V, lbX, lbfit, g);
        k_eval<<<nt, kBlock, a->lds_bytes, s>>>(d, static_cast<const tile *>(a->tiles.ptr), t0, make_ctx(a, g));
        PGH_HIP(hipGetLastError());
    }
    return PGH_OK;
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
    uint64_t xoff = 0, foff = 0, woff = 0, toff = 0, eoff = 0;
    std::vector<seg_desc> seg_list;
    std::vector<uint64_t> dists;
    a->tile_begin.push_back(0);
    for (uint64_t k = 0; k < n_islands; ++k) {
        const pgh_island_spec &sp = islands[k];
        prob p;
        PGH_TRY(resolve(&sp.udp, p));
        if (p.nf != 1u) return fail(PGH_UNSUPPORTED, "pso_gen needs a single-objective problem");
        if (p.shift) return fail(PGH_UNSUPPORTED, "the archipelago does not take translated problems");
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
        if (q.variant != 5u || q.neighb_type != 2u) {
            return fail(PGH_UNSUPPORTED, "the device pso_gen implements variant 5 with the lbest topology");
        }
        if (sp.pop == 0u) return fail(PGH_INVALID_ARGUMENT, "pso_gen does not work on an empty population");
        const uint32_t radius = q.neighb_param / 2u;
        if (radius == 0u || sp.pop < radius || sp.pop > (uint64_t(1) << 24)) {
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
        const uint64_t elems = sp.pop * p.nx;
        di.xoff = xoff;
        di.foff = foff;
        di.woff = woff;
        di.toff = toff;
        di.eoff = eoff;
        xoff += align_up(elems, 32u); // 256-byte aligned rows blocks
        foff += align_up(sp.pop, 32u);
        // the words of a batch of generations; a lone island of up to 2^22
        // words per batch gets room for a second batch, which the one-shot
        // evolve fills while it evolves the first (k_gens_one_tile)
        const uint64_t wbatch = align_up(elems * 4u * kGenBatch, 64u);
        const bool second = n_islands == 1u && wbatch <= (uint64_t(1) << 22);
        woff += second ? 2u * wbatch : wbatch;
        if (second) a->words_half = wbatch;
        toff += static_cast<uint64_t>(q.gen) * sp.pop;
        eoff += elems;
        a->max_gens = std::max(a->max_gens, q.gen);
        a->memory.push_back(q.memory ? 1u : 0u);
        // evaluation tiles
        uint32_t per = 256u;
        if (p.kind == PGH_LENNARD_JONES) {
            di.lanes = static_cast<uint32_t>(lj_lanes(p.dim));
            per = lj_groups_for(p.dim, false);
            if (per == 0u) return fail(PGH_UNSUPPORTED, "lennard_jones cluster too large for one workgroup");
            a->lds_bytes =
                std::max<size_t>(a->lds_bytes, static_cast<size_t>(per) * lj_stride(p.dim, false) * sizeof(double));
        } else if (p.kind != PGH_ROSENBROCK) {
            di.lanes = static_cast<uint32_t>(sep_lanes(p.kind, p.nx));
            per = 256u / di.lanes;
        } else {
            di.lanes = 1u;
        }
        // the generator's segments (k_rng): as few as keep each segment's
        // share of a batch within kSegWords words (a power of two, at most
        // one segment per generation of the batch)
        {
            const uint64_t W = 4u * elems;
            const uint32_t batch = std::min<uint32_t>(kGenBatch, std::max<uint32_t>(q.gen, 1u));
            di.seg0 = static_cast<uint32_t>(seg_list.size());
            di.segs = 1u;
            while (di.segs < batch && ((batch + di.segs - 1u) / di.segs) * W > kSegWords) di.segs *= 2u;
            // gps * segs = kGenBatch when a call has a batch or more of
            // generations; otherwise the single batch's generations
            di.gps = (batch + di.segs - 1u) / di.segs;
            di.j_n0 = di.j_tree = di.j_main = 0u;
            if (di.segs > 1u) {
                const uint64_t share = static_cast<uint64_t>(di.gps) * W;
                di.j_n0 = static_cast<uint32_t>(dists.size());
                dists.push_back(2u * elems);
                di.j_tree = static_cast<uint32_t>(dists.size());
                for (uint32_t b = 0; (di.segs - 1u) >> b; ++b) dists.push_back(share << b);
                di.j_main = static_cast<uint32_t>(dists.size());
                dists.push_back((di.segs - 1u) * share);
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
    const uint64_t head_bytes = arena_bytes;
    place(a->X, xoff * 8u);
    place(a->V, xoff * 8u);
    place(a->lbX, xoff * 8u);
    place(a->fit, foff * 8u);
    place(a->lbfit, foff * 8u);
    place(a->words, std::max<uint64_t>(woff, 1u) * 4u);
    place(a->seg_states, seg_list.size() * sizeof(mt_state));
    PGH_TRY(a->arena.reserve(arena_bytes));
    auto *base = static_cast<uint8_t *>(a->arena.ptr);
    for (arena_view *v : {&a->desc, &a->tiles, &a->segs, &a->jump_table, &a->states, &a->X, &a->V, &a->lbX, &a->fit,
                          &a->lbfit, &a->words, &a->seg_states}) {
        v->ptr = base + v->off;
    }
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
