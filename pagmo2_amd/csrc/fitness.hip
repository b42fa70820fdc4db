// fitness.hip -- batch fitness and batch gradient kernels (the hip_bfe path).
//
// Data layout: dvs is the row-major n_dvs x nx matrix that pagmo packs before
// every bfe call (src/algorithms/pso_gen.cpp:420-425, src/problem.cpp:383-392);
// fvs is n_dvs x nf, grad is n_dvs x nx (dense, nf == 1).
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "fitness_dev.hpp"
#include "pgh_internal.hpp"

namespace pgh
{
using namespace dev;

// Individuals per Lennard-Jones workgroup set by pgh_lj_set_groups or found by
// pgh_lj_tune, keyed by 2 * atoms + (gradient kernel); lj_group_count
// otherwise.  Any count from 1 to lj_group_count gives the same results: each
// individual is evaluated by its own L-lane group in the same order.
namespace
{
std::mutex g_lj_mu;
std::map<uint32_t, uint32_t> g_lj_groups;
} // namespace

uint32_t lj_groups_for(uint32_t atoms, bool grad)
{
    {
        std::lock_guard<std::mutex> lk(g_lj_mu);
        const auto it = g_lj_groups.find(2u * atoms + (grad ? 1u : 0u));
        if (it != g_lj_groups.end()) return it->second;
    }
    return lj_group_count(atoms, grad);
}

namespace
{

constexpr int kBlock = 256;

// ---- separable problems: an L-lane group per individual -------------------
// For rastrigin and ackley an individual with an element outside the streaming
// cosine's range comes back marked; one lane then evaluates it again with the
// reference's expression and an accurate cosine (rare).  That path stays out
// of line and the kernel is held to 64 VGPRs (8 waves per SIMD), so that its
// register needs do not lower the streaming loop's occupancy; it may spill.
template <int KIND, bool ALIGNED>
__device__ __noinline__ double separable_cold(const double *__restrict__ x, uint32_t nx)
{
    return separable_eval_exact<KIND, 1, ALIGNED>(x, nx, 0);
}

constexpr int sep_min_waves(int kind)
{
    return (kind == RASTRIGIN || kind == ACKLEY) ? 8 : 1;
}

template <int KIND, int L, bool ALIGNED>
__global__ __launch_bounds__(kBlock, sep_min_waves(KIND)) void k_separable(const double *__restrict__ dvs, uint64_t n,
                                                                           uint32_t nx, double *__restrict__ fvs)
{
    constexpr uint32_t G = kBlock / L;
    const uint64_t ind = static_cast<uint64_t>(blockIdx.x) * G + threadIdx.x / L;
    const int r = static_cast<int>(threadIdx.x % L);
    // a trailing partial group re-reads the last row so that every lane of the
    // group takes part in the shuffles; its result is dropped
    const uint64_t row = ind < n ? ind : n - 1u;
    double f = separable_eval<KIND, L, ALIGNED>(dvs + row * nx, nx, r);
    if constexpr (KIND == RASTRIGIN || KIND == ACKLEY) {
        if (r == 0 && is_recompute_mark(f)) f = separable_cold<KIND, ALIGNED>(dvs + row * nx, nx);
    }
    if (ind < n && r == 0) fvs[ind] = f;
}

// Rastrigin fitness and gradient in one pass over 16-byte aligned rows: the
// fitness is k_separable's (the same routine and reduction, hence the same
// bits as a bfe call), the gradient is written by the lanes as they add
// their pairs.  A row with an element beyond the streaming routines' range
// gets its fitness and its gradient again from the general routines, out of
// line.
template <int L>
__global__ __launch_bounds__(kBlock, sep_min_waves(RASTRIGIN)) void k_rastrigin_fg(const double *__restrict__ dvs,
                                                                                   uint64_t n, uint32_t nx,
                                                                                   double *__restrict__ fvs,
                                                                                   double *__restrict__ grad)
{
    constexpr uint32_t G = kBlock / L;
    const uint64_t ind = static_cast<uint64_t>(blockIdx.x) * G + threadIdx.x / L;
    const int r = static_cast<int>(threadIdx.x % L);
    // a trailing partial group evaluates the last row again, writing the same
    // gradient values to it, and drops its fitness
    const uint64_t row = ind < n ? ind : n - 1u;
    double f = separable_eval<RASTRIGIN, L, true, true>(dvs + row * nx, nx, r, grad + row * nx);
    if (is_recompute_mark(f)) {
        // group-uniform: every lane of the group holds the mark
        rastrigin_grad_row(dvs + row * nx, nx, static_cast<uint32_t>(r), L, grad + row * nx);
        if (r == 0) f = separable_cold<RASTRIGIN, true>(dvs + row * nx, nx);
    }
    if (ind < n && r == 0 && fvs) fvs[ind] = f;
}

// ---- rosenbrock over rows of 64 and more: 8 rows per wavefront ------------
// A wavefront takes 8 rows through column chunks of 64: for each row the
// lanes read 64 consecutive elements (contiguous loads), and compute in
// parallel each element's term of the sum and, with GRAD, its gradient
// (stored contiguously, non-temporal: nothing reads it back).  The terms go
// to the wavefront's LDS tile, where lane r (r < 8) then adds row r's terms in
// the reference's order (rosenbrock.cpp:62-64), so the sum is the reference's
// bit for bit; the next chunk's loads are issued before that sum.  Each
// wavefront uses its own tile: no workgroup barrier.
constexpr uint32_t kRWRows = 8, kRWStride = 65; // odd stride: the 8 summing lanes hit distinct banks

template <bool GRAD>
__global__ __launch_bounds__(kBlock) void k_rosen_wave(const double *__restrict__ dvs, uint64_t n, uint32_t nx,
                                                       double *__restrict__ fvs, double *__restrict__ grad)
{
    constexpr uint32_t W = 64, kWaves = kBlock / W;
    __shared__ double terms[kWaves * kRWRows * kRWStride];
    const uint32_t wave = threadIdx.x / W, lane = threadIdx.x % W;
    const uint64_t row0 = (static_cast<uint64_t>(blockIdx.x) * kWaves + wave) * kRWRows;
    if (row0 >= n) return; // wavefront-uniform
    const uint32_t rows = static_cast<uint32_t>(min(static_cast<uint64_t>(kRWRows), n - row0));
    double *tile = terms + wave * kRWRows * kRWStride;
    const double *__restrict__ base = dvs + row0 * nx;
    double xm[kRWRows], xi[kRWRows], xn[kRWRows];
    // Clamped indices: lanes past the row's end take element nx - 1, and rows
    // past the batch the last row, so that every load stays in the batch with
    // no branch; such a lane stores nothing and its term is not summed.
    const auto load = [&](uint32_t c0) {
        const uint32_t ii = min(c0 + lane, nx - 1u), in = min(ii + 1u, nx - 1u), im = ii == 0u ? 0u : ii - 1u;
#pragma unroll
        for (uint32_t r = 0; r < kRWRows; ++r) {
            const double *x = base + static_cast<uint64_t>(min(r, rows - 1u)) * nx;
            xi[r] = x[ii];
            xn[r] = x[in];
            if constexpr (GRAD) xm[r] = x[im];
        }
    };
    load(0u);
    double acc = 0.;
    for (uint32_t c0 = 0; c0 < nx; c0 += W) {
        const uint32_t ii = min(c0 + lane, nx - 1u);
#pragma unroll
        for (uint32_t r = 0; r < kRWRows; ++r) {
            const double d = xi[r] * xi[r] - xn[r];
            tile[r * kRWStride + lane] = 100. * d * d + (xi[r] - 1) * (xi[r] - 1);
            if constexpr (GRAD) {
                const double gv = rosenbrock_grad3(xm[r], xi[r], xn[r], nx, ii);
                if (r < rows && c0 + lane < nx) {
                    __builtin_nontemporal_store(gv, grad + (row0 + r) * nx + ii);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (c0 + W < nx) load(c0 + W);
        // terms i = c0 .. c0 + cnt - 1 of row `lane`, in order
        const uint32_t cnt = c0 + 1u < nx ? min(W, nx - 1u - c0) : 0u;
        if (lane < rows) {
            const double *t = tile + lane * kRWStride;
            for (uint32_t j = 0; j < cnt; ++j) acc += t[j];
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (lane < rows && fvs) fvs[row0 + lane] = acc;
}

// ---- rosenbrock over even rows of 16-byte aligned matrices: 8 rows per ------
// wavefront, column chunks of 128 K.  Lane l reads elements c0 + 128 k + 2l
// and c0 + 128 k + 2l + 1 (k < K) of each row with 16-byte loads (contiguous
// KiBs per row and instruction, which stream from HBM better than
// k_rosen_wave's 512-byte rows); the neighbours x[i - 1] and x[i + 1] come
// from the adjacent lanes, across pieces k and k + 1 from lanes 63 and 0, the
// chunk's last one from one more load and the previous chunk's from a carried
// register.  Terms, their in-order sum by lane r < 8 and the gradient are
// k_rosen_wave's, bit for bit.
// (At 65536 x 1000 with the gradient: 8 rows 0.217 ms, 4 rows 0.235, 16 rows
// 0.222, 2 rows 0.289; K = 2 no faster than K = 1, with 104 VGPRs against 80.)
constexpr uint32_t kRPRows = 8, kRPK = 1, kRPCols = 128u * kRPK, kRPStride = kRPCols + 1u;

__device__ __forceinline__ double shfl_dbl(double v, int src)
{
    return __shfl(v, src, 64);
}

template <bool GRAD>
__global__ __launch_bounds__(kBlock) void k_rosen_pairs(const double *__restrict__ dvs, uint64_t n, uint32_t nx,
                                                        double *__restrict__ fvs, double *__restrict__ grad)
{
    constexpr uint32_t W = 64, kWaves = kBlock / W;
    __shared__ double terms[kWaves * kRPRows * kRPStride];
    const uint32_t wave = threadIdx.x / W, lane = threadIdx.x % W;
    const uint64_t row0 = (static_cast<uint64_t>(blockIdx.x) * kWaves + wave) * kRPRows;
    if (row0 >= n) return; // wavefront-uniform
    const uint32_t rows = static_cast<uint32_t>(min(static_cast<uint64_t>(kRPRows), n - row0));
    double *tile = terms + wave * kRPRows * kRPStride;
    const uint32_t half = nx / 2u; // pairs per row (nx even)
    const double *__restrict__ base = dvs + row0 * nx;
    double2 v[kRPRows][kRPK];
    double ext[kRPRows], carry[kRPRows];
    // Clamped indices: lanes past the row's end take its last pair, rows past
    // the batch the last row, so that every load stays in the batch; such
    // lanes store nothing and their terms are not summed.
    const auto load = [&](uint32_t c0) {
        const uint32_t ce = min(c0 + kRPCols, nx - 1u);
#pragma unroll
        for (uint32_t r = 0; r < kRPRows; ++r) {
            const double *x = base + static_cast<uint64_t>(min(r, rows - 1u)) * nx;
#pragma unroll
            for (uint32_t k = 0; k < kRPK; ++k) {
                v[r][k] = reinterpret_cast<const double2 *>(x)[min(c0 / 2u + 64u * k + lane, half - 1u)];
            }
            ext[r] = x[ce];
        }
    };
#pragma unroll
    for (uint32_t r = 0; r < kRPRows; ++r) carry[r] = 0.;
    load(0u);
    double acc = 0.;
    const int up_lane = static_cast<int>((lane + W - 1u) % W), down_lane = static_cast<int>((lane + 1u) % W);
    for (uint32_t c0 = 0; c0 < nx; c0 += kRPCols) {
#pragma unroll
        for (uint32_t r = 0; r < kRPRows; ++r) {
#pragma unroll
            for (uint32_t k = 0; k < kRPK; ++k) {
                const uint32_t i = c0 + 128u * k + 2u * lane; // this lane's first element of piece k
                const double a = v[r][k].x, b = v[r][k].y;
                const double down = shfl_dbl(a, down_lane); // x[i + 2] for lane < 63
                // lane 63: x[i + 2] is lane 0's of piece k + 1, or the chunk's
                // next element (every lane takes part in the shuffles)
                const double after = k + 1u < kRPK ? shfl_dbl(v[r][k + 1u < kRPK ? k + 1u : k].x, 0) : ext[r];
                const double xn = lane + 1u < W ? down : after;
                const double da = a * a - b, db = b * b - xn;
                tile[r * kRPStride + 128u * k + 2u * lane] = 100. * da * da + (a - 1) * (a - 1);
                tile[r * kRPStride + 128u * k + 2u * lane + 1u] = 100. * db * db + (b - 1) * (b - 1);
                if constexpr (GRAD) {
                    const double up = shfl_dbl(b, up_lane); // x[i - 1] for lane > 0
                    // lane 0: x[i - 1] is lane 63's of piece k - 1, or the carried one
                    const double prev = k > 0u ? shfl_dbl(v[r][k > 0u ? k - 1u : 0u].y, 63) : carry[r];
                    const double xm = lane ? up : prev;
                    if (r < rows && i < nx) {
                        store_pair_nt(reinterpret_cast<double2 *>(grad + (row0 + r) * nx + i),
                                      rosenbrock_grad3(xm, a, b, nx, i), rosenbrock_grad3(a, b, xn, nx, i + 1u));
                    }
                }
            }
            if constexpr (GRAD) carry[r] = shfl_dbl(v[r][kRPK - 1u].y, 63); // the chunk's last element
        }
        __builtin_amdgcn_wave_barrier();
        if (c0 + kRPCols < nx) load(c0 + kRPCols);
        // terms i = c0 .. c0 + cnt - 1 of row `lane`, in order
        const uint32_t cnt = c0 + 1u < nx ? min(kRPCols, nx - 1u - c0) : 0u;
        if (lane < rows) {
            const double *t = tile + lane * kRPStride;
            for (uint32_t j = 0; j < cnt; ++j) acc += t[j];
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (lane < rows && fvs) fvs[row0 + lane] = acc;
}

// ---- multi-objective sphere: one lane per individual ------------------------
__global__ __launch_bounds__(kBlock) void k_mo_sphere(const double *__restrict__ dvs, uint64_t n, uint32_t nx,
                                                      uint32_t nobj, const double *__restrict__ centers,
                                                      double *__restrict__ fvs)
{
    const uint64_t ind = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (ind >= n) return;
    const double *x = dvs + ind * nx;
    for (uint32_t k = 0; k < nobj; ++k) {
        fvs[ind * nobj + k] = mo_sphere_eval(x, nx, centers + static_cast<uint64_t>(k) * nx);
    }
}

// ---- one lane per individual, rows staged in LDS ----------------------------
// A workgroup copies its `rows` consecutive individuals (a contiguous span of
// rows * nx doubles) into LDS with coalesced loads, at an odd row stride so
// that the lanes' reads of one column fall in distinct banks, and each lane
// then evaluates its own row in the same sequential order as the direct
// kernels below (which serve rows too long for the tile).  Without the
// staging every lane walks its own row and a wavefront's loads touch 64
// cache lines each, which the L1 cannot hold for short rows.
constexpr uint32_t kRowTileBytes = 32768;

constexpr uint32_t row_stride(uint32_t nx)
{
    return nx | 1u;
}

// Rows per staged workgroup (a multiple of 64, at most 256; 0: not staged,
// which includes nx = 1, whose reciprocal below does not fit 32 bits).
constexpr uint32_t staged_rows(uint32_t nx)
{
    if (nx < 2u) return 0u;
    const uint32_t r = kRowTileBytes / (8u * row_stride(nx));
    return r >= 256u ? 256u : (r / 64u) * 64u;
}

struct row_task {
    uint32_t nx, nf, dim, prob_id, alpha;
    uint32_t rows;  // individuals per workgroup (staged_rows)
    uint32_t magic; // floor(2^32 / nx) + 1: e / nx == umulhi(e, magic) for e * nx < 2^32
    const double *centers;
};

constexpr uint32_t kStageUnroll = 4;

// Element e of the workgroup's span into its LDS slot.
__device__ __forceinline__ void stage_put(double *__restrict__ tile, uint32_t e, double v, uint32_t nx,
                                          uint32_t stride, uint32_t magic)
{
    const uint32_t r = __umulhi(e, magic);
    tile[r * stride + (e - r * nx)] = v;
}

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_rows_staged(const double *__restrict__ dvs, uint64_t n, row_task t,
                                                       double *__restrict__ fvs)
{
    extern __shared__ double tile[];
    const uint32_t nx = t.nx, stride = row_stride(nx);
    const uint64_t r0 = static_cast<uint64_t>(blockIdx.x) * t.rows;
    const uint32_t cnt = static_cast<uint32_t>(min(static_cast<uint64_t>(t.rows), n - r0));
    const uint32_t total = cnt * nx;
    // the span starts at an even element (rows is a multiple of 64) of a
    // 16-byte aligned matrix: pairs of elements are read with 16-byte loads,
    // kStageUnroll of them in flight per lane, then placed in LDS
    const double2 *src = reinterpret_cast<const double2 *>(dvs + r0 * nx);
    const uint32_t pairs = total / 2u;
    for (uint32_t base = threadIdx.x; base < pairs; base += kBlock * kStageUnroll) {
        double2 v[kStageUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kStageUnroll; ++u) {
            const uint32_t q = base + u * kBlock;
            if (q < pairs) v[u] = src[q];
        }
#pragma unroll
        for (uint32_t u = 0; u < kStageUnroll; ++u) {
            const uint32_t q = base + u * kBlock;
            if (q < pairs) {
                stage_put(tile, 2u * q, v[u].x, nx, stride, t.magic);
                stage_put(tile, 2u * q + 1u, v[u].y, nx, stride, t.magic);
            }
        }
    }
    if ((total & 1u) && threadIdx.x == 0) {
        stage_put(tile, total - 1u, dvs[r0 * nx + total - 1u], nx, stride, t.magic);
    }
    __syncthreads();
    const uint32_t i = threadIdx.x;
    if (i >= cnt) return;
    const double *x = tile + i * stride;
    const uint64_t ind = r0 + i;
    if constexpr (KIND == ROSENBROCK) {
        fvs[ind] = rosenbrock_eval(x, t.dim);
    } else if constexpr (KIND == MO_SPHERE) {
        for (uint32_t k = 0; k < t.nf; ++k) {
            fvs[ind * t.nf + k] = mo_sphere_eval(x, nx, t.centers + static_cast<uint64_t>(k) * nx);
        }
    } else if constexpr (KIND == ZDT) {
        double f0, f1;
        zdt_eval(t.prob_id, x, nx, f0, f1);
        fvs[2u * ind] = f0;
        fvs[2u * ind + 1u] = f1;
    } else {
        dtlz_eval(t.prob_id, x, nx, t.nf, t.alpha, fvs + ind * t.nf);
    }
}

// Staged when the row fits the tile and the matrix is 16-byte aligned.
bool use_staged(const prob &p, const double *dvs)
{
    return staged_rows(p.nx) != 0u && reinterpret_cast<uintptr_t>(dvs) % 16u == 0u;
}

template <int KIND>
pgh_status launch_rows_staged(const prob &p, const double *dvs, uint64_t n, double *fvs, hipStream_t s)
{
    const uint32_t rows = staged_rows(p.nx);
    const uint32_t magic = static_cast<uint32_t>((uint64_t(1) << 32) / p.nx + 1u);
    const row_task t{p.nx, p.nf, p.dim, p.prob_id, p.alpha, rows, magic, p.centers};
    const size_t lds = static_cast<size_t>(rows) * row_stride(p.nx) * sizeof(double);
    k_rows_staged<KIND><<<dim3(static_cast<uint32_t>((n + rows - 1u) / rows)), dim3(kBlock), lds, s>>>(dvs, n, t,
                                                                                                        fvs);
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

// ---- ZDT and DTLZ: one lane per individual ----------------------------------
__global__ __launch_bounds__(kBlock) void k_zdt(const double *__restrict__ dvs, uint64_t n, uint32_t nx,
                                                uint32_t id, double *__restrict__ fvs)
{
    const uint64_t ind = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (ind >= n) return;
    double f0, f1;
    zdt_eval(id, dvs + ind * nx, nx, f0, f1);
    fvs[2u * ind] = f0;
    fvs[2u * ind + 1u] = f1;
}

__global__ __launch_bounds__(kBlock) void k_dtlz(const double *__restrict__ dvs, uint64_t n, uint32_t nx,
                                                 uint32_t M, uint32_t id, uint32_t alpha, double *__restrict__ fvs)
{
    const uint64_t ind = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (ind >= n) return;
    dtlz_eval(id, dvs + ind * nx, nx, M, alpha, fvs + ind * M);
}

// ---- lennard_jones: G individuals per workgroup staged in LDS --------------
// Up to three owned slots per lane the energy fits 64 VGPRs: 8 wavefronts per
// SIMD, so that a batch of 16384 individuals on 32-lane groups (8 wavefronts
// per SIMD) is resident at once.
template <int L, uint32_t APL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(APL >= 1 && APL <= 3 ? 8 : 1))) void k_lj(
    const double *__restrict__ dvs, uint64_t n, uint32_t atoms, uint32_t G, double *__restrict__ fvs)
{
    extern __shared__ double lds[];
    const uint32_t nx = 3u * atoms - 6u;
    const uint64_t ind0 = static_cast<uint64_t>(blockIdx.x) * G;
    const uint32_t count = static_cast<uint32_t>(min(static_cast<uint64_t>(G), n - ind0));
    lj_stage(dvs + ind0 * nx, count, atoms, false, lds);
    __syncthreads();
    const uint32_t g = threadIdx.x / L;
    const int r = static_cast<int>(threadIdx.x % L);
    const uint32_t gg = g < count ? g : count - 1u;
    const double e = lj_energy<L, APL>(lds + gg * lj_stride(atoms, false), atoms, r);
    if (g < count && r == 0) fvs[ind0 + g] = e;
}

template <int L, uint32_t APL>
__global__ __launch_bounds__(kBlock) void k_lj_grad(const double *__restrict__ dvs, uint64_t n, uint32_t atoms,
                                                    uint32_t G, double *__restrict__ grad, double *__restrict__ fvs)
{
    extern __shared__ double lds[];
    const uint32_t nx = 3u * atoms - 6u;
    const uint64_t ind0 = static_cast<uint64_t>(blockIdx.x) * G;
    const uint32_t count = static_cast<uint32_t>(min(static_cast<uint64_t>(G), n - ind0));
    lj_stage(dvs + ind0 * nx, count, atoms, true, lds);
    __syncthreads();
    const uint32_t g = threadIdx.x / L;
    const int r = static_cast<int>(threadIdx.x % L);
    if (g >= count) return; // whole groups retire together: no shuffle partner is lost
    const double e = lj_energy_grad<L, APL>(lds + g * lj_stride(atoms, true), atoms, r, grad + (ind0 + g) * nx);
    if (fvs && r == 0) fvs[ind0 + g] = e;
}

// ---- elementwise gradients ---------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_rastrigin_grad(const double *__restrict__ dvs, uint64_t count,
                                                           double *__restrict__ grad)
{
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= count) return;
    grad[t] = rastrigin_grad(dvs[t]);
}

template <int KIND, int L>
pgh_status launch_sep_l(const double *dvs, uint64_t n, uint32_t nx, double *fvs, hipStream_t s)
{
    constexpr uint64_t G = kBlock / L;
    const uint64_t blocks = (n + G - 1u) / G;
    const bool aligned = (nx % 2u == 0u) && (reinterpret_cast<uintptr_t>(dvs) % 16u == 0u);
    const dim3 grid(static_cast<uint32_t>(blocks));
    if (aligned) {
        k_separable<KIND, L, true><<<grid, dim3(kBlock), 0, s>>>(dvs, n, nx, fvs);
    } else {
        k_separable<KIND, L, false><<<grid, dim3(kBlock), 0, s>>>(dvs, n, nx, fvs);
    }
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

template <int KIND>
pgh_status launch_sep(const double *dvs, uint64_t n, uint32_t nx, double *fvs, hipStream_t s)
{
    switch (sep_lanes(KIND, nx)) {
        case 64:
            return launch_sep_l<KIND, 64>(dvs, n, nx, fvs, s);
        case 32:
            return launch_sep_l<KIND, 32>(dvs, n, nx, fvs, s);
        case 16:
            return launch_sep_l<KIND, 16>(dvs, n, nx, fvs, s);
        default:
            return launch_sep_l<KIND, 4>(dvs, n, nx, fvs, s);
    }
}

template <int L>
pgh_status launch_rastrigin_fg_l(const double *dvs, uint64_t n, uint32_t nx, double *fvs, double *grad, hipStream_t s)
{
    constexpr uint64_t G = kBlock / L;
    k_rastrigin_fg<L><<<dim3(static_cast<uint32_t>((n + G - 1u) / G)), dim3(kBlock), 0, s>>>(dvs, n, nx, fvs, grad);
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

pgh_status launch_rastrigin_fg(const double *dvs, uint64_t n, uint32_t nx, double *fvs, double *grad, hipStream_t s)
{
    switch (sep_lanes(RASTRIGIN, nx)) {
        case 64:
            return launch_rastrigin_fg_l<64>(dvs, n, nx, fvs, grad, s);
        case 32:
            return launch_rastrigin_fg_l<32>(dvs, n, nx, fvs, grad, s);
        case 16:
            return launch_rastrigin_fg_l<16>(dvs, n, nx, fvs, grad, s);
        default:
            return launch_rastrigin_fg_l<4>(dvs, n, nx, fvs, grad, s);
    }
}

pgh_status launch_rosen_rows(const double *dvs, uint64_t n, uint32_t nx, double *fvs, double *grad, hipStream_t s)
{
    // even rows of 16-byte aligned matrices: 16-byte loads and stores
    const bool pairs = nx % 2u == 0u && reinterpret_cast<uintptr_t>(dvs) % 16u == 0u &&
                       (!grad || reinterpret_cast<uintptr_t>(grad) % 16u == 0u);
    if (pairs) {
        constexpr uint64_t rpb = kBlock / 64u * kRPRows;
        const dim3 grid(static_cast<uint32_t>((n + rpb - 1u) / rpb));
        if (grad) k_rosen_pairs<true><<<grid, dim3(kBlock), 0, s>>>(dvs, n, nx, fvs, grad);
        else k_rosen_pairs<false><<<grid, dim3(kBlock), 0, s>>>(dvs, n, nx, fvs, nullptr);
        PGH_HIP(hipGetLastError());
        return PGH_OK;
    }
    constexpr uint64_t rows_per_block = kBlock / 64u * kRWRows;
    const dim3 grid(static_cast<uint32_t>((n + rows_per_block - 1u) / rows_per_block));
    if (grad) k_rosen_wave<true><<<grid, dim3(kBlock), 0, s>>>(dvs, n, nx, fvs, grad);
    else k_rosen_wave<false><<<grid, dim3(kBlock), 0, s>>>(dvs, n, nx, fvs, nullptr);
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

template <int L, uint32_t APL>
pgh_status launch_lj_la(const double *dvs, uint64_t n, uint32_t atoms, double *fvs, double *grad, hipStream_t s)
{
    const bool with_grad = grad != nullptr;
    const uint32_t G = lj_groups_for(atoms, with_grad);
    const uint64_t blocks = (n + G - 1u) / G;
    const size_t lds = static_cast<size_t>(G) * lj_stride(atoms, with_grad) * sizeof(double);
    if (grad) {
        k_lj_grad<L, APL><<<dim3(static_cast<uint32_t>(blocks)), dim3(G * L), lds, s>>>(dvs, n, atoms, G, grad,
                                                                                        fvs);
    } else {
        k_lj<L, APL><<<dim3(static_cast<uint32_t>(blocks)), dim3(G * L), lds, s>>>(dvs, n, atoms, G, fvs);
    }
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

template <int L>
pgh_status launch_lj_l(const double *dvs, uint64_t n, uint32_t atoms, double *fvs, double *grad, hipStream_t s)
{
    switch (lj_apl(atoms, L)) {
        case 1:
            return launch_lj_la<L, 1>(dvs, n, atoms, fvs, grad, s);
        case 2:
            return launch_lj_la<L, 2>(dvs, n, atoms, fvs, grad, s);
        case 3:
            return launch_lj_la<L, 3>(dvs, n, atoms, fvs, grad, s);
        case 4:
            return launch_lj_la<L, 4>(dvs, n, atoms, fvs, grad, s);
        default:
            break;
    }
    if constexpr (L >= 32) {
        switch (lj_apl(atoms, L)) {
            case 5:
                return launch_lj_la<L, 5>(dvs, n, atoms, fvs, grad, s);
            case 6:
                return launch_lj_la<L, 6>(dvs, n, atoms, fvs, grad, s);
            case 7:
                return launch_lj_la<L, 7>(dvs, n, atoms, fvs, grad, s);
            case 8:
                return launch_lj_la<L, 8>(dvs, n, atoms, fvs, grad, s);
            default:
                break;
        }
    }
    return launch_lj_la<L, 0>(dvs, n, atoms, fvs, grad, s);
}

pgh_status launch_lj(const double *dvs, uint64_t n, uint32_t atoms, double *fvs, double *grad, hipStream_t s)
{
    if (lj_group_count(atoms, grad != nullptr) == 0u) {
        return fail(PGH_UNSUPPORTED, "lennard_jones with " + std::to_string(atoms)
                                         + " atoms does not fit the LDS of one workgroup");
    }
    switch (lj_lanes(atoms, grad != nullptr)) {
        case 32:
            return launch_lj_l<32>(dvs, n, atoms, fvs, grad, s);
        case 16:
            return launch_lj_l<16>(dvs, n, atoms, fvs, grad, s);
        case 8:
            return launch_lj_l<8>(dvs, n, atoms, fvs, grad, s);
        case 4:
            return launch_lj_l<4>(dvs, n, atoms, fvs, grad, s);
        case 2:
            return launch_lj_l<2>(dvs, n, atoms, fvs, grad, s);
        default:
            return launch_lj_l<1>(dvs, n, atoms, fvs, grad, s);
    }
}

uint32_t blocks_for(uint64_t count)
{
    return static_cast<uint32_t>((count + kBlock - 1u) / kBlock);
}

} // namespace

namespace
{
__global__ __launch_bounds__(kBlock) void k_shift(double *__restrict__ x, uint64_t count, uint32_t nx,
                                                  const double *__restrict__ t)
{
    const uint64_t e = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (e < count) x[e] = x[e] - t[e % nx];
}
} // namespace

pgh_status launch_shift(double *x, uint64_t count, uint32_t nx, const double *t, hipStream_t s)
{
    if (count == 0u) return PGH_OK;
    if (count > (uint64_t(1) << 39)) return fail(PGH_OVERFLOW, "batch too large for one launch");
    k_shift<<<dim3(blocks_for(count)), dim3(kBlock), 0, s>>>(x, count, nx, t);
    PGH_HIP(hipGetLastError());
    return PGH_OK;
}

// Grid sizes stay below 2^31 workgroups: at most 2^31 individuals per call and
// 2^39 matrix elements for the elementwise gradient kernels.
pgh_status check_launch_size(uint64_t n, uint32_t nx)
{
    if (n > (uint64_t(1) << 31) || n * nx > (uint64_t(1) << 39)) {
        return fail(PGH_OVERFLOW, "batch too large for one launch: " + std::to_string(n) + " x "
                                      + std::to_string(nx));
    }
    return PGH_OK;
}

pgh_status launch_fitness(const prob &p, const double *dvs, uint64_t n, double *fvs, hipStream_t s)
{
    if (n == 0u) return PGH_OK;
    if (check_launch_size(n, p.nx) != PGH_OK) return PGH_OVERFLOW;
    switch (p.kind) {
        case ROSENBROCK:
            if (use_staged(p, dvs)) return launch_rows_staged<ROSENBROCK>(p, dvs, n, fvs, s);
            return launch_rosen_rows(dvs, n, p.nx, fvs, nullptr, s);
        case RASTRIGIN:
            return launch_sep<RASTRIGIN>(dvs, n, p.nx, fvs, s);
        case ACKLEY:
            return launch_sep<ACKLEY>(dvs, n, p.nx, fvs, s);
        case SCHWEFEL:
            return launch_sep<SCHWEFEL>(dvs, n, p.nx, fvs, s);
        case GRIEWANK:
            return launch_sep<GRIEWANK>(dvs, n, p.nx, fvs, s);
        case LENNARD_JONES:
            return launch_lj(dvs, n, p.dim, fvs, nullptr, s);
        case MO_SPHERE:
            if (use_staged(p, dvs)) return launch_rows_staged<MO_SPHERE>(p, dvs, n, fvs, s);
            k_mo_sphere<<<dim3(blocks_for(n)), dim3(kBlock), 0, s>>>(dvs, n, p.nx, p.nf, p.centers, fvs);
            PGH_HIP(hipGetLastError());
            return PGH_OK;
        case ZDT:
            if (use_staged(p, dvs)) return launch_rows_staged<ZDT>(p, dvs, n, fvs, s);
            k_zdt<<<dim3(blocks_for(n)), dim3(kBlock), 0, s>>>(dvs, n, p.nx, p.prob_id, fvs);
            PGH_HIP(hipGetLastError());
            return PGH_OK;
        case DTLZ:
            if (use_staged(p, dvs)) return launch_rows_staged<DTLZ>(p, dvs, n, fvs, s);
            k_dtlz<<<dim3(blocks_for(n)), dim3(kBlock), 0, s>>>(dvs, n, p.nx, p.nf, p.prob_id, p.alpha, fvs);
            PGH_HIP(hipGetLastError());
            return PGH_OK;
    }
    return fail(PGH_UNSUPPORTED, "no fitness kernel for this problem kind");
}

pgh_status launch_gradient(const prob &p, const double *dvs, uint64_t n, double *grad, double *fvs, hipStream_t s)
{
    if (n == 0u) return PGH_OK;
    if (check_launch_size(n, p.nx) != PGH_OK) return PGH_OVERFLOW;
    switch (p.kind) {
        case ROSENBROCK:
            // fitness (when asked) and gradient in one pass
            return launch_rosen_rows(dvs, n, p.nx, fvs, grad, s);
        case RASTRIGIN:
            if (p.nx % 2u == 0u && reinterpret_cast<uintptr_t>(dvs) % 16u == 0u &&
                reinterpret_cast<uintptr_t>(grad) % 16u == 0u) {
                return launch_rastrigin_fg(dvs, n, p.nx, fvs, grad, s);
            }
            k_rastrigin_grad<<<dim3(blocks_for(n * p.nx)), dim3(kBlock), 0, s>>>(dvs, n * p.nx, grad);
            PGH_HIP(hipGetLastError());
            return fvs ? launch_fitness(p, dvs, n, fvs, s) : PGH_OK;
        case LENNARD_JONES:
            return launch_lj(dvs, n, p.dim, fvs, grad, s);
    }
    return fail(PGH_UNSUPPORTED, "The gradient has not been implemented in the user-defined problem");
}

} // namespace pgh

using namespace pgh;

extern "C" {

pgh_status pgh_lj_set_groups(uint64_t atoms, int grad, uint32_t groups)
try {
    if (atoms < 3u || atoms > 4096u) return fail(PGH_INVALID_ARGUMENT, "atoms out of range: " + std::to_string(atoms));
    const uint32_t a = static_cast<uint32_t>(atoms);
    const uint32_t most = lj_group_count(a, grad != 0);
    if (groups > most) {
        return fail(PGH_INVALID_ARGUMENT, "at most " + std::to_string(most) + " individuals per workgroup for "
                                              + std::to_string(atoms) + " atoms, got " + std::to_string(groups));
    }
    std::lock_guard<std::mutex> lk(g_lj_mu);
    if (groups == 0u) g_lj_groups.erase(2u * a + (grad ? 1u : 0u));
    else g_lj_groups[2u * a + (grad ? 1u : 0u)] = groups;
    return PGH_OK;
} catch (...) {
    return api_exception();
}

pgh_status pgh_lj_tune(const pgh_udp *udp, const double *dvs, uint64_t n_dvs, double *fvs, double *grad,
                       uint32_t reps, void *stream, uint32_t *chosen)
try {
    prob p;
    PGH_TRY(resolve(udp, p));
    if (p.kind != PGH_LENNARD_JONES) return fail(PGH_INVALID_ARGUMENT, "pgh_lj_tune needs a lennard_jones problem");
    if (reps == 0u) return fail(PGH_INVALID_ARGUMENT, "the tuner needs at least one timed repetition");
    if (n_dvs == 0u || !dvs || (!fvs && !grad)) return fail(PGH_INVALID_ARGUMENT, "the tuner needs a batch to time");
    if (chosen) *chosen = 0u;
    PGH_TRY(check_launch_size(n_dvs, p.nx));
    const bool with_grad = grad != nullptr;
    const uint32_t most = lj_group_count(p.dim, with_grad);
    if (most == 0u) return fail(PGH_UNSUPPORTED, "lennard_jones cluster too large for one workgroup");
    std::vector<uint32_t> cands;
    for (uint32_t g = most; g >= 1u && cands.size() < 4u; g >>= 1) cands.push_back(g);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PGH_HIP(hipEventCreate(&e0));
    const hipError_t ce = hipEventCreate(&e1);
    if (ce != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(ce, "hipEventCreate");
    }
    float best = 0.f;
    uint32_t best_g = 0u;
    pgh_status st = PGH_OK;
    for (uint32_t g : cands) {
        if ((st = pgh_lj_set_groups(p.dim, with_grad, g)) != PGH_OK) break;
        if ((st = launch_lj(dvs, n_dvs, p.dim, fvs, grad, s)) != PGH_OK) break; // warm-up
        if ((st = hip_status(hipEventRecord(e0, s), "hipEventRecord")) != PGH_OK) break;
        for (uint32_t r = 0; r < reps && st == PGH_OK; ++r) st = launch_lj(dvs, n_dvs, p.dim, fvs, grad, s);
        if (st != PGH_OK) break;
        if ((st = hip_status(hipEventRecord(e1, s), "hipEventRecord")) != PGH_OK) break;
        if ((st = hip_status(hipEventSynchronize(e1), "hipEventSynchronize")) != PGH_OK) break;
        float ms = 0.f;
        if ((st = hip_status(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) != PGH_OK) break;
        if (best_g == 0u || ms < best) {
            best = ms;
            best_g = g;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != PGH_OK) {
        (void)pgh_lj_set_groups(p.dim, with_grad, 0u);
        return st;
    }
    PGH_TRY(pgh_lj_set_groups(p.dim, with_grad, best_g));
    if (chosen) *chosen = best_g;
    return launch_lj(dvs, n_dvs, p.dim, fvs, grad, s);
} catch (...) {
    return api_exception();
}

} // extern "C"
