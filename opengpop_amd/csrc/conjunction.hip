// Conjunction screening and refinement for gfx950.
//
// Screening (replaces UpdateClosestApproachKernel, closest_approach/src/device.cu:56-91): the
// reference updates a per-pair running minimum in HBM after every step (492 B of state per
// pair). Here the trajectory of every object is stored once ([step][component][object]) and a
// pair tile (64 x 64 objects) walks the whole time axis with its running minima in registers;
// object positions are staged through LDS 8 steps at a time. Per pair the state is the smallest
// squared distance and the last decision window (of one or more stages) in which its square
// root dropped.
//
// Decision rule, identical to the reference's "dist < closest" on sqrt'ed distances with the
// first occurrence kept; the squared distance is formed as fma(dz, dz, fma(dy, dy, dx * dx))
// exactly as in the host oracle, so identical trajectories give identical bits.
//
// Refinement (CalculateClosestApproachesKernel, device.cu:93-173; interpolation.cuh:27-152):
// one thread per candidate pair; 8th-degree Hermite interpolation through the samples at
// steps k-1, k, k+1 and a secant root of (rA - rB).(vA - vB).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "conjunction.hpp"

namespace ogp {

namespace {
constexpr int TA = 64, TB = 64, PA = 4, PB = 4, STAGE = 8, THREADS = 256;
// elements of one stage (STAGE steps x 3 components x (TA + TB) objects) per thread
constexpr int STAGE_ELEMS = STAGE * 3 * (TA + TB) / THREADS;
static_assert(STAGE * 3 * (TA + TB) % THREADS == 0, "stage must divide over the block");
// d2 < best * kShrink guarantees sqrt(d2) < sqrt(best) for correctly rounded square roots
// (the two roots then differ by more than one rounding each), so the roots are only formed for
// samples within a few ulps of the running minimum.
constexpr double kShrink = 1.0 - 0x1p-50;

// v_min_f64 without the quieting v_max_f64 the compiler puts in front of a loop-carried operand:
// every operand here is the result of arithmetic or the constant inf, never a signalling NaN,
// and with one quiet NaN operand the instruction returns the other operand.
__device__ __forceinline__ double min_f64(double a, double b)
{
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// A stage is 48 rows of 64 objects: row r holds tile r / 24 (0 = A, 1 = B), step (r % 24) / 3
// and component r % 3. Wave w of the block loads rows w, w + 4, ..., one object per lane, so the
// row arithmetic is wave-uniform and each load is 512 contiguous bytes. Values past the end of
// the time axis or of a set are 0 and never screened.
static_assert(TA == 64 && TB == 64 && THREADS == 256 && STAGE_ELEMS * 4 == STAGE * 3 * 2,
    "stage rows assume 64-object tiles and 4 waves");
__device__ __forceinline__ double stage_load(const double* __restrict__ traj, int64_t N,
    int64_t sstride, int64_t nsteps, int64_t k0, int row, int lane, int64_t ta0, int64_t tb0,
    int64_t alim, int64_t blim)
{
    const int which = row / 24, r = row % 24;
    const int64_t k = k0 + r / 3;
    const int64_t obj = (which == 0 ? ta0 : tb0) + lane;
    if (k < nsteps && obj < (which == 0 ? alim : blim))
        return traj[k * sstride + (r % 3) * N + obj];
    return 0.0;
}

// Smallest and largest of the wave's values (every lane gets them).
__device__ __forceinline__ void wave_min_max(double& lo, double& hi)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, m));
        hi = fmax(hi, __shfl_xor(hi, m));
    }
}

// Whether radius ranges [alo, ahi] and [blo, bhi] lie more than the threshold apart.
__device__ __forceinline__ bool radially_apart(double alo, double ahi, double blo, double bhi,
    double threshold)
{
    return fmax(alo - bhi, blo - ahi) > threshold + kRadialMargin;
}
}

// One block screens a TA x TB tile of pairs over the whole time axis. Thread (ty, tx) owns A
// objects ty*PA .. ty*PA+3 and B objects {2tx, 2tx+1, 32+2tx, 32+2tx+1}: neighbouring lanes read
// neighbouring 16-byte pairs of the staged B positions (no LDS bank conflicts). The next stage's
// positions are loaded into registers while the current stage is screened.
//
// The reference's decision is deferred instead of being made per sample. Its minimum after any
// number of samples has the square root of the smallest d2 so far, and its step is the first
// sample whose root equals that. So per sample only the minimum over a window of SPAN samples
// is kept (one v_min_f64). At the end of each window the thread records whether the root
// strictly dropped in it: certainly when the window's minimum is below mn * kShrink, not at all
// when it is not below mn, and by exact roots in the rare case in between. After the time axis
// the step is the first sample, in the last window where the root dropped, whose root equals
// sqrt(mn); that scan reads the trajectory again only for pairs that are written out. A NaN
// distance is never taken (the minimum keeps the other operand), as under the reference's
// comparison.
#ifndef OGP_SCREEN_WAVES_PER_EU
#define OGP_SCREEN_WAVES_PER_EU 2
#endif
#ifndef OGP_SCREEN_UNROLL
#define OGP_SCREEN_UNROLL 1
#endif
// WIN stages per decision window (SPAN samples). Longer windows cost less per sample; shorter
// ones reread fewer samples for the pairs written out. Measured on 36 M pairs x 4,320 steps:
// 1, 2 and 4 stages gave 3.20, 3.23 and 3.30 x 10^12 pair-steps/s when only candidates below a
// threshold are written, and 100, 107 and 114 ms for the whole screen when every pair is.
template <int WIN>
__global__ void __launch_bounds__(THREADS)
__attribute__((amdgpu_waves_per_eu(OGP_SCREEN_WAVES_PER_EU)))
screen_kernel(const double* __restrict__ traj,
    int64_t N, int64_t sstride, int64_t nsteps, int64_t a0, int64_t na, int64_t b0, int64_t nb,
    int64_t pair0, int64_t row_len, double threshold, double* __restrict__ min_dist,
    int32_t* __restrict__ min_step, unsigned long long* __restrict__ cand_count, int64_t cand_cap,
    uint32_t* __restrict__ cand_a, uint32_t* __restrict__ cand_b, int32_t* __restrict__ cand_k,
    double* __restrict__ cand_d, const double* __restrict__ rlo, const double* __restrict__ rhi,
    const uint32_t* __restrict__ ids)
{
    constexpr int SPAN = STAGE * WIN; // samples per decision window
    __shared__ double sA[STAGE][3][TA];
    __shared__ double sB[STAGE][3][TB];
    const int tid = threadIdx.x;
    const int tx = tid % (TB / PB); // 16 columns of B sub-tiles
    const int ty = tid / (TB / PB); // 16 rows of A sub-tiles
    const int64_t ta0 = a0 + (int64_t)blockIdx.y * TA;
    const int64_t tb0 = b0 + (int64_t)blockIdx.x * TB;
    const int64_t alim = a0 + na, blim = b0 + nb;
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const int lane = tid % 64;
    if (rlo) { // radially sorted input: the tile's radius ranges (wave 0: A, wave 1: B)
        __shared__ double bnd[4];
        if (wave < 2) {
            const int64_t o = (wave == 0 ? ta0 : tb0) + lane;
            const bool ok = o < (wave == 0 ? alim : blim);
            double lo = ok ? rlo[o] : __builtin_inf(), hi = ok ? rhi[o] : -__builtin_inf();
            wave_min_max(lo, hi);
            if (lane == 0) {
                bnd[2 * wave] = lo;
                bnd[2 * wave + 1] = hi;
            }
        }
        __syncthreads();
        if (radially_apart(bnd[0], bnd[1], bnd[2], bnd[3], threshold))
            return; // the whole block: no pair of the tile comes within the threshold
    }
    int bcol[PB];
#pragma unroll
    for (int j = 0; j < PB; ++j)
        bcol[j] = (j / 2) * 32 + 2 * tx + (j % 2);
    double mn[PA][PB];     // smallest d2 before the current window
    double wm[PA][PB];     // smallest d2 in the current window
    int32_t sk[PA][PB];    // last window in which sqrt(mn) strictly dropped (-1: none)
#pragma unroll
    for (int i = 0; i < PA; ++i)
#pragma unroll
        for (int j = 0; j < PB; ++j) {
            mn[i][j] = __builtin_inf();
            sk[i][j] = -1;
        }
    double pre[STAGE_ELEMS];
#pragma unroll
    for (int q = 0; q < STAGE_ELEMS; ++q)
        pre[q] = stage_load(traj, N, sstride, nsteps, 0, 4 * q + wave, lane, ta0, tb0, alim, blim);
    for (int64_t k0 = 0; k0 < nsteps; k0 += STAGE) {
        __syncthreads(); // the previous stage has been screened by every thread
#pragma unroll
        for (int q = 0; q < STAGE_ELEMS; ++q) {
            const int row = 4 * q + wave;
            double* dst = row < 24 ? &sA[0][0][0] : &sB[0][0][0];
            dst[(row % 24) * 64 + lane] = pre[q];
        }
        __syncthreads();
        if (k0 + STAGE < nsteps) { // flies while this stage is screened
#pragma unroll
            for (int q = 0; q < STAGE_ELEMS; ++q)
                pre[q] = stage_load(traj, N, sstride, nsteps, k0 + STAGE, 4 * q + wave, lane, ta0,
                    tb0, alim, blim);
        }
        if (k0 % SPAN == 0) {
#pragma unroll
            for (int i = 0; i < PA; ++i)
#pragma unroll
                for (int j = 0; j < PB; ++j)
                    wm[i][j] = __builtin_inf();
        }
        const int smax = (int)((nsteps - k0) < STAGE ? (nsteps - k0) : STAGE);
        auto screen_step = [&](int st) {
            double ax[PA], ay[PA], az[PA], bx[PB], by[PB], bz[PB];
#pragma unroll
            for (int i = 0; i < PA; ++i) {
                ax[i] = sA[st][0][ty * PA + i];
                ay[i] = sA[st][1][ty * PA + i];
                az[i] = sA[st][2][ty * PA + i];
            }
#pragma unroll
            for (int j = 0; j < PB; ++j) {
                bx[j] = sB[st][0][bcol[j]];
                by[j] = sB[st][1][bcol[j]];
                bz[j] = sB[st][2][bcol[j]];
            }
#pragma unroll
            for (int i = 0; i < PA; ++i)
#pragma unroll
                for (int j = 0; j < PB; ++j) {
                    const double dx = ax[i] - bx[j];
                    const double dy = ay[i] - by[j];
                    const double dz = az[i] - bz[j];
                    const double d2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                    wm[i][j] = min_f64(wm[i][j], d2);
                }
        };
#if OGP_SCREEN_UNROLL
        if (smax == STAGE) { // full stage: the next step's reads can overlap this step's math
#pragma unroll
            for (int st = 0; st < STAGE; ++st)
                screen_step(st);
        } else
#endif
        {
            for (int st = 0; st < smax; ++st)
                screen_step(st);
        }
        if ((k0 + STAGE) % SPAN == 0 || k0 + STAGE >= nsteps) { // end of a window
            const int32_t win = (int32_t)(k0 / SPAN);
            bool unsure = false;
#pragma unroll
            for (int i = 0; i < PA; ++i)
#pragma unroll
                for (int j = 0; j < PB; ++j) {
                    const bool drop = wm[i][j] < mn[i][j] * kShrink;
                    unsure |= (wm[i][j] < mn[i][j]) & !drop;
                    sk[i][j] = drop ? win : sk[i][j];
                }
            if (unsure) { // a drop by a few ulps at most somewhere: compare the roots exactly
#pragma unroll
                for (int i = 0; i < PA; ++i)
#pragma unroll
                    for (int j = 0; j < PB; ++j)
                        if (wm[i][j] < mn[i][j] && __builtin_sqrt(wm[i][j]) < __builtin_sqrt(mn[i][j]))
                            sk[i][j] = win;
            }
#pragma unroll
            for (int i = 0; i < PA; ++i)
#pragma unroll
                for (int j = 0; j < PB; ++j)
                    mn[i][j] = min_f64(mn[i][j], wm[i][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < PA; ++i)
#pragma unroll
        for (int j = 0; j < PB; ++j) {
            const int64_t ia = ta0 + ty * PA + i;
            const int64_t jb = tb0 + bcol[j];
            if (ia >= alim || jb >= blim)
                continue;
            // pairs that never took a sample keep the reference's initial DBL_MAX
            const double d = sk[i][j] < 0 ? 1.7976931348623157e308 : __builtin_sqrt(mn[i][j]);
            const bool cand = threshold > 0.0 && d < threshold;
            if (!min_dist && !cand)
                continue;
            int32_t k = -1;
            if (sk[i][j] >= 0) { // first sample of that window whose root is the minimum
                const int64_t kb = (int64_t)sk[i][j] * SPAN;
                const int64_t ke = kb + SPAN < nsteps ? kb + SPAN : nsteps;
                for (int64_t kk = kb; kk < ke; ++kk) {
                    const double* p = traj + kk * sstride;
                    const double dx = p[ia] - p[jb];
                    const double dy = p[N + ia] - p[N + jb];
                    const double dz = p[2 * N + ia] - p[2 * N + jb];
                    const double d2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                    if (__builtin_sqrt(d2) == d) {
                        k = (int32_t)kk;
                        break;
                    }
                }
            }
            const int64_t p = pair0 + (ia - a0) * row_len + (jb - b0);
            if (min_dist) {
                min_dist[p] = d;
                min_step[p] = k;
            }
            if (cand) {
                const unsigned long long slot = atomicAdd(cand_count, 1ull);
                if (slot < (unsigned long long)cand_cap) {
                    cand_a[slot] = ids ? ids[ia] : (uint32_t)ia;
                    cand_b[slot] = ids ? ids[jb] : (uint32_t)jb;
                    cand_k[slot] = k;
                    cand_d[slot] = d;
                }
            }
        }
}

// Narrow tiles for a small set A: thread t of block (x, y) screens B object b0 + 256 x + t
// against the NA objects a0 + NA y .. a0 + NA y + NA - 1. Each B object's positions are read
// once per A tile, straight from HBM into the thread's registers (the next stage's loads fly
// while the current one is screened); the NA objects' positions are staged through LDS. Where
// set A holds a few objects, a 64 x 64 tile would leave most of its pairs empty; here the
// kernel streams set B's trajectory about once. The decision rule and the squared distance
// are screen_kernel's, so the minima and steps are the same bits.
constexpr int NA = 4, NB_THREADS = 256;
template <int WIN>
__global__ void __launch_bounds__(NB_THREADS)
screen_narrow_kernel(const double* __restrict__ traj, int64_t N, int64_t sstride,
    int64_t nsteps, int64_t a0, int64_t na, int64_t b0, int64_t nb, int64_t pair0,
    int64_t row_len, double threshold, double* __restrict__ min_dist,
    int32_t* __restrict__ min_step, unsigned long long* __restrict__ cand_count, int64_t cand_cap,
    uint32_t* __restrict__ cand_a, uint32_t* __restrict__ cand_b, int32_t* __restrict__ cand_k,
    double* __restrict__ cand_d, const double* __restrict__ rlo, const double* __restrict__ rhi,
    const uint32_t* __restrict__ ids)
{
    constexpr int SPAN = STAGE * WIN;
    __shared__ double sA[STAGE][3][NA];
    const int tid = threadIdx.x;
    const int64_t ta0 = a0 + (int64_t)blockIdx.y * NA;
    const int64_t alim = a0 + na, blim = b0 + nb;
    const int64_t jb = b0 + (int64_t)blockIdx.x * NB_THREADS + tid;
    const bool bvalid = jb < blim;
    const int64_t jl = bvalid ? jb : b0; // lanes past set B read object b0 and write nothing
    if (rlo) { // radially sorted input: the block's radius ranges (A: NA objects, B: its threads)
        __shared__ double bnd[NB_THREADS / 64][2];
        double lo = bvalid ? rlo[jb] : __builtin_inf(), hi = bvalid ? rhi[jb] : -__builtin_inf();
        wave_min_max(lo, hi);
        if (tid % 64 == 0) {
            bnd[tid / 64][0] = lo;
            bnd[tid / 64][1] = hi;
        }
        __syncthreads();
        double alo = __builtin_inf(), ahi = -__builtin_inf();
        for (int i = 0; i < NA; ++i)
            if (ta0 + i < alim) {
                alo = fmin(alo, rlo[ta0 + i]);
                ahi = fmax(ahi, rhi[ta0 + i]);
            }
        double blo = bnd[0][0], bhi = bnd[0][1];
        for (int w = 1; w < NB_THREADS / 64; ++w) {
            blo = fmin(blo, bnd[w][0]);
            bhi = fmax(bhi, bnd[w][1]);
        }
        if (radially_apart(alo, ahi, blo, bhi, threshold))
            return; // the whole block: no pair of its objects comes within the threshold
    }
    // A stage element e (< STAGE * 3 * NA): step e / (3 NA), component (e / NA) % 3, object e % NA
    auto a_load = [&](int64_t k0, int e) {
        const int64_t k = k0 + e / (3 * NA);
        const int c = (e / NA) % 3;
        const int64_t ia = ta0 + e % NA;
        return k < nsteps && ia < alim ? traj[k * sstride + c * N + ia] : 0.0;
    };
    auto b_load = [&](int64_t k0, double* dst) {
#pragma unroll
        for (int st = 0; st < STAGE; ++st)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int64_t k = k0 + st;
                dst[3 * st + c] = k < nsteps ? traj[k * sstride + c * N + jl] : 0.0;
            }
    };
    double mn[NA], wm[NA];
    int32_t sk[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        mn[i] = __builtin_inf();
        sk[i] = -1;
    }
    double cur[3 * STAGE], pre[3 * STAGE];
    double apre = tid < STAGE * 3 * NA ? a_load(0, tid) : 0.0;
    b_load(0, pre);
    for (int64_t k0 = 0; k0 < nsteps; k0 += STAGE) {
        __syncthreads(); // the previous stage has been screened by every thread
        if (tid < STAGE * 3 * NA)
            (&sA[0][0][0])[tid] = apre;
#pragma unroll
        for (int q = 0; q < 3 * STAGE; ++q)
            cur[q] = pre[q];
        __syncthreads();
        if (k0 + STAGE < nsteps) { // flies while this stage is screened
            if (tid < STAGE * 3 * NA)
                apre = a_load(k0 + STAGE, tid);
            b_load(k0 + STAGE, pre);
        }
        if (k0 % SPAN == 0) {
#pragma unroll
            for (int i = 0; i < NA; ++i)
                wm[i] = __builtin_inf();
        }
        const int smax = (int)((nsteps - k0) < STAGE ? (nsteps - k0) : STAGE);
        auto screen_step = [&](int st) {
#pragma unroll
            for (int i = 0; i < NA; ++i) {
                const double dx = sA[st][0][i] - cur[3 * st];
                const double dy = sA[st][1][i] - cur[3 * st + 1];
                const double dz = sA[st][2][i] - cur[3 * st + 2];
                const double d2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                wm[i] = min_f64(wm[i], d2);
            }
        };
        if (smax == STAGE) {
#pragma unroll
            for (int st = 0; st < STAGE; ++st)
                screen_step(st);
        } else {
#pragma unroll
            for (int st = 0; st < STAGE; ++st)
                if (st < smax)
                    screen_step(st);
        }
        if ((k0 + STAGE) % SPAN == 0 || k0 + STAGE >= nsteps) { // end of a window
            const int32_t win = (int32_t)(k0 / SPAN);
            bool unsure = false;
#pragma unroll
            for (int i = 0; i < NA; ++i) {
                const bool drop = wm[i] < mn[i] * kShrink;
                unsure |= (wm[i] < mn[i]) & !drop;
                sk[i] = drop ? win : sk[i];
            }
            if (unsure) {
#pragma unroll
                for (int i = 0; i < NA; ++i)
                    if (wm[i] < mn[i] && __builtin_sqrt(wm[i]) < __builtin_sqrt(mn[i]))
                        sk[i] = win;
            }
#pragma unroll
            for (int i = 0; i < NA; ++i)
                mn[i] = min_f64(mn[i], wm[i]);
        }
    }
    if (!bvalid)
        return;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        const int64_t ia = ta0 + i;
        if (ia >= alim)
            continue;
        const double d = sk[i] < 0 ? 1.7976931348623157e308 : __builtin_sqrt(mn[i]);
        const bool cand = threshold > 0.0 && d < threshold;
        if (!min_dist && !cand)
            continue;
        int32_t k = -1;
        if (sk[i] >= 0) { // first sample of that window whose root is the minimum
            const int64_t kb = (int64_t)sk[i] * SPAN;
            const int64_t ke = kb + SPAN < nsteps ? kb + SPAN : nsteps;
            for (int64_t kk = kb; kk < ke; ++kk) {
                const double* p = traj + kk * sstride;
                const double dx = p[ia] - p[jb];
                const double dy = p[N + ia] - p[N + jb];
                const double dz = p[2 * N + ia] - p[2 * N + jb];
                const double d2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                if (__builtin_sqrt(d2) == d) {
                    k = (int32_t)kk;
                    break;
                }
            }
        }
        const int64_t p = pair0 + (ia - a0) * row_len + (jb - b0);
        if (min_dist) {
            min_dist[p] = d;
            min_step[p] = k;
        }
        if (cand) {
            const unsigned long long slot = atomicAdd(cand_count, 1ull);
            if (slot < (unsigned long long)cand_cap) {
                cand_a[slot] = ids ? ids[ia] : (uint32_t)ia;
                cand_b[slot] = ids ? ids[jb] : (uint32_t)jb;
                cand_k[slot] = k;
                cand_d[slot] = d;
            }
        }
    }
}

void launch_screen(hipStream_t st, const double* traj, int64_t N, int64_t nsteps, int64_t a0,
    int64_t na, int64_t b0, int64_t nb, int64_t pair0, double threshold, double* min_dist,
    int32_t* min_step, unsigned long long* cand_count, int64_t cand_cap, uint32_t* cand_a,
    uint32_t* cand_b, int32_t* cand_k, double* cand_d, int shape, int64_t row_len,
    const ScreenSorted& sorted)
{
    if (na <= 0 || nb <= 0 || nsteps <= 0)
        return;
    if (row_len <= 0)
        row_len = nb;
    // the stored trajectory (9 N doubles per step, positions first) or the sorted positions
    const bool srt = sorted.pos != nullptr && threshold > 0.0 && min_dist == nullptr;
    const double* src = srt ? sorted.pos : traj;
    const int64_t sstride = srt ? 3 * N : 9 * N;
    const double* rlo = srt ? sorted.rlo : nullptr;
    const double* rhi = srt ? sorted.rhi : nullptr;
    const uint32_t* ids = srt ? sorted.ids : nullptr;
    // every pair written out: short windows; candidates only: long ones
    if (shape == kScreenNarrow) {
        const dim3 grid((unsigned)((nb + NB_THREADS - 1) / NB_THREADS), (unsigned)((na + NA - 1) / NA));
        if (min_dist)
            hipLaunchKernelGGL(screen_narrow_kernel<1>, grid, dim3(NB_THREADS), 0, st, src, N,
                sstride, nsteps, a0, na, b0, nb, pair0, row_len, threshold, min_dist, min_step,
                cand_count, cand_cap, cand_a, cand_b, cand_k, cand_d, rlo, rhi, ids);
        else
            hipLaunchKernelGGL(screen_narrow_kernel<4>, grid, dim3(NB_THREADS), 0, st, src, N,
                sstride, nsteps, a0, na, b0, nb, pair0, row_len, threshold, min_dist, min_step,
                cand_count, cand_cap, cand_a, cand_b, cand_k, cand_d, rlo, rhi, ids);
        return;
    }
    const dim3 grid((unsigned)((nb + TB - 1) / TB), (unsigned)((na + TA - 1) / TA));
    if (min_dist)
        hipLaunchKernelGGL(screen_kernel<1>, grid, dim3(THREADS), 0, st, src, N, sstride, nsteps,
            a0, na, b0, nb, pair0, row_len, threshold, min_dist, min_step, cand_count, cand_cap,
            cand_a, cand_b, cand_k, cand_d, rlo, rhi, ids);
    else
        hipLaunchKernelGGL(screen_kernel<4>, grid, dim3(THREADS), 0, st, src, N, sstride, nsteps,
            a0, na, b0, nb, pair0, row_len, threshold, min_dist, min_step, cand_count, cand_cap,
            cand_a, cand_b, cand_k, cand_d, rlo, rhi, ids);
}

// ------------------------------------------------------------------------------------------
// Radially sorted input for threshold screens
// ------------------------------------------------------------------------------------------
// Squared radii of the samples of step chunk y (of kRadialChunks) for objects 64 x .. 64 x + 63:
// the smallest into partial[y][i], the largest into partial[kRadialChunks + y][i]. A NaN sample
// is passed over, as the screen passes over a NaN distance.
__global__ void radial_partial_kernel(const double* __restrict__ traj, int64_t N, int64_t nsteps,
    double* __restrict__ partial)
{
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= N)
        return;
    const int64_t per = (nsteps + kRadialChunks - 1) / kRadialChunks;
    const int64_t k0 = (int64_t)blockIdx.y * per;
    const int64_t k1 = k0 + per < nsteps ? k0 + per : nsteps;
    double lo = __builtin_inf(), hi = -__builtin_inf();
    for (int64_t k = k0; k < k1; ++k) {
        const double* p = traj + k * 9 * N;
        const double x = p[i], y = p[N + i], z = p[2 * N + i];
        const double r2 = __builtin_fma(z, z, __builtin_fma(y, y, x * x));
        lo = fmin(lo, r2);
        hi = fmax(hi, r2);
    }
    partial[(int64_t)blockIdx.y * N + i] = lo;
    partial[(int64_t)(kRadialChunks + blockIdx.y) * N + i] = hi;
}

__global__ void radial_reduce_kernel(const double* __restrict__ partial, int64_t N,
    double* __restrict__ rlo, double* __restrict__ rhi)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    double lo = __builtin_inf(), hi = -__builtin_inf();
    for (int y = 0; y < kRadialChunks; ++y) {
        lo = fmin(lo, partial[(int64_t)y * N + i]);
        hi = fmax(hi, partial[(int64_t)(kRadialChunks + y) * N + i]);
    }
    rlo[i] = __builtin_sqrt(lo); // inf and -inf (no finite sample) leave every tile skippable
    rhi[i] = hi >= 0.0 ? __builtin_sqrt(hi) : hi;
}

// Slot s of every step gets object ids[s]'s position; steps blockIdx.y, + gridDim.y, ...
__global__ void gather_sorted_kernel(const double* __restrict__ traj, int64_t N, int64_t nsteps,
    const uint32_t* __restrict__ ids, const double* __restrict__ rlo,
    const double* __restrict__ rhi, double* __restrict__ pos, double* __restrict__ slo,
    double* __restrict__ shi)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N)
        return;
    const int64_t o = ids[s];
    if (blockIdx.y == 0) {
        slo[s] = rlo[o];
        shi[s] = rhi[o];
    }
    for (int64_t k = blockIdx.y; k < nsteps; k += gridDim.y)
        for (int c = 0; c < 3; ++c)
            pos[(k * 3 + c) * N + s] = traj[(k * 9 + c) * N + o];
}

void launch_radial_bounds(hipStream_t st, const double* traj, int64_t N, int64_t nsteps,
    double* partial, double* rlo, double* rhi)
{
    if (N <= 0 || nsteps <= 0)
        return;
    hipLaunchKernelGGL(radial_partial_kernel, dim3((unsigned)((N + 63) / 64), kRadialChunks),
        dim3(64), 0, st, traj, N, nsteps, partial);
    hipLaunchKernelGGL(radial_reduce_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st,
        partial, N, rlo, rhi);
}

void launch_gather_sorted(hipStream_t st, const double* traj, int64_t N, int64_t nsteps,
    const uint32_t* ids, const double* rlo, const double* rhi, double* pos, double* slo,
    double* shi)
{
    if (N <= 0 || nsteps <= 0)
        return;
    const unsigned gy = (unsigned)std::min<int64_t>(nsteps, 4096);
    hipLaunchKernelGGL(gather_sorted_kernel, dim3((unsigned)((N + 255) / 256), gy), dim3(256), 0,
        st, traj, N, nsteps, ids, rlo, rhi, pos, slo, shi);
}

// ------------------------------------------------------------------------------------------
// Refinement
// ------------------------------------------------------------------------------------------
namespace {

// Newton-form Hermite interpolation through (p, v, a) at 0, dt, 2 dt
// (interpolation.cuh:27-98).
struct Hermite {
    double c[9];
    double dt;
    __device__ Hermite(const double* q, double dt_)
        : dt(dt_)
    {
        const double t[9] = { 0.0, 0.0, 0.0, dt, dt, dt, 2.0 * dt, 2.0 * dt, 2.0 * dt };
        for (int i = 0; i < 9; ++i)
            c[i] = q[3 * (i / 3)];
        for (int j = 1; j < 9; ++j)
            for (int i = 8; i >= j; --i) {
                if (t[i] == t[i - j]) {
                    if (j == 1)
                        c[i] = q[3 * (i / 3) + 1];
                    else if (j == 2)
                        c[i] = q[3 * (i / 3) + 2] / 2.0;
                } else {
                    c[i] = (c[i] - c[i - 1]) / (t[i] - t[i - j]);
                }
            }
    }
    __device__ double val(double t) const
    {
        const double z[8] = { 0.0, 0.0, 0.0, dt, dt, dt, 2.0 * dt, 2.0 * dt };
        double res = c[0], term = 1.0;
        for (int i = 1; i < 9; ++i) {
            term *= (t - z[i - 1]);
            res += c[i] * term;
        }
        return res;
    }
    __device__ double der(double t) const
    {
        const double z[8] = { 0.0, 0.0, 0.0, dt, dt, dt, 2.0 * dt, 2.0 * dt };
        double res = 0.0, Nn = 1.0, dN = 0.0;
        for (int i = 0; i < 9; ++i) {
            res += c[i] * dN;
            if (i < 8) {
                dN = dN * (t - z[i]) + Nn;
                Nn *= (t - z[i]);
            }
        }
        return res;
    }
};

__device__ double approach(const Hermite* A, const Hermite* B, double t)
{
    double f = 0.0;
    for (int c = 0; c < 3; ++c)
        f += (A[c].val(t) - B[c].val(t)) * (A[c].der(t) - B[c].der(t));
    return f;
}

} // namespace

// Output of candidate q: distance, time, position of a, position of b. field_stride == 0: eight
// consecutive doubles per candidate; field_stride = S > 0: field blocks (distance[S], time[S],
// pos_a[S][3], pos_b[S][3]).
struct RefineOut {
    double* out;
    int64_t S;
    __device__ double& dist(int64_t q) const { return S ? out[q] : out[8 * q]; }
    __device__ double& time(int64_t q) const { return S ? out[S + q] : out[8 * q + 1]; }
    __device__ double& pa(int64_t q, int c) const { return S ? out[2 * S + 3 * q + c] : out[8 * q + 2 + c]; }
    __device__ double& pb(int64_t q, int c) const { return S ? out[5 * S + 3 * q + c] : out[8 * q + 5 + c]; }
};

__global__ void refine_kernel(const double* __restrict__ traj, int64_t N, int64_t nsteps, double h,
    double t0, int64_t n, const uint32_t* __restrict__ ia, const uint32_t* __restrict__ ib,
    const int32_t* __restrict__ ik, const double* __restrict__ id, RefineOut o)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n)
        return;
    const int64_t a = ia[q], b = ib[q];
    const int64_t k = ik[q];
    if (k <= 0 || k >= nsteps - 1) {
        // The reference reports time 0 (first sample) or total * h (last sample) and leaves
        // the positions unset; the sampled positions are returned here.
        o.dist(q) = id[q];
        o.time(q) = t0 + (k <= 0 ? 0.0 : (double)nsteps * h);
        for (int c = 0; c < 3; ++c) {
            o.pa(q, c) = traj[((k < 0 ? 0 : k) * 9 + c) * N + a];
            o.pb(q, c) = traj[((k < 0 ? 0 : k) * 9 + c) * N + b];
        }
        return;
    }
    double qa[3][9], qb[3][9];
    for (int j = 0; j < 3; ++j)
        for (int c = 0; c < 3; ++c) {
            const int64_t base = (k - 1 + j) * 9;
            qa[c][3 * j + 0] = traj[(base + c) * N + a];
            qa[c][3 * j + 1] = traj[(base + 3 + c) * N + a];
            qa[c][3 * j + 2] = traj[(base + 6 + c) * N + a];
            qb[c][3 * j + 0] = traj[(base + c) * N + b];
            qb[c][3 * j + 1] = traj[(base + 3 + c) * N + b];
            qb[c][3 * j + 2] = traj[(base + 6 + c) * N + b];
        }
    const Hermite HA[3] = { Hermite(qa[0], h), Hermite(qa[1], h), Hermite(qa[2], h) };
    const Hermite HB[3] = { Hermite(qb[0], h), Hermite(qb[1], h), Hermite(qb[2], h) };
    // FindRoot (interpolation.cuh:125-152): secant from (0, h), tol 1e-7, <= 100 iterations
    double x0 = 0.0, x1 = h;
    double f0 = approach(HA, HB, x0), f1 = approach(HA, HB, x1);
    double t = x1;
    for (int it = 0; it < 100; ++it) {
        if (fabs(f1 - f0) < 1e-14) {
            t = x1;
            break;
        }
        const double xn = x1 - f1 * (x1 - x0) / (f1 - f0);
        const double fn = approach(HA, HB, xn);
        if (fabs(xn - x1) < 1e-7 || fabs(fn) < 1e-7) {
            t = xn;
            break;
        }
        x0 = x1;
        f0 = f1;
        x1 = xn;
        f1 = fn;
        t = x1;
    }
    double pa[3], pb[3];
    for (int c = 0; c < 3; ++c) {
        pa[c] = HA[c].val(t);
        pb[c] = HB[c].val(t);
    }
    const double dx = pa[0] - pb[0], dy = pa[1] - pb[1], dz = pa[2] - pb[2];
    o.dist(q) = sqrt(dx * dx + dy * dy + dz * dz);
    o.time(q) = t0 + t + (double)k * h;
    for (int c = 0; c < 3; ++c) {
        o.pa(q, c) = pa[c];
        o.pb(q, c) = pb[c];
    }
}

void launch_refine(hipStream_t st, const double* traj, int64_t N, int64_t nsteps, double h,
    double t0, int64_t n, const uint32_t* a, const uint32_t* b, const int32_t* k, const double* d,
    double* out, int64_t field_stride)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(refine_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, traj, N,
        nsteps, h, t0, n, a, b, k, d, RefineOut { out, field_stride });
}

// ------------------------------------------------------------------------------------------
// Scratch of the allocation-free candidate search (ogp_conj_find_noalloc): the candidate
// counter, the candidate queue and the refinement results in one file-scope device array.
// Host code never passes `scratch` itself to a kernel (in host code the name is only a stand-in
// for the symbol): noalloc_scratch_setup asks the runtime for its device address, once per
// device, and the launches take that address.
// ------------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

__global__ void zero_counter_kernel(unsigned long long* counter) { counter[threadIdx.x] = 0; }
} // namespace

hipError_t noalloc_scratch_setup(int device, void** out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

void launch_zero_counter(hipStream_t st, unsigned long long* counter)
{
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, st, counter);
}

} // namespace ogp
