// Host-visible launchers of the conjunction kernels (conjunction.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ogp {

// Screens the pair rectangle A = [a0, a0 + na) x B = [b0, b0 + nb) over steps [0, nsteps) of
// the stored trajectory traj[step][9][N] (components 0-2 are positions). Pair (ia, jb) gets
// index pair0 + (ia - a0) * row_len + (jb - b0) (partition_parameters.cuh:58-77; row_len <= 0
// means nb). min_dist / min_step may be null; with threshold > 0 pairs below the threshold are
// appended to the candidate arrays (at most cand_cap; *cand_count counts all).
// shape: kScreenSquare (64 x 64 pair tiles) or kScreenNarrow (4 A objects x 256 B objects);
// the results are the same bits.
enum { kScreenSquare = 1, kScreenNarrow = 2 };
// A threshold screen on radially sorted input: the positions of every object reordered within
// its set (pos[step][3][N], 3 N doubles per step), each slot's smallest and largest sampled
// radius, and the object index each slot holds. A tile whose A and B objects' radius ranges lie
// more than the threshold (plus kRadialMargin) apart holds no pair closer than the threshold at
// any sample (|rA - rB| >= ||rA| - |rB||) and is skipped; candidates are reported with the
// object indices. Only with threshold > 0 and no per-pair minima.
struct ScreenSorted {
    const double* pos = nullptr; // null: screen the stored trajectory, every tile
    const double* rlo = nullptr;
    const double* rhi = nullptr;
    const uint32_t* ids = nullptr;
};
// Far above the rounding of the radii and of the squared distances (~1e-9 m at 10^7 m).
constexpr double kRadialMargin = 1e-3;
void launch_screen(hipStream_t st, const double* traj, int64_t N, int64_t nsteps, int64_t a0,
    int64_t na, int64_t b0, int64_t nb, int64_t pair0, double threshold, double* min_dist,
    int32_t* min_step, unsigned long long* cand_count, int64_t cand_cap, uint32_t* cand_a,
    uint32_t* cand_b, int32_t* cand_k, double* cand_d, int shape = kScreenSquare,
    int64_t row_len = 0, const ScreenSorted& sorted = ScreenSorted {});

// The smallest and largest radius of each object over the samples [0, nsteps) of traj:
// partial[2][chunks][N] is scratch for kRadialChunks step chunks.
constexpr int kRadialChunks = 16;
void launch_radial_bounds(hipStream_t st, const double* traj, int64_t N, int64_t nsteps,
    double* partial, double* rlo, double* rhi);
// pos[k][c][s] = traj[k][c][ids[s]] (positions) and the bounds in slot order.
void launch_gather_sorted(hipStream_t st, const double* traj, int64_t N, int64_t nsteps,
    const uint32_t* ids, const double* rlo, const double* rhi, double* pos, double* slo,
    double* shi);

// Refines n candidates (a, b, sampled minimum step k and distance d). out: distance, time
// (t0 + seconds), position of a, position of b at the TCA; field_stride = 0: 8 doubles per
// candidate; field_stride = S: blocks distance[S], time[S], pos_a[S][3], pos_b[S][3].
void launch_refine(hipStream_t st, const double* traj, int64_t N, int64_t nsteps, double h,
    double t0, int64_t n, const uint32_t* a, const uint32_t* b, const int32_t* k, const double* d,
    double* out, int64_t field_stride = 0);

// Allocation-free candidate search (ogp_conj_find_noalloc): one file-scope device array per
// device holds, at these byte offsets, the candidate counter, kNoAllocMaxCandidates candidates
// (a, b, step, sampled distance) and their refinement results (field blocks of that stride).
constexpr int64_t kNoAllocMaxCandidates = 1 << 16;
constexpr size_t kNoAllocCounter = 0;
constexpr size_t kNoAllocA = 256;
constexpr size_t kNoAllocB = kNoAllocA + 4 * kNoAllocMaxCandidates;
constexpr size_t kNoAllocK = kNoAllocB + 4 * kNoAllocMaxCandidates;
constexpr size_t kNoAllocD = kNoAllocK + 4 * kNoAllocMaxCandidates;
constexpr size_t kNoAllocOut = kNoAllocD + 8 * kNoAllocMaxCandidates;
constexpr size_t kNoAllocBytes = kNoAllocOut + 8 * 8 * kNoAllocMaxCandidates;
static_assert(kNoAllocD % 8 == 0 && kNoAllocOut % 8 == 0, "8-byte fields");
// The array's device address on `device` (which must be the current device): looked up with
// hipGetSymbolAddress on the first call for that device and kept.
hipError_t noalloc_scratch_setup(int device, void** out);
void launch_zero_counter(hipStream_t st, unsigned long long* counter);

} // namespace ogp
