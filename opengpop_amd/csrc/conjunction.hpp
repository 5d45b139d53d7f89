// Host-visible launchers of the conjunction kernels (conjunction.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ogp {

// Screens the pair rectangle A = [a0, a0 + na) x B = [b0, b0 + nb) over steps [0, nsteps) of
// the stored trajectory traj[step][9][N] (components 0-2 are positions). Pair (ia, jb) gets
// index pair0 + (ia - a0) * nb + (jb - b0) (partition_parameters.cuh:58-77). min_dist /
// min_step may be null; with threshold > 0 pairs below the threshold are appended to the
// candidate arrays (at most cand_cap; *cand_count counts all).
void launch_screen(hipStream_t st, const double* traj, int64_t N, int64_t nsteps, int64_t a0,
    int64_t na, int64_t b0, int64_t nb, int64_t pair0, double threshold, double* min_dist,
    int32_t* min_step, uint32_t* cand_count, int64_t cand_cap, uint32_t* cand_a,
    uint32_t* cand_b, int32_t* cand_k, double* cand_d);

// Refines n candidates (a, b, sampled minimum step k and distance d). out: 8 doubles per
// candidate = distance, time (t0 + seconds), position of a, position of b at the TCA.
void launch_refine(hipStream_t st, const double* traj, int64_t N, int64_t nsteps, double h,
    double t0, int64_t n, const uint32_t* a, const uint32_t* b, const int32_t* k, const double* d,
    double* out);

} // namespace ogp
