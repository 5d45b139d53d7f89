// Host-visible launchers of the propagation kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace ogp {

// Frame datasets on the device (time_points.cu:73-124 uploads the same tables).
struct FrameDev {
    const double* eop;      // 4 doubles per record: mjd, xp, yp, ut1-utc
    int64_t n_eop;
    const double* leap_mjd; // leap-second table
    const double* leap_sec;
    int n_leap;
    const double* nut[4];   // lon j=0, lon j=1, obl j=0, obl j=1: 16 doubles per term
    int n_nut[4];           // (sin, cos, 14 multipliers)
};

// Structure-of-arrays state of all objects. Component c of object i lives at [c * N + i];
// history slot q (acceleration at time index j with q = j mod 9) at [(q * 3 + c) * N + i].
struct PropState {
    double* r;
    double* v;
    double* s;
    double* S;
    double* hist;
    const int64_t* nsteps;
    int64_t N;
};

// Per-object startup work areas ([slot][component][object]).
struct StartupScratch {
    double* node_r; // 9 x 3
    double* node_v; // 9 x 3
    double* acc;    // 9 x 3
    double* sm;     // 9 x 3
    double* Sm;     // 9 x 3
    double* Kx;     // 10 x 3
    double* Kv;     // 10 x 3
};

// Monte-Carlo population (perturb.hip): n satellites at r0 with velocity v0 + strength * N(0, 1)
// per component, normals from Philox4x32-10 keyed by (k0, k1) with the satellite index as
// counter; with `append`, (obj_r, obj_v) is written as entry n. AoS output ([i][3]) when
// stride == 0, else SoA with component c of entry i at [c * stride + i] (stride >= n + append).
struct PerturbArgs {
    double r0[3], v0[3], obj_r[3], obj_v[3];
    double strength;
    uint32_t k0, k1;
    int64_t n;
    int append;
    double* r;
    double* v;
    int64_t stride;
};
void launch_perturb(hipStream_t st, const PerturbArgs& a);

// Kernel variants: force kind (ForceKind; kGravitySplit for the split-precision gravity).
// Per node: the ITRS->GCRS matrix (9, column-major), MJD(UTC) and, when d_tb is not null, the
// Sun and Moon GCRS positions (6) of the third-body sources.
void launch_time_nodes(hipStream_t st, const FrameDev& fd, const int* ymdhm, double sec_shift,
    const double* d_offsets, int64_t len, double* d_M, double* d_mjd, double* d_tb = nullptr);

// Gravity of n AoS positions: in the ITRS (d_M null), or in the GCRS at the node of matrix d_M
// plus the sources of `src` with that node's Sun/Moon positions d_tb and AoS velocities d_v
// (null = zero).
void launch_gravity_aos(hipStream_t st, bool split, const GravityDev& g, const double* d_M,
    const double* d_r, int64_t n, double* d_a, int64_t out_stride, int block,
    const SourcesDev& src = SourcesDev {}, const double* d_tb = nullptr,
    const double* d_v = nullptr);

// Startup and GJ8 kernels (propagate.hip). `waves` = W waves per group of 64 objects
// (1, 2, 4 or 8; analytic test forces always use 1). Results do not depend on W.
// fixed_point_exit: the startup corrections of a 64-object group stop once the corrected nodes
// repeat the nodes of the last acceleration matrix bit for bit for every object that runs (the
// results are those of all ncorr corrections).
void launch_startup(hipStream_t st, int kind, const ForceCtx& f, int rk8_variant,
    const PropState& ps, const StartupScratch& sc, int64_t obj0, int64_t nobj, int ncorr,
    double* traj, int64_t traj_steps, int waves, bool fixed_point_exit = false);

// GJ8 steps k0 .. k1-1 (k0 >= 5) of objects [obj0, obj0 + nobj). With traj != null, steps
// traj_first .. traj_first + traj_steps - 1 are stored at rows 0.. of traj[row][9][N].
// lockstep = 2 puts two object groups in each workgroup (gravity kinds with 4 or 8 waves; else
// one). fixed_point_exit (gravity kinds, one group per workgroup): a step of a group of 64
// objects ends its corrector evaluations once the point to evaluate next is, for every object
// that stores the step, bit for bit the point just evaluated (every further evaluation would
// repeat the last one, so the results are those of all ncorr evaluations). evals (optional)
// accumulates, over the steps, the evaluations of each step times the objects that store it.
// concurrent_objects: the objects whose chunks run on the device at the same time as this one
// (this launch's and those of the other slices of its step-size class; 0 = this launch's
// alone); it chooses the register budget of the 8-wave gravity instances.
void launch_chunk(hipStream_t st, int kind, const ForceCtx& f, const PropState& ps,
    int64_t obj0, int64_t nobj, int64_t k0, int64_t k1, int ncorr, double* traj,
    int64_t traj_first, int64_t traj_steps, int waves, int lockstep = 1,
    bool fixed_point_exit = false, unsigned long long* evals = nullptr,
    int64_t concurrent_objects = 0);

// The kParts balanced contiguous column ranges of the harmonic sums (fills kParts + 1 bounds,
// parts[0] = 0, parts[kParts] = L + 1, non-decreasing).
void column_parts(int L, int* parts);

void launch_kat_predict(hipStream_t st, const double* acc27, const double* s3, double* S3,
    int64_t n, double h, double* r3, double* v3);
void launch_kat_finalize(hipStream_t st, double* r3, double* v3, const double* acc27,
    const double* s3, const double* S3, int64_t n, double h);
// Predictor::Predict without its force evaluation (prediction.cu:49-68): the predicted r, v and S
// of AoS [n][27] histories, then the history shifted by one column (column 8 is left for the new
// acceleration).
void launch_predict_aos(hipStream_t st, double* acc27, const double* s3, double* S3, int64_t n,
    double h, double* r3, double* v3);
// [slot][component][object] (n objects) -> [object][slot][component]
void launch_slots_to_aos(hipStream_t st, const double* src, double* dst, int64_t n, int nslots);
// flags[i] = code for every object i whose SoA r or v (stride n) holds a non-finite value and
// whose flag is still 0 (fresh: every flag is taken as 0 and rewritten); *count counts the
// objects newly flagged.
void launch_finite_soa(hipStream_t st, const double* r, const double* v, int64_t n, int32_t* flags,
    uint32_t* count, int32_t code, bool fresh);

// key[i] = semi-major axis (vis-viva, as float; 0 where not finite) of SoA states r, v (stride
// n) for gravitational parameter gm.
void launch_orbit_key_soa(hipStream_t st, const double* r, const double* v, int64_t n, double gm,
    float* key);

// Permutations between the caller's order and the internal (class-sorted) order.
// mode 0: AoS src -> SoA dst gather (dst[p] = src[perm[p]]); 1: SoA -> AoS scatter;
// 2: SoA -> SoA gather; 3: SoA -> SoA scatter.
void launch_permute(hipStream_t st, int mode, const double* src, const int64_t* perm, double* dst,
    int64_t n);

// SoA <-> AoS helpers for the boundary (runner.cu:27-33 hands AoS arrays).
void launch_aos_to_soa(hipStream_t st, const double* aos, double* soa, int64_t n);
void launch_soa_to_aos(hipStream_t st, const double* soa, double* aos, int64_t n);

// Derivatives (adjoint.hip). The startup's derivative runs on 6 tangent lanes per object (lane
// col * n + o: the derivative of object o's state with respect to initial component col), with
// the Jacobians a kGravityRecord startup of the objects recorded (kRecordedTangent).
void launch_gravity_jac(hipStream_t st, const GravityDev& g, const double* r, int64_t n,
    double* a, double* G);
void launch_tangent_init(hipStream_t st, const int64_t* steps, int64_t n, const PropState& pt,
    int64_t* ns_out);
void launch_tangent_columns(hipStream_t st, const PropState& pt, int64_t n, double* Zd,
    double* RVd);
void launch_checkpoint(hipStream_t st, const PropState& ps, double* ck, bool to_checkpoint);
// Steps k0 .. k1-1 of the plain state ps with the gravity Jacobian of every corrector iterate
// stored (6 unique entries); the state is advanced as by the chunk kernel. `waves` (4 below 8,
// else 8) splits the harmonic columns as in the chunk kernel; results do not depend on it.
// fixed_point_exit: as launch_chunk; the iterates after the fixed point store its Jacobian.
void launch_gj8_record(hipStream_t st, const ForceCtx& f, const PropState& ps, int64_t k0,
    int64_t k1, int ncorr, double* Gbuf, int waves, bool fixed_point_exit = false);
void launch_gj8_reverse(hipStream_t st, const int64_t* nsteps, int64_t n, double h, int64_t k0,
    int64_t k1, int ncorr, const double* Gbuf, int R, const double* seeds, double* lam);
void launch_adjoint_combine(hipStream_t st, const int64_t* nsteps, int64_t n, int R,
    const double* lam, const double* Zd, const double* RVd, const double* seeds, double* out);
// Forward tangent pass over steps k0 .. k1-1 with the Jacobians recorded for those steps:
// Zd / RVd (tangent_columns layout) are advanced in place.
void launch_gj8_tangent(hipStream_t st, const int64_t* nsteps, int64_t n, double h, int64_t k0,
    int64_t k1, int ncorr, const double* Gbuf, double* Zd, double* RVd);
void launch_stm_final(hipStream_t st, const int64_t* nsteps, int64_t n, const double* RVd,
    double* stm);
int adjoint_max_corrections();

} // namespace ogp
