// FP64 host oracle for the OpenGPOP propagation and conjunction path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, the smoke check of
// __graft_entry__.py and the cpu_baseline leg of bench.py may load it. The product
// (opengpop_amd/) never links, loads or calls it.
//
// It restates, in plain C++ and in the reference's own operation order, the equations of:
//   lib/force_model/src/acceleration/gravity_source_device.cu (Gottlieb gravity)
//   lib/force_model/src/acceleration/gravity_source.cu:61-262 (sums over m, then n)
//   lib/force_model/src/utility/ref_sys_trans.cu (frames), time_points.cu:26-207 (time nodes)
//   lib/force_model/src/utility/dataset_loaders.cu, include/force_model/types.cuh (data)
//   lib/force_model/src/force_model.cu:47-109 (step/stage -> time-node mapping)
//   lib/common/gpu_constants/gpu_constants/gpu_constants.cuh (RK8, GJ8, summed-Adams tables)
//   lib/gpu_startup/src/gpu_startup.cu + src/device/*.cu (RK8 and startup corrector)
//   lib/gpu_integrator/src/*.cu, lib/gpu_propagator/src/*.cu (GJ8 predictor / corrector)
//   lib/closest_approach/** (partition, screening, Hermite + secant refinement)
//   lib/perturbation_generator/src/perturbation_generator.cu (Monte-Carlo population)
// Every summation that the reference hands to cuBLAS or Thrust is done here sequentially in
// index order; the reference's own GPU reduction order is not reproducible.
//
// Pinning of each function (see tests/test_oracle_*.py): reference golden vectors and
// known-answer tests, closed forms, published SOFA test-suite values, and invariants.
#pragma once
#include <cstdint>

extern "C" {

// ---- datasets ---------------------------------------------------------------------------
// Loads JGM3.gfc, the EOP table, the two nutation tables and the leap-second list from
// data_dir. degree < 0 keeps the file's max_degree. Returns 0 on success.
int orc_open(const char* data_dir, int degree);
void orc_close(void);
const char* orc_error(void);
int orc_degree(void);
double orc_gm(void);
double orc_radius(void);

// ---- time and frames (ref_sys_trans.cu, types.cuh) ---------------------------------------
double orc_calendar_to_mjd(int y, int mo, int d, int h, int mi, double sec);
double orc_ntp_to_mjd(long long ntp);
void orc_ftable(double t, double* F14);
void orc_precession(double t, double* eps_gamma_phi_psi);
void orc_nutation(double t, double* dpsi_deps);
double orc_gast(double mjd_ut1, double mjd_tt, double dpsi, double eps_a, const double* F14);
void orc_rot(int axis, double angle, double* M9);
void orc_eop(double mjd_utc, double* rec4);
double orc_leap_seconds(double mjd_utc);
// Per-node ITRS->GCRS matrices (column-major, 9 per node) and MJD(UTC), time_points.cu:26-71.
// times: node offsets in seconds from the epoch (already shifted by the caller).
int orc_time_nodes(const int* ymdhm, double sec, const double* times, int len, double* M,
    double* mjd);

// ---- Monte-Carlo population (perturbation_generator.cu:21-79) ----------------------------
// The reference's normals come from cuRAND; the product uses Philox4x32-10 (Salmon, Moraes,
// Dror and Shaw, SC 2011) with Box-Muller, restated here. orc_philox4x32_10 is one 10-round
// block (pinned by the published known-answer vectors); orc_perturbed writes n satellites (AoS)
// at r0 with velocity v0 + strength * z, z = the normals of counter blocks (i, 0) and (i, 1).
void orc_philox4x32_10(const uint32_t* ctr4, const uint32_t* key2, uint32_t* out4);
void orc_perturbed(const double* r0, const double* v0, long n, double strength, uint64_t seed,
    double* r, double* v);

// ---- acceleration sources beyond gravity (iacceleration_source.cuh:19-33) ------------------
// Sun and Moon point masses and cannonball drag, added after gravity in every force evaluation
// of the propagations that follow orc_set_sources (mask bits 1 Sun, 2 Moon, 4 drag; C_D A / m
// in m^2/kg). orc_sun_moon: Sun and Moon GCRS positions (m) at an MJD of TT; orc_density: the
// exponential atmosphere (kg/m^3) at an altitude (m); orc_sources_accel adds the sources'
// accelerations at one point to a[3] (tb6 = that time's Sun and Moon positions).
void orc_set_sources(int mask, double drag_bc);
void orc_sun_moon(double mjd_tt, double* out6);
double orc_density(double h_m);
void orc_sources_accel(int mask, double drag_bc, const double* tb6, const double* r, const double* v,
    double* a);

// ---- gravity (gravity_source*.cu) ------------------------------------------------------
int orc_gravity_itrs(const double* r, long n, double* a);
// Non-central potential U - GM/r (degrees 2..L) at an ITRS point by an independent textbook recursion
// (Holmes-Featherstone style on normalized associated Legendre functions); used only to pin
// orc_gravity_itrs through a = grad U.
double orc_potential_itrs(const double* r);

// ---- integrator pieces (known-answer checks) ----------------------------------------------
// gpu_integrator/src/device/dev_functions.cu:34-50 + prediction.cu:53-59 (without the E).
void orc_gj_predict(const double* acc27, const double* s3, double* S3, double h, double* r3,
    double* v3);
// gpu_propagator/src/device/dev_functions.cu:52-69 (FinalizeCorrectionKernel): r3/v3 hold
// the partial sums on entry.
void orc_gj_finalize_correction(
    double* r3, double* v3, const double* acc27, const double* s3, const double* S3, double h);

// Force kinds: 0 gravity (needs orc_open), 1 none, 2 const (1,2,3), 3 poly7, 4 poly8,
// 5 poly9, 6 exp(-t), 7 damped oscillator a = -2r - 2v.
// quirks: bit 0 = reference RK8 row 7 (-234), bit 1 = reference backward stage times.
// epoch: Y,M,D,h,m + sec (only used for gravity).

// StartupCorrector pieces on given matrices (startup_test.cu): mode 0 = CalculateS0,
// 1 = one Correct(), 2 = CalculateAccelerationMatrix(true). Matrices are [9][3].
int orc_startup_piece(int mode, int kind, double h, double* pos27, double* vel27,
    double* acc27, double* s27, double* S27);

// RK8 startup prediction only (4 forward + 4 backward steps) into pos27/vel27
int orc_rk8_nodes(int kind, int quirks, const int* ymdhm, double sec, double h,
    const double* r0, const double* v0, double* pos27, double* vel27);

// Full GpuStartup::Startup for one object: outputs node matrices, s, S and acc [9][3].
int orc_startup(int kind, int quirks, const int* ymdhm, double sec, double h, int n_corr,
    const double* r0, const double* v0, double* pos27, double* vel27, double* s27,
    double* S27, double* acc27);

// Propagate n objects (AoS r/v in, out) through steps[i] steps of size h[i] with n_corr
// corrections, exactly the GpuPropagator::Step sequence. If traj != nullptr it receives,
// for objects with equal step counts, per-step states [step][obj][9] = r, v, a (a as the
// conjunction path sees it: acceleration-matrix column 8 after the step).
// threads <= 0 uses the OpenMP default. Returns 0 on success.
int orc_propagate(int kind, int quirks, const int* ymdhm, double sec, long n, const double* r0,
    const double* v0, const double* h, const int* steps, int n_corr, double* r_out,
    double* v_out, double* traj, int threads);

// ---- conjunction (closest_approach/**) ----------------------------------------------------
// PartitionParameters::Calculate. out arrays sized n_ids. Returns numSets or -1 if invalid.
int orc_partition(const unsigned* set_ids, int n_ids, unsigned* set_offsets,
    unsigned* num_higher, unsigned* map_offsets, unsigned* num_pairs);
int orc_partition_idx(const unsigned* set_ids, int n_ids, unsigned a, unsigned b);
int orc_partition_pair(const unsigned* set_ids, int n_ids, unsigned idx, unsigned* a,
    unsigned* b);
// HermitePolynomial (interpolation.cuh:27-98): coefficients and value/derivative.
void orc_hermite(const double* pva9, double dt, double* coeffs9);
double orc_hermite_eval(const double* coeffs9, double dt, double t, int deriv);
// FindRoot (interpolation.cuh:127-152) on the polynomial sum_i coef[i] x^i, from (x0, x1).
double orc_find_root_poly(const double* coef, int ncoef, double x0, double x1, double tol);
// ApproachFunc (interpolation.cuh:100-125) of objects A and B, each given per axis as 9 Hermite
// data (p, v, a at 0, dt, 2 dt): f(t) when root == 0, else FindRoot(f, t, x1, tol).
double orc_approach_func(const double* qa27, const double* qb27, double dt, double t, double x1,
    double tol, int root);
// Screening over stored trajectories traj[step][obj][3 or 9] (stride comps):
// for every cross-set pair p: min distance (first occurrence) and its step.
int orc_screen(const double* traj, int comps, int n_obj, int n_steps, const unsigned* set_ids,
    double* min_dist, int* min_step);
// The same rule in a blocked loop order for large screens: min_dist for every pair, min_step for
// the pairs below threshold (every pair when threshold <= 0), -1 for the others.
int orc_screen_blocked(const double* traj, int comps, int n_obj, int n_steps,
    const unsigned* set_ids, double threshold, double* min_dist, int* min_step, int threads);
// Refine pair (a, b) whose sampled minimum is at step k: TCA and miss distance and the
// positions of A and B at TCA (CalculateClosestApproachesKernel, device.cu:93-173).
int orc_refine(const double* traj, int n_obj, int n_steps, double h, unsigned a, unsigned b,
    int k, double min_dist, double* out8);
}
