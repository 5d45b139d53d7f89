/* OpenGPOP-MI355X: C ABI of the propagation and conjunction engine.
 *
 * One context = one device (device 0) + its HIP streams (one for the calls, four for the
 * concurrent object slices of a propagation, all at the default priority) + the loaded
 * datasets. A context is not thread-safe; separate contexts are independent. Every call is
 * synchronous at return, like the reference (which synchronises after nearly every launch).
 * Errors never exit or throw
 * across this boundary: every entry point returns OGP_OK or an error code, and
 * ogp_last_error(ctx) returns the message (ogp_last_error(NULL) the last context-less one).
 *
 * Units and layouts follow the reference: metres, m/s, seconds, UTC calendar epoch, and
 * array-of-structures xyz triples per object (runner.cu:27-33). Column-major 3x3 matrices.
 */
#ifndef OGP_H
#define OGP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OGP_OK = 0,
    OGP_ERR_ARGUMENT = 1, /* invalid sizes, set ids, step sizes, degree above file maximum */
    OGP_ERR_DATA = 2,     /* missing or malformed dataset file */
    OGP_ERR_DEVICE = 3,   /* HIP runtime error (includes out-of-memory) */
    OGP_ERR_STATE = 4,    /* call out of sequence (e.g. step before init) */
    OGP_ERR_NONFINITE = 5 /* non-finite input state */
};

/* Force kinds. OGP_FORCE_GRAVITY is the product force model; the analytic kinds reproduce the
 * reference's test forces (unit_tests/include/gpu_test_helpers.cuh) with the reference mock's
 * time mapping (unit_tests/include/mock_force_model.cuh:69-73, :99). */
enum {
    OGP_FORCE_GRAVITY = 0,
    OGP_FORCE_NONE = 1,
    OGP_FORCE_CONST = 2,
    OGP_FORCE_POLY7 = 3,
    OGP_FORCE_POLY8 = 4,
    OGP_FORCE_POLY9 = 5,
    OGP_FORCE_EXP = 6,
    OGP_FORCE_DAMPED = 7
};

/* Precision of the gravity harmonics. */
enum { OGP_PRECISION_FP64 = 0, OGP_PRECISION_SPLIT = 1 };

/* Reference behaviours kept by default (bit set = reference behaviour). */
enum {
    OGP_QUIRK_RK8_ROW7 = 1,       /* RK8 a[6][0] = -234/20 (gpu_constants.cuh:31) */
    OGP_QUIRK_BACKWARD_TIMES = 2, /* backward RK8 stages at (s + c_k) h (force_model.cu:69-71) */
    OGP_QUIRKS_REFERENCE = 3
};

typedef struct ogp_config {
    const char* data_dir;  /* folder with JGM3.gfc, EOP, nutation and leap-second files */
    int degree;            /* gravity degree/order; < 0 = the file's max_degree */
    int precision;         /* OGP_PRECISION_* */
    int split_degree;      /* split mode: degrees <= split_degree stay FP64 (default 2) */
    int n_corrections;     /* GJ8 / startup corrector iterations; reference: 7 */
    int quirks;            /* OGP_QUIRK_* bits */
    int force;             /* OGP_FORCE_* */
    int year, month, day, hour, minute; /* epoch, UTC (TimeData, Parameters.hpp:19-26) */
    double second;
} ogp_config;

typedef struct ogp_context ogp_context;

/* Fills cfg with the reference defaults (degree 70, FP64, 7 corrections, reference quirks,
 * gravity, epoch 2025-01-01T00:00:00). */
void ogp_default_config(ogp_config* cfg);

/* Replaces ForceModel::PreInitialize (force_model.cu:23-45): device selection and dataset
 * loading (FMTimePoints::Initialize time_points.cu:73-124, GravitySource::Initialize
 * gravity_source.cu:61-135). */
int ogp_create(const ogp_config* cfg, ogp_context** out);
void ogp_destroy(ogp_context* ctx);
const char* ogp_last_error(const ogp_context* ctx);
/* Size of the library's device allocations (bytes). */
int64_t ogp_device_bytes(const ogp_context* ctx);

/* ---- force model boundary --------------------------------------------------------------- */

/* GravitySource::ComputeAcceleration with an identity frame (the reference's gravity tests,
 * fm_gravity_source_test.cu:43-47): ITRS positions -> ITRS accelerations, host arrays. */
int ogp_gravity_itrs(ogp_context* ctx, const double* r, int64_t n, double* a);

/* ForceModel::ComputePropagationAcceleration (force_model.cu:91-109): GCRS positions at
 * propagation node `node` of a run with step `step_s` (node time (node - 4) * step_s relative
 * to the epoch) -> GCRS accelerations: gravity plus the enabled sources (ogp_set_sources).
 * out_stride = 3 * reference outStride (doubles between consecutive objects' outputs).
 * v (AoS velocities) may be null unless drag is enabled. Host arrays. */
int ogp_accel(ogp_context* ctx, const double* r, const double* v, int64_t n, double step_s,
    int64_t node, double* a, int64_t out_stride);

/* Acceleration sources added to gravity: the reference's IAccelerationSource extension point
 * (iacceleration_source.cuh:19-33), whose outputs ForceModel sums (force_model.cu:73-86); the
 * reference implements gravity only and names drag and third bodies as further sources. Here
 * the sources run inside the integrator kernels after gravity, added in the order Sun, Moon,
 * drag, in every propagation, conjunction and ogp_accel call of the context:
 *   OGP_SOURCE_SUN, OGP_SOURCE_MOON: point masses (GM 1.32712440018e20 and 4.9028e12 m^3/s^2)
 *     at low-precision analytic positions (Montenbruck and Gill 2000, eq. 3.43-3.49) computed
 *     once per time node, a += GM ((s - r)/|s - r|^3 - s/|s|^3);
 *   OGP_SOURCE_DRAG: a cannonball of ballistic coefficient drag_bc = C_D A / m (m^2/kg) in an
 *     exponential atmosphere (Vallado, Table 8-4) rotating with the Earth,
 *     a += -1/2 drag_bc rho |v_r| v_r, v_r = v - omega_E z x r.
 * mask = 0 (the default) is the reference's gravity-only model, bit for bit. The derivative
 * entry points (ogp_stm, ogp_adjoint_*) require mask = 0. Gravity contexts only. */
enum {
    OGP_SOURCE_SUN = 1,
    OGP_SOURCE_MOON = 2,
    OGP_SOURCE_DRAG = 4
};
typedef struct ogp_sources {
    int mask;       /* OGP_SOURCE_* bits */
    double drag_bc; /* C_D A / m of every object, m^2/kg */
} ogp_sources;
int ogp_set_sources(ogp_context* ctx, const ogp_sources* sources);
int ogp_get_sources(const ogp_context* ctx, ogp_sources* sources);

/* Per-object ballistic coefficients for OGP_SOURCE_DRAG (no counterpart in the reference):
 * bc[i] = C_D A / m (m^2/kg) of object i in the caller's order. While drag is enabled they replace
 * ogp_sources.drag_bc in every later call of the context on n objects (propagation, stepwise
 * propagator and closest-approach search from their init calls, ogp_startup, ogp_accel); a call
 * on another number of objects is then an error. bc = NULL or n = 0 returns to the single
 * coefficient. */
int ogp_set_drag_coefficients(ogp_context* ctx, const double* bc, int64_t n);

/* Device-pointer forms of the force-model boundary, for callers that keep their state in HBM
 * (the reference's IForceModel takes device pointers, iforce_model.cuh:19-33). Arrays are AoS
 * [n][3] device buffers; nothing crosses PCIe.
 *
 * ForceModel::ComputePropagationAcceleration (force_model.cu:91-109): as ogp_accel on device
 * arrays; d_a[i * out_stride + c] receives object i's component c. */
int ogp_accel_device(ogp_context* ctx, const double* d_r, const double* d_v, int64_t n,
    double step_s, int64_t node, double* d_a, int64_t out_stride);

/* ForceModel::ComputeGenerationAcceleration (force_model.cu:59-89): the acceleration at RK8
 * stage k_index (0..9) of startup step step_num (+-1 .. +-4) of a run with step step_s, written
 * into column k_index of the caller's K matrix d_K [n][10][3]. The stage's time node is the
 * reference's ((step_num > 0 ? step_num - 1 : step_num) + 4) * 10 + k_index (with
 * OGP_QUIRK_BACKWARD_TIMES, the reference's mirrored times of backward stages). */
int ogp_accel_generation(ogp_context* ctx, const double* d_r, const double* d_v, int64_t n,
    double step_s, int step_num, int k_index, double* d_K);

/* FMTimePoints::CalculateParameters + GetParametersDev (time_points.cu:126-207): the
 * ITRS->GCRS matrix (column-major, 9 per node) and MJD(UTC) of `len` nodes at offsets
 * offsets_s (seconds) from (epoch - before * step_s). */
int ogp_time_nodes(ogp_context* ctx, double step_s, int before, const double* offsets_s,
    int64_t len, double* M, double* mjd);

/* ---- propagation ----------------------------------------------------------------------- */

/* The runner's propagate() (exec/runner.cu:11-70) for n objects with per-object step sizes
 * and step counts (uneven spans, mixed regimes): GpuStartup::Startup followed by
 * GpuPropagator::Step up to steps[i] for each object. Host arrays in/out. */
int ogp_propagate(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    const double* step_s, const int64_t* steps, double* r_out, double* v_out);

/* Same on device-resident SoA arrays (x[n], y[n], z[n] blocks of the given pointers);
 * nothing crosses PCIe. */
int ogp_propagate_device(ogp_context* ctx, const double* d_r0_soa, const double* d_v0_soa,
    int64_t n, const double* step_s, const int64_t* steps, double* d_r_soa, double* d_v_soa);

/* Per-object failure flags (no counterpart in the reference, which exits the process on any
 * error, gpu_utils.cuh:24-48). As ogp_propagate / ogp_propagate_device, but instead of failing
 * the call: an object whose initial state holds a non-finite value is skipped (its output is its
 * input) and flagged OGP_OBJECT_NONFINITE_INPUT; one whose final state is non-finite is flagged
 * OGP_OBJECT_NONFINITE_RESULT. status (host, n entries, caller order) receives the flags; every
 * other object's result is bit-identical to a run without the flagged objects. (The unchecked
 * entry points reject a non-finite initial state with OGP_ERR_NONFINITE; ogp_propagate_device
 * checks it on the device.) */
enum {
    OGP_OBJECT_OK = 0,
    OGP_OBJECT_NONFINITE_INPUT = 1,
    OGP_OBJECT_NONFINITE_RESULT = 2
};
int ogp_propagate_checked(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    const double* step_s, const int64_t* steps, double* r_out, double* v_out, int32_t* status);
int ogp_propagate_device_checked(ogp_context* ctx, const double* d_r0_soa, const double* d_v0_soa,
    int64_t n, const double* step_s, const int64_t* steps, double* d_r_soa, double* d_v_soa,
    int32_t* status);

/* PerturbationGenerator::generatePerturbedSatellites (perturbation_generator.cu:40-79,
 * perturbation_generator.cuh:24-43): n satellites at the epoch position r0 with velocity
 * v0 + strength * N(0, 1) per component, followed by (obj_r, obj_v) as entry n when both are
 * non-null (n + 1 entries). The normals come from Philox4x32-10 keyed by `seed`, with the
 * satellite index as counter, through the Box-Muller transform; the reference draws them with
 * cuRAND (seed 1234), so the two streams differ while the distribution is the same. Entry i
 * depends only on (seed, i). Host AoS arrays of n (+ 1) x 3 doubles. */
int ogp_perturbed_satellites(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    double strength, uint64_t seed, const double* obj_r, const double* obj_v, double* r_out,
    double* v_out);

/* The same into device-resident SoA arrays: component c of entry i at [c * stride + i]
 * (stride >= n + 1 with the appended object, else >= n), the layout ogp_propagate_device
 * takes with stride = number of objects. */
int ogp_perturbed_satellites_device(ogp_context* ctx, const double* r0, const double* v0,
    int64_t n, double strength, uint64_t seed, const double* obj_r, const double* obj_v,
    double* d_r_soa, double* d_v_soa, int64_t stride);

/* Timing of the last propagate call, from HIP events on the context stream. */
typedef struct ogp_timing {
    double total_ms;       /* all device work of the call (nodes + startup + steps) */
    double nodes_ms;       /* time-node (frame) kernels */
    double startup_ms;     /* RK8 + startup corrector */
    double steps_ms;       /* GJ8 chunk kernels */
    double h2d_ms, d2h_ms; /* host copies (host-array entry points only) */
    double step_launches;  /* number of GJ8 chunk launches */
    double object_steps;   /* sum over objects of GJ8 + startup-copy steps */
    double gj_object_steps; /* object-steps executed by the GJ8 chunk kernels */
    double launch_ms_sum;   /* sum of the GJ8 chunk launches' own durations; launches of
                               concurrent object slices overlap, so this exceeds steps_ms */
    double evaluations;     /* gravity evaluations the GJ8 chunk kernels ran, counted per object
                               that stores the step (at most n_corrections per object-step) */
} ogp_timing;
int ogp_last_timing(const ogp_context* ctx, ogp_timing* t);

/* Launch shape of the propagation kernels (no counterpart in the reference, which launches
 * 256-thread blocks per helper kernel): `waves` = waves per group of 64 objects that share one
 * object group's harmonic columns (1, 2, 4, 8; 0 = autotuned), chunk_steps = GJ8 steps per
 * launch (0 = 64). Results are bit-identical for every choice. With waves = 0 the context times
 * two chunks per candidate (in mirrored order) the first time it meets a group count (per power
 * of two) and keeps the fastest; ogp_set_tuning clears that table. A shape that gives fewer waves than the device has
 * SIMDs is not a candidate unless it is the widest, so small populations take 8 waves without
 * trials. ogp_get_tuning reports the choice for n objects. */
int ogp_set_tuning(ogp_context* ctx, int waves, int64_t chunk_steps);
int ogp_get_tuning(const ogp_context* ctx, int64_t n, int* waves, int64_t* chunk_steps);

/* Object groups per workgroup of the GJ8 chunks (no counterpart in the reference): 2 runs two
 * groups of 64 objects in lockstep in one workgroup (with 4 or 8 waves per group), so the waves
 * that sum the same harmonic columns for different groups read the same table lines at about
 * the same time; 1 (and 0, the default) runs one. With 2 the tuner chooses the waves among the
 * paired shapes. Results are bit-identical for every choice. ogp_get_lockstep reports the
 * choice for n objects. */
int ogp_set_lockstep(ogp_context* ctx, int groups);
int ogp_get_lockstep(const ogp_context* ctx, int64_t n, int* groups);

/* Corrector evaluations of a GJ8 step (no counterpart in the reference, which always runs
 * n_corrections, gpu_propagator.cu:102-117, and n_corrections startup corrections,
 * gpu_startup.cu:63-65). With on = 1 (the default) a step of a group of 64 objects stops
 * evaluating once the state to evaluate next is, for every object of the group that stores the
 * step, bit for bit the state just evaluated: each further evaluation and correction would repeat
 * the last one exactly, so every result is bit-identical to running all n_corrections
 * evaluations (typically 2-4 of 7 are run). The startup corrections stop alike once the corrected
 * nodes repeat the nodes of the last acceleration matrix, and the derivative passes record the
 * fixed point's Jacobian for the remaining iterates. 0 always runs n_corrections. Lockstep groups
 * (ogp_set_lockstep 2) always run n_corrections. */
int ogp_set_fixed_point_exit(ogp_context* ctx, int on);

/* Object slices per step-size class (no counterpart in the reference): each class is cut into
 * `slices` contiguous slices whose GJ8 chunk sequences run concurrently on streams of their own,
 * so the end of one launch overlaps the launches of other slices. 0 = automatic (2 for a class
 * of at least 4096 objects, unless it is the call's only class and has no more 64-object groups
 * than the device has compute units; else 1), at most 16. Results are bit-identical for every
 * choice. */
int ogp_set_slices(ogp_context* ctx, int slices);

/* GpuPropagator (gpu_propagator.cuh:26-59, gpu_propagator.cu:28-154): a stepwise propagator
 * of n objects with one step size. Each propagator owns its integrator state (s, S, the
 * acceleration history), so any number of them, and any other call of the context, may be
 * interleaved. It belongs to its context: destroy it first (a context destroyed before it
 * frees its device memory, and its later calls return OGP_ERR_STATE).
 *
 * ogp_propagator_create: the GpuPropagator constructor. d_r, d_v are the caller's AoS [n][3]
 * device buffers: the first step reads the epoch state from them, and every step writes the
 * state at that step back into them, in place (gpu_propagator.cu:128-129 advances them).
 * ogp_propagator_create_host: the same from host arrays; the propagator then owns the device
 * buffers.
 * ogp_propagator_step: GpuPropagator::Step(hostPositions, hostVelocities, devAcceleration):
 * the startup on the first call, the startup nodes for steps 1-4, a GJ8 step from step 5; then
 * optionally the state to host arrays and the acceleration column the reference hands to the
 * conjunction screen (gpu_propagator.cu:68-72) to a device array d_acc [n][3]. Each may be null.
 * ogp_propagator_acceleration: that acceleration column of the last step, to a host array.
 * ogp_propagator_copy_state_to_host: GpuPropagator::CopyStateToHost.
 * ogp_propagator_destroy: GpuPropagator::FreeAllMemory and the destructor. */
typedef struct ogp_propagator ogp_propagator;
int ogp_propagator_create(ogp_context* ctx, double* d_r, double* d_v, int64_t n, double step_s,
    int64_t max_steps, ogp_propagator** out);
int ogp_propagator_create_host(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    double step_s, int64_t max_steps, ogp_propagator** out);
int ogp_propagator_step(ogp_propagator* p, double* host_r, double* host_v, double* d_acc);
int ogp_propagator_acceleration(const ogp_propagator* p, double* host_a);
int ogp_propagator_copy_state_to_host(const ogp_propagator* p, double* host_r, double* host_v);
void ogp_propagator_destroy(ogp_propagator* p);

/* The context's own stepwise propagator (one per context, host arrays): ogp_prop_init
 * replaces it, ogp_prop_step performs one GpuPropagator::Step and returns r, v and the
 * acceleration column as ogp_propagator_step / _acceleration do; the arrays hold the n objects
 * of the last ogp_prop_init. Its state is its own, as above. */
int ogp_prop_init(ogp_context* ctx, const double* r0, const double* v0, int64_t n, double step_s,
    int64_t max_steps);
int ogp_prop_step(ogp_context* ctx, double* r_out, double* v_out, double* a_out);

/* GpuStartup::Startup (gpu_startup.cu:55-68) for n objects: node matrices, s, S and the
 * acceleration matrix, each [n][9][3] (host). */
int ogp_startup(ogp_context* ctx, const double* r0, const double* v0, int64_t n, double step_s,
    double* pos, double* vel, double* s, double* S, double* acc);
/* The same on device arrays (the FullStartupData of gpu_startup.cuh:95-120): d_r0, d_v0 AoS
 * [n][3]; outputs [n][9][3] device arrays, each may be null. */
int ogp_startup_device(ogp_context* ctx, const double* d_r0, const double* d_v0, int64_t n,
    double step_s, double* d_pos, double* d_vel, double* d_s, double* d_S, double* d_acc);

/* Predictor::Predict(stepNum) (prediction.cu:49-74) on the device arrays Predictor::SetPointers
 * names: AoS d_r, d_v [n][3] receive the predicted state, d_S [n][3] is advanced, d_s [n][3] is
 * read, and the history d_acc [n][9][3] is shifted by one column with the acceleration at
 * propagation node step_num + 4 (gravity plus the enabled sources) written into column 8. */
int ogp_predict_device(ogp_context* ctx, double* d_r, double* d_v, double* d_acc,
    const double* d_s, double* d_S, int64_t n, double step_s, int64_t step_num);

/* Known-answer hooks that run the device's own predictor and corrector arithmetic on given
 * inputs (integrator_prediction_test.cu, integrator_test.cu:207-245). n objects, host arrays
 * [n][27] / [n][3]. */
int ogp_kat_predict(ogp_context* ctx, const double* acc27, const double* s3, double* S3,
    int64_t n, double h, double* r3, double* v3);
int ogp_kat_finalize_correction(ogp_context* ctx, double* r3, double* v3, const double* acc27,
    const double* s3, const double* S3, int64_t n, double h);

/* ---- conjunction (ClosestApproach, closest_approach.cuh:36-88) --------------------------- */

/* ClosestApproach::Initialize: set ids must start at 0, be non-decreasing with steps of +1
 * (partition_parameters.cuh:94-135); at least two sets. A context holds one search: calling
 * it again starts a new one (objects, trajectory and results are replaced). */
int ogp_conj_init(ogp_context* ctx, const double* r, const double* v, const uint32_t* set_ids,
    int64_t n, double step_s);
/* ClosestApproach::Propagate: continues from the current state; may be called repeatedly. */
int ogp_conj_propagate(ogp_context* ctx, int64_t n_steps);
/* ClosestApproach::FindClosest. threshold_m <= 0 refines every cross-set pair (reference
 * behaviour); threshold_m > 0 screens and refines only pairs whose sampled minimum distance is
 * below the threshold. Returns the number of result records in *count. */
int ogp_conj_find(ogp_context* ctx, double threshold_m, int64_t* count);
/* Copies the records of the last ogp_conj_find, sorted by pair index: object indices (a < b),
 * sampled minimum step, refined miss distance, TCA (s from the epoch) and positions at TCA. */
int ogp_conj_results(ogp_context* ctx, uint32_t* obj_a, uint32_t* obj_b, int32_t* step,
    double* distance, double* time, double* pos_a, double* pos_b);
/* Allocation-free form of ogp_conj_find with threshold_m > 0 (no counterpart in the reference;
 * for callers that cannot allocate device memory: tight serving loops, code that must not call
 * hipMalloc). It calls no device allocator: the candidate counter, the candidate queue and the
 * refinement results live in a file-scope device array of the library (5.25 MiB per device,
 * held until the process ends), which limits a call to OGP_NOALLOC_MAX_CANDIDATES candidates.
 * The records (as ogp_conj_results gives them, sorted by pair) go straight into the caller's
 * host arrays of `capacity` entries (pos_a, pos_b: 3 per entry); *count receives the number of
 * candidates, also when it exceeds the capacity or the limit, which is an OGP_ERR_ARGUMENT.
 * The trajectory is the one ogp_conj_propagate stored. Every call on a device shares that
 * device's array: calls that overlap on one device (from several threads or contexts) race. */
#define OGP_NOALLOC_MAX_CANDIDATES 65536
int ogp_conj_find_noalloc(ogp_context* ctx, double threshold_m, int64_t capacity,
    uint32_t* obj_a, uint32_t* obj_b, int32_t* step, double* distance, double* time,
    double* pos_a, double* pos_b, int64_t* count);
/* Radial filter of threshold screens (no counterpart in the reference, which screens every
 * pair). With the filter, each set's objects are sorted by their smallest and largest sampled
 * radius and every pair tile whose two radius ranges lie more than the threshold apart is
 * skipped (for such pairs |rA - rB| >= ||rA| - |rB|| exceeds the threshold at every sample, so
 * none is a candidate). mode 0: screen every tile; 1 (the default): filter when there are
 * hundreds of pairs per object, where skipping outweighs the sort; 2: always filter. The
 * candidates, their steps and distances and the refined results are the same in every mode;
 * ogp_conj_screen and threshold_m <= 0 always screen every pair. */
int ogp_set_radial_filter(ogp_context* ctx, int mode);

/* Tile shape of the conjunction screen (no counterpart in the reference, which runs one
 * 256-thread block per pair range): 1 = square tiles of 64 x 64 pairs, 2 = narrow tiles of 4
 * objects of the lower set against 256 of the higher ones, which stream the higher set's
 * trajectory once per 4 lower objects (for a small lower set, e.g. a few objects against a
 * catalogue). 0 (the default) = tuned: the first time a context screens a lower set of at most
 * 64 objects of a size bucket (per power of two), both shapes screen it against a prefix of the
 * higher set and the faster one is kept; larger lower sets take square tiles. Results are
 * bit-identical for every choice. ogp_get_screen_shape reports the shape for a lower set of na
 * objects. */
int ogp_set_screen_shape(ogp_context* ctx, int shape);
int ogp_get_screen_shape(const ogp_context* ctx, int64_t na, int* shape);
/* Sampled screening minimum of every cross-set pair (pair index order). */
int ogp_conj_screen(ogp_context* ctx, double* min_dist, int32_t* min_step);
int ogp_conj_num_pairs(const ogp_context* ctx, int64_t* num_pairs);
/* Copies the stored trajectory [steps][n][9] = r, v, a per step (host). */
int ogp_conj_trajectory(ogp_context* ctx, double* out, int64_t* steps);
typedef struct ogp_conj_timing {
    double propagate_ms, screen_ms, refine_ms;
    double pair_steps, pairs, candidates;
    /* pair-steps the screen kernels evaluated: pair_steps less the pairs of the tiles the radial
     * filter skipped (ogp_set_radial_filter) */
    double screened_pair_steps;
} ogp_conj_timing;
int ogp_conj_last_timing(const ogp_context* ctx, ogp_conj_timing* t);

/* ---- derivatives (no counterpart in the reference) ------------------------------------------
 * Gravity-only FP64 contexts (OGP_FORCE_GRAVITY, OGP_PRECISION_FP64); one step size per call,
 * per-object step counts; host arrays. Jacobians are [n][6][6] row-major: row = final component
 * (x, y, z, vx, vy, vz), column = initial component. */

/* Gravity acceleration and its Jacobian G[i][j] = d a_i / d r_j in the ITRS ([n][3], [n][3][3]);
 * the acceleration is bit-identical to ogp_gravity_itrs. */
int ogp_gravity_jacobian_itrs(ogp_context* ctx, const double* r, int64_t n, double* a, double* G);

/* HBM budget for the gravity Jacobians the derivative passes record (no counterpart in the
 * reference): 6 doubles per object, corrector iterate and step. Up to the budget the adjoint's
 * forward pass records the whole span (beyond it, or with checkpoint_steps > 0, it keeps
 * checkpoints and records each segment again in the reverse sweep), and ogp_stm records and
 * consumes segments of at most that size. Results do not depend on it. 0 = 16 GiB (default). */
int ogp_set_derivative_memory(ogp_context* ctx, int64_t bytes);

/* State-transition matrix from the variational equations of the discrete scheme (forward mode):
 * the six tangents are carried through the same RK8 startup and GJ8 steps as the state, with
 * the gravity Jacobian of every corrector iterate evaluated once per object. Also returns the
 * final state (bit-identical to ogp_propagate). */
int ogp_stm(ogp_context* ctx, const double* r0, const double* v0, int64_t n, double step_s,
    const int64_t* steps, double* r_out, double* v_out, double* stm);

/* d(r, v)_final / d(r, v)_initial by the discrete adjoint of the GJ8 steps (reverse mode)
 * composed with the forward-mode derivative of the startup. Also returns the final state.
 * checkpoint_steps = 0: the forward pass records the gravity Jacobians of the whole span when
 * they fit in 16 GiB (6 doubles per object, step and corrector iterate), else checkpoints every
 * 64 steps; checkpoint_steps > 0: checkpoints every that many steps, each segment recorded again
 * in the reverse sweep. Results do not depend on the choice. */
int ogp_adjoint_jacobian(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    double step_s, const int64_t* steps, int64_t checkpoint_steps, double* r_out, double* v_out,
    double* jac);

/* Vector-Jacobian product by the same adjoint: given the gradients g_r, g_v ([n][3]) of a scalar
 * with respect to the final state, returns its gradients with respect to the initial state. */
int ogp_adjoint_vjp(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    double step_s, const int64_t* steps, int64_t checkpoint_steps, const double* g_r,
    const double* g_v, double* g_r0, double* g_v0);

/* Device time of the last derivative call: forward pass (variational startup, checkpointed GJ8
 * steps), adjoint pass (recording + reverse + assembly), or the whole variational run. */
typedef struct ogp_adjoint_timing {
    double forward_ms, adjoint_ms, stm_ms;
} ogp_adjoint_timing;
int ogp_adjoint_last_timing(const ogp_context* ctx, ogp_adjoint_timing* t);

#ifdef __cplusplus
}
#endif
#endif
