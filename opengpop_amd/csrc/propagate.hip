// RK8 startup and GJ8 step kernels for gfx950.
//
// Layout: a workgroup of W waves (W = 1, 2, 4 or 8) holds one group of 64 objects; lane l of
// every wave owns object 64 * blockIdx.x + l. The spherical-harmonic columns of each force
// evaluation are divided between the W waves and their partial sums meet in LDS. Wave 0 (the
// lead) alone runs the integrator, its global loads and stores, the fold of the parts and the
// corrector; it hands the geometry of each evaluation point to the other waves through LDS, and
// they only sum their parts (device.hpp, gravity_gcrs_lead / gravity_helper). Only the lead
// reads or writes the state and the startup work areas in global memory.
//
// Every workgroup barrier sits in workgroup-uniform control flow: the loops run to the group's
// last step (the same in every wave, which hold the same objects), and a lane whose span has
// ended, or that lies past the launch's objects, takes part in the evaluations on a state it
// does not store.
#include "gj8.hpp"
#include "kernels.hpp"

#include <math.h>

#include <algorithm>
#include <vector>

namespace ogp {

namespace {

// Equal bit patterns (unlike ==, tells +0 from -0).
__device__ __forceinline__ bool same_bits(double a, double b)
{
    return __double_as_longlong(a) == __double_as_longlong(b);
}

// pred holds in every lane of the wave (the same answer in every lane).
__device__ __forceinline__ bool wave_all(bool pred) { return __ballot(!pred) == 0; }

// One object's rows of a startup work area ([slot][component][object]). Stores happen only for
// lanes that run (`on`); the others read and discard.
struct Slots {
    double* base;
    int64_t N, i;
    bool on;
    __device__ double get(int slot, int c) const { return base[(slot * 3 + c) * N + i]; }
    __device__ void set(int slot, int c, double x) const
    {
        if (on)
            base[(slot * 3 + c) * N + i] = x;
    }
};

// Force evaluations of the startup: the lead form for gravity kinds with W > 1 (the helpers
// run gravity_helper for each), the one-wave form otherwise.
template <int KIND, int W>
__device__ __forceinline__ void eval_gen(const ForceCtx& f, const double* x, const double* v,
    int s, int k, double* a, Exchange& ex)
{
    if constexpr (W > 1 && (KIND == kGravity || KIND == kGravitySplit))
        force_gen_lead<KIND, W>(f, x, v, s, k, a, ex.red, ex.pos);
    else
        force_gen<KIND, W>(f, x, v, s, k, a, ex);
}
template <int KIND, int W>
__device__ __forceinline__ void eval_prop(const ForceCtx& f, const double* x, const double* v,
    int64_t node, double* a, Exchange& ex)
{
    if constexpr (W > 1 && (KIND == kGravity || KIND == kGravitySplit))
        force_prop_lead<KIND, W>(f, x, v, node, a, ex.red, ex.pos);
    else
        force_prop<KIND, W>(f, x, v, node, a, ex);
}

// RK8::step (gpu_startup.cu:108-144, reset_temp_vectors.cu, device_functions.cu,
// copy_prev_state.cu): stage states from the previous node, stage accelerations into Kv,
// stage velocities into Kx, then node(4 + step) = prev + h * sum b_k K_k.
template <int KIND, int W>
__device__ void rk8_step(const ForceCtx& f, const double (*a)[9], double h, const Slots& nr,
    const Slots& nv, const Slots& Kx, const Slots& Kv, int step, Exchange& ex)
{
    const double hs = step > 0 ? h : -h;
    const int prev = 4 + step + (step > 0 ? -1 : 1);
    double x0[3], v0[3], tX[3], tV[3], out[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        x0[c] = nr.get(prev, c);
        v0[c] = nv.get(prev, c);
    }
    eval_gen<KIND, W>(f, x0, v0, step, 0, out, ex);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Kx.set(0, c, v0[c]);
        Kv.set(0, c, out[c]);
    }
#pragma unroll 1
    for (int k = 1; k <= 9; ++k) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double sx = 0.0, sv = 0.0;
#pragma unroll 1
            for (int j = 0; j < k; ++j) {
                sx = fma(Kx.get(j, c), a[k - 1][j], sx);
                sv = fma(Kv.get(j, c), a[k - 1][j], sv);
            }
            tX[c] = fma(hs, sx, x0[c]);
            tV[c] = fma(hs, sv, v0[c]);
        }
        eval_gen<KIND, W>(f, tX, tV, step, k, out, ex);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Kv.set(k, c, out[c]);
            Kx.set(k, c, tV[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double sv = 0.0, sx = 0.0;
        for (int k = 0; k < 10; ++k) {
            sv = fma(Kv.get(k, c), c_rk8b[k], sv);
            sx = fma(Kx.get(k, c), c_rk8b[k], sx);
        }
        nv.set(4 + step, c, fma(hs, sv, v0[c]));
        nr.set(4 + step, c, fma(hs, sx, x0[c]));
    }
}

// StartupCorrector::CalculateAccelerationMatrix (gpu_startup.cu:196-224)
template <int KIND, int W>
__device__ void acc_matrix(const ForceCtx& f, const Slots& nr, const Slots& nv, const Slots& acc,
    bool epoch, Exchange& ex)
{
    ++ex.mat; // matrix ex.mat - 1 (recorded Jacobians' slots)
#pragma unroll 1
    for (int j = 0; j < 9; ++j) {
        if (j == 4 && !epoch)
            continue;
        double x[3], v[3], o[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x[c] = nr.get(j, c);
            v[c] = nv.get(j, c);
        }
        eval_prop<KIND, W>(f, x, v, j, o, ex);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            acc.set(j, c, o[c]);
    }
}

// StartupCorrector::Correct (gpu_startup.cu:247-294) with CalculateS0 (:170-195) and
// CalculateAllSKernel (calculate_all_s.cu:19-47). With fixed_point_exit, when the corrected
// nodes are bit for bit the nodes the acceleration matrix was evaluated at (the positions, and
// the velocities for forces that read them), for every lane that runs, the matrix is not
// evaluated again and true is returned: a correction depends only on the matrix and the fixed
// node 4, so every later correction would repeat this one exactly (the caller stops).
template <int KIND, int W>
__device__ bool startup_correct(const ForceCtx& f, double h, const Slots& nr, const Slots& nv,
    const Slots& acc, const Slots& sm, const Slots& Sm, Exchange& ex, bool fixed_point_exit)
{
    // the nodes the next acceleration matrix reads: positions, and velocities for forces that
    // depend on them (drag, the damped test force)
    constexpr bool kGrav = KIND == kGravity || KIND == kGravitySplit || KIND == kGravityRecord;
    const bool vel = !kGrav || (f.src.mask & kSourceDrag) != 0;
    bool same = true;
    const double invh = 1.0 / h, invh2 = 1.0 / (h * h), hh = h * h;
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        double ss = 0.0, sS = 0.0;
        for (int k = 0; k < 9; ++k) {
            ss = fma(acc.get(k, c), c_sa[4][k], ss);
            sS = fma(acc.get(k, c), c_gj[4][k], sS);
        }
        sm.set(4, c, fma(invh, nv.get(4, c), -ss));
        Sm.set(4, c, fma(invh2, nr.get(4, c), -sS));
        for (int i = 1; i <= 4; ++i) {
            sm.set(4 + i, c, sm.get(4 + i - 1, c) + (acc.get(4 + i - 1, c) + acc.get(4 + i, c)) / 2);
            sm.set(4 - i, c, sm.get(4 - i + 1, c) - (acc.get(4 - i + 1, c) + acc.get(4 - i, c)) / 2);
            Sm.set(4 + i, c, sm.get(4 + i - 1, c) + Sm.get(4 + i - 1, c) + acc.get(4 + i - 1, c) / 2);
            Sm.set(4 - i, c, -sm.get(4 - i + 1, c) + Sm.get(4 - i + 1, c) + acc.get(4 - i + 1, c) / 2);
        }
        for (int j = 0; j < 9; ++j) {
            if (j == 4)
                continue;
            double sv = 0.0, sp = 0.0;
            for (int k = 0; k < 9; ++k) {
                sv = fma(acc.get(k, c), c_sa[j][k], sv);
                sp = fma(acc.get(k, c), c_gj[j][k], sp);
            }
            const double vn = (sm.get(j, c) + sv) * h, rn = (Sm.get(j, c) + sp) * hh;
            same = same && (!vel || same_bits(vn, nv.get(j, c))) && same_bits(rn, nr.get(j, c));
            nv.set(j, c, vn);
            nr.set(j, c, rn);
        }
    }
    if (fixed_point_exit && wave_all(same || !nr.on))
        return true;
    acc_matrix<KIND, W>(f, nr, nv, acc, false, ex);
    return false;
}

template <int W> struct ExchangeArea {
    static constexpr int size = W > 1 ? kParts * 4 * 64 : 1;
};

// Lane l's object in object group `group` of the launch. Lanes past the launch's objects read
// object obj0 (a valid index) and count as having no steps.
struct Lane {
    bool valid;
    int64_t i;
};
__device__ __forceinline__ Lane lane_object(int64_t obj0, int64_t nobj, int64_t group)
{
    const int64_t t = group * 64 + threadIdx.x % 64;
    return Lane { t < nobj, obj0 + (t < nobj ? t : 0) };
}

} // namespace

// GpuStartup::Startup (gpu_startup.cu:55-68): 4 forward and 4 backward RK8 steps, the
// acceleration matrix and ncorr corrections; then the hand-over to the GJ8 steps
// (gpu_propagator.cu:31-53, copy_s.cu:19-30). A lane with no step (no GpuPropagator::Step call)
// leaves its state untouched.
// Register budget: 128 VGPRs (4 resident waves per SIMD) for gravity, whose lead wave then
// spills part of its RK8 work area (100,000 objects at degree 70: 18.7 against 25.0 ms with
// 256 VGPRs and 2 waves per SIMD; 20,000: 5.4 against 8.6 ms); 256 for the recording startup,
// whose gradient evaluation holds 13 sums per part, and for the analytic test forces.
template <int KIND, int W>
__global__ void __launch_bounds__(64 * W)
__attribute__((amdgpu_waves_per_eu(KIND == kGravity || KIND == kGravitySplit ? 4 : 2)))
startup_kernel(ForceCtx f0, int rk8_variant, PropState ps,
    StartupScratch sc, int64_t obj0, int64_t nobj, int ncorr, double* traj, int64_t traj_steps,
    int fixed_point_exit)
{
    constexpr bool kLead = W > 1 && (KIND == kGravity || KIND == kGravitySplit);
    constexpr bool kRecLead = W > 1 && KIND == kGravityRecord; // gradient parts, 53 KB exchange
    __shared__ double red[kRecLead ? kJacExchange : ExchangeArea<W>::size];
    __shared__ double pos[W > 1 ? kPointLen * 64 : 1]; // the lead's geometry (gravity_gcrs_lead)
    __shared__ int more; // the lead's control word for its helpers (gravity_helper_loop)
    Exchange ex { red, pos };
    const Lane ln = lane_object(obj0, nobj, blockIdx.x);
    const int64_t i = ln.i;
    const ForceCtx f = object_ctx(f0, i);
    const int64_t N = ps.N;
    const int64_t ns = ln.valid ? ps.nsteps[i] : 0;
    if (__builtin_amdgcn_readfirstlane((int)std::min<int64_t>(wave_max(ns), 1)) <= 0)
        return; // no lane of the group runs (the same in every wave of the workgroup)
    const bool act = ns > 0;
    ex.on = act;
    if constexpr (kLead || kRecLead) {
        const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
        if (w > 0) {
            // one part sum per force evaluation of the lead: 8 RK8 steps of 10 stages, the
            // acceleration matrix (9 nodes) and up to ncorr corrections of 8 nodes each
            if constexpr (kRecLead)
                gravity_record_helper_loop<W>(f.grav, red, pos, &more, w);
            else
                gravity_helper_loop<KIND == kGravitySplit, W>(f.grav, red, pos, &more, w);
            return;
        }
        if (threadIdx.x == 0)
            more = kRecLead ? 2 : 1; // every recording evaluation evaluates the gradient
    }
    const Slots nr { sc.node_r, N, i, act }, nv { sc.node_v, N, i, act }, acc { sc.acc, N, i, act };
    const Slots sm { sc.sm, N, i, act }, Sm { sc.Sm, N, i, act }, Kx { sc.Kx, N, i, act },
        Kv { sc.Kv, N, i, act };
    const double h = f.h;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        nr.set(4, c, ps.r[c * N + i]);
        nv.set(4, c, ps.v[c * N + i]);
    }
    const double(*a)[9] = c_rk8a[rk8_variant];
#pragma unroll 1
    for (int s = 1; s <= 4; ++s) {
        rk8_step<KIND, W>(f, a, h, nr, nv, Kx, Kv, s, ex);
        rk8_step<KIND, W>(f, a, h, nr, nv, Kx, Kv, -s, ex);
    }
    acc_matrix<KIND, W>(f, nr, nv, acc, true, ex);
#pragma unroll 1
    for (int it = 0; it < ncorr; ++it)
        if (startup_correct<KIND, W>(f, h, nr, nv, acc, sm, Sm, ex, fixed_point_exit != 0))
            break;
    if constexpr (kLead || kRecLead) {
        // release the helpers (gravity_helper_loop), in wave-uniform control flow
        if (threadIdx.x == 0)
            more = 0;
        __syncthreads();
    }
    if constexpr (KIND == kGravityRecord) {
        // matrices skipped at the corrections' fixed point are at the last matrix's nodes: the
        // same Jacobians (the tangent startup runs every correction)
        if (act)
            for (int m = ex.mat; m <= ncorr; ++m)
                for (int j = 0; j < 9; ++j)
                    if (j != 4)
#pragma unroll
                        for (int e = 0; e < 6; ++e)
                            grec_at(f, grec_slot_prop(m, j), e) =
                                grec_at(f, grec_slot_prop(ex.mat - 1, j), e);
    }
    if (!act)
        return;
    const int last = (int)(ns < 4 ? ns : 4);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ps.s[c * N + i] = sm.get(8, c);
        ps.S[c * N + i] = Sm.get(8, c);
        ps.r[c * N + i] = nr.get(4 + last, c);
        ps.v[c * N + i] = nv.get(4 + last, c);
        for (int j = -4; j <= 4; ++j)
            ps.hist[(ring_slot(j) * 3 + c) * N + i] = acc.get(j + 4, c);
    }
    if (traj) {
        // steps 1-4 output the startup nodes; the acceleration handed to the conjunction
        // screen is the acceleration-matrix column 8 (gpu_propagator.cu:68-72)
        for (int k = 1; k <= last && k <= traj_steps; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                traj[((int64_t)(k - 1) * 9 + c) * N + i] = nr.get(4 + k, c);
                traj[((int64_t)(k - 1) * 9 + 3 + c) * N + i] = nv.get(4 + k, c);
                traj[((int64_t)(k - 1) * 9 + 6 + c) * N + i] = acc.get(8, c);
            }
    }
}

// GJ8 steps k0 .. k1-1 (each k >= 5), persistent over the chunk: Predictor::Predict
// (prediction.cu:49-74, dev_functions.cu:20-50) and the corrector (gpu_propagator.cu:86-119,
// dev_functions.cu:36-69). The acceleration history is a ring of 9 slots in HBM.
// G object groups share a workgroup in lockstep (G x W waves, one exchange area per group): the
// barriers are the workgroup's, so the waves summing the same harmonic part of different groups
// read the same table lines at about the same time. The results do not depend on G.
// SRC: gravity with acceleration sources (f.src.mask != 0), an instance of its own so that the
// gravity-only kernel's registers are allocated for the gravity-only path.
// WPE: resident waves per SIMD the register allocation aims for. 4 (128 VGPRs) by default; the
// 8-wave gravity shapes also come with 6 (80 VGPRs: three workgroups per CU instead of two, a
// few spills), which the launcher takes when more than two object groups per CU run at once,
// counted over the slices of a step-size class (100,000 objects at degree 70: 12.4 against 13.7
// ms per launch; 20,000 at degree 20, about one group per CU: 3% slower, so it keeps 4).
template <int KIND, int W, int G, bool SRC = false, int WPE = 4>
__global__ void __launch_bounds__(64 * W * G)
__attribute__((amdgpu_waves_per_eu(WPE)))
gj8_chunk_kernel(ForceCtx f0, PropState ps, int64_t obj0,
    int64_t nobj, int64_t k0, int64_t k1, int ncorr, double* traj, int64_t traj_first,
    int64_t traj_steps, int fixed_point_exit, unsigned long long* evals)
{
    static_assert(G == 1 || W > 1, "lockstep groups need the lead/helper form");
    __shared__ double red[G * ExchangeArea<W>::size];
    __shared__ double pos[W > 1 ? G * kPointLen * 64 : 1]; // the leads' geometry
    __shared__ int64_t group_end[G];
    __shared__ int more; // one group: the lead's control word for its helpers
    // gravity without sources, 4 or 8 waves: what the one correction after a step's evaluations
    // needs (the corrector rows T1, T2, a_{k-1}, s_{k-1} and S), parked per lane while the lead
    // evaluates, so that the evaluations have the registers instead of spilling them (lane-
    // private: written and read by one wave). 7.5 KB per group: with 1 or 2 waves it would cost
    // resident workgroups, so those shapes keep the values in registers.
    constexpr bool kPark = (KIND == kGravity || KIND == kGravitySplit) && W >= 4;
    __shared__ double park[kPark ? G * 15 * 64 : 1];
    const int wg = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
    const int g = wg / W, w = wg % W; // object group of the workgroup, wave within the group
    double* gred = red + (W > 1 ? g * ExchangeArea<W>::size : 0);
    double* gpos = pos + (W > 1 ? g * kPointLen * 64 : 0);
    Exchange ex { gred, gpos };
    const Lane ln = lane_object(obj0, nobj, (int64_t)blockIdx.x * G + g);
    const int64_t i = ln.i;
    const ForceCtx f = object_ctx(f0, i);
    const int64_t N = ps.N;
    const int64_t ns = ln.valid ? ps.nsteps[i] : -1;
    // the last step of the chunk over the workgroup's groups, the same in every wave
    int64_t kend = std::min<int64_t>(k1, wave_max(ns) + 1);
    if constexpr (G > 1) {
        if (w == 0 && threadIdx.x % 64 == 0)
            group_end[g] = kend;
        __syncthreads();
        for (int q = 0; q < G; ++q)
            kend = std::max<int64_t>(kend, group_end[q]);
    }
    kend = __builtin_amdgcn_readfirstlane((int)kend);
    if (kend <= k0)
        return;
    // One group per workgroup: a step may end its corrector evaluations at a fixed point (below),
    // so the helpers follow the lead's control word. Lockstep groups pass common barriers and
    // run every step's ncorr evaluations.
    const bool fp_exit = G == 1 && fixed_point_exit != 0;
    if constexpr (W > 1) {
        // Waves 1 .. W-1 of a group only sum their harmonic parts; its wave 0 integrates. Two
        // barriers per evaluation in every wave, and with one group a last barrier after which
        // the helpers read more == 0.
        if (w > 0) {
            if constexpr (G == 1) {
                gravity_helper_loop<KIND == kGravitySplit, W>(f.grav, gred, gpos, &more, w);
            } else {
                for (int64_t k = k0; k < kend; ++k)
                    for (int it = 0; it < ncorr; ++it)
                        gravity_helper<KIND == kGravitySplit, W>(f.grav, gred, gpos, w);
            }
            return;
        }
        if constexpr (G == 1) {
            if (threadIdx.x == 0)
                more = 1;
        }
    }
    // Evaluations of the live objects of the group (lanes that store step k), for the FLOP count.
    unsigned long long done = 0;
    const double h = f.h, hh = h * h;
    double r[3], v[3], s[3], S[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r[c] = ps.r[c * N + i];
        v[c] = ps.v[c * N + i];
        s[c] = ps.s[c * N + i];
        S[c] = ps.S[c * N + i];
    }
    auto store_state = [&]() {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ps.r[c * N + i] = r[c];
            ps.v[c * N + i] = v[c];
            ps.s[c * N + i] = s[c];
            ps.S[c * N + i] = S[c];
        }
    };
    for (int64_t k = k0; k < kend; ++k) {
        // Every lane runs every step up to kend; a lane past its span (k > ns) integrates on
        // a history it no longer updates and stores nothing: its final state was stored at its
        // last step.
        const bool live = k <= ns;
        int nev = 0;
        GJ8Sums q;
        gj8_sums(ps.hist, N, i, k, q);
        double sprev[3], anew[3];
        gj8_predict(q, h, s, S, r, v, sprev);
        // P E (E C)^ncorr: the predictor's evaluation is the first corrector evaluation.
        if constexpr (KIND == kGravity || KIND == kGravitySplit) {
            // Every evaluation of the step is at node k + 4, one matrix M. Without sources the
            // corrected position rotated into the ITRS is M^T r = hh (M^T (T2 + S) + g88 g_ITRS),
            // so each evaluation point follows from the previous evaluation's ITRS gravity in two
            // operations; and since a correction depends only on the latest acceleration (s, v
            // and r from sprev, S and the sums), the step corrects once, after its last
            // evaluation. The corrector and both rotations leave the serial path of the
            // evaluations (gj8_record_kernel runs the same arithmetic, bit for bit). With
            // sources, which need each corrected GCRS position and velocity, every evaluation
            // is corrected and rotated.
            const uniform_dptr M = (uniform_dptr)(f.prop_M + 9 * (k + 4));
            auto gravity = [&](const double* ri, double* gi) {
                if constexpr (W > 1)
                    gravity_itrs_lead<KIND == kGravitySplit, W>(ri, f.grav, ex.red, ex.pos, gi);
                else
                    gravity_itrs<KIND == kGravitySplit>(ri[0], ri[1], ri[2], f.grav, gi[0], gi[1],
                        gi[2]);
            };
            double ri[3], gi[3];
            to_itrs(M, r, ri);
            if constexpr (!SRC) {
                double base[3], t[3];
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    t[c] = q.T2[c] + S[c];
                to_itrs(M, t, base);
                double* pk = park + (kPark ? g * 15 * 64 + threadIdx.x % 64 : 0);
                if constexpr (kPark) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        pk[c * 64] = q.T1[c];
                        pk[(3 + c) * 64] = q.T2[c];
                        pk[(6 + c) * 64] = q.alast[c];
                        pk[(9 + c) * 64] = sprev[c];
                        pk[(12 + c) * 64] = S[c];
                    }
                }
                for (int it = 0; it < ncorr; ++it) {
                    gravity(ri, gi);
                    ++nev;
                    bool same = true;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double rn = hh * __builtin_fma(c_gj[8][8], gi[c], base[c]);
                        same = same && same_bits(rn, ri[c]);
                        ri[c] = rn;
                    }
                    // The next point is the one just evaluated: every further evaluation would
                    // return gi again, bit for bit, so the step's result is already final.
                    if (fp_exit && wave_all(same || !live))
                        break;
                }
                to_gcrs(M, gi, anew);
                if constexpr (kPark) {
                    GJ8Sums qc;
                    double sp[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        qc.T1[c] = pk[c * 64];
                        qc.T2[c] = pk[(3 + c) * 64];
                        qc.alast[c] = pk[(6 + c) * 64];
                        sp[c] = pk[(9 + c) * 64];
                        S[c] = pk[(12 + c) * 64];
                    }
                    gj8_correct(qc, anew, sp, S, h, hh, s, r, v);
                } else {
                    gj8_correct(q, anew, sprev, S, h, hh, s, r, v);
                }
            } else {
                // With drag the velocity settles an iteration after the position: with the
                // fixed-point exit, an evaluation whose ITRS point repeats the last one (in every
                // lane that stores the step) keeps that gravity, bit for bit the same, and
                // recomputes only the sources.
                double rg[3] = { 0.0, 0.0, 0.0 };
                for (int it = 0; it < ncorr; ++it) {
                    bool same_point = fp_exit && it > 0;
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        same_point = same_point && same_bits(ri[c], rg[c]);
                    if (!wave_all(same_point || !live)) {
                        gravity(ri, gi);
                        ++nev;
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            rg[c] = ri[c];
                    }
                    to_gcrs(M, gi, anew);
                    add_sources(f.src, f.prop_tb + 6 * (k + 4), r, v, anew);
                    const double r0[3] = { r[0], r[1], r[2] }, v0[3] = { v[0], v[1], v[2] };
                    gj8_correct(q, anew, sprev, S, h, hh, s, r, v);
                    to_itrs(M, r, ri);
                    // The corrected state is the one just evaluated: the next acceleration, and
                    // every correction after it, would repeat this one bit for bit. A correction
                    // depends on the step's sums and the latest acceleration only, so with forces
                    // of the position alone (gravity, Sun, Moon) the position decides; drag reads
                    // the velocity as well.
                    const bool vel = (f.src.mask & kSourceDrag) != 0;
                    bool same = true;
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        same = same && same_bits(r[c], r0[c]) && (!vel || same_bits(v[c], v0[c]));
                    if (fp_exit && wave_all(same || !live))
                        break;
                }
            }
        } else {
            for (int it = 0; it < ncorr; ++it) {
                eval_prop<KIND, W>(f, r, v, k + 4, anew, ex);
                gj8_correct(q, anew, sprev, S, h, hh, s, r, v);
            }
            nev = ncorr;
        }
        done += (unsigned long long)nev * (unsigned long long)__popcll(__ballot(live));
        if (k > ns)
            continue;
        const int wslot = ring_slot(k);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            ps.hist[(wslot * 3 + c) * N + i] = anew[c];
        if (traj && k >= traj_first && k < traj_first + traj_steps) {
            const int64_t row = k - traj_first;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                traj[(row * 9 + c) * N + i] = r[c];
                traj[(row * 9 + 3 + c) * N + i] = v[c];
                traj[(row * 9 + 6 + c) * N + i] = anew[c];
            }
        }
        if (k == ns)
            store_state();
    }
    if constexpr (W > 1 && G == 1) {
        // release the helpers (gravity_helper_loop)
        if (threadIdx.x == 0)
            more = 0;
        __syncthreads();
    }
    if (ns >= kend)
        store_state();
    if (evals && threadIdx.x % 64 == 0)
        atomicAdd(evals, done);
}

// Instantiation over force kinds (analytic test forces use W = 1 only).
#define OGP_DISPATCH(kind, waves, LAUNCH)                                                          \
    do {                                                                                           \
        if (kind == kGravity || kind == kGravitySplit) {                                           \
            const bool split_ = kind == kGravitySplit;                                             \
            switch (waves) {                                                                       \
            case 2: if (split_) LAUNCH(kGravitySplit, 2); else LAUNCH(kGravity, 2); break;         \
            case 4: if (split_) LAUNCH(kGravitySplit, 4); else LAUNCH(kGravity, 4); break;         \
            case 8: if (split_) LAUNCH(kGravitySplit, 8); else LAUNCH(kGravity, 8); break;         \
            default: if (split_) LAUNCH(kGravitySplit, 1); else LAUNCH(kGravity, 1); break;        \
            }                                                                                      \
        } else {                                                                                   \
            switch (kind) {                                                                        \
            case kNone: LAUNCH(kNone, 1); break;                                                   \
            case kConst: LAUNCH(kConst, 1); break;                                                 \
            case kPoly7: LAUNCH(kPoly7, 1); break;                                                 \
            case kPoly8: LAUNCH(kPoly8, 1); break;                                                 \
            case kPoly9: LAUNCH(kPoly9, 1); break;                                                 \
            case kExp: LAUNCH(kExp, 1); break;                                                     \
            case kDamped: LAUNCH(kDamped, 1); break;                                               \
            default: break;                                                                        \
            }                                                                                      \
        }                                                                                          \
    } while (0)

int effective_waves(int kind, int waves)
{
    if (kind != kGravity && kind != kGravitySplit)
        return 1;
    return waves == 2 || waves == 4 || waves == 8 ? waves : 1;
}

void launch_startup(hipStream_t st, int kind, const ForceCtx& f, int rk8_variant,
    const PropState& ps, const StartupScratch& sc, int64_t obj0, int64_t nobj, int ncorr,
    double* traj, int64_t traj_steps, int waves, bool fixed_point_exit)
{
    if (nobj <= 0)
        return;
    const dim3 grid((unsigned)((nobj + 63) / 64));
    const int W_ = effective_waves(kind, waves);
#define L_(K, Wv)                                                                                  \
    hipLaunchKernelGGL((startup_kernel<K, Wv>), grid, dim3(64 * (Wv)), 0, st, f, rk8_variant, ps,   \
        sc, obj0, nobj, ncorr, traj, traj_steps, fixed_point_exit ? 1 : 0)
    if (kind == kGravityRecord) {
        // 4 or 8 waves: the 53 KB gradient exchange leaves a 2-wave shape one wave per SIMD
        if (waves >= 8)
            L_(kGravityRecord, 8);
        else
            L_(kGravityRecord, 4);
    } else if (kind == kRecordedTangent) {
        L_(kRecordedTangent, 1);
    } else {
        OGP_DISPATCH(kind, W_, L_);
    }
#undef L_
}

// Compute units of the current device (queried once; one device per process).
int device_cus()
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess
            && hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            cus = n > 0 ? n : 1;
        else
            cus = 1;
    }
    return cus;
}

template <int K, int Wv, int G, bool S>
void launch_chunk_instance(bool occ6, dim3 grid, hipStream_t st, const ForceCtx& f,
    const PropState& ps, int64_t obj0, int64_t nobj, int64_t k0, int64_t k1, int ncorr,
    double* traj, int64_t traj_first, int64_t traj_steps, int fp, unsigned long long* evals)
{
    if constexpr ((K == kGravity || K == kGravitySplit) && Wv == 8 && G == 1) {
        if (occ6) {
            hipLaunchKernelGGL((gj8_chunk_kernel<K, Wv, G, S, 6>), grid, dim3(64 * Wv * G), 0, st,
                f, ps, obj0, nobj, k0, k1, ncorr, traj, traj_first, traj_steps, fp, evals);
            return;
        }
    }
    hipLaunchKernelGGL((gj8_chunk_kernel<K, Wv, G, S, 4>), grid, dim3(64 * Wv * G), 0, st, f, ps,
        obj0, nobj, k0, k1, ncorr, traj, traj_first, traj_steps, fp, evals);
}

template <int K, int Wv, int G>
void launch_chunk_kernel(bool src, bool occ6, dim3 grid, hipStream_t st, const ForceCtx& f,
    const PropState& ps, int64_t obj0, int64_t nobj, int64_t k0, int64_t k1, int ncorr,
    double* traj, int64_t traj_first, int64_t traj_steps, int fp, unsigned long long* evals)
{
    if constexpr (K == kGravity || K == kGravitySplit) {
        if (src) {
            launch_chunk_instance<K, Wv, G, true>(occ6, grid, st, f, ps, obj0, nobj, k0, k1,
                ncorr, traj, traj_first, traj_steps, fp, evals);
            return;
        }
    }
    launch_chunk_instance<K, Wv, G, false>(occ6, grid, st, f, ps, obj0, nobj, k0, k1, ncorr,
        traj, traj_first, traj_steps, fp, evals);
}

void launch_chunk(hipStream_t st, int kind, const ForceCtx& f, const PropState& ps,
    int64_t obj0, int64_t nobj, int64_t k0, int64_t k1, int ncorr, double* traj,
    int64_t traj_first, int64_t traj_steps, int waves, int lockstep, bool fixed_point_exit,
    unsigned long long* evals, int64_t concurrent_objects)
{
    if (nobj <= 0 || k1 <= k0)
        return;
    const int W_ = effective_waves(kind, waves);
    const int G_ = W_ >= 4 && lockstep == 2 ? 2 : 1;
    const int64_t ngroups = (nobj + 63) / 64;
    const dim3 grid((unsigned)((ngroups + G_ - 1) / G_));
    const int fp_ = fixed_point_exit ? 1 : 0;
    const bool src_ = f.src.mask != 0; // gravity with sources: the SRC instances
    // more than two 8-wave groups per CU at once (this launch's and those running beside it):
    // the 6-waves-per-SIMD register allocation, which lets a CU hold three of them
    const int64_t resident = (std::max(nobj, concurrent_objects) + 63) / 64;
    const bool occ6 = resident > 2 * (int64_t)device_cus();
#define L_(K, Wv)                                                                                  \
    launch_chunk_kernel<K, Wv, 1>(src_, occ6, grid, st, f, ps, obj0, nobj, k0, k1, ncorr, traj,   \
        traj_first, traj_steps, fp_, evals)
#define L2_(K, Wv)                                                                                 \
    launch_chunk_kernel<K, Wv, 2>(src_, occ6, grid, st, f, ps, obj0, nobj, k0, k1, ncorr, traj,   \
        traj_first, traj_steps, fp_, evals)
    if (G_ == 2) {
        const bool split_ = kind == kGravitySplit;
        if (W_ == 4) {
            if (split_) L2_(kGravitySplit, 4); else L2_(kGravity, 4);
        } else {
            if (split_) L2_(kGravitySplit, 8); else L2_(kGravity, 8);
        }
    } else {
        OGP_DISPATCH(kind, W_, L_);
    }
#undef L2_
#undef L_
}

void column_parts(int L, int* parts)
{
    static_assert(kParts == 8, "the pairs below assume 8 parts");
    // Cost model, in tenths of a full record (nine FP64 instructions), fitted to the time of
    // each part measured on gfx950, alone on a CU (degrees 10 to 30) and in 4- and 8-wave groups
    // of the 20,000-object propagations at degrees 20 and 70: a column's set-up and epilogue 68,
    // a part's closed-form start at column m > 0 38 plus 19 per halving of m, and a record of
    // the zonal column 10 at degree 20 rising to 14 at degree 70 (it issues fewer instructions
    // than a full record but waits as long for each batch of records). One wave issues at most
    // one FP64 instruction per ~7.5 cycles and a SIMD one per ~4.5, in favour of the older
    // wave: while two waves share a SIMD the older one runs rho times as fast as the younger,
    // rho = R / 10 measured about 1.3 at degree 20 and 1.15 at degree 70.
    //
    // Wave s of a 4-wave group sums parts 2s and 2s + 1, alone on its SIMD; in an 8-wave group
    // the same two parts go to the older and the younger wave of SIMD s (device.hpp,
    // wave_part), which finish together when the older one holds rho / (1 + rho) of the pair's
    // work. So the columns are cut into 4 contiguous chunks (the pairs; every cut is tried) and
    // each chunk into its two parts at the split that minimises the 8-wave time of the pair, and
    // the chunk cut minimises a mix of the 8-, 4- and 2-wave times relative to their ideals. All
    // arithmetic is on integers, so the parts of a degree are the same on every host.
    const int64_t R = L <= 20 ? 13 : L >= 70 ? 12 : 13 - (L - 20 + 25) / 50;
    const int64_t Z = L <= 20 ? 10 : L >= 70 ? 14 : 10 + (4 * (L - 20) + 25) / 50;
    const int N = L + 1;
    std::vector<int64_t> pre(N + 1, 0);
    for (int m = 0; m < N; ++m)
        pre[m + 1] = pre[m] + (m == 0 ? Z : 10) * (L - m + 1) + 68;
    auto part_cost = [&](int a, int b) -> int64_t {
        if (b <= a)
            return 0;
        int halvings = 0;
        for (int e = a - 1; e > 0; e >>= 1)
            ++halvings;
        return pre[b] - pre[a] + (a > 0 ? 38 + 19 * halvings : 0);
    };
    // best split of chunk [x, y): (R x 8-wave time, chunk cost, split), lexicographic minimum
    struct Inner {
        int64_t t8, cost;
        int split;
    };
    std::vector<Inner> inner((size_t)(N + 1) * (N + 1));
    for (int x = 0; x <= N; ++x)
        for (int y = x; y <= N; ++y) {
            Inner best { 0, 0, -1 };
            for (int sp = x; sp <= y; ++sp) {
                const int64_t a = part_cost(x, sp), b = part_cost(sp, y);
                const Inner t { R * a + std::max<int64_t>(0, R * b - 10 * a), a + b, sp };
                if (best.split < 0 || t.t8 < best.t8 || (t.t8 == best.t8 && t.cost < best.cost))
                    best = t;
            }
            inner[(size_t)x * (N + 1) + y] = best;
        }
    // objective f / S (S the total cost): f = 10 (a8 + a4) + 3 a2 + 20 (max(a8, a4, a2) - S R^2)
    // with a8, a4, a2 the slowest SIMD (8 waves), wave (4 waves) and wave (2 waves) scaled so
    // that S R^2 is a perfect balance
    int64_t best_f = -1, best_S = 1;
    int bounds[3] = { 0, 0, 0 };
    for (int b2 = 0; b2 <= N; ++b2)
        for (int b4 = b2; b4 <= N; ++b4)
            for (int b6 = b4; b6 <= N; ++b6) {
                const int cut[5] = { 0, b2, b4, b6, N };
                int64_t sums[4], S = 0, M8 = 0, M4 = 0;
                for (int c = 0; c < 4; ++c) {
                    const Inner& t = inner[(size_t)cut[c] * (N + 1) + cut[c + 1]];
                    sums[c] = t.cost;
                    S += t.cost;
                    M8 = std::max(M8, t.t8);
                    M4 = std::max(M4, t.cost);
                }
                const int64_t M2 = std::max(sums[0] + sums[1], sums[2] + sums[3]);
                const int64_t a8 = 4 * (R + 10) * M8, a4 = 4 * M4 * R * R, a2 = 2 * M2 * R * R;
                const int64_t f = 10 * (a8 + a4) + 3 * a2
                    + 20 * (std::max(a8, std::max(a4, a2)) - S * R * R);
                if (S > 0 && (best_f < 0 || f * best_S < best_f * S)) {
                    best_f = f;
                    best_S = S;
                    bounds[0] = b2;
                    bounds[1] = b4;
                    bounds[2] = b6;
                }
            }
    const int cut[5] = { 0, bounds[0], bounds[1], bounds[2], N };
    for (int c = 0; c < 4; ++c) {
        parts[2 * c] = cut[c];
        parts[2 * c + 1] = inner[(size_t)cut[c] * (N + 1) + cut[c + 1]].split;
    }
    parts[kParts] = N;
}

} // namespace ogp
