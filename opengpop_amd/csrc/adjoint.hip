// Derivatives of the propagation: gravity Jacobian, variational-equation (tangent) layout, and
// the discrete adjoint of the GJ8 steps. No counterpart in the reference.
//
// d(r, v)_final / d(r, v)_initial is assembled from two pieces:
//  * the startup (RK8 + startup corrector, gpu_startup.cu) is differentiated forward: the startup
//    of the objects runs with the gravity gradient evaluation and records the Jacobian of every
//    force evaluation (kGravityRecord, device.hpp); the startup kernel then runs again on six
//    tangent lanes per object with those Jacobians (kRecordedTangent), giving the derivatives of
//    the post-startup state Z = (s, S, a_{-4..4}) w.r.t. the six initial components;
//  * the GJ8 steps are differentiated backward. Each step's gravity Jacobian is taken at the
//    point of its last corrector evaluation (the corrector's fixed point; the earlier iterates
//    lie within the predictor's error of it, where the gradient differs by ~1e-11 of itself) and
//    used for every iterate of the step. When the Jacobians of the span fit in HBM, the forward
//    pass itself records them (gj8_record_kernel, bit-identical trajectory); otherwise the
//    forward pass stores the state (s, S and the
//    acceleration ring, 33 doubles per object) every C steps and, per segment, last to first, a
//    recording pass re-runs the steps from the checkpoint. The reverse pass carries the
//    cotangents of (s, S, ring) back through each segment. Cotangents of Z times dZ/dX give the
//    Jacobian rows.
//
// Reverse step (transpose of gj8_sums / gj8_predict / gj8_correct, gj8.hpp): with e_it the
// acceleration at iterate r_it (it = 0 .. K-1), r_{it+1} = hh((T2 + gj88 e_it) + S'),
// v_{it+1} = h((T1 + sa88 e_it) + s'_it), s'_it = s + (e_it + a_{k-1})/2, S' = S + s + a_{k-1}/2,
// r_0 = hh(Pr + S'); the predicted velocity does not reach the gravity, so Pv carries no
// cotangent. The cotangent of e_it is mapped to r_it by the step's G, symmetric (a Hessian).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "gj8.hpp"
#include "kernels.hpp"

namespace ogp {

namespace {

constexpr int kMaxCorr = 16;

} // namespace

// One object per lane, all parts (gravity_itrs_jac); 256 VGPRs, where 128 spill the fold.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
gravity_jac_kernel(GravityDev g, const double* __restrict__ r, int64_t n,
    double* __restrict__ a, double* __restrict__ G)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double ai[3], Gi[3][3];
    gravity_itrs_jac(r[3 * i], r[3 * i + 1], r[3 * i + 2], g, ai, Gi);
    for (int c = 0; c < 3; ++c) {
        a[3 * i + c] = ai[c];
        for (int j = 0; j < 3; ++j)
            G[9 * i + 3 * c + j] = Gi[c][j];
    }
}

// The recorded steps' gravity Jacobians (after gj8_record_kernel): slot (k - k0, o) of Gbuf
// holds the ITRS point of object o's last evaluation of step k in its first three entries; it is
// replaced by the GCRS Jacobian G = M Gi M^T (6 unique entries), M the node's ITRS->GCRS matrix.
// One lane per (object, step): block (x, y) = (object group x, step k0 + y), so M is
// wave-uniform. Slots of steps past an object's span hold nothing and are skipped (the sweeps do
// not read them).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
gravity_gradient_kernel(GravityDev g, const double* __restrict__ prop_M,
    const int64_t* __restrict__ nsteps, int64_t n, int64_t k0, double* __restrict__ Gbuf)
{
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    const int64_t k = k0 + blockIdx.y;
    if (i >= n || k > nsteps[i])
        return;
    double* slot = Gbuf + (int64_t)blockIdx.y * 6 * n + i;
    const double ri[3] = { slot[0], slot[n], slot[2 * n] };
    const uniform_dptr M = (uniform_dptr)(prop_M + 9 * (k + 4));
    double gi[3], Gi[3][3], a[3], G[3][3];
    gravity_itrs_jac(ri[0], ri[1], ri[2], g, gi, Gi);
    jac_to_gcrs(M, gi, Gi, a, G);
    const double g6[6] = { G[0][0], G[0][1], G[0][2], G[1][1], G[1][2], G[2][2] };
    for (int e = 0; e < 6; ++e)
        slot[e * n] = g6[e];
}

// Tangent lanes of the startup's derivative: lane col * n + o holds the derivative of object o's
// state with respect to its initial component col (x, y, z, vx, vy, vz): it starts at that unit
// vector and runs the startup with the recorded Jacobians of object o (kRecordedTangent).
__global__ void tangent_init_kernel(const int64_t* __restrict__ steps, int64_t n, PropState pt,
    int64_t* __restrict__ ns_out)
{
    const int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= 6 * n)
        return;
    const int col = (int)(L / n);
    const int64_t o = L % n;
    for (int c = 0; c < 3; ++c) {
        pt.r[c * pt.N + L] = col == c ? 1.0 : 0.0;
        pt.v[c * pt.N + L] = col == 3 + c ? 1.0 : 0.0;
    }
    ns_out[L] = steps[o];
}

// Forward tangent pass over GJ8 steps k0 .. k1-1: the variational equations of the discrete
// scheme, driven by the Jacobians the record pass stored for the same steps. Lane (o, col) carries
// the derivative with respect to initial component col of the object's (s, S, ring), in place in
// Zd[(col * 33 + e) * n + o], and of its (r, v) in RVd[(col * 6 + e) * n + o]. The integrator is
// linear in these, so the step arithmetic is gj8.hpp's, applied to the derivatives; the
// acceleration of a derivative is G dr (G symmetric, six stored entries). Block b holds objects
// 64 (b / 6) .. +63 for column b % 6: the six blocks of an object group are adjacent in the grid
// and read the same Jacobians close together in time.
__global__ void __launch_bounds__(64) gj8_tangent_kernel(const int64_t* __restrict__ nsteps,
    int64_t n, double h, int64_t k0, int64_t k1, int ncorr, const double* __restrict__ Gbuf,
    double* __restrict__ Zd, double* __restrict__ RVd)
{
    const int col = (int)(blockIdx.x % 6);
    const int64_t o = (int64_t)(blockIdx.x / 6) * 64 + threadIdx.x;
    if (o >= n)
        return;
    const int64_t ns = nsteps[o];
    if (ns < k0)
        return;
    const int64_t klast = std::min<int64_t>(k1 - 1, ns);
    double* Z = Zd + (int64_t)col * 33 * n;
    double* RV = RVd + (int64_t)col * 6 * n;
    double* ring = Z + 6 * n;
    const double hh = h * h;
    double s[3], S[3], r[3], v[3];
    for (int c = 0; c < 3; ++c) {
        s[c] = Z[c * n + o];
        S[c] = Z[(3 + c) * n + o];
        r[c] = RV[c * n + o];
        v[c] = RV[(3 + c) * n + o];
    }
    // the ring in registers for the whole segment: A[j] = a_{k-9+j} when step k begins
    double A[9][3];
    for (int j = 0; j < 9; ++j)
        for (int c = 0; c < 3; ++c)
            A[j][c] = ring[(ring_slot(k0 - 9 + j) * 3 + c) * n + o];
    // the step's Jacobian, loaded one step ahead
    auto load_g = [&](int64_t k, double* g) {
        const double* gp = Gbuf + (((k - k0) * 6) * n + o);
        for (int e = 0; e < 6; ++e)
            g[e] = gp[e * n];
    };
    double g[6], gn[6];
    load_g(k0, g);
    for (int64_t k = k0; k <= klast; ++k) {
        if (k < klast)
            load_g(k + 1, gn);
        GJ8Sums q;
        gj8_sums_of([&](int j, int c) { return A[j][c]; }, q);
        double sprev[3], anew[3];
        gj8_predict(q, h, s, S, r, v, sprev);
        const double gxx = g[0], gxy = g[1], gxz = g[2], gyy = g[3], gyz = g[4], gzz = g[5];
        for (int it = 0; it < ncorr; ++it) {
            anew[0] = gxx * r[0] + gxy * r[1] + gxz * r[2];
            anew[1] = gxy * r[0] + gyy * r[1] + gyz * r[2];
            anew[2] = gxz * r[0] + gyz * r[1] + gzz * r[2];
            gj8_correct(q, anew, sprev, S, h, hh, s, r, v);
        }
        for (int j = 0; j < 8; ++j)
            for (int c = 0; c < 3; ++c)
                A[j][c] = A[j + 1][c];
        for (int c = 0; c < 3; ++c)
            A[8][c] = anew[c];
        for (int e = 0; e < 6; ++e)
            g[e] = gn[e];
    }
    // A[j] = a_{klast-8+j}, at its ring slot
    for (int j = 0; j < 9; ++j)
        for (int c = 0; c < 3; ++c)
            ring[(ring_slot(klast - 8 + j) * 3 + c) * n + o] = A[j][c];
    for (int c = 0; c < 3; ++c) {
        Z[c * n + o] = s[c];
        Z[(3 + c) * n + o] = S[c];
        RV[c * n + o] = r[c];
        RV[(3 + c) * n + o] = v[c];
    }
}

// The 6x6 matrix stm[o][row][col] from the derivatives of the final (r, v) (zero steps: the
// identity).
__global__ void stm_final_kernel(const int64_t* __restrict__ nsteps, int64_t n,
    const double* __restrict__ RVd, double* __restrict__ stm)
{
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n)
        return;
    const int64_t ns = nsteps[o];
    for (int col = 0; col < 6; ++col)
        for (int e = 0; e < 6; ++e)
            stm[36 * o + 6 * e + col] = ns == 0 ? (e == col ? 1.0 : 0.0) : RVd[(col * 6 + e) * n + o];
}

// After the tangent startup: the derivatives of the post-startup state of object o with respect
// to initial component col, Zd[(col * 33 + e) * n + o] with e = 0..2 s, 3..5 S, 6 + 3 slot + c
// the ring, and of (r, v), RVd[(col * 6 + e) * n + o] (the final derivative of spans of at most
// four steps), from tangent lane col * n + o.
__global__ void tangent_columns_kernel(PropState pt, int64_t n, double* __restrict__ Zd,
    double* __restrict__ RVd)
{
    const int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= 6 * n)
        return;
    const int col = (int)(L / n);
    const int64_t o = L % n, N = pt.N;
    for (int c = 0; c < 3; ++c) {
        Zd[(col * 33 + c) * n + o] = pt.s[c * N + L];
        Zd[(col * 33 + 3 + c) * n + o] = pt.S[c * N + L];
        for (int sl = 0; sl < 9; ++sl)
            Zd[(col * 33 + 6 + sl * 3 + c) * n + o] = pt.hist[(sl * 3 + c) * N + L];
        RVd[(col * 6 + c) * n + o] = pt.r[c * N + L];
        RVd[(col * 6 + 3 + c) * n + o] = pt.v[c * N + L];
    }
}

// Copies (s, S, ring) between the state and a checkpoint slot ck[e * n + o], e = 0..32.
__global__ void checkpoint_kernel(PropState ps, double* __restrict__ ck, int to_checkpoint)
{
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = ps.N;
    if (o >= n)
        return;
    for (int c = 0; c < 3; ++c) {
        double* p[2] = { &ps.s[c * n + o], &ps.S[c * n + o] };
        for (int q = 0; q < 2; ++q) {
            if (to_checkpoint)
                ck[(3 * q + c) * n + o] = *p[q];
            else
                *p[q] = ck[(3 * q + c) * n + o];
        }
        for (int sl = 0; sl < 9; ++sl) {
            double& hv = ps.hist[(sl * 3 + c) * n + o];
            if (to_checkpoint)
                ck[(6 + sl * 3 + c) * n + o] = hv;
            else
                hv = ck[(6 + sl * 3 + c) * n + o];
        }
    }
}

// GJ8 steps k0 .. k1-1 exactly as gj8_chunk_kernel<kGravity> (the same plain evaluations and
// arithmetic, gj8.hpp, so the trajectory is bit-identical; with fixed_point_exit each step stops
// at the corrector's fixed point as the chunk kernel does), storing per step the ITRS point of
// the step's last evaluation in the first three entries of its Jacobian slot,
// Gbuf[((k - k0) * 6 + e) * n + o]; gravity_gradient_kernel then replaces it with the step's
// GCRS gravity Jacobian there (e = xx, xy, xz, yy, yz, zz). The derivative sweeps use it for
// every corrector iterate of the step: the iterates converge to that point (from the
// predictor's point, a few 1e-5 m away for LEO at 60 s), where the gravity gradient changes by
// about 1e-11 of itself, so the derivatives are those of the scheme with each step's Jacobian
// taken at its fixed point. The point, and so G, is the same with or without the fixed-point
// exit (the last evaluation of all ncorr is at the fixed point). The Jacobians are evaluated
// apart from the steps because they do not feed the trajectory: in the steps they would lengthen
// each step's serial chain, while all of a segment's Jacobians evaluate in one launch that
// fills the chip.
template <int W>
__global__ void __launch_bounds__(64 * W)
__attribute__((amdgpu_waves_per_eu(4))) // 128 VGPRs: 4 resident waves per SIMD
gj8_record_kernel(ForceCtx f, PropState ps, int64_t k0,
    int64_t k1, int ncorr, double* __restrict__ Gbuf, int fixed_point_exit)
{
    __shared__ double red[W > 1 ? kParts * 4 * 64 : 1];
    __shared__ double pos[W > 1 ? kPointLen * 64 : 1];
    // the one correction's inputs, parked per lane during a step's evaluations (as in
    // gj8_chunk_kernel: the evaluations get the registers instead of spilling them)
    __shared__ double park[15 * 64];
    __shared__ int more; // the lead's control word for its helpers
    // Every barrier sits in workgroup-uniform control flow: the loops run to the group's last
    // step (the same in every wave, which hold the same objects); a lane past its span, or past
    // the objects, integrates on a history it does not update and stores nothing.
    const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x % 64;
    const int64_t n = ps.N;
    const bool valid = t < n;
    const int64_t i = valid ? t : 0;
    const int64_t ns = valid ? ps.nsteps[i] : -1;
    const int64_t kend = __builtin_amdgcn_readfirstlane(
        (int)std::min<int64_t>(k1, wave_max(ns) + 1));
    if (kend <= k0)
        return;
    if constexpr (W > 1) {
        // The W waves of a workgroup hold the same 64 objects and share the harmonic columns.
        // Waves 1 .. W-1 only sum their parts (gravity_helper_loop) for every evaluation wave 0
        // runs; wave 0 integrates, stores the evaluation points and writes the ring and the
        // state. Two barriers per evaluation, one more at the end.
        const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
        if (w > 0) {
            gravity_helper_loop<false, W>(f.grav, red, pos, &more, w);
            return;
        }
    }
    auto announce = [&](int what) {
        if constexpr (W > 1) {
            if (threadIdx.x == 0)
                more = what;
        }
    };
    const double h = f.h, hh = h * h;
    double r[3], v[3], s[3], S[3];
    for (int c = 0; c < 3; ++c) {
        s[c] = ps.s[c * n + i];
        S[c] = ps.S[c * n + i];
        r[c] = ps.r[c * n + i];
        v[c] = ps.v[c * n + i];
    }
    // the state after the segment (or at the object's last step): the record pass can then
    // serve as the forward pass
    auto store_state = [&]() {
        for (int c = 0; c < 3; ++c) {
            ps.s[c * n + i] = s[c];
            ps.S[c * n + i] = S[c];
            ps.r[c * n + i] = r[c];
            ps.v[c * n + i] = v[c];
        }
    };
    for (int64_t k = k0; k < kend; ++k) {
        const bool act = k <= ns;
        GJ8Sums q;
        gj8_sums(ps.hist, n, i, k, q);
        double sprev[3], anew[3];
        gj8_predict(q, h, s, S, r, v, sprev);
        // the evaluation points and the one correction of gj8_chunk_kernel without sources, bit
        // for bit: the predictor's point rotated into the ITRS, then hh (M^T (T2 + S) + g88 g_ITRS)
        // after each evaluation (contexts with sources have no derivatives)
        const uniform_dptr M = (uniform_dptr)(f.prop_M + 9 * (k + 4));
        double ri[3], rl[3], base[3], t[3], gi[3];
        to_itrs(M, r, ri);
        for (int c = 0; c < 3; ++c)
            t[c] = q.T2[c] + S[c];
        to_itrs(M, t, base);
        double* pk = park + threadIdx.x % 64;
        for (int c = 0; c < 3; ++c) {
            pk[c * 64] = q.T1[c];
            pk[(3 + c) * 64] = q.T2[c];
            pk[(6 + c) * 64] = q.alast[c];
            pk[(9 + c) * 64] = sprev[c];
            pk[(12 + c) * 64] = S[c];
        }
        for (int it = 0; it < ncorr; ++it) {
            announce(1);
            if constexpr (W > 1)
                gravity_itrs_lead<false, W>(ri, f.grav, red, pos, gi);
            else
                gravity_itrs<false>(ri[0], ri[1], ri[2], f.grav, gi[0], gi[1], gi[2]);
            bool same = true;
            for (int c = 0; c < 3; ++c) {
                const double rn = hh * __builtin_fma(c_gj[8][8], gi[c], base[c]);
                same = same && __double_as_longlong(rn) == __double_as_longlong(ri[c]);
                rl[c] = ri[c];
                ri[c] = rn;
            }
            if (fixed_point_exit && __ballot(!(same || !act)) == 0)
                break; // every later evaluation would repeat this one
        }
        to_gcrs(M, gi, anew);
        if (act) {
            // the point of the step's last evaluation, for gravity_gradient_kernel
            for (int c = 0; c < 3; ++c)
                Gbuf[((k - k0) * 6 + c) * n + i] = rl[c];
        }
        {
            GJ8Sums qc;
            double sp[3];
            for (int c = 0; c < 3; ++c) {
                qc.T1[c] = pk[c * 64];
                qc.T2[c] = pk[(3 + c) * 64];
                qc.alast[c] = pk[(6 + c) * 64];
                sp[c] = pk[(9 + c) * 64];
                S[c] = pk[(12 + c) * 64];
            }
            gj8_correct(qc, anew, sp, S, h, hh, s, r, v);
        }
        if (!act)
            continue;
        const int wslot = ring_slot(k);
        for (int c = 0; c < 3; ++c)
            ps.hist[(wslot * 3 + c) * n + i] = anew[c];
        if (k == ns)
            store_state();
    }
    if constexpr (W > 1) {
        // release the helpers (gravity_helper_loop)
        announce(0);
        __syncthreads();
    }
    if (ns >= kend)
        store_state();
}

// Reverse pass over steps k1-1 .. k0 for R cotangent rows. lam[(rho * 33 + e) * n + o] holds the
// cotangents of (s, S, ring) after the step being reversed; seeds[(rho * 6 + e) * n + o] the
// cotangent of the final (r, v), injected at the object's last step. One lane per (object, row):
// block b reverses row b % R of objects 64 (b / R) .. +63, so the R blocks of an object group are
// adjacent in the grid and read the same recorded Jacobians close together in time.
__global__ void __launch_bounds__(64) gj8_reverse_kernel(const int64_t* __restrict__ nsteps,
    int64_t n, double h, int64_t k0, int64_t k1, int ncorr, const double* __restrict__ Gbuf,
    int R, const double* __restrict__ seeds, double* __restrict__ lam)
{
    const int rho = (int)(blockIdx.x % R);
    const int64_t i = (int64_t)(blockIdx.x / R) * 64 + threadIdx.x;
    if (i >= n)
        return;
    const int64_t ns = nsteps[i];
    const int64_t kfirst = std::min<int64_t>(k1 - 1, ns); // steps past the span change nothing
    if (kfirst < k0)
        return;
    const double hh = h * h, sa88 = c_sa[8][8], gj88 = c_gj[8][8];
    double* L = lam + (int64_t)rho * 33 * n;
    // The cotangents in registers for the whole segment: lsp (s), lS (S) and the ring A, where
    // A[j] holds the cotangent of a_{k-8+j} when step k begins (A[8]: the entry a_k that step k
    // wrote); L holds them between launches, the ring at ring_slot of its acceleration's index.
    double lsp[3], lS[3], A[9][3];
    for (int c = 0; c < 3; ++c) {
        lsp[c] = L[c * n + i];
        lS[c] = L[(3 + c) * n + i];
    }
    for (int j = 0; j < 9; ++j)
        for (int c = 0; c < 3; ++c)
            A[j][c] = L[(6 + ring_slot(kfirst - 8 + j) * 3 + c) * n + i];
    // the step's Jacobian, loaded one step ahead
    auto load_g = [&](int64_t k, double* g) {
        const double* gp = Gbuf + (((k - k0) * 6) * n + i);
        for (int e = 0; e < 6; ++e)
            g[e] = gp[e * n];
    };
    double g[6], gn[6];
    load_g(kfirst, g);
    for (int64_t k = kfirst; k >= k0; --k) {
        if (k > k0)
            load_g(k - 1, gn);
        const double gxx = g[0], gxy = g[1], gxz = g[2], gyy = g[3], gyz = g[4], gzz = g[5];
        // the cotangent of iterate it's acceleration gathers only in that iterate's reversal
        // (and, for the last iterate, from the ring entry a_k it became)
        double lk[3];
        double lr[3], lv[3], lT1[3], lT2[3], ls[3], lal[3];
        for (int c = 0; c < 3; ++c) {
            lk[c] = A[8][c]; // the new ring entry a_k is the last iterate's acceleration
            lr[c] = k == ns ? seeds[(rho * 6 + c) * n + i] : 0.0;
            lv[c] = k == ns ? seeds[(rho * 6 + 3 + c) * n + i] : 0.0;
            lT1[c] = lT2[c] = ls[c] = lal[c] = 0.0;
        }
        for (int it = ncorr - 1; it >= 0; --it) {
            double le[3];
            for (int c = 0; c < 3; ++c) {
                le[c] = it == ncorr - 1 ? lk[c] : 0.0;
                const double t = hh * lr[c]; // r_{it+1} = hh ((T2 + gj88 e) + S')
                lT2[c] += t;
                le[c] += gj88 * t;
                lS[c] += t;
                const double u = h * lv[c]; // v_{it+1} = h ((T1 + sa88 e) + s'_it)
                lT1[c] += u;
                le[c] += sa88 * u;
                lsp[c] += u;
                ls[c] += lsp[c]; // s'_it = s + (e + a_{k-1}) / 2
                le[c] += lsp[c] / 2;
                lal[c] += lsp[c] / 2;
            }
            lr[0] = gxx * le[0] + gxy * le[1] + gxz * le[2];
            lr[1] = gxy * le[0] + gyy * le[1] + gyz * le[2];
            lr[2] = gxz * le[0] + gyz * le[1] + gzz * le[2];
            for (int c = 0; c < 3; ++c)
                lv[c] = lsp[c] = 0.0;
        }
        double lPr[3];
        for (int c = 0; c < 3; ++c) {
            lPr[c] = hh * lr[c]; // r_0 = hh (Pr + S')
            lS[c] += hh * lr[c];
            ls[c] += lS[c]; // S' = S + s + a_{k-1} / 2
            lal[c] += lS[c] / 2;
        }
        // Pr, T1, T2 and a_{k-1} over a_{k-9} .. a_{k-1}: entry j of the new ring is a_{k-9+j},
        // the old entry j - 1 (a_{k-9} starts from zero in the slot a_k leaves)
        for (int j = 8; j >= 0; --j) {
            for (int c = 0; c < 3; ++c) {
                double add = c_gj[9][j] * lPr[c];
                if (j >= 1)
                    add += c_sa[8][j - 1] * lT1[c] + c_gj[8][j - 1] * lT2[c];
                if (j == 8)
                    add += lal[c];
                A[j][c] = (j >= 1 ? A[j - 1][c] : 0.0) + add;
            }
        }
        for (int c = 0; c < 3; ++c)
            lsp[c] = ls[c];
        for (int e = 0; e < 6; ++e)
            g[e] = gn[e];
    }
    // the ring now begins step k0 - 1: A[j] is the cotangent of a_{k0-9+j}
    for (int c = 0; c < 3; ++c) {
        L[c * n + i] = lsp[c];
        L[(3 + c) * n + i] = lS[c];
    }
    for (int j = 0; j < 9; ++j)
        for (int c = 0; c < 3; ++c)
            L[(6 + ring_slot(k0 - 9 + j) * 3 + c) * n + i] = A[j][c];
}

// Rows of the Jacobian (or the vector-Jacobian product when R = 1): out[(o * R + rho) * 6 + col].
// Spans longer than four steps: cotangents of Z times dZ/dX; shorter spans: the seeds times
// d(r, v)/dX after the startup (zero steps: the identity).
__global__ void adjoint_combine_kernel(const int64_t* __restrict__ nsteps, int64_t n, int R,
    const double* __restrict__ lam, const double* __restrict__ Zd, const double* __restrict__ RVd,
    const double* __restrict__ seeds, double* __restrict__ out)
{
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n)
        return;
    const int64_t ns = nsteps[o];
    for (int rho = 0; rho < R; ++rho)
        for (int col = 0; col < 6; ++col) {
            double acc = 0.0;
            if (ns > 4) {
                for (int e = 0; e < 33; ++e)
                    acc += lam[(rho * 33 + e) * n + o] * Zd[(col * 33 + e) * n + o];
            } else {
                for (int e = 0; e < 6; ++e) {
                    const double d = ns == 0 ? (e == col ? 1.0 : 0.0) : RVd[(col * 6 + e) * n + o];
                    acc += seeds[(rho * 6 + e) * n + o] * d;
                }
            }
            out[(o * R + rho) * 6 + col] = acc;
        }
}

// ---- launchers ---------------------------------------------------------------------------

void launch_gravity_jac(hipStream_t st, const GravityDev& g, const double* r, int64_t n,
    double* a, double* G)
{
    if (n > 0)
        hipLaunchKernelGGL(gravity_jac_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, g, r,
            n, a, G);
}

void launch_tangent_init(hipStream_t st, const int64_t* steps, int64_t n, const PropState& pt,
    int64_t* ns_out)
{
    if (n > 0)
        hipLaunchKernelGGL(tangent_init_kernel, dim3((unsigned)((6 * n + 255) / 256)), dim3(256), 0,
            st, steps, n, pt, ns_out);
}

void launch_gj8_tangent(hipStream_t st, const int64_t* nsteps, int64_t n, double h, int64_t k0,
    int64_t k1, int ncorr, const double* Gbuf, double* Zd, double* RVd)
{
    if (n > 0 && k1 > k0)
        hipLaunchKernelGGL(gj8_tangent_kernel, dim3((unsigned)(6 * ((n + 63) / 64))), dim3(64), 0,
            st, nsteps, n, h, k0, k1, ncorr, Gbuf, Zd, RVd);
}

void launch_stm_final(hipStream_t st, const int64_t* nsteps, int64_t n, const double* RVd,
    double* stm)
{
    if (n > 0)
        hipLaunchKernelGGL(stm_final_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
            nsteps, n, RVd, stm);
}

void launch_tangent_columns(hipStream_t st, const PropState& pt, int64_t n, double* Zd,
    double* RVd)
{
    if (n > 0)
        hipLaunchKernelGGL(tangent_columns_kernel, dim3((unsigned)((6 * n + 255) / 256)), dim3(256),
            0, st, pt, n, Zd, RVd);
}

void launch_checkpoint(hipStream_t st, const PropState& ps, double* ck, bool to_checkpoint)
{
    if (ps.N > 0)
        hipLaunchKernelGGL(checkpoint_kernel, dim3((unsigned)((ps.N + 255) / 256)), dim3(256), 0, st,
            ps, ck, to_checkpoint ? 1 : 0);
}

void launch_gj8_record(hipStream_t st, const ForceCtx& f, const PropState& ps, int64_t k0,
    int64_t k1, int ncorr, double* Gbuf, int waves, bool fixed_point_exit)
{
    if (ps.N <= 0 || k1 <= k0)
        return;
    const dim3 grid((unsigned)((ps.N + 63) / 64));
#define L_(Wv)                                                                                     \
    hipLaunchKernelGGL(gj8_record_kernel<Wv>, grid, dim3(64 * (Wv)), 0, st, f, ps, k0, k1, ncorr, \
        Gbuf, fixed_point_exit ? 1 : 0)
    if (waves >= 8)
        L_(8);
    else
        L_(4);
#undef L_
    // the segment's Jacobians, behind the steps on the same stream; at most 65,535 steps per
    // launch (the grid's y limit)
    for (int64_t y0 = 0; y0 < k1 - k0; y0 += 65535) {
        const unsigned ny = (unsigned)std::min<int64_t>(65535, k1 - k0 - y0);
        hipLaunchKernelGGL(gravity_gradient_kernel, dim3((unsigned)((ps.N + 63) / 64), ny), dim3(64),
            0, st, f.grav, f.prop_M, ps.nsteps, ps.N, k0 + y0, Gbuf + y0 * 6 * ps.N);
    }
}

void launch_gj8_reverse(hipStream_t st, const int64_t* nsteps, int64_t n, double h, int64_t k0,
    int64_t k1, int ncorr, const double* Gbuf, int R, const double* seeds, double* lam)
{
    if (n > 0 && k1 > k0 && R > 0)
        hipLaunchKernelGGL(gj8_reverse_kernel, dim3((unsigned)(R * ((n + 63) / 64))), dim3(64), 0,
            st, nsteps, n, h, k0, k1, ncorr, Gbuf, R, seeds, lam);
}

void launch_adjoint_combine(hipStream_t st, const int64_t* nsteps, int64_t n, int R,
    const double* lam, const double* Zd, const double* RVd, const double* seeds, double* out)
{
    if (n > 0)
        hipLaunchKernelGGL(adjoint_combine_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
            st, nsteps, n, R, lam, Zd, RVd, seeds, out);
}

int adjoint_max_corrections() { return kMaxCorr; }

} // namespace ogp
