// Device-side building blocks shared by the kernels: integrator coefficient tables, the
// analytic test forces, frame rotation helpers and the spherical-harmonic gravity evaluation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ogp {

// Integrator tables (Berry & Healy 2004 ordinate form, 8th order), same values as
// lib/common/gpu_constants/gpu_constants/gpu_constants.cuh:22-98. Row 9 is the predictor,
// rows 0-8 the mid-corrector rows j = -4..4. Stored [row][k].
extern __constant__ double c_gj[10][9];
extern __constant__ double c_sa[10][9];
// RK8 (10 stages): a[i][j] for stage i+1, b, nodes c. a[6][0] is selected per context (quirk).
extern __constant__ double c_rk8b[10];
extern __constant__ double c_rk8c[10];

// Wave-uniform reads: pointers in the constant address space make the compiler emit scalar
// (SMEM) loads, so a value needed by all 64 lanes travels once per wave through the scalar
// cache instead of once per lane through LDS or the vector memory path.
typedef const __attribute__((address_space(4))) double* uniform_dptr;
typedef const __attribute__((address_space(4))) float* uniform_fptr;

// ------------------------------------------------------------------------------------------
// Gravity field (Gottlieb 1993, normalized derived Legendre functions)
// ------------------------------------------------------------------------------------------
struct GravityDev {
    const double* rec;  // 8 doubles per record, see datasets.hpp (GravityModel::records)
    const float* recf;  // the same records in FP32
    const double* diag; // Pbar_mm / Pbar_{m-1,m-1}
    int L;              // degree and order
    int nsplit;         // split precision: degrees <= nsplit in FP64, above in FP32
    double gm, radius;
};

// Fused multiply-add in the precision of the operands (no promotion of FP32 to FP64).
__device__ __forceinline__ float fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }

// One column m of records [k0, k1): k0 is the diagonal n = m (accumulated only), the rest run
// the recursion P_n = (xi * eps * rho) P_{n-1} - (eta * rho^2) P_{n-2} on the scaled functions.
template <typename T>
__device__ __forceinline__ void column_terms(const void* table, int k0, int k1, T er, T rr2,
    T& P, T& Pp, T& A1, T& A2, T& B1, T& B2, T& E1, T& E2, bool first_is_diag)
{
    typedef const __attribute__((address_space(4))) T* tptr;
    tptr q = (tptr)table + 8 * (int64_t)k0;
    int k = k0;
    if (first_is_diag) {
        A1 = fmat(P, q[2], A1);
        A2 = fmat(P, q[3], A2);
        B1 = fmat(P, q[4], B1);
        B2 = fmat(P, q[5], B2);
        E1 = fmat(P, q[6], E1);
        E2 = fmat(P, q[7], E2);
        q += 8;
        ++k;
    }
#pragma unroll 2
    for (; k < k1; ++k, q += 8) {
        const T xi = q[0], eta = q[1];
        const T a = xi * er;
        const T b = eta * rr2;
        const T Pn = fmat(a, P, -b * Pp);
        Pp = P;
        P = Pn;
        A1 = fmat(P, q[2], A1);
        A2 = fmat(P, q[3], A2);
        B1 = fmat(P, q[4], B1);
        B2 = fmat(P, q[5], B2);
        E1 = fmat(P, q[6], E1);
        E2 = fmat(P, q[7], E2);
    }
}

// Acceleration in the ITRS at (x, y, z), same quantity as gravity_source_device.cu:63-239.
//
// Differences in formulation (same mathematics, different summation order):
//  * the sums over m and n run column by column (m outer), with (R/r)^n folded into the
//    recursion so no per-degree partial sums are stored;
//  * column m >= 1 is scaled by cos(phi)^(m-1) and the longitude terms use cos(m lambda),
//    sin(m lambda): the scaled functions stay O(sqrt(2n+1)) everywhere, which keeps the FP32
//    part of the split-precision path free of overflow and cancellation near the poles;
//  * the H term pairing (n, m-1) with Pbar_n^m is accumulated while column m is visited.
// SPLIT: degrees n <= nsplit in FP64, n > nsplit in FP32; the per-column partial sums are
// promoted to FP64 before they meet the longitude terms.
template <bool SPLIT>
__device__ __forceinline__ void gravity_itrs(
    double x, double y, double z, const GravityDev& g, double& ax, double& ay, double& az)
{
    const double pxy2 = __builtin_fma(y, y, x * x);
    const double r2 = __builtin_fma(z, z, pxy2);
    const double r = __builtin_sqrt(r2);
    const double ir = 1.0 / r;
    const double eps = z * ir;
    const double rho = g.radius * ir;
    const double er = eps * rho;
    const double rr2 = rho * rho;
    const double pxy = __builtin_sqrt(pxy2);
    const double cphi = pxy * ir;
    double U1 = 1.0, W1 = 0.0;
    if (pxy > 0.0) {
        const double ip = 1.0 / pxy;
        U1 = x * ip;
        W1 = y * ip;
    }
    const float erf_ = (float)er, rr2f = (float)rr2;

    double G = 0.0, H = 0.0, J = 0.0, K = 0.0;
    double Um = 1.0, Wm = 0.0;   // cos(m lambda), sin(m lambda)
    double Um1 = 1.0, Wm1 = 0.0; // (m - 1)
    double diagP = 1.0;          // scaled diagonal cos(phi)^(m-1) (R/r)^m Pbar_mm
    const uniform_dptr diag = (uniform_dptr)g.diag;
    const int L = g.L;
    int kcol = 0; // first record of column m
    for (int m = 0; m <= L; ++m) {
        if (m == 1)
            diagP = rho * diag[1];
        else if (m >= 2)
            diagP *= cphi * rho * diag[m];
        const int ncol = L - m + 1;
        double A1, A2, B1, B2, E1, E2;
        if (!SPLIT) {
            double P = diagP, Pp = 0.0;
            A1 = A2 = B1 = B2 = E1 = E2 = 0.0;
            column_terms<double>(g.rec, kcol, kcol + ncol, er, rr2, P, Pp, A1, A2, B1, B2, E1,
                E2, true);
        } else {
            double P = diagP, Pp = 0.0;
            A1 = A2 = B1 = B2 = E1 = E2 = 0.0;
            int k32 = kcol; // first FP32 record
            if (m <= g.nsplit) {
                const int n64 = (g.nsplit < L ? g.nsplit : L) - m + 1;
                column_terms<double>(g.rec, kcol, kcol + n64, er, rr2, P, Pp, A1, A2, B1, B2,
                    E1, E2, true);
                k32 = kcol + n64;
            }
            if (k32 < kcol + ncol) {
                float Pf = (float)P, Ppf = (float)Pp;
                float a1 = 0.f, a2 = 0.f, b1 = 0.f, b2 = 0.f, e1 = 0.f, e2 = 0.f;
                column_terms<float>(g.recf, k32, kcol + ncol, erf_, rr2f, Pf, Ppf, a1, a2, b1,
                    b2, e1, e2, k32 == kcol);
                A1 += (double)a1;
                A2 += (double)a2;
                B1 += (double)b1;
                B2 += (double)b2;
                E1 += (double)e1;
                E2 += (double)e2;
            }
        }
        if (m == 0) {
            G += B1;
        } else {
            const double dm = (double)m;
            G += cphi * __builtin_fma(Um, B1, Wm * B2);
            J += dm * __builtin_fma(Um1, A1, Wm1 * A2);
            K -= dm * __builtin_fma(Wm1, A1, -Um1 * A2);
            H += __builtin_fma(Um1, E1, Wm1 * E2);
        }
        // advance the longitude terms: (U + iW)_{m+1} = (U + iW)_m (U1 + i W1)
        Um1 = Um;
        Wm1 = Wm;
        const double un = __builtin_fma(Um, U1, -Wm * W1);
        Wm = __builtin_fma(Wm, U1, Um * W1);
        Um = un;
        kcol += ncol;
    }
    // gravity_source_device.cu:219-238
    const double Lam = (G + 1.0) + z * H * ir;
    const double coef = -g.gm * ir * ir;
    ax = coef * __builtin_fma(Lam * x, ir, -J);
    ay = coef * __builtin_fma(Lam * y, ir, -K);
    az = coef * __builtin_fma(Lam * z, ir, -H);
}

// a_GCRS = M g(M^T r) with M the column-major ITRS->GCRS matrix of the node
// (gravity_source.cu:176-178, :208-210). M is wave-uniform.
template <bool SPLIT>
__device__ __forceinline__ void gravity_gcrs(
    const double* Mg, const double* r, const GravityDev& g, double* a)
{
    const uniform_dptr M = (uniform_dptr)Mg;
    double ri[3], gi[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        ri[i] = __builtin_fma(M[i * 3 + 2], r[2], __builtin_fma(M[i * 3 + 1], r[1], M[i * 3] * r[0]));
    gravity_itrs<SPLIT>(ri[0], ri[1], ri[2], g, gi[0], gi[1], gi[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        a[i] = __builtin_fma(M[6 + i], gi[2], __builtin_fma(M[3 + i], gi[1], M[i] * gi[0]));
}

// ------------------------------------------------------------------------------------------
// Analytic test forces (unit_tests/include/gpu_test_helpers.cuh) and their time mapping
// (unit_tests/include/mock_force_model.cuh:69-73 and :99)
// ------------------------------------------------------------------------------------------
enum ForceKind { kGravity = 0, kNone = 1, kConst = 2, kPoly7 = 3, kPoly8 = 4, kPoly9 = 5,
    kExp = 6, kDamped = 7, kGravitySplit = 100 };

template <int KIND>
__device__ __forceinline__ void analytic_force(const double* p, const double* v, double t, double* a)
{
    double val = 0.0;
    if (KIND == kConst) {
        a[0] = 1.0;
        a[1] = 2.0;
        a[2] = 3.0;
        return;
    }
    if (KIND == kDamped) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
            a[i] = -2.0 * p[i] - 2.0 * v[i];
        return;
    }
    if (KIND == kPoly7)
        val = -3 * 2 * t + 7 * 6 * t * t * t * t * t;
    else if (KIND == kPoly8)
        val = 56 * pow(t, 6) - 42 * pow(t, 5) + 60 * pow(t, 4) - 100 * pow(t, 3) - 120 * pow(t, 2)
            + 6 * t - 2;
    else if (KIND == kPoly9)
        val = 72 * pow(t, 7) + 56 * pow(t, 6) - 42 * pow(t, 5) + 60 * pow(t, 4) - 100 * pow(t, 3)
            - 120 * pow(t, 2) + 6 * t - 2;
    else if (KIND == kExp)
        val = exp(-t);
    a[0] = a[1] = a[2] = val;
}

// Everything a force evaluation needs for one step-size class.
struct ForceCtx {
    GravityDev grav;
    const double* gen_M;  // 80 RK8 stage nodes x 9
    const double* prop_M; // propagation nodes x 9 (node j at time (j - 4) h + t0)
    double h;
};

// ForceModel::ComputeGenerationAcceleration (force_model.cu:59-89): RK8 stage k of startup
// step s (s = +-1..4).
template <int KIND>
__device__ __forceinline__ void force_gen(const ForceCtx& f, const double* x, const double* v,
    int s, int k, double* a)
{
    if (KIND == kGravity || KIND == kGravitySplit) {
        const int idx = ((s > 0 ? s - 1 : s) + 4) * 10 + k;
        gravity_gcrs<KIND == kGravitySplit>(f.gen_M + 9 * idx, x, f.grav, a);
    } else {
        const int dir = s > 0 ? -1 : 1;
        const double t = (double)(s + dir) * f.h + (double)(-dir) * c_rk8c[k] * f.h;
        analytic_force<KIND>(x, v, t, a);
    }
}

// ForceModel::ComputePropagationAcceleration (force_model.cu:91-109) at node j.
template <int KIND>
__device__ __forceinline__ void force_prop(
    const ForceCtx& f, const double* x, const double* v, int64_t node, double* a)
{
    if (KIND == kGravity || KIND == kGravitySplit) {
        gravity_gcrs<KIND == kGravitySplit>(f.prop_M + 9 * node, x, f.grav, a);
    } else {
        const double t = (double)(node - 4) * f.h;
        analytic_force<KIND>(x, v, t, a);
    }
}

} // namespace ogp
