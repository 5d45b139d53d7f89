// Device-side building blocks shared by the kernels: integrator coefficient tables, the
// analytic test forces, frame rotation helpers and the spherical-harmonic gravity evaluation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "coefficients.hpp"
#include "datasets.hpp"

namespace ogp {

// Wave-uniform reads: pointers in the constant address space make the compiler emit scalar
// (SMEM) loads, so a value needed by all 64 lanes travels once per wave through the scalar
// cache instead of once per lane through LDS or the vector memory path.
typedef const __attribute__((address_space(4))) double* uniform_dptr;
typedef const __attribute__((address_space(4))) float* uniform_fptr;
typedef const __attribute__((address_space(4))) int* uniform_iptr;

// ------------------------------------------------------------------------------------------
// Gravity field (Gottlieb 1993, normalized derived Legendre functions)
// ------------------------------------------------------------------------------------------
constexpr int kMaxWaves = 8; // waves sharing one object group (column split)
// The columns m = 0..L form kParts contiguous parts of about equal cost. Each part is summed
// from zero and the part sums are added in part order, whatever the number of waves that share
// the work, so the acceleration is bit-identical for every launch shape.
constexpr int kParts = 8;

struct GravityDev {
    const double* rec;  // kRecordLen doubles per record, see datasets.hpp (GravityModel::records)
    const float* recf;  // the same records in FP32
    const double* jext; // kJacExtLen doubles per record: the gravity gradient's extension
    const double* diag; // Pbar_mm / Pbar_{m-1,m-1}, m = 0..L, and two zero entries after
    const int* parts;   // kParts + 1 column bounds: part q = columns [parts[q], parts[q + 1])
    const double* part_diag; // per part q: D_m = prod_{j=1..m} diag[j] at m = parts[q] (D_0 = 1)
    int L;              // degree and order
    int nsplit;         // split precision: degrees <= nsplit in FP64, above in FP32
    double gm, radius;
};

// Partial harmonic sums of one column range (gravity_source_device.cu:122-210 restated).
// V is double (FP32 parts of the split precision run on float values inside the columns).
template <typename V> struct GravSumsT {
    V G, H, J, K;
};
using GravSums = GravSumsT<double>;

// Fused multiply-add in the precision of the operands (no promotion of FP32 to FP64).
__device__ __forceinline__ float fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float tofloat(double a) { return (float)a; }

// 1 / sqrt(a) for a positive normal a: the hardware estimate (v_rsq_f64, about 24 correct bits)
// and two Newton steps y += y (1/2 - (a/2) y^2). Measured on 2^20 values over 22 decades: below
// one ulp, correctly rounded in 86% of cases. Seven dependent operations, where sqrt followed
// by a division is about seventeen.
__device__ __forceinline__ double drsqrt(double a)
{
    const double y0 = __builtin_amdgcn_rsq(a);
    const double h = 0.5 * a;
    const double y1 = __builtin_fma(y0, __builtin_fma(-h * y0, y0, 0.5), y0);
    return __builtin_fma(y1, __builtin_fma(-h * y1, y1, 0.5), y1);
}

#ifndef OGP_RECORD_BATCH
#define OGP_RECORD_BATCH 4 // gravity records per scalar-load wait
#endif

// Zero records after the last one of each table: a column's last batch is fetched whole.
constexpr int kRecordPad = OGP_RECORD_BATCH - 1;

// The gravity gradient's column sums beside the plain ones (JAC below), with V_n = n U_n:
//   NA1, NA2, NB1, NB2, NE1, NE2 : sum V_n r[1..6] (the radial derivatives of the J/K, Gamma and
//                                  H sums: (R/r)^n differentiates to -n (R/r)^n / r)
//   T1, T2                       : sum U_n tc, U_n ts (second derivative in sin(phi) of the H
//                                  terms of column m - 2)
struct JacColumn {
    double NA1, NA2, NB1, NB2, NE1, NE2, T1, T2;
};

// One column m of records [k0, k1): k0 is the diagonal n = m (accumulated only), the rest run
// the recursion U_n = (xi'_n * eps * rho) U_{n-1} - rho^2 U_{n-2} on the scaled, rescaled
// functions (the rescaling constants are folded into the coefficients, see GravityModel).
// Nine instructions per record: two multiplications and seven FMAs.
// T: type of the table entries; V: type of the per-object values.
//
// Records are fetched BATCH at a time into scalar registers and then consumed, one wait for the
// scalar loads per batch. The diagonal record opens the first batch, and the last batch is
// fetched whole even when it holds fewer records (the loads past k1 read the next column's
// records or the table's kRecordPad zero records, and their values are not used), so a column
// of n records waits ceil(n / BATCH) times. At a few waves per SIMD each wait is exposed, and
// the short columns of low degrees would otherwise wait once per record of their last batch.
// ZONAL: column m = 0, whose records feed the Gamma sum B1 only (its J, K and H terms carry the
// factor m, and it has no sine terms): one FMA per record instead of six.
// JAC: the gradient's sums jc as well (FP64), from the extension records ext; the plain sums
// are formed by the same operations as without JAC, so they are bit-identical. Nine more
// instructions per record (one multiplication and eight FMAs; two for ZONAL).
template <typename T, typename V, bool ZONAL = false, bool JAC = false,
    int BATCH = OGP_RECORD_BATCH>
__device__ __forceinline__ void column_terms(const void* table, int k0, int k1, V er, V rr2,
    V& P, V& Pp, V& A1, V& A2, V& B1, V& B2, V& E1, V& E2, bool first_is_diag,
    const double* ext = nullptr, JacColumn* jc = nullptr)
{
    static_assert(BATCH - 1 <= kRecordPad, "a whole batch must stay inside the padded table");
    constexpr int R = kRecordLen;
    constexpr int X = JAC ? kJacExtLen : 1;
    typedef const __attribute__((address_space(4))) T* tptr;
    typedef const __attribute__((address_space(4))) double* xptr;
    tptr q = (tptr)table + R * (int64_t)k0;
    xptr qx = JAC ? (xptr)ext + X * (int64_t)k0 : nullptr;
    auto accumulate = [&](const T* r, const double* x) {
        B1 = fmat(P, r[3], B1);
        if constexpr (!ZONAL) {
            A1 = fmat(P, r[1], A1);
            A2 = fmat(P, r[2], A2);
            B2 = fmat(P, r[4], B2);
            E1 = fmat(P, r[5], E1);
            E2 = fmat(P, r[6], E2);
        }
        if constexpr (JAC) {
            const V Vn = x[0] * P;
            jc->NB1 = fmat(Vn, r[3], jc->NB1);
            if constexpr (!ZONAL) {
                jc->NA1 = fmat(Vn, r[1], jc->NA1);
                jc->NA2 = fmat(Vn, r[2], jc->NA2);
                jc->NB2 = fmat(Vn, r[4], jc->NB2);
                jc->NE1 = fmat(Vn, r[5], jc->NE1);
                jc->NE2 = fmat(Vn, r[6], jc->NE2);
                jc->T1 = fmat(P, x[1], jc->T1);
                jc->T2 = fmat(P, x[2], jc->T2);
            }
        }
    };
    auto step = [&](const T* r, const double* x) {
        const V Pn = fmat(r[0] * er, P, -(rr2 * Pp));
        Pp = P;
        P = Pn;
        accumulate(r, x);
    };
    auto fetch = [&](T (*rec)[R], double (*xr)[X]) {
#pragma unroll
        for (int b = 0; b < BATCH; ++b) {
#pragma unroll
            for (int e = 0; e < R; ++e)
                rec[b][e] = q[R * b + e];
            if constexpr (JAC) {
#pragma unroll
                for (int e = 0; e < X; ++e)
                    xr[b][e] = qx[X * b + e];
            }
        }
    };
    auto advance = [&]() {
        q += R * BATCH;
        if constexpr (JAC)
            qx += X * BATCH;
    };
    int k = k0;
    if (first_is_diag) {
        T rec[BATCH][R];
        double xr[BATCH][X];
        fetch(rec, xr);
        accumulate(rec[0], xr[0]);
        if (k1 - k0 >= BATCH) {
#pragma unroll
            for (int b = 1; b < BATCH; ++b)
                step(rec[b], xr[b]);
        } else {
#pragma unroll
            for (int b = 1; b < BATCH - 1; ++b)
                if (b < k1 - k0)
                    step(rec[b], xr[b]);
        }
        k += BATCH;
        advance();
    }
    for (; k + BATCH <= k1; k += BATCH) {
        T rec[BATCH][R];
        double xr[BATCH][X];
        fetch(rec, xr);
#pragma unroll
        for (int b = 0; b < BATCH; ++b)
            step(rec[b], xr[b]);
        advance();
    }
    if (k < k1) {
        T rec[BATCH][R];
        double xr[BATCH][X];
        fetch(rec, xr);
        step(rec[0], xr[0]);
#pragma unroll
        for (int b = 1; b < BATCH - 1; ++b)
            if (b < k1 - k)
                step(rec[b], xr[b]);
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
// Geometry shared by all columns of one evaluation point.
template <typename V> struct GravPointT {
    V ir, eps, rho, er, rr2, cphi, U1, W1;
};
using GravPoint = GravPointT<double>;

template <typename V>
__device__ __forceinline__ GravPointT<V> grav_point(V x, V y, V z, double radius)
{
    // 1/r and 1/|(x, y)| by reciprocal square roots (drsqrt): this geometry is on the serial
    // path of every evaluation, between the corrector and the harmonic sums
    GravPointT<V> p;
    const V pxy2 = fmat(y, y, x * x);
    const V r2 = fmat(z, z, pxy2);
    p.ir = drsqrt(r2);
    p.eps = z * p.ir;
    p.rho = radius * p.ir;
    p.er = p.eps * p.rho;
    p.rr2 = p.rho * p.rho;
    p.cphi = 0.0;
    p.U1 = 1.0;
    p.W1 = 0.0;
    if (pxy2 > 0.0) {
        const V ip = drsqrt(pxy2);
        p.cphi = (pxy2 * ip) * p.ir;
        p.U1 = x * ip;
        p.W1 = y * ip;
    }
    return p;
}

// gravity_source_device.cu:219-238 on the complete sums.
template <typename V>
__device__ __forceinline__ void grav_assemble(V x, V y, V z, const GravPointT<V>& p,
    const GravSumsT<V>& s, double gm, V& ax, V& ay, V& az)
{
    const V Lam = (s.G + 1.0) + z * s.H * p.ir;
    const V coef = -gm * p.ir * p.ir;
    ax = coef * fmat(Lam * x, p.ir, -s.J);
    ay = coef * fmat(Lam * y, p.ir, -s.K);
    az = coef * fmat(Lam * z, p.ir, -s.H);
}

// The column recurrences' state at the first column m0 of a part, in closed form:
// (cos m0 lambda, sin m0 lambda) and (cos (m0-1) lambda, sin (m0-1) lambda) as powers of
// (U1, W1) by repeated squaring, and the scaled diagonal rho (cos(phi) rho)^(m0-1) D_m0. Every
// part starts from these values, whatever part precedes it and whichever wave sums it, so the
// sums do not depend on the launch shape; a wave reaches its first part in O(log m0)
// operations and no table loads.
template <typename V>
__device__ __forceinline__ void part_start(const GravPointT<V>& p, int m0, double Dm0, V& Um,
    V& Wm, V& Um1, V& Wm1, V& diagP)
{
    if (m0 == 0) {
        Um = Um1 = 1.0;
        Wm = Wm1 = 0.0;
        diagP = 1.0;
        return;
    }
    V pu = 1.0, pw = 0.0, cr = 1.0;                   // z^(m0-1), (cphi rho)^(m0-1)
    V bu = p.U1, bw = p.W1, cb = p.cphi * p.rho;      // z^(2^i), (cphi rho)^(2^i)
    for (int e = m0 - 1; e > 0;) {                   // e is wave-uniform
        if (e & 1) {
            // a branch, not both products and selects: e is the same in every lane
            __asm__ volatile("");
            const V nu = fmat(pu, bu, -pw * bw);
            pw = fmat(pw, bu, pu * bw);
            pu = nu;
            cr = cr * cb;
        }
        e >>= 1;
        if (e == 0)
            break;
        // double angle on the unit circle: cos 2x = 2 cos^2 x - 1, sin 2x = 2 sin x cos x
        const V twob = bu + bu;
        bw = twob * bw;
        bu = fmat(twob, bu, -1.0);
        cb = cb * cb;
    }
    Um1 = pu;
    Wm1 = pw;
    Um = fmat(pu, p.U1, -pw * p.W1);
    Wm = fmat(pw, p.U1, pu * p.W1);
    diagP = p.rho * cr * Dm0;
}

// The gravity gradient's part sums (JAC evaluations): the plain sums s (bit-identical to the
// plain evaluation's), the radial sums NJ, NK, NH (n-weighted J, K, H) and NF (n (n + 1)
// weighted potential terms), and the second derivatives of the potential's angular sum in the
// direction cosines (x/r, y/r, z/r) = (s, t, sin phi) taken as independent variables (the
// formulation of Pines 1973): Tss (= -Ttt), Tst, Tse, Tte, Tee, each still to be divided by
// cos(phi) once (they pair the functions of column m with the longitude terms of column m - 2).
struct GravJacSums {
    GravSums s;
    double NJ, NK, NH, NF, Tss, Tst, Tse, Tte, Tee;
};
constexpr int kJacSums = 13; // doubles of a GravJacSums

// Sums of parts [q0, q1) (q0 < q1, wave-uniform); sink(q, sums) receives the sums of part q, in
// part order. Each part starts from part_start's closed-form values.
// JAC (FP64 only): the gravity gradient's sums as well; sink then receives GravJacSums.
template <bool SPLIT, bool JAC = false, typename V, class Sink>
__device__ __forceinline__ void grav_parts(const GravPointT<V>& p, const GravityDev& g, int q0,
    int q1, Sink&& sink)
{
    static_assert(!(JAC && SPLIT), "the gradient is evaluated in FP64");
    const V er = p.er, rr2 = p.rr2, cphi = p.cphi, rho = p.rho, U1 = p.U1, W1 = p.W1;
    const float erf_ = tofloat(er), rr2f = tofloat(rr2);
    V Um, Wm;   // cos(m lambda), sin(m lambda)
    V Um1, Wm1; // (m - 1)
    V Um2, Wm2; // (m - 2), JAC only
    V diagP;    // scaled diagonal cos(phi)^(m-1) (R/r)^m Pbar_mm
    const uniform_dptr diag = (uniform_dptr)g.diag;
    const uniform_iptr parts = (uniform_iptr)g.parts;
    const uniform_dptr part_diag = (uniform_dptr)g.part_diag;
    const int L = g.L;
    for (int q = q0; q < q1; ++q) {
        const int mb = parts[q];
        part_start(p, mb, part_diag[q], Um, Wm, Um1, Wm1, diagP);
        if constexpr (JAC) {
            // cos and sin of (mb - 2) lambda; used from column 2 on, which a part that starts
            // at column 0 or 1 reaches through the advance below
            Um2 = fmat(Um1, U1, Wm1 * W1);
            Wm2 = fmat(Wm1, U1, -(Um1 * W1));
        }
        int m = mb;
        int kcol = m * (L + 1) - m * (m - 1) / 2; // first record of column m
        // Gc: the Gamma sum of the columns m >= 1 before its factor cos(phi), applied once
        V G = 0.0, Gc = 0.0, H = 0.0, J = 0.0, K = 0.0;
        // the gradient's part sums (JAC); NFc: the columns m >= 1 of NF before the cos(phi)
        V NJ = 0.0, NK = 0.0, NH = 0.0, NF = 0.0, NFc = 0.0;
        V Tss = 0.0, Tst = 0.0, Tse = 0.0, Tte = 0.0, Tee = 0.0;
        // the next column's diagonal ratio, loaded one column ahead so that its wait is the
        // wait of the current column's first record batch
        double ratio = diag[mb + 1];
        for (const int mend = parts[q + 1]; m < mend; ++m) {
            if (m != mb) { // the part's first diagonal comes from part_start
                if (m == 1)
                    diagP = rho * ratio;
                else
                    diagP *= cphi * rho * ratio;
                ratio = diag[m + 1];
            }
            const int ncol = L - m + 1;
            V A1, A2, B1, B2, E1, E2;
            JacColumn jc;
            if constexpr (JAC) {
                V P = diagP, Pp = 0.0;
                A1 = A2 = B1 = B2 = E1 = E2 = 0.0;
                jc = JacColumn { 0, 0, 0, 0, 0, 0, 0, 0 };
                // batches of 2: the plain and extension records of a batch, 22 doubles, stay
                // within the scalar registers
                if (m == 0)
                    column_terms<double, V, true, true, 2>(g.rec, kcol, kcol + ncol, er, rr2, P,
                        Pp, A1, A2, B1, B2, E1, E2, true, g.jext, &jc);
                else
                    column_terms<double, V, false, true, 2>(g.rec, kcol, kcol + ncol, er, rr2, P,
                        Pp, A1, A2, B1, B2, E1, E2, true, g.jext, &jc);
            } else if constexpr (!SPLIT) {
                V P = diagP, Pp = 0.0;
                A1 = A2 = B1 = B2 = E1 = E2 = 0.0;
                if (m == 0)
                    column_terms<double, V, true>(g.rec, kcol, kcol + ncol, er, rr2, P, Pp, A1, A2,
                        B1, B2, E1, E2, true);
                else
                    column_terms<double>(g.rec, kcol, kcol + ncol, er, rr2, P, Pp, A1, A2, B1, B2,
                        E1, E2, true);
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
                    if (m == 0)
                        column_terms<float, float, true>(g.recf, k32, kcol + ncol, erf_, rr2f, Pf,
                            Ppf, a1, a2, b1, b2, e1, e2, k32 == kcol);
                    else
                        column_terms<float>(g.recf, k32, kcol + ncol, erf_, rr2f, Pf, Ppf, a1, a2,
                            b1, b2, e1, e2, k32 == kcol);
                    A1 += (double)a1;
                    A2 += (double)a2;
                    B1 += (double)b1;
                    B2 += (double)b2;
                    E1 += (double)e1;
                    E2 += (double)e2;
                }
            }
            if (m == 0) {
                G = B1;
                if constexpr (JAC)
                    NF = jc.NB1; // (1 + n) c_n C: r[3] - r[1] with r[1] = 0 in column 0
            } else {
                // A1 and A2 carry the factor m (GravityModel::records)
                Gc = fmat(Um, B1, fmat(Wm, B2, Gc));
                J = fmat(Um1, A1, fmat(Wm1, A2, J));
                K = fmat(Um1, A2, fmat(-Wm1, A1, K));
                H = fmat(Um1, E1, fmat(Wm1, E2, H));
                if constexpr (JAC) {
                    // n (n + 1) c_n C = n ((1 + n + m) - m) c_n C = n (r[3] - r[1])
                    NFc = fmat(Um, jc.NB1 - jc.NA1, fmat(Wm, jc.NB2 - jc.NA2, NFc));
                    NJ = fmat(Um1, jc.NA1, fmat(Wm1, jc.NA2, NJ));
                    NK = fmat(Um1, jc.NA2, fmat(-Wm1, jc.NA1, NK));
                    NH = fmat(Um1, jc.NE1, fmat(Wm1, jc.NE2, NH));
                    if (m >= 2) {
                        // d2/ds2 and d2/ds dt of the J, K terms: m (m - 1) c_n C = (m - 1) r[1];
                        // d/d sin(phi) of the J, K terms of column m - 1, carried here:
                        // (m - 1) zeta_{n,m-1} c_n C_{n,m-1} = (m - 1) r[5]
                        const double f = (double)(m - 1);
                        const V a1 = f * A1, a2 = f * A2, e1 = f * E1, e2 = f * E2;
                        Tss = fmat(Um2, a1, fmat(Wm2, a2, Tss));
                        Tst = fmat(Um2, a2, fmat(-Wm2, a1, Tst));
                        Tse = fmat(Um2, e1, fmat(Wm2, e2, Tse));
                        Tte = fmat(Um2, e2, fmat(-Wm2, e1, Tte));
                        Tee = fmat(Um2, jc.T1, fmat(Wm2, jc.T2, Tee));
                    }
                }
            }
            // advance the longitude terms: (U + iW)_{m+1} = (U + iW)_m (U1 + i W1)
            if constexpr (JAC) {
                Um2 = Um1;
                Wm2 = Wm1;
            }
            Um1 = Um;
            Wm1 = Wm;
            const V un = fmat(Um, U1, -Wm * W1);
            Wm = fmat(Wm, U1, Um * W1);
            Um = un;
            kcol += ncol;
        }
        if constexpr (JAC)
            sink(q, GravJacSums { GravSumsT<V> { fmat(cphi, Gc, G), H, J, K }, NJ, NK, NH,
                                     fmat(cphi, NFc, NF), Tss, Tst, Tse, Tte, Tee });
        else
            sink(q, GravSumsT<V> { fmat(cphi, Gc, G), H, J, K });
    }
}

// The parts wave w of a W-wave group sums: [wave_part(w), wave_part(w) + kParts / W). The
// hardware places the waves of a workgroup on the CU's 4 SIMDs in turn, so waves w and w + 4 of
// an 8-wave group share a SIMD; they take parts 2s and 2s + 1 (s = w mod 4), the pair that a
// 4-wave group gives its wave s, and the part boundaries balance both shapes at once. Fewer
// waves take consecutive parts. Which wave sums a part does not change the result.
template <int W> __device__ __forceinline__ int wave_part(int w)
{
    if constexpr (W == 8)
        return (w & 3) * 2 + (w >> 2);
    else
        return w * (kParts / W);
}

// Adds the part sums, received in part order q = 0 .. kParts - 1, as
// ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7)): the fold every launch shape performs, three
// dependent additions deep instead of seven.
template <typename V> struct PartFoldT {
    GravSumsT<V> t { 0.0, 0.0, 0.0, 0.0 }, quad, pair;
    static __device__ __forceinline__ void add(GravSumsT<V>& a, const GravSumsT<V>& b)
    {
        a.G += b.G;
        a.H += b.H;
        a.J += b.J;
        a.K += b.K;
    }
    __device__ __forceinline__ void operator()(int q, const GravSumsT<V>& s)
    {
        static_assert(kParts == 8, "the fold tree is written for 8 parts");
        if (q % 2 == 0) {
            pair = s;
            return;
        }
        add(pair, s);
        if (q % 4 == 1) {
            quad = pair;
            return;
        }
        add(quad, pair);
        if (q == 3)
            t = quad;
        else
            add(t, quad);
    }
};
using PartFold = PartFoldT<double>;

// The kParts part sums of an LDS exchange (red[(q * 4 + c) * 64 + lane], c = G, H, J, K),
// folded. The reads are issued four parts at a time, one LDS wait per four parts.
__device__ __forceinline__ GravSums fold_exchange(const double* red, int lane)
{
    PartFold f;
#pragma unroll
    for (int h = 0; h < kParts; h += 4) {
        GravSums s[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double* o = red + (h + j) * 4 * 64 + lane;
            s[j] = GravSums { o[0], o[64], o[128], o[192] };
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            f(h + j, s[j]);
    }
    return f.t;
}

template <bool SPLIT>
__device__ __forceinline__ void gravity_itrs(
    double x, double y, double z, const GravityDev& g, double& ax, double& ay, double& az)
{
    const GravPoint p = grav_point(x, y, z, g.radius);
    PartFold f;
    grav_parts<SPLIT>(p, g, 0, kParts, f);
    grav_assemble(x, y, z, p, f.t, g.gm, ax, ay, az);
}

// ------------------------------------------------------------------------------------------
// Gravity gradient: the acceleration and its Jacobian G[i][j] = d a_i / d r_j in the ITRS
// ------------------------------------------------------------------------------------------
// With u = (s, t, e) = r / |r| and the potential U = (gm / r) (1 + sum_n (R/r)^n Phi_n(u)), where
// each Phi_n is a polynomial in s, t and e taken as independent variables (Pines 1973), the
// acceleration is a = -(gm / r^2) (Lam u - Q) with Q = (J, K, H) the first derivatives of the
// angular sum and Lam = 1 + Gamma + e H (grav_assemble). Differentiating once more, with Hs the
// Hessian of the angular sum in u (Tss, Tst, Tse; Tst, -Tss, Tte; Tse, Tte, Tee), NQ = (NJ, NK,
// NH) and NF the n- and n (n + 1)-weighted sums (the radial derivative of (R/r)^n):
//   G = -(gm / r^3) (Lam I - Hs + u W^T + W u^T - (3 Lam + sig) u u^T),
//   W = 2 Q + NQ + Hs u,  sig = NF + u.NQ + u.W.
// The derivatives of the derived Legendre functions are those of the next columns
// (d/de Pbar_n^m = zeta_{n,m} Pbar_n^{m+1}), so every second-derivative sum is formed while the
// column recursion visits column m, from the J/K and H records ((m - 1) times their column sums)
// and the extension records (GravityModel::jac_ext). The functions of column m carry
// cos(phi)^(m-1) and the second derivatives pair them with the longitude terms of column m - 2,
// whose factor is cos(phi)^(m-2): the T sums are divided by cos(phi) once, after the fold.
// Exactly on the polar axis (x = y = 0, cos(phi) = 0) the columns m >= 2 vanish and their
// second-derivative terms are omitted.
//
// The plain sums G, H, J, K are formed and folded by the operations of the plain evaluation, so
// the acceleration is bit-identical to gravity_itrs<false> (and gravity_gcrs<false, W>).

__device__ __forceinline__ void jac_sums_to_array(const GravJacSums& s, double* v)
{
    v[0] = s.s.G;
    v[1] = s.s.H;
    v[2] = s.s.J;
    v[3] = s.s.K;
    v[4] = s.NJ;
    v[5] = s.NK;
    v[6] = s.NH;
    v[7] = s.NF;
    v[8] = s.Tss;
    v[9] = s.Tst;
    v[10] = s.Tse;
    v[11] = s.Tte;
    v[12] = s.Tee;
}

// The part sums added in part order with the fold tree of PartFoldT: the plain sums fold exactly
// as in the plain evaluation.
struct JacFold {
    double t[kJacSums], quad[kJacSums], pair[kJacSums];
    __device__ __forceinline__ void operator()(int q, const double* v)
    {
        static_assert(kParts == 8, "the fold tree is written for 8 parts");
#pragma unroll
        for (int c = 0; c < kJacSums; ++c) {
            if (q % 2 == 0) {
                pair[c] = v[c];
                continue;
            }
            pair[c] += v[c];
            if (q % 4 == 1) {
                quad[c] = pair[c];
                continue;
            }
            quad[c] += pair[c];
            if (q == 3)
                t[c] = quad[c];
            else
                t[c] += quad[c];
        }
    }
};

// The folded sums v (kJacSums, JacFold order) at the point (x, y, z) with geometry p: the
// acceleration a (grav_assemble's) and G.
__device__ __forceinline__ void grav_jac_assemble(double x, double y, double z, const GravPoint& p,
    const double* v, double gm, double* a, double (*G)[3])
{
    const GravSums s { v[0], v[1], v[2], v[3] };
    grav_assemble(x, y, z, p, s, gm, a[0], a[1], a[2]);
    const double ic = p.cphi > 0.0 ? 1.0 / p.cphi : 0.0;
    const double Tss = v[8] * ic, Tst = v[9] * ic, Tse = v[10] * ic, Tte = v[11] * ic,
                 Tee = v[12] * ic;
    const double u[3] = { x * p.ir, y * p.ir, p.eps };
    const double Q[3] = { s.J, s.K, s.H };
    const double NQ[3] = { v[4], v[5], v[6] };
    const double Hs[3][3] = { { Tss, Tst, Tse }, { Tst, -Tss, Tte }, { Tse, Tte, Tee } };
    const double Lam = (s.G + 1.0) + z * s.H * p.ir;
    double Wv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double Hu = fmat(Hs[i][2], u[2], fmat(Hs[i][1], u[1], Hs[i][0] * u[0]));
        Wv[i] = fmat(2.0, Q[i], NQ[i] + Hu);
    }
    const double uNQ = fmat(u[2], NQ[2], fmat(u[1], NQ[1], u[0] * NQ[0]));
    const double uW = fmat(u[2], Wv[2], fmat(u[1], Wv[1], u[0] * Wv[0]));
    const double kuu = fmat(3.0, Lam, v[7] + uNQ + uW);
    const double c3 = -gm * p.ir * p.ir * p.ir;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const double e = (i == j ? Lam : 0.0) - Hs[i][j]
                + fmat(u[i], Wv[j], fmat(u[j], Wv[i], -(kuu * u[i]) * u[j]));
            G[i][j] = c3 * e;
            G[j][i] = G[i][j];
        }
}

// One wave: every part, folded in registers.
__device__ __forceinline__ void gravity_itrs_jac(
    double x, double y, double z, const GravityDev& g, double* a, double (*G)[3])
{
    const GravPoint p = grav_point(x, y, z, g.radius);
    JacFold f;
    grav_parts<false, true>(p, g, 0, kParts, [&](int q, const GravJacSums& s) {
        double v[kJacSums];
        jac_sums_to_array(s, v);
        f(q, v);
    });
    grav_jac_assemble(x, y, z, p, f.t, g.gm, a, G);
}

// Column-split form for a workgroup of W waves holding the same 64 objects: each part's
// kJacSums sums go to LDS (red[(q * kJacSums + c) * 64 + lane], red holding kJacExchange
// doubles, 53 KB), and the folding wave adds the kParts slots in part order.
constexpr int kJacExchange = kParts * kJacSums * 64;

// Wave w's part sums (parts [wave_part(w), wave_part(w) + kParts / W)) into their LDS slots.
template <int W>
__device__ __forceinline__ void jac_parts_to_lds(const GravPoint& p, const GravityDev& g,
    double* red, int w)
{
    const int lane = threadIdx.x % 64;
    grav_parts<false, true>(p, g, wave_part<W>(w), wave_part<W>(w) + kParts / W,
        [&](int q, const GravJacSums& s) {
            double v[kJacSums];
            jac_sums_to_array(s, v);
#pragma unroll
            for (int c = 0; c < kJacSums; ++c)
                red[(q * kJacSums + c) * 64 + lane] = v[c];
        });
}

// The kParts exchanged part sums, folded in part order, then the ITRS value and Jacobian.
__device__ __forceinline__ void jac_fold_assemble(double x, double y, double z,
    const GravPoint& p, const GravityDev& g, const double* red, double* a, double (*G)[3])
{
    const int lane = threadIdx.x % 64;
    JacFold f;
#pragma unroll 1
    for (int q = 0; q < kParts; ++q) {
        double v[kJacSums];
#pragma unroll
        for (int c = 0; c < kJacSums; ++c)
            v[c] = red[(q * kJacSums + c) * 64 + lane];
        f(q, v);
    }
    grav_jac_assemble(x, y, z, p, f.t, g.gm, a, G);
}

// Column-split evaluation for a workgroup of W waves that all hold the same 64 objects: wave w
// sums parts [wave_part(w), wave_part(w) + kParts / W), every part sum goes to its LDS slot, and
// every wave adds the kParts slots in part order, which is the fold of the one-wave evaluation,
// so the result does not depend on W. red holds kParts * 4 * 64 doubles; the first barrier keeps
// the writes of this evaluation behind the reads of the previous one (the waves arrive together:
// they run the same integrator arithmetic between evaluations).
template <bool SPLIT, int W>
__device__ __forceinline__ void gravity_itrs_split(double x, double y, double z,
    const GravityDev& g, double* red, double& ax, double& ay, double& az)
{
    const GravPoint p = grav_point(x, y, z, g.radius);
    if (W == 1) {
        PartFold f;
        grav_parts<SPLIT>(p, g, 0, kParts, f);
        grav_assemble(x, y, z, p, f.t, g.gm, ax, ay, az);
        return;
    }
    static_assert(kParts % W == 0, "waves must divide the part count");
    // The wave index is uniform; readfirstlane keeps it, the part range and every record
    // address in scalar registers.
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
    const int lane = threadIdx.x % 64;
    __syncthreads();
    grav_parts<SPLIT>(p, g, wave_part<W>(w), wave_part<W>(w) + kParts / W,
        [&](int q, const GravSums& s) {
            red[(q * 4 + 0) * 64 + lane] = s.G;
            red[(q * 4 + 1) * 64 + lane] = s.H;
            red[(q * 4 + 2) * 64 + lane] = s.J;
            red[(q * 4 + 3) * 64 + lane] = s.K;
        });
    __syncthreads();
    grav_assemble(x, y, z, p, fold_exchange(red, lane), g.gm, ax, ay, az);
}

// a_GCRS = M g(M^T r) with M the column-major ITRS->GCRS matrix of the node
// (gravity_source.cu:176-178, :208-210). M is wave-uniform.
template <bool SPLIT, int W = 1>
__device__ __forceinline__ void gravity_gcrs(const double* Mg, const double* r,
    const GravityDev& g, double* a, double* red = nullptr)
{
    const uniform_dptr M = (uniform_dptr)Mg;
    double ri[3], gi[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        ri[i] = __builtin_fma(M[i * 3 + 2], r[2], __builtin_fma(M[i * 3 + 1], r[1], M[i * 3] * r[0]));
    if (W == 1)
        gravity_itrs<SPLIT>(ri[0], ri[1], ri[2], g, gi[0], gi[1], gi[2]);
    else
        gravity_itrs_split<SPLIT, W>(ri[0], ri[1], ri[2], g, red, gi[0], gi[1], gi[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        a[i] = __builtin_fma(M[6 + i], gi[2], __builtin_fma(M[3 + i], gi[1], M[i] * gi[0]));
}

// Lead/helper form of gravity_gcrs for a workgroup of W > 1 waves in which only wave 0 (the
// lead) integrates. The lead rotates r into the ITRS, computes the evaluation point's geometry
// (grav_point) and leaves it in pos (kPointLen x 64 doubles) before the first barrier; every
// wave then sums its parts from that geometry (gravity_helper for waves 1 .. W-1) and after the
// second barrier the lead alone folds the parts, assembles and rotates back. The arithmetic is
// that of gravity_gcrs<SPLIT, W>, so the result is bit-identical to it. Barrier hazards: pos is
// rewritten only after the second barrier, behind the helpers' reads; red is rewritten only
// after the next first barrier, behind the lead's fold.
constexpr int kPointLen = 8; // doubles of a GravPoint
__device__ __forceinline__ void put_point(double* pos, int lane, const GravPoint& p)
{
    const double v[kPointLen] = { p.ir, p.eps, p.rho, p.er, p.rr2, p.cphi, p.U1, p.W1 };
#pragma unroll
    for (int e = 0; e < kPointLen; ++e)
        pos[e * 64 + lane] = v[e];
}
__device__ __forceinline__ GravPoint get_point(const double* pos, int lane)
{
    GravPoint p;
    p.ir = pos[lane];
    p.eps = pos[64 + lane];
    p.rho = pos[128 + lane];
    p.er = pos[192 + lane];
    p.rr2 = pos[256 + lane];
    p.cphi = pos[320 + lane];
    p.U1 = pos[384 + lane];
    p.W1 = pos[448 + lane];
    return p;
}

// r_ITRS = M^T r and a_GCRS = M g for the column-major ITRS->GCRS matrix M (wave-uniform), in
// the operation order of gravity_gcrs.
__device__ __forceinline__ void to_itrs(const uniform_dptr M, const double* r, double* ri)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
        ri[i] = __builtin_fma(M[i * 3 + 2], r[2], __builtin_fma(M[i * 3 + 1], r[1], M[i * 3] * r[0]));
}
__device__ __forceinline__ void to_gcrs(const uniform_dptr M, const double* gi, double* a)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
        a[i] = __builtin_fma(M[6 + i], gi[2], __builtin_fma(M[3 + i], gi[1], M[i] * gi[0]));
}

// The lead wave's part of a lead/helper evaluation at the ITRS point ri: the geometry to pos,
// the first barrier, its own parts, the second barrier, the fold and assembly into gi (ITRS).
template <bool SPLIT, int W>
__device__ __forceinline__ void gravity_itrs_lead(const double* ri, const GravityDev& g,
    double* red, double* pos, double* gi)
{
    static_assert(W > 1 && kParts % W == 0, "waves must divide the part count");
    const int lane = threadIdx.x % 64;
    const GravPoint p = grav_point(ri[0], ri[1], ri[2], g.radius);
    put_point(pos, lane, p);
    __syncthreads();
    grav_parts<SPLIT>(p, g, 0, kParts / W, [&](int q, const GravSums& s) {
        red[(q * 4 + 0) * 64 + lane] = s.G;
        red[(q * 4 + 1) * 64 + lane] = s.H;
        red[(q * 4 + 2) * 64 + lane] = s.J;
        red[(q * 4 + 3) * 64 + lane] = s.K;
    });
    __syncthreads();
    grav_assemble(ri[0], ri[1], ri[2], p, fold_exchange(red, lane), g.gm, gi[0], gi[1], gi[2]);
}

template <bool SPLIT, int W>
__device__ __forceinline__ void gravity_gcrs_lead(const double* Mg, const double* r,
    const GravityDev& g, double* red, double* pos, double* a)
{
    const uniform_dptr M = (uniform_dptr)Mg;
    double ri[3], gi[3];
    to_itrs(M, r, ri);
    gravity_itrs_lead<SPLIT, W>(ri, g, red, pos, gi);
    to_gcrs(M, gi, a);
}

// The part of one gravity_gcrs_lead evaluation that wave w (1 <= w < W) runs.
template <bool SPLIT, int W>
__device__ __forceinline__ void gravity_helper(const GravityDev& g, double* red,
    const double* pos, int w)
{
    const int lane = threadIdx.x % 64;
    __syncthreads();
    const GravPoint p = get_point(pos, lane);
    grav_parts<SPLIT>(p, g, wave_part<W>(w), wave_part<W>(w) + kParts / W,
        [&](int q, const GravSums& s) {
            red[(q * 4 + 0) * 64 + lane] = s.G;
            red[(q * 4 + 1) * 64 + lane] = s.H;
            red[(q * 4 + 2) * 64 + lane] = s.J;
            red[(q * 4 + 3) * 64 + lane] = s.K;
        });
    __syncthreads();
}

// Wave w's share of every evaluation the lead runs, for as many evaluations as the lead decides:
// after each first barrier the wave reads the lead's control word `more` (LDS), and leaves when
// it is 0. The lead sets more = 1 before its first evaluation and more = 0, followed by one
// barrier, after its last; it never rewrites `more` between a first barrier and the following
// second barrier, where the helpers read it.
template <bool SPLIT, int W>
__device__ __forceinline__ void gravity_helper_loop(const GravityDev& g, double* red,
    const double* pos, const int* more, int w)
{
    const int lane = threadIdx.x % 64;
    for (;;) {
        __syncthreads();
        if (__builtin_amdgcn_readfirstlane(*more) == 0)
            return;
        const GravPoint p = get_point(pos, lane);
        grav_parts<SPLIT>(p, g, wave_part<W>(w), wave_part<W>(w) + kParts / W,
            [&](int q, const GravSums& s) {
                red[(q * 4 + 0) * 64 + lane] = s.G;
                red[(q * 4 + 1) * 64 + lane] = s.H;
                red[(q * 4 + 2) * 64 + lane] = s.J;
                red[(q * 4 + 3) * 64 + lane] = s.K;
            });
        __syncthreads();
    }
}

// a = M gi and G = M Gi M^T for the column-major ITRS->GCRS matrix M.
__device__ __forceinline__ void jac_to_gcrs(const uniform_dptr M, const double* gi,
    const double (*Gi)[3], double* a, double (*G)[3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
        a[i] = __builtin_fma(M[6 + i], gi[2], __builtin_fma(M[3 + i], gi[1], M[i] * gi[0]));
    // column-major M: M(i, j) = M[j * 3 + i]; T = Gi M^T, G = M T
    double T[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            T[i][j] = Gi[i][0] * M[j] + Gi[i][1] * M[3 + j] + Gi[i][2] * M[6 + j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            G[i][j] = M[i] * T[0][j] + M[3 + i] * T[1][j] + M[6 + i] * T[2][j];
}

// Lead/helper form of the gravity gradient for W > 1 (the startup's recording), as
// gravity_gcrs_lead: wave 0 computes the evaluation point's geometry and leaves it in pos before
// the first barrier, every wave sums its parts there (gravity_record_helper_loop for waves
// 1 .. W-1), and wave 0 alone folds and assembles. The
// acceleration is bit-identical to gravity_itrs_lead<false, W>; the barrier hazards are those of
// gravity_gcrs_lead.
//
// gravity_itrs_jac_lead is the lead wave's part at the ITRS point ri (the geometry to pos, its
// parts, the fold and assembly): gi and its Jacobian Gi in the ITRS.
template <int W>
__device__ __forceinline__ void gravity_itrs_jac_lead(const double* ri, const GravityDev& g,
    double* red, double* pos, double* gi, double (*Gi)[3])
{
    static_assert(W > 1 && kParts % W == 0, "waves must divide the part count");
    const int lane = threadIdx.x % 64;
    const GravPoint p = grav_point(ri[0], ri[1], ri[2], g.radius);
    put_point(pos, lane, p);
    __syncthreads();
    jac_parts_to_lds<W>(p, g, red, 0);
    __syncthreads();
    jac_fold_assemble(ri[0], ri[1], ri[2], p, g, red, gi, Gi);
}

// The helpers' share of the startup's recording, whose lead evaluates the gravity gradient at
// every force evaluation (gravity_itrs_jac_lead): after each first barrier the control word
// `more` is nonzero while the lead evaluates (the gradient's part sums at the geometry in pos to
// red) and 0 at the end (after the lead's last barrier). The lead writes `more` before the first
// barrier of each evaluation; the helpers read it between that barrier and the second.
template <int W>
__device__ __forceinline__ void gravity_record_helper_loop(const GravityDev& g, double* red,
    const double* pos, const int* more, int w)
{
    const int lane = threadIdx.x % 64;
    for (;;) {
        __syncthreads();
        if (__builtin_amdgcn_readfirstlane(*more) == 0)
            return;
        jac_parts_to_lds<W>(get_point(pos, lane), g, red, w);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// Analytic test forces (unit_tests/include/gpu_test_helpers.cuh) and their time mapping
// (unit_tests/include/mock_force_model.cuh:69-73 and :99)
// ------------------------------------------------------------------------------------------
// The startup's derivative (the variational equations of the RK8 startup and its corrector):
//  * kGravityRecord runs the startup of the objects themselves with the gravity gradient
//    evaluation (its acceleration bit-identical to kGravity) and records, for every force
//    evaluation, the GCRS Jacobian G = dg/dr (6 unique entries) at the evaluation's slot
//    (grec_slot_gen / grec_slot_prop);
//  * kRecordedTangent runs the same startup code on tangent states (the derivatives of the
//    state with respect to one initial component each), whose acceleration at an evaluation is
//    G dr with the G recorded for that evaluation of their object. The integrator is linear in
//    states and accelerations, so this is the discrete startup's exact derivative.
enum ForceKind { kGravity = 0, kNone = 1, kConst = 2, kPoly7 = 3, kPoly8 = 4, kPoly9 = 5,
    kExp = 6, kDamped = 7, kGravitySplit = 100, kGravityRecord = 102, kRecordedTangent = 103 };

// Slots of the recorded startup Jacobians: the 80 RK8 stage evaluations, then acceleration
// matrix m (0 = the epoch matrix, 1 .. ncorr = the corrections') at node j.
constexpr int kGenSlots = 80;
__device__ __forceinline__ int64_t grec_slot_gen(int s, int k)
{
    return ((s > 0 ? s - 1 : s) + 4) * 10 + k;
}
__device__ __forceinline__ int64_t grec_slot_prop(int m, int64_t node)
{
    return kGenSlots + 9 * m + node;
}

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
// ------------------------------------------------------------------------------------------
// Acceleration sources added to gravity. The reference's ForceModel sums a list of
// IAccelerationSource outputs with a matrix-vector product (iacceleration_source.cuh:19-33,
// force_model.cu:73-86); only gravity is implemented there. Here every source is a device
// function of the state and of its per-node data, evaluated inside the integrator kernels after
// gravity and added in a fixed order (Sun, Moon, drag), so no per-source buffer exists. A new
// source is one more function called from add_sources with its own mask bit and node data.
// ------------------------------------------------------------------------------------------
enum SourceBit { kSourceSun = 1, kSourceMoon = 2, kSourceDrag = 4 };
constexpr int kSourceMask = kSourceSun | kSourceMoon | kSourceDrag;
constexpr double kGmSun = 1.32712440018e20; // m^3 / s^2
constexpr double kGmMoon = 4.9028e12;       // m^3 / s^2
constexpr double kOmegaEarth = 7.292115e-5; // rad / s, rotation of the atmosphere about GCRS z
constexpr double kAtmRadius = 6378137.0;    // m, altitude reference of the atmosphere
constexpr int kAtmBands = 28;

// Exponential atmosphere (Vallado, Fundamentals of Astrodynamics and Applications, 4th ed.,
// Table 8-4): base altitude [km], base density [kg/m^3], scale height [km] per band.
static __constant__ const double kAtmosphere[kAtmBands][3] = {
    { 0, 1.225, 7.249 }, { 25, 3.899e-2, 6.349 }, { 30, 1.774e-2, 6.682 },
    { 40, 3.972e-3, 7.554 }, { 50, 1.057e-3, 8.382 }, { 60, 3.206e-4, 7.714 },
    { 70, 8.770e-5, 6.549 }, { 80, 1.905e-5, 5.799 }, { 90, 3.396e-6, 5.382 },
    { 100, 5.297e-7, 5.877 }, { 110, 9.661e-8, 7.263 }, { 120, 2.438e-8, 9.473 },
    { 130, 8.484e-9, 12.636 }, { 140, 3.845e-9, 16.149 }, { 150, 2.070e-9, 22.523 },
    { 180, 5.464e-10, 29.740 }, { 200, 2.789e-10, 37.105 }, { 250, 7.248e-11, 45.546 },
    { 300, 2.418e-11, 53.628 }, { 350, 9.518e-12, 53.298 }, { 400, 3.725e-12, 58.515 },
    { 450, 1.585e-12, 60.828 }, { 500, 6.967e-13, 63.822 }, { 600, 1.454e-13, 71.835 },
    { 700, 3.614e-14, 88.667 }, { 800, 1.170e-14, 124.64 }, { 900, 5.245e-15, 181.05 },
    { 1000, 3.019e-15, 268.00 }
};

struct SourcesDev {
    int mask;              // SourceBit set; 0 = gravity only
    double drag_bc;        // C_D A / m, m^2 / kg (per object inside the integrator kernels)
    const double* bc_obj;  // per-object C_D A / m in the state's order, or null = drag_bc for all
};


// Point mass at s (GCRS, m): a += GM ((s - r) / |s - r|^3 - s / |s|^3).
__device__ __forceinline__ void third_body(double gm, const double* s, const double* r, double* a)
{
    const double d0 = s[0] - r[0], d1 = s[1] - r[1], d2 = s[2] - r[2];
    const double dd = fma(d2, d2, fma(d1, d1, d0 * d0));
    const double ss = fma(s[2], s[2], fma(s[1], s[1], s[0] * s[0]));
    const double kd = gm / (dd * sqrt(dd)), ks = gm / (ss * sqrt(ss));
    a[0] += fma(kd, d0, -ks * s[0]);
    a[1] += fma(kd, d1, -ks * s[1]);
    a[2] += fma(kd, d2, -ks * s[2]);
}

// Density at altitude h (m) above the reference sphere: the band whose base is the highest one
// at or below h (the first band below the surface, the last one above 1000 km).
__device__ __forceinline__ double atm_density(double h)
{
    const double km = h * 1e-3;
    int b = 0;
    for (int q = 1; q < kAtmBands; ++q)
        b = km >= kAtmosphere[q][0] ? q : b;
    return kAtmosphere[b][1] * exp((kAtmosphere[b][0] - km) / kAtmosphere[b][2]);
}

// Drag on a cannonball in an atmosphere rotating with the Earth:
// a += -1/2 (C_D A / m) rho |v_r| v_r with v_r = v - omega x r, omega = (0, 0, omega_E).
__device__ __forceinline__ void drag(double bc, const double* r, const double* v, double* a)
{
    const double rr = sqrt(fma(r[2], r[2], fma(r[1], r[1], r[0] * r[0])));
    const double vr0 = fma(kOmegaEarth, r[1], v[0]), vr1 = fma(-kOmegaEarth, r[0], v[1]),
                 vr2 = v[2];
    const double vn = sqrt(fma(vr2, vr2, fma(vr1, vr1, vr0 * vr0)));
    const double k = -0.5 * bc * atm_density(rr - kAtmRadius) * vn;
    a[0] = fma(k, vr0, a[0]);
    a[1] = fma(k, vr1, a[1]);
    a[2] = fma(k, vr2, a[2]);
}

// The sources' epilogue at one node: tb = that node's Sun and Moon GCRS positions (6 doubles,
// wave-uniform).
__device__ __forceinline__ void add_sources(const SourcesDev& src, const double* tb_node,
    const double* r, const double* v, double* a)
{
    if (src.mask == 0)
        return;
    const uniform_dptr tb = (uniform_dptr)tb_node;
    if (src.mask & kSourceSun) {
        const double s[3] = { tb[0], tb[1], tb[2] };
        third_body(kGmSun, s, r, a);
    }
    if (src.mask & kSourceMoon) {
        const double s[3] = { tb[3], tb[4], tb[5] };
        third_body(kGmMoon, s, r, a);
    }
    if (src.mask & kSourceDrag)
        drag(src.drag_bc, r, v, a);
}

struct ForceCtx {
    GravityDev grav;
    const double* gen_M;  // 80 RK8 stage nodes x 9
    const double* prop_M; // propagation nodes x 9 (node j at time (j - 4) h + t0)
    const double* gen_tb; // Sun and Moon GCRS positions of the same nodes, 6 per node
    const double* prop_tb;
    SourcesDev src;
    double h;
    // recorded startup Jacobians (kGravityRecord writes, kRecordedTangent reads): entry e of slot
    // q of object o at grec[(q * 6 + e) * grec_n + o]; object_ctx points grec at the object
    double* grec = nullptr;
    int64_t grec_n = 1;
};

// The force context of object i (index into the state): drag with that object's coefficient
// when per-object coefficients are given; the recorded Jacobians of object i mod grec_n (the
// tangent lanes of object o are o, o + n, ..., o + 5n).
__device__ __forceinline__ ForceCtx object_ctx(const ForceCtx& f, int64_t i)
{
    ForceCtx o = f;
    if (f.src.bc_obj)
        o.src.drag_bc = f.src.bc_obj[i];
    if (f.grec)
        o.grec = f.grec + i % f.grec_n;
    return o;
}

// Per-workgroup exchange areas of the column-split evaluation (unused when W == 1): the part
// sums, and the lead's geometry of the evaluation point (lead/helper form). Per lane: mat counts
// the startup's acceleration matrices begun (the recorded Jacobians' slots), on says whether the
// lane records (its object runs).
struct Exchange {
    double* red;
    double* pos = nullptr;
    int mat = 0;
    bool on = true;
};

// Entry e of recorded slot q of the object (ForceCtx::grec after object_ctx).
__device__ __forceinline__ double& grec_at(const ForceCtx& f, int64_t q, int e)
{
    return f.grec[(q * 6 + e) * f.grec_n];
}

// kGravityRecord: the gravity and its gradient at the GCRS position x for node matrix Mg (W > 1:
// the lead's part of a lead/helper evaluation, whose helpers run gravity_record_helper_loop),
// the acceleration into a (bit-identical to gravity_gcrs<false, W>) and G into slot q.
template <int W>
__device__ __forceinline__ void record_gravity(const ForceCtx& f, const double* Mg,
    const double* x, int64_t q, double* a, Exchange& ex)
{
    const uniform_dptr M = (uniform_dptr)Mg;
    double ri[3], gi[3], Gi[3][3], G[3][3];
    to_itrs(M, x, ri);
    if constexpr (W > 1)
        gravity_itrs_jac_lead<W>(ri, f.grav, ex.red, ex.pos, gi, Gi);
    else
        gravity_itrs_jac(ri[0], ri[1], ri[2], f.grav, gi, Gi);
    jac_to_gcrs(M, gi, Gi, a, G);
    if (ex.on) {
        const double g6[6] = { G[0][0], G[0][1], G[0][2], G[1][1], G[1][2], G[2][2] };
#pragma unroll
        for (int e = 0; e < 6; ++e)
            grec_at(f, q, e) = g6[e];
    }
}

// kRecordedTangent: a = G dx with the G of slot q (symmetric, 6 stored entries).
__device__ __forceinline__ void recorded_tangent(const ForceCtx& f, int64_t q, const double* x,
    double* a)
{
    const double gxx = grec_at(f, q, 0), gxy = grec_at(f, q, 1), gxz = grec_at(f, q, 2),
                 gyy = grec_at(f, q, 3), gyz = grec_at(f, q, 4), gzz = grec_at(f, q, 5);
    a[0] = gxx * x[0] + gxy * x[1] + gxz * x[2];
    a[1] = gxy * x[0] + gyy * x[1] + gyz * x[2];
    a[2] = gxz * x[0] + gyz * x[1] + gzz * x[2];
}

// ForceModel::ComputeGenerationAcceleration (force_model.cu:59-89): RK8 stage k of startup
// step s (s = +-1..4).
template <int KIND, int W = 1>
__device__ __forceinline__ void force_gen(const ForceCtx& f, const double* x, const double* v,
    int s, int k, double* a, Exchange& ex)
{
    if (KIND == kGravity || KIND == kGravitySplit) {
        const int idx = ((s > 0 ? s - 1 : s) + 4) * 10 + k;
        gravity_gcrs<KIND == kGravitySplit, W>(f.gen_M + 9 * idx, x, f.grav, a, ex.red);
        add_sources(f.src, f.gen_tb + 6 * idx, x, v, a);
    } else if (KIND == kGravityRecord) {
        const int idx = ((s > 0 ? s - 1 : s) + 4) * 10 + k;
        record_gravity<W>(f, f.gen_M + 9 * idx, x, grec_slot_gen(s, k), a, ex);
    } else if (KIND == kRecordedTangent) {
        recorded_tangent(f, grec_slot_gen(s, k), x, a);
    } else {
        const int dir = s > 0 ? -1 : 1;
        const double t = (double)(s + dir) * f.h + (double)(-dir) * c_rk8c[k] * f.h;
        analytic_force<KIND>(x, v, t, a);
    }
}

// ForceModel::ComputePropagationAcceleration (force_model.cu:91-109) at node j.
template <int KIND, int W = 1>
__device__ __forceinline__ void force_prop(const ForceCtx& f, const double* x, const double* v,
    int64_t node, double* a, Exchange& ex)
{
    if (KIND == kGravity || KIND == kGravitySplit) {
        gravity_gcrs<KIND == kGravitySplit, W>(f.prop_M + 9 * node, x, f.grav, a, ex.red);
        add_sources(f.src, f.prop_tb + 6 * node, x, v, a);
    } else if (KIND == kGravityRecord) {
        record_gravity<W>(f, f.prop_M + 9 * node, x, grec_slot_prop(ex.mat - 1, node), a, ex);
    } else if (KIND == kRecordedTangent) {
        recorded_tangent(f, grec_slot_prop(ex.mat - 1, node), x, a);
    } else {
        const double t = (double)(node - 4) * f.h;
        analytic_force<KIND>(x, v, t, a);
    }
}

// force_gen for the lead wave of a lead/helper workgroup (gravity kinds, W > 1); bit-identical
// to force_gen<KIND, W>.
template <int KIND, int W>
__device__ __forceinline__ void force_gen_lead(const ForceCtx& f, const double* x, const double* v,
    int s, int k, double* a, double* red, double* pos)
{
    static_assert(KIND == kGravity || KIND == kGravitySplit, "gravity kinds only");
    const int idx = ((s > 0 ? s - 1 : s) + 4) * 10 + k;
    gravity_gcrs_lead<KIND == kGravitySplit, W>(f.gen_M + 9 * idx, x, f.grav, red, pos, a);
    add_sources(f.src, f.gen_tb + 6 * idx, x, v, a);
}

// The largest x over the 64 lanes of the wave, returned in every lane; every lane must be active.
__device__ __forceinline__ int64_t wave_max(int64_t x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int64_t y = __shfl_xor(x, off, 64);
        x = x > y ? x : y;
    }
    return x;
}

// force_prop for the lead wave of a lead/helper workgroup (gravity kinds, W > 1); bit-identical
// to force_prop<KIND, W>.
template <int KIND, int W>
__device__ __forceinline__ void force_prop_lead(const ForceCtx& f, const double* x,
    const double* v, int64_t node, double* a, double* red, double* pos)
{
    static_assert(KIND == kGravity || KIND == kGravitySplit, "gravity kinds only");
    gravity_gcrs_lead<KIND == kGravitySplit, W>(f.prop_M + 9 * node, x, f.grav, red, pos, a);
    add_sources(f.src, f.prop_tb + 6 * node, x, v, a);
}

} // namespace ogp
