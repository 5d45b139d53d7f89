// Propagation kernels for gfx950: time nodes, gravity, RK8 startup and the GJ8 step chunks.
#include "kernels.hpp"

#include <math.h>

namespace ogp {


// =========================================================================================
// Time nodes: ITRS->GCRS matrix per node (time_points.cu:26-71, ref_sys_trans.cu)
// =========================================================================================
namespace frames {

constexpr double kPi = 3.14159265358979324;
constexpr double kArcsPerRad = 3600.0 * 180.0 / kPi;
constexpr double kMjdJ2000 = 51544.5;

__device__ double calendar_to_mjd(int year, int mon, int day, int h, int min, double sec)
{
    if (mon <= 2) {
        --year;
        mon += 12;
    }
    int B;
    const long key = 10000L * year + 100L * mon + day;
    if (key <= 15821004L)
        B = -2 + ((year + 4716) / 4) - 1179;
    else if (key >= 15821015L)
        B = (year / 400) - (year / 100) + (year / 4);
    else
        return -1.0;
    const long mjd = 365L * year - 679004L + B + (int)(30.6001 * (mon + 1)) + day;
    const double frac = (h + min / 60.0 + sec / 3600.0) / 24.0;
    return (double)mjd + frac;
}

struct Eop {
    double xp, yp, ut1_utc;
};

// Binary search and linear interpolation, clamped at the table ends (types.cuh:79-130).
__device__ Eop eop_at(const FrameDev& fd, double t)
{
    const double* e = fd.eop;
    const int64_t n = fd.n_eop;
    if (t <= e[0])
        return { e[1], e[2], e[3] };
    if (t >= e[4 * (n - 1)])
        return { e[4 * (n - 1) + 1], e[4 * (n - 1) + 2], e[4 * (n - 1) + 3] };
    int64_t lo = 0, hi = n - 1;
    while (lo <= hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        const double c = e[4 * mid];
        if (c == t)
            return { e[4 * mid + 1], e[4 * mid + 2], e[4 * mid + 3] };
        if (c < t)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    const double* a = e + 4 * hi;
    const double* b = e + 4 * lo;
    const double f = (t - a[0]) / (b[0] - a[0]);
    return { a[1] + f * (b[1] - a[1]), a[2] + f * (b[2] - a[2]), a[3] + f * (b[3] - a[3]) };
}

__device__ double leap_at(const FrameDev& fd, double t)
{
    for (int i = fd.n_leap; i-- > 0;)
        if (t >= fd.leap_mjd[i])
            return fd.leap_sec[i];
    return fd.leap_sec[0];
}

// IERS 2010 eq. 5.43 and 5.44 (Delaunay and planetary arguments), mod 2 pi.
__device__ void fund_args(double t, double* F)
{
    const double l[5] = { 134.96340251 * 3600.0, 1717915923.2178, 31.8792, 0.051635, -0.00024470 };
    const double lp[5] = { 357.52910918 * 3600.0, 129596581.0481, -0.5532, 0.000136, -0.00001149 };
    const double f[5] = { 93.27209062 * 3600.0, 1739527262.8478, -12.7512, -0.001037, 0.00000417 };
    const double d[5] = { 297.85019547 * 3600.0, 1602961601.2090, -6.3706, 0.006593, -0.00003169 };
    const double om[5] = { 125.04455501 * 3600.0, -6962890.5431, 7.4722, 0.007702, -0.00005939 };
    const double* ls[5] = { l, lp, f, d, om };
    for (int q = 0; q < 5; ++q) {
        double v = ls[q][4];
        for (int i = 3; i >= 0; --i)
            v = v * t + ls[q][i];
        F[q] = fmod(v / kArcsPerRad, 2.0 * M_PI);
    }
    const double pl[9][3] = { { 4.402608842, 2608.7903141574, 0.0 },
        { 3.176146697, 1021.3285546211, 0.0 }, { 1.753470314, 628.3075849991, 0.0 },
        { 6.203480913, 334.0612426700, 0.0 }, { 0.599546497, 52.9690962641, 0.0 },
        { 0.874016757, 21.3299104960, 0.0 }, { 5.481293872, 7.4781598567, 0.0 },
        { 5.311886287, 3.8133035638, 0.0 }, { 0.0, 0.02438175, 0.00000538691 } };
    for (int q = 0; q < 9; ++q)
        F[5 + q] = fmod((pl[q][2] * t + pl[q][1]) * t + pl[q][0], 2.0 * M_PI);
}

// IERS 2010 eq. 5.40 (epsilon_A, gamma_bar, phi_bar, psi_bar), radians.
__device__ void precession(double t, double* o)
{
    const double c[4][6] = {
        { 84381.406, -46.836769, -0.0001831, 0.00200340, -0.000000576, -0.0000000434 },
        { -0.052928, 10.556378, 0.4932044, -0.00031238, -0.000002788, 0.0000000260 },
        { 84381.412819, -46.811016, 0.0511268, 0.00053289, -0.000000440, -0.0000000176 },
        { -0.041775, 5038.481484, 1.5584175, -0.00018522, -0.000026452, -0.0000000148 } };
    for (int q = 0; q < 4; ++q) {
        double v = c[q][5];
        for (int i = 4; i >= 0; --i)
            v = v * t + c[q][i];
        o[q] = v / kArcsPerRad;
    }
}

// Greenwich apparent sidereal time, equinox method (ref_sys_trans.cu:75-119).
__device__ double gast(double mjd_ut1, double mjd_tt, double dpsi, double epsA, const double* F)
{
    const double T = (mjd_tt - kMjdJ2000) / 36525.0;
    const double Du = mjd_ut1 - kMjdJ2000;
    double gmst = 86400.0 * (0.7790572732640 + 0.00273781191135448 * Du + (Du - floor(Du)));
    const double c[6] = { 0.00096707, 307.47710227, 0.092772113, -0.0000000293, -0.00000199707,
        -2.453e-9 };
    double poly = c[5];
    for (int i = 4; i >= 0; --i)
        poly = poly * T + c[i];
    gmst = (gmst + poly) * (M_PI / 43200.0);
    const double l = F[0], lp = F[1], f = F[2], D = F[3], Om = F[4];
    double ee = -0.00000087 * T * sin(Om);
    ee -= 0.00000063 * sin(l - Om);
    ee -= 0.00000063 * sin(l + Om);
    ee -= 0.00000126 * sin(lp - Om);
    ee -= 0.00000141 * sin(lp + Om);
    ee -= 0.00000172 * sin(3.0 * Om);
    ee += 0.00000198 * sin(2.0 * f + Om);
    ee += 0.00000202 * sin(2.0 * f + 3.0 * Om);
    ee -= 0.00000455 * sin(2.0 * f - 2.0 * D + 2.0 * Om);
    ee += 0.00001121 * sin(2.0 * f - 2.0 * D + Om);
    ee += 0.00001175 * sin(2.0 * f - 2.0 * D + 3.0 * Om);
    ee += 0.00006352 * sin(2.0 * Om);
    ee += 0.00264096 * sin(Om);
    ee = ee * ((M_PI / 43200.0) / 15.0) + dpsi * cos(epsA);
    return fmod(gmst + ee, 2.0 * M_PI);
}

// Rotation matrices in the SOFA sign convention, column-major (ref_sys_trans.cu:24-67).
__device__ void rot(int axis, double a, double* p)
{
    const double c = cos(a), s = sin(a);
    for (int i = 0; i < 9; ++i)
        p[i] = 0.0;
    const int i1 = (axis + 1) % 3, i2 = (axis + 2) % 3;
    p[axis * 3 + axis] = 1.0;
    p[i1 * 3 + i1] = c;
    p[i2 * 3 + i2] = c;
    // element (i1, i2) = s, (i2, i1) = -s for x and z; y follows the same cyclic rule
    p[i2 * 3 + i1] = s;
    p[i1 * 3 + i2] = -s;
}

__device__ void mat3(const double* A, const double* B, double* C)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k)
                s = fma(A[k * 3 + i], B[j * 3 + k], s);
            C[j * 3 + i] = s;
        }
}

// Partial nutation sum over terms i = lane, lane + 64, ... (largest index first, like the
// reference's smallest-first order, ref_sys_trans.cu:241-254), then a wave reduction.
__device__ double series(const double* terms, int n, const double* F, int lane)
{
    constexpr double scale = 1.0 / (kArcsPerRad * 1e6);
    double acc = 0.0;
    for (int i = n - 1 - lane; i >= 0; i -= 64) {
        const double* q = terms + 16 * i;
        double arg = 0.0;
        for (int k = 13; k >= 0; --k)
            arg = fma(F[k], q[2 + k], arg);
        double s, c;
        sincos(arg, &s, &c);
        acc = fma(q[0] * scale, s, acc);
        acc = fma(q[1] * scale, c, acc);
    }
    for (int off = 32; off >= 1; off >>= 1)
        acc += __shfl_xor(acc, off, 64);
    return acc;
}

// Sun and Moon positions (m) in the mean equator and equinox of J2000 (taken as the GCRS, whose
// axes differ by the 23 mas frame bias) at T Julian centuries of TT from J2000: the low-precision
// series of Montenbruck and Gill, Satellite Orbits (2000), eq. 3.43-3.49, referred to the J2000
// ecliptic (about 0.1 deg and 0.1% in distance; the third-body accelerations they feed are
// 1e-7 of gravity).
__device__ void sun_moon(double T, double* out6)
{
    constexpr double D = kPi / 180.0, AS = D / 3600.0;
    constexpr double kEps = 23.43929111 * D; // obliquity of the J2000 ecliptic
    const double ce = cos(kEps), se = sin(kEps);
    auto put = [&](double lam, double beta, double r, double* o) {
        const double cb = cos(beta);
        const double x = r * cb * cos(lam), y = r * cb * sin(lam), z = r * sin(beta);
        o[0] = x;
        o[1] = ce * y - se * z;
        o[2] = se * y + ce * z;
    };
    // Sun
    const double Ms = D * (357.5256 + 35999.049 * T);
    const double lam_s = D * 282.94 + Ms + AS * (6892.0 * sin(Ms) + 72.0 * sin(2.0 * Ms));
    const double r_s = (149.619 - 2.499 * cos(Ms) - 0.021 * cos(2.0 * Ms)) * 1e9;
    put(lam_s, 0.0, r_s, out6);
    // Moon
    const double L0 = D * (218.31617 + 481267.88088 * T - 1.3972 * T);
    const double l = D * (134.96292 + 477198.86753 * T);
    const double lp = D * (357.52543 + 35999.04944 * T);
    const double F = D * (93.27283 + 483202.01873 * T);
    const double Dm = D * (297.85027 + 445267.11135 * T);
    const double lam_m = L0
        + AS * (22640.0 * sin(l) + 769.0 * sin(2.0 * l) - 4586.0 * sin(l - 2.0 * Dm)
            + 2370.0 * sin(2.0 * Dm) - 668.0 * sin(lp) - 412.0 * sin(2.0 * F)
            - 212.0 * sin(2.0 * l - 2.0 * Dm) - 206.0 * sin(l + lp - 2.0 * Dm)
            + 192.0 * sin(l + 2.0 * Dm) - 165.0 * sin(lp - 2.0 * Dm) + 148.0 * sin(l - lp)
            - 125.0 * sin(Dm) - 110.0 * sin(l + lp) - 55.0 * sin(2.0 * F - 2.0 * Dm));
    const double beta_m = AS
        * (18520.0 * sin(F + lam_m - L0 + AS * (412.0 * sin(2.0 * F) + 541.0 * sin(lp)))
            - 526.0 * sin(F - 2.0 * Dm) + 44.0 * sin(l + F - 2.0 * Dm)
            - 31.0 * sin(-l + F - 2.0 * Dm) - 25.0 * sin(-2.0 * l + F)
            - 23.0 * sin(lp + F - 2.0 * Dm) + 21.0 * sin(-l + F) + 11.0 * sin(-lp + F - 2.0 * Dm));
    const double r_m = 1e3
        * (385000.0 - 20905.0 * cos(l) - 3699.0 * cos(2.0 * Dm - l) - 2956.0 * cos(2.0 * Dm)
            - 570.0 * cos(2.0 * l) + 246.0 * cos(2.0 * l - 2.0 * Dm) - 205.0 * cos(lp - 2.0 * Dm)
            - 171.0 * cos(l + 2.0 * Dm) - 152.0 * cos(l + lp - 2.0 * Dm));
    put(lam_m, beta_m, r_m, out6 + 3);
}

} // namespace frames

__global__ void __launch_bounds__(64) time_nodes_kernel(FrameDev fd, int y, int mo, int d, int hh,
    int mi, double sec_shift, const double* offsets, int64_t len, double* M, double* mjd_out,
    double* tb)
{
    using namespace frames;
    const int64_t node = blockIdx.x;
    const int lane = threadIdx.x;
    if (node >= len)
        return;
    const double mjd = calendar_to_mjd(y, mo, d, hh, mi, sec_shift + offsets[node]);
    const Eop e = eop_at(fd, mjd);
    const double mjd_ut1 = mjd + e.ut1_utc / 86400.0;
    const double mjd_tt = mjd + (leap_at(fd, mjd) + 32.184) / 86400.0;
    const double t = (mjd_tt - kMjdJ2000) / 36525.0;
    double F[14];
    fund_args(t, F);
    const double lon0 = series(fd.nut[0], fd.n_nut[0], F, lane);
    const double lon1 = series(fd.nut[1], fd.n_nut[1], F, lane);
    const double obl0 = series(fd.nut[2], fd.n_nut[2], F, lane);
    const double obl1 = series(fd.nut[3], fd.n_nut[3], F, lane);
    if (lane != 0)
        return;
    const double dpsi = lon0 + lon1 * t;
    const double deps = obl0 + obl1 * t;
    double pr[4];
    precession(t, pr);
    double Q1[9], Q2[9], Q3[9], Q4[9], W1[9], W2[9], W3[9], R[9], A[9], B[9], Q[9], W[9];
    rot(2, -pr[1], Q1);
    rot(0, -pr[2], Q2);
    rot(2, pr[3] + dpsi, Q3);
    rot(0, pr[0] + deps, Q4);
    const double sp = -47e-6 * t / kArcsPerRad;
    rot(2, -sp, W1);
    rot(1, e.xp, W2);
    rot(0, e.yp, W3);
    rot(2, -gast(mjd_ut1, mjd_tt, dpsi, pr[0], F), R);
    mat3(Q1, Q2, A);
    mat3(Q3, Q4, B);
    mat3(A, B, Q);
    mat3(W1, W2, A);
    mat3(A, W3, W);
    mat3(Q, R, A);
    mat3(A, W, B);
    for (int i = 0; i < 9; ++i)
        M[node * 9 + i] = B[i];
    mjd_out[node] = mjd;
    if (tb) {
        double o[6];
        sun_moon(t, o);
        for (int i = 0; i < 6; ++i)
            tb[node * 6 + i] = o[i];
    }
}

void launch_time_nodes(hipStream_t st, const FrameDev& fd, const int* ymdhm, double sec_shift,
    const double* d_offsets, int64_t len, double* d_M, double* d_mjd, double* d_tb)
{
    if (len <= 0)
        return;
    hipLaunchKernelGGL(time_nodes_kernel, dim3((unsigned)len), dim3(64), 0, st, fd, ymdhm[0],
        ymdhm[1], ymdhm[2], ymdhm[3], ymdhm[4], sec_shift, d_offsets, len, d_M, d_mjd, d_tb);
}

// =========================================================================================
// Standalone gravity evaluation (GravitySource::ComputeAcceleration)
// =========================================================================================
template <bool SPLIT>
__global__ void __launch_bounds__(256) gravity_aos_kernel(GravityDev g, const double* M, const double* r, int64_t n,
    double* a, int64_t out_stride, SourcesDev src, const double* tb, const double* v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double p[3] = { r[3 * i], r[3 * i + 1], r[3 * i + 2] };
    double o[3];
    if (M) {
        gravity_gcrs<SPLIT>(M, p, g, o);
        const double w[3] = { v ? v[3 * i] : 0.0, v ? v[3 * i + 1] : 0.0, v ? v[3 * i + 2] : 0.0 };
        SourcesDev so = src;
        if (src.bc_obj)
            so.drag_bc = src.bc_obj[i];
        add_sources(so, tb, p, w, o);
    } else {
        gravity_itrs<SPLIT>(p[0], p[1], p[2], g, o[0], o[1], o[2]);
    }
    a[i * out_stride] = o[0];
    a[i * out_stride + 1] = o[1];
    a[i * out_stride + 2] = o[2];
}

void launch_gravity_aos(hipStream_t st, bool split, const GravityDev& g, const double* d_M,
    const double* d_r, int64_t n, double* d_a, int64_t out_stride, int block,
    const SourcesDev& src, const double* d_tb, const double* d_v)
{
    if (n <= 0)
        return;
    const dim3 grid((unsigned)((n + block - 1) / block));
    if (split)
        hipLaunchKernelGGL(gravity_aos_kernel<true>, grid, dim3(block), 0, st, g, d_M, d_r, n,
            d_a, out_stride, src, d_tb, d_v);
    else
        hipLaunchKernelGGL(gravity_aos_kernel<false>, grid, dim3(block), 0, st, g, d_M, d_r, n,
            d_a, out_stride, src, d_tb, d_v);
}

// =========================================================================================
// Known-answer hooks: the predictor and corrector arithmetic of gj8_chunk_kernel on given
// inputs (AoS [n][27] histories, [n][3] vectors).
// =========================================================================================
// SHIFT: Predictor::Predict's MoveAccelerationsKernel too (prediction.cu:63-65,
// dev_functions.cu:20-32): columns 0-7 of the history take columns 1-8 after the sums.
template <bool SHIFT>
__global__ void kat_predict_kernel(double* acc27, const double* s3, double* S3, int64_t n,
    double h, double* r3, double* v3)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    for (int c = 0; c < 3; ++c) {
        double Pv = 0.0, Pr = 0.0;
        for (int q = 0; q < 9; ++q) {
            const double a = acc27[27 * i + 3 * q + c];
            Pv = fma(c_sa[9][q], a, Pv);
            Pr = fma(c_gj[9][q], a, Pr);
        }
        const double alast = acc27[27 * i + 24 + c];
        const double s = s3[3 * i + c];
        const double S = S3[3 * i + c] + (s + alast / 2);
        S3[3 * i + c] = S;
        v3[3 * i + c] = ((Pv + s) + alast / 2) * h;
        r3[3 * i + c] = ((Pr + S) * h) * h;
    }
    if (SHIFT)
        for (int q = 0; q < 8; ++q)
            for (int c = 0; c < 3; ++c)
                acc27[27 * i + 3 * q + c] = acc27[27 * i + 3 * (q + 1) + c];
}

__global__ void kat_finalize_kernel(double* r3, double* v3, const double* acc27, const double* s3,
    const double* S3, int64_t n, double h)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    for (int c = 0; c < 3; ++c) {
        const double a = acc27[27 * i + 24 + c];
        v3[3 * i + c] = h * ((v3[3 * i + c] + a * c_sa[8][8]) + s3[3 * i + c]);
        r3[3 * i + c] = (h * h) * ((r3[3 * i + c] + a * c_gj[8][8]) + S3[3 * i + c]);
    }
}

void launch_kat_predict(hipStream_t st, const double* acc27, const double* s3, double* S3,
    int64_t n, double h, double* r3, double* v3)
{
    hipLaunchKernelGGL(kat_predict_kernel<false>, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st,
        const_cast<double*>(acc27), s3, S3, n, h, r3, v3);
}

void launch_predict_aos(hipStream_t st, double* acc27, const double* s3, double* S3, int64_t n,
    double h, double* r3, double* v3)
{
    if (n > 0)
        hipLaunchKernelGGL(kat_predict_kernel<true>, dim3((unsigned)((n + 63) / 64)), dim3(64), 0,
            st, acc27, s3, S3, n, h, r3, v3);
}

void launch_kat_finalize(hipStream_t st, double* r3, double* v3, const double* acc27,
    const double* s3, const double* S3, int64_t n, double h)
{
    hipLaunchKernelGGL(kat_finalize_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, r3,
        v3, acc27, s3, S3, n, h);
}

__global__ void aos_to_soa_kernel(const double* aos, double* soa, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    soa[i] = aos[3 * i];
    soa[n + i] = aos[3 * i + 1];
    soa[2 * n + i] = aos[3 * i + 2];
}

__global__ void soa_to_aos_kernel(const double* soa, double* aos, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    aos[3 * i] = soa[i];
    aos[3 * i + 1] = soa[n + i];
    aos[3 * i + 2] = soa[2 * n + i];
}

__global__ void gather_aos_kernel(const double* aos, const int64_t* perm, double* soa, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int64_t i = perm[p];
    soa[p] = aos[3 * i];
    soa[n + p] = aos[3 * i + 1];
    soa[2 * n + p] = aos[3 * i + 2];
}

__global__ void scatter_aos_kernel(const double* soa, const int64_t* perm, double* aos, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int64_t i = perm[p];
    aos[3 * i] = soa[p];
    aos[3 * i + 1] = soa[n + p];
    aos[3 * i + 2] = soa[2 * n + p];
}

__global__ void gather_soa_kernel(const double* src, const int64_t* perm, double* dst, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int64_t i = perm[p];
    for (int c = 0; c < 3; ++c)
        dst[c * n + p] = src[c * n + i];
}

__global__ void scatter_soa_kernel(const double* src, const int64_t* perm, double* dst, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int64_t i = perm[p];
    for (int c = 0; c < 3; ++c)
        dst[c * n + i] = src[c * n + p];
}

void launch_permute(hipStream_t st, int mode, const double* src, const int64_t* perm, double* dst,
    int64_t n)
{
    if (n <= 0)
        return;
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
    if (mode == 0)
        hipLaunchKernelGGL(gather_aos_kernel, g, b, 0, st, src, perm, dst, n);
    else if (mode == 1)
        hipLaunchKernelGGL(scatter_aos_kernel, g, b, 0, st, src, perm, dst, n);
    else if (mode == 2)
        hipLaunchKernelGGL(gather_soa_kernel, g, b, 0, st, src, perm, dst, n);
    else
        hipLaunchKernelGGL(scatter_soa_kernel, g, b, 0, st, src, perm, dst, n);
}

void launch_aos_to_soa(hipStream_t st, const double* aos, double* soa, int64_t n)
{
    if (n > 0)
        hipLaunchKernelGGL(aos_to_soa_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
            aos, soa, n);
}

__global__ void slots_to_aos_kernel(const double* src, double* dst, int64_t n, int nslots)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    for (int q = 0; q < nslots; ++q)
        for (int c = 0; c < 3; ++c)
            dst[(i * nslots + q) * 3 + c] = src[(q * 3 + c) * n + i];
}

void launch_slots_to_aos(hipStream_t st, const double* src, double* dst, int64_t n, int nslots)
{
    if (n > 0)
        hipLaunchKernelGGL(slots_to_aos_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
            st, src, dst, n, nslots);
}

__global__ void finite_soa_kernel(const double* r, const double* v, int64_t n, int32_t* flags,
    uint32_t* count, int32_t code, int fresh)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    bool ok = true;
    for (int c = 0; c < 3; ++c)
        ok = ok && isfinite(r[c * n + i]) && isfinite(v[c * n + i]);
    const int32_t prev = fresh ? 0 : flags[i];
    if (!ok && prev == 0)
        atomicAdd(count, 1u);
    flags[i] = !ok && prev == 0 ? code : prev;
}

void launch_finite_soa(hipStream_t st, const double* r, const double* v, int64_t n, int32_t* flags,
    uint32_t* count, int32_t code, bool fresh)
{
    if (n > 0)
        hipLaunchKernelGGL(finite_soa_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
            r, v, n, flags, count, code, fresh ? 1 : 0);
}

// Semi-major axis a = 1 / (2 / |r| - |v|^2 / GM) of SoA states (vis-viva), 0 where it is not
// finite: the grouping key of the propagation (orbit_size in capi.cpp).
__global__ void orbit_key_soa_kernel(const double* r, const double* v, int64_t n, double gm,
    float* key)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double r2 = 0.0, v2 = 0.0;
    for (int c = 0; c < 3; ++c) {
        r2 += r[c * n + i] * r[c * n + i];
        v2 += v[c * n + i] * v[c * n + i];
    }
    const double a = 1.0 / (2.0 / sqrt(r2) - v2 / gm);
    key[i] = isfinite(a) ? (float)a : 0.0f;
}

void launch_orbit_key_soa(hipStream_t st, const double* r, const double* v, int64_t n, double gm,
    float* key)
{
    if (n > 0)
        hipLaunchKernelGGL(orbit_key_soa_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
            st, r, v, n, gm, key);
}

void launch_soa_to_aos(hipStream_t st, const double* soa, double* aos, int64_t n)
{
    if (n > 0)
        hipLaunchKernelGGL(soa_to_aos_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
            soa, aos, n);
}

} // namespace ogp
