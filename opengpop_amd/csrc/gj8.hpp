// Gauss-Jackson 8th-order summed-form step arithmetic (Predictor::Predict, prediction.cu:49-74,
// dev_functions.cu:20-50; the corrector of gpu_propagator.cu:86-119, dev_functions.cu:36-69),
// shared by the propagation kernels and the adjoint's recording kernel so that both perform the
// same operations in the same order. The acceleration history is a ring of 9 slots
// (slot = step mod 9) in HBM, SoA: hist[(slot * 3 + c) * N + i].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "coefficients.hpp"

namespace ogp {

__device__ __forceinline__ int ring_slot(int64_t j) { return (int)(((j % 9) + 9) % 9); }

// Sums over the 9 stored accelerations a_{k-9} .. a_{k-1} needed by step k.
struct GJ8Sums {
    double Pv[3], Pr[3]; // predictor rows (summed-Adams and GJ row 9)
    double T1[3], T2[3]; // corrector rows 8 without the new acceleration
    double alast[3];     // a_{k-1}
};

// get(j, c): component c of a_{k-9+j}.
template <class Get> __device__ __forceinline__ void gj8_sums_of(Get&& get, GJ8Sums& q)
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
        q.Pv[c] = q.Pr[c] = q.T1[c] = q.T2[c] = 0.0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double a = get(j, c);
            q.Pv[c] = fma(c_sa[9][j], a, q.Pv[c]);
            q.Pr[c] = fma(c_gj[9][j], a, q.Pr[c]);
            if (j >= 1) {
                q.T1[c] = fma(c_sa[8][j - 1], a, q.T1[c]);
                q.T2[c] = fma(c_gj[8][j - 1], a, q.T2[c]);
            }
            if (j == 8)
                q.alast[c] = a;
        }
    }
}

// The ring in memory: a_j at slot ring_slot(j), hist[(slot * 3 + c) * N + i]. The slots of
// a_{k-9} .. a_{k-1} follow ring_slot(k - 9) cyclically (one 64-bit remainder per step).
__device__ __forceinline__ void gj8_sums(const double* hist, int64_t N, int64_t i, int64_t k,
    GJ8Sums& q)
{
    const int s0 = ring_slot(k - 9);
    gj8_sums_of([&](int j, int c) {
        const int s = s0 + j >= 9 ? s0 + j - 9 : s0 + j;
        return hist[(s * 3 + c) * N + i];
    }, q);
}

// Predictor: S_{k} = S_{k-1} + s_{k-1} + a_{k-1}/2, predicted v and r; sprev keeps s_{k-1}.
__device__ __forceinline__ void gj8_predict(const GJ8Sums& q, double h, const double* s,
    double* S, double* r, double* v, double* sprev)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        S[c] += s[c] + q.alast[c] / 2;
        v[c] = ((q.Pv[c] + s[c]) + q.alast[c] / 2) * h;
        r[c] = ((q.Pr[c] + S[c]) * h) * h;
        sprev[c] = s[c];
    }
}

// One corrector pass with the new acceleration anew.
__device__ __forceinline__ void gj8_correct(const GJ8Sums& q, const double* anew,
    const double* sprev, const double* S, double h, double hh, double* s, double* r, double* v)
{
    const double sa88 = c_sa[8][8], gj88 = c_gj[8][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s[c] = sprev[c] + (anew[c] + q.alast[c]) / 2;
        v[c] = h * ((q.T1[c] + anew[c] * sa88) + s[c]);
        r[c] = hh * ((q.T2[c] + anew[c] * gj88) + S[c]);
    }
}

} // namespace ogp
