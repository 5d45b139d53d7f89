// The adapter a maintainer of the reference's C++ code base would add to call libogp through
// include/ogp.h instead of the CUDA libraries: the reference's class names and call shapes
// (force_model.cuh:28-87, gpu_startup.cuh:24-120, prediction.cuh:23-46,
// gpu_propagator.cuh:26-59, closest_approach.cuh:36-88, types.cuh:113-169) on the C ABI.
// Header-only C++17, links -logp; tests/integration/adapter_test.cpp compiles and drives it.
#pragma once
#include "ogp.h"

#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace ogp_adapter {

// Parameters.hpp:19-26
struct TimeData {
    int Year, Month, Day, Hour, Min;
    double Sec;
};

// ForceModel::PreInitialize + the dataset loaders (force_model.cu:23-45)
struct OgpContext {
    ogp_context* ctx = nullptr;
    explicit OgpContext(const TimeData& start, int degree = 70, const char* data_dir = "data")
    {
        ogp_config cfg;
        ogp_default_config(&cfg); // degree 70, FP64, 7 corrections, the reference's quirks
        cfg.degree = degree;
        cfg.data_dir = data_dir;
        cfg.year = start.Year;
        cfg.month = start.Month;
        cfg.day = start.Day;
        cfg.hour = start.Hour;
        cfg.minute = start.Min;
        cfg.second = start.Sec;
        if (ogp_create(&cfg, &ctx) != OGP_OK)
            throw std::runtime_error(ogp_last_error(nullptr));
    }
    ~OgpContext() { ogp_destroy(ctx); }
    OgpContext(const OgpContext&) = delete;
    OgpContext& operator=(const OgpContext&) = delete;
    void check(int rc) const
    {
        if (rc != OGP_OK)
            throw std::runtime_error(ogp_last_error(ctx));
    }
};

using Vec3s = std::vector<std::array<double, 3>>;

// runner.cu:11-70: GpuStartup + GpuPropagator::Step x totalSteps for all objects at once
inline void propagate(OgpContext& c, const Vec3s& r0, const Vec3s& v0, double h, int64_t steps,
    Vec3s& r, Vec3s& v)
{
    const int64_t n = (int64_t)r0.size();
    std::vector<double> hs(n, h);
    std::vector<int64_t> ns(n, steps);
    r.resize(n);
    v.resize(n);
    c.check(ogp_propagate(c.ctx, &r0[0][0], &v0[0][0], n, hs.data(), ns.data(), &r[0][0],
        &v[0][0]));
}

// IForceModel (iforce_model.cuh:19-33) on the caller's device arrays. PreInitialize and the
// Initialize* calls have nothing to do: the context builds and caches the time nodes itself.
struct ForceModel {
    OgpContext& c;
    unsigned n;
    double h;
    ForceModel(OgpContext& c, unsigned n, double h)
        : c(c), n(n), h(h)
    {
    }
    void ComputeGenerationAcceleration(double* dPos, double* dVel, double* dK, int stepNum,
        int kIndex)
    {
        c.check(ogp_accel_generation(c.ctx, dPos, dVel, n, h, stepNum, kIndex, dK));
    }
    void ComputePropagationAcceleration(double* dPos, double* dVel, double* dAcc, int stepNum,
        int returnColumnIndex)
    {
        c.check(ogp_accel_device(c.ctx, dPos, dVel, n, h, stepNum, dAcc + 3 * returnColumnIndex,
            27));
    }
};

// GpuStartup::Startup -> FullStartupData (gpu_startup.cuh:95-120): [n][9][3] device matrices,
// allocated here as the reference's startup allocates them
struct FullStartupData {
    double *devPositionMatrix = nullptr, *devVelocityMatrix = nullptr, *devSmallS = nullptr,
           *devBigS = nullptr, *devAccelerationMatrix = nullptr;
};
struct GpuStartup {
    OgpContext& c;
    double *dR0, *dV0;
    unsigned n;
    double h;
    GpuStartup(OgpContext& c, double* dR0, double* dV0, unsigned n, double h)
        : c(c), dR0(dR0), dV0(dV0), n(n), h(h)
    {
    }
    FullStartupData Startup()
    {
        FullStartupData d;
        double** all[5] = { &d.devPositionMatrix, &d.devVelocityMatrix, &d.devSmallS, &d.devBigS,
            &d.devAccelerationMatrix };
        for (double** p : all)
            if (hipMalloc((void**)p, sizeof(double) * 27 * n) != hipSuccess)
                throw std::runtime_error("hipMalloc failed");
        c.check(ogp_startup_device(c.ctx, dR0, dV0, n, h, d.devPositionMatrix,
            d.devVelocityMatrix, d.devSmallS, d.devBigS, d.devAccelerationMatrix));
        return d;
    }
    static void Free(FullStartupData& d)
    {
        double* all[5] = { d.devPositionMatrix, d.devVelocityMatrix, d.devSmallS, d.devBigS,
            d.devAccelerationMatrix };
        for (double* p : all)
            (void)hipFree(p);
        d = FullStartupData {};
    }
};

// Predictor::{SetPointers, Predict} (prediction.cuh:23-46)
struct Predictor {
    OgpContext& c;
    int n;
    double h;
    double *r = nullptr, *v = nullptr, *acc = nullptr, *s = nullptr, *S = nullptr;
    Predictor(OgpContext& c, int n, double h)
        : c(c), n(n), h(h)
    {
    }
    void SetPointers(double* dPos, double* dVel, double* dAccMatrix, double* dSmallS,
        double* dBigS)
    {
        r = dPos;
        v = dVel;
        acc = dAccMatrix;
        s = dSmallS;
        S = dBigS;
    }
    void Predict(int stepNum) { c.check(ogp_predict_device(c.ctx, r, v, acc, s, S, n, h, stepNum)); }
};

// GpuPropagator (gpu_propagator.cuh:26-59): advances the caller's device buffers in place
struct GpuPropagator {
    OgpContext& c;
    ogp_propagator* p = nullptr;
    GpuPropagator(OgpContext& c, double* dR, double* dV, unsigned n, double h, int maxSteps)
        : c(c)
    {
        c.check(ogp_propagator_create(c.ctx, dR, dV, n, h, maxSteps, &p));
    }
    void Step(double* hR = nullptr, double* hV = nullptr, double* dAcc = nullptr)
    {
        c.check(ogp_propagator_step(p, hR, hV, dAcc));
    }
    void CopyStateToHost(double* hR, double* hV) const
    {
        c.check(ogp_propagator_copy_state_to_host(p, hR, hV));
    }
    void FreeAllMemory()
    {
        ogp_propagator_destroy(p);
        p = nullptr;
    }
    ~GpuPropagator() { FreeAllMemory(); }
};

// ClosestApproach::{Initialize, Propagate, FindClosest} + ClosestApproachData::Get
struct Approach {
    double distance, time;
    double positionA[3], positionB[3];
};
struct ClosestApproach {
    OgpContext& c;
    double h;
    std::vector<uint32_t> a, b;
    std::vector<int32_t> k;
    std::vector<double> d, t, pa, pb;
    ClosestApproach(OgpContext& c, double h)
        : c(c), h(h)
    {
    }
    void Initialize(const Vec3s& r, const Vec3s& v, const std::vector<unsigned>& sets)
    {
        c.check(ogp_conj_init(c.ctx, &r[0][0], &v[0][0], sets.data(), (int64_t)r.size(), h));
    }
    void Propagate(int64_t steps) { c.check(ogp_conj_propagate(c.ctx, steps)); }
    int64_t FindClosest(double threshold_m = 0.0) // 0: every cross-set pair
    {
        int64_t cnt = 0;
        c.check(ogp_conj_find(c.ctx, threshold_m, &cnt));
        a.resize(cnt);
        b.resize(cnt);
        k.resize(cnt);
        d.resize(cnt);
        t.resize(cnt);
        pa.resize(3 * cnt);
        pb.resize(3 * cnt);
        c.check(ogp_conj_results(c.ctx, a.data(), b.data(), k.data(), d.data(), t.data(),
            pa.data(), pb.data()));
        return cnt; // records sorted by (a, b)
    }
    // types.cuh:153-169: (i, j) in either order; -1 records for pairs that were not screened
    Approach Get(unsigned i, unsigned j) const
    {
        Approach r {};
        if (i == j)
            return r;
        const bool swap = i > j;
        const unsigned lo = swap ? j : i, hi = swap ? i : j;
        for (size_t q = 0; q < a.size(); ++q)
            if (a[q] == lo && b[q] == hi) {
                r.distance = d[q];
                r.time = t[q];
                for (int x = 0; x < 3; ++x) {
                    r.positionA[x] = swap ? pb[3 * q + x] : pa[3 * q + x];
                    r.positionB[x] = swap ? pa[3 * q + x] : pb[3 * q + x];
                }
                return r;
            }
        r.distance = r.time = -1.0;
        for (int x = 0; x < 3; ++x)
            r.positionA[x] = r.positionB[x] = -1.0;
        return r;
    }
};

} // namespace ogp_adapter
