// Drives tests/integration/ogp_adapter.hpp (the reference-side adapter over include/ogp.h) the
// way the reference's runner and unit tests call its classes. Exit status 0 and "adapter ok"
// when every check holds. Usage: adapter_test DATA_DIR
#include "ogp_adapter.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

using namespace ogp_adapter;

static int failures = 0;
#define EXPECT(cond)                                                                               \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);                 \
            ++failures;                                                                            \
        }                                                                                          \
    } while (0)

static void circular(int n, Vec3s& r, Vec3s& v)
{
    const double gm = 3.986004415e14;
    r.resize(n);
    v.resize(n);
    for (int q = 0; q < n; ++q) {
        const double a = 6.9e6 + 1.5e4 * q, th = 0.37 * q, inc = 0.1 + 0.05 * q;
        const double s = std::sqrt(gm / a);
        r[q] = { a * std::cos(th), a * std::sin(th) * std::cos(inc), a * std::sin(th) * std::sin(inc) };
        v[q] = { -s * std::sin(th), s * std::cos(th) * std::cos(inc), s * std::cos(th) * std::sin(inc) };
    }
}

static double* device_copy(const Vec3s& x)
{
    double* d = nullptr;
    if (hipMalloc((void**)&d, sizeof(double) * 3 * x.size()) != hipSuccess)
        throw std::runtime_error("hipMalloc failed");
    if (hipMemcpy(d, &x[0][0], sizeof(double) * 3 * x.size(), hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("hipMemcpy failed");
    return d;
}

int main(int argc, char** argv)
{
    const char* data = argc > 1 ? argv[1] : "data";
    try {
        OgpContext c(TimeData { 2025, 1, 1, 0, 0, 0.0 }, 20, data);
        const int n = 12, steps = 20;
        const double h = 60.0;
        Vec3s r0, v0, rb, vb;
        circular(n, r0, v0);

        // runner.cu: the whole span in one call
        propagate(c, r0, v0, h, steps, rb, vb);

        // GpuPropagator on the caller's device buffers, one Step at a time
        double* dR = device_copy(r0);
        double* dV = device_copy(v0);
        double* dAcc = nullptr;
        (void)hipMalloc((void**)&dAcc, sizeof(double) * 3 * n);
        {
            GpuPropagator p(c, dR, dV, n, h, steps);
            Vec3s hr(n), hv(n);
            for (int k = 0; k < steps; ++k)
                p.Step(&hr[0][0], &hv[0][0], dAcc);
            EXPECT(std::memcmp(&hr[0][0], &rb[0][0], sizeof(double) * 3 * n) == 0);
            EXPECT(std::memcmp(&hv[0][0], &vb[0][0], sizeof(double) * 3 * n) == 0);
            Vec3s cr(n), cv(n);
            p.CopyStateToHost(&cr[0][0], &cv[0][0]);
            EXPECT(std::memcmp(&cr[0][0], &hr[0][0], sizeof(double) * 3 * n) == 0);
            Vec3s dr(n);
            (void)hipMemcpy(&dr[0][0], dR, sizeof(double) * 3 * n, hipMemcpyDeviceToHost);
            EXPECT(std::memcmp(&dr[0][0], &hr[0][0], sizeof(double) * 3 * n) == 0);
        }

        // IForceModel: generation stage 0 of step 1 and propagation node 4 are the epoch
        double* dR0 = device_copy(r0);
        double* dV0 = device_copy(v0);
        double *dK = nullptr, *dA = nullptr;
        (void)hipMalloc((void**)&dK, sizeof(double) * 30 * n);
        (void)hipMalloc((void**)&dA, sizeof(double) * 27 * n);
        ForceModel fm(c, n, h);
        fm.ComputeGenerationAcceleration(dR0, dV0, dK, 1, 0);
        fm.ComputePropagationAcceleration(dR0, dV0, dA, 4, 8);
        std::vector<double> K(30 * n), A(27 * n);
        (void)hipMemcpy(K.data(), dK, sizeof(double) * 30 * n, hipMemcpyDeviceToHost);
        (void)hipMemcpy(A.data(), dA, sizeof(double) * 27 * n, hipMemcpyDeviceToHost);
        for (int q = 0; q < n; ++q)
            for (int x = 0; x < 3; ++x) {
                EXPECT(K[30 * q + x] == A[27 * q + 24 + x]);
                EXPECT(std::isfinite(K[30 * q + x]) && std::fabs(K[30 * q + x]) < 10.0);
            }

        // GpuStartup + Predictor::Predict(5): the history's column 8 after the prediction is the
        // acceleration of the predicted state at node 9
        GpuStartup su(c, dR0, dV0, n, h);
        FullStartupData sd = su.Startup();
        double* dPr = nullptr;
        double* dPv = nullptr;
        (void)hipMalloc((void**)&dPr, sizeof(double) * 3 * n);
        (void)hipMalloc((void**)&dPv, sizeof(double) * 3 * n);
        Predictor pr(c, n, h);
        // the predictor's s and S are the startup's column 8 (copy_s.cu)
        double *ds = nullptr, *dS = nullptr;
        (void)hipMalloc((void**)&ds, sizeof(double) * 3 * n);
        (void)hipMalloc((void**)&dS, sizeof(double) * 3 * n);
        std::vector<double> s27(27 * n), S27(27 * n), s3(3 * n), S3(3 * n);
        (void)hipMemcpy(s27.data(), sd.devSmallS, sizeof(double) * 27 * n, hipMemcpyDeviceToHost);
        (void)hipMemcpy(S27.data(), sd.devBigS, sizeof(double) * 27 * n, hipMemcpyDeviceToHost);
        for (int q = 0; q < n; ++q)
            for (int x = 0; x < 3; ++x) {
                s3[3 * q + x] = s27[27 * q + 24 + x];
                S3[3 * q + x] = S27[27 * q + 24 + x];
            }
        (void)hipMemcpy(ds, s3.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice);
        (void)hipMemcpy(dS, S3.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice);
        pr.SetPointers(dPr, dPv, sd.devAccelerationMatrix, ds, dS);
        pr.Predict(5);
        std::vector<double> acc(27 * n), pred(3 * n);
        (void)hipMemcpy(acc.data(), sd.devAccelerationMatrix, sizeof(double) * 27 * n,
            hipMemcpyDeviceToHost);
        (void)hipMemcpy(pred.data(), dPr, sizeof(double) * 3 * n, hipMemcpyDeviceToHost);
        fm.ComputePropagationAcceleration(dPr, dPv, dA, 9, 0);
        (void)hipMemcpy(A.data(), dA, sizeof(double) * 27 * n, hipMemcpyDeviceToHost);
        for (int q = 0; q < n; ++q)
            for (int x = 0; x < 3; ++x) {
                EXPECT(acc[27 * q + 24 + x] == A[27 * q + x]);
                EXPECT(std::isfinite(pred[3 * q + x]));
            }
        GpuStartup::Free(sd);

        // ClosestApproach: two sets, every cross-set pair, Get in both orders
        std::vector<unsigned> sets(n, 0);
        for (int q = n / 2; q < n; ++q)
            sets[q] = 1;
        ClosestApproach ca(c, h);
        ca.Initialize(r0, v0, sets);
        ca.Propagate(30);
        const int64_t cnt = ca.FindClosest();
        EXPECT(cnt == (n / 2) * (n - n / 2));
        const Approach ab = ca.Get(0, n - 1), ba = ca.Get(n - 1, 0);
        EXPECT(ab.distance > 0 && ab.distance == ba.distance);
        EXPECT(ab.positionA[0] == ba.positionB[0]);
        EXPECT(ca.Get(0, 1).distance == -1.0);

        // an error comes back as an exception with the library's message, not an exit
        bool threw = false;
        try {
            fm.ComputeGenerationAcceleration(dR0, dV0, dK, 0, 0);
        } catch (const std::runtime_error& e) {
            threw = std::strstr(e.what(), "stepNum") != nullptr;
        }
        EXPECT(threw);
        double* bufs[] = { dR, dV, dAcc, dR0, dV0, dK, dA, dPr, dPv, ds, dS };
        for (double* b : bufs)
            (void)hipFree(b);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("adapter ok\n");
    return 0;
}
