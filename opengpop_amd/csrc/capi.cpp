// C ABI of the engine (include/ogp.h): context, datasets on the device, propagation and
// conjunction orchestration. Host code; all device work is issued on the context's stream.
#include "../../include/ogp.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "conjunction.hpp"
#include "datasets.hpp"
#include "kernels.hpp"

using namespace ogp;

static_assert(kNoAllocMaxCandidates == OGP_NOALLOC_MAX_CANDIDATES,
    "the header's limit is the scratch array's");

namespace {

thread_local std::string g_error;

struct ApiError : std::runtime_error {
    int code;
    ApiError(int c, const std::string& m)
        : std::runtime_error(m)
        , code(c)
    {
    }
};

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw ApiError(OGP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(x) hip_check((x), #x)

void arg_check(bool ok, const std::string& msg)
{
    if (!ok)
        throw ApiError(OGP_ERR_ARGUMENT, msg);
}

// Device allocation owned by the context.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int64_t* counter = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) {
            (void)hipFree(p);
            if (counter)
                *counter -= (int64_t)bytes;
        }
        p = nullptr;
        bytes = 0;
    }
    // grows (never shrinks) to at least `need` bytes; contents are not preserved
    void reserve(size_t need, int64_t* ctr)
    {
        if (need <= bytes && p)
            return;
        release();
        counter = ctr;
        size_t b = std::max<size_t>(need, 256);
        HIPC(hipMalloc(&p, b));
        bytes = b;
        if (counter)
            *counter += (int64_t)b;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Time-node tables of one step-size class (FMTimePoints, time_points.cu:126-207).
struct NodeTable {
    DevBuf gen_M, prop_M, prop_mjd, gen_mjd, offsets;
    DevBuf gen_tb, prop_tb; // Sun and Moon GCRS positions per node (6 doubles)
    int64_t n_prop = 0;
};

struct Event {
    hipEvent_t e = nullptr;
    Event() { HIPC(hipEventCreate(&e)); }
    ~Event()
    {
        if (e)
            (void)hipEventDestroy(e);
    }
};

} // namespace

struct ogp_propagator;
// Frees the device memory of every propagator of a context that is being destroyed and marks
// them detached (defined with ogp_propagator below).
void detach_propagators(ogp_context* ctx);

struct ogp_context {
    ogp_config cfg {};
    std::string data_dir;
    std::string error;
    hipStream_t stream = nullptr;
    int64_t dev_bytes = 0;

    GravityModel grav;
    FrameData frames;
    DevBuf d_rec, d_recf, d_jext, d_diag, d_parts, d_part_diag, d_eop, d_leap_mjd, d_leap_sec, d_nut[4];
    GravityDev gdev {};
    FrameDev fdev {};
    int kind = kGravity;  // kernel variant of the force
    int rk8_variant = 0;  // 0 = reference row 7
    int waves_override = 0; // waves per 64-object group; 0 = chosen from the object count
    int64_t chunk = 64;     // GJ8 steps per launch

    // node tables keyed by (h, t0)
    std::map<std::pair<double, double>, std::unique_ptr<NodeTable>> nodes;

    // propagation state (internal, class-sorted order)
    DevBuf st_r, st_v, st_s, st_S, st_hist, st_ns, perm, io_a, io_b, io_c;
    DevBuf d_flags, d_count; // per-object flags of the input / result checks
    DevBuf sc[7];
    ogp_timing timing {};

    // stepwise propagators: every live one (ogp_propagator_*), and the context's own
    // (ogp_prop_init / ogp_prop_step), which is one of them
    std::vector<ogp_propagator*> props;
    ogp_propagator* sw = nullptr;

    // conjunction
    int64_t cj_n = 0, cj_steps = 0, cj_cap_steps = 0;
    double cj_h = 0;
    std::vector<uint32_t> cj_sets;
    std::vector<int64_t> cj_set_off;
    DevBuf cj_traj, cj_r, cj_v, cj_cnt, cj_a, cj_b, cj_k, cj_d, cj_out, cj_mind, cj_mink;
    // radially sorted input of threshold screens (sort_for_screen)
    DevBuf cj_sorted, cj_rpart, cj_rlo, cj_rhi, cj_slo, cj_shi, cj_ids;
    int radial_filter = 1; // 0 off, 1 when it pays (radial_filter_pays), 2 always
    double cj_screened = 0; // pairs the last screen evaluated
    std::vector<uint32_t> res_a, res_b;
    std::vector<int32_t> res_k;
    std::vector<double> res_out;
    ogp_conj_timing cj_timing {};

    // derivatives
    DevBuf ad_in, ad_r, ad_v, ad_s, ad_S, ad_hist, ad_ns, ad_Zd, ad_RVd, ad_ck, ad_G, ad_grec,
        ad_lam, ad_seeds, ad_out;
    ogp_adjoint_timing ad_timing {};

    // Screen tile shape per set-rectangle (conjunction.hpp: kScreenSquare, kScreenNarrow),
    // tuned per ceil(log2(set-A size)) for set A of at most 64 objects (one square tile); the
    // override of ogp_set_screen_shape (0 = tuned).
    std::map<int, int> tuned_screen;
    int screen_override = 0;
    static int screen_bucket(int64_t na)
    {
        int b = 0;
        while ((int64_t(1) << b) < na)
            ++b;
        return b;
    }
    // The shape for a rectangle with set A of na objects: a set A of more than 64 objects fills
    // square tiles (a narrow tile would stream set B once per 4 A objects); below that the tuned
    // choice, else narrow up to 8 objects.
    int screen_shape(int64_t na) const
    {
        if (screen_override > 0)
            return screen_override;
        if (na > 64)
            return kScreenSquare;
        auto it = tuned_screen.find(screen_bucket(na));
        if (it != tuned_screen.end())
            return it->second;
        return na <= 8 ? kScreenNarrow : kScreenSquare;
    }

    // Streams of the concurrent GJ8 chunk sequences (object slices), default priority.
    static constexpr int kLaneStreams = 4;
    hipStream_t lane_streams[kLaneStreams] = {};
    int slices_override = 0; // object slices per class; 0 = chosen from the class size
    // GJ8 corrector evaluations end at a bitwise fixed point (ogp_set_fixed_point_exit), and the
    // device count of the object-evaluations the chunk kernels ran in the last call
    bool fp_exit = true;
    DevBuf d_evals;
    // orbit-size keys of the device-resident inputs (plan_groups)
    DevBuf d_key;
    // The last-step window within which plan_groups orders objects by orbit size: 16 steps, or
    // the largest divisor of the chunk length below it (100,000 mixed objects: 619 ms per
    // propagation against 635 ms ordered by step count alone, 622-626 ms with windows of 8, 32
    // and 64 steps; 50,000 LEO over 7 days 7% faster with any window)
    int64_t sort_window() const { return std::gcd<int64_t>(chunk, 16); }
    // HBM for the gravity Jacobians recorded by the derivative passes (ogp_set_derivative_memory)
    double record_bytes = 16.0 * (1 << 30);

    ~ogp_context()
    {
        detach_propagators(this);
        for (hipStream_t s : lane_streams)
            if (s)
                (void)hipStreamDestroy(s);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    void sync() { HIPC(hipStreamSynchronize(stream)); }
    void check_launch() { HIPC(hipGetLastError()); }

    // Launch shape of the GJ8 chunks: W waves per 64-object group and G groups per workgroup
    // in lockstep, encoded as W | (G - 1) << 4. Autotuned per floor(log2(number of groups)).
    // Results do not depend on the shape (the harmonic parts are summed in a fixed order), so
    // the tuner may switch between chunks of one propagation.
    static int make_shape(int W, int G) { return W | ((G - 1) << 4); }
    static int shape_waves(int shape) { return shape & 15; }
    static int shape_groups(int shape) { return (shape >> 4) + 1; }
    std::map<int, int> tuned_shape;
    int lockstep_override = 0; // groups per workgroup; 0 = tuned (1 when waves are fixed)

    static int group_bucket(int64_t n)
    {
        int64_t groups = (n + 63) / 64;
        int b = 0;
        while (groups > 1) {
            groups >>= 1;
            ++b;
        }
        return b;
    }

    int simds = 1024; // SIMDs of the device (4 per CU), read at creation
    void* noalloc = nullptr; // device address of the allocation-free search's scratch array

    // Shape of the chunk launches at n objects: the overrides, else the tuned value for this
    // size, else enough waves for about 1.2 resident per SIMD (which hides most of the latency
    // of the scalar coefficient loads) and one group per workgroup.
    int shape_for(int64_t n) const
    {
        if (kind != kGravity && kind != kGravitySplit)
            return make_shape(1, 1);
        if (waves_override > 0)
            return make_shape(waves_override,
                waves_override >= 4 && lockstep_override == 2 ? 2 : 1);
        auto it = tuned_shape.find(group_bucket(n));
        if (it != tuned_shape.end())
            return it->second;
        const int64_t groups = (n + 63) / 64;
        int w = 1;
        while (w < kMaxWaves && groups * w < simds + simds / 5)
            w *= 2;
        return make_shape(w, lockstep_override == 2 && w >= 4 ? 2 : 1);
    }
    // Waves per 64-object group (every kernel with the column split; only the chunks take
    // lockstep groups).
    int waves_for(int64_t n) const { return shape_waves(shape_for(n)); }

    // Shapes the tuner tries at n objects. One that cannot give every SIMD a wave
    // (groups * W < simds) leaves most of the chip idle and is not tried, unless it is the
    // widest. Lockstep pairs of groups are tried only when asked for (ogp_set_lockstep(2)): on
    // the ragged mixed 100k workload they measured 12% slower (1,343 against 1,185 ms), and as
    // candidates they only added noise to the choice. Since simds and the bucket bounds are
    // powers of two, the set is the same for every size in a tuning bucket.
    std::vector<int> tuning_candidates(int64_t n) const
    {
        const int64_t groups = (n + 63) / 64;
        std::vector<int> c;
        for (int w = 1; w <= kMaxWaves; w *= 2)
            if (groups * w >= simds || w == kMaxWaves)
                c.push_back(make_shape(w, lockstep_override == 2 && w >= 4 ? 2 : 1));
        return c;
    }

    bool needs_tuning(int64_t n) const
    {
        return (kind == kGravity || kind == kGravitySplit) && waves_override == 0 && n > 0
            && tuned_shape.find(group_bucket(n)) == tuned_shape.end();
    }

    // Acceleration sources added to gravity (ogp_set_sources) and the per-object drag
    // coefficients (ogp_set_drag_coefficients, caller order). bc_obj: device copy in state order
    // for the launches of the current call (null = drag_bc for every object); the bulk, stepwise
    // and conjunction states keep their own copies.
    ogp_sources sources {};
    std::vector<double> drag_bc_obj;
    DevBuf d_bc, cj_bc;
    const double* bc_obj = nullptr;
    const double* cj_bc_ptr = nullptr;
    SourcesDev sources_dev() const
    {
        SourcesDev d {};
        if (kind == kGravity || kind == kGravitySplit) {
            d.mask = sources.mask;
            d.drag_bc = sources.drag_bc;
            d.bc_obj = (sources.mask & kSourceDrag) ? bc_obj : nullptr;
        }
        return d;
    }
    // Device copy of the per-object drag coefficients for a call on n objects, in state order
    // (perm[p] = caller index at state position p; null = the caller's order), or null when none
    // are set or drag is off. A different object count is an error while drag is on.
    const double* stage_bc(DevBuf& buf, int64_t n, const int64_t* perm)
    {
        if (drag_bc_obj.empty() || !(sources.mask & kSourceDrag) || n <= 0
            || (kind != kGravity && kind != kGravitySplit))
            return nullptr;
        if ((int64_t)drag_bc_obj.size() != n)
            throw ApiError(OGP_ERR_ARGUMENT, "per-object drag coefficients were set for "
                    + std::to_string(drag_bc_obj.size()) + " objects; this call has "
                    + std::to_string(n) + " (ogp_set_drag_coefficients)");
        std::vector<double> h(n);
        for (int64_t p = 0; p < n; ++p)
            h[p] = drag_bc_obj[perm ? perm[p] : p];
        buf.reserve(sizeof(double) * n, &dev_bytes);
        HIPC(hipMemcpy(buf.p, h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        return buf.as<double>();
    }
    // The derivative paths linearise gravity alone.
    void require_gravity_only(const char* what) const
    {
        if (sources.mask != 0)
            throw ApiError(OGP_ERR_STATE,
                std::string(what) + " covers gravity only; clear the sources (ogp_set_sources)");
    }

    ForceCtx force_ctx(const NodeTable* nt, double h) const
    {
        ForceCtx f {};
        f.grav = gdev;
        f.gen_M = nt ? nt->gen_M.as<double>() : nullptr;
        f.prop_M = nt ? nt->prop_M.as<double>() : nullptr;
        f.gen_tb = nt ? nt->gen_tb.as<double>() : nullptr;
        f.prop_tb = nt ? nt->prop_tb.as<double>() : nullptr;
        f.src = nt ? sources_dev() : SourcesDev {};
        f.h = h;
        return f;
    }

    // Node offsets of the 80 RK8 stage nodes (time_points.cu:35-46 with beforeEpoch = 4);
    // without OGP_QUIRK_BACKWARD_TIMES the backward stages use their true times.
    std::vector<double> gen_offsets(double h) const
    {
        static const double c[10] = { 0.0, 4.0 / 27.0, 2.0 / 9.0, 1.0 / 3.0, 1.0 / 2.0,
            2.0 / 3.0, 1.0 / 6.0, 1.0, 5.0 / 6.0, 1.0 };
        std::vector<double> t(80);
        for (int idx = 0; idx < 80; ++idx) {
            const int blk = idx / 10, k = idx % 10, s = blk - 4;
            if ((cfg.quirks & OGP_QUIRK_BACKWARD_TIMES) || s >= 0)
                t[idx] = ((double)blk + c[k]) * h;
            else
                t[idx] = ((double)(s + 1) - c[k]) * h + 4 * h;
        }
        return t;
    }

    // Ensures the node table for step h (time origin t0 seconds after the epoch) covers
    // propagation nodes 0 .. 4 + max_steps.
    NodeTable* node_table(double h, int64_t max_steps, double t0 = 0.0)
    {
        if (cfg.force != OGP_FORCE_GRAVITY)
            return nullptr;
        auto key = std::make_pair(h, t0);
        auto& slot = nodes[key];
        const int64_t need = 4 + std::max<int64_t>(max_steps, 4) + 1;
        if (slot && slot->n_prop >= need)
            return slot.get();
        if (!slot)
            slot = std::make_unique<NodeTable>();
        NodeTable& nt = *slot;
        const int ymdhm[5] = { cfg.year, cfg.month, cfg.day, cfg.hour, cfg.minute };
        const double shifted = (cfg.second - (double)4 * h) + t0;
        std::vector<double> off = gen_offsets(h);
        const int64_t len = std::max<int64_t>(need, 80);
        nt.offsets.reserve(sizeof(double) * len, &dev_bytes);
        nt.gen_M.reserve(sizeof(double) * 9 * 80, &dev_bytes);
        nt.gen_mjd.reserve(sizeof(double) * 80, &dev_bytes);
        nt.gen_tb.reserve(sizeof(double) * 6 * 80, &dev_bytes);
        HIPC(hipMemcpyAsync(nt.offsets.p, off.data(), sizeof(double) * 80, hipMemcpyHostToDevice,
            stream));
        launch_time_nodes(stream, fdev, ymdhm, shifted, nt.offsets.as<double>(), 80,
            nt.gen_M.as<double>(), nt.gen_mjd.as<double>(), nt.gen_tb.as<double>());
        check_launch();
        sync(); // offsets buffer is reused below
        off.resize(need);
        for (int64_t i = 0; i < need; ++i)
            off[i] = (double)i * h;
        nt.prop_M.reserve(sizeof(double) * 9 * need, &dev_bytes);
        nt.prop_mjd.reserve(sizeof(double) * need, &dev_bytes);
        nt.prop_tb.reserve(sizeof(double) * 6 * need, &dev_bytes);
        HIPC(hipMemcpyAsync(nt.offsets.p, off.data(), sizeof(double) * need, hipMemcpyHostToDevice,
            stream));
        launch_time_nodes(stream, fdev, ymdhm, shifted, nt.offsets.as<double>(), need,
            nt.prop_M.as<double>(), nt.prop_mjd.as<double>(), nt.prop_tb.as<double>());
        check_launch();
        sync();
        nt.n_prop = need;
        return &nt;
    }

    PropState prop_state(int64_t N)
    {
        PropState ps;
        ps.r = st_r.as<double>();
        ps.v = st_v.as<double>();
        ps.s = st_s.as<double>();
        ps.S = st_S.as<double>();
        ps.hist = st_hist.as<double>();
        ps.nsteps = st_ns.as<int64_t>();
        ps.N = N;
        return ps;
    }

    void reserve_state(int64_t N)
    {
        const size_t v3 = sizeof(double) * 3 * N;
        st_r.reserve(v3, &dev_bytes);
        st_v.reserve(v3, &dev_bytes);
        st_s.reserve(v3, &dev_bytes);
        st_S.reserve(v3, &dev_bytes);
        st_hist.reserve(9 * v3, &dev_bytes);
        st_ns.reserve(sizeof(int64_t) * N, &dev_bytes);
        perm.reserve(sizeof(int64_t) * N, &dev_bytes);
        reserve_scratch(N);
    }

    // startup work areas for N objects (shared by every startup of the context; each call
    // that runs a startup completes it before it returns)
    void reserve_scratch(int64_t N)
    {
        const size_t v3 = sizeof(double) * 3 * N;
        const int slots[7] = { 9, 9, 9, 9, 9, 10, 10 };
        for (int q = 0; q < 7; ++q)
            sc[q].reserve(slots[q] * v3, &dev_bytes);
    }

    StartupScratch scratch()
    {
        StartupScratch s;
        s.node_r = sc[0].as<double>();
        s.node_v = sc[1].as<double>();
        s.acc = sc[2].as<double>();
        s.sm = sc[3].as<double>();
        s.Sm = sc[4].as<double>();
        s.Kx = sc[5].as<double>();
        s.Kv = sc[6].as<double>();
        return s;
    }

    struct Group {
        double h;
        int64_t first, count, max_steps;
    };

    // Startup and GJ8 steps for all groups of the internal state (class-sorted order).
    // traj (optional) receives steps 1..traj_steps of the single group. ns (optional, host,
    // class-sorted order, plan_groups: the objects that run at a chunk start form a prefix of
    // their class) trims every chunk launch to the prefix of objects that still have steps in
    // it.
    void run_groups(const std::vector<Group>& groups, int64_t N, double t0, double* traj,
        int64_t traj_steps, const int64_t* ns = nullptr)
    {
        Event e0, e1, e2, e3;
        const PropState ps = prop_state(N);
        d_evals.reserve(sizeof(unsigned long long), &dev_bytes);
        HIPC(hipMemsetAsync(d_evals.p, 0, sizeof(unsigned long long), stream));
        // nodes
        std::vector<NodeTable*> tabs;
        HIPC(hipEventRecord(e0.e, stream));
        for (const Group& g : groups)
            tabs.push_back(node_table(g.h, g.max_steps, t0));
        HIPC(hipEventRecord(e1.e, stream));
        // A class with less than a third of the object-steps of the largest class starts late:
        // its startup and chunks wait until the largest class's remaining object-steps fall to
        // 3 times its own. It then fills the chip while the largest class runs down its longest
        // spans (few objects, launches that leave CUs idle) instead of competing with it while
        // it fills the chip (mixed 100k workload: 1,093 against 1,107 ms; starting at half or
        // 0.3 of the largest class's objects still active measured 1,097 and 1,096 ms, at 0.2
        // 1,112 ms). Every class integrates the same arithmetic whenever it starts.
        std::vector<double> work(groups.size(), 0.0);
        std::vector<char> started(groups.size(), 1);
        size_t qbig = 0;
        std::vector<double> big_prefix; // object-steps of the largest class's first j objects
        if (ns && groups.size() > 1) {
            for (size_t q = 0; q < groups.size(); ++q) {
                for (int64_t j = 0; j < groups[q].count; ++j)
                    work[q] += (double)ns[groups[q].first + j];
                if (work[q] > work[qbig])
                    qbig = q;
            }
            // (a class with no GJ8 step, only its startup, is never deferred)
            for (size_t q = 0; q < groups.size(); ++q)
                started[q] = q == qbig || 3.0 * work[q] >= work[qbig] || groups[q].max_steps < 5;
            big_prefix.assign(groups[qbig].count + 1, 0.0);
            for (int64_t j = 0; j < groups[qbig].count; ++j)
                big_prefix[j + 1] = big_prefix[j] + (double)ns[groups[qbig].first + j];
        }
        for (size_t q = 0; q < groups.size(); ++q) {
            if (!started[q])
                continue;
            const Group& g = groups[q];
            const int W = waves_for(g.count);
            launch_startup(stream, kind, force_ctx(tabs[q], g.h), rk8_variant, ps, scratch(),
                g.first, g.count, cfg.n_corrections, traj, traj_steps, W, fp_exit);
            check_launch();
        }
        HIPC(hipEventRecord(e2.e, stream));
        double launches = 0;

        // GJ8 chunks. Each class is cut into contiguous object slices whose chunk sequences run
        // on streams of their own: slices do not depend on each other, so the end of one slice's
        // launch, where few waves remain on the chip, overlaps launches of the other slices and
        // classes. The host issues the chunks of all slices of a class in lockstep (one k0 per
        // class); the waves per group follow the class's whole active count at k0.
        struct Lane {
            size_t q;
            int64_t first, count;
            hipStream_t st;
        };
        std::vector<Lane> lanes;
        for (size_t q = 0; q < groups.size(); ++q) {
            const Group& g = groups[q];
            const int64_t ngroups = (g.count + 63) / 64;
            // two slices for a class of at least 4,096 objects, except when it is the only
            // class and has no more object groups than the device has CUs: its launches then
            // leave no CU for a second slice to fill, and concurrent launches only compete
            // (10,000 objects at degree 20: 39 ms with one slice, 45 ms with two; 20,000:
            // 157 ms with one, 142 ms with two). Beside other classes, two slices of a small
            // class still pay (mixed 100k workload: 1,176 against 1,197 ms).
            const bool alone_small = groups.size() == 1 && ngroups <= simds / 4;
            int S = slices_override > 0 ? slices_override
                                        : (g.count >= 4096 && !alone_small ? 2 : 1);
            S = (int)std::max<int64_t>(1, std::min<int64_t>(S, ngroups));
            for (int s = 0; s < S; ++s) {
                const int64_t a = 64 * (ngroups * s / S), b = std::min(g.count, 64 * (ngroups * (s + 1) / S));
                lanes.push_back(Lane { q, g.first + a, b - a,
                    lane_streams[lanes.size() % kLaneStreams] });
            }
        }
        std::vector<ForceCtx> fctx;
        for (size_t q = 0; q < groups.size(); ++q)
            fctx.push_back(force_ctx(tabs[q], groups[q].h));
        auto active_in = [&](int64_t first, int64_t count, int64_t k0) {
            if (!ns)
                return count;
            const int64_t* b = ns + first;
            return (int64_t)(std::partition_point(b, b + count,
                                 [k0](int64_t s) { return s >= k0; }) - b);
        };
        // every launch is bracketed by events on its own stream, for its own duration
        std::vector<std::pair<std::unique_ptr<Event>, std::unique_ptr<Event>>> spans;
        auto launch = [&](hipStream_t st, size_t q, int64_t first, int64_t active, int64_t k0,
                          int shape, int64_t class_active) {
            const int64_t k1 = std::min<int64_t>(k0 + chunk, groups[q].max_steps + 1);
            spans.emplace_back(std::make_unique<Event>(), std::make_unique<Event>());
            HIPC(hipEventRecord(spans.back().first->e, st));
            launch_chunk(st, kind, fctx[q], ps, first, active, k0, k1, cfg.n_corrections, traj, 1,
                traj_steps, shape_waves(shape), shape_groups(shape), fp_exit,
                d_evals.as<unsigned long long>(), class_active);
            check_launch();
            HIPC(hipEventRecord(spans.back().second->e, st));
            launches += 1;
            return k1 - k0;
        };
        // join: `to` waits for the work issued so far on every stream of `from` (the events live
        // until the end of the call)
        std::vector<std::unique_ptr<Event>> marks;
        auto join = [&](hipStream_t to, const std::vector<hipStream_t>& from) {
            for (hipStream_t s : from) {
                if (s == to)
                    continue;
                marks.push_back(std::make_unique<Event>());
                HIPC(hipEventRecord(marks.back()->e, s));
                HIPC(hipStreamWaitEvent(to, marks.back()->e, 0));
            }
        };
        std::vector<hipStream_t> used;
        for (const Lane& l : lanes)
            if (std::find(used.begin(), used.end(), l.st) == used.end())
                used.push_back(l.st);
        for (hipStream_t s : used)
            join(s, { stream });

        std::vector<int64_t> k0(groups.size(), 5);
        for (bool more = true; more;) {
            more = false;
            for (size_t q = 0; q < groups.size(); ++q) {
                const Group& g = groups[q];
                if (k0[q] > g.max_steps)
                    continue;
                more = true;
                if (!started[q]) {
                    // the largest class's object-steps from k0 on: its active objects (a prefix,
                    // spans descending) each have ns - k0 + 1 steps left
                    const Group& gb = groups[qbig];
                    const int64_t a = k0[qbig] <= gb.max_steps
                        ? active_in(gb.first, gb.count, k0[qbig]) : 0;
                    if (big_prefix[a] - (double)a * (double)(k0[qbig] - 1) > 3.0 * work[q])
                        continue;
                    // after the largest class's chunks issued so far: the startup on the first
                    // stream of the class, then its chunks on all of them
                    std::vector<hipStream_t> mine, big;
                    for (const Lane& l : lanes) {
                        if (l.q == q)
                            mine.push_back(l.st);
                        else if (l.q == qbig)
                            big.push_back(l.st);
                    }
                    for (hipStream_t st : mine)
                        join(st, big);
                    launch_startup(mine[0], kind, fctx[q], rk8_variant, ps, scratch(), g.first,
                        g.count, cfg.n_corrections, traj, traj_steps, waves_for(g.count),
                        fp_exit);
                    check_launch();
                    for (hipStream_t st : mine)
                        join(st, { mine[0] });
                    started[q] = 1;
                }
                const int64_t active = active_in(g.first, g.count, k0[q]);
                // one chunk of every slice of the class with the given launch shape; returns
                // the steps done and appends (launch index, object-steps) of each launch to `out`
                auto class_chunk = [&](int shape, std::vector<std::pair<size_t, double>>* out) {
                    int64_t steps = 0;
                    for (const Lane& l : lanes) {
                        if (l.q != q)
                            continue;
                        const int64_t a = active_in(l.first, l.count, k0[q]);
                        if (a > 0) {
                            steps = launch(l.st, q, l.first, a, k0[q], shape, active);
                            if (out)
                                out->emplace_back(spans.size() - 1, (double)a * (double)steps);
                        }
                    }
                    k0[q] += steps > 0 ? steps : chunk;
                };
                const std::vector<int> cand = needs_tuning(active) ? tuning_candidates(active)
                                                                   : std::vector<int>();
                if (cand.size() == 1)
                    tuned_shape[group_bucket(active)] = cand[0];
                if (cand.size() > 1
                    && g.max_steps - k0[q] + 1 >= 2 * (int64_t)cand.size() * chunk) {
                    // the next chunks of the class, two per candidate in mirrored order (A B C
                    // C B A, which cancels a steady drift), run as every other chunk does (the
                    // slices concurrently on their streams); a candidate's rate is the sum over
                    // its launches of object-steps / launch duration, and the fastest is kept.
                    // The chunks do real work (results do not depend on the shape).
                    const int nc = (int)cand.size();
                    std::vector<std::vector<std::pair<size_t, double>>> trial(nc);
                    for (int c = 0; c < nc; ++c)
                        class_chunk(cand[c], &trial[c]);
                    for (int c = nc - 1; c >= 0; --c)
                        class_chunk(cand[c], &trial[c]);
                    int best = 0;
                    double best_rate = 0.0;
                    for (int c = 0; c < nc; ++c) {
                        double rate = 0.0;
                        for (const auto& t : trial[c]) {
                            const auto& sp = spans[t.first];
                            HIPC(hipEventSynchronize(sp.second->e));
                            float ms = 0;
                            HIPC(hipEventElapsedTime(&ms, sp.first->e, sp.second->e));
                            rate += t.second / std::max(1e-3, (double)ms);
                        }
                        if (rate > best_rate) {
                            best_rate = rate;
                            best = cand[c];
                        }
                    }
                    tuned_shape[group_bucket(active)] = best;
                    continue;
                }
                class_chunk(shape_for(active), nullptr);
            }
        }
        join(stream, used);
        HIPC(hipEventRecord(e3.e, stream));
        sync();
        float ms0 = 0, ms1 = 0, ms2 = 0;
        HIPC(hipEventElapsedTime(&ms0, e0.e, e1.e));
        HIPC(hipEventElapsedTime(&ms1, e1.e, e2.e));
        HIPC(hipEventElapsedTime(&ms2, e2.e, e3.e));
        timing.nodes_ms = ms0;
        timing.startup_ms = ms1;
        timing.steps_ms = ms2;
        timing.total_ms = (double)ms0 + ms1 + ms2;
        timing.step_launches = launches;
        double sum = 0.0;
        for (const auto& sp : spans) {
            float ms = 0;
            HIPC(hipEventElapsedTime(&ms, sp.first->e, sp.second->e));
            sum += ms;
        }
        timing.launch_ms_sum = sum;
        unsigned long long evals = 0;
        HIPC(hipMemcpy(&evals, d_evals.p, sizeof(evals), hipMemcpyDeviceToHost));
        timing.evaluations = (double)evals;
    }
};

namespace {

int fail(ogp_context* ctx, int code, const std::string& msg)
{
    g_error = msg;
    if (ctx)
        ctx->error = msg;
    return code;
}

template <class F> int guarded(ogp_context* ctx, F&& fn)
{
    try {
        fn();
        return OGP_OK;
    } catch (const ApiError& e) {
        return fail(ctx, e.code, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(ctx, OGP_ERR_ARGUMENT, e.what());
    } catch (const std::bad_alloc&) {
        return fail(ctx, OGP_ERR_DEVICE, "host out of memory");
    } catch (const std::exception& e) {
        return fail(ctx, OGP_ERR_DATA, e.what());
    }
}

// Copies a host table to the device and waits: callers pass temporaries that are destroyed
// right after the call.
template <typename T> void upload(ogp_context* c, DevBuf& b, const std::vector<T>& v)
{
    b.reserve(sizeof(T) * std::max<size_t>(v.size(), 1), &c->dev_bytes);
    if (!v.empty()) {
        HIPC(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, c->stream));
        c->sync();
    }
}

void check_finite(const double* p, int64_t n, const char* what)
{
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i]))
            throw ApiError(OGP_ERR_NONFINITE, std::string("non-finite value in ") + what
                    + " at index " + std::to_string(i));
}

} // namespace

// =========================================================================================
extern "C" {

void ogp_default_config(ogp_config* cfg)
{
    if (!cfg)
        return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->data_dir = "data";
    cfg->degree = -1;
    cfg->precision = OGP_PRECISION_FP64;
    cfg->split_degree = 2;
    cfg->n_corrections = 7;
    cfg->quirks = OGP_QUIRKS_REFERENCE;
    cfg->force = OGP_FORCE_GRAVITY;
    cfg->year = 2025;
    cfg->month = 1;
    cfg->day = 1;
}

int ogp_create(const ogp_config* cfg, ogp_context** out)
{
    if (!cfg || !out)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null config or output pointer");
    *out = nullptr;
    auto ctx = std::make_unique<ogp_context>();
    int rc = guarded(ctx.get(), [&] {
        ogp_context& c = *ctx;
        c.cfg = *cfg;
        c.data_dir = cfg->data_dir ? cfg->data_dir : "data";
        c.cfg.data_dir = c.data_dir.c_str();
        arg_check(cfg->n_corrections >= 1, "n_corrections must be at least 1");
        arg_check(cfg->force >= OGP_FORCE_GRAVITY && cfg->force <= OGP_FORCE_DAMPED,
            "unknown force kind");
        arg_check(cfg->precision == OGP_PRECISION_FP64 || cfg->precision == OGP_PRECISION_SPLIT,
            "unknown precision");
        arg_check(cfg->month >= 1 && cfg->month <= 12 && cfg->day >= 1 && cfg->day <= 31,
            "invalid epoch date");
        int ndev = 0;
        HIPC(hipGetDeviceCount(&ndev));
        if (ndev < 1)
            throw ApiError(OGP_ERR_DEVICE, "no HIP device");
        HIPC(hipSetDevice(0));
        hip_check(noalloc_scratch_setup(0, &c.noalloc), "hipGetSymbolAddress(scratch)");
        int cus = 0;
        HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
        c.simds = std::max(4, 4 * cus);
        HIPC(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        for (hipStream_t& s : c.lane_streams)
            HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        c.rk8_variant = (cfg->quirks & OGP_QUIRK_RK8_ROW7) ? 0 : 1;
        if (cfg->force == OGP_FORCE_GRAVITY) {
            c.kind = cfg->precision == OGP_PRECISION_SPLIT ? kGravitySplit : kGravity;
            c.grav = load_gravity(c.data_dir + "/JGM3.gfc", cfg->degree);
            c.frames = load_frames(c.data_dir);
            // zero records after each table (column_terms fetches a column's last batch
            // whole) and zero entries after the diagonal ratios (read one column ahead, also
            // by an empty part that starts at column L + 1)
            std::vector<double> rec = c.grav.records;
            std::vector<float> recf = c.grav.records_f;
            rec.resize(rec.size() + kRecordPad * kRecordLen, 0.0);
            recf.resize(recf.size() + kRecordPad * kRecordLen, 0.0f);
            std::vector<double> jext = c.grav.jac_ext;
            jext.resize(jext.size() + kRecordPad * kJacExtLen, 0.0);
            std::vector<double> diag = c.grav.diag_ratio;
            diag.resize(diag.size() + 2, 0.0);
            upload(&c, c.d_rec, rec);
            upload(&c, c.d_recf, recf);
            upload(&c, c.d_jext, jext);
            upload(&c, c.d_diag, diag);
            c.gdev.rec = c.d_rec.as<double>();
            c.gdev.recf = c.d_recf.as<float>();
            c.gdev.jext = c.d_jext.as<double>();
            c.gdev.diag = c.d_diag.as<double>();
            std::vector<int> parts(kParts + 1);
            column_parts(c.grav.L, parts.data());
            upload(&c, c.d_parts, parts);
            c.gdev.parts = c.d_parts.as<int>();
            // D_m = prod_{j=1..m} diag[j] at each part's first column (the closed-form start of
            // the diagonal recurrence; parts past the last column have no columns to start)
            std::vector<double> part_diag(kParts, 0.0);
            for (int q = 0; q < kParts; ++q) {
                const int m0 = std::min(parts[q], c.grav.L);
                double D = 1.0;
                for (int j = 1; j <= m0; ++j)
                    D *= c.grav.diag_ratio[j];
                part_diag[q] = D;
            }
            upload(&c, c.d_part_diag, part_diag);
            c.gdev.part_diag = c.d_part_diag.as<double>();
            c.gdev.L = c.grav.L;
            c.gdev.nsplit = cfg->precision == OGP_PRECISION_SPLIT ? std::max(cfg->split_degree, 1)
                                                                  : c.grav.L;
            c.gdev.gm = c.grav.gm;
            c.gdev.radius = c.grav.radius;
            std::vector<double> eop;
            for (const EopRecord& r : c.frames.eop) {
                eop.push_back(r.mjd);
                eop.push_back(r.xp);
                eop.push_back(r.yp);
                eop.push_back(r.ut1_utc);
            }
            upload(&c, c.d_eop, eop);
            upload(&c, c.d_leap_mjd, c.frames.leap_mjd);
            upload(&c, c.d_leap_sec, c.frames.leap_sec);
            const std::vector<NutationTerm>* tabs[4] = { &c.frames.lon0, &c.frames.lon1,
                &c.frames.obl0, &c.frames.obl1 };
            for (int q = 0; q < 4; ++q) {
                std::vector<double> packed;
                for (const NutationTerm& t : *tabs[q]) {
                    packed.push_back(t.sin_uas);
                    packed.push_back(t.cos_uas);
                    for (int k = 0; k < 14; ++k)
                        packed.push_back((double)t.mult[k]);
                }
                upload(&c, c.d_nut[q], packed);
                c.fdev.nut[q] = c.d_nut[q].as<double>();
                c.fdev.n_nut[q] = (int)tabs[q]->size();
            }
            c.fdev.eop = c.d_eop.as<double>();
            c.fdev.n_eop = (int64_t)c.frames.eop.size();
            c.fdev.leap_mjd = c.d_leap_mjd.as<double>();
            c.fdev.leap_sec = c.d_leap_sec.as<double>();
            c.fdev.n_leap = (int)c.frames.leap_mjd.size();
        } else {
            c.kind = cfg->force;
        }
        c.sync();
    });
    if (rc != OGP_OK)
        return rc;
    *out = ctx.release();
    return OGP_OK;
}

void ogp_destroy(ogp_context* ctx) { delete ctx; }

const char* ogp_last_error(const ogp_context* ctx)
{
    return ctx ? ctx->error.c_str() : g_error.c_str();
}

int64_t ogp_device_bytes(const ogp_context* ctx) { return ctx ? ctx->dev_bytes : 0; }

int ogp_gravity_itrs(ogp_context* ctx, const double* r, int64_t n, double* a)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n >= 0 && (n == 0 || (r && a)), "invalid gravity arguments");
        if (ctx->cfg.force != OGP_FORCE_GRAVITY)
            throw ApiError(OGP_ERR_STATE, "context was created without the gravity field");
        if (n == 0)
            return;
        check_finite(r, 3 * n, "positions");
        ctx->io_a.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
        ctx->io_b.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
        HIPC(hipMemcpyAsync(ctx->io_a.p, r, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
            ctx->stream));
        launch_gravity_aos(ctx->stream, ctx->kind == kGravitySplit, ctx->gdev, nullptr,
            ctx->io_a.as<double>(), n, ctx->io_b.as<double>(), 3, 64);
        ctx->check_launch();
        HIPC(hipMemcpyAsync(a, ctx->io_b.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
            ctx->stream));
        ctx->sync();
    });
}

int ogp_accel(ogp_context* ctx, const double* r, const double* v, int64_t n, double step_s,
    int64_t node, double* a, int64_t out_stride)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n >= 0 && step_s > 0 && node >= 0 && out_stride >= 3, "invalid accel arguments");
        if (ctx->cfg.force != OGP_FORCE_GRAVITY)
            throw ApiError(OGP_ERR_STATE, "context was created without the gravity field");
        arg_check(v || !(ctx->sources.mask & kSourceDrag), "drag needs the velocities");
        if (n == 0)
            return;
        ctx->bc_obj = ctx->stage_bc(ctx->d_bc, n, nullptr);
        const SourcesDev src = ctx->sources_dev();
        check_finite(r, 3 * n, "positions");
        if (v)
            check_finite(v, 3 * n, "velocities");
        NodeTable* nt = ctx->node_table(step_s, node);
        const int64_t out_len = out_stride * (n - 1) + 3;
        ctx->io_a.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
        ctx->io_b.reserve(sizeof(double) * out_len, &ctx->dev_bytes);
        HIPC(hipMemcpyAsync(ctx->io_a.p, r, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
            ctx->stream));
        HIPC(hipMemcpyAsync(ctx->io_b.p, a, sizeof(double) * out_len, hipMemcpyHostToDevice,
            ctx->stream));
        if (v) {
            ctx->io_c.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
            HIPC(hipMemcpyAsync(ctx->io_c.p, v, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                ctx->stream));
        }
        launch_gravity_aos(ctx->stream, ctx->kind == kGravitySplit, ctx->gdev,
            nt->prop_M.as<double>() + 9 * node, ctx->io_a.as<double>(), n, ctx->io_b.as<double>(),
            out_stride, 64, src, nt->prop_tb.as<double>() + 6 * node,
            v ? ctx->io_c.as<double>() : nullptr);
        ctx->check_launch();
        HIPC(hipMemcpyAsync(a, ctx->io_b.p, sizeof(double) * out_len, hipMemcpyDeviceToHost,
            ctx->stream));
        ctx->sync();
    });
}

int ogp_accel_device(ogp_context* ctx, const double* d_r, const double* d_v, int64_t n,
    double step_s, int64_t node, double* d_a, int64_t out_stride)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n >= 0 && std::isfinite(step_s) && step_s > 0 && node >= 0 && out_stride >= 3,
            "invalid accel arguments");
        arg_check(n == 0 || (d_r && d_a), "null array");
        if (ctx->cfg.force != OGP_FORCE_GRAVITY)
            throw ApiError(OGP_ERR_STATE, "context was created without the gravity field");
        arg_check(d_v || !(ctx->sources.mask & kSourceDrag), "drag needs the velocities");
        if (n == 0)
            return;
        ctx->bc_obj = ctx->stage_bc(ctx->d_bc, n, nullptr);
        NodeTable* nt = ctx->node_table(step_s, node);
        launch_gravity_aos(ctx->stream, ctx->kind == kGravitySplit, ctx->gdev,
            nt->prop_M.as<double>() + 9 * node, d_r, n, d_a, out_stride, 64, ctx->sources_dev(),
            nt->prop_tb.as<double>() + 6 * node, d_v);
        ctx->check_launch();
        ctx->sync();
    });
}

int ogp_accel_generation(ogp_context* ctx, const double* d_r, const double* d_v, int64_t n,
    double step_s, int step_num, int k_index, double* d_K)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n >= 0 && std::isfinite(step_s) && step_s > 0, "invalid accel arguments");
        // force_model.cu:62-65 rejects the stage index; the step is one of the 4 + 4 RK8 steps
        arg_check(k_index >= 0 && k_index < 10, "kIndex must be between 0 and 9");
        arg_check(step_num != 0 && step_num >= -4 && step_num <= 4,
            "stepNum must be one of -4 .. -1, 1 .. 4");
        arg_check(n == 0 || (d_r && d_K), "null array");
        if (ctx->cfg.force != OGP_FORCE_GRAVITY)
            throw ApiError(OGP_ERR_STATE, "context was created without the gravity field");
        arg_check(d_v || !(ctx->sources.mask & kSourceDrag), "drag needs the velocities");
        if (n == 0)
            return;
        ctx->bc_obj = ctx->stage_bc(ctx->d_bc, n, nullptr);
        NodeTable* nt = ctx->node_table(step_s, 4);
        const int idx = ((step_num > 0 ? step_num - 1 : step_num) + 4) * 10 + k_index;
        launch_gravity_aos(ctx->stream, ctx->kind == kGravitySplit, ctx->gdev,
            nt->gen_M.as<double>() + 9 * idx, d_r, n, d_K + 3 * k_index, 30, 64,
            ctx->sources_dev(), nt->gen_tb.as<double>() + 6 * idx, d_v);
        ctx->check_launch();
        ctx->sync();
    });
}

int ogp_time_nodes(ogp_context* ctx, double step_s, int before, const double* offsets_s,
    int64_t len, double* M, double* mjd)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(len >= 0 && (len == 0 || (offsets_s && M && mjd)), "invalid node arguments");
        if (ctx->cfg.force != OGP_FORCE_GRAVITY)
            throw ApiError(OGP_ERR_STATE, "context was created without the frame datasets");
        if (len == 0)
            return;
        const ogp_config& c = ctx->cfg;
        const int ymdhm[5] = { c.year, c.month, c.day, c.hour, c.minute };
        const double shifted = c.second - (double)before * step_s;
        ctx->io_a.reserve(sizeof(double) * len, &ctx->dev_bytes);
        ctx->io_b.reserve(sizeof(double) * 9 * len, &ctx->dev_bytes);
        ctx->io_c.reserve(sizeof(double) * len, &ctx->dev_bytes);
        HIPC(hipMemcpyAsync(ctx->io_a.p, offsets_s, sizeof(double) * len, hipMemcpyHostToDevice,
            ctx->stream));
        launch_time_nodes(ctx->stream, ctx->fdev, ymdhm, shifted, ctx->io_a.as<double>(), len,
            ctx->io_b.as<double>(), ctx->io_c.as<double>());
        ctx->check_launch();
        HIPC(hipMemcpyAsync(M, ctx->io_b.p, sizeof(double) * 9 * len, hipMemcpyDeviceToHost,
            ctx->stream));
        HIPC(hipMemcpyAsync(mjd, ctx->io_c.p, sizeof(double) * len, hipMemcpyDeviceToHost,
            ctx->stream));
        ctx->sync();
    });
}

} // extern "C"

namespace {

// Groups objects by step size (one time-node table and one launch sequence per class) and
// orders each class by descending step count so that the lanes of a wave end together. With
// key (the semi-major axis, orbit_size), objects whose last steps fall in the same `sub`-step
// window of their chunks are ordered by key instead: a group of 64 objects runs each step's
// corrector evaluations until its slowest object reaches the fixed point, and the count an object
// needs rises as its orbit shrinks (a stronger gravity gradient), so groups of similar orbits run
// fewer evaluations. The windows nest in the chunks (sub divides chunk), so the objects that
// still run at a chunk start remain a prefix of the class.
std::vector<ogp_context::Group> plan_groups(const double* step_s, const int64_t* steps, int64_t n,
    std::vector<int64_t>& perm, const float* key = nullptr, int64_t sub = 0)
{
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0);
    // One 64-bit order key per object, sorted by a stable radix sort (100,000 objects: about
    // 4 ms on the host, against 17 ms for a comparison sort with the same order):
    // step-size class (ascending step size), last-step window (descending), orbit size in km
    // (ascending), last step within the window (descending); objects with fewer than 5 steps
    // (no GJ8 step) after every window, by descending step count.
    const int64_t w = key && sub > 0 ? sub : 1;
    std::map<double, uint64_t> cls;
    {
        double last = 0.0;
        bool have = false;
        for (int64_t i = 0; i < n; ++i)
            if (!have || step_s[i] != last) {
                cls.emplace(step_s[i], 0);
                last = step_s[i];
                have = true;
            }
        uint64_t c = 0;
        for (auto& kv : cls)
            kv.second = c++;
    }
    int64_t wmax = -1;
    for (int64_t i = 0; i < n; ++i)
        wmax = std::max<int64_t>(wmax, steps[i] >= 5 ? (steps[i] - 5) / w : -1);
    auto bits_for = [](uint64_t v) {
        int b = 0;
        while (b < 64 && (v >> b) != 0)
            ++b;
        return b;
    };
    const int cb = bits_for(cls.size() - 1), wb = bits_for((uint64_t)(wmax + 1)),
              ab = key ? 20 : 0, sb = std::max(3, bits_for((uint64_t)(w - 1)));
    if (n > 0 && cb + wb + ab + sb <= 64) {
        std::vector<uint64_t> k(n);
        auto it = cls.begin();
        for (int64_t i = 0; i < n; ++i) {
            if (it->first != step_s[i])
                it = cls.find(step_s[i]);
            const int64_t win = steps[i] >= 5 ? (steps[i] - 5) / w : -1;
            uint64_t ak = 0;
            if (key && win >= 0) {
                const float km = key[i] * 1e-3f;
                ak = km > 0.0f ? (uint64_t)std::min(km, 1048575.0f) : 0;
            }
            const uint64_t sd = win >= 0 ? (uint64_t)(w - 1 - (steps[i] - 5) % w)
                                         : (uint64_t)(4 - std::max<int64_t>(steps[i], 0));
            k[i] = ((((it->second << wb) | (uint64_t)(wmax - win)) << ab | ak) << sb) | sd;
        }
        std::vector<uint64_t> k2(n);
        std::vector<int64_t> p2(n);
        constexpr int D = 11; // bits per pass
        std::vector<uint32_t> cnt((1u << D) + 1);
        for (int sh = 0; sh < cb + wb + ab + sb; sh += D) {
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (int64_t j = 0; j < n; ++j)
                ++cnt[((k[j] >> sh) & ((1u << D) - 1)) + 1];
            for (uint32_t d = 0; d < (1u << D); ++d)
                cnt[d + 1] += cnt[d];
            for (int64_t j = 0; j < n; ++j) {
                const uint32_t at = cnt[(k[j] >> sh) & ((1u << D) - 1)]++;
                k2[at] = k[j];
                p2[at] = perm[j];
            }
            k.swap(k2);
            perm.swap(p2);
        }
    } else if (n > 0) {
        // keys too wide for 64 bits (step counts beyond 2^40): the same order by comparisons
        std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) {
            if (step_s[a] != step_s[b])
                return step_s[a] < step_s[b];
            const int64_t wa = steps[a] >= 5 ? (steps[a] - 5) / w : -1;
            const int64_t wb_ = steps[b] >= 5 ? (steps[b] - 5) / w : -1;
            if (wa != wb_)
                return wa > wb_;
            if (key && wa >= 0) {
                const float ka = std::max(0.0f, std::min(key[a] * 1e-3f, 1048575.0f));
                const float kb = std::max(0.0f, std::min(key[b] * 1e-3f, 1048575.0f));
                if ((uint64_t)ka != (uint64_t)kb)
                    return (uint64_t)ka < (uint64_t)kb;
            }
            return steps[a] > steps[b];
        });
    }
    std::vector<ogp_context::Group> groups;
    for (int64_t p = 0; p < n; ++p) {
        const int64_t i = perm[p];
        if (groups.empty() || groups.back().h != step_s[i])
            groups.push_back({ step_s[i], p, 0, 0 });
        groups.back().count++;
        groups.back().max_steps = std::max(groups.back().max_steps, steps[i]);
    }
    return groups;
}

// Semi-major axis from the vis-viva equation (0 where it is not finite), the grouping key of
// plan_groups.
float orbit_size(const double* r, const double* v, double gm)
{
    const double rr = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double a = 1.0 / (2.0 / rr - v2 / gm);
    return std::isfinite(a) ? (float)a : 0.0f;
}

void validate_prop(const double* step_s, const int64_t* steps, int64_t n)
{
    arg_check(n >= 0, "negative object count");
    for (int64_t i = 0; i < n; ++i) {
        arg_check(std::isfinite(step_s[i]) && step_s[i] > 0,
            "step size must be positive and finite (object " + std::to_string(i) + ")");
        arg_check(steps[i] >= 0, "negative step count (object " + std::to_string(i) + ")");
    }
}

// Object flags for the failure-flag entry points (ogp.h: OGP_OBJECT_*).
bool finite6(const double* r, const double* v)
{
    return std::isfinite(r[0]) && std::isfinite(r[1]) && std::isfinite(r[2])
        && std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

void record_counts(ogp_context* ctx, const int64_t* steps, int64_t n)
{
    double os = 0, gj = 0;
    for (int64_t i = 0; i < n; ++i) {
        os += (double)steps[i];
        gj += (double)std::max<int64_t>(steps[i] - 4, 0);
    }
    ctx->timing.object_steps = os;
    ctx->timing.gj_object_steps = gj;
}

// The runner's propagate() from host arrays. With status, objects with a non-finite initial
// state are skipped (zero steps) and flagged, and non-finite results are flagged; without it a
// non-finite initial state fails the call.
void propagate_host(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    const double* step_s, const int64_t* steps_in, double* r_out, double* v_out, int32_t* status)
{
    arg_check(n == 0 || (r0 && v0 && step_s && steps_in && r_out && v_out), "null array");
    validate_prop(step_s, steps_in, n);
    if (n == 0)
        return;
    std::vector<int64_t> steps(steps_in, steps_in + n);
    if (status) {
        for (int64_t i = 0; i < n; ++i) {
            status[i] = finite6(r0 + 3 * i, v0 + 3 * i) ? OGP_OBJECT_OK : OGP_OBJECT_NONFINITE_INPUT;
            if (status[i] != OGP_OBJECT_OK)
                steps[i] = 0;
        }
    } else {
        check_finite(r0, 3 * n, "positions");
        check_finite(v0, 3 * n, "velocities");
    }
    std::vector<int64_t> perm;
    std::vector<float> key;
    if (ctx->cfg.force == OGP_FORCE_GRAVITY) {
        key.resize(n);
        for (int64_t i = 0; i < n; ++i)
            key[i] = orbit_size(r0 + 3 * i, v0 + 3 * i, ctx->grav.gm);
    }
    auto groups = plan_groups(step_s, steps.data(), n, perm, key.empty() ? nullptr : key.data(),
        ctx->sort_window());
    std::vector<int64_t> ns(n);
    for (int64_t p = 0; p < n; ++p)
        ns[p] = steps[perm[p]];
    ctx->bc_obj = ctx->stage_bc(ctx->d_bc, n, perm.data());
    ctx->reserve_state(n);
    ctx->io_a.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
    ctx->io_b.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
    Event h0, h1, d0, d1;
    HIPC(hipEventRecord(h0.e, ctx->stream));
    HIPC(hipMemcpyAsync(ctx->perm.p, perm.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
        ctx->stream));
    HIPC(hipMemcpyAsync(ctx->st_ns.p, ns.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
        ctx->stream));
    HIPC(hipMemcpyAsync(ctx->io_a.p, r0, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
        ctx->stream));
    HIPC(hipMemcpyAsync(ctx->io_b.p, v0, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
        ctx->stream));
    launch_permute(ctx->stream, 0, ctx->io_a.as<double>(), ctx->perm.as<int64_t>(),
        ctx->st_r.as<double>(), n);
    launch_permute(ctx->stream, 0, ctx->io_b.as<double>(), ctx->perm.as<int64_t>(),
        ctx->st_v.as<double>(), n);
    ctx->check_launch();
    HIPC(hipEventRecord(h1.e, ctx->stream));
    ctx->run_groups(groups, n, 0.0, nullptr, 0, ns.data());
    HIPC(hipEventRecord(d0.e, ctx->stream));
    launch_permute(ctx->stream, 1, ctx->st_r.as<double>(), ctx->perm.as<int64_t>(),
        ctx->io_a.as<double>(), n);
    launch_permute(ctx->stream, 1, ctx->st_v.as<double>(), ctx->perm.as<int64_t>(),
        ctx->io_b.as<double>(), n);
    ctx->check_launch();
    HIPC(hipMemcpyAsync(r_out, ctx->io_a.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
        ctx->stream));
    HIPC(hipMemcpyAsync(v_out, ctx->io_b.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
        ctx->stream));
    HIPC(hipEventRecord(d1.e, ctx->stream));
    ctx->sync();
    float a = 0, b = 0;
    HIPC(hipEventElapsedTime(&a, h0.e, h1.e));
    HIPC(hipEventElapsedTime(&b, d0.e, d1.e));
    ctx->timing.h2d_ms = a;
    ctx->timing.d2h_ms = b;
    record_counts(ctx, steps.data(), n);
    if (status)
        for (int64_t i = 0; i < n; ++i)
            if (status[i] == OGP_OBJECT_OK && !finite6(r_out + 3 * i, v_out + 3 * i))
                status[i] = OGP_OBJECT_NONFINITE_RESULT;
}

// The same on device-resident SoA arrays. The initial state is checked on the device: without
// status a non-finite value fails the call before any propagation, with status the object is
// skipped and flagged (and non-finite results are flagged after the propagation).
void propagate_device(ogp_context* ctx, const double* d_r0, const double* d_v0, int64_t n,
    const double* step_s, const int64_t* steps_in, double* d_r, double* d_v, int32_t* status)
{
    arg_check(n == 0 || (d_r0 && d_v0 && step_s && steps_in && d_r && d_v), "null array");
    validate_prop(step_s, steps_in, n);
    if (n == 0)
        return;
    std::vector<int64_t> steps(steps_in, steps_in + n);
    ctx->d_flags.reserve(sizeof(int32_t) * n, &ctx->dev_bytes);
    ctx->d_count.reserve(sizeof(uint32_t), &ctx->dev_bytes);
    const uint32_t zero = 0;
    HIPC(hipMemcpyAsync(ctx->d_count.p, &zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream));
    launch_finite_soa(ctx->stream, d_r0, d_v0, n, ctx->d_flags.as<int32_t>(),
        ctx->d_count.as<uint32_t>(), OGP_OBJECT_NONFINITE_INPUT, true);
    ctx->check_launch();
    uint32_t bad = 0;
    HIPC(hipMemcpyAsync(&bad, ctx->d_count.p, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream));
    std::vector<float> key;
    if (ctx->cfg.force == OGP_FORCE_GRAVITY) {
        key.resize(n);
        ctx->d_key.reserve(sizeof(float) * n, &ctx->dev_bytes);
        launch_orbit_key_soa(ctx->stream, d_r0, d_v0, n, ctx->grav.gm, ctx->d_key.as<float>());
        ctx->check_launch();
        HIPC(hipMemcpyAsync(key.data(), ctx->d_key.p, sizeof(float) * n, hipMemcpyDeviceToHost,
            ctx->stream));
    }
    ctx->sync();
    std::vector<int32_t> flags;
    if (bad > 0 || status) {
        flags.resize(n);
        if (bad > 0)
            HIPC(hipMemcpy(flags.data(), ctx->d_flags.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    }
    if (bad > 0) {
        const int64_t first = std::find_if(flags.begin(), flags.end(),
                                  [](int32_t f) { return f != 0; }) - flags.begin();
        if (!status)
            throw ApiError(OGP_ERR_NONFINITE,
                "non-finite value in the initial state of object " + std::to_string(first));
        for (int64_t i = 0; i < n; ++i)
            if (flags[i] != 0)
                steps[i] = 0;
    }
    std::vector<int64_t> perm;
    auto groups = plan_groups(step_s, steps.data(), n, perm, key.empty() ? nullptr : key.data(),
        ctx->sort_window());
    std::vector<int64_t> ns(n);
    for (int64_t p = 0; p < n; ++p)
        ns[p] = steps[perm[p]];
    ctx->bc_obj = ctx->stage_bc(ctx->d_bc, n, perm.data());
    ctx->reserve_state(n);
    HIPC(hipMemcpyAsync(ctx->perm.p, perm.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
        ctx->stream));
    HIPC(hipMemcpyAsync(ctx->st_ns.p, ns.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
        ctx->stream));
    launch_permute(ctx->stream, 2, d_r0, ctx->perm.as<int64_t>(), ctx->st_r.as<double>(), n);
    launch_permute(ctx->stream, 2, d_v0, ctx->perm.as<int64_t>(), ctx->st_v.as<double>(), n);
    ctx->check_launch();
    ctx->run_groups(groups, n, 0.0, nullptr, 0, ns.data());
    launch_permute(ctx->stream, 3, ctx->st_r.as<double>(), ctx->perm.as<int64_t>(), d_r, n);
    launch_permute(ctx->stream, 3, ctx->st_v.as<double>(), ctx->perm.as<int64_t>(), d_v, n);
    ctx->check_launch();
    if (status) {
        launch_finite_soa(ctx->stream, d_r, d_v, n, ctx->d_flags.as<int32_t>(),
            ctx->d_count.as<uint32_t>(), OGP_OBJECT_NONFINITE_RESULT, false);
        ctx->check_launch();
        HIPC(hipMemcpyAsync(status, ctx->d_flags.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
            ctx->stream));
    }
    ctx->sync();
    record_counts(ctx, steps.data(), n);
    ctx->timing.h2d_ms = ctx->timing.d2h_ms = 0;
}

} // namespace

extern "C" {

int ogp_propagate(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    const double* step_s, const int64_t* steps, double* r_out, double* v_out)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] { propagate_host(ctx, r0, v0, n, step_s, steps, r_out, v_out, nullptr); });
}

int ogp_propagate_checked(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    const double* step_s, const int64_t* steps, double* r_out, double* v_out, int32_t* status)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n == 0 || status, "null status array");
        propagate_host(ctx, r0, v0, n, step_s, steps, r_out, v_out, status);
    });
}

int ogp_propagate_device(ogp_context* ctx, const double* d_r0, const double* d_v0, int64_t n,
    const double* step_s, const int64_t* steps, double* d_r, double* d_v)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] { propagate_device(ctx, d_r0, d_v0, n, step_s, steps, d_r, d_v, nullptr); });
}

int ogp_propagate_device_checked(ogp_context* ctx, const double* d_r0, const double* d_v0,
    int64_t n, const double* step_s, const int64_t* steps, double* d_r, double* d_v,
    int32_t* status)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n == 0 || status, "null status array");
        propagate_device(ctx, d_r0, d_v0, n, step_s, steps, d_r, d_v, status);
    });
}

// ---- Monte-Carlo populations -----------------------------------------------------------
namespace {

PerturbArgs perturb_args(const double* r0, const double* v0, int64_t n, double strength,
    uint64_t seed, const double* obj_r, const double* obj_v)
{
    arg_check(r0 && v0, "null epoch state");
    arg_check(n >= 0, "n must be >= 0");
    arg_check(std::isfinite(strength), "strength must be finite");
    PerturbArgs a {};
    for (int c = 0; c < 3; ++c) {
        a.r0[c] = r0[c];
        a.v0[c] = v0[c];
    }
    // generatePerturbedSatellites appends the object only when both of its arrays are given
    a.append = obj_r && obj_v ? 1 : 0;
    if (a.append)
        for (int c = 0; c < 3; ++c) {
            a.obj_r[c] = obj_r[c];
            a.obj_v[c] = obj_v[c];
        }
    a.strength = strength;
    a.k0 = (uint32_t)seed;
    a.k1 = (uint32_t)(seed >> 32);
    a.n = n;
    return a;
}

} // namespace

int ogp_perturbed_satellites(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    double strength, uint64_t seed, const double* obj_r, const double* obj_v, double* r_out,
    double* v_out)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        PerturbArgs a = perturb_args(r0, v0, n, strength, seed, obj_r, obj_v);
        const int64_t count = n + a.append;
        arg_check(count == 0 || (r_out && v_out), "null output array");
        if (count == 0)
            return;
        const size_t bytes = sizeof(double) * 3 * count;
        ctx->io_a.reserve(bytes, &ctx->dev_bytes);
        ctx->io_b.reserve(bytes, &ctx->dev_bytes);
        a.r = ctx->io_a.as<double>();
        a.v = ctx->io_b.as<double>();
        a.stride = 0;
        launch_perturb(ctx->stream, a);
        ctx->check_launch();
        HIPC(hipMemcpyAsync(r_out, a.r, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipMemcpyAsync(v_out, a.v, bytes, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
    });
}

int ogp_perturbed_satellites_device(ogp_context* ctx, const double* r0, const double* v0,
    int64_t n, double strength, uint64_t seed, const double* obj_r, const double* obj_v,
    double* d_r_soa, double* d_v_soa, int64_t stride)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        PerturbArgs a = perturb_args(r0, v0, n, strength, seed, obj_r, obj_v);
        const int64_t count = n + a.append;
        arg_check(count == 0 || (d_r_soa && d_v_soa), "null output array");
        arg_check(stride >= count && stride >= 1, "stride must be at least the number of entries");
        if (count == 0)
            return;
        a.r = d_r_soa;
        a.v = d_v_soa;
        a.stride = stride;
        launch_perturb(ctx->stream, a);
        ctx->check_launch();
        ctx->sync();
    });
}

int ogp_set_tuning(ogp_context* ctx, int waves, int64_t chunk_steps)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    if (!(waves == 0 || waves == 1 || waves == 2 || waves == 4 || waves == 8) || chunk_steps < 0)
        return fail(ctx, OGP_ERR_ARGUMENT, "waves must be 0, 1, 2, 4 or 8; chunk_steps >= 0");
    ctx->waves_override = waves;
    ctx->chunk = chunk_steps > 0 ? chunk_steps : 64;
    ctx->tuned_shape.clear();
    return OGP_OK;
}

int ogp_set_screen_shape(ogp_context* ctx, int shape)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    if (shape < 0 || shape > 2)
        return fail(ctx, OGP_ERR_ARGUMENT, "shape must be 0 (tuned), 1 (square) or 2 (narrow)");
    ctx->screen_override = shape;
    ctx->tuned_screen.clear();
    return OGP_OK;
}

int ogp_set_radial_filter(ogp_context* ctx, int on)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    if (on < 0 || on > 2)
        return fail(ctx, OGP_ERR_ARGUMENT, "radial filter must be 0 (off), 1 (automatic) or 2 (always)");
    ctx->radial_filter = on;
    return OGP_OK;
}

int ogp_get_screen_shape(const ogp_context* ctx, int64_t na, int* shape)
{
    if (!ctx || !shape)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null argument");
    *shape = ctx->screen_shape(na);
    return OGP_OK;
}

int ogp_set_lockstep(ogp_context* ctx, int groups)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    if (groups < 0 || groups > 2)
        return fail(ctx, OGP_ERR_ARGUMENT, "groups must be 0 (tuned), 1 or 2");
    ctx->lockstep_override = groups;
    ctx->tuned_shape.clear();
    return OGP_OK;
}

int ogp_get_lockstep(const ogp_context* ctx, int64_t n, int* groups)
{
    if (!ctx || !groups)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null argument");
    *groups = ogp_context::shape_groups(ctx->shape_for(n));
    return OGP_OK;
}

int ogp_set_slices(ogp_context* ctx, int slices)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    if (slices < 0 || slices > 16)
        return fail(ctx, OGP_ERR_ARGUMENT, "slices must be 0 (automatic) .. 16");
    ctx->slices_override = slices;
    return OGP_OK;
}

int ogp_set_fixed_point_exit(ogp_context* ctx, int on)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    if (on != 0 && on != 1)
        return fail(ctx, OGP_ERR_ARGUMENT, "fixed-point exit must be 0 or 1");
    ctx->fp_exit = on != 0;
    return OGP_OK;
}

int ogp_set_derivative_memory(ogp_context* ctx, int64_t bytes)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    if (bytes < 0)
        return fail(ctx, OGP_ERR_ARGUMENT, "bytes must be >= 0 (0 = the default 16 GiB)");
    ctx->record_bytes = bytes > 0 ? (double)bytes : 16.0 * (1 << 30);
    return OGP_OK;
}

int ogp_set_sources(ogp_context* ctx, const ogp_sources* sources)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(sources != nullptr, "null sources");
        arg_check((sources->mask & ~kSourceMask) == 0, "unknown source bit");
        arg_check(std::isfinite(sources->drag_bc) && sources->drag_bc >= 0.0,
            "drag_bc must be finite and >= 0");
        if (sources->mask != 0 && ctx->cfg.force != OGP_FORCE_GRAVITY)
            throw ApiError(OGP_ERR_STATE, "sources are added to the gravity force model only");
        ctx->sources = *sources;
    });
}

int ogp_set_drag_coefficients(ogp_context* ctx, const double* bc, int64_t n)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n >= 0, "negative object count");
        if (!bc || n == 0) {
            ctx->drag_bc_obj.clear();
            return;
        }
        for (int64_t i = 0; i < n; ++i)
            arg_check(std::isfinite(bc[i]) && bc[i] >= 0.0,
                "drag coefficients must be finite and >= 0 (object " + std::to_string(i) + ")");
        ctx->drag_bc_obj.assign(bc, bc + n);
    });
}

int ogp_get_sources(const ogp_context* ctx, ogp_sources* sources)
{
    if (!ctx || !sources)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null argument");
    *sources = ctx->sources;
    return OGP_OK;
}

int ogp_get_tuning(const ogp_context* ctx, int64_t n, int* waves, int64_t* chunk_steps)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    if (waves)
        *waves = ctx->waves_for(n);
    if (chunk_steps)
        *chunk_steps = ctx->chunk;
    return OGP_OK;
}

int ogp_last_timing(const ogp_context* ctx, ogp_timing* t)
{
    if (!ctx || !t)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null argument");
    *t = ctx->timing;
    return OGP_OK;
}

// ---- stepwise propagator --------------------------------------------------------------
} // extern "C"

// GpuPropagator (gpu_propagator.cu:28-154) with its own integrator state. The caller's AoS
// buffers r_aos / v_aos hold the state at the current step (advanced in place); the SoA state
// (r, v, s, S, acceleration ring) and the trajectory rows of the last step live here.
struct ogp_propagator {
    ogp_context* ctx = nullptr; // null once the context has been destroyed
    int64_t n = 0, step = 0, max_steps = 0;
    double h = 0;
    double* r_aos = nullptr;
    double* v_aos = nullptr;
    DevBuf own_r, own_v;                  // device buffers of a host-created propagator
    DevBuf r, v, s, S, hist, ns, traj, bc, acc_tmp;
    const double* bc_ptr = nullptr;
    int64_t last_row = -1; // trajectory row of the last step (r, v, a SoA)

    void release()
    {
        DevBuf* all[] = { &own_r, &own_v, &r, &v, &s, &S, &hist, &ns, &traj, &bc, &acc_tmp };
        for (DevBuf* b : all)
            b->release();
    }
    ~ogp_propagator()
    {
        release();
        if (ctx) {
            auto& v = ctx->props;
            v.erase(std::remove(v.begin(), v.end(), this), v.end());
            if (ctx->sw == this)
                ctx->sw = nullptr;
        }
    }
    PropState state()
    {
        PropState ps;
        ps.r = r.as<double>();
        ps.v = v.as<double>();
        ps.s = s.as<double>();
        ps.S = S.as<double>();
        ps.hist = hist.as<double>();
        ps.nsteps = ns.as<int64_t>();
        ps.N = n;
        return ps;
    }
};

void detach_propagators(ogp_context* ctx)
{
    delete ctx->sw; // the context's own propagator dies with it
    for (ogp_propagator* p : ctx->props) {
        p->release();
        p->ctx = nullptr;
    }
    ctx->props.clear();
}

namespace {

ogp_propagator* new_propagator(ogp_context* ctx, int64_t n, double step_s, int64_t max_steps)
{
    arg_check(n > 0, "GpuPropagator needs at least one object");
    arg_check(std::isfinite(step_s) && step_s > 0, "step size must be positive");
    arg_check(max_steps >= 1, "max_steps must be at least 1");
    auto p = std::make_unique<ogp_propagator>();
    p->n = n;
    p->h = step_s;
    p->max_steps = max_steps;
    int64_t* ctr = &ctx->dev_bytes;
    const size_t v3 = sizeof(double) * 3 * n;
    p->r.reserve(v3, ctr);
    p->v.reserve(v3, ctr);
    p->s.reserve(v3, ctr);
    p->S.reserve(v3, ctr);
    p->hist.reserve(9 * v3, ctr);
    p->ns.reserve(sizeof(int64_t) * n, ctr);
    p->traj.reserve(9 * 4 * v3 / 3, ctr);
    p->acc_tmp.reserve(v3, ctr);
    std::vector<int64_t> ns(n, max_steps);
    HIPC(hipMemcpy(p->ns.p, ns.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
    p->bc_ptr = ctx->stage_bc(p->bc, n, nullptr);
    p->ctx = ctx;
    ctx->props.push_back(p.get());
    return p.release();
}

ogp_context* live_ctx(const ogp_propagator* p)
{
    if (!p->ctx)
        throw ApiError(OGP_ERR_STATE, "the propagator's context has been destroyed");
    return p->ctx;
}

// GpuPropagator::Step (gpu_propagator.cu:28-77)
void propagator_step(ogp_propagator& p, double* host_r, double* host_v, double* d_acc)
{
    ogp_context* ctx = live_ctx(&p);
    const int64_t n = p.n, k = p.step + 1;
    if (k > p.max_steps)
        throw ApiError(OGP_ERR_STATE, "step beyond max_steps of the propagator");
    NodeTable* nt = ctx->node_table(p.h, p.max_steps);
    const int W = ctx->waves_for(n);
    ctx->bc_obj = p.bc_ptr;
    const ForceCtx f = ctx->force_ctx(nt, p.h);
    const PropState ps = p.state();
    double* traj = p.traj.as<double>();
    hipStream_t st = ctx->stream;
    if (k == 1) {
        // GpuStartup::Startup reads the epoch state from the caller's buffers
        launch_aos_to_soa(st, p.r_aos, ps.r, n);
        launch_aos_to_soa(st, p.v_aos, ps.v, n);
        ctx->reserve_scratch(n);
        launch_startup(st, ctx->kind, f, ctx->rk8_variant, ps, ctx->scratch(), 0, n,
            ctx->cfg.n_corrections, traj, 4, W, ctx->fp_exit);
        ctx->check_launch();
    }
    if (k <= 4) {
        p.last_row = k - 1;
    } else {
        launch_chunk(st, ctx->kind, f, ps, 0, n, k, k + 1, ctx->cfg.n_corrections, traj, k, 1, W, 1,
            ctx->fp_exit);
        ctx->check_launch();
        p.last_row = 0;
    }
    const double* row = traj + p.last_row * 9 * n;
    launch_soa_to_aos(st, row, p.r_aos, n);
    launch_soa_to_aos(st, row + 3 * n, p.v_aos, n);
    if (d_acc)
        launch_soa_to_aos(st, row + 6 * n, d_acc, n);
    ctx->check_launch();
    ctx->sync();
    p.step = k;
    if (host_r)
        HIPC(hipMemcpy(host_r, p.r_aos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    if (host_v)
        HIPC(hipMemcpy(host_v, p.v_aos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
}

void propagator_acceleration(const ogp_propagator& p, double* host_a)
{
    ogp_context* ctx = live_ctx(&p);
    if (p.last_row < 0)
        throw ApiError(OGP_ERR_STATE, "no step has been taken");
    const int64_t n = p.n;
    launch_soa_to_aos(ctx->stream, p.traj.as<double>() + (p.last_row * 9 + 6) * n,
        p.acc_tmp.as<double>(), n);
    ctx->check_launch();
    ctx->sync();
    HIPC(hipMemcpy(host_a, p.acc_tmp.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
}

int fail_prop(const ogp_propagator* p, int code, const std::string& msg)
{
    return fail(p ? p->ctx : nullptr, code, msg);
}

template <class F> int guarded_prop(ogp_propagator* p, F&& fn)
{
    if (!p)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null propagator");
    try {
        fn();
        return OGP_OK;
    } catch (const ApiError& e) {
        return fail_prop(p, e.code, e.what());
    } catch (const std::exception& e) {
        return fail_prop(p, OGP_ERR_DEVICE, e.what());
    }
}

} // namespace

extern "C" {

int ogp_propagator_create(ogp_context* ctx, double* d_r, double* d_v, int64_t n, double step_s,
    int64_t max_steps, ogp_propagator** out)
{
    if (!ctx || !out)
        return fail(ctx, OGP_ERR_ARGUMENT, "null context or output pointer");
    *out = nullptr;
    return guarded(ctx, [&] {
        arg_check(d_r && d_v, "null state buffer");
        ogp_propagator* p = new_propagator(ctx, n, step_s, max_steps);
        p->r_aos = d_r;
        p->v_aos = d_v;
        *out = p;
    });
}

int ogp_propagator_create_host(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    double step_s, int64_t max_steps, ogp_propagator** out)
{
    if (!ctx || !out)
        return fail(ctx, OGP_ERR_ARGUMENT, "null context or output pointer");
    *out = nullptr;
    return guarded(ctx, [&] {
        arg_check(r0 && v0, "null state array");
        arg_check(n > 0, "GpuPropagator needs at least one object");
        check_finite(r0, 3 * n, "positions");
        check_finite(v0, 3 * n, "velocities");
        std::unique_ptr<ogp_propagator> p(new_propagator(ctx, n, step_s, max_steps));
        p->own_r.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
        p->own_v.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
        HIPC(hipMemcpy(p->own_r.p, r0, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(p->own_v.p, v0, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        p->r_aos = p->own_r.as<double>();
        p->v_aos = p->own_v.as<double>();
        *out = p.release();
    });
}

int ogp_propagator_step(ogp_propagator* p, double* host_r, double* host_v, double* d_acc)
{
    return guarded_prop(p, [&] { propagator_step(*p, host_r, host_v, d_acc); });
}

int ogp_propagator_acceleration(const ogp_propagator* p, double* host_a)
{
    return guarded_prop(const_cast<ogp_propagator*>(p), [&] {
        arg_check(host_a != nullptr, "null output array");
        propagator_acceleration(*p, host_a);
    });
}

int ogp_propagator_copy_state_to_host(const ogp_propagator* p, double* host_r, double* host_v)
{
    return guarded_prop(const_cast<ogp_propagator*>(p), [&] {
        live_ctx(p);
        if (host_r)
            HIPC(hipMemcpy(host_r, p->r_aos, sizeof(double) * 3 * p->n, hipMemcpyDeviceToHost));
        if (host_v)
            HIPC(hipMemcpy(host_v, p->v_aos, sizeof(double) * 3 * p->n, hipMemcpyDeviceToHost));
    });
}

void ogp_propagator_destroy(ogp_propagator* p) { delete p; }

int ogp_prop_init(ogp_context* ctx, const double* r0, const double* v0, int64_t n, double step_s,
    int64_t max_steps)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    ogp_propagator* p = nullptr;
    const int rc = ogp_propagator_create_host(ctx, r0, v0, n, step_s, max_steps, &p);
    if (rc != OGP_OK)
        return rc;
    delete ctx->sw;
    ctx->sw = p;
    return OGP_OK;
}

int ogp_prop_step(ogp_context* ctx, double* r_out, double* v_out, double* a_out)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    if (!ctx->sw)
        return fail(ctx, OGP_ERR_STATE, "ogp_prop_step before ogp_prop_init");
    int rc = ogp_propagator_step(ctx->sw, r_out, v_out, nullptr);
    if (rc == OGP_OK && a_out)
        rc = ogp_propagator_acceleration(ctx->sw, a_out);
    return rc;
}

int ogp_startup(ogp_context* ctx, const double* r0, const double* v0, int64_t n, double step_s,
    double* pos, double* vel, double* s, double* S, double* acc)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n > 0 && r0 && v0 && step_s > 0, "invalid startup arguments");
        check_finite(r0, 3 * n, "positions");
        check_finite(v0, 3 * n, "velocities");
        ctx->reserve_state(n);
        ctx->io_a.reserve(sizeof(double) * 27 * n, &ctx->dev_bytes);
        HIPC(hipMemcpyAsync(ctx->io_a.p, r0, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
            ctx->stream));
        launch_aos_to_soa(ctx->stream, ctx->io_a.as<double>(), ctx->st_r.as<double>(), n);
        ctx->sync();
        HIPC(hipMemcpyAsync(ctx->io_a.p, v0, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
            ctx->stream));
        launch_aos_to_soa(ctx->stream, ctx->io_a.as<double>(), ctx->st_v.as<double>(), n);
        std::vector<int64_t> ns(n, 4);
        HIPC(hipMemcpyAsync(ctx->st_ns.p, ns.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
            ctx->stream));
        NodeTable* nt = ctx->node_table(step_s, 4);
        const int W = ctx->waves_for(n);
        ctx->bc_obj = ctx->stage_bc(ctx->d_bc, n, nullptr);
        launch_startup(ctx->stream, ctx->kind, ctx->force_ctx(nt, step_s), ctx->rk8_variant,
            ctx->prop_state(n), ctx->scratch(), 0, n, ctx->cfg.n_corrections, nullptr, 0, W,
            ctx->fp_exit);
        ctx->check_launch();
        ctx->sync();
        // scratch [slot][c][obj] -> caller [obj][slot][c]
        double* outs[5] = { pos, vel, acc, s, S };
        std::vector<double> host(27 * n);
        for (int q = 0; q < 5; ++q) {
            if (!outs[q])
                continue;
            HIPC(hipMemcpy(host.data(), ctx->sc[q].p, sizeof(double) * 27 * n, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i)
                for (int sl = 0; sl < 9; ++sl)
                    for (int c = 0; c < 3; ++c)
                        outs[q][27 * i + 3 * sl + c] = host[(sl * 3 + c) * n + i];
        }
    });
}

int ogp_startup_device(ogp_context* ctx, const double* d_r0, const double* d_v0, int64_t n,
    double step_s, double* d_pos, double* d_vel, double* d_s, double* d_S, double* d_acc)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n > 0 && d_r0 && d_v0 && std::isfinite(step_s) && step_s > 0,
            "invalid startup arguments");
        ctx->reserve_state(n);
        ctx->d_flags.reserve(sizeof(int32_t) * n, &ctx->dev_bytes);
        ctx->d_count.reserve(sizeof(uint32_t), &ctx->dev_bytes);
        launch_aos_to_soa(ctx->stream, d_r0, ctx->st_r.as<double>(), n);
        launch_aos_to_soa(ctx->stream, d_v0, ctx->st_v.as<double>(), n);
        const uint32_t zero = 0;
        uint32_t bad = 0;
        HIPC(hipMemcpyAsync(ctx->d_count.p, &zero, sizeof(zero), hipMemcpyHostToDevice,
            ctx->stream));
        launch_finite_soa(ctx->stream, ctx->st_r.as<double>(), ctx->st_v.as<double>(), n,
            ctx->d_flags.as<int32_t>(), ctx->d_count.as<uint32_t>(), 1, true);
        ctx->check_launch();
        HIPC(hipMemcpyAsync(&bad, ctx->d_count.p, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        if (bad > 0)
            throw ApiError(OGP_ERR_NONFINITE, "non-finite value in the initial state of "
                    + std::to_string(bad) + " object(s)");
        std::vector<int64_t> ns(n, 4);
        HIPC(hipMemcpyAsync(ctx->st_ns.p, ns.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
            ctx->stream));
        NodeTable* nt = ctx->node_table(step_s, 4);
        const int W = ctx->waves_for(n);
        ctx->bc_obj = ctx->stage_bc(ctx->d_bc, n, nullptr);
        launch_startup(ctx->stream, ctx->kind, ctx->force_ctx(nt, step_s), ctx->rk8_variant,
            ctx->prop_state(n), ctx->scratch(), 0, n, ctx->cfg.n_corrections, nullptr, 0, W,
            ctx->fp_exit);
        // scratch [slot][c][obj] -> caller [obj][slot][c]
        double* outs[5] = { d_pos, d_vel, d_acc, d_s, d_S };
        for (int q = 0; q < 5; ++q)
            if (outs[q])
                launch_slots_to_aos(ctx->stream, ctx->sc[q].as<double>(), outs[q], n, 9);
        ctx->check_launch();
        ctx->sync();
    });
}

int ogp_predict_device(ogp_context* ctx, double* d_r, double* d_v, double* d_acc,
    const double* d_s, double* d_S, int64_t n, double step_s, int64_t step_num)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n >= 0 && std::isfinite(step_s) && step_s > 0 && step_num >= 0,
            "invalid predictor arguments");
        arg_check(n == 0 || (d_r && d_v && d_acc && d_s && d_S), "null array");
        if (ctx->cfg.force != OGP_FORCE_GRAVITY)
            throw ApiError(OGP_ERR_STATE, "context was created without the gravity field");
        if (n == 0)
            return;
        ctx->bc_obj = ctx->stage_bc(ctx->d_bc, n, nullptr);
        NodeTable* nt = ctx->node_table(step_s, step_num);
        const int64_t node = step_num + 4;
        launch_predict_aos(ctx->stream, d_acc, d_s, d_S, n, step_s, d_r, d_v);
        launch_gravity_aos(ctx->stream, ctx->kind == kGravitySplit, ctx->gdev,
            nt->prop_M.as<double>() + 9 * node, d_r, n, d_acc + 24, 27, 64, ctx->sources_dev(),
            nt->prop_tb.as<double>() + 6 * node, d_v);
        ctx->check_launch();
        ctx->sync();
    });
}

int ogp_kat_predict(ogp_context* ctx, const double* acc27, const double* s3, double* S3,
    int64_t n, double h, double* r3, double* v3)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n > 0, "n must be positive");
        DevBuf a, s, S, r, v;
        a.reserve(sizeof(double) * 27 * n, nullptr);
        s.reserve(sizeof(double) * 3 * n, nullptr);
        S.reserve(sizeof(double) * 3 * n, nullptr);
        r.reserve(sizeof(double) * 3 * n, nullptr);
        v.reserve(sizeof(double) * 3 * n, nullptr);
        HIPC(hipMemcpy(a.p, acc27, sizeof(double) * 27 * n, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(s.p, s3, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(S.p, S3, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        launch_kat_predict(ctx->stream, a.as<double>(), s.as<double>(), S.as<double>(), n, h,
            r.as<double>(), v.as<double>());
        ctx->check_launch();
        ctx->sync();
        HIPC(hipMemcpy(S3, S.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(r3, r.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(v3, v.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    });
}

int ogp_kat_finalize_correction(ogp_context* ctx, double* r3, double* v3, const double* acc27,
    const double* s3, const double* S3, int64_t n, double h)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        arg_check(n > 0, "n must be positive");
        DevBuf a, s, S, r, v;
        a.reserve(sizeof(double) * 27 * n, nullptr);
        s.reserve(sizeof(double) * 3 * n, nullptr);
        S.reserve(sizeof(double) * 3 * n, nullptr);
        r.reserve(sizeof(double) * 3 * n, nullptr);
        v.reserve(sizeof(double) * 3 * n, nullptr);
        HIPC(hipMemcpy(a.p, acc27, sizeof(double) * 27 * n, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(s.p, s3, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(S.p, S3, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(r.p, r3, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(v.p, v3, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        launch_kat_finalize(ctx->stream, r.as<double>(), v.as<double>(), a.as<double>(),
            s.as<double>(), S.as<double>(), n, h);
        ctx->check_launch();
        ctx->sync();
        HIPC(hipMemcpy(r3, r.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(v3, v.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    });
}

// ---- conjunction ------------------------------------------------------------------------
int ogp_conj_init(ogp_context* ctx, const double* r, const double* v, const uint32_t* set_ids,
    int64_t n, double step_s)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        // closest_approach.cu:50-81 and partition_parameters.cuh:94-135. A context holds one
        // search; initialising again replaces it (objects, trajectory and results).
        ctx->cj_n = 0;
        ctx->cj_steps = 0;
        ctx->res_a.clear();
        ctx->res_b.clear();
        ctx->res_k.clear();
        ctx->res_out.clear();
        ctx->cj_timing = ogp_conj_timing {};
        arg_check(n > 0 && r && v && set_ids, "positions array is empty");
        arg_check(std::isfinite(step_s) && step_s > 0, "step size must be positive");
        arg_check(set_ids[0] == 0, "set ids must start at 0");
        arg_check(set_ids[0] != set_ids[n - 1], "invalid setIndices: at least 2 sets are required");
        std::vector<int64_t> off { 0 };
        for (int64_t i = 1; i < n; ++i) {
            arg_check(set_ids[i] == set_ids[i - 1] || set_ids[i] == set_ids[i - 1] + 1,
                "set ids must be non-decreasing in steps of 1");
            if (set_ids[i] != set_ids[i - 1])
                off.push_back(i);
        }
        check_finite(r, 3 * n, "positions");
        check_finite(v, 3 * n, "velocities");
        ctx->cj_sets.assign(set_ids, set_ids + n);
        ctx->cj_set_off = off;
        ctx->cj_r.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
        ctx->cj_v.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
        ctx->io_a.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
        HIPC(hipMemcpyAsync(ctx->io_a.p, r, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
            ctx->stream));
        launch_aos_to_soa(ctx->stream, ctx->io_a.as<double>(), ctx->cj_r.as<double>(), n);
        ctx->sync();
        HIPC(hipMemcpyAsync(ctx->io_a.p, v, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
            ctx->stream));
        launch_aos_to_soa(ctx->stream, ctx->io_a.as<double>(), ctx->cj_v.as<double>(), n);
        ctx->check_launch();
        ctx->sync();
        ctx->cj_bc_ptr = ctx->stage_bc(ctx->cj_bc, n, nullptr);
        ctx->cj_n = n;
        ctx->cj_h = step_s;
        ctx->cj_steps = 0;
        // capacity of the existing trajectory store, in steps of this object count
        ctx->cj_cap_steps = (int64_t)(ctx->cj_traj.bytes / (sizeof(double) * 9 * (size_t)n));
    });
}

int ogp_conj_propagate(ogp_context* ctx, int64_t n_steps)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        if (ctx->cj_n <= 0)
            throw ApiError(OGP_ERR_STATE, "ClosestApproach::Propagate before Initialize");
        arg_check(n_steps > 0, "numSteps must be positive");
        const int64_t n = ctx->cj_n;
        const int64_t total = ctx->cj_steps + n_steps;
        // grow the trajectory store, keeping the samples already taken
        if (total > ctx->cj_cap_steps) {
            const int64_t cap = std::max<int64_t>(total, ctx->cj_cap_steps * 2);
            DevBuf nb;
            nb.reserve(sizeof(double) * 9 * n * cap, &ctx->dev_bytes);
            if (ctx->cj_steps > 0)
                HIPC(hipMemcpyAsync(nb.p, ctx->cj_traj.p, sizeof(double) * 9 * n * ctx->cj_steps,
                    hipMemcpyDeviceToDevice, ctx->stream));
            ctx->sync();
            std::swap(ctx->cj_traj.p, nb.p);
            std::swap(ctx->cj_traj.bytes, nb.bytes);
            std::swap(ctx->cj_traj.counter, nb.counter);
            ctx->cj_cap_steps = cap;
        }
        // every phase restarts the startup from the current state (a new GpuPropagator per
        // Propagate call, closest_approach.cu:134-135); its time origin continues from the
        // previous phase
        ctx->reserve_state(n);
        HIPC(hipMemcpyAsync(ctx->st_r.p, ctx->cj_r.p, sizeof(double) * 3 * n,
            hipMemcpyDeviceToDevice, ctx->stream));
        HIPC(hipMemcpyAsync(ctx->st_v.p, ctx->cj_v.p, sizeof(double) * 3 * n,
            hipMemcpyDeviceToDevice, ctx->stream));
        std::vector<int64_t> ns(n, n_steps);
        HIPC(hipMemcpyAsync(ctx->st_ns.p, ns.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
            ctx->stream));
        ctx->sync();
        const double t0 = (double)ctx->cj_steps * ctx->cj_h;
        std::vector<ogp_context::Group> g { { ctx->cj_h, 0, n, n_steps } };
        double* traj = ctx->cj_traj.as<double>() + 9 * n * ctx->cj_steps;
        ctx->bc_obj = ctx->cj_bc_ptr;
        ctx->run_groups(g, n, t0, traj, n_steps);
        HIPC(hipMemcpyAsync(ctx->cj_r.p, ctx->st_r.p, sizeof(double) * 3 * n,
            hipMemcpyDeviceToDevice, ctx->stream));
        HIPC(hipMemcpyAsync(ctx->cj_v.p, ctx->st_v.p, sizeof(double) * 3 * n,
            hipMemcpyDeviceToDevice, ctx->stream));
        ctx->sync();
        // a later phase starts at another time origin and never reuses this phase's table:
        // drop it, so that a loop of short phases does not grow the cache (the epoch's own
        // table, t0 = 0, is shared with the other propagation calls and kept)
        if (t0 != 0.0)
            ctx->nodes.erase(std::make_pair(ctx->cj_h, t0));
        ctx->cj_steps = total;
        ctx->cj_timing.propagate_ms = ctx->timing.total_ms;
    });
}

int ogp_conj_num_pairs(const ogp_context* ctx, int64_t* num_pairs)
{
    if (!ctx || !num_pairs)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null argument");
    int64_t np = 0;
    const auto& off = ctx->cj_set_off;
    for (size_t s = 0; s + 1 < off.size(); ++s)
        np += (off[s + 1] - off[s]) * (ctx->cj_n - off[s + 1]);
    *num_pairs = np;
    return OGP_OK;
}

} // extern "C"

namespace {

// The screen tuner: the first time the context meets a set A of this size bucket (at most 64
// objects, beside at least 1,024 B objects), it screens A against a prefix of set B (up to 16,384
// objects) with each tile shape, writing nothing, times both on the context stream and keeps the
// faster. The shapes give the same bits, so the choice only changes the speed.
void tune_screen(ogp_context* ctx, int64_t a0, int64_t na, int64_t b0, int64_t nb)
{
    if (ctx->screen_override > 0 || na > 64 || nb < 1024
        || ctx->tuned_screen.count(ogp_context::screen_bucket(na)))
        return;
    const int64_t pb = std::min<int64_t>(nb, 16384);
    const int shapes[2] = { kScreenSquare, kScreenNarrow };
    float ms[2] = { 0, 0 };
    for (int q = 0; q < 2; ++q) {
        Event e0, e1;
        HIPC(hipEventRecord(e0.e, ctx->stream));
        launch_screen(ctx->stream, ctx->cj_traj.as<double>(), ctx->cj_n, ctx->cj_steps, a0, na, b0,
            pb, 0, 0.0, nullptr, nullptr, ctx->cj_cnt.as<unsigned long long>(), 0, nullptr,
            nullptr, nullptr, nullptr, shapes[q]);
        ctx->check_launch();
        HIPC(hipEventRecord(e1.e, ctx->stream));
        HIPC(hipEventSynchronize(e1.e));
        HIPC(hipEventElapsedTime(&ms[q], e0.e, e1.e));
    }
    ctx->tuned_screen[ogp_context::screen_bucket(na)] = ms[1] < ms[0] ? kScreenNarrow : kScreenSquare;
}

// Radially sorted input of a threshold screen (ScreenSorted): every object's smallest and
// largest sampled radius, the objects of each set ordered by the lower radius in 100 km bins
// and by the upper radius within a bin (so that a 64-object tile spans a narrow range of both),
// and the positions gathered in that order. slo / shi receive the bounds in slot order.
ScreenSorted sort_for_screen(ogp_context* ctx, std::vector<double>& slo, std::vector<double>& shi)
{
    const int64_t n = ctx->cj_n, nsteps = ctx->cj_steps;
    ctx->cj_rpart.reserve(sizeof(double) * 2 * kRadialChunks * n, &ctx->dev_bytes);
    ctx->cj_rlo.reserve(sizeof(double) * n, &ctx->dev_bytes);
    ctx->cj_rhi.reserve(sizeof(double) * n, &ctx->dev_bytes);
    ctx->cj_slo.reserve(sizeof(double) * n, &ctx->dev_bytes);
    ctx->cj_shi.reserve(sizeof(double) * n, &ctx->dev_bytes);
    ctx->cj_ids.reserve(sizeof(uint32_t) * n, &ctx->dev_bytes);
    ctx->cj_sorted.reserve(sizeof(double) * 3 * n * nsteps, &ctx->dev_bytes);
    launch_radial_bounds(ctx->stream, ctx->cj_traj.as<double>(), n, nsteps,
        ctx->cj_rpart.as<double>(), ctx->cj_rlo.as<double>(), ctx->cj_rhi.as<double>());
    ctx->check_launch();
    std::vector<double> lo(n), hi(n);
    HIPC(hipMemcpyAsync(lo.data(), ctx->cj_rlo.p, sizeof(double) * n, hipMemcpyDeviceToHost,
        ctx->stream));
    HIPC(hipMemcpyAsync(hi.data(), ctx->cj_rhi.p, sizeof(double) * n, hipMemcpyDeviceToHost,
        ctx->stream));
    ctx->sync();
    std::vector<uint32_t> ids(n);
    std::iota(ids.begin(), ids.end(), 0u);
    std::vector<int64_t> bounds = ctx->cj_set_off;
    bounds.push_back(n);
    for (size_t s = 0; s + 1 < bounds.size(); ++s)
        std::sort(ids.begin() + bounds[s], ids.begin() + bounds[s + 1], [&](uint32_t x, uint32_t y) {
            const double bx = std::floor(lo[x] / 1e5), by = std::floor(lo[y] / 1e5);
            if (bx != by)
                return bx < by;
            return hi[x] != hi[y] ? hi[x] < hi[y] : x < y;
        });
    slo.resize(n);
    shi.resize(n);
    for (int64_t s = 0; s < n; ++s) {
        slo[s] = lo[ids[s]];
        shi[s] = hi[ids[s]];
    }
    HIPC(hipMemcpyAsync(ctx->cj_ids.p, ids.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice,
        ctx->stream));
    launch_gather_sorted(ctx->stream, ctx->cj_traj.as<double>(), n, nsteps,
        ctx->cj_ids.as<uint32_t>(), ctx->cj_rlo.as<double>(), ctx->cj_rhi.as<double>(),
        ctx->cj_sorted.as<double>(), ctx->cj_slo.as<double>(), ctx->cj_shi.as<double>());
    ctx->check_launch();
    ctx->sync(); // ids leaves host memory
    ScreenSorted srt;
    srt.pos = ctx->cj_sorted.as<double>();
    srt.rlo = ctx->cj_slo.as<double>();
    srt.rhi = ctx->cj_shi.as<double>();
    srt.ids = ctx->cj_ids.as<uint32_t>();
    return srt;
}

// The pairs of rectangle A x B that the screen kernels of this shape evaluate on sorted input:
// the tiles the kernels do not skip (the same tiles, bounds and comparison as the kernels).
double screened_pairs(const std::vector<double>& slo, const std::vector<double>& shi, int64_t a0,
    int64_t na, int64_t b0, int64_t nb, int shape, double threshold)
{
    const int64_t ta = shape == kScreenNarrow ? 4 : 64, tb = shape == kScreenNarrow ? 256 : 64;
    auto range = [&](int64_t first, int64_t count, double& lo, double& hi) {
        lo = std::numeric_limits<double>::infinity();
        hi = -lo;
        for (int64_t s = first; s < first + count; ++s) {
            lo = std::fmin(lo, slo[s]);
            hi = std::fmax(hi, shi[s]);
        }
    };
    std::vector<double> blo, bhi;
    for (int64_t j = 0; j < nb; j += tb) {
        double lo, hi;
        range(b0 + j, std::min(tb, nb - j), lo, hi);
        blo.push_back(lo);
        bhi.push_back(hi);
    }
    double pairs = 0;
    for (int64_t i = 0; i < na; i += ta) {
        const int64_t ca = std::min(ta, na - i);
        double alo, ahi;
        range(a0 + i, ca, alo, ahi);
        for (size_t t = 0; t < blo.size(); ++t)
            if (!(std::fmax(alo - bhi[t], blo[t] - ahi) > threshold + kRadialMargin))
                pairs += (double)ca * (double)std::min(tb, nb - (int64_t)t * tb);
    }
    return pairs;
}

// Whether sorting for the radial filter pays: the sort reads the trajectory twice and writes
// the positions once (about 100 B per object and step, plus a host sort), while the screen
// costs about 0.3 ns per pair and step and the filter skipped 45% of the pairs of 20,000 LEO
// objects. It pays with hundreds of pairs per object: 36 M pairs of 20,000 objects over 4,320
// steps screen in 33 instead of 47 ms; one object against 19,999 took 6.0 ms sorted against
// 1.3 ms unsorted.
bool radial_filter_pays(int64_t pairs, int64_t n, int64_t nsteps)
{
    return pairs >= 256 * n && (double)pairs * (double)nsteps >= 2e10;
}

// Screens every set rectangle; all_pairs writes the per-pair minima, threshold collects
// candidates. Returns the number of candidates.
int64_t conj_screen(ogp_context* ctx, double threshold, bool all_pairs)
{
    const int64_t n = ctx->cj_n, nsteps = ctx->cj_steps;
    int64_t np = 0;
    ogp_conj_num_pairs(ctx, &np);
    const auto& off = ctx->cj_set_off;
    if (all_pairs) {
        ctx->cj_mind.reserve(sizeof(double) * std::max<int64_t>(np, 1), &ctx->dev_bytes);
        ctx->cj_mink.reserve(sizeof(int32_t) * std::max<int64_t>(np, 1), &ctx->dev_bytes);
    }
    int64_t cap = std::max<int64_t>(1 << 16, ctx->cj_a.bytes / sizeof(uint32_t));
    // threshold screens skip the tiles of radially separated objects (ogp_set_radial_filter)
    ScreenSorted srt;
    std::vector<double> slo, shi;
    if (!all_pairs && threshold > 0.0
        && (ctx->radial_filter == 2
            || (ctx->radial_filter == 1 && radial_filter_pays(np, n, nsteps))))
        srt = sort_for_screen(ctx, slo, shi);
    for (int attempt = 0; attempt < 2; ++attempt) {
        ctx->cj_cnt.reserve(sizeof(uint32_t), &ctx->dev_bytes);
        ctx->cj_a.reserve(sizeof(uint32_t) * cap, &ctx->dev_bytes);
        ctx->cj_b.reserve(sizeof(uint32_t) * cap, &ctx->dev_bytes);
        ctx->cj_k.reserve(sizeof(int32_t) * cap, &ctx->dev_bytes);
        ctx->cj_d.reserve(sizeof(double) * cap, &ctx->dev_bytes);
        const unsigned long long zero = 0;
        HIPC(hipMemcpyAsync(ctx->cj_cnt.p, &zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream));
        int64_t pair0 = 0;
        ctx->cj_screened = 0;
        for (size_t s = 0; s + 1 < off.size(); ++s) {
            const int64_t a0 = off[s], na = off[s + 1] - off[s];
            const int64_t b0 = off[s + 1], nb = n - off[s + 1];
            tune_screen(ctx, a0, na, b0, nb);
            const int shape = ctx->screen_shape(na);
            launch_screen(ctx->stream, ctx->cj_traj.as<double>(), n, nsteps, a0, na, b0, nb, pair0,
                all_pairs ? 0.0 : threshold, all_pairs ? ctx->cj_mind.as<double>() : nullptr,
                all_pairs ? ctx->cj_mink.as<int32_t>() : nullptr,
                ctx->cj_cnt.as<unsigned long long>(), cap,
                ctx->cj_a.as<uint32_t>(), ctx->cj_b.as<uint32_t>(), ctx->cj_k.as<int32_t>(),
                ctx->cj_d.as<double>(), shape, 0, srt);
            ctx->check_launch();
            ctx->cj_screened += srt.pos ? screened_pairs(slo, shi, a0, na, b0, nb, shape, threshold)
                                        : (double)na * (double)nb;
            pair0 += na * nb;
        }
        unsigned long long cnt = 0;
        HIPC(hipMemcpyAsync(&cnt, ctx->cj_cnt.p, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        if (all_pairs)
            return np;
        if (cnt <= (unsigned long long)cap)
            return (int64_t)cnt;
        cap = (int64_t)cnt; // too many candidates for the buffer: grow once and rescreen
    }
    throw ApiError(OGP_ERR_DEVICE, "candidate buffer overflow");
}

} // namespace

extern "C" {

int ogp_conj_screen(ogp_context* ctx, double* min_dist, int32_t* min_step)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        if (ctx->cj_steps <= 0)
            throw ApiError(OGP_ERR_STATE, "screen before Propagate");
        const int64_t np = conj_screen(ctx, 0.0, true);
        HIPC(hipMemcpy(min_dist, ctx->cj_mind.p, sizeof(double) * np, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(min_step, ctx->cj_mink.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
    });
}

int ogp_conj_find(ogp_context* ctx, double threshold_m, int64_t* count)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        if (ctx->cj_steps <= 0)
            throw ApiError(OGP_ERR_STATE, "FindClosest before Propagate");
        arg_check(count != nullptr, "null count");
        const int64_t n = ctx->cj_n;
        Event e0, e1, e2;
        HIPC(hipEventRecord(e0.e, ctx->stream));
        int64_t nc;
        if (threshold_m > 0) {
            nc = conj_screen(ctx, threshold_m, false);
        } else {
            // every cross-set pair becomes a candidate (reference: all pairs refined)
            const int64_t np = conj_screen(ctx, 0.0, true);
            std::vector<double> d(np);
            std::vector<int32_t> k(np);
            HIPC(hipMemcpy(d.data(), ctx->cj_mind.p, sizeof(double) * np, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(k.data(), ctx->cj_mink.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
            std::vector<uint32_t> a(np), b(np);
            int64_t p = 0;
            const auto& off = ctx->cj_set_off;
            for (size_t s = 0; s + 1 < off.size(); ++s)
                for (int64_t i = off[s]; i < off[s + 1]; ++i)
                    for (int64_t j = off[s + 1]; j < n; ++j, ++p) {
                        a[p] = (uint32_t)i;
                        b[p] = (uint32_t)j;
                    }
            ctx->cj_a.reserve(sizeof(uint32_t) * np, &ctx->dev_bytes);
            ctx->cj_b.reserve(sizeof(uint32_t) * np, &ctx->dev_bytes);
            ctx->cj_k.reserve(sizeof(int32_t) * np, &ctx->dev_bytes);
            ctx->cj_d.reserve(sizeof(double) * np, &ctx->dev_bytes);
            HIPC(hipMemcpy(ctx->cj_a.p, a.data(), sizeof(uint32_t) * np, hipMemcpyHostToDevice));
            HIPC(hipMemcpy(ctx->cj_b.p, b.data(), sizeof(uint32_t) * np, hipMemcpyHostToDevice));
            HIPC(hipMemcpy(ctx->cj_k.p, k.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice));
            HIPC(hipMemcpy(ctx->cj_d.p, d.data(), sizeof(double) * np, hipMemcpyHostToDevice));
            nc = np;
        }
        HIPC(hipEventRecord(e1.e, ctx->stream));
        ctx->cj_out.reserve(sizeof(double) * 8 * std::max<int64_t>(nc, 1), &ctx->dev_bytes);
        launch_refine(ctx->stream, ctx->cj_traj.as<double>(), n, ctx->cj_steps, ctx->cj_h, 0.0, nc,
            ctx->cj_a.as<uint32_t>(), ctx->cj_b.as<uint32_t>(), ctx->cj_k.as<int32_t>(),
            ctx->cj_d.as<double>(), ctx->cj_out.as<double>());
        ctx->check_launch();
        HIPC(hipEventRecord(e2.e, ctx->stream));
        ctx->sync();
        ctx->res_a.resize(nc);
        ctx->res_b.resize(nc);
        ctx->res_k.resize(nc);
        ctx->res_out.resize(8 * nc);
        if (nc > 0) {
            HIPC(hipMemcpy(ctx->res_a.data(), ctx->cj_a.p, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(ctx->res_b.data(), ctx->cj_b.p, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(ctx->res_k.data(), ctx->cj_k.p, sizeof(int32_t) * nc, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(ctx->res_out.data(), ctx->cj_out.p, sizeof(double) * 8 * nc,
                hipMemcpyDeviceToHost));
        }
        // order by pair (a, b): the candidate list of the threshold screen is filled by
        // atomics and has no fixed order
        std::vector<int64_t> ord(nc);
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) {
            return ctx->res_a[x] != ctx->res_a[y] ? ctx->res_a[x] < ctx->res_a[y]
                                                  : ctx->res_b[x] < ctx->res_b[y];
        });
        auto permute = [&](auto& vec, int w) {
            auto cp = vec;
            for (int64_t q = 0; q < nc; ++q)
                for (int j = 0; j < w; ++j)
                    vec[q * w + j] = cp[ord[q] * w + j];
        };
        permute(ctx->res_a, 1);
        permute(ctx->res_b, 1);
        permute(ctx->res_k, 1);
        permute(ctx->res_out, 8);
        float s1 = 0, s2 = 0;
        HIPC(hipEventElapsedTime(&s1, e0.e, e1.e));
        HIPC(hipEventElapsedTime(&s2, e1.e, e2.e));
        ctx->cj_timing.screen_ms = s1;
        ctx->cj_timing.refine_ms = s2;
        int64_t np = 0;
        ogp_conj_num_pairs(ctx, &np);
        ctx->cj_timing.pairs = (double)np;
        ctx->cj_timing.pair_steps = (double)np * (double)ctx->cj_steps;
        ctx->cj_timing.candidates = (double)nc;
        ctx->cj_timing.screened_pair_steps = ctx->cj_screened * (double)ctx->cj_steps;
        *count = nc;
    });
}

int ogp_conj_find_noalloc(ogp_context* ctx, double threshold_m, int64_t capacity, uint32_t* a,
    uint32_t* b, int32_t* step, double* distance, double* time, double* pos_a, double* pos_b,
    int64_t* count)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        if (ctx->cj_steps <= 0)
            throw ApiError(OGP_ERR_STATE, "FindClosest before Propagate");
        arg_check(count != nullptr, "null count");
        arg_check(std::isfinite(threshold_m) && threshold_m > 0, "threshold must be positive");
        arg_check(capacity >= 0, "negative capacity");
        arg_check(capacity == 0 || (a && b && step && distance && time && pos_a && pos_b),
            "null output array");
        if (!ctx->noalloc)
            throw ApiError(OGP_ERR_DEVICE, "scratch array not set up on this device");
        char* base = static_cast<char*>(ctx->noalloc);
        auto* counter = reinterpret_cast<unsigned long long*>(base + kNoAllocCounter);
        auto* ca = reinterpret_cast<uint32_t*>(base + kNoAllocA);
        auto* cb = reinterpret_cast<uint32_t*>(base + kNoAllocB);
        auto* ck = reinterpret_cast<int32_t*>(base + kNoAllocK);
        auto* cd = reinterpret_cast<double*>(base + kNoAllocD);
        auto* out = reinterpret_cast<double*>(base + kNoAllocOut);
        const int64_t n = ctx->cj_n, nsteps = ctx->cj_steps;
        const auto& off = ctx->cj_set_off;
        launch_zero_counter(ctx->stream, counter);
        int64_t pair0 = 0;
        for (size_t s = 0; s + 1 < off.size(); ++s) {
            const int64_t a0 = off[s], na = off[s + 1] - off[s];
            const int64_t b0 = off[s + 1], nb = n - off[s + 1];
            launch_screen(ctx->stream, ctx->cj_traj.as<double>(), n, nsteps, a0, na, b0, nb, pair0,
                threshold_m, nullptr, nullptr, counter, kNoAllocMaxCandidates, ca, cb, ck, cd,
                ctx->screen_shape(na));
            pair0 += na * nb;
        }
        ctx->check_launch();
        unsigned long long cnt = 0;
        HIPC(hipMemcpyAsync(&cnt, counter, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        *count = (int64_t)cnt;
        if (cnt > (unsigned long long)kNoAllocMaxCandidates)
            throw ApiError(OGP_ERR_ARGUMENT, std::to_string(cnt) + " candidates exceed the "
                    + std::to_string(kNoAllocMaxCandidates) + " of the allocation-free search "
                    "(lower the threshold, or use ogp_conj_find)");
        if ((int64_t)cnt > capacity)
            throw ApiError(OGP_ERR_ARGUMENT, std::to_string(cnt) + " candidates exceed the "
                    "capacity of the output arrays (" + std::to_string(capacity) + ")");
        const int64_t nc = (int64_t)cnt;
        if (nc == 0)
            return;
        launch_refine(ctx->stream, ctx->cj_traj.as<double>(), n, nsteps, ctx->cj_h, 0.0, nc, ca,
            cb, ck, cd, out, kNoAllocMaxCandidates);
        ctx->check_launch();
        const int64_t S = kNoAllocMaxCandidates;
        HIPC(hipMemcpyAsync(a, ca, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipMemcpyAsync(b, cb, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipMemcpyAsync(step, ck, sizeof(int32_t) * nc, hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipMemcpyAsync(distance, out, sizeof(double) * nc, hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipMemcpyAsync(time, out + S, sizeof(double) * nc, hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipMemcpyAsync(pos_a, out + 2 * S, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost,
            ctx->stream));
        HIPC(hipMemcpyAsync(pos_b, out + 5 * S, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost,
            ctx->stream));
        ctx->sync();
        // order by pair (a, b) in place (the queue is filled by atomics in no fixed order):
        // heapsort over the caller's arrays, which needs no memory of its own
        auto less = [&](int64_t x, int64_t y) {
            return a[x] != a[y] ? a[x] < a[y] : b[x] < b[y];
        };
        auto swap = [&](int64_t x, int64_t y) {
            std::swap(a[x], a[y]);
            std::swap(b[x], b[y]);
            std::swap(step[x], step[y]);
            std::swap(distance[x], distance[y]);
            std::swap(time[x], time[y]);
            for (int c = 0; c < 3; ++c) {
                std::swap(pos_a[3 * x + c], pos_a[3 * y + c]);
                std::swap(pos_b[3 * x + c], pos_b[3 * y + c]);
            }
        };
        auto sift = [&](int64_t root, int64_t end) {
            for (;;) {
                int64_t child = 2 * root + 1;
                if (child >= end)
                    return;
                if (child + 1 < end && less(child, child + 1))
                    ++child;
                if (!less(root, child))
                    return;
                swap(root, child);
                root = child;
            }
        };
        for (int64_t i = nc / 2 - 1; i >= 0; --i)
            sift(i, nc);
        for (int64_t end = nc - 1; end > 0; --end) {
            swap(0, end);
            sift(0, end);
        }
    });
}

int ogp_conj_results(ogp_context* ctx, uint32_t* a, uint32_t* b, int32_t* step, double* distance,
    double* time, double* pos_a, double* pos_b)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    const size_t nc = ctx->res_a.size();
    for (size_t q = 0; q < nc; ++q) {
        if (a)
            a[q] = ctx->res_a[q];
        if (b)
            b[q] = ctx->res_b[q];
        if (step)
            step[q] = ctx->res_k[q];
        const double* o = &ctx->res_out[8 * q];
        if (distance)
            distance[q] = o[0];
        if (time)
            time[q] = o[1];
        for (int c = 0; c < 3; ++c) {
            if (pos_a)
                pos_a[3 * q + c] = o[2 + c];
            if (pos_b)
                pos_b[3 * q + c] = o[5 + c];
        }
    }
    return OGP_OK;
}

int ogp_conj_trajectory(ogp_context* ctx, double* out, int64_t* steps)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        if (steps)
            *steps = ctx->cj_steps;
        if (!out || ctx->cj_steps == 0)
            return;
        const int64_t n = ctx->cj_n, ns = ctx->cj_steps;
        std::vector<double> host(9 * n * ns);
        HIPC(hipMemcpy(host.data(), ctx->cj_traj.p, sizeof(double) * host.size(), hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < ns; ++k)
            for (int c = 0; c < 9; ++c)
                for (int64_t i = 0; i < n; ++i)
                    out[(k * n + i) * 9 + c] = host[(k * 9 + c) * n + i];
    });
}

int ogp_conj_last_timing(const ogp_context* ctx, ogp_conj_timing* t)
{
    if (!ctx || !t)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null argument");
    *t = ctx->cj_timing;
    return OGP_OK;
}

} // extern "C"

// ---- derivatives ---------------------------------------------------------------------------
namespace {

void check_derivative_inputs(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    double step_s, const int64_t* steps)
{
    if (ctx->kind != kGravity)
        throw ApiError(OGP_ERR_ARGUMENT,
            "derivatives need the FP64 gravity force model (OGP_FORCE_GRAVITY, OGP_PRECISION_FP64)");
    ctx->require_gravity_only("the derivative");
    arg_check(n > 0 && r0 && v0 && steps, "null array or no objects");
    arg_check(std::isfinite(step_s) && step_s > 0, "step size must be positive and finite");
    arg_check(ctx->cfg.n_corrections >= 1 && ctx->cfg.n_corrections <= adjoint_max_corrections(),
        "derivatives support 1 to " + std::to_string(adjoint_max_corrections()) + " corrections");
    for (int64_t i = 0; i < n; ++i)
        arg_check(steps[i] >= 0, "negative step count (object " + std::to_string(i) + ")");
    check_finite(r0, 3 * n, "positions");
    check_finite(v0, 3 * n, "velocities");
}

// Recorded gravity Jacobians of a whole span up to ctx->record_bytes are kept in HBM and the
// record pass is the forward pass; beyond it (or with an explicit checkpoint interval) the
// forward pass keeps checkpoints and every segment is recorded again in the reverse sweep. The
// forward-mode STM records and consumes segments of at most that size.

// The startup of the objects (the compact state pc, n objects, ctx->ad_* buffers) with the
// Jacobian of every force evaluation recorded, then the same startup on six tangent lanes per
// object (the st_* buffers) driven by those Jacobians, and their derivatives of the post-startup
// state: Zd (s, S, ring) and RVd (r, v), see launch_tangent_columns.
PropState derivative_start(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    double h, const int64_t* steps, ForceCtx* f)
{
    const int64_t max_steps = *std::max_element(steps, steps + n);
    const int ncorr = ctx->cfg.n_corrections;
    const size_t v3 = sizeof(double) * 3 * n;
    ctx->ad_in.reserve(2 * v3, &ctx->dev_bytes);
    ctx->ad_r.reserve(v3, &ctx->dev_bytes);
    ctx->ad_v.reserve(v3, &ctx->dev_bytes);
    ctx->ad_s.reserve(v3, &ctx->dev_bytes);
    ctx->ad_S.reserve(v3, &ctx->dev_bytes);
    ctx->ad_hist.reserve(9 * v3, &ctx->dev_bytes);
    ctx->ad_ns.reserve(sizeof(int64_t) * n, &ctx->dev_bytes);
    ctx->ad_Zd.reserve(sizeof(double) * 6 * 33 * n, &ctx->dev_bytes);
    ctx->ad_RVd.reserve(sizeof(double) * 36 * n, &ctx->dev_bytes);
    const int64_t slots = 80 + 9 * (ncorr + 1); // the startup's force evaluations
    ctx->ad_grec.reserve(sizeof(double) * 6 * slots * n, &ctx->dev_bytes);
    PropState pc;
    pc.r = ctx->ad_r.as<double>();
    pc.v = ctx->ad_v.as<double>();
    pc.s = ctx->ad_s.as<double>();
    pc.S = ctx->ad_S.as<double>();
    pc.hist = ctx->ad_hist.as<double>();
    pc.nsteps = ctx->ad_ns.as<int64_t>();
    pc.N = n;
    double* d_in = ctx->ad_in.as<double>();
    HIPC(hipMemcpyAsync(d_in, r0, v3, hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(d_in + 3 * n, v0, v3, hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(ctx->ad_ns.p, steps, sizeof(int64_t) * n, hipMemcpyHostToDevice,
        ctx->stream));
    launch_aos_to_soa(ctx->stream, d_in, pc.r, n);
    launch_aos_to_soa(ctx->stream, d_in + 3 * n, pc.v, n);
    NodeTable* nt = ctx->node_table(h, max_steps, 0.0);
    *f = ctx->force_ctx(nt, h);
    ForceCtx fr = *f;
    fr.grec = ctx->ad_grec.as<double>();
    fr.grec_n = n;
    // the tangent lanes and the startup work areas for them (the recording startup needs n)
    ctx->reserve_state(6 * n);
    launch_startup(ctx->stream, kGravityRecord, fr, ctx->rk8_variant, pc, ctx->scratch(), 0, n,
        ncorr, nullptr, 0, ctx->waves_for(n), ctx->fp_exit);
    ctx->check_launch();
    const PropState pt = ctx->prop_state(6 * n);
    launch_tangent_init(ctx->stream, pc.nsteps, n, pt, ctx->st_ns.as<int64_t>());
    // every correction (the derivative of all ncorr, whose values the fixed-point exit repeats)
    launch_startup(ctx->stream, kRecordedTangent, fr, ctx->rk8_variant, pt, ctx->scratch(), 0,
        6 * n, ncorr, nullptr, 0, 1, false);
    launch_tangent_columns(ctx->stream, pt, n, ctx->ad_Zd.as<double>(), ctx->ad_RVd.as<double>());
    ctx->check_launch();
    return pc;
}

// Final state of the compact layout, SoA on the device, to AoS host arrays (either may be null).
void copy_state_out(ogp_context* ctx, const PropState& pc, double* r_out, double* v_out)
{
    const int64_t n = pc.N;
    std::vector<double> rs(3 * n), vs(3 * n);
    HIPC(hipMemcpyAsync(rs.data(), pc.r, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipMemcpyAsync(vs.data(), pc.v, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    for (int64_t o = 0; o < n; ++o)
        for (int c = 0; c < 3; ++c) {
            if (r_out)
                r_out[3 * o + c] = rs[c * n + o];
            if (v_out)
                v_out[3 * o + c] = vs[c * n + o];
        }
}

// Reverse mode for R cotangent rows seeded by seeds_host[(rho * 6 + e) * n + o]; writes the
// final state (AoS) and out_host[(o * R + rho) * 6 + col]. C = checkpoint interval in steps,
// 0 = automatic.
void adjoint_run(ogp_context* ctx, const double* r0, const double* v0, int64_t n, double h,
    const int64_t* steps, int64_t C, int R, const std::vector<double>& seeds_host,
    double* r_out, double* v_out, double* out_host)
{
    Event e0, e1, e2;
    HIPC(hipEventRecord(e0.e, ctx->stream));
    ForceCtx f;
    const PropState pc = derivative_start(ctx, r0, v0, n, h, steps, &f);
    const int ncorr = ctx->cfg.n_corrections;
    const int64_t max_steps = *std::max_element(steps, steps + n);
    const int64_t gj_steps = std::max<int64_t>(max_steps - 4, 0);
    const bool record_all = C == 0
        && (double)gj_steps * 6 * (double)n * sizeof(double) <= ctx->record_bytes;
    if (record_all)
        C = std::max<int64_t>(gj_steps, 1);
    else if (C == 0)
        C = 64;
    const int64_t nseg = max_steps >= 5 ? (max_steps - 4 + C - 1) / C : 0;
    const int W = ctx->waves_for(n);
    ctx->ad_G.reserve(sizeof(double) * C * 6 * n, &ctx->dev_bytes);
    if (record_all) {
        // forward pass = record pass over the whole span
        if (nseg > 0)
            launch_gj8_record(ctx->stream, f, pc, 5, max_steps + 1, ncorr, ctx->ad_G.as<double>(), W,
                ctx->fp_exit);
        ctx->check_launch();
    } else {
        // forward pass over the GJ8 steps, (s, S, ring) stored at the start of every segment
        ctx->ad_ck.reserve(sizeof(double) * 33 * n * std::max<int64_t>(nseg, 1), &ctx->dev_bytes);
        for (int64_t q = 0; q < nseg; ++q) {
            const int64_t k0 = 5 + q * C, k1 = std::min<int64_t>(k0 + C, max_steps + 1);
            launch_checkpoint(ctx->stream, pc, ctx->ad_ck.as<double>() + q * 33 * n, true);
            launch_chunk(ctx->stream, kGravity, f, pc, 0, n, k0, k1, ncorr, nullptr, 1, 0, W, 1,
                ctx->fp_exit);
            ctx->check_launch();
        }
    }
    HIPC(hipEventRecord(e1.e, ctx->stream));
    copy_state_out(ctx, pc, r_out, v_out);
    // reverse pass, segment by segment from the last
    ctx->ad_lam.reserve(sizeof(double) * R * 33 * n, &ctx->dev_bytes);
    ctx->ad_seeds.reserve(sizeof(double) * R * 6 * n, &ctx->dev_bytes);
    ctx->ad_out.reserve(sizeof(double) * R * 6 * n, &ctx->dev_bytes);
    HIPC(hipMemsetAsync(ctx->ad_lam.p, 0, sizeof(double) * R * 33 * n, ctx->stream));
    HIPC(hipMemcpyAsync(ctx->ad_seeds.p, seeds_host.data(), sizeof(double) * R * 6 * n,
        hipMemcpyHostToDevice, ctx->stream));
    for (int64_t q = nseg - 1; q >= 0; --q) {
        const int64_t k0 = 5 + q * C, k1 = std::min<int64_t>(k0 + C, max_steps + 1);
        if (!record_all) {
            launch_checkpoint(ctx->stream, pc, ctx->ad_ck.as<double>() + q * 33 * n, false);
            launch_gj8_record(ctx->stream, f, pc, k0, k1, ncorr, ctx->ad_G.as<double>(), W,
                ctx->fp_exit);
        }
        launch_gj8_reverse(ctx->stream, pc.nsteps, n, h, k0, k1, ncorr, ctx->ad_G.as<double>(), R,
            ctx->ad_seeds.as<double>(), ctx->ad_lam.as<double>());
        ctx->check_launch();
    }
    launch_adjoint_combine(ctx->stream, pc.nsteps, n, R, ctx->ad_lam.as<double>(),
        ctx->ad_Zd.as<double>(), ctx->ad_RVd.as<double>(), ctx->ad_seeds.as<double>(),
        ctx->ad_out.as<double>());
    ctx->check_launch();
    HIPC(hipEventRecord(e2.e, ctx->stream));
    HIPC(hipMemcpyAsync(out_host, ctx->ad_out.p, sizeof(double) * R * 6 * n, hipMemcpyDeviceToHost,
        ctx->stream));
    ctx->sync();
    float a = 0, b = 0;
    HIPC(hipEventElapsedTime(&a, e0.e, e1.e));
    HIPC(hipEventElapsedTime(&b, e1.e, e2.e));
    ctx->ad_timing.forward_ms = a;
    ctx->ad_timing.adjoint_ms = b;
}

} // namespace

extern "C" {

int ogp_gravity_jacobian_itrs(ogp_context* ctx, const double* r, int64_t n, double* a, double* G)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        if (ctx->cfg.force != OGP_FORCE_GRAVITY)
            throw ApiError(OGP_ERR_ARGUMENT, "the context has no gravity model");
        arg_check(n >= 0 && (n == 0 || (r && a && G)), "null array");
        if (n == 0)
            return;
        check_finite(r, 3 * n, "positions");
        ctx->io_a.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
        ctx->io_b.reserve(sizeof(double) * 3 * n, &ctx->dev_bytes);
        ctx->io_c.reserve(sizeof(double) * 9 * n, &ctx->dev_bytes);
        HIPC(hipMemcpyAsync(ctx->io_a.p, r, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
            ctx->stream));
        launch_gravity_jac(ctx->stream, ctx->gdev, ctx->io_a.as<double>(), n, ctx->io_b.as<double>(),
            ctx->io_c.as<double>());
        ctx->check_launch();
        HIPC(hipMemcpyAsync(a, ctx->io_b.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
            ctx->stream));
        HIPC(hipMemcpyAsync(G, ctx->io_c.p, sizeof(double) * 9 * n, hipMemcpyDeviceToHost,
            ctx->stream));
        ctx->sync();
    });
}

int ogp_stm(ogp_context* ctx, const double* r0, const double* v0, int64_t n, double step_s,
    const int64_t* steps, double* r_out, double* v_out, double* stm)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        check_derivative_inputs(ctx, r0, v0, n, step_s, steps);
        arg_check(r_out && v_out && stm, "null output array");
        // forward mode: each segment of GJ8 steps is recorded (the trajectory, the propagation's
        // arithmetic, and each step's Jacobian at its corrector fixed point) and then the six
        // tangent columns are carried through it; one segment when the whole span's Jacobians fit
        Event e0, e1;
        HIPC(hipEventRecord(e0.e, ctx->stream));
        ForceCtx f;
        const PropState pc = derivative_start(ctx, r0, v0, n, step_s, steps, &f);
        const int ncorr = ctx->cfg.n_corrections;
        const int64_t max_steps = *std::max_element(steps, steps + n);
        const int64_t gj_steps = std::max<int64_t>(max_steps - 4, 0);
        const double per_step = 6.0 * (double)n * sizeof(double);
        const int64_t C = std::max<int64_t>(1,
            std::min<int64_t>(gj_steps, (int64_t)(ctx->record_bytes / per_step)));
        const int W = ctx->waves_for(n);
        ctx->ad_G.reserve(sizeof(double) * C * 6 * n, &ctx->dev_bytes);
        for (int64_t k0 = 5; k0 <= max_steps; k0 += C) {
            const int64_t k1 = std::min<int64_t>(k0 + C, max_steps + 1);
            launch_gj8_record(ctx->stream, f, pc, k0, k1, ncorr, ctx->ad_G.as<double>(), W,
                ctx->fp_exit);
            launch_gj8_tangent(ctx->stream, pc.nsteps, n, step_s, k0, k1, ncorr,
                ctx->ad_G.as<double>(), ctx->ad_Zd.as<double>(), ctx->ad_RVd.as<double>());
            ctx->check_launch();
        }
        ctx->io_c.reserve(sizeof(double) * 36 * n, &ctx->dev_bytes);
        launch_stm_final(ctx->stream, pc.nsteps, n, ctx->ad_RVd.as<double>(), ctx->io_c.as<double>());
        ctx->check_launch();
        HIPC(hipEventRecord(e1.e, ctx->stream));
        HIPC(hipMemcpyAsync(stm, ctx->io_c.p, sizeof(double) * 36 * n, hipMemcpyDeviceToHost,
            ctx->stream));
        copy_state_out(ctx, pc, r_out, v_out);
        float ms = 0;
        HIPC(hipEventElapsedTime(&ms, e0.e, e1.e));
        ctx->ad_timing.stm_ms = ms;
    });
}

int ogp_adjoint_jacobian(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    double step_s, const int64_t* steps, int64_t checkpoint_steps, double* r_out, double* v_out,
    double* jac)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        check_derivative_inputs(ctx, r0, v0, n, step_s, steps);
        arg_check(jac != nullptr, "null output array");
        arg_check(checkpoint_steps >= 0, "negative checkpoint interval");
        const int64_t C = checkpoint_steps;
        std::vector<double> seeds(36 * n, 0.0);
        for (int rho = 0; rho < 6; ++rho)
            for (int64_t o = 0; o < n; ++o)
                seeds[(rho * 6 + rho) * n + o] = 1.0;
        adjoint_run(ctx, r0, v0, n, step_s, steps, C, 6, seeds, r_out, v_out, jac);
    });
}

int ogp_adjoint_vjp(ogp_context* ctx, const double* r0, const double* v0, int64_t n,
    double step_s, const int64_t* steps, int64_t checkpoint_steps, const double* g_r,
    const double* g_v, double* g_r0, double* g_v0)
{
    if (!ctx)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null context");
    return guarded(ctx, [&] {
        check_derivative_inputs(ctx, r0, v0, n, step_s, steps);
        arg_check(g_r && g_v && g_r0 && g_v0, "null array");
        arg_check(checkpoint_steps >= 0, "negative checkpoint interval");
        check_finite(g_r, 3 * n, "position gradients");
        check_finite(g_v, 3 * n, "velocity gradients");
        const int64_t C = checkpoint_steps;
        std::vector<double> seeds(6 * n), out(6 * n);
        for (int64_t o = 0; o < n; ++o)
            for (int c = 0; c < 3; ++c) {
                seeds[c * n + o] = g_r[3 * o + c];
                seeds[(3 + c) * n + o] = g_v[3 * o + c];
            }
        adjoint_run(ctx, r0, v0, n, step_s, steps, C, 1, seeds, nullptr, nullptr, out.data());
        for (int64_t o = 0; o < n; ++o)
            for (int c = 0; c < 3; ++c) {
                g_r0[3 * o + c] = out[6 * o + c];
                g_v0[3 * o + c] = out[6 * o + 3 + c];
            }
    });
}

int ogp_adjoint_last_timing(const ogp_context* ctx, ogp_adjoint_timing* t)
{
    if (!ctx || !t)
        return fail(nullptr, OGP_ERR_ARGUMENT, "null argument");
    *t = ctx->ad_timing;
    return OGP_OK;
}

} // extern "C"
