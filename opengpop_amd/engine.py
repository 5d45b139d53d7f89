"""Python mirror of the reference's library entry points over the C ABI.

Names follow the reference: ForceModel (force_model.cuh:28-87), GpuPropagator
(gpu_propagator.cuh:26-59), GpuStartup (gpu_startup.cuh:24-120), ClosestApproach
(closest_approach.cuh:36-88) and ClosestApproachData::Get (types.cuh:153-169). All compute runs
in libogp.so on the GPU; errors raise OgpError with the library's message.
"""
import ctypes

import numpy as np

from . import _capi as C
from ._capi import DATA_DIR, check, f64, lib, ptr


class Context:
    """One device context: datasets, configuration and the HIP stream (include/ogp.h)."""

    def __init__(self, degree=-1, precision="fp64", n_corrections=7,
                 quirks=C.QUIRKS_REFERENCE, force=C.FORCE_GRAVITY,
                 epoch=(2025, 1, 1, 0, 0, 0.0), data_dir=DATA_DIR, split_degree=2):
        L = lib()
        cfg = C.OgpConfig()
        L.ogp_default_config(ctypes.byref(cfg))
        self._dir = data_dir.encode()
        cfg.data_dir = self._dir
        cfg.degree = degree
        cfg.precision = C.PRECISION_SPLIT if precision == "split" else C.PRECISION_FP64
        cfg.split_degree = split_degree
        cfg.n_corrections = n_corrections
        cfg.quirks = quirks
        cfg.force = force
        (cfg.year, cfg.month, cfg.day, cfg.hour, cfg.minute), cfg.second = epoch[:5], float(epoch[5])
        h = ctypes.c_void_p()
        check(None, L.ogp_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.epoch = tuple(epoch)
        self.force = force

    def close(self):
        if getattr(self, "h", None):
            lib().ogp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def device_bytes(self):
        return lib().ogp_device_bytes(self.h)

    # -- force model -----------------------------------------------------------------
    def gravity_itrs(self, r):
        r = f64(r).reshape(-1, 3)
        a = np.zeros_like(r)
        check(self.h, lib().ogp_gravity_itrs(self.h, ptr(r), len(r), ptr(a)))
        return a

    def accel(self, r, step, node, out_stride=3, v=None):
        """ForceModel::ComputePropagationAcceleration at propagation node `node`: gravity plus the
        enabled sources (v, AoS velocities, is needed for drag)."""
        r = f64(r).reshape(-1, 3)
        n = len(r)
        vv = None if v is None else f64(v).reshape(-1, 3)
        out = np.zeros(max(out_stride * (n - 1) + 3, 3))
        check(self.h, lib().ogp_accel(self.h, ptr(r), ptr(vv), n, step, node, ptr(out),
                                      out_stride))
        return out

    def time_nodes(self, step, before, offsets):
        offsets = f64(offsets)
        M = np.zeros((len(offsets), 9))
        mjd = np.zeros(len(offsets))
        check(self.h, lib().ogp_time_nodes(self.h, step, before, ptr(offsets), len(offsets),
                                           ptr(M), ptr(mjd)))
        return M, mjd

    # -- propagation -----------------------------------------------------------------
    def propagate(self, r0, v0, step, steps):
        r0 = f64(r0).reshape(-1, 3)
        v0 = f64(v0).reshape(-1, 3)
        n = len(r0)
        h = np.ascontiguousarray(np.broadcast_to(np.asarray(step, np.float64), (n,)))
        ns = np.ascontiguousarray(np.broadcast_to(np.asarray(steps, np.int64), (n,)))
        r = np.zeros_like(r0)
        v = np.zeros_like(v0)
        check(self.h, lib().ogp_propagate(self.h, ptr(r0), ptr(v0), n, ptr(h), ptr(ns), ptr(r),
                                          ptr(v)))
        return r, v

    def propagate_checked(self, r0, v0, step, steps):
        """ogp_propagate_checked: (r, v, status) with per-object flags (OBJECT_*) instead of a
        failed call for non-finite states."""
        r0 = f64(r0).reshape(-1, 3)
        v0 = f64(v0).reshape(-1, 3)
        n = len(r0)
        h = np.ascontiguousarray(np.broadcast_to(np.asarray(step, np.float64), (n,)))
        ns = np.ascontiguousarray(np.broadcast_to(np.asarray(steps, np.int64), (n,)))
        r = np.zeros_like(r0)
        v = np.zeros_like(v0)
        st = np.zeros(n, np.int32)
        check(self.h, lib().ogp_propagate_checked(self.h, ptr(r0), ptr(v0), n, ptr(h), ptr(ns),
                                                  ptr(r), ptr(v), ptr(st)))
        return r, v, st

    def propagate_device(self, r0_soa, v0_soa, n, step, steps, r_soa, v_soa, status=False):
        """ogp_propagate_device on device-resident SoA arrays (x[n], y[n], z[n] blocks).
        Array arguments are device addresses: ints or objects with data_ptr() (torch tensors).
        step and steps are host values (scalars or per-object arrays). With status=True the
        checked entry point runs and the per-object flags are returned."""
        addr = [a.data_ptr() if hasattr(a, "data_ptr") else int(a)
                for a in (r0_soa, v0_soa, r_soa, v_soa)]
        h = np.ascontiguousarray(np.broadcast_to(np.asarray(step, np.float64), (n,)))
        ns = np.ascontiguousarray(np.broadcast_to(np.asarray(steps, np.int64), (n,)))
        if status:
            st = np.zeros(n, np.int32)
            check(self.h, lib().ogp_propagate_device_checked(self.h, addr[0], addr[1], n, ptr(h),
                                                             ptr(ns), addr[2], addr[3], ptr(st)))
            return st
        check(self.h, lib().ogp_propagate_device(self.h, addr[0], addr[1], n, ptr(h), ptr(ns),
                                                 addr[2], addr[3]))
        return None

    # -- device-pointer forms of the force model, startup and predictor -------------------
    @staticmethod
    def _addr(a):
        if a is None:
            return None
        return a.data_ptr() if hasattr(a, "data_ptr") else int(a)

    def accel_device(self, d_r, d_v, n, step, node, d_a, out_stride=3):
        """ForceModel::ComputePropagationAcceleration on device AoS arrays."""
        A = self._addr
        check(self.h, lib().ogp_accel_device(self.h, A(d_r), A(d_v), n, step, node, A(d_a),
                                             out_stride))

    def accel_generation(self, d_r, d_v, n, step, step_num, k_index, d_K):
        """ForceModel::ComputeGenerationAcceleration: column k_index of the device K matrix
        [n][10][3] at RK8 stage k_index of startup step step_num."""
        A = self._addr
        check(self.h, lib().ogp_accel_generation(self.h, A(d_r), A(d_v), n, step, step_num,
                                                 k_index, A(d_K)))

    def startup_device(self, d_r0, d_v0, n, step, d_pos=None, d_vel=None, d_s=None, d_S=None,
                       d_acc=None):
        """GpuStartup::Startup on device arrays: outputs [n][9][3] (each optional)."""
        A = self._addr
        check(self.h, lib().ogp_startup_device(self.h, A(d_r0), A(d_v0), n, step, A(d_pos),
                                               A(d_vel), A(d_s), A(d_S), A(d_acc)))

    def predict_device(self, d_r, d_v, d_acc, d_s, d_S, n, step, step_num):
        """Predictor::Predict(step_num) on the device arrays of Predictor::SetPointers."""
        A = self._addr
        check(self.h, lib().ogp_predict_device(self.h, A(d_r), A(d_v), A(d_acc), A(d_s), A(d_S),
                                               n, step, step_num))

    def perturbed_satellites(self, r0, v0, n, strength, seed=1234, obj_r=None, obj_v=None):
        """PerturbationGenerator::generatePerturbedSatellites: (r, v) host arrays [n (+1)][3]
        of n satellites at r0 with velocity v0 + strength * N(0, 1) per component, plus
        (obj_r, obj_v) as the last entry when both are given."""
        r0 = f64(r0).reshape(3)
        v0 = f64(v0).reshape(3)
        append = obj_r is not None and obj_v is not None
        orr = f64(obj_r).reshape(3) if append else None
        ov = f64(obj_v).reshape(3) if append else None
        count = max(n + (1 if append else 0), 0)  # a negative n is reported by the library
        r = np.zeros((count, 3))
        v = np.zeros((count, 3))
        check(self.h, lib().ogp_perturbed_satellites(self.h, ptr(r0), ptr(v0), n, strength, seed,
                                                     ptr(orr), ptr(ov), ptr(r), ptr(v)))
        return r, v

    def perturbed_satellites_device(self, r0, v0, n, strength, r_soa, v_soa, stride=None,
                                    seed=1234, obj_r=None, obj_v=None):
        """ogp_perturbed_satellites_device into device SoA arrays (device addresses: ints or
        objects with data_ptr()); stride defaults to the number of entries."""
        r0 = f64(r0).reshape(3)
        v0 = f64(v0).reshape(3)
        append = obj_r is not None and obj_v is not None
        orr = f64(obj_r).reshape(3) if append else None
        ov = f64(obj_v).reshape(3) if append else None
        if stride is None:
            stride = n + (1 if append else 0)
        addr = [a.data_ptr() if hasattr(a, "data_ptr") else int(a) for a in (r_soa, v_soa)]
        check(self.h, lib().ogp_perturbed_satellites_device(
            self.h, ptr(r0), ptr(v0), n, strength, seed, ptr(orr), ptr(ov), addr[0], addr[1],
            stride))

    def set_tuning(self, waves=0, chunk_steps=0):
        check(self.h, lib().ogp_set_tuning(self.h, waves, chunk_steps))

    def set_sources(self, mask=0, drag_bc=0.0):
        """Acceleration sources added to gravity: SOURCE_SUN | SOURCE_MOON | SOURCE_DRAG bits,
        drag_bc = C_D A / m (m^2/kg). mask 0 = gravity only."""
        s = C.OgpSources(mask, drag_bc)
        check(self.h, lib().ogp_set_sources(self.h, ctypes.byref(s)))

    def set_drag_coefficients(self, bc=None):
        """Per-object C_D A / m (m^2/kg) for drag, in object order; None returns to the single
        coefficient of set_sources."""
        if bc is None:
            check(self.h, lib().ogp_set_drag_coefficients(self.h, None, 0))
            return
        b = f64(bc).reshape(-1)
        check(self.h, lib().ogp_set_drag_coefficients(self.h, ptr(b), len(b)))

    def get_sources(self):
        s = C.OgpSources()
        check(self.h, lib().ogp_get_sources(self.h, ctypes.byref(s)))
        return s.mask, s.drag_bc

    def set_screen_shape(self, shape=0):
        """Conjunction screen tiles: 0 tuned, 1 square (64 x 64), 2 narrow (4 x 256)."""
        check(self.h, lib().ogp_set_screen_shape(self.h, shape))

    def set_radial_filter(self, mode=1):
        """Threshold screens skip the pair tiles whose objects' sampled radius ranges lie more
        than the threshold apart: 0 off, 1 when there are hundreds of pairs per object (the
        default), 2 always. Results are the same in every mode."""
        check(self.h, lib().ogp_set_radial_filter(self.h, int(mode)))

    def get_screen_shape(self, na):
        g = ctypes.c_int()
        check(self.h, lib().ogp_get_screen_shape(self.h, na, ctypes.byref(g)))
        return g.value

    def set_lockstep(self, groups=0):
        """Object groups per workgroup of the GJ8 chunks: 0 tuned, 1 or 2."""
        check(self.h, lib().ogp_set_lockstep(self.h, groups))

    def get_lockstep(self, n):
        g = ctypes.c_int()
        check(self.h, lib().ogp_get_lockstep(self.h, n, ctypes.byref(g)))
        return g.value

    def set_slices(self, slices=0):
        """Object slices per step-size class run on concurrent streams (0 = automatic)."""
        check(self.h, lib().ogp_set_slices(self.h, slices))

    def set_fixed_point_exit(self, on=True):
        """End a GJ8 step's corrector evaluations once the next state to evaluate repeats the
        last one bit for bit (results unchanged; the default), or always run n_corrections."""
        check(self.h, lib().ogp_set_fixed_point_exit(self.h, 1 if on else 0))

    def set_derivative_memory(self, nbytes=0):
        """HBM budget for the Jacobians recorded by the derivative passes (0 = 16 GiB)."""
        check(self.h, lib().ogp_set_derivative_memory(self.h, int(nbytes)))

    def get_tuning(self, n):
        w = ctypes.c_int()
        ch = C._I64()
        check(self.h, lib().ogp_get_tuning(self.h, n, ctypes.byref(w), ctypes.byref(ch)))
        return w.value, ch.value

    def timing(self):
        t = C.OgpTiming()
        check(self.h, lib().ogp_last_timing(self.h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in C.OgpTiming._fields_}

    # -- derivatives ------------------------------------------------------------------
    def gravity_jacobian_itrs(self, r):
        """(a [n][3], G [n][3][3]) with G[i][j] = d a_i / d r_j, ITRS."""
        r = f64(r).reshape(-1, 3)
        a = np.zeros_like(r)
        G = np.zeros((len(r), 3, 3))
        check(self.h, lib().ogp_gravity_jacobian_itrs(self.h, ptr(r), len(r), ptr(a), ptr(G)))
        return a, G

    def _steps(self, n, steps):
        return np.ascontiguousarray(np.broadcast_to(np.asarray(steps, np.int64), (n,)))

    def stm(self, r0, v0, step, steps):
        """Final state and the state-transition matrix [n][6][6] from the variational
        equations."""
        r0 = f64(r0).reshape(-1, 3)
        v0 = f64(v0).reshape(-1, 3)
        n = len(r0)
        ns = self._steps(n, steps)
        r, v, m = np.zeros_like(r0), np.zeros_like(v0), np.zeros((n, 6, 6))
        check(self.h, lib().ogp_stm(self.h, ptr(r0), ptr(v0), n, float(step), ptr(ns), ptr(r),
                                    ptr(v), ptr(m)))
        return r, v, m

    def adjoint_jacobian(self, r0, v0, step, steps, checkpoint_steps=0):
        """Final state and d(r, v)_final / d(r, v)_initial [n][6][6] by the discrete adjoint."""
        r0 = f64(r0).reshape(-1, 3)
        v0 = f64(v0).reshape(-1, 3)
        n = len(r0)
        ns = self._steps(n, steps)
        r, v, J = np.zeros_like(r0), np.zeros_like(v0), np.zeros((n, 6, 6))
        check(self.h, lib().ogp_adjoint_jacobian(self.h, ptr(r0), ptr(v0), n, float(step), ptr(ns),
                                                 checkpoint_steps, ptr(r), ptr(v), ptr(J)))
        return r, v, J

    def adjoint_vjp(self, r0, v0, step, steps, g_r, g_v, checkpoint_steps=0):
        """Gradients w.r.t. (r0, v0) of a scalar whose gradients w.r.t. the final state are
        (g_r, g_v)."""
        r0 = f64(r0).reshape(-1, 3)
        v0 = f64(v0).reshape(-1, 3)
        n = len(r0)
        ns = self._steps(n, steps)
        g_r = f64(g_r).reshape(n, 3)
        g_v = f64(g_v).reshape(n, 3)
        gr0, gv0 = np.zeros_like(r0), np.zeros_like(v0)
        check(self.h, lib().ogp_adjoint_vjp(self.h, ptr(r0), ptr(v0), n, float(step), ptr(ns),
                                            checkpoint_steps, ptr(g_r), ptr(g_v), ptr(gr0),
                                            ptr(gv0)))
        return gr0, gv0

    def adjoint_timing(self):
        t = C.OgpAdjointTiming()
        check(self.h, lib().ogp_adjoint_last_timing(self.h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in C.OgpAdjointTiming._fields_}

    def startup(self, r0, v0, step):
        """GpuStartup::Startup: node matrices, s, S and accelerations, each [n][9][3]."""
        r0 = f64(r0).reshape(-1, 3)
        v0 = f64(v0).reshape(-1, 3)
        n = len(r0)
        out = [np.zeros((n, 9, 3)) for _ in range(5)]
        check(self.h, lib().ogp_startup(self.h, ptr(r0), ptr(v0), n, step, *[ptr(o) for o in out]))
        return dict(zip(("pos", "vel", "s", "S", "acc"), out))

    def kat_predict(self, acc27, s3, S3, h):
        acc = f64(acc27).reshape(-1, 27)
        n = len(acc)
        s = f64(s3).reshape(n, 3)
        S = np.array(S3, dtype=np.float64).reshape(n, 3)
        r = np.zeros((n, 3))
        v = np.zeros((n, 3))
        check(self.h, lib().ogp_kat_predict(self.h, ptr(acc), ptr(s), ptr(S), n, h, ptr(r), ptr(v)))
        return r, v, S

    def kat_finalize_correction(self, r3, v3, acc27, s3, S3, h):
        acc = f64(acc27).reshape(-1, 27)
        n = len(acc)
        r = np.array(r3, dtype=np.float64).reshape(n, 3)
        v = np.array(v3, dtype=np.float64).reshape(n, 3)
        s = f64(s3).reshape(n, 3)
        S = f64(S3).reshape(n, 3)
        check(self.h, lib().ogp_kat_finalize_correction(self.h, ptr(r), ptr(v), ptr(acc), ptr(s),
                                                        ptr(S), n, h))
        return r, v


class GpuPropagator:
    """Stepwise propagator (GpuPropagator, gpu_propagator.cu:28-154) with its own state: any
    number of them may share a context. From host arrays (the propagator owns its device state),
    or with device=True from the caller's AoS [n][3] device buffers r0, v0 (ints or objects with
    data_ptr(), e.g. torch tensors), which every step advances in place."""

    def __init__(self, ctx, r0, v0, step, max_steps, n=None, device=False):
        self.ctx = ctx
        h = ctypes.c_void_p()
        if device:
            self.n = int(n)
            addr = [a.data_ptr() if hasattr(a, "data_ptr") else int(a) for a in (r0, v0)]
            check(ctx.h, lib().ogp_propagator_create(ctx.h, addr[0], addr[1], self.n, step,
                                                     max_steps, ctypes.byref(h)))
        else:
            r0 = f64(r0).reshape(-1, 3)
            v0 = f64(v0).reshape(-1, 3)
            self.n = len(r0)
            check(ctx.h, lib().ogp_propagator_create_host(ctx.h, ptr(r0), ptr(v0), self.n, step,
                                                          max_steps, ctypes.byref(h)))
        self.h = h

    def step(self, d_acc=None):
        """One GpuPropagator::Step. Returns (r, v, a) as host arrays; with d_acc (a device
        address) the acceleration column is also written there, as the reference does."""
        r = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        a = np.zeros((self.n, 3))
        dacc = None if d_acc is None else (d_acc.data_ptr() if hasattr(d_acc, "data_ptr")
                                           else int(d_acc))
        check(self.ctx.h, lib().ogp_propagator_step(self.h, ptr(r), ptr(v), dacc))
        check(self.ctx.h, lib().ogp_propagator_acceleration(self.h, ptr(a)))
        return r, v, a

    def copy_state_to_host(self):
        r = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        check(self.ctx.h, lib().ogp_propagator_copy_state_to_host(self.h, ptr(r), ptr(v)))
        return r, v

    def close(self):
        if getattr(self, "h", None):
            lib().ogp_propagator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ClosestApproach:
    """ClosestApproach::{Initialize, Propagate, FindClosest} (closest_approach.cu:47-210)."""

    def __init__(self, ctx, step):
        self.ctx = ctx
        self.step = step

    def initialize(self, positions, velocities, set_indices):
        r = f64(positions).reshape(-1, 3)
        v = f64(velocities).reshape(-1, 3)
        s = np.ascontiguousarray(set_indices, dtype=np.uint32)
        check(self.ctx.h, lib().ogp_conj_init(self.ctx.h, ptr(r), ptr(v), ptr(s), len(r),
                                              self.step))
        self.sets = s

    def propagate(self, num_steps):
        check(self.ctx.h, lib().ogp_conj_propagate(self.ctx.h, num_steps))

    def find_closest(self, threshold=0.0):
        """Returns ClosestApproachData; threshold <= 0 refines every cross-set pair."""
        cnt = C._I64()
        check(self.ctx.h, lib().ogp_conj_find(self.ctx.h, threshold, ctypes.byref(cnt)))
        n = cnt.value
        a = np.zeros(n, np.uint32)
        b = np.zeros(n, np.uint32)
        k = np.zeros(n, np.int32)
        d = np.zeros(n)
        t = np.zeros(n)
        pa = np.zeros((n, 3))
        pb = np.zeros((n, 3))
        check(self.ctx.h, lib().ogp_conj_results(self.ctx.h, ptr(a), ptr(b), ptr(k), ptr(d),
                                                 ptr(t), ptr(pa), ptr(pb)))
        return ClosestApproachData(a, b, k, d, t, pa, pb)

    def find_closest_noalloc(self, threshold, capacity=65536):
        """ogp_conj_find_noalloc: the threshold search of find_closest with no device
        allocation (at most 65,536 candidates)."""
        a = np.zeros(capacity, np.uint32)
        b = np.zeros(capacity, np.uint32)
        k = np.zeros(capacity, np.int32)
        d = np.zeros(capacity)
        t = np.zeros(capacity)
        pa = np.zeros((capacity, 3))
        pb = np.zeros((capacity, 3))
        cnt = C._I64()
        check(self.ctx.h, lib().ogp_conj_find_noalloc(self.ctx.h, threshold, capacity, ptr(a),
                                                      ptr(b), ptr(k), ptr(d), ptr(t), ptr(pa),
                                                      ptr(pb), ctypes.byref(cnt)))
        n = cnt.value
        return ClosestApproachData(a[:n], b[:n], k[:n], d[:n], t[:n], pa[:n], pb[:n])

    def screen(self):
        np_ = C._I64()
        check(self.ctx.h, lib().ogp_conj_num_pairs(self.ctx.h, ctypes.byref(np_)))
        d = np.zeros(np_.value)
        k = np.zeros(np_.value, np.int32)
        check(self.ctx.h, lib().ogp_conj_screen(self.ctx.h, ptr(d), ptr(k)))
        return d, k

    def trajectory(self):
        steps = C._I64()
        check(self.ctx.h, lib().ogp_conj_trajectory(self.ctx.h, None, ctypes.byref(steps)))
        out = np.zeros((steps.value, len(self.sets), 9))
        check(self.ctx.h, lib().ogp_conj_trajectory(self.ctx.h, ptr(out), ctypes.byref(steps)))
        return out

    def timing(self):
        t = C.OgpConjTiming()
        check(self.ctx.h, lib().ogp_conj_last_timing(self.ctx.h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in C.OgpConjTiming._fields_}


class ClosestApproachData:
    """Results of FindClosest; get(i, j) mirrors ClosestApproachData::Get (types.cuh:153-169)."""

    def __init__(self, a, b, k, d, t, pa, pb):
        self.a, self.b, self.step, self.distance, self.time = a, b, k, d, t
        self.pos_a, self.pos_b = pa, pb
        self._index = None  # (a, b) -> record, built on the first get()

    def __len__(self):
        return len(self.a)

    def get(self, i, j):
        if i == j:
            return dict(distance=0.0, time=0.0, pos_a=np.zeros(3), pos_b=np.zeros(3))
        swap = i > j
        if self._index is None:
            self._index = {(int(x), int(y)): q for q, (x, y) in enumerate(zip(self.a, self.b))}
        q = self._index.get((min(i, j), max(i, j)))
        if q is None:
            return dict(distance=-1.0, time=-1.0, pos_a=-np.ones(3), pos_b=-np.ones(3))
        pa, pb = self.pos_a[q], self.pos_b[q]
        # positionA belongs to the first argument (the reference stores (A, B) of the pair)
        return dict(distance=self.distance[q], time=self.time[q],
                    pos_a=pb if swap else pa, pos_b=pa if swap else pb)
