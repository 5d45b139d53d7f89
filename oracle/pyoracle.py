"""ctypes wrapper around the FP64 host oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in __graft_entry__.py and by
the cpu_baseline leg of bench.py, and by nothing in the product package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
DATA_DIR = os.path.join(os.path.dirname(HERE), "data")

# Force kinds shared with the product's analytic test forces.
GRAVITY, NONE, CONST, POLY7, POLY8, POLY9, EXP, DAMPED = range(8)
# quirk bits: reference RK8 row 7, reference backward stage times
QUIRK_RK8_ROW7 = 1
QUIRK_BACKWARD_TIMES = 2
REFERENCE_QUIRKS = QUIRK_RK8_ROW7 | QUIRK_BACKWARD_TIMES

_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int


def _ptr(a):
    return a.ctypes.data if a is not None else None


# The C library holds one dataset at a time; every Oracle re-opens its own degree before a call
# that depends on it.
_loaded = {"key": None}


class Oracle:
    def __init__(self, degree=-1, data_dir=DATA_DIR, lib_path=LIB_PATH):
        if not os.path.exists(lib_path):
            raise FileNotFoundError(f"oracle not built: {lib_path} (run make -C oracle)")
        L = ctypes.CDLL(lib_path)
        self.L = L
        L.orc_open.argtypes = [ctypes.c_char_p, _I]
        L.orc_error.restype = ctypes.c_char_p
        for name in ("orc_gm", "orc_radius"):
            getattr(L, name).restype = _D
        L.orc_calendar_to_mjd.restype = _D
        L.orc_calendar_to_mjd.argtypes = [_I] * 5 + [_D]
        L.orc_ntp_to_mjd.restype = _D
        L.orc_ntp_to_mjd.argtypes = [ctypes.c_longlong]
        L.orc_ftable.argtypes = [_D, _P]
        L.orc_precession.argtypes = [_D, _P]
        L.orc_nutation.argtypes = [_D, _P]
        L.orc_gast.restype = _D
        L.orc_gast.argtypes = [_D, _D, _D, _D, _P]
        L.orc_rot.argtypes = [_I, _D, _P]
        L.orc_eop.argtypes = [_D, _P]
        L.orc_leap_seconds.restype = _D
        L.orc_leap_seconds.argtypes = [_D]
        L.orc_time_nodes.argtypes = [_P, _D, _P, _I, _P, _P]
        L.orc_gravity_itrs.argtypes = [_P, ctypes.c_long, _P]
        L.orc_potential_itrs.restype = _D
        L.orc_philox4x32_10.argtypes = [_P, _P, _P]
        L.orc_set_sources.argtypes = [_I, _D]
        L.orc_sun_moon.argtypes = [_D, _P]
        L.orc_density.restype = _D
        L.orc_density.argtypes = [_D]
        L.orc_sources_accel.argtypes = [_I, _D, _P, _P, _P, _P]
        L.orc_perturbed.argtypes = [_P, _P, ctypes.c_long, _D, ctypes.c_uint64, _P, _P]
        L.orc_potential_itrs.argtypes = [_P]
        L.orc_gj_predict.argtypes = [_P, _P, _P, _D, _P, _P]
        L.orc_gj_finalize_correction.argtypes = [_P, _P, _P, _P, _P, _D]
        L.orc_startup_piece.argtypes = [_I, _I, _D, _P, _P, _P, _P, _P]
        L.orc_rk8_nodes.argtypes = [_I, _I, _P, _D, _D, _P, _P, _P, _P]
        L.orc_startup.argtypes = [_I, _I, _P, _D, _D, _I, _P, _P, _P, _P, _P, _P, _P]
        L.orc_propagate.argtypes = [_I, _I, _P, _D, ctypes.c_long, _P, _P, _P, _P, _I, _P, _P,
                                    _P, _I]
        L.orc_partition.argtypes = [_P, _I, _P, _P, _P, _P]
        L.orc_partition_idx.argtypes = [_P, _I, ctypes.c_uint, ctypes.c_uint]
        L.orc_partition_pair.argtypes = [_P, _I, ctypes.c_uint, _P, _P]
        L.orc_hermite.argtypes = [_P, _D, _P]
        L.orc_hermite_eval.restype = _D
        L.orc_hermite_eval.argtypes = [_P, _D, _D, _I]
        L.orc_find_root_poly.restype = _D
        L.orc_find_root_poly.argtypes = [_P, _I, _D, _D, _D]
        L.orc_approach_func.restype = _D
        L.orc_approach_func.argtypes = [_P, _P, _D, _D, _D, _D, _I]
        L.orc_screen.argtypes = [_P, _I, _I, _I, _P, _P, _P]
        L.orc_screen_blocked.argtypes = [_P, _I, _I, _I, _P, _D, _P, _P, _I]
        L.orc_refine.argtypes = [_P, _I, _I, _D, ctypes.c_uint, ctypes.c_uint, _I, _D, _P]
        self.loaded = False
        if degree is not None:
            self.open(degree, data_dir)

    # -- datasets ------------------------------------------------------------------------
    def open(self, degree=-1, data_dir=DATA_DIR):
        if self.L.orc_open(data_dir.encode(), degree) != 0:
            raise RuntimeError(self.L.orc_error().decode())
        self.loaded = True
        self._key = (data_dir, degree)
        _loaded["key"] = self._key

    def _ensure(self):
        if self.loaded and _loaded["key"] != self._key:
            self.open(self._key[1], self._key[0])

    @property
    def degree(self):
        self._ensure()
        return self.L.orc_degree()

    @property
    def gm(self):
        self._ensure()
        return self.L.orc_gm()

    @property
    def radius(self):
        self._ensure()
        return self.L.orc_radius()

    # -- frames --------------------------------------------------------------------------
    def calendar_to_mjd(self, y, mo, d, h, mi, sec):
        return self.L.orc_calendar_to_mjd(y, mo, d, h, mi, sec)

    def ntp_to_mjd(self, ntp):
        return self.L.orc_ntp_to_mjd(ntp)

    def ftable(self, t):
        F = np.zeros(14)
        self.L.orc_ftable(t, _ptr(F))
        return F

    def precession(self, t):
        o = np.zeros(4)
        self.L.orc_precession(t, _ptr(o))
        return o

    def nutation(self, t):
        self._ensure()
        o = np.zeros(2)
        self.L.orc_nutation(t, _ptr(o))
        return o

    def gast(self, mjd_ut1, mjd_tt, dpsi, eps_a, F):
        F = np.ascontiguousarray(F, dtype=np.float64)
        return self.L.orc_gast(mjd_ut1, mjd_tt, dpsi, eps_a, _ptr(F))

    def rot(self, axis, angle):
        M = np.zeros(9)
        self.L.orc_rot(axis, angle, _ptr(M))
        return M

    def eop(self, mjd):
        self._ensure()
        o = np.zeros(4)
        self.L.orc_eop(mjd, _ptr(o))
        return o

    def leap_seconds(self, mjd):
        self._ensure()
        return self.L.orc_leap_seconds(mjd)

    def time_nodes(self, epoch, shifted_sec, offsets):
        self._ensure()
        ymdhm = np.array(epoch[:5], dtype=np.int32)
        offsets = np.ascontiguousarray(offsets, dtype=np.float64)
        M = np.zeros((len(offsets), 9))
        mjd = np.zeros(len(offsets))
        if self.L.orc_time_nodes(_ptr(ymdhm), shifted_sec, _ptr(offsets), len(offsets), _ptr(M),
                                 _ptr(mjd)) != 0:
            raise RuntimeError(self.L.orc_error().decode())
        return M, mjd

    # -- gravity -------------------------------------------------------------------------
    def gravity_itrs(self, r):
        self._ensure()
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
        a = np.zeros_like(r)
        if self.L.orc_gravity_itrs(_ptr(r), len(r), _ptr(a)) != 0:
            raise RuntimeError(self.L.orc_error().decode())
        return a

    def potential_noncentral(self, r):
        self._ensure()
        r = np.ascontiguousarray(r, dtype=np.float64)
        return self.L.orc_potential_itrs(_ptr(r))

    # -- integrator pieces -----------------------------------------------------------------
    def set_sources(self, mask=0, drag_bc=0.0):
        """Sources added to gravity in the propagations that follow (bits 1 Sun, 2 Moon, 4 drag)."""
        self.L.orc_set_sources(mask, drag_bc)

    def sun_moon(self, mjd_tt):
        out = np.zeros(6)
        self.L.orc_sun_moon(mjd_tt, _ptr(out))
        return out[:3], out[3:]

    def density(self, h_m):
        return self.L.orc_density(h_m)

    def sources_accel(self, mask, drag_bc, tb6, r, v):
        tb6 = np.ascontiguousarray(tb6, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        a = np.zeros(3)
        self.L.orc_sources_accel(mask, drag_bc, _ptr(tb6), _ptr(r), _ptr(v), _ptr(a))
        return a

    def philox4x32_10(self, ctr, key):
        """One Philox4x32-10 block: four uint32 counter words and two key words."""
        c = np.ascontiguousarray(ctr, dtype=np.uint32)
        k = np.ascontiguousarray(key, dtype=np.uint32)
        out = np.zeros(4, np.uint32)
        self.L.orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
        return out

    def perturbed(self, r0, v0, n, strength, seed=1234):
        """(r, v) [n][3] of the Monte-Carlo population (generator's normals restated)."""
        r0 = np.ascontiguousarray(r0, dtype=np.float64)
        v0 = np.ascontiguousarray(v0, dtype=np.float64)
        r = np.zeros((n, 3))
        v = np.zeros((n, 3))
        self.L.orc_perturbed(_ptr(r0), _ptr(v0), n, strength, seed, _ptr(r), _ptr(v))
        return r, v

    def gj_predict(self, acc27, s3, S3, h):
        acc = np.ascontiguousarray(acc27, dtype=np.float64)
        s = np.ascontiguousarray(s3, dtype=np.float64)
        S = np.array(S3, dtype=np.float64)
        r = np.zeros(3)
        v = np.zeros(3)
        self.L.orc_gj_predict(_ptr(acc), _ptr(s), _ptr(S), h, _ptr(r), _ptr(v))
        return r, v, S

    def gj_finalize_correction(self, r3, v3, acc27, s3, S3, h):
        r = np.array(r3, dtype=np.float64)
        v = np.array(v3, dtype=np.float64)
        acc = np.ascontiguousarray(acc27, dtype=np.float64)
        s = np.ascontiguousarray(s3, dtype=np.float64)
        S = np.ascontiguousarray(S3, dtype=np.float64)
        self.L.orc_gj_finalize_correction(_ptr(r), _ptr(v), _ptr(acc), _ptr(s), _ptr(S), h)
        return r, v

    def startup_piece(self, mode, kind, h, pos27, vel27):
        self._ensure()
        pos = np.array(pos27, dtype=np.float64)
        vel = np.array(vel27, dtype=np.float64)
        acc, s, S = np.zeros(27), np.zeros(27), np.zeros(27)
        if self.L.orc_startup_piece(mode, kind, h, _ptr(pos), _ptr(vel), _ptr(acc), _ptr(s),
                                    _ptr(S)) != 0:
            raise RuntimeError(self.L.orc_error().decode())
        return dict(pos=pos, vel=vel, acc=acc, s=s, S=S)

    def rk8_nodes(self, kind, h, r0, v0, quirks=REFERENCE_QUIRKS, epoch=(2025, 1, 1, 0, 0, 0.0)):
        self._ensure()
        ymdhm = np.array(epoch[:5], dtype=np.int32)
        r0 = np.ascontiguousarray(r0, dtype=np.float64)
        v0 = np.ascontiguousarray(v0, dtype=np.float64)
        pos, vel = np.zeros(27), np.zeros(27)
        if self.L.orc_rk8_nodes(kind, quirks, _ptr(ymdhm), float(epoch[5]), h, _ptr(r0),
                                _ptr(v0), _ptr(pos), _ptr(vel)) != 0:
            raise RuntimeError(self.L.orc_error().decode())
        return pos.reshape(9, 3), vel.reshape(9, 3)

    def startup(self, kind, h, r0, v0, n_corr=7, quirks=REFERENCE_QUIRKS,
                epoch=(2025, 1, 1, 0, 0, 0.0)):
        self._ensure()
        ymdhm = np.array(epoch[:5], dtype=np.int32)
        r0 = np.ascontiguousarray(r0, dtype=np.float64)
        v0 = np.ascontiguousarray(v0, dtype=np.float64)
        out = [np.zeros(27) for _ in range(5)]
        if self.L.orc_startup(kind, quirks, _ptr(ymdhm), float(epoch[5]), h, n_corr, _ptr(r0),
                              _ptr(v0), *[_ptr(o) for o in out]) != 0:
            raise RuntimeError(self.L.orc_error().decode())
        return dict(zip(("pos", "vel", "s", "S", "acc"), [o.reshape(9, 3) for o in out]))

    def propagate(self, kind, r0, v0, h, steps, n_corr=7, quirks=REFERENCE_QUIRKS,
                  epoch=(2025, 1, 1, 0, 0, 0.0), trajectory=False, threads=0):
        self._ensure()
        r0 = np.ascontiguousarray(r0, dtype=np.float64).reshape(-1, 3)
        v0 = np.ascontiguousarray(v0, dtype=np.float64).reshape(-1, 3)
        n = len(r0)
        h = np.broadcast_to(np.asarray(h, dtype=np.float64), (n,)).copy()
        steps = np.broadcast_to(np.asarray(steps, dtype=np.int32), (n,)).copy()
        ymdhm = np.array(epoch[:5], dtype=np.int32)
        r = np.zeros_like(r0)
        v = np.zeros_like(v0)
        traj = None
        if trajectory:
            if len(set(steps.tolist())) != 1:
                raise ValueError("trajectory output needs equal step counts")
            traj = np.zeros((int(steps[0]), n, 9))
        rc = self.L.orc_propagate(kind, quirks, _ptr(ymdhm), float(epoch[5]), n, _ptr(r0),
                                  _ptr(v0), _ptr(h), _ptr(steps), n_corr, _ptr(r), _ptr(v),
                                  _ptr(traj), threads)
        if rc != 0:
            raise RuntimeError(self.L.orc_error().decode())
        return (r, v, traj) if trajectory else (r, v)

    # -- conjunction ---------------------------------------------------------------------
    def partition(self, set_ids):
        ids = np.ascontiguousarray(set_ids, dtype=np.uint32)
        n = len(ids)
        so, nh, mo = (np.zeros(max(n, 1), dtype=np.uint32) for _ in range(3))
        npairs = np.zeros(1, dtype=np.uint32)
        k = self.L.orc_partition(_ptr(ids), n, _ptr(so), _ptr(nh), _ptr(mo), _ptr(npairs))
        if k < 0:
            return None
        return dict(num_sets=k, set_offsets=so[:k].tolist(), num_higher=nh[:k].tolist(),
                    map_offsets=mo[:k].tolist(), num_pairs=int(npairs[0]))

    def partition_idx(self, set_ids, a, b):
        ids = np.ascontiguousarray(set_ids, dtype=np.uint32)
        return self.L.orc_partition_idx(_ptr(ids), len(ids), a, b)

    def partition_pair(self, set_ids, idx):
        ids = np.ascontiguousarray(set_ids, dtype=np.uint32)
        a = np.zeros(1, dtype=np.uint32)
        b = np.zeros(1, dtype=np.uint32)
        ok = self.L.orc_partition_pair(_ptr(ids), len(ids), idx, _ptr(a), _ptr(b))
        return (bool(ok == 1), int(a[0]), int(b[0]))

    def hermite(self, pva9, dt):
        q = np.ascontiguousarray(pva9, dtype=np.float64)
        c = np.zeros(9)
        self.L.orc_hermite(_ptr(q), dt, _ptr(c))
        return c

    def hermite_eval(self, coeffs, dt, t, deriv=False):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        return self.L.orc_hermite_eval(_ptr(c), dt, t, 1 if deriv else 0)

    def screen(self, traj, set_ids):
        traj = np.ascontiguousarray(traj, dtype=np.float64)
        nsteps, nobj, comps = traj.shape
        ids = np.ascontiguousarray(set_ids, dtype=np.uint32)
        part = self.partition(ids)
        npairs = part["num_pairs"]
        mind = np.zeros(npairs)
        mink = np.zeros(npairs, dtype=np.int32)
        if self.L.orc_screen(_ptr(traj), comps, nobj, nsteps, _ptr(ids), _ptr(mind),
                             _ptr(mink)) != 0:
            raise RuntimeError("invalid set ids")
        return mind, mink

    def screen_blocked(self, traj, set_ids, threshold=0.0, threads=0):
        """orc_screen's rule in a blocked loop order: (min_dist, min_step) per pair, with the step
        only for pairs below threshold (every pair when threshold <= 0) and -1 for the others."""
        traj = np.ascontiguousarray(traj, dtype=np.float64)
        nsteps, nobj, comps = traj.shape
        ids = np.ascontiguousarray(set_ids, dtype=np.uint32)
        npairs = self.partition(ids)["num_pairs"]
        mind = np.zeros(npairs)
        mink = np.zeros(npairs, dtype=np.int32)
        if self.L.orc_screen_blocked(_ptr(traj), comps, nobj, nsteps, _ptr(ids), threshold,
                                     _ptr(mind), _ptr(mink), threads) != 0:
            raise RuntimeError("invalid set ids")
        return mind, mink

    def find_root_poly(self, coef, x0, x1, tol=1e-7):
        c = np.ascontiguousarray(coef, dtype=np.float64)
        return self.L.orc_find_root_poly(_ptr(c), len(c), x0, x1, tol)

    def approach_func(self, qa, qb, dt, t):
        a = np.ascontiguousarray(qa, dtype=np.float64).reshape(27)
        b = np.ascontiguousarray(qb, dtype=np.float64).reshape(27)
        return self.L.orc_approach_func(_ptr(a), _ptr(b), dt, t, 0.0, 0.0, 0)

    def approach_root(self, qa, qb, dt, x0, x1, tol=1e-7):
        a = np.ascontiguousarray(qa, dtype=np.float64).reshape(27)
        b = np.ascontiguousarray(qb, dtype=np.float64).reshape(27)
        return self.L.orc_approach_func(_ptr(a), _ptr(b), dt, x0, x1, tol, 1)

    def refine(self, traj, h, a, b, k, min_dist):
        traj = np.ascontiguousarray(traj, dtype=np.float64)
        nsteps, nobj, _ = traj.shape
        out = np.zeros(8)
        self.L.orc_refine(_ptr(traj), nobj, nsteps, h, a, b, k, min_dist, _ptr(out))
        return dict(distance=out[0], time=out[1], pos_a=out[2:5].copy(), pos_b=out[5:8].copy())
