"""ctypes binding of the FP64 CPU oracle (TEST INFRASTRUCTURE).

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may import this
module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
sys.path.insert(0, str(HERE.parent / "esm-sampling_amd"))
from esmsampling import _lib as L  # noqa: E402  (the esm_model struct layout only)

NK, NV, NREN, NPARTS = L.NK, L.NV, L.NREN, L.NPARTS
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)

_lib = None


NFAM = 15   # move families of the statistics (include/esm.h ESM_NFAM)


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `python oracle/build.py`")
        lib = C.CDLL(str(LIB_PATH))
        M = C.POINTER(L.EsmModel)
        lib.orc_energy.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        lib.orc_scores.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        lib.orc_langevin.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                     _dp, _ip]
        lib.orc_log_measure.argtypes = [M, _dp, C.c_int]
        lib.orc_log_measure.restype = C.c_double
        lib.orc_log_measure_grad.argtypes = [M, _dp, C.c_int, _dp]
        lib.orc_log_measure_grad.restype = None
        lib.orc_philox.argtypes = [_u32p, _u32p, _u32p]
        lib.orc_philox.restype = None
        lib.orc_weather_init.argtypes = [M, C.c_int, _u32p, C.c_uint32, _ip, _dp, _dp]
        lib.orc_weather_step.argtypes = [M, C.c_int, _dp, _dp, C.c_int, C.c_int, _u32p, C.c_uint32, _ip, _dp, _dp,
                                         _dp, _dp, _ip]
        lib.orc_weather_ffbs.argtypes = [M, C.c_int, _dp, _ip, _ip, _u32p, C.c_uint32, _dp]
        lib.orc_categorical.argtypes = [_dp, C.c_int, C.c_double]
        lib.orc_categorical.restype = C.c_int
        lib.orc_regime_window_path.argtypes = [M, _ip, C.c_int, C.c_int, _dp]
        lib.orc_regime_window_path.restype = None
        i64p = C.POINTER(C.c_int64)
        lib.orc_sweeps.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, _u32p, C.c_uint32, C.c_double,
                                   C.c_double, _dp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                   C.c_int, i64p, C.c_int]
        lib.orc_sx_slot.argtypes = [M, C.c_int, C.c_int, _ip, _ip, _ip, _dp, _ip, _ip, _dp, _ip, _ip]
        lib.orc_sx_partner.argtypes = [M, _dp, C.c_double, C.c_int, C.c_int, C.c_double, _ip, _ip]
        lib.orc_sx_sizes.argtypes = [M, C.c_int, _ip, _ip]
        lib.orc_sx_move.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_double, _dp, _dp, _dp, _ip]
        lib.orc_ref_sweeps.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, _u32p, C.c_uint32, C.c_double,
                                       C.c_double, _dp, i64p, C.c_int]
        lib.orc_slong_cost.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       _dp, C.c_int, _dp]
        lib.orc_slong_draw_many.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.c_double, _dp, _dp, C.c_int, _dp, _ip]
        lib.orc_slong_apply.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_double]
        lib.orc_sms_cost.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _ip,
                                     C.c_int, _dp, C.c_int, _dp]
        lib.orc_sms_draw_many.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _ip,
                                          C.c_int, C.c_double, _dp, _dp, C.c_int, _dp, _ip]
        lib.orc_sms_apply.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _ip,
                                      C.c_int, C.c_double]
        lib.orc_tables_info.argtypes = [M, _ip, _ip]
        lib.orc_template_nodes.argtypes = [M, C.c_int, _ip, _ip, _ip, _dp, _ip]
        lib.orc_sparse_draw.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_double, C.c_double, C.c_double, _dp, _ip]
        lib.orc_sparse_cost.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp,
                                        C.c_int, _dp]
        lib.orc_sparse_apply.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_double]
        lib.orc_sparse_draw_many.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_double, _dp, _dp, C.c_int, _dp, _ip]
        lib.orc_colouring.argtypes = [M, _ip]
        lib.orc_move_draw.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.c_double, C.c_double, _dp, _ip]
        lib.orc_move_cost.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      _dp, C.c_int, _dp]
        lib.orc_move_apply.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_double]
        lib.orc_follow.argtypes = [M, C.c_int, _dp, _dp, _dp, _dp, C.c_double]
        lib.orc_logseg.argtypes = [C.c_double, C.c_double]
        lib.orc_logseg.restype = C.c_double
        lib.orc_trexp_offset.argtypes = [C.c_double, C.c_double, C.c_double]
        lib.orc_trexp_offset.restype = C.c_double
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Oracle:
    def __init__(self, sys):
        self.sys = sys
        self.cm, self.keep = L.make_cmodel(sys)
        self.lib = load()

    def energy(self, v, k, x, want_caches=True):
        s = self.sys
        R = v.shape[-1]
        v, k, x = (np.ascontiguousarray(a, dtype=np.float64) for a in (v, k, x))
        parts = np.empty((R, NPARTS))
        ph = np.empty((s.T, R))
        r = np.empty((s.N, s.T, R)) if want_caches else None
        F = np.empty((s.L, s.T, R)) if want_caches else None
        rc = self.lib.orc_energy(C.byref(self.cm), R, _p(v), _p(k), _p(x), _p(parts), _p(ph), _p(r), _p(F))
        assert rc == 0
        return parts, ph, r, F

    def scores(self, v, k, x, measure="utilization"):
        """(sk [R, NK, N], fk [NK, N], |terms| of sk): ``capacity_scores`` with the reference
        measure ``measure`` ("utilization" or "physical": sk gains ``log_measure_grad``)."""
        s = self.sys
        R = v.shape[-1]
        v, k, x = (np.ascontiguousarray(a, dtype=np.float64) for a in (v, k, x))
        sk = np.empty((R, NK, s.N))
        fk = np.empty((NK, s.N))
        sa = np.empty((R, NK, s.N))
        assert self.lib.orc_scores(C.byref(self.cm), R, _p(v), _p(k), _p(x), _p(sk), _p(fk), _p(sa)) == 0
        g = self.log_measure_grad(k, measure)
        return sk + g[None], fk, sa + np.abs(g)[None]

    def log_measure(self, k, measure="physical"):
        k = np.ascontiguousarray(k, dtype=np.float64)
        return self.lib.orc_log_measure(C.byref(self.cm), _p(k), int(measure == "physical"))

    def log_measure_grad(self, k, measure="physical"):
        k = np.ascontiguousarray(k, dtype=np.float64)
        g = np.empty((NK, self.sys.N))
        self.lib.orc_log_measure_grad(C.byref(self.cm), _p(k), int(measure == "physical"), _p(g))
        return g

    def langevin(self, k, sk, fk, weights, xi, eta, kappa_s, kmin=1.0, clip=0.02, box=True):
        R = sk.shape[0]
        arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (k, sk, fk, weights, xi,
                                                                   np.broadcast_to(eta, (NK,)),
                                                                   np.broadcast_to(kappa_s, (NK,)))]
        knew = np.empty((NK, self.sys.N))
        nc = C.c_int32(0)
        assert self.lib.orc_langevin(C.byref(self.cm), R, *[_p(a) for a in arr], float(kmin), float(clip), int(box),
                                     _p(knew), C.byref(nc)) == 0
        return knew, int(nc.value)


    def weather_init(self, R, key=(1, 0), counter=0):
        s = self.sys
        m = np.empty((s.T, R), dtype=np.int32)
        z = np.empty((NREN * s.N, s.T, R))
        x = np.empty((NREN, s.N, s.T, R))
        k = (C.c_uint32 * 2)(*key)
        assert self.lib.orc_weather_init(C.byref(self.cm), R, k, int(counter), m.ctypes.data_as(_ip), _p(z), _p(x)) == 0
        return m, z, x

    def weather_step(self, m, z, x, v, kcap, theta, win_prob, window=60, compensate=True, key=(1, 0), counter=0):
        """In place on copies: returns (m, z, x, v, moved)."""
        R = m.shape[1]
        m = np.ascontiguousarray(m, dtype=np.int32).copy()
        z, x, v = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in (z, x, v))
        kcap = np.ascontiguousarray(kcap, dtype=np.float64)
        th = np.ascontiguousarray(np.broadcast_to(theta, (R,)), dtype=np.float64)
        wp = np.ascontiguousarray(np.broadcast_to(win_prob, (R,)), dtype=np.float64)
        moved = np.zeros(R, dtype=np.int32)
        k = (C.c_uint32 * 2)(*key)
        assert self.lib.orc_weather_step(C.byref(self.cm), R, _p(th), _p(wp), int(window), int(compensate), k,
                                         int(counter), m.ctypes.data_as(_ip), _p(z), _p(x), _p(v), _p(kcap),
                                         moved.ctypes.data_as(_ip)) == 0
        return m, z, x, v, moved

    def weather_ffbs(self, m, z, mask=None, key=(1, 0), counter=0):
        """Centred FFBS of the regimes (``ffbs``, src/weather.jl:160-174) of the masked replicas
        given z, on a copy: returns (m, log p(z) per replica from the forward filter)."""
        R = m.shape[1]
        m = np.ascontiguousarray(m, dtype=np.int32).copy()
        z = np.ascontiguousarray(z, dtype=np.float64)
        mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
        ll = np.full(R, np.nan)
        kk = (C.c_uint32 * 2)(*key)
        rc = self.lib.orc_weather_ffbs(C.byref(self.cm), R, _p(z), m.ctypes.data_as(_ip),
                                       None if mk is None else mk.ctypes.data_as(_ip), kk, int(counter), _p(ll))
        assert rc == 0, rc
        return m, ll

    def sweeps(self, v, k, nsweeps, key=(1, 0), counter=0, maxnats=40.0, x=None, r=None, F=None, threads=1,
               extra=4.0, weights=None, atoms=True, p_storage_ms=0.5, p_storage_long=1.0, long_moves=32,
               long_hours=24.0, p_xpair=0.0, xlong_moves=0):
        """Operation sweeps in place on copies; returns (v, r, F, stats [NFAM, 5]). ``p_xpair``: the
        weight of the cross-node storage pair family sxpair in the local sweep, ``xlong_moves``:
        the interval pairs sxlong per storage unit in the long-range sweep (0, 0: off)."""
        R = v.shape[-1]
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        k = np.ascontiguousarray(k, dtype=np.float64)
        if r is None or F is None:
            _, _, r, F = self.energy(v, k, x)
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        F = np.ascontiguousarray(F, dtype=np.float64).copy()
        st = np.zeros(NFAM * 5, dtype=np.int64)
        kk = (C.c_uint32 * 2)(*key)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        rc = self.lib.orc_sweeps(C.byref(self.cm), R, _p(v), _p(k), _p(r), _p(F), int(nsweeps), kk, int(counter),
                                 float(maxnats), float(extra), _p(w), int(bool(atoms)), float(p_storage_ms),
                                 float(p_storage_long), int(long_moves), float(long_hours), float(p_xpair),
                                 int(xlong_moves), st.ctypes.data_as(C.POINTER(C.c_int64)), int(threads))
        assert rc == 0, rc
        return v, r, F, st.reshape(NFAM, 5)

    def ref_sweeps(self, v, k, nsweeps, key=(1, 0), counter=0, maxnats=40.0, x=None, r=None, F=None, threads=1,
                   extra=4.0, weights=None):
        """Sweeps in the reference's own schedule (``local_sweep!`` with the plain families: per
        replica and step, the atoms and the family draws in one shuffled list; no storage sweeps),
        in place on copies; returns (v, r, F, stats [NFAM, 5])."""
        R = v.shape[-1]
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        k = np.ascontiguousarray(k, dtype=np.float64)
        if r is None or F is None:
            _, _, r, F = self.energy(v, k, x)
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        F = np.ascontiguousarray(F, dtype=np.float64).copy()
        st = np.zeros(NFAM * 5, dtype=np.int64)
        kk = (C.c_uint32 * 2)(*key)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        rc = self.lib.orc_ref_sweeps(C.byref(self.cm), R, _p(v), _p(k), _p(r), _p(F), int(nsweeps), kk, int(counter),
                                     float(maxnats), float(extra), _p(w), st.ctypes.data_as(C.POINTER(C.c_int64)),
                                     int(threads))
        assert rc == 0, rc
        return v, r, F, st.reshape(NFAM, 5)

    # -- cross-node storage pairs (sxpair: L = 1, W = 1; sxlong: W = 64) -------------------------
    def sx_slot(self, n, k):
        """Slot k of node n's storage-pair templates: dict(partner, th (nodes, values), leak (nodes,
        requirement change per unit), foot) or None past the last slot."""
        N = self.sys.N
        part, nth, nlk, nft = (C.c_int(), C.c_int(), C.c_int(), C.c_int())
        thn, lkn, ft = (np.zeros(N, dtype=np.int32) for _ in range(3))
        thv, lkg = np.zeros(N), np.zeros(N)
        rc = self.lib.orc_sx_slot(C.byref(self.cm), int(n), int(k), C.byref(part), C.byref(nth),
                                  thn.ctypes.data_as(_ip), _p(thv), C.byref(nlk), lkn.ctypes.data_as(_ip), _p(lkg),
                                  C.byref(nft), ft.ctypes.data_as(_ip))
        if rc == -2:
            return None
        assert rc == 0, rc
        return dict(partner=part.value, th=(thn[:nth.value].copy(), thv[:nth.value].copy()),
                    leak=(lkn[:nlk.value].copy(), lkg[:nlk.value].copy()), foot=ft[:nft.value].copy())

    def sx_sizes(self):
        """[slots, 4] per storage-pair slot: kinks, lines, balances, node–line entries per end."""
        cap = 8 * 256 * 16
        out = np.zeros(4 * cap, dtype=np.int32)
        n = C.c_int()
        assert self.lib.orc_sx_sizes(C.byref(self.cm), cap, C.byref(n), out.ctypes.data_as(_ip)) == 0
        return out[:4 * min(n.value, cap)].reshape(-1, 4).copy()

    def sx_partner(self, k, sa, na, u, hmin=24.0):
        """The partner (sb, nb) of unit (sa, na) drawn with the uniform u, or None."""
        sb, nb = C.c_int(), C.c_int()
        k = np.ascontiguousarray(k, dtype=np.float64)
        rc = self.lib.orc_sx_partner(C.byref(self.cm), _p(k), float(hmin), int(sa), int(na), float(u), C.byref(sb),
                                     C.byref(nb))
        assert rc >= 0, rc
        return (sb.value, nb.value) if rc == 1 else None

    def _sx(self, v, k, r, F, j, A, B, a, L, W, op, ts, maxnats=40.0, useg=None, uoff=None):
        R = v.shape[-1]
        ts = np.ascontiguousarray(ts, dtype=np.float64)
        n = len(ts) if useg is None else len(useg)
        out = np.zeros(max(n, 1))
        st = np.zeros(max(n, 1), dtype=np.int32)
        us = None if useg is None else np.ascontiguousarray(useg, dtype=np.float64)
        uo = None if uoff is None else np.ascontiguousarray(uoff, dtype=np.float64)
        rc = self.lib.orc_sx_move(C.byref(self.cm), R, *self._mv_args(v, k, r, F), int(j), int(A[0]), int(A[1]),
                                  int(B[0]), int(B[1]), int(a), int(L), int(W), int(op), _p(ts), n, float(maxnats),
                                  _p(us), _p(uo), _p(out), st.ctypes.data_as(_ip))
        assert rc == 0, rc
        return out[:n], st[:n]

    def sx_cost(self, v, k, r, F, j, A, B, a, L, ts, W=1):
        """ΔC(t) of the pair move of units A = (carrier, node) and B at instants a+1 … a+L."""
        return self._sx(v, k, r, F, j, A, B, a, L, W, 0, ts)[0]

    def sx_draw_many(self, v, k, r, F, j, A, B, a, L, useg, uoff, W=1, maxnats=40.0):
        return self._sx(v, k, r, F, j, A, B, a, L, W, 1, np.zeros(1), maxnats, useg, uoff)

    def sx_apply(self, v, k, r, F, j, A, B, a, L, t, W=1):
        self._sx(v, k, r, F, j, A, B, a, L, W, 2, np.array([float(t)]))

    # -- storage multiscale plateaus -----------------------------------------------------------
    def sms_cost(self, v, k, r, F, j, L, b0, si, nodes, ts):
        R = v.shape[-1]
        nd = np.ascontiguousarray(nodes, dtype=np.int32)
        ts = np.ascontiguousarray(ts, dtype=np.float64)
        out = np.empty_like(ts)
        rc = self.lib.orc_sms_cost(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, L, b0, si,
                                   nd.ctypes.data_as(_ip), len(nd), _p(ts), len(ts), _p(out))
        assert rc == 0, rc
        return out

    def sms_draw_many(self, v, k, r, F, j, L, b0, si, nodes, useg, uoff, maxnats=40.0):
        R = v.shape[-1]
        nd = np.ascontiguousarray(nodes, dtype=np.int32)
        useg = np.ascontiguousarray(useg, dtype=np.float64)
        uoff = np.ascontiguousarray(uoff, dtype=np.float64)
        t = np.empty(len(useg))
        st = np.empty(len(useg), dtype=np.int32)
        rc = self.lib.orc_sms_draw_many(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, L, b0, si,
                                        nd.ctypes.data_as(_ip), len(nd), float(maxnats), _p(useg), _p(uoff), len(useg),
                                        _p(t), st.ctypes.data_as(_ip))
        assert rc == 0, rc
        return t, st

    def sms_apply(self, v, k, r, F, j, L, b0, si, nodes, t):
        R = v.shape[-1]
        nd = np.ascontiguousarray(nodes, dtype=np.int32)
        rc = self.lib.orc_sms_apply(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, L, b0, si,
                                    nd.ctypes.data_as(_ip), len(nd), float(t))
        assert rc == 0, rc

    # -- long-range storage moves (L = T: the level move) -----------------------------------------
    def slong_cost(self, v, k, r, F, j, si, n, a, L, ts):
        R = v.shape[-1]
        ts = np.ascontiguousarray(ts, dtype=np.float64)
        out = np.empty_like(ts)
        rc = self.lib.orc_slong_cost(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, si, n, a, L, _p(ts), len(ts),
                                     _p(out))
        assert rc == 0, rc
        return out

    def slong_draw_many(self, v, k, r, F, j, si, n, a, L, useg, uoff, maxnats=40.0):
        R = v.shape[-1]
        useg = np.ascontiguousarray(useg, dtype=np.float64)
        uoff = np.ascontiguousarray(uoff, dtype=np.float64)
        t = np.empty(len(useg))
        st = np.empty(len(useg), dtype=np.int32)
        rc = self.lib.orc_slong_draw_many(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, si, n, a, L,
                                          float(maxnats), _p(useg), _p(uoff), len(useg), _p(t), st.ctypes.data_as(_ip))
        assert rc == 0, rc
        return t, st

    def slong_apply(self, v, k, r, F, j, si, n, a, L, t):
        R = v.shape[-1]
        rc = self.lib.orc_slong_apply(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, si, n, a, L, float(t))
        assert rc == 0, rc

    # -- table-driven random moves (kind 0 cell, 1 xloc, 2 xlocc, 3 xpair) ----------------------
    def tables_info(self):
        """dict(ntpl, xp0, cl0, nlev, cells per level, xp_off [N + 1])."""
        info = np.zeros(8, dtype=np.int32)
        xp = np.zeros(self.sys.N + 1, dtype=np.int32)
        assert self.lib.orc_tables_info(C.byref(self.cm), info.ctypes.data_as(_ip), xp.ctypes.data_as(_ip)) == 0
        return dict(ntpl=int(info[0]), xp0=int(info[1]), cl0=int(info[2]), nlev=int(info[3]),
                    cells=[int(c) for c in info[4:4 + int(info[3])]], xp_off=xp)

    def template_nodes(self, tid):
        """(touched nodes, leak nodes, their requirement change per unit t) of template tid."""
        N = self.sys.N
        nodes = np.zeros(N, dtype=np.int32)
        kk = np.zeros(N, dtype=np.int32)
        kg = np.zeros(N)
        nn, nk = C.c_int32(), C.c_int32()
        assert self.lib.orc_template_nodes(C.byref(self.cm), int(tid), nodes.ctypes.data_as(_ip), C.byref(nn),
                                           kk.ctypes.data_as(_ip), _p(kg), C.byref(nk)) == 0
        return nodes[:nn.value].copy(), kk[:nk.value].copy(), kg[:nk.value].copy()

    def sparse_draw(self, v, k, r, F, j, kind, tid, tau, useg, uoff, maxnats=40.0):
        R = v.shape[-1]
        t = C.c_double()
        st = C.c_int32()
        rc = self.lib.orc_sparse_draw(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, kind, tid, tau,
                                      float(maxnats), float(useg), float(uoff), C.byref(t), C.byref(st))
        assert rc == 0, rc
        return t.value, st.value

    def sparse_draw_many(self, v, k, r, F, j, kind, tid, tau, useg, uoff, maxnats=40.0):
        R = v.shape[-1]
        useg = np.ascontiguousarray(useg, dtype=np.float64)
        uoff = np.ascontiguousarray(uoff, dtype=np.float64)
        t = np.empty(len(useg))
        st = np.empty(len(useg), dtype=np.int32)
        rc = self.lib.orc_sparse_draw_many(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, kind, tid, tau,
                                           float(maxnats), _p(useg), _p(uoff), len(useg), _p(t), st.ctypes.data_as(_ip))
        assert rc == 0, rc
        return t, st

    def sparse_cost(self, v, k, r, F, j, kind, tid, tau, ts):
        R = v.shape[-1]
        ts = np.ascontiguousarray(ts, dtype=np.float64)
        out = np.empty_like(ts)
        rc = self.lib.orc_sparse_cost(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, kind, tid, tau, _p(ts),
                                      len(ts), _p(out))
        assert rc == 0, rc
        return out

    def sparse_apply(self, v, k, r, F, j, kind, tid, tau, t):
        R = v.shape[-1]
        rc = self.lib.orc_sparse_apply(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, kind, tid, tau, float(t))
        assert rc == 0, rc

    def follow(self, v, k, x, knew, kmin=1.0):
        """follow_capacities! on copies: returns (v, k)."""
        R = v.shape[-1]
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        k = np.ascontiguousarray(k, dtype=np.float64).copy()
        x = np.ascontiguousarray(x, dtype=np.float64)
        kn = np.ascontiguousarray(knew, dtype=np.float64)
        assert self.lib.orc_follow(C.byref(self.cm), R, _p(v), _p(k), _p(x), _p(kn), float(kmin)) == 0
        return v, k

    def colouring(self):
        col = np.zeros(self.sys.N, dtype=np.int32)
        nc = self.lib.orc_colouring(C.byref(self.cm), col.ctypes.data_as(_ip))
        return col, nc

    def _mv_args(self, v, k, r, F):
        return [_p(v), _p(k), _p(r), _p(F)]

    def move_draw(self, v, k, r, F, j, fam, s, n, tau, useg, uoff, maxnats=40.0):
        R = v.shape[-1]
        t = C.c_double()
        st = C.c_int32()
        rc = self.lib.orc_move_draw(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, fam, s, n, tau,
                                    float(maxnats), float(useg), float(uoff), C.byref(t), C.byref(st))
        assert rc == 0, rc
        return t.value, st.value

    def move_cost(self, v, k, r, F, j, fam, s, n, tau, ts):
        R = v.shape[-1]
        ts = np.ascontiguousarray(ts, dtype=np.float64)
        out = np.empty_like(ts)
        rc = self.lib.orc_move_cost(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, fam, s, n, tau, _p(ts),
                                    len(ts), _p(out))
        assert rc == 0, rc
        return out

    def move_apply(self, v, k, r, F, j, fam, s, n, tau, t):
        R = v.shape[-1]
        rc = self.lib.orc_move_apply(C.byref(self.cm), R, *self._mv_args(v, k, r, F), j, fam, s, n, tau, float(t))
        assert rc == 0, rc

    def regime_window_path(self, path, a, b, us):
        path = np.ascontiguousarray(path, dtype=np.int32).copy()
        us = np.ascontiguousarray(us, dtype=np.float64)
        self.lib.orc_regime_window_path(C.byref(self.cm), path.ctypes.data_as(_ip), int(a), int(b), _p(us))
        return path


def categorical(p, u):
    lib = load()
    p = np.ascontiguousarray(p, dtype=np.float64)
    return lib.orc_categorical(_p(p), len(p), float(u))


def philox(key, ctr):
    lib = load()
    k = (C.c_uint32 * 2)(*key)
    c = (C.c_uint32 * 4)(*ctr)
    o = (C.c_uint32 * 4)()
    lib.orc_philox(k, c, o)
    return list(o)


def weather_features(load, x, kref, m, K):
    """The replica weather features (``replica_features``, src/replicas.jl:270-273, over
    ``weather_features``, :250-267), numpy FP64, replica by replica.

    load [N, T] (the system's load; here not a weather series), x [3, N, T, R] the complete-level
    weather after the data map, kref [6, N] the reference capacities (fields 0-2 read), m [T, R]
    regimes (0-based), K regimes. Per replica: D = Σ_n max(load, 0) (:253), for each weather
    carrier A = clamp(x, 0, 1), cf = mean(A) and R += kref' A (:256-261), res = (D − R)/1e3
    (:262), w7 = max_t mean(res[t … t+6 cyclic]) (:264), then
    [mean(D)/1e3, cf…, mean(res), max(res), w7, count(m == k)/T…, count(diff(m) != 0)] (:271-273).
    """
    load = np.asarray(load, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    kref = np.asarray(kref, dtype=np.float64)
    m = np.asarray(m)
    T, R = m.shape
    D = np.maximum(load, 0.0).sum(axis=0)
    rows = []
    for j in range(R):
        Rt = np.zeros(T)
        cf = []
        for c in range(NREN):
            A = np.clip(x[c, :, :, j], 0.0, 1.0)
            cf.append(A.mean())
            Rt += kref[c] @ A
        res = (D - Rt) / 1e3
        idx = (np.arange(T)[:, None] + np.arange(7)[None, :]) % T
        w7 = res[idx].mean(axis=1).max()
        mj = m[:, j]
        rows.append([D.mean() / 1e3, *cf, res.mean(), res.max(), w7,
                     *[np.count_nonzero(mj == k) / T for k in range(K)], float(np.count_nonzero(np.diff(mj) != 0))])
    return np.array(rows)
