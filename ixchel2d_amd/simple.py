"""Python host for the SIMPLE iteration on one MI355X.

``Cavity`` owns one device state (``ixh_state``) and exposes the phases of
upstream's time-step loop body (source/main.f90:391-1271) under the names of
the C-ABI; ``Cavity.run`` is the whole loop with upstream's convergence
control.  Every method goes through libixchel2d.so; nothing here computes a
field on the CPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np

from . import _lib
from .case import Case

_DP = C.POINTER(C.c_double)


def _f64(a) -> np.ndarray:
    return np.require(np.asarray(a, dtype=np.float64), requirements=["F", "A", "W"])


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def params_of(case: Case, solver: str = "pcr", precision: str = "fp64") -> _lib.Params:
    p = case.params
    d = p.derived()
    return _lib.Params(p.dt, p.rel_pres, p.rel_vel, p.rel_ener, d["gamma_momen"], d["gamma_energ"],
                       d["Ri"], d["Riy"], p.conv_u, p.conv_t, p.conv_p, p.conv_resi, p.conv_paso,
                       p.iter_ecuaci_max, p.iter_simple_max, _lib.SOLVER_ID[solver],
                       _lib.PRECISION_ID[precision])


class Cavity:
    """One cavity resident on the GPU."""

    def __init__(self, case: Case, solver: str = "pcr", precision: str = "fp64", upload: bool = True):
        L = _lib.load()
        self.L = L
        self.case = case
        m = case.mesh
        self.mi, self.nj = m.mi, m.nj
        self._geo = [np.ascontiguousarray(x, dtype=np.float64) for x in (m.xu, m.yv, m.xp, m.yp)]
        grid = _lib.Grid(m.mi, m.nj, *[_ptr(x) for x in self._geo])
        self._params = params_of(case, solver, precision)
        h = C.c_void_p()
        _lib.check(L.ixh_create(C.byref(grid), C.byref(self._params), C.byref(h)))
        self.h = h
        for var, bset in case.bcs.items():
            for side, (typ, val) in bset.expand(m).items():
                t = np.ascontiguousarray(typ, dtype=np.int32)
                v = np.ascontiguousarray(val, dtype=np.float64)
                _lib.check(L.ixh_set_bc(self.h, _lib.VAR_ID[var], _lib.SIDE_ID[side],
                                        t.ctypes.data_as(C.POINTER(C.c_int32)), _ptr(v)))
        for var, (sc, sp) in case.sources.items():
            self.set_sources(var, sc, sp)
        if upload:
            u, v, t, p = case.initial_fields()
            for name, arr in (("u_ant", u), ("v_ant", v), ("t_ant", t), ("u", u), ("v", v), ("t", t), ("p", p)):
                self.set(name, arr)

    def close(self):
        if getattr(self, "h", None):
            self.L.ixh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- state
    def shape(self, name: str):
        mi, nj = self.mi, self.nj
        if name in ("u", "u_ant", "au", "resu"):
            return (mi, nj + 1)
        if name in ("v", "v_ant", "av"):
            return (mi + 1, nj)
        return (mi + 1, nj + 1)

    def set(self, name: str, arr) -> None:
        a = _f64(arr)
        if a.shape != self.shape(name):
            raise ValueError(f"{name}: shape {a.shape} != {self.shape(name)}")
        _lib.check(self.L.ixh_upload_field(self.h, _lib.FIELD_ID[name], _ptr(a), a.shape[0]))

    def get(self, name: str) -> np.ndarray:
        out = np.zeros(self.shape(name), dtype=np.float64, order="F")
        _lib.check(self.L.ixh_download_field(self.h, _lib.FIELD_ID[name], _ptr(out), out.shape[0]))
        return out

    def fields(self) -> Dict[str, np.ndarray]:
        return {n: self.get(n) for n in _lib.FIELD_ID}

    def set_bc(self, var: str, side: str, types, values) -> None:
        """Per-node boundary types (IXH_BC_*) and values of one side."""
        t = np.ascontiguousarray(types, dtype=np.int32)
        v = np.ascontiguousarray(values, dtype=np.float64)
        _lib.check(self.L.ixh_set_bc(self.h, _lib.VAR_ID[var], _lib.SIDE_ID[side],
                                     t.ctypes.data_as(C.POINTER(C.c_int32)), _ptr(v)))

    def set_sources(self, var: str, sc, sp) -> None:
        if sc is None and sp is None:
            _lib.check(self.L.ixh_set_sources(self.h, _lib.VAR_ID[var], None, None))
            return
        a, b = _f64(sc), _f64(sp)
        _lib.check(self.L.ixh_set_sources(self.h, _lib.VAR_ID[var], _ptr(a), _ptr(b)))

    def set_params(self, **kw) -> None:
        for k, v in kw.items():
            if k == "solver":
                v = _lib.SOLVER_ID[v]
            setattr(self._params, k, v)
        _lib.check(self.L.ixh_set_params(self.h, C.byref(self._params)))

    def set_stream(self, stream_handle: Optional[int]) -> None:
        _lib.check(self.L.ixh_set_stream(self.h, C.c_void_p(stream_handle) if stream_handle else None))

    def synchronize(self) -> None:
        _lib.check(self.L.ixh_synchronize(self.h))

    # --------------------------------------------------------------- phases
    def momentum_u(self, direction: int) -> None:
        _lib.check(self.L.ixh_momentum_u(self.h, direction))

    def momentum_v(self, direction: int) -> Optional[float]:
        e = C.c_double()
        _lib.check(self.L.ixh_momentum_v(self.h, direction, C.byref(e) if direction == _lib.DIR_X else None))
        return e.value if direction == _lib.DIR_X else None

    def pcorr_zero(self) -> None:
        _lib.check(self.L.ixh_pcorr_zero(self.h))

    def pcorr(self, direction: int):
        e, b = C.c_double(), C.c_double()
        if direction == _lib.DIR_X:
            _lib.check(self.L.ixh_pcorr(self.h, direction, C.byref(e), C.byref(b)))
            return e.value, b.value
        _lib.check(self.L.ixh_pcorr(self.h, direction, None, None))
        return None

    def correct(self) -> None:
        _lib.check(self.L.ixh_correct(self.h))

    def energy(self, direction: int) -> Optional[float]:
        e = C.c_double()
        _lib.check(self.L.ixh_energy(self.h, direction, C.byref(e) if direction == _lib.DIR_X else None))
        return e.value if direction == _lib.DIR_X else None

    def residual_u(self) -> float:
        r = C.c_double()
        _lib.check(self.L.ixh_residual_u(self.h, C.byref(r)))
        return r.value

    def lag(self) -> None:
        _lib.check(self.L.ixh_lag(self.h))

    def run(self, nsteps: int, cap: Optional[int] = None) -> Dict[str, np.ndarray]:
        """``nsteps`` time steps of the upstream loop; one record per SIMPLE
        iteration."""
        cap = max(1, nsteps * 4 if cap is None else cap)
        hist = (_lib.Hist * cap)()
        n = C.c_int32()
        _lib.check(self.L.ixh_simple_iterations(self.h, nsteps, hist, cap, C.byref(n)))
        return _records(hist, 0, min(n.value, cap))

    def nusselt_y(self, i0: Optional[int] = None, i1: Optional[int] = None):
        m = self.case.mesh
        a, b = C.c_double(), C.c_double()
        _lib.check(self.L.ixh_nusselt_y(self.h, m.placa_min if i0 is None else i0,
                                        m.placa_max if i1 is None else i1, C.byref(a), C.byref(b)))
        return a.value, b.value

    def nusselt_x(self):
        a, b = C.c_double(), C.c_double()
        _lib.check(self.L.ixh_nusselt_x(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ------------------------------------------------------------- adjoint
    def adjoint_record(self, nsteps: int, every: int = 20) -> None:
        """Run ``nsteps`` single-pass steps keeping a checkpoint every
        ``every`` steps for ``adjoint_energy``."""
        _lib.check(self.L.ixh_adjoint_record(self.h, nsteps, every))

    def adjoint_energy(self, tbar, coupled: bool = False):
        """Backward pass of the recorded steps for dJ/dT(final) = ``tbar``:
        returns ({side: dJ/d(boundary values)}, dJ/dT(initial)).  With
        ``coupled`` the adjoint runs through the whole step (exact under
        buoyancy); otherwise through the energy steps with the velocities
        held fixed."""
        tb = _f64(tbar)
        if tb.shape != self.shape("t"):
            raise ValueError(f"tbar: shape {tb.shape} != {self.shape('t')}")
        mi, nj = self.mi, self.nj
        g = {"a": np.zeros(nj + 1), "b": np.zeros(mi + 1), "c": np.zeros(nj + 1), "d": np.zeros(mi + 1)}
        t0 = np.zeros(self.shape("t"), order="F")
        fn = self.L.ixh_adjoint_coupled if coupled else self.L.ixh_adjoint_energy
        _lib.check(fn(self.h, _ptr(tb), tb.shape[0], _ptr(g["a"]), _ptr(g["b"]), _ptr(g["c"]), _ptr(g["d"]),
                      _ptr(t0), t0.shape[0]))
        return g, t0

    # -------------------------------------------------------------- timing
    def set_profiling(self, on: bool) -> None:
        _lib.check(self.L.ixh_set_profiling(self.h, 1 if on else 0))

    def kernel_times(self) -> Dict[str, tuple]:
        ms = (C.c_double * 32)()
        cnt = (C.c_int32 * 32)()
        nk = C.c_int32()
        _lib.check(self.L.ixh_kernel_times(self.h, ms, cnt, 32, C.byref(nk)))
        return {self.L.ixh_kernel_name(k).decode(): (ms[k], cnt[k]) for k in range(nk.value)}


# ixh_hist as a numpy record (5 doubles, 6 int32: 64 bytes)
_HIST_DTYPE = np.dtype([(f, "<f8") for f in ("err_v", "err_p", "maxbo", "err_t", "residuo")] +
                       [(f, "<i4") for f in ("passes_mom", "passes_p", "passes_t", "simple_iter", "logged", "pad_")])
assert _HIST_DTYPE.itemsize == C.sizeof(_lib.Hist)


def _records(hist, k0: int, k: int) -> Dict[str, np.ndarray]:
    """Records k0 .. k0+k-1 of an ixh_hist array as one array per field."""
    a = np.frombuffer(hist, dtype=_HIST_DTYPE, count=k, offset=k0 * _HIST_DTYPE.itemsize)
    rec = {f: a[f].astype(np.float64) for f in ("err_v", "err_p", "maxbo", "err_t", "residuo")}
    for f in ("passes_mom", "passes_p", "passes_t", "simple_iter", "logged"):
        rec[f] = a[f].astype(np.int32)
    return rec


class Batch:
    """Many cavities advanced together (``ixh_batch``): one launch per phase
    covers the lines of every cavity.  The cavities stay usable on their own
    (fields, Nusselt numbers) and must outlive the batch."""

    def __init__(self, cavities, order: str = "xcd", graph: bool = False):
        L = _lib.load()
        self.L = L
        self.cavities = list(cavities)
        arr = (C.c_void_p * len(self.cavities))(*[c.h.value for c in self.cavities])
        h = C.c_void_p()
        _lib.check(L.ixh_batch_create(arr, len(self.cavities), C.byref(h)))
        self.h = h
        self.set_schedule(order, graph)

    def set_schedule(self, order: str = "xcd", graph: bool = False) -> None:
        _lib.check(self.L.ixh_batch_set_schedule(self.h, _lib.ORDER_ID[order], 1 if graph else 0))

    def set_shape(self, line_threads: int = 64, lines_per_group: int = _lib.BATCH_LINE_GROUP,
                  residual_with_t: bool = False) -> None:
        """Threads per line (64 or 256), lines per workgroup in the sweeps
        (1 or ``_lib.BATCH_LINE_GROUP``) and whether the u residual runs in
        the T x-sweep's launch (``ixh_batch_set_shape``)."""
        _lib.check(self.L.ixh_batch_set_shape(self.h, line_threads, lines_per_group, 1 if residual_with_t else 0))

    def autotune(self, trial_steps: int = 10):
        """Time every schedule the batch allows (ordering, graph, threads per
        line, lines per workgroup, residual placement), keep the fastest.
        Returns (schedule, table): schedule = dict(order, graph,
        line_threads, lines_per_group, residual_with_t), table = {(order,
        graph, line_threads, lines_per_group, residual_with_t): ms per step}
        over the schedules tried."""
        o, g, lt, lg, rt = (C.c_int32() for _ in range(5))
        ms = (C.c_double * _lib.BATCH_SCHEDULES)()
        _lib.check(self.L.ixh_batch_autotune(self.h, trial_steps, C.byref(o), C.byref(g), C.byref(lt),
                                             C.byref(lg), C.byref(rt), ms))
        names = {v: k for k, v in _lib.ORDER_ID.items()}
        table = {}
        for k in range(_lib.BATCH_SCHEDULES):
            if ms[k] >= 0.0:
                table[(names[k >> 4], bool((k >> 3) & 1), 256 if (k >> 2) & 1 else 64,
                       _lib.BATCH_LINE_GROUP if (k >> 1) & 1 else 1, bool(k & 1))] = ms[k]
        sched = {"order": names[o.value], "graph": bool(g.value), "line_threads": lt.value,
                 "lines_per_group": lg.value, "residual_with_t": bool(rt.value)}
        return sched, table

    def run(self, nsteps: int, cap: Optional[int] = None):
        """``nsteps`` single-pass time steps of every cavity; one history dict
        per cavity (the first ``cap`` steps)."""
        cap = max(1, nsteps if cap is None else cap)
        n = len(self.cavities)
        hist = (_lib.Hist * (n * cap))()
        _lib.check(self.L.ixh_batch_iterations(self.h, nsteps, hist, cap))
        k = min(nsteps, cap)
        return [_records(hist, c * cap, k) for c in range(n)]

    def close(self):
        if getattr(self, "h", None):
            self.L.ixh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tridiagonal(a, b, c, r, solver: str = "pcr"):
    """Batched ``tridiagonal(a,b,c,r,n)`` on the device; the last axis is the
    system.  Returns (solution, overwritten b)."""
    L = _lib.load()
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64).copy()
    c = np.ascontiguousarray(c, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64).copy()
    n = r.shape[-1]
    nsys = r.size // n if n else 0
    _lib.check(L.ixh_tridiagonal(_ptr(a), _ptr(b), _ptr(c), _ptr(r), n, nsys, _lib.SOLVER_ID[solver]))
    return r, b
