"""ctypes wrapper over the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FIELDS = ["u", "v", "t", "p", "u_ant", "v_ant", "t_ant", "pcorr", "bo", "au", "av", "resu"]
FIELD_ID = {n: k for k, n in enumerate(FIELDS)}
VAR_ID = {"u": 0, "v": 1, "t": 2}
SIDE_ID = {"a": 0, "b": 1, "c": 2, "d": 3}


class Hist(C.Structure):
    _fields_ = [("err_v", C.c_double), ("err_p", C.c_double), ("maxbo", C.c_double),
                ("err_t", C.c_double), ("residuo", C.c_double), ("passes_mom", C.c_int32),
                ("passes_p", C.c_int32), ("passes_t", C.c_int32), ("simple_iter", C.c_int32),
                ("logged", C.c_int32), ("pad_", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.ox_create.restype = C.c_void_p
        L.ox_create.argtypes = [C.c_int32, C.c_int32, dp, dp, dp, dp]
        L.ox_destroy.argtypes = [C.c_void_p]
        L.ox_set_params.argtypes = [C.c_void_p, dp, C.c_int32, C.c_int32]
        L.ox_set_bc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, ip, dp]
        L.ox_set_sources.argtypes = [C.c_void_p, C.c_int32, dp, dp]
        L.ox_set_field.argtypes = [C.c_void_p, C.c_int32, dp]
        L.ox_get_field.argtypes = [C.c_void_p, C.c_int32, dp]
        L.ox_field_size.restype = C.c_int64
        L.ox_field_size.argtypes = [C.c_void_p, C.c_int32]
        for name in ["ox_momentum_u_y", "ox_momentum_u_x", "ox_momentum_v_y", "ox_momentum_v_x",
                     "ox_pcorr_zero", "ox_pcorr_y", "ox_pcorr_x", "ox_correct", "ox_energy_y",
                     "ox_energy_x", "ox_lag"]:
            getattr(L, name).argtypes = [C.c_void_p]
        L.ox_residual_u.restype = C.c_double
        L.ox_residual_u.argtypes = [C.c_void_p]
        L.ox_run.restype = C.c_int32
        L.ox_run.argtypes = [C.c_void_p, C.c_int32, C.POINTER(Hist), C.c_int32]
        L.ox_tridiagonal.argtypes = [dp, dp, dp, dp, C.c_int32]
        L.ox_nusselt_y.argtypes = [C.c_void_p, C.c_int32, C.c_int32, dp, dp]
        L.ox_nusselt_x.argtypes = [C.c_void_p, dp, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a) -> np.ndarray:
    return np.require(np.asarray(a, dtype=np.float64), requirements=["F", "A", "W"])


def tridiagonal(a, b, c, r):
    """Thomas solve; returns (solution, eliminated diagonal)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64).copy()
    c = np.ascontiguousarray(c, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64).copy()
    lib().ox_tridiagonal(_dp(a), _dp(b), _dp(c), _dp(r), len(r))
    return r, b


class Oracle:
    """One cavity on the CPU, in upstream's Fortran layout."""

    def __init__(self, case, params_override: Optional[dict] = None):
        L = lib()
        m = case.mesh
        self.case = case
        self.mi, self.nj = m.mi, m.nj
        self._geo = [np.ascontiguousarray(x, dtype=np.float64) for x in (m.xu, m.yv, m.xp, m.yp)]
        self.h = L.ox_create(m.mi, m.nj, *[_dp(x) for x in self._geo])
        p = case.params
        d = p.derived()
        sc = dict(dt=p.dt, rel_pres=p.rel_pres, rel_vel=p.rel_vel, rel_ener=p.rel_ener,
                  gamma_momen=d["gamma_momen"], gamma_energ=d["gamma_energ"], Ri=d["Ri"],
                  Riy=d["Riy"], conv_u=p.conv_u, conv_t=p.conv_t, conv_p=p.conv_p,
                  conv_resi=p.conv_resi, conv_paso=p.conv_paso)
        iem, ism = p.iter_ecuaci_max, p.iter_simple_max
        if params_override:
            iem = params_override.pop("iter_ecuaci_max", iem)
            ism = params_override.pop("iter_simple_max", ism)
            sc.update(params_override)
        arr = np.array([sc[k] for k in ("dt", "rel_pres", "rel_vel", "rel_ener", "gamma_momen",
                                         "gamma_energ", "Ri", "Riy", "conv_u", "conv_t", "conv_p",
                                         "conv_resi", "conv_paso")], dtype=np.float64)
        L.ox_set_params(self.h, _dp(arr), iem, ism)
        for var, bset in case.bcs.items():
            for side, (typ, val) in bset.expand(m).items():
                typ = np.ascontiguousarray(typ, dtype=np.int32)
                val = np.ascontiguousarray(val, dtype=np.float64)
                L.ox_set_bc(self.h, VAR_ID[var], SIDE_ID[side],
                            typ.ctypes.data_as(C.POINTER(C.c_int32)), _dp(val))
        for var, (sc_arr, sp_arr) in case.sources.items():
            a, b = _f64(sc_arr), _f64(sp_arr)
            L.ox_set_sources(self.h, VAR_ID[var], _dp(a), _dp(b))
        u, v, t, pp = case.initial_fields()
        for name, arr in (("u", u), ("v", v), ("t", t), ("p", pp), ("u_ant", u), ("v_ant", v),
                          ("t_ant", t)):
            self.set(name, arr)

    def set_bc(self, var, side, types, values):
        """Per-node boundary types and values of one side."""
        L = lib()
        typ = np.ascontiguousarray(types, dtype=np.int32)
        val = np.ascontiguousarray(values, dtype=np.float64)
        L.ox_set_bc(self.h, VAR_ID[var], SIDE_ID[side], typ.ctypes.data_as(C.POINTER(C.c_int32)), _dp(val))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().ox_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def shape(self, name):
        mi, nj = self.mi, self.nj
        if name in ("u", "u_ant", "au", "resu"):
            return (mi, nj + 1)
        if name in ("v", "v_ant", "av"):
            return (mi + 1, nj)
        return (mi + 1, nj + 1)

    def set(self, name, arr):
        a = _f64(arr)
        assert a.shape == self.shape(name), (name, a.shape, self.shape(name))
        lib().ox_set_field(self.h, FIELD_ID[name], _dp(a))

    def get(self, name) -> np.ndarray:
        out = np.zeros(self.shape(name), dtype=np.float64, order="F")
        lib().ox_get_field(self.h, FIELD_ID[name], _dp(out))
        return out

    def fields(self) -> Dict[str, np.ndarray]:
        return {n: self.get(n) for n in FIELDS}

    def phase(self, name: str):
        return getattr(lib(), "ox_" + name)(self.h)

    def run(self, nsteps: int, cap: Optional[int] = None):
        cap = cap or max(16, nsteps * 4)
        hist = (Hist * cap)()
        n = lib().ox_run(self.h, nsteps, hist, cap)
        n = min(n, cap)
        keys = ["err_v", "err_p", "maxbo", "err_t", "residuo"]
        rec = {k: np.array([getattr(hist[i], k) for i in range(n)]) for k in keys}
        for k in ["passes_mom", "passes_p", "passes_t", "simple_iter", "logged"]:
            rec[k] = np.array([getattr(hist[i], k) for i in range(n)], dtype=np.int32)
        return rec

    def nusselt_y(self, i0: Optional[int] = None, i1: Optional[int] = None):
        m = self.case.mesh
        i0 = m.placa_min if i0 is None else i0
        i1 = m.placa_max if i1 is None else i1
        if not (1 <= i0 and i1 + 2 <= self.mi + 1):
            raise ValueError("Nusselt panel range outside the mesh")
        a, b = C.c_double(), C.c_double()
        lib().ox_nusselt_y(self.h, i0, i1, C.byref(a), C.byref(b))
        return a.value, b.value

    def nusselt_x(self):
        a, b = C.c_double(), C.c_double()
        lib().ox_nusselt_x(self.h, C.byref(a), C.byref(b))
        return a.value, b.value
