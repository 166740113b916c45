"""Case description for the SIMPLE iteration: mesh, parameters, boundary
conditions, initial fields and source terms, plus the Ixchel2D case-file
formats (so the same case can be fed to the upstream program and to this
library bit-for-bit).

Index conventions follow Ixchel2D: ``mi`` nodes of the u-mesh along x
(``xu(1..mi)``, the cell faces), ``nj`` v-nodes along y (``yv(1..nj)``),
pressure/temperature nodes ``xp(1..mi+1)``, ``yp(1..nj+1)`` including one
ghost node half a cell outside each wall.  Arrays are kept in Fortran order
(first index fastest), extents ``u(mi, nj+1)``, ``v(mi+1, nj)``,
``T, p (mi+1, nj+1)`` (source/mod_ec_momento.f90:27-41,
source/mod_ec_energia.f90:23-24).
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

SIDES = ("a", "b", "c", "d")  # west, south, east, north
BC_NONE, BC_DIRI, BC_NEUM = 0, 1, 2
_BC_CODE = {"diri": BC_DIRI, "neum": BC_NEUM}


# --------------------------------------------------------------------------
# mesh
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Mesh:
    mi: int
    nj: int
    xu: np.ndarray  # (mi,)
    yv: np.ndarray  # (nj,)
    xp: np.ndarray  # (mi+1,)
    yp: np.ndarray  # (nj+1,)
    placa_min: int = 2
    placa_max: int = 2
    ao: float = 1.0

    @property
    def width(self) -> float:
        return float(self.xu[-1] - self.xu[0])

    @property
    def height(self) -> float:
        return float(self.yv[-1] - self.yv[0])


def _faces(n: int, length: float, kind: str, beta: float) -> np.ndarray:
    xi = np.arange(n, dtype=np.float64) / float(n - 1)
    if kind == "unif":
        return length * xi
    if kind == "tanh":
        # symmetric tanh clustering toward both walls
        return 0.5 * length * (1.0 + np.tanh(beta * (2.0 * xi - 1.0)) / math.tanh(beta))
    raise ValueError(f"unknown mesh kind {kind!r}")


def _nodes_from_faces(f: np.ndarray) -> np.ndarray:
    """Pressure nodes at face midpoints plus one ghost node extrapolated half a
    cell outside each end (malla/malla_irr.f90:66,77,90,95,101,106)."""
    n = f.shape[0]
    p = np.empty(n + 1)
    p[0] = (3.0 * f[0] - f[1]) / 2.0
    p[1:n] = (f[1:] + f[:-1]) / 2.0
    p[n] = (3.0 * f[n - 1] - f[n - 2]) / 2.0
    return p


def make_mesh(mi: int, nj: int, width: float = 1.0, height: float = 1.0,
              kind: str = "unif", beta: float = 1.2,
              placa: Optional[Tuple[int, int]] = None) -> Mesh:
    """Staggered mesh in the malla_irr layout.  ``kind`` is ``"unif"`` or
    ``"tanh"`` (x = W/2 (1 + tanh(beta (2 xi - 1)) / tanh beta), same in y).
    ``placa`` are the Nusselt integration limits written into the mesh-file
    header (default: the whole bottom wall, 2 .. mi-1 rounded to an even
    panel count)."""
    if mi < 4 or nj < 4:
        raise ValueError("mesh needs mi, nj >= 4")
    xu = _faces(mi, width, kind, beta)
    yv = _faces(nj, height, kind, beta)
    xp = _nodes_from_faces(xu)
    yp = _nodes_from_faces(yv)
    if placa is None:
        lo = 2
        hi = mi - 1
        if (hi - lo) % 2:
            hi -= 1
        placa = (lo, hi)
    return Mesh(mi, nj, xu, yv, xp, yp, placa[0], placa[1], width)


# --------------------------------------------------------------------------
# boundary conditions
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Segment:
    x0: float
    x1: float
    kind: str  # 'diri' | 'neum'
    value: float


def expand_side(segments: Sequence[Segment], coords: np.ndarray, nn: int) -> Tuple[np.ndarray, np.ndarray]:
    """Per-node (type, value) along one side, length ``nn`` (1-based node k at
    index k-1).  Follows the index rule of lectura_cond_frontera
    (source/mod_cond_frontera.f90:125-149): the first segment starts at node 2,
    each later one at the first node whose coordinate exceeds the previous
    segment's x1, the last ends at node nn-1; segment ranges are inclusive and
    a later segment overwrites the node it shares with the previous one
    (impone_cond_frontera, :254-420)."""
    nd = len(segments)
    if nd < 1:
        raise ValueError("each side needs at least one segment")
    ind = [2] + [0] * nd
    for j, seg in enumerate(segments):
        kk = nn + 1
        for k in range(1, nn + 1):
            if coords[k - 1] > seg.x1:
                kk = k
                break
        ind[j + 1] = kk
    ind[nd] = nn - 1
    typ = np.zeros(nn, dtype=np.int32)
    val = np.zeros(nn, dtype=np.float64)
    for j, seg in enumerate(segments):
        code = _BC_CODE.get(seg.kind)
        if code is None:
            raise ValueError(f"unsupported boundary type {seg.kind!r}")
        for k in range(ind[j], ind[j + 1] + 1):
            if not 1 <= k <= nn:
                raise ValueError("boundary segment index outside the side")
            typ[k - 1] = code
            val[k - 1] = seg.value
    if np.any(typ[1:nn - 1] == BC_NONE):
        raise ValueError("boundary segments leave interior nodes of a side unset")
    return typ, val


@dataclasses.dataclass
class BoundarySet:
    """Segments per side for one variable (cond_frontera{u,v,t}.dat)."""
    var: str
    sides: Dict[str, List[Segment]]

    def expand(self, mesh: Mesh) -> Dict[str, Tuple[np.ndarray, np.ndarray]]:
        if self.var == "u":
            xx, yy, mm, nn = mesh.xu, mesh.yp, mesh.mi, mesh.nj + 1
        elif self.var == "v":
            xx, yy, mm, nn = mesh.xp, mesh.yv, mesh.mi + 1, mesh.nj
        else:
            xx, yy, mm, nn = mesh.xp, mesh.yp, mesh.mi + 1, mesh.nj + 1
        out = {}
        for s in SIDES:
            coords, n = (yy, nn) if s in ("a", "c") else (xx, mm)
            out[s] = expand_side(self.sides[s], coords, n)
        return out


def uniform_bc(var: str, a: Tuple[str, float], b: Tuple[str, float],
               c: Tuple[str, float], d: Tuple[str, float]) -> BoundarySet:
    big = 1.0e30
    return BoundarySet(var, {s: [Segment(-big, big, k, v)] for s, (k, v) in zip(SIDES, (a, b, c, d))})


# --------------------------------------------------------------------------
# parameters and case
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Params:
    """The 22 scalars of parametros.dat (source/main.f90:164-189)."""
    adimen: str = "natural"
    Ra: float = 1.0e5           # Rayleigh ('natural') or Reynolds ('mixta')
    Pr: float = 0.71
    Ri_1: float = 0.0
    dt: float = 1.0
    paq_itera: int = 10
    itermax: int = 10
    rel_pres: float = 0.3
    rel_vel: float = 0.7
    rel_ener: float = 0.7
    conv_u: float = 1.0e30
    conv_t: float = 1.0e30
    conv_p: float = 1.0e30
    conv_resi: float = 1.0e30
    conv_paso: float = 1.0e30
    iter_ecuaci_max: int = -1
    iter_simple_max: int = -1
    flujo_ini: str = "archivo"
    tempe_ini: str = "archivo"

    def derived(self) -> Dict[str, float]:
        """Non-dimensional coefficients (source/main.f90:221-244)."""
        if self.adimen == "natural":
            return dict(gamma_momen=math.sqrt(self.Pr / self.Ra),
                        gamma_energ=math.sqrt(1.0 / (self.Pr * self.Ra)),
                        Ri=0.0, Riy=-1.0)
        if self.adimen == "mixta":
            return dict(gamma_momen=1.0 / self.Ra, gamma_energ=1.0 / (self.Ra * self.Pr),
                        Ri=0.0, Riy=self.Ri_1)
        raise ValueError("adimen must be 'natural' or 'mixta'")

    def rec(self) -> str:
        """The 5-digit Re/Ra tag used in upstream file names
        (source/main.f90:230,241, mod_postproceso.f90:375-399)."""
        n = math.ceil(math.sqrt(self.Ra)) if self.adimen == "natural" else math.ceil(self.Ra)
        return f"{n % 100000:05d}"


@dataclasses.dataclass
class Case:
    mesh: Mesh
    params: Params
    bcs: Dict[str, BoundarySet]
    u0: np.ndarray  # (mi, nj+1)  Fortran order
    v0: np.ndarray  # (mi+1, nj)
    t0: np.ndarray  # (mi+1, nj+1)
    p0: np.ndarray  # (mi+1, nj+1)
    # optional per-node source terms (immersed-boundary forcing); None = 0
    sources: Dict[str, Tuple[np.ndarray, np.ndarray]] = dataclasses.field(default_factory=dict)
    name: str = "case"

    def initial_fields(self):
        """Fields after condicion_inicial_flujo/tempe
        (source/mod_ec_momento.f90:59-89, mod_ec_energia.f90:46-80)."""
        u, v, t = self.u0.copy(order="F"), self.v0.copy(order="F"), self.t0.copy(order="F")
        if self.params.flujo_ini == "nulo":
            u[:] = 0.0
            v[:] = 0.0
        if self.params.tempe_ini == "baja":
            t[:] = 0.0
        elif self.params.tempe_ini == "alta":
            t[:] = 1.0
        return u, v, t, self.p0.copy(order="F")


def _zeros(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


def lid_driven_case(n: int = 66, Re: float = 100.0, iters: int = 2000, dt: float = 1.0,
                    rel_vel: float = 0.7, rel_pres: float = 0.3) -> Case:
    """Lid-driven square cavity: top lid u=1, no-slip elsewhere, energy
    equation decoupled (Ri=0) with T=0 everywhere."""
    mesh = make_mesh(n, n, 1.0, 1.0, "unif")
    prm = Params(adimen="mixta", Ra=Re, Pr=0.71, Ri_1=0.0, dt=dt,
                 paq_itera=max(10, min(iters, 1000)), itermax=iters,
                 rel_pres=rel_pres, rel_vel=rel_vel, rel_ener=0.7,
                 flujo_ini="nulo", tempe_ini="baja")
    bcs = {
        "u": uniform_bc("u", ("diri", 0.0), ("diri", 0.0), ("diri", 0.0), ("diri", 1.0)),
        "v": uniform_bc("v", ("diri", 0.0), ("diri", 0.0), ("diri", 0.0), ("diri", 0.0)),
        "t": uniform_bc("t", ("diri", 0.0), ("diri", 0.0), ("diri", 0.0), ("diri", 0.0)),
    }
    mi, nj = mesh.mi, mesh.nj
    return Case(mesh, prm, bcs, _zeros((mi, nj + 1)), _zeros((mi + 1, nj)),
                _zeros((mi + 1, nj + 1)), _zeros((mi + 1, nj + 1)), name=f"lid{n}")


def natural_convection_case(mi: int = 301, nj: int = 701, height: float = 2.33,
                            Ra: float = 1.0e5, Pr: float = 0.71, beta: float = 1.2,
                            iters: int = 5000, dt: float = 1.0, seed: int = 0,
                            noise: float = 1.0e-3, mesh_kind: str = "tanh",
                            rel_vel: float = 0.7, rel_pres: float = 0.3,
                            rel_ener: float = 0.7) -> Case:
    """Differentially heated cavity: hot wall x=0 (T=1), cold wall x=W (T=0),
    adiabatic top and bottom, no-slip everywhere; T0 = 0.5 + U(-noise, noise)."""
    mesh = make_mesh(mi, nj, 1.0, height, mesh_kind, beta)
    prm = Params(adimen="natural", Ra=Ra, Pr=Pr, Ri_1=0.0, dt=dt,
                 paq_itera=max(10, min(iters, 1000)), itermax=iters,
                 rel_pres=rel_pres, rel_vel=rel_vel, rel_ener=rel_ener,
                 flujo_ini="nulo", tempe_ini="archivo")
    bcs = {
        "u": uniform_bc("u", ("diri", 0.0), ("diri", 0.0), ("diri", 0.0), ("diri", 0.0)),
        "v": uniform_bc("v", ("diri", 0.0), ("diri", 0.0), ("diri", 0.0), ("diri", 0.0)),
        "t": uniform_bc("t", ("diri", 1.0), ("neum", 0.0), ("diri", 0.0), ("neum", 0.0)),
    }
    rng = np.random.default_rng(seed)
    t0 = np.asfortranarray(0.5 + rng.uniform(-noise, noise, size=(mi + 1, nj + 1)))
    return Case(mesh, prm, bcs, _zeros((mi, nj + 1)), _zeros((mi + 1, nj)),
                t0, _zeros((mi + 1, nj + 1)), name=f"nc{mi}x{nj}")


def heated_lid_case(n: int = 34, Re: float = 100.0, iters: int = 100, dt: float = 0.05,
                    seed: int = 0) -> Case:
    """Lid-driven square cavity carrying a passive temperature (no
    buoyancy): hot wall x=0 (T=1), cold wall x=W (T=0), adiabatic lid and
    bottom, T0 = 0.5 + U(-1e-3, 1e-3).  The flow does not depend on T."""
    case = lid_driven_case(n, Re=Re, iters=iters, dt=dt)
    case.bcs["t"] = uniform_bc("t", ("diri", 1.0), ("neum", 0.0), ("diri", 0.0), ("neum", 0.0))
    case.params.tempe_ini = "archivo"
    rng = np.random.default_rng(seed)
    case.t0 = np.asfortranarray(0.5 + rng.uniform(-1e-3, 1e-3, size=(n + 1, n + 1)))
    case.name = f"hlid{n}"
    return case


# --------------------------------------------------------------------------
# immersed bodies (large-source penalty, source/mod_front_inmersa.f90:54-179)
# --------------------------------------------------------------------------
PENALTY = 10.0e50


def cylinder_sources(case: Case, xc: float, yc: float, radius: float,
                     t_body: Optional[float] = 1.0) -> None:
    """Mark a circular solid.  As in definir_cuerpo, the mask is tested at the
    pressure-node coordinates (xp(ii), yp(jj)) for the u, v and T arrays alike,
    inside nodes get Sp = -1e51 and Sc = 1e51*1e-11 for u, v (so u, v ~ 1e-11)
    and Sc = 1e51*T_body for T."""
    m = case.mesh
    mi, nj = m.mi, m.nj
    X, Y = np.meshgrid(m.xp, m.yp, indexing="ij")
    inside = (X - xc) ** 2 + (Y - yc) ** 2 <= radius * radius  # (mi+1, nj+1)
    def put(var, shape, sc_val):
        sc = np.zeros(shape, order="F")
        sp = np.zeros(shape, order="F")
        msk = inside[: shape[0], : shape[1]]
        sp[msk] = -PENALTY
        sc[msk] = PENALTY * sc_val
        old = case.sources.get(var)
        if old is not None:
            sc = np.where(msk, sc, old[0])
            sp = np.where(msk, sp, old[1])
        case.sources[var] = (np.asfortranarray(sc), np.asfortranarray(sp))
    put("u", (mi, nj + 1), 10.0e-12)
    put("v", (mi + 1, nj), 10.0e-12)
    if t_body is not None:
        put("t", (mi + 1, nj + 1), t_body)


def upstream_body_mask(mesh: Mesh, opcion: str) -> np.ndarray:
    """The (mi+1, nj+1) nodes inside one of upstream's immersed bodies
    (source/mod_front_inmersa.f90:54-179), tested at the pressure-node
    coordinates (xp(ii), yp(jj)) with the same floating-point expressions as
    upstream's tests:
      'cuadr': the square of side 0.1 centred at (6.0, 0.05) (:99-128);
      'trian': the triangle 0 <= y <= 0.1 between the lines x = 2y + 5.8
               and x = 6.2 - 2y (recta_mxb1 / recta_mxb2, :74-97, 192-252);
      'chime': the rod of side 0.3 centred at (8.5, 0.5) (:130-175)."""
    X, Y = np.meshgrid(mesh.xp, mesh.yp, indexing="ij")
    if opcion in ("cuadr", "chime"):
        yv, xv, hh = (0.05, 6.0, 0.1) if opcion == "cuadr" else (0.5, 8.5, 0.3)
        return (yv - hh / 2.0 <= Y) & (Y <= yv + hh / 2.0) & (xv - hh / 2.0 <= X) & (X <= xv + hh / 2.0)
    if opcion == "trian":
        yv, hh = 0.1, 0.1
        m1, m2 = 1.0 / 2.0, -1.0 / 2.0
        b1 = -(0.1 - m1 * 6.0) / m1
        b2 = -(0.1 - m2 * 6.0) / m2
        return (yv - hh <= Y) & (Y <= yv) & (Y / m1 + b1 <= X) & (X <= Y / m2 + b2)
    raise ValueError(f"body {opcion!r}: 'cuadr', 'trian' or 'chime'")


def upstream_body_sources(case: Case, opcion: str) -> None:
    """The sources of upstream's immersed bodies (upstream_body_mask): u and
    v are forced to rest inside, Sp = -1e51 and Sc = 1e51 * 1e-11; 'chime'
    also holds T = 1 there, Sp = -1e51 and Sc = 1e51 * 1 (:164-165).  The
    'chime' branch also sets a_u, a_v = 1e41 on the rod (:158, 162); every
    x-assembly rewrites a_u, a_v on interior nodes before anything reads
    them, so those values never take effect and are not set here."""
    m = case.mesh
    mi, nj = m.mi, m.nj
    inside = upstream_body_mask(m, opcion)
    for var, shape in (("u", (mi, nj + 1)), ("v", (mi + 1, nj))):
        msk = inside[: shape[0], : shape[1]]
        sc = np.zeros(shape, order="F")
        sp = np.zeros(shape, order="F")
        sp[msk] = -PENALTY
        sc[msk] = PENALTY * 10.0e-12
        case.sources[var] = (np.asfortranarray(sc), np.asfortranarray(sp))
    if opcion == "chime":
        sc = np.zeros((mi + 1, nj + 1), order="F")
        sp = np.zeros((mi + 1, nj + 1), order="F")
        sp[inside] = -PENALTY
        sc[inside] = PENALTY * 1.0
        case.sources["t"] = (np.asfortranarray(sc), np.asfortranarray(sp))


# --------------------------------------------------------------------------
# batches of uneven cavities
# --------------------------------------------------------------------------
BATCH_SIZES = ((66, 66, 1.0), (130, 130, 1.0), (258, 258, 1.0), (130, 302, 2.33))


def batch_cases(n: int = 256, seed: int = 0, iters: int = 1000, lid_dt: float = 1.0,
                nc_dt: float = 0.01) -> List[Case]:
    """Mixed-size batch on uniform meshes: sizes 65^2, 129^2, 257^2 and
    129 x 301 pressure cells (mi, nj = cells + 1; the last is a tall cavity
    of height 2.33), drawn uniformly from `seed`.  Even indices are lid-driven
    cavities with Re log-uniform in [100, 3200]; odd ones natural convection
    with Ra log-uniform in [1e3, 1e6], Pr = 0.71.  Cavities 3 and 11 of every
    16 hold a circular cylinder (radius 0.05..0.15 of the width, random
    centre at least 0.05 from the walls) held at T = 1."""
    rng = np.random.default_rng(seed)
    out: List[Case] = []
    for c in range(n):
        mi, nj, h = BATCH_SIZES[int(rng.integers(len(BATCH_SIZES)))]
        if c % 2 == 0:
            re = float(10.0 ** rng.uniform(2.0, math.log10(3200.0)))
            case = lid_driven_case(mi, Re=re, iters=iters, dt=lid_dt)
            if nj != mi or h != 1.0:
                mesh = make_mesh(mi, nj, 1.0, h, "unif")
                case = Case(mesh, case.params, case.bcs, _zeros((mi, nj + 1)), _zeros((mi + 1, nj)),
                            _zeros((mi + 1, nj + 1)), _zeros((mi + 1, nj + 1)))
            case.name = f"b{c:03d}_lid{mi}x{nj}_Re{re:.0f}"
        else:
            ra = float(10.0 ** rng.uniform(3.0, 6.0))
            case = natural_convection_case(mi, nj, height=h, Ra=ra, iters=iters, dt=nc_dt,
                                           seed=int(rng.integers(1 << 30)), mesh_kind="unif")
            case.name = f"b{c:03d}_nc{mi}x{nj}_Ra{ra:.3g}"
        if c % 16 in (3, 11):
            r = float(rng.uniform(0.05, 0.15))
            xc = float(rng.uniform(r + 0.05, 1.0 - r - 0.05))
            yc = float(rng.uniform(r + 0.05, h - r - 0.05))
            cylinder_sources(case, xc, yc, r, t_body=1.0)
            case.name += f"_cyl{xc:.2f},{yc:.2f},{r:.2f}"
        out.append(case)
    return out


# --------------------------------------------------------------------------
# upstream file formats
# --------------------------------------------------------------------------
def _d23(x: float) -> str:
    # 17 significant digits so that every double round-trips exactly through
    # the (3D23.15)/(4D23.15) reads of mod_malla.f90:137-162
    s = f"{x:23.16E}"
    if len(s) != 23:
        s = f"{x:23.15E}"
        if len(s) != 23:
            raise ValueError(f"value {x!r} does not fit a D23 field")
    return s


def write_mesh_files(case: Case, directory: str, tag: str = "Rxxx") -> Tuple[str, str, str]:
    m = case.mesh
    mi, nj = m.mi, m.nj
    base = f"out_n{nj:05d}m{mi:05d}_{tag}"
    hdr = f" {m.placa_min} {m.placa_max} 0 {m.ao!r}\n"
    names = (base + "u.dat", base + "v.dat", base + "p.dat")
    with open(os.path.join(directory, names[0]), "w") as f:
        f.write(hdr)
        for j in range(nj + 1):
            for i in range(mi):
                f.write(_d23(m.xu[i]) + _d23(m.yp[j]) + _d23(case.u0[i, j]) + "\n")
    with open(os.path.join(directory, names[1]), "w") as f:
        f.write(hdr)
        for j in range(nj):
            for i in range(mi + 1):
                f.write(_d23(m.xp[i]) + _d23(m.yv[j]) + _d23(case.v0[i, j]) + "\n")
    with open(os.path.join(directory, names[2]), "w") as f:
        f.write(hdr)
        for j in range(nj + 1):
            for i in range(mi + 1):
                f.write(_d23(m.xp[i]) + _d23(m.yp[j]) + _d23(case.t0[i, j]) + _d23(case.p0[i, j]) + "\n")
    return names


def write_bc_file(bset: BoundarySet, path: str) -> None:
    with open(path, "w") as f:
        for s in SIDES:
            segs = bset.sides[s]
            f.write(f"{bset.var} {s} {len(segs)}\n")
            for seg in segs:
                f.write(f"{seg.x0!r} {seg.x1!r} {seg.kind} {seg.value!r}\n")


def write_reference_case(case: Case, directory: str, postprocesar: bool = True) -> None:
    """Write parametros.dat, cond_frontera{u,v,t}.dat, the three mesh/initial
    field files and the output directory upstream expects."""
    if case.sources:
        raise ValueError("upstream reads no source terms from files")
    os.makedirs(directory, exist_ok=True)
    names = write_mesh_files(case, directory)
    p = case.params
    m = case.mesh
    lines = [p.adimen, repr(p.Ra), repr(p.Pr), repr(p.Ri_1), repr(p.dt), str(p.paq_itera),
             str(p.itermax), repr(p.rel_pres), repr(p.rel_vel), repr(p.rel_ener), repr(p.conv_u),
             repr(p.conv_t), repr(p.conv_p), repr(p.conv_resi), repr(p.conv_paso),
             str(p.iter_ecuaci_max), str(p.iter_simple_max), names[0], names[1], names[2],
             p.flujo_ini, p.tempe_ini, ".true." if postprocesar else ".false."]
    with open(os.path.join(directory, "parametros.dat"), "w") as f:
        f.write("\n".join(lines) + "\n")
    for var, fname in (("u", "cond_fronterau.dat"), ("v", "cond_fronterav.dat"), ("t", "cond_fronterat.dat")):
        write_bc_file(case.bcs[var], os.path.join(directory, fname))
    os.makedirs(os.path.join(directory, f"n{m.nj:05d}m{m.mi:05d}R{p.rec()}"), exist_ok=True)


def read_reference_outputs(case: Case, directory: str) -> Dict[str, np.ndarray]:
    """Exact FP64 fields from upstream's stream-access .bin dumps
    (source/mod_postproceso.f90:310-365) plus the fort.102 history
    ('SIMPLE', iter, maxbo, residuo per line, main.f90:1204,1212) and the
    nuss_sim file (time, Nu0, -Nu1, T_mid, ..., main.f90:1242-1246)."""
    m = case.mesh
    mi, nj = m.mi, m.nj
    base = os.path.join(directory, f"out_n{nj:05d}m{mi:05d}_R{case.params.rec()}")
    ub = np.fromfile(base + "u.bin", dtype="<f8").reshape(nj + 1, mi, 3)
    vb = np.fromfile(base + "v.bin", dtype="<f8").reshape(nj, mi + 1, 3)
    pb = np.fromfile(base + "p.bin", dtype="<f8").reshape(nj + 1, mi + 1, 2)
    out = dict(u=np.asfortranarray(ub[:, :, 2].T), v=np.asfortranarray(vb[:, :, 2].T),
               t=np.asfortranarray(pb[:, :, 0].T), p=np.asfortranarray(pb[:, :, 1].T))
    hist = []
    f102 = os.path.join(directory, "fort.102")
    if os.path.exists(f102):
        with open(f102) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "SIMPLE":
                    hist.append((float(parts[2].replace("D", "E")), float(parts[3].replace("D", "E"))))
    out["history"] = np.array(hist, dtype=np.float64).reshape(-1, 2)
    nus = os.path.join(directory, f"nuss_sim_n{nj:05d}m{mi:05d}_R{case.params.rec()}.dat")
    rows = []
    if os.path.exists(nus):
        with open(nus) as f:
            for line in f:
                s = line.replace("D", "E")
                vals = [s[k:k + 23] for k in range(0, 6 * 23, 23)]
                try:
                    rows.append([float(x) for x in vals[:3]])
                except ValueError:
                    pass
    out["nusselt"] = np.array(rows, dtype=np.float64).reshape(-1, 3)
    return out
