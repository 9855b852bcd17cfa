"""ctypes mirror of include/km_hydro.h (structs, constants, layout helpers)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

KM_OK, KM_EINVAL, KM_EHIP, KM_ENOMEM, KM_ENODEV, KM_ESTATE = 0, 1, 2, 3, 4, 5
EOS_INIT = {"dens_pres": 0, "dens_ener": 1, "dens_temp": 2}
KM_TAG_NOALLOC_MAX_BLOCKS = 65536  # include/km_hydro.h

RECON = {"fog": 0, "plm": 1, "ppm": 2, "wenoz": 3}
LIMITER = {"minmod": 0, "van_leer": 1, "mc": 2}
RIEMANN = {"hll": 0, "hllc": 1, "hlld": 2}
MHD = {"off": 0, "ct": 1}
GEOMETRY = {"cartesian": 0, "cylindrical": 1}
PRECISION = {"fp64": 0, "fp32_mixed": 1, "fp64_fma": 2}
BC = {"periodic": 0, "outflow": 1, "reflect": 2, "axisymmetric": 3}
PATH_TASKS, PATH_FUSED = 0, 1

KM_DENS, KM_MOM0, KM_ENER, KM_MAGC0 = 0, 1, 4, 5


def ncons(mhd: int) -> int:
    return 8 if mhd == 1 else 5


def nvar(mhd: int, nspecies: int = 1) -> int:
    """cell vars of a state pack: the hydro/MHD slots plus the mass fractions when there is
    more than one species"""
    return (16 if mhd == 1 else 12) + (nspecies if nspecies > 1 else 0)


def mfrac(mhd: int, s: int) -> int:
    return (16 if mhd == 1 else 12) + s


def slot(name: str, mhd: int = 0) -> int:
    """Cell-pack slot of a field name (DENS, MOM0..2, ENER, MAGC0..2, VEL0..2, PRES,
    BMOD, EINT, TEMP, DIVB)."""
    nc = ncons(mhd)
    fixed = {"DENS": 0, "MOM0": 1, "MOM1": 2, "MOM2": 3, "ENER": 4, "MAGC0": 5, "MAGC1": 6,
             "MAGC2": 7}
    if name in fixed:
        return fixed[name]
    rel = {"VEL0": 0, "VEL1": 1, "VEL2": 2, "PRES": 3, "BMOD": 4, "EINT": 5, "TEMP": 6,
           "DIVB": 7}
    return nc + rel[name]


class KmCfg(C.Structure):
    _fields_ = [
        ("recon", C.c_int32), ("limiter", C.c_int32), ("riemann", C.c_int32),
        ("recon_vars", C.c_int32), ("strategy", C.c_int32), ("emf_averaging", C.c_int32),
        ("fluid", C.c_int32), ("mhd", C.c_int32), ("geometry", C.c_int32),
        ("nstages", C.c_int32), ("precision", C.c_int32), ("ndim", C.c_int32),
        ("gamma", C.c_double), ("abar", C.c_double), ("cfl", C.c_double),
        ("nspecies", C.c_int32), ("mode_init", C.c_int32), ("species0", C.c_char * 32),
        ("dt_factor", C.c_double), ("dt_init", C.c_double), ("dt_ceil", C.c_double),
        ("dt_floor", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("dt_user", C.c_double), ("dt_force", C.c_double),
        ("dt_init_force", C.c_int32), ("dt_min_cycle_limit", C.c_int32),
        ("dt_max_cycle_limit", C.c_int32), ("pack_size", C.c_int32),
    ]


class KmBlock(C.Structure):
    _fields_ = [
        ("xmin", C.c_double * 3), ("dx", C.c_double * 3), ("area", C.c_double * 3),
        ("vol", C.c_double), ("level", C.c_int32), ("lx", C.c_int32 * 3),
    ]


PP = C.POINTER(C.c_void_p)


class KmPack(C.Structure):
    _fields_ = [
        ("nblocks", C.c_int32), ("ndim", C.c_int32), ("ng", C.c_int32),
        ("nx", C.c_int32 * 3), ("nvar", C.c_int32), ("nflux", C.c_int32),
        ("blocks", C.c_void_p), ("cell", C.c_void_p),
        ("flux", C.c_void_p * 3), ("face", C.c_void_p * 3), ("edge", C.c_void_p * 3),
    ]


class KmMeshDesc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32), ("ng", C.c_int32), ("nx", C.c_int32 * 3), ("nb", C.c_int32 * 3),
        ("xmin", C.c_double * 3), ("xmax", C.c_double * 3), ("bc", C.c_int32 * 6),
    ]


KM_AMR_DEREFINE, KM_AMR_SAME, KM_AMR_REFINE = -1, 0, 1
KM_AMR_MAX_CRIT = 4


class KmAmrCriterion(C.Structure):
    _fields_ = [
        ("var", C.c_int32), ("max_level", C.c_int32), ("refine_tol", C.c_double),
        ("derefine_tol", C.c_double), ("filter", C.c_double),
    ]


class KmAmrAdapt(C.Structure):
    _fields_ = [
        ("numlevel", C.c_int32), ("derefine_count", C.c_int32), ("ncrit", C.c_int32),
        ("crit", KmAmrCriterion * KM_AMR_MAX_CRIT),
    ]


def amr_criterion(var: int, max_level: int, refine_tol: float = 0.8,
                  derefine_tol: float = 0.2, filter: float = 0.01) -> KmAmrCriterion:
    """One <kamayan/refinementN> Loehner criterion (defaults of grid.cpp:107-113)."""
    c = KmAmrCriterion()
    c.var, c.max_level = var, max_level
    c.refine_tol, c.derefine_tol, c.filter = refine_tol, derefine_tol, filter
    return c


def amr_adapt(numlevel: int, criteria, derefine_count: int = 10) -> KmAmrAdapt:
    a = KmAmrAdapt()
    a.numlevel, a.derefine_count, a.ncrit = numlevel, derefine_count, len(criteria)
    for i, c in enumerate(criteria):
        a.crit[i] = c
    return a


@dataclass
class Layout:
    """Shapes of the block-pack arrays of a mesh: nb is the (root) block grid; a block-tree
    mesh sets nleaves to its leaf count."""

    ndim: int
    ng: int
    nx: tuple
    nb: tuple
    nleaves: int = 0

    @property
    def nblocks(self) -> int:
        return self.nleaves if self.nleaves else self.nb[0] * self.nb[1] * self.nb[2]

    @property
    def E(self) -> tuple:
        e1 = self.nx[0] + 2 * self.ng
        e2 = self.nx[1] + 2 * self.ng if self.ndim > 1 else 1
        e3 = self.nx[2] + 2 * self.ng if self.ndim > 2 else 1
        return (e1, e2, e3)

    @property
    def S(self) -> tuple:
        e1, e2, e3 = self.E
        return (e1 + 1, e2 + (self.ndim > 1), e3 + (self.ndim > 2))

    def cell_shape(self, nv: int) -> tuple:
        e1, e2, e3 = self.E
        return (self.nblocks, nv, e3, e2, e1)

    def stag_shape(self) -> tuple:
        s1, s2, s3 = self.S
        return (self.nblocks, s3, s2, s1)

    def interior(self) -> tuple:
        """slices (k, j, i) of the interior of a cell array."""
        g = self.ng
        sl = [slice(g, g + self.nx[0])]
        sl.append(slice(g, g + self.nx[1]) if self.ndim > 1 else slice(0, 1))
        sl.append(slice(g, g + self.nx[2]) if self.ndim > 2 else slice(0, 1))
        return (sl[2], sl[1], sl[0])

    def lo(self, d: int) -> int:
        return self.ng if d < self.ndim else 0

    def hi(self, d: int) -> int:
        return self.ng + self.nx[d] - 1 if d < self.ndim else 0


def mesh_desc(layout: Layout, xmin=(0.0, 0.0, 0.0), xmax=(1.0, 1.0, 1.0),
              bc=("periodic",) * 6) -> KmMeshDesc:
    m = KmMeshDesc()
    m.ndim = layout.ndim
    m.ng = layout.ng
    for d in range(3):
        m.nx[d] = layout.nx[d] if d < layout.ndim else 1
        m.nb[d] = layout.nb[d] if d < layout.ndim else 1
        m.xmin[d] = xmin[d]
        m.xmax[d] = xmax[d]
    for d in range(6):
        m.bc[d] = BC[bc[d]] if isinstance(bc[d], str) else int(bc[d])
    return m


def f64ptr(a: np.ndarray) -> C.c_void_p:
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)
