"""Python access to the CPU FP64 oracle (libkm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
CPU-baseline leg of bench.py, as the checker. The product (kamayan_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

from kamayan_amd._abi import (KmAmrAdapt, KmAmrCriterion, KmBlock, KmCfg, KmMeshDesc, KmPack,
                              Layout, f64ptr, nvar)

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libkm_oracle.so"

_lib = None


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    vp, i32, dbl = C.c_void_p, C.c_int32, C.c_double
    P = C.POINTER
    sig = {
        "km_cpu_reconstruct": [i32, i32, P(dbl), P(dbl), P(dbl)],
        "km_cpu_riemann": [i32, i32, i32, P(dbl), P(dbl), P(dbl)],
        "km_cpu_eos": [i32, dbl, dbl, P(dbl)],
        "km_cpu_fast_speed": [i32, i32, P(dbl), P(dbl)],
        "km_cpu_prim2cons": [i32, P(dbl), P(dbl)],
        "km_cpu_prim2flux": [i32, i32, P(dbl), P(dbl)],
        "km_cpu_calculate_fluxes": [P(KmCfg), P(KmPack), vp],
        "km_cpu_calculate_emf": [P(KmCfg), P(KmPack), vp],
        "km_cpu_fluxes_to_dudt": [P(KmCfg), P(KmPack), P(KmPack), vp],
        "km_cpu_add_source_terms": [P(KmCfg), P(KmPack), P(KmPack), vp],
        "km_cpu_prepare_conserved_stage": [P(KmCfg), P(KmPack), vp],
        "km_cpu_apply_dudt": [P(KmCfg), P(KmPack), P(KmPack), P(KmPack), P(KmPack), dbl, dbl,
                              vp],
        "km_cpu_prepare_primitive": [P(KmCfg), P(KmPack), vp],
        "km_cpu_prepare_conserved": [P(KmCfg), P(KmPack), vp],
        "km_cpu_estimate_dt": [P(KmCfg), P(KmPack), P(dbl), vp],
        "km_cpu_fill_divb": [P(KmCfg), P(KmPack), vp],
        "km_cpu_history_vortex_errors": [P(KmCfg), P(KmPack), P(KmMeshDesc), P(dbl), C.c_int32,
                                         P(dbl)],
        "km_cpu_ghost_fill_uniform": [P(KmCfg), P(KmPack), P(i32), P(i32), i32, i32, i32],
        "km_cpu_mesh_blocks": [P(KmMeshDesc), vp],
        "km_cpu_sim_create": [P(KmCfg), P(KmMeshDesc), P(vp)],
        "km_cpu_sim_destroy": [vp],
        "km_cpu_sim_set_primitive": [vp, vp, P(vp)],
        "km_cpu_sim_get_state": [vp, vp, P(vp)],
        "km_cpu_sim_advance": [vp, i32, dbl],
        "km_cpu_sim_info": [vp, P(dbl), P(dbl), P(i32)],
        "km_cpu_sim_set_dt": [vp, dbl],
        "km_cpu_amr_blocks": [P(KmMeshDesc), P(i32), i32, vp],
        "km_cpu_amr_ghost_fill": [P(KmCfg), P(KmPack), P(KmMeshDesc), P(i32), i32, i32, i32],
        "km_cpu_flux_correct": [P(KmCfg), P(KmPack), P(KmMeshDesc), P(i32), i32],
        "km_cpu_amr_ghost_faces": [P(KmCfg), P(KmPack), P(KmMeshDesc), P(i32), i32],
        "km_cpu_sim_create_amr": [P(KmCfg), P(KmMeshDesc), P(i32), i32, P(vp)],
        "km_cpu_amr_tag_loehner": [P(KmPack), P(KmAmrCriterion), P(i32), i32],
        "km_cpu_amr_update_tree": [P(KmMeshDesc), i32, i32, P(i32), i32, P(i32), P(i32),
                                   P(i32), P(i32), P(i32), P(i32), P(i32)],
        "km_cpu_amr_remesh": [P(KmPack), P(KmPack), P(i32), i32, i32],
        "km_cpu_amr_remesh_geom": [P(KmPack), P(KmPack), P(i32), i32, i32, P(KmMeshDesc), i32],
        "km_cpu_sim_set_adaptive": [vp, P(KmAmrAdapt)],
        "km_cpu_sim_amr_leaves": [vp, P(i32), P(i32)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def reconstruct(recon: int, limiter: int, s5) -> tuple:
    s = np.ascontiguousarray(s5, dtype=np.float64)
    vm, vp = C.c_double(), C.c_double()
    lib().km_cpu_reconstruct(recon, limiter, _dp(s), C.byref(vm), C.byref(vp))
    return vm.value, vp.value


def riemann(solver: int, mhd: int, d: int, vl, vr) -> np.ndarray:
    a = np.zeros(10)
    b = np.zeros(10)
    a[: len(vl)] = vl
    b[: len(vr)] = vr
    f = np.zeros(8)
    lib().km_cpu_riemann(solver, mhd, d, _dp(a), _dp(b), _dp(f))
    return f[: (8 if mhd else 5)]


def eos(mode: int, gamma: float, abar: float, state) -> np.ndarray:
    s = np.ascontiguousarray(state, dtype=np.float64).copy()
    lib().km_cpu_eos(mode, gamma, abar, _dp(s))
    return s


def fast_speed(mhd: int, d: int, V) -> float:
    a = np.zeros(10)
    a[: len(V)] = V
    out = C.c_double()
    lib().km_cpu_fast_speed(mhd, d, _dp(a), C.byref(out))
    return out.value


def prim2cons(mhd: int, V) -> np.ndarray:
    a = np.zeros(10)
    a[: len(V)] = V
    u = np.zeros(8)
    lib().km_cpu_prim2cons(mhd, _dp(a), _dp(u))
    return u[: (8 if mhd else 5)]


def prim2flux(mhd: int, d: int, V) -> np.ndarray:
    a = np.zeros(10)
    a[: len(V)] = V
    f = np.zeros(8)
    lib().km_cpu_prim2flux(mhd, d, _dp(a), _dp(f))
    return f[: (8 if mhd else 5)]


def block_table(desc: KmMeshDesc, nblocks: int):
    arr = (KmBlock * nblocks)()
    lib().km_cpu_mesh_blocks(C.byref(desc), C.cast(arr, C.c_void_p))
    return arr


def _i32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def amr_blocks(desc: KmMeshDesc, leaves: np.ndarray):
    leaves = np.ascontiguousarray(leaves, dtype=np.int32)
    arr = (KmBlock * len(leaves))()
    call("km_cpu_amr_blocks", C.byref(desc), _i32(leaves), len(leaves), C.cast(arr, C.c_void_p))
    return arr


def amr_tag_loehner(pack: KmPack, crit: KmAmrCriterion, tags: np.ndarray, reset: bool):
    """AMRLoehner tags (in place) of a host pack."""
    assert tags.dtype == np.int32
    call("km_cpu_amr_tag_loehner", C.byref(pack), C.byref(crit), _i32(tags), 1 if reset else 0)


def amr_update_tree(desc: KmMeshDesc, numlevel: int, derefine_count: int, leaves, tags,
                    deref: np.ndarray):
    """SetRefinement + UpdateMeshBlockTree; deref (int32) is updated in place. Returns
    (new_leaves, new_deref, origin[n, 9], changed)."""
    leaves = np.ascontiguousarray(leaves, dtype=np.int32)
    tags = np.ascontiguousarray(tags, dtype=np.int32)
    cap = 8 * len(leaves) + 64
    nl = np.zeros((cap, 4), np.int32)
    nd = np.zeros(cap, np.int32)
    org = np.zeros((cap, 9), np.int32)
    n, ch = C.c_int32(cap), C.c_int32(0)
    call("km_cpu_amr_update_tree", C.byref(desc), numlevel, derefine_count, _i32(leaves),
         len(leaves), _i32(tags), _i32(deref), _i32(nl), _i32(nd), _i32(org), C.byref(n),
         C.byref(ch))
    k = n.value
    return nl[:k].copy(), nd[:k].copy(), org[:k].copy(), bool(ch.value)


def amr_remesh(old: KmPack, new: KmPack, origin: np.ndarray, v0: int, v1: int, desc=None,
               geometry: int = 0):
    """remesh of cell slots [v0, v1) (and faces when both packs have them); geometry 1
    (cylindrical) takes the r-weighted operators and needs the mesh description"""
    origin = np.ascontiguousarray(origin, dtype=np.int32)
    call("km_cpu_amr_remesh_geom", C.byref(old), C.byref(new), _i32(origin), v0, v1,
         C.byref(desc) if desc is not None else None, geometry)


def initial_refinement(cfg, desc: KmMeshDesc, layout: Layout, problem_fill, adapt: KmAmrAdapt,
                       leaves=None):
    """The oracle's Parthenon initial-refinement loop: problem_fill(leaves) -> cell (the
    library's generator, host code), ghost fill through the oracle's simulation, Loehner
    tags and tree update, until the tree stops changing. Returns (leaves, cell)."""
    nd = layout.ndim
    if leaves is None:
        leaves = np.array([(0, i, j, k) for k in range(layout.nb[2] if nd > 2 else 1)
                           for j in range(layout.nb[1] if nd > 1 else 1)
                           for i in range(layout.nb[0])], dtype=np.int32)
    for _ in range(max(adapt.numlevel, 1)):
        lay = Layout(layout.ndim, layout.ng, layout.nx, layout.nb, nleaves=len(leaves))
        cell = problem_fill(leaves, lay)
        sim = CpuSimulation(cfg, desc, lay, leaves=leaves)
        sim.set_primitive(cell)
        state, _ = sim.state()
        sim.close()
        hp = HostPack(lay, state.shape[1], blocks=amr_blocks(desc, leaves))
        hp.cell[:] = state
        tags = np.full(len(leaves), -1, np.int32)
        for c in range(adapt.ncrit):
            amr_tag_loehner(hp.pack, adapt.crit[c], tags, c == 0)
        nl, _, _, changed = amr_update_tree(desc, adapt.numlevel, 1 << 30, leaves, tags,
                                            np.zeros(len(leaves), np.int32))
        if not changed:
            return leaves, cell
        leaves = nl
    lay = Layout(layout.ndim, layout.ng, layout.nx, layout.nb, nleaves=len(leaves))
    return leaves, problem_fill(leaves, lay)


class HostPack:
    """numpy block-pack arrays and host pointer tables for a km_pack_t."""

    def __init__(self, layout: Layout, nv: int, nflux: int = 0, faces: bool = False,
                 edges: bool = False, blocks=None):
        self.layout = layout
        nb = layout.nblocks
        self.cell = np.zeros(layout.cell_shape(nv))
        self.flux = [np.zeros(layout.cell_shape(nflux)) if (nflux and d < layout.ndim) else None
                     for d in range(3)]
        self.face = [np.zeros(layout.stag_shape()) if faces else None for d in range(3)]
        self.edge = [np.zeros(layout.stag_shape()) if edges else None for d in range(3)]
        self._keep = []
        p = self.pack = KmPack()
        p.nblocks = nb
        p.ndim = layout.ndim
        p.ng = layout.ng
        for d in range(3):
            p.nx[d] = layout.nx[d] if d < layout.ndim else 1
        p.nvar = nv
        p.nflux = nflux
        p.cell = self._table(self.cell.reshape(nb * nv, -1))
        for d in range(3):
            if self.flux[d] is not None:
                p.flux[d] = self._table(self.flux[d].reshape(nb * nflux, -1))
            if self.face[d] is not None:
                p.face[d] = self._table(self.face[d].reshape(nb, -1))
            if self.edge[d] is not None:
                p.edge[d] = self._table(self.edge[d].reshape(nb, -1))
        self.blocks = blocks
        if blocks is not None:
            p.blocks = C.cast(blocks, C.c_void_p)

    def _table(self, rows: np.ndarray) -> int:
        t = np.array([rows[r].ctypes.data for r in range(rows.shape[0])], dtype=np.uint64)
        self._keep.append(t)
        return t.ctypes.data


def call(name: str, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise RuntimeError(f"{name} returned {rc}")


class CpuSimulation:
    """km_cpu_sim_*: the oracle's whole-mesh RK driver."""

    def __init__(self, cfg: KmCfg, desc: KmMeshDesc, layout: Layout, leaves=None):
        self.cfg, self.desc, self.layout = cfg, desc, layout
        self.h = C.c_void_p()
        if leaves is not None:
            self._leaves = np.ascontiguousarray(leaves, dtype=np.int32)
            call("km_cpu_sim_create_amr", C.byref(cfg), C.byref(desc), _i32(self._leaves),
                 len(self._leaves), C.byref(self.h))
        else:
            call("km_cpu_sim_create", C.byref(cfg), C.byref(desc), C.byref(self.h))

    def set_primitive(self, cell: np.ndarray, faces=None):
        fp = (C.c_void_p * 3)()
        if faces is not None:
            for d in range(3):
                fp[d] = faces[d].ctypes.data
        call("km_cpu_sim_set_primitive", self.h, f64ptr(cell), fp)

    def set_adaptive(self, adapt: KmAmrAdapt):
        call("km_cpu_sim_set_adaptive", self.h, C.byref(adapt))

    def leaves(self) -> np.ndarray:
        n = C.c_int32(0)
        lib().km_cpu_sim_amr_leaves(self.h, _i32(np.zeros(4, np.int32)), C.byref(n))
        out = np.zeros((max(n.value, 1), 4), np.int32)
        call("km_cpu_sim_amr_leaves", self.h, _i32(out), C.byref(n))
        self.layout = Layout(self.layout.ndim, self.layout.ng, self.layout.nx, self.layout.nb,
                             nleaves=n.value)
        return out[: n.value]

    def state(self):
        mhd = self.cfg.mhd
        nv = nvar(mhd, self.cfg.nspecies)
        if getattr(self, "_leaves", None) is not None:
            self.leaves()
        cell = np.zeros(self.layout.cell_shape(nv))
        fp = (C.c_void_p * 3)()
        faces = None
        if mhd == 1:
            faces = [np.zeros(self.layout.stag_shape()) for _ in range(3)]
            for d in range(3):
                fp[d] = faces[d].ctypes.data
        call("km_cpu_sim_get_state", self.h, f64ptr(cell), fp)
        return cell, faces

    def advance(self, ncycles: int, tlim: float = 0.0):
        call("km_cpu_sim_advance", self.h, ncycles, tlim)

    def set_dt(self, dt: float):
        call("km_cpu_sim_set_dt", self.h, dt)

    def info(self):
        t, dt, n = C.c_double(), C.c_double(), C.c_int32()
        call("km_cpu_sim_info", self.h, C.byref(t), C.byref(dt), C.byref(n))
        return t.value, dt.value, n.value

    def close(self):
        if self.h:
            lib().km_cpu_sim_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_threads(n: int) -> None:
    os.environ["OMP_NUM_THREADS"] = str(n)


def history_vortex_errors(cfg: KmCfg, pack: KmPack, desc: KmMeshDesc, params=None) -> np.ndarray:
    """ErrorHistory of the isentropic vortex's history variables for a host pack."""
    out = (C.c_double * 8)()
    arr = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    pp = arr.ctypes.data_as(C.POINTER(C.c_double)) if arr is not None else None
    call("km_cpu_history_vortex_errors", C.byref(cfg), C.byref(pack), C.byref(desc), pp,
         0 if arr is None else len(arr), out)
    return np.array(out[: 8 if cfg.mhd == 1 else 5])
