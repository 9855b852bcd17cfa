"""kamayan_amd -- MI355X (gfx950) hydro/MHD flux-and-update path of kamayan.

Python is only the test/benchmark harness here: every computation runs in
libkm_hydro.so (C ABI declared in include/km_hydro.h). Loading fails loudly when the
library has not been built; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

from ._abi import (BC, KM_OK, LIMITER, MHD, PATH_FUSED, PATH_TASKS, PRECISION, RECON,
                   RIEMANN, KmAmrAdapt, KmAmrCriterion, KmBlock, KmCfg, KmMeshDesc, KmPack,
                   Layout, amr_adapt, amr_criterion, f64ptr, mesh_desc, mfrac, ncons, nvar,
                   slot)

HERE = pathlib.Path(__file__).resolve().parent
# KAMAYAN_AMD_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = pathlib.Path(os.environ.get("KAMAYAN_AMD_LIB", HERE / "libkm_hydro.so"))

__all__ = [
    "lib", "Config", "Simulation", "DevicePack", "problem_fill", "Layout", "mesh_desc",
    "KmCfg", "KmPack", "KmBlock", "KmMeshDesc", "ncons", "nvar", "mfrac", "slot", "PATH_TASKS",
    "PATH_FUSED", "LIB_PATH", "device_count", "block_table", "AmrAdapt", "amr_adapt",
    "amr_criterion", "amr_update_tree", "amr_tag_loehner", "amr_remesh", "initial_refinement",
    "history_vortex_errors",
]

_lib = None


class KmError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """The loaded libkm_hydro.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() or "
                          "python kamayan_amd/build.py")
    L = C.CDLL(str(LIB_PATH))
    vp, i32, dbl, u64 = C.c_void_p, C.c_int32, C.c_double, C.c_uint64
    P = C.POINTER
    sig = {
        "km_cfg_default": [P(KmCfg)],
        "km_cfg_set": [P(KmCfg), C.c_char_p, C.c_char_p],
        "km_cfg_parse_deck": [P(KmCfg), C.c_char_p],
        "km_cfg_min_ghost": [P(KmCfg)],
        "km_calculate_fluxes": [P(KmCfg), P(KmPack), vp],
        "km_calculate_emf": [P(KmCfg), P(KmPack), vp],
        "km_fluxes_to_dudt": [P(KmCfg), P(KmPack), P(KmPack), vp],
        "km_add_source_terms": [P(KmCfg), P(KmPack), P(KmPack), vp],
        "km_prepare_conserved_stage": [P(KmCfg), P(KmPack), vp],
        "km_apply_dudt": [P(KmCfg), P(KmPack), P(KmPack), P(KmPack), P(KmPack), dbl, dbl, vp],
        "km_prepare_primitive": [P(KmCfg), P(KmPack), vp],
        "km_prepare_conserved": [P(KmCfg), P(KmPack), vp],
        "km_estimate_dt": [P(KmCfg), P(KmPack), vp, vp],
        "km_fill_divb": [P(KmCfg), P(KmPack), vp],
        "km_ghost_plan_uniform": [P(KmPack), P(i32), P(i32), P(vp)],
        "km_ghost_plan_destroy": [vp],
        "km_ghost_fill": [vp, P(KmPack), i32, i32, i32, vp],
        "km_sim_create": [P(KmCfg), P(KmMeshDesc), i32, P(vp)],
        "km_sim_create_ensemble": [P(KmCfg), P(KmMeshDesc), i32, i32, P(vp)],
        "km_sim_ensemble_info": [vp, P(dbl), P(dbl), i32],
        "km_sim_destroy": [vp],
        "km_sim_set_primitive": [vp, vp, P(vp)],
        "km_sim_get_state": [vp, vp, P(vp)],
        "km_sim_advance": [vp, i32, dbl],
        "km_sim_info": [vp, P(dbl), P(dbl), P(i32)],
        "km_sim_set_dt": [vp, dbl],
        "km_sim_profile": [vp, i32],
        "km_sim_kernel_times": [vp, P(dbl), P(i32), i32],
        "km_sim_set_fused_kernel": [vp, i32, i32],
        "km_sim_fused_kernel": [vp, i32, P(i32)],
        "km_problem_fill": [C.c_char_p, P(KmCfg), P(KmMeshDesc), u64, P(dbl), i32, vp, P(vp)],
        "km_problem_fill_blocks": [C.c_char_p, P(KmCfg), P(KmMeshDesc), vp, i32, u64, P(dbl),
                                   i32, vp, P(vp)],
        "km_mesh_blocks": [P(KmMeshDesc), vp],
        "km_amr_blocks": [P(KmMeshDesc), P(i32), i32, vp],
        "km_amr_refine_ball": [P(KmMeshDesc), i32, P(dbl), dbl, P(i32), P(i32)],
        "km_ghost_plan_amr": [P(KmPack), P(KmMeshDesc), P(i32), i32, P(vp)],
        "km_ghost_plan_amr_geom": [P(KmPack), P(KmMeshDesc), P(i32), i32, i32, P(vp)],
        "km_flux_correct": [vp, P(KmPack), vp],
        "km_sim_create_amr": [P(KmCfg), P(KmMeshDesc), P(i32), i32, i32, P(vp)],
        "km_device_count": [P(i32)],
        "km_malloc": [P(vp), C.c_size_t],
        "km_free": [vp],
        "km_memcpy_h2d": [vp, vp, C.c_size_t],
        "km_memcpy_d2h": [vp, vp, C.c_size_t],
        "km_synchronize": [],
        "km_copy_device": [vp, vp, C.c_size_t, vp],
        "km_amr_adapt_default": [P(KmAmrAdapt)],
        "km_amr_adapt_set": [P(KmAmrAdapt), i32, C.c_char_p, C.c_char_p],
        "km_amr_tag_loehner": [P(KmPack), P(KmAmrCriterion), vp, i32, vp],
        "km_amr_tag_loehner_noalloc": [P(KmPack), P(KmAmrCriterion), vp, i32, vp],
        "km_amr_update_tree": [P(KmMeshDesc), i32, i32, P(i32), i32, P(i32), P(i32), P(i32),
                               P(i32), P(i32), P(i32), P(i32)],
        "km_amr_remesh": [P(KmPack), P(KmPack), vp, i32, i32, vp],
        "km_amr_remesh_geom": [P(KmPack), P(KmPack), vp, i32, i32, P(KmMeshDesc), i32, vp],
        "km_sim_set_adaptive": [vp, P(KmAmrAdapt)],
        "km_sim_amr_leaves": [vp, P(i32), P(i32)],
        "km_history_vortex_errors": [P(KmCfg), P(KmPack), P(KmMeshDesc), P(dbl), i32, P(dbl),
                                     vp],
        "km_sim_set_history": [vp, C.c_char_p, dbl, C.c_char_p, P(dbl), i32],
        "km_stage_plan_uniform": [P(KmCfg), P(KmMeshDesc), i32, P(vp)],
        "km_stage_plan_amr": [P(KmCfg), P(KmMeshDesc), P(i32), i32, P(vp)],
        "km_stage_plan_destroy": [vp],
        "km_stage_plan_set_kernel": [vp, i32, i32],
        "km_stage_plan_kernel": [vp, i32, P(i32)],
        "km_stage_fused": [P(KmCfg), vp, P(KmPack), P(KmPack), P(KmPack), dbl, vp, vp, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.km_last_error.restype = C.c_char_p
    L.km_last_error.argtypes = []
    L.km_version.restype = C.c_char_p
    L.km_version.argtypes = []
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != KM_OK:
        msg = lib().km_last_error().decode()
        raise KmError(f"{what} failed (code {rc}): {msg}")


def device_count() -> int:
    n = C.c_int32(0)
    lib().km_device_count(C.byref(n))
    return n.value


class Config:
    """km_cfg_t built from the reference's input-deck keys."""

    def __init__(self, deck: str | None = None, **keys):
        self.c = KmCfg()
        check(lib().km_cfg_default(C.byref(self.c)), "km_cfg_default")
        if deck:
            check(lib().km_cfg_parse_deck(C.byref(self.c), deck.encode()), "parse deck")
        for k, v in keys.items():
            self.set(k.replace("__", "/"), v)

    def set(self, key: str, value) -> "Config":
        check(lib().km_cfg_set(C.byref(self.c), key.encode(), str(value).encode()), key)
        return self

    def __getattr__(self, name):
        return getattr(self.__dict__["c"], name)


def block_table(layout: Layout, xmin, xmax) -> np.ndarray:
    m = mesh_desc(layout, xmin, xmax)
    arr = (KmBlock * layout.nblocks)()
    check(lib().km_mesh_blocks(C.byref(m), C.cast(arr, C.c_void_p)), "km_mesh_blocks")
    return arr


def _i32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def refine_ball(desc: KmMeshDesc, lmax: int, center, radius: float) -> np.ndarray:
    """Leaves ([n, 4] int32: level, lx1, lx2, lx3) of the root grid refined to lmax
    where blocks meet the ball, 2:1 balanced."""
    c = np.ascontiguousarray(center, dtype=np.float64)
    n = C.c_int32(0)
    lib().km_amr_refine_ball(C.byref(desc), lmax, c.ctypes.data_as(C.POINTER(C.c_double)),
                             radius, None, C.byref(n))
    out = np.zeros((n.value, 4), dtype=np.int32)
    n2 = C.c_int32(n.value)
    check(lib().km_amr_refine_ball(C.byref(desc), lmax, c.ctypes.data_as(C.POINTER(C.c_double)),
                                   radius, _i32(out), C.byref(n2)), "km_amr_refine_ball")
    return out


def amr_blocks(desc: KmMeshDesc, leaves: np.ndarray):
    leaves = np.ascontiguousarray(leaves, dtype=np.int32)
    arr = (KmBlock * len(leaves))()
    check(lib().km_amr_blocks(C.byref(desc), _i32(leaves), len(leaves), C.cast(arr, C.c_void_p)),
          "km_amr_blocks")
    return arr


class AmrAdapt:
    """km_amr_adapt_t built from the reference's deck keys (parthenon/mesh/numlevel,
    parthenon/mesh/derefine_count, kamayan/refinementN/...)."""

    def __init__(self, mhd: int = 0, deck_keys=None):
        self.mhd = mhd
        self.c = KmAmrAdapt()
        check(lib().km_amr_adapt_default(C.byref(self.c)), "km_amr_adapt_default")
        for k, v in (deck_keys or {}).items():
            self.set(k, v)

    def set(self, key: str, value) -> "AmrAdapt":
        check(lib().km_amr_adapt_set(C.byref(self.c), self.mhd, key.encode(),
                                     str(value).encode()), key)
        return self


def amr_update_tree(desc: KmMeshDesc, numlevel: int, derefine_count: int, leaves, tags,
                    deref: np.ndarray):
    """Refinement flags + tree update on the host (km_amr_update_tree); deref (int32) is
    updated in place. Returns (new_leaves, new_deref, origin[n, 9], changed)."""
    leaves = np.ascontiguousarray(leaves, dtype=np.int32)
    tags = np.ascontiguousarray(tags, dtype=np.int32)
    assert deref.dtype == np.int32 and deref.flags["C_CONTIGUOUS"]
    cap = 8 * len(leaves) + 64
    nl = np.zeros((cap, 4), np.int32)
    nd = np.zeros(cap, np.int32)
    org = np.zeros((cap, 9), np.int32)
    n, ch = C.c_int32(cap), C.c_int32(0)
    check(lib().km_amr_update_tree(C.byref(desc), numlevel, derefine_count, _i32(leaves),
                                   len(leaves), _i32(tags), _i32(deref), _i32(nl), _i32(nd),
                                   _i32(org), C.byref(n), C.byref(ch)), "km_amr_update_tree")
    k = n.value
    return nl[:k].copy(), nd[:k].copy(), org[:k].copy(), bool(ch.value)


def amr_tag_loehner(pack: "DevicePack", crit: KmAmrCriterion, tags_dev: int, reset: bool):
    check(lib().km_amr_tag_loehner(C.byref(pack.pack), C.byref(crit), C.c_void_p(tags_dev),
                                   1 if reset else 0, None), "km_amr_tag_loehner")


def amr_remesh(old: "DevicePack", new: "DevicePack", origin_dev: int, v0: int, v1: int,
               desc: KmMeshDesc = None, geometry: int = 0):
    """km_amr_remesh_geom: cell slots [v0, v1) and, when both packs carry them, the faces;
    geometry 1 (cylindrical) needs the mesh description"""
    check(lib().km_amr_remesh_geom(C.byref(old.pack), C.byref(new.pack), C.c_void_p(origin_dev),
                                   v0, v1, C.byref(desc) if desc is not None else None,
                                   geometry, None), "km_amr_remesh")


def _params(params):
    if params is None:
        return None, 0
    arr = np.ascontiguousarray(params, dtype=np.float64)
    return arr, len(arr)


def history_vortex_errors(cfg: Config, pack: "DevicePack", desc: KmMeshDesc, params=None):
    """The isentropic vortex's history variables (ErrorHistory of DENS, VEL0..2, PRES and,
    with MHD, MAGC0..2) of the primitives in a device pack."""
    arr, n = _params(params)
    out = (C.c_double * 8)()
    pp = arr.ctypes.data_as(C.POINTER(C.c_double)) if arr is not None else None
    check(lib().km_history_vortex_errors(C.byref(cfg.c), C.byref(pack.pack), C.byref(desc), pp,
                                         n, out, None), "km_history_vortex_errors")
    return np.array(out[: 8 if cfg.c.mhd == 1 else 5])


NO_MERGE = 1 << 30  # a derefine count no run reaches: the initial loop only refines


def initial_refinement(cfg: Config, desc: KmMeshDesc, layout: Layout, problem: str, adapt,
                       seed: int = 0, params=None, leaves=None):
    """Parthenon's initial refinement (Mesh::Initialize with an adaptive mesh): fill the
    problem on the leaves, fill ghost zones, tag with the criteria and refine, re-running
    the problem generator on the new leaves until the tree stops changing (at most
    numlevel - 1 rounds). Starts from the root grid unless `leaves` is given. Returns
    (leaves, cell, faces) of the final tree."""
    a = adapt.c if isinstance(adapt, AmrAdapt) else adapt
    nd = layout.ndim
    if leaves is None:
        leaves = np.array([(0, i, j, k) for k in range(layout.nb[2] if nd > 2 else 1)
                           for j in range(layout.nb[1] if nd > 1 else 1)
                           for i in range(layout.nb[0])], dtype=np.int32)
    for _ in range(max(a.numlevel, 1)):
        lay = Layout(layout.ndim, layout.ng, layout.nx, layout.nb, nleaves=len(leaves))
        cell, faces = problem_fill(problem, cfg, desc, lay, seed=seed, params=params,
                                   leaves=leaves)
        sim = Simulation(cfg, desc, lay, path=PATH_TASKS, leaves=leaves)
        sim.set_primitive(cell, faces)
        state, _ = sim.state()
        sim.close()
        dev = DevicePack(lay, nvar(cfg.c.mhd, cfg.c.nspecies), blocks=amr_blocks(desc, leaves))
        dev.upload(cell=state)
        tags_d = dev._alloc(4 * len(leaves))
        for c in range(a.ncrit):
            amr_tag_loehner(dev, a.crit[c], tags_d, c == 0)
        tags = np.full(len(leaves), -1, np.int32)
        if a.ncrit:
            check(lib().km_memcpy_d2h(_i32(tags), C.c_void_p(tags_d), 4 * len(leaves)), "d2h")
        dev.close()
        nl, _, _, changed = amr_update_tree(desc, a.numlevel, NO_MERGE, leaves, tags,
                                            np.zeros(len(leaves), np.int32))
        if not changed:
            return leaves, cell, faces
        leaves = nl
    lay = Layout(layout.ndim, layout.ng, layout.nx, layout.nb, nleaves=len(leaves))
    cell, faces = problem_fill(problem, cfg, desc, lay, seed=seed, params=params, leaves=leaves)
    return leaves, cell, faces


def problem_fill(name: str, cfg: Config, desc: KmMeshDesc, layout: Layout, seed: int = 0,
                 params=None, leaves=None):
    """Primitive initial state (cell array, face arrays or None) of a problem generator;
    with `leaves`, over the leaf blocks of a block tree."""
    mhd = cfg.c.mhd
    cell = np.zeros(layout.cell_shape(nvar(mhd, cfg.c.nspecies)))
    faces = None
    face_ptrs = (C.c_void_p * 3)()
    if mhd == 1:
        faces = [np.zeros(layout.stag_shape()) for _ in range(3)]
        for d in range(3):
            face_ptrs[d] = faces[d].ctypes.data
    par = None
    npar = 0
    if params is not None:
        params = np.ascontiguousarray(params, dtype=np.float64)
        par = params.ctypes.data_as(C.POINTER(C.c_double))
        npar = len(params)
    if leaves is not None:
        blocks = amr_blocks(desc, leaves)
        check(lib().km_problem_fill_blocks(name.encode(), C.byref(cfg.c), C.byref(desc),
                                           C.cast(blocks, C.c_void_p), len(leaves), seed, par,
                                           npar, f64ptr(cell), face_ptrs), f"problem {name}")
    else:
        check(lib().km_problem_fill(name.encode(), C.byref(cfg.c), C.byref(desc), seed, par,
                                    npar, f64ptr(cell), face_ptrs), f"problem {name}")
    return cell, faces


class Simulation:
    """Whole-mesh driver on the GPU (km_sim_*). With `descs` (a list of mesh descriptors
    sharing cells per block), an ensemble of independent runs advanced together; its
    buffers hold the instances' blocks one after the other and `layout` must describe
    the total block count."""

    def __init__(self, cfg: Config, desc: KmMeshDesc, layout: Layout, path: int = PATH_TASKS,
                 descs=None, leaves=None):
        self.cfg, self.desc, self.layout = cfg, desc, layout
        self.h = C.c_void_p()
        self.ninst = 1
        if leaves is not None:
            self._leaves = np.ascontiguousarray(leaves, dtype=np.int32)
            check(lib().km_sim_create_amr(C.byref(cfg.c), C.byref(desc), _i32(self._leaves),
                                          len(self._leaves), path, C.byref(self.h)),
                  "km_sim_create_amr")
        elif descs is not None:
            arr = (KmMeshDesc * len(descs))(*descs)
            self.ninst = len(descs)
            check(lib().km_sim_create_ensemble(C.byref(cfg.c), arr, len(descs), path,
                                               C.byref(self.h)), "km_sim_create_ensemble")
        else:
            check(lib().km_sim_create(C.byref(cfg.c), C.byref(desc), path, C.byref(self.h)),
                  "km_sim_create")

    def close(self):
        if self.h:
            lib().km_sim_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_primitive(self, cell: np.ndarray, faces=None):
        fp = (C.c_void_p * 3)()
        if faces is not None:
            for d in range(3):
                fp[d] = faces[d].ctypes.data
        check(lib().km_sim_set_primitive(self.h, f64ptr(cell), fp), "set_primitive")

    def set_adaptive(self, adapt):
        """Adaptive refinement (AmrAdapt or KmAmrAdapt) of a block-tree simulation."""
        c = adapt.c if isinstance(adapt, AmrAdapt) else adapt
        check(lib().km_sim_set_adaptive(self.h, C.byref(c)), "set_adaptive")

    def set_history(self, path, dt: float, problem: str = "isentropic_vortex", params=None):
        """Write a history file (Parthenon's hst output) with a row now and at every output
        time dt; `problem` names the history variables ("isentropic_vortex" or "")."""
        arr, n = _params(params)
        pp = arr.ctypes.data_as(C.POINTER(C.c_double)) if arr is not None else None
        check(lib().km_sim_set_history(self.h, str(path).encode(), dt,
                                       problem.encode() if problem else None, pp, n),
              "set_history")

    def leaves(self) -> np.ndarray:
        """Current leaves; also resizes `layout` to the current leaf count."""
        n = C.c_int32(0)
        lib().km_sim_amr_leaves(self.h, None, C.byref(n))
        out = np.zeros((max(n.value, 1), 4), np.int32)
        check(lib().km_sim_amr_leaves(self.h, _i32(out), C.byref(n)), "amr_leaves")
        L = self.layout
        self.layout = Layout(L.ndim, L.ng, L.nx, L.nb, nleaves=n.value)
        return out[: n.value]

    def state(self):
        mhd = self.cfg.c.mhd
        if getattr(self, "_leaves", None) is not None:
            self.leaves()
        cell = np.zeros(self.layout.cell_shape(nvar(mhd, self.cfg.c.nspecies)))
        fp = (C.c_void_p * 3)()
        faces = None
        if mhd == 1:
            faces = [np.zeros(self.layout.stag_shape()) for _ in range(3)]
            for d in range(3):
                fp[d] = faces[d].ctypes.data
        check(lib().km_sim_get_state(self.h, f64ptr(cell), fp), "get_state")
        return cell, faces

    def advance(self, ncycles: int, tlim: float = 0.0):
        check(lib().km_sim_advance(self.h, ncycles, tlim), "advance")

    def set_dt(self, dt: float):
        check(lib().km_sim_set_dt(self.h, dt), "set_dt")

    def info(self):
        t, dt, n = C.c_double(), C.c_double(), C.c_int32()
        check(lib().km_sim_info(self.h, C.byref(t), C.byref(dt), C.byref(n)), "info")
        return t.value, dt.value, n.value

    def ensemble_info(self):
        """(times, dts) arrays of all instances."""
        t = (C.c_double * self.ninst)()
        d = (C.c_double * self.ninst)()
        check(lib().km_sim_ensemble_info(self.h, t, d, self.ninst), "ensemble_info")
        return np.array(t[:]), np.array(d[:])

    def profile(self, on: bool = True):
        check(lib().km_sim_profile(self.h, 1 if on else 0), "profile")

    def set_fused_kernel(self, dir: int, kern: int):
        """Kernel of the fused sweep along x(dir+1) (dir -1: all): -1 autotune, 0 march,
        k pencil-tile shape k."""
        check(lib().km_sim_set_fused_kernel(self.h, dir, kern), "set_fused_kernel")

    def fused_kernel(self, dir: int) -> int:
        v = C.c_int32()
        check(lib().km_sim_fused_kernel(self.h, dir, C.byref(v)), "fused_kernel")
        return v.value

    KERNEL_KINDS = ("sweep_x1", "sweep_x2", "sweep_x3_update", "ghost_fill", "stage")

    def kernel_times(self):
        """{kind: (total ms, launches)} of the kernels recorded since profile(True)."""
        ms = (C.c_double * 5)()
        cnt = (C.c_int32 * 5)()
        check(lib().km_sim_kernel_times(self.h, ms, cnt, 5), "kernel_times")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.KERNEL_KINDS)}


class DevicePack:
    """Device copies of block-pack arrays plus the pointer tables a km_pack_t needs."""

    def __init__(self, layout: Layout, nv: int, nflux: int = 0, faces: bool = False,
                 edges: bool = False, blocks=None):
        self.layout, self.nv, self.nflux = layout, nv, nflux
        self._allocs = []
        self.pack = KmPack()
        p = self.pack
        p.nblocks = layout.nblocks
        p.ndim = layout.ndim
        p.ng = layout.ng
        for d in range(3):
            p.nx[d] = layout.nx[d] if d < layout.ndim else 1
        p.nvar = nv
        p.nflux = nflux
        e1, e2, e3 = layout.E
        self.ncell = e1 * e2 * e3
        s1, s2, s3 = layout.S
        self.nstag = s1 * s2 * s3
        nb = layout.nblocks
        self.cell = self._alloc(nb * nv * self.ncell * 8)
        p.cell = self._table([self.cell + 8 * self.ncell * t for t in range(nb * nv)])
        self.flux = [None] * 3
        if nflux:
            for d in range(layout.ndim):
                self.flux[d] = self._alloc(nb * nflux * self.ncell * 8)
                p.flux[d] = self._table([self.flux[d] + 8 * self.ncell * t
                                         for t in range(nb * nflux)])
        self.face = [None] * 3
        if faces:
            for d in range(3):
                self.face[d] = self._alloc(nb * self.nstag * 8)
                p.face[d] = self._table([self.face[d] + 8 * self.nstag * b for b in range(nb)])
        self.edge = [None] * 3
        if edges:
            for d in range(3):
                self.edge[d] = self._alloc(nb * self.nstag * 8)
                p.edge[d] = self._table([self.edge[d] + 8 * self.nstag * b for b in range(nb)])
        if blocks is not None:
            bptr = self._alloc(C.sizeof(blocks))
            check(lib().km_memcpy_h2d(C.c_void_p(bptr), C.cast(blocks, C.c_void_p),
                                      C.sizeof(blocks)), "h2d")
            p.blocks = bptr

    def _alloc(self, nbytes: int) -> int:
        ptr = C.c_void_p()
        check(lib().km_malloc(C.byref(ptr), max(nbytes, 8)), "km_malloc")
        self._allocs.append(ptr.value)
        zeros = np.zeros(max(nbytes, 8) // 8)
        check(lib().km_memcpy_h2d(ptr, f64ptr(zeros), zeros.nbytes), "h2d")
        return ptr.value

    def _table(self, ptrs) -> int:
        arr = np.array(ptrs, dtype=np.uint64)
        dev = self._alloc(arr.nbytes)
        check(lib().km_memcpy_h2d(C.c_void_p(dev), C.c_void_p(arr.ctypes.data), arr.nbytes),
              "h2d")
        return dev

    def upload(self, cell=None, flux=None, face=None, edge=None):
        def put(dev, a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            check(lib().km_memcpy_h2d(C.c_void_p(dev), f64ptr(a), a.nbytes), "h2d")
        if cell is not None:
            put(self.cell, cell)
        for d in range(3):
            if flux is not None and flux[d] is not None and self.flux[d]:
                put(self.flux[d], flux[d])
            if face is not None and face[d] is not None and self.face[d]:
                put(self.face[d], face[d])
            if edge is not None and edge[d] is not None and self.edge[d]:
                put(self.edge[d], edge[d])

    def download(self):
        L = self.layout

        def get(dev, shape):
            a = np.empty(shape)
            check(lib().km_memcpy_d2h(f64ptr(a), C.c_void_p(dev), a.nbytes), "d2h")
            return a
        check(lib().km_synchronize(), "sync")
        cell = get(self.cell, L.cell_shape(self.nv))
        flux = [get(self.flux[d], L.cell_shape(self.nflux)) if self.flux[d] else None
                for d in range(3)]
        face = [get(self.face[d], L.stag_shape()) if self.face[d] else None for d in range(3)]
        edge = [get(self.edge[d], L.stag_shape()) if self.edge[d] else None for d in range(3)]
        return cell, flux, face, edge

    def close(self):
        for p in self._allocs:
            lib().km_free(C.c_void_p(p))
        self._allocs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A small device buffer of float64 or int32 values (dt, CFL estimates, tags)."""

    def __init__(self, n: int, dtype=np.float64):
        self.n, self.dtype = n, np.dtype(dtype)
        ptr = C.c_void_p()
        check(lib().km_malloc(C.byref(ptr), max(n * self.dtype.itemsize, 8)), "km_malloc")
        self.ptr = ptr.value
        self.set(np.zeros(n, self.dtype))

    def set(self, values):
        a = np.ascontiguousarray(values, dtype=self.dtype)
        check(lib().km_memcpy_h2d(C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data), a.nbytes),
              "h2d")

    def get(self) -> np.ndarray:
        a = np.empty(self.n, self.dtype)
        check(lib().km_synchronize(), "sync")
        check(lib().km_memcpy_d2h(C.c_void_p(a.ctypes.data), C.c_void_p(self.ptr), a.nbytes),
              "d2h")
        return a

    def close(self):
        if self.ptr:
            lib().km_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StagePlan:
    """km_stage_plan_* and km_stage_fused: one RK stage of the fused kernels on
    caller-owned packs (DevicePack). `descs`: the uniform mesh(es) of a single-level mesh or
    ensemble; `desc` + `leaves`: a block tree."""

    def __init__(self, cfg: Config, descs=None, desc=None, leaves=None):
        self.h = C.c_void_p()
        if leaves is not None:
            self._leaves = np.ascontiguousarray(leaves, dtype=np.int32)
            check(lib().km_stage_plan_amr(C.byref(cfg.c), C.byref(desc), _i32(self._leaves),
                                          len(self._leaves), C.byref(self.h)),
                  "km_stage_plan_amr")
        else:
            arr = (KmMeshDesc * len(descs))(*descs)
            check(lib().km_stage_plan_uniform(C.byref(cfg.c), arr, len(descs), C.byref(self.h)),
                  "km_stage_plan_uniform")

    def set_kernel(self, dir: int, kern: int):
        check(lib().km_stage_plan_set_kernel(self.h, dir, kern), "km_stage_plan_set_kernel")

    def kernel(self, dir: int) -> int:
        v = C.c_int32()
        check(lib().km_stage_plan_kernel(self.h, dir, C.byref(v)), "km_stage_plan_kernel")
        return v.value

    def stage(self, cfg: Config, base: "DevicePack", u0: "DevicePack", u1: "DevicePack",
              beta: float, dt: DeviceArray, est: DeviceArray = None, stream=None):
        check(lib().km_stage_fused(C.byref(cfg.c), self.h, C.byref(base.pack), C.byref(u0.pack),
                                   C.byref(u1.pack), C.c_double(beta), C.c_void_p(dt.ptr),
                                   C.c_void_p(est.ptr) if est is not None else None, stream),
              "km_stage_fused")

    def close(self):
        if self.h:
            lib().km_stage_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
