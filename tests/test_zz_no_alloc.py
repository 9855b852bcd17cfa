"""km_amr_tag_loehner_noalloc (the tagging entry point that allocates nothing: its
per-block maxima live in a library-held device array) against km_amr_tag_loehner through
the C ABI: the same tags, bit for bit, for several criteria with and without reset, on a
3-D block tree and then on a smaller 2-D one (the second call reuses the array the first
filled, so a maximum that is not cleared would show)."""
import ctypes as C

import numpy as np
import pytest

from kamayan_amd._abi import amr_criterion, slot
from tests import _cases

pytestmark = pytest.mark.gpu


def _state(oracle, ndim, nroot):
    cfg, desc, lay, leaves, cell = _cases.amr_case(
        "sedov", ndim=ndim, nroot=nroot, nblock=8, ng=4, lmax=1, center=(0.0, 0.0, 0.0),
        radius=0.2, bc="outflow", xmin=(-0.5,) * 3, xmax=(0.5,) * 3, gamma=1.4,
        params=[1.0, 0.12, 1.0, 1e-5])
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell)
    ref.advance(2)
    state, _ = ref.state()
    return desc, lay, leaves, state


def test_noalloc_tags_match_the_allocating_entry_point(gpu, oracle):
    km = gpu
    L = km.lib()
    for ndim, nroot in ((3, 4), (2, 4)):  # 3-D first, then the smaller 2-D pack
        desc, lay, leaves, state = _state(oracle, ndim, nroot)
        n = len(leaves)
        print(f"no-alloc tags: {ndim}-D, {n} leaves", flush=True)
        assert n <= km._abi.KM_TAG_NOALLOC_MAX_BLOCKS
        dev = km.DevicePack(lay, 12, blocks=km.amr_blocks(desc, leaves))
        dev.upload(cell=state)
        t_a, t_b = km.DeviceArray(n, np.int32), km.DeviceArray(n, np.int32)
        seen = set()
        for var, rtol, dtol in [(slot("PRES"), 0.8, 0.2), (slot("DENS"), 0.3, 0.05),
                                (slot("VEL0"), 0.5, 0.1)]:
            crit = amr_criterion(var, 2, refine_tol=rtol, derefine_tol=dtol)
            for reset in (1, 0):
                km.check(L.km_amr_tag_loehner(C.byref(dev.pack), C.byref(crit),
                                              C.c_void_p(t_a.ptr), reset, None), "tags")
                km.check(L.km_amr_tag_loehner_noalloc(C.byref(dev.pack), C.byref(crit),
                                                      C.c_void_p(t_b.ptr), reset, None),
                         "tags without allocation")
                a, b = t_a.get(), t_b.get()
                assert np.array_equal(a, b), (ndim, var, reset, np.argwhere(a != b)[:4])
                seen |= set(a.tolist())
        assert len(seen) > 1  # the criteria tag some blocks differently
        for x in (dev, t_a, t_b):
            x.close()


def test_noalloc_refuses_more_blocks_than_its_array(gpu):
    km = gpu
    from kamayan_amd._abi import KmPack
    pack = KmPack()
    pack.nblocks = km._abi.KM_TAG_NOALLOC_MAX_BLOCKS + 1
    pack.ndim, pack.ng, pack.nvar = 3, 4, 12
    crit = amr_criterion(slot("DENS"), 2)
    rc = km.lib().km_amr_tag_loehner_noalloc(C.byref(pack), C.byref(crit), C.c_void_p(16), 1,
                                             None)
    assert rc == km._abi.KM_EINVAL
    assert b"scratch limit" in km.lib().km_last_error()
