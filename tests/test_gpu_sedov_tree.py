"""The block-tree benchmark configuration at its full shape: the 3-D Sedov blast on the
3,760-leaf density-gradient tree (numlevel 4, 16^3 blocks, 15.4 M zones;
tests/golden/sedov3d_density_tree.npz), PLM-MC + HLLC, RK2.

- the fused path in FP64 equals the oracle bit for bit after two cycles (the second one
  runs the fused stages; the first starts from the initial primitives on the task path);
- the task path launched pack by pack, one block per pack (parthenon/mesh/pack_size 1),
  equals the one-pack task path bit for bit;
- the fused FP32-mixed form stays within 1e-5 relative L1 per conserved variable of the
  fused FP64 form after three cycles."""
import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

NEXCH = 11  # exchanged cell fields (conserved and primitives without the temperature)


@pytest.fixture(scope="module")
def c4(gpu):
    return bench._sedov_setup(gpu, "fp64")


def _run(km, setup, ncycles, path, precision="fp64", pack_size=None):
    _, desc, lay, leaves, cell = setup
    cfg, _, _, _, _ = bench._sedov_setup(km, precision, pack_size) if (
        precision != "fp64" or pack_size) else setup
    sim = km.Simulation(cfg, desc, lay, path=path, leaves=leaves)
    sim.set_primitive(cell)
    sim.advance(ncycles)
    g, _ = sim.state()
    info = sim.info()
    sim.close()
    return g, info


def _first_diff(a, b):
    bad = np.argwhere(a != b)
    return f"{len(bad)} differ, first {tuple(bad[0])}" if len(bad) else ""


def test_fused_fp64_bitwise_with_oracle(gpu, oracle, c4):
    cfg, desc, lay, leaves, cell = c4
    print(f"C4 tree, {len(leaves)} leaves, fused FP64, 2 cycles", flush=True)
    g, info = _run(gpu, c4, 2, gpu.PATH_FUSED)
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell)
    ref.advance(2)
    r, _ = ref.state()
    rinfo = ref.info()
    ref.close()
    sl = (slice(None), slice(0, NEXCH))
    assert np.array_equal(g[sl], r[sl]), _first_diff(g[sl], r[sl])
    assert info == rinfo


def test_per_block_packs_bitwise(gpu, c4):
    print("C4 tree, task path, one pack vs one pack per block, 2 cycles", flush=True)
    g1, i1 = _run(gpu, c4, 2, gpu.PATH_TASKS)
    gb, ib = _run(gpu, c4, 2, gpu.PATH_TASKS, pack_size=1)
    sl = (slice(None), slice(0, NEXCH))
    assert np.array_equal(g1[sl], gb[sl]), _first_diff(g1[sl], gb[sl])
    assert i1 == ib


def test_fused_fp32_mixed_within_1e5(gpu, c4):
    _, _, lay, _, _ = c4
    print("C4 tree, fused FP32-mixed vs FP64, 3 cycles", flush=True)
    g64, _ = _run(gpu, c4, 3, gpu.PATH_FUSED)
    g32, _ = _run(gpu, c4, 3, gpu.PATH_FUSED, "fp32_mixed")
    l1 = bench._rel_l1(g32, g64, lay)
    print("rel L1", l1)
    assert np.all(np.array(l1) < 1e-5), l1
