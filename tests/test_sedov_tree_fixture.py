"""The block tree and blast state the block-tree benchmark runs on
(tests/golden/sedov3d_density_tree.npz, written by make_sedov_density_tree.py from an
adaptive run of the sedov problem refined on the Loehner density estimator), checked
without a GPU:

- the leaves form a 3-level refinement of the 4^3 root blocks (numlevel 4) with 2,000 to
  5,000 blocks of uneven level, and their volumes tile the domain exactly;
- the recorded blast follows the Sedov-Taylor solution: the shock radius is
  xi0 (E t^2 / rho)^(1/5) with xi0 = 1.033 for gamma = 1.4 in three dimensions (Sedov,
  Similarity and Dimensional Methods in Mechanics, 1959, ch. IV; the value Taylor
  tabulated for air), within 3 % (twice or half the energy moves it by 15 %), and the
  shell's peak density stays below the strong-shock limit (gamma+1)/(gamma-1) = 6;
- parthenon/mesh/pack_size parses as -1 or a positive block count."""
import os

import numpy as np
import pytest

import kamayan_amd as km

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "sedov3d_density_tree.npz")


@pytest.fixture(scope="module")
def tree():
    return np.load(FIXTURE)


def test_tree_levels_and_size(tree):
    leaves = tree["leaves"]
    assert int(tree["numlevel"]) == 4
    assert 2000 <= len(leaves) <= 5000
    per = np.bincount(leaves[:, 0], minlength=4)
    assert per[0] == 0 and (per[1:] > 0).all(), per
    assert len(set(per[1:].tolist())) == 3  # uneven
    # every leaf lies inside the root grid at its level
    for L in range(4):
        sel = leaves[:, 0] == L
        assert (leaves[sel, 1:] >= 0).all() and (leaves[sel, 1:] < 4 << L).all()


def test_tree_tiles_the_domain(tree):
    root = km.Layout(3, 4, (16,) * 3, (4,) * 3)
    desc = km.mesh_desc(root, xmin=(-0.5,) * 3, xmax=(0.5,) * 3, bc=("outflow",) * 6)
    blocks = km.amr_blocks(desc, tree["leaves"])
    vol = sum(b.vol * 16 ** 3 for b in blocks)
    assert abs(vol - 1.0) < 1e-12
    # no two leaves overlap: mark the finest-level cells each covers
    cover = np.zeros((32,) * 3, np.int32)
    for L, i, j, k in tree["leaves"]:
        s = 1 << (3 - L)
        cover[k * s:(k + 1) * s, j * s:(j + 1) * s, i * s:(i + 1) * s] += 1
    assert (cover == 1).all()


def test_blast_follows_sedov_taylor(tree):
    r, dens, t = tree["r"], tree["dens"], float(tree["time"])
    xi0, energy, rho0 = 1.033, 1.0, 1.0
    r_st = xi0 * (energy * t * t / rho0) ** 0.2
    r_shock = r[np.where(dens > 1.5)[0].max()]
    assert abs(r_shock / r_st - 1.0) < 0.03, (r_shock, r_st)
    gamma = 1.4
    assert 3.0 < dens.max() < (gamma + 1) / (gamma - 1)
    assert abs(dens[r > 1.2 * r_st] - 1.0).max() < 1e-6  # undisturbed ahead of the shock


def test_pack_size_key():
    assert km.Config().c.pack_size == -1
    assert km.Config(**{"parthenon/mesh/pack_size": 1}).c.pack_size == 1
    assert km.Config(**{"parthenon/mesh/pack_size": -1}).c.pack_size == -1
    for bad in (0, -2, "x"):
        with pytest.raises(km.KmError):
            km.Config(**{"parthenon/mesh/pack_size": bad})
