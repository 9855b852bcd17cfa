"""Host-side checks of the cylindrical option surface (no GPU work): geometry/geometry
parses, and block-tree meshes refuse the axisymmetric boundary."""
from kamayan_amd import Layout, mesh_desc
import pytest

from tests import _cases


def test_cylindrical_parses(km):
    cfg, _, _, _, _ = _cases.vortex_cyl()
    assert cfg.c.geometry == 1
    with pytest.raises(km.KmError, match="invalid value"):
        km.Config(**{"geometry/geometry": "spherical"})


def test_block_trees_take_the_axisymmetric_boundary_at_the_inner_x1_edge_only(km):
    lay = Layout(2, 2, (8, 8, 1), (2, 2, 1))
    desc = mesh_desc(lay, bc=("axisymmetric", "outflow", "periodic", "periodic", "periodic",
                              "periodic"))
    leaves = km.refine_ball(mesh_desc(lay), 1, (0.5, 0.5, 0.5), 0.1)
    assert len(km.amr_blocks(desc, leaves)) == len(leaves)
    bad = mesh_desc(lay, bc=("outflow", "axisymmetric", "periodic", "periodic", "periodic",
                             "periodic"))
    with pytest.raises(km.KmError, match="inner x1 boundary"):
        km.amr_blocks(bad, leaves)
    lay3 = Layout(3, 2, (8, 8, 8), (2, 2, 2))
    bad3 = mesh_desc(lay3, bc=("axisymmetric", "outflow") + ("periodic",) * 4)
    with pytest.raises(km.KmError, match="inner x1 boundary"):
        km.amr_blocks(bad3, km.refine_ball(mesh_desc(lay3), 1, (0.5, 0.5, 0.5), 0.1))
