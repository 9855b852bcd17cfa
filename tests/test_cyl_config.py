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


def test_block_trees_refuse_the_axisymmetric_boundary(km):
    lay = Layout(2, 2, (8, 8, 1), (2, 2, 1))
    desc = mesh_desc(lay, bc=("axisymmetric", "outflow", "periodic", "periodic", "periodic",
                              "periodic"))
    with pytest.raises(km.KmError, match="block-tree meshes take"):
        km.amr_blocks(desc, km.refine_ball(mesh_desc(lay), 1, (0.5, 0.5, 0.5), 0.1))
