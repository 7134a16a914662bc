"""Device-side normal encoding and map files written from the engine: the
device codes equal the host encoder (itself equal to the numpy restatement
of NormalMap::Encode), and the files read back to the engine's maps."""
import numpy as np
import pytest

from sensemap_amd import io as pmio
from sensemap_amd import patch_match as pmm
from sensemap_amd import scene
from tests import helpers

pytestmark = pytest.mark.gpu


def test_device_normal_codes_and_map_files(tmp_path):
    prob = scene.to_problem(helpers.small_planes(seed=3))
    e = pmm.PatchMatchACMMCuda(helpers.c1_options(num_iterations=3), prob)
    e.Run()
    normal = e.GetNormalMap()
    codes = e.GetNormalCodes()
    assert np.array_equal(codes, pmio.encode_normals(normal))
    dp, np_ = str(tmp_path / "d.bin"), str(tmp_path / "n.bin")
    e.WriteMaps(dp, np_)
    assert np.array_equal(pmio.read_depth_map(dp), e.GetDepthMap())
    assert np.array_equal(pmio.read_mat(np_, np.uint32)[0], codes)
    e.close()
