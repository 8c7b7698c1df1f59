import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle"), GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name)))


class _Scene:
    """SceneData-like view over scene_* arrays of a fixture."""

    def __init__(self, d):
        self.objects = d["scene_objects"].reshape(-1, 4)
        self.tris = d["scene_tris"].reshape(-1, 9)
        self.spheres = d["scene_spheres"].reshape(-1, 4)
        self.materials = d["scene_materials"].reshape(-1, 12)
        self.background = d["scene_background"]
        self.camera = d["scene_camera"]
        self.env = None
        if "scene_env_pixels" in d:
            w, h = (int(x) for x in d["scene_env_wh"])
            self.env = d["scene_env_pixels"].reshape(h, w, 3)
        self.point_lights = None
        self.name = "fixture"
        self.tri_flags = self.tri_normals = self.tri_uvs = self.tri_tangents = None
        self.mat_tex = self.textures = None
        if "scene_tri_flags" in d:
            a = d["scene_tri_attr"].reshape(-1, 27)
            self.tri_flags = d["scene_tri_flags"].astype(np.uint8)
            self.tri_normals, self.tri_uvs, self.tri_tangents = a[:, :9], a[:, 9:15], a[:, 15:]
        if "scene_mat_tex" in d:
            self.mat_tex = d["scene_mat_tex"].reshape(-1, 4)
            whc = d["scene_tex_whc"].reshape(-1, 3)
            px, off, self.textures = d["scene_texels"], 0, []
            for w, h, c in whc:
                self.textures.append(px[off:off + w * h * c].reshape(h, w, c))
                off += w * h * c


def fixture_scene(d):
    return _Scene(d)


@pytest.fixture(scope="session")
def golden():
    return load_golden
