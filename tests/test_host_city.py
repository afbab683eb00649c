"""The procedural city generator (the wavefront-batching workload): sizes,
the per-mesh triangle cap, outward winding of every surface, downward-facing
lamps, materials and the four cameras of uneven resolution.  Host only."""
import numpy as np

from tpt import scenes


def small():
    return scenes.city(grid=12, rocks=300, lights=64, sky_size=(64, 32), tex_size=32, build=False)


def normals(mesh):
    v = mesh.positions[mesh.indices].astype(np.float64)
    return np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])


def test_counts_and_cameras():
    c = small()
    tris = [m.triangle_count() for m in c.scene.meshes]
    towers = sum(tris[:-3])
    assert towers == 22 * 12 * 12  # 8 side quads + 6-triangle roof per tower
    assert tris[-3] == 20 * 300 and tris[-2] == 2 and tris[-1] == 64 == c.emissive_triangles
    assert len(c.materials) == 9 and len(c.scene.instances) == len(c.scene.meshes)
    assert [(w, h) for _, w, h in c.cameras] == [(1920, 1080), (1280, 720), (640, 360), (3840, 2160)]
    assert c.scene.environment is not None


def test_towers_face_outward_and_roofs_up():
    c = small()
    for inst in c.scene.instances[:-3]:
        m = c.scene.meshes[inst.mesh_index]
        n = normals(m)
        v = m.positions[m.indices].astype(np.float64)
        cen = v.mean(1)
        # each tower's 22 triangles: its axis is the mean of its 16 vertices
        t = len(m.indices) // 22
        axis = m.positions.reshape(t, 16, 3).astype(np.float64).mean(1)
        ax = np.repeat(axis, 22, axis=0)
        side = np.tile(np.r_[np.ones(16, bool), np.zeros(6, bool)], t)
        out = ((cen - ax)[:, [0, 2]] * n[:, [0, 2]]).sum(1)
        assert np.all(out[side] > 0)
        assert np.all(n[~side][:, 1] > 0)


def test_ground_up_lamps_down_and_uvs():
    c = small()
    ground, lamps = c.scene.meshes[-2], c.scene.meshes[-1]
    assert np.all(normals(ground)[:, 1] > 0) and ground.uvs is not None
    assert np.all(normals(lamps)[:, 1] < 0)
    lamp_mat = c.scene.instances[-1].material_index
    assert c.materials[lamp_mat].emission_luminance > 0
    textured = [i for i in c.scene.instances if i.material_index == 7]
    assert textured and all(c.scene.meshes[i.mesh_index].uvs is not None for i in textured)


def test_full_size_fits_the_per_mesh_cap():
    c = scenes.city(build=False, sky_size=(64, 32), tex_size=32)
    tris = [m.triangle_count() for m in c.scene.meshes]
    assert 9_000_000 < sum(tris) < 10_500_000
    assert max(tris) < 1 << 24 and c.emissive_triangles == 2048
