"""The NFF scene loader (naivetrace_amd/csrc/scene_file.cpp) through the C ABI, on the CPU.

Fixture scenes are the reference's own test scenes (tests/golden/scenes, copied by
tests/golden/copy_scenes.py); the expected values come from the files' contents: COLLADA vertex
coordinates in centimetres (asset unit 0.01 m), node translations, the NFF's shader definitions.
"""
import os
import textwrap

import numpy as np
import pytest

from naivetrace_amd import abi
from naivetrace_amd.scene import rgb_to_spectrum, spectrum_to_rgb
from naivetrace_amd.scene_file import SceneFile
import oracle_lib

SCENES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


def arrays(sf):
    d = sf.to_abi()
    nv, nt, ns = d.num_vertices, d.num_triangles, d.num_spheres
    v = np.ctypeslib.as_array(d.vertices, (nv * 3,)).reshape(nv, 3) if nv else np.zeros((0, 3))
    n = np.ctypeslib.as_array(d.normals, (nv * 3,)).reshape(nv, 3) if nv else np.zeros((0, 3))
    idx = np.ctypeslib.as_array(d.indices, (nt * 3,)).reshape(nt, 3) if nt else np.zeros((0, 3), int)
    tm = np.ctypeslib.as_array(d.tri_material, (nt,)) if nt else np.zeros(0, int)
    sp = np.ctypeslib.as_array(d.spheres, (ns * 4,)).reshape(ns, 4) if ns else np.zeros((0, 4))
    sm = np.ctypeslib.as_array(d.sphere_material, (ns,)) if ns else np.zeros(0, int)
    mats = [d.materials[i] for i in range(d.num_materials)]
    shaders = [d.shaders[i] for i in range(d.num_shaders)]
    return d, v, n, idx, tm, sp, sm, mats, shaders


def nff(text, tmp_path=None, name="scene.nff"):
    text = textwrap.dedent(text)
    if tmp_path is None:
        return SceneFile.from_text(text)
    p = tmp_path / name
    p.write_text(text)
    return SceneFile(str(p))


def refl(shader):
    return np.array(shader.reflectance[:])


def test_cubelight_unit_cube_from_collada():
    sf = SceneFile(os.path.join(SCENES, "cubelight.nff"))
    d, v, n, idx, tm, *_ , mats, shaders = arrays(sf)
    assert (sf.width, sf.height) == (120, 120)
    assert d.camera.kind == abi.CAMERA_PINHOLE and d.camera.fov_deg == 80
    np.testing.assert_allclose([d.camera.dir[k] for k in range(3)], [0, -0.5, -1])
    assert d.num_triangles == 12
    # unitcube.dae: vertices at +-50 in centimetres -> the unit cube.
    np.testing.assert_allclose(v.min(axis=0), [-0.5] * 3, atol=1e-6)
    np.testing.assert_allclose(v.max(axis=0), [0.5] * 3, atol=1e-6)
    # Shader "black" (diffuse, zero albedo) and the area light of the enclosing scope.
    m = mats[tm[0]]
    assert np.all(tm == tm[0]) and m.emitter == 0 and shaders[m.shader].kind == abi.SHADER_DIFFUSE
    assert np.all(refl(shaders[m.shader]) == 0)
    np.testing.assert_allclose(d.area_emitters[0].spectrum[:], 5 * rgb_to_spectrum([1, 1, 1]), rtol=1e-6)
    # Normals: unit length, facing out of the cube (per-corner normals of the file).
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
    c = v[idx].mean(axis=1)
    assert np.all(np.einsum("ij,ij->i", n[idx[:, 0]], c) > 0)


def test_cornell_box_collada_materials_and_yaml_transform():
    sf = SceneFile(os.path.join(SCENES, "cornelbox_translucentsphere.nff"))
    d, v, n, idx, tm, sp, sm, mats, shaders = arrays(sf)
    # Spheres: glass shells of ior 1.4 and 0.71 around (0, 0.5, -0.55).
    np.testing.assert_allclose(sp, [[0, 0.5, -0.55, 0.3], [0, 0.5, -0.55, 0.29]], rtol=1e-6)
    assert [shaders[mats[m].shader].kind for m in sm] == [abi.SHADER_DIELECTRIC] * 2
    np.testing.assert_allclose([shaders[mats[m].shader].ior_ratio for m in sm], [1.4, 0.71], rtol=1e-6)
    emissive = np.array([mats[m].emitter >= 0 for m in tm])
    box = ~emissive
    # The box (translate 0 50 -60 cm; faces at x, y = +-50 cm, z = +-60 cm): x in [-0.5, 0.5],
    # y in [0, 1], z in [-1.2, 0].
    np.testing.assert_allclose(v[idx[box]].reshape(-1, 3).min(axis=0), [-0.5, 0, -1.2], atol=1e-6)
    np.testing.assert_allclose(v[idx[box]].reshape(-1, 3).max(axis=0), [0.5, 1, 0], atol=1e-6)
    # Materials by the COLLADA material names: white (0.8), red (1 0 0 x 0.2), green, blue.
    want = {tuple(np.round(rgb_to_spectrum(np.array(c) * k), 6)) for c, k in
            (([1, 1, 1], 0.8), ([1, 0, 0], 0.2), ([0, 1, 0], 0.2), ([1, 1, 1], 0.2))}
    got = {tuple(np.round(refl(shaders[mats[m].shader]), 6)) for m in set(tm[box])}
    assert got == want
    # The light: unitrectangle.obj under pos [0, 0.99, -0.55], hpb [0, 0, 180] deg, scale [0.4, 1, 0.4]:
    # a 0.4 x 0.4 square at y = 0.99 facing down.
    lv = v[idx[emissive]].reshape(-1, 3)
    np.testing.assert_allclose(lv[:, 1], 0.99, atol=1e-6)
    np.testing.assert_allclose(lv.min(axis=0)[[0, 2]], [-0.2, -0.75], atol=1e-6)
    np.testing.assert_allclose(lv.max(axis=0)[[0, 2]], [0.2, -0.35], atol=1e-6)
    t = v[idx[emissive]]
    fn = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    assert np.all(fn[:, 1] < 0)
    # Dome light: lddome 0 0 1 with RGB (0.02, 0.02, 0.05).
    e = d.env_lights[0]
    assert e.kind == abi.ENV_DOME and [e.dir[k] for k in range(3)] == [0, 0, 1]
    np.testing.assert_allclose(e.spectrum[:], rgb_to_spectrum([0.02, 0.02, 0.05]), rtol=1e-5)


def test_all_fixture_scenes_load():
    for f in sorted(os.listdir(SCENES)):
        if f.endswith(".nff"):
            sf = SceneFile(os.path.join(SCENES, f))
            d = sf.to_abi()
            assert d.num_triangles + d.num_spheres > 0, f
            for m in range(d.num_materials):
                mat = d.materials[m]
                assert -1 <= mat.shader < d.num_shaders and -1 <= mat.medium < d.num_media


def test_scopes_keep_their_changes():
    # Shader b is defined only inside the scope: activating it afterwards fails like the reference.
    with pytest.raises(ValueError, match="Shader b not defined"):
        nff("""
            diffuse a 1 0 0 0.5
            {
            diffuse b 0 1 0 0.5
            }
            shader b
            """)
    sf = nff("""
        diffuse a 1 0 0 0.5
        {
        diffuse b 0 1 0 0.5
        transform 5 0 0
        s 0 0 0 1
        }
        s 1 0 0 1
        """)
    d, *_, sp, sm, mats, shaders = arrays(sf)
    np.testing.assert_allclose(sp[:, :3], [[5, 0, 0], [1, 0, 0]])
    np.testing.assert_allclose(refl(shaders[mats[sm[0]].shader]), rgb_to_spectrum([0, 0.5, 0]), rtol=1e-6)
    np.testing.assert_allclose(refl(shaders[mats[sm[1]].shader]), rgb_to_spectrum([0.5, 0, 0]), rtol=1e-6)


def test_transform_rotation_scale_and_polygon_normals():
    # translate (1, 2, 3), heading pi/2 about y, scale x by 2: (1, 0, 0) -> (1, 2, 3) + R_y(pi/2) (2, 0, 0) = (1, 2, 1).
    sf = nff("""
        transform 1 2 3 1.5707963267948966 0 0 2 1 1
        p 3
        0 0 0
        1 0 0
        0 1 0
        p 3
        0 0 0  0 0 1
        1 0 0  0 0 1
        0 1 0  1 0 1
        """)
    d, v, n, idx, *_ = arrays(sf)
    np.testing.assert_allclose(v[idx[0]], [[1, 2, 3], [1, 2, 1], [1, 3, 3]], atol=1e-6)
    # First polygon: flat normal of the transformed triangle.  Second: given normals through the
    # inverse transpose (z -> R_y z = x; scale 2 in x shrinks x components).
    fn = np.cross(v[idx[0, 1]] - v[idx[0, 0]], v[idx[0, 2]] - v[idx[0, 0]])
    np.testing.assert_allclose(n[idx[0]], np.tile(fn / np.linalg.norm(fn), (3, 1)), atol=1e-6)
    np.testing.assert_allclose(n[idx[1, 0]], [1, 0, 0], atol=1e-6)
    exp = np.array([0.5, 0, 1.0])          # (1, 0, 1) scaled by (1/2, 1, 1) in the local frame
    exp = np.array([exp[2], 0, -exp[0]])   # heading pi/2: x -> -z, z -> x
    np.testing.assert_allclose(n[idx[1, 2]], exp / np.linalg.norm(exp), atol=1e-6)


def test_media_priorities_and_phase_functions():
    sf = nff("""
        medium m1 0.2 0.01
        pf henleygreenstein 0.5
        medium m2 1 2 3 0.1 0.2 0.3
        pf rayleigh
        vacuummedium clean
        shader invisible
        medium m1
        s 0 0 0 1
        """)
    d, *_, sp, sm, mats, shaders = arrays(sf)
    media = [d.media[i] for i in range(d.num_media)]
    assert [m.kind for m in media] == [abi.MEDIUM_HOMOGENEOUS, abi.MEDIUM_HOMOGENEOUS, abi.MEDIUM_VACUUM]
    # Priority = number of media known when defined ("default" counts).
    assert [m.priority for m in media] == [1, 2, 3]
    assert media[0].phase_kind == abi.PHASE_HENYEY_GREENSTEIN and abs(media[0].phase_g - 0.5) < 1e-7
    assert media[1].phase_kind == abi.PHASE_RAYLEIGH
    np.testing.assert_allclose(media[1].sigma_s[:], rgb_to_spectrum([1, 2, 3]), rtol=1e-6)
    np.testing.assert_allclose(media[0].sigma_a[:], rgb_to_spectrum([0.01] * 3), rtol=1e-6)
    m = mats[sm[0]]
    assert m.medium == 0 and shaders[m.shader].kind == abi.SHADER_INVISIBLE


def test_lights():
    sf = nff("""
        lsun 1 0 0 200
        lsun 0 -1 0 50 2.5
        lddirection 0 2 -1 2 2 1
        lddome 0 0 1 0.2 0.2 0.2
        l 1 2 3
        l 1 2 3 1 0.5 0.25 4
        larea lamp parallel 1 1 1 7
        s 0 0 0 1
        """)
    d, *_, sp, sm, mats, shaders = arrays(sf)
    e = [d.env_lights[i] for i in range(d.num_env_lights)]
    # NFF vectors give the direction light travels; the ABI takes the direction towards the light.
    assert e[0].kind == abi.ENV_SUN and [e[0].dir[k] for k in range(3)] == [-1, 0, 0]
    assert e[0].total_power == 200 and abs(e[0].opening_angle_deg - 0.26) < 1e-12
    assert e[1].opening_angle_deg == 2.5 and e[1].dir[1] == 1
    np.testing.assert_allclose([e[2].dir[k] for k in range(3)], -np.array([0, 2, -1]) / np.sqrt(5))
    assert e[2].kind == abi.ENV_DIRECTIONAL and e[3].kind == abi.ENV_DOME
    np.testing.assert_allclose(e[2].spectrum[:], rgb_to_spectrum([2, 2, 1]), rtol=1e-6)
    pts = [d.point_lights[i] for i in range(d.num_point_lights)]
    np.testing.assert_allclose(pts[0].spectrum[:], rgb_to_spectrum([1, 1, 1]), rtol=1e-6)
    np.testing.assert_allclose(pts[1].spectrum[:], rgb_to_spectrum([4, 2, 1]), rtol=1e-6)
    assert d.area_emitters[mats[sm[0]].emitter].kind == abi.AREA_PARALLEL


def test_obj_negative_indices_usemtl_and_fans(tmp_path):
    (tmp_path / "quad.obj").write_text(textwrap.dedent("""
        v 0 0 0
        v 1 0 0
        v 1 0 1
        v 0 0 1
        vn 0 1 0
        usemtl red
        f -4//1 -3//1 -2//1 -1//1
        usemtl DefaultMaterial
        f 1 2 3
        """))
    sf = nff("""
        diffuse red 1 0 0 0.5
        diffuse other 0 0 1 0.5
        m quad.obj
        s 0 0 0 1
        """, tmp_path)
    d, v, n, idx, tm, sp, sm, mats, shaders = arrays(sf)
    assert d.num_triangles == 3
    # The quad is a fan of two triangles with the given normal; usemtl red picks shader red.
    np.testing.assert_allclose(n[idx[:2]].reshape(-1, 3), np.tile([0, 1, 0], (6, 1)))
    np.testing.assert_allclose(refl(shaders[mats[tm[0]].shader]), rgb_to_spectrum([0.5, 0, 0]), rtol=1e-6)
    # A name look-up activates that shader in the mesh's own copy of the scope
    # (src/parsenff.cxx:47-68, 922-930): "DefaultMaterial" then takes red as well ...
    np.testing.assert_allclose(refl(shaders[mats[tm[2]].shader]), rgb_to_spectrum([0.5, 0, 0]), rtol=1e-6)
    # ... while the file's scope keeps "other" active.
    np.testing.assert_allclose(refl(shaders[mats[sm[0]].shader]), rgb_to_spectrum([0, 0, 0.5]), rtol=1e-6)
    fn = np.cross(v[idx[2, 1]] - v[idx[2, 0]], v[idx[2, 2]] - v[idx[2, 0]])
    np.testing.assert_allclose(n[idx[2, 0]], fn / np.linalg.norm(fn))


def test_collada_nodes_units_and_up_axis(tmp_path):
    dae = """<?xml version="1.0"?>
    <COLLADA version="1.4.1">
      <asset><unit meter="0.5"/><up_axis>Z_UP</up_axis></asset>
      <library_geometries><geometry id="g"><mesh>
        <source id="p"><float_array count="9">0 0 0 2 0 0 0 2 0</float_array>
          <technique_common><accessor count="3" stride="3"/></technique_common></source>
        <vertices id="vx"><input semantic="POSITION" source="#p"/></vertices>
        <triangles count="1"><input offset="0" semantic="VERTEX" source="#vx"/><p>0 1 2</p></triangles>
      </mesh></geometry></library_geometries>
      <library_visual_scenes><visual_scene id="s">
        <node><translate>0 0 4</translate><instance_geometry url="#g"/>
          <node><scale>1 1 3</scale><rotate>0 0 1 90</rotate><instance_geometry url="#g"/></node>
        </node>
      </visual_scene></library_visual_scenes>
    </COLLADA>"""
    (tmp_path / "tri.dae").write_text(dae)
    sf = nff("m tri.dae\n", tmp_path)
    d, v, n, idx, *_ = arrays(sf)
    # Z_UP -> Y_UP maps (x, y, z) -> (x, z, -y); unit 0.5.  First instance: translate z by 4.
    np.testing.assert_allclose(v[idx[0]], 0.5 * np.array([[0, 4, 0], [2, 4, 0], [0, 4, -2]]), atol=1e-6)
    # Second: rotate 90 deg about z, then scale z by 3 (no effect on z = 0 points), then translate.
    np.testing.assert_allclose(v[idx[1]], 0.5 * np.array([[0, 4, 0], [0, 4, -2], [-2, 4, 0]]), atol=1e-6)


def test_errors_name_the_line():
    with pytest.raises(ValueError, match=r"^2: Unknown directive \[bogus 1 2 3\]"):
        nff("s 0 0 0 1\nbogus 1 2 3\n")
    with pytest.raises(ValueError, match="textures are not supported"):
        nff("diffuse d 1 1 1 1\ndiffusetexture checker.png\n")
    with pytest.raises(RuntimeError, match="could not open data.json"):
        nff("tabulatedatmosphere a 0 0 0 6300 data.json\n")
    with pytest.raises(ValueError, match="Bad number of arguments to transform"):
        nff("transform 1 2\n")
    with pytest.raises(ValueError, match="yaml"):
        nff("yaml{\ntransform:\n  pos: [1, 2]\n}yaml\n")
    with pytest.raises(RuntimeError, match="could not open"):
        SceneFile("/nonexistent/scene.nff")


# A diffuse plane (albedo 0.6) under a uniform dome (radiance 1.5): every camera path sees
# albedo x dome radiance per wavelength, so the image mean is M (refl * L) over the 36 bins.
PLANE_UNDER_DOME = """
v
from 0 1 0
at 0 0 -1
up 0 1 0
resolution 32 32
angle 30
lddome 0 1 0 1.5 1.5 1.5
diffuse plane 1 1 1 0.6
p 4
-10000 0 10000
10000 0 10000
10000 0 -10000
-10000 0 -10000
"""


def plane_under_dome_expected():
    return spectrum_to_rgb(rgb_to_spectrum([0.6] * 3) * rgb_to_spectrum([1.5] * 3))


def test_nff_plane_under_dome_oracle_closed_form():
    sf = nff(PLANE_UNDER_DOME)
    assert (sf.width, sf.height) == (32, 32)
    p = abi.default_params(32, 32, max_depth=3, seed=3)
    spp = 64
    rgb, sq = oracle_lib.render(sf, p, spp)
    mean = rgb.sum(axis=(0, 1)) / (32 * 32 * spp)
    np.testing.assert_allclose(mean, plane_under_dome_expected(), rtol=5e-3)
