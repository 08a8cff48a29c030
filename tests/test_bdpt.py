"""Bidirectional path tracing (BdptWorker, src/renderingalgorithms_pathtracing.hxx:341-700) in the
oracle, pinned by closed forms and by the forward path tracer `pt` (tests/test_pt.py), which
estimates the same image.

* Furnace box: L_e / (1 - rho) with TTGPU_MIS_CONSISTENT at every depth.  With
  TTGPU_MIS_REFERENCE the reference's Russian-roulette placement (the vertex reached when the walk
  ends keeps its unscaled weight, while surviving walks scale it by 1 / p) biases vertices beyond
  node 5 upwards: the estimate equals the consistent one up to depth 5 and exceeds it beyond.
* An emitter filling the view: L_e in every pixel (light tracing splats and MIS with camera rays).
* A plane under a dome: rho L (light subpaths start on a disc across the scene's box).
* The furnace with a glass sphere and a lossless mirror sphere: still L_e / (1 - rho) (refraction in
  importance transport, delta vertices excluded from connections).
* Cornell box, and a wide view with a point light and a dome: per-pixel agreement with pt.
"""
import numpy as np
import pytest

from naivetrace_amd import abi, configs
from naivetrace_amd.scene import flat, spectrum_to_rgb
from test_oracle import furnace_scene
from test_pt import batched, emitter_wall_scene


def _params(w, h, depth, algo=abi.ALGO_BDPT, mis=abi.MIS_CONSISTENT, seed=5):
    p = abi.default_params(w, h, max_depth=depth, seed=seed, algo=algo)
    p.mis_mode = mis
    return p


def _image(mean, se, w, h):
    return mean.mean(axis=(0, 1)), np.sqrt((se ** 2).sum(axis=(0, 1))) / (w * h)


@pytest.mark.parametrize("depth", [3, 8, 60])
def test_bdpt_furnace_closed_form(oracle, depth):
    rho = 0.5
    mean, se = batched(oracle, furnace_scene(rho), _params(4, 4, depth), 256)
    img, img_se = _image(mean, se, 4, 4)
    # max_ray_depth d keeps paths of up to d + 1 nodes: camera and d surface vertices.
    expected = spectrum_to_rgb(flat(sum(rho ** k for k in range(depth))))
    assert np.all(np.abs(img - expected) < 4 * img_se + 2e-3 * expected), (img, img_se, expected)


def test_bdpt_reference_russian_roulette_bias(oracle):
    s = furnace_scene(0.5)
    for depth, biased in ((5, False), (8, True)):
        ref, ref_se = _image(*batched(oracle, s, _params(4, 4, depth, mis=abi.MIS_REFERENCE), 512), 4, 4)
        con, con_se = _image(*batched(oracle, s, _params(4, 4, depth, mis=abi.MIS_CONSISTENT), 512), 4, 4)
        diff = (ref - con) / np.sqrt(ref_se ** 2 + con_se ** 2)
        if biased:
            assert np.all(diff > 6), diff
        else:
            assert np.all(np.abs(diff) < 4), diff


def test_bdpt_visible_emitter_every_pixel(oracle):
    L = 2.0
    w, h = 8, 6
    mean, se = batched(oracle, emitter_wall_scene(L), _params(w, h, 4), 256)
    expected = spectrum_to_rgb(flat(L))
    assert np.all(np.abs(mean - expected) < 5 * se + 2e-3 * expected)


def test_bdpt_plane_under_dome_closed_form(oracle):
    from naivetrace_amd.scene import Scene
    s = Scene()
    rho, L = 0.6, 1.5
    m = s.add_material(shader=s.add_shader(abi.SHADER_DIFFUSE, reflectance=flat(rho)))
    big = 1e4
    s.add_quad([-big, 0, big], [big, 0, big], [big, 0, -big], [-big, 0, -big], m)
    s.add_dome([0, 1, 0], flat(L))
    s.set_camera([0, 1, 0], [0, 0, -0.2], [0, 0, -1], 30.0)
    img, img_se = _image(*batched(oracle, s, _params(4, 4, 3), 512), 4, 4)
    expected = spectrum_to_rgb(flat(rho * L))
    assert np.all(np.abs(img - expected) < 4 * img_se + 2e-3), (img, img_se, expected)


def specular_furnace(rho=0.5):
    """The emitting furnace box with a glass sphere and a lossless mirror sphere inside: in an
    isotropic radiance field L a lossless dielectric and a perfect mirror leave the radiance in air
    at L, so every pixel still sees L_e / (1 - rho).  Light subpaths refract through the glass in
    importance transport (no 1 / eta^2 radiance scaling), and connections skip delta vertices."""
    from naivetrace_amd.scene import Scene
    s = furnace_scene(rho)
    mirror = s.add_material(shader=s.add_shader(abi.SHADER_MIRROR, reflectance=flat(1.0)))
    glass = s.add_material(shader=s.add_shader(abi.SHADER_DIELECTRIC, ior=1.5))
    s.add_sphere([-0.4, 0.1, 0.5], 0.3, mirror)
    s.add_sphere([0.35, -0.2, 0.45], 0.35, glass)
    return s


@pytest.mark.parametrize("algo", [abi.ALGO_BDPT, abi.ALGO_PT])
def test_specular_furnace_closed_form(oracle, algo):
    rho = 0.5
    mean, se = batched(oracle, specular_furnace(rho), _params(4, 4, 40, algo=algo), 512)
    img, img_se = _image(mean, se, 4, 4)
    expected = spectrum_to_rgb(flat(sum(rho ** k for k in range(40))))
    print(f"specular furnace algo {algo}: {img} +- {img_se}, expected {expected}", flush=True)
    assert np.all(np.abs(img - expected) < 4 * img_se + 3e-3 * expected), (img, img_se, expected)


def pointlit_dome(w, h):
    from golden.make_oracle_images import pointlit_dome as build
    return build(w, h)


@pytest.mark.parametrize("name,builder", [("cornell", lambda w, h: configs.cornell_box(w, h)), ("pointlit_dome", pointlit_dome)])
def test_bdpt_agrees_with_pt(oracle, name, builder):
    w = h = 16
    s = builder(w, h)
    mb, sb = batched(oracle, s, _params(w, h, 8), 128)
    mp, sp = batched(oracle, s, _params(w, h, 8, algo=abi.ALGO_PT, seed=9), 256)
    z = (mb - mp) / np.sqrt(sb ** 2 + sp ** 2 + 1e-30)
    frac3 = float(np.mean(np.abs(z) > 3))
    print(f"{name}: bdpt vs pt mean z {z.mean():+.3f}, |z|>3 {frac3:.4f}", flush=True)
    assert frac3 < 0.04
    ib, ib_se = _image(mb, sb, w, h)
    ip, ip_se = _image(mp, sp, w, h)
    assert np.all(np.abs(ib - ip) < 4 * np.sqrt(ib_se ** 2 + ip_se ** 2) + 1e-3 * ip), (ib, ip)
