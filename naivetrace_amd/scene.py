"""Host-side scene description: numpy arrays flattened into the ttgpu_scene ABI.

A `Scene` collects what ToyTrace's loader fills into its `Scene`
(src/scene.hxx:134-315): geometry with a material per primitive, shaders,
media, area emitters, environment and point lights and the camera.  Spectra
are 36-bin arrays (src/spectral.hxx:270-274); RGB inputs go through the Smits
conversion exactly like the YAML loader's ParseSpectrum
(src/parse_yaml_scene.cxx:195-207).
"""
import ctypes as C

import numpy as np

from . import abi
from .spectral_data import BIN_CENTERS, SMITS, SPEC_TO_RGB

NBINS = abi.NBINS
_SMITS = np.array(SMITS)
_TO_RGB = np.array(SPEC_TO_RGB)
WAVELENGTHS = np.array(BIN_CENTERS)


def rgb_to_spectrum(rgb):
    """Smits (2000) RGB -> 36-bin spectrum, src/spectral.cxx:81-116."""
    r, g, b = (float(x) for x in rgb)
    W, R, G, B, Y, Cy, M = range(7)

    def lin(a, ia, bb, ib, c, ic):
        return a * _SMITS[ia] + bb * _SMITS[ib] + c * _SMITS[ic]
    if r <= g and r <= b:
        return lin(r, W, g - r, Cy, b - g, B) if g <= b else lin(r, W, b - r, Cy, g - b, G)
    if g <= r and g <= b:
        return lin(g, W, r - g, M, b - r, B) if r <= b else lin(g, W, b - g, M, r - b, R)
    return lin(b, W, r - b, Y, g - r, G) if r <= g else lin(b, W, g - b, Y, r - g, R)


def spectrum_to_rgb(spec):
    """36-bin spectrum -> linear RGB, src/spectral.cxx:118-128."""
    return _TO_RGB @ np.asarray(spec, dtype=np.float64)


def flat(v):
    return np.full(NBINS, float(v))


def _spec(s):
    a = np.asarray(s, dtype=np.float64)
    if a.shape == (3,):
        a = rgb_to_spectrum(a)
    if a.shape != (NBINS,):
        raise ValueError("spectrum must have 36 bins or be an RGB triple")
    return abi.Spectrum(*[float(x) for x in a])


class Scene:
    def __init__(self):
        self.vertices = []
        self.normals = []
        self.indices = []
        self.tri_material = []
        self.spheres = []
        self.sphere_material = []
        self._nvert = 0
        self.shaders = []
        self.media = []
        self.materials = []
        self.emitters = []
        self.envs = []
        self._env_textures = []
        self.points = []
        self.camera = abi.Camera()
        self.empty_space_medium = -1
        self._has_normals = False
        self._keep = []

    # ---- tables -----------------------------------------------------------
    def add_shader(self, kind, reflectance=0.8, alpha=0.0, ior=1.5, ior_lambda_coeff=0.0, specular=0.0, alpha_min=None):
        s = abi.Shader()
        s.kind = kind
        s.alpha = alpha
        s.alpha_min = alpha if alpha_min is None else alpha_min
        s.ior_ratio = ior
        s.ior_lambda_coeff = ior_lambda_coeff
        s.specular = specular
        s.reflectance = _spec(reflectance if np.ndim(reflectance) else flat(reflectance))
        self.shaders.append(s)
        return len(self.shaders) - 1

    def add_homogeneous_medium(self, sigma_s, sigma_a, phase=abi.PHASE_UNIFORM, g=0.0, priority=None):
        m = abi.Medium()
        m.kind = abi.MEDIUM_HOMOGENEOUS
        m.priority = len(self.media) if priority is None else priority
        m.phase_kind = phase
        m.phase_g = g
        m.sigma_s = _spec(sigma_s if np.ndim(sigma_s) else flat(sigma_s))
        m.sigma_a = _spec(sigma_a if np.ndim(sigma_a) else flat(sigma_a))
        self.media.append(m)
        return len(self.media) - 1

    def add_atmosphere(self, center, radius, height, rayleigh_h, mie_h, mie_g, rayleigh_sigma_s, mie_sigma_s, mie_sigma_a,
                       priority=None):
        m = abi.Medium()
        m.kind = abi.MEDIUM_ATMOSPHERE
        m.priority = len(self.media) if priority is None else priority
        m.phase_kind = abi.PHASE_RAYLEIGH
        m.planet_center[:] = [float(x) for x in center]
        m.planet_radius = radius
        m.atmosphere_height = height
        m.rayleigh_scale_height = rayleigh_h
        m.mie_scale_height = mie_h
        m.mie_g = mie_g
        m.rayleigh_sigma_s = _spec(rayleigh_sigma_s)
        m.mie_sigma_s = _spec(mie_sigma_s)
        m.mie_sigma_a = _spec(mie_sigma_a)
        self.media.append(m)
        return len(self.media) - 1

    def add_tabulated_atmosphere(self, center, radius, h0, dh, sigma_t, sigma_s_molecules, sigma_s_aerosols, mie_g=0.7,
                                 priority=None):
        """Atmosphere with coefficient tables over altitude (rows h0 + i * dh, 36 bins each)."""
        tabs = [np.ascontiguousarray(np.asarray(t, dtype=np.float32).reshape(-1, NBINS))
                for t in (sigma_t, sigma_s_molecules, sigma_s_aerosols)]
        if not (tabs[0].shape == tabs[1].shape == tabs[2].shape):
            raise ValueError("the three tables need the same number of altitudes")
        self._keep.extend(tabs)
        m = abi.Medium()
        m.kind = abi.MEDIUM_ATMOSPHERE_TABULATED
        m.priority = len(self.media) if priority is None else priority
        m.phase_kind = abi.PHASE_RAYLEIGH
        m.planet_center[:] = [float(x) for x in center]
        m.planet_radius = radius
        m.mie_g = mie_g
        m.table_size = len(tabs[0])
        m.table_h0, m.table_dh = h0, dh
        fp = C.POINTER(C.c_float)
        m.table_sigma_t, m.table_sigma_s_molecules, m.table_sigma_s_aerosols = (t.ctypes.data_as(fp) for t in tabs)
        self.media.append(m)
        return len(self.media) - 1

    def add_area_emitter(self, spectrum, kind=abi.AREA_UNIFORM):
        e = abi.AreaEmitter()
        e.kind = kind
        e.spectrum = _spec(spectrum if np.ndim(spectrum) else flat(spectrum))
        self.emitters.append(e)
        return len(self.emitters) - 1

    def add_material(self, shader=-1, medium=-1, outer_medium=-1, emitter=-1):
        m = abi.Material(shader, medium, outer_medium, emitter)
        self.materials.append(m)
        return len(self.materials) - 1

    # ---- lights -----------------------------------------------------------
    def add_sun(self, direction, total_power, opening_angle_deg=0.25):
        e = abi.EnvLight()
        e.kind = abi.ENV_SUN
        e.dir[:] = [float(x) for x in direction]
        e.total_power = total_power
        e.opening_angle_deg = opening_angle_deg
        self.envs.append(e)

    def add_dome(self, up, spectrum):
        e = abi.EnvLight()
        e.kind = abi.ENV_DOME
        e.dir[:] = [float(x) for x in up]
        e.spectrum = _spec(spectrum if np.ndim(spectrum) else flat(spectrum))
        self.envs.append(e)

    def add_directional(self, direction, spectrum):
        e = abi.EnvLight()
        e.kind = abi.ENV_DIRECTIONAL
        e.dir[:] = [float(x) for x in direction]
        e.spectrum = _spec(spectrum if np.ndim(spectrum) else flat(spectrum))
        self.envs.append(e)

    def add_env_map(self, up, texture_rgb):
        """texture_rgb: (h, w, 3) float32; row y covers polar angles
        [pi*(1-(y+1)/h), pi*(1-y/h)] from `up` (src/texture.hxx PixelToUvBounds)."""
        tex = np.ascontiguousarray(texture_rgb, dtype=np.float32)
        if tex.ndim != 3 or tex.shape[2] != 3:
            raise ValueError("texture must be (h, w, 3)")
        self._env_textures.append(tex)
        e = abi.EnvLight()
        e.kind = abi.ENV_MAP
        e.dir[:] = [float(x) for x in up]
        e.tex_height, e.tex_width = tex.shape[0], tex.shape[1]
        e.texture_rgb = tex.ctypes.data_as(C.POINTER(C.c_float))
        self.envs.append(e)

    def add_point_light(self, pos, spectrum):
        p = abi.PointLight()
        p.pos[:] = [float(x) for x in pos]
        p.spectrum = _spec(spectrum if np.ndim(spectrum) else flat(spectrum))
        self.points.append(p)

    # ---- geometry ---------------------------------------------------------
    def add_mesh(self, vertices, indices, material, normals=None):
        v = np.asarray(vertices, dtype=np.float32).reshape(-1, 3)
        i = np.asarray(indices, dtype=np.uint32).reshape(-1, 3)
        mat = np.broadcast_to(np.asarray(material, dtype=np.int32), (len(i),))
        if normals is not None:
            self._has_normals = True
            n = np.asarray(normals, dtype=np.float32).reshape(-1, 3)
        else:
            n = None
        self.vertices.append(v)
        self.normals.append(n)
        self.indices.append(i)
        self.tri_material.append(np.array(mat))
        self._nvert += len(v)

    def add_quad(self, a, b, c, d, material):
        """Quad a,b,c,d (counter-clockwise seen from the side the normal cross(b-a, c-a) points to)."""
        self.add_mesh([a, b, c, d], [[0, 1, 2], [0, 2, 3]], material)

    def add_box(self, lo, hi, material):
        """Axis-aligned box, 12 outward-facing triangles."""
        x0, y0, z0 = lo
        x1, y1, z1 = hi
        v = [[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
             [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]]
        f = [[0, 3, 2], [0, 2, 1], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
             [3, 7, 6], [3, 6, 2], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]]
        self.add_mesh(v, f, material)

    def add_sphere(self, center, radius, material):
        self.spheres.append([float(center[0]), float(center[1]), float(center[2]), float(radius)])
        self.sphere_material.append(int(material))

    def set_camera(self, pos, look_at, up, fov_deg, kind=abi.CAMERA_PINHOLE):
        cam = abi.Camera()
        cam.kind = kind
        cam.pos[:] = [float(x) for x in pos]
        cam.dir[:] = [float(a) - float(b) for a, b in zip(look_at, pos)]
        cam.up[:] = [float(x) for x in up]
        cam.fov_deg = fov_deg
        self.camera = cam

    @property
    def num_triangles(self):
        return int(sum(len(i) for i in self.indices))

    # ---- ABI --------------------------------------------------------------
    def to_abi(self):
        """Returns a SceneDesc; the arrays it points to stay owned by this object."""
        P = C.POINTER
        d = abi.SceneDesc()
        if self.vertices:
            vs, ns, ids, off = [], [], [], 0
            for v, n, i in zip(self.vertices, self.normals, self.indices):
                if self._has_normals and n is None:
                    # Flat shading inside a scene with smooth meshes: one vertex per corner.
                    tri = v[i]                                       # (t, 3, 3)
                    fn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
                    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
                    v = tri.reshape(-1, 3)
                    n = np.repeat(fn, 3, axis=0).astype(np.float32)
                    i = np.arange(len(v), dtype=np.uint32).reshape(-1, 3)
                vs.append(v)
                ns.append(n)
                ids.append(i + off)
                off += len(v)
            self._v = np.ascontiguousarray(np.concatenate(vs), dtype=np.float32)
            self._i = np.ascontiguousarray(np.concatenate(ids), dtype=np.uint32)
            self._m = np.ascontiguousarray(np.concatenate(self.tri_material), dtype=np.int32)
            if self._has_normals:
                self._n = np.ascontiguousarray(np.concatenate(ns), dtype=np.float32)
            d.num_vertices = len(self._v)
            d.vertices = self._v.ctypes.data_as(P(C.c_float))
            d.normals = self._n.ctypes.data_as(P(C.c_float)) if self._has_normals else None
            d.num_triangles = len(self._i)
            d.indices = self._i.ctypes.data_as(P(C.c_uint32))
            d.tri_material = self._m.ctypes.data_as(P(C.c_int32))
        if self.spheres:
            self._s = np.ascontiguousarray(self.spheres, dtype=np.float32)
            self._sm = np.ascontiguousarray(self.sphere_material, dtype=np.int32)
            d.num_spheres = len(self._s)
            d.spheres = self._s.ctypes.data_as(P(C.c_float))
            d.sphere_material = self._sm.ctypes.data_as(P(C.c_int32))

        def arr(T, items):
            if not items:
                return 0, None
            a = (T * len(items))(*items)
            self._keep.append(a)
            return len(items), C.cast(a, P(T))
        d.num_shaders, d.shaders = arr(abi.Shader, self.shaders)
        d.num_media, d.media = arr(abi.Medium, self.media)
        d.empty_space_medium = self.empty_space_medium
        d.num_materials, d.materials = arr(abi.Material, self.materials)
        d.num_area_emitters, d.area_emitters = arr(abi.AreaEmitter, self.emitters)
        d.num_env_lights, d.env_lights = arr(abi.EnvLight, self.envs)
        d.num_point_lights, d.point_lights = arr(abi.PointLight, self.points)
        d.camera = self.camera
        self._desc = d
        return d
