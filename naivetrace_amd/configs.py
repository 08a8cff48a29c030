"""Procedural scenes of the benchmark configurations (seeded, no asset files).

cornell_box()        Plain PT + NEE, Cornell-style box: 5 walls (10 tris) + 2 axis-aligned
                     boxes (24 tris), 0.25 x 0.25 area light with a flat spectrum, white 0.8,
                     red/green as spectral steps at 600/550 nm, pinhole 256^2, 64 spp, depth 8.
terrain_instances()  Plain PT at scale: displaced 512x512 heightfield (524,288 tris) and 200
                     geodesic icospheres of frequency 11 (484,000 tris) = 1,008,288 tris; GGX
                     metal / glass / mirror / Lambert materials; procedural 1024x512 HDR sky with
                     a sun disc; 1920x1080, 256 spp, depth 16.
medium_sphere()      Glass sphere (eta 1.5) filled with a homogeneous chromatic medium, HG g=0.7.
atmosphere()         Planet R=6371 km, Rayleigh (H=8 km) + Mie (H=1.2 km, g=0.76), sun at 5 deg.
cornell_hidden_light() C1 geometry with a baffle hiding the light from direct view (guiding).
"""
import numpy as np

from . import abi
from .scene import Scene, flat, NBINS, WAVELENGTHS


def step_spectrum(lo_value, hi_value, threshold_nm, above=True):
    """Bin-wise step: hi_value where the bin centre is >= threshold (above=True) or <= threshold."""
    sel = WAVELENGTHS >= threshold_nm if above else WAVELENGTHS <= threshold_nm
    return np.where(sel, hi_value, lo_value).astype(np.float64)


CORNELL_LIGHT_POWER = 400.0  # flat area-light spectrum value (L = 400 / (2 pi) per bin)


def cornell_box(width=256, height=256, light_power=CORNELL_LIGHT_POWER, baffle=False):
    s = Scene()
    white = s.add_shader(abi.SHADER_DIFFUSE, reflectance=flat(0.8))
    red = s.add_shader(abi.SHADER_DIFFUSE, reflectance=step_spectrum(0.05, 0.8, 600.0, above=True))
    green = s.add_shader(abi.SHADER_DIFFUSE, reflectance=step_spectrum(0.05, 0.8, 550.0, above=False))
    black = s.add_shader(abi.SHADER_DIFFUSE, reflectance=flat(0.0))
    m_white = s.add_material(shader=white)
    m_red = s.add_material(shader=red)
    m_green = s.add_material(shader=green)
    em = s.add_area_emitter(flat(light_power))
    m_light = s.add_material(shader=black, emitter=em)
    # Room: x in [-1,1], y in [0,2], z in [-1,1]; open at z=+1.  Normals point inwards.
    s.add_quad([-1, 0, 1], [1, 0, 1], [1, 0, -1], [-1, 0, -1], m_white)     # floor (+y)
    s.add_quad([-1, 2, -1], [1, 2, -1], [1, 2, 1], [-1, 2, 1], m_white)     # ceiling (-y)
    s.add_quad([-1, 0, -1], [1, 0, -1], [1, 2, -1], [-1, 2, -1], m_white)   # back (+z)
    s.add_quad([-1, 0, 1], [-1, 0, -1], [-1, 2, -1], [-1, 2, 1], m_red)     # left (+x)
    s.add_quad([1, 0, -1], [1, 0, 1], [1, 2, 1], [1, 2, -1], m_green)       # right (-x)
    s.add_box([0.1, 0.0, -0.1], [0.7, 0.6, 0.5], m_white)                   # short box
    s.add_box([-0.7, 0.0, -0.7], [-0.1, 1.2, -0.1], m_white)                # tall box
    # 0.25 x 0.25 light just below the ceiling, emitting downwards (normal -y).
    h = 0.125
    y = 1.995
    s.add_quad([-h, y, -h], [h, y, -h], [h, y, h], [-h, y, h], m_light)
    if baffle:
        # A black-backed white panel under the light hides it from the camera: indirect light dominates.
        s.add_quad([-0.35, 1.8, 0.35], [0.35, 1.8, 0.35], [0.35, 1.8, -0.35], [-0.35, 1.8, -0.35], m_white)
        s.add_quad([-0.35, 1.8, -0.35], [0.35, 1.8, -0.35], [0.35, 1.8, 0.35], [-0.35, 1.8, 0.35], m_white)
    s.set_camera([0.0, 1.0, 3.4], [0.0, 1.0, 0.0], [0.0, 1.0, 0.0], 40.0)
    s.width, s.height = width, height
    return s


def cornell_hidden_light(width=256, height=256):
    return cornell_box(width, height, baffle=True)


# ---------------------------------------------------------------------------
# Config 2: terrain + icosphere instances + HDR sky
# ---------------------------------------------------------------------------
def _value_noise(n, rng, octaves=5):
    h = np.zeros((n, n))
    amp = 1.0
    for o in range(octaves):
        k = 2 ** (o + 2) + 1
        g = rng.random((k, k))
        xs = np.linspace(0, k - 1, n)
        i0 = np.clip(np.floor(xs).astype(int), 0, k - 2)
        f = xs - i0
        f = f * f * (3 - 2 * f)
        a = g[i0][:, i0] * (1 - f)[None, :] + g[i0][:, i0 + 1] * f[None, :]
        b = g[i0 + 1][:, i0] * (1 - f)[None, :] + g[i0 + 1][:, i0 + 1] * f[None, :]
        h += amp * (a * (1 - f)[:, None] + b * f[:, None])
        amp *= 0.5
    return h


def heightfield(cells=512, size=100.0, height=12.0, seed=1):
    rng = np.random.default_rng(seed)
    n = cells + 1
    h = _value_noise(n, rng) * height
    xs = np.linspace(-size / 2, size / 2, n)
    X, Z = np.meshgrid(xs, xs)
    v = np.stack([X, h, Z], axis=-1).reshape(-1, 3)
    i = np.arange(cells)
    a = (i[:, None] * n + i[None, :]).reshape(-1)
    b, c, d = a + 1, a + n, a + n + 1
    tris = np.concatenate([np.stack([a, c, b], 1), np.stack([b, c, d], 1)])
    return v.astype(np.float32), tris.astype(np.uint32), h


def icosphere(freq):
    """Geodesic sphere: every icosahedron face subdivided into freq^2 triangles (20*freq^2 total)."""
    t = (1 + 5 ** 0.5) / 2
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    F = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    verts = {}
    out_v, out_f = [], []

    def vid(p):
        key = tuple(np.round(p / np.linalg.norm(p), 9))
        if key not in verts:
            verts[key] = len(out_v)
            out_v.append(p / np.linalg.norm(p))
        return verts[key]
    for a, b, c in F:
        A, B, Cc = V[a], V[b], V[c]
        grid = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                p = A + (B - A) * i / freq + (Cc - A) * j / freq
                grid[(i, j)] = vid(p)
        for i in range(freq):
            for j in range(freq - i):
                out_f.append([grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]])
                if j < freq - i - 1:
                    out_f.append([grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]])
    return np.array(out_v), np.array(out_f, dtype=np.uint32)


def sky_texture(w=1024, h=512, sun_dir=(0.3, 0.6, 0.2)):
    """Procedural HDR sky: blue-to-white gradient over the upper hemisphere, dim ground, sun disc.

    Row y covers polar angles [pi*(1-(y+1)/h), pi*(1-y/h)] measured from the up axis (+y)."""
    ys = (np.arange(h) + 0.5) / h
    theta = np.pi * (1 - ys)                              # polar angle from up
    xs = (np.arange(w) + 0.5) / w
    phi = 2 * np.pi * xs
    T, Pp = np.meshgrid(theta, phi, indexing="ij")
    up = np.clip(np.cos(T), 0, 1)
    zenith = np.array([0.25, 0.45, 1.0])
    horizon = np.array([0.9, 0.95, 1.0])
    sky = horizon[None, None] * (1 - up[..., None]) + zenith[None, None] * up[..., None]
    ground = np.array([0.12, 0.1, 0.08])
    tex = np.where((np.cos(T) > 0)[..., None], sky * 1.2, ground[None, None])
    # Sun disc in the texture's own frame: local z = up, x/y from the azimuth (KartesianToSpherical).
    sd = np.asarray(sun_dir, dtype=np.float64)
    sd /= np.linalg.norm(sd)
    st, sp = np.arccos(sd[2]), np.arctan2(sd[1], sd[0]) % (2 * np.pi)
    cosang = np.sin(T) * np.sin(st) * np.cos(Pp - sp) + np.cos(T) * np.cos(st)
    tex = tex + np.where(cosang > np.cos(np.radians(1.5)), 800.0, 0.0)[..., None] * np.array([1.0, 0.95, 0.85])
    return tex.astype(np.float32)


def terrain_instances(width=1920, height=1080, seed=1, cells=512, n_spheres=200, freq=11, sky=(1024, 512)):
    rng = np.random.default_rng(seed)
    s = Scene()
    size = 100.0
    v, f, hmap = heightfield(cells, size, 12.0, seed)
    # Materials drawn uniformly from GGX metal (alpha U(0.05,0.5)), glass 1.5, mirror, Lambert (albedo U(0.1,0.9)).
    def random_material():
        k = rng.integers(4)
        if k == 0:
            sh = s.add_shader(abi.SHADER_GGX_METAL, reflectance=flat(rng.uniform(0.5, 0.95)), alpha=rng.uniform(0.05, 0.5))
        elif k == 1:
            sh = s.add_shader(abi.SHADER_DIELECTRIC, ior=1.5)
        elif k == 2:
            sh = s.add_shader(abi.SHADER_MIRROR, reflectance=flat(0.9))
        else:
            sh = s.add_shader(abi.SHADER_DIFFUSE, reflectance=flat(rng.uniform(0.1, 0.9)))
        return s.add_material(shader=sh)
    ground = s.add_material(shader=s.add_shader(abi.SHADER_DIFFUSE, reflectance=flat(rng.uniform(0.1, 0.9))))
    s.add_mesh(v, f, ground)
    sv, sf = icosphere(freq)
    n = cells + 1
    for _ in range(n_spheres):
        r = rng.uniform(0.5, 3.0)
        x, z = rng.uniform(-size * 0.4, size * 0.4, 2)
        gi = int((x + size / 2) / size * cells)
        gj = int((z + size / 2) / size * cells)
        y = hmap[min(gj, n - 1), min(gi, n - 1)] + r * rng.uniform(0.2, 1.0)
        s.add_mesh(sv * r + np.array([x, y, z]), sf, random_material(), normals=sv)
    s.add_env_map([0, 1, 0], sky_texture(*sky))
    s.set_camera([0.0, 30.0, 70.0], [0.0, 5.0, 0.0], [0.0, 1.0, 0.0], 60.0)
    s.width, s.height = width, height
    return s


# ---------------------------------------------------------------------------
# Config 3: participating media
# ---------------------------------------------------------------------------
def medium_sphere(width=256, height=256, seed=1):
    rng = np.random.default_rng(seed)
    s = Scene()
    sig_s = rng.uniform(0.5, 2.0, NBINS)
    sig_a = rng.uniform(0.01, 0.2, NBINS)
    med = s.add_homogeneous_medium(sig_s, sig_a, phase=abi.PHASE_HENYEY_GREENSTEIN, g=0.7)
    glass = s.add_shader(abi.SHADER_DIELECTRIC, ior=1.5)
    m_sphere = s.add_material(shader=glass, medium=med)
    white = s.add_material(shader=s.add_shader(abi.SHADER_DIFFUSE, reflectance=flat(0.8)))
    s.add_quad([-4, 0, 4], [4, 0, 4], [4, 0, -4], [-4, 0, -4], white)
    s.add_sphere([0.0, 1.0, 0.0], 1.0, m_sphere)
    em = s.add_area_emitter(flat(60.0))
    m_light = s.add_material(shader=s.add_shader(abi.SHADER_DIFFUSE, reflectance=flat(0.0)), emitter=em)
    s.add_quad([-0.5, 4, -0.5], [0.5, 4, -0.5], [0.5, 4, 0.5], [-0.5, 4, 0.5], m_light)
    s.add_dome([0, 1, 0], flat(0.2))
    s.set_camera([0.0, 1.5, 5.0], [0.0, 0.9, 0.0], [0.0, 1.0, 0.0], 45.0)
    s.width, s.height = width, height
    return s


def rayleigh_sigma_s_per_km():
    """Rayleigh sea-level scattering per km (src/atmosphere.cxx:157-176 constants)."""
    Na, rho, m, ior = 6.022e23, 1.225, 29e-3, 1.00028
    N = rho * Na / m
    pref = 8. * np.pi ** 3 / 3. * (ior ** 2 - 1.) ** 2 / N * 1.e27 * 1.e12
    return pref / WAVELENGTHS ** 4


def atmosphere(width=1024, height=512, sun_elevation_deg=5.0, mie_g=0.76):
    """Planet of radius 6371 km (scene unit = km) with an exponential atmosphere, viewed from the
    ground through an upward fisheye; sun at `sun_elevation_deg`."""
    s = Scene()
    R = 6371.0
    center = [0.0, -R, 0.0]
    med = s.add_atmosphere(center, R, 100.0, 8.0, 1.2, mie_g, rayleigh_sigma_s_per_km(),
                           flat(20.e-3), flat(2.22e-3))
    # The atmosphere boundary is a sphere with the medium inside; the ground is a diffuse sphere.
    shell = s.add_material(medium=med)
    s.add_sphere(center, R + 100.0, shell)
    ground = s.add_material(shader=s.add_shader(abi.SHADER_DIFFUSE, reflectance=flat(0.3)))
    s.add_sphere(center, R, ground)
    el = np.radians(sun_elevation_deg)
    s.add_sun([np.cos(el), np.sin(el), 0.0], total_power=20.0, opening_angle_deg=0.5)
    s.set_camera([0.0, 0.002, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], 180.0, kind=abi.CAMERA_FISHEYE)
    s.width, s.height = width, height
    return s


# ---------------------------------------------------------------------------
# Media scenes with closed-form answers (transmittance through a pure absorber)
# ---------------------------------------------------------------------------
def absorber_slab(width=16, height=16, seed=3, thickness=0.5, dome=1.0, fov_deg=2.0):
    """Camera looking along -z through an invisible slab |z| < thickness/2 filled with a purely
    absorbing chromatic medium, at a uniform dome light behind it.  Every camera path escapes
    or is absorbed, so each pixel's expected radiance is sum_b M[:, b] * dome * exp(-sigma_a[b] * d)
    with d the chord of the pixel's ray (d = thickness / |dir_z|)."""
    rng = np.random.default_rng(seed)
    s = Scene()
    sig_a = rng.uniform(0.2, 2.0, NBINS)
    med = s.add_homogeneous_medium(np.zeros(NBINS), sig_a)
    boundary = s.add_material(medium=med)
    h = 0.5 * thickness
    s.add_box([-50.0, -50.0, -h], [50.0, 50.0, h], boundary)
    s.add_dome([0.0, 0.0, -1.0], flat(dome))
    s.set_camera([0.0, 0.0, 10.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0], fov_deg)
    s.width, s.height = width, height
    s.absorber_sigma_a = sig_a
    return s


def tabulated_absorber_shell(width=8, height=8, seed=6, radius=10.0, fov_deg=20.0):
    """A tabulated atmosphere (TabulatedConstituents) with altitude-constant, purely absorbing
    chromatic tables inside an invisible sphere of `radius` around the camera, which looks up at a
    uniform dome.  Every ray leaves through the sphere after `radius`, so each pixel's expected
    radiance is sum_b M[:, b] exp(-sigma_a[b] * radius)."""
    rng = np.random.default_rng(seed)
    s = Scene()
    sig_a = rng.uniform(0.02, 0.2, NBINS)
    n = 21
    st = np.tile(sig_a, (n, 1))
    zero = np.zeros((n, NBINS))
    # Planet radius 5: altitudes inside the sphere span [-5, 5], within the table's [-10, 10].
    med = s.add_tabulated_atmosphere([0.0, 0.0, 0.0], 5.0, -10.0, 1.0, st, zero, zero)
    s.add_sphere([0.0, 0.0, 0.0], radius, s.add_material(medium=med))
    s.add_dome([0.0, 1.0, 0.0], flat(1.0))
    s.set_camera([0.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], fov_deg)
    s.width, s.height = width, height
    s.absorber_sigma_a = sig_a
    s.shell_radius = radius
    return s


def point_lit_plane_behind_absorber(width=16, height=16, seed=4, thickness=0.4, power=50.0, albedo=0.5):
    """Lambertian plane z = 0 lit by a point light at (0, 0, 2) through a purely absorbing slab
    0.8 < z < 0.8 + thickness (an invisible volume boundary).  The camera sits below the slab at
    z = 0.5 looking straight down, so its rays never cross the slab; only the NEE shadow ray does.
    Direct light at the plane's centre:  albedo/pi * power/(4 pi r^2) * exp(-sigma_a * thickness)."""
    rng = np.random.default_rng(seed)
    s = Scene()
    sig_a = rng.uniform(0.5, 3.0, NBINS)
    med = s.add_homogeneous_medium(np.zeros(NBINS), sig_a)
    boundary = s.add_material(medium=med)
    s.add_box([-20.0, -20.0, 0.8], [20.0, 20.0, 0.8 + thickness], boundary)
    diffuse = s.add_material(shader=s.add_shader(abi.SHADER_DIFFUSE, reflectance=flat(albedo)))
    s.add_quad([-20, -20, 0], [20, -20, 0], [20, 20, 0], [-20, 20, 0], diffuse)
    s.add_point_light([0.0, 0.0, 2.0], flat(power))
    s.set_camera([0.0, 0.0, 0.5], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0], 1.0)
    s.width, s.height = width, height
    s.absorber_sigma_a = sig_a
    return s


def fog_box(width=32, height=32, seed=5):
    """Camera inside an invisible box of scattering fog (HG g=0.5) over a diffuse floor, with a
    glass sphere holding a denser, higher-priority medium, an area light and a point light.
    Exercises the medium tracker (initial medium at the camera, priorities, nested media),
    volume boundaries, phase-function NEE and scattering."""
    rng = np.random.default_rng(seed)
    s = Scene()
    fog = s.add_homogeneous_medium(rng.uniform(0.2, 0.5, NBINS), rng.uniform(0.02, 0.1, NBINS),
                                   phase=abi.PHASE_HENYEY_GREENSTEIN, g=0.5, priority=0)
    ink = s.add_homogeneous_medium(rng.uniform(1.0, 3.0, NBINS), rng.uniform(0.1, 0.5, NBINS),
                                   phase=abi.PHASE_UNIFORM, priority=1)
    s.add_box([-3.0, -0.5, -3.0], [3.0, 3.0, 3.0], s.add_material(medium=fog))
    floor = s.add_material(shader=s.add_shader(abi.SHADER_DIFFUSE, reflectance=flat(0.7)))
    s.add_quad([-2.5, 0, 2.5], [2.5, 0, 2.5], [2.5, 0, -2.5], [-2.5, 0, -2.5], floor)
    glass = s.add_shader(abi.SHADER_DIELECTRIC, ior=1.33)
    s.add_sphere([0.0, 0.8, 0.0], 0.6, s.add_material(shader=glass, medium=ink))
    em = s.add_area_emitter(flat(30.0))
    m_light = s.add_material(shader=s.add_shader(abi.SHADER_DIFFUSE, reflectance=flat(0.0)), emitter=em)
    s.add_quad([-0.5, 2.7, 0.5], [0.5, 2.7, 0.5], [0.5, 2.7, -0.5], [-0.5, 2.7, -0.5], m_light)
    s.add_point_light([1.2, 2.0, 0.8], flat(40.0))
    s.set_camera([0.0, 1.2, 2.5], [0.0, 0.8, 0.0], [0.0, 1.0, 0.0], 50.0)
    s.width, s.height = width, height
    return s
