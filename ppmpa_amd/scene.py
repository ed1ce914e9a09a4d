"""Scene description: materials, shapes, lights and camera in the reference's
terms, and the synthetic scenes of the benchmark configurations.

Reference types: Material (material.rs:11-16), Surface (surface.rs:16-66),
Shape (geometry.rs:66-115), Object (object.rs:10-13), Light (light.rs:16-39),
Camera (camera.rs:23-50, read_camera camera.rs:108-206), M_AIR
(scene.rs:13-18).  A Scene converts to the ppm_scene_desc of include/ppm.h.
"""
import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

from . import _abi as A

Vec = Tuple[float, float, float]


def _v(x) -> Vec:
    return (float(x[0]), float(x[1]), float(x[2]))


def color_normalize(c: Sequence[float]) -> Vec:
    """Color::normalize (physics.rs:61-71): clip negatives, scale to sum 1."""
    r, g, b = (max(0.0, float(x)) for x in c)
    mag = r + g + b
    if mag == 0.0:
        return (1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0)
    return (r / mag, g / mag, b / mag)


@dataclass
class Material:
    """Material + surface parameters (material.rs:11-16, surface.rs:16-41).

    color1/color2 are reflectance/specular_refl for Simple and
    albedo_diff/albedo_spec for TS; diffuseness is scatterness for TS.
    """
    surface: int = A.SURF_SIMPLE
    color1: Vec = (0.0, 0.0, 0.0)
    color2: Vec = (0.0, 0.0, 0.0)
    diffuseness: float = 0.0
    metalness: float = 0.0
    roughness: float = 0.0
    emittance: Vec = (0.0, 0.0, 0.0)
    transmittance: Vec = (0.0, 0.0, 0.0)
    ior: Vec = (0.0, 0.0, 0.0)


def simple(reflectance, specular_refl, diffuseness, metalness, roughness, *, emittance=(0, 0, 0),
           ior=(0, 0, 0), transmittance=(0, 0, 0)) -> Material:
    """Surface::new_simple (surface.rs:45-54) inside a Material."""
    return Material(A.SURF_SIMPLE, _v(reflectance), _v(specular_refl), float(diffuseness), float(metalness),
                    float(roughness), _v(emittance), _v(transmittance), _v(ior))


def ts(albedo_diff, albedo_spec, scatterness, metalness, roughness, *, emittance=(0, 0, 0), ior=(0, 0, 0),
       transmittance=(0, 0, 0)) -> Material:
    """Surface::new_ts (surface.rs:56-66) inside a Material."""
    return Material(A.SURF_TS, _v(albedo_diff), _v(albedo_spec), float(scatterness), float(metalness),
                    float(roughness), _v(emittance), _v(transmittance), _v(ior))


# Encodings of the benchmark materials in the reference's Simple model.
def lambertian(rho) -> Material:
    if not isinstance(rho, (tuple, list, np.ndarray)):
        rho = (rho, rho, rho)
    return simple(rho, (0, 0, 0), 1.0, 0.0, 0.0)


def mirror() -> Material:
    # Schlick with F0 = 1 always reflects (physics.rs:316-318)
    return simple((0, 0, 0), (1, 1, 1), 0.0, 1.0, 0.0)


def glass(ior=1.5) -> Material:
    f0 = ((ior - 1.0) / (ior + 1.0)) ** 2  # 0.04 at n = 1.5
    return simple((0, 0, 0), (f0, f0, f0), 0.0, 0.0, 0.0, ior=(ior, ior, ior))


@dataclass
class Shape:
    kind: int
    material: int
    pos: Vec = (0.0, 0.0, 0.0)
    nvec: Vec = (0.0, 0.0, 0.0)
    dir1: Vec = (0.0, 0.0, 0.0)
    dir2: Vec = (0.0, 0.0, 0.0)
    dist: float = 0.0
    radius: float = 0.0


def plain(nvec, dist, material) -> Shape:
    """Shape::Plain: points x with nvec.x + dist = 0 (geometry.rs:70-73,170-177)."""
    return Shape(A.SHAPE_PLAIN, material, nvec=_v(nvec), dist=float(dist))


def sphere(center, radius, material) -> Shape:
    return Shape(A.SHAPE_SPHERE, material, pos=_v(center), radius=float(radius))


def _poly(kind, p0, p1, p2, material, nvec=None) -> Shape:
    p0, p1, p2 = np.asarray(p0, float), np.asarray(p1, float), np.asarray(p2, float)
    d1, d2 = p1 - p0, p2 - p0
    if nvec is None:
        n = np.cross(d1, d2)
        nn = np.linalg.norm(n)
        if nn == 0.0:
            raise ValueError("degenerate polygon")
        nvec = n / nn
    return Shape(kind, material, pos=_v(p0), nvec=_v(nvec), dir1=_v(d1), dir2=_v(d2))


def polygon(p0, p1, p2, material) -> Shape:
    """Shape::new_polygon (geometry.rs:93-103)."""
    return _poly(A.SHAPE_POLYGON, p0, p1, p2, material)


def parallelogram(p0, p1, p2, material, nvec=None) -> Shape:
    """Shape::new_parallelogram (geometry.rs:105-115)."""
    return _poly(A.SHAPE_PARALLELOGRAM, p0, p1, p2, material, nvec)


@dataclass
class Light:
    kind: int
    color: Vec
    flux: float
    pos: Vec = (0.0, 0.0, 0.0)
    nvec: Vec = (0.0, 0.0, 0.0)
    dir1: Vec = (0.0, 0.0, 0.0)
    dir2: Vec = (0.0, 0.0, 0.0)
    dir: Vec = (0.0, 0.0, 0.0)


def parallelogram_light(color, flux, pos, nvec, dir1, dir2) -> Light:
    return Light(A.LIGHT_PARALLELOGRAM, color_normalize(color), float(flux), _v(pos), _v(nvec), _v(dir1), _v(dir2))


def point_light(color, flux, pos) -> Light:
    return Light(A.LIGHT_POINT, color_normalize(color), float(flux), _v(pos))


def sun_light(color, flux, pos, nvec, dir1, dir2, direction) -> Light:
    return Light(A.LIGHT_SUN, color_normalize(color), float(flux), _v(pos), _v(nvec), _v(dir1), _v(dir2),
                 _v(direction))


@dataclass
class Camera:
    xreso: int
    yreso: int
    eye: Vec
    target: Vec
    upper: Vec = (0.0, 1.0, 0.0)
    focus: float = 7.0
    focal_len: float = 0.05    # metres (camera.rs:142 divides mm by 1000)
    f_number: float = 4.0
    progressive: bool = True
    antialias: bool = True
    blur: bool = True
    pfilter: int = A.FILTER_NONE

    def screen(self):
        """Derived vectors of read_camera (camera.rs:151-164)."""
        eye = np.asarray(self.eye, float)
        ez = np.asarray(self.target, float) - eye
        ez = ez / np.linalg.norm(ez)
        ex = np.cross(np.asarray(self.upper, float), ez)
        ex = ex / np.linalg.norm(ex)
        ey = np.cross(ex, ez)
        ey = ey / np.linalg.norm(ey)
        step = (self.focus * 0.035 / self.focal_len) / self.xreso
        esx, esy = step * ex, step * ey
        ea = self.focal_len / self.f_number
        lx, ly = float(self.xreso // 2), float(self.yreso // 2)
        origin = self.focus * ez - (lx - 0.5) * esx - (ly - 0.5) * esy
        return dict(eye_pos=eye, origin=origin, esx=esx, esy=esy, eex=ea * ex, eey=ea * ey)

    def desc(self) -> A.CameraDesc:
        s = self.screen()
        d = A.CameraDesc()
        d.xreso, d.yreso = int(self.xreso), int(self.yreso)
        d.progressive, d.antialias, d.blur = int(self.progressive), int(self.antialias), int(self.blur)
        d.pfilter = int(self.pfilter)
        for k, v in s.items():
            setattr(d, k, A.D3(*[float(x) for x in v]))
        return d


def pinhole(xreso, yreso, eye, target, fov_deg=60.0, upper=(0.0, 1.0, 0.0), antialias=True) -> Camera:
    """A pinhole camera with horizontal field of view fov_deg: the reference's
    camera with blur off and focus*0.035/focal_len = 2 tan(fov/2)."""
    focal_len = 0.035 / (2.0 * math.tan(math.radians(fov_deg) / 2.0))
    return Camera(xreso, yreso, _v(eye), _v(target), _v(upper), focus=1.0, focal_len=focal_len, f_number=4.0,
                  progressive=True, antialias=antialias, blur=False)


def read_camera() -> Camera:
    """read_camera's built-in defaults (camera.rs:109-128); the file is ignored
    by the reference as well."""
    return Camera(256, 256, (1.0, 2.0, -4.5), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), focus=7.0, focal_len=0.05,
                  f_number=4.0, progressive=True, antialias=True, blur=True)


# numpy mirror of ppm_shape_desc (same field offsets, 120 bytes)
SHAPE_DTYPE = np.dtype([("kind", "<i4"), ("material", "<i4"), ("pos", "<f8", 3), ("nvec", "<f8", 3),
                        ("dir1", "<f8", 3), ("dir2", "<f8", 3), ("dist", "<f8"), ("radius", "<f8")], align=True)


@dataclass
class Mesh:
    """Triangles (n x 3 x 3 vertex coordinates) of one material: each one a
    Shape::Polygon (geometry.rs:93-103) with the unit normal (v1-v0)x(v2-v0)."""
    tris: np.ndarray
    material: int


@dataclass
class Scene:
    materials: List[Material] = field(default_factory=list)
    shapes: List[Shape] = field(default_factory=list)
    lights: List[Light] = field(default_factory=list)
    camera: Camera = None
    bounds_min: Vec = (0.0, 0.0, 0.0)
    bounds_max: Vec = (1.0, 1.0, 1.0)
    meshes: List[Mesh] = field(default_factory=list)  # their triangles follow `shapes` in scene order
    bvh: int = A.SCENE_BVH_AUTO

    def add_material(self, m: Material) -> int:
        self.materials.append(m)
        return len(self.materials) - 1

    @property
    def n_shapes(self) -> int:
        return len(self.shapes) + sum(len(m.tris) for m in self.meshes)

    def shape_array(self) -> np.ndarray:
        """All shapes in scene order as a SHAPE_DTYPE array."""
        out = np.zeros(self.n_shapes, SHAPE_DTYPE)
        for i, s in enumerate(self.shapes):
            out[i] = (int(s.kind), int(s.material), s.pos, s.nvec, s.dir1, s.dir2, float(s.dist), float(s.radius))
        k = len(self.shapes)
        for m in self.meshes:
            t = np.asarray(m.tris, np.float64).reshape(-1, 3, 3)
            d1, d2 = t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]
            n = np.cross(d1, d2)
            ln = np.linalg.norm(n, axis=1)
            if np.any(ln == 0.0):
                raise ValueError("degenerate triangle in mesh")
            sl = out[k:k + len(t)]
            sl["kind"], sl["material"] = A.SHAPE_POLYGON, m.material
            sl["pos"], sl["dir1"], sl["dir2"], sl["nvec"] = t[:, 0], d1, d2, n / ln[:, None]
            k += len(t)
        return out

    def desc(self):
        """ppm_scene_desc; the returned tuple keeps the arrays alive."""
        arr = self.shape_array()
        if len(arr) == 0:
            arr = np.zeros(1, SHAPE_DTYPE)
        assert arr.dtype.itemsize == C.sizeof(A.ShapeDesc)
        shapes = arr
        mats = (A.MaterialDesc * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            d = mats[i]
            d.emittance, d.transmittance, d.ior = A.D3(*m.emittance), A.D3(*m.transmittance), A.D3(*m.ior)
            d.surface = int(m.surface)
            d.color1, d.color2 = A.D3(*m.color1), A.D3(*m.color2)
            d.diffuseness, d.metalness, d.roughness = float(m.diffuseness), float(m.metalness), float(m.roughness)
        lights = (A.LightDesc * max(1, len(self.lights)))()
        for i, l in enumerate(self.lights):
            d = lights[i]
            d.kind = int(l.kind)
            d.color, d.flux = A.D3(*l.color), float(l.flux)
            d.pos, d.nvec, d.dir1, d.dir2, d.dir = (A.D3(*l.pos), A.D3(*l.nvec), A.D3(*l.dir1), A.D3(*l.dir2),
                                                    A.D3(*l.dir))
        sd = A.SceneDesc()
        sd.n_shapes, sd.n_materials, sd.n_lights = self.n_shapes, len(self.materials), len(self.lights)
        sd.flags = int(self.bvh)
        sd.shapes = shapes.ctypes.data_as(C.POINTER(A.ShapeDesc))
        sd.materials = C.cast(mats, C.POINTER(A.MaterialDesc))
        sd.lights = C.cast(lights, C.POINTER(A.LightDesc))
        sd.camera = self.camera.desc()
        sd.bounds_min, sd.bounds_max = A.D3(*self.bounds_min), A.D3(*self.bounds_max)
        return sd, (shapes, mats, lights)


# ---------------------------------------------------------------------------
# Synthetic scenes of the benchmark configurations.

BOX_BOUNDS = ((-0.01, -0.01, -0.01), (1.01, 1.01, 1.01))
TUNNEL_BOUNDS = ((-0.01, -0.01, -1.0), (1.01, 1.01, 1.01))


def cornell_box(sc: Scene, left=(0.75, 0.25, 0.25), right=(0.25, 0.75, 0.25), grey=0.75, flux=100.0,
                light_size=0.25, walls="quad"):
    """The 1x1x1 m box of five Lambertian walls (floor y=0, ceiling y=1, back
    z=1, left x=0, right x=1; open at z=0) with a light_size^2 ceiling area
    light of `flux` watts of white light at y = 0.999.

    walls="quad": each wall is a 1 x 1 m Parallelogram (geometry.rs:84-89,
    the reference's finite planar shape), so photons leaving through the
    open front escape as from a real box.  walls="plane": infinite Plain
    shapes (geometry.rs:70-73), which turn the open front into an endless
    tunnel that photons wander and are stored in.

    The light is a ParallelogramLight plus an emissive Parallelogram object at
    the same place: the reference's direct-light test needs the shadow ray to
    hit the light's own object (tracer.rs:281-286, scene.rs:434-441).  The
    object's material follows the reference's light material (scene.rs:226-238:
    Simple, specular 0.8, ior 0) with emittance 2*flux/area, so that its
    emitted radiance emittance/(2 pi) (tracer.rs:175) is the Lambertian
    flux/(pi area).
    """
    g = sc.add_material(lambertian(grey))
    l = sc.add_material(lambertian(left))
    r = sc.add_material(lambertian(right))
    if walls == "plane":
        sc.shapes += [
            plain((0, 1, 0), 0.0, g),    # floor
            plain((0, -1, 0), 1.0, g),   # ceiling
            plain((0, 0, -1), 1.0, g),   # back wall
            plain((1, 0, 0), 0.0, l),    # left wall
            plain((-1, 0, 0), 1.0, r),   # right wall
        ]
        sc.bounds_min, sc.bounds_max = TUNNEL_BOUNDS
    else:
        # corners ordered so that dir1 x dir2 points into the box
        sc.shapes += [
            parallelogram((0, 0, 0), (0, 0, 1), (1, 0, 0), g),  # floor
            parallelogram((0, 1, 0), (1, 1, 0), (0, 1, 1), g),  # ceiling
            parallelogram((0, 0, 1), (0, 1, 1), (1, 0, 1), g),  # back wall
            parallelogram((0, 0, 0), (0, 1, 0), (0, 0, 1), l),  # left wall
            parallelogram((1, 0, 0), (1, 0, 1), (1, 1, 0), r),  # right wall
        ]
        sc.bounds_min, sc.bounds_max = BOX_BOUNDS
    s = light_size
    x0 = 0.5 - s / 2
    p0 = (x0, 0.999, x0)
    area = s * s
    em = 2.0 * flux / area
    lm = sc.add_material(simple((0, 0, 0), (0.8, 0.8, 0.8), 0.0, 0.0, 0.0, emittance=(em, em, em)))
    sc.shapes.append(parallelogram(p0, (x0 + s, 0.999, x0), (x0, 0.999, x0 + s), lm, nvec=(0, -1, 0)))
    sc.lights.append(parallelogram_light((1, 1, 1), flux, p0, (0, -1, 0), (s, 0, 0), (0, 0, s)))
    return sc


def box_camera(xreso, yreso, antialias=True) -> Camera:
    # 60 degree horizontal FOV: the open face (width 1 at z = 0) fills the view
    # from z = -0.5/tan(30 deg).
    return pinhole(xreso, yreso, (0.5, 0.5, -0.87), (0.5, 0.5, 1.0), 60.0, antialias=antialias)


def two_spheres_scene(xreso=256, yreso=256, walls="quad") -> Scene:
    """Cornell box + a diffuse sphere (r 0.2, albedo 0.75) + a perfect-mirror
    sphere (r 0.2); pinhole camera, 60 deg FOV."""
    sc = cornell_box(Scene(), walls=walls)
    d = sc.add_material(lambertian(0.75))
    m = sc.add_material(mirror())
    sc.shapes += [sphere((0.3, 0.2, 0.65), 0.2, d), sphere((0.72, 0.2, 0.45), 0.2, m)]
    sc.camera = box_camera(xreso, yreso)
    return sc


def caustic_scene(xreso=1024, yreso=768, walls="quad") -> Scene:
    """The box of two_spheres_scene with a diffuse, a mirror and a glass (IOR 1.5) sphere
    (r 0.2 each); the glass sphere sits under the light to cast a caustic."""
    sc = cornell_box(Scene(), walls=walls)
    d = sc.add_material(lambertian(0.75))
    m = sc.add_material(mirror())
    gl = sc.add_material(glass(1.5))
    sc.shapes += [sphere((0.25, 0.2, 0.75), 0.2, d), sphere((0.75, 0.2, 0.75), 0.2, m),
                  sphere((0.5, 0.2, 0.3), 0.2, gl)]
    sc.camera = box_camera(xreso, yreso)
    return sc


def icosphere(subdiv: int, center=(0.0, 0.0, 0.0), radius=1.0, noise=0.0, seed=1):
    """Triangles of an icosahedron subdivided `subdiv` times (20 * 4^subdiv
    faces, outward winding) projected onto a sphere; each vertex is moved
    radially by a factor 1 + U(-noise, +noise) drawn from a Philox stream
    keyed by `seed`.  Returns (vertices (V x 3), faces (F x 3 int))."""
    g = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([(-1, g, 0), (1, g, 0), (-1, -g, 0), (1, -g, 0), (0, -1, g), (0, 1, g), (0, -1, -g), (0, 1, -g),
                  (g, 0, -1), (g, 0, 1), (-g, 0, -1), (-g, 0, 1)], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], np.int64)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        es = np.sort(e, axis=1)
        uniq, inv = np.unique(es, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = base + inv[:nf], base + inv[nf:2 * nf], base + inv[2 * nf:]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    if noise:
        rng = np.random.Generator(np.random.Philox(key=int(seed)))
        v = v * (1.0 + rng.uniform(-noise, noise, (len(v), 1)))
    return np.asarray(center, np.float64) + radius * v, f


def mesh_scene(xreso=1920, yreso=1080, subdiv=7, n_spheres=64, seed=1, walls="quad") -> Scene:
    """The Cornell box of two_spheres_scene holding a displaced icosphere
    (subdivision `subdiv`: 327,680 triangles at 7; radius 0.22 m about
    (0.5, 0.3, 0.6), radial noise +-2%, diffuse albedo 0.75) and `n_spheres`
    spheres of radius U(0.02, 0.08) at seeded non-overlapping places inside
    the box, each diffuse (0.75), mirror or glass (IOR 1.5) with equal
    probability.  Camera as in the other box scenes, 1920x1080."""
    sc = cornell_box(Scene(), walls=walls)
    dm = sc.add_material(lambertian(0.75))
    mm = sc.add_material(mirror())
    gm = sc.add_material(glass(1.5))
    ctr, rad = np.array([0.5, 0.3, 0.6]), 0.22
    vert, face = icosphere(subdiv, ctr, rad, noise=0.02, seed=seed)
    rng = np.random.Generator(np.random.Philox(key=int(seed) + 1))
    placed = []
    while len(placed) < n_spheres:
        r = rng.uniform(0.02, 0.08)
        c = rng.uniform(r + 0.01, 1.0 - r - 0.01, 3)
        if np.linalg.norm(c - ctr) < rad * 1.02 + r + 0.005:
            continue
        if any(np.linalg.norm(c - q) < r + s + 0.005 for q, s, _ in placed):
            continue
        placed.append((c, r, (dm, mm, gm)[int(rng.integers(0, 3))]))
    for c, r, m in placed:
        sc.shapes.append(sphere(c, r, m))
    sc.meshes.append(Mesh(vert[face], dm))
    sc.camera = box_camera(xreso, yreso)
    return sc


def batch_scenes(n=64, seed=1, max_side=512, min_side=128):
    """The batched workload: n Cornell-box-like rooms (1 x 1 x 1 m, quad
    walls of random RGB albedo U(0.1, 0.9)^3 each, the 100 W ceiling light),
    each holding k spheres, k log-uniform in [2, 512], of radius
    U(0.5, 1) * min(0.15, 0.25 k^(-1/3)) at non-overlapping places (at most
    200 tries per sphere), each diffuse (0.75), mirror or glass (IOR 1.5) with
    equal probability.  Per room: a square image of side U{min_side..max_side},
    photons per pass log-uniform in [50k, 2M], passes U{32..256}, r_1 =
    0.01 m * sqrt(200k / photons), alpha = 2/3, max depth 8, seed = seed * 1000
    + room.  Draws come from a Philox stream keyed by `seed`.

    Returns [(scene, cfg dict, passes)]."""
    rng = np.random.Generator(np.random.Philox(key=int(seed) + 7))
    out = []
    for i in range(n):
        sc = Scene()
        walls = [lambertian(tuple(rng.uniform(0.1, 0.9, 3))) for _ in range(5)]
        wm = [sc.add_material(m) for m in walls]
        sc.shapes += [
            parallelogram((0, 0, 0), (0, 0, 1), (1, 0, 0), wm[0]),
            parallelogram((0, 1, 0), (1, 1, 0), (0, 1, 1), wm[1]),
            parallelogram((0, 0, 1), (0, 1, 1), (1, 0, 1), wm[2]),
            parallelogram((0, 0, 0), (0, 1, 0), (0, 0, 1), wm[3]),
            parallelogram((1, 0, 0), (1, 0, 1), (1, 1, 0), wm[4]),
        ]
        sc.bounds_min, sc.bounds_max = BOX_BOUNDS
        s = 0.25
        x0 = 0.5 - s / 2
        p0 = (x0, 0.999, x0)
        em = 2.0 * 100.0 / (s * s)
        lm = sc.add_material(simple((0, 0, 0), (0.8, 0.8, 0.8), 0.0, 0.0, 0.0, emittance=(em, em, em)))
        sc.shapes.append(parallelogram(p0, (x0 + s, 0.999, x0), (x0, 0.999, x0 + s), lm, nvec=(0, -1, 0)))
        sc.lights.append(parallelogram_light((1, 1, 1), 100.0, p0, (0, -1, 0), (s, 0, 0), (0, 0, s)))
        mats = (sc.add_material(lambertian(0.75)), sc.add_material(mirror()), sc.add_material(glass(1.5)))
        k = int(round(np.exp(rng.uniform(np.log(2), np.log(512)))))
        rmax = min(0.15, 0.25 * k ** (-1.0 / 3.0))
        placed = []
        for _ in range(k):
            for _try in range(200):
                r = rng.uniform(0.5, 1.0) * rmax
                c = rng.uniform(r + 0.005, 1.0 - r - 0.005, 3)
                c[1] = min(c[1], 0.999 - r - 0.01)
                if all(np.linalg.norm(c - q) > r + t for q, t in placed):
                    placed.append((c, r))
                    sc.shapes.append(sphere(c, r, mats[int(rng.integers(0, 3))]))
                    break
        side = int(rng.integers(min_side, max_side + 1))
        sc.camera = box_camera(side, side)
        n_photon = int(round(np.exp(rng.uniform(np.log(50_000), np.log(2_000_000)))))
        passes = int(rng.integers(32, 257))
        cfg = dict(n_photon=n_photon, radius0=0.01 * np.sqrt(200_000 / n_photon), alpha=2.0 / 3.0, max_depth=8,
                   seed=int(seed) * 1000 + i)
        out.append((sc, cfg, passes))
    return out


N_INVERSE_MATERIALS = 32


def inverse_albedos(seed):
    """32 RGB albedos drawn from U(0.1, 0.9) by a Philox stream keyed by seed."""
    rng = np.random.Generator(np.random.Philox(key=int(seed) + 11))
    return rng.uniform(0.1, 0.9, (N_INVERSE_MATERIALS, 3))


def inverse_scene(albedo=None, xreso=512, yreso=512, seed=1, walls="quad") -> Scene:
    """The differentiable workload's scene: the caustic scene's box, mirror and
    glass spheres, with 32 Lambertian materials as the unknowns: materials
    0-2 are the walls (floor, ceiling and back share 0; left 1; right 2) and
    3-31 are 29 spheres of radius 0.05 at seeded non-overlapping places.
    albedo: 32 x 3 (default U(0.1, 0.9) from `seed`).  Materials 0-31 come
    first, so a gradient over the scene's materials starts with them."""
    alb = inverse_albedos(seed) if albedo is None else np.asarray(albedo, np.float64).reshape(N_INVERSE_MATERIALS, 3)
    sc = Scene()
    mats = [sc.add_material(lambertian(tuple(alb[i]))) for i in range(N_INVERSE_MATERIALS)]
    if walls == "plane":
        sc.shapes += [plain((0, 1, 0), 0.0, mats[0]), plain((0, -1, 0), 1.0, mats[0]), plain((0, 0, -1), 1.0, mats[0]),
                      plain((1, 0, 0), 0.0, mats[1]), plain((-1, 0, 0), 1.0, mats[2])]
        sc.bounds_min, sc.bounds_max = TUNNEL_BOUNDS
    else:
        sc.shapes += [
            parallelogram((0, 0, 0), (0, 0, 1), (1, 0, 0), mats[0]),
            parallelogram((0, 1, 0), (1, 1, 0), (0, 1, 1), mats[0]),
            parallelogram((0, 0, 1), (0, 1, 1), (1, 0, 1), mats[0]),
            parallelogram((0, 0, 0), (0, 1, 0), (0, 0, 1), mats[1]),
            parallelogram((1, 0, 0), (1, 0, 1), (1, 1, 0), mats[2]),
        ]
        sc.bounds_min, sc.bounds_max = BOX_BOUNDS
    s = 0.25
    x0 = 0.5 - s / 2
    p0 = (x0, 0.999, x0)
    em = 2.0 * 100.0 / (s * s)
    lm = sc.add_material(simple((0, 0, 0), (0.8, 0.8, 0.8), 0.0, 0.0, 0.0, emittance=(em, em, em)))
    sc.shapes.append(parallelogram(p0, (x0 + s, 0.999, x0), (x0, 0.999, x0 + s), lm, nvec=(0, -1, 0)))
    sc.lights.append(parallelogram_light((1, 1, 1), 100.0, p0, (0, -1, 0), (s, 0, 0), (0, 0, s)))
    mm = sc.add_material(mirror())
    gm = sc.add_material(glass(1.5))
    big = [(np.array([0.75, 0.2, 0.75]), 0.2), (np.array([0.5, 0.2, 0.3]), 0.2)]
    sc.shapes += [sphere(big[0][0], 0.2, mm), sphere(big[1][0], 0.2, gm)]
    rng = np.random.Generator(np.random.Philox(key=int(seed) + 13))
    placed = []
    while len(placed) < N_INVERSE_MATERIALS - 3:
        c = rng.uniform(0.06, 0.94, 3)
        c[1] = rng.uniform(0.05, 0.8)
        if all(np.linalg.norm(c - q) > 0.05 + r + 0.01 for q, r in big + [(p, 0.05) for p in placed]):
            placed.append(c)
    for i, c in enumerate(placed):
        sc.shapes.append(sphere(c, 0.05, mats[3 + i]))
    sc.camera = box_camera(xreso, yreso)
    return sc


def reference_room() -> Scene:
    """A room in the style of the reference's built-in scene (scene.rs:20-447):
    4 x 4 x 11 m walls of Plain shapes, a 1.33 m ceiling parallelogram light of
    5 W with its emissive object (scene.rs:226-238,434-441), and balls of the
    reference's Simple/TS materials; camera = read_camera's defaults."""
    sc = Scene()
    mwall = sc.add_material(simple((0.5, 0.5, 0.5), (0.8, 0.8, 0.8), 1.0, 0.0, 0.0))
    mwallb = sc.add_material(simple((0.1, 0.1, 0.4), (0.8, 0.0, 0.8), 1.0, 0.0, 0.0))
    mwallr = sc.add_material(simple((0.4, 0.1, 0.1), (0.0, 0.0, 0.0), 1.0, 0.0, 0.0))
    mparal = sc.add_material(simple((0, 0, 0), (0.8, 0.8, 0.8), 0.0, 0.0, 0.0, emittance=(0.15, 0.15, 0.15)))
    silver = sc.add_material(simple((0, 0, 0), (0.78, 0.78, 0.78), 0.0, 1.0, 0.7))
    tsball = sc.add_material(ts((0, 0, 0), (0.78, 0.78, 0.78), 0.0, 1.0, 0.5))
    glassb = sc.add_material(ts((1, 1, 1), (0.05, 0.05, 0.05), 0.0, 0.0, 0.8, ior=(1.5, 1.5, 1.5)))
    sc.shapes += [
        plain((0, 1, 0), 0.0, mwall), plain((0, -1, 0), 4.0, mwall),
        plain((-1, 0, 0), 2.0, mwallb), plain((1, 0, 0), 2.0, mwallr),
        plain((0, 0, -1), 6.0, mwall), plain((0, 0, 1), 5.0, mwall),
        sphere((-0.8, 0.5, 2.5), 0.4, silver), sphere((0.0, 0.5, 2.5), 0.4, tsball),
        sphere((0.8, 0.5, 2.5), 0.4, glassb),
        parallelogram((-0.67, 3.99, 2.33), (0.67, 3.99, 2.33), (-0.67, 3.99, 3.67), mparal, nvec=(0, -1, 0)),
    ]
    sc.lights.append(parallelogram_light((1, 1, 1), 5.0, (-0.67, 3.99, 2.33), (0, -1, 0), (1.33, 0, 0),
                                         (0, 0, 1.33)))
    sc.camera = read_camera()
    sc.bounds_min, sc.bounds_max = (-2.01, -0.01, -5.01), (2.01, 4.01, 6.01)
    return sc


# Pass schedules of the two workloads: the small CPU-runnable one and the
# 2M-photon, 512-pass caustic scene the throughput is quoted on.
CONFIGS = {
    "two_spheres": dict(scene=two_spheres_scene, n_photon=200_000, passes=16, radius0=0.01, alpha=2.0 / 3.0, max_depth=8,
                    seed=1),
    "caustic": dict(scene=caustic_scene, n_photon=2_000_000, passes=512, radius0=0.005, alpha=2.0 / 3.0,
                    max_depth=16, seed=1),
    # the same box with infinite Plain walls, the reference's own wall shape
    # (geometry.rs:170-177): the open front lets photons escape into the
    # planes' far reaches, where they are stored too
    "caustic_plane": dict(scene=lambda xreso=1024, yreso=768: caustic_scene(xreso, yreso, walls="plane"),
                          n_photon=2_000_000, passes=512, radius0=0.005, alpha=2.0 / 3.0, max_depth=16, seed=1),
    "mesh": dict(scene=mesh_scene, n_photon=8_000_000, passes=256, radius0=0.005, alpha=2.0 / 3.0, max_depth=16,
                 seed=1),
    "inverse": dict(scene=inverse_scene, n_photon=1_000_000, passes=64, radius0=0.005, alpha=2.0 / 3.0, max_depth=16,
                    seed=1),
}
