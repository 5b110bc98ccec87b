"""Seeded procedural scenes of the benchmark configurations.

Each returns a dict with numpy arrays in the C-ABI layouts:
  positions  float32 [T, 9]      (v0, v1, v2 per triangle)
  normals    float32 [T, 9] or None
  mat_ids    uint32  [T]
  meshes     list of (offset, count)      (one mesh for single-level scenes)
  materials  float32 [M, 17]    (trt_material layout; the last column holds the
                                 model id as uint32 bits)
  instances  float32 [I, 14] or None  (3x4 transform, mesh, material_offset as uint32 bits)
  camera     dict(position, target, up, fov, aperture, focal_dist, distortion,
                  bokeh_sides, bokeh_rotation)
  render     dict(width, height, spp, max_depth, rr_depth, seed, ray_epsilon, clamp)

The reference renders scenes loaded from JSON/OBJ files (Resources/scenes/cornell.json);
these generators build the benchmark scenes procedurally instead (its model files are
not available).
"""
import math

import numpy as np

MODEL_DISNEY = 0
MODEL_LAMBERT = 1


def material(base=(0.8, 0.8, 0.8), roughness=0.5, emission=(0, 0, 0), metallic=0.0, specular=0.5,
             specular_tint=0.0, sheen=0.0, sheen_tint=0.5, clearcoat=0.0, clearcoat_gloss=1.0, subsurface=0.0,
             anisotropic=0.0, model=MODEL_DISNEY):
    m = np.zeros(17, np.float32)
    m[0:3] = base
    m[3] = roughness
    m[4:7] = emission
    m[7] = metallic
    m[8:16] = [specular, specular_tint, sheen, sheen_tint, clearcoat, clearcoat_gloss, subsurface, anisotropic]
    m[16:17] = np.array([model], np.uint32).view(np.float32)
    return m


def _quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); normal = (p1-p0) x (p2-p0)."""
    return [np.concatenate([p0, p1, p2]), np.concatenate([p0, p2, p3])]


def _box(center, size, angle_y):
    """Axis-aligned box rotated about y; outward-facing triangles (12)."""
    hx, hy, hz = size[0] / 2, size[1] / 2, size[2] / 2
    c, s = math.cos(angle_y), math.sin(angle_y)

    def P(x, y, z):
        return np.array([center[0] + c * x + s * z, center[1] + y, center[2] - s * x + c * z], np.float32)

    v = {k: P(*k) for k in [(sx * hx, sy * hy, sz * hz) for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]}

    def V(sx, sy, sz):
        return v[(sx * hx, sy * hy, sz * hz)]

    tris = []
    tris += _quad(V(-1, 1, -1), V(-1, 1, 1), V(1, 1, 1), V(1, 1, -1))      # top (+y)
    tris += _quad(V(-1, -1, -1), V(1, -1, -1), V(1, -1, 1), V(-1, -1, 1))  # bottom (-y)
    tris += _quad(V(-1, -1, -1), V(-1, 1, -1), V(1, 1, -1), V(1, -1, -1))  # front (-z)
    tris += _quad(V(-1, -1, 1), V(1, -1, 1), V(1, 1, 1), V(-1, 1, 1))      # back (+z)
    tris += _quad(V(-1, -1, -1), V(-1, -1, 1), V(-1, 1, 1), V(-1, 1, -1))  # left (-x)
    tris += _quad(V(1, -1, -1), V(1, 1, -1), V(1, 1, 1), V(1, -1, 1))      # right (+x)
    return tris


def cornell(width=256, height=256, spp=16, max_depth=4, rr_depth=3, seed=0):
    """Procedural Cornell box in [0,1]^3: 5 wall quads + 2 rotated boxes + a 0.25 x 0.25
    ceiling light of radiance (17, 12, 4) = 36 triangles. Lambert walls (red / green /
    white), one white Lambert box and one GGX metal box (roughness 0.2)."""
    P = lambda *a: np.array(a, np.float32)  # noqa: E731
    tris, mats = [], []
    # material ids: 0 white, 1 red, 2 green, 3 light, 4 metal
    white = material((0.725, 0.71, 0.68), roughness=1.0, model=MODEL_LAMBERT)
    red = material((0.63, 0.065, 0.05), roughness=1.0, model=MODEL_LAMBERT)
    green = material((0.14, 0.45, 0.091), roughness=1.0, model=MODEL_LAMBERT)
    light = material((0.78, 0.78, 0.78), roughness=1.0, emission=(17.0, 12.0, 4.0), model=MODEL_LAMBERT)
    metal = material((0.95, 0.93, 0.88), roughness=0.2, metallic=1.0, model=MODEL_DISNEY)
    materials = np.stack([white, red, green, light, metal])

    def add(tlist, mid):
        for t in tlist:
            tris.append(t)
            mats.append(mid)

    add(_quad(P(0, 0, 0), P(0, 0, 1), P(1, 0, 1), P(1, 0, 0)), 0)      # floor (normal +y)
    add(_quad(P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1)), 0)      # ceiling (normal -y)
    add(_quad(P(0, 0, 1), P(0, 1, 1), P(1, 1, 1), P(1, 0, 1)), 0)      # back wall (normal -z)
    add(_quad(P(0, 0, 0), P(0, 1, 0), P(0, 1, 1), P(0, 0, 1)), 1)      # x = 0 wall, red
    add(_quad(P(1, 0, 0), P(1, 0, 1), P(1, 1, 1), P(1, 1, 0)), 2)      # x = 1 wall, green
    add(_box((0.33, 0.3, 0.62), (0.28, 0.6, 0.28), 0.3), 0)            # tall white box
    add(_box((0.68, 0.15, 0.36), (0.28, 0.3, 0.28), -0.3), 4)          # short metal box
    y = 0.999
    add(_quad(P(0.375, y, 0.375), P(0.625, y, 0.375), P(0.625, y, 0.625), P(0.375, y, 0.625)), 3)  # light, -y
    positions = np.stack(tris).astype(np.float32)
    assert positions.shape[0] == 36
    return dict(
        positions=positions, normals=None, mat_ids=np.array(mats, np.uint32), meshes=[(0, 36)],
        materials=materials, instances=None,
        camera=dict(position=(0.5, 0.5, -1.4), target=(0.5, 0.5, 0.5), up=(0, 1, 0), fov=math.radians(40.0),
                    aperture=0.0, focal_dist=1.0, distortion=0.0, bokeh_sides=0, bokeh_rotation=0.0),
        render=dict(width=width, height=height, spp=spp, max_depth=max_depth, rr_depth=rr_depth, seed=seed,
                    ray_epsilon=1e-3, clamp=0.0),
    )


PARAM_LO = np.array([0.0, 0.0, 0.0, 0.05, 0.0], np.float32)
PARAM_HI = np.array([1.0, 1.0, 1.0, 1.0, 1.0], np.float32)


def material_params(materials):
    """[M, 5] differentiable parameters (base colour rgb, roughness, metallic)."""
    m = np.asarray(materials, np.float32)
    return np.concatenate([m[:, 0:3], m[:, 3:4], m[:, 7:8]], axis=1).copy()


def cornell_diff(width=512, height=512, spp=64, max_depth=4, seed=3, sigma=0.1, rr_depth=None):
    """Differentiable Cornell box: the Cornell geometry with one Disney material per
    triangle (36 x 5 = 180 parameters: base colour rgb, roughness, metallic). Returns
    (scene, target_params): the target is the scene's parameters perturbed by
    N(0, sigma) and clipped to their valid ranges. Russian roulette starts at rr_depth
    (default: the last vertex, max_depth - 1, where no extension is traced); its survival
    probability comes from the sampling parameters, so with those held fixed the
    estimator is a smooth function of the evaluation parameters either way."""
    rr = max_depth - 1 if rr_depth is None else rr_depth
    sc = cornell(width=width, height=height, spp=spp, max_depth=max_depth, rr_depth=rr, seed=seed)
    base = sc["materials"][sc["mat_ids"]].copy()
    rng = np.random.default_rng(seed)
    for i, m in enumerate(base):
        m[16:17] = np.array([MODEL_DISNEY], np.uint32).view(np.float32)
        if m[4:7].sum() == 0:  # non-emitters get varied roughness / metallic
            m[3] = rng.uniform(0.3, 0.9)
            m[7] = 1.0 if m[7] > 0.5 else rng.uniform(0.0, 0.3)
    sc["materials"] = base
    sc["mat_ids"] = np.arange(36, dtype=np.uint32)
    p = material_params(base)
    target = np.clip(p + rng.normal(0.0, sigma, p.shape).astype(np.float32), PARAM_LO, PARAM_HI)
    return sc, target


def _unit_vectors(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def soup(n_tris=1_000_000, seed=1, width=1920, height=1080, spp=8, max_depth=8, rr_depth=3, n_emissive=64):
    """Random triangle soup: centres uniform in the unit cube, vertices at c + l*u_k with
    l log-uniform in [1e-3, 1e-1] and u_k uniform on the sphere. Per-triangle Disney
    parameters: baseColor U[0,1]^3, roughness U[0.05,1], metallic in {0,1} with p=0.3.
    `n_emissive` randomly chosen triangles emit grey radiance P/(pi A) with power P
    log-uniform in [1, 100]."""
    rng = np.random.default_rng(seed)
    c = rng.random((n_tris, 3))
    ell = np.exp(rng.uniform(math.log(1e-3), math.log(1e-1), n_tris))
    pos = np.empty((n_tris, 9), np.float32)
    for k in range(3):
        pos[:, 3 * k:3 * k + 3] = (c + ell[:, None] * _unit_vectors(rng, n_tris)).astype(np.float32)
    mats = np.zeros((n_tris, 17), np.float32)
    mats[:, 0:3] = rng.random((n_tris, 3))
    mats[:, 3] = rng.uniform(0.05, 1.0, n_tris)
    mats[:, 7] = (rng.random(n_tris) < 0.3).astype(np.float32)
    mats[:, 8] = 0.5       # specular
    mats[:, 11] = 0.5      # sheen tint
    mats[:, 13] = 1.0      # clearcoat gloss
    mats[:, 16] = np.array([MODEL_DISNEY], np.uint32).view(np.float32)[0]
    em = rng.choice(n_tris, size=n_emissive, replace=False)
    p = pos[em].astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(p[:, 3:6] - p[:, 0:3], p[:, 6:9] - p[:, 0:3]), axis=1)
    power = np.exp(rng.uniform(0.0, math.log(100.0), n_emissive))
    rad = power / (math.pi * np.maximum(area, 1e-12))
    mats[em, 4:7] = rad[:, None].astype(np.float32)
    return dict(
        positions=pos, normals=None, mat_ids=np.arange(n_tris, dtype=np.uint32), meshes=[(0, n_tris)],
        materials=mats, instances=None, emissive=em,
        camera=dict(position=(0.5, 0.5, -1.2), target=(0.5, 0.5, 0.5), up=(0, 1, 0), fov=math.radians(50.0),
                    aperture=0.0, focal_dist=1.0, distortion=0.0, bokeh_sides=0, bokeh_rotation=0.0),
        render=dict(width=width, height=height, spp=spp, max_depth=max_depth, rr_depth=rr_depth, seed=seed,
                    ray_epsilon=1e-5, clamp=0.0),
    )


def icosphere(subdiv=5):
    """Unit icosphere by repeated 4-way subdivision of an icosahedron (20 * 4^subdiv
    triangles); returns positions [T,9] and smooth normals [T,9]."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts)
    F = np.array(faces)
    pos = V[F].reshape(-1, 9).astype(np.float32)
    return pos, pos.copy()


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def instanced(n_instances=1000, subdiv=5, seed=2, width=3840, height=2160, spp=4, max_depth=6, rr_depth=3,
              emissive_per_instance=10, sigma=2.0):
    """Two-level scene: `n_instances` instances of one subdivided icosphere under random
    rotation, scale U[0.5, 2] and translation U[0, 100]^3. Each instance owns a material
    palette (base Disney material + `emissive_per_instance` emissive slots assigned to
    fixed triangles of the mesh), giving n_instances * emissive_per_instance emissive
    triangles with log-normal power (sigma = 2)."""
    rng = np.random.default_rng(seed)
    pos, nrm = icosphere(subdiv)
    T = pos.shape[0]
    mat_ids = np.zeros(T, np.uint32)
    em_tris = rng.choice(T, size=emissive_per_instance, replace=False)
    mat_ids[em_tris] = np.arange(1, emissive_per_instance + 1, dtype=np.uint32)
    pal = emissive_per_instance + 1
    mats = np.zeros((n_instances * pal, 17), np.float32)
    inst = np.zeros((n_instances, 14), np.float32)
    p = pos[em_tris].astype(np.float64)
    area_obj = 0.5 * np.linalg.norm(np.cross(p[:, 3:6] - p[:, 0:3], p[:, 6:9] - p[:, 0:3]), axis=1)
    for i in range(n_instances):
        R = _random_rotation(rng)
        s = rng.uniform(0.5, 2.0)
        tr = rng.uniform(0.0, 100.0, 3)
        M = np.zeros((3, 4))
        M[:, :3] = R * s
        M[:, 3] = tr
        inst[i, :12] = M.reshape(-1)
        inst[i, 12:14] = np.array([0, i * pal], np.uint32).view(np.float32)
        base = material(tuple(rng.random(3)), roughness=float(rng.uniform(0.05, 1.0)),
                        metallic=float(rng.random() < 0.3))
        mats[i * pal] = base
        power = np.exp(rng.normal(0.0, sigma, emissive_per_instance))
        rad = power / (math.pi * area_obj * s * s)
        for k in range(emissive_per_instance):
            m = base.copy()
            m[4:7] = rad[k]
            mats[i * pal + 1 + k] = m
    return dict(
        positions=pos, normals=nrm, mat_ids=mat_ids, meshes=[(0, T)], materials=mats, instances=inst,
        camera=dict(position=(50.0, 50.0, -90.0), target=(50.0, 50.0, 50.0), up=(0, 1, 0), fov=math.radians(60.0),
                    aperture=0.0, focal_dist=1.0, distortion=0.0, bokeh_sides=0, bokeh_rotation=0.0),
        render=dict(width=width, height=height, spp=spp, max_depth=max_depth, rr_depth=rr_depth, seed=seed,
                    ray_epsilon=1e-4, clamp=0.0),
    )


def from_file(path, width=256, height=256, rr_depth=3, seed=0):
    """Reads a scene file in the reference's JSON format (and its OBJ models) on the
    host, through the library's parser (trt_scene_file_read; no device use), into the
    scene dict used by Renderer.UpdateScene and the oracle. Returns (scene, info,
    warnings): the file's camera and render settings (multisample -> spp, maxdepth,
    ray-epsilon) go into the dict; film size, seed and RR depth are the arguments."""
    import ctypes
    import os

    from . import _lib
    lib = _lib.load()
    info = _lib.TrtSceneFileInfo()
    msg = ctypes.create_string_buffer(1 << 16)
    p = os.fsencode(path)
    st = lib.trt_scene_file_read(p, ctypes.byref(info), None, None, None, None, None, None, None, None, msg, len(msg))
    if st != 0:
        raise _lib.TrtError(st, msg.value.decode(errors="replace"))
    T, M, K, I = info.n_triangles, info.n_models, info.n_materials, info.n_instances
    pos = np.zeros((T, 9), np.float32)
    nrm = np.zeros((T, 9), np.float32)
    mids = np.zeros(T, np.uint32)
    toff = np.zeros(M, np.uint32)
    tcnt = np.zeros(M, np.uint32)
    moff = np.zeros(M, np.uint32)
    mats = (_lib.TrtMaterial * max(K, 1))()
    insts = (_lib.TrtInstance * max(I, 1))()
    fp = lambda a: a.ctypes.data_as(_lib.c_float_p)  # noqa: E731
    up = lambda a: a.ctypes.data_as(_lib.c_u32_p)  # noqa: E731
    st = lib.trt_scene_file_read(p, ctypes.byref(info), fp(pos), fp(nrm), up(mids), up(toff), up(tcnt), up(moff),
                                 mats, insts, msg, len(msg))
    if st != 0:
        raise _lib.TrtError(st, msg.value.decode(errors="replace"))
    mat_arr = np.frombuffer(bytes(mats), np.float32).reshape(-1, 17)[:K].copy()
    inst_arr = np.frombuffer(bytes(insts), np.float32).reshape(-1, 14)[:I].copy()
    d = info.as_dict()
    cam = d["camera"] if info.has_camera else dict(position=(0.0, 0.0, -1.0), target=(0.0, 0.0, 0.0),
                                                      up=(0.0, 1.0, 0.0), fov=math.radians(90.0))
    sky = dict(enabled=bool(info.sky_enabled), draw_sun=bool(info.sky_draw_sun), dir=tuple(info.sky_dir),
               angular_diameter=float(info.sky_angular_diameter), intensity=float(info.sky_intensity),
               turbidity=float(info.sky_turbidity))
    scene = dict(positions=pos, normals=nrm, mat_ids=mids, meshes=[(int(o), int(c)) for o, c in zip(toff, tcnt)],
                 materials=mat_arr, instances=inst_arr, camera=cam, sky=sky,
                 render=dict(width=width, height=height, spp=int(info.spp), max_depth=int(info.max_depth),
                             rr_depth=rr_depth, seed=seed, ray_epsilon=float(info.ray_epsilon), clamp=0.0,
                             ao_dist=float(info.ao_dist), zdepth_max=float(info.zdepth_max)))
    warnings = [w for w in msg.value.decode(errors="replace").split("\n") if w]
    return scene, d, warnings
