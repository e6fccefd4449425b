"""Procedural scenes of the SMS workloads, built from a seed.

A ``Scene`` owns numpy arrays for every buffer a shape points at and turns itself into an
``SmsSceneDesc`` for ``sms_ctx_create`` (and for the CPU oracle, which takes the same
description).  Nothing here runs the solver.

Workloads (sizes as used by the benchmark and the parity tests):
  * sphere_caustic  -- glass sphere r=1, IOR 1.5 at (0,1.5,0) over a diffuse plane
    (albedo 0.5), sphere area light r=0.1, radiance 100 at (0,5,0), pinhole 45 deg.
  * glint_plane     -- 1x1 plane with a seed-generated RGB8 normal map of line scratches and
    Gaussian bumps on a mirror conductor, under a point light; footprint-seeded glints.
  * pool            -- 512x512 grid mesh displaced by Gerstner waves (IOR 1.33) over a floor.
"""
import ctypes as C
import math

import numpy as np

from . import abi


def _vec(ctype_arr, values):
    for i, v in enumerate(values):
        ctype_arr[i] = float(v)


class Scene:
    def __init__(self):
        self.shapes = []        # dicts
        self.emitters = []      # dicts
        self.normalmaps = []    # (res, uint8 or float32 array)
        self.encoded_normalmaps = set()   # indices of maps whose frames decode interpolated RGB
        self.camera = None
        self._keep = []

    # ---------------------------------------------------------------- builders
    def add_sphere(self, center, radius, bsdf, flags=0, eta=1.5, k=0.0, reflectance=(0.5, 0.5, 0.5),
                   emitter=-1, roughness=0.0):
        self.shapes.append(dict(type=abi.SHAPE_SPHERE, bsdf=bsdf, flags=flags, center=center,
                                radius=radius, eta=eta, k=k, reflectance=reflectance,
                                emitter=emitter, roughness=roughness))
        return len(self.shapes) - 1

    def add_rectangle(self, to_world, bsdf, flags=0, eta=1.5, k=0.0, reflectance=(0.5, 0.5, 0.5),
                      normalmap=-1, tiles=1.0, roughness=0.0):
        self.shapes.append(dict(type=abi.SHAPE_RECTANGLE, bsdf=bsdf, flags=flags,
                                to_world=np.asarray(to_world, np.float32).reshape(12), eta=eta, k=k,
                                reflectance=reflectance, normalmap=normalmap, tiles=tiles, roughness=roughness))
        return len(self.shapes) - 1

    def add_mesh(self, positions, faces, normals=None, texcoords=None, bsdf=abi.BSDF_DIFFUSE, flags=0,
                 eta=1.5, k=0.0, reflectance=(0.5, 0.5, 0.5), roughness=0.0):
        self.shapes.append(dict(type=abi.SHAPE_MESH, bsdf=bsdf, flags=flags, eta=eta, k=k,
                                reflectance=reflectance, roughness=roughness,
                                positions=np.ascontiguousarray(positions, np.float32),
                                faces=np.ascontiguousarray(faces, np.uint32),
                                normals=None if normals is None else np.ascontiguousarray(normals, np.float32),
                                texcoords=None if texcoords is None else np.ascontiguousarray(texcoords, np.float32)))
        return len(self.shapes) - 1

    def add_area_light(self, shape, radiance, caustic_ss=True, caustic_ms=True):
        self.emitters.append(dict(type=abi.EMITTER_AREA, shape=shape, radiance=radiance,
                                  caustic_ss=caustic_ss, caustic_ms=caustic_ms))
        self.shapes[shape]["emitter"] = len(self.emitters) - 1
        return len(self.emitters) - 1

    def add_point_light(self, position, intensity, caustic_ss=True, caustic_ms=True):
        self.emitters.append(dict(type=abi.EMITTER_POINT, shape=-1, position=position,
                                  intensity=intensity, caustic_ss=caustic_ss, caustic_ms=caustic_ms))
        return len(self.emitters) - 1

    def add_directional_light(self, direction, irradiance, caustic_ss=True, caustic_ms=True):
        """Directional emitter: light travelling along `direction` with `irradiance`
        (src/emitters/directional.cpp)."""
        d = np.asarray(direction, np.float64)
        d = d / np.linalg.norm(d)
        self.emitters.append(dict(type=abi.EMITTER_DIRECTIONAL, shape=-1, direction=d, irradiance=irradiance,
                                  caustic_ss=caustic_ss, caustic_ms=caustic_ms))
        return len(self.emitters) - 1

    def add_normalmap(self, rgb, encoded=False):
        """RGB normal map: uint8 (code values / 255) or float32 (values as an OpenEXR file
        stores them), square power-of-two.  encoded: frames interpolate the RGB and decode
        2 rgb - 1 (NormalmapBSDF without use_slopes) instead of interpolating slopes."""
        rgb = np.ascontiguousarray(rgb, np.float32 if np.asarray(rgb).dtype.kind == "f" else np.uint8)
        assert rgb.ndim == 3 and rgb.shape[0] == rgb.shape[1] and rgb.shape[2] == 3
        self.normalmaps.append((rgb.shape[0], rgb))
        if encoded:
            self.encoded_normalmaps.add(len(self.normalmaps) - 1)
        return len(self.normalmaps) - 1

    def set_camera(self, origin, target, up, fov_deg, width, height):
        self.camera = dict(origin=origin, target=target, up=up, fov_deg=fov_deg, width=width, height=height)

    # ---------------------------------------------------------------- C description
    def desc(self):
        keep = []
        shapes = (abi.SmsShapeDesc * max(1, len(self.shapes)))()
        for i, s in enumerate(self.shapes):
            d = shapes[i]
            d.type, d.bsdf, d.flags, d.emitter = s["type"], s["bsdf"], s["flags"], s.get("emitter", -1)
            _vec(d.center, s.get("center", (0, 0, 0)))
            d.radius = s.get("radius", 1.0)
            _vec(d.to_world, s.get("to_world", np.zeros(12)))
            d.eta, d.k = s.get("eta", 1.5), s.get("k", 0.0)
            _vec(d.reflectance, s.get("reflectance", (0.5, 0.5, 0.5)))
            d.normalmap, d.tiles = s.get("normalmap", -1), s.get("tiles", 1.0)
            d.roughness = s.get("roughness", 0.0)
            if s["type"] == abi.SHAPE_MESH:
                d.n_vertices = s["positions"].shape[0]
                d.n_faces = s["faces"].shape[0]
                d.positions = s["positions"].ctypes.data_as(C.POINTER(C.c_float))
                d.faces = s["faces"].ctypes.data_as(C.POINTER(C.c_uint32))
                if s["normals"] is not None:
                    d.normals = s["normals"].ctypes.data_as(C.POINTER(C.c_float))
                if s["texcoords"] is not None:
                    d.texcoords = s["texcoords"].ctypes.data_as(C.POINTER(C.c_float))
        emitters = (abi.SmsEmitterDesc * max(1, len(self.emitters)))()
        for i, e in enumerate(self.emitters):
            d = emitters[i]
            d.type, d.shape = e["type"], e["shape"]
            _vec(d.radiance, e.get("radiance", (0, 0, 0)))
            _vec(d.position, e.get("position", (0, 0, 0)))
            _vec(d.intensity, e.get("intensity", (0, 0, 0)))
            _vec(d.direction, e.get("direction", (0, 0, 0)))
            _vec(d.irradiance, e.get("irradiance", (0, 0, 0)))
            d.caustic_ss, d.caustic_ms = int(e["caustic_ss"]), int(e["caustic_ms"])
        nms = (abi.SmsNormalmapDesc * max(1, len(self.normalmaps)))()
        for i, (res, rgb) in enumerate(self.normalmaps):
            nms[i].res = res
            nms[i].encoded_normals = int(i in self.encoded_normalmaps)
            if rgb.dtype == np.float32:
                nms[i].rgb32f = rgb.ctypes.data_as(C.POINTER(C.c_float))
            else:
                nms[i].rgb = rgb.ctypes.data_as(C.POINTER(C.c_uint8))
        desc = abi.SmsSceneDesc()
        desc.n_shapes, desc.shapes = len(self.shapes), shapes
        desc.n_emitters, desc.emitters = len(self.emitters), emitters
        desc.n_normalmaps, desc.normalmaps = len(self.normalmaps), nms
        if self.camera:
            c = self.camera
            _vec(desc.camera.origin, c["origin"])
            _vec(desc.camera.target, c["target"])
            _vec(desc.camera.up, c["up"])
            desc.camera.fov_deg = c["fov_deg"]
            desc.camera.width, desc.camera.height = c["width"], c["height"]
        self._keep = [shapes, emitters, nms, desc]
        return desc


def plane_to_world(center, half_size, normal_axis="y"):
    """Rectangle to_world (row-major 3x4) mapping local x->+x, local y->-z, local z->+y."""
    s = float(half_size)
    cx, cy, cz = center
    return np.array([[s, 0, 0, cx],
                     [0, 0, 1, cy],
                     [0, -s, 0, cz]], np.float32)


def mirror_wall(normalmap=None, nm_res=64, n_scratches=16, seed=1, tiles=1.0):
    """Reflective single-scatter caustic from a normal-mapped mirror: a vertical 1x1 conductor
    wall (eta=0, k=1) at x = 0.8 facing -x, its normal map as glint_plane's, casts the light of
    a sphere area light (r 0.1, radiance 50) at (-0.4, 1.4, 0) onto a diffuse floor (y = 0,
    2x2, receiver).  Camera above the floor.  Shapes: 0 floor, 1 wall, 2 light."""
    sc = Scene()
    rgb = normalmap if normalmap is not None else scratch_bump_normalmap(nm_res, n_scratches, seed=seed)
    nm = sc.add_normalmap(rgb)
    sc.add_rectangle(plane_to_world((0, 0, 0), 1.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER)
    wall = np.array([[0, 0, -1, 0.8], [0, 0.5, 0, 0.6], [0.5, 0, 0, 0.0]], np.float32)
    sc.add_rectangle(wall, abi.BSDF_NORMALMAP, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=0.0, k=1.0, normalmap=nm,
                     tiles=tiles)
    light = sc.add_sphere((-0.4, 1.4, 0.0), 0.1, abi.BSDF_DIFFUSE)
    sc.add_area_light(light, (50.0, 50.0, 50.0))
    sc.set_camera((-1.2, 2.2, 0.0), (0.2, 0.0, 0.0), (0, 1, 0), 50.0, 64, 64)
    return sc


# ==================================================================== sphere caustic
def sphere_caustic(width=256, height=256, center_y=1.5, light_radius=0.1, light_radiance=100.0):
    """Glass sphere (r=1, IOR 1.5) at (0,center_y,0) over a diffuse plane (albedo 0.5), sphere
    area light r=0.1 with radiance 100 at (0,5,0); camera at (0,3,6) looking at (0,0.5,0), 45 deg.

    With the sphere wholly above the plane every single-scatter refraction path is rejected by
    the light-visibility test of the estimator (the light sees the solution vertex only through
    the sphere), so the estimate is zero; walks themselves converge normally.  ``sphere_dome``
    puts the centre on the plane, as the reference's own sphere scene does
    (results/Figure_14_15_MainComparison/sphere_sms.xml:116-128), which gives a non-zero
    single-scatter caustic inside the dome."""
    sc = Scene()
    sc.add_sphere((0.0, center_y, 0.0), 1.0, abi.BSDF_DIELECTRIC, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=1.5)
    sc.add_rectangle(plane_to_world((0, 0, 0), 10.0), abi.BSDF_DIFFUSE,
                     flags=abi.FLAG_CAUSTIC_RECEIVER, reflectance=(0.5, 0.5, 0.5))
    light = sc.add_sphere((0.0, 5.0, 0.0), light_radius, abi.BSDF_DIFFUSE, reflectance=(0.0, 0.0, 0.0))
    sc.add_area_light(light, (light_radiance,) * 3)
    sc.set_camera((0, 3, 6), (0, 0.5, 0), (0, 1, 0), 45.0, width, height)
    return sc


def sphere_dome(width=256, height=256, light_radius=0.1, light_radiance=100.0):
    """The sphere-caustic scene with the sphere centre on the receiver plane (a glass dome)."""
    return sphere_caustic(width, height, center_y=0.0, light_radius=light_radius, light_radiance=light_radiance)


# ==================================================================== water pool
def gerstner_grid(n=512, extent=1.0, n_waves=64, seed=3, amp_range=(1e-3, 2e-2), steepness=0.5):
    """(n+1)^2 grid over [-extent, extent]^2 (xz) displaced by ``n_waves`` Gerstner waves with
    amplitudes log-uniform in ``amp_range`` and random direction and phase; returns positions,
    analytic vertex normals, grid uvs and 2 n^2 triangles (FP64 evaluation, stored as float32)."""
    rng = np.random.default_rng(seed)
    amp = np.exp(rng.uniform(math.log(amp_range[0]), math.log(amp_range[1]), n_waves))
    ang = rng.uniform(0, 2 * np.pi, n_waves)
    dirs = np.stack([np.cos(ang), np.sin(ang)], -1)
    wavelen = np.clip(amp * rng.uniform(80, 160, n_waves), 0.05, 2.0)
    k = 2 * np.pi / wavelen
    phase = rng.uniform(0, 2 * np.pi, n_waves)
    q = steepness / (k * amp * n_waves)
    t = np.linspace(-extent, extent, n + 1)
    X0, Z0 = np.meshgrid(t, t, indexing="xy")
    P = np.stack([X0, np.zeros_like(X0), Z0], -1)
    dPdx = np.zeros_like(P)
    dPdz = np.zeros_like(P)
    dPdx[..., 0] = 1.0
    dPdz[..., 2] = 1.0
    for i in range(n_waves):
        th = k[i] * (dirs[i, 0] * X0 + dirs[i, 1] * Z0) + phase[i]
        c, s = np.cos(th), np.sin(th)
        P[..., 0] += q[i] * amp[i] * dirs[i, 0] * c
        P[..., 2] += q[i] * amp[i] * dirs[i, 1] * c
        P[..., 1] += amp[i] * s
        for d, ax in ((dPdx, 0), (dPdz, 1)):
            dth = k[i] * dirs[i, ax]
            d[..., 0] += -q[i] * amp[i] * dirs[i, 0] * s * dth
            d[..., 2] += -q[i] * amp[i] * dirs[i, 1] * s * dth
            d[..., 1] += amp[i] * c * dth
    N = np.cross(dPdz, dPdx)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    uu, vv = np.meshgrid(np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1), indexing="xy")
    uv = np.stack([uu, vv], -1)
    idx = np.arange((n + 1) * (n + 1)).reshape(n + 1, n + 1)
    a, b = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel()
    c_, d = idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
    faces = np.concatenate([np.stack([a, c_, b], -1), np.stack([b, c_, d], -1)])
    return (P.reshape(-1, 3).astype(np.float32), N.reshape(-1, 3).astype(np.float32),
            uv.reshape(-1, 2).astype(np.float32), faces.astype(np.uint32))


def pool(width=1024, height=1024, n=512, n_waves=64, seed=3, light_radius=0.05, light_radiance=2000.0):
    """Water surface (IOR 1.33, caustic caster) over a diffuse floor 1 unit below (receiver),
    sphere light of radius 0.05 above; single-scatter refractive caustics on the floor."""
    sc = Scene()
    P, N, UV, F = gerstner_grid(n, 1.0, n_waves, seed)
    sc.add_mesh(P, F, normals=N, texcoords=UV, bsdf=abi.BSDF_DIELECTRIC, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=1.33)
    sc.add_rectangle(plane_to_world((0, -1.0, 0), 1.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER,
                     reflectance=(0.5, 0.5, 0.5))
    light = sc.add_sphere((0.25, 2.5, 0.1), light_radius, abi.BSDF_DIFFUSE, reflectance=(0.0, 0.0, 0.0))
    sc.add_area_light(light, (light_radiance,) * 3)
    sc.set_camera((0.0, 2.2, 2.6), (0.0, -1.0, 0.0), (0, 1, 0), 40.0, width, height)
    return sc


# ==================================================================== double-refraction slab
def value_noise(n, seed, octaves=(8, 16, 32), gains=(1.0, 0.5, 0.25)):
    """(n+1)^2 smooth value noise in [-1, 1]: random lattice values, smoothstep-interpolated,
    summed over octaves and normalised by the maximum magnitude."""
    rng = np.random.default_rng(seed)
    t = np.linspace(0.0, 1.0, n + 1)
    h = np.zeros((n + 1, n + 1))
    for cells, g in zip(octaves, gains):
        lat = rng.uniform(-1, 1, (cells + 1, cells + 1))
        x = t * cells
        i = np.minimum(np.floor(x).astype(int), cells - 1)
        f = x - i
        f = f * f * (3 - 2 * f)
        a = lat[np.ix_(i, i)]
        b = lat[np.ix_(i, i + 1)]
        c = lat[np.ix_(i + 1, i)]
        d = lat[np.ix_(i + 1, i + 1)]
        fy, fx = f[:, None], f[None, :]
        h += g * ((a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy)
    return h / np.max(np.abs(h))


def height_mesh(h, extent, y0, amplitude, flip):
    """Grid mesh over [-extent, extent]^2 at height y0 + amplitude*h with central-difference
    vertex normals (pointing down when ``flip``) and grid uvs."""
    n = h.shape[0] - 1
    t = np.linspace(-extent, extent, n + 1)
    X, Z = np.meshgrid(t, t, indexing="xy")
    Y = y0 + amplitude * h
    P = np.stack([X, Y, Z], -1)
    gy, gx = np.gradient(Y, t, t)
    N = np.stack([-gx, np.ones_like(Y), -gy], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    uu, vv = np.meshgrid(np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1), indexing="xy")
    idx = np.arange((n + 1) * (n + 1)).reshape(n + 1, n + 1)
    a, b = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel()
    c, d = idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
    faces = np.concatenate([np.stack([a, c, b], -1), np.stack([b, c, d], -1)])
    if flip:
        N = -N
        faces = faces[:, [0, 2, 1]]
    return (P.reshape(-1, 3).astype(np.float32), N.reshape(-1, 3).astype(np.float32),
            np.stack([uu, vv], -1).reshape(-1, 2).astype(np.float32), faces.astype(np.uint32))


def slab(width=1024, height=1024, n=256, thickness=0.2, amplitude=0.01, seed=5, caster="top"):
    """Glass slab (IOR 1.5) whose top and bottom faces carry independent value-noise
    heightfields, one the multi-scatter caster (`caster`), the other a bouncer; diffuse floor
    below, sphere light above.  Chains floor -> bottom face -> top face -> light have two
    specular vertices.  The path integrators' chain records (path_sms_ms.cpp:85-138,
    path_filtered_ms.cpp:248-299) expect the caster next to the receiver: caster="bottom"."""
    sc = Scene()
    top = height_mesh(value_noise(n, seed), 0.5, 0.5 + thickness / 2, amplitude, flip=False)
    bot = height_mesh(value_noise(n, seed + 1), 0.5, 0.5 - thickness / 2, amplitude, flip=True)
    ft, fb = ((abi.FLAG_CAUSTIC_CASTER_MS, abi.FLAG_CAUSTIC_BOUNCER) if caster == "top"
              else (abi.FLAG_CAUSTIC_BOUNCER, abi.FLAG_CAUSTIC_CASTER_MS))
    sc.add_mesh(top[0], top[3], normals=top[1], texcoords=top[2], bsdf=abi.BSDF_DIELECTRIC, flags=ft, eta=1.5)
    sc.add_mesh(bot[0], bot[3], normals=bot[1], texcoords=bot[2], bsdf=abi.BSDF_DIELECTRIC, flags=fb, eta=1.5)
    sc.add_rectangle(plane_to_world((0, 0, 0), 2.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER,
                     reflectance=(0.5, 0.5, 0.5))
    light = sc.add_sphere((0.1, 2.0, -0.1), 0.05, abi.BSDF_DIFFUSE, reflectance=(0.0, 0.0, 0.0))
    sc.add_area_light(light, (2000.0, 2000.0, 2000.0))
    sc.set_camera((0.0, 1.8, 2.2), (0.0, 0.2, 0.0), (0, 1, 0), 40.0, width, height)
    return sc


# ==================================================================== glint normal map
def scratch_bump_normalmap(res=2048, n_scratches=4096, n_bumps=None, seed=1, slope_std=0.3):
    """RGB8 normal map: Gaussian bumps (sigma 2-8 texels) scaled to a slope standard deviation
    of ``slope_std``, plus ``n_scratches`` line scratches of width 1-3 texels and uniform
    orientation (a rounded groove across the line).  Encoding n -> byte round((n/2+1/2)*255)."""
    rng = np.random.default_rng(seed)
    sx = np.zeros((res, res), np.float64)
    sy = np.zeros((res, res), np.float64)
    if n_bumps is None:
        n_bumps = max(1, (res * res) // 96)
    # --- Gaussian bumps: slope field of A*exp(-r^2/(2 sigma^2)), splatted on a 6-sigma window
    sig = rng.uniform(2.0, 8.0, n_bumps)
    cxs = rng.uniform(0, res, n_bumps)
    cys = rng.uniform(0, res, n_bumps)
    amp = rng.normal(0.0, 1.0, n_bumps) * sig
    for s_val in np.unique(np.round(sig).astype(int)):
        sel = np.nonzero(np.round(sig).astype(int) == s_val)[0]
        r = int(3 * s_val) + 1
        oy, ox = np.mgrid[-r:r + 1, -r:r + 1]
        for i in sel:
            x0, y0 = int(cxs[i]), int(cys[i])
            dx = ox + (x0 - cxs[i]) + 0.5
            dy = oy + (y0 - cys[i]) + 0.5
            g = amp[i] * np.exp(-(dx * dx + dy * dy) / (2 * sig[i] ** 2)) / (sig[i] ** 2)
            xs = np.clip(x0 + ox, 0, res - 1)
            ys = np.clip(y0 + oy, 0, res - 1)
            np.add.at(sx, (ys, xs), -g * dx)
            np.add.at(sy, (ys, xs), -g * dy)
    std = math.sqrt(float(np.mean(sx * sx + sy * sy)) / 2.0) + 1e-12
    sx *= slope_std / std
    sy *= slope_std / std
    # --- scratches: rounded grooves, slope across the line in [-a, a]
    theta = rng.uniform(0, np.pi, n_scratches)
    length = rng.uniform(32, 512, n_scratches)
    width = rng.uniform(1.0, 3.0, n_scratches)
    cx = rng.uniform(0, res, n_scratches)
    cy = rng.uniform(0, res, n_scratches)
    a = rng.uniform(0.3, 0.8, n_scratches)
    for i in range(n_scratches):
        dxl, dyl = math.cos(theta[i]), math.sin(theta[i])
        px, py = -dyl, dxl
        t = np.arange(-length[i] / 2, length[i] / 2, 0.5)
        u = np.arange(-width[i] / 2, width[i] / 2 + 1e-9, 0.5)
        T, U = np.meshgrid(t, u)
        X = np.floor(cx[i] + T * dxl + U * px).astype(np.int64)
        Y = np.floor(cy[i] + T * dyl + U * py).astype(np.int64)
        ok = (X >= 0) & (X < res) & (Y >= 0) & (Y < res)
        prof = a[i] * (2.0 * U / width[i])
        sx[Y[ok], X[ok]] = prof[ok] * px
        sy[Y[ok], X[ok]] = prof[ok] * py
    n = np.stack([-sx, -sy, np.ones_like(sx)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    rgb = np.clip(np.round((n * 0.5 + 0.5) * 255.0), 0, 255).astype(np.uint8)
    return rgb


def glint_plane(width=512, height=512, nm_res=2048, n_scratches=4096, seed=1, normalmap=None, area_light=None,
                point_light=True, eta=0.0, k=1.0):
    """1x1 normal-mapped mirror plane (conductor eta=0, k=1 -> F=1) at y=0 under a point light;
    area_light = (centre, radius, radiance) adds a spherical area light (diffuse sphere)."""
    sc = Scene()
    rgb = normalmap if normalmap is not None else scratch_bump_normalmap(nm_res, n_scratches, seed=seed)
    nm = sc.add_normalmap(rgb)
    sc.add_rectangle(plane_to_world((0, 0, 0), 0.5), abi.BSDF_NORMALMAP, flags=abi.FLAG_GLINTY,
                     eta=eta, k=k, normalmap=nm, tiles=1.0)
    if point_light:
        sc.add_point_light((0.35, 1.2, -0.3), (3.0, 3.0, 3.0))
    if area_light is not None:
        centre, radius, radiance = area_light
        light = sc.add_sphere(centre, radius, abi.BSDF_DIFFUSE)
        sc.add_area_light(light, radiance)
    sc.set_camera((0.0, 0.95, 0.9), (0.0, 0.0, 0.02), (0, 1, 0), 40.0, width, height)
    return sc
