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
        self.normalmaps = []    # (res, uint8 array)
        self.camera = None
        self._keep = []

    # ---------------------------------------------------------------- builders
    def add_sphere(self, center, radius, bsdf, flags=0, eta=1.5, k=0.0, reflectance=(0.5, 0.5, 0.5),
                   emitter=-1):
        self.shapes.append(dict(type=abi.SHAPE_SPHERE, bsdf=bsdf, flags=flags, center=center,
                                radius=radius, eta=eta, k=k, reflectance=reflectance,
                                emitter=emitter))
        return len(self.shapes) - 1

    def add_rectangle(self, to_world, bsdf, flags=0, eta=1.5, k=0.0, reflectance=(0.5, 0.5, 0.5),
                      normalmap=-1, tiles=1.0):
        self.shapes.append(dict(type=abi.SHAPE_RECTANGLE, bsdf=bsdf, flags=flags,
                                to_world=np.asarray(to_world, np.float32).reshape(12), eta=eta, k=k,
                                reflectance=reflectance, normalmap=normalmap, tiles=tiles))
        return len(self.shapes) - 1

    def add_mesh(self, positions, faces, normals=None, texcoords=None, bsdf=abi.BSDF_DIFFUSE, flags=0,
                 eta=1.5, k=0.0, reflectance=(0.5, 0.5, 0.5)):
        self.shapes.append(dict(type=abi.SHAPE_MESH, bsdf=bsdf, flags=flags, eta=eta, k=k,
                                reflectance=reflectance,
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

    def add_normalmap(self, rgb):
        rgb = np.ascontiguousarray(rgb, np.uint8)
        assert rgb.ndim == 3 and rgb.shape[0] == rgb.shape[1] and rgb.shape[2] == 3
        self.normalmaps.append((rgb.shape[0], rgb))
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
            d.caustic_ss, d.caustic_ms = int(e["caustic_ss"]), int(e["caustic_ms"])
        nms = (abi.SmsNormalmapDesc * max(1, len(self.normalmaps)))()
        for i, (res, rgb) in enumerate(self.normalmaps):
            nms[i].res = res
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


# ==================================================================== sphere caustic
def sphere_caustic(width=256, height=256, center_y=1.5):
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
    light = sc.add_sphere((0.0, 5.0, 0.0), 0.1, abi.BSDF_DIFFUSE, reflectance=(0.0, 0.0, 0.0))
    sc.add_area_light(light, (100.0, 100.0, 100.0))
    sc.set_camera((0, 3, 6), (0, 0.5, 0), (0, 1, 0), 45.0, width, height)
    return sc


def sphere_dome(width=256, height=256):
    """The sphere-caustic scene with the sphere centre on the receiver plane (a glass dome)."""
    return sphere_caustic(width, height, center_y=0.0)


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


def glint_plane(width=512, height=512, nm_res=2048, n_scratches=4096, seed=1, normalmap=None):
    """1x1 normal-mapped mirror plane (conductor eta=0, k=1 -> F=1) at y=0 under a point light."""
    sc = Scene()
    rgb = normalmap if normalmap is not None else scratch_bump_normalmap(nm_res, n_scratches, seed=seed)
    nm = sc.add_normalmap(rgb)
    sc.add_rectangle(plane_to_world((0, 0, 0), 0.5), abi.BSDF_NORMALMAP, flags=abi.FLAG_GLINTY,
                     eta=0.0, k=1.0, normalmap=nm, tiles=1.0)
    sc.add_point_light((0.35, 1.2, -0.3), (3.0, 3.0, 3.0))
    sc.set_camera((0.0, 0.95, 0.9), (0.0, 0.0, 0.02), (0, 1, 0), 40.0, width, height)
    return sc
