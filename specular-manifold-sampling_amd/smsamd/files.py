"""Scene and asset files through the library's host-side readers (sms_obj_load, sms_image_load,
sms_scene_load): Wavefront OBJ meshes, OpenEXR / PNG images and Mitsuba 2 scene files, as
numpy arrays and as a scenes.Scene the rest of the harness renders.  Nothing here touches the
GPU."""
import ctypes as C

import numpy as np

from . import abi, scenes

_ERR_LEN = 1024


def _lib():
    from . import gpu
    return gpu.lib()


class LoadError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"status {status}: {message}")
        self.status = status


def _arr(ptr, n, dtype):
    if not ptr or n == 0:
        return None
    ct = {np.float32: C.c_float, np.uint32: C.c_uint32}[dtype]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()


def load_obj(path, to_world=None, flip_tex_coords=True, face_normals=False):
    """dict(positions (n,3), normals (n,3) or None, texcoords (n,2) or None, faces (m,3))."""
    err = C.create_string_buffer(_ERR_LEN)
    out = C.POINTER(abi.SmsMeshData)()
    tw = None if to_world is None else np.ascontiguousarray(to_world, np.float32).reshape(16)
    flags = (abi.OBJ_FLIP_TEXCOORDS if flip_tex_coords else 0) | (abi.OBJ_FACE_NORMALS if face_normals else 0)
    st = _lib().sms_obj_load(str(path).encode(), None if tw is None else tw.ctypes.data, flags, C.byref(out), err,
                             _ERR_LEN)
    if st:
        raise LoadError(st, err.value.decode())
    m = out.contents
    try:
        nv, nf = int(m.n_vertices), int(m.n_faces)
        r = {"positions": _arr(m.positions, 3 * nv, np.float32).reshape(nv, 3),
             "normals": None, "texcoords": None,
             "faces": _arr(m.faces, 3 * nf, np.uint32).reshape(nf, 3)}
        if m.normals:
            r["normals"] = _arr(m.normals, 3 * nv, np.float32).reshape(nv, 3)
        if m.texcoords:
            r["texcoords"] = _arr(m.texcoords, 2 * nv, np.float32).reshape(nv, 2)
        return r
    finally:
        _lib().sms_mesh_data_free(out)


def load_image(path):
    """(height, width, 3) float32 RGB."""
    err = C.create_string_buffer(_ERR_LEN)
    out = C.POINTER(abi.SmsImage)()
    st = _lib().sms_image_load(str(path).encode(), C.byref(out), err, _ERR_LEN)
    if st:
        raise LoadError(st, err.value.decode())
    im = out.contents
    try:
        return _arr(im.pixels, im.width * im.height * 3, np.float32).reshape(im.height, im.width, 3)
    finally:
        _lib().sms_image_free(out)


def load_scene(path, defines=None, substitute=False):
    """(scenes.Scene, info) of a Mitsuba 2 scene file; info = dict(integrator, properties
    {name: value}, spp, filter_gaussian, filter_stddev, notes)."""
    lib = _lib()
    defs = [f"{k}={v}".encode() for k, v in (defines or {}).items()]
    arr = (C.c_char_p * max(1, len(defs)))(*defs)
    err = C.create_string_buffer(_ERR_LEN)
    h = C.c_void_p()
    st = lib.sms_scene_load(str(path).encode(), arr, len(defs), abi.LOAD_SUBSTITUTE if substitute else 0,
                            C.byref(h), err, _ERR_LEN)
    if st:
        raise LoadError(st, err.value.decode())
    try:
        d = lib.sms_scene_file_desc(h).contents
        sc = scenes.Scene()
        for i in range(d.n_shapes):
            s = d.shapes[i]
            sh = dict(type=s.type, bsdf=s.bsdf, flags=s.flags, emitter=s.emitter, center=tuple(s.center),
                      radius=s.radius, to_world=np.array(s.to_world[:], np.float32), eta=s.eta, k=s.k,
                      reflectance=tuple(s.reflectance), normalmap=s.normalmap, tiles=s.tiles, roughness=s.roughness)
            if s.type == abi.SHAPE_MESH:
                nv, nf = s.n_vertices, s.n_faces
                sh["positions"] = _arr(s.positions, 3 * nv, np.float32).reshape(nv, 3)
                sh["faces"] = _arr(s.faces, 3 * nf, np.uint32).reshape(nf, 3)
                sh["normals"] = _arr(s.normals, 3 * nv, np.float32).reshape(nv, 3) if s.normals else None
                sh["texcoords"] = _arr(s.texcoords, 2 * nv, np.float32).reshape(nv, 2) if s.texcoords else None
            sc.shapes.append(sh)
        for i in range(d.n_emitters):
            e = d.emitters[i]
            sc.emitters.append(dict(type=e.type, shape=e.shape, radiance=tuple(e.radiance), position=tuple(e.position),
                                    intensity=tuple(e.intensity), direction=tuple(e.direction),
                                    irradiance=tuple(e.irradiance), caustic_ss=bool(e.caustic_ss),
                                    caustic_ms=bool(e.caustic_ms)))
        for i in range(d.n_normalmaps):
            nm = d.normalmaps[i]
            r = nm.res
            if nm.rgb:
                rgb = np.ctypeslib.as_array(C.cast(nm.rgb, C.POINTER(C.c_uint8)), shape=(r * r * 3,)).copy()
            else:
                rgb = _arr(nm.rgb32f, r * r * 3, np.float32)
            sc.normalmaps.append((r, rgb.reshape(r, r, 3)))
            if nm.encoded_normals:
                sc.encoded_normalmaps.add(i)
        c = d.camera
        sc.set_camera(tuple(c.origin), tuple(c.target), tuple(c.up), c.fov_deg, c.width, c.height)
        info = abi.SmsSceneInfo()
        lib.sms_scene_file_info(h, C.byref(info))
        props = {}
        for i in range(info.n_properties):
            p = abi.SmsProperty()
            lib.sms_scene_file_property(h, i, C.byref(p))
            v = {abi.PROP_BOOL: lambda: bool(p.value), abi.PROP_INT: lambda: int(p.value),
                 abi.PROP_FLOAT: lambda: float(p.value), abi.PROP_STRING: lambda: p.text.decode()}[p.type]()
            props[p.name.decode()] = v
        notes = [lib.sms_scene_file_note(h, i).decode() for i in range(info.n_notes)]
        return sc, {"integrator": info.integrator.decode(), "properties": props, "spp": info.spp,
                    "filter_gaussian": bool(info.filter_gaussian), "filter_stddev": info.filter_stddev,
                    "notes": notes}
    finally:
        lib.sms_scene_file_free(h)
