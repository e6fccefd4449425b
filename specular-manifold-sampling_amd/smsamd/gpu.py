"""Python handle on an ``sms_ctx`` (the gfx950 library through its C ABI).

Every call goes to the HIP library; there is no CPU fallback.  Host-array methods use the
``*_host`` entry points; ``DeviceBuffer`` keeps inputs resident in HBM for timed runs.
"""
import ctypes as C

import numpy as np

from . import abi

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = abi.load_library()
    return _LIB


class SmsError(RuntimeError):
    pass


def _check(status, ctx=None, what=""):
    if status != abi.SMS_OK:
        msg = abi.STATUS_NAMES.get(status, str(status))
        if ctx is not None:
            err = lib().sms_ctx_last_error(ctx)
            msg += ": " + (err.decode() if err else "")
        raise SmsError(f"{what}: {msg}")


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = nbytes
        p = C.c_void_p()
        _check(lib().sms_device_alloc(ctx.h, nbytes, C.byref(p)), ctx.h, "sms_device_alloc")
        self.ptr = p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib().sms_copy_h2d(self.ctx.h, self.ptr, arr.ctypes.data, arr.nbytes), self.ctx.h, "h2d")

    def download(self, arr):
        assert arr.nbytes <= self.nbytes and arr.flags["C_CONTIGUOUS"]
        _check(lib().sms_copy_d2h(self.ctx.h, arr.ctypes.data, self.ptr, arr.nbytes), self.ctx.h, "d2h")
        return arr

    def free(self):
        if self.ptr:
            lib().sms_device_free(self.ctx.h, self.ptr)
            self.ptr = None


class Context:
    """sms_ctx_create(scene, cfg) on the current HIP device."""

    def __init__(self, scene, cfg):
        self.scene = scene
        self.desc = scene.desc()
        self.cfg = cfg
        h = C.c_void_p()
        _check(lib().sms_ctx_create(C.byref(self.desc), C.byref(cfg), C.byref(h)), None, "sms_ctx_create")
        self.h = h.value

    def close(self):
        if self.h:
            lib().sms_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, cfg):
        self.cfg = cfg
        _check(lib().sms_ctx_set_config(self.h, C.byref(cfg)), self.h, "set_config")

    def set_schedule(self, **kw):
        """Change some of the sms_schedule fields (abi.SmsSchedule.KEYS); the others keep their
        current values."""
        cur = self.get_schedule()
        unknown = set(kw) - set(cur)
        if unknown:
            raise ValueError(f"unknown schedule keys {sorted(unknown)}")
        cur.update(kw)
        s = abi.SmsSchedule.from_dict(cur)
        _check(lib().sms_ctx_set_schedule(self.h, C.byref(s)), self.h, "set_schedule")

    def get_schedule(self):
        s = abi.SmsSchedule()
        _check(lib().sms_ctx_get_schedule(self.h, C.byref(s)), self.h, "get_schedule")
        return s.as_dict()

    def device_info(self):
        name = C.create_string_buffer(256)
        ncu = C.c_int32()
        _check(lib().sms_ctx_device_info(self.h, name, 256, C.byref(ncu)), self.h, "device_info")
        return name.value.decode(), ncu.value

    def sync(self):
        _check(lib().sms_ctx_sync(self.h), self.h, "sync")

    def last_kernel_ms(self):
        ms = (C.c_float * 2)()
        _check(lib().sms_ctx_last_kernel_ms(self.h, ms), self.h, "last_kernel_ms")
        return float(ms[0]), float(ms[1])

    # ---------------------------------------------------------------- host-buffer API
    def ss_walks(self, win):
        out = np.zeros(len(win), abi.SS_WALK_OUT)
        _check(lib().sms_ss_walk_batch_host(self.h, win.ctypes.data, len(win), out.ctypes.data), self.h, "ss_walks")
        return out

    def glint_walks(self, win):
        out = np.zeros(len(win), abi.GLINT_WALK_OUT)
        _check(lib().sms_glint_walk_batch_host(self.h, win.ctypes.data, len(win), out.ctypes.data), self.h,
               "glint_walks")
        return out

    def ss_estimate(self, q, seed):
        out = np.zeros((len(q), 3), np.float32)
        st = abi.SmsStats()
        _check(lib().sms_ss_estimate_batch_host(self.h, q.ctypes.data, len(q), seed, out.ctypes.data, C.byref(st)),
               self.h, "ss_estimate")
        return out, st.as_dict()

    def glint_estimate(self, q, seed):
        out = np.zeros((len(q), 3), np.float32)
        st = abi.SmsStats()
        _check(lib().sms_glint_estimate_batch_host(self.h, q.ctypes.data, len(q), seed, out.ctypes.data,
                                                   C.byref(st)), self.h, "glint_estimate")
        return out, st.as_dict()

    def ms_walks(self, win):
        out = np.zeros(len(win), abi.MS_WALK_OUT)
        _check(lib().sms_ms_walk_batch_host(self.h, win.ctypes.data, len(win), out.ctypes.data), self.h, "ms_walks")
        return out

    def ms_estimate(self, q, seed):
        out = np.zeros((len(q), 3), np.float32)
        st = abi.SmsStats()
        _check(lib().sms_ms_estimate_batch_host(self.h, q.ctypes.data, len(q), seed, out.ctypes.data, C.byref(st)),
               self.h, "ms_estimate")
        return out, st.as_dict()

    def glint_queries(self, spp, seed):
        cam = self.scene.camera
        n = cam["width"] * cam["height"] * spp
        buf = DeviceBuffer(self, n * abi.GLINT_QUERY.itemsize)
        try:
            _check(lib().sms_glint_make_queries(self.h, spp, seed, buf.ptr), self.h, "glint_queries")
            q = np.zeros(n, abi.GLINT_QUERY)
            buf.download(q)
        finally:
            buf.free()
        return q

    # ---------------------------------------------------------------- device-resident API
    def glint_queries_device(self, spp, seed):
        cam = self.scene.camera
        n = cam["width"] * cam["height"] * spp
        buf = DeviceBuffer(self, n * abi.GLINT_QUERY.itemsize)
        _check(lib().sms_glint_make_queries(self.h, spp, seed, buf.ptr), self.h, "glint_queries")
        self.sync()
        return buf, n

    def glint_estimate_device(self, qbuf, n, seed, outbuf):
        st = abi.SmsStats()
        _check(lib().sms_glint_estimate_batch(self.h, qbuf.ptr, n, seed, outbuf.ptr, C.byref(st)), self.h,
               "glint_estimate")
        return st.as_dict()

    def glint_autotune_device(self, qbuf, n, seed, outbuf):
        """Pick the fastest glint schedule on these queries (sms_glint_autotune); returns it."""
        s = abi.SmsSchedule()
        _check(lib().sms_glint_autotune(self.h, qbuf.ptr, n, seed, outbuf.ptr, C.byref(s)), self.h, "glint_autotune")
        return s.as_dict()

    def ss_estimate_device(self, qbuf, n, seed, outbuf):
        st = abi.SmsStats()
        _check(lib().sms_ss_estimate_batch(self.h, qbuf.ptr, n, seed, outbuf.ptr, C.byref(st)), self.h,
               "ss_estimate")
        return st.as_dict()
