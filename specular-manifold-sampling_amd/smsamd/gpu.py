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

    def ss_queries(self, spp, seed):
        """Camera pass of the single-scatter integrator: (queries, path weights) on the host."""
        qb, wb, n = self.ss_queries_device(spp, seed)
        try:
            q = np.zeros(n, abi.SS_QUERY)
            w = np.zeros(n, np.float32)
            qb.download(q)
            wb.download(w)
        finally:
            qb.free()
            wb.free()
        return q, w

    # ---------------------------------------------------------------- device-resident API
    def ss_queries_device(self, spp, seed, qbuf=None, wbuf=None):
        cam = self.scene.camera
        n = cam["width"] * cam["height"] * spp
        qb = qbuf or DeviceBuffer(self, n * abi.SS_QUERY.itemsize)
        wb = wbuf or DeviceBuffer(self, n * 4)
        _check(lib().sms_ss_make_queries(self.h, spp, seed, qb.ptr, wb.ptr), self.h, "ss_queries")
        return qb, wb, n

    def film_accumulate_device(self, rgbbuf, wbuf, pixels, spp, imgbuf, imgsqbuf=None, imgnbuf=None):
        _check(lib().sms_film_accumulate(self.h, rgbbuf.ptr, wbuf.ptr, pixels, spp, imgbuf.ptr,
                                         imgsqbuf.ptr if imgsqbuf is not None else None,
                                         imgnbuf.ptr if imgnbuf is not None else None), self.h, "film_accumulate")

    # ---------------------------------------------------------------- path integrators
    def path_samples_device(self, pc, spp, seed, rgbbuf, posbuf):
        """One pass of the path integrator pc into device buffers (n x 3 radiance, n x 2 film
        positions, n = w*h*spp); returns the pass's SMS statistics."""
        st = abi.SmsStats()
        _check(lib().sms_path_samples(self.h, C.byref(pc), spp, seed, rgbbuf.ptr, posbuf.ptr, C.byref(st)), self.h,
               "path_samples")
        return st.as_dict()

    def path_samples(self, pc, spp, seed):
        cam = self.scene.camera
        n = cam["width"] * cam["height"] * spp
        rgb = np.zeros((n, 3), np.float32)
        pos = np.zeros((n, 2), np.float32)
        bufs = [DeviceBuffer(self, rgb.nbytes), DeviceBuffer(self, pos.nbytes)]
        try:
            st = self.path_samples_device(pc, spp, seed, bufs[0], bufs[1])
            bufs[0].download(rgb)
            bufs[1].download(pos)
        finally:
            for b in bufs:
                b.free()
        return rgb, pos, st

    def film_splat_device(self, rgbbuf, posbuf, spp, stddev, imgbuf, wbuf):
        _check(lib().sms_film_splat(self.h, rgbbuf.ptr, posbuf.ptr, spp, stddev, imgbuf.ptr, wbuf.ptr), self.h,
               "film_splat")

    def film_splat(self, rgb, pos, spp, stddev):
        cam = self.scene.camera
        px = cam["width"] * cam["height"]
        image = np.zeros((px, 3), np.float32)
        weight = np.zeros(px, np.float32)
        rgb = np.ascontiguousarray(rgb, np.float32)
        pos = np.ascontiguousarray(pos, np.float32)
        bufs = [DeviceBuffer(self, rgb.nbytes), DeviceBuffer(self, pos.nbytes), DeviceBuffer(self, image.nbytes),
                DeviceBuffer(self, weight.nbytes)]
        try:
            bufs[0].upload(rgb)
            bufs[1].upload(pos)
            bufs[2].upload(image)
            bufs[3].upload(weight)
            self.film_splat_device(bufs[0], bufs[1], spp, stddev, bufs[2], bufs[3])
            bufs[2].download(image)
            bufs[3].download(weight)
        finally:
            for b in bufs:
                b.free()
        return image, weight

    def glint_queries_device(self, spp, seed, qbuf=None):
        """Camera pass of a glint frame into device memory (qbuf, or a new buffer): (buffer, n)."""
        cam = self.scene.camera
        n = cam["width"] * cam["height"] * spp
        buf = qbuf if qbuf is not None else DeviceBuffer(self, n * abi.GLINT_QUERY.itemsize)
        _check(lib().sms_glint_make_queries(self.h, spp, seed, buf.ptr), self.h, "glint_queries")
        if qbuf is None:
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

    def ss_autotune_device(self, qbuf, n, seed, outbuf):
        """Pick the fastest unbiased single-scatter schedule on these queries (sms_ss_autotune)."""
        s = abi.SmsSchedule()
        _check(lib().sms_ss_autotune(self.h, qbuf.ptr, n, seed, outbuf.ptr, C.byref(s)), self.h, "ss_autotune")
        return s.as_dict()

    def ms_autotune_device(self, qbuf, n, seed, outbuf):
        """Pick the fastest unbiased multi-scatter schedule on these queries (sms_ms_autotune)."""
        s = abi.SmsSchedule()
        _check(lib().sms_ms_autotune(self.h, qbuf.ptr, n, seed, outbuf.ptr, C.byref(s)), self.h, "ms_autotune")
        return s.as_dict()

    def ss_estimate_device(self, qbuf, n, seed, outbuf):
        st = abi.SmsStats()
        _check(lib().sms_ss_estimate_batch(self.h, qbuf.ptr, n, seed, outbuf.ptr, C.byref(st)), self.h,
               "ss_estimate")
        return st.as_dict()

    def ms_estimate_device(self, qbuf, n, seed, outbuf):
        st = abi.SmsStats()
        _check(lib().sms_ms_estimate_batch(self.h, qbuf.ptr, n, seed, outbuf.ptr, C.byref(st)), self.h,
               "ms_estimate")
        return st.as_dict()


    # ---------------------------------------------------------------- learned seed proposal
    def set_proposal(self, mlp, shape, defensive=0.2):
        _check(lib().sms_ctx_set_proposal(self.h, mlp.h if mlp is not None else None, shape, defensive), self.h,
               "sms_ctx_set_proposal")

    def set_proposal_anchor(self, radius):
        """Anchored single-scatter mixtures (uv box of half side `radius` around each query's
        anchor); 0: mixtures over the caster's uv square."""
        _check(lib().sms_ctx_set_proposal_anchor(self.h, radius), self.h, "sms_ctx_set_proposal_anchor")

    def ss_proposal_anchors(self, queries, seed):
        """Anchors (n x 2, NaN: none) of single-scatter queries (structured array) for an
        estimate with `seed`."""
        n = len(queries)
        qb = DeviceBuffer(self, max(queries.nbytes, 4))
        ab = DeviceBuffer(self, max(n * 8, 4))
        try:
            qb.upload(queries)
            _check(lib().sms_ss_proposal_anchors(self.h, qb.ptr, n, seed, ab.ptr), self.h, "sms_ss_proposal_anchors")
            a = np.zeros((n, 2), np.float32)
            if n:
                ab.download(a)
        finally:
            qb.free()
            ab.free()
        return a

    def set_training_capture(self, capacity, mode=abi.CAPTURE_SOLUTIONS):
        _check(lib().sms_ctx_set_training_capture(self.h, capacity, mode), self.h, "sms_ctx_set_training_capture")

    def training_count(self):
        n = C.c_int64(0)
        _check(lib().sms_ctx_training_samples(self.h, C.byref(n), None, None, None), self.h,
               "sms_ctx_training_samples")
        return int(n.value)

    def training_samples_device(self):
        """(x buffer, uv buffer, weight buffer, count) of the captured samples, in device memory."""
        n = self.training_count()
        xb = DeviceBuffer(self, max(n * abi.MLP_INPUTS * 4, 4))
        ub = DeviceBuffer(self, max(n * 8, 4))
        wb = DeviceBuffer(self, max(n * 4, 4))
        m = C.c_int64(0)
        _check(lib().sms_ctx_training_samples(self.h, C.byref(m), xb.ptr, ub.ptr, wb.ptr), self.h,
               "sms_ctx_training_samples")
        assert m.value == n
        return xb, ub, wb, n

    def training_samples(self):
        xb, ub, wb, n = self.training_samples_device()
        try:
            x = np.zeros((n, abi.MLP_INPUTS), np.float32)
            uv = np.zeros((n, 2), np.float32)
            w = np.zeros(n, np.float32)
            if n:
                xb.download(x)
                ub.download(uv)
                wb.download(w)
        finally:
            xb.free()
            ub.free()
            wb.free()
        return x, uv, w

    def clear_training(self):
        _check(lib().sms_ctx_clear_training(self.h), self.h, "sms_ctx_clear_training")

    def _per_query(self, fn, q, seed, width, what):
        n = len(q)
        qb = DeviceBuffer(self, max(q.nbytes, 4))
        ob = DeviceBuffer(self, max(n * width * 4, 4))
        try:
            qb.upload(q)
            _check(fn(self.h, qb.ptr, n, seed, ob.ptr), self.h, what)
            out = np.zeros((n, width), np.float32)
            ob.download(out)
        finally:
            qb.free()
            ob.free()
        return out

    def _graph(self, fn, *args):
        h = C.c_void_p()
        _check(fn(self.h, *args, C.byref(h)), self.h, fn.__name__)
        return Graph(self, h.value)

    def glint_graph(self, qbuf, n, outbuf):
        return self._graph(lib().sms_glint_graph_create, qbuf.ptr, n, outbuf.ptr)

    def glint_frame_graph(self, spp, outbuf):
        """Graph of the whole glint frame (camera queries generated where they are read)."""
        return self._graph(lib().sms_glint_frame_graph_create, spp, outbuf.ptr)

    def glint_estimate_frame_device(self, spp, seed, outbuf):
        st = abi.SmsStats()
        _check(lib().sms_glint_estimate_frame(self.h, spp, seed, outbuf.ptr, C.byref(st)), self.h,
               "glint_estimate_frame")
        return st.as_dict()

    def glint_estimate_frame(self, spp, seed):
        cam = self.scene.camera
        n = cam["width"] * cam["height"] * spp
        out = np.zeros((n, 3), np.float32)
        buf = DeviceBuffer(self, out.nbytes)
        try:
            st = self.glint_estimate_frame_device(spp, seed, buf)
            buf.download(out)
        finally:
            buf.free()
        return out, st

    def ss_graph(self, qbuf, n, outbuf):
        return self._graph(lib().sms_ss_graph_create, qbuf.ptr, n, outbuf.ptr)

    def ms_graph(self, qbuf, n, outbuf):
        return self._graph(lib().sms_ms_graph_create, qbuf.ptr, n, outbuf.ptr)

    def ss_pass_graph(self, spp, qbuf, wbuf, rgbbuf, imgbuf, imgsqbuf=None, imgnbuf=None):
        return self._graph(lib().sms_ss_pass_graph_create, spp, qbuf.ptr, wbuf.ptr, rgbbuf.ptr, imgbuf.ptr,
                           imgsqbuf.ptr if imgsqbuf is not None else None,
                           imgnbuf.ptr if imgnbuf is not None else None)

    def proposal_inputs(self, q, seed):
        return self._per_query(lib().sms_ss_proposal_inputs, q, seed, abi.MLP_INPUTS, "sms_ss_proposal_inputs")

    def proposal_outputs(self, q, seed):
        return self._per_query(lib().sms_ss_proposal_outputs, q, seed, abi.MLP_OUTPUTS, "sms_ss_proposal_outputs")

    def ms_proposal_inputs(self, q, seed):
        return self._per_query(lib().sms_ms_proposal_inputs, q, seed, abi.MLP_INPUTS, "sms_ms_proposal_inputs")

    def ms_proposal_outputs(self, q, seed):
        return self._per_query(lib().sms_ms_proposal_outputs, q, seed, abi.MLP_OUTPUTS, "sms_ms_proposal_outputs")

    def glint_proposal_inputs(self, q, seed):
        return self._per_query(lib().sms_glint_proposal_inputs, q, seed, abi.MLP_INPUTS, "sms_glint_proposal_inputs")

    def glint_proposal_outputs(self, q, seed):
        return self._per_query(lib().sms_glint_proposal_outputs, q, seed, abi.MLP_OUTPUTS,
                               "sms_glint_proposal_outputs")


class Graph:
    """A captured frame launch sequence (sms_graph_*): launch(seed) replays it asynchronously."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h

    def launch(self, seed):
        _check(lib().sms_graph_launch(self.h, seed), self.ctx.h, "sms_graph_launch")

    def stats(self, reset=False):
        st = abi.SmsStats()
        _check(lib().sms_graph_stats(self.h, C.byref(st), 1 if reset else 0), self.ctx.h, "sms_graph_stats")
        return st.as_dict()

    def close(self):
        if self.h:
            lib().sms_graph_destroy(self.h)
            self.h = None


class Mlp:
    """The learned seed proposal's network (sms_mlp_*) on a context's device."""

    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib().sms_mlp_create(ctx.h, seed, C.byref(h)), ctx.h, "sms_mlp_create")
        self.h = h.value

    def close(self):
        if self.h:
            lib().sms_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(abi.MLP_PARAMS, np.float32)
        _check(lib().sms_mlp_get_params(self.h, p.ctypes.data), self.ctx.h, "sms_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (abi.MLP_PARAMS,)
        _check(lib().sms_mlp_set_params(self.h, p.ctypes.data), self.ctx.h, "sms_mlp_set_params")

    def infer_device(self, xbuf, n, outbuf):
        _check(lib().sms_mlp_infer(self.h, xbuf.ptr, n, outbuf.ptr), self.ctx.h, "sms_mlp_infer")

    def infer(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, abi.MLP_INPUTS)
        n = len(x)
        xb = DeviceBuffer(self.ctx, max(x.nbytes, 4))
        ob = DeviceBuffer(self.ctx, max(n * abi.MLP_OUTPUTS * 4, 4))
        try:
            xb.upload(x)
            self.infer_device(xb, n, ob)
            out = np.zeros((n, abi.MLP_OUTPUTS), np.float32)
            ob.download(out)
        finally:
            xb.free()
            ob.free()
        return out

    def grad_device(self, xbuf, uvbuf, n, want_grad=False, want_loss=True, offset=0, wbuf=None):
        """Gradient of the (weighted) mean mixture NLL over samples [offset, offset + n) of device
        buffers; wbuf None: unit weights."""
        g = np.zeros(abi.MLP_PARAMS, np.float32) if want_grad else None
        loss = C.c_float(0.0)
        _check(lib().sms_mlp_grad(self.h, xbuf.ptr + 4 * abi.MLP_INPUTS * offset, uvbuf.ptr + 8 * offset,
                                  wbuf.ptr + 4 * offset if wbuf is not None else None, n,
                                  g.ctypes.data if want_grad else None, C.byref(loss) if want_loss else None),
               self.ctx.h, "sms_mlp_grad")
        return (float(loss.value) if want_loss else None), g

    def grad(self, x, uv, w=None):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, abi.MLP_INPUTS)
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        bufs = [DeviceBuffer(self.ctx, x.nbytes), DeviceBuffer(self.ctx, uv.nbytes)]
        try:
            bufs[0].upload(x)
            bufs[1].upload(uv)
            if w is not None:
                w = np.ascontiguousarray(w, np.float32).reshape(-1)
                bufs.append(DeviceBuffer(self.ctx, w.nbytes))
                bufs[2].upload(w)
            return self.grad_device(bufs[0], bufs[1], len(x), want_grad=True, wbuf=bufs[2] if w is not None else None)
        finally:
            for b in bufs:
                b.free()

    def adam(self, lr):
        _check(lib().sms_mlp_adam(self.h, lr), self.ctx.h, "sms_mlp_adam")

    def train_device(self, xbuf, uvbuf, records, batch, offsets, lr, wbuf=None, want_losses=False):
        """len(offsets) training steps (gradient over the batch window at each offset of the
        device sample buffers, then Adam), queued without host synchronisation; returns the
        per-step losses if asked (then waits)."""
        offs = np.ascontiguousarray(offsets, np.int64)
        if len(offs) and (offs.min() < 0 or offs.max() + batch > records):
            raise ValueError("batch window outside the sample buffers")
        for b, per in ((xbuf, 4 * abi.MLP_INPUTS), (uvbuf, 8), (wbuf, 4)):
            if b is not None and b.nbytes < records * per:
                raise ValueError("sample buffer smaller than the record count")
        losses = np.zeros(len(offs), np.float32) if want_losses else None
        _check(lib().sms_mlp_train(self.h, xbuf.ptr, uvbuf.ptr, wbuf.ptr if wbuf is not None else None, batch,
                                   offs.ctypes.data if len(offs) else None, len(offs), lr,
                                   losses.ctypes.data if want_losses else None), self.ctx.h, "sms_mlp_train")
        return losses

