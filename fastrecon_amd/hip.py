"""ctypes binding of libfastrecon_hip.so (C ABI in include/fastrecon_hip.h).

This is plumbing for Python callers (tests, bench): every compute call goes
through the native library.  There is no Python or CPU fallback: if the
library is missing, loading it raises.

A process that also uses PyTorch's GPU side must import torch before loading
this library: torch ships its own HIP and HSA runtime libraries under the same
sonames as ROCm's, the copy loaded first serves both, and torch finds no
device when ROCm's copy came first (bench.py and fastrecon_amd.ldcf.evaluate
follow this order).
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfastrecon_hip.so")

FR_SAMPLES_C64, FR_SAMPLES_C16 = 0, 1
FR_LAYOUT_COIL_MAJOR, FR_LAYOUT_SAMPLE_MAJOR = 0, 1

# every symbol the header declares (checked by the CPU test suite)
EXPORTS = [
    "fr_kb_lut", "fr_traj_normalize", "fr_traj_transform",
    "fr_ctx_create", "fr_ctx_destroy", "fr_last_error", "fr_ctx_synchronize", "fr_ctx_stream",
    "fr_samples_create", "fr_samples_destroy", "fr_samples_set_traj", "fr_samples_set_dcf",
    "fr_samples_set_coil", "fr_samples_update", "fr_samples_set_data_device", "fr_samples_set_dcf_device",
    "fr_samples_set_traj_device", "fr_samples_get_dcf",
    "fr_plan_create", "fr_plan_destroy", "fr_plan_get_info",
    "fr_grids_create", "fr_grids_create_frames", "fr_samples_create_frames", "fr_grids_destroy", "fr_grids_download", "fr_grids_upload", "fr_grids_device_ptr",
    "fr_grid_execute", "fr_grid_execute_noalloc", "fr_degrid_execute",
    "fr_fft_create", "fr_fft_destroy", "fr_fft_execute",
    "fr_fftshift", "fr_sos", "fr_fft_sos", "fr_deapod_map", "fr_lowfilter", "fr_normalize", "fr_crop", "fr_sos_map",
    "fr_grids_copy",
    "fr_dcf_pipe_menon",
    "fr_dcf_identity_scale",
    "fr_graph_create", "fr_graph_launch", "fr_graph_destroy", "fr_autotune",
    "fr_tune_key_of", "fr_tune_cache_lookup", "fr_tune_cache_store", "fr_autotune_search",
    "fr_ldcf_features", "fr_mlp_param_count", "fr_mlp_create", "fr_mlp_destroy", "fr_mlp_set_params",
    "fr_mlp_get_params", "fr_mlp_set_base", "fr_mlp_get_base", "fr_mlp_forward", "fr_mlp_dcf", "fr_ldcf_dcf",
    "fr_mlp_grad", "fr_mlp_train_step",
    "fr_malloc", "fr_free", "fr_memcpy_d2h", "fr_memcpy_h2d",
]


class HipError(RuntimeError):
    pass


class fr_tune(C.Structure):
    _fields_ = [("tile_x", C.c_int), ("tile_y", C.c_int), ("tile_z", C.c_int), ("slabs", C.c_int),
                ("accumulate", C.c_int), ("split_taps", C.c_int),
                ("exact_order", C.c_int), ("items_per_wave", C.c_int), ("xcd_run", C.c_int),
                ("stores", C.c_int), ("waves_per_simd", C.c_int), ("tile_writes", C.c_int)]


class fr_tune_key(C.Structure):
    _fields_ = [("arch", C.c_char * 32), ("cus", C.c_int), ("dim", C.c_int), ("N", C.c_uint), ("M", C.c_int64),
                ("C", C.c_int), ("type", C.c_int), ("frames", C.c_int), ("kwidth", C.c_float), ("L", C.c_int)]


class fr_plan_info(C.Structure):
    _fields_ = [("samples", C.c_int64), ("taps", C.c_int64), ("records", C.c_int64), ("streams", C.c_int64),
                ("tile_x", C.c_int), ("tile_y", C.c_int), ("tile_z", C.c_int), ("slabs", C.c_int),
                ("tiles", C.c_int), ("lds_bytes", C.c_int), ("plan_bytes", C.c_int64), ("build_ms", C.c_double),
                ("items", C.c_int64), ("heavy_units", C.c_int64), ("max_unit_taps", C.c_int64),
                ("exact_order", C.c_int), ("accumulate", C.c_int), ("scratch_bytes", C.c_int64)]


_lib = None
_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)


def load():
    """Load the native library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipError(f"{LIB_PATH} is not built (run fastrecon_amd/build.py); there is no fallback")
    L = C.CDLL(LIB_PATH)
    sig = {
        "fr_kb_lut": [C.c_float, C.c_float, C.c_int, _vp, _vp],
        "fr_traj_normalize": [_vp, C.c_int64, C.c_int, C.c_int, C.c_uint, _vp],
        "fr_traj_transform": [_vp, C.c_int64, C.c_int, C.c_int, C.c_float, C.c_float],
        "fr_ctx_create": [C.c_int, _pp],
        "fr_ctx_destroy": [_vp],
        "fr_ctx_synchronize": [_vp],
        "fr_samples_create": [_vp, C.c_int, C.c_int64, C.c_int, C.c_int, _pp],
        "fr_samples_destroy": [_vp],
        "fr_samples_set_traj": [_vp, _vp, C.c_int],
        "fr_samples_set_dcf": [_vp, _vp],
        "fr_samples_set_coil": [_vp, C.c_int, _vp],
        "fr_samples_update": [_vp, C.c_int, C.c_int64, C.c_int64, _vp],
        "fr_samples_set_data_device": [_vp, _vp, C.c_int],
        "fr_samples_set_dcf_device": [_vp, _vp],
        "fr_samples_set_traj_device": [_vp, _vp, C.c_int],
        "fr_samples_get_dcf": [_vp, _vp],
        "fr_plan_create": [_vp, _vp, C.c_uint, C.c_float, _vp, C.c_int, _vp, _pp],
        "fr_plan_destroy": [_vp],
        "fr_plan_get_info": [_vp, C.POINTER(fr_plan_info)],
        "fr_grids_create": [_vp, C.c_int, C.c_uint, C.c_int, _pp],
        "fr_grids_create_frames": [_vp, C.c_uint, C.c_int, C.c_int, _pp],
        "fr_samples_create_frames": [_vp, C.c_int, _vp, C.c_int, C.c_int, _pp],
        "fr_graph_create": [_vp, _vp, _vp, _vp, C.c_uint, C.c_uint, C.c_uint, C.c_int, _vp, _vp, _pp],
        "fr_graph_launch": [_vp],
        "fr_autotune": [_vp, _vp, C.c_uint, C.c_float, _vp, C.c_int, _vp, C.c_int, C.c_int, C.POINTER(fr_tune),
                        C.POINTER(C.c_float), _vp],
        "fr_tune_key_of": [_vp, _vp, C.c_uint, C.c_float, C.c_int, C.POINTER(fr_tune_key)],
        "fr_tune_cache_lookup": [C.c_char_p, C.POINTER(fr_tune_key), C.POINTER(fr_tune), C.POINTER(C.c_float),
                                 C.POINTER(C.c_float), C.POINTER(C.c_int)],
        "fr_tune_cache_store": [C.c_char_p, C.POINTER(fr_tune_key), C.POINTER(fr_tune), C.c_float, C.c_float],
        "fr_autotune_search": [_vp, _vp, C.c_uint, C.c_float, _vp, C.c_int, C.c_int, C.c_char_p, C.POINTER(fr_tune),
                               C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int)],
        "fr_grids_destroy": [_vp],
        "fr_grids_download": [_vp, C.c_int, _vp],
        "fr_grids_upload": [_vp, C.c_int, _vp],
        "fr_grid_execute": [_vp, _vp, _vp],
        "fr_grid_execute_noalloc": [_vp, _vp, _vp],
        "fr_degrid_execute": [_vp, _vp, _vp, _vp],
        "fr_fft_create": [_vp, C.c_int, C.c_uint, C.c_int, C.c_int, _pp],
        "fr_fft_destroy": [_vp],
        "fr_fft_execute": [_vp, _vp],
        "fr_fftshift": [_vp, _vp],
        "fr_sos": [_vp, _vp, C.c_uint, C.c_uint, C.c_uint, C.c_int, _vp, _vp],
        "fr_fft_sos": [_vp, _vp, C.c_uint, C.c_uint, C.c_uint, C.c_int, _vp, _vp],
        "fr_deapod_map": [_vp, C.c_int, C.c_uint, C.c_float, _vp, C.c_int, _vp, C.c_uint, C.c_uint, C.c_uint, _vp],
        "fr_dcf_pipe_menon": [_vp, _vp, C.c_int, C.c_int],
        "fr_dcf_identity_scale": [_vp, _vp, C.c_int, C.POINTER(C.c_float)],
        "fr_lowfilter": [_vp, _vp, C.c_int],
        "fr_normalize": [_vp, _vp],
        "fr_crop": [_vp, _vp, C.c_uint, C.c_uint, C.c_uint, _vp],
        "fr_sos_map": [_vp, _vp, _vp, C.c_uint, C.c_uint, C.c_uint, C.c_int, _vp],
        "fr_grids_copy": [_vp, _vp, _vp],
        "fr_ldcf_features": [_vp, _vp, C.c_uint, _vp],
        "fr_mlp_param_count": [],
        "fr_mlp_create": [_vp, _pp],
        "fr_mlp_set_params": [_vp, _vp, C.c_float, C.c_float],
        "fr_mlp_get_params": [_vp, _vp, C.POINTER(C.c_float), C.POINTER(C.c_float)],
        "fr_mlp_set_base": [_vp, C.c_int],
        "fr_mlp_get_base": [_vp],
        "fr_mlp_forward": [_vp, _vp, C.c_int64, _vp],
        "fr_mlp_dcf": [_vp, _vp, _vp],
        "fr_ldcf_dcf": [_vp, _vp, C.c_uint],
        "fr_mlp_grad": [_vp, _vp, _vp, _vp, C.c_int64, _vp, C.POINTER(C.c_float)],
        "fr_mlp_train_step": [_vp, _vp, _vp, _vp, C.c_int64, C.c_float, C.POINTER(C.c_float)],
        "fr_malloc": [_vp, C.c_size_t, _pp],
        "fr_free": [_vp, _vp],
        "fr_memcpy_d2h": [_vp, _vp, _vp, C.c_size_t],
        "fr_memcpy_h2d": [_vp, _vp, _vp, C.c_size_t],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.fr_last_error.argtypes = [_vp]
    L.fr_last_error.restype = C.c_char_p
    L.fr_ctx_stream.argtypes = [_vp]
    L.fr_ctx_stream.restype = _vp
    L.fr_grids_device_ptr.argtypes = [_vp]
    L.fr_grids_device_ptr.restype = _vp
    L.fr_mlp_destroy.argtypes = [_vp]
    L.fr_mlp_destroy.restype = None
    L.fr_graph_destroy.argtypes = [_vp]
    L.fr_graph_destroy.restype = None
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(_vp)


# ----------------------------------------------------------- host helpers
def kb_lut(kwidth=4.0, og=2.0, L=512):
    out = np.zeros(L, np.float32)
    beta = np.zeros(1, np.float32)
    st = load().fr_kb_lut(kwidth, og, L, _ptr(out), _ptr(beta))
    if st:
        raise HipError(f"fr_kb_lut failed ({st})")
    return out, float(beta[0])


def normalize_traj(traj, N):
    """ReconData::normalizeTraj on a host (M, dim) float32 copy."""
    t = np.array(traj, np.float32, copy=True, order="C")
    M, dim = t.shape
    b = np.zeros(2 * dim, np.float32)
    st = load().fr_traj_normalize(_ptr(t), M, dim, dim, N, _ptr(b))
    if st:
        raise HipError(f"fr_traj_normalize failed ({st})")
    return t, b


# --------------------------------------------------------------- handles
_contexts = weakref.WeakSet()


@atexit.register
def _close_all():
    """At interpreter exit, every open context closes its handles and then
    itself, in that order, before module teardown destroys objects in an
    arbitrary order."""
    for ctx in list(_contexts):
        try:
            ctx.close()
        except Exception:
            pass


class Context:
    def __init__(self, device=0):
        self.L = load()
        self._children = weakref.WeakSet()
        h = _vp()
        st = self.L.fr_ctx_create(device, C.byref(h))
        if st:
            raise HipError(f"fr_ctx_create({device}) failed with status {st} (is a GPU visible?)")
        self.h = h
        _contexts.add(self)

    def adopt(self, child):
        """Handles are destroyed before their context (they use its stream)."""
        self._children.add(child)

    def check(self, st, what):
        if st:
            raise HipError(f"{what} failed ({st}): {self.L.fr_last_error(self.h).decode()}")

    @property
    def stream(self):
        return self.L.fr_ctx_stream(self.h)

    def sync(self):
        self.check(self.L.fr_ctx_synchronize(self.h), "fr_ctx_synchronize")

    def close(self):
        if self.h:
            for child in list(self._children):
                child.close()
            self.L.fr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Samples:
    def __init__(self, ctx, dim, M, coils, sample_type=FR_SAMPLES_C64, frame_samples=None):
        """frame_samples: per-frame sample counts for batched 2D frames (dim 2; M is their sum)."""
        self.ctx, self.dim, self.M, self.C = ctx, dim, int(M), int(coils)
        self.frames = 0
        h = _vp()
        if frame_samples is not None:
            fs = np.ascontiguousarray(frame_samples, np.int64)
            assert dim == 2 and int(fs.sum()) == self.M
            self.frames = len(fs)
            ctx.check(ctx.L.fr_samples_create_frames(ctx.h, self.frames, _ptr(fs), self.C, sample_type, C.byref(h)),
                      "fr_samples_create_frames")
        else:
            ctx.check(ctx.L.fr_samples_create(ctx.h, dim, self.M, self.C, sample_type, C.byref(h)),
                      "fr_samples_create")
        self.h = h
        ctx.adopt(self)

    def set_traj(self, traj):
        t = np.ascontiguousarray(traj, np.float32)
        assert t.shape == (self.M, self.dim)
        self.ctx.check(self.ctx.L.fr_samples_set_traj(self.h, _ptr(t), self.dim), "fr_samples_set_traj")

    def set_dcf(self, dcf):
        d = np.ascontiguousarray(dcf, np.float32)
        assert d.shape == (self.M,)
        self.ctx.check(self.ctx.L.fr_samples_set_dcf(self.h, _ptr(d)), "fr_samples_set_dcf")

    def get_dcf(self):
        d = np.zeros(self.M, np.float32)
        self.ctx.check(self.ctx.L.fr_samples_get_dcf(self.h, _ptr(d)), "fr_samples_get_dcf")
        return d

    def set_coil(self, c, data):
        d = np.ascontiguousarray(data, np.complex64)
        assert d.shape == (self.M,)
        self.ctx.check(self.ctx.L.fr_samples_set_coil(self.h, c, _ptr(d)), "fr_samples_set_coil")

    def set_coils(self, data):
        for c in range(self.C):
            self.set_coil(c, data[c])

    def update(self, c, first, data):
        d = np.ascontiguousarray(data, np.complex64)
        self.ctx.check(self.ctx.L.fr_samples_update(self.h, c, first, d.shape[0], _ptr(d)), "fr_samples_update")

    def set_data_device(self, ptr, layout):
        self.ctx.check(self.ctx.L.fr_samples_set_data_device(self.h, _vp(ptr), layout), "fr_samples_set_data_device")

    def set_dcf_device(self, ptr):
        self.ctx.check(self.ctx.L.fr_samples_set_dcf_device(self.h, _vp(ptr)), "fr_samples_set_dcf_device")

    def set_traj_device(self, ptr, stride):
        self.ctx.check(self.ctx.L.fr_samples_set_traj_device(self.h, _vp(ptr), stride), "fr_samples_set_traj_device")

    def close(self):
        if self.h:
            self.ctx.L.fr_samples_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    def __init__(self, ctx, samples, N, kwidth, lut, tune=None):
        self.ctx, self.N = ctx, int(N)
        lut = np.ascontiguousarray(lut, np.float32)
        t = None
        if tune:
            t = fr_tune(**{k: int(v) for k, v in tune.items()})
        h = _vp()
        ctx.check(ctx.L.fr_plan_create(ctx.h, samples.h, self.N, float(kwidth), _ptr(lut), len(lut),
                                       C.byref(t) if t is not None else None, C.byref(h)), "fr_plan_create")
        self.h = h
        ctx.adopt(self)

    def info(self):
        i = fr_plan_info()
        self.ctx.check(self.ctx.L.fr_plan_get_info(self.h, C.byref(i)), "fr_plan_get_info")
        return {k: getattr(i, k) for k, _ in fr_plan_info._fields_}

    def close(self):
        if self.h:
            self.ctx.L.fr_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Grids:
    def __init__(self, ctx, dim, N, coils, frames=0):
        """frames > 0: batched 2D frames, [coil][frame][N^2]."""
        self.ctx, self.dim, self.N, self.C, self.frames = ctx, dim, int(N), int(coils), int(frames)
        self.cells = self.N ** dim * max(self.frames, 1)
        h = _vp()
        if self.frames:
            assert dim == 2
            ctx.check(ctx.L.fr_grids_create_frames(ctx.h, self.N, self.C, self.frames, C.byref(h)),
                      "fr_grids_create_frames")
        else:
            ctx.check(ctx.L.fr_grids_create(ctx.h, dim, self.N, self.C, C.byref(h)), "fr_grids_create")
        self.h = h
        ctx.adopt(self)

    @property
    def ptr(self):
        return self.ctx.L.fr_grids_device_ptr(self.h)

    def download(self, c):
        out = np.zeros(self.cells, np.complex64)
        self.ctx.check(self.ctx.L.fr_grids_download(self.h, c, _ptr(out)), "fr_grids_download")
        return out

    def download_all(self):
        return np.stack([self.download(c) for c in range(self.C)])

    def upload(self, c, data):
        d = np.ascontiguousarray(data, np.complex64)
        self.ctx.check(self.ctx.L.fr_grids_upload(self.h, c, _ptr(d)), "fr_grids_upload")

    def close(self):
        if self.h:
            self.ctx.L.fr_grids_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FFT:
    def __init__(self, ctx, dim, N, batch, sign=+1):
        self.ctx = ctx
        h = _vp()
        ctx.check(ctx.L.fr_fft_create(ctx.h, dim, int(N), int(batch), int(sign), C.byref(h)), "fr_fft_create")
        self.h = h
        ctx.adopt(self)

    def __call__(self, grids):
        self.ctx.check(self.ctx.L.fr_fft_execute(self.h, grids.h), "fr_fft_execute")

    def close(self):
        if self.h:
            self.ctx.L.fr_fft_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx, self.nbytes = ctx, int(nbytes)
        p = _vp()
        ctx.check(ctx.L.fr_malloc(ctx.h, self.nbytes, C.byref(p)), "fr_malloc")
        self.ptr = p.value
        ctx.adopt(self)

    def to_host(self, dtype, count=None):
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else count
        out = np.zeros(n, dtype)
        self.ctx.check(self.ctx.L.fr_memcpy_d2h(self.ctx.h, _ptr(out), _vp(self.ptr), out.nbytes), "fr_memcpy_d2h")
        return out

    def from_host(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        self.ctx.check(self.ctx.L.fr_memcpy_h2d(self.ctx.h, _vp(self.ptr), _ptr(a), a.nbytes), "fr_memcpy_h2d")

    def close(self):
        if self.ptr and self.ctx.h:
            self.ctx.L.fr_free(self.ctx.h, _vp(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DevicePtr:
    """A byte offset into a DeviceBuffer, usable wherever a buffer's .ptr is read."""

    def __init__(self, buf, offset):
        assert 0 <= offset <= buf.nbytes
        self.buf, self.ptr = buf, buf.ptr + int(offset)


# ----------------------------------------------------------------- calls
def grid_execute(plan, samples, grids):
    plan.ctx.check(plan.ctx.L.fr_grid_execute(plan.h, samples.h, grids.h), "fr_grid_execute")


def grid_execute_noalloc(plan, samples, grids):
    """fr_grid_execute_noalloc: the same pass with no device allocation (the
    slice partials in the library's static per-GPU scratch)."""
    plan.ctx.check(plan.ctx.L.fr_grid_execute_noalloc(plan.h, samples.h, grids.h), "fr_grid_execute_noalloc")


def degrid_execute(plan, samples, grids, out_buf):
    plan.ctx.check(plan.ctx.L.fr_degrid_execute(plan.h, samples.h, grids.h, _vp(out_buf.ptr)), "fr_degrid_execute")


def fftshift(ctx, grids):
    ctx.check(ctx.L.fr_fftshift(ctx.h, grids.h), "fr_fftshift")


def sos(ctx, grids, rc, shift=True, deapod=None, out=None):
    rcx, rcy, rcz = rc
    if grids.dim == 2:
        rcz = max(grids.frames, 1)       # batched frames: one image per frame
    n = rcx * rcy * rcz
    if out is None:
        out = DeviceBuffer(ctx, 4 * n)
    ctx.check(ctx.L.fr_sos(ctx.h, grids.h, rcx, rcy, rcz, 1 if shift else 0,
                           _vp(deapod.ptr) if deapod is not None else None, _vp(out.ptr)), "fr_sos")
    return out


def fft_sos(fft, grids, rc, shift=True, deapod=None, out=None):
    """FFT + shift + crop + SOS in one call (fr_fft_sos); the grids' contents
    afterwards are unspecified (untransformed on the cropped path)."""
    ctx = fft.ctx
    rcx, rcy, rcz = rc
    if grids.dim == 2:
        rcz = max(grids.frames, 1)
    n = rcx * rcy * rcz
    if out is None:
        out = DeviceBuffer(ctx, 4 * n)
    ctx.check(ctx.L.fr_fft_sos(fft.h, grids.h, rcx, rcy, rcz, 1 if shift else 0,
                               _vp(deapod.ptr) if deapod is not None else None, _vp(out.ptr)), "fr_fft_sos")
    return out


def deapod_map(ctx, dim, N, kwidth, lut, centre, rc):
    rcx, rcy, rcz = rc
    if dim == 2:
        rcz = 1
    out = DeviceBuffer(ctx, 4 * rcx * rcy * rcz)
    lut = np.ascontiguousarray(lut, np.float32)
    c = np.ascontiguousarray(centre, np.float32)
    ctx.check(ctx.L.fr_deapod_map(ctx.h, dim, N, float(kwidth), _ptr(lut), len(lut), _ptr(c), rcx, rcy, rcz,
                                  _vp(out.ptr)), "fr_deapod_map")
    return out


def pipe_menon(plan, samples, iters=10, init_ones=True):
    plan.ctx.check(plan.ctx.L.fr_dcf_pipe_menon(plan.h, samples.h, iters, 1 if init_ones else 0),
                   "fr_dcf_pipe_menon")


def dcf_identity_scale(plan, samples, tile_stride=1):
    """Scale the samples' DCF in place by alpha = sum(w) / |G w|^2 (the
    Pipe-Menon fixed-point identity; |G w|^2 estimated from a sample of the
    grid with stride tile_stride when > 1, fr_dcf_identity_scale); returns
    alpha."""
    a = C.c_float(0.0)
    plan.ctx.check(plan.ctx.L.fr_dcf_identity_scale(plan.h, samples.h, int(tile_stride), C.byref(a)),
                   "fr_dcf_identity_scale")
    return float(a.value)


def lowfilter(ctx, grids, res):
    ctx.check(ctx.L.fr_lowfilter(ctx.h, grids.h, int(res)), "fr_lowfilter")


def normalize(ctx, grids):
    ctx.check(ctx.L.fr_normalize(ctx.h, grids.h), "fr_normalize")


def crop(ctx, grids, rc):
    rcx, rcy, rcz = rc
    if grids.dim == 2:
        rcz = 1
    out = DeviceBuffer(ctx, 8 * rcx * rcy * rcz * grids.C)
    ctx.check(ctx.L.fr_crop(ctx.h, grids.h, rcx, rcy, rcz, _vp(out.ptr)), "fr_crop")
    return out


def sos_map(ctx, grids, cmap, rc, shift=True):
    rcx, rcy, rcz = rc
    if grids.dim == 2:
        rcz = 1
    out = DeviceBuffer(ctx, 8 * rcx * rcy * rcz)
    ctx.check(ctx.L.fr_sos_map(ctx.h, grids.h, cmap.h, rcx, rcy, rcz, 1 if shift else 0, _vp(out.ptr)),
              "fr_sos_map")
    return out


def grids_copy(ctx, src, dst):
    ctx.check(ctx.L.fr_grids_copy(ctx.h, src.h, dst.h), "fr_grids_copy")



# ------------------------------------------------------ learned density compensation
def ldcf_features(ctx, samples, N):
    """Per-sample feature rows (M x 8 floats, device) of the samples' normalised trajectory."""
    out = DeviceBuffer(ctx, 32 * max(samples.M, 1))
    ctx.check(ctx.L.fr_ldcf_features(ctx.h, samples.h, int(N), _vp(out.ptr)), "fr_ldcf_features")
    return out


class MLP:
    """The 6-64-64-1 SiLU network (BF16 weight matrices, FP32 master copy and Adam state on the device)."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = _vp()
        ctx.check(ctx.L.fr_mlp_create(ctx.h, C.byref(h)), "fr_mlp_create")
        self.h = h
        self.nparam = ctx.L.fr_mlp_param_count()
        ctx.adopt(self)

    def set_params(self, params, mu=0.0, sigma=1.0):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == self.nparam
        self.ctx.check(self.ctx.L.fr_mlp_set_params(self.h, _ptr(p), float(mu), float(sigma)), "fr_mlp_set_params")

    def get_params(self):
        p = np.zeros(self.nparam, np.float32)
        mu, sg = C.c_float(), C.c_float()
        self.ctx.check(self.ctx.L.fr_mlp_get_params(self.h, _ptr(p), C.byref(mu), C.byref(sg)), "fr_mlp_get_params")
        return p, mu.value, sg.value

    @property
    def base(self):
        """Base feature of the DCF head (3..5: the DCF is predicted relative
        to 1 / (1 + n_k)), or -1."""
        return int(self.ctx.L.fr_mlp_get_base(self.h))

    @base.setter
    def base(self, feature):
        self.ctx.check(self.ctx.L.fr_mlp_set_base(self.h, int(feature)), "fr_mlp_set_base")

    def forward(self, feat, M):
        y = DeviceBuffer(self.ctx, 4 * max(M, 1))
        self.ctx.check(self.ctx.L.fr_mlp_forward(self.h, _vp(feat.ptr), int(M), _vp(y.ptr)), "fr_mlp_forward")
        return y

    def dcf(self, feat, samples):
        self.ctx.check(self.ctx.L.fr_mlp_dcf(self.h, _vp(feat.ptr), samples.h), "fr_mlp_dcf")

    def dcf_from_traj(self, samples, N):
        """Features + network in one pass (fr_ldcf_dcf): the samples' DCF from their trajectory."""
        self.ctx.check(self.ctx.L.fr_ldcf_dcf(self.h, samples.h, int(N)), "fr_ldcf_dcf")

    def grad(self, feat, targ, B, idx=None):
        g = np.zeros(self.nparam, np.float32)
        loss = C.c_float()
        self.ctx.check(self.ctx.L.fr_mlp_grad(self.h, _vp(feat.ptr), _vp(targ.ptr), _vp(idx.ptr) if idx else None,
                                              int(B), _ptr(g), C.byref(loss)), "fr_mlp_grad")
        return g, loss.value

    def train_step(self, feat, targ, B, lr, idx=None, want_loss=True):
        loss = C.c_float()
        self.ctx.check(self.ctx.L.fr_mlp_train_step(self.h, _vp(feat.ptr), _vp(targ.ptr),
                                                    _vp(idx.ptr) if idx else None, int(B), float(lr),
                                                    C.byref(loss) if want_loss else None), "fr_mlp_train_step")
        return loss.value if want_loss else None

    def close(self):
        if self.h:
            self.ctx.L.fr_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass



# ------------------------------------------------------------ graphs, autotuning
class Graph:
    """One captured reconstruction pass (gridding [+ FFT [+ shift/crop/SOS]])."""

    def __init__(self, plan, samples, grids, fft=None, rc=(0, 0, 0), shift=True, deapod=None, out=None):
        self.ctx = plan.ctx
        rcx, rcy, rcz = rc
        h = _vp()
        self.ctx.check(self.ctx.L.fr_graph_create(plan.h, samples.h, grids.h, fft.h if fft else None, rcx, rcy, rcz,
                                                  1 if shift else 0, _vp(deapod.ptr) if deapod else None,
                                                  _vp(out.ptr) if out else None, C.byref(h)), "fr_graph_create")
        self.h = h
        self.ctx.adopt(self)

    def launch(self):
        self.ctx.check(self.ctx.L.fr_graph_launch(self.h), "fr_graph_launch")

    def close(self):
        if self.h:
            self.ctx.L.fr_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


AUTOTUNE_DEFAULT_CANDIDATES = 8     # fr_autotune's built-in list (fr_tune_structural_candidates)


def _tune_dict(t):
    return {k: getattr(t, k) for k, _ in fr_tune._fields_}


def autotune_search(ctx, samples, N, kwidth, lut, reps=3, cache_path=None):
    """Staged autotuning search (fr_autotune_search), with an optional cache
    file.  Returns (best tune dict, best ms, default ms, timed configurations;
    0 on a cache hit)."""
    lut = np.ascontiguousarray(lut, np.float32)
    best = fr_tune()
    best_ms, def_ms, timed = C.c_float(), C.c_float(), C.c_int()
    path = os.fsencode(cache_path) if cache_path else None
    ctx.check(ctx.L.fr_autotune_search(ctx.h, samples.h, int(N), float(kwidth), _ptr(lut), len(lut), int(reps), path,
                                       C.byref(best), C.byref(best_ms), C.byref(def_ms), C.byref(timed)),
              "fr_autotune_search")
    return _tune_dict(best), best_ms.value, def_ms.value, timed.value


def tune_key(ctx, samples, N, kwidth, L):
    k = fr_tune_key()
    ctx.check(ctx.L.fr_tune_key_of(ctx.h, samples.h, int(N), float(kwidth), int(L), C.byref(k)), "fr_tune_key_of")
    return k


def tune_cache_lookup(lib, path, key):
    """Host-only cache read: (tune dict, best ms, default ms) or None on a miss."""
    t, bm, dm, found = fr_tune(), C.c_float(), C.c_float(), C.c_int()
    r = lib.fr_tune_cache_lookup(os.fsencode(path), C.byref(key), C.byref(t), C.byref(bm), C.byref(dm),
                                 C.byref(found))
    if r != 0:
        raise HipError(f"fr_tune_cache_lookup failed ({r})")
    return (_tune_dict(t), bm.value, dm.value) if found.value else None


def tune_cache_store(lib, path, key, tune, best_ms, default_ms):
    t = fr_tune(**{k: int(v) for k, v in tune.items()})
    r = lib.fr_tune_cache_store(os.fsencode(path), C.byref(key), C.byref(t), float(best_ms), float(default_ms))
    if r != 0:
        raise HipError(f"fr_tune_cache_store failed ({r})")


def autotune(ctx, samples, N, kwidth, lut, candidates=None, reps=3):
    """Fastest fr_tune (as a dict) for these samples; also each candidate's ms."""
    lut = np.ascontiguousarray(lut, np.float32)
    cands = None
    n = 0
    if candidates:
        n = len(candidates)
        cands = (fr_tune * n)(*[fr_tune(**{k: int(v) for k, v in c.items()}) for c in candidates])
    best = fr_tune()
    best_ms = C.c_float()
    ms = np.zeros(max(n, AUTOTUNE_DEFAULT_CANDIDATES), np.float32)
    ctx.check(ctx.L.fr_autotune(ctx.h, samples.h, int(N), float(kwidth), _ptr(lut), len(lut), cands, n, int(reps),
                                C.byref(best), C.byref(best_ms), _ptr(ms)), "fr_autotune")
    return _tune_dict(best), best_ms.value, ms
