"""ctypes wrapper around oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the checker.  It may be imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
in the product package.  Every function restates the reference CPU path
(pengwg/fastrecon); the C source cites the reference file:line for each.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["sh", os.path.join(_HERE, "build_oracle.sh")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c = ctypes
        L.or_i0.restype = c.c_double
        L.or_i0.argtypes = [c.c_double]
        L.or_kb_lut.argtypes = [c.c_float, c.c_float, c.c_int, _f32p, c.POINTER(c.c_float)]
        L.or_traj_bounds.argtypes = [_f32p, c.c_int64, c.c_int, c.c_int, _f32p]
        L.or_normalize_traj.argtypes = [_f32p, c.c_int64, c.c_int, c.c_int, _f32p, c.c_uint]
        L.or_grid_f32.argtypes = [c.c_int, c.c_uint, c.c_int64, c.c_int, _f32p, c.c_int, _f32p,
                                  _f32p, c.c_float, _f32p, c.c_int, _f32p, c.c_int]
        L.or_grid_f64.argtypes = [c.c_int, c.c_uint, c.c_int64, c.c_int, _f32p, c.c_int, _f32p,
                                  _f32p, c.c_float, _f32p, c.c_int, _f64p, c.c_int]
        L.or_degrid_f64.argtypes = [c.c_int, c.c_uint, c.c_int64, c.c_int, _f32p, c.c_int,
                                    _f64p, c.c_float, _f32p, c.c_int, _f64p]
        L.or_degrid_f32.argtypes = [c.c_int, c.c_uint, c.c_int64, c.c_int, _f32p, c.c_int,
                                    _f32p, c.c_float, _f32p, c.c_int, _f32p]
        L.or_fft_c2c.argtypes = [c.c_int, c.c_uint, c.c_int, _f64p, c.c_int]
        L.or_fftshift_f32.argtypes = [c.c_int, c.c_uint, c.c_int, _f32p]
        L.or_fftshift_f64.argtypes = [c.c_int, c.c_uint, c.c_int, _f64p]
        L.or_sos_f32.argtypes = [c.c_int, c.c_uint, c.c_int, _f32p, c.c_uint, c.c_uint, c.c_uint, _f32p]
        L.or_sos_f64.argtypes = [c.c_int, c.c_uint, c.c_int, _f64p, c.c_uint, c.c_uint, c.c_uint, _f64p]
        L.or_pipe_menon_f64.argtypes = [c.c_int, c.c_uint, c.c_int64, _f32p, c.c_int, c.c_float,
                                        _f32p, c.c_int, c.c_int, _f64p]
        L.or_lowfilter_f32.argtypes = [c.c_int, c.c_uint, c.c_int, _f32p, c.c_int]
        L.or_normalize_f32.argtypes = [c.c_int, c.c_uint, c.c_int, _f32p]
        L.or_sos_map_f32.argtypes = [c.c_int, c.c_uint, c.c_int, _f32p, _f32p, c.c_uint, c.c_uint,
                                     c.c_uint, _f32p]
        _lib = L
    return _lib


def i0(x):
    return lib().or_i0(float(x))


def kb_lut(kwidth=4.0, og=2.0, L=512):
    out = np.zeros(L, np.float32)
    beta = ctypes.c_float()
    lib().or_kb_lut(kwidth, og, L, out, ctypes.byref(beta))
    return out, beta.value


def traj_bounds(traj):
    traj = np.ascontiguousarray(traj, np.float32)
    M, dim = traj.shape
    b = np.zeros(2 * dim, np.float32)
    lib().or_traj_bounds(traj, M, dim, dim, b)
    return b


def normalize_traj(traj, N):
    """Returns (normalised copy, updated bounds) -- ReconData::normalizeTraj."""
    t = np.array(traj, np.float32, copy=True, order="C")
    M, dim = t.shape
    b = traj_bounds(t)
    lib().or_normalize_traj(t, M, dim, dim, b, N)
    return t, b


def _cells(dim, N):
    return N ** 3 if dim == 3 else N * N


def grid_f32(traj, dcf, data, N, kwidth, lut, nthreads=1):
    """traj (M,dim) normalised; data (C,M) complex64 -> (C, N^dim) complex64."""
    traj = np.ascontiguousarray(traj, np.float32)
    M, dim = traj.shape
    data = np.ascontiguousarray(data, np.complex64)
    C = data.shape[0]
    out = np.zeros(C * _cells(dim, N) * 2, np.float32)
    lib().or_grid_f32(dim, N, M, C, traj, dim, np.ascontiguousarray(dcf, np.float32),
                      data.view(np.float32).reshape(-1), kwidth,
                      np.ascontiguousarray(lut, np.float32), len(lut), out, nthreads)
    return out.view(np.complex64).reshape(C, -1)


def grid_f64(traj, dcf, data, N, kwidth, lut, nthreads=1):
    traj = np.ascontiguousarray(traj, np.float32)
    M, dim = traj.shape
    data = np.ascontiguousarray(data, np.complex64)
    C = data.shape[0]
    out = np.zeros(C * _cells(dim, N) * 2, np.float64)
    lib().or_grid_f64(dim, N, M, C, traj, dim, np.ascontiguousarray(dcf, np.float32),
                      data.view(np.float32).reshape(-1), kwidth,
                      np.ascontiguousarray(lut, np.float32), len(lut), out, nthreads)
    return out.view(np.complex128).reshape(C, -1)


def degrid_f64(traj, grid, N, kwidth, lut):
    traj = np.ascontiguousarray(traj, np.float32)
    M, dim = traj.shape
    grid = np.ascontiguousarray(grid, np.complex128)
    C = grid.shape[0]
    out = np.zeros(C * M * 2, np.float64)
    lib().or_degrid_f64(dim, N, M, C, traj, dim, grid.view(np.float64).reshape(-1), kwidth,
                        np.ascontiguousarray(lut, np.float32), len(lut), out)
    return out.view(np.complex128).reshape(C, M)


def degrid_f32(traj, grid, N, kwidth, lut):
    """FP32 degridding in tap order (fr_oracle.c or_degrid_f32)."""
    traj = np.ascontiguousarray(traj, np.float32)
    M, dim = traj.shape
    grid = np.ascontiguousarray(grid, np.complex64)
    C = grid.shape[0]
    out = np.zeros(C * M * 2, np.float32)
    lib().or_degrid_f32(dim, N, M, C, traj, dim, grid.view(np.float32).reshape(-1), kwidth,
                        np.ascontiguousarray(lut, np.float32), len(lut), out)
    return out.view(np.complex64).reshape(C, M)


def fft_c2c(dim, N, data, sign=+1):
    """data (B, N^dim) complex -> FP64 unnormalised FFT (copy)."""
    a = np.array(data, np.complex128, copy=True, order="C")
    B = a.shape[0]
    lib().or_fft_c2c(dim, N, B, a.view(np.float64).reshape(-1), sign)
    return a


def fftshift(dim, N, data):
    a = np.array(data, copy=True, order="C")
    B = a.shape[0]
    if a.dtype == np.complex64:
        lib().or_fftshift_f32(dim, N, B, a.view(np.float32).reshape(-1))
    else:
        a = a.astype(np.complex128)
        lib().or_fftshift_f64(dim, N, B, a.view(np.float64).reshape(-1))
    return a


def sos(dim, N, data, rc):
    rcx, rcy, rcz = rc
    if dim == 2:
        rcz = 1
    a = np.ascontiguousarray(data)
    C = a.shape[0]
    if a.dtype == np.complex64:
        out = np.zeros(rcx * rcy * rcz, np.float32)
        lib().or_sos_f32(dim, N, C, a.view(np.float32).reshape(-1), rcx, rcy, rcz, out)
    else:
        a = a.astype(np.complex128)
        out = np.zeros(rcx * rcy * rcz, np.float64)
        lib().or_sos_f64(dim, N, C, a.view(np.float64).reshape(-1), rcx, rcy, rcz, out)
    return out


def pipe_menon(traj, N, kwidth, lut, iters=10, w0=None):
    traj = np.ascontiguousarray(traj, np.float32)
    M, dim = traj.shape
    w = np.ones(M, np.float64) if w0 is None else np.array(w0, np.float64)
    lib().or_pipe_menon_f64(dim, N, M, traj, dim, kwidth,
                            np.ascontiguousarray(lut, np.float32), len(lut), iters, w)
    return w


def recon_f32(traj_norm, dcf, data, N, kwidth, lut, rc, nthreads=1):
    """Reference CPU pipeline (Main.cpp:116-181): grid -> FFT(+1) -> shift -> SOS.

    The FFT runs in FP64 (FFTW is absent here); gridding and SOS keep the
    reference's FP32 semantics."""
    M, dim = np.asarray(traj_norm).shape
    g = grid_f32(traj_norm, dcf, data, N, kwidth, lut, nthreads)
    f = fft_c2c(dim, N, g, +1).astype(np.complex64)
    f = fftshift(dim, N, f)
    return sos(dim, N, f, rc)


def lowfilter(dim, N, data, res):
    """ImageFilter::lowFilter (ImageFilter.cpp:56-93) on C x N^dim complex64."""
    a = np.ascontiguousarray(data, dtype=np.complex64).copy()
    lib().or_lowfilter_f32(dim, N, a.shape[0], a.view(np.float32).reshape(-1), res)
    return a


def normalize(dim, N, data):
    """ImageFilter::normalize (ImageFilter.cpp:95-118)."""
    a = np.ascontiguousarray(data, dtype=np.complex64).copy()
    lib().or_normalize_f32(dim, N, a.shape[0], a.view(np.float32).reshape(-1))
    return a


def sos_map(dim, N, data, cmap, rc):
    """ImageFilter::SOS(map, size) (ImageFilter.cpp:170-209): complex64 crop."""
    rcx, rcy, rcz = rc
    if dim == 2:
        rcz = 1
    a = np.ascontiguousarray(data, dtype=np.complex64)
    m = np.ascontiguousarray(cmap, dtype=np.complex64)
    out = np.zeros(rcx * rcy * rcz, np.complex64)
    lib().or_sos_map_f32(dim, N, a.shape[0], a.view(np.float32).reshape(-1),
                         m.view(np.float32).reshape(-1), rcx, rcy, rcz, out.view(np.float32))
    return out


def pils_f32(traj_norm, dcf, data, N, kwidth, lut, rc, nthreads=1):
    """Reference PILS pipeline (Main.cpp:116-170): the coil map is the
    gridded k-space low-pass filtered (res 22), FFT'd, shifted and
    normalised; the image is sum_c img_c conj(map_c); the file holds |.|."""
    M, dim = np.asarray(traj_norm).shape
    g = grid_f32(traj_norm, dcf, data, N, kwidth, lut, nthreads)
    img = fftshift(dim, N, fft_c2c(dim, N, g, +1).astype(np.complex64))
    m = lowfilter(dim, N, g, 22)
    m = fftshift(dim, N, fft_c2c(dim, N, m, +1).astype(np.complex64))
    m = normalize(dim, N, m)
    return np.abs(sos_map(dim, N, img, m, rc)).astype(np.float32)
