"""ctypes binding of libmbx.so, the gfx950 kernel library (C ABI in include/mbx.h).

The library is built in-tree by __graft_entry__.build() (mbirjax_amd/_build.py).
Nothing here falls back to another implementation: if libmbx.so is missing or
no GPU is visible, every compute entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmbx.so")

MBX_OK = 0
MBX_EINVAL = -1
MBX_ELAUNCH = -2
MBX_EUNSUPPORTED = -3
MBX_EWORKSPACE = -4

GEOM_PARALLEL = 0
GEOM_CONE = 1
GEOM_TRANSLATION = 2
DTYPE_F32 = 0
DTYPE_BF16 = 1
DTYPE_F64 = 2
DTYPE_F16 = 3
LAYOUT_VCR = 0
LAYOUT_VRC = 1

# Every exported symbol of include/mbx.h.
EXPORTS = (
    "mbx_version", "mbx_last_error", "mbx_back_workspace_bytes", "mbx_forward_workspace_bytes",
    "mbx_forward_project", "mbx_forward_plan", "mbx_forward_apply", "mbx_back_project", "mbx_sino_transpose",
    "mbx_qggmrf_grad_hess",
    "mbx_dot_workspace_bytes", "mbx_dot", "mbx_sino_axpy", "mbx_vcd_workspace_bytes", "mbx_vcd_subset_update",
    "mbx_denoise_workspace_bytes", "mbx_denoise_subset_update", "mbx_median_filter3d",
    "mbx_noalloc_scratch_bytes", "mbx_back_project_noalloc", "mbx_vcd_subset_update_planned",
    "mbx_back_apply_workspace_bytes", "mbx_back_apply",
)


class MbxGeom(ctypes.Structure):
    """mirror of struct mbx_geom (include/mbx.h)"""
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("num_views", ctypes.c_int32),
        ("num_det_rows", ctypes.c_int32),
        ("num_det_channels", ctypes.c_int32),
        ("num_recon_rows", ctypes.c_int32),
        ("num_recon_cols", ctypes.c_int32),
        ("num_recon_slices", ctypes.c_int32),
        ("psf_radius", ctypes.c_int32),
        ("bp_psf_radius", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("delta_det_row", ctypes.c_double),
        ("delta_det_channel", ctypes.c_double),
        ("det_row_offset", ctypes.c_double),
        ("det_channel_offset", ctypes.c_double),
        ("source_detector_dist", ctypes.c_double),
        ("magnification", ctypes.c_double),
        ("delta_voxel", ctypes.c_double),
        ("recon_slice_offset", ctypes.c_double),
        ("delta_recon_row", ctypes.c_double),
    ]


class MbxQggmrf(ctypes.Structure):
    """mirror of struct mbx_qggmrf (include/mbx.h)"""
    _fields_ = [("b", ctypes.c_double * 6), ("sigma_x", ctypes.c_double), ("p", ctypes.c_double),
                ("q", ctypes.c_double), ("T", ctypes.c_double)]


_lib = None


def lib():
    """Load libmbx.so (once).  Raises RuntimeError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmbx.so is not built (%s); run `python -c \"import __graft_entry__ as g; "
                               "g.build()\"` first" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, f32, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double
        G = ctypes.POINTER(MbxGeom)
        Q = ctypes.POINTER(MbxQggmrf)
        sig = {
            "mbx_version": (i32, []),
            "mbx_last_error": (ctypes.c_char_p, []),
            "mbx_back_workspace_bytes": (i64, [G, i32, i32]),
            "mbx_forward_workspace_bytes": (i64, [G, i32, i32]),
            "mbx_forward_project": (i32, [G, vp, vp, vp, i32, vp, i32, vp, i32, i32, i32, vp, vp, vp, i64, vp]),
            "mbx_forward_plan": (i32, [G, vp, vp, i32, vp, i32, vp, i64, vp]),
            "mbx_forward_apply": (i32, [G, vp, i64, vp, i32, vp, i32, vp, i32, i32, i32, vp, vp, vp]),
            "mbx_back_project": (i32, [G, vp, vp, i32, i32, vp, vp, i32, vp, i32, i32, f32, i32, vp, vp, i64, vp]),
            "mbx_sino_transpose": (i32, [vp, i32, i32, vp, i32, i32, i32, i32, i32, vp]),
            "mbx_qggmrf_grad_hess": (i32, [vp, i32, i32, i32, vp, i32, Q, vp, vp, vp]),
            "mbx_dot_workspace_bytes": (i64, []),
            "mbx_dot": (i32, [vp, i32, vp, i32, vp, i64, vp, vp, i64, vp]),
            "mbx_sino_axpy": (i32, [vp, i32, vp, i32, vp, i64, vp]),
            "mbx_vcd_workspace_bytes": (i64, [G, i32, i32]),
            "mbx_vcd_subset_update": (i32, [G, vp, vp, vp, i32, vp, vp, vp, i32, f64, Q, vp, f64, i32, i32, vp, vp,
                                            i64, vp]),
            "mbx_vcd_subset_update_planned": (i32, [G, vp, vp, vp, i32, vp, vp, vp, i32, f64, Q, vp, f64, i32, i32,
                                                    vp, i64, vp, vp, i64, vp]),
            "mbx_denoise_workspace_bytes": (i64, [i32, i32]),
            "mbx_denoise_subset_update": (i32, [vp, vp, i32, i32, i32, vp, i32, f64, Q, vp, vp, i64, vp]),
            "mbx_median_filter3d": (i32, [vp, i32, i32, i32, i32, vp, vp]),
            "mbx_noalloc_scratch_bytes": (i64, []),
            "mbx_back_project_noalloc": (i32, [G, vp, vp, i32, i32, vp, vp, i32, vp, i32, i32, f32, i32, vp, vp]),
            "mbx_back_apply_workspace_bytes": (i64, [G, i32, i32]),
            "mbx_back_apply": (i32, [G, vp, i64, vp, i32, i32, vp, i32, vp, i32, i32, f32, i32, vp, vp, i64, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    """Turn a libmbx status into an exception carrying mbx_last_error()."""
    if rc == MBX_OK:
        return
    msg = lib().mbx_last_error().decode(errors="replace")
    if rc in (MBX_EINVAL, MBX_EWORKSPACE):
        raise ValueError("libmbx: " + msg)
    raise RuntimeError("libmbx error %d: %s" % (rc, msg))


def require_gpu():
    """The product has no CPU path: fail loudly when no GPU is visible."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("mbirjax_amd needs an AMD Instinct GPU (HIP); none is visible")
    lib()


def ptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def np_geom_struct(kind, sinogram_shape, recon_shape, psf_radius, bp_psf_radius, delta_det_row, delta_det_channel,
                   det_row_offset, det_channel_offset, source_detector_dist, magnification, delta_voxel,
                   recon_slice_offset, delta_recon_row=0.0):
    V, R, C = (int(v) for v in sinogram_shape)
    rows, cols, nz = (int(v) for v in recon_shape)
    return MbxGeom(int(kind), V, R, C, rows, cols, nz, int(psf_radius), int(bp_psf_radius), 0,
                   float(delta_det_row), float(delta_det_channel), float(det_row_offset), float(det_channel_offset),
                   float(source_detector_dist), float(magnification), float(delta_voxel), float(recon_slice_offset),
                   float(delta_recon_row))


def qggmrf_struct(b, sigma_x, p, q, T):
    s = MbxQggmrf()
    for i in range(6):
        s.b[i] = float(b[i])
    s.sigma_x, s.p, s.q, s.T = float(sigma_x), float(p), float(q), float(T)
    return s


def as_i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))
