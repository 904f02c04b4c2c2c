"""ORACLE -- test infrastructure only (see oracle/__init__.py).

ctypes front end of mbx_oracle.c (FP64 parity oracle, FP32 CPU-baseline port)
and a restatement of the geometry helpers:
  ConeBeamModel.get_psf_radius        mbirjax/cone_beam.py:138-172
  ConeBeamModel.auto_set_recon_size   mbirjax/cone_beam.py:174-187
  ParallelBeamModel.get_psf_radius    mbirjax/parallel_beam.py:115-123
  ParallelBeamModel.auto_set_recon_size mbirjax/parallel_beam.py:125-136
  TranslationModel.get_psf_radius     mbirjax/translation_model.py:109-140
  TranslationModel.auto_set_recon_size mbirjax/translation_model.py:142-186
  TomographyModel.__init__ delta_voxel default  mbirjax/tomography_model.py:59-64

Parity pins (tests/test_oracle.py): adjoint identity to ~1e-12, Hessian ==
squared column norms of a brute-force dense A, parallel-beam closed forms at
0 and 90 degrees, view-subset additivity, cone with sdd -> inf equal to
parallel beam.  Cone-beam geometry pins (tests/test_oracle_geometry.py), from
the physical scanner and not from the projector's formulas: the footprint
centroid equals the pinhole image of the voxel centre, the plateau equals the
chord of the central ray through the voxel, the footprint mass equals the
voxel volume times the Jacobian of the central projection (both to the
separable model's second-order gap), and offset shifts move the image by
whole cells; each of eight planted restatement mistakes (1/cos(phi), offset
signs, magnification in W_r, slice-offset sign, cos(alpha), 1/cos(theta_p),
magnification sign, rotation sense) fails at least one of them.  The
reference's own FBP/FDK accuracy test (tests/test_fbp_fdk.py) passes at its
tolerances.  Projector values are not pinned by reference golden vectors (the
reference holds none; JAX cannot run here).
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class OGeom(ctypes.Structure):
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


_libs = {}


def _lib(real):
    key = np.dtype(real).name
    if key not in _libs:
        name = {"float64": "liboracle_f64.so", "float32": "liboracle_f32.so"}[key]
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError("oracle library %s is not built; run oracle/build.sh" % path)
        lib = ctypes.CDLL(path)
        p = ctypes.c_void_p
        lib.oracle_forward.argtypes = [ctypes.POINTER(OGeom), p, p, p, ctypes.c_int, p, ctypes.c_int, p]
        lib.oracle_back.argtypes = [ctypes.POINTER(OGeom), p, p, p, ctypes.c_int, p, ctypes.c_int, ctypes.c_int, p]
        lib.oracle_set_mutation.argtypes = [ctypes.c_int]
        lib.oracle_get_mutation.restype = ctypes.c_int
        _libs[key] = lib
    return _libs[key]


class planted_mistake:
    """Context manager that plants one numbered restatement mistake in the
    FP64 oracle (see the list in mbx_oracle.c) and removes it on exit.  Only
    the tests that show the geometry pins catch such mistakes use it."""

    def __init__(self, number):
        self.number = int(number)

    def __enter__(self):
        lib = _lib(np.float64)
        assert lib.oracle_get_mutation() == 0
        lib.oracle_set_mutation(self.number)
        return self

    def __exit__(self, *exc):
        _lib(np.float64).oracle_set_mutation(0)
        return False


class Geometry:
    """Plain description of one scanner + recon grid (no JAX, no torch)."""

    def __init__(self, kind, sinogram_shape, angles, source_detector_dist=None, source_iso_dist=None,
                 recon_shape=None, delta_det_channel=1.0, delta_det_row=1.0, det_channel_offset=0.0,
                 det_row_offset=0.0, delta_voxel=None, recon_slice_offset=0.0, delta_recon_row=None):
        """kind: "parallel", "cone" or "translation"; angles: (V,) angles, or
        for "translation" the (V, 3) translation vectors (x, y, z)."""
        self.kind = kind
        self.sinogram_shape = tuple(int(v) for v in sinogram_shape)
        # The reference holds the view parameters as float32 JAX arrays
        # (jnp.asarray, mbirjax/cone_beam.py:56, parallel_beam.py:50,
        # translation_model.py:63): evaluate the geometry in FP64 at those
        # float32 values.
        self.angles = np.asarray(angles, dtype=np.float64).astype(np.float32).astype(np.float64)
        if kind == "translation":
            self.angles = self.angles.reshape(-1, 3)
        self.delta_det_channel = float(delta_det_channel)
        self.delta_det_row = float(delta_det_row)
        self.det_channel_offset = float(det_channel_offset)
        self.det_row_offset = float(det_row_offset)
        self.recon_slice_offset = float(recon_slice_offset)
        if kind in ("cone", "translation"):
            self.source_detector_dist = float(source_detector_dist)
            self.source_iso_dist = float(source_iso_dist)
        else:
            self.source_detector_dist = math.inf
            self.source_iso_dist = math.inf
        mag = self.magnification()
        self.delta_voxel = float(delta_voxel) if delta_voxel is not None else self.delta_det_channel / mag
        self.delta_recon_row = float(delta_recon_row) if delta_recon_row is not None else self.delta_voxel
        if recon_shape is not None:
            self.recon_shape = tuple(int(v) for v in recon_shape)
        elif kind == "translation":
            self.recon_shape, self.delta_recon_row = self.auto_translation_recon()
        else:
            self.recon_shape = self.auto_recon_shape()

    def magnification(self):
        if self.kind == "parallel" or math.isinf(self.source_detector_dist):
            return 1.0
        return self.source_detector_dist / self.source_iso_dist

    def auto_translation_recon(self):
        """(recon_shape, delta_recon_row) as TranslationModel.auto_set_recon_size."""
        V, R, C = self.sinogram_shape
        mag = self.magnification()
        box = np.array([self.delta_det_channel * C, self.delta_det_row * R])
        half_cone = math.atan2(box.max(), 2 * self.source_detector_dist)
        drow = math.sqrt(2) * (self.delta_voxel / math.tan(half_cone))
        t = self.angles.astype(np.float32)
        cube = t.max(axis=0) - t.min(axis=0)
        rbox = np.ceil((np.array([cube[0], cube[2]]) + box / mag / 2) / self.delta_voxel)
        rows = math.ceil(V * C * R / (rbox[0] * rbox[1]))
        rows = min(rows, math.floor((self.source_iso_dist - cube[1]) / drow))
        return (int(rows), int(rbox[0]), int(rbox[1])), drow

    def auto_recon_shape(self):
        _, R, C = self.sinogram_shape
        mag = self.magnification()
        if self.kind == "cone":
            rows = int(np.round(C * ((self.delta_det_channel / self.delta_voxel) / mag)))
            slices = int(np.round(R * ((self.delta_det_row / self.delta_voxel) / mag)))
        else:
            rows = int(np.ceil(C * self.delta_det_channel / (self.delta_voxel * mag)))
            slices = int(np.round(R * ((self.delta_det_row / self.delta_voxel) / mag)))
        return (rows, rows, slices)

    def psf_radii(self):
        if self.kind == "translation":
            d_det = min(self.delta_det_row, self.delta_det_channel)
            closest = self.source_iso_dist - 0.5 * self.recon_shape[0] * self.delta_recon_row - self.angles[:, 1].max()
            max_mag = self.source_detector_dist / closest
            psf = int(np.ceil(np.ceil(self.delta_voxel * max_mag / d_det) / 2))
            return psf, psf
        if self.kind != "cone":
            return int(np.ceil(np.ceil(self.delta_voxel / self.delta_det_channel) / 2)), 1
        d_det = min(self.delta_det_row, self.delta_det_channel)
        sdd = self.source_detector_dist
        if math.isinf(sdd):
            max_mag = min_mag = 1.0
        else:
            sid = sdd / self.magnification()
            ext = max(self.recon_shape[0], self.recon_shape[1]) * self.delta_voxel
            max_mag = sdd / (sid - ext)
            min_mag = sdd / (sid + ext)
        psf = int(np.ceil(np.ceil(self.delta_voxel * max_mag / d_det) / 2))
        bp = int(np.ceil(np.ceil(d_det / (min_mag * self.delta_voxel)) / 2))
        return psf, bp

    def ogeom(self):
        V, R, C = self.sinogram_shape
        rows, cols, nz = self.recon_shape
        psf, bp = self.psf_radii()
        kind = {"parallel": 0, "cone": 1, "translation": 2}[self.kind]
        return OGeom(kind, V, R, C, rows, cols, nz, psf, bp, 0,
                     self.delta_det_row, self.delta_det_channel, self.det_row_offset, self.det_channel_offset,
                     self.source_detector_dist, self.magnification(), self.delta_voxel, self.recon_slice_offset,
                     self.delta_recon_row)


def _views(views, n_default):
    if views is None:
        return None, n_default
    v = np.ascontiguousarray(np.asarray(views, dtype=np.int32))
    return v, v.size


def forward(geom, voxel_values, pixel_indices, views=None, real=np.float64):
    """Sparse forward projection -> (num view slots, R, C) in `real`."""
    V, R, C = geom.sinogram_shape
    nz = geom.recon_shape[2]
    vox = np.ascontiguousarray(np.asarray(voxel_values, dtype=real).reshape(-1, nz))
    pix = np.ascontiguousarray(np.asarray(pixel_indices, dtype=np.int32).reshape(-1))
    assert vox.shape[0] == pix.size
    vv, nv = _views(views, V)
    ang = np.ascontiguousarray(geom.angles, dtype=np.float64)
    out = np.zeros((nv, R, C), dtype=real)
    g = geom.ogeom()
    _lib(real).oracle_forward(ctypes.byref(g), ang.ctypes.data, vox.ctypes.data, pix.ctypes.data, pix.size,
                              vv.ctypes.data if vv is not None else None, nv, out.ctypes.data)
    return out


def back(geom, sinogram, pixel_indices, views=None, coeff_power=1, real=np.float64):
    """Sparse back projection -> (npix, Nz) in `real`.  With views, sinogram
    holds every view and views select both the data view and its angle."""
    V, R, C = geom.sinogram_shape
    nz = geom.recon_shape[2]
    sino = np.ascontiguousarray(np.asarray(sinogram, dtype=real))
    assert sino.shape[1:] == (R, C)
    pix = np.ascontiguousarray(np.asarray(pixel_indices, dtype=np.int32).reshape(-1))
    vv, nv = _views(views, sino.shape[0])
    ang = np.ascontiguousarray(geom.angles, dtype=np.float64)
    out = np.zeros((pix.size, nz), dtype=real)
    g = geom.ogeom()
    _lib(real).oracle_back(ctypes.byref(g), ang.ctypes.data, sino.ctypes.data, pix.ctypes.data, pix.size,
                           vv.ctypes.data if vv is not None else None, nv, int(coeff_power), out.ctypes.data)
    return out


def dense_matrix(geom):
    """Brute-force dense A (rows: sinogram entries (V,R,C) flattened; columns:
    voxels (rows*cols, Nz) flattened), built one voxel impulse at a time.
    Only for tiny grids."""
    V, R, C = geom.sinogram_shape
    rows, cols, nz = geom.recon_shape
    npix = rows * cols
    pix = np.arange(npix, dtype=np.int32)
    A = np.zeros((V * R * C, npix * nz))
    for p in range(npix):
        for k in range(nz):
            x = np.zeros((1, nz))
            x[0, k] = 1.0
            A[:, p * nz + k] = forward(geom, x, pix[p:p + 1]).reshape(-1)
    return A
