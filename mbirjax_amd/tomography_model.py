"""TomographyModel: the mbirjax model API on MI355X kernels.

Public methods keep the signatures and semantics of mbirjax/tomography_model.py
(forward_project 531, back_project 553, sparse_forward_project 576,
sparse_back_project 624, compute_hessian_diagonal 679, set_params 710, recon 966,
vcd_recon 1086, prox_map 1660, gen_weights 1727, get_forward_model_loss 1633).
Arrays come back as torch tensors on the GPU; inputs may be NumPy arrays or
torch tensors on any device.

Compute goes through libmbx (include/mbx.h).  Internally the error sinogram
lives in the channel-major (views, channels, rows) layout that the projectors
stream; user-facing sinograms are (views, rows, channels) and are converted by
mbx_sino_transpose at the boundary.
"""
import datetime
import inspect
import math
import os
import warnings
from collections import namedtuple

import numpy as np
import torch

from . import _lib
from . import qggmrf as _qggmrf
from . import utilities
from . import vcd_utils
from .parameter_handler import ParameterHandler

recon_param_names = ['num_iterations', 'granularity', 'partition_sequence', 'fm_rmse', 'prior_loss',
                     'regularization_params', 'stop_threshold_change_pct', 'alpha_values']
ReconParams = namedtuple('ReconParams', recon_param_names)

_F32 = _lib.DTYPE_F32
_BF16 = _lib.DTYPE_BF16
_F16 = _lib.DTYPE_F16

# precision mode -> (error-sinogram storage, delta-sinogram storage)
PRECISIONS = {'fp32': (_F32, _F32), 'mixed': (_F32, _BF16), 'fp16': (_F16, _BF16), 'bf16': (_BF16, _BF16)}


def _torch_dtype(dt):
    return {_F32: torch.float32, _BF16: torch.bfloat16, _F16: torch.float16}[dt]


def _dtype_code(t):
    return {torch.float32: _F32, torch.bfloat16: _BF16, torch.float16: _F16}[t.dtype]


def _as_host_f32(a):
    if isinstance(a, torch.Tensor):
        return a.detach().to('cpu', torch.float32).numpy()
    return np.asarray(a, dtype=np.float32)


class TomographyModel(ParameterHandler):
    """Base model; subclasses provide the geometry (ConeBeamModel, ParallelBeamModel)."""

    DIRECT_RECON_VIEW_BATCH_SIZE = 100
    GEOM_KIND = None

    def __init__(self, sinogram_shape, **kwargs):
        super().__init__()
        self.set_params(no_compile=True, no_warning=True, sinogram_shape=tuple(int(v) for v in sinogram_shape),
                        **kwargs)
        if self.get_params('delta_voxel') is None:
            delta_voxel = self.get_params('delta_det_channel') / self.get_magnification()
            self.set_params(no_compile=True, no_warning=True, delta_voxel=delta_voxel)
        self.use_ror_mask = True
        self.auto_set_recon_size(sinogram_shape, no_compile=True, no_warning=True)
        self.set_params(geometry_type=str(type(self)))
        self.verify_valid_params()
        self.view_batch_size_for_vmap = 512
        self.pixel_batch_size_for_vmap = 2048
        self.transfer_pixel_batch_size = 100 * self.pixel_batch_size_for_vmap
        self.use_gpu = 'full'
        self.main_device = self.sinogram_device = self.worker = None
        self._geom = None
        self._angles_dev = None
        self.create_projectors()

    @classmethod
    def get_required_param_names(cls):
        names = [p.name for p in inspect.signature(cls.__init__).parameters.values()
                 if p.kind not in (inspect.Parameter.VAR_POSITIONAL, inspect.Parameter.VAR_KEYWORD)]
        return names[1:] if names and names[0] == 'self' else names

    # ------------------------------------------------------------------ setup
    def set_devices_and_batch_sizes(self):
        """Everything runs on the current HIP device (no CPU placement tiers)."""
        if torch.cuda.is_available():
            dev = torch.device('cuda', torch.cuda.current_device())
            self.main_device = self.sinogram_device = self.worker = dev

    def create_projectors(self):
        """Re-plan the projectors after a geometry change (mbirjax/tomography_model.py:475-484)."""
        self.set_devices_and_batch_sizes()
        gp = self.get_geometry_parameters()
        sinogram_shape, recon_shape = self.get_params(['sinogram_shape', 'recon_shape'])
        g = gp._asdict()
        self._geom = _lib.np_geom_struct(
            self.GEOM_KIND, sinogram_shape, recon_shape, g['psf_radius'], g.get('bp_psf_radius', 1),
            g.get('delta_det_row', 1.0), g['delta_det_channel'], g.get('det_row_offset', 0.0),
            g['det_channel_offset'], g.get('source_detector_dist', math.inf), g.get('magnification', 1.0),
            g['delta_voxel'], g.get('recon_slice_offset', 0.0), g.get('delta_recon_row', 0.0))
        self._angles_dev = None

    def set_params(self, no_warning=False, no_compile=False, **kwargs):
        recompile = super().set_params(no_warning=no_warning, no_compile=no_compile, **kwargs)
        if recompile:
            self.create_projectors()

    def verify_valid_params(self):
        super().verify_valid_params()
        if self.get_params('use_gpu') not in ['automatic', 'full', 'sinograms', 'projections', 'none']:
            raise ValueError("use_gpu must be one of 'automatic', 'full', 'sinograms', 'projections', 'none'")

    def get_geometry_parameters(self):
        raise NotImplementedError

    def auto_set_recon_size(self, sinogram_shape, no_compile=True, no_warning=False):
        raise NotImplementedError('auto_set_recon_size must be implemented by each specific geometry model.')

    # ------------------------------------------------------------------ device helpers
    def _device(self):
        _lib.require_gpu()
        if self.main_device is None:
            self.set_devices_and_batch_sizes()
        return self.main_device

    def _angles(self):
        dev = self._device()
        if self._angles_dev is None:
            name = self.get_params('view_params_name')
            ang = self.get_params(name)
            self._angles_dev = torch.as_tensor(_as_host_f32(ang).reshape(-1), device=dev).contiguous()
        return self._angles_dev

    def _tensor(self, a, dtype=torch.float32):
        dev = self._device()
        if isinstance(a, torch.Tensor):
            return a.detach().to(device=dev, dtype=dtype).contiguous()
        return torch.as_tensor(np.asarray(a), device=dev).to(dtype).contiguous()

    def _indices(self, idx, bound=None, what='pixel_indices', device_check=True):
        """int32 device copy of an index array; with bound, every entry must lie in [0, bound).

        Host arrays (NumPy, lists, CPU tensors) are checked on the host before
        the upload.  A tensor already on the GPU is checked there only with
        device_check, which waits for the device (two reductions read back);
        the projectors pass device_check=False for pixel indices, which the
        kernels use for geometry alone: an index outside the recon contributes
        nothing (zero channel weights) and never addresses memory."""
        dev = self._device()
        if bound is None:
            rows, cols = self.get_params('recon_shape')[:2]
            bound = rows * cols
        on_device = isinstance(idx, torch.Tensor) and idx.device.type != 'cpu'
        if on_device:
            t = idx.detach().to(device=dev, dtype=torch.int32).reshape(-1).contiguous()
        else:
            h = np.asarray(idx.detach().numpy() if isinstance(idx, torch.Tensor) else idx).reshape(-1)
            if h.size > 0:
                lo, hi = int(h.min()), int(h.max())
                if lo < 0 or hi >= bound:
                    raise ValueError('{} must lie in [0, {}); got range [{}, {}]'.format(what, bound, lo, hi))
            return torch.as_tensor(_lib.as_i32(h), device=dev)
        if device_check and t.numel() > 0:
            lo, hi = int(t.min().item()), int(t.max().item())
            if lo < 0 or hi >= bound:
                raise ValueError('{} must lie in [0, {}); got range [{}, {}]'.format(what, bound, lo, hi))
        return t

    @staticmethod
    def _output(t, output_device):
        if output_device is None:
            return t
        return t.to(output_device)

    # ------------------------------------------------------------------ layout conversion
    def _transpose(self, src, src_layout, dst_layout, dst_dtype=_F32):
        V = src.shape[0]
        R, C = self.get_params('sinogram_shape')[1:]
        shape = (V, C, R) if dst_layout == _lib.LAYOUT_VCR else (V, R, C)
        dst = torch.empty(shape, dtype=_torch_dtype(dst_dtype), device=src.device)
        src_dt = _dtype_code(src)
        _lib.check(_lib.lib().mbx_sino_transpose(_lib.ptr(src), src_dt, src_layout, _lib.ptr(dst), dst_dtype,
                                                 dst_layout, V, R, C, _lib.stream()))
        return dst

    def to_channel_major(self, sinogram, dtype=_F32):
        """(V, R, C) user sinogram -> (V, C, R) device tensor."""
        s = self._tensor(sinogram)
        if s.dim() != 3 or tuple(s.shape[1:]) != tuple(self.get_params('sinogram_shape')[1:]):
            raise ValueError('sinogram must have shape (views, {}, {}), got {}'.format(
                *self.get_params('sinogram_shape')[1:], tuple(s.shape)))
        return self._transpose(s, _lib.LAYOUT_VRC, _lib.LAYOUT_VCR, dtype)

    def from_channel_major(self, sino_vcr):
        return self._transpose(sino_vcr, _lib.LAYOUT_VCR, _lib.LAYOUT_VRC, _F32)

    # ------------------------------------------------------------------ projector cores
    def _fwd(self, vox, pix, views=None, layout=_lib.LAYOUT_VRC, out_dtype=_F32, weights=None, want_sumsq=False,
             out=None, accumulate=False):
        """Forward project cylinders `vox` (npix, Nz) at pixels `pix` (device tensors)."""
        V, R, C = self.get_params('sinogram_shape')
        nv = V if views is None else int(views.numel())
        if out is None:
            shape = (nv, C, R) if layout == _lib.LAYOUT_VCR else (nv, R, C)
            out = torch.empty(shape, dtype=_torch_dtype(out_dtype), device=vox.device)
        sumsq = torch.zeros(1, dtype=torch.float64, device=vox.device) if want_sumsq else None
        wsb = int(_lib.lib().mbx_forward_workspace_bytes(self._geom, int(pix.numel()), nv))
        ws = torch.empty(max(wsb, 256), dtype=torch.uint8, device=vox.device)
        _lib.check(_lib.lib().mbx_forward_project(
            self._geom, _lib.ptr(self._angles()), _lib.ptr(vox), _lib.ptr(pix), int(pix.numel()), _lib.ptr(views),
            nv, _lib.ptr(out), out_dtype, layout, int(accumulate), _lib.ptr(weights), _lib.ptr(sumsq),
            _lib.ptr(ws), wsb, _lib.stream()))
        return (out, sumsq) if want_sumsq else out

    def _fwd_plan(self, pix, views=None):
        """Geometry table, bins and forward records of the pixel set pix (device
        int32), reusable by _fwd_apply while pix and views are unchanged."""
        V = self.get_params('sinogram_shape')[0]
        nv = V if views is None else int(views.numel())
        nb = int(_lib.lib().mbx_forward_workspace_bytes(self._geom, int(pix.numel()), nv))
        buf = torch.empty(max(nb, 256), dtype=torch.uint8, device=pix.device)
        _lib.check(_lib.lib().mbx_forward_plan(self._geom, _lib.ptr(self._angles()), _lib.ptr(pix), int(pix.numel()),
                                               _lib.ptr(views), nv, _lib.ptr(buf), nb, _lib.stream()))
        return ForwardPlan(buf, nb, int(pix.numel()), nv)

    def _fwd_apply(self, plan, vox, views=None, layout=_lib.LAYOUT_VCR, out_dtype=_F32, out=None, accumulate=False):
        """Forward projection of cylinders vox (npix, Nz) with a plan from _fwd_plan."""
        V, R, C = self.get_params('sinogram_shape')
        nv = V if views is None else int(views.numel())
        if int(vox.shape[0]) != plan.npix or nv != plan.nviews:
            raise ValueError('forward plan of {} pixels and {} views applied to {} pixels and {} views'.format(
                plan.npix, plan.nviews, int(vox.shape[0]), nv))
        if out is None:
            shape = (nv, C, R) if layout == _lib.LAYOUT_VCR else (nv, R, C)
            out = torch.empty(shape, dtype=_torch_dtype(out_dtype), device=vox.device)
        _lib.check(_lib.lib().mbx_forward_apply(self._geom, _lib.ptr(plan.buf), plan.nbytes, _lib.ptr(vox),
                                                int(vox.shape[0]), _lib.ptr(views), nv, _lib.ptr(out), out_dtype,
                                                layout, int(accumulate), None, None, _lib.stream()))
        return out

    def _back_apply(self, plan, sino_vcr, views=None, coeff_power=1, weights=None, scale=1.0, out=None,
                    accumulate=False):
        """Back projection of a channel-major sinogram onto the pixel set of a
        plan from _fwd_plan (its geometry table reused): _back without the table."""
        nz = self.get_params('recon_shape')[2]
        V = self.get_params('sinogram_shape')[0]
        nv = V if views is None else int(views.numel())
        if nv != plan.nviews:
            raise ValueError('plan of {} views applied to {} views'.format(plan.nviews, nv))
        if out is None:
            out = torch.empty((plan.npix, nz), dtype=torch.float32, device=sino_vcr.device)
        wsb = int(_lib.lib().mbx_back_apply_workspace_bytes(self._geom, plan.npix, nv))
        ws = torch.empty(wsb, dtype=torch.uint8, device=sino_vcr.device) if wsb > 0 else None
        _lib.check(_lib.lib().mbx_back_apply(
            self._geom, _lib.ptr(plan.buf), plan.nbytes, _lib.ptr(sino_vcr), _dtype_code(sino_vcr),
            int(sino_vcr.shape[0]), _lib.ptr(weights), plan.npix, _lib.ptr(views), nv, int(coeff_power),
            float(scale), int(accumulate), _lib.ptr(out), _lib.ptr(ws), wsb, _lib.stream()))
        return out

    def _fwd_rows(self, vox, pix, views=None):
        """(V', R, C) float32 projection: the channel-major forward, then one
        transpose (bitwise the same sums; the channel-major kernel's sliding
        window beats the row-major register tile by more than the transpose costs)."""
        return self.from_channel_major(self._fwd(vox, pix, views, layout=_lib.LAYOUT_VCR))

    def _back(self, sino_vcr, pix, views=None, coeff_power=1, weights=None, scale=1.0, out=None, accumulate=False):
        """Back project a channel-major sinogram (f32, bf16 or f16) to cylinders (npix, Nz)."""
        nz = self.get_params('recon_shape')[2]
        npix = int(pix.numel())
        nv = sino_vcr.shape[0] if views is None else int(views.numel())
        if out is None:
            out = torch.empty((npix, nz), dtype=torch.float32, device=sino_vcr.device)
        wsb = int(_lib.lib().mbx_back_workspace_bytes(self._geom, npix, nv))
        ws = torch.empty(wsb, dtype=torch.uint8, device=sino_vcr.device) if wsb > 0 else None
        _lib.check(_lib.lib().mbx_back_project(
            self._geom, _lib.ptr(self._angles()), _lib.ptr(sino_vcr), _dtype_code(sino_vcr), int(sino_vcr.shape[0]),
            _lib.ptr(weights), _lib.ptr(pix), npix, _lib.ptr(views), nv, int(coeff_power), float(scale),
            int(accumulate), _lib.ptr(out), _lib.ptr(ws), wsb, _lib.stream()))
        return out

    def _dot(self, a, b=None, w=None):
        """FP64 sum(a*b*w) on the device (deterministic); returns a Python float."""
        out = torch.zeros(1, dtype=torch.float64, device=a.device)
        ws = torch.empty(int(_lib.lib().mbx_dot_workspace_bytes()), dtype=torch.uint8, device=a.device)
        adt = _dtype_code(a)
        bdt = adt if b is None else _dtype_code(b)
        _lib.check(_lib.lib().mbx_dot(_lib.ptr(a), adt, _lib.ptr(b), bdt, _lib.ptr(w), int(a.numel()),
                                      _lib.ptr(out), _lib.ptr(ws), int(ws.numel()), _lib.stream()))
        return float(out.item())

    # ------------------------------------------------------------------ public projections
    def sparse_forward_project(self, voxel_values, pixel_indices, view_indices=None, output_device=None):
        """(V', R, C) projection of the cylinders voxel_values (npix, Nz) at pixel_indices."""
        nz = self.get_params('recon_shape')[2]
        vox = self._tensor(voxel_values)
        pix = self._indices(pixel_indices, device_check=False)
        if vox.dim() != 2 or vox.shape[0] != pix.numel() or vox.shape[1] != nz:
            raise ValueError('voxel_values must have shape[0:2] = (num_indices, num_slices)')
        V = self.get_params('sinogram_shape')[0]
        views = None if view_indices is None else self._indices(view_indices, V, 'view_indices')
        return self._output(self._fwd_rows(vox, pix, views), output_device)

    def sparse_back_project(self, sinogram, pixel_indices, view_indices=None, coeff_power=1, output_device=None):
        """(npix, Nz) back projection; view_indices select rows of the passed sinogram
        (and their angles), as in mbirjax/tomography_model.py:661-663."""
        pix = self._indices(pixel_indices, device_check=False)
        sino = self.to_channel_major(sinogram)
        V = self.get_params('sinogram_shape')[0]
        views = None if view_indices is None else self._indices(view_indices, min(V, sino.shape[0]),
                                                                   'view_indices')
        if views is None and sino.shape[0] != V:
            raise ValueError('sinogram must hold all views when view_indices is None')
        return self._output(self._back(sino, pix, views, coeff_power=coeff_power), output_device)

    def get_voxels_at_indices(self, recon, indices):
        nz = self.get_params('recon_shape')[2]
        r = self._tensor(recon).reshape(-1, nz)
        return r[self._indices(indices).long()]

    def forward_project(self, recon):
        recon_shape = self.get_params('recon_shape')
        full = vcd_utils.gen_full_indices(recon_shape, use_ror_mask=self.use_ror_mask)
        vox = self.get_voxels_at_indices(recon, full)
        return self.sparse_forward_project(vox, full)

    def back_project(self, sinogram):
        recon_shape = self.get_params('recon_shape')
        full = vcd_utils.gen_full_indices(recon_shape, use_ror_mask=self.use_ror_mask)
        cyl = self.sparse_back_project(sinogram, full)
        recon = torch.zeros((recon_shape[0] * recon_shape[1], recon_shape[2]), dtype=torch.float32,
                            device=cyl.device)
        recon[torch.as_tensor(full, device=cyl.device).long()] = cyl
        return recon.reshape(recon_shape)

    def compute_hessian_diagonal(self, weights=None, output_device=None):
        """diag(A^T W A) at every rows*cols pixel (coeff_power=2 back projection)."""
        sinogram_shape, recon_shape = self.get_params(['sinogram_shape', 'recon_shape'])
        dev = self._device()
        if weights is None:
            w = torch.ones((sinogram_shape[0], sinogram_shape[2], sinogram_shape[1]), dtype=torch.float32, device=dev)
        else:
            if tuple(weights.shape) != tuple(sinogram_shape):
                raise ValueError('Weights must be constant or an array compatible with sinogram'
                                 '\nGot weights.shape = {}, but sinogram.shape = {}'.format(weights.shape,
                                                                                           sinogram_shape))
            w = self.to_channel_major(weights)
        h = self._hessian_vcr(w)
        return self._output(h.reshape(recon_shape), output_device)

    def _hessian_vcr(self, w_vcr):
        rows, cols, nz = self.get_params('recon_shape')
        pix = torch.arange(rows * cols, dtype=torch.int32, device=w_vcr.device)
        return self._back(w_vcr, pix, coeff_power=2)

    # ------------------------------------------------------------------ regularisation
    def auto_set_regularization_params(self, sinogram, weights=None):
        """sigma_y, sigma_x, sigma_prox from <= 20 sub-sampled views, with the
        reference's float32 NumPy arithmetic (mbirjax/tomography_model.py:756-932)."""
        if self.get_params('auto_regularize_flag'):
            V = sinogram.shape[0]
            step = V // np.minimum(20, V)
            small = _as_host_f32(sinogram[::step])
            small_w = 1 if weights is None else _as_host_f32(weights[::step])
            indicator = self._get_sino_indicator(small)
            self.auto_set_sigma_y(small, indicator, small_w)
            recon_std = self._get_estimate_of_recon_std(small, indicator)
            self.auto_set_sigma_x(recon_std)
            self.auto_set_sigma_prox(recon_std)
        names = ['sigma_y', 'sigma_x', 'sigma_prox']
        return namedtuple('RegularizationParams', names)(*[float(v) for v in self.get_params(names)])._asdict()

    def auto_set_sigma_y(self, sinogram, sino_indicator, weights=1):
        snr_db = self.get_params('snr_db')
        mag = self.get_magnification()
        delta_voxel, delta_det_channel = self.get_params(['delta_voxel', 'delta_det_channel'])
        signal_rms = float(np.average(weights * sinogram ** 2, None, sino_indicator) ** 0.5)
        rel_noise_std = 10 ** (-snr_db / 20)
        rel_pitch = delta_voxel / (delta_det_channel / mag)
        sigma_y = np.float32(rel_noise_std * signal_rms * (rel_pitch ** 0.5))
        self.set_params(no_warning=True, sigma_y=sigma_y, auto_regularize_flag=True)

    def auto_set_sigma_x(self, recon_std):
        sigma_x = np.float32(0.2 * (2 ** self.get_params('sharpness')) * recon_std)
        self.set_params(no_warning=True, sigma_x=sigma_x, auto_regularize_flag=True)

    def auto_set_sigma_prox(self, recon_std):
        sigma_prox = np.float32(0.2 * (2 ** self.get_params('sharpness')) * recon_std)
        self.set_params(no_warning=True, sigma_prox=sigma_prox, auto_regularize_flag=True)

    @staticmethod
    def _get_sino_indicator(sinogram):
        percent_noise_floor = 5.0
        return np.int8(sinogram > (0.01 * percent_noise_floor) * np.mean(np.fabs(sinogram)))

    def _get_estimate_of_recon_std(self, sinogram, sino_indicator):
        delta_det_channel, delta_voxel, recon_shape = self.get_params(['delta_det_channel', 'delta_voxel',
                                                                       'recon_shape'])
        mag = self.get_magnification()
        typical = np.average(np.abs(sinogram), weights=sino_indicator)
        path_space = (2 * recon_shape[0] * recon_shape[1]) / (recon_shape[0] + recon_shape[1]) * delta_voxel
        path_sino = sinogram.shape[-1] * delta_det_channel / mag
        return typical / np.minimum(path_space, path_sino)

    # ------------------------------------------------------------------ direct recon
    def direct_filter(self, sinogram, filter_name="ramp", view_batch_size=DIRECT_RECON_VIEW_BATCH_SIZE):
        warnings.warn('direct_filter not implemented for TomographyModel.')
        return torch.zeros_like(self._tensor(sinogram))

    def direct_recon(self, sinogram, filter_name="ramp", view_batch_size=DIRECT_RECON_VIEW_BATCH_SIZE):
        return self.back_project(self.direct_filter(sinogram, filter_name=filter_name,
                                                    view_batch_size=view_batch_size))

    @staticmethod
    def _ramp_conv_rows(sino, recon_filter):
        """'valid' convolution of every detector row (last axis, length C) with
        the (2C-1)-tap filter: out[i] = sum_j row[j] h(i - j).  Up to 2048
        channels this is one float32 GEMM with the C x C Toeplitz matrix
        H[j, i] = h(i - j) (512^3 benchmark: 4 ms, against 27 ms for the rFFT
        route and 2.4 s for the FFT library's first call in a process);
        beyond that, or when float32 GEMMs may use TF32, an rFFT."""
        C = sino.shape[-1]
        if C <= 2048 and not torch.backends.cuda.matmul.allow_tf32:
            i = torch.arange(C, device=sino.device)
            H = recon_filter[(i[None, :] - i[:, None]) + (C - 1)].to(sino.dtype).contiguous()
            return torch.matmul(sino, H)
        L = 1 << int(math.ceil(math.log2(3 * C - 2)))
        fr = torch.fft.rfft(sino, n=L, dim=-1)
        fh = torch.fft.rfft(recon_filter, n=L)
        full = torch.fft.irfft(fr * fh, n=L, dim=-1)
        return full[..., C - 1:2 * C - 1].contiguous()

    # ------------------------------------------------------------------ VCD
    def recon(self, sinogram, weights=None, init_recon=None, max_iterations=15, stop_threshold_change_pct=0.2,
              first_iteration=0, compute_prior_loss=False, num_iterations=None, logfile_path='./logs/recon.log',
              print_logs=True, precision='fp32'):
        """MBIR reconstruction with multi-granular VCD (mbirjax/tomography_model.py:966-1084).

        precision: 'fp32' (plain); 'mixed' (error sinogram accumulated in float32,
        delta sinogram A*delta stored in bfloat16, FP64 step-size and cost
        reductions); 'fp16' (as 'mixed' with the error sinogram stored in IEEE
        float16, the problem scaled by a power of two so that it stays in range:
        half the residual traffic of 'mixed' within NRMSE 1e-3 of FP64); or
        'bf16' (experimental: the error sinogram in bfloat16, whose 8-bit
        significand drops late updates below half an ulp of the residual --
        NRMSE ~3.4e-3 of FP64 on the 512^3 benchmark's sub-problem, above the
        1e-3 bar the other modes meet).
        """
        if num_iterations is not None:
            warnings.warn('num_iterations has been deprecated and will be removed in a future release.\n'
                          'In the current run, the value of num_iterations will be used to set max_iterations.')
            max_iterations = num_iterations
        if first_iteration == 0 or self.logger is None:
            self.setup_logger(logfile_path=logfile_path, print_logs=print_logs)
        self.logger.info('GPU used for: full (MI355X kernels)')
        if compute_prior_loss:
            self.logger.warning('Computing the prior loss on every iteration uses significant memory and computing '
                                'power.\nSet compute_prior_loss=False for most applications aside from debugging and '
                                'demos.')
        regularization_params = self.auto_set_regularization_params(sinogram, weights=weights)
        recon_shape, granularity = self.get_params(['recon_shape', 'granularity'])
        partitions = vcd_utils.gen_set_of_pixel_partitions(recon_shape, granularity, use_ror_mask=self.use_ror_mask)
        partition_sequence = vcd_utils.gen_partition_sequence(self.get_params('partition_sequence'),
                                                              max_iterations=max_iterations)
        partition_sequence = partition_sequence[first_iteration:]
        recon, loss_vectors = self.vcd_recon(sinogram, partitions, partition_sequence, stop_threshold_change_pct,
                                             weights=weights, init_recon=init_recon,
                                             compute_prior_loss=compute_prior_loss, first_iteration=first_iteration,
                                             precision=precision)
        partition_sequence = [int(v) for v in partition_sequence]
        fm_rmse = [float(v) for v in loss_vectors[0]]
        prior_loss = [float(v) for v in loss_vectors[1]] if compute_prior_loss else [0]
        stop_pct = [100 * float(v) for v in loss_vectors[2]]
        alpha_values = [float(v) for v in loss_vectors[3]]
        recon_params = ReconParams(len(fm_rmse), granularity, partition_sequence, fm_rmse, prior_loss,
                                   regularization_params, stop_pct, alpha_values)._asdict()
        if logfile_path:
            self.logger.info('Logs written to {}'.format(os.path.abspath(logfile_path)))
        for h in list(self.logger.handlers):
            h.flush()
        notes = 'Reconstruction completed: {}\n\n'.format(datetime.datetime.now())
        return recon, self.get_recon_dict(recon_params, notes=notes)

    def get_recon_dict(self, recon_params=None, notes=None, save_log=True, save_model=True):
        d = dict()
        d['recon_params'] = "# Recon params not saved." if recon_params is None else recon_params
        d['recon_log'] = "# Log info not saved." if (self.log_buffer is None or not save_log) else \
            self.log_buffer.getvalue()
        d['notes'] = '# No notes saved' if notes is None else notes
        d['model_params'] = self.params.copy() if save_model else '# Model not saved'
        return d

    def _qggmrf_params(self):
        nbr, sigma_x, p, q, T = self.get_params(['qggmrf_nbr_wts', 'sigma_x', 'p', 'q', 'T'])
        return (_qggmrf.get_b_from_nbr_wts(nbr), float(sigma_x), float(p), float(q), float(T))

    def vcd_setup(self, sinogram, weights=None, init_recon=None, prox_input=None, precision='fp32',
                  max_subset_size=None, channel_major=False):
        """Initial VCD state on the device: e = y - alpha0 A x0 with the optimal
        scale alpha0 of the initial recon, the forward-model Hessian diagonal and
        the subset-update workspace (mbirjax/tomography_model.py:1110-1192).
        With channel_major=True, sinogram (and weights) are already (views, C, R)
        float32 device tensors."""
        if precision not in PRECISIONS:
            raise ValueError("precision must be one of {}".format(sorted(PRECISIONS)))
        if precision == 'fp16' and prox_input is not None:
            # the proximal prior's curvature is sigma_prox**2 (tomography_model.py:1390 of the reference),
            # which does not scale with the problem, so the scaled float16 run would be another update
            raise ValueError("precision 'fp16' is not available for the proximal map")
        dev = self._device()
        recon_shape = self.get_params('recon_shape')
        rows, cols, nz = recon_shape
        V, R, C = self.get_params('sinogram_shape')
        if channel_major:
            if tuple(sinogram.shape) != (V, C, R):
                raise ValueError('channel-major sinogram must have shape {}'.format((V, C, R)))
            y = self._tensor(sinogram)
            w = None if weights is None else self._tensor(weights)
        else:
            y = self.to_channel_major(sinogram)
            w = None if weights is None else self.to_channel_major(weights)
        if init_recon is None:
            self._log('info', 'Starting direct recon for initial reconstruction')
            init_recon = self.direct_recon(self.from_channel_major(y))
        zero_init = isinstance(init_recon, int) and init_recon == 0
        if zero_init:
            init_recon = torch.zeros(recon_shape, dtype=torch.float32, device=dev)
        init_recon = self._tensor(init_recon)
        if tuple(init_recon.shape) != tuple(recon_shape):
            msg = "init_recon does not have the correct shape. \nExpected {}, but got shape {} for init_recon " \
                  "shape.".format(recon_shape, tuple(init_recon.shape))
            self._log('error', msg)
            raise ValueError(msg)

        self._log('info', 'Initializing error sinogram')
        full = vcd_utils.gen_full_indices(recon_shape, use_ror_mask=self.use_ror_mask)
        full_d = torch.as_tensor(full, device=dev)
        e_dtype, delta_dtype = PRECISIONS[precision]
        e32 = y.clone()
        alpha0 = 1.0
        if not zero_init:  # (A 0 = 0: alpha0 = 1 and e = y)
            Ax = self._fwd(init_recon.reshape(-1, nz)[full_d.long()].contiguous(), full_d, layout=_lib.LAYOUT_VCR)
            nrm = self._dot(Ax, None, w)
            alpha0 = self._dot(Ax, y, w) / nrm if nrm > 0 else 1.0
            alpha_dev = torch.tensor([float(np.float32(alpha0))], dtype=torch.float64, device=dev)
            _lib.check(_lib.lib().mbx_sino_axpy(_lib.ptr(e32), _F32, _lib.ptr(Ax), _F32, _lib.ptr(alpha_dev),
                                                int(e32.numel()), _lib.stream()))
            del Ax
        # float16 storage: solve the problem scaled by a power of two s (data,
        # recon, sigma_y, sigma_x and sigma_prox all divided by s) that puts the
        # largest data or residual magnitude in [2^11, 2^12), far below the
        # float16 overflow at 65504 and far above its subnormals.  Division by a
        # power of two is exact, and the VCD update is equivariant under the
        # scaling (delta scales by 1/s, alpha and the costs are unchanged), so
        # only the storage rounding differs from the unscaled run.
        scale = 1.0
        if e_dtype == _F16:
            peak = max(float(torch.max(torch.abs(y)).item()), float(torch.max(torch.abs(e32)).item()))
            if peak > 0 and math.isfinite(peak):
                scale = 2.0 ** (math.floor(math.log2(peak)) - 11)
                e32.mul_(1.0 / scale)
        if e_dtype != _F32:
            e = torch.empty(e32.shape, dtype=_torch_dtype(e_dtype), device=dev)
            _lib.check(_lib.lib().mbx_sino_transpose(_lib.ptr(e32), _F32, _lib.LAYOUT_VCR, _lib.ptr(e), e_dtype,
                                                     _lib.LAYOUT_VCR, V, R, C, _lib.stream()))
            del e32
        else:
            e = e32
        flat_recon = (np.float32(alpha0) * init_recon).reshape(-1, nz).contiguous()
        if scale != 1.0:
            flat_recon.mul_(1.0 / scale)

        if prox_input is not None:
            prox_input = self._tensor(prox_input)
            if tuple(prox_input.shape) != tuple(recon_shape):
                msg = "prox_input does not have the correct size. \nExpected {}, but got shape {} for prox_input " \
                      "shape.".format(recon_shape, tuple(prox_input.shape))
                self._log('error', msg)
                raise ValueError(msg)
            prox_input = prox_input.reshape(-1, nz).contiguous()
            if scale != 1.0:
                prox_input = prox_input * (1.0 / scale)

        self._log('info', 'Computing Hessian diagonal')
        w_h = w if w is not None else torch.ones_like(y)
        fm_hessian = self._hessian_vcr(w_h)
        del w_h, y
        if max_subset_size is None:
            max_subset_size = rows * cols
        return VCDState(self, e, e_dtype, delta_dtype, w, flat_recon, fm_hessian, prox_input, precision,
                        max_subset_size, scale)

    def _log(self, level, msg):
        if self.logger is not None:
            getattr(self.logger, level)(msg)

    def vcd_recon(self, sinogram, partitions, partition_sequence, stop_threshold_change_pct, weights=None,
                  init_recon=None, prox_input=None, compute_prior_loss=False, first_iteration=0, precision='fp32'):
        """Multi-granular VCD on the device (mbirjax/tomography_model.py:1086-1255).
        Returns (recon (rows, cols, Nz), (fm_rmse, prior_loss, nmae_update, alpha_values))."""
        recon_shape = self.get_params('recon_shape')
        nS_max = max(int(p.shape[1]) for p in partitions)
        state = self.vcd_setup(sinogram, weights=weights, init_recon=init_recon, prox_input=prox_input,
                               precision=precision, max_subset_size=nS_max)
        dev = state.e.device
        parts_dev = [torch.as_tensor(np.ascontiguousarray(p, dtype=np.int32), device=dev) for p in partitions]
        verbose = self.get_params('verbose')
        self._log('info', 'Starting VCD iterations')
        max_iters = int(np.size(partition_sequence))
        fm_rmse = np.zeros(max_iters)
        pm_loss = np.zeros(max_iters)
        nmae_update = np.zeros(max_iters)
        alpha_values = np.zeros(max_iters)
        num_iters = 0
        # partitions swept more than once keep their subsets' projector setups
        uses = np.bincount(np.asarray(partition_sequence, dtype=np.int64), minlength=len(parts_dev))
        for i in range(max_iters):
            k = int(partition_sequence[i])
            part = parts_dev[k]
            order = np.random.permutation(part.shape[0])
            ell1, alpha_sum = state.iterate(part, order, cache=bool(uses[k] > 1))
            fm_rmse[i] = state.forward_model_rmse()
            nmae_update[i] = ell1 / float(torch.sum(torch.abs(state.flat_recon), dtype=torch.float64).item())
            alpha_values[i] = alpha_sum / part.shape[0]
            if verbose >= 1:
                msg = '\nAfter iteration {} of a max of {}: Pct change={:.4f}, Forward loss={:.4f}'.format(
                    i + first_iteration, max_iters + first_iteration, 100 * nmae_update[i], fm_rmse[i])
                if compute_prior_loss:
                    n_e, n_x = state.e.numel(), state.flat_recon.numel()
                    pm_loss[i] = _qggmrf.qggmrf_loss(state.flat_recon.reshape(recon_shape), state.qp) / n_x
                    total = (fm_rmse[i] * n_e + pm_loss[i] * n_x) / (0.5 * (n_e + n_x))
                    msg += ', Prior loss={:.4f}, Weighted total loss={:.4f}'.format(pm_loss[i], total)
                self._log('info', msg)
                self._log('info', 'Relative step size (alpha)={:.2f}, Error sino RMSE={:.4f}'.format(
                    alpha_values[i], state.error_rmse()))
                self._log('info', 'Number subsets = {}'.format(part.shape[0]))
            num_iters += 1
            if nmae_update[i] < stop_threshold_change_pct / 100:
                self._log('warning', 'Change threshold stopping condition reached')
                break
        self.last_error_sinogram = state.error_sinogram()
        return self.reshape_recon(state.recon()), (fm_rmse[:num_iters], pm_loss[:num_iters],
                                                      nmae_update[:num_iters], alpha_values[:num_iters])

    @staticmethod
    def get_forward_model_loss(error_sinogram, sigma_y, weights=None, normalize=True):
        e = error_sinogram if isinstance(error_sinogram, torch.Tensor) else torch.as_tensor(np.asarray(error_sinogram))
        e = e.to(torch.float64)
        if weights is None:
            w, avg = 1.0, 1.0
        else:
            w = weights if isinstance(weights, torch.Tensor) else torch.as_tensor(np.asarray(weights))
            w = w.to(device=e.device, dtype=torch.float64)
            avg = torch.mean(w)
        if normalize:
            return float(torch.sqrt((1.0 / sigma_y ** 2) * torch.mean((e * e) * (w / avg))))
        return float((1.0 / (2 * sigma_y ** 2)) * torch.sum((e * e) * w))

    def prox_map(self, prox_input, sinogram, weights=None, init_recon=None, stop_threshold_change_pct=0.2,
                 max_iterations=3, first_iteration=0, logfile_path='./logs/recon.log', print_logs=True):
        """Proximal map with prior (1/2 sigma_prox^2)||x - prox_input||^2 (mbirjax/tomography_model.py:1660-1697)."""
        recon_shape, granularity = self.get_params(['recon_shape', 'granularity'])
        partitions = vcd_utils.gen_set_of_pixel_partitions(recon_shape, granularity)
        seq = vcd_utils.gen_partition_sequence(self.get_params('partition_sequence'),
                                               max_iterations=max_iterations)[first_iteration:]
        if first_iteration == 0 or self.logger is None:
            self.setup_logger(logfile_path=logfile_path, print_logs=print_logs)
        return self.vcd_recon(sinogram, partitions, seq, stop_threshold_change_pct, weights=weights,
                              init_recon=init_recon, prox_input=prox_input, first_iteration=first_iteration)

    def gen_weights(self, sinogram, weight_type):
        """Noise-model weights (mbirjax/tomography_model.py:1727-1764)."""
        s = self._tensor(sinogram)
        if weight_type == 'unweighted':
            return torch.ones_like(s)
        if weight_type == 'transmission':
            return torch.exp(-s)
        if weight_type == 'transmission_root':
            return torch.exp(-s / 2)
        if weight_type == 'emission':
            return 1.0 / (torch.abs(s) + 0.1)
        raise Exception("gen_weights: undefined weight_type {}".format(weight_type))

    def gen_modified_3d_sl_phantom(self):
        """Low-dynamic-range 3D Shepp-Logan phantom of the model's recon shape
        (mbirjax/tomography_model.py:1766-1775)."""
        return utilities.generate_3d_shepp_logan_low_dynamic_range(self.get_params('recon_shape'))

    def reshape_recon(self, recon):
        return recon.reshape(self.get_params('recon_shape'))

    def scale_recon_shape(self, row_scale=1.0, col_scale=1.0, slice_scale=1.0):
        r, c, s = self.get_params('recon_shape')
        nr, nc, ns = int(r * row_scale), int(c * col_scale), int(s * slice_scale)
        self.set_params(recon_shape=(nr, nc, ns))
        return nr - r, nc - c, ns - s

    # ------------------------------------------------------------------ save / load
    def to_file(self, filename):
        return self.save_params(self.params, filename)

    @classmethod
    def from_file(cls, source):
        required_names = cls.get_required_param_names()
        required, params = ParameterHandler.load_param_dict(source, required_names, values_only=True)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            model = cls(**required)
            model.set_params(**params)
        return model


class ForwardPlan:
    """A device buffer from mbx_forward_plan with the counts it was built for."""

    def __init__(self, buf, nbytes, npix, nviews):
        self.buf, self.nbytes, self.npix, self.nviews = buf, int(nbytes), int(npix), int(nviews)


class VCDState:
    """Device state of one VCD reconstruction: the recon (rows*cols, Nz), the
    error sinogram e = y - A x in channel-major layout (float32, float16 or
    bfloat16 by precision), weights, Hessian diagonal and workspace.  With
    float16 storage the state holds the problem scaled by 1/scale (a power of
    two): recon() and error_sinogram() undo it.  iterate() runs one partition
    through mbx_vcd_subset_update with no host syncs; with cache=True each
    subset's projector setup (geometry table, bins, forward records:
    mbx_forward_plan) is built on its first update and reused by
    mbx_vcd_subset_update_planned, within half the device memory that was free
    when the state was made."""

    def __init__(self, model, e, e_dtype, delta_dtype, w, flat_recon, fm_hessian, prox_input, precision,
                 max_subset_size, scale=1.0):
        self.model = model
        self.e = e
        self.e_dtype = e_dtype
        self.delta_dtype = delta_dtype
        self.w = w
        self.flat_recon = flat_recon
        self.fm_hessian = fm_hessian
        self.prox_input = prox_input
        self.precision = precision
        self.scale = float(scale)
        sigma_y = float(model.get_params('sigma_y')) / self.scale
        self.fm_constant = 1.0 / (sigma_y ** 2.0)
        b, sigma_x, p, q, T = model._qggmrf_params()
        self.qp = (b, sigma_x / self.scale, p, q, T)
        self.qstruct = _lib.qggmrf_struct(*self.qp)
        self.sigma_prox = float(model.get_params('sigma_prox')) / self.scale
        self.positivity = bool(model.get_params('positivity_flag'))
        self.wsb = int(_lib.lib().mbx_vcd_workspace_bytes(model._geom, int(max_subset_size), self.delta_dtype))
        self.ws = torch.empty(self.wsb, dtype=torch.uint8, device=e.device)
        self.stats = torch.zeros(4, dtype=torch.float64, device=e.device)
        self.avg_w = 1.0 if w is None else float(torch.mean(w, dtype=torch.float64).item())
        # forward plans of subsets: (partition data pointer, row) -> (partition, its version, plan or None)
        self.plans = {}
        self.plan_bytes = 0
        self.plan_budget = torch.cuda.mem_get_info(e.device)[0] // 2

    def recon(self):
        """The flat recon (rows*cols, Nz) in the caller's units."""
        return self.flat_recon if self.scale == 1.0 else self.flat_recon * self.scale

    def error_sinogram(self):
        """The error sinogram (views, C, R) as float32 in the caller's units."""
        e = self.e.float()
        return e if self.scale == 1.0 else e * self.scale

    def update_subset(self, subset, plan=None):
        """One subset update (vcd_subset_updater); subset is an int32 device tensor,
        plan (optional) its ForwardPlan over all views."""
        m = self.model
        common = (m._geom, _lib.ptr(m._angles()), _lib.ptr(self.flat_recon), _lib.ptr(self.e), self.e_dtype,
                  _lib.ptr(self.w), _lib.ptr(self.fm_hessian), _lib.ptr(subset), int(subset.numel()),
                  self.fm_constant, self.qstruct, _lib.ptr(self.prox_input), self.sigma_prox, int(self.positivity),
                  self.delta_dtype)
        if plan is None:
            _lib.check(_lib.lib().mbx_vcd_subset_update(*common, _lib.ptr(self.stats), _lib.ptr(self.ws), self.wsb,
                                                        _lib.stream()))
        else:
            _lib.check(_lib.lib().mbx_vcd_subset_update_planned(*common, _lib.ptr(plan.buf), plan.nbytes,
                                                                _lib.ptr(self.stats), _lib.ptr(self.ws), self.wsb,
                                                                _lib.stream()))

    def _plan(self, part, s):
        """The forward plan of subset part[s] (built on its first use), or None when
        the subset spans more than one geometry-table chunk or the cache is full."""
        key = (part.data_ptr(), int(s))
        hit = self.plans.get(key)
        if hit is not None and hit[0] is part and hit[1] == part._version:
            return hit[2]
        if hit is not None and hit[2] is not None:
            self.plan_bytes -= hit[2].nbytes
        self.plans.pop(key, None)
        subset = part[int(s)]
        V = self.model.get_params('sinogram_shape')[0]
        nb = int(_lib.lib().mbx_forward_workspace_bytes(self.model._geom, int(subset.numel()), V))
        if self.plan_bytes + nb > self.plan_budget:
            return None
        try:
            plan = self.model._fwd_plan(subset)
        except RuntimeError:  # more pixels than one geometry-table chunk
            plan = None
        self.plans[key] = (part, part._version, plan)
        if plan is not None:
            self.plan_bytes += plan.nbytes
        return plan

    def iterate(self, part, order, cache=True):
        """All subsets of a partition (K, nS) in the given order (vcd_partition_iterator);
        returns (sum |alpha delta|, sum alpha) -- one host sync.  cache: reuse
        each subset's projector setup across calls with this partition tensor
        (bitwise the same results)."""
        self.stats.zero_()
        for s in order:
            self.update_subset(part[int(s)], self._plan(part, s) if cache else None)
        ell1, alpha_sum, _, _ = self.stats.cpu().tolist()
        return ell1, alpha_sum

    def forward_model_rmse(self):
        n = self.e.numel()
        return math.sqrt(self.fm_constant * self.model._dot(self.e, None, self.w) / (n * self.avg_w))

    def error_rmse(self):
        return self.scale * math.sqrt(self.model._dot(self.e) / self.e.numel())
