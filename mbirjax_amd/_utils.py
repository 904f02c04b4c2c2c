"""Parameter defaults of the TomographyModel API (names, values and recompile
flags of mbirjax/_utils.py:27-58, kept identical so that set_params/get_params
behave the same)."""
import copy
from dataclasses import dataclass
from typing import Any

FILE_FORMAT_NUMBER = 1.0

_geometry_types_for_tests = ['parallel', 'cone', 'translation']


@dataclass
class Param:
    val: Any
    recompile_flag: bool = True

    def __repr__(self):
        return f"Param(val={self.val}, recompile_flag={self.recompile_flag})"


_forward_model_defaults_dict = {
    'geometry_type': Param(None, False),
    'file_format': Param(FILE_FORMAT_NUMBER, False),
    'sinogram_shape': Param(None, True),
    'delta_det_channel': Param(1.0, True),
    'delta_det_row': Param(1.0, True),
    'det_row_offset': Param(0.0, True),
    'det_channel_offset': Param(0.0, True),
    'sigma_y': Param(1.0, False),
}

_recon_model_defaults_dict = {
    'recon_shape': Param(None, True),
    'delta_voxel': Param(None, True),
    'sigma_x': Param(1.0, False),
    'sigma_prox': Param(1.0, False),
    'p': Param(2.0, False),
    'q': Param(1.2, False),
    'T': Param(1.0, False),
    'qggmrf_nbr_wts': Param([1.0, 1.0, 1.0], False),  # row, col, slice neighbour weights
}

_reconstruction_defaults_dict = {
    'auto_regularize_flag': Param(True, False),
    'positivity_flag': Param(False, False),
    'snr_db': Param(30.0, False),
    'sharpness': Param(1.0, False),
    'granularity': Param([1, 2, 4, 8, 16, 32, 64, 128, 256], False),
    'partition_sequence': Param([0, 2, 4, 6, 7], False),
    'verbose': Param(1, False),
    'use_gpu': Param('automatic', True),
}

headings = ['Forward model parameters', 'Recon parameters', 'Reconstruction parameters']
dicts = [_forward_model_defaults_dict, _recon_model_defaults_dict, _reconstruction_defaults_dict]

recon_defaults_dict = dict()
for _d in dicts:
    recon_defaults_dict = {**recon_defaults_dict, **_d}

all_param_keys = list(recon_defaults_dict.keys())


def get_default_params():
    return copy.deepcopy(recon_defaults_dict)
