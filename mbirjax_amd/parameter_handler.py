"""ParameterHandler: the parameter store behind set_params / get_params.

Same semantics as mbirjax/parameter_handler.py:433-559: unknown names raise
ValueError (unless no_warning), numpy scalars are normalised to Python
scalars, setting sigma_y/sigma_x/sigma_prox disables auto-regularisation and
setting sharpness/snr_db re-enables it, and the return value says whether a
geometry parameter changed (the model then re-plans its projectors).
Models are saved as YAML ({name: {val, recompile_flag}}) with PyYAML's safe
loader.
"""
import io
import logging
import os
import warnings

import numpy as np
import yaml

from . import _utils
from ._utils import Param


class ParameterHandler:
    array_prefix = ':ARRAY:'

    def __init__(self):
        self.params = _utils.get_default_params()
        self.logger = None
        self.log_buffer = None

    # ------------------------------------------------------------------ logging
    def setup_logger(self, *, logfile_path="./logs/recon.log", print_logs=True):
        verbose = self.get_params('verbose')
        level = logging.WARNING if verbose < 1 else (logging.INFO if verbose < 2 else logging.DEBUG)
        logger = logging.getLogger(self.__class__.__name__)
        logger.setLevel(level)
        logger.propagate = False
        for h in list(logger.handlers):
            try:
                h.flush()
            finally:
                h.close()
                logger.removeHandler(h)
        self.log_buffer = io.StringIO()
        bh = logging.StreamHandler(self.log_buffer)
        bh.setLevel(level)
        bh.setFormatter(logging.Formatter('%(message)s'))
        logger.addHandler(bh)
        if print_logs:
            ch = logging.StreamHandler()
            ch.setLevel(level)
            ch.setFormatter(logging.Formatter('%(message)s'))
            logger.addHandler(ch)
        if logfile_path:
            d = os.path.dirname(os.path.abspath(logfile_path))
            os.makedirs(d, exist_ok=True)
            fh = logging.FileHandler(logfile_path, mode='w')
            fh.setLevel(level)
            fh.setFormatter(logging.Formatter('%(message)s'))
            logger.addHandler(fh)
        self.logger = logger

    def print_params(self):
        verbose, view_params_name = self.get_params(['verbose', 'view_params_name'])
        print("----")
        for key, entry in self.params.items():
            if verbose < 3 and key == view_params_name:
                continue
            if verbose < 3:
                print("{} = {}".format(key, entry.val))
            else:
                print("{} = {}, recompile_flag = {}".format(key, entry.val, entry.recompile_flag))
        print("----")

    # ------------------------------------------------------------------ values
    @staticmethod
    def normalize_scalar(val):
        if isinstance(val, np.generic):
            return val.item()
        if isinstance(val, (list, tuple)):
            return type(val)(ParameterHandler.normalize_scalar(v) for v in val)
        return val

    def set_params(self, no_warning=False, no_compile=False, **kwargs):
        recompile = False
        regularization_parameter_change = False
        meta_parameter_change = False
        for key, val in kwargs.items():
            recompile_flag = True
            if key in self.params.keys():
                recompile_flag = self.params[key].recompile_flag
            elif not no_warning:
                msg = '{} is not a recognized parameter\nValid parameters are: \n'.format(key)
                for valid_key in self.params.keys():
                    msg += '   {}\n'.format(valid_key)
                raise ValueError(msg)
            self.params[key] = Param(ParameterHandler.normalize_scalar(val), recompile_flag)
            if recompile_flag:
                recompile = True
            elif key in ["sigma_y", "sigma_x", "sigma_prox"]:
                regularization_parameter_change = True
            elif key in ["sharpness", "snr_db"]:
                meta_parameter_change = True
        if not no_warning:
            self.verify_valid_params()
        if regularization_parameter_change and not no_warning:
            self.set_params(auto_regularize_flag=False)
            warnings.warn('You are directly setting regularization parameters, sigma_x, sigma_y or sigma_prox. '
                          'This is an advanced feature that will disable auto-regularization.')
        if meta_parameter_change and self.get_params('auto_regularize_flag') is False:
            self.set_params(auto_regularize_flag=True)
            if not no_warning:
                warnings.warn('You have re-enabled auto-regularization by setting sharpness or snr_db. '
                              'It was previously disabled')
        return bool(recompile and not no_compile)

    @staticmethod
    def get_params_from_dict(param_dict, parameter_names):
        if isinstance(parameter_names, str):
            if parameter_names in param_dict.keys():
                return param_dict[parameter_names].val
            raise NameError('"{}" is not a recognized argument'.format(parameter_names))
        values = []
        for name in parameter_names:
            if name in param_dict.keys():
                values.append(param_dict[name].val)
            else:
                raise NameError('"{}" is not a recognized argument'.format(name))
        return values

    def get_params(self, parameter_names):
        return ParameterHandler.get_params_from_dict(self.params, parameter_names)

    def get_magnification(self):
        raise NotImplementedError('get_magnification is not implemented.')

    def verify_valid_params(self):
        sinogram_shape = self.get_params('sinogram_shape')
        if len(sinogram_shape) != 3:
            raise ValueError("sinogram_shape must be (views, rows, channels). \n"
                             "Got {} for sinogram shape.".format(sinogram_shape))
        geometry_type = self.get_params('geometry_type')
        if geometry_type != str(type(self)):
            raise ValueError('Parameters are associated with {}, but the current model is {}'.format(
                geometry_type, str(type(self))))

    # ------------------------------------------------------------------ YAML
    @staticmethod
    def _to_yaml_value(val):
        if hasattr(val, 'detach'):
            val = val.detach().cpu().numpy()
        if isinstance(val, np.ndarray):
            return ParameterHandler.array_prefix + yaml.safe_dump({'shape': list(val.shape), 'dtype': str(val.dtype),
                                                                   'data': val.ravel().tolist()})
        if isinstance(val, tuple):
            return {'__tuple__': [ParameterHandler._to_yaml_value(v) for v in val]}
        if isinstance(val, list):
            return [ParameterHandler._to_yaml_value(v) for v in val]
        if isinstance(val, float) and not np.isfinite(val):
            return {'__float__': repr(val)}
        return ParameterHandler.normalize_scalar(val)

    @staticmethod
    def _from_yaml_value(val):
        if isinstance(val, str) and val.startswith(ParameterHandler.array_prefix):
            d = yaml.safe_load(val[len(ParameterHandler.array_prefix):])
            return np.array(d['data'], dtype=d['dtype']).reshape(d['shape'])
        if isinstance(val, dict) and '__tuple__' in val:
            return tuple(ParameterHandler._from_yaml_value(v) for v in val['__tuple__'])
        if isinstance(val, dict) and '__float__' in val:
            return float(val['__float__'])
        if isinstance(val, list):
            return [ParameterHandler._from_yaml_value(v) for v in val]
        return val

    @staticmethod
    def save_params(params, filename=None):
        out = {k: {'val': ParameterHandler._to_yaml_value(p.val), 'recompile_flag': bool(p.recompile_flag)}
               for k, p in params.items()}
        text = yaml.safe_dump(out, sort_keys=False)
        if filename is None:
            return text
        if not (filename.endswith('.yml') or filename.endswith('.yaml')):
            raise ValueError('Filename must end in .yaml or .yml: ' + filename)
        with open(filename, 'w') as f:
            f.write(text)
        return None

    @staticmethod
    def load_param_dict(source, required_param_names=None, values_only=True):
        if hasattr(source, 'read'):
            text = source.read()
        elif isinstance(source, str) and (source.endswith('.yml') or source.endswith('.yaml')) and \
                os.path.exists(source):
            with open(source) as f:
                text = f.read()
        else:
            text = source
        raw = yaml.safe_load(text)
        params = {k: Param(ParameterHandler._from_yaml_value(v['val']), v['recompile_flag']) for k, v in raw.items()}
        if required_param_names is None:
            return params
        required = {}
        for name in required_param_names:
            required[name] = params.pop(name)
        if values_only:
            required = {k: v.val for k, v in required.items()}
            params = {k: v.val for k, v in params.items()}
        return required, params

    @staticmethod
    def compare_parameter_handlers(ph1, ph2, atol=1e-6, verbose=False):
        p1, p2 = ph1.params, ph2.params
        if set(p1.keys()) != set(p2.keys()):
            return False
        for k in p1:
            a, b = p1[k].val, p2[k].val
            try:
                if isinstance(a, (str, type(None), bool)) or isinstance(b, (str, type(None), bool)):
                    same = a == b
                else:
                    aa = np.asarray(a.detach().cpu().numpy() if hasattr(a, 'detach') else a, dtype=np.float64)
                    bb = np.asarray(b.detach().cpu().numpy() if hasattr(b, 'detach') else b, dtype=np.float64)
                    same = aa.shape == bb.shape and np.allclose(aa, bb, atol=atol, equal_nan=True)
            except (TypeError, ValueError):
                same = a == b
            if not same:
                if verbose:
                    print('Parameter {} differs: {} vs {}'.format(k, a, b))
                return False
        return True
