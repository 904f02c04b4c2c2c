"""ParameterHandler: the parameter store behind set_params / get_params.

Same semantics as mbirjax/parameter_handler.py:433-559: unknown names raise
ValueError (unless no_warning), numpy scalars are normalised to Python
scalars, setting sigma_y/sigma_x/sigma_prox disables auto-regularisation and
setting sharpness/snr_db re-enables it, and the return value says whether a
geometry parameter changed (the model then re-plans its projectors).
Models are saved as YAML in the reference's format ({name: {val, recompile_flag}},
arrays as ':ARRAY:' strings of %.7f values with a 'shape' entry) and read with
PyYAML's safe loader.
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
        if self._class_name(geometry_type) != self._class_name(str(type(self))):
            raise ValueError('Parameters are associated with {}, but the current model is {}'.format(
                geometry_type, str(type(self))))

    @staticmethod
    def _class_name(type_str):
        """'ConeBeamModel' from "<class 'mbirjax.cone_beam.ConeBeamModel'>" (or from the
        same string naming this package): files written by either load here."""
        return str(type_str).strip("<>'").split('.')[-1].strip("'")

    # ------------------------------------------------------------------ YAML
    # The on-disk format is the reference's (mbirjax/parameter_handler.py:195-249):
    # {name: {val, recompile_flag}} with every array written as the string
    # ':ARRAY:' followed by its values in %.7f and a 'shape' entry beside 'val',
    # numpy scalars as Python scalars and lists as tuples (YAML sequences).

    @staticmethod
    def _plain(val):
        """numpy scalars as Python scalars, lists and tuples as tuples (recursively)."""
        if isinstance(val, np.generic):
            return val.item()
        if isinstance(val, (list, tuple)):
            return tuple(ParameterHandler._plain(v) for v in val)
        return val

    @staticmethod
    def serialize_parameter(param):
        """A parameter as the YAML entry the reference writes for it."""
        val = param.val
        if hasattr(val, 'detach'):
            val = val.detach().cpu().numpy()
        if isinstance(val, np.ndarray):
            return {'val': ParameterHandler.array_prefix + " ".join("%.7f" % x for x in val.ravel()),
                    'shape': tuple(int(n) for n in val.shape), 'recompile_flag': bool(param.recompile_flag)}
        return {'val': ParameterHandler._plain(val), 'recompile_flag': bool(param.recompile_flag)}

    @staticmethod
    def deserialize_parameter(entry):
        """A YAML entry (as the reference writes it) as a Param: ':ARRAY:' strings
        become float64 arrays of the entry's shape, sequences become tuples."""
        val = entry['val']
        flag = entry.get('recompile_flag', True)
        if isinstance(val, str) and val.startswith(ParameterHandler.array_prefix):
            body = val[len(ParameterHandler.array_prefix):].replace('[', ' ').replace(']', ' ')
            arr = np.array(body.split(), dtype=np.float64)
            if 'shape' in entry:
                arr = arr.reshape(tuple(entry['shape']))
            return Param(arr, flag)
        return Param(ParameterHandler._plain(val), flag)

    @staticmethod
    def save_params(params, filename=None):
        """YAML text of the parameters (returned), or written to filename (.yml or .yaml)."""
        out = {k: ParameterHandler.serialize_parameter(p) for k, p in params.items()}
        text = yaml.safe_dump(out, sort_keys=False, default_flow_style=False)
        if filename is None:
            return text
        if not (filename.lower().endswith('.yml') or filename.lower().endswith('.yaml')):
            raise ValueError('Filename must end in .yaml or .yml: ' + filename)
        d = os.path.dirname(filename)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(filename, 'w') as f:
            f.write(text)
        return None

    @staticmethod
    def load_param_dict(source, required_param_names=None, values_only=True):
        """Parameters from a YAML file name (.yml or .yaml), YAML text or a file-like
        object; with required_param_names, (required, others) as the reference's
        load_param_dict returns them."""
        if hasattr(source, 'read'):
            text = source.read()
        elif isinstance(source, str) and os.path.exists(source):
            if not source.lower().endswith(('.yml', '.yaml')):
                raise ValueError('Filename must end in .yml or .yaml')
            with open(source) as f:
                text = f.read()
        elif isinstance(source, str):
            text = source
        else:
            raise TypeError('Invalid source type for load_param_dict. Must be str or file-like object.')
        raw = yaml.safe_load(text)
        params = {k: ParameterHandler.deserialize_parameter(v) for k, v in raw.items()}
        if required_param_names is None:
            return params
        required = {}
        for name in required_param_names:
            if name not in params:
                raise ValueError("The parameters do not hold '{}', which this model requires".format(name))
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
