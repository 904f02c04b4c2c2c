"""ORACLE -- test infrastructure only (see oracle/__init__.py).

NumPy restatement of the VCD pixel partitions (mbirjax/vcd_utils.py):
  get_2d_ror_mask             vcd_utils.py:66-102
  gen_pixel_partition         vcd_utils.py:176-223
  gen_set_of_pixel_partitions vcd_utils.py:105-124
  gen_partition_sequence      vcd_utils.py:284-307
  gen_full_indices            vcd_utils.py:310-318
The reference draws from NumPy's legacy global RNG in the order
permutation(indices) -> choice(non-final indices, replace=False) per
granularity, so a seeded restatement reproduces its partitions exactly.
Pinned by tests/test_oracle.py: every subset is sorted, the subsets cover the
ROR exactly once apart from the padding drawn from non-final subsets, no
subset repeats a pixel, and the ROR pixel counts at the benchmark shapes equal
the closed-form circle counts.
"""
import warnings

import numpy as np


def get_2d_ror_mask(recon_shape):
    rows, cols = recon_shape[:2]
    rc, cc = (rows - 1) / 2, (cols - 1) / 2
    radius = max(rc, cc)
    col_coords = np.arange(cols) - cc
    row_coords = np.arange(rows) - rc
    coords = np.meshgrid(col_coords, row_coords)
    return coords[0] ** 2 + coords[1] ** 2 <= radius ** 2


def gen_pixel_partition(recon_shape, num_subsets, use_ror_mask=True):
    rows, cols = recon_shape[:2]
    indices = np.arange(rows * cols, dtype=np.int32)
    if use_ror_mask:
        indices = indices[get_2d_ror_mask(recon_shape).flatten() == 1]
    if num_subsets > len(indices):
        num_subsets = len(indices)
        warnings.warn("number of subsets reduced to the number of ROR pixels")
    per = int(np.ceil(len(indices) / num_subsets))
    extra = num_subsets * per - len(indices)
    indices = np.random.permutation(indices)
    non_final = (num_subsets - 1) * per
    extra_indices = np.random.choice(indices[:non_final], size=extra, replace=False)
    indices = np.concatenate((indices, extra_indices))
    return np.sort(indices.reshape(num_subsets, indices.size // num_subsets), axis=1)


def gen_set_of_pixel_partitions(recon_shape, granularity, use_ror_mask=True):
    return [gen_pixel_partition(recon_shape, k, use_ror_mask=use_ror_mask) for k in granularity]


def gen_partition_sequence(partition_sequence, max_iterations):
    seq = np.array(partition_sequence)
    if max_iterations > seq.size:
        return np.concatenate((seq, np.full(max_iterations - seq.size, seq[-1])))
    return seq[:max_iterations]


def gen_full_indices(recon_shape, use_ror_mask=True):
    return gen_pixel_partition(recon_shape, 1, use_ror_mask=use_ror_mask)[0]
