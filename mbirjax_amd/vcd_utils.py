"""Pixel partitions for VCD (host side, NumPy).

The draws from NumPy's legacy global RNG follow mbirjax/vcd_utils.py:105-318
call for call (one permutation of the ROR pixels, then a choice without
replacement of the padding from the non-final subsets, per granularity), so a
seeded run visits exactly the reference's subsets in the reference's order.
"""
import warnings

import numpy as np


def get_2d_ror_mask(recon_shape, *, crop_radius_pixels=0, crop_radius_fraction=0.0):
    """Largest circle inscribed on the longer edge of recon_shape[:2]
    (mbirjax/vcd_utils.py:66-102)."""
    if crop_radius_pixels != 0 and crop_radius_fraction != 0.0:
        raise ValueError('Only one of crop_radius_pixels and crop_radius_fraction can be nonzero.')
    n_rows, n_cols = recon_shape[:2]
    r_c, c_c = (n_rows - 1) / 2, (n_cols - 1) / 2
    radius = max(r_c, c_c)
    crop = max(int(radius * crop_radius_fraction), crop_radius_pixels, 0)
    radius -= crop
    dy = (np.arange(n_rows) - r_c)[:, None]
    dx = (np.arange(n_cols) - c_c)[None, :]
    return dx * dx + dy * dy <= radius * radius


def gen_pixel_partition(recon_shape, num_subsets, use_ror_mask=True):
    """(num_subsets, ceil(P/num_subsets)) int32 array of flat pixel indices,
    each row sorted (mbirjax/vcd_utils.py:176-223)."""
    n_rows, n_cols = recon_shape[:2]
    pixels = np.arange(n_rows * n_cols, dtype=np.int32)
    if use_ror_mask:
        pixels = pixels[get_2d_ror_mask(recon_shape).ravel()]
    if num_subsets > pixels.size:
        num_subsets = pixels.size
        warnings.warn('\nThe number of partition subsets is greater than the number of pixels in the region of '
                      'reconstruction.  \nReducing the number of subsets to equal the number of indices.')
    per_subset = int(np.ceil(pixels.size / num_subsets))
    n_pad = num_subsets * per_subset - pixels.size
    shuffled = np.random.permutation(pixels)
    padding = np.random.choice(shuffled[:(num_subsets - 1) * per_subset], size=n_pad, replace=False)
    table = np.concatenate((shuffled, padding)).reshape(num_subsets, per_subset)
    return np.sort(table, axis=1).astype(np.int32)


def gen_set_of_pixel_partitions(recon_shape, granularity, output_device=None, use_ror_mask=True):
    """One partition per entry of granularity (mbirjax/vcd_utils.py:105-124).
    Returned as host arrays; the model uploads them."""
    return [gen_pixel_partition(recon_shape, int(k), use_ror_mask=use_ror_mask) for k in granularity]


def gen_partition_sequence(partition_sequence, max_iterations):
    """Extend partition_sequence with its last entry to max_iterations
    (mbirjax/vcd_utils.py:284-307)."""
    seq = np.asarray(partition_sequence)
    if max_iterations > seq.size:
        return np.concatenate((seq, np.full(max_iterations - seq.size, seq[-1])))
    return seq[:max_iterations]


def gen_full_indices(recon_shape, use_ror_mask=True):
    """All ROR pixels, sorted (mbirjax/vcd_utils.py:310-318).  Like the
    reference this draws from the global RNG."""
    return gen_pixel_partition(recon_shape, 1, use_ror_mask=use_ror_mask)[0]
