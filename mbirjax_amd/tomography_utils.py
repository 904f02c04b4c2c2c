"""Direct-reconstruction filter (mbirjax/tomography_utils.py:4-35)."""
import torch


def generate_direct_recon_filter(num_channels, filter_name="ramp", device=None):
    """Space-domain ramp filter of 2*num_channels-1 taps:
    h(n) = sinc(n)/2 - sinc(n/2)^2/4 (a ramp in frequency)."""
    supported = ["ramp"]
    if filter_name not in supported:
        raise ValueError(f"Unsupported filter. Supported filters are: {', '.join(supported)}.")
    n = torch.arange(-num_channels + 1, num_channels, dtype=torch.float64, device=device)
    h = 0.5 * torch.sinc(n) - 0.25 * torch.sinc(n / 2) ** 2
    return h.to(torch.float32)
