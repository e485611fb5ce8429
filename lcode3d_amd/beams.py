"""Device-side synthetic beams: the cosine-Gaussian bunch of the reference's
example config (config_example.py:21-28), evaluated with torch on the GPU
instead of with NumPy on the host (lcode.py:1040) so that a run needs no
host->device beam copy per slice."""

import math

import torch


def cosine_gaussian(config, xi_i, amplitude=.05, sigma=1.0, shift_x=0.0,
                    shift_y=0.0, device=None):
    """rho_b(xi, x, y) = A exp(-r^2 / (2 sigma^2)) (1 - cos(xi sqrt(pi/2)))
    for xi = -xi_i * xi_step_size >= -2 sqrt(2 pi), else the scalar 0."""
    xi = -xi_i * config.xi_step_size
    if xi < -2 * math.sqrt(2 * math.pi):
        return 0
    n = config.grid_steps
    dev = device or torch.device("cuda", torch.cuda.current_device())
    g = (torch.arange(n, device=dev, dtype=torch.float64) - n // 2) \
        * config.grid_step_size
    x, y = g[:, None] - shift_x, g[None, :] - shift_y
    r = torch.sqrt(x * x + y * y)
    return (amplitude * torch.exp(-.5 * (r / sigma) ** 2)
            * (1 - math.cos(xi * math.sqrt(math.pi / 2))))
