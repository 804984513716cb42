"""Source waveforms for point pressure sources."""
from __future__ import annotations

import numpy as np


def differentiated_gaussian(n_steps: int, sigma: float = 8.0, n0: float = 48.0,
                            amplitude: float = 1.0) -> np.ndarray:
    """s(n) = -A (t - n0)/sigma * exp(-(t - n0)^2 / (2 sigma^2)), t = n + 1.

    Sample n is added after the pressure update of step n, i.e. at time level
    n + 1.  Evaluated in float64, returned as float32.
    """
    t = np.arange(n_steps, dtype=np.float64) + 1.0
    u = (t - n0) / sigma
    return (-amplitude * u * np.exp(-0.5 * u * u)).astype(np.float32)
