"""Small host-side helpers with the reference's numeric conventions
(pymdp/maths.py:15-79).  Used only for argument preparation, never for the
hot-path arithmetic, which runs in the HIP kernels."""
from __future__ import annotations

import torch

# pymdp/maths.py:15  MINVAL = jnp.finfo(float).eps
MINVAL = 2.220446049250313e-16


def log_stable(x: torch.Tensor) -> torch.Tensor:
    """pymdp/maths.py:64-79."""
    return torch.log(torch.clamp(x, min=MINVAL))
