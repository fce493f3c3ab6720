"""pymdp_amd: MI355X-native (gfx950) FPI + EFE hot path of pymdp.

Drop-in for pymdp's `Agent.infer_states` / `Agent.infer_policies` (and the
functions behind them), computed by hand-written HIP kernels in
`pymdp_amd/_lib/libpymdp_hip.so` (C ABI: include/pymdp_hip.h).  Tensors are
torch tensors on a HIP device; torch provides device memory and streams.
"""
from . import ffi
from .maths import MINVAL, log_stable

__version__ = "0.1.0"

__all__ = ["ffi", "MINVAL", "log_stable", "__version__"]
