"""State inference (mirror of the one-step branch of pymdp/inference.py).

`update_posterior_states` with method in {"fpi", "exact"} runs the fused HIP
FPI; the sequence methods (mmp, vmp) and smoothing (ovf) are not on this
hot path and raise NotImplementedError.
"""
from __future__ import annotations

from typing import Sequence

import torch

from .ffi import infer_states_ffi

EXACT_METHOD = "exact"
ONE_STEP_METHODS = {"fpi", EXACT_METHOD}
SEQUENCE_METHODS = {"mmp", "vmp"}


def update_posterior_states(A: Sequence[torch.Tensor], B, obs: Sequence[torch.Tensor], past_actions=None,
                            prior: Sequence[torch.Tensor] | None = None, qs_hist=None,
                            A_dependencies: Sequence[Sequence[int]] | None = None, B_dependencies=None,
                            num_iter: int = 16, method: str = "fpi", distr_obs: bool = True,
                            inference_horizon=None, valid_steps=None, return_info: bool = False):
    """pymdp/inference.py:338-485, one-step methods, batched over agents.

    obs[m]: [batch, O_m] (distr_obs) or [batch] indices; prior[f]: [batch, S_f].
    Returns qs[f] of shape [batch, 1, S_f] (time axis appended,
    pymdp/inference.py:274-295), or appended to qs_hist when given.
    """
    if method not in ONE_STEP_METHODS:
        raise NotImplementedError(f"method {method!r} is not on the HIP hot path (fpi, exact)")
    if return_info:
        raise NotImplementedError("return_info VFE diagnostics are not on the HIP hot path")
    if prior is None:
        raise ValueError("prior must be provided")
    F = len(prior)
    if A_dependencies is None:
        A_dependencies = [list(range(F)) for _ in A]
    # pymdp/inference.py:43-53: the current observation is the last one.
    cur = []
    for o in obs:
        if distr_obs:
            cur.append(o if o.dim() <= 2 else o[:, -1])
        else:
            cur.append(o if o.dim() <= 1 else o[:, -1])
    qs = infer_states_ffi(A, cur, prior, A_dependencies, num_iter, distr_obs=distr_obs,
                          method_exact=(method == EXACT_METHOD))
    qs = [q[:, None, :] for q in qs]
    if qs_hist is not None:
        qs = [torch.cat([h, q], 1) for h, q in zip(qs_hist, qs)]
        if inference_horizon is not None:
            qs = [q[:, -inference_horizon:] for q in qs]
    return qs
