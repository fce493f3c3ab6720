"""State inference (mirror of pymdp/inference.py:338-485).

`update_posterior_states` with method in {"fpi", "exact"} runs the fused HIP
FPI (and, with return_info, the VFE kernel); "mmp" / "vmp" run the sequence
kernel over the (horizon-truncated) window.  Online filtering / smoothing
("ovf") and the sequence VFE are not built and raise NotImplementedError.
"""
from __future__ import annotations

from typing import Sequence

import torch

from .ffi import fpi_vfe_ffi, infer_states_ffi
from .ffi._seq import sequence_infer_ffi

EXACT_METHOD = "exact"
ONE_STEP_METHODS = {"fpi", EXACT_METHOD}
SEQUENCE_METHODS = {"mmp", "vmp"}


def update_posterior_states(A: Sequence[torch.Tensor], B, obs: Sequence[torch.Tensor], past_actions=None,
                            prior: Sequence[torch.Tensor] | None = None, qs_hist=None,
                            A_dependencies: Sequence[Sequence[int]] | None = None, B_dependencies=None,
                            num_iter: int = 16, method: str = "fpi", distr_obs: bool = True,
                            inference_horizon=None, valid_steps=None, return_info: bool = False):
    """pymdp/inference.py:338-485, batched over agents.

    One-step methods: obs[m] [batch, O_m] (distr_obs) or [batch] indices (the
    last step of a history is used); returns qs[f] [batch, 1, S_f] (time axis
    appended, pymdp/inference.py:274-295) or appended to qs_hist, and with
    return_info also {"vfe_t", "vfe", "vfe_components"} ([batch] each).
    Sequence methods: obs[m] [batch, T] indices or [batch, T, O_m];
    past_actions [batch, T-1, F]; returns qs[f] [batch, T, S_f].
    """
    if prior is None:
        raise ValueError("prior must be provided")
    F = len(prior)
    if A_dependencies is None:
        A_dependencies = [list(range(F)) for _ in A]
    if method in SEQUENCE_METHODS:
        if return_info:
            raise NotImplementedError("return_info for sequence methods (sequence VFE) is not built")
        if B is None:
            raise ValueError(f"Sequence inference method `{method}` requires `B`.")
        if inference_horizon is not None:
            # pymdp/inference.py:56-73
            if inference_horizon < 1:
                raise ValueError("`inference_horizon` must be >= 1 when provided")
            obs = [o[:, -inference_horizon:] for o in obs]
            if past_actions is not None:
                h = max(inference_horizon - 1, 0)
                past_actions = past_actions[:, :0] if h == 0 else past_actions[:, -h:]
        return sequence_infer_ffi(method, A, B, obs, prior, A_dependencies, B_dependencies,
                                  past_actions=past_actions, valid_steps=valid_steps, num_iter=num_iter,
                                  distr_obs=distr_obs)
    if method not in ONE_STEP_METHODS:
        raise NotImplementedError(f"method {method!r} is not built (fpi, exact, mmp, vmp)")
    # pymdp/inference.py:43-53: the current observation is the last one.
    cur = []
    for o in obs:
        if distr_obs:
            cur.append(o if o.dim() <= 2 else o[:, -1])
        else:
            cur.append(o if o.dim() <= 1 else o[:, -1])
    qs = infer_states_ffi(A, cur, prior, A_dependencies, num_iter, distr_obs=distr_obs,
                          method_exact=(method == EXACT_METHOD))
    info = None
    if return_info:
        v = fpi_vfe_ffi(A, cur, prior, torch.cat(qs, -1), A_dependencies, distr_obs=distr_obs)
        zero = torch.zeros_like(v[:, 3])
        info = {"vfe_t": v[:, 3], "vfe": v[:, 3], "vfe_components": {
            "neg_entropy_t": v[:, 0], "prior_cross_entropy_t": v[:, 1], "transition_cross_entropy_t": zero,
            "accuracy_t": v[:, 2], "state_vfe": v[:, 3], "parameter_kl_A": zero, "parameter_kl_B": zero,
            "parameter_kl": zero}}
    qs = [q[:, None, :] for q in qs]
    if qs_hist is not None:
        qs = [torch.cat([h, q], 1) for h, q in zip(qs_hist, qs)]
        if inference_horizon is not None:
            qs = [q[:, -inference_horizon:] for q in qs]
    return (qs, info) if return_info else qs
