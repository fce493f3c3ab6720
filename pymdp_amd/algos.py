"""Fixed-point iteration entry points (mirror of pymdp/algos.py:53-155).

Same names and argument meaning as the reference; the computation is the
fused gfx950 kernel (log-likelihood build + num_iter Jacobi sweeps + final
softmax in one launch).  Inputs are torch tensors on a HIP device, either for
one agent (no leading axis) or for a batch (leading agent axis on obs and
prior; A with or without it).
"""
from __future__ import annotations

from typing import Sequence

import torch

from .ffi import infer_states_ffi


def _batched(A, obs, prior, A_dependencies, distr_obs):
    unbatched = prior[0].dim() == 1
    if unbatched:
        prior = [p[None] for p in prior]
        obs = [o[None] if distr_obs else o.reshape(1) for o in obs]
    return unbatched, obs, prior


def run_factorized_fpi(A: Sequence[torch.Tensor], obs: Sequence[torch.Tensor], prior: Sequence[torch.Tensor],
                       A_dependencies: Sequence[Sequence[int]], num_iter: int = 1,
                       distr_obs: bool = True) -> list[torch.Tensor]:
    """pymdp/algos.py:88-155.  F == 1 is the one-pass closed form."""
    unbatched, obs_b, prior_b = _batched(A, obs, prior, A_dependencies, distr_obs)
    qs = infer_states_ffi(A, obs_b, prior_b, A_dependencies, num_iter, distr_obs=distr_obs)
    return [q[0] for q in qs] if unbatched else qs


def run_vanilla_fpi(A: Sequence[torch.Tensor], obs: Sequence[torch.Tensor], prior: Sequence[torch.Tensor],
                    num_iter: int = 1, distr_obs: bool = True) -> list[torch.Tensor]:
    """pymdp/algos.py:53-86: every modality depends on every factor; the
    summed log-likelihood over the full joint is iterated."""
    F = len(prior)
    deps = [list(range(F)) for _ in A]
    unbatched, obs_b, prior_b = _batched(A, obs, prior, deps, distr_obs)
    if F == 1:
        # the vanilla scan has no closed-form shortcut; with one factor every
        # sweep gives the same log_q, so num_iter >= 1 equals one sweep and
        # num_iter == 0 leaves the uniform start.
        if num_iter == 0:
            qs = [torch.full_like(p, 1.0 / p.shape[-1]) for p in prior_b]
            return [q[0] for q in qs] if unbatched else qs
    qs = infer_states_ffi(A, obs_b, prior_b, deps, num_iter, distr_obs=distr_obs)
    return [q[0] for q in qs] if unbatched else qs
