"""Policy construction, EFE and action selection (mirror of pymdp/control.py).

The EFE over all policies and the policy posterior run in the gfx950
action-history pipeline (`ffi.update_posterior_policies_ffi`); policy
construction is host-side setup; action selection is a device argmax /
categorical draw over the posterior.  All functions are batched over agents
(leading axis), the way `Agent.infer_policies` vmaps the reference.
"""
from __future__ import annotations

import itertools
from typing import Sequence

import numpy as np
import torch

from .ffi._efe import update_posterior_policies_ffi
from .maths import log_stable


class Policies:
    """pymdp/control.py:22-40: a policy array [P, T, F] with its sizes."""

    def __init__(self, policy_arr):
        arr = policy_arr.detach().cpu().numpy() if isinstance(policy_arr, torch.Tensor) else np.asarray(policy_arr)
        self.policy_arr = np.ascontiguousarray(arr.astype(np.int32))
        self.num_policies = self.policy_arr.shape[0]
        self.horizon = self.policy_arr.shape[1]

    def __getitem__(self, idx):
        return self.policy_arr[idx]

    def __len__(self):
        return self.num_policies


def construct_policies(num_states: Sequence[int], num_controls: Sequence[int] | None = None, policy_len: int = 1,
                       control_fac_idx: Sequence[int] | None = None) -> np.ndarray:
    """pymdp/control.py:152-195: all control sequences of length policy_len,
    [P, policy_len, F] int32, in itertools.product order."""
    num_factors = len(num_states)
    if control_fac_idx is None:
        if num_controls is not None:
            control_fac_idx = [f for f, n_c in enumerate(num_controls) if n_c > 1]
        else:
            control_fac_idx = list(range(num_factors))
    if num_controls is None:
        num_controls = [num_states[c] if c in control_fac_idx else 1 for c in range(num_factors)]
    x = list(num_controls) * policy_len
    pols = np.array(list(itertools.product(*[range(i) for i in x])), dtype=np.int32)
    return pols.reshape(len(pols), policy_len, num_factors)


def update_posterior_policies(policy_matrix, qs_init, A, B, C, E, pA, pB, A_dependencies, B_dependencies,
                              gamma=16.0, use_utility=True, use_states_info_gain=True, use_param_info_gain=False,
                              ffi_cache_params_static=False):
    """pymdp/control.py:198-349 (batched): returns (q_pi, neg_efe), [batch, P]."""
    if use_param_info_gain and pA is None and pB is None:
        raise ValueError("use_param_info_gain=True requires at least one of pA or pB.")
    return update_posterior_policies_ffi(
        policy_matrix, qs_init, A, B, C, E, A_dependencies, B_dependencies, gamma=gamma,
        use_utility=use_utility, use_states_info_gain=use_states_info_gain,
        use_param_info_gain=use_param_info_gain, use_inductive=False,
        ffi_cache_params_static=ffi_cache_params_static, pA=pA, pB=pB)


def update_posterior_policies_inductive(policy_matrix, qs_init, A, B, C, E, pA, pB, A_dependencies, B_dependencies,
                                        I, gamma=16.0, inductive_epsilon=1e-3, use_utility=True,
                                        use_states_info_gain=True, use_param_info_gain=False, use_inductive=True,
                                        ffi_cache_params_static=False):
    """pymdp/control.py:795-963 (batched): returns (q_pi, neg_efe), [batch, P]."""
    if use_param_info_gain and pA is None and pB is None:
        raise ValueError("use_param_info_gain=True requires at least one of pA or pB.")
    return update_posterior_policies_ffi(
        policy_matrix, qs_init, A, B, C, E, A_dependencies, B_dependencies, gamma=gamma, I=I,
        inductive_epsilon=inductive_epsilon, use_utility=use_utility, use_states_info_gain=use_states_info_gain,
        use_param_info_gain=use_param_info_gain, use_inductive=use_inductive,
        ffi_cache_params_static=ffi_cache_params_static, pA=pA, pB=pB)


def compute_expected_state(qs_prior: Sequence[torch.Tensor], B: Sequence[torch.Tensor], u_t: torch.Tensor,
                           B_dependencies=None) -> list[torch.Tensor]:
    """pymdp/control.py:351-389, batched: qs_prior[f] [batch, S_f],
    B[f] [batch, S_f, S_deps..., U_f], u_t [batch, F]."""
    if B_dependencies is None:
        B_dependencies = [[f] for f in range(len(B))]
    out = []
    batch = qs_prior[0].shape[0]
    ar = torch.arange(batch, device=qs_prior[0].device)
    for f, (B_f, deps) in enumerate(zip(B, B_dependencies)):
        Bb = B_f if B_f.dim() == len(deps) + 3 else B_f.expand((batch,) + tuple(B_f.shape))
        Bu = Bb[ar, ..., u_t[:, f].long()]  # [batch, S_f, S_deps...]
        x = Bu
        for d in reversed(deps):
            x = (x * qs_prior[d].reshape((batch,) + (1,) * (x.dim() - 2) + (-1,))).sum(-1)
        out.append(x)
    return out


def get_marginals(q_pi: torch.Tensor, policies, num_controls: Sequence[int]) -> list[torch.Tensor]:
    """pymdp/control.py:42-67 (batched q_pi [batch, P])."""
    pol = torch.as_tensor(np.asarray(policies)[:, 0, :], device=q_pi.device, dtype=torch.long)
    out = []
    for f, n in enumerate(num_controls):
        onehot = torch.nn.functional.one_hot(pol[:, f], n).to(q_pi.dtype)  # [P, n]
        out.append(q_pi @ onehot)
    return out


def sample_policy(policies, q_pi: torch.Tensor, action_selection: str = "deterministic", alpha=16.0,
                  generator: torch.Generator | None = None) -> torch.Tensor:
    """pymdp/control.py:115-150 (batched): first-step multi-action of the
    selected policy.  Deterministic = first-index argmax; stochastic =
    categorical over alpha * log_stable(q_pi) with a torch generator (the
    reference's jax.random stream is not reproducible without JAX)."""
    pol = torch.as_tensor(np.asarray(policies)[:, 0, :], device=q_pi.device, dtype=torch.long)
    if action_selection == "deterministic":
        idx = torch.argmax(q_pi, dim=-1)
    elif action_selection == "stochastic":
        a = torch.as_tensor(alpha, device=q_pi.device, dtype=q_pi.dtype).reshape(-1, 1)
        probs = torch.softmax(log_stable(q_pi) * a, dim=-1)
        idx = torch.multinomial(probs, 1, generator=generator)[:, 0]
    else:
        raise NotImplementedError(action_selection)
    return pol[idx]


def sample_action(policies, num_controls, q_pi: torch.Tensor, action_selection: str = "deterministic",
                  alpha=16.0, generator: torch.Generator | None = None) -> torch.Tensor:
    """pymdp/control.py:69-113 (batched, marginal mode)."""
    marg = get_marginals(q_pi, policies, num_controls)
    if action_selection == "deterministic":
        return torch.stack([torch.argmax(m, -1) for m in marg], -1)
    if action_selection == "stochastic":
        a = torch.as_tensor(alpha, device=q_pi.device, dtype=q_pi.dtype).reshape(-1, 1)
        return torch.stack([torch.multinomial(torch.softmax(log_stable(m) * a, -1), 1, generator=generator)[:, 0]
                            for m in marg], -1)
    raise NotImplementedError(action_selection)
