"""Agent (mirror of pymdp/agent.py, the hot-path methods).

Keeps the reference's constructor arguments and the `infer_states`,
`infer_policies`, `sample_action`, `update_empirical_prior`,
`multiaction_probabilities` and `infer_parameters` methods with their
argument meaning and return shapes.  Model tensors are torch tensors on a HIP device; a tensor given
without the leading batch axis is shared by every agent (no broadcast copy
is made: the kernels read it with batch stride 0).

Sequence inference ("mmp", "vmp") runs the window kernel
(`pymdp_hip_sequence_infer`).  Not provided: online filtering / smoothing
("ovf") and the smoothing-based learning it enables, and multi-action
flattening (`B_action_dependencies`).  `infer_states(...,
return_info=True)` returns the VFE diagnostics of the posterior
(`pymdp_hip_fpi_vfe`).
`infer_parameters` (Dirichlet learning of A and B, pymdp/agent.py:416-620)
runs the fused learning kernels of `pymdp_amd.learning`.  Parameter information gain
(`use_param_info_gain` with `pA` / `pB`) is computed by the neg-EFE kernels.
"""
from __future__ import annotations

import copy
import math
from typing import Sequence

import numpy as np
import torch

from . import control, learning
from .control import Policies, construct_policies
from .ffi._efe import update_posterior_policies_ffi
from .ffi._fpi import fpi_vfe_ffi, infer_states_ffi
from .ffi._seq import sequence_infer_ffi


def _resolve_a(num_factors, num_modalities, A_dependencies):
    if A_dependencies is None:
        return [list(range(num_factors)) for _ in range(num_modalities)]
    return [list(d) for d in A_dependencies]


def _resolve_b(num_factors, B_dependencies):
    if B_dependencies is None:
        return [[f] for f in range(num_factors)]
    return [list(d) for d in B_dependencies]


class Agent:
    """pymdp/agent.py:19-620 (hot-path subset)."""

    def __init__(self, A, B, C=None, D=None, E=None, pA=None, pB=None, H=None, I=None, A_dependencies=None,
                 B_dependencies=None, B_action_dependencies=None, num_controls=None, control_fac_idx=None,
                 policy_len: int = 1, policies=None, gamma=1.0, alpha=1.0, inductive_depth: int = 1,
                 inductive_threshold=0.1, inductive_epsilon=1e-3, use_utility=True, use_states_info_gain=True,
                 use_param_info_gain=False, use_inductive=False, categorical_obs=False, preprocess_fn=None,
                 action_selection="deterministic", sampling_mode="full", inference_algo="fpi",
                 inference_horizon=None, num_iter: int = 16, batch_size: int = 1, learning_mode="online",
                 learn_A=False, learn_B=False, learn_C=False, learn_D=False, learn_E=False, device=None):
        if B_action_dependencies is not None:
            raise NotImplementedError("B_action_dependencies (multi-action flattening) is not supported")
        if use_param_info_gain and pA is None and pB is None:
            raise ValueError("use_param_info_gain=True requires at least one of pA or pB.")
        if learn_A and pA is None:
            raise ValueError("pA is required for A learning")
        if learn_B and pB is None:
            raise ValueError("pB is required for B learning")
        dev = torch.device(device) if device is not None else A[0].device
        self.device = dev
        self.num_modalities = len(A)
        self.num_factors = len(B)
        self.A_dependencies = _resolve_a(self.num_factors, self.num_modalities, A_dependencies)
        self.B_dependencies = _resolve_b(self.num_factors, B_dependencies)
        self.batch_size = int(batch_size)
        # num_states from B and B_dependencies (pymdp/agent.py:406-420)
        self.num_states = []
        for f, deps in enumerate(self.B_dependencies):
            if f not in deps:
                raise ValueError(f"num_states cannot be inferred from B, B_dependencies if the dynamics of hidden "
                                 f"state {f} is not conditioned on itself")
            i = deps.index(f)
            self.num_states.append(int(B[f].shape[-(len(deps) + 1 - i)]))
        self.num_controls = [int(b.shape[-1]) for b in B]

        def _dev(x):
            return torch.as_tensor(x, dtype=torch.float32, device=dev).contiguous()

        A = [_dev(a) for a in A]
        B = [_dev(b) for b in B]
        for m, a in enumerate(A):
            core = len(self.A_dependencies[m]) + 1
            if a.dim() == core + 1 and a.shape[0] == 1 and self.batch_size > 1:
                A[m] = a[0].contiguous()  # shared
            elif a.dim() == core + 1 and a.shape[0] != self.batch_size:
                raise ValueError(f"Batch size {self.batch_size} does not match the first dimension of A[{m}] "
                                 f"with shape {tuple(a.shape)}")
        for f, b in enumerate(B):
            core = len(self.B_dependencies[f]) + 2
            if b.dim() == core + 1 and b.shape[0] == 1 and self.batch_size > 1:
                B[f] = b[0].contiguous()
            elif b.dim() == core + 1 and b.shape[0] != self.batch_size:
                raise ValueError(f"Batch size {self.batch_size} does not match the first dimension of B[{f}] "
                                 f"with shape {tuple(b.shape)}")
        self.num_obs = [int(a.shape[-len(self.A_dependencies[m]) - 1]) for m, a in enumerate(A)]
        self.num_iter = num_iter
        self.inference_algo = inference_algo.lower()
        if self.inference_algo not in ("fpi", "exact", "mmp", "vmp"):
            raise NotImplementedError(f"inference_algo {inference_algo!r} is not built (fpi, exact, mmp, vmp)")
        self.inference_horizon = inference_horizon
        self.policy_len = policy_len
        self.action_selection = action_selection
        self.sampling_mode = sampling_mode
        self.use_utility = use_utility
        self.use_states_info_gain = use_states_info_gain
        self.use_param_info_gain = use_param_info_gain
        self.use_inductive = use_inductive
        self.inductive_depth = inductive_depth
        self.categorical_obs = categorical_obs
        self.preprocess_fn = preprocess_fn
        if control_fac_idx is None:
            self.control_fac_idx = [f for f in range(self.num_factors) if self.num_controls[f] > 1]
        else:
            self.control_fac_idx = list(control_fac_idx)
        if policies is None:
            pol = construct_policies(self.num_states, self.num_controls, self.policy_len, self.control_fac_idx)
            self.policies = Policies(pol)
        else:
            self.policies = policies if isinstance(policies, Policies) else Policies(policies)
        bs = self.batch_size
        if C is None:
            C = [torch.full((bs, o), 1.0 / o, device=dev) for o in self.num_obs]
        C = [_dev(c) for c in C]
        if D is None:
            D = [torch.full((bs, s), 1.0 / s, device=dev) for s in self.num_states]
        D = [_dev(d) for d in D]
        D = [d.expand(bs, -1).contiguous() if d.dim() == 1 or d.shape[0] == 1 else d for d in D]
        P = self.policies.num_policies
        if E is None:
            E = torch.full((bs, P), 1.0 / P, device=dev)
        E = _dev(E)
        if E.dim() == 1 or E.shape[0] == 1:
            E = E.reshape(1, -1).expand(bs, -1).contiguous()
        self.A, self.B, self.C, self.D, self.E = A, B, C, D, E
        self.pA = [_dev(x) for x in pA] if pA is not None else None
        self.pB = [_dev(x) for x in pB] if pB is not None else None
        self.H = H
        self.learning_mode = learning_mode
        self.learn_A, self.learn_B = bool(learn_A), bool(learn_B)
        self.learn_C, self.learn_D, self.learn_E = bool(learn_C), bool(learn_D), bool(learn_E)
        self.gamma = torch.as_tensor(gamma, dtype=torch.float32, device=dev).expand(bs).contiguous()
        self.alpha = torch.as_tensor(alpha, dtype=torch.float32, device=dev).expand(bs).contiguous()
        self.inductive_threshold = torch.as_tensor(inductive_threshold, dtype=torch.float32,
                                                   device=dev).expand(bs).contiguous()
        self.inductive_epsilon = torch.as_tensor(inductive_epsilon, dtype=torch.float32,
                                                 device=dev).expand(bs).contiguous()
        if use_inductive and self.inductive_depth < 1:
            raise ValueError("`inductive_depth` must be >= 1 when `use_inductive` is True")
        if use_inductive and H is not None:
            # per agent, as the reference's vmap of generate_I_matrix over
            # (H, B, inductive_threshold) (pymdp/agent.py:366-376)
            I = []
            for f in range(self.num_factors):
                h_f = torch.as_tensor(H[f], dtype=torch.float32, device=dev)
                b_f = B[f].to(dev)
                I.append(torch.stack([
                    _generate_I(h_f if h_f.dim() == 1 else h_f[b],
                                b_f if b_f.dim() == len(self.B_dependencies[f]) + 2 else b_f[b],
                                float(self.inductive_threshold[b]), self.inductive_depth)
                    for b in range(bs)]))
        elif use_inductive and I is not None:
            I = [_dev(i) for i in I]
        else:
            I = [torch.zeros((bs, 1, s), device=dev) for s in self.num_states]
        self.I = I

    # ------------------------------------------------------------------
    def make_categorical(self, observations):
        """pymdp/agent.py:659-673."""
        return [torch.nn.functional.one_hot(torch.as_tensor(o, device=self.device).long(), self.num_obs[m]).float()
                for m, o in enumerate(observations)]

    def infer_states(self, observations, empirical_prior, *, past_actions=None, qs_hist=None, valid_steps=None,
                     mask=None, preprocess_fn=None, return_info=False):
        """pymdp/agent.py:675-823: posterior qs[f] of shape [batch, 1, S_f]
        (appended to qs_hist when given); with return_info also the VFE
        diagnostics of the new posterior (pymdp/inference.py:456-483):
        {"vfe_t", "vfe", "vfe_components"}, each entry [batch]."""
        if not isinstance(empirical_prior, (list, tuple)):
            raise ValueError("`empirical_prior` must be a list/tuple with one entry per hidden-state factor.")
        if len(empirical_prior) != self.num_factors:
            raise ValueError(f"`empirical_prior` has {len(empirical_prior)} factor(s), expected {self.num_factors}")
        prior = [p if p.dim() == 2 else p.reshape(self.batch_size, -1) for p in empirical_prior]
        if self.inference_algo in ("mmp", "vmp"):
            return self._infer_states_sequence(observations, prior, past_actions, valid_steps, return_info)
        fn = preprocess_fn or self.preprocess_fn
        if fn is not None:
            obs, distr = fn(observations), True
        elif self.categorical_obs:
            obs, distr = list(observations), True
        else:
            # index observations: the kernel applies one_hot itself
            obs = [torch.as_tensor(o, device=self.device).reshape(self.batch_size) for o in observations]
            distr = False
        A = self.A
        if mask is not None:
            if not distr:
                obs = self.make_categorical(obs)
                distr = True
            A = list(A)
            obs = list(obs)
            for i, m in enumerate(mask):
                m = torch.as_tensor(m, dtype=torch.float32, device=self.device)
                obs[i] = m * obs[i] + (1 - m) * torch.ones_like(obs[i]) / self.num_obs[i]
                A[i] = (m * A[i] + (1 - m) * torch.ones_like(A[i]) / self.num_obs[i]).contiguous()
        if distr:
            obs = [o if o.dim() == 2 else o[:, -1] for o in obs]
        qs = infer_states_ffi(A, obs, prior, self.A_dependencies, self.num_iter, distr_obs=distr,
                              method_exact=(self.inference_algo == "exact"), batch=self.batch_size)
        info = None
        if return_info:
            v = fpi_vfe_ffi(A, obs, prior, torch.cat(qs, -1), self.A_dependencies, distr_obs=distr,
                            batch=self.batch_size)
            zero = torch.zeros_like(v[:, 3])
            info = {"vfe_t": v[:, 3], "vfe": v[:, 3], "vfe_components": {
                "neg_entropy_t": v[:, 0], "prior_cross_entropy_t": v[:, 1], "transition_cross_entropy_t": zero,
                "accuracy_t": v[:, 2], "state_vfe": v[:, 3], "parameter_kl_A": zero, "parameter_kl_B": zero,
                "parameter_kl": zero}}
        qs = [q[:, None, :] for q in qs]
        if qs_hist is not None:
            qs = [torch.cat([h, q], 1) for h, q in zip(qs_hist, qs)]
            if self.inference_horizon is not None:
                qs = [q[:, -self.inference_horizon:] for q in qs]
        if return_info:
            return qs, info
        return qs

    def _infer_states_sequence(self, observations, prior, past_actions, valid_steps, return_info):
        """Sequence branch of pymdp/agent.py:675-823: observations[m] is the
        window [batch, T] (indices) or [batch, T, O_m] (categorical_obs);
        returns qs[f] [batch, T, S_f]."""
        if return_info:
            raise NotImplementedError("return_info for sequence methods (sequence VFE) is not built")
        if self.categorical_obs:
            obs, distr = [torch.as_tensor(o, device=self.device).float() for o in observations], True
        else:
            obs = [torch.as_tensor(o, device=self.device).reshape(self.batch_size, -1) for o in observations]
            distr = False
        if past_actions is not None:
            past_actions = torch.as_tensor(past_actions, device=self.device)
        if self.inference_horizon is not None:
            h = self.inference_horizon
            obs = [o[:, -h:] for o in obs]
            if past_actions is not None:
                past_actions = past_actions[:, :0] if h <= 1 else past_actions[:, -(h - 1):]
        return sequence_infer_ffi(self.inference_algo, self.A, self.B, obs, prior, self.A_dependencies,
                                  self.B_dependencies, past_actions=past_actions, valid_steps=valid_steps,
                                  num_iter=self.num_iter, distr_obs=distr, batch=self.batch_size)

    def update_empirical_prior(self, action, qs):
        """pymdp/agent.py:825-851 (sequence methods keep D)."""
        if self.inference_algo in ("mmp", "vmp"):
            return self.D
        qs_last = [q[:, -1] for q in qs]
        return control.compute_expected_state(qs_last, self.B, torch.as_tensor(action, device=self.device),
                                              self.B_dependencies)

    def infer_policies(self, qs):
        """pymdp/agent.py:853-906: (q_pi, neg_efe), each [batch, P]."""
        latest = [q[:, -1] for q in qs]
        return update_posterior_policies_ffi(
            self.policies.policy_arr, latest, self.A, self.B, self.C, self.E, self.A_dependencies,
            self.B_dependencies, gamma=self.gamma, I=self.I, inductive_epsilon=self.inductive_epsilon,
            use_utility=self.use_utility, use_states_info_gain=self.use_states_info_gain,
            use_param_info_gain=self.use_param_info_gain, use_inductive=self.use_inductive,
            ffi_cache_params_static=not (self.learn_A or self.learn_B), pA=self.pA, pB=self.pB)

    def infer_parameters(self, beliefs_A, observations, actions, beliefs_B=None, lr_pA=1.0, lr_pB=1.0,
                         **kwargs):
        """pymdp/agent.py:416-620 for the filtering algorithms ("fpi",
        "exact"): returns a new Agent whose pA, A (learn_A) and pB, B
        (learn_B) are the Dirichlet posteriors and their expected values.
        beliefs [batch, T, S_f]; observations [batch, T] per modality
        (or [batch, T, O_m] with categorical_obs); actions [batch, T, F] or
        [batch, F]."""
        agent = copy.copy(self)
        seq = beliefs_A if beliefs_B is None else beliefs_B
        if actions is not None:
            actions = torch.as_tensor(actions, device=self.device)
            if actions.dim() == 2:
                actions = actions[:, None, :]
        T = int(seq[0].shape[1])
        act1 = actions[:, :max(T - 1, 0)] if actions is not None else None
        if self.learn_A:
            lr = lr_pA if not isinstance(lr_pA, torch.Tensor) else lr_pA.expand(self.batch_size)
            qA, E_qA = learning.update_obs_likelihood_dirichlet(
                self.pA, self.A, observations, beliefs_A, A_dependencies=self.A_dependencies,
                categorical_obs=self.categorical_obs, num_obs=self.num_obs, lr=lr)
            agent.pA, agent.A = qA, E_qA
        if self.learn_B:
            # some transition of the batch is valid (pymdp/agent.py:486-499)
            can_update = (act1 is not None and T > 1 and int(act1.shape[1]) == T - 1
                          and bool((act1 >= 0).all(-1).any()))
            if can_update:
                joints = [[seq[f][:, 1:]] + [seq[d][:, :-1] for d in self.B_dependencies[f]]
                          for f in range(self.num_factors)]
                lr = lr_pB if not isinstance(lr_pB, torch.Tensor) else lr_pB.expand(self.batch_size)
                qB, E_qB = learning.update_state_transition_dirichlet(
                    self.pB, self.B, joints, act1, num_controls=self.num_controls, lr=lr, mask_invalid=True)
                agent.pB, agent.B = qB, E_qB
                if self.use_inductive and self.H is not None:
                    # inductive matrices from the learned B (pymdp/agent.py:606-615)
                    I = []
                    for f in range(self.num_factors):
                        h = torch.as_tensor(self.H[f], dtype=torch.float32, device=self.device)
                        I.append(torch.stack([
                            _generate_I(h if h.dim() == 1 else h[b], E_qB[f][b], float(self.inductive_threshold[b]),
                                        self.inductive_depth) for b in range(self.batch_size)]))
                    agent.I = I
        return agent

    def multiaction_probabilities(self, q_pi):
        """pymdp/agent.py:908-937."""
        pol0 = torch.as_tensor(self.policies.policy_arr[:, 0], device=self.device)
        if self.sampling_mode == "marginal":
            marg = control.get_marginals(q_pi, self.policies.policy_arr, self.num_controls)
            out = marg[0]
            for m in marg[1:]:
                out = (out[:, :, None] * m[:, None, :]).reshape(out.shape[0], -1)
            return out
        uniq = torch.unique(pol0, dim=0)
        locs = (pol0[None, :, :] == uniq[:, None, :]).all(-1).to(q_pi.dtype)  # [U, P]
        return q_pi @ locs.T

    def sample_action(self, q_pi, rng_key=None):
        """pymdp/agent.py:939-967.  `rng_key` is a torch.Generator for
        stochastic selection."""
        if rng_key is None and self.action_selection == "stochastic":
            raise ValueError("Please provide a random number generator key to sample actions stochastically")
        if self.sampling_mode == "marginal":
            return control.sample_action(self.policies.policy_arr, self.num_controls, q_pi, self.action_selection,
                                         self.alpha, rng_key)
        return control.sample_policy(self.policies.policy_arr, q_pi, self.action_selection, self.alpha, rng_key)


def _generate_I(H_f: torch.Tensor, B_f: torch.Tensor, threshold: float, depth: int) -> torch.Tensor:
    """pymdp/control.py:965-1019 for one factor (factor-local B): row i marks
    the states that can reach the goal set H_f in i backward steps through
    transitions of probability above `threshold` under some action."""
    if depth < 1:
        raise ValueError("`depth` must be >= 1")
    reach = ((B_f > threshold).float().sum(-1) > 0).float()
    pred = reach.T
    rows = [H_f]
    cur = H_f
    for _ in range(depth - 1):
        cur = ((pred @ cur) > 0).float()
        rows.append(cur)
    return torch.stack(rows, 0)
