"""Seeded synthetic workloads of the FPI / EFE hot path.

  plain_fpi_small   64 agents, S=[4,3], O=[5,6], A Dirichlet(1), D uniform,
                    uniform one-hot observations, 16 sweeps (CPU-runnable case)
  plain_fpi_large   1024 agents, S=[16,16,8,8] (16384 joint states), 6
                    modalities of 32 outcomes, A Dirichlet(0.5) (2 MiB per
                    modality per agent), D Dirichlet(1), observations sampled
                    from the generative model, 16 sweeps
  plain_efe         64 agents, S=[32,16,8], U=[8,4,1], horizon 3 (32768
                    policies), 4 modalities of 64 outcomes, A Dirichlet(0.5),
                    B Dirichlet(0.2), C ~ N(0,1), gamma 16, E uniform
  ragged            4096 agents of different shapes: per agent F ~ U{1..4}
                    factors of S ~ U{2..64} states and U ~ U{1..8} controls,
                    M ~ U{1..5} modalities of O ~ U{2..64} outcomes, each
                    depending on r ~ U{1..min(F,3)} distinct factors, horizon
                    T ~ U{1,2}; an agent is redrawn while its policy count
                    (prod U)^T exceeds 4096 or its A exceeds 4 MiB.  A and B
                    Dirichlet(0.5) columns, factor-local B, C ~ N(0,1),
                    D Dirichlet(1), observation indices uniform, 16 sweeps,
                    gamma 16, E uniform
  grad_fpi_efe      512 agents, S=[16,8], U=[16,4], horizon 1, 3 modalities
                    of 16 outcomes, softmax-parameterised A, B, C, D

Every generator takes (seed, device, scale) where scale < 1 shrinks the
agent count for CPU oracle checks while keeping the per-agent shapes.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import control, utils


def plain_fpi_small(seed: int = 0, device="cpu", batch: int = 64) -> dict:
    g = utils._gen(seed, device)
    S, O = [4, 3], [5, 6]
    deps = [[0, 1], [0, 1]]
    A = utils.random_A_array(O, S, deps, batch=batch, alpha=1.0, generator=g, device=device)
    D = [torch.full((batch, s), 1.0 / s, device=device) for s in S]
    obs = [torch.randint(0, o, (batch,), generator=g, device=device) for o in O]
    return dict(name="plain_fpi_small", S=S, O=O, A=A, D=D, obs=obs, A_dependencies=deps, num_iter=16,
                batch=batch)


def plain_fpi_large(seed: int = 0, device="cpu", batch: int = 1024) -> dict:
    g = utils._gen(seed, device)
    S, O = [16, 16, 8, 8], [32] * 6
    deps = [[0, 1, 2, 3] for _ in O]
    A = utils.random_A_array(O, S, deps, batch=batch, alpha=0.5, generator=g, device=device)
    D = utils.random_factorized_categorical(S, batch=batch, alpha=1.0, generator=g, device=device)
    obs = utils.sample_observations(A, D, deps, g)
    return dict(name="plain_fpi_large", S=S, O=O, A=A, D=D, obs=obs, A_dependencies=deps, num_iter=16,
                batch=batch)


def learning_large(seed: int = 0, device="cpu", batch: int = 1024) -> dict:
    """Dirichlet learning of A on the large plain-FPI model: 1024 agents,
    S=[16,16,8,8], 6 modalities x 32 outcomes (2 MiB of counts per modality
    per agent), pA ~ U[1, 2), one time step of Dirichlet(1) beliefs and
    uniform observation indices."""
    g = utils._gen(seed, device)
    S, O = [16, 16, 8, 8], [32] * 6
    deps = [[0, 1, 2, 3] for _ in O]
    pA = [1.0 + torch.rand((batch, o) + tuple(S), generator=g, device=device) for o in O]
    qs = [q[:, None, :] for q in utils.random_factorized_categorical(S, batch=batch, alpha=1.0, generator=g,
                                                                      device=device)]
    obs = [torch.randint(0, o, (batch, 1), generator=g, device=device) for o in O]
    return dict(name="learning_large", S=S, O=O, pA=pA, qs=qs, obs=obs, A_dependencies=deps, batch=batch)


def plain_efe(seed: int = 0, device="cpu", batch: int = 64, S=(32, 16, 8), U=(8, 4, 1), T: int = 3,
              O=(64, 64, 64, 64)) -> dict:
    g = utils._gen(seed, device)
    S, U, O = list(S), list(U), list(O)
    F = len(S)
    A_deps = [list(range(F)) for _ in O]
    B_deps = [[f] for f in range(F)]
    A = utils.random_A_array(O, S, A_deps, batch=batch, alpha=0.5, generator=g, device=device)
    B = utils.random_B_array(S, U, B_deps, batch=batch, alpha=0.2, generator=g, device=device)
    C = [torch.randn((batch, o), generator=g, device=device) for o in O]
    qs = utils.random_factorized_categorical(S, batch=batch, alpha=1.0, generator=g, device=device)
    return dict(name="plain_efe", S=S, U=U, O=O, T=T, A=A, B=B, C=C, qs=qs, A_dependencies=A_deps,
                B_dependencies=B_deps, gamma=16.0, batch=batch,
                num_policies=int(math.prod(U) ** T))


def ragged(seed: int = 0, device="cpu", n_agents: int = 4096, max_policies: int = 4096,
           max_A_bytes: int = 4 << 20, state_range=(2, 64), obs_range=(2, 64), max_factors: int = 4,
           max_controls: int = 8, max_modalities: int = 5, horizons=(1, 2), num_iter: int = 16,
           gamma: float = 16.0, tensors: bool = True) -> dict:
    """Independent agents of different shapes (see the module docstring).
    Shapes come from a NumPy PCG64 stream, tensors from a torch generator on
    `device`, both seeded by `seed`; tensors=False keeps only the shapes
    (same shapes, policies and observations)."""
    rng = np.random.default_rng(seed)
    g = utils._gen(seed, device)
    agents = []
    for _ in range(n_agents):
        while True:
            F = int(rng.integers(1, max_factors + 1))
            S = [int(x) for x in rng.integers(state_range[0], state_range[1] + 1, F)]
            U = [int(x) for x in rng.integers(1, max_controls + 1, F)]
            M = int(rng.integers(1, max_modalities + 1))
            O = [int(x) for x in rng.integers(obs_range[0], obs_range[1] + 1, M)]
            T = int(rng.choice(list(horizons)))
            A_deps = []
            for _m in range(M):
                r = int(rng.integers(1, min(F, 3) + 1))
                A_deps.append(sorted(int(x) for x in rng.choice(F, r, replace=False)))
            P = math.prod(U) ** T
            A_bytes = sum(4 * O[m] * math.prod(S[d] for d in A_deps[m]) for m in range(M))
            if P <= max_policies and A_bytes <= max_A_bytes:
                break
        B_deps = [[f] for f in range(F)]
        A = B = C = D = None
        if tensors:
            A = [utils.dirichlet([O[m]] + [S[d] for d in A_deps[m]], 0.5, 0, g, device) for m in range(M)]
            B = [utils.dirichlet([S[f], S[f], U[f]], 0.5, 0, g, device) for f in range(F)]
            C = [torch.randn(O[m], generator=g, device=device) for m in range(M)]
            D = [utils.dirichlet([S[f]], 1.0, 0, g, device) for f in range(F)]
        obs = [int(rng.integers(0, O[m])) for m in range(M)]
        policies = control.construct_policies(S, U, T)
        agents.append(dict(S=S, U=U, O=O, T=T, A=A, B=B, C=C, D=D, E=None, obs=obs, A_dependencies=A_deps,
                           B_dependencies=B_deps, policies=policies, gamma=float(gamma)))
    return dict(name="ragged", agents=agents, num_iter=num_iter, n_agents=n_agents)


def grad_fpi_efe(seed: int = 0, device="cpu", batch: int = 512, T: int = 32, S=(16, 8), U=(16, 4),
                 O=(16, 16, 16)) -> dict:
    """Model fitting through FPI + EFE: per agent, logits of A [O_m, S...]
    (every modality depends on every factor), B [S_f, S_f, U_f]
    (factor-local), C [O_m] and D [S_f], all N(0,1); policies of horizon 1
    over U (P = prod U); T-step trajectories of observation indices and
    taken policy indices, uniform."""
    g = utils._gen(seed, device)
    S, U, O = list(S), list(U), list(O)
    F = len(S)
    A_deps = [list(range(F)) for _ in O]
    B_deps = [[f] for f in range(F)]
    A_logits = [torch.randn((batch, o, *S), generator=g, device=device) for o in O]
    B_logits = [torch.randn((batch, s, s, u), generator=g, device=device) for s, u in zip(S, U)]
    C_logits = [torch.randn((batch, o), generator=g, device=device) for o in O]
    D_logits = [torch.randn((batch, s), generator=g, device=device) for s in S]
    policies = control.construct_policies(S, U, 1)
    obs = torch.stack([torch.randint(0, o, (T, batch), generator=g, device=device) for o in O], -1).to(torch.int32)
    actions = torch.randint(0, policies.shape[0], (T, batch), generator=g, device=device)
    return dict(name="grad_fpi_efe", S=S, U=U, O=O, T=T, batch=batch, A_logits=A_logits, B_logits=B_logits,
                C_logits=C_logits, D_logits=D_logits, policies=policies, obs=obs, actions=actions,
                A_dependencies=A_deps, B_dependencies=B_deps, gamma=1.0, num_iter=16)
