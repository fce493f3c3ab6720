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
  grad_fpi_efe      512 agents, S=[16,8], U=[16,4], horizon 1, 3 modalities
                    of 16 outcomes, softmax-parameterised A, B, C, D

Every generator takes (seed, device, scale) where scale < 1 shrinks the
agent count for CPU oracle checks while keeping the per-agent shapes.
"""
from __future__ import annotations

import math

import torch

from . import utils


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
