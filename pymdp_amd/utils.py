"""Seeded random model generators (torch), after pymdp/utils.py:147-259.

The reference draws with jax.random; this package has no JAX, so the same
distributions (Dirichlet columns over the leading axis of A and B, Dirichlet
categorical vectors) are drawn with a seeded torch.Generator on the target
device.  Parity tests feed the *same* generated arrays to the kernels and to
the CPU oracle, so no cross-library bit-equality of random streams is needed.
"""
from __future__ import annotations

from typing import Sequence

import torch


def _gen(seed: int, device) -> torch.Generator:
    g = torch.Generator(device=torch.device(device))
    g.manual_seed(int(seed))
    return g


def dirichlet(shape: Sequence[int], alpha: float, dim: int, generator: torch.Generator,
              device="cpu") -> torch.Tensor:
    """Dirichlet(alpha) vectors along axis `dim` of a float32 tensor of `shape`
    (normalised Gamma(alpha) draws)."""
    conc = torch.full(tuple(shape), float(alpha), dtype=torch.float32, device=device)
    x = torch._standard_gamma(conc, generator=generator)
    del conc
    x /= x.sum(dim=dim, keepdim=True)
    return x


def random_A_array(num_obs: Sequence[int], num_states: Sequence[int], A_dependencies=None, *,
                   batch: int | None = None, alpha: float = 1.0, generator=None, device="cpu"):
    """pymdp/utils.py:163-194: A[m] of shape (batch?, O_m, S[deps]...), columns ~ Dirichlet(alpha)."""
    if A_dependencies is None:
        A_dependencies = [list(range(len(num_states))) for _ in num_obs]
    out = []
    for m, n_o in enumerate(num_obs):
        shp = (n_o,) + tuple(num_states[d] for d in A_dependencies[m])
        if batch is not None:
            shp = (batch,) + shp
        out.append(dirichlet(shp, alpha, 1 if batch is not None else 0, generator, device))
    return out


def random_B_array(num_states: Sequence[int], num_controls: Sequence[int], B_dependencies=None, *,
                   batch: int | None = None, alpha: float = 1.0, generator=None, device="cpu"):
    """pymdp/utils.py:196-259: B[f] of shape (batch?, S_f, S[deps]..., U_f), columns ~ Dirichlet."""
    if B_dependencies is None:
        B_dependencies = [[f] for f in range(len(num_states))]
    out = []
    for f, ns in enumerate(num_states):
        shp = (ns,) + tuple(num_states[d] for d in B_dependencies[f]) + (num_controls[f],)
        if batch is not None:
            shp = (batch,) + shp
        out.append(dirichlet(shp, alpha, 1 if batch is not None else 0, generator, device))
    return out


def random_factorized_categorical(dims: Sequence[int], *, batch: int | None = None, alpha: float = 1.0,
                                  generator=None, device="cpu"):
    """pymdp/utils.py:147-161: Dirichlet(alpha) categorical vectors."""
    out = []
    for d in dims:
        shp = (d,) if batch is None else (batch, d)
        out.append(dirichlet(shp, alpha, -1, generator, device))
    return out


def sample_observations(A: Sequence[torch.Tensor], D: Sequence[torch.Tensor], A_dependencies,
                        generator: torch.Generator) -> list[torch.Tensor]:
    """Seeded categorical observations from the generative model: a true
    state s_f ~ D_f per agent, then o_m ~ A_m[:, s_deps] (batched A)."""
    batch = D[0].shape[0]
    s = [torch.multinomial(d, 1, generator=generator)[:, 0] for d in D]
    obs = []
    ar = torch.arange(batch, device=D[0].device)
    for m, a in enumerate(A):
        idx = (ar, slice(None)) + tuple(s[f] for f in A_dependencies[m])
        col = a[idx]  # [batch, O_m]
        obs.append(torch.multinomial(col, 1, generator=generator)[:, 0])
    return obs
