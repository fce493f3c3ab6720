"""pymdp_hip_infer_states_noalloc (fixed device scratch, no allocation)
against pymdp_hip_infer_states, through the C ABI: bitwise-identical
posteriors on two calls in a row, the second with fewer agents, for the
streaming FPI variant (whose workspace is the fixed scratch) and for a small
model; a call needing more than the fixed scratch is rejected."""
import numpy as np
import pytest
import torch

from pymdp_amd import _native, utils
from pymdp_amd.ffi._fpi import _pack_infer_args

pytestmark = pytest.mark.gpu


def _call(entry, A, obs, D, deps, num_iter, batch):
    k = _pack_infer_args(A, obs, D, deps, num_iter, False, batch)
    totS = int(k["lp_offsets"][-1])
    out = torch.full((batch, totS), float("nan"), dtype=torch.float32, device=D[0].device)
    a = k["i64"]
    rc = getattr(_native.lib(), entry)(
        k["Ap"], a["As"].ptr, k["obs_t"].data_ptr(), k["is_index"], k["Pp"], a["Ps"].ptr, out.data_ptr(), batch,
        a["S"].ptr, k["F"], a["O"].ptr, k["M"], a["oo"].ptr, a["lo"].ptr, a["df"].ptr, a["do"].ptr,
        int(num_iter), 0, _native.stream_handle(D[0].device))
    return rc, out


def _case(dev, S, O, deps, batch, seed):
    g = torch.Generator(device=dev).manual_seed(seed)
    A = utils.random_A_array(O, S, deps, batch=batch, alpha=0.5, generator=g, device=dev)
    D = utils.random_factorized_categorical(S, batch=batch, generator=g, device=dev)
    obs = [torch.randint(0, o, (batch,), generator=g, device=dev) for o in O]
    return A, obs, D


# (S, O, deps, first batch, second batch): a rank-8 tuple of 32768 joint
# states (the streaming variant, ~700 KB of scratch per agent) and a small
# model (no scratch)
CASES = [
    ([4, 4, 4, 4, 4, 4, 4, 2], [3, 2], [[0, 1, 2, 3, 4, 5, 6, 7], [7, 0]], 80, 33),
    ([4, 3], [5, 6], [[0, 1], [1]], 64, 7),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_matches_infer_states(dev, case):
    S, O, deps, b1, b2 = CASES[case]
    print("case", case, S, O, deps, b1, b2, flush=True)
    for i, batch in enumerate((b1, b2)):
        A, obs, D = _case(dev, S, O, deps, batch, 100 + 10 * case + i)
        rc_ref, ref = _call("pymdp_hip_infer_states", A, obs, D, deps, 8, batch)
        _native.check(rc_ref)
        rc, out = _call("pymdp_hip_infer_states_noalloc", A, obs, D, deps, 8, batch)
        _native.check(rc)
        if case == 0:  # the streaming variant ran, i.e. the fixed scratch was used
            assert _native.lib().pymdp_hip_last_launch() == b"fpi_kernel<0,512>"
        torch.cuda.synchronize()
        assert torch.isfinite(out).all()
        assert torch.equal(out, ref), (case, batch, float((out - ref).abs().max()))


def test_noalloc_rejects_oversize(dev):
    S, O, deps = [4, 4, 4, 4, 4, 4, 4, 2], [3], [[0, 1, 2, 3, 4, 5, 6, 7]]
    g = torch.Generator(device=dev).manual_seed(5)
    # shared A and prior (stride 0): only the observations and q scale with batch
    A = utils.random_A_array(O, S, deps, batch=None, alpha=0.5, generator=g, device=dev)
    D = utils.random_factorized_categorical(S, batch=None, generator=g, device=dev)
    batch = 1024  # ~700 MB of workspace, beyond the 64 MiB scratch
    obs = [torch.randint(0, O[0], (batch,), generator=g, device=dev)]
    rc, _ = _call("pymdp_hip_infer_states_noalloc", A, obs, D, deps, 8, batch)
    assert rc == _native.PYMDP_HIP_EINVAL
    msg = _native.lib().pymdp_hip_last_error().decode()
    assert msg.startswith("fpi_ffi: ") and "scratch" in msg, msg
