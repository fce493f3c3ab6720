"""The C ABI library builds, loads and exports every declared symbol; invalid
arguments are rejected on the host with the reference's error prefixes
(no GPU needed: validation happens before any device call)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from pymdp_amd import _native

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    text = (ROOT / "include" / "pymdp_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pymdp_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    syms = _declared_symbols()
    assert syms, "no declarations parsed"
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _native.SIGNATURES, f"{s} has no ctypes signature"


def test_version_string():
    assert b"gfx950" in _native.lib().pymdp_hip_version()


def _fpi_call(S, deps, ll_sizes=None, num_iter=16, lp_count=None, ll_count=None):
    S = np.asarray(S, dtype=np.int64)
    F = len(S)
    M = len(deps)
    ll_sizes = ll_sizes or [int(np.prod([S[d] for d in dl])) for dl in deps]
    ll_off = np.concatenate([[0], np.cumsum(ll_sizes)]).astype(np.int64)
    lp_off = np.concatenate([[0], np.cumsum(S)]).astype(np.int64)
    flat = np.array([d for dl in deps for d in dl], dtype=np.int64)
    doff = np.concatenate([[0], np.cumsum([len(d) for d in deps])]).astype(np.int64)
    a = [_native.I64(x) for x in (S, ll_off, lp_off, flat, doff)]
    lp_count = int(lp_off[-1]) if lp_count is None else lp_count
    ll_count = int(ll_off[-1]) if ll_count is None else ll_count
    rc = _native.lib().pymdp_hip_fpi(None, None, None, ll_count, lp_count, lp_count, a[0].ptr, F, a[1].ptr, M,
                                     a[2].ptr, a[3].ptr, a[4].ptr, num_iter, None)
    return rc, _native.lib().pymdp_hip_last_error().decode()


def test_fpi_rejects_duplicate_dependency():
    rc, msg = _fpi_call([3, 4], [[0, 0]], ll_sizes=[9])
    assert rc == _native.PYMDP_HIP_EINVAL
    assert msg.startswith("fpi_ffi: ") and "duplicate" in msg


def test_fpi_rejects_rank_over_8():
    rc, msg = _fpi_call([2] * 9, [list(range(9))])
    assert rc == _native.PYMDP_HIP_EINVAL and "rank" in msg


def test_fpi_rejects_nonpositive_num_iter():
    rc, msg = _fpi_call([3, 4], [[0, 1]], num_iter=0)
    assert rc == _native.PYMDP_HIP_EINVAL and "num_iter" in msg


def test_fpi_rejects_bad_ll_offsets():
    rc, msg = _fpi_call([3, 4], [[0, 1]], ll_sizes=[11])
    assert rc == _native.PYMDP_HIP_EINVAL and "ll_offsets" in msg


def test_fpi_rejects_indivisible_batch():
    rc, msg = _fpi_call([3, 4], [[0, 1]], lp_count=8)
    assert rc == _native.PYMDP_HIP_EINVAL and "not divisible" in msg


def test_fpi_rejects_out_of_range_factor():
    rc, msg = _fpi_call([3, 4], [[0, 2]], ll_sizes=[12])
    assert rc == _native.PYMDP_HIP_EINVAL and "out of range" in msg


def test_check_maps_einval_to_valueerror():
    _fpi_call([3, 4], [[0, 0]], ll_sizes=[9])
    with pytest.raises(ValueError, match="fpi_ffi: "):
        _native.check(_native.PYMDP_HIP_EINVAL)


def test_can_handle_fpi_mirrors_reference_gate():
    import torch
    from pymdp_amd.ffi import can_handle_fpi
    prior = [torch.ones(3), torch.ones(4)]
    assert can_handle_fpi(prior, [[0, 1]], 16)
    assert not can_handle_fpi(prior, [[0, 0]], 16)      # duplicate
    assert not can_handle_fpi(prior, [[0, 1]], 0)       # num_iter <= 0
    assert not can_handle_fpi(prior, [[0, 5]], 16)      # out of range
    assert not can_handle_fpi(prior, [[]], 16)          # rank 0
