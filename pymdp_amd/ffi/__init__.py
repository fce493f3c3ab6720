"""Native boundary of pymdp_amd (mirror of the reference's pymdp/ffi package).

FPI:  can_handle_fpi, fpi_ffi, infer_states_ffi, fpi_vfe_ffi   (pymdp/ffi/_fpi.py)
EFE:  can_handle, neg_efe_all_policies_ffi, flags      (pymdp/ffi/_efe.py)
"""
from ._core import (MAX_FFI_DEP_RANK, dep_indices_in_range, dep_list_rank_ok, flatten_dep_lists,
                    flatten_ragged_f32, is_available)
from ._fpi import can_handle_fpi, fpi_ffi, fpi_vfe_ffi, infer_states_ffi

__all__ = [
    "MAX_FFI_DEP_RANK", "dep_indices_in_range", "dep_list_rank_ok", "flatten_dep_lists",
    "flatten_ragged_f32", "is_available", "can_handle_fpi", "fpi_ffi", "infer_states_ffi", "fpi_vfe_ffi",
]
