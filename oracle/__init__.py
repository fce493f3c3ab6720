"""CPU oracle for the FPI + EFE hot path.

TEST INFRASTRUCTURE ONLY. Nothing in the product package (`pymdp_amd`)
imports, links or executes anything in this folder. Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` use it,
and only as the checker / the timed CPU baseline, never as a code path of
the GPU implementation.

Contents
--------
numpy_ref.py   NumPy restatement of the reference's JAX math (FP32 mirrors
               the reference JAX path, FP64 is the high-precision oracle).
dedup_ref.py   NumPy restatement of the factor-history (action-history
               dedup) neg-EFE algorithm of the reference's native layer.
gen.py         Seeded synthetic generators for the five workloads.
cpu_baseline.c C/OpenMP restatement of the reference's CPU FFI algorithms
               (FPI and factor-history neg-EFE with JAX-path numerics), the
               timed CPU baseline.
Makefile       Builds cpu_baseline.c into oracle/_build/libpymdp_cpu_ref.so.

The reference's own C++ (`pymdp/ffi/**`) is not buildable here: every
translation unit includes `xla/ffi/api/ffi.h` (via
`pymdp/ffi/common/error_helpers.h:10`), which ships with jaxlib and is
absent from this image, so there is no `oracle/_ref` build.
"""
