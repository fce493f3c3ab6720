"""CPU oracle for the FPI + EFE hot path.

TEST INFRASTRUCTURE ONLY. Nothing in the product package (`pymdp_amd`)
imports, links or executes anything in this folder. Only `tests/`,
`__graft_entry__.smoke()` and `bench.py` use it: as the checker of the GPU
results, and in `bench.py` as the timed baselines the GPU path is compared
with (the CPU baseline; the torch-autograd gradient), never as a code path
of the GPU implementation.

Contents
--------
numpy_ref.py   NumPy restatement of the reference's JAX math (FP32 mirrors
               the reference JAX path, FP64 is the high-precision oracle).
dedup_ref.py   NumPy restatement of the factor-history (action-history
               dedup) neg-EFE algorithm of the reference's native layer.
grad_ref.py    FP64 finite-difference gradients of the trajectory loss.
seq_ref.py     Sequence inference restatements (MMP, VMP, history windows).
autograd_ref.py Pure-torch forward of the trajectory loss, differentiated by
               torch autograd: the comparison baseline of the HIP adjoints
               (timed on the GPU and, as the CPU baseline, on host cores).
cpu_baseline.c C/OpenMP restatement of the reference's CPU FFI algorithms
               (FPI and factor-history neg-EFE + q_pi + action, JAX-path
               numerics, FP32), the timed CPU baseline.
cpu_ref.py     ctypes wrapper of cpu_baseline.c (prebuilt x86-64-v3 library,
               or a -march=native build for the host that times it).
Makefile       Builds cpu_baseline.c into oracle/_build/libpymdp_cpu_ref.so
               (`make native`: -march=native).

The seeded synthetic workloads themselves are generated by
pymdp_amd/workloads.py (on the device), the same inputs for the product and
for these checkers.

The reference's own C++ (`pymdp/ffi/**`) is not buildable here: every
translation unit includes `xla/ffi/api/ffi.h` (via
`pymdp/ffi/common/error_helpers.h:10`), which ships with jaxlib and is
absent from this image, so there is no `oracle/_ref` build.
"""
