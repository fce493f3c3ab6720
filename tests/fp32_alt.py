"""A second FP32 restatement of run_factorized_fpi (pymdp/algos.py:88-155)
with a different summation order: modalities in reverse, and each
leave-one-out contraction done one axis at a time from the last axis as a
matrix-vector product.  Same algorithm and numerics (MINVAL clip, Jacobi
sweeps from log_q = 0, final softmax) as oracle.numpy_ref.run_factorized_fpi;
only the FP32 rounding differs.  Used to measure how far two FP32
implementations of the reference path can legitimately disagree."""
import numpy as np

from oracle import numpy_ref as R


def run_factorized_fpi_alt32(A, obs, prior, deps, num_iter, ll_rounding="float32"):
    """ll_rounding = "float32": ln A in FP32 (NumPy's float32 log), as the
    oracle's FP32 path; "rounded64": ln A computed in FP64 and rounded once to
    FP32 (the correctly rounded FP32 log), a perturbation of at most one ulp
    of each log-likelihood term."""
    dt = np.float32
    if len(prior) == 1:
        return R.run_factorized_fpi(A, obs, prior, deps, num_iter, dt)
    if ll_rounding == "rounded64":
        lls = [R.compute_log_likelihood_single_modality(o, a, np.float64).astype(dt) for o, a in zip(obs, A)]
    else:
        lls = [R.compute_log_likelihood_single_modality(o, a, dt) for o, a in zip(obs, A)]
    lp = [R.log_stable(p, dt) for p in prior]
    log_q = [np.zeros_like(x) for x in lp]
    for _ in range(num_iter):
        q = [R.softmax(x).astype(dt) for x in log_q]
        acc = [np.zeros_like(x) for x in lp]
        for m in reversed(range(len(lls))):
            for i, f in enumerate(deps[m]):
                x = lls[m]
                for j in reversed(range(len(deps[m]))):
                    if j != i:
                        x = np.moveaxis(x, j, -1) @ q[deps[m][j]]
                acc[f] = acc[f] + x.astype(dt)
        log_q = [a + x for a, x in zip(acc, lp)]
    return [R.softmax(x) for x in log_q]
