// Host-side action-history plan for the neg-EFE kernels.
//
// Per-(t, f) history dedup follows the algorithm of the reference's
// pymdp/ffi/neg_efe/factor_history_tree.h:57-126: at level t a history of
// factor f is identified by (the level-(t-1) histories of its B-dependency
// parents, u_f(t)); every policy maps to one history per (t, f).  The
// predicted marginal of factor f under a history is then computed once,
// however many policies share it.
//
// Modality scoring is organised for the gfx950 kernels (this part is not in
// the reference): for a modality with dependencies (d_0, ..., d_{R-1}) the
// first dependency is the "outer" one and the rest form an "inner tuple".
// At level t the realised inner tuples are deduplicated over policies; the
// predicted outcome of (inner tuple n, outer history h0) is
//     qo[n, h0] = W[n] . q_{d_0}[t][h0],   W[n] = A x_{inner} (x q_inner[n])
// so W is one GEMM over all inner tuples and the outcomes are a second,
// batched GEMM over the outer factor's histories.  Policy p reads its score
// at n(p) * H_{d_0}(t) + h_{d_0}(p).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pmh {

struct FactorLevelPlan {
  int n = 0;                     // histories of this factor at this level
  int rank = 0;                  // |B_dependencies[f]|
  std::vector<int32_t> parents;  // [n * rank] parent history ids at level t-1 (0 at t=0)
  std::vector<int32_t> action;   // [n]
  std::vector<int32_t> p2h;      // [P]
};

struct GroupLevelPlan {
  int n_in = 0;                  // realised inner tuples at this level
  int H0 = 0;                    // histories of the outer dependency at this level
  std::vector<int32_t> tuples;   // [n_in * (R-1)] inner history ids (positions 1..R-1)
  std::vector<int32_t> p2score;  // [P] n(p) * H0 + h0(p)
};

struct DepGroup {
  std::vector<int64_t> deps;     // dependency tuple shared by its modalities
  std::vector<int> modalities;   // modality ids in the group
};

struct EfeHostPlan {
  int64_t P = 0, T = 0, F = 0, M = 0;
  std::vector<int64_t> S, U, O;
  std::vector<std::vector<int64_t>> A_deps, B_deps;
  std::vector<DepGroup> groups;
  std::vector<int> mod_group;                    // [M]
  std::vector<std::vector<FactorLevelPlan>> fac;  // [T][F]
  std::vector<std::vector<GroupLevelPlan>> grp;   // [T][G]
  uint64_t key = 0;
};

// Builds the plan; returns "" on success or an error message (without prefix).
std::string build_efe_host_plan(const int32_t* policy, int64_t P, int64_t T, int64_t F, const int64_t* S,
                                const int64_t* U, const int64_t* O, int64_t M, const int64_t* A_dep_flat,
                                const int64_t* A_dep_off, const int64_t* B_dep_flat, const int64_t* B_dep_off,
                                EfeHostPlan* out);

// FNV-1a 64 over raw bytes (full content, not a sample).
uint64_t fnv1a64(const void* data, size_t bytes, uint64_t h = 14695981039346656037ull);

}  // namespace pmh
