// Host-side action-history plan (see history.hpp).
#include "history.hpp"

#include <unordered_map>

namespace pmh {

uint64_t fnv1a64(const void* data, size_t bytes, uint64_t h) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < bytes; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

namespace {

struct VecHash {
  size_t operator()(const std::vector<int32_t>& v) const {
    return (size_t)fnv1a64(v.data(), v.size() * sizeof(int32_t));
  }
};

using IdMap = std::unordered_map<std::vector<int32_t>, int32_t, VecHash>;

}  // namespace

std::string build_efe_host_plan(const int32_t* policy, int64_t P, int64_t T, int64_t F, const int64_t* S,
                                const int64_t* U, const int64_t* O, int64_t M, const int64_t* A_dep_flat,
                                const int64_t* A_dep_off, const int64_t* B_dep_flat, const int64_t* B_dep_off,
                                EfeHostPlan* out) {
  EfeHostPlan& pl = *out;
  pl = EfeHostPlan{};
  if (P <= 0 || T <= 0 || F <= 0 || M <= 0) return "policy_matrix shape is inconsistent with the model";
  pl.P = P;
  pl.T = T;
  pl.F = F;
  pl.M = M;
  pl.S.assign(S, S + F);
  pl.U.assign(U, U + F);
  pl.O.assign(O, O + M);
  for (int64_t f = 0; f < F; ++f) {
    if (S[f] <= 0) return "S[f] must be positive";
    if (U[f] <= 0) return "U[f] must be positive";
    const int64_t a = B_dep_off[f], e = B_dep_off[f + 1];
    if (e - a < 1 || e - a > 8) return "B_dependencies[f] rank must be in [1, 8]";
    std::vector<int64_t> d(B_dep_flat + a, B_dep_flat + e);
    for (auto x : d)
      if (x < 0 || x >= F) return "B_dependencies factor index out of range";
    pl.B_deps.push_back(d);
  }
  pl.mod_group.assign(M, -1);
  for (int64_t m = 0; m < M; ++m) {
    if (O[m] <= 0) return "O[m] must be positive";
    const int64_t a = A_dep_off[m], e = A_dep_off[m + 1];
    if (e - a < 1 || e - a > 8) return "A_dependencies[m] rank must be in [1, 8]";
    std::vector<int64_t> d(A_dep_flat + a, A_dep_flat + e);
    for (size_t i = 0; i < d.size(); ++i) {
      if (d[i] < 0 || d[i] >= F) return "A_dependencies factor index out of range";
      for (size_t j = 0; j < i; ++j)
        if (d[j] == d[i]) return "duplicate factor in A_dependencies";
    }
    pl.A_deps.push_back(d);
    int g = -1;
    for (size_t k = 0; k < pl.groups.size(); ++k)
      if (pl.groups[k].deps == d) g = (int)k;
    if (g < 0) {
      pl.groups.push_back(DepGroup{d, {}});
      g = (int)pl.groups.size() - 1;
    }
    pl.groups[g].modalities.push_back((int)m);
    pl.mod_group[m] = g;
  }

  // ---- factor histories (pymdp/ffi/neg_efe/factor_history_tree.h:57-126) ----
  pl.fac.assign(T, std::vector<FactorLevelPlan>(F));
  IdMap ids;
  std::vector<int32_t> key;
  for (int64_t t = 0; t < T; ++t) {
    for (int64_t f = 0; f < F; ++f) {
      FactorLevelPlan& lv = pl.fac[t][f];
      const auto& deps = pl.B_deps[f];
      lv.rank = (int)deps.size();
      lv.p2h.assign(P, -1);
      ids.clear();
      for (int64_t p = 0; p < P; ++p) {
        const int32_t u = policy[(p * T + t) * F + f];
        if (u < 0 || u >= U[f]) return "policy action out of bounds";
        key.clear();
        for (auto d : deps) key.push_back(t == 0 ? 0 : pl.fac[t - 1][d].p2h[p]);
        key.push_back(u);
        auto it = ids.find(key);
        int32_t id;
        if (it == ids.end()) {
          id = lv.n++;
          ids.emplace(key, id);
          for (size_t i = 0; i < deps.size(); ++i) lv.parents.push_back(key[i]);
          lv.action.push_back(u);
        } else {
          id = it->second;
        }
        lv.p2h[p] = id;
      }
    }
  }

  // ---- inner tuples per (t, dependency group) --------------------------------
  const size_t G = pl.groups.size();
  pl.grp.assign(T, std::vector<GroupLevelPlan>(G));
  for (int64_t t = 0; t < T; ++t) {
    for (size_t g = 0; g < G; ++g) {
      const auto& deps = pl.groups[g].deps;
      GroupLevelPlan& gl = pl.grp[t][g];
      gl.H0 = pl.fac[t][deps[0]].n;
      gl.p2score.assign(P, 0);
      ids.clear();
      for (int64_t p = 0; p < P; ++p) {
        key.clear();
        for (size_t i = 1; i < deps.size(); ++i) key.push_back(pl.fac[t][deps[i]].p2h[p]);
        auto it = ids.find(key);
        int32_t n;
        if (it == ids.end()) {
          n = gl.n_in++;
          ids.emplace(key, n);
          for (auto v : key) gl.tuples.push_back(v);
        } else {
          n = it->second;
        }
        const int64_t sidx = (int64_t)n * gl.H0 + pl.fac[t][deps[0]].p2h[p];
        if (sidx > INT32_MAX) return "too many history combinations";
        gl.p2score[p] = (int32_t)sidx;
      }
    }
  }

  // content key of everything the plan depends on
  uint64_t h = fnv1a64(policy, sizeof(int32_t) * (size_t)(P * T * F));
  h = fnv1a64(S, sizeof(int64_t) * F, h);
  h = fnv1a64(U, sizeof(int64_t) * F, h);
  h = fnv1a64(O, sizeof(int64_t) * M, h);
  h = fnv1a64(A_dep_off, sizeof(int64_t) * (M + 1), h);
  h = fnv1a64(A_dep_flat, sizeof(int64_t) * A_dep_off[M], h);
  h = fnv1a64(B_dep_off, sizeof(int64_t) * (F + 1), h);
  h = fnv1a64(B_dep_flat, sizeof(int64_t) * B_dep_off[F], h);
  const int64_t dims[4] = {P, T, F, M};
  pl.key = fnv1a64(dims, sizeof(dims), h);
  return "";
}

}  // namespace pmh
