// frontend.cpp -- host-side circuit preparation: single-qubit gate merging,
// level partitioning, and the fusion planner with its cost model.  Plain
// C++ on integer arrays; no device work.  See include/dmsim.h (dm_merge_*,
// dm_partition, dm_plan_fusion, dm_mlp_predict_host, dm_cost_block_seconds).

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <deque>
#include <map>
#include <set>
#include <tuple>
#include <vector>

#include "dmsim.h"

namespace dmsim_internal {
int set_error(int code, const char *msg);  // dmsim.hip: the message dm_last_error reports
}

namespace {

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return dmsim_internal::set_error(code, buf);
}

// ---------------------------------------------------------------------------
// Merging of single-qubit rotations (basicaertools.py:129-307: U3_merge,
// mergeU, merge_gates, single_gate_merge).  Runs of u1/u3 on one qubit
// between separators fold into one gate, the fold reusing the first record
// of the run; every separator flushes all pending runs in qubit order.
// ---------------------------------------------------------------------------

// Rz(alpha) Ry(beta) Rz(gamma) equal (up to phase) to Ry(th1) Rz(xi) Ry(th2):
// the expressions of U3_merge (basicaertools.py:162-178), evaluated in the
// same order.
void zyz_from_yzy(double xi, double th1, double th2, double *beta, double *alpha, double *gamma) {
  const double s_x = sin(xi * 0.5), c_x = cos(xi * 0.5);
  const double s_p = sin((th1 + th2) * 0.5), c_p = cos((th1 + th2) * 0.5);
  const double s_m = sin((th1 - th2) * 0.5), c_m = cos((th1 - th2) * 0.5);
  const double t_sum = atan2(s_x * c_m, c_x * c_p);
  const double t_dif = atan2(-s_x * s_m, c_x * s_p);
  const double a = c_x * c_p, b = s_x * c_m;
  *beta = 2 * acos(sqrt(a * a + b * b));
  *alpha = t_sum + t_dif;
  *gamma = t_sum - t_dif;
}

struct Rot {
  int src;       // record the result reuses
  int kind;      // DM_INST_U1 or DM_INST_U3
  double p[3];
};

// fold of two consecutive rotations on one qubit (the earlier one first)
Rot fold(const Rot &a, const Rot &b) {
  Rot out = a;
  if (a.kind == DM_INST_U1 && b.kind == DM_INST_U1) {
    out.p[0] = a.p[0] + b.p[0];
  } else if (a.kind == DM_INST_U1 && b.kind == DM_INST_U3) {
    out.kind = DM_INST_U3;
    out.p[0] = b.p[0];
    out.p[1] = b.p[1];
    out.p[2] = b.p[2] + a.p[0];
  } else if (a.kind == DM_INST_U3 && b.kind == DM_INST_U1) {
    out.p[0] = a.p[0];
    out.p[1] = a.p[1] + b.p[0];
    out.p[2] = a.p[2];
  } else {
    double beta, alpha, gamma;
    zyz_from_yzy(b.p[2] + a.p[1], b.p[0], a.p[0], &beta, &alpha, &gamma);
    out.p[0] = beta;
    out.p[1] = b.p[1] + alpha;
    out.p[2] = a.p[2] + gamma;
  }
  return out;
}

// ---------------------------------------------------------------------------
// Fusion planner (no reference counterpart).  Gates are noisy 1- or 2-qubit
// PTMs; a block is a set of gates that can run together (see dm_plan_fusion).
// ---------------------------------------------------------------------------

constexpr int kLowDigits = 3;       // every sequential tile pass holds the 3 lowest positions
constexpr int kMinTileQubits = 6;   // registers below this use blocks of at most 2 qubits
constexpr double kTimeWeight = 0.25;
constexpr int kNearLimit = 6, kNearWindow = 48;
constexpr int kFeatures = 34, kHidden = 64, kMaxPos = 20;
enum { K_PLAIN1Q = 0, K_PLAINCX = 1, K_SEQ = 2, K_DENSE = 3 };

struct Gate {
  int nq;
  int q[2];
  uint32_t mask;
};

struct Block {
  std::vector<int> qubits;   // in order of first use by the members
  std::vector<int> members;  // gate indices, ascending
};

inline int popc(uint32_t x) { return __builtin_popcount(x); }

uint32_t low_mask(int n) {
  uint32_t m = 0;
  for (int q = std::max(0, n - kLowDigits); q < n; ++q) m |= 1u << q;
  return m;
}

uint32_t mask_of(const std::vector<int> &qs) {
  uint32_t m = 0;
  for (int q : qs) m |= 1u << q;
  return m;
}

// The largest block (at most kmax qubits outside the free mask fm) seeded at
// the first pending gate: a gate joins when none of its qubits is blocked and
// the union stays within kmax; a gate that cannot join blocks its qubits.
// With fm the scan stops once the counted qubits are final.
void grow(const std::vector<Gate> &g, const std::vector<char> &done, int start, int kmax, int n, uint32_t fm,
          std::vector<int> *qs, std::vector<int> *members) {
  const uint32_t full = (n >= 32) ? ~0u : ((1u << n) - 1);
  uint32_t qm = 0, blocked = 0;
  qs->clear();
  members->clear();
  for (int i = start; i < (int)g.size(); ++i) {
    if (done[i]) continue;
    const uint32_t gm = g[i].mask;
    if (gm & blocked) {
      blocked |= gm;
    } else {
      const uint32_t nw = gm & ~qm & ~fm;
      if ((int)qs->size() + popc(nw) <= kmax) {
        if (nw) {
          for (int j = 0; j < g[i].nq; ++j)
            if ((nw >> g[i].q[j]) & 1) qs->push_back(g[i].q[j]);
          qm |= nw;
        }
        members->push_back(i);
      } else {
        blocked |= gm;
      }
    }
    if (fm) {
      if ((int)qs->size() >= kmax || (blocked | qm | fm) == full) break;
    } else if (qm && !(qm & ~blocked) && ((int)qs->size() >= kmax || (blocked | qm) == full)) {
      break;
    }
  }
}

// Gates of the block on qubit set S seeded at the first pending gate.
void absorb(const std::vector<Gate> &g, const std::vector<char> &done, int start, uint32_t S,
            std::vector<int> *members) {
  uint32_t blocked = 0;
  members->clear();
  for (int i = start; i < (int)g.size(); ++i) {
    if (done[i]) continue;
    const uint32_t gm = g[i].mask;
    if ((gm & blocked) || (gm & ~S)) {
      blocked |= gm;
      if (!(S & ~blocked)) break;
    } else {
      members->push_back(i);
    }
  }
}

// Up to `limit` other qubits, in order of first appearance, of the pending
// two-qubit gates within `window` gates that connect to the seed qubits.
std::vector<int> near_qubits(const std::vector<Gate> &g, const std::vector<char> &done, int start,
                             const std::vector<int> &seed, int limit, int window) {
  uint32_t reach = mask_of(seed);
  std::vector<int> out;
  int seen = 0;
  for (int i = start; i < (int)g.size(); ++i) {
    if (done[i]) continue;
    ++seen;
    if (seen > window || (int)out.size() >= limit) break;
    if (g[i].nq == 2 && (((reach >> g[i].q[0]) & 1) || ((reach >> g[i].q[1]) & 1)))
      for (int j = 0; j < 2; ++j)
        if (!((reach >> g[i].q[j]) & 1)) {
          reach |= 1u << g[i].q[j];
          out.push_back(g[i].q[j]);
        }
  }
  if ((int)out.size() > limit) out.resize(limit);
  return out;
}

// Qubits of the tile of a sequential pass over T (outside the low positions):
// T, the low positions, filled up to 6 digits (|T| <= 2, and |T| = 3 in
// FP64) or 7 (|T| = 4, and |T| = 3 in FP32: runs of 4^4 coefficients).
uint32_t tile_mask(int n, const std::vector<int> &T, bool fp32) {
  const int m = std::min(n, T.size() <= 2 || (T.size() == 3 && !fp32) ? 6 : 7);
  uint32_t S = mask_of(T) | low_mask(n);
  for (int p = 0; popc(S) < m; ++p) S |= 1u << (n - 1 - p);
  return S;
}

std::vector<int> used_qubits(const std::vector<Gate> &g, const std::vector<int> &members) {
  std::vector<int> qs;
  uint32_t seen = 0;
  for (int i : members)
    for (int j = 0; j < g[i].nq; ++j)
      if (!((seen >> g[i].q[j]) & 1)) {
        seen |= 1u << g[i].q[j];
        qs.push_back(g[i].q[j]);
      }
  return qs;
}

template <typename F>
void combinations(const std::vector<int> &pool, int k, F &&emit) {
  std::vector<int> idx(k);
  for (int i = 0; i < k; ++i) idx[i] = i;
  const int m = (int)pool.size();
  if (k > m) return;
  std::vector<int> pick(k);
  while (true) {
    for (int i = 0; i < k; ++i) pick[i] = pool[idx[i]];
    emit(pick);
    int i = k - 1;
    while (i >= 0 && idx[i] == m - k + i) --i;
    if (i < 0) return;
    ++idx[i];
    for (int j = i + 1; j < k; ++j) idx[j] = idx[j - 1] + 1;
  }
}

// Distinct candidate blocks seeded at `start`: the greedy growth for every
// k <= kmax; every seed-containing qubit set of at most kmax qubits drawn
// from the nearby qubits; with `tiles`, for each set T outside the low
// positions (greedy growth counting only such qubits, and the seed's high
// qubits with nearby ones), every gate the whole tile of T admits.
std::vector<Block> candidate_blocks(const std::vector<Gate> &g, const std::vector<char> &done, int start, int kmax,
                                    int n, bool tiles, bool fp32) {
  std::vector<Block> out;
  std::set<std::vector<int>> keys;
  auto add = [&](const std::vector<int> &members) {
    if (members.empty() || keys.count(members)) return;
    keys.insert(members);
    out.push_back(Block{used_qubits(g, members), members});
  };
  std::vector<int> seed(g[start].q, g[start].q + g[start].nq);
  std::vector<int> qs, mem;
  add({start});
  for (int k = 2; k <= kmax; ++k) {
    grow(g, done, start, k, n, 0, &qs, &mem);
    add(mem);
  }
  const std::vector<int> near = near_qubits(g, done, start, seed, kNearLimit, kNearWindow);
  for (int extra = 1; extra <= kmax - (int)seed.size(); ++extra)
    combinations(near, extra, [&](const std::vector<int> &T) {
      std::vector<int> set = seed;
      set.insert(set.end(), T.begin(), T.end());
      absorb(g, done, start, mask_of(set), &mem);
      add(mem);
    });
  if (!tiles) return out;
  const uint32_t low = low_mask(n);
  std::set<std::vector<int>> sets;  // sorted qubit lists, in lexicographic order
  for (int k = 1; k <= kmax; ++k) {
    grow(g, done, start, k, n, low, &qs, &mem);
    std::vector<int> s = qs;
    std::sort(s.begin(), s.end());
    sets.insert(s);
  }
  std::vector<int> hi;
  for (int q : seed)
    if (!((low >> q) & 1)) hi.push_back(q);
  std::vector<int> near2;
  for (int q : near_qubits(g, done, start, seed, kNearLimit + kLowDigits, kNearWindow))
    if (!((low >> q) & 1) && near2.size() < 6) near2.push_back(q);
  for (int extra = 0; extra <= kmax - (int)hi.size(); ++extra) {
    if (extra == 0) {
      std::vector<int> s = hi;
      std::sort(s.begin(), s.end());
      sets.insert(s);
      continue;
    }
    combinations(near2, extra, [&](const std::vector<int> &X) {
      std::vector<int> s = hi;
      s.insert(s.end(), X.begin(), X.end());
      std::sort(s.begin(), s.end());
      s.erase(std::unique(s.begin(), s.end()), s.end());
      sets.insert(s);
    });
  }
  for (const std::vector<int> &T : sets) {
    absorb(g, done, start, tile_mask(n, T, fp32), &mem);
    add(mem);
  }
  return out;
}

// --- cost model --------------------------------------------------------------

// w: the lowest positions a tile of m digits holds when it is `positions`
// completed with the lowest other positions (runs of 4^w coefficients).
int run_digits(int n, uint32_t positions, int m) {
  uint32_t inside = positions;
  for (int p = 0; popc(inside) < m && p < n; ++p) inside |= 1u << p;
  int w = 0;
  while ((inside >> w) & 1) ++w;
  return w;
}

// Tile digits of a sequential pass over `positions` (k qubits): 7 where 6
// would leave runs of 4^2 coefficients, and in FP32 where 7 turns runs of
// 4^3 into runs of 4^4 (measured faster on MI355X); else 6.
int tile_for(int n, uint32_t positions, int k, bool fp32) {
  const int w6 = run_digits(n, positions, 6), w7 = run_digits(n, positions, 7);
  if (n >= 7 && (k > 6 || (w6 < 3 && w7 > w6) || (fp32 && w6 == 3 && w7 == 4))) return 7;
  return 6;
}

float mlp_forward(const float *p, int F, int H, const float *x) {
  const float *W1 = p, *b1 = W1 + (size_t)F * H, *W2 = b1 + H, *b2 = W2 + (size_t)H * H, *W3 = b2 + H,
              *b3 = W3 + H;
  std::vector<float> h1(H), h2(H);
  for (int j = 0; j < H; ++j) {
    float acc = b1[j];
    for (int i = 0; i < F; ++i) acc += x[i] * W1[(size_t)i * H + j];
    h1[j] = acc > 0.f ? acc : 0.f;
  }
  for (int j = 0; j < H; ++j) {
    float acc = b2[j];
    for (int i = 0; i < H; ++i) acc += h1[i] * W2[(size_t)i * H + j];
    h2[j] = acc > 0.f ? acc : 0.f;
  }
  float y = b3[0];
  for (int i = 0; i < H; ++i) y += h2[i] * W3[i];
  return y;
}

struct Stats {
  int n1, n2, n2low, d2;
};

struct CostModel {
  const float *params = nullptr;
  int F = kFeatures, H = kHidden;
  int fp64 = 0;
  std::map<std::tuple<int, uint32_t, int, int, int, int>, double> cache;

  // the feature vector of one pass (kind, block qubits, gate statistics)
  void features(int kind, int n, const std::vector<int> &qubits, const Stats &st, float *x) const {
    std::vector<int> pos;
    uint32_t pm = 0;
    for (int q : qubits) {
      pos.push_back(n - 1 - q);
      pm |= 1u << (n - 1 - q);
    }
    std::sort(pos.begin(), pos.end());
    for (int i = 0; i < F; ++i) x[i] = 0.f;
    x[kind] = 1.f;
    x[4] = (float)((double)pos.size() / 6.0);
    x[5] = (float)(log2(1.0 + st.n1) / 4.0);
    x[6] = (float)(log2(1.0 + st.n2) / 4.0);
    int w;
    if (kind == K_SEQ) w = run_digits(n, pm, tile_for(n, pm, (int)qubits.size(), !fp64));
    else if (kind == K_DENSE) w = run_digits(n, pm, 6);
    else w = pos[0] > 0 ? 1 : 0;
    x[7] = (float)(w / 6.0);
    x[8] = fp64 ? 1.f : 0.f;
    x[9] = (float)(n / 20.0);
    x[10] = (float)((double)pos.front() / (double)n);
    x[11] = (float)((double)pos.back() / (double)n);
    for (int p : pos)
      if (p < kMaxPos) x[12 + p] = 1.f;
    x[32] = (float)(log2(1.0 + st.n2low) / 4.0);
    x[33] = (float)(st.d2 / 7.0);
  }

  double predict(int kind, int n, const std::vector<int> &qubits, const Stats &st) {
    const auto key = std::make_tuple(kind, mask_of(qubits), st.n1, st.n2, st.n2low, st.d2);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    float x[kFeatures];
    features(kind, n, qubits, st, x);
    const float y = mlp_forward(params, F, H, x);
    const double gib = ldexp(1.0, 2 * n + (fp64 ? 3 : 2) - 30);
    const double t = exp((double)y) * gib;
    cache.emplace(key, t);
    return t;
  }

  // predicted seconds of a block: a plain kernel for one qubit or one gate,
  // else the cheaper of the sequential and dense tile forms (2-4 qubits)
  double block_time(const std::vector<Gate> &g, int n, const Block &b) {
    Stats st{0, 0, 0, 0};
    const uint32_t low = low_mask(n);
    uint32_t q2 = 0;
    for (int i : b.members) {
      if (g[i].nq == 1) {
        ++st.n1;
      } else {
        ++st.n2;
        if (g[i].mask & low) ++st.n2low;
        q2 |= g[i].mask;
      }
    }
    st.d2 = popc(q2);
    if (b.qubits.size() == 1) return predict(K_PLAIN1Q, n, b.qubits, st);
    if (b.members.size() == 1) return predict(K_PLAINCX, n, b.qubits, st);
    double t = predict(K_SEQ, n, b.qubits, st);
    if (b.qubits.size() >= 2 && b.qubits.size() <= 4) t = std::min(t, predict(K_DENSE, n, b.qubits, st));
    return t;
  }
};

struct PlanCtx {
  const std::vector<Gate> &g;
  int n, km;
  bool tiles;
  CostModel *cost;
  std::vector<int> widths;
};

void state_after(const std::vector<Gate> &g, const std::vector<char> &done, int first, const std::vector<int> &mem,
                 std::vector<char> *d2, int *f2) {
  *d2 = done;
  for (int j : mem) (*d2)[j] = 1;
  int f = first;
  while (f < (int)g.size() && (*d2)[f]) ++f;
  *f2 = f;
}

// indices of the k best scores, ties in index order
std::vector<int> top_k(const std::vector<double> &score, int k) {
  std::vector<int> order(score.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return score[a] > score[b]; });
  if ((int)order.size() > k) order.resize(k);
  return order;
}

std::vector<double> scores(PlanCtx &c, const std::vector<Block> &cands, double per_gate) {
  std::vector<double> s(cands.size());
  for (size_t i = 0; i < cands.size(); ++i)
    s[i] = (double)cands[i].members.size() * per_gate - kTimeWeight * c.cost->block_time(c.g, c.n, cands[i]);
  return s;
}

// best total saving of the blocks that can follow, searched over the
// widths[depth..] best candidates at each step; past the last width the best
// score alone
double best_after(PlanCtx &c, const std::vector<char> &done, int first, int depth, double per_gate) {
  if (first >= (int)c.g.size()) return 0.0;
  const std::vector<Block> cands = candidate_blocks(c.g, done, first, c.km, c.n, c.tiles, !c.cost->fp64);
  const std::vector<double> s = scores(c, cands, per_gate);
  if (depth >= (int)c.widths.size()) return *std::max_element(s.begin(), s.end());
  double best = 0.0;
  bool any = false;
  for (int i : top_k(s, c.widths[depth])) {
    std::vector<char> d2;
    int f2;
    state_after(c.g, done, first, cands[i].members, &d2, &f2);
    const double v = s[i] + best_after(c, d2, f2, depth + 1, per_gate);
    if (!any || v > best) {
      best = v;
      any = true;
    }
  }
  return best;
}

}  // namespace

extern "C" {

int dm_merge_rotations(int ninst, const int32_t *kind, const int32_t *qubit, const double *params, int merge,
                       int32_t *out_src, int32_t *out_kind, double *out_params, int *nout) {
  if (ninst < 0 || (ninst > 0 && (!kind || !qubit || !params)) || !out_src || !out_kind || !out_params || !nout)
    return fail(DM_ERR_ARG, "dm_merge_rotations: bad arguments");
  int k = 0;
  auto emit = [&](const Rot &r) {
    out_src[k] = r.src;
    out_kind[k] = r.kind;
    for (int j = 0; j < 3; ++j) out_params[3 * k + j] = r.p[j];
    ++k;
  };
  std::map<int, std::vector<Rot>> pending;  // qubit -> run, flushed in qubit order
  auto flush = [&]() {
    for (auto &kv : pending) {
      Rot acc = kv.second[0];
      for (size_t j = 1; j < kv.second.size(); ++j) acc = fold(acc, kv.second[j]);
      emit(acc);
    }
    pending.clear();
  };
  for (int i = 0; i < ninst; ++i) {
    const int kd = kind[i];
    if (kd == DM_INST_ID) continue;
    if (kd == DM_INST_U1 || kd == DM_INST_U3) {
      Rot r{i, kd, {params[3 * i], params[3 * i + 1], params[3 * i + 2]}};
      if (!merge) emit(r);
      else pending[qubit[i]].push_back(r);
      continue;
    }
    if (kd != DM_INST_SEPARATOR) return fail(DM_ERR_ARG, "dm_merge_rotations: instruction %d has kind %d", i, kd);
    if (merge) flush();
    Rot r{i, kd, {0, 0, 0}};
    emit(r);
  }
  if (merge) flush();
  *nout = k;
  return DM_OK;
}

int dm_partition(int n, int ninst, const int32_t *kind, const int32_t *q0, const int32_t *q1, int32_t *out_ops,
                 int32_t *out_level_start, int *nlevels, int *count) {
  if (n < 1 || n > 30 || ninst < 0 || (ninst > 0 && (!kind || !q0 || !q1)) || !out_ops || !out_level_start ||
      !nlevels || !count)
    return fail(DM_ERR_ARG, "dm_partition: bad arguments");
  // entries of the per-qubit stacks: an instruction index, or a placeholder
  // (-1 - kind) a measure / reset group leaves on the other qubits.  A Bell /
  // Expect / Ensemble measure opens a level of its own only at the head of a
  // segment; elsewhere it is a measure.
  std::vector<std::vector<int>> levels_all;
  int total = 0;
  std::vector<int> kk(kind, kind + ninst);
  for (int j = 0; j < ninst; ++j)
    if (kk[j] == DM_INST_OWN_LEVEL && j > 0 && kind[j - 1] != DM_INST_BARRIER) kk[j] = DM_INST_MEASURE;
  auto name_of = [&](int e) { return e >= 0 ? kk[e] : -1 - e; };
  int i = 0;
  while (i <= ninst) {
    // one barrier-free segment [s, i)
    const int s = i;
    while (i < ninst && kind[i] != DM_INST_BARRIER) ++i;
    const int e = i;
    ++i;  // past the barrier
    if (e == s) continue;
    if (kk[s] == DM_INST_OWN_LEVEL) {
      std::vector<int> lv;
      for (int j = s; j < e; ++j) lv.push_back(j);
      levels_all.push_back(lv);
      total += 1;
      continue;
    }
    std::vector<std::deque<int>> st(n);
    for (int j = s; j < e; ++j) {
      const int kd = kk[j];
      if (kd == DM_INST_ROT || kd == DM_INST_CX) {
        if (q0[j] < 0 || q0[j] >= n || (kd == DM_INST_CX && (q1[j] < 0 || q1[j] >= n || q1[j] == q0[j])))
          return fail(DM_ERR_ARG, "dm_partition: instruction %d has bad qubits", j);
        st[q0[j]].push_back(j);
        if (kd == DM_INST_CX) st[q1[j]].push_back(j);
        continue;
      }
      if (kd != DM_INST_MEASURE && kd != DM_INST_RESET)
        return fail(DM_ERR_ARG, "dm_partition: instruction %d has kind %d", j, kd);
      const int hold = kd == DM_INST_MEASURE ? DM_INST_HOLD_MEASURE : DM_INST_HOLD_RESET;
      const int q = q0[j];
      if (q < 0 || q >= n) return fail(DM_ERR_ARG, "dm_partition: instruction %d has bad qubits", j);
      if (!st[q].empty() && name_of(st[q].back()) == hold) {
        st[q].back() = j;
      } else {
        st[q].push_back(j);
        for (int o = 0; o < n; ++o)
          if (o != q) st[o].push_back(-1 - hold);
      }
    }
    size_t depth = 0;
    for (auto &d : st) depth = std::max(depth, d.size());
    std::vector<std::vector<int>> levels(depth);
    int left = e - s, level = 0, stall = 0;
    while (left) {
      const int before = left;
      uint32_t used = 0;  // n <= 30
      if (level == (int)levels.size()) levels.emplace_back();
      for (int q = 0; q < n; ++q) {
        if (st[q].empty()) continue;
        const int head = st[q].front();
        const int nm = name_of(head);
        if (nm == DM_INST_HOLD_MEASURE || nm == DM_INST_HOLD_RESET) continue;
        if (nm == DM_INST_ROT) {
          if ((used >> q) & 1) continue;
          levels[level].push_back(head);
          used |= 1u << q;
          st[q].pop_front();
          --left;
        } else if (nm == DM_INST_CX) {
          const int other = q0[head] == q ? q1[head] : q0[head];
          if (((used >> q) & 1) || ((used >> other) & 1)) continue;
          if (st[other].front() == head) {
            used |= (1u << q) | (1u << other);
            levels[level].push_back(head);
            st[q].pop_front();
            st[other].pop_front();
            --left;
          }
        } else {
          const int hold = nm == DM_INST_MEASURE ? DM_INST_HOLD_MEASURE : DM_INST_HOLD_RESET;
          bool all = true;
          for (int x = 0; x < n && all; ++x)
            all = st[x].empty() || name_of(st[x].front()) == nm || name_of(st[x].front()) == hold;
          if (all) {
            if (!levels[level].empty()) {
              used = 0;
              ++level;
              if (level == (int)levels.size()) levels.emplace_back();
            }
            for (int x = 0; x < n; ++x) {
              if (st[x].empty()) continue;
              if (name_of(st[x].front()) == nm) {
                used |= 1u << x;
                levels[level].push_back(st[x].front());
                --left;
              }
              st[x].pop_front();
            }
            break;
          }
        }
        if (!left) break;
      }
      ++level;
      stall = left == before ? stall + 1 : 0;
      if (stall > 2) return fail(DM_ERR_ARG, "partition cannot place the remaining %d instructions", left);
    }
    for (auto &lv : levels) levels_all.push_back(lv);
    total += level;
  }
  int k = 0;
  out_level_start[0] = 0;
  for (size_t l = 0; l < levels_all.size(); ++l) {
    for (int op : levels_all[l]) out_ops[k++] = op;
    out_level_start[l + 1] = k;
  }
  *nlevels = (int)levels_all.size();
  *count = total;
  return DM_OK;
}

int dm_mlp_predict_host(const float *params, int n_features, int hidden, const float *X, int n, float *out) {
  if (!params || n_features < 1 || hidden < 1 || n < 0 || (n > 0 && (!X || !out)))
    return fail(DM_ERR_ARG, "dm_mlp_predict_host: bad arguments");
  for (int r = 0; r < n; ++r) out[r] = mlp_forward(params, n_features, hidden, X + (size_t)r * n_features);
  return DM_OK;
}

int dm_plan_fusion(int n, int ngates, const int32_t *gate_qubits, int method, int kmax, int tiles,
                   const int32_t *lookahead, int nlookahead, const float *mlp_params, int precision,
                   int32_t *out_members, int32_t *out_block_start, int32_t *out_qubits, int32_t *out_qubit_start,
                   int *nblocks) {
  if (n < 1 || n > 20 || ngates < 0 || (ngates > 0 && !gate_qubits) || kmax < 1 || kmax > 7 || method < 0 ||
      method > 2 || nlookahead < 0 || (nlookahead > 0 && !lookahead) || !out_members || !out_block_start ||
      !out_qubits || !out_qubit_start || !nblocks)
    return fail(DM_ERR_ARG, "dm_plan_fusion: bad arguments");
  if (method == DM_PLAN_LEARNED && !mlp_params) return fail(DM_ERR_ARG, "dm_plan_fusion: learned planning needs a model");
  std::vector<Gate> g(ngates);
  for (int i = 0; i < ngates; ++i) {
    const int a = gate_qubits[2 * i], b = gate_qubits[2 * i + 1];
    if (a < 0 || a >= n || b >= n || b == a) return fail(DM_ERR_ARG, "dm_plan_fusion: gate %d has bad qubits", i);
    g[i].nq = b < 0 ? 1 : 2;
    g[i].q[0] = a;
    g[i].q[1] = b < 0 ? a : b;
    g[i].mask = (1u << a) | (b < 0 ? 0u : 1u << b);
    // a block of at most kmax qubits must be able to hold every gate, or the
    // plan could not advance past it
    if (g[i].nq > kmax) return fail(DM_ERR_ARG, "dm_plan_fusion: kmax=%d cannot hold gate %d on 2 qubits", kmax, i);
  }
  const bool allow_mfma = n >= kMinTileQubits;
  const int km = allow_mfma ? kmax : std::min(kmax, 2);
  CostModel cost;
  cost.params = mlp_params;
  cost.fp64 = precision == DM_FP64;
  PlanCtx ctx{g, n, km, tiles && allow_mfma && method == DM_PLAN_LEARNED, &cost,
              std::vector<int>(lookahead, lookahead + nlookahead)};
  std::vector<char> done(ngates, 0);
  int first = 0, nb = 0, nm = 0, nq = 0;
  double spent_t = 0.0;
  long spent_g = 0;
  out_block_start[0] = 0;
  out_qubit_start[0] = 0;
  std::vector<int> qs, mem;
  while (true) {
    while (first < ngates && done[first]) ++first;
    if (first >= ngates) break;
    Block pick;
    if (method == DM_PLAN_GREEDY) {
      grow(g, done, first, km, n, 0, &qs, &mem);
      pick = Block{qs, mem};
    } else {
      std::vector<Block> cands = candidate_blocks(g, done, first, km, n, ctx.tiles, precision != DM_FP64);
      if (method == DM_PLAN_SEARCH) {
        size_t best = 0;
        for (size_t i = 1; i < cands.size(); ++i)
          if (cands[i].members.size() > cands[best].members.size()) best = i;
        pick = cands[best];
      } else {
        // predicted net saving: the gates a block takes, valued at the plan's
        // running cost per gate, minus its weighted own time
        std::vector<double> t(cands.size());
        for (size_t i = 0; i < cands.size(); ++i) t[i] = cost.block_time(g, n, cands[i]);
        double per_gate;
        if (spent_g) {
          per_gate = spent_t / (double)spent_g;
        } else {
          size_t j = 0;
          for (size_t i = 1; i < cands.size(); ++i)
            if (cands[i].members.size() > cands[j].members.size()) j = i;
          per_gate = t[j] / (double)cands[j].members.size();
        }
        std::vector<double> score(cands.size());
        for (size_t i = 0; i < cands.size(); ++i)
          score[i] = (double)cands[i].members.size() * per_gate - kTimeWeight * t[i];
        const std::vector<int> &w = ctx.widths;
        const std::vector<int> top = top_k(score, std::max(1, w.empty() ? 1 : w[0]));
        int choice = top[0];
        if (w.size() >= 2 || (!w.empty() && w[0] > 1)) {
          double best = 0.0;
          bool any = false;
          for (int i : top) {
            std::vector<char> d2;
            int f2;
            state_after(g, done, first, cands[i].members, &d2, &f2);
            const double v = score[i] + best_after(ctx, d2, f2, 1, per_gate);
            if (!any || v > best) {
              best = v;
              choice = i;
              any = true;
            }
          }
        }
        spent_t += t[choice];
        spent_g += (long)cands[choice].members.size();
        pick = cands[choice];
      }
    }
    // every block takes at least one pending gate, so there are at most
    // ngates blocks; an empty one would never advance
    if (pick.members.empty()) return fail(DM_ERR_ARG, "dm_plan_fusion: no block takes gate %d", first);
    for (int i : pick.members) {
      done[i] = 1;
      out_members[nm++] = i;
    }
    for (int q : pick.qubits) out_qubits[nq++] = q;
    ++nb;
    out_block_start[nb] = nm;
    out_qubit_start[nb] = nq;
  }
  *nblocks = nb;
  return DM_OK;
}

int dm_cost_block_seconds(int n, int ngates, const int32_t *gate_qubits, int nblocks, const int32_t *members,
                          const int32_t *block_start, const int32_t *qubits, const int32_t *qubit_start,
                          const float *mlp_params, int precision, double *out) {
  if (n < 1 || n > 20 || ngates < 0 || nblocks < 0 || !mlp_params || (nblocks > 0 && (!members || !block_start ||
      !qubits || !qubit_start || !out)) || (ngates > 0 && !gate_qubits))
    return fail(DM_ERR_ARG, "dm_cost_block_seconds: bad arguments");
  std::vector<Gate> g(ngates);
  for (int i = 0; i < ngates; ++i) {
    const int a = gate_qubits[2 * i], b = gate_qubits[2 * i + 1];
    if (a < 0 || a >= n || b >= n || b == a) return fail(DM_ERR_ARG, "dm_cost_block_seconds: gate %d has bad qubits", i);
    g[i].nq = b < 0 ? 1 : 2;
    g[i].q[0] = a;
    g[i].q[1] = b < 0 ? a : b;
    g[i].mask = (1u << a) | (b < 0 ? 0u : 1u << b);
  }
  CostModel cost;
  cost.params = mlp_params;
  cost.fp64 = precision == DM_FP64;
  for (int b = 0; b < nblocks; ++b) {
    Block blk;
    for (int j = block_start[b]; j < block_start[b + 1]; ++j) {
      if (members[j] < 0 || members[j] >= ngates) return fail(DM_ERR_ARG, "dm_cost_block_seconds: bad member");
      blk.members.push_back(members[j]);
    }
    for (int j = qubit_start[b]; j < qubit_start[b + 1]; ++j) blk.qubits.push_back(qubits[j]);
    if (blk.members.empty() || blk.qubits.empty()) return fail(DM_ERR_ARG, "dm_cost_block_seconds: empty block");
    out[b] = cost.block_time(g, n, blk);
  }
  return DM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batch circuit programs (no reference counterpart): the op records of one
// circuit of a batch, built on the host from its gate list.  A circuit has a
// barrier after every gate, so every gate is followed by one memory-noise
// level (dm_simulator.py:1172-1177, _add_decoherence_and_amp_decay :397-425).
// ---------------------------------------------------------------------------

namespace {

// A noisy gate: a 4x4 PTM on one qubit, or a CX with optional per-pair
// scales (the depolarising channel on both qubits).
struct NoisyOp {
  int nq;
  int q[2];
  bool scaled;
  double R[16];  // 1 qubit: row-major 4x4 PTM; CX: 16 per-pair scales
};

void mat4_mul(const double *A, const double *B, double *C) {  // C = A B (C may alias B)
  double t[16];
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k) {
      double s = 0.0;
      for (int j = 0; j < 4; ++j) s += A[4 * i + j] * B[4 * j + k];
      t[4 * i + k] = s;
    }
  memcpy(C, t, sizeof(t));
}

// M <- R M, R the rotation of the component pair (k0, k1) by angle:
// (k0, k1) -> (c k0 - s k1, c k1 + s k0) (rot_gate_dm_matrix,
// basicaertools.py:93-126, without rotation error).
void rotate4(int k0, int k1, double angle, double *M) {
  double R[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  const double c = cos(angle + 0.0), s = sin(angle + 0.0);
  R[4 * k0 + k0] = c;
  R[4 * k0 + k1] = -s;
  R[4 * k1 + k1] = c;
  R[4 * k1 + k0] = s;
  mat4_mul(R, M, M);
}

// u3(theta, phi, lambda) = rz(phi) ry(theta) rz(lambda) and u1(lambda) =
// rz(lambda) as 4x4 PTMs (single_gate_dm_matrix, basicaertools.py:69-90);
// component pairs rz: (X, Y), ry: (Z, X).
void rotation_ptm(int kind, const double *a, double *M) {
  static const double I4[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  memcpy(M, I4, sizeof(I4));
  if (kind == DM_INST_U1) {
    rotate4(1, 2, a[0], M);
  } else {
    rotate4(1, 2, a[2], M);
    rotate4(3, 1, a[0], M);
    rotate4(1, 2, a[1], M);
  }
}

// Tile digits of a batched pass over the block's positions: 6, or in FP32
// 7 where 6 would leave runs of only 4^2 coefficients that 7 lengthens, or
// runs of 4^3 that 7 makes 4^4.
int batch_tile_digits(int n, uint32_t positions, int nq, bool fp32) {
  if (fp32 && n >= 7 && nq > 1) {
    const int w6 = run_digits(n, positions, 6), w7 = run_digits(n, positions, 7);
    if ((w6 < 3 && w7 > w6) || (w6 == 3 && w7 == 4)) return 7;
  }
  return 6;
}

struct ProgramOut {
  std::vector<dm_op> ops;
  std::vector<double> vals;
  void add(int kind, int k, const int *qs, int nq, const double *v, int nv) {
    dm_op o;
    memset(&o, 0, sizeof(o));
    o.kind = kind;
    o.k = k;
    for (int j = 0; j < nq; ++j) o.qubits[j] = qs[j];
    o.param = (int64_t)vals.size();
    ops.push_back(o);
    vals.insert(vals.end(), v, v + nv);
  }
  void plain(const NoisyOp &g) {  // a gate on its own kernel
    if (g.nq == 1)
      add(DM_OP_1Q, 1, g.q, 1, g.R + 4, 12);
    else if (g.scaled)
      add(DM_OP_CX_SCALED, 2, g.q, 2, g.R, 16);
    else
      add(DM_OP_CX, 2, g.q, 2, nullptr, 0);
  }
};

}  // namespace

extern "C" int dm_batch_circuit_program(int n, int ngates, const int32_t *kind, const int32_t *q0,
                                        const int32_t *q1, const double *angles, double decay_factor,
                                        double decoherence_factor, double thermal_factor,
                                        double gate_depolarization, int fused, int precision, dm_op *ops,
                                        int max_ops, double *values, int64_t max_values, int *nops,
                                        int64_t *nvalues) {
  if (n < 1 || n > 12 || ngates < 0 || (ngates > 0 && (!kind || !q0 || !q1 || !angles)) || !nops || !nvalues ||
      max_ops < 0 || max_values < 0 || (precision != DM_FP64 && precision != DM_FP32))
    return fail(DM_ERR_ARG, "dm_batch_circuit_program: bad arguments");
  if (fused && n < 6) return fail(DM_ERR_ARG, "dm_batch_circuit_program: fused programs need n >= 6 (n=%d)", n);
  // at most n + 1 records per gate; op counts are int
  if ((int64_t)ngates * (n + 1) > (int64_t)INT32_MAX / 2)
    return fail(DM_ERR_ARG, "dm_batch_circuit_program: %d gates are too many", ngates);
  const double p = gate_depolarization, lam = 1.0 - p;
  std::vector<NoisyOp> g;
  g.reserve((size_t)ngates * (fused ? n + 1 : 1));
  ProgramOut out;
  // memory noise: x,y <- sqrt(g) f (x,y); z <- g z + (1-g)(2p-1) a_I
  double mem[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  const double od = sqrt(decay_factor) * decoherence_factor;
  mem[5] = mem[10] = od;
  mem[15] = decay_factor;
  mem[12] = (1 - decay_factor) * (2 * thermal_factor - 1);
  const bool mem_identity = decay_factor == 1.0 && decoherence_factor == 1.0;
  const double level[3] = {decoherence_factor, thermal_factor, decay_factor};
  for (int i = 0; i < ngates; ++i) {
    NoisyOp o;
    memset(&o, 0, sizeof(o));
    if (kind[i] == DM_INST_U1 || kind[i] == DM_INST_U3) {
      if (q0[i] < 0 || q0[i] >= n) return fail(DM_ERR_ARG, "dm_batch_circuit_program: gate %d has bad qubits", i);
      o.nq = 1;
      o.q[0] = q0[i];
      rotation_ptm(kind[i], angles + 3 * i, o.R);
      if (p != 0.0)  // depolarising channel after the gate: X, Y, Z rows times 1-p
        for (int e = 4; e < 16; ++e) o.R[e] *= lam;
    } else if (kind[i] == DM_INST_CX) {
      if (q0[i] < 0 || q0[i] >= n || q1[i] < 0 || q1[i] >= n || q0[i] == q1[i])
        return fail(DM_ERR_ARG, "dm_batch_circuit_program: gate %d has bad qubits", i);
      o.nq = 2;
      o.q[0] = q0[i];
      o.q[1] = q1[i];
      o.scaled = p != 0.0;
      if (o.scaled)  // (1-p)^(number of non-identity components of the pair)
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b) o.R[4 * a + b] = pow(lam, (double)((a != 0) + (b != 0)));
    } else {
      return fail(DM_ERR_ARG, "dm_batch_circuit_program: gate %d is not u1, u3 or cx (kind %d)", i, kind[i]);
    }
    if (!fused) {
      out.plain(o);
      out.add(DM_OP_MEMORY_NOISE, 0, nullptr, 0, level, 3);
      continue;
    }
    g.push_back(o);
    if (!mem_identity)
      for (int q = 0; q < n; ++q) {
        NoisyOp m;
        memset(&m, 0, sizeof(m));
        m.nq = 1;
        m.q[0] = q;
        memcpy(m.R, mem, sizeof(mem));
        g.push_back(m);
      }
  }
  if (fused && !g.empty()) {
    const int ng = (int)g.size();
    std::vector<int32_t> gq(2 * (size_t)ng), members(ng), bstart(ng + 1), bq(2 * (size_t)ng), qstart(ng + 1);
    for (int i = 0; i < ng; ++i) {
      gq[2 * i] = g[i].q[0];
      gq[2 * i + 1] = g[i].nq == 2 ? g[i].q[1] : -1;
    }
    int nb = 0;
    if (int rc = dm_plan_fusion(n, ng, gq.data(), DM_PLAN_GREEDY, 4, 0, nullptr, 0, nullptr, precision,
                                members.data(), bstart.data(), bq.data(), qstart.data(), &nb))
      return rc;
    std::vector<double> rec;
    for (int b = 0; b < nb; ++b) {
      const int m0 = bstart[b], m1 = bstart[b + 1];
      const int *qs = bq.data() + qstart[b];
      const int nq = qstart[b + 1] - qstart[b];
      if (m1 - m0 == 1) {
        out.plain(g[members[m0]]);
        continue;
      }
      if (nq == 1) {  // a run on one qubit: the product of its PTMs
        double M[16];
        memcpy(M, g[members[m0]].R, sizeof(M));
        for (int j = m0 + 1; j < m1; ++j) mat4_mul(g[members[j]].R, M, M);
        out.add(DM_OP_1Q, 1, qs, 1, M + 4, 12);
        continue;
      }
      uint32_t pos = 0;
      for (int j = 0; j < nq; ++j) pos |= 1u << (n - 1 - qs[j]);
      const int tile = batch_tile_digits(n, pos, nq, precision == DM_FP32);
      // sub-operations: each run of 1-qubit PTMs on one qubit folded into one,
      // emitted before the next CX on its qubit (or at the end, by qubit)
      std::vector<std::pair<int, std::vector<double>>> subs;  // (-1 - qubit) for 1q, else member index
      double pend[kMaxPos][16];
      bool has[kMaxPos] = {};
      auto flush = [&](int q) {
        if (!has[q]) return;
        subs.emplace_back(-1 - q, std::vector<double>(pend[q] + 4, pend[q] + 16));
        has[q] = false;
      };
      for (int j = m0; j < m1; ++j) {
        const NoisyOp &o = g[members[j]];
        if (o.nq == 1) {
          const int q = o.q[0];
          if (has[q]) {
            mat4_mul(o.R, pend[q], pend[q]);
          } else {
            memcpy(pend[q], o.R, sizeof(o.R));
            has[q] = true;
          }
          continue;
        }
        flush(o.q[0]);
        flush(o.q[1]);
        subs.emplace_back(members[j], o.scaled ? std::vector<double>(o.R, o.R + 16) : std::vector<double>());
      }
      for (int q = 0; q < n; ++q) flush(q);
      rec.clear();
      rec.push_back(tile);
      rec.push_back(nq);
      for (int j = 0; j < nq; ++j) rec.push_back(qs[j]);
      rec.push_back((double)subs.size());
      for (const auto &sv : subs) {
        if (sv.first < 0) {
          const double h[7] = {1, 1, (double)(-1 - sv.first), 0, 0, 0, 12};
          rec.insert(rec.end(), h, h + 7);
        } else {
          const NoisyOp &o = g[sv.first];
          const double h[7] = {2, 2, (double)o.q[0], (double)o.q[1], 0, 0, (double)sv.second.size()};
          rec.insert(rec.end(), h, h + 7);
        }
        rec.insert(rec.end(), sv.second.begin(), sv.second.end());
      }
      out.add(DM_OP_FUSED, 0, nullptr, 0, rec.data(), (int)rec.size());
    }
  }
  *nops = (int)out.ops.size();
  *nvalues = (int64_t)out.vals.size();
  if (!ops && !values) return DM_OK;  // size query
  if (!ops || !values || *nops > max_ops || *nvalues > max_values)
    return fail(DM_ERR_ARG, "dm_batch_circuit_program: output holds %d ops / %lld values, need %d / %lld", max_ops,
                (long long)max_values, *nops, (long long)*nvalues);
  if (!out.ops.empty()) memcpy(ops, out.ops.data(), out.ops.size() * sizeof(dm_op));
  if (!out.vals.empty()) memcpy(values, out.vals.data(), out.vals.size() * sizeof(double));
  return DM_OK;
}
