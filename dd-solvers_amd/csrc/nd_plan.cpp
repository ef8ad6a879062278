// Symbolic analysis of the sparse block LDL^T solver (host code): nested-dissection
// ordering of a symmetric sparse pattern and every table the device factorisation and the
// two device sweeps read. Built once per pattern at setup; nothing here touches the GPU.
//
// Ordering. Recursive bisection: on each connected piece, a breadth-first search from its
// smallest-numbered node finds a pseudo-peripheral node (the last one reached); the level
// sets of a second search from there are the separator candidates, and the separator is
// the smallest level set leaving at least `balance` of the remaining nodes on either side
// (else the most balanced one). The two sides are dissected in turn down to `leaf_size`
// nodes; disconnected pieces become sibling subtrees. Searches visit neighbours in
// increasing index order, so the ordering is a function of the pattern alone.
//
// Fronts. Tree nodes at depth < d_top (the largest depth bound whose nodes hold at most
// `top_max` unknowns together) merge into one dense top front numbered last; the other
// nodes are numbered in post-order of their subtrees, so every front's unknowns J are one
// contiguous range. A front's coupling set S is the set of ancestor unknowns that J touches
// after elimination: J's own neighbours above J plus its children's coupling sets above J.
// Fronts are then renumbered by height (leaves 0), one device launch per height.
//
// Tables (element (r, c) of a front's stacked factor [Finv; W] at foff + c * m + r, m =
// nJ + nS; the row-major copy of W at wtoff): child maps (positions of a child's coupling
// set in its parent's front), forward slots (soff, wbase), backward buffers and push lists
// (moff, push_dst, push_src, push_ptr), and the per-height work items of the two sweeps.
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "dd_solvers_hip.h"

namespace dds {
void set_error(const char* fmt, ...);
}

namespace {

using i64 = int64_t;
using i32 = int32_t;
using V64 = std::vector<i64>;
using V32 = std::vector<i32>;

// LDS capacity of the sweep kernels in elements (f_J forward, x_S backward): 60 KiB of FP64
constexpr i64 kLdsCap = 7680;
constexpr double kBalance = 0.47;

struct NdNode {
  V32 J;  // original indices of the unknowns eliminated here, ascending
  i64 depth;
  i64 parent;
  V64 children;
};

// Separator level: the smallest level set leaving both sides >= kBalance of the rest, else
// the most balanced one (ties: the lowest level).
i64 pick_level(const V32& dist) {
  const i64 nl = (i64)dist.back() + 1;
  V64 cnt(nl, 0);
  for (i32 d : dist) ++cnt[d];
  const i64 n = (i64)dist.size();
  i64 before = cnt[0];
  i64 best_ok = -1, best_any = -1;
  double best_bal = -1.0;
  for (i64 l = 1; l + 1 < nl; ++l) {
    const i64 after = n - before - cnt[l];
    const i64 den = std::max<i64>(n - cnt[l], 1);
    const double bal = (double)std::min(before, after) / (double)den;
    if (bal >= kBalance && (best_ok < 0 || cnt[l] < cnt[best_ok])) best_ok = l;
    if (bal > best_bal) {
      best_bal = bal;
      best_any = l;
    }
    before += cnt[l];
  }
  return best_ok >= 0 ? best_ok : best_any;
}

// Outcome of dissecting one piece: a tree node (a leaf, or a piece no level set splits), a
// separator with its two sides, or a disconnected piece's components (no node of its own).
struct Part {
  enum Kind { kNode, kSplit, kPieces } kind = kNode;
  V32 J;                               // kNode: the piece; kSplit: the separator
  std::vector<std::unique_ptr<Part>> sub;  // kSplit: {A, B}; kPieces: in order
};

struct Job {
  V32 ids;  // the piece's nodes, ascending
  Part* out;
};

// Recursive bisection of every piece, pieces in parallel (they share no node, so the node
// labels `tag` and search stamps `seen` of two pieces never overlap; stamps are globally
// unique). The tree is then numbered sequentially in the order a depth-first dissection
// would create its nodes, so the result does not depend on the thread count.
class Dissector {
 public:
  Dissector(i64 n, const i64* ptr, const i32* idx, i64 leaf)
      : ptr_(ptr), idx_(idx), leaf_(leaf), tag_(n), seen_(n, 0) {
    for (auto& t : tag_) t.store(-1, std::memory_order_relaxed);
  }

  std::unique_ptr<Part> run(int threads) {
    auto root = std::make_unique<Part>();
    V32 all(tag_.size());
    std::iota(all.begin(), all.end(), 0);
    queue_.push_back(Job{std::move(all), root.get()});
    pending_ = 1;
    std::vector<std::thread> pool;
    for (int i = 1; i < threads; ++i) pool.emplace_back([this] { worker(); });
    worker();
    for (auto& t : pool) t.join();
    return root;
  }

 private:
  // pieces at least this large go back to the shared queue; smaller ones stay on the thread
  static constexpr i64 kShare = 8192;
  const i64* ptr_;
  const i32* idx_;
  i64 leaf_;
  std::vector<std::atomic<i32>> tag_;
  std::vector<i32> seen_;
  std::atomic<i32> next_tag_{0}, next_stamp_{0};
  std::mutex mu_;
  std::condition_variable cv_;
  std::vector<Job> queue_;
  i64 pending_ = 0;  // jobs queued or running

  struct Search {
    V32 order, dist;
    i32 stamp = 0;
  };

  i32 tag(i32 v) const { return tag_[v].load(std::memory_order_relaxed); }
  void set_tag(i32 v, i32 t) { tag_[v].store(t, std::memory_order_relaxed); }

  // breadth-first search over the nodes tagged tg: visit order and distance of each
  void bfs(Search& s, i32 src, i32 tg) {
    s.stamp = next_stamp_.fetch_add(1, std::memory_order_relaxed) + 1;
    s.order.clear();
    s.dist.clear();
    s.order.push_back(src);
    s.dist.push_back(0);
    seen_[src] = s.stamp;
    for (size_t h = 0; h < s.order.size(); ++h) {
      const i32 v = s.order[h];
      const i32 d = s.dist[h] + 1;
      for (i64 k = ptr_[v]; k < ptr_[v + 1]; ++k) {
        const i32 u = idx_[k];
        if (u == v || tag(u) != tg || seen_[u] == s.stamp) continue;
        seen_[u] = s.stamp;
        s.order.push_back(u);
        s.dist.push_back(d);
      }
    }
  }

  // one piece: fills job.out and returns the pieces still to dissect
  void split(Job& job, Search& s, std::vector<Job>& more) {
    Part& out = *job.out;
    V32& ids = job.ids;
    const i64 m = (i64)ids.size();
    if (m <= leaf_) {
      out.kind = Part::kNode;
      out.J = std::move(ids);
      return;
    }
    const i32 tg = next_tag_.fetch_add(1, std::memory_order_relaxed);
    for (i32 v : ids) set_tag(v, tg);
    bfs(s, ids[0], tg);
    if ((i64)s.order.size() < m) {
      // disconnected: components in the order of their smallest node, each its own subtree
      out.kind = Part::kPieces;
      const i32 first = s.stamp;
      V32 p0;
      for (i32 v : ids)
        if (seen_[v] == first) p0.push_back(v);
      std::vector<V32> pieces;
      pieces.push_back(std::move(p0));
      for (i32 v : ids) {
        if (seen_[v] >= first) continue;  // in a component already found
        bfs(s, v, tg);
        V32 pc(s.order.begin(), s.order.end());
        std::sort(pc.begin(), pc.end());
        pieces.push_back(std::move(pc));
      }
      for (i32 v : ids) set_tag(v, -1);
      for (auto& pc : pieces) {
        out.sub.push_back(std::make_unique<Part>());
        more.push_back(Job{std::move(pc), out.sub.back().get()});
      }
      return;
    }
    bfs(s, s.order.back(), tg);
    if (s.dist.back() < 2) {  // no level set splits it: keep it whole
      for (i32 v : ids) set_tag(v, -1);
      out.kind = Part::kNode;
      out.J = std::move(ids);
      return;
    }
    const i64 L = pick_level(s.dist);
    for (size_t k = 0; k < s.order.size(); ++k) set_tag(s.order[k], (i32)(-2 - s.dist[k]));
    V32 A, S, B;
    for (i32 v : ids) {
      const i64 d = -2 - (i64)tag(v);
      if (d < L) A.push_back(v);
      else if (d == L) S.push_back(v);
      else B.push_back(v);
    }
    for (i32 v : ids) set_tag(v, -1);
    out.kind = Part::kSplit;
    out.J = std::move(S);
    out.sub.push_back(std::make_unique<Part>());
    out.sub.push_back(std::make_unique<Part>());
    more.push_back(Job{std::move(A), out.sub[0].get()});
    more.push_back(Job{std::move(B), out.sub[1].get()});
  }

  void worker() {
    Search s;
    std::vector<Job> mine, more;
    for (;;) {
      Job job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [this] { return !queue_.empty() || pending_ == 0; });
        if (queue_.empty()) return;  // pending_ == 0: everything is done
        job = std::move(queue_.back());
        queue_.pop_back();
      }
      mine.clear();
      mine.push_back(std::move(job));
      i64 shared = 0;
      std::vector<Job> out_jobs;
      while (!mine.empty()) {
        Job j = std::move(mine.back());
        mine.pop_back();
        more.clear();
        split(j, s, more);
        for (auto& c : more) {
          if ((i64)c.ids.size() >= kShare) out_jobs.push_back(std::move(c));
          else mine.push_back(std::move(c));
        }
      }
      {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& c : out_jobs) queue_.push_back(std::move(c));
        shared = (i64)out_jobs.size();
        pending_ += shared - 1;
      }
      if (shared > 1) cv_.notify_all();
      else if (shared == 1) cv_.notify_one();
      else if (pending_ == 0) cv_.notify_all();
    }
  }
};

// number the tree in the order of a depth-first dissection (side A before side B,
// components in order)
void number(Part& p, i64 depth, i64 parent, std::vector<NdNode>& nodes) {
  if (p.kind == Part::kPieces) {
    for (auto& c : p.sub) number(*c, depth, parent, nodes);
    return;
  }
  nodes.push_back(NdNode{std::move(p.J), depth, parent, {}});
  const i64 me = (i64)nodes.size() - 1;
  if (parent >= 0) nodes[parent].children.push_back(me);
  if (p.kind == Part::kSplit)
    for (auto& c : p.sub) number(*c, depth + 1, me, nodes);
}

int plan_threads() {
  int t = (int)std::thread::hardware_concurrency();
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) t = std::min(t > 0 ? t : v, v);
  }
  return std::max(1, std::min(t, 16));
}

std::vector<NdNode> dissect(i64 n, const i64* ptr, const i32* idx, i64 leaf_size) {
  Dissector d(n, ptr, idx, leaf_size);
  std::unique_ptr<Part> root = d.run(plan_threads());
  std::vector<NdNode> nodes;
  number(*root, 0, -1, nodes);
  return nodes;
}

i64 lower_bound_idx(const i64* a, i64 n, i64 v) { return (i64)(std::lower_bound(a, a + n, v) - a); }

}  // namespace

// The plan and its tables, owned by the library until dds_nd_plan_free.
struct dds_nd_plan {
  i64 scalars[DDS_ND_N_SCALARS] = {};
  V64 perm, fj0, fnj, fns, level_ptr, foff, wtoff, soff, wbase, moff, push_dst, push_ptr,
      parent, child_ptr, child_idx, s_idx, ap_ptr, ap_src, fwd_ptr, fwd_items, fwd_lds, bwd_ptr,
      bwd_items, bwd_lds, uoff2;
  V32 maps, push_src, ap_idx;
};

namespace {

int build_plan(dds_nd_plan& P, i64 n, const i64* ptr, const i32* idx, i64 leaf_size,
               i64 top_max) {
  std::vector<NdNode> nodes = dissect(n, ptr, idx, leaf_size);
  const i64 nn = (i64)nodes.size();
  i64 dmax = 0;
  for (auto& nd : nodes) dmax = std::max(dmax, nd.depth + 1);
  V64 by_depth(dmax + 1, 0);
  for (auto& nd : nodes) by_depth[nd.depth] += (i64)nd.J.size();
  // largest d whose nodes of depth < d hold at most top_max unknowns
  i64 d_top = 0, acc = 0;
  while (d_top < dmax && acc + by_depth[d_top] <= top_max) acc += by_depth[d_top++];
  std::vector<char> top(nn);
  for (i64 i = 0; i < nn; ++i) top[i] = nodes[i].depth < d_top;

  // numbering: lower subtrees in post-order, then the top nodes
  V64 post;
  post.reserve(nn);
  {
    std::vector<std::pair<i64, bool>> st;
    for (i64 r = 0; r < nn; ++r) {
      if (top[r] || !(nodes[r].parent < 0 || top[nodes[r].parent])) continue;
      st.push_back({r, false});
      while (!st.empty()) {
        auto [i, done] = st.back();
        st.pop_back();
        if (done) {
          post.push_back(i);
          continue;
        }
        st.push_back({i, true});
        for (auto c = nodes[i].children.rbegin(); c != nodes[i].children.rend(); ++c)
          st.push_back({*c, false});
      }
    }
  }
  V64 j0(nn, 0);
  P.perm.assign(n, 0);
  i64 pos = 0;
  i64 nT = 0;
  auto place = [&](i64 i) {
    j0[i] = pos;
    for (i32 v : nodes[i].J) P.perm[pos++] = v;
  };
  for (i64 i : post) place(i);
  for (i64 i = 0; i < nn; ++i)
    if (top[i]) {
      place(i);
      nT += (i64)nodes[i].J.size();
    }
  if (pos != n) {
    dds::set_error("dds_nd_plan_build: the dissection lost unknowns (%lld of %lld)",
                   (long long)pos, (long long)n);
    return DDS_EINVAL;
  }
  V64 iperm(n);
  for (i64 i = 0; i < n; ++i) iperm[P.perm[i]] = i;

  // the permuted pattern Ap = A[perm][:, perm], rows sorted, and each entry's source in A
  P.ap_ptr.assign(n + 1, 0);
  for (i64 i = 0; i < n; ++i) P.ap_ptr[i + 1] = P.ap_ptr[i] + (ptr[P.perm[i] + 1] - ptr[P.perm[i]]);
  const i64 nnz = P.ap_ptr[n];
  P.ap_idx.assign(nnz, 0);
  P.ap_src.assign(nnz, 0);
  {
    std::vector<std::pair<i64, i64>> row;
    for (i64 i = 0; i < n; ++i) {
      const i64 o = P.perm[i];
      row.clear();
      for (i64 k = ptr[o]; k < ptr[o + 1]; ++k) row.push_back({iperm[idx[k]], k});
      std::sort(row.begin(), row.end());
      i64 q = P.ap_ptr[i];
      for (auto& e : row) {
        P.ap_idx[q] = (i32)e.first;
        P.ap_src[q] = e.second;
        ++q;
      }
    }
  }

  // fronts: lower nodes (post-order), then the top front (all top nodes merged)
  const i64 F = (i64)post.size() + (nT ? 1 : 0);
  V64 fj0(F, 0), fnj(F, 0);
  std::vector<V64> S(F), children(F);
  V64 front_of(nn, -1);
  for (i64 t = 0; t < (i64)post.size(); ++t) front_of[post[t]] = t;
  if (nT)
    for (i64 i = 0; i < nn; ++i)
      if (top[i]) front_of[i] = F - 1;
  for (i64 t = 0; t < (i64)post.size(); ++t) {
    const i64 i = post[t];
    fj0[t] = j0[i];
    fnj[t] = (i64)nodes[i].J.size();
    for (i64 c : nodes[i].children) children[t].push_back(front_of[c]);
    if (nodes[i].parent >= 0 && top[nodes[i].parent]) children[F - 1].push_back(t);
  }
  if (nT) {
    fj0[F - 1] = n - nT;
    fnj[F - 1] = nT;
  }
  {
    std::vector<i64> mark(n, -1);
    for (i64 t = 0; t < (i64)post.size(); ++t) {
      const i64 a = fj0[t], e = a + fnj[t];
      V64& s = S[t];
      auto take = [&](i64 v) {
        if (v >= e && mark[v] != t) {
          mark[v] = t;
          s.push_back(v);
        }
      };
      for (i64 k = P.ap_ptr[a]; k < P.ap_ptr[e]; ++k) take(P.ap_idx[k]);
      for (i64 c : children[t])
        for (i64 v : S[c]) take(v);
      std::sort(s.begin(), s.end());
    }
  }
  // heights: leaves 0, parents above all their children (children precede parents)
  V64 height(F, 0);
  for (i64 t = 0; t < F; ++t)
    for (i64 c : children[t]) height[t] = std::max(height[t], height[c] + 1);
  V64 order(F);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](i64 a, i64 b) { return height[a] < height[b]; });
  V64 rank(F);
  for (i64 i = 0; i < F; ++i) rank[order[i]] = i;
  P.fj0.resize(F);
  P.fnj.resize(F);
  P.fns.resize(F);
  std::vector<V64> So(F), Co(F);
  V64 hgt(F);
  for (i64 i = 0; i < F; ++i) {
    const i64 o = order[i];
    P.fj0[i] = fj0[o];
    P.fnj[i] = fnj[o];
    P.fns[i] = (i64)S[o].size();
    So[i] = std::move(S[o]);
    for (i64 c : children[o]) Co[i].push_back(rank[c]);
    hgt[i] = height[o];
  }
  const i64 n_levels = F ? hgt[F - 1] + 1 : 0;
  P.level_ptr.assign(n_levels + 1, 0);
  for (i64 i = 0; i < F; ++i) ++P.level_ptr[hgt[i] + 1];
  for (i64 h = 0; h < n_levels; ++h) P.level_ptr[h + 1] += P.level_ptr[h];
  const i64 n_lower = F - (nT ? 1 : 0);
  for (i64 t = 0; t < n_lower; ++t) {
    if (P.fnj[t] > kLdsCap || P.fns[t] > kLdsCap) {
      dds::set_error("a front exceeds %lld unknowns (nJ %lld, nS %lld); lower leaf_size or "
                     "raise top_max", (long long)kLdsCap, (long long)P.fnj[t], (long long)P.fns[t]);
      return DDS_EINVAL;
    }
  }

  V64 m(F);
  for (i64 t = 0; t < F; ++t) m[t] = P.fnj[t] + P.fns[t];
  P.foff.assign(F + 1, 0);
  for (i64 t = 0; t < F; ++t) P.foff[t + 1] = P.foff[t] + m[t] * P.fnj[t];
  P.wtoff.assign(F + 1, P.foff[F]);
  for (i64 t = 0; t < F; ++t) P.wtoff[t + 1] = P.wtoff[t] + P.fns[t] * P.fnj[t];
  // child maps: positions of each child's coupling set in its parent's front
  std::vector<V64> cmap(F);
  P.parent.assign(F, -1);
  V64 rank_in_parent(F, 0);
  for (i64 t = 0; t < F; ++t) {
    const i64 a = P.fj0[t], nj = P.fnj[t];
    for (size_t q = 0; q < Co[t].size(); ++q) {
      const i64 c = Co[t][q];
      V64& p = cmap[c];
      p.resize(So[c].size());
      for (size_t i = 0; i < So[c].size(); ++i) {
        const i64 v = So[c][i];
        if (v < a) {
          dds::set_error("dds_nd_plan_build: coupling set below its parent front");
          return DDS_EINVAL;
        }
        if (v < a + nj) {
          p[i] = v - a;
        } else {
          const i64 k = lower_bound_idx(So[t].data(), (i64)So[t].size(), v);
          if (k >= (i64)So[t].size() || So[t][k] != v) {
            dds::set_error("dds_nd_plan_build: coupling set not nested");
            return DDS_EINVAL;
          }
          p[i] = nj + k;
        }
      }
      P.parent[c] = t;
      rank_in_parent[c] = (i64)q;
    }
  }
  P.child_ptr.assign(F + 1, 0);
  for (i64 t = 0; t < F; ++t) P.child_ptr[t + 1] = P.child_ptr[t] + (i64)Co[t].size();
  P.child_idx.clear();
  for (i64 t = 0; t < F; ++t) P.child_idx.insert(P.child_idx.end(), Co[t].begin(), Co[t].end());
  // forward: child c adds its update vector into slot rank(c) of its parent's front
  P.soff.assign(F + 1, 0);
  for (i64 t = 0; t < F; ++t) P.soff[t + 1] = P.soff[t] + (i64)Co[t].size() * m[t];
  P.wbase.assign(F, -1);
  for (i64 c = 0; c < F; ++c)
    if (P.parent[c] >= 0) P.wbase[c] = P.soff[P.parent[c]] + rank_in_parent[c] * m[P.parent[c]];
  P.moff.assign(F + 1, 0);
  for (i64 t = 0; t < F; ++t) P.moff[t + 1] = P.moff[t] + P.fns[t];
  P.maps.assign(P.moff[F], 0);
  P.s_idx.assign(P.moff[F], 0);
  for (i64 c = 0; c < F; ++c) {
    if ((i64)cmap[c].size() != P.fns[c]) {
      dds::set_error("dds_nd_plan_build: a front couples to unknowns without a parent");
      return DDS_EINVAL;
    }
    for (size_t i = 0; i < cmap[c].size(); ++i) P.maps[P.moff[c] + i] = (i32)cmap[c][i];
    for (size_t i = 0; i < So[c].size(); ++i) P.s_idx[P.moff[c] + i] = So[c][i];
  }
  // backward: x at a front's coupling set is pushed into its buffer xs[moff:+nS] by its
  // parent; the parent's push list sorted by source position (stable)
  P.push_ptr.assign(F + 1, 0);
  P.push_dst.clear();
  P.push_src.clear();
  {
    std::vector<std::pair<i64, i64>> pr;  // (src, dst)
    for (i64 t = 0; t < F; ++t) {
      pr.clear();
      for (i64 c : Co[t])
        for (i64 i = 0; i < P.fns[c]; ++i) pr.push_back({cmap[c][i], P.moff[c] + i});
      std::stable_sort(pr.begin(), pr.end(),
                       [](const std::pair<i64, i64>& x, const std::pair<i64, i64>& y) {
                         return x.first < y.first;
                       });
      for (auto& e : pr) {
        P.push_dst.push_back(e.second);
        P.push_src.push_back((i32)e.first);
      }
      P.push_ptr[t + 1] = (i64)P.push_dst.size();
    }
  }
  P.uoff2.assign(F + 1, 0);
  for (i64 t = 0; t < F; ++t) P.uoff2[t + 1] = P.uoff2[t] + P.fns[t] * P.fns[t];

  // work items of the lower heights (the top front is solved densely)
  const i64 nlev = n_levels - (nT ? 1 : 0);
  P.fwd_ptr.assign(nlev + 1, 0);
  P.bwd_ptr.assign(nlev + 1, 0);
  P.fwd_lds.assign(nlev, 0);
  P.bwd_lds.assign(nlev, 0);
  P.fwd_items.clear();
  P.bwd_items.clear();
  for (i64 h = 0; h < nlev; ++h) {
    const i64 lo = P.level_ptr[h], hi = P.level_ptr[h + 1];
    i64 whole = 0;
    for (i64 t = lo; t < hi; ++t) whole += (m[t] + 255) / 256;
    const bool split = whole < 4096;
    i64 lf = 0, lb = 0;
    for (i64 t = lo; t < hi; ++t) {
      const i64 j = P.fj0[t], nj = P.fnj[t], ns = P.fns[t], mt = m[t];
      lf = std::max(lf, nj);
      lb = std::max(lb, ns);
      const i64 nch = (i64)Co[t].size();
      const i64 R = !split ? 256 : nj >= 32 ? 64 : nj >= 12 ? 128 : 256;
      for (i64 r0 = 0; r0 < mt; r0 += R) {
        const i64 rec[12] = {j, nj, ns, P.foff[t], P.soff[t], nch, P.wbase[t], P.moff[t],
                             r0, std::min(R, mt - r0), R, 0};
        P.fwd_items.insert(P.fwd_items.end(), rec, rec + 12);
      }
      const i64 k_lo = P.push_ptr[t], k_hi = P.push_ptr[t + 1];
      const i32* src = P.push_src.data() + k_lo;
      const i64 ns_push = k_hi - k_lo;
      i64 C = nj > 32 ? 64 : nj > 16 ? 32 : 16;
      if (split) C = std::min<i64>(C, ns <= 128 ? 64 : ns <= 256 ? 32 : 16);
      for (i64 c0 = 0; c0 < nj; c0 += C) {
        const i64 nc = std::min(C, nj - c0);
        const i64 k0 = k_lo + (i64)(std::lower_bound(src, src + ns_push, (i32)c0) - src);
        const i64 k1 = c0 + nc == nj
                           ? k_hi
                           : k_lo + (i64)(std::lower_bound(src, src + ns_push, (i32)(c0 + nc)) - src);
        const i64 rec[12] = {j, nj, ns, P.wtoff[t], P.moff[t], c0, nc, k0, k1, C, 0, 0};
        P.bwd_items.insert(P.bwd_items.end(), rec, rec + 12);
      }
    }
    P.fwd_ptr[h + 1] = (i64)P.fwd_items.size() / 12;
    P.bwd_ptr[h + 1] = (i64)P.bwd_items.size() / 12;
    P.fwd_lds[h] = lf;
    P.bwd_lds[h] = lb;
  }

  i64* s = P.scalars;
  s[DDS_ND_N] = n;
  s[DDS_ND_FRONTS] = F;
  s[DDS_ND_TOP] = nT;
  s[DDS_ND_DTOP] = d_top;
  s[DDS_ND_LEVELS] = n_levels;
  s[DDS_ND_LOWER_LEVELS] = nlev;
  s[DDS_ND_FAC_SIZE] = P.wtoff[F];
  s[DDS_ND_SLOT_SIZE] = P.soff[F];
  s[DDS_ND_XS_SIZE] = P.moff[F];
  s[DDS_ND_NNZ] = nnz;
  s[DDS_ND_TREE_NODES] = nn;
  return DDS_OK;
}

}  // namespace

extern "C" int dds_nd_plan_build(int64_t n, const int64_t* indptr, const int32_t* indices,
                                 int64_t leaf_size, int64_t top_max, dds_nd_plan** plan) {
  if (plan == nullptr || n < 1 || indptr == nullptr || leaf_size < 1 || top_max < 0) {
    dds::set_error("dds_nd_plan_build: bad arguments (n %lld, leaf_size %lld, top_max %lld)",
                   (long long)n, (long long)leaf_size, (long long)top_max);
    return DDS_EINVAL;
  }
  *plan = nullptr;
  if (n > INT32_MAX) {
    dds::set_error("dds_nd_plan_build: n = %lld exceeds the int32 index range", (long long)n);
    return DDS_EINVAL;
  }
  if (indptr[0] != 0 || (indptr[n] > 0 && indices == nullptr)) {
    dds::set_error("dds_nd_plan_build: bad row pointers");
    return DDS_EINVAL;
  }
  for (int64_t i = 0; i < n; ++i) {
    if (indptr[i + 1] < indptr[i]) {
      dds::set_error("dds_nd_plan_build: row pointers decrease at row %lld", (long long)i);
      return DDS_EINVAL;
    }
    for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) {
      if (indices[k] < 0 || indices[k] >= n || (k > indptr[i] && indices[k] <= indices[k - 1])) {
        dds::set_error("dds_nd_plan_build: row %lld: column indices must be sorted, unique and "
                       "in [0, n)", (long long)i);
        return DDS_EINVAL;
      }
    }
  }
  dds_nd_plan* P = new dds_nd_plan();
  int rc;
  try {
    rc = build_plan(*P, n, indptr, indices, leaf_size, top_max);
  } catch (const std::bad_alloc&) {
    dds::set_error("dds_nd_plan_build: out of host memory");
    rc = DDS_EINVAL;
  }
  if (rc != DDS_OK) {
    delete P;
    return rc;
  }
  *plan = P;
  return DDS_OK;
}

namespace {
template <typename T>
int64_t copy_out(const std::vector<T>& v, void* dst) {
  if (dst != nullptr && !v.empty()) std::copy(v.begin(), v.end(), static_cast<T*>(dst));
  return (int64_t)v.size();
}
}  // namespace

extern "C" int64_t dds_nd_plan_field(const dds_nd_plan* P, int field, void* dst) {
  if (P == nullptr) return -1;
  switch (field) {
    case DDS_ND_SCALARS:
      if (dst) std::copy(P->scalars, P->scalars + DDS_ND_N_SCALARS, static_cast<int64_t*>(dst));
      return DDS_ND_N_SCALARS;
    case DDS_ND_PERM: return copy_out(P->perm, dst);
    case DDS_ND_FJ0: return copy_out(P->fj0, dst);
    case DDS_ND_FNJ: return copy_out(P->fnj, dst);
    case DDS_ND_FNS: return copy_out(P->fns, dst);
    case DDS_ND_LEVEL_PTR: return copy_out(P->level_ptr, dst);
    case DDS_ND_FOFF: return copy_out(P->foff, dst);
    case DDS_ND_WTOFF: return copy_out(P->wtoff, dst);
    case DDS_ND_SOFF: return copy_out(P->soff, dst);
    case DDS_ND_WBASE: return copy_out(P->wbase, dst);
    case DDS_ND_MOFF: return copy_out(P->moff, dst);
    case DDS_ND_MAPS: return copy_out(P->maps, dst);
    case DDS_ND_PUSH_DST: return copy_out(P->push_dst, dst);
    case DDS_ND_PUSH_SRC: return copy_out(P->push_src, dst);
    case DDS_ND_PUSH_PTR: return copy_out(P->push_ptr, dst);
    case DDS_ND_PARENT: return copy_out(P->parent, dst);
    case DDS_ND_CHILD_PTR: return copy_out(P->child_ptr, dst);
    case DDS_ND_CHILD_IDX: return copy_out(P->child_idx, dst);
    case DDS_ND_S_IDX: return copy_out(P->s_idx, dst);
    case DDS_ND_AP_PTR: return copy_out(P->ap_ptr, dst);
    case DDS_ND_AP_IDX: return copy_out(P->ap_idx, dst);
    case DDS_ND_AP_SRC: return copy_out(P->ap_src, dst);
    case DDS_ND_FWD_PTR: return copy_out(P->fwd_ptr, dst);
    case DDS_ND_FWD_ITEMS: return copy_out(P->fwd_items, dst);
    case DDS_ND_FWD_LDS: return copy_out(P->fwd_lds, dst);
    case DDS_ND_BWD_PTR: return copy_out(P->bwd_ptr, dst);
    case DDS_ND_BWD_ITEMS: return copy_out(P->bwd_items, dst);
    case DDS_ND_BWD_LDS: return copy_out(P->bwd_lds, dst);
    case DDS_ND_UOFF2: return copy_out(P->uoff2, dst);
    default:
      dds::set_error("dds_nd_plan_field: unknown field %d", field);
      return -1;
  }
}

extern "C" void dds_nd_plan_free(dds_nd_plan* P) { delete P; }
