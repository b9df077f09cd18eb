// Row blocks of the fused assembly (k_assemble_rows).  Each actor's vertices
// are split by recursive coordinate bisection of their rest positions into
// blocks of at most kBlockRows vertices whose BSR rows hold at most a given
// number of slots (kBlockSlots by default).  A workgroup owns one block: it
// computes every tet and boundary triangle that touches one of its rows (a
// tet shared by blocks is computed by each of them), sums their
// contributions to its rows in LDS and writes the rows' BSR values,
// residuals and preconditioner blocks, so no per-element records pass
// through memory.  A block lists its rows (ascending vertex ids) with the
// first LDS slot of each, and its tets and triangles with the block-local
// row of each of their vertices (kNotInBlock for a vertex of another block).
// Plain C++ (no HIP), so host tests can build it.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace sdx {

constexpr int kBlockRows = 64;
constexpr int kBlockSlots = 1024;
constexpr uint32_t kNotInBlock = 0xFFu;

struct RowBlocks {
  explicit RowBlocks(int slot_cap = kBlockSlots) : slot_cap_(slot_cap) {}
  // per block, with one more entry at the end for the starts
  std::vector<int32_t> row0{0}, tet0{0}, tri0{0};
  std::vector<int32_t> env;
  // per block row: global vertex, first LDS slot within the block
  std::vector<int32_t> rows, rbase;
  // per block tet / triangle: global index, local rows (8 bits per vertex)
  std::vector<int32_t> tets, tet_rows, tris, tri_rows;
  int max_slots = 0;

  int count() const { return int(env.size()); }

  // Append the blocks of one actor: rest positions X [3 nV], tets [4 nT] and
  // boundary triangles [3 nF] with actor-local vertices, the BSR row lengths
  // [nV]; v0, t0, f0: the actor's first global vertex, tet and triangle.
  void add(int env_id, const std::vector<double>& X, int nV, const std::vector<int32_t>& atets,
           const std::vector<int32_t>& atris, const std::vector<int32_t>& row_len, int v0, int t0, int f0) {
    if (nV <= 0) return;
    std::vector<int32_t> idx(static_cast<size_t>(nV));
    for (int i = 0; i < nV; ++i) idx[size_t(i)] = i;
    std::vector<std::pair<int, int>> parts;
    split(idx, 0, nV, X, row_len, slot_cap_, parts);
    const std::vector<int32_t> vt_ptr = incidence(nV, atets, 4, vt_);
    const std::vector<int32_t> vf_ptr = incidence(nV, atris, 3, vf_);
    std::vector<int32_t> local(static_cast<size_t>(nV), -1);
    std::vector<int32_t> stamp_t(atets.size() / 4, -1), stamp_f(atris.size() / 3, -1);
    for (size_t p = 0; p < parts.size(); ++p) {
      const int lo = parts[p].first, hi = parts[p].second;
      std::vector<int32_t> br(idx.begin() + lo, idx.begin() + hi);
      std::sort(br.begin(), br.end());
      int slots = 0;
      for (size_t i = 0; i < br.size(); ++i) {
        local[size_t(br[i])] = int32_t(i);
        rows.push_back(v0 + br[i]);
        rbase.push_back(slots);
        slots += row_len[size_t(br[i])];
      }
      max_slots = std::max(max_slots, slots);
      // the tets and triangles touching the block, ascending
      std::vector<int32_t> bt, bf;
      for (int v : br) {
        for (int k = vt_ptr[size_t(v)]; k < vt_ptr[size_t(v) + 1]; ++k) {
          const int t = vt_[size_t(k)];
          if (stamp_t[size_t(t)] != int32_t(p)) {
            stamp_t[size_t(t)] = int32_t(p);
            bt.push_back(t);
          }
        }
        for (int k = vf_ptr[size_t(v)]; k < vf_ptr[size_t(v) + 1]; ++k) {
          const int f = vf_[size_t(k)];
          if (stamp_f[size_t(f)] != int32_t(p)) {
            stamp_f[size_t(f)] = int32_t(p);
            bf.push_back(f);
          }
        }
      }
      std::sort(bt.begin(), bt.end());
      std::sort(bf.begin(), bf.end());
      for (int t : bt) {
        tets.push_back(t0 + t);
        tet_rows.push_back(int32_t(pack(local, &atets[size_t(4 * t)], 4)));
      }
      for (int f : bf) {
        tris.push_back(f0 + f);
        tri_rows.push_back(int32_t(pack(local, &atris[size_t(3 * f)], 3)));
      }
      for (int v : br) local[size_t(v)] = -1;
      env.push_back(env_id);
      row0.push_back(int32_t(rows.size()));
      tet0.push_back(int32_t(tets.size()));
      tri0.push_back(int32_t(tris.size()));
      if (tets.size() > size_t(INT32_MAX) || tris.size() > size_t(INT32_MAX))
        throw std::invalid_argument("mesh too large for its row blocks");
    }
  }

 private:
  // Recursive coordinate bisection of idx[lo, hi): split along the longest
  // extent into parts proportional to the numbers of blocks each side needs.
  static void split(std::vector<int32_t>& idx, int lo, int hi, const std::vector<double>& X,
                    const std::vector<int32_t>& row_len, int cap, std::vector<std::pair<int, int>>& out) {
    const int n = hi - lo;
    int64_t slots = 0;
    for (int i = lo; i < hi; ++i) slots += row_len[size_t(idx[size_t(i)])];
    if (n <= kBlockRows && slots <= cap) {
      out.emplace_back(lo, hi);
      return;
    }
    if (n == 1) throw std::invalid_argument("a vertex has more BSR blocks than a row block holds");
    const int64_t nb = std::max<int64_t>(
        2, std::max<int64_t>((n + kBlockRows - 1) / kBlockRows, (slots + cap - 1) / cap));
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = lo; i < hi; ++i)
      for (int d = 0; d < 3; ++d) {
        const double x = X[size_t(3 * idx[size_t(i)] + d)];
        mn[d] = std::min(mn[d], x);
        mx[d] = std::max(mx[d], x);
      }
    int axis = 0;
    for (int d = 1; d < 3; ++d)
      if (mx[d] - mn[d] > mx[axis] - mn[axis]) axis = d;
    const int64_t left_nb = nb / 2;
    int k = lo + int((int64_t(n) * left_nb + nb / 2) / nb);
    k = std::min(std::max(k, lo + 1), hi - 1);
    auto key = [&](int32_t v) { return X[size_t(3 * v + axis)]; };
    std::nth_element(idx.begin() + lo, idx.begin() + k, idx.begin() + hi, [&](int32_t a, int32_t b) {
      return key(a) < key(b) || (key(a) == key(b) && a < b);
    });
    split(idx, lo, k, X, row_len, cap, out);
    split(idx, k, hi, X, row_len, cap, out);
  }

  // Vertex -> element incidence (CSR) of elements with `w` vertices each.
  static std::vector<int32_t> incidence(int nV, const std::vector<int32_t>& el, int w, std::vector<int32_t>& list) {
    std::vector<int32_t> ptr(size_t(nV) + 1, 0);
    for (int32_t v : el) ptr[size_t(v) + 1]++;
    for (int i = 0; i < nV; ++i) ptr[size_t(i) + 1] += ptr[size_t(i)];
    list.assign(el.size(), 0);
    std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
    for (size_t k = 0; k < el.size(); ++k) list[size_t(pos[size_t(el[k])]++)] = int32_t(k / size_t(w));
    return ptr;
  }

  static uint32_t pack(const std::vector<int32_t>& local, const int32_t* vs, int n) {
    uint32_t r = 0;
    for (int a = 0; a < n; ++a) {
      const int32_t l = local[size_t(vs[a])];
      r |= uint32_t(l < 0 ? kNotInBlock : uint32_t(l)) << (8 * a);
    }
    return r;
  }

  int slot_cap_;
  std::vector<int32_t> vt_, vf_;
};

}  // namespace sdx
