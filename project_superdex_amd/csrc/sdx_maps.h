// Host-side maps of the assembly: the BSR pattern of an actor, the stable
// counting sort that orders records by destination, and the element patches
// with their descriptions.  Plain C++ (no HIP), so host tests can build them.
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "sdx_layout.h"

namespace sdx {

struct Pattern {
  std::vector<int32_t> ptr, col;
};

// BSR pattern of one actor: node-to-node adjacency through tets, sorted.
inline Pattern BuildPattern(int nV, const std::vector<int32_t>& tets) {
  const int nT = int(tets.size() / 4);
  std::vector<std::vector<int32_t>> adj(static_cast<size_t>(nV));
  for (int i = 0; i < nV; ++i) adj[size_t(i)].push_back(i);
  for (int t = 0; t < nT; ++t)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) adj[size_t(tets[4 * t + a])].push_back(tets[4 * t + b]);
  Pattern p;
  p.ptr.assign(size_t(nV) + 1, 0);
  for (int i = 0; i < nV; ++i) {
    auto& v = adj[size_t(i)];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    p.ptr[size_t(i) + 1] = p.ptr[size_t(i)] + int32_t(v.size());
  }
  p.col.reserve(size_t(p.ptr.back()));
  for (int i = 0; i < nV; ++i) p.col.insert(p.col.end(), adj[size_t(i)].begin(), adj[size_t(i)].end());
  return p;
}

inline int FindSlot(const Pattern& p, int i, int j) {
  auto b = p.col.begin() + p.ptr[size_t(i)], e = p.col.begin() + p.ptr[size_t(i) + 1];
  auto it = std::lower_bound(b, e, j);
  if (it == e || *it != j) throw std::runtime_error("BSR slot not in pattern");
  return int(it - p.col.begin());
}

// Stable counting sort of (key, code) pairs into CSR form.
inline void CountingCsr(int nKeys, const std::vector<int32_t>& keys, const std::vector<uint32_t>& codes,
                 std::vector<int32_t>& ptr, std::vector<uint32_t>& out) {
  ptr.assign(size_t(nKeys) + 1, 0);
  for (int32_t k : keys) ptr[size_t(k) + 1]++;
  for (int i = 0; i < nKeys; ++i) ptr[size_t(i) + 1] += ptr[size_t(i)];
  out.resize(codes.size());
  std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
  for (size_t i = 0; i < keys.size(); ++i) out[size_t(pos[size_t(keys[i])]++)] = codes[i];
}

// Element patches (their layout: the description of a patch in sdx_layout.h).
// add() takes the tets of one patch (actor-local vertex indices; the actor's
// first vertex v0, first BSR slot b0 and pattern P) and emits the record keys
// and codes of its local vertices (code: running local-vertex index) and
// local pairs (code: running local-pair index); finish() writes the
// descriptions once the records' positions are known.
class PatchBuilder {
 public:
  int count() const { return int(t0_.size()); }
  size_t n_local_vertices() const { return lv_total_; }
  size_t n_local_pairs() const { return le_total_; }

  void add(const Pattern& P, int b0, int v0, int gt0, const int32_t* tv, int nt, std::vector<int32_t>& vtx_keys,
           std::vector<uint32_t>& vtx_codes, std::vector<int32_t>& edge_keys, std::vector<uint32_t>& edge_codes) {
    vid_.resize(size_t(4 * nt));
    for (int k = 0; k < 4 * nt; ++k) vid_[size_t(k)] = v0 + tv[k];
    lv_.assign(vid_.begin(), vid_.end());
    std::sort(lv_.begin(), lv_.end());
    lv_.erase(std::unique(lv_.begin(), lv_.end()), lv_.end());
    slot_.resize(size_t(6 * nt));
    tr_.resize(size_t(6 * nt));
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < 6; ++k) {
        const int va = tv[4 * t + kPairA[k]], vb = tv[4 * t + kPairB[k]];
        slot_[size_t(6 * t + k)] = b0 + FindSlot(P, std::min(va, vb), std::max(va, vb));
        tr_[size_t(6 * t + k)] = va > vb ? 1 : 0;
      }
    le_.assign(slot_.begin(), slot_.end());
    std::sort(le_.begin(), le_.end());
    le_.erase(std::unique(le_.begin(), le_.end()), le_.end());
    const int nLV = int(lv_.size()), nLE = int(le_.size()), nS = nLV + nLE;
    if (nLV > 1023 || nLE > 1023 || nt > kPatchTets) throw std::logic_error("element patch too large");
    // lists: counting sort of the contributions by local slot, stable in
    // (tet, a) / (tet, k) order
    std::vector<int32_t> cnt(size_t(nS) + 1, 0), loc(size_t(10 * nt));
    for (int k = 0; k < 4 * nt; ++k) {
      loc[size_t(k)] = int(std::lower_bound(lv_.begin(), lv_.end(), vid_[size_t(k)]) - lv_.begin());
      cnt[size_t(loc[size_t(k)]) + 1]++;
    }
    for (int k = 0; k < 6 * nt; ++k) {
      loc[size_t(4 * nt + k)] =
          nLV + int(std::lower_bound(le_.begin(), le_.end(), slot_[size_t(k)]) - le_.begin());
      cnt[size_t(loc[size_t(4 * nt + k)]) + 1]++;
    }
    for (int j = 0; j < nS; ++j) cnt[size_t(j) + 1] += cnt[size_t(j)];
    const size_t l0 = lists_.size();
    for (int j = 0; j <= nS; ++j) lists_.push_back(uint16_t(cnt[size_t(j)]));
    const size_t e0 = lists_.size();
    lists_.resize(e0 + size_t(10 * nt));
    for (int k = 0; k < 10 * nt; ++k) {
      const uint16_t v = k < 4 * nt ? uint16_t(k)
                                    : uint16_t((k - 4 * nt) | (tr_[size_t(k - 4 * nt)] ? 0x8000 : 0));
      lists_[e0 + size_t(cnt[size_t(loc[size_t(k)])]++)] = v;
    }
    for (int j = 0; j < nLV; ++j) {
      vtx_keys.push_back(lv_[size_t(j)]);
      vtx_codes.push_back(uint32_t(lv_total_ + size_t(j)));
    }
    for (int j = 0; j < nLE; ++j) {
      edge_keys.push_back(le_[size_t(j)]);
      edge_codes.push_back(uint32_t(le_total_ + size_t(j)));
    }
    t0_.push_back(gt0);
    nt_.push_back(nt);
    n_lv_.push_back(nLV);
    n_le_.push_back(nLE);
    list0_.push_back(l0);
    lv_total_ += size_t(nLV);
    le_total_ += size_t(nLE);
  }

  // Descriptions of every patch, in order: patch p occupies words
  // [meta_ptr[p], meta_ptr[p + 1]) of meta, its pair part starts at word
  // meta_mid[p].
  void finish(const std::vector<int32_t>& pos_lv, const std::vector<int32_t>& pos_le, std::vector<int32_t>& t0,
              std::vector<int32_t>& meta_ptr, std::vector<int32_t>& meta_mid, std::vector<uint32_t>& meta) const {
    const int n = count();
    t0.assign(t0_.begin(), t0_.end());
    t0.push_back(n ? t0_.back() + nt_.back() : 0);
    meta_ptr.assign(1, 0);
    meta_mid.clear();
    meta.clear();
    size_t lvo = 0, leo = 0;
    std::vector<uint16_t> h;
    auto put16 = [&]() {
      for (size_t k = 0; k < h.size(); k += 2)
        meta.push_back(uint32_t(h[k]) | (k + 1 < h.size() ? uint32_t(h[k + 1]) << 16 : 0u));
    };
    auto pad4 = [&](size_t w0) {
      while ((meta.size() - w0) % 4) meta.push_back(0);
    };
    for (int p = 0; p < n; ++p) {
      const int nt = nt_[size_t(p)], nLV = n_lv_[size_t(p)], nLE = n_le_[size_t(p)], nS = nLV + nLE;
      const uint16_t* starts = &lists_[list0_[size_t(p)]];  // nS + 1 starts into the entries
      const uint16_t* ent = starts + nS + 1;                  // 4 nt vertex entries, then 6 nt pair entries
      const size_t w0 = meta.size();
      // vertex part: header, records, starts and entries
      meta.push_back(uint32_t(nLV) | uint32_t(nLE) << 10 | uint32_t(nt) << 20);
      for (int j = 0; j < nLV; ++j) meta.push_back(uint32_t(pos_lv[lvo + size_t(j)]));
      h.assign(starts, starts + nLV + 1);
      h.insert(h.end(), ent, ent + 4 * nt);
      put16();
      pad4(w0);
      if (meta.size() - w0 != size_t(PatchVertexWords(nLV, nt))) throw std::logic_error("patch vertex part size");
      meta_mid.push_back(int32_t(meta.size()));
      // pair part: records, starts (from the first pair entry) and entries
      const size_t w1 = meta.size();
      for (int j = 0; j < nLE; ++j) meta.push_back(uint32_t(pos_le[leo + size_t(j)]));
      h.clear();
      for (int j = 0; j <= nLE; ++j) h.push_back(uint16_t(starts[nLV + j] - 4 * nt));
      h.insert(h.end(), ent + 4 * nt, ent + 10 * nt);
      put16();
      pad4(w1);
      if (meta.size() - w0 > size_t(kPatchMetaWords)) throw std::logic_error("element patch description too long");
      if (meta.size() > size_t(INT32_MAX)) throw std::invalid_argument("mesh too large for its patch descriptions");
      meta_ptr.push_back(int32_t(meta.size()));
      lvo += size_t(nLV);
      leo += size_t(nLE);
    }
  }

 private:
  std::vector<int32_t> vid_, lv_, slot_, le_;
  std::vector<uint8_t> tr_;
  std::vector<int32_t> t0_, nt_, n_lv_, n_le_;
  std::vector<size_t> list0_;
  std::vector<uint16_t> lists_;
  size_t lv_total_ = 0, le_total_ = 0;
};

}  // namespace sdx
