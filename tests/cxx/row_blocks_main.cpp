// Host check of the row blocks of the fused assembly (sdx_rowblocks.h)
// against brute force.
//
// Input (argv[1]): meshes, each "nV nT nF" followed by 3 nV rest coordinates,
// 4 nT tet vertices and 3 nF boundary-triangle vertices.  All meshes are
// added to one RowBlocks side by side (global vertex, tet and triangle
// offsets, as a world lays out its actors), with the slot cap of argv[2].
// Checks: every vertex lies in exactly one block of its actor; a block has at
// most kBlockRows rows and at most `cap` slots, rows ascending with LDS bases
// the prefix sums of their row lengths; a block's tets and triangles are
// exactly those touching one of its rows, ascending, with the right local
// row (or kNotInBlock) for every vertex; and integer-valued contributions of
// every (tet, a, b) summed block by block into the block rows' slots, the
// way the fused kernel does (own rows only, column found in the row), equal
// a direct accumulation into the BSR pattern exactly.  Prints "OK", the
// block count and the tet redundancy (block tets / tets), or the first
// mismatch (exit 1).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "sdx_maps.h"
#include "sdx_rowblocks.h"

using namespace sdx;

namespace {

int fail(const char* what, long a, long b) {
  std::printf("FAIL %s (%ld, %ld)\n", what, a, b);
  return 1;
}

struct Mesh {
  int nV = 0, nT = 0, nF = 0;
  std::vector<double> X;
  std::vector<int32_t> tets, tris;
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return fail("usage", 0, 0);
  const int cap = std::atoi(argv[2]);
  std::ifstream in(argv[1]);
  std::vector<Mesh> meshes;
  Mesh m;
  while (in >> m.nV >> m.nT >> m.nF) {
    m.X.resize(size_t(3) * size_t(m.nV));
    m.tets.resize(size_t(4) * size_t(m.nT));
    m.tris.resize(size_t(3) * size_t(m.nF));
    for (auto& x : m.X) in >> x;
    for (auto& x : m.tets) in >> x;
    for (auto& x : m.tris) in >> x;
    meshes.push_back(m);
  }
  if (meshes.empty()) return fail("no meshes", 0, 0);
  RowBlocks rb(cap);
  std::vector<int> v0s, t0s, f0s;
  std::vector<Pattern> pats;
  int v0 = 0, t0 = 0, f0 = 0;
  for (size_t e = 0; e < meshes.size(); ++e) {
    const Mesh& a = meshes[e];
    Pattern P = BuildPattern(a.nV, a.tets);
    std::vector<int32_t> len(static_cast<size_t>(a.nV));
    for (int i = 0; i < a.nV; ++i) len[size_t(i)] = P.ptr[size_t(i) + 1] - P.ptr[size_t(i)];
    rb.add(int(e), a.X, a.nV, a.tets, a.tris, len, v0, t0, f0);
    v0s.push_back(v0);
    t0s.push_back(t0);
    f0s.push_back(f0);
    pats.push_back(P);
    v0 += a.nV;
    t0 += a.nT;
    f0 += a.nF;
  }
  const int nV = v0;
  // every vertex in exactly one block of its own actor
  std::vector<int> blk_of(static_cast<size_t>(nV), -1), local(static_cast<size_t>(nV), -1);
  long redundant = 0;
  for (int b = 0; b < rb.count(); ++b) {
    const int e = rb.env[size_t(b)];
    const Mesh& a = meshes[size_t(e)];
    const Pattern& P = pats[size_t(e)];
    const int r0 = rb.row0[size_t(b)], r1 = rb.row0[size_t(b) + 1];
    if (r1 <= r0 || r1 - r0 > kBlockRows) return fail("block rows", b, r1 - r0);
    int slots = 0;
    for (int r = r0; r < r1; ++r) {
      const int v = rb.rows[size_t(r)];
      if (v < v0s[size_t(e)] || v >= v0s[size_t(e)] + a.nV) return fail("row outside its actor", b, v);
      if (r > r0 && v <= rb.rows[size_t(r) - 1]) return fail("rows not ascending", b, v);
      if (blk_of[size_t(v)] != -1) return fail("vertex in two blocks", v, b);
      blk_of[size_t(v)] = b;
      local[size_t(v)] = r - r0;
      if (rb.rbase[size_t(r)] != slots) return fail("LDS base", r, rb.rbase[size_t(r)]);
      const int lv = v - v0s[size_t(e)];
      slots += P.ptr[size_t(lv) + 1] - P.ptr[size_t(lv)];
    }
    if (slots > cap || slots > rb.max_slots) return fail("block slots", b, slots);
    // the tets touching the block, ascending, with their local rows
    std::vector<int32_t> want;
    for (int t = 0; t < a.nT; ++t)
      for (int k = 0; k < 4; ++k)
        if (blk_of[size_t(v0s[size_t(e)] + a.tets[size_t(4 * t + k)])] == b) {
          want.push_back(t0s[size_t(e)] + t);
          break;
        }
    const int q0 = rb.tet0[size_t(b)], q1 = rb.tet0[size_t(b) + 1];
    if (int(want.size()) != q1 - q0) return fail("block tet count", b, q1 - q0);
    redundant += q1 - q0;
    for (int q = q0; q < q1; ++q) {
      if (rb.tets[size_t(q)] != want[size_t(q - q0)]) return fail("block tet", b, rb.tets[size_t(q)]);
      const int t = rb.tets[size_t(q)] - t0s[size_t(e)];
      for (int k = 0; k < 4; ++k) {
        const int v = v0s[size_t(e)] + a.tets[size_t(4 * t + k)];
        const uint32_t lr = (uint32_t(rb.tet_rows[size_t(q)]) >> (8 * k)) & kNotInBlock;
        const uint32_t expect = blk_of[size_t(v)] == b ? uint32_t(local[size_t(v)]) : kNotInBlock;
        if (lr != expect) return fail("tet local row", q, long(lr));
      }
    }
    std::vector<int32_t> wantf;
    for (int f = 0; f < a.nF; ++f)
      for (int k = 0; k < 3; ++k)
        if (blk_of[size_t(v0s[size_t(e)] + a.tris[size_t(3 * f + k)])] == b) {
          wantf.push_back(f0s[size_t(e)] + f);
          break;
        }
    const int g0 = rb.tri0[size_t(b)], g1 = rb.tri0[size_t(b) + 1];
    if (int(wantf.size()) != g1 - g0) return fail("block triangle count", b, g1 - g0);
    for (int g = g0; g < g1; ++g) {
      if (rb.tris[size_t(g)] != wantf[size_t(g - g0)]) return fail("block triangle", b, rb.tris[size_t(g)]);
      const int f = rb.tris[size_t(g)] - f0s[size_t(e)];
      for (int k = 0; k < 3; ++k) {
        const int v = v0s[size_t(e)] + a.tris[size_t(3 * f + k)];
        const uint32_t lr = (uint32_t(rb.tri_rows[size_t(g)]) >> (8 * k)) & kNotInBlock;
        const uint32_t expect = blk_of[size_t(v)] == b ? uint32_t(local[size_t(v)]) : kNotInBlock;
        if (lr != expect) return fail("triangle local row", g, long(lr));
      }
    }
  }
  for (int v = 0; v < nV; ++v)
    if (blk_of[size_t(v)] < 0) return fail("vertex in no block", v, 0);
  // integer-valued contributions c(t, a, b) summed per block row slot vs a direct accumulation
  for (size_t e = 0; e < meshes.size(); ++e) {
    const Mesh& a = meshes[e];
    const Pattern& P = pats[e];
    std::vector<double> direct(P.col.size(), 0.0), fused(P.col.size(), 0.0);
    auto contrib = [](int t, int i, int j) { return double((t * 131 + i * 17 + j * 5) % 1009 + 1); };
    for (int t = 0; t < a.nT; ++t)
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          direct[size_t(FindSlot(P, a.tets[size_t(4 * t + i)], a.tets[size_t(4 * t + j)]))] += contrib(t, i, j);
    for (int b = 0; b < rb.count(); ++b) {
      if (rb.env[size_t(b)] != int(e)) continue;
      const int r0 = rb.row0[size_t(b)];
      for (int q = rb.tet0[size_t(b)]; q < rb.tet0[size_t(b) + 1]; ++q) {
        const int t = rb.tets[size_t(q)] - t0s[e];
        for (int i = 0; i < 4; ++i) {
          const uint32_t lr = (uint32_t(rb.tet_rows[size_t(q)]) >> (8 * i)) & kNotInBlock;
          if (lr == kNotInBlock) continue;
          const int row = rb.rows[size_t(r0) + lr] - v0s[e];
          for (int j = 0; j < 4; ++j) {
            // the column's position in the row, as the kernel searches it
            const int c = a.tets[size_t(4 * t + j)];
            int pos = -1;
            for (int k = P.ptr[size_t(row)]; k < P.ptr[size_t(row) + 1]; ++k)
              if (P.col[size_t(k)] == c) pos = k;
            if (pos < 0) return fail("column not in its row", row, c);
            fused[size_t(pos)] += contrib(t, i, j);
          }
        }
      }
    }
    for (size_t k = 0; k < direct.size(); ++k)
      if (direct[k] != fused[k]) return fail("slot sum", long(e), long(k));
  }
  int nT = t0;
  std::printf("OK blocks %d redundancy %.3f max_slots %d\n", rb.count(), double(redundant) / double(nT > 0 ? nT : 1),
              rb.max_slots);
  return 0;
}
