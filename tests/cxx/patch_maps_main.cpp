// Host check of the assembly maps (sdx_maps.h): element patches, their
// descriptions and the record order, against a brute-force assembly.
//
// Input (argv[1]): a text file of meshes, each "nV nT" followed by 4 nT
// vertex indices.  Every (tet, a) and (tet, pair) gets integer-valued
// contributions (exact in double), which are summed the way the element
// kernel reads a patch description (stage -> per local slot -> record) and
// then per vertex / upper slot over the record runs, as the gathers do; the
// sums must equal a direct accumulation over the tets exactly.  Prints "OK"
// and the record counts, or the first mismatch (exit 1).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <vector>

#include "sdx_maps.h"

using namespace sdx;

namespace {

int fail(const char* what, long a, long b) {
  std::printf("FAIL %s (%ld, %ld)\n", what, a, b);
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return fail("usage", 0, 0);
  std::ifstream in(argv[1]);
  std::vector<std::vector<int32_t>> meshes;
  std::vector<int> nVs;
  int nV, nT;
  while (in >> nV >> nT) {
    std::vector<int32_t> t(size_t(4) * size_t(nT));
    for (auto& x : t) in >> x;
    meshes.push_back(t);
    nVs.push_back(nV);
  }
  // one world of the actors side by side, as the world builds it
  PatchBuilder patches;
  std::vector<int32_t> vtx_keys, edge_keys;
  std::vector<uint32_t> vtx_codes, edge_codes;
  std::vector<int32_t> row_ptr(1, 0), col, tets, env_p0;
  int v0 = 0, t0 = 0;
  std::vector<Pattern> pats;
  for (size_t e = 0; e < meshes.size(); ++e) {
    const std::vector<int32_t>& A = meshes[e];
    const int av = nVs[e], at = int(A.size() / 4);
    Pattern P = BuildPattern(av, A);
    const int b0 = int(col.size());
    for (int i = 0; i < av; ++i) {
      for (int k = P.ptr[size_t(i)]; k < P.ptr[size_t(i) + 1]; ++k) col.push_back(v0 + P.col[size_t(k)]);
      row_ptr.push_back(int32_t(col.size()));
    }
    for (int32_t x : A) tets.push_back(v0 + x);
    env_p0.push_back(patches.count());
    for (int pt0 = 0; pt0 < at; pt0 += kPatchTets)
      patches.add(P, b0, v0, t0 + pt0, &A[size_t(4 * pt0)], std::min(kPatchTets, at - pt0), vtx_keys, vtx_codes,
                  edge_keys, edge_codes);
    v0 += av;
    t0 += at;
  }
  const int NV = v0, NT = t0, NB = int(col.size());
  std::vector<int32_t> vtx_ptr, eptr;
  std::vector<uint32_t> vtx_code, ecode;
  CountingCsr(NV, vtx_keys, vtx_codes, vtx_ptr, vtx_code);
  CountingCsr(NB, edge_keys, edge_codes, eptr, ecode);
  std::vector<int32_t> pos_lv(patches.n_local_vertices()), pos_le(patches.n_local_pairs());
  for (size_t q = 0; q < vtx_code.size(); ++q) pos_lv[vtx_code[q]] = int32_t(q);
  for (size_t q = 0; q < ecode.size(); ++q) pos_le[ecode[q]] = int32_t(q);
  std::vector<int32_t> patch_t0, meta_ptr, meta_mid;
  std::vector<uint32_t> meta;
  patches.finish(pos_lv, pos_le, patch_t0, meta_ptr, meta_mid, meta);
  if (patch_t0.back() != NT) return fail("patch tets", patch_t0.back(), NT);

  // contributions: residual 12, diagonal blocks 4 x 6, off-diagonal 6 x 9 per tet
  std::mt19937 rng(7);
  std::uniform_int_distribution<int> val(-1000, 1000);
  std::vector<double> cres(size_t(12) * NT), cdiag(size_t(24) * NT), coff(size_t(54) * NT);
  for (auto& x : cres) x = val(rng);
  for (auto& x : cdiag) x = val(rng);
  for (auto& x : coff) x = val(rng);

  // the kernel's reading of every patch description
  std::vector<double> Rv(size_t(3) * vtx_code.size(), NAN), Dv(size_t(6) * vtx_code.size(), NAN),
      Ev(size_t(9) * ecode.size(), NAN);
  for (int p = 0; p < patches.count(); ++p) {
    const uint32_t* m = &meta[size_t(meta_ptr[size_t(p)])];
    const int words = meta_ptr[size_t(p) + 1] - meta_ptr[size_t(p)];
    if (words % 4 || words > kPatchMetaWords) return fail("description size", p, words);
    if (meta_ptr[size_t(p)] % 4) return fail("description alignment", p, meta_ptr[size_t(p)]);
    const int pt0 = patch_t0[size_t(p)], nt = patch_t0[size_t(p) + 1] - pt0;
    const int nLV = int(m[0] & 1023u), nLE = int((m[0] >> 10) & 1023u);
    if (int(m[0] >> 20) != nt) return fail("patch tets in header", p, int(m[0] >> 20));
    // vertex part
    const int vw = PatchVertexWords(nLV, nt);
    if (meta_mid[size_t(p)] - meta_ptr[size_t(p)] != vw) return fail("vertex part size", p, vw);
    const uint32_t* rec = m + 1;
    const uint16_t* ls = reinterpret_cast<const uint16_t*>(m + 1 + nLV);
    const uint16_t* ent = ls + nLV + 1;
    if (1 + nLV + (nLV + 1 + 4 * nt + 1) / 2 > vw) return fail("vertex part overrun", p, vw);
    if (ls[0] != 0 || ls[nLV] != 4 * nt) return fail("vertex list starts", p, ls[nLV]);
    // pair part
    const uint32_t* prec = m + vw;
    const uint16_t* pls = reinterpret_cast<const uint16_t*>(prec + nLE);
    const uint16_t* pent = pls + nLE + 1;
    if (vw + nLE + (nLE + 1 + 6 * nt + 1) / 2 > words) return fail("pair part overrun", p, words);
    if (pls[0] != 0 || pls[nLE] != 6 * nt) return fail("pair list starts", p, pls[nLE]);
    std::vector<int> seen(size_t(10 * nt), 0);
    for (int j = 0; j < nLV; ++j) {
      double r[3] = {0, 0, 0}, d[6] = {0, 0, 0, 0, 0, 0};
      for (int q = ls[j]; q < ls[j + 1]; ++q) {
        const int c = ent[q], tl = c >> 2, a = c & 3;
        if (c >= 4 * nt) return fail("vertex entry", p, c);
        seen[size_t(c)]++;
        for (int k = 0; k < 3; ++k) r[k] += cres[size_t(12) * (pt0 + tl) + size_t(3 * a + k)];
        for (int k = 0; k < 6; ++k) d[k] += cdiag[size_t(24) * (pt0 + tl) + size_t(6 * a + k)];
      }
      const size_t rp = rec[j];
      if (rp >= vtx_code.size()) return fail("vertex record", p, long(rp));
      for (int k = 0; k < 3; ++k) Rv[3 * rp + size_t(k)] = r[k];
      for (int k = 0; k < 6; ++k) Dv[6 * rp + size_t(k)] = d[k];
    }
    for (int j = 0; j < nLE; ++j) {
      double o[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int q = pls[j]; q < pls[j + 1]; ++q) {
        const int c = pent[q], id = c & 0x7fff, tl = id / 6, k = id % 6;
        if (id >= 6 * nt) return fail("pair entry", p, c);
        seen[size_t(4 * nt + id)]++;
        const double* s = &coff[size_t(54) * (pt0 + tl) + size_t(9 * k)];
        for (int r = 0; r < 3; ++r)
          for (int cc = 0; cc < 3; ++cc) o[3 * r + cc] += (c >> 15) ? s[3 * cc + r] : s[3 * r + cc];
      }
      const size_t rp = prec[j];
      if (rp >= ecode.size()) return fail("pair record", p, long(rp));
      for (int k = 0; k < 9; ++k) Ev[9 * rp + size_t(k)] = o[k];
    }
    for (int k = 0; k < 10 * nt; ++k)
      if (seen[size_t(k)] != 1) return fail("contribution listed", p, k);
  }
  for (double x : Rv)
    if (std::isnan(x)) return fail("unwritten vertex record", 0, 0);
  for (double x : Ev)
    if (std::isnan(x)) return fail("unwritten pair record", 0, 0);

  // brute force: every tet's contributions straight to its vertices / slots
  std::vector<double> bres(size_t(3) * NV, 0.0), bdiag(size_t(6) * NV, 0.0), boff(size_t(9) * NB, 0.0);
  auto slot = [&](int i, int j) {
    for (int k = row_ptr[size_t(i)]; k < row_ptr[size_t(i) + 1]; ++k)
      if (col[size_t(k)] == j) return k;
    return -1;
  };
  for (int t = 0; t < NT; ++t)
    for (int a = 0; a < 4; ++a) {
      const int v = tets[size_t(4 * t + a)];
      for (int k = 0; k < 3; ++k) bres[size_t(3 * v + k)] += cres[size_t(12 * t + 3 * a + k)];
      for (int k = 0; k < 6; ++k) bdiag[size_t(6 * v + k)] += cdiag[size_t(24 * t + 6 * a + k)];
    }
  for (int t = 0; t < NT; ++t)
    for (int k = 0; k < 6; ++k) {
      const int va = tets[size_t(4 * t + kPairA[k])], vb = tets[size_t(4 * t + kPairB[k])];
      const int sl = slot(std::min(va, vb), std::max(va, vb));
      if (sl < 0) return fail("pattern", va, vb);
      const double* s = &coff[size_t(54 * t + 9 * k)];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) boff[size_t(9 * sl + 3 * r + c)] += va > vb ? s[3 * c + r] : s[3 * r + c];
    }
  // the gathers: each vertex / upper slot sums its run of records
  for (int v = 0; v < NV; ++v) {
    double r[3] = {0, 0, 0}, d[6] = {0, 0, 0, 0, 0, 0};
    for (int q = vtx_ptr[size_t(v)]; q < vtx_ptr[size_t(v) + 1]; ++q) {
      for (int k = 0; k < 3; ++k) r[k] += Rv[size_t(3 * q + k)];
      for (int k = 0; k < 6; ++k) d[k] += Dv[size_t(6 * q + k)];
    }
    for (int k = 0; k < 3; ++k)
      if (r[k] != bres[size_t(3 * v + k)]) return fail("residual", v, k);
    for (int k = 0; k < 6; ++k)
      if (d[k] != bdiag[size_t(6 * v + k)]) return fail("diagonal block", v, k);
  }
  for (int sl = 0; sl < NB; ++sl) {
    double o[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = eptr[size_t(sl)]; q < eptr[size_t(sl) + 1]; ++q)
      for (int k = 0; k < 9; ++k) o[k] += Ev[size_t(9 * q + k)];
    for (int k = 0; k < 9; ++k)
      if (o[k] != boff[size_t(9 * sl + k)]) return fail("off-diagonal block", sl, k);
  }
  std::printf("OK patches %d vertex_records %zu pair_records %zu tets %d\n", patches.count(), vtx_code.size(),
              ecode.size(), NT);
  return 0;
}
