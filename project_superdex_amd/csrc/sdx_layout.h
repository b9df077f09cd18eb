// Layout constants shared by the device kernels and the host-side maps.
#pragma once

namespace sdx {

// Off-diagonal local pairs (a, b), a < b, of a tet in the order of their index k.
constexpr int kPairA[6] = {0, 0, 0, 1, 1, 2};
constexpr int kPairB[6] = {1, 2, 3, 2, 3, 3};
// Local pairs of a boundary triangle.
constexpr int kTriPairA[3] = {0, 0, 1};
constexpr int kTriPairB[3] = {1, 2, 2};

// Element patches: the tets of an environment are cut into runs of up to
// kPatchTets consecutive tets.  The description of a patch has two parts of
// 32-bit words, each a multiple of 4 words long, the vertex part first (all a
// residual-only assembly reads), at most kPatchMetaWords in all:
//   vertex part: word 0 = nLV | nLE << 10 | tets << 20 (local vertices,
//     local vertex pairs, tets); words 1 .. nLV = the record of each local
//     vertex; then 16-bit values: the starts ls[0 .. nLV] and the 4 tets
//     entries tet * 4 + a (tet within the patch); local vertex j sums the
//     entries [ls[j], ls[j + 1]);
//   pair part: words 0 .. nLE - 1 = the record of each local pair; then
//     16-bit values: the starts ls[0 .. nLE] and the 6 tets entries tet * 6 +
//     k, bit 15 set where the tet's pair k runs from the larger vertex to the
//     smaller (its block is the transpose of the record's).
// Every (tet, a) and (tet, k) of the patch is listed once, in increasing
// order within each slot.  Local vertices and pairs are sorted by vertex and
// by BSR slot.
constexpr int kPatchTets = 64;
// bound: vertex part 1 + 4 P + (4 P + 1 + 4 P + 1) / 2, pair part
// 6 P + (6 P + 1 + 6 P + 1) / 2, each rounded up to 4
constexpr int kPatchMetaWords = 1288;

// Words of the vertex part of a patch with nlv local vertices and nt tets.
constexpr int PatchVertexWords(int nlv, int nt) { return (1 + nlv + (nlv + 1 + 4 * nt + 1) / 2 + 3) / 4 * 4; }

}  // namespace sdx
