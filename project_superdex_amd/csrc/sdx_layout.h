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
// kPatchTets consecutive tets.  Description of a patch (32-bit words, a
// multiple of 4 of them, at most kPatchMetaWords):
//   word 0 ............ nLV | nLE << 10 | tets << 20 (local vertices, pairs)
//   words 1 .. nS ..... record of each local vertex, then of each local pair
//                       (nS = nLV + nLE)
//   then, as 16-bit values, list starts ls[0 .. nS] (2 per word, from word
//   1 + nS) and the list entries (from word 1 + nS + (nS + 2) / 2): slot j
//   sums entries [ls[j], ls[j + 1]); a vertex entry is tet * 4 + a (tet
//   within the patch), a pair entry is tet * 6 + k, bit 15 set where the
//   tet's pair k runs from the larger vertex to the smaller (its block is the
//   transpose of the record's).  Every (tet, a) and (tet, k) of the patch is
//   listed once, in increasing order within each slot.  Local vertices and
//   pairs are sorted by vertex and by BSR slot.
constexpr int kPatchTets = 64;
constexpr int kPatchMetaWords = 1284;  // bound: 1 + 10 P + (10 P + 2) / 2 + 5 P, rounded up to 4

}  // namespace sdx
