// Byte and JSON forms of a world's dynamic state, in the container layout of
// the reference's state capture (mochi_physics/src/mochi_capture.cpp:38-75,
// 456-560): a 24-byte header ("mcap", the numbers of context and entity
// component blocks, 8 zero bytes, the byte count after the header), the
// context component blocks, the entity component blocks, and a 4-byte footer
// ("mend").  Every block starts with a 24-byte header: the component's type
// id, a hash of the entities it covers, its data size and its entity count.
//
// The components of a world (type ids are FNV-1a 64 of the names):
//   context  "sdx::CWarmStartGate"                  12 int32 counters of the warm-start gate
//   actors   "mochi::CDisplacementSlice<R>"         per actor: uint64 n = 3 nV, n reals
//            "mochi::CVelocitySlice<R>"             per actor: uint64 n = 3 nV, n reals
//            "sdx::CNewtonStepHistory<R>"           per actor: uint64 n = 8 * 3 nV, n reals
//   rigid    "sdx::CColliderTarget"                 per collider: target and stage-start
//                                                   target (6 doubles)
// with R the world's real type (float or double).  The entity hash of the
// actor blocks covers every actor's mesh, material and permanent pins, that of
// the collider block every collider's kind, attachment and shape: a world
// restores only the states of worlds that step the same way.
//
// Host code only: no device calls, usable without a GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace sdx::capture {

// Component kinds (bit flags), as in include/sdx_fem.h's SDX_STATE_*.
enum : uint32_t {
  kDisplacements = 1u,
  kVelocities = 2u,
  kWarmStart = 4u,  // Newton-step history and the gate counters
  kColliderTargets = 8u,
};

constexpr int kGateCounters = 12;
constexpr int kHistPlanes = 8;  // two previous Newton steps for each of four Newton iterations

// The state of a world, with each actor's arrays concatenated in actor order
// (reals as raw bytes of the world's real type).
struct StateParts {
  uint32_t real_bytes = 4;
  uint64_t actor_hash = 0, collider_hash = 0;
  std::vector<int64_t> actor_dofs;  // 3 nV of every actor
  std::vector<uint8_t> u, v, hist;  // hist: per actor, kHistPlanes planes of its 3 nV reals
  std::vector<double> colliders;    // 6 per collider
  int32_t gate[kGateCounters] = {};
};

uint64_t fnv1a(const void* p, size_t n, uint64_t h = 1469598103934665603ull);

// The byte form of a state: its size, and the bytes written to out.
size_t encoded_size(const StateParts& s);
void encode_into(const StateParts& s, uint8_t* out);
std::vector<uint8_t> encode(const StateParts& s);

// Parse a byte form.  Throws std::invalid_argument when the bytes are not a
// well-formed state of this library (header, block sizes, footer).  `present`
// receives the component kinds the bytes hold.
StateParts decode(const uint8_t* data, size_t n, uint32_t* present = nullptr);

// Equal values of every component (NaN equal to NaN, 0 equal to -0):
// IsEqualState on two byte forms (mochi_capture.cpp:736-826).
bool equal(const uint8_t* a, size_t na, const uint8_t* b, size_t nb);

// A human-readable JSON document of a byte form, shaped like the reference's
// CaptureStateToJson (mochi_capture.cpp:592-710): "actors" named by type
// ("Soft", "Soft2", ..., "Rigid", ...) with "_handle", "_name", "_type" and
// "components", and "scene" holding the context components.  Reals are
// written with enough digits to read back exactly; non-finite values as null.
std::string to_json(const uint8_t* data, size_t n, bool pretty);

}  // namespace sdx::capture
