// Byte and JSON forms of a world's dynamic state (container layout of
// mochi_physics/src/mochi_capture.cpp:38-75): encode, decode, compare, JSON.
#include "sdx_capture.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

namespace sdx::capture {

namespace {

// mochi_capture.cpp:58-69: label, block counts, 8 zero bytes (an unused
// field and the alignment of the next), bytes after the header
struct CaptureHeader {
  char label[4];
  uint16_t num_ctx;
  uint16_t num_entity;
  uint32_t unused[2];
  uint64_t data_size;
};
static_assert(sizeof(CaptureHeader) == 24, "capture header is 24 bytes");

// mochi_capture.cpp:38-47: type id, entity hash, data size, entity count
struct ComponentHeader {
  uint64_t id;
  uint64_t entity_hash;
  uint32_t data_size;
  uint32_t count;
};
static_assert(sizeof(ComponentHeader) == 24, "component header is 24 bytes");

constexpr char kLabel[4] = {'m', 'c', 'a', 'p'};
constexpr char kFooter[4] = {'m', 'e', 'n', 'd'};

uint64_t id_of(const char* name) { return fnv1a(name, std::strlen(name)); }

const char* real_name(uint32_t rb) { return rb == 8 ? "double" : "float"; }

std::string actor_component(const char* base, uint32_t rb) { return std::string(base) + "<" + real_name(rb) + ">"; }

const char* const kDisp = "mochi::CDisplacementSlice";
const char* const kVel = "mochi::CVelocitySlice";
const char* const kHist = "sdx::CNewtonStepHistory";
const char* const kGate = "sdx::CWarmStartGate";
const char* const kColl = "sdx::CColliderTarget";

// Which component a type id names, and the real size of actor components.
struct Known {
  uint32_t kind = 0;  // 0: unknown
  uint32_t real_bytes = 0;
  int which = -1;  // actor components: 0 displacement, 1 velocity, 2 history
};

Known lookup(uint64_t id) {
  for (uint32_t rb : {4u, 8u}) {
    if (id == id_of(actor_component(kDisp, rb).c_str())) return {kDisplacements, rb, 0};
    if (id == id_of(actor_component(kVel, rb).c_str())) return {kVelocities, rb, 1};
    if (id == id_of(actor_component(kHist, rb).c_str())) return {kWarmStart, rb, 2};
  }
  if (id == id_of(kColl)) return {kColliderTargets, 0, -1};
  if (id == id_of(kGate)) return {kWarmStart, 0, -1};
  return {};
}

class Writer {
 public:
  explicit Writer(uint8_t* out) : out_(out) {}
  void bytes(const void* p, size_t n) {
    if (n) std::memcpy(out_ + at_, p, n);
    at_ += n;
  }
  template <class V>
  void pod(const V& v) {
    bytes(&v, sizeof(V));
  }
  size_t pos() const { return at_; }
  template <class V>
  void pod_at(size_t at, const V& v) {
    std::memcpy(out_ + at, &v, sizeof(V));
  }

 private:
  uint8_t* out_;
  size_t at_ = 0;
};

class Reader {
 public:
  Reader(const uint8_t* p, size_t n) : p_(p), n_(n) {}
  void bytes(void* dst, size_t n) {
    if (n > n_ - at_) throw std::invalid_argument("state bytes end early");
    std::memcpy(dst, p_ + at_, n);
    at_ += n;
  }
  template <class V>
  V pod() {
    V v;
    bytes(&v, sizeof(V));
    return v;
  }
  size_t pos() const { return at_; }
  size_t left() const { return n_ - at_; }

 private:
  const uint8_t* p_;
  size_t n_;
  size_t at_ = 0;
};

// One entity block of per-actor arrays: per entity its length, then its reals.
void write_actor_block(Writer& w, const char* base, const StateParts& s, const std::vector<uint8_t>& data, int64_t mult) {
  ComponentHeader h{};
  h.id = id_of(actor_component(base, s.real_bytes).c_str());
  h.entity_hash = s.actor_hash;
  h.count = uint32_t(s.actor_dofs.size());
  uint64_t size = 0;
  for (int64_t d : s.actor_dofs) size += 8 + uint64_t(d * mult) * s.real_bytes;
  if (size > UINT32_MAX) throw std::invalid_argument("state component larger than the format's 4 GiB block limit");
  h.data_size = uint32_t(size);
  w.pod(h);
  size_t off = 0;
  for (int64_t d : s.actor_dofs) {
    const uint64_t n = uint64_t(d * mult);
    w.pod(n);
    w.bytes(data.data() + off, n * s.real_bytes);
    off += n * s.real_bytes;
  }
}

void read_actor_block(Reader& r, const ComponentHeader& h, uint32_t rb, int64_t mult, StateParts& s,
                      std::vector<uint8_t>& data) {
  const size_t start = r.pos();
  std::vector<int64_t> dofs;
  data.clear();
  for (uint32_t e = 0; e < h.count; ++e) {
    const uint64_t n = r.pod<uint64_t>();
    if (n % uint64_t(mult) != 0 || n / uint64_t(mult) % 3 != 0 || n > h.data_size || n * rb > r.left())
      throw std::invalid_argument("state component holds an array of a wrong length");
    dofs.push_back(int64_t(n / uint64_t(mult)));
    const size_t off = data.size();
    data.resize(off + n * rb);
    r.bytes(data.data() + off, n * rb);
  }
  if (r.pos() - start != h.data_size) throw std::invalid_argument("state component data size mismatch");
  if (s.actor_dofs.empty() && s.real_bytes == 0) {
    s.actor_dofs = dofs;
    s.real_bytes = rb;
    s.actor_hash = h.entity_hash;
  } else if (dofs != s.actor_dofs || rb != s.real_bytes || h.entity_hash != s.actor_hash) {
    throw std::invalid_argument("state components disagree on their actors");
  }
}

bool same_reals(const std::vector<uint8_t>& a, const std::vector<uint8_t>& b, uint32_t rb) {
  if (a.size() != b.size()) return false;
  const size_t n = a.size() / rb;
  for (size_t i = 0; i < n; ++i) {
    double x, y;
    if (rb == 8) {
      std::memcpy(&x, a.data() + 8 * i, 8);
      std::memcpy(&y, b.data() + 8 * i, 8);
    } else {
      float fx, fy;
      std::memcpy(&fx, a.data() + 4 * i, 4);
      std::memcpy(&fy, b.data() + 4 * i, 4);
      x = fx;
      y = fy;
    }
    if (!(x == y || (std::isnan(x) && std::isnan(y)))) return false;
  }
  return true;
}

// JSON text, compact or indented by two spaces per level.  Every value
// written after a key belongs to it; other values are array items.
class Json {
 public:
  explicit Json(bool pretty) : pretty_(pretty) {}
  void key(const std::string& k) {
    prelude();
    quoted(k);
    out_ += pretty_ ? ": " : ":";
    keyed_ = true;
  }
  void open(char c) {
    prelude();
    out_ += c;
    ++depth_;
    first_ = true;
  }
  void close(char c) {
    --depth_;
    if (!first_) newline();
    out_ += c;
    first_ = false;
  }
  void str(const std::string& s) {
    prelude();
    quoted(s);
  }
  void integer(int64_t v) {
    prelude();
    out_ += std::to_string(v);
  }
  // a flat array of numbers on one line
  void reals(const uint8_t* p, size_t n, uint32_t rb) {
    prelude();
    out_ += '[';
    for (size_t i = 0; i < n; ++i) {
      if (i) out_ += pretty_ ? ", " : ",";
      if (rb == 8) {
        double x;
        std::memcpy(&x, p + 8 * i, 8);
        number(x, 17);
      } else {
        float x;
        std::memcpy(&x, p + 4 * i, 4);
        number(x, 9);
      }
    }
    out_ += ']';
  }
  void ints(const int32_t* p, size_t n) {
    prelude();
    out_ += '[';
    for (size_t i = 0; i < n; ++i) {
      if (i) out_ += pretty_ ? ", " : ",";
      out_ += std::to_string(p[i]);
    }
    out_ += ']';
  }
  std::string text() const { return out_; }

 private:
  void prelude() {
    if (keyed_) {
      keyed_ = false;
      return;
    }
    if (!first_) out_ += ',';
    newline();
    first_ = false;
  }
  void newline() {
    if (!pretty_ || out_.empty()) return;
    out_ += '\n';
    out_.append(size_t(2 * depth_), ' ');
  }
  void quoted(const std::string& s) {
    out_ += '"';
    for (char c : s) {
      if (c == '"' || c == '\\') out_ += '\\';
      out_ += c;
    }
    out_ += '"';
  }
  void number(double v, int digits) {
    if (!std::isfinite(v)) {
      out_ += "null";
      return;
    }
    char buf[40];
    std::snprintf(buf, sizeof(buf), "%.*g", digits, v);
    out_ += buf;
  }
  bool pretty_;
  std::string out_;
  int depth_ = 0;
  bool first_ = true, keyed_ = false;
};

}  // namespace

uint64_t fnv1a(const void* p, size_t n, uint64_t h) {
  const auto* b = static_cast<const uint8_t*>(p);
  for (size_t k = 0; k < n; ++k) {
    h ^= b[k];
    h *= 1099511628211ull;
  }
  return h;
}

size_t encoded_size(const StateParts& s) {
  size_t n = sizeof(CaptureHeader) + sizeof(ComponentHeader) + sizeof(s.gate) + 4;
  if (!s.actor_dofs.empty()) {
    n += 3 * sizeof(ComponentHeader) + 3 * 8 * s.actor_dofs.size();
    for (int64_t d : s.actor_dofs) n += size_t(2 + kHistPlanes) * size_t(d) * s.real_bytes;
  }
  if (!s.colliders.empty()) n += sizeof(ComponentHeader) + s.colliders.size() * sizeof(double);
  return n;
}

void encode_into(const StateParts& s, uint8_t* out) {
  if (s.real_bytes != 4 && s.real_bytes != 8) throw std::invalid_argument("reals are 4 or 8 bytes");
  Writer w(out);
  CaptureHeader hd{};
  std::memcpy(hd.label, kLabel, 4);
  w.pod(hd);
  const size_t start = w.pos();
  // context components
  ComponentHeader g{};
  g.id = id_of(kGate);
  g.data_size = uint32_t(sizeof(s.gate));
  w.pod(g);
  w.bytes(s.gate, sizeof(s.gate));
  hd.num_ctx = 1;
  // entity components, each only when some entity has it
  if (!s.actor_dofs.empty()) {
    write_actor_block(w, kDisp, s, s.u, 1);
    write_actor_block(w, kVel, s, s.v, 1);
    write_actor_block(w, kHist, s, s.hist, kHistPlanes);
    hd.num_entity += 3;
  }
  const size_t nc = s.colliders.size() / 6;
  if (nc) {
    ComponentHeader h{};
    h.id = id_of(kColl);
    h.entity_hash = s.collider_hash;
    h.count = uint32_t(nc);
    h.data_size = uint32_t(nc * 6 * sizeof(double));
    w.pod(h);
    w.bytes(s.colliders.data(), nc * 6 * sizeof(double));
    hd.num_entity += 1;
  }
  w.bytes(kFooter, 4);
  hd.data_size = uint64_t(w.pos() - start);
  w.pod_at(0, hd);
  if (w.pos() != encoded_size(s)) throw std::logic_error("state encoding size mismatch");
}

std::vector<uint8_t> encode(const StateParts& s) {
  std::vector<uint8_t> out(encoded_size(s));
  encode_into(s, out.data());
  return out;
}

StateParts decode(const uint8_t* data, size_t n, uint32_t* present) {
  if (!data || n < sizeof(CaptureHeader)) throw std::invalid_argument("not a state: too short");
  Reader r(data, n);
  const CaptureHeader hd = r.pod<CaptureHeader>();
  if (std::memcmp(hd.label, kLabel, 4) != 0) throw std::invalid_argument("invalid capture data header");
  if (hd.data_size != n - sizeof(CaptureHeader))
    throw std::invalid_argument("capture data is malformed: its size does not match its header");
  StateParts s;
  s.real_bytes = 0;
  uint32_t got = 0;
  for (uint32_t b = 0; b < uint32_t(hd.num_ctx) + hd.num_entity; ++b) {
    const bool ctx = b < hd.num_ctx;
    const ComponentHeader h = r.pod<ComponentHeader>();
    const Known k = lookup(h.id);
    if (!k.kind || ctx != (h.id == id_of(kGate))) throw std::invalid_argument("unknown component type in state");
    if (ctx) {
      if (h.data_size != sizeof(s.gate)) throw std::invalid_argument("state component data size mismatch");
      r.bytes(s.gate, sizeof(s.gate));
      got |= kWarmStart;
    } else if (k.which >= 0) {
      std::vector<uint8_t>& dst = k.which == 0 ? s.u : (k.which == 1 ? s.v : s.hist);
      read_actor_block(r, h, k.real_bytes, k.which == 2 ? kHistPlanes : 1, s, dst);
      got |= k.kind;
    } else {
      if (uint64_t(h.data_size) != uint64_t(h.count) * 6 * sizeof(double))
        throw std::invalid_argument("state component data size mismatch");
      s.colliders.resize(size_t(h.count) * 6);
      r.bytes(s.colliders.data(), s.colliders.size() * sizeof(double));
      s.collider_hash = h.entity_hash;
      got |= kColliderTargets;
    }
  }
  char foot[4];
  r.bytes(foot, 4);
  if (std::memcmp(foot, kFooter, 4) != 0)
    throw std::invalid_argument("capture data is malformed: expected the footer tag at the end");
  if (r.pos() != n) throw std::invalid_argument("capture data is malformed: bytes after the footer");
  // actor arrays come as a set: displacement, velocity and history together
  if ((got & (kDisplacements | kVelocities)) && (s.u.size() != s.v.size() || s.hist.size() != kHistPlanes * s.u.size()))
    throw std::invalid_argument("state lacks some of its actors' components");
  if (s.real_bytes == 0) s.real_bytes = 4;
  if (present) *present = got;
  return s;
}

bool equal(const uint8_t* a, size_t na, const uint8_t* b, size_t nb) {
  if (na == nb && na && std::memcmp(a, b, na) == 0) return true;
  uint32_t pa = 0, pb = 0;
  const StateParts A = decode(a, na, &pa), B = decode(b, nb, &pb);
  if (pa != pb || A.real_bytes != B.real_bytes || A.actor_hash != B.actor_hash || A.collider_hash != B.collider_hash ||
      A.actor_dofs != B.actor_dofs || std::memcmp(A.gate, B.gate, sizeof(A.gate)) != 0)
    return false;
  if (A.colliders.size() != B.colliders.size()) return false;
  for (size_t i = 0; i < A.colliders.size(); ++i) {
    const double x = A.colliders[i], y = B.colliders[i];
    if (!(x == y || (std::isnan(x) && std::isnan(y)))) return false;
  }
  return same_reals(A.u, B.u, A.real_bytes) && same_reals(A.v, B.v, A.real_bytes) &&
         same_reals(A.hist, B.hist, A.real_bytes);
}

std::string to_json(const uint8_t* data, size_t n, bool pretty) {
  uint32_t present = 0;
  const StateParts s = decode(data, n, &present);
  const uint32_t rb = s.real_bytes;
  Json j(pretty);
  j.open('{');
  const size_t na = s.actor_dofs.size(), nc = s.colliders.size() / 6;
  if (na + nc) {
    j.key("actors");
    j.open('{');
    size_t off = 0;
    for (size_t a = 0; a < na; ++a) {
      const int64_t d = s.actor_dofs[a];
      j.key(a ? "Soft" + std::to_string(a + 1) : std::string("Soft"));
      j.open('{');
      j.key("_handle");
      j.integer(int64_t(a));
      j.key("_name");
      j.str("");
      j.key("_type");
      j.str("Soft");
      j.key("components");
      j.open('{');
      j.key(actor_component(kDisp, rb));
      j.reals(s.u.data() + off * rb, size_t(d), rb);
      j.key(actor_component(kVel, rb));
      j.reals(s.v.data() + off * rb, size_t(d), rb);
      j.key(actor_component(kHist, rb));
      j.open('[');
      for (int p = 0; p < kHistPlanes; ++p) {
        j.reals(s.hist.data() + (off * kHistPlanes + size_t(p) * size_t(d)) * rb, size_t(d), rb);
      }
      j.close(']');
      j.close('}');
      j.close('}');
      off += size_t(d);
    }
    for (size_t c = 0; c < nc; ++c) {
      j.key(c ? "Rigid" + std::to_string(c + 1) : std::string("Rigid"));
      j.open('{');
      j.key("_handle");
      j.integer(int64_t(c));
      j.key("_name");
      j.str("");
      j.key("_type");
      j.str("Rigid");
      j.key("components");
      j.open('{');
      j.key(kColl);
      j.open('{');
      const double* p = s.colliders.data() + 6 * c;
      const uint8_t* pb = reinterpret_cast<const uint8_t*>(p);
      j.key("target");
      j.reals(pb, 3, 8);
      j.key("stage_start_target");
      j.reals(pb + 3 * sizeof(double), 3, 8);
      j.close('}');
      j.close('}');
      j.close('}');
    }
    j.close('}');
  }
  j.key("scene");
  j.open('{');
  j.key(kGate);
  j.open('{');
  j.key("skip");
  j.ints(s.gate, 4);
  j.key("seen");
  j.ints(s.gate + 4, 4);
  j.key("net_skip");
  j.ints(s.gate + 8, 4);
  j.close('}');
  j.close('}');
  j.close('}');
  return j.text();
}

}  // namespace sdx::capture
