// World state of the MI355X soft-body FEM step (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sdx_fem.h"
#include "sdx_kernels.h"
#include "sdx_mlp.h"

namespace sdx {

struct DeviceError : std::runtime_error {
  explicit DeviceError(const std::string& m) : std::runtime_error(m) {}
};

extern thread_local std::string g_last_error;
sdx_status set_error(sdx_status code, const std::string& msg);
void hip_check(hipError_t e, const char* what);

// Runs fn() and maps every exception to an sdx_status with its message, so
// nothing thrown crosses the extern "C" boundary.
template <class F>
sdx_status guarded(F&& fn) {
  try {
    return fn();
  } catch (const DeviceError& e) {
    return set_error(SDX_ERR_DEVICE, e.what());
  } catch (const std::bad_alloc&) {
    return set_error(SDX_ERR_OUT_OF_MEMORY, "out of device or host memory");
  } catch (const std::out_of_range& e) {
    return set_error(SDX_ERR_INVALID_ARGUMENT, e.what());
  } catch (const std::invalid_argument& e) {
    return set_error(SDX_ERR_INVALID_ARGUMENT, e.what());
  } catch (const std::exception& e) {
    return set_error(SDX_ERR_STATE, e.what());
  } catch (...) {
    return set_error(SDX_ERR_STATE, "unknown exception");
  }
}

struct DBufBase {
  virtual ~DBufBase() = default;
  virtual void free() = 0;
};

// Device buffer owned by a world; registers itself for release.
template <class T>
struct DBuf : DBufBase {
  T* p = nullptr;
  size_t n = 0;
  hipStream_t* stream = nullptr;
  DBuf(std::vector<DBufBase*>& reg, hipStream_t* s) : stream(s) { reg.push_back(this); }
  ~DBuf() override { free(); }
  void alloc(size_t count) {
    free();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
    if (e == hipErrorOutOfMemory) throw std::bad_alloc();
    hip_check(e, "hipMalloc");
    n = count;
  }
  void free() override {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
  void upload(const std::vector<T>& h) {
    if (!p || n < h.size()) alloc(h.size());
    if (!h.empty())
      hip_check(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, *stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(*stream), "hipStreamSynchronize");
  }
  // The host vector must stay alive until the stream reaches this copy.
  void upload_async(const std::vector<T>& h, hipStream_t s) {
    if (!p || n < h.size()) throw std::logic_error("device buffer too small");
    if (!h.empty())
      hip_check(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  }
  // Raw transfers of the first n values (n <= the buffer's size), waited for.
  void download_raw(void* dst, size_t n) const {
    if (n > this->n) throw std::logic_error("device buffer too small");
    if (n) hip_check(hipMemcpyAsync(dst, p, n * sizeof(T), hipMemcpyDeviceToHost, *stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(*stream), "hipStreamSynchronize");
  }
  void upload_raw(const void* src, size_t n) {
    if (n > this->n) throw std::logic_error("device buffer too small");
    if (n) hip_check(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyHostToDevice, *stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(*stream), "hipStreamSynchronize");
  }
  void download(std::vector<T>& h) const {
    if (!h.empty())
      hip_check(hipMemcpyAsync(h.data(), p, h.size() * sizeof(T), hipMemcpyDeviceToHost, *stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(*stream), "hipStreamSynchronize");
  }
};

struct HostActor {
  int nV = 0, nT = 0;
  std::vector<double> X;
  std::vector<int32_t> tets;
  sdx_soft_params params{};
  std::vector<double> E, nu;  // per-element fields (empty: homogeneous)
  std::vector<uint8_t> pinned;
  std::vector<double> pin_value;
  // permanent boundary conditions (sdx_pin_nodes); pinned/pin_value also hold the
  // clearable ones of sdx_set_boundary_conditions
  std::vector<uint8_t> perm;
  std::vector<double> perm_value;
  std::vector<double> u, v;
};

struct HostCollider {
  int kind = 0;  // 0 plane, 1 sphere
  double p[3] = {0, 0, 0}, n[3] = {0, 1, 0}, r = 0;
  int actor = -1;
  double p0[3] = {0, 0, 0};  // position at stage start: the previous step's target
};

struct HostTiles {
  std::vector<int32_t> env, row0, row1, first;
};

class WorldBase {
 public:
  WorldBase(int device, int precision) : device_(device), precision_(precision) {
    hip_check(hipSetDevice(device), "hipSetDevice");
    hip_check(hipStreamCreate(&stream_), "hipStreamCreate");
    // coherent and mapped: the compaction kernel writes the active count here
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&pinned_active_), kBurstSlots * sizeof(int32_t),
                            hipHostMallocMapped | hipHostMallocCoherent),
              "hipHostMalloc");
    for (int k = 0; k < kBurstSlots; ++k) pinned_active_[k] = 0;
    hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&pinned_active_dev_), pinned_active_, 0),
              "hipHostGetDevicePointer");
    for (auto& ev : burst_done_) hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
  }
  virtual ~WorldBase() {
    for (auto& kv : timers_)
      for (auto& ev : kv.events) {
        (void)hipEventDestroy(ev.start);
        (void)hipEventDestroy(ev.stop);
      }
    for (auto& ev : burst_done_)
      if (ev) (void)hipEventDestroy(ev);
    if (pinned_active_) (void)hipHostFree(pinned_active_);
    if (stage_) (void)hipHostFree(stage_);
    if (dl_) (void)hipHostFree(dl_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  // Per-step host <-> device transfers (masks, collider targets,
  // per-environment scalars) go through pinned memory: from pinned memory a
  // copy is asynchronous DMA, from pageable memory the runtime stages it while
  // the host waits.  Uploads take space in a pinned ring; when the ring is
  // full the stream is synchronised (every earlier copy has run) and the ring
  // starts over.
  template <class U>
  void put(DBuf<U>& d, const std::vector<U>& h) {
    if (!d.p || d.n < h.size()) throw std::logic_error("device buffer too small");
    if (h.empty()) return;
    const size_t bytes = h.size() * sizeof(U);
    void* src = stage_take(bytes);
    std::memcpy(src, h.data(), bytes);
    hip_check(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync");
  }
  // Synchronous download of h.size() values through pinned memory.
  template <class U>
  void get(DBuf<U>& d, std::vector<U>& h) {
    if (h.empty()) return;
    const size_t bytes = h.size() * sizeof(U);
    if (dl_cap_ < bytes) {
      if (dl_) hip_check(hipHostFree(dl_), "hipHostFree");
      dl_ = nullptr;
      dl_cap_ = std::max<size_t>(bytes, size_t(1) << 16);
      hip_check(hipHostMalloc(reinterpret_cast<void**>(&dl_), dl_cap_), "hipHostMalloc");
    }
    hip_check(hipMemcpyAsync(dl_, d.p, bytes, hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    std::memcpy(h.data(), dl_, bytes);
  }
  // Queued downloads: fetch() records a copy of h.size() values of d into h,
  // fetch_all() issues every recorded copy through pinned memory and waits
  // once.  h must not be resized before fetch_all().
  template <class U>
  void fetch(DBuf<U>& d, std::vector<U>& h) {
    if (h.empty()) return;
    if (!d.p || d.n < h.size()) throw std::logic_error("device buffer too small");
    fetches_.push_back(Fetch{d.p, h.data(), h.size() * sizeof(U)});
  }
  void fetch_all() {
    size_t total = 0;
    for (const Fetch& f : fetches_) total += (f.bytes + 255) & ~size_t(255);
    if (total == 0) return;
    if (dl_cap_ < total) {  // no copy into dl_ is pending: every download waits
      if (dl_) hip_check(hipHostFree(dl_), "hipHostFree");
      dl_ = nullptr;
      dl_cap_ = std::max<size_t>(total, size_t(1) << 16);
      hip_check(hipHostMalloc(reinterpret_cast<void**>(&dl_), dl_cap_), "hipHostMalloc");
    }
    size_t off = 0;
    for (const Fetch& f : fetches_) {
      hip_check(hipMemcpyAsync(dl_ + off, f.src, f.bytes, hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
      off += (f.bytes + 255) & ~size_t(255);
    }
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    off = 0;
    for (const Fetch& f : fetches_) {
      std::memcpy(f.dst, dl_ + off, f.bytes);
      off += (f.bytes + 255) & ~size_t(255);
    }
    fetches_.clear();
  }
  int device() const { return device_; }
  int precision() const { return precision_; }

  int add_actor(HostActor a) {
    invalidate();
    actors_.push_back(std::move(a));
    return int(actors_.size()) - 1;
  }
  HostActor& actor(int a) {
    if (a < 0 || a >= int(actors_.size())) throw std::out_of_range("no such actor");
    return actors_[size_t(a)];
  }
  int num_actors() const { return int(actors_.size()); }
  int add_collider(const HostCollider& c) {
    colliders_.push_back(c);
    colliders_dirty_ = true;
    return int(colliders_.size()) - 1;
  }
  HostCollider& collider(int c) {
    if (c < 0 || c >= int(colliders_.size())) throw std::out_of_range("no such collider");
    colliders_dirty_ = true;
    return colliders_[size_t(c)];
  }
  const HostCollider& collider_view(int c) const {
    if (c < 0 || c >= int(colliders_.size())) throw std::out_of_range("no such collider");
    return colliders_[size_t(c)];
  }
  // Bring device state back to host mirrors and drop device buffers; the next
  // step rebuilds them (topology or material change).
  void invalidate() {
    pull_state();
    release_device();
  }

  virtual void finalize() = 0;
  virtual void step(double dt, const sdx_solver_params& sp, sdx_step_stats* stats, int n_stats) = 0;
  virtual void pull_state() = 0;
  virtual void release_device() = 0;
  virtual void get_state(int actor, int which, double* out) = 0;
  virtual void set_state(int actor, int which, const double* in) = 0;
  virtual void assemble_actor(int actor, double dt, const sdx_solver_params& sp, double* res, int32_t* ptr,
                              int32_t* col, double* vals) = 0;
  virtual double objective_actor(int actor, double dt, const sdx_solver_params& sp) = 0;
  // Clearable boundary conditions (AddBoundaryConditionDofsWorld / ClearBoundaryConditions):
  // DoF 3 node + axis held at a world coordinate; with clear, the earlier clearable ones go.
  virtual void set_boundary_conditions(int actor, const int32_t* dofs, const double* values, int n,
                                       bool clear) = 0;
  // Contact query (CQueryActorContactForces): with it on, every step ends by
  // evaluating each (actor, collider) pair's contact at the step's result.
  virtual void set_contact_query(bool on) = 0;
  // Force on the actor's samples from the collider (3), torque of the reaction on
  // the collider about its position (3), least signed distance of the samples in
  // contact (+inf: none), as of the last step.
  virtual void contact_wrench(int actor, int collider, double out[7]) = 0;
  virtual void info(int64_t out[8]) = 0;
  virtual void actor_bsr_size(int actor, int32_t* nrows, int32_t* nnzb) = 0;
  virtual void set_schedule(const sdx_schedule& s) = 0;
  // Learned warm start: a snapshot of the network parameters (FP32, kMlpParams
  // values; NULL removes the model).
  virtual void set_ws_params(const float* params) = 0;
  // Training samples: start collecting (capacity rows, per_env vertices per
  // environment and solve; capacity 0 stops and frees), read them back.
  virtual void collect(int64_t capacity, int per_env) = 0;
  virtual int64_t collected(float* x, float* t, int64_t cap_rows) = 0;
  // Device rows of the collected samples (after a stream synchronisation).
  virtual void collected_device(const float** x, const float** t, int64_t* n) = 0;
  virtual double autotune(double dt, const sdx_solver_params& sp, int trials) = 0;
  // State snapshots (Scene::CaptureState / RestoreState / ReleaseState /
  // IsEqualState / CaptureStateToBytes / RestoreStateFromBytes,
  // mochi_physics/include/mochi_physics/cpp_api/mochi_scene.h:79-93): the
  // displacements and velocities of every actor, the collider targets and
  // the Newton-step history of the learned warm start.  Handles keep a copy
  // on the device and are dropped with the device buffers when the topology
  // changes; the byte form is self-describing and checked against the world
  // on restore.
  virtual int32_t capture_state() = 0;
  virtual void restore_state(int32_t handle, bool release) = 0;
  virtual void release_state(int32_t handle) = 0;  // 0: every handle
  virtual bool equal_states(int32_t a, int32_t b) = 0;
  virtual size_t state_size() = 0;
  virtual void capture_bytes(uint8_t* out, size_t cap) = 0;
  virtual void restore_bytes(const uint8_t* in, size_t n, uint32_t excluded) = 0;
  const sdx_schedule& schedule() const { return sched_; }

  // kernel timing: class 0 CG matvec, 1 CG update, 2 element, 3 BSR gather,
  // 4 whole-solve CG (env_cg).
  // on > 0: time every on-th launch of an assembly class, and every launch of
  // the first CG burst of every on-th solve (other bursts replay as graphs);
  // 0: off.
  void timing_enable(int k, int on) {
    auto& t = timers_[size_t(k)];
    t.enabled = on > 0;
    t.every = on > 0 ? on : 1;
    t.counter = 0;
    t.launches = 0;
    t.ms = 0.0;
    t.used = 0;
  }
  std::pair<double, double> timing_read(int k) {
    flush_timers();
    return {double(timers_[size_t(k)].launches), timers_[size_t(k)].ms};
  }
  sdx_schedule sched_{8, 1, 16, 128, 1, 0, 16, 0, 1, 0, 1};
  bool tuned_ = false;  // the schedule was autotuned for the current topology

 protected:
  struct Timer {
    bool enabled = false;
    int every = 1;
    int64_t counter = 0;
    size_t used = 0;
    int64_t launches = 0;
    double ms = 0.0;
    std::vector<KernelEvents> events;
  };
  // Next event pair of class k (nullptr when the class is not timed).
  const KernelEvents* timer_events(int k) {
    Timer& t = timers_[size_t(k)];
    if (!t.enabled) return nullptr;
    if (t.used == t.events.size()) {
      KernelEvents ev{};
      hip_check(hipEventCreate(&ev.start), "hipEventCreate");
      hip_check(hipEventCreate(&ev.stop), "hipEventCreate");
      t.events.push_back(ev);
    }
    return &t.events[t.used++];
  }
  // Sampled launch of an assembly class: events for every `every`-th launch.
  const KernelEvents* sampled_events(int k) {
    Timer& t = timers_[size_t(k)];
    if (!t.enabled || (t.counter++ % t.every) != 0) return nullptr;
    return timer_events(k);
  }
  // Whether CG burst m of the current solve is launched eagerly with its
  // launches timed: the first burst (every environment still active; with
  // environment groups, each group's first burst) of every `every`-th solve.
  bool burst_timed(int m) {
    if (m != 0 || !(timers_[0].enabled || timers_[1].enabled)) return false;
    Timer& t = timers_[0];
    return (t.counter++ % t.every) == 0;
  }
  // Events of the whole-solve CG launch of every `every`-th solve.
  const KernelEvents* env_timer_events() {
    Timer& t = timers_[4];
    if (!t.enabled || (t.counter++ % t.every) != 0) return nullptr;
    return timer_events(4);
  }
  // Call after a stream synchronisation: folds recorded event pairs into totals.
  void flush_timers() {
    for (auto& t : timers_) {
      if (!t.enabled || t.used == 0) continue;
      hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
      for (size_t i = 0; i < t.used; ++i) {
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, t.events[i].start, t.events[i].stop), "hipEventElapsedTime");
        t.ms += double(ms);
      }
      t.launches += int64_t(t.used);
      t.used = 0;
    }
  }

  int device_;
  int precision_;
  hipStream_t stream_ = nullptr;
  // CG bursts in flight ahead of the host's check of the active count, and the
  // ring of count slots / events they use (at least kBurstAhead + 2)
  static constexpr int kBurstAhead = 2, kBurstSlots = 4;
  int32_t* pinned_active_ = nullptr;  // active-count slots, one per burst mod kBurstSlots
  int32_t* pinned_active_dev_ = nullptr;  // the same slots as a device pointer
  char* stage_ = nullptr;  // pinned upload ring
  size_t stage_cap_ = 0, stage_off_ = 0;
  char* dl_ = nullptr;  // pinned download buffer
  size_t dl_cap_ = 0;
  struct Fetch {
    const void* src;
    void* dst;
    size_t bytes;
  };
  std::vector<Fetch> fetches_;
  void* stage_take(size_t bytes) {
    const size_t need = (bytes + 255) & ~size_t(255);
    if (stage_off_ + need > stage_cap_) {
      hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");  // every staged copy has run
      if (need > stage_cap_) {
        if (stage_) hip_check(hipHostFree(stage_), "hipHostFree");
        stage_ = nullptr;
        stage_cap_ = std::max<size_t>(need, size_t(4) << 20);
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&stage_), stage_cap_), "hipHostMalloc");
      }
      stage_off_ = 0;
    }
    void* p = stage_ + stage_off_;
    stage_off_ += need;
    return p;
  }
  hipEvent_t burst_done_[kBurstSlots] = {};
  std::vector<HostActor> actors_;
  std::vector<HostCollider> colliders_;
  bool colliders_dirty_ = true;
  bool finalized_ = false;
  Timer timers_[5];
};

template <class T>
class World final : public WorldBase {
 public:
  World(int device, int precision);
  ~World() override;
  void finalize() override;
  void step(double dt, const sdx_solver_params& sp, sdx_step_stats* stats, int n_stats) override;
  void pull_state() override;
  void release_device() override;
  void get_state(int actor, int which, double* out) override;
  void set_state(int actor, int which, const double* in) override;
  void assemble_actor(int actor, double dt, const sdx_solver_params& sp, double* res, int32_t* ptr, int32_t* col,
                      double* vals) override;
  double objective_actor(int actor, double dt, const sdx_solver_params& sp) override;
  void set_boundary_conditions(int actor, const int32_t* dofs, const double* values, int n, bool clear) override;
  void set_contact_query(bool on) override;
  void contact_wrench(int actor, int collider, double out[7]) override;
  void info(int64_t out[8]) override;
  void actor_bsr_size(int actor, int32_t* nrows, int32_t* nnzb) override;
  void set_schedule(const sdx_schedule& s) override;
  double autotune(double dt, const sdx_solver_params& sp, int trials) override;
  int32_t capture_state() override;
  void restore_state(int32_t handle, bool release) override;
  void release_state(int32_t handle) override;
  bool equal_states(int32_t a, int32_t b) override;
  size_t state_size() override;
  void capture_bytes(uint8_t* out, size_t cap) override;
  void restore_bytes(const uint8_t* in, size_t n, uint32_t excluded) override;
  void set_ws_params(const float* params) override;
  void collect(int64_t capacity, int per_env) override;
  int64_t collected(float* x, float* t, int64_t cap_rows) override;
  void collected_device(const float** x, const float** t, int64_t* n) override;

 private:
  struct DevTiles {
    DBuf<int32_t> env, row0, row1, first;
    Tiles view{};
    DevTiles(std::vector<DBufBase*>& r, hipStream_t* s) : env(r, s), row0(r, s), row1(r, s), first(r, s) {}
  };
  void push_state();
  void upload_tiles(const HostTiles& h, DevTiles& d);
  void alloc_lists();
  // workgroups of the persistent CG kernels: about 1.6 x the matvec workgroups
  // (5 waves per SIMD) and 2 x the update workgroups that 256 CUs hold at once;
  // on 512 batched cubes these measured 3% faster than grids of half the size
  static constexpr int kNewtonHist = 4;
  int hist_k_ = 0;  // Newton iteration whose history the warm start uses
  bool stiff_damp_ = false;  // some actor has stiffness damping (the element kernel's damping variant)
  // contact query: per device collider entry (an environment's copy of a collider)
  // its host collider id, the first entry of each environment, and the last step's
  // results (kWrench values and the least distance per entry)
  bool contact_query_ = false, contact_ready_ = false;
  std::vector<int32_t> coll_entry_id_, coll_first_h_;
  std::vector<double> wrench_h_;
  std::vector<long long> wrench_dmin_h_;
  T* hist1() { return hist_.p + size_t(2 * std::min(hist_k_, kNewtonHist - 1)) * size_t(3 * nV_); }
  T* hist2() { return hist1() + size_t(3 * nV_); }
  static constexpr int kPersistentGridA = 8192, kPersistentGridV = 4096;
  void upload_colliders();
  void set_env_mask(const std::vector<int32_t>& mask);
  // Newton state of one environment during a step
  struct EnvState {
    bool active = true;
    int iters = 0;
    double res0 = 0, res = 0;
    int conv = SDX_CONV_NONE;
    int lin_iters = 0, ls_iters = 0, lin_conv = SDX_CONV_NONE, refreshes = 0;
    double lin_rel = 0, lin_true_rel = 0;
    int ws_tried = 0, ws_accepted = 0, ws_network = 0;
    double ws_ratio = 0, ws_ratio_hist = 0;
    double dx_norm = std::numeric_limits<double>::infinity();
    bool have_dx = false;
    int psd = SDX_PSD_ALWAYS;  // PSD mode (IfFailAlways turns Always after a failed try)
    double merit = 0;          // objective at u (objective-based line searches)
    double merit0 = 0;         // and where the Newton iteration started
    double prev_res = 0;       // |res| at the start of the previous Newton iteration
    double lin_res = 0;        // final residual norm of the last linear solve (its monitored norm)
    double eta = 0;            // relative tolerance of the current linear solve
    double res_start = 0;      // |res| where the Newton iteration started
    bool res_known = false;    // res_ and res hold the residual at the current u
  };
  void line_search(double dt, const sdx_solver_params& sp, const std::vector<int32_t>& take,
                   const std::vector<int32_t>& keep_final, std::vector<EnvState>& S, std::vector<uint8_t>& improved,
                   std::vector<int>& ls_it);
  // project: PSD-project the element Hessians of this assembly
  void assemble(double dt, const sdx_solver_params& sp, bool wantH, int project);
  // the DResidual (and residual) of the environments in mask, projected where proj[e]
  void assemble_h(double dt, const sdx_solver_params& sp, const std::vector<int32_t>& mask,
                  const std::vector<uint8_t>& proj);
  std::vector<double> read_env_sq();
  // queue the objective of the environments in the current mask into env_obj_
  void queue_energy(double dt, const sdx_solver_params& sp);
  // queue the per-environment dot a.b into env_dot_
  void queue_env_dot(const T* a, const T* b);
  struct WarmStats {
    std::vector<int32_t> accepted;  // learned warm start kept (per environment)
    std::vector<double> ratio;      // |b - A x0| / |b| of the learned guess
    std::vector<double> ratio_hist; // the same for the previous Newton steps alone
  };
  std::vector<EnvCG> cg_solve(const sdx_krylov_params& kp, const std::vector<int32_t>& solve,
                              std::vector<double>* true_rel, const std::vector<int32_t>* warm,
                              WarmStats* learned = nullptr, const std::vector<double>* env_tol = nullptr);
  // Runs CG on the device; false if no environment had anything to solve (h
  // then holds the final states), else the states are on the device (env_cg_).
  // timing_ok: its first burst may be a timed sample (a whole solve, not a
  // continuation of a few environments).
  bool cg_run(const CgConfig& cfg, const std::vector<int32_t>& solve, const std::vector<int32_t>& warm,
              bool decide, std::vector<EnvCG>& h, bool timing_ok);
  void ws_features();
  // FP64 ||b - A x||^2 and ||b||^2 of every environment with mask[e] into
  // norms2_ (queued); true_rel_from turns them into relative residuals.
  void issue_true_residual(const std::vector<int32_t>& mask);
  std::vector<double> true_rel_from(const std::vector<int32_t>& mask, const std::vector<double>& q) const;
  // assembly groups (sdx_schedule.asm_groups): contiguous environment ranges
  // assembled one after the other through record buffers sized for the
  // largest; one group keeps the records of every environment
  struct AsmGroup {
    AsmRange r;
    int64_t rv0, re0;  // first vertex-sorted and pair-sorted record of the group
    int e0, e1;        // its environments
  };
  std::vector<AsmGroup> asm_plan_;
  int asm_plan_n_ = 0;
  // row blocks of the fused assembly (sdx_schedule.fused_asm), built on first use
  bool rb_ready_ = false;
  RowBlockView rbv_{};
  bool plan_row_blocks();
  static constexpr double kAsmGroupBytes = 96e6;
  int asm_groups(const sdx_schedule& sc) const;
  void plan_assembly(int n);
  T* group_records(T* buf, int W, int64_t rec0) const;
  std::vector<int32_t> env_t0_, env_f0_, env_rv0_, env_re0_, env_p0_;
  int nP_ = 0;  // element patches
  void issue_cg(const CgConfig& cfg, int m, int par, int k0, int n, bool timed, int cls);
  void issue_cg_tail(const CgConfig& cfg, int par, int k0, int n);
  void plan_groups(int n);
  int cg_groups(const sdx_schedule& sc) const;
  static constexpr double kGroupBytes = 200e6;
  hipGraphExec_t cg_graph(const CgConfig& cfg, int m, int par, int B, int cls);
  void drop_graphs();
  template <class A>
  A* xcg() { return reinterpret_cast<A*>(xcg_.p); }

  std::vector<hipGraphExec_t> graphs_;
  // environment groups of the grouped CG bursts (cg_groups): tile ranges
  // [group_a_[g], group_a_[g + 1]) of the matvec tiles and likewise of the
  // update tiles, the largest group's tile counts, the number of groups they
  // were planned for (0: none) and the grouped bursts of the current solve
  std::vector<int32_t> group_a_, group_v_;
  int group_max_a_ = 0, group_max_v_ = 0, group_plan_n_ = 0, group_bursts_ = 0;
  // matvec tiles of the largest environment (0: not yet counted), and the
  // persistent grid classes (grids of 1, 1/4, 1/16, 1/64 of the full size)
  int max_env_a_ = 0;
  int env_max_rows_ = 0;   // rows of the largest environment
  int num_cus_ = 0;        // compute units of the device
  int env_lds_limit_ = 0;  // LDS bytes a workgroup may take on the device
  // whether the environments fit the whole-solve CG, and whether this solve
  // runs as one whole-solve launch (sdx_schedule.env_cg)
  bool env_cg_fits() const;
  bool env_cg_applies() const;
  static constexpr int kGridClasses = 4;
  // fused tail bursts (k_cg_tail): workgroups per launch, the barrier counter,
  // the flag a barrier sets when its wait runs out, and whether the current
  // solve used them
  static constexpr int kTailGrid = 64;
  bool tail_used_ = false;
  CgConfig graph_cfg_{};
  int graph_burst_ = 0;
  // the grouped bursts of a solve (every environment group) as one graph
  hipGraphExec_t grouped_graph_ = nullptr;
  int grouped_graph_groups_ = 0, grouped_graph_bursts_ = 0;
  // drops the captured graphs when the CG configuration or burst length changed
  void graphs_for(const CgConfig& cfg, int B);

  std::vector<DBufBase*> bufs_;
  int nEnv_ = 0, nV_ = 0, nT_ = 0, nB_ = 0, nF_ = 0;
  std::vector<int32_t> env_v0_;
  std::vector<int32_t> env_b0_;
  std::vector<int32_t> h_row_ptr_, h_col_;
  HostTiles tiles_v_, tiles_a_;
  MeshView mesh_{};
  FieldPtrs<T> fields_{};
#define SDX_BUF(type, name) DBuf<type> name{bufs_, &stream_};
  SDX_BUF(int32_t, d_rb_row0_)
  SDX_BUF(int32_t, d_rb_tet0_)
  SDX_BUF(int32_t, d_rb_tri0_)
  SDX_BUF(int32_t, d_rb_env_)
  SDX_BUF(int32_t, d_rb_rows_)
  SDX_BUF(int32_t, d_rb_rbase_)
  SDX_BUF(int32_t, d_rb_tets_)  // (tet, local rows) pairs
  SDX_BUF(int32_t, d_rb_tris_)  // (triangle, local rows) pairs
  SDX_BUF(int32_t, d_tets_)
  SDX_BUF(int32_t, d_tet_env_)
  SDX_BUF(int32_t, d_tris_)
  SDX_BUF(int32_t, d_tri_env_)
  SDX_BUF(int32_t, d_row_env_)
  SDX_BUF(int32_t, d_row_ptr_)
  SDX_BUF(int32_t, d_col_)
  SDX_BUF(int16_t, d_col16_)  // columns as 16-bit offsets from their rows (empty if one does not fit)
  SDX_BUF(uint32_t, tail_bar_)
  SDX_BUF(int32_t, tail_abort_)
  SDX_BUF(int32_t, d_slot_row_)
  SDX_BUF(int32_t, d_diag_slot_)
  SDX_BUF(int32_t, d_eptr_)
  SDX_BUF(int32_t, d_slot_desc_)     // [4 nB] gather description of each BSR slot
  SDX_BUF(int32_t, d_patch_t0_)
  SDX_BUF(int32_t, d_patch_meta_ptr_)
  SDX_BUF(int32_t, d_patch_meta_mid_)
  SDX_BUF(uint32_t, d_patch_meta_)
  SDX_BUF(int32_t, d_pos_vc_)
  SDX_BUF(int32_t, d_pos_ec_)
  SDX_BUF(int32_t, d_vtx_ptr_)
  SDX_BUF(uint8_t, d_pinned_)
  SDX_BUF(uint8_t, d_coll_flags_)
  SDX_BUF(T, d_pin_value_)
  SDX_BUF(double, d_contact_)
  SDX_BUF(double, d_coll_)
  SDX_BUF(int32_t, d_coll_first_)
  SDX_BUF(T, d_X_)
  SDX_BUF(T, d_geom_)
  SDX_BUF(T, u_)
  SDX_BUF(T, v_)
  SDX_BUF(T, u0_)
  SDX_BUF(T, v0_)
  SDX_BUF(T, res_)
  SDX_BUF(T, dx_)
  SDX_BUF(double, xcg_)  // CG iterate in the accumulator type (float in FP32 mode)
  SDX_BUF(T, xd_)        // mixed precision: FP32 increments of the iterate since the last refresh
  SDX_BUF(T, sol0_)
  SDX_BUF(T, r_)
  SDX_BUF(T, z_)
  SDX_BUF(T, p_)
  SDX_BUF(T, sv_)
  SDX_BUF(T, w_)
  SDX_BUF(double, rref_)
  SDX_BUF(double, rref2_)  // b - A v, b - A w of the two previous Newton steps (learned warm start)
  SDX_BUF(double, rref3_)
  SDX_BUF(T, cres_)   // element and contact residual records, vertex-sorted [4 nT + 3 nF][3]
  SDX_BUF(T, cdiag_)  // diagonal-block records, vertex-sorted [4 nT + 3 nF][6]
  SDX_BUF(T, coff_)   // off-diagonal-block records, pair-sorted [6 nT + 3 nF][9]
  SDX_BUF(T, vals_)
  SDX_BUF(T, dinv_)
  SDX_BUF(EnvCG, env_cg_)
  SDX_BUF(double, acc_)
  SDX_BUF(int32_t, solve_mask_)
  SDX_BUF(int32_t, warm_mask_)
  // Newton steps of earlier time steps, per Newton iteration index k (the
  // last kNewtonHist - 1 share a slot): [2 kNewtonHist][3 nV], slot 2k the
  // step of iteration k one time step back, slot 2k + 1 two time steps back
  SDX_BUF(T, hist_)
  SDX_BUF(float, ws_y_)          // Jacobi-scaled residual
  SDX_BUF(float, ws_F_)          // feature rows [nV][16]
  SDX_BUF(double, ws_scale_)     // per-environment scale of features and output
  SDX_BUF(double, ws_ratio_)     // relative residual of the learned guess
  SDX_BUF(double, ws_ratio_hist_)  // and of the previous Newton steps alone
  SDX_BUF(double, ws_alpha_)     // least-residual multiple of the learned guess
  SDX_BUF(double, ws_beta_)      // and of the two previous Newton steps
  SDX_BUF(double, ws_gamma_)
  SDX_BUF(int32_t, d_env_v0_)    // first vertex of each environment [nEnv+1]
  SDX_BUF(int32_t, d_env_order_)  // environments by rows, largest first (whole-solve CG queue)
  SDX_BUF(int32_t, env_next_)     // whole-solve CG queue position
  SDX_BUF(int32_t, list_a_)      // work list of active matvec tiles
  SDX_BUF(int32_t, list_v_)      // work list of active vector tiles
  SDX_BUF(int32_t, list_count_)  // list lengths, two parities each
  SDX_BUF(int32_t, num_active_)
  SDX_BUF(int32_t, env_mask_)
  std::vector<int32_t> env_mask_h_;  // host copy of env_mask_
  SDX_BUF(int32_t, patch_list_)      // element patches of the masked environments (partial assemblies)
  SDX_BUF(int32_t, ls_mask_)
  SDX_BUF(double, ls_scale_)
  SDX_BUF(int32_t, reset_)
  SDX_BUF(double, env_sq_)
  SDX_BUF(double, norms2_)
  SDX_BUF(double, env_tol_)  // per-environment relative tolerance of the linear solve
  SDX_BUF(double, env_obj_)  // per-environment objective (merit)
  SDX_BUF(double, env_dot_)  // per-environment dot product
  SDX_BUF(T, res0_)          // residual at the start of a Newton iteration (gradient-descent fallback)
  SDX_BUF(double, wrench_)     // contact query: kWrench values per collider entry
  SDX_BUF(long long, wrench_dmin_)
#undef SDX_BUF
  DevTiles dt_v_{bufs_, &stream_};
  DevTiles dt_a_{bufs_, &stream_};
  // kept across topology changes: the warm-start network and collected samples
  std::vector<DBufBase*> keep_;
  DBuf<float> ws_params_{keep_, &stream_};
  DBuf<uint16_t> ws_wpack_{keep_, &stream_};
  DBuf<float> ws_bpack_{keep_, &stream_};
  MlpBf16 ws_w_{};
  bool ws_on_ = false;
  // Gates of the warm start, per Newton iteration index (the last
  // kNewtonHist - 1 share one); both act from the index's third warm-started
  // solve on (the first two run before its history of previous steps is filled):
  // - skip: steps left that start from zero because most guesses of its last
  //   warm-started solve were rejected (fewer than kWarmMinKept of the
  //   environments kept theirs);
  // - net_skip: steps left that combine the previous Newton steps alone,
  //   without running the network, because in its last solve with the network
  //   the network's guess left more than kNetMinGain of the residual that the
  //   previous steps alone leave (geometric mean over the environments);
  // - seen: warm-started solves per index.
  struct WarmGate {
    int32_t skip[kNewtonHist] = {0, 0, 0, 0};
    int32_t seen[kNewtonHist] = {0, 0, 0, 0};
    int32_t net_skip[kNewtonHist] = {0, 0, 0, 0};
  };
  WarmGate wg_;
  static constexpr double kWarmMinKept = 0.5;
  static constexpr double kNetMinGain = 0.5;
  static constexpr int kWarmSkipSteps = 8;
  // state snapshots under their handles
  struct Snapshot {
    std::vector<DBufBase*> reg;
    DBuf<T> u, v, hist;
    std::vector<double> coll_p;  // collider targets and stage-start positions, 6 per collider
    WarmGate wg;
    explicit Snapshot(hipStream_t* s) : u(reg, s), v(reg, s), hist(reg, s) {}
  };
  std::map<int32_t, std::unique_ptr<Snapshot>> snaps_;
  int32_t next_snap_ = 1;
  Snapshot& snapshot(int32_t handle);
  std::unique_ptr<Snapshot> take_snapshot();
  void put_snapshot(const Snapshot& s);
  DBuf<float> coll_x_{keep_, &stream_};
  DBuf<float> coll_t_{keep_, &stream_};
  int64_t coll_cap_ = 0, coll_n_ = 0;
  int coll_per_env_ = 0;
  uint32_t coll_salt_ = 0;
};

}  // namespace sdx
