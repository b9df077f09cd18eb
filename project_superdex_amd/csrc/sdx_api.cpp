// C ABI of include/sdx_fem.h: argument checking, error codes and dispatch
// to World<T> / the standalone PCG solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/sdx_fem.h"
#include "sdx_capture.h"
#include "sdx_kernels.h"
#include "sdx_mlp_host.h"
#include "sdx_world.h"

using sdx::set_error;

struct sdx_world {
  std::unique_ptr<sdx::WorldBase> impl;
};

struct sdx_mlp {
  std::unique_ptr<sdx::Mlp> impl;
};

namespace {

using sdx::guarded;

#define SDX_REQUIRE(cond, msg) \
  do {                         \
    if (!(cond)) return set_error(SDX_ERR_INVALID_ARGUMENT, msg); \
  } while (0)

}  // namespace

extern "C" {

const char* sdx_last_error(void) { return sdx::g_last_error.c_str(); }

void sdx_default_krylov(sdx_krylov_params* k) {
  k->abs_tol = 1e-9;
  k->rel_tol = 1e-5;
  k->rel_div_tol = 1e10;
  k->max_iter = 500;
  k->norm_type = SDX_NORM_PRECONDITIONED_RESIDUAL_L2;
  k->preconditioner = SDX_PREC_BLOCK_JACOBI;
  k->abort_if_not_spd = 0;
  k->initial_guess_zero = 1;
  k->precision_mode = SDX_PREC_MODE_FP32;
  k->refresh_every = 0;
  k->reserved = 0;
}

void sdx_default_solver_params(sdx_solver_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->newton_max_iter = 4;
  p->psd_proj_mode = SDX_PSD_ALWAYS;
  p->newton_abs_tol = 1e-3;
  p->newton_rel_tol = 1e-6;
  p->rel_step_tol = 10.0 * double(std::numeric_limits<float>::epsilon());
  p->explosion_control = 1;
  p->line_search_type = SDX_LS_RESIDUAL_NORM;
  p->abs_div_tol = 1e9;
  p->rel_div_tol = 1e4;
  p->line_search_max_iter = 4;
  p->linear_tolerance_strategy = SDX_LINTOL_EW2;
  p->line_search_alpha = 0.5;
  p->line_search_wolfe1 = 1e-4;
  p->line_search_wolfe2 = 0.9;
  p->line_search_max_rel_increase = 0.0;
  p->gradient_descent_fallback = 0;
  sdx_default_krylov(&p->linear);
  p->gravity[0] = 0.0;
  p->gravity[1] = -9.8;
  p->gravity[2] = 0.0;
  p->warm_start = SDX_WARM_START_ZERO;
  p->schedule = 0;
}

void sdx_default_soft_params(sdx_soft_params* s) {
  s->youngs_modulus = 1e5;
  s->poisson_ratio = 0.45;
  s->density = 1000.0;
  s->mass_damping = 0.0;
  s->stiffness_damping = 0.0;
  s->penalty_coefficient = 1e9;
  s->penalty_smoothing_half_distance = 0.005;
  s->penalty_threshold = 0.001;
  s->max_alignment_normals = 0.0;
  s->coulomb_friction_coefficient = 0.5;
  s->viscous_friction_coefficient = 0.0;
  s->friction_falloff_vel = 0.01;
  s->normal_viscous_damping_coefficient = 0.0;
  s->stiffness_damping_geometric_term = 0;
  s->reserved = 0;
}

int32_t sdx_abi_struct_size(int32_t which) {
  switch (which) {
    case 0: return int32_t(sizeof(sdx_krylov_params));
    case 1: return int32_t(sizeof(sdx_linear_status));
    case 2: return int32_t(sizeof(sdx_solver_params));
    case 3: return int32_t(sizeof(sdx_step_stats));
    case 4: return int32_t(sizeof(sdx_soft_params));
    case 5: return int32_t(sizeof(sdx_schedule));
    case 6: return int32_t(sizeof(sdx_train_params));
    default: return -1;
  }
}

sdx_status sdx_world_create(int32_t device, int32_t precision_mode, sdx_world** out) {
  SDX_REQUIRE(out, "out is NULL");
  SDX_REQUIRE(precision_mode >= SDX_PREC_MODE_FP32 && precision_mode <= SDX_PREC_MODE_FP64, "bad precision mode");
  return guarded([&]() -> sdx_status {
    int n = 0;
    sdx::hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    SDX_REQUIRE(device >= 0 && device < n, "no such HIP device");
    auto* w = new sdx_world;
    if (precision_mode == SDX_PREC_MODE_FP64)
      w->impl.reset(new sdx::World<double>(device, precision_mode));
    else
      w->impl.reset(new sdx::World<float>(device, precision_mode));
    *out = w;
    return SDX_OK;
  });
}

sdx_status sdx_world_destroy(sdx_world* world) {
  delete world;
  return SDX_OK;
}

sdx_status sdx_add_soft_tet_mesh(sdx_world* world, const double* X, int32_t n_verts, const int32_t* tets,
                                 int32_t n_tets, const sdx_soft_params* params, int32_t* out_actor) {
  SDX_REQUIRE(world && X && tets && params, "NULL argument");
  SDX_REQUIRE(n_verts > 0 && n_tets > 0, "empty mesh");
  SDX_REQUIRE(params->youngs_modulus > 0 && params->poisson_ratio > -1 && params->poisson_ratio < 0.5,
              "invalid Young's modulus or Poisson ratio");
  SDX_REQUIRE(params->density > 0, "density must be positive");
  SDX_REQUIRE(params->penalty_coefficient > 0 && params->penalty_smoothing_half_distance >= 0,
              "invalid contact parameters");
  SDX_REQUIRE(params->max_alignment_normals >= -1 && params->max_alignment_normals <= 1,
              "max_alignment_normals must be in [-1, 1]");
  SDX_REQUIRE(params->coulomb_friction_coefficient >= 0 && params->viscous_friction_coefficient >= 0 &&
                  params->friction_falloff_vel >= 0 && params->normal_viscous_damping_coefficient >= 0,
              "friction coefficients and falloff velocity must not be negative");
  // material_params_utils.cpp:107-109
  SDX_REQUIRE(std::isfinite(params->mass_damping) && params->mass_damping >= 0 &&
                  std::isfinite(params->stiffness_damping) && params->stiffness_damping >= 0,
              "damping coefficients must be finite and non-negative");
  SDX_REQUIRE(params->stiffness_damping_geometric_term == 0 || params->stiffness_damping_geometric_term == 1,
              "stiffness_damping_geometric_term must be 0 or 1");
  for (int64_t k = 0; k < 4 * int64_t(n_tets); ++k)
    SDX_REQUIRE(tets[k] >= 0 && tets[k] < n_verts, "tet node index out of range");
  return guarded([&]() -> sdx_status {
    sdx::HostActor a;
    a.nV = n_verts;
    a.nT = n_tets;
    a.X.assign(X, X + 3 * size_t(n_verts));
    a.tets.assign(tets, tets + 4 * size_t(n_tets));
    a.params = *params;
    a.pinned.assign(3 * size_t(n_verts), 0);
    a.pin_value.assign(3 * size_t(n_verts), 0.0);
    a.perm.assign(3 * size_t(n_verts), 0);
    a.perm_value.assign(3 * size_t(n_verts), 0.0);
    a.u.assign(3 * size_t(n_verts), 0.0);
    a.v.assign(3 * size_t(n_verts), 0.0);
    const int id = world->impl->add_actor(std::move(a));
    if (out_actor) *out_actor = id;
    return SDX_OK;
  });
}

sdx_status sdx_set_element_materials(sdx_world* world, int32_t actor, const double* E, const double* nu,
                                     int32_t n_tets) {
  SDX_REQUIRE(world && E && nu, "NULL argument");
  return guarded([&]() -> sdx_status {
    sdx::HostActor& a = world->impl->actor(actor);
    SDX_REQUIRE(n_tets == a.nT, "n_tets does not match the actor");
    for (int t = 0; t < n_tets; ++t)
      SDX_REQUIRE(E[t] > 0 && nu[t] > -1 && nu[t] < 0.5, "invalid per-element material");
    world->impl->invalidate();
    sdx::HostActor& b = world->impl->actor(actor);
    b.E.assign(E, E + n_tets);
    b.nu.assign(nu, nu + n_tets);
    return SDX_OK;
  });
}

sdx_status sdx_pin_nodes(sdx_world* world, int32_t actor, const int32_t* nodes, int32_t n) {
  SDX_REQUIRE(world && (nodes || n == 0), "NULL argument");
  return guarded([&]() -> sdx_status {
    world->impl->invalidate();
    sdx::HostActor& a = world->impl->actor(actor);
    for (int i = 0; i < n; ++i) SDX_REQUIRE(nodes[i] >= 0 && nodes[i] < a.nV, "node index out of range");
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        const size_t k = 3 * size_t(nodes[i]) + size_t(d);
        a.pinned[k] = a.perm[k] = 1;
        a.pin_value[k] = a.perm_value[k] = a.u[k];
      }
    return SDX_OK;
  });
}

sdx_status sdx_set_boundary_conditions(sdx_world* world, int32_t actor, const int32_t* dofs,
                                       const double* values_world, int32_t n, int32_t clear) {
  SDX_REQUIRE(world && n >= 0 && ((dofs && values_world) || n == 0), "NULL argument");
  return guarded([&]() -> sdx_status {
    std::vector<int32_t> sorted(dofs, dofs + n);
    std::sort(sorted.begin(), sorted.end());
    SDX_REQUIRE(std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end(),
                "boundary-condition DoF indices must be unique");
    world->impl->set_boundary_conditions(actor, dofs, values_world, n, clear != 0);
    return SDX_OK;
  });
}

sdx_status sdx_register_contact_query(sdx_world* world, int32_t enable) {
  SDX_REQUIRE(world, "NULL argument");
  return guarded([&]() -> sdx_status {
    world->impl->set_contact_query(enable != 0);
    return SDX_OK;
  });
}

sdx_status sdx_get_contact_wrench(sdx_world* world, int32_t actor, int32_t collider, double out[7]) {
  SDX_REQUIRE(world && out, "NULL argument");
  return guarded([&]() -> sdx_status {
    world->impl->contact_wrench(actor, collider, out);
    return SDX_OK;
  });
}

sdx_status sdx_add_rigid_plane(sdx_world* world, const double p[3], const double n[3], int32_t actor,
                               int32_t* out_collider) {
  SDX_REQUIRE(world && p && n, "NULL argument");
  const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  SDX_REQUIRE(l > 0 && std::isfinite(l), "plane normal must be non-zero");
  return guarded([&]() -> sdx_status {
    SDX_REQUIRE(actor >= -1 && actor < world->impl->num_actors(), "no such actor");
    sdx::HostCollider c;
    c.kind = 0;
    for (int d = 0; d < 3; ++d) {
      c.p[d] = p[d];
      c.p0[d] = p[d];  // stage start of the first step: where it was created
      c.n[d] = n[d] / l;
    }
    c.actor = actor;
    const int id = world->impl->add_collider(c);
    if (out_collider) *out_collider = id;
    return SDX_OK;
  });
}

sdx_status sdx_add_rigid_sphere(sdx_world* world, const double c[3], double r, int32_t actor,
                                int32_t* out_collider) {
  SDX_REQUIRE(world && c, "NULL argument");
  SDX_REQUIRE(r > 0, "sphere radius must be positive");
  return guarded([&]() -> sdx_status {
    SDX_REQUIRE(actor >= -1 && actor < world->impl->num_actors(), "no such actor");
    sdx::HostCollider s;
    s.kind = 1;
    for (int d = 0; d < 3; ++d) s.p[d] = s.p0[d] = c[d];  // stage start of the first step: where it was created
    s.r = r;
    s.actor = actor;
    const int id = world->impl->add_collider(s);
    if (out_collider) *out_collider = id;
    return SDX_OK;
  });
}

sdx_status sdx_set_rigid_position(sdx_world* world, int32_t collider, const double pos[3]) {
  SDX_REQUIRE(world && pos, "NULL argument");
  return guarded([&]() -> sdx_status {
    sdx::HostCollider& c = world->impl->collider(collider);
    for (int d = 0; d < 3; ++d) c.p[d] = pos[d];
    return SDX_OK;
  });
}

sdx_status sdx_get_rigid_position(sdx_world* world, int32_t collider, double out[3]) {
  SDX_REQUIRE(world && out, "NULL argument");
  return guarded([&]() -> sdx_status {
    const sdx::HostCollider& c = world->impl->collider_view(collider);
    for (int d = 0; d < 3; ++d) out[d] = c.p[d];
    return SDX_OK;
  });
}

sdx_status sdx_get_rigid_positions(sdx_world* world, int32_t n, const int32_t* colliders, double* positions) {
  SDX_REQUIRE(world && n >= 0 && (n == 0 || (colliders && positions)), "NULL argument");
  return guarded([&]() -> sdx_status {
    for (int32_t k = 0; k < n; ++k) {
      const sdx::HostCollider& c = world->impl->collider_view(colliders[k]);
      for (int d = 0; d < 3; ++d) positions[3 * size_t(k) + size_t(d)] = c.p[d];
    }
    return SDX_OK;
  });
}

sdx_status sdx_set_rigid_positions(sdx_world* world, int32_t n, const int32_t* colliders, const double* positions) {
  SDX_REQUIRE(world && n >= 0 && (n == 0 || (colliders && positions)), "NULL argument");
  return guarded([&]() -> sdx_status {
    for (int32_t k = 0; k < n; ++k) world->impl->collider(colliders[k]);  // throws on a bad index
    for (int32_t k = 0; k < n; ++k) {
      sdx::HostCollider& c = world->impl->collider(colliders[k]);
      for (int d = 0; d < 3; ++d) c.p[d] = positions[3 * size_t(k) + size_t(d)];
    }
    return SDX_OK;
  });
}

static sdx_status validate_solver_params(const sdx_solver_params& p) {
  const sdx_solver_params* params = &p;
  SDX_REQUIRE(params->newton_max_iter >= 0 && params->linear.max_iter > 0, "invalid iteration limits");
  SDX_REQUIRE(params->line_search_type >= SDX_LS_NONE && params->line_search_type <= SDX_LS_ARMIJO_OR_RESIDUAL_NORM,
              "unknown line search type");
  SDX_REQUIRE(params->line_search_type == SDX_LS_NONE || params->line_search_max_iter >= 1,
              "a line search needs line_search_max_iter >= 1 (use SDX_LS_NONE for none)");
  // NewtonSolver::SetParams (newton_solver.cpp:262-280) only warns about these; a step
  // factor, or Wolfe coefficients, outside these ranges never end a line search well
  SDX_REQUIRE(params->line_search_alpha > 0.0 && params->line_search_alpha < 1.0, "line_search_alpha must be in (0, 1)");
  SDX_REQUIRE(params->line_search_wolfe1 > 0.0 && params->line_search_wolfe1 < 1.0,
              "line_search_wolfe1 must be in (0, 1)");
  SDX_REQUIRE(params->line_search_wolfe2 > params->line_search_wolfe1 && params->line_search_wolfe2 < 1.0,
              "line_search_wolfe2 must be in (line_search_wolfe1, 1)");
  SDX_REQUIRE(params->line_search_max_rel_increase >= 0.0 && std::isfinite(params->line_search_max_rel_increase),
              "line_search_max_rel_increase must be finite and >= 0");
  SDX_REQUIRE(params->psd_proj_mode >= SDX_PSD_NEVER && params->psd_proj_mode <= SDX_PSD_IF_FAIL_ALWAYS,
              "unknown PSD projection mode");
  SDX_REQUIRE(params->linear_tolerance_strategy >= SDX_LINTOL_CONSTANT &&
                  params->linear_tolerance_strategy <= SDX_LINTOL_EW3,
              "unknown linear tolerance strategy");
  SDX_REQUIRE(params->linear.preconditioner >= SDX_PREC_NONE && params->linear.preconditioner <= SDX_PREC_BLOCK_JACOBI,
              "unsupported preconditioner");
  SDX_REQUIRE(params->linear.norm_type >= 0 && params->linear.norm_type <= 2, "unsupported norm type");
  SDX_REQUIRE(params->warm_start == SDX_WARM_START_ZERO || params->warm_start == SDX_WARM_START_LEARNED ||
                  params->warm_start == SDX_WARM_START_HISTORY,
              "unknown warm start");
  SDX_REQUIRE(params->schedule == 0 || params->schedule == -1, "schedule must be 0 or -1");
  return SDX_OK;
}

sdx_status sdx_step(sdx_world* world, double dt, const sdx_solver_params* params, sdx_step_stats* stats,
                    int32_t n_stats) {
  SDX_REQUIRE(world && params, "NULL argument");
  SDX_REQUIRE(dt > 0 && std::isfinite(dt), "dt must be positive");
  {
    const sdx_status v = validate_solver_params(*params);
    if (v != SDX_OK) return v;
  }
  return guarded([&]() -> sdx_status {
    world->impl->step(dt, *params, stats, stats ? n_stats : 0);
    return SDX_OK;
  });
}

static sdx_status get_state(sdx_world* world, int32_t actor, int which, double* out, int32_t n) {
  SDX_REQUIRE(world && out, "NULL argument");
  return guarded([&]() -> sdx_status {
    SDX_REQUIRE(n == 3 * world->impl->actor(actor).nV, "size must be 3*n_verts");
    world->impl->get_state(actor, which, out);
    return SDX_OK;
  });
}

static sdx_status set_state(sdx_world* world, int32_t actor, int which, const double* in, int32_t n) {
  SDX_REQUIRE(world && in, "NULL argument");
  return guarded([&]() -> sdx_status {
    SDX_REQUIRE(n == 3 * world->impl->actor(actor).nV, "size must be 3*n_verts");
    world->impl->set_state(actor, which, in);
    return SDX_OK;
  });
}

sdx_status sdx_get_displacements(sdx_world* w, int32_t a, double* out, int32_t n) { return get_state(w, a, 0, out, n); }
sdx_status sdx_set_displacements(sdx_world* w, int32_t a, const double* in, int32_t n) { return set_state(w, a, 0, in, n); }
sdx_status sdx_get_velocities(sdx_world* w, int32_t a, double* out, int32_t n) { return get_state(w, a, 1, out, n); }
sdx_status sdx_set_velocities(sdx_world* w, int32_t a, const double* in, int32_t n) { return set_state(w, a, 1, in, n); }
sdx_status sdx_get_positions_world(sdx_world* w, int32_t a, double* out, int32_t n) { return get_state(w, a, 2, out, n); }

sdx_status sdx_capture_state(sdx_world* world, int32_t* handle) {
  SDX_REQUIRE(world && handle, "NULL argument");
  return guarded([&]() -> sdx_status {
    *handle = world->impl->capture_state();
    return SDX_OK;
  });
}

sdx_status sdx_restore_state(sdx_world* world, int32_t handle, int32_t release_immediately) {
  SDX_REQUIRE(world, "NULL argument");
  SDX_REQUIRE(handle > 0, "state handles are positive");
  return guarded([&]() -> sdx_status {
    world->impl->restore_state(handle, release_immediately != 0);
    return SDX_OK;
  });
}

sdx_status sdx_release_state(sdx_world* world, int32_t handle) {
  SDX_REQUIRE(world, "NULL argument");
  SDX_REQUIRE(handle >= 0, "state handles are positive (0: every handle)");
  return guarded([&]() -> sdx_status {
    world->impl->release_state(handle);
    return SDX_OK;
  });
}

sdx_status sdx_is_equal_state(sdx_world* world, int32_t a, int32_t b, int32_t* equal) {
  SDX_REQUIRE(world && equal, "NULL argument");
  return guarded([&]() -> sdx_status {
    *equal = world->impl->equal_states(a, b) ? 1 : 0;
    return SDX_OK;
  });
}

sdx_status sdx_state_size(sdx_world* world, int64_t* bytes) {
  SDX_REQUIRE(world && bytes, "NULL argument");
  return guarded([&]() -> sdx_status {
    *bytes = int64_t(world->impl->state_size());
    return SDX_OK;
  });
}

sdx_status sdx_capture_state_to_bytes(sdx_world* world, uint8_t* out, int64_t capacity) {
  SDX_REQUIRE(world && out, "NULL argument");
  SDX_REQUIRE(capacity >= 0, "capacity must not be negative");
  return guarded([&]() -> sdx_status {
    world->impl->capture_bytes(out, size_t(capacity));
    return SDX_OK;
  });
}

sdx_status sdx_restore_state_from_bytes(sdx_world* world, const uint8_t* data, int64_t size) {
  return sdx_restore_partial_state_from_bytes(world, data, size, 0);
}

sdx_status sdx_restore_partial_state_from_bytes(sdx_world* world, const uint8_t* data, int64_t size,
                                                uint32_t excluded) {
  SDX_REQUIRE(world && data, "NULL argument");
  SDX_REQUIRE(size >= 0, "size must not be negative");
  SDX_REQUIRE((excluded & ~uint32_t(SDX_STATE_DISPLACEMENTS | SDX_STATE_VELOCITIES | SDX_STATE_WARM_START |
                                     SDX_STATE_COLLIDER_TARGETS)) == 0,
              "excluded holds bits that name no state component");
  return guarded([&]() -> sdx_status {
    world->impl->restore_bytes(data, size_t(size), excluded);
    return SDX_OK;
  });
}

sdx_status sdx_capture_state_to_file(sdx_world* world, const char* path) {
  SDX_REQUIRE(world && path, "NULL argument");
  return guarded([&]() -> sdx_status {
    std::vector<uint8_t> b(world->impl->state_size());
    world->impl->capture_bytes(b.data(), b.size());
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(SDX_ERR_INVALID_ARGUMENT, "cannot open the state file for writing");
    const size_t put = std::fwrite(b.data(), 1, b.size(), f);
    const int closed = std::fclose(f);
    if (put != b.size() || closed != 0) return set_error(SDX_ERR_STATE, "writing the state file failed");
    return SDX_OK;
  });
}

sdx_status sdx_is_equal_state_bytes(const uint8_t* a, int64_t size_a, const uint8_t* b, int64_t size_b,
                                    int32_t* equal) {
  SDX_REQUIRE(a && b && equal, "NULL argument");
  SDX_REQUIRE(size_a >= 0 && size_b >= 0, "size must not be negative");
  return guarded([&]() -> sdx_status {
    *equal = sdx::capture::equal(a, size_t(size_a), b, size_t(size_b)) ? 1 : 0;
    return SDX_OK;
  });
}

static sdx_status json_out(const std::string& j, char* out, int64_t capacity, int64_t* length) {
  *length = int64_t(j.size());
  if (!out && capacity == 0) return SDX_OK;
  SDX_REQUIRE(out, "NULL output buffer with a nonzero capacity");
  SDX_REQUIRE(capacity > int64_t(j.size()), "JSON buffer too small (a call with out = NULL gives the length)");
  std::memcpy(out, j.c_str(), j.size() + 1);
  return SDX_OK;
}

sdx_status sdx_state_bytes_to_json(const uint8_t* data, int64_t size, int32_t pretty, char* out, int64_t capacity,
                                   int64_t* length) {
  SDX_REQUIRE(data && length, "NULL argument");
  SDX_REQUIRE(size >= 0 && capacity >= 0, "sizes must not be negative");
  return guarded([&]() -> sdx_status {
    return json_out(sdx::capture::to_json(data, size_t(size), pretty != 0), out, capacity, length);
  });
}

sdx_status sdx_capture_state_to_json(sdx_world* world, int32_t pretty, char* out, int64_t capacity,
                                     int64_t* length) {
  SDX_REQUIRE(world && length, "NULL argument");
  SDX_REQUIRE(capacity >= 0, "capacity must not be negative");
  return guarded([&]() -> sdx_status {
    std::vector<uint8_t> b(world->impl->state_size());
    world->impl->capture_bytes(b.data(), b.size());
    return json_out(sdx::capture::to_json(b.data(), b.size(), pretty != 0), out, capacity, length);
  });
}

sdx_status sdx_world_info(sdx_world* world, int64_t out[8]) {
  SDX_REQUIRE(world && out, "NULL argument");
  return guarded([&]() -> sdx_status {
    world->impl->info(out);
    return SDX_OK;
  });
}

sdx_status sdx_actor_bsr_size(sdx_world* world, int32_t actor, int32_t* n_block_rows, int32_t* nnzb) {
  SDX_REQUIRE(world, "NULL argument");
  return guarded([&]() -> sdx_status {
    world->impl->actor_bsr_size(actor, n_block_rows, nnzb);
    return SDX_OK;
  });
}

sdx_status sdx_assemble(sdx_world* world, int32_t actor, double dt, const sdx_solver_params* params,
                        double* out_residual, int32_t* out_row_ptr, int32_t* out_col_idx, double* out_values) {
  SDX_REQUIRE(world && params, "NULL argument");
  SDX_REQUIRE(dt > 0, "dt must be positive");
  return guarded([&]() -> sdx_status {
    world->impl->assemble_actor(actor, dt, *params, out_residual, out_row_ptr, out_col_idx, out_values);
    return SDX_OK;
  });
}

sdx_status sdx_objective(sdx_world* world, int32_t actor, double dt, const sdx_solver_params* params,
                         double* out_objective) {
  SDX_REQUIRE(world && params && out_objective, "NULL argument");
  SDX_REQUIRE(dt > 0, "dt must be positive");
  return guarded([&]() -> sdx_status {
    *out_objective = world->impl->objective_actor(actor, dt, *params);
    return SDX_OK;
  });
}

sdx_status sdx_kernel_timing(sdx_world* world, int32_t kernel, int32_t enable, double out[2]) {
  SDX_REQUIRE(world, "NULL argument");
  SDX_REQUIRE(kernel >= 0 && kernel < 5, "kernel class must be 0..4");
  return guarded([&]() -> sdx_status {
    if (enable >= 0) world->impl->timing_enable(kernel, enable);
    if (out) {
      auto r = world->impl->timing_read(kernel);
      out[0] = r.first;
      out[1] = r.second;
    }
    return SDX_OK;
  });
}

sdx_status sdx_set_cg_burst(sdx_world* world, int32_t burst) {
  SDX_REQUIRE(world && burst > 0 && burst <= 256, "burst must be in 1..256");
  sdx_schedule sc = world->impl->schedule();
  sc.cg_burst = burst;
  return guarded([&]() -> sdx_status {
    world->impl->set_schedule(sc);
    return SDX_OK;
  });
}

sdx_status sdx_set_schedule(sdx_world* world, const sdx_schedule* sc) {
  SDX_REQUIRE(world && sc, "NULL argument");
  SDX_REQUIRE(sc->cg_burst > 0 && sc->cg_burst <= 256, "cg_burst must be in 1..256");
  SDX_REQUIRE(sc->use_graphs == 0 || sc->use_graphs == 1, "use_graphs must be 0 or 1");
  SDX_REQUIRE(sc->spmv_lanes == 4 || sc->spmv_lanes == 8 || sc->spmv_lanes == 16, "spmv_lanes must be 4, 8 or 16");
  SDX_REQUIRE(sc->element_block == 64 || sc->element_block == 128 || sc->element_block == 256,
              "element_block must be 64, 128 or 256");
  SDX_REQUIRE(sc->active_lists == 0 || sc->active_lists == 1, "active_lists must be 0 or 1");
  SDX_REQUIRE(sc->cg_groups >= 0 && sc->cg_groups <= 256, "cg_groups must be in 0..256");
  SDX_REQUIRE(sc->group_iters >= 0 && sc->group_iters <= 4096, "group_iters must be in 0..4096");
  SDX_REQUIRE(sc->fused_tail == 0 || sc->fused_tail == 1, "fused_tail must be 0 or 1");
  SDX_REQUIRE(sc->asm_groups >= 0 && sc->asm_groups <= 4096, "asm_groups must be in 0..4096");
  SDX_REQUIRE(sc->fused_asm == 0 || sc->fused_asm == 1, "fused_asm must be 0 or 1");
  SDX_REQUIRE(sc->env_cg >= 0 && sc->env_cg <= 2, "env_cg must be 0, 1 or 2");
  return guarded([&]() -> sdx_status {
    world->impl->set_schedule(*sc);
    return SDX_OK;
  });
}

sdx_status sdx_get_schedule(sdx_world* world, sdx_schedule* out) {
  SDX_REQUIRE(world && out, "NULL argument");
  *out = world->impl->schedule();
  return SDX_OK;
}

sdx_status sdx_autotune(sdx_world* world, double dt, const sdx_solver_params* params, int32_t trials,
                        sdx_schedule* out, double* out_ms) {
  SDX_REQUIRE(world && params, "NULL argument");
  SDX_REQUIRE(dt > 0, "dt must be positive");
  SDX_REQUIRE(trials >= 1 && trials <= 16, "trials must be in 1..16");
  const sdx_status v = validate_solver_params(*params);
  if (v != SDX_OK) return v;
  return guarded([&]() -> sdx_status {
    const double ms = world->impl->autotune(dt, *params, trials);
    if (out) *out = world->impl->schedule();
    if (out_ms) *out_ms = ms;
    return SDX_OK;
  });
}

void sdx_default_train_params(sdx_train_params* p) {
  p->epochs = 10;
  p->batch = 4096;
  p->lr = 1e-3f;
  p->beta1 = 0.9f;
  p->beta2 = 0.999f;
  p->eps = 1e-8f;
  p->seed = 1;
}

static sdx_status check_train(const sdx_train_params* p) {
  SDX_REQUIRE(p, "NULL train params");
  SDX_REQUIRE(p->epochs > 0 && p->epochs <= 100000, "epochs must be in 1..100000");
  SDX_REQUIRE(p->batch > 0 && p->batch <= (1 << 20), "batch must be in 1..2^20");
  SDX_REQUIRE(p->lr > 0 && p->beta1 >= 0 && p->beta1 < 1 && p->beta2 >= 0 && p->beta2 < 1 && p->eps > 0,
              "invalid Adam parameters");
  return SDX_OK;
}

static sdx::TrainConfig train_config(const sdx_train_params* p) {
  sdx::TrainConfig c;
  c.epochs = p->epochs;
  c.batch = p->batch;
  c.lr = p->lr;
  c.beta1 = p->beta1;
  c.beta2 = p->beta2;
  c.eps = p->eps;
  c.seed = p->seed;
  return c;
}

sdx_status sdx_mlp_create(int32_t device, uint64_t seed, sdx_mlp** out) {
  SDX_REQUIRE(out, "out is NULL");
  return guarded([&]() -> sdx_status {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
      return set_error(SDX_ERR_DEVICE, "no such HIP device");
    auto m = std::make_unique<sdx_mlp>();
    m->impl = std::make_unique<sdx::Mlp>(device, seed);
    *out = m.release();
    return SDX_OK;
  });
}

sdx_status sdx_mlp_destroy(sdx_mlp* mlp) {
  delete mlp;
  return SDX_OK;
}

sdx_status sdx_mlp_set_params(sdx_mlp* mlp, const float* params, int32_t n) {
  SDX_REQUIRE(mlp && params, "NULL argument");
  SDX_REQUIRE(n == SDX_MLP_PARAMS, "n must be SDX_MLP_PARAMS");
  for (int32_t i = 0; i < n; ++i) SDX_REQUIRE(std::isfinite(params[i]), "parameters must be finite");
  return guarded([&]() -> sdx_status {
    mlp->impl->set_params(params);
    return SDX_OK;
  });
}

sdx_status sdx_mlp_get_params(sdx_mlp* mlp, float* params, int32_t n) {
  SDX_REQUIRE(mlp && params, "NULL argument");
  SDX_REQUIRE(n == SDX_MLP_PARAMS, "n must be SDX_MLP_PARAMS");
  return guarded([&]() -> sdx_status {
    mlp->impl->get_params(params);
    return SDX_OK;
  });
}

sdx_status sdx_mlp_forward(sdx_mlp* mlp, const float* x, int64_t n_rows, int32_t bf16, float* y) {
  SDX_REQUIRE(mlp && x && y, "NULL argument");
  SDX_REQUIRE(n_rows > 0 && n_rows < (int64_t(1) << 31), "n_rows must be in 1..2^31-1");
  return guarded([&]() -> sdx_status {
    mlp->impl->forward(x, n_rows, bf16 != 0, y);
    return SDX_OK;
  });
}

sdx_status sdx_mlp_loss_grad(sdx_mlp* mlp, const float* x, const float* t, int32_t n_rows, float* grad,
                             double* loss) {
  SDX_REQUIRE(mlp && x && t, "NULL argument");
  SDX_REQUIRE(n_rows > 0 && n_rows <= (1 << 22), "n_rows must be in 1..2^22");
  return guarded([&]() -> sdx_status {
    const double l = mlp->impl->loss_grad(x, t, n_rows, grad);
    if (loss) *loss = l;
    return SDX_OK;
  });
}

sdx_status sdx_mlp_train(sdx_mlp* mlp, const float* x, const float* t, int64_t n_rows,
                         const sdx_train_params* params, double* epoch_loss) {
  SDX_REQUIRE(mlp && x && t, "NULL argument");
  SDX_REQUIRE(n_rows > 0 && n_rows < (int64_t(1) << 31), "n_rows must be in 1..2^31-1");
  const sdx_status v = check_train(params);
  if (v != SDX_OK) return v;
  return guarded([&]() -> sdx_status {
    sdx::hip_check(hipSetDevice(mlp->impl->device()), "hipSetDevice");
    float *dx = nullptr, *dt = nullptr;
    sdx::hip_check(hipMalloc(reinterpret_cast<void**>(&dx), sizeof(float) * size_t(n_rows) * SDX_MLP_ROW),
                   "hipMalloc");
    const hipError_t e2 = hipMalloc(reinterpret_cast<void**>(&dt), sizeof(float) * size_t(n_rows) * SDX_MLP_OUT);
    if (e2 != hipSuccess) {
      (void)hipFree(dx);
      sdx::hip_check(e2, "hipMalloc");
    }
    std::vector<double> losses;
    try {
      sdx::hip_check(hipMemcpy(dx, x, sizeof(float) * size_t(n_rows) * SDX_MLP_ROW, hipMemcpyHostToDevice), "memcpy");
      sdx::hip_check(hipMemcpy(dt, t, sizeof(float) * size_t(n_rows) * SDX_MLP_OUT, hipMemcpyHostToDevice), "memcpy");
      losses = mlp->impl->train(dx, dt, n_rows, train_config(params));
    } catch (...) {
      (void)hipFree(dx);
      (void)hipFree(dt);
      throw;
    }
    (void)hipFree(dx);
    (void)hipFree(dt);
    if (epoch_loss)
      for (size_t i = 0; i < losses.size(); ++i) epoch_loss[i] = losses[i];
    return SDX_OK;
  });
}

sdx_status sdx_mlp_train_collected(sdx_mlp* mlp, sdx_world* world, const sdx_train_params* params,
                                   double* epoch_loss) {
  SDX_REQUIRE(mlp && world, "NULL argument");
  SDX_REQUIRE(mlp->impl->device() == world->impl->device(), "model and world must be on the same device");
  const sdx_status v = check_train(params);
  if (v != SDX_OK) return v;
  return guarded([&]() -> sdx_status {
    const float *x = nullptr, *t = nullptr;
    int64_t n = 0;
    world->impl->collected_device(&x, &t, &n);
    SDX_REQUIRE(n > 0, "the world has collected no samples");
    const std::vector<double> losses = mlp->impl->train(x, t, n, train_config(params));
    if (epoch_loss)
      for (size_t i = 0; i < losses.size(); ++i) epoch_loss[i] = losses[i];
    return SDX_OK;
  });
}

sdx_status sdx_world_set_warm_start_model(sdx_world* world, sdx_mlp* mlp) {
  SDX_REQUIRE(world, "NULL argument");
  SDX_REQUIRE(!mlp || mlp->impl->device() == world->impl->device(), "model and world must be on the same device");
  return guarded([&]() -> sdx_status {
    if (!mlp) {
      world->impl->set_ws_params(nullptr);
      return SDX_OK;
    }
    std::vector<float> p(SDX_MLP_PARAMS);
    mlp->impl->get_params(p.data());
    world->impl->set_ws_params(p.data());
    return SDX_OK;
  });
}

sdx_status sdx_world_collect(sdx_world* world, int64_t capacity, int32_t per_env) {
  SDX_REQUIRE(world, "NULL argument");
  SDX_REQUIRE(capacity >= 0 && capacity < (int64_t(1) << 31), "capacity must be in 0..2^31-1");
  SDX_REQUIRE(capacity == 0 || (per_env > 0 && per_env <= 65536), "per_env must be in 1..65536");
  return guarded([&]() -> sdx_status {
    world->impl->collect(capacity, per_env);
    return SDX_OK;
  });
}

sdx_status sdx_world_collected(sdx_world* world, int64_t* n_rows, float* x, float* t, int64_t cap_rows) {
  SDX_REQUIRE(world && n_rows, "NULL argument");
  SDX_REQUIRE(cap_rows >= 0, "cap_rows must be non-negative");
  return guarded([&]() -> sdx_status {
    *n_rows = world->impl->collected(x, t, cap_rows);
    return SDX_OK;
  });
}

}  // extern "C"
