// Host side of the MI355X soft-body FEM step: world set-up, the Newton /
// line-search / CG drivers, and the C ABI of include/sdx_fem.h.
//
// Reference call surface replaced (paths under superdex_physics/libraries/mochi/):
//   Scene::Step ............... mochi_physics/src/mochi_scene.cpp:594
//   StepIslandNewtonAsync ..... mochi_physics/src/mochi_solve.cpp:1090-1160
//   NewtonSolver<T>::Solve .... mochi_core/src/solvers/newton_solver.cpp:603-876
//   LineSearch::ResidualNorm .. mochi_core/include/mochi_core/solvers/line_search.h:300-335
//   LinearSolver::Solve (CG) .. mochi_core/include/mochi_core/solvers/linear_solver.h:321-496
//   soft stage integration .... mochi_physics/src/mochi_soft.cpp:94-163
#include <chrono>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sdx_fem.h"
#include "sdx_kernels.h"
#include "sdx_maps.h"
#include <type_traits>

#include "sdx_world.h"
#include "sdx_capture.h"
#include "sdx_rowblocks.h"

namespace sdx {

thread_local std::string g_last_error;

sdx_status set_error(sdx_status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ set-up helpers

namespace {

// Boundary triangles: tet faces used once, oriented with the normal away from
// the opposite vertex.
void BuildBoundary(const std::vector<double>& X, const std::vector<int32_t>& tets, std::vector<int32_t>& tris) {
  struct Face {
    std::array<int32_t, 3> key, nodes;
    int32_t opp;
  };
  const int nT = int(tets.size() / 4);
  static const int fv[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  std::vector<Face> faces(size_t(4) * size_t(nT));
  for (int t = 0; t < nT; ++t)
    for (int f = 0; f < 4; ++f) {
      Face& F = faces[size_t(4 * t + f)];
      for (int i = 0; i < 3; ++i) F.nodes[size_t(i)] = tets[size_t(4 * t + fv[f][i])];
      F.opp = tets[size_t(4 * t + f)];
      F.key = F.nodes;
      std::sort(F.key.begin(), F.key.end());
    }
  std::sort(faces.begin(), faces.end(), [](const Face& a, const Face& b) { return a.key < b.key; });
  for (size_t i = 0; i < faces.size();) {
    size_t j = i;
    while (j < faces.size() && faces[j].key == faces[i].key) ++j;
    if (j - i == 1) {
      Face F = faces[i];
      double e1[3], e2[3], w[3];
      for (int d = 0; d < 3; ++d) {
        e1[d] = X[size_t(3 * F.nodes[1] + d)] - X[size_t(3 * F.nodes[0] + d)];
        e2[d] = X[size_t(3 * F.nodes[2] + d)] - X[size_t(3 * F.nodes[0] + d)];
        w[d] = X[size_t(3 * F.opp + d)] - X[size_t(3 * F.nodes[0] + d)];
      }
      const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                           e1[0] * e2[1] - e1[1] * e2[0]};
      if (n[0] * w[0] + n[1] * w[1] + n[2] * w[2] > 0) std::swap(F.nodes[1], F.nodes[2]);
      tris.insert(tris.end(), F.nodes.begin(), F.nodes.end());
    }
    i = j;
  }
}

void BuildTiles(const std::vector<int32_t>& env_v0, int nEnv, int rows_per_tile, HostTiles& t) {
  t.env.clear();
  t.row0.clear();
  t.row1.clear();
  t.first.assign(size_t(nEnv) + 1, 0);
  for (int e = 0; e < nEnv; ++e) {
    t.first[size_t(e)] = int32_t(t.env.size());
    for (int r = env_v0[size_t(e)]; r < env_v0[size_t(e) + 1]; r += rows_per_tile) {
      t.env.push_back(e);
      t.row0.push_back(r);
      t.row1.push_back(std::min(r + rows_per_tile, env_v0[size_t(e) + 1]));
    }
  }
  t.first[size_t(nEnv)] = int32_t(t.env.size());
}

}  // namespace

// ------------------------------------------------------------------ World<T>

template <class T>
World<T>::World(int device, int precision) : WorldBase(device, precision) {}

template <class T>
World<T>::~World() {
  release_device();
}

template <class T>
void World<T>::release_device() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  snaps_.clear();  // snapshots hold the old topology's sizes
  drop_graphs();
  tuned_ = false;
  contact_ready_ = false;
  rb_ready_ = false;
  for (auto* b : bufs_) b->free();
  finalized_ = false;
}

template <class T>
void World<T>::pull_state() {
  if (!finalized_) return;
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  std::vector<T> hu(size_t(3 * nV_)), hv(size_t(3 * nV_));
  u_.download(hu);
  v_.download(hv);
  for (size_t a = 0; a < actors_.size(); ++a) {
    auto& A = actors_[a];
    for (int k = 0; k < 3 * A.nV; ++k) {
      A.u[size_t(k)] = double(hu[size_t(3 * env_v0_[a] + k)]);
      A.v[size_t(k)] = double(hv[size_t(3 * env_v0_[a] + k)]);
    }
  }
}

template <class T>
void World<T>::push_state() {
  std::vector<T> hu(size_t(3 * nV_)), hv(size_t(3 * nV_));
  for (size_t a = 0; a < actors_.size(); ++a) {
    auto& A = actors_[a];
    for (int k = 0; k < 3 * A.nV; ++k) {
      hu[size_t(3 * env_v0_[a] + k)] = T(A.u[size_t(k)]);
      hv[size_t(3 * env_v0_[a] + k)] = T(A.v[size_t(k)]);
    }
  }
  u_.upload(hu);
  v_.upload(hv);
}

template <class T>
void World<T>::finalize() {
  if (finalized_) return;
  if (actors_.empty()) throw std::invalid_argument("world has no soft actors");
  hip_check(hipSetDevice(device_), "hipSetDevice");
  const int nEnv = int(actors_.size());
  nEnv_ = nEnv;
  stiff_damp_ = false;
  env_v0_.assign(size_t(nEnv) + 1, 0);
  std::vector<int32_t> env_t0(size_t(nEnv) + 1, 0);
  for (int e = 0; e < nEnv; ++e) {
    env_v0_[size_t(e) + 1] = env_v0_[size_t(e)] + actors_[size_t(e)].nV;
    env_t0[size_t(e) + 1] = env_t0[size_t(e)] + actors_[size_t(e)].nT;
  }
  nV_ = env_v0_.back();
  nT_ = env_t0.back();
  // global pattern, tets, geometry
  std::vector<int32_t> row_ptr(1, 0), col, slot_row, diag_slot(static_cast<size_t>(nV_)), row_env(static_cast<size_t>(nV_));
  std::vector<int32_t> tets(size_t(4) * size_t(nT_)), tet_env(static_cast<size_t>(nT_));
  std::vector<T> geom(size_t(kGeomPlanes) * size_t(nT_));
  std::vector<T> X(size_t(3) * size_t(nV_));
  std::vector<uint8_t> pinned(size_t(3) * size_t(nV_), 0);
  std::vector<T> pin_value(size_t(3) * size_t(nV_), T(0));
  std::vector<int32_t> tris, tri_env;
  std::vector<int32_t> edge_keys;
  std::vector<uint32_t> edge_codes;
  std::vector<int32_t> vtx_keys;
  std::vector<uint32_t> vtx_codes;
  std::vector<double> contact(size_t(kContactParams) * size_t(nEnv));
  PatchBuilder patches;
  env_p0_.assign(size_t(nEnv) + 1, 0);
  for (int e = 0; e < nEnv; ++e) {
    env_p0_[size_t(e)] = patches.count();
    auto& A = actors_[size_t(e)];
    const int v0 = env_v0_[size_t(e)], t0 = env_t0[size_t(e)];
    Pattern P = BuildPattern(A.nV, A.tets);
    const int b0 = int(col.size());
    for (int i = 0; i < A.nV; ++i) {
      row_env[size_t(v0 + i)] = e;
      for (int k = P.ptr[size_t(i)]; k < P.ptr[size_t(i) + 1]; ++k) {
        col.push_back(v0 + P.col[size_t(k)]);
        slot_row.push_back(v0 + i);
        if (P.col[size_t(k)] == i) diag_slot[size_t(v0 + i)] = b0 + k;
      }
      row_ptr.push_back(int32_t(col.size()));
    }
    for (int k = 0; k < 3 * A.nV; ++k) {
      X[size_t(3 * v0 + k)] = T(A.X[size_t(k)]);
      pinned[size_t(3 * v0 + k)] = A.pinned[size_t(k)];
      pin_value[size_t(3 * v0 + k)] = T(A.pin_value[size_t(k)]);
    }
    const double lam0 = A.params.youngs_modulus * A.params.poisson_ratio /
                        ((1 + A.params.poisson_ratio) * (1 - 2 * A.params.poisson_ratio));
    const double mu0 = A.params.youngs_modulus / (2 * (1 + A.params.poisson_ratio));
    for (int t = 0; t < A.nT; ++t) {
      const int gt = t0 + t;
      tet_env[size_t(gt)] = e;
      const int32_t* tv = &A.tets[size_t(4 * t)];
      for (int a = 0; a < 4; ++a) tets[size_t(4 * gt + a)] = v0 + tv[a];
      // dMap columns X_f - X_0, dBasis = dMap^{-T} dN_param (finite_element.h:95-125)
      double D[9];
      for (int d = 0; d < 3; ++d)
        for (int f = 0; f < 3; ++f)
          D[d * 3 + f] = A.X[size_t(3 * tv[f + 1] + d)] - A.X[size_t(3 * tv[0] + d)];
      const double det = D[0] * (D[4] * D[8] - D[5] * D[7]) - D[1] * (D[3] * D[8] - D[5] * D[6]) +
                         D[2] * (D[3] * D[7] - D[4] * D[6]);
      if (!(det > 0)) throw std::invalid_argument("tet with non-positive volume in actor " + std::to_string(e));
      const double cof[9] = {D[4] * D[8] - D[5] * D[7], D[5] * D[6] - D[3] * D[8], D[3] * D[7] - D[4] * D[6],
                             D[2] * D[7] - D[1] * D[8], D[0] * D[8] - D[2] * D[6], D[1] * D[6] - D[0] * D[7],
                             D[1] * D[5] - D[2] * D[4], D[2] * D[3] - D[0] * D[5], D[0] * D[4] - D[1] * D[3]};
      // dN_f (f = 1..3) = row f-1 of dMap^{-1} = column of cof / det
      for (int f = 0; f < 3; ++f)
        for (int r = 0; r < 3; ++r) geom[size_t(kGdN + 3 * f + r) * size_t(nT_) + size_t(gt)] = T(cof[r * 3 + f] / det);
      const double vol = det / 6.0;
      double mu = mu0, lam = lam0;
      if (!A.E.empty()) {
        const double E = A.E[size_t(t)], nu = A.nu[size_t(t)];
        lam = E * nu / ((1 + nu) * (1 - 2 * nu));
        mu = E / (2 * (1 + nu));
      }
      geom[size_t(kGVol) * size_t(nT_) + size_t(gt)] = T(vol);
      geom[size_t(kGMu) * size_t(nT_) + size_t(gt)] = T(mu);
      geom[size_t(kGLambda) * size_t(nT_) + size_t(gt)] = T(lam);
      geom[size_t(kGMass) * size_t(nT_) + size_t(gt)] = T(A.params.density * vol);
      geom[size_t(kGAlpha) * size_t(nT_) + size_t(gt)] = T(A.params.mass_damping);
      geom[size_t(kGBeta) * size_t(nT_) + size_t(gt)] = T(A.params.stiffness_damping);
      geom[size_t(kGBetaGeo) * size_t(nT_) + size_t(gt)] = T(A.params.stiffness_damping_geometric_term ? 1 : 0);
      if (A.params.stiffness_damping > 0) stiff_damp_ = true;
    }
    // element patches: runs of kPatchTets consecutive tets, each with its
    // local vertices and vertex pairs (upper slots), sorted, one record each
    for (int pt0 = 0; pt0 < A.nT; pt0 += kPatchTets)
      patches.add(P, b0, v0, t0 + pt0, &A.tets[size_t(4 * pt0)], std::min(kPatchTets, A.nT - pt0), vtx_keys,
                  vtx_codes, edge_keys, edge_codes);
    std::vector<int32_t> atris;
    BuildBoundary(A.X, A.tets, atris);
    const int f0 = int(tri_env.size());
    for (size_t f = 0; f < atris.size() / 3; ++f) {
      const int gf = f0 + int(f);
      tri_env.push_back(e);
      // contact records join the element records of the same vertex / pair
      // (flag bits 31 / 30 mark contact codes)
      for (int a = 0; a < 3; ++a) {
        tris.push_back(v0 + atris[3 * f + size_t(a)]);
        vtx_keys.push_back(v0 + atris[3 * f + size_t(a)]);
        vtx_codes.push_back((uint32_t(gf) * 3u + uint32_t(a)) | 0x80000000u);
      }
      for (int k = 0; k < 3; ++k) {
        const int va = atris[3 * f + size_t(kTriPairA[k])], vb = atris[3 * f + size_t(kTriPairB[k])];
        edge_keys.push_back(b0 + FindSlot(P, std::min(va, vb), std::max(va, vb)));
        edge_codes.push_back((uint32_t(gf) * 3u + uint32_t(k)) | 0x40000000u);
      }
    }
    double* cp = &contact[size_t(kContactParams) * size_t(e)];
    cp[0] = A.params.penalty_coefficient;
    cp[1] = A.params.penalty_smoothing_half_distance;
    cp[2] = A.params.penalty_threshold;
    cp[3] = A.params.max_alignment_normals;
    cp[4] = A.params.coulomb_friction_coefficient;
    cp[5] = A.params.viscous_friction_coefficient;
    cp[6] = A.params.friction_falloff_vel;
    cp[7] = A.params.normal_viscous_damping_coefficient;
  }
  env_p0_[size_t(nEnv)] = patches.count();
  nB_ = int(col.size());
  nF_ = int(tri_env.size());
  h_row_ptr_ = row_ptr;
  h_col_ = col;
  if (int64_t(nT_) * 6 + int64_t(nF_) * 3 >= (int64_t(1) << 30) || nV_ >= (1 << kItemRowBits))
    throw std::invalid_argument("mesh too large for 32-bit record codes");
  std::vector<int32_t> eptr, vtx_ptr;
  std::vector<uint32_t> ecode, vtx_code;
  CountingCsr(nB_, edge_keys, edge_codes, eptr, ecode);
  CountingCsr(nV_, vtx_keys, vtx_codes, vtx_ptr, vtx_code);
  // per-environment starts of tets, triangles and records (assembly groups)
  env_t0_ = env_t0;
  env_f0_.assign(size_t(nEnv) + 1, nF_);
  for (int f = nF_ - 1; f >= 0; --f) env_f0_[size_t(tri_env[size_t(f)])] = f;
  for (int e = nEnv - 1; e >= 0; --e) env_f0_[size_t(e)] = std::min(env_f0_[size_t(e)], env_f0_[size_t(e) + 1]);
  env_rv0_.assign(size_t(nEnv) + 1, 0);
  env_re0_.assign(size_t(nEnv) + 1, 0);
  for (int e = 0; e <= nEnv; ++e) {
    env_rv0_[size_t(e)] = vtx_ptr[size_t(env_v0_[size_t(e)])];
    env_re0_[size_t(e)] = eptr[size_t(row_ptr[size_t(env_v0_[size_t(e)])])];
  }
  // record of every patch vertex / pair and contact contribution: the
  // vertex-sorted position of (patch, local vertex) / (triangle, a) for
  // residuals and diagonal blocks, the pair-sorted position of (patch, local
  // pair) / (triangle, pair) for off-diagonal blocks
  std::vector<int32_t> pos_lv(patches.n_local_vertices()), pos_vc(size_t(3) * size_t(nF_));
  for (size_t q = 0; q < vtx_code.size(); ++q) {
    const uint32_t code = vtx_code[q];
    if (code >> 31)
      pos_vc[size_t(code & 0x7fffffffu)] = int32_t(q);
    else
      pos_lv[size_t(code)] = int32_t(q);
  }
  std::vector<int32_t> pos_le(patches.n_local_pairs()), pos_ec(size_t(3) * size_t(nF_));
  for (size_t q = 0; q < ecode.size(); ++q) {
    const uint32_t code = ecode[q];
    if (code & 0x40000000u)
      pos_ec[size_t(code & 0x3fffffffu)] = int32_t(q);
    else
      pos_le[size_t(code)] = int32_t(q);
  }
  std::vector<int32_t> patch_t0, patch_meta_ptr, patch_meta_mid;
  std::vector<uint32_t> patch_meta;
  patches.finish(pos_lv, pos_le, patch_t0, patch_meta_ptr, patch_meta_mid, patch_meta);
  nP_ = patches.count();
  // gather description of every BSR slot: its record range (the vertex's
  // diagonal-block records; an upper slot's pair records; a lower slot's
  // mirror's, read transposed), row with the Dirichlet flags of its row and
  // column DoFs, environment and kind
  std::vector<int32_t> slot_desc(size_t(4) * size_t(nB_));
  for (int i = 0; i < nV_; ++i)
    for (int k = row_ptr[size_t(i)]; k < row_ptr[size_t(i) + 1]; ++k) {
      const int j = col[size_t(k)];
      int r0, r1, kind;
      if (j == i) {
        r0 = vtx_ptr[size_t(i)];
        r1 = vtx_ptr[size_t(i) + 1];
        kind = kSlotDiag;
      } else if (j > i) {
        r0 = eptr[size_t(k)];
        r1 = eptr[size_t(k) + 1];
        kind = kSlotUpper;
      } else {
        auto b = col.begin() + row_ptr[size_t(j)], e2 = col.begin() + row_ptr[size_t(j) + 1];
        const size_t mirror = size_t(std::lower_bound(b, e2, i) - col.begin());  // slot (j, i)
        r0 = eptr[mirror];
        r1 = eptr[mirror + 1];
        kind = kSlotLower;
      }
      int pins = 0;
      for (int d = 0; d < 3; ++d)
        pins |= (pinned[size_t(3 * i + d)] ? 1 : 0) << d | (pinned[size_t(3 * j + d)] ? 1 : 0) << (3 + d);
      int32_t* sd = &slot_desc[size_t(4) * size_t(k)];
      sd[0] = r0;
      sd[1] = r1;
      sd[2] = i | (pins << kItemRowBits);
      sd[3] = int32_t(uint32_t(row_env[size_t(i)]) | (uint32_t(kind) << 30));
    }
  BuildTiles(env_v0_, nEnv, kVecThreads, tiles_v_);
  BuildTiles(env_v0_, nEnv, kSpmvThreads / sched_.spmv_lanes, tiles_a_);
  max_env_a_ = 0;
  group_plan_n_ = 0;

  // upload
  d_tets_.upload(tets);
  d_tet_env_.upload(tet_env);
  d_tris_.upload(tris);
  d_tri_env_.upload(tri_env);
  d_row_env_.upload(row_env);
  d_row_ptr_.upload(row_ptr);
  d_col_.upload(col);
  {  // the CG matvec reads 16-bit column offsets when every block's fits
    std::vector<int16_t> c16(col.size());
    bool fits = true;
    for (size_t i = 0; i + 1 < row_ptr.size() && fits; ++i)
      for (int32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
        const int64_t d = int64_t(col[size_t(k)]) - int64_t(i);
        if (d < -32768 || d > 32767) {
          fits = false;
          break;
        }
        c16[size_t(k)] = int16_t(d);
      }
    if (fits && !c16.empty())
      d_col16_.upload(c16);
    else
      d_col16_.free();
  }
  d_slot_row_.upload(slot_row);
  d_diag_slot_.upload(diag_slot);
  d_eptr_.upload(eptr);
  d_slot_desc_.upload(slot_desc);
  d_patch_t0_.upload(patch_t0);
  d_patch_meta_ptr_.upload(patch_meta_ptr);
  d_patch_meta_mid_.upload(patch_meta_mid);
  d_patch_meta_.upload(patch_meta);
  patch_list_.alloc(std::max<size_t>(size_t(nP_), 1));
  d_pos_vc_.upload(pos_vc);
  d_pos_ec_.upload(pos_ec);
  d_vtx_ptr_.upload(vtx_ptr);
  d_pinned_.upload(pinned);
  d_pin_value_.upload(pin_value);
  d_contact_.upload(contact);
  d_X_.upload(X);
  d_geom_.upload(geom);
  upload_tiles(tiles_v_, dt_v_);
  upload_tiles(tiles_a_, dt_a_);
  alloc_lists();
  const size_t n3 = size_t(3) * size_t(nV_);
  u_.alloc(n3);
  v_.alloc(n3);
  u0_.alloc(n3);
  v0_.alloc(n3);
  res_.alloc(n3);
  dx_.alloc(n3);
  xcg_.alloc(n3);
  xd_.alloc(n3);
  hip_check(hipMemsetAsync(xd_.p, 0, xd_.bytes(), stream_), "hipMemsetAsync");
  sol0_.alloc(n3);
  r_.alloc(n3);
  z_.alloc(size_t(4) * size_t(nV_));  // padded to 4 components per vertex
  p_.alloc(n3);
  sv_.alloc(n3);
  w_.alloc(n3);
  rref_.alloc(n3);
  rref2_.alloc(n3);
  rref3_.alloc(n3);
  asm_plan_n_ = 0;
  plan_assembly(asm_groups(sched_));  // record buffers and assembly group ranges
  vals_.alloc(size_t(9) * size_t(nB_));
  dinv_.alloc(size_t(6) * size_t(nV_));
  env_cg_.alloc(static_cast<size_t>(nEnv));
  acc_.alloc(size_t(kCgSets) * size_t(nEnv) * size_t(kSlots) * size_t(kSlotStride));
  solve_mask_.alloc(static_cast<size_t>(nEnv));
  warm_mask_.alloc(static_cast<size_t>(nEnv));
  list_count_.alloc(4);
  tail_bar_.alloc(1);
  tail_abort_.alloc(1);
  hist_.alloc(size_t(2 * kNewtonHist) * n3);
  hip_check(hipMemsetAsync(hist_.p, 0, hist_.bytes(), stream_), "hipMemsetAsync");
  ws_y_.alloc(n3);
  ws_F_.alloc(size_t(kMlpInPad) * size_t(nV_));
  ws_scale_.alloc(static_cast<size_t>(nEnv));
  ws_ratio_.alloc(static_cast<size_t>(nEnv));
  ws_ratio_hist_.alloc(static_cast<size_t>(nEnv));
  ws_alpha_.alloc(static_cast<size_t>(nEnv));
  ws_beta_.alloc(static_cast<size_t>(nEnv));
  ws_gamma_.alloc(static_cast<size_t>(nEnv));
  d_env_v0_.upload(env_v0_);
  {  // whole-solve CG queue: environments by rows, largest first (ties by index)
    std::vector<int32_t> order(static_cast<size_t>(nEnv));
    env_max_rows_ = 0;
    for (int e = 0; e < nEnv; ++e) {
      order[size_t(e)] = e;
      env_max_rows_ = std::max(env_max_rows_, env_v0_[size_t(e) + 1] - env_v0_[size_t(e)]);
    }
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
      return env_v0_[size_t(a) + 1] - env_v0_[size_t(a)] > env_v0_[size_t(b) + 1] - env_v0_[size_t(b)];
    });
    d_env_order_.upload(order);
    env_next_.alloc(1);
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&num_cus_, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    env_lds_limit_ = env_lds_limit();
  }
  num_active_.alloc(1);
  env_mask_.alloc(static_cast<size_t>(nEnv));
  ls_mask_.alloc(static_cast<size_t>(nEnv));
  ls_scale_.alloc(static_cast<size_t>(nEnv));
  reset_.alloc(static_cast<size_t>(nEnv));
  env_sq_.alloc(static_cast<size_t>(nEnv));
  norms2_.alloc(size_t(2) * size_t(nEnv));
  env_tol_.alloc(static_cast<size_t>(nEnv));
  env_obj_.alloc(static_cast<size_t>(nEnv));
  env_dot_.alloc(static_cast<size_t>(nEnv));
  res0_.alloc(n3);
  finalized_ = true;
  push_state();
  colliders_dirty_ = true;
  mesh_.nV = nV_;
  mesh_.nT = nT_;
  mesh_.nF = nF_;
  mesh_.nB = nB_;
  mesh_.nEnv = nEnv;
  mesh_.tets = d_tets_.p;
  mesh_.tet_env = d_tet_env_.p;
  mesh_.tris = d_tris_.p;
  mesh_.tri_env = d_tri_env_.p;
  mesh_.row_env = d_row_env_.p;
  mesh_.row_ptr = d_row_ptr_.p;
  mesh_.col = d_col_.p;
  mesh_.slot_row = d_slot_row_.p;
  mesh_.diag_slot = d_diag_slot_.p;
  mesh_.eptr = d_eptr_.p;
  mesh_.slot_desc = reinterpret_cast<const int4*>(d_slot_desc_.p);
  mesh_.nP = nP_;
  mesh_.patch_t0 = d_patch_t0_.p;
  mesh_.patch_meta_ptr = d_patch_meta_ptr_.p;
  mesh_.patch_meta_mid = d_patch_meta_mid_.p;
  mesh_.patch_meta = d_patch_meta_.p;
  mesh_.pos_vc = d_pos_vc_.p;
  mesh_.pos_ec = d_pos_ec_.p;
  mesh_.coll_flags = d_coll_flags_.p;
  mesh_.vtx_ptr = d_vtx_ptr_.p;
  mesh_.pinned = d_pinned_.p;
  mesh_.env_mask = env_mask_.p;
  fields_.X = d_X_.p;
  fields_.geom = d_geom_.p;
  fields_.contact = d_contact_.p;
  upload_colliders();
}

template <class T>
void World<T>::alloc_lists() {
  list_a_.alloc(std::max<size_t>(tiles_a_.env.size(), 1));
  list_v_.alloc(std::max<size_t>(tiles_v_.env.size(), 1));
}

template <class T>
void World<T>::upload_tiles(const HostTiles& h, DevTiles& d) {
  d.env.upload(h.env);
  d.row0.upload(h.row0);
  d.row1.upload(h.row1);
  d.first.upload(h.first);
  d.view.env = d.env.p;
  d.view.row0 = d.row0.p;
  d.view.row1 = d.row1.p;
  d.view.first = d.first.p;
  d.view.n = int32_t(h.env.size());
}

// The colliders of each environment (those of every actor and its own), in
// collider order: own colliders bucketed by actor, then merged per
// environment with the shared ones, in O(environments + colliders).
template <class T>
void World<T>::upload_colliders() {
  const int nC = int(colliders_.size());
  std::vector<int32_t> shared, own_first(static_cast<size_t>(nEnv_) + 1, 0), own(static_cast<size_t>(nC));
  for (int i = 0; i < nC; ++i) {
    const int a = colliders_[size_t(i)].actor;
    if (a < 0)
      shared.push_back(i);
    else if (a < nEnv_)
      ++own_first[size_t(a) + 1];
  }
  for (int e = 0; e < nEnv_; ++e) own_first[size_t(e) + 1] += own_first[size_t(e)];
  {
    std::vector<int32_t> fill(own_first.begin(), own_first.end() - 1);
    for (int i = 0; i < nC; ++i) {
      const int a = colliders_[size_t(i)].actor;
      if (a >= 0 && a < nEnv_) own[size_t(fill[size_t(a)]++)] = i;
    }
  }
  std::vector<int32_t> first(size_t(nEnv_) + 1, 0);
  std::vector<double> cd;
  cd.reserve(size_t(kCollStride) * (size_t(own_first[size_t(nEnv_)]) + shared.size() * size_t(nEnv_)));
  coll_entry_id_.clear();
  auto emit = [&](int i) {
    const HostCollider& c = colliders_[size_t(i)];
    coll_entry_id_.push_back(i);
    cd.push_back(double(c.kind));
    cd.insert(cd.end(), c.p, c.p + 3);
    cd.insert(cd.end(), c.n, c.n + 3);
    cd.push_back(c.r);
    cd.insert(cd.end(), c.p0, c.p0 + 3);
    cd.push_back(0.0);
  };
  for (int e = 0; e < nEnv_; ++e) {
    first[size_t(e)] = int32_t(cd.size() / kCollStride);
    size_t a = 0, b = size_t(own_first[size_t(e)]);
    const size_t b1 = size_t(own_first[size_t(e) + 1]);
    while (a < shared.size() || b < b1) {
      if (b >= b1 || (a < shared.size() && shared[a] < own[b]))
        emit(shared[a++]);
      else
        emit(own[b++]);
    }
  }
  first[size_t(nEnv_)] = int32_t(cd.size() / kCollStride);
  coll_first_h_ = first;
  if (cd.empty()) cd.assign(kCollStride, 0.0);
  if (d_coll_.n < cd.size()) d_coll_.alloc(cd.size());
  put(d_coll_, cd);
  if (d_coll_first_.n < first.size()) d_coll_first_.alloc(first.size());
  put(d_coll_first_, first);
  mesh_.coll_first = d_coll_first_.p;
  mesh_.coll = d_coll_.p;
  colliders_dirty_ = false;
}

template <class T>
void World<T>::set_env_mask(const std::vector<int32_t>& mask) {
  put(env_mask_, mask);
  env_mask_h_ = mask;
}

// Assemble residual (and BSR + preconditioner if wantH) for the environments
// in the current mask; per-environment ||res||^2 lands in env_sq_.  Assembly
// group by assembly group: a group's element and contact kernels write its
// records, its gathers read them back at once, so with several groups the
// records of one group (sized to stay in the last-level cache) are all that
// is live, and the next group reuses the same buffers.
template <class T>
void World<T>::assemble(double dt, const sdx_solver_params& sp, bool wantH, int project) {
  const int64_t set0 = int64_t(nEnv_) * kSlots * kSlotStride;
  const bool dd = precision_ != SDX_PREC_MODE_FP32;
  if (dd)
    launch_fill<double>(stream_, set0, 0.0, acc_.p);
  else
    launch_fill<T>(stream_, set0, T(0), reinterpret_cast<T*>(acc_.p));
  if (wantH && sched_.fused_asm && plan_row_blocks()) {  // residual, matrix and preconditioner in one kernel
    if (dd)
      launch_assemble_rows<T, double>(stream_, mesh_, rbv_, fields_, u_.p, u0_.p, v0_.p, dt, sp.gravity, project,
                                      sp.linear.preconditioner, vals_.p, dinv_.p, res_.p, acc_.p, stiff_damp_,
                                      sampled_events(2));
    else
      launch_assemble_rows<T, T>(stream_, mesh_, rbv_, fields_, u_.p, u0_.p, v0_.p, dt, sp.gravity, project,
                                 sp.linear.preconditioner, vals_.p, dinv_.p, res_.p, reinterpret_cast<T*>(acc_.p),
                                 stiff_damp_, sampled_events(2));
    if (dd)
      launch_acc_finalize<double>(stream_, nEnv_, 1, acc_.p, env_sq_.p);
    else
      launch_acc_finalize<T>(stream_, nEnv_, 1, reinterpret_cast<T*>(acc_.p), env_sq_.p);
    return;
  }
  const bool shared = asm_plan_.size() > 1;
  for (const AsmGroup& g : asm_plan_) {
    AsmRange r = g.r;
    r.contact_always = shared ? 1 : 0;
    T* Rv = group_records(cres_.p, 3, g.rv0);
    T* Dv = group_records(cdiag_.p, 6, g.rv0);
    T* Ev = group_records(coff_.p, 9, g.re0);
    // where the mask leaves out most of the group's patches, the element
    // kernel runs over a list of the others (a masked-out patch still loads
    // its description and tets before it finds its environment masked)
    AsmRange re = r;
    const int32_t* plist = nullptr;
    int64_t on = 0;
    if (env_mask_h_.size() == size_t(nEnv_))
      for (int e = g.e0; e < g.e1; ++e)
        if (env_mask_h_[size_t(e)]) on += env_p0_[size_t(e) + 1] - env_p0_[size_t(e)];
    if (on * 4 < int64_t(r.p1 - r.p0) * 3) {
      std::vector<int32_t> list;
      list.reserve(size_t(on));
      for (int e = g.e0; e < g.e1; ++e)
        if (env_mask_h_[size_t(e)])
          for (int q = env_p0_[size_t(e)]; q < env_p0_[size_t(e) + 1]; ++q) list.push_back(q);
      put(patch_list_, list);
      plist = patch_list_.p;
      re.p0 = 0;
      re.p1 = int32_t(list.size());
    }
    launch_element<T>(stream_, mesh_, re, fields_, u_.p, u0_.p, v0_.p, dt, sp.gravity, wantH ? 1 : 0, project, Rv,
                      Dv, Ev, sched_.element_block, sampled_events(2), stiff_damp_, plist);
    launch_contact<T>(stream_, mesh_, r, fields_, u_.p, u0_.p, T(dt), wantH ? 1 : 0, Rv, Dv, Ev);
    if (wantH)
      launch_gather_bsr<T>(stream_, mesh_, r, Dv, Ev, sp.linear.preconditioner, vals_.p, dinv_.p, sampled_events(3));
    if (dd)
      launch_gather_res<T, double>(stream_, mesh_, r, dt_v_.view, Rv, res_.p, acc_.p);
    else
      launch_gather_res<T, T>(stream_, mesh_, r, dt_v_.view, Rv, res_.p, reinterpret_cast<T*>(acc_.p));
  }
  if (dd)
    launch_acc_finalize<double>(stream_, nEnv_, 1, acc_.p, env_sq_.p);
  else
    launch_acc_finalize<T>(stream_, nEnv_, 1, reinterpret_cast<T*>(acc_.p), env_sq_.p);
}

// Row blocks of every actor (sdx_rowblocks.h), uploaded once per topology:
// blocks of at most 64 vertices and 1024 BSR slots (512 in FP64), so a
// block's LDS stays within 64 KB.  Returns false (the assembly then takes
// the record path) when some vertex couples to more vertices than one block
// holds.
template <class T>
bool World<T>::plan_row_blocks() {
  if (rb_ready_) return rbv_.n > 0;
  const int cap = sizeof(T) == 8 ? kBlockSlots / 2 : kBlockSlots;
  rb_ready_ = true;
  rbv_ = RowBlockView{};
  for (int v = 0; v < nV_; ++v)
    if (h_row_ptr_[size_t(v) + 1] - h_row_ptr_[size_t(v)] > cap) return false;
  RowBlocks rb(cap);
  for (int e = 0; e < nEnv_; ++e) {
    const HostActor& A = actors_[size_t(e)];
    const int v0 = env_v0_[size_t(e)];
    std::vector<int32_t> row_len(static_cast<size_t>(A.nV));
    for (int i = 0; i < A.nV; ++i) row_len[size_t(i)] = h_row_ptr_[size_t(v0 + i) + 1] - h_row_ptr_[size_t(v0 + i)];
    std::vector<int32_t> atris;
    BuildBoundary(A.X, A.tets, atris);
    if (int(atris.size() / 3) != env_f0_[size_t(e) + 1] - env_f0_[size_t(e)])
      throw std::logic_error("boundary triangles of an actor changed since its upload");
    rb.add(e, A.X, A.nV, A.tets, atris, row_len, v0, env_t0_[size_t(e)], env_f0_[size_t(e)]);
  }
  if (rows_lds_bytes_host<T>(rb.max_slots) > size_t(65536))
    throw std::invalid_argument("row blocks too large for the fused assembly's LDS");
  auto pairs = [](const std::vector<int32_t>& a, const std::vector<int32_t>& b) {
    std::vector<int32_t> out(2 * a.size());
    for (size_t k = 0; k < a.size(); ++k) {
      out[2 * k] = a[k];
      out[2 * k + 1] = b[k];
    }
    return out;
  };
  d_rb_row0_.upload(rb.row0);
  d_rb_tet0_.upload(rb.tet0);
  d_rb_tri0_.upload(rb.tri0);
  d_rb_env_.upload(rb.env);
  d_rb_rows_.upload(rb.rows);
  d_rb_rbase_.upload(rb.rbase);
  d_rb_tets_.upload(pairs(rb.tets, rb.tet_rows));
  d_rb_tris_.upload(rb.tris.empty() ? std::vector<int32_t>(2, 0) : pairs(rb.tris, rb.tri_rows));
  rbv_.n = rb.count();
  rbv_.max_slots = rb.max_slots;
  rbv_.row0 = d_rb_row0_.p;
  rbv_.tet0 = d_rb_tet0_.p;
  rbv_.tri0 = d_rb_tri0_.p;
  rbv_.env = d_rb_env_.p;
  rbv_.rows = d_rb_rows_.p;
  rbv_.rbase = d_rb_rbase_.p;
  rbv_.tets = reinterpret_cast<const int2*>(d_rb_tets_.p);
  rbv_.tris = reinterpret_cast<const int2*>(d_rb_tris_.p);
  return rbv_.n > 0;
}

// DResidual of the environments in mask, in up to two assemblies: the
// projected ones (proj[e]) and the others.  Leaves env_mask_ at the last
// subset assembled.
template <class T>
void World<T>::assemble_h(double dt, const sdx_solver_params& sp, const std::vector<int32_t>& mask,
                          const std::vector<uint8_t>& proj) {
  for (int pr = 1; pr >= 0; --pr) {
    std::vector<int32_t> m(size_t(nEnv_), 0);
    bool any = false;
    for (int e = 0; e < nEnv_; ++e)
      if (mask[size_t(e)] && int(proj[size_t(e)] != 0) == pr) {
        m[size_t(e)] = 1;
        any = true;
      }
    if (!any) continue;
    set_env_mask(m);
    assemble(dt, sp, true, pr);
  }
}

template <class T>
void World<T>::queue_energy(double dt, const sdx_solver_params& sp) {
  launch_fill<double>(stream_, int64_t(nEnv_) * kSlots * kSlotStride, 0.0, acc_.p);
  launch_energy<T>(stream_, mesh_, fields_, u_.p, u0_.p, v0_.p, dt, sp.gravity, acc_.p);
  launch_acc_finalize<double>(stream_, nEnv_, 1, acc_.p, env_obj_.p);
}

template <class T>
void World<T>::queue_env_dot(const T* a, const T* b) {
  launch_fill<double>(stream_, int64_t(nEnv_) * kSlots * kSlotStride, 0.0, acc_.p);
  launch_env_dot<T>(stream_, dt_v_.view, nEnv_, a, b, acc_.p);
  launch_acc_finalize<double>(stream_, nEnv_, 1, acc_.p, env_dot_.p);
}

// Record pointer of a group whose first record is rec0 (W values each) for
// the buffer `buf`: shifted so that record rec0 lands at buf + off, with off
// (< one 16-byte vector) chosen so that the gathers' 16-byte-aligned loads
// of global record positions stay aligned.  Kernels index it with global
// record positions of the group only, so every access lies in [buf, buf +
// off + records of the group + one vector).
template <class T>
T* World<T>::group_records(T* buf, int W, int64_t rec0) const {
  constexpr int64_t kW = 16 / int64_t(sizeof(T));
  const int64_t off = (rec0 * W) % kW;
  return reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(buf) +
                              uintptr_t(off - rec0 * W) * uintptr_t(sizeof(T)));
}

// Assembly groups asked for by the schedule; asm_groups 0 sizes them to
// about kAsmGroupBytes of element and contact records each.
template <class T>
int World<T>::asm_groups(const sdx_schedule& sc) const {
  int n = sc.asm_groups;
  if (n == 0) {
    const double bytes = double(env_rv0_.back()) * 9.0 * sizeof(T) + double(env_re0_.back()) * 9.0 * sizeof(T);
    n = int(std::ceil(bytes / kAsmGroupBytes));
  }
  return std::min(std::max(1, n), std::max(1, nEnv_));
}

// n contiguous environment ranges of about equal record counts, their tet,
// triangle, item and tile ranges, and record buffers for the largest.
template <class T>
void World<T>::plan_assembly(int n) {
  if (asm_plan_n_ == n && !asm_plan_.empty()) return;
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  asm_plan_n_ = n;
  asm_plan_.clear();
  const int64_t total = int64_t(env_rv0_.back()) * 9 + int64_t(env_re0_.back()) * 9;
  int64_t max_rv = 0, max_re = 0;
  int e = 0;
  for (int g = 1; g <= n && e < nEnv_; ++g) {
    const int64_t goal = total * g / n;
    int e1 = e + 1;  // at least one environment per group
    auto recs = [&](int k) { return int64_t(env_rv0_[size_t(k)]) * 9 + int64_t(env_re0_[size_t(k)]) * 9; };
    while (e1 < nEnv_ && recs(e1 + 1) <= goal) ++e1;
    if (g == n) e1 = nEnv_;
    AsmGroup G{};
    G.r.t0 = env_t0_[size_t(e)];
    G.r.t1 = env_t0_[size_t(e1)];
    G.r.p0 = env_p0_[size_t(e)];
    G.r.p1 = env_p0_[size_t(e1)];
    G.r.f0 = env_f0_[size_t(e)];
    G.r.f1 = env_f0_[size_t(e1)];
    G.r.s0 = h_row_ptr_[size_t(env_v0_[size_t(e)])];
    G.r.s1 = h_row_ptr_[size_t(env_v0_[size_t(e1)])];
    G.r.tile0 = tiles_v_.first[size_t(e)];
    G.r.tile1 = tiles_v_.first[size_t(e1)];
    G.rv0 = env_rv0_[size_t(e)];
    G.re0 = env_re0_[size_t(e)];
    G.e0 = e;
    G.e1 = e1;
    // every range within its array: the kernels index with these unchecked
    if (G.r.t0 < 0 || G.r.t1 > nT_ || G.r.t0 > G.r.t1 || G.r.p0 < 0 || G.r.p1 > nP_ || G.r.p0 > G.r.p1 ||
        G.r.f0 < 0 || G.r.f1 > nF_ || G.r.f0 > G.r.f1 ||
        G.r.s0 < 0 || G.r.s1 > nB_ || G.r.s0 > G.r.s1 ||
        G.r.tile0 < 0 || G.r.tile1 > int(tiles_v_.env.size()) || G.r.tile0 > G.r.tile1)
      throw std::logic_error("assembly group range out of bounds");
    max_rv = std::max<int64_t>(max_rv, int64_t(env_rv0_[size_t(e1)]) - G.rv0);
    max_re = std::max<int64_t>(max_re, int64_t(env_re0_[size_t(e1)]) - G.re0);
    asm_plan_.push_back(G);
    e = e1;
  }
  // the group's records, its alignment offset and one 16-byte vector past
  // them (the gathers' vector loads)
  constexpr size_t kW = 16 / sizeof(T);
  cres_.alloc(size_t(3 * max_rv) + 2 * kW);
  cdiag_.alloc(size_t(6 * max_rv) + 2 * kW);
  coff_.alloc(size_t(9 * max_re) + 2 * kW);
  // every triangle's contact records start out unwritten (both flags set)
  d_coll_flags_.upload(std::vector<uint8_t>(std::max<size_t>(size_t(nF_), 1), uint8_t(3)));
  mesh_.coll_flags = d_coll_flags_.p;
}

template <class T>
std::vector<double> World<T>::read_env_sq() {
  std::vector<double> h(static_cast<size_t>(nEnv_));
  get(env_sq_, h);
  flush_timers();
  return h;
}

// Launch CG burst m, iterations [k0, k0 + n), on the stream; with `timed`,
// every matvec and update launch records its own dispatch events.  With
// work lists, the caller compacted the tiles of the environments still
// active into the lists of parity `par`, and the kernels are persistent over
// them: grids sized for the largest group in the grouped bursts m <
// group_bursts_, else for all tiles, divided by 4^cls (grid class cls: the
// caller's bound on the active tiles; any grid is correct, a persistent
// kernel strides over its list).
template <class T>
void World<T>::issue_cg(const CgConfig& cfg, int m, int par, int k0, int n, bool timed, int cls) {
  const bool dd = precision_ != SDX_PREC_MODE_FP32;
  T* accf = reinterpret_cast<T*>(acc_.p);
  WorkList wa{nullptr, nullptr}, wv{nullptr, nullptr};
  int ga = 0, gv = 0;
  if (sched_.active_lists) {
    wa = WorkList{list_a_.p, list_count_.p + par};
    wv = WorkList{list_v_.p, list_count_.p + 2 + par};
    const bool grouped = m < group_bursts_;
    ga = std::max(8, std::min(grouped ? group_max_a_ : dt_a_.view.n, kPersistentGridA >> (2 * cls)));
    gv = std::max(8, std::min(grouped ? group_max_v_ : dt_v_.view.n, kPersistentGridV >> (2 * cls)));
  }
  for (int k = k0; k < k0 + n; ++k) {
    const KernelEvents* e0 = timed ? timer_events(0) : nullptr;
    if (dd)
      launch_cg_spmv<T, double>(stream_, dt_a_.view, wa, ga, cfg, k, sched_.spmv_lanes, d_row_ptr_.p, d_col_.p,
                                vals_.p, nB_, z_.p, xcg<double>(), res_.p, w_.p, rref_.p, env_cg_.p, acc_.p,
                                num_active_.p, e0, d_col16_.p);
    else
      launch_cg_spmv<T, T>(stream_, dt_a_.view, wa, ga, cfg, k, sched_.spmv_lanes, d_row_ptr_.p, d_col_.p, vals_.p,
                           nB_, z_.p, xcg<T>(), res_.p, w_.p, rref_.p, env_cg_.p, accf, num_active_.p, e0,
                           d_col16_.p);
    const KernelEvents* e1 = timed ? timer_events(1) : nullptr;
    if (dd)
      launch_cg_update<T, double>(stream_, dt_v_.view, wv, gv, cfg, k, dinv_.p, nV_, w_.p, rref_.p, xcg<double>(),
                                  r_.p, z_.p, p_.p, sv_.p, env_cg_.p, acc_.p, num_active_.p, e1);
    else
      launch_cg_update<T, T>(stream_, dt_v_.view, wv, gv, cfg, k, dinv_.p, nV_, w_.p, rref_.p, xcg<T>(), r_.p, z_.p,
                             p_.p, sv_.p, env_cg_.p, accf, num_active_.p, e1);
  }
}

// CG iterations [k0, k0 + n) over the lists of parity par as one cooperative
// launch of kTailGrid workgroups (k_cg_tail).
template <class T>
void World<T>::issue_cg_tail(const CgConfig& cfg, int par, int k0, int n) {
  const WorkList wa{list_a_.p, list_count_.p + par}, wv{list_v_.p, list_count_.p + 2 + par};
  if (precision_ != SDX_PREC_MODE_FP32)
    launch_cg_tail<T, double>(stream_, dt_a_.view, dt_v_.view, wa, wv, kTailGrid, cfg, k0, n, sched_.spmv_lanes,
                              d_row_ptr_.p, d_col_.p, d_col16_.p, vals_.p, nB_, dinv_.p, nV_, xcg<double>(), res_.p,
                              w_.p, rref_.p, r_.p, z_.p, p_.p, sv_.p, env_cg_.p, acc_.p, num_active_.p, tail_bar_.p,
                              tail_abort_.p);
  else
    launch_cg_tail<T, T>(stream_, dt_a_.view, dt_v_.view, wa, wv, kTailGrid, cfg, k0, n, sched_.spmv_lanes,
                         d_row_ptr_.p, d_col_.p, d_col16_.p, vals_.p, nB_, dinv_.p, nV_, xcg<T>(), res_.p, w_.p,
                         rref_.p, r_.p, z_.p, p_.p, sv_.p, env_cg_.p, reinterpret_cast<T*>(acc_.p), num_active_.p,
                         tail_bar_.p, tail_abort_.p);
}

// CG burst m (iterations [m B, m B + B)) over the lists of parity par as a
// replayable graph, captured on first use for the current configuration.
template <class T>
void World<T>::graphs_for(const CgConfig& cfg, int B) {
  if (std::memcmp(&graph_cfg_, &cfg, sizeof(CgConfig)) != 0 || graph_burst_ != B) {
    drop_graphs();
    graph_cfg_ = cfg;
    graph_burst_ = B;
  }
}

template <class T>
hipGraphExec_t World<T>::cg_graph(const CgConfig& cfg, int m, int par, int B, int cls) {
  graphs_for(cfg, B);
  const size_t key = (size_t(m) * kGridClasses + size_t(cls)) * 2 + size_t(par);
  if (key >= graphs_.size()) graphs_.resize(key + 1, nullptr);
  if (!graphs_[key]) {
    hipGraph_t g = nullptr;
    hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    issue_cg(cfg, m, par, m * B, B, false, cls);
    hip_check(hipStreamEndCapture(stream_, &g), "hipStreamEndCapture");
    hipGraphExec_t ex = nullptr;
    const hipError_t err = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hip_check(err, "hipGraphInstantiate");
    graphs_[key] = ex;
  }
  return graphs_[key];
}

// Environment groups asked for by the schedule; cg_groups 0 sizes them to
// about kGroupBytes of matrix values and column indices each, which with the
// group's vectors stays within the 256 MB last-level cache.
template <class T>
int World<T>::cg_groups(const sdx_schedule& sc) const {
  int n = sc.cg_groups;
  if (n == 0) {
    const double bytes = double(nB_) * double(9 * sizeof(T) + sizeof(int32_t));
    n = int(std::lround(bytes / kGroupBytes));
  }
  return std::min(std::max(1, n), std::max(1, nEnv_));
}

// Environment groups of the grouped CG bursts: `n` contiguous ranges of
// about equal matvec tile counts, as tile ranges of the matvec and the
// update tiles; also the largest group's tile counts.
template <class T>
void World<T>::plan_groups(int n) {
  drop_graphs();  // their grids and ranges follow the groups
  group_plan_n_ = n;
  group_a_.assign(1, 0);
  group_v_.assign(1, 0);
  group_max_a_ = group_max_v_ = 0;
  const std::vector<int32_t>& fa = tiles_a_.first;
  const std::vector<int32_t>& fv = tiles_v_.first;
  const int64_t total = fa[size_t(nEnv_)];
  int e = 0;
  for (int g = 1; g <= n && e < nEnv_; ++g) {
    const int64_t goal = total * g / n;
    int e1 = e + 1;  // at least one environment per group
    while (e1 < nEnv_ && fa[size_t(e1)] < goal) ++e1;
    if (g == n) e1 = nEnv_;
    group_a_.push_back(fa[size_t(e1)]);
    group_v_.push_back(fv[size_t(e1)]);
    group_max_a_ = std::max(group_max_a_, fa[size_t(e1)] - fa[size_t(e)]);
    group_max_v_ = std::max(group_max_v_, fv[size_t(e1)] - fv[size_t(e)]);
    e = e1;
  }
}

template <class T>
void World<T>::drop_graphs() {
  for (auto& g : graphs_)
    if (g) (void)hipGraphExecDestroy(g);
  graphs_.clear();
  if (grouped_graph_) (void)hipGraphExecDestroy(grouped_graph_);
  grouped_graph_ = nullptr;
}

// Whether every environment's rows fit a workgroup's LDS (whole-solve CG),
// in a world of several environments (a single one keeps the bursts, which
// spread its rows over every CU).
template <class T>
bool World<T>::env_cg_fits() const {
  return nEnv_ > 1 && env_max_rows_ > 0 && env_lds_bytes<T, double>(env_max_rows_, false) <= size_t(env_lds_limit_);
}

// Whether a solve runs as one whole-solve CG launch: asked for by the
// schedule, and env_cg_fits.
template <class T>
bool World<T>::env_cg_applies() const {
  return sched_.env_cg > 0 && env_cg_fits();
}

// One CG run for the environments with solve[e]; those with warm[e] keep
// dx_ as the initial guess.  Returns the per-environment CG state.
template <class T>
bool World<T>::cg_run(const CgConfig& cfg, const std::vector<int32_t>& solve, const std::vector<int32_t>& warm,
                      bool decide, std::vector<EnvCG>& h, bool timing_ok) {
  h.assign(static_cast<size_t>(nEnv_), EnvCG{});
  int active = 0, n_warm = 0;
  for (int e = 0; e < nEnv_; ++e) {
    std::memset(&h[size_t(e)], 0, sizeof(EnvCG));
    h[size_t(e)].status = solve[size_t(e)] ? kCgActive : kCgIdle;
    active += solve[size_t(e)] ? 1 : 0;
    n_warm += solve[size_t(e)] && warm[size_t(e)] ? 1 : 0;
  }
  if (active == 0) return false;
  put(solve_mask_, solve);
  std::vector<int32_t> na(1, active);
  put(num_active_, na);
  const int32_t* wm = n_warm ? warm_mask_.p : nullptr;
  launch_fill<int32_t>(stream_, 4, 0, list_count_.p);
  // decide (the learned guess u in x): the combination of u with the two
  // previous Newton steps v, w of least residual |b - A x0|, kept only where
  // it removes most of the residual (on device)
  if (n_warm && decide) launch_fill<double>(stream_, int64_t(nEnv_) * kSlots * kSlotStride, 0.0, acc_.p);
  const T* v = decide ? hist1() : nullptr;
  const T* w = decide ? hist2() : nullptr;
  const double* rv = decide ? rref2_.p : nullptr;
  const double* rw = decide ? rref3_.p : nullptr;
  const double* wa = decide ? ws_alpha_.p : nullptr;
  const double* wb = decide ? ws_beta_.p : nullptr;
  const double* wc = decide ? ws_gamma_.p : nullptr;
  if (precision_ != SDX_PREC_MODE_FP32) {
    if (n_warm) {
      put(warm_mask_, warm);
      if (decide) {
        launch_residual64_multi<T, double>(stream_, dt_a_.view, sched_.spmv_lanes, nEnv_, warm_mask_.p, d_row_ptr_.p,
                                           d_col_.p, vals_.p, nB_, xcg<double>(), v, w, res_.p,
                                           rref_.p, rref2_.p, rref3_.p, acc_.p);
        launch_ws_decide_multi(stream_, nEnv_, warm_mask_.p, acc_.p, warm_mask_.p, ws_ratio_.p, ws_ratio_hist_.p,
                               ws_alpha_.p,
                               ws_beta_.p, ws_gamma_.p);
      } else {
        launch_residual64<T, double>(stream_, dt_a_.view, sched_.spmv_lanes, nEnv_, warm_mask_.p, d_row_ptr_.p,
                                     d_col_.p, vals_.p, nB_, xcg<double>(), res_.p, rref_.p, nullptr);
      }
    }
    launch_cg_reset<double>(stream_, nEnv_, solve_mask_.p, env_cg_.p, acc_.p);
    launch_cg_init<T, double>(stream_, dt_v_.view, cfg, res_.p, wm, rref_.p, wa, v, rv, wb, w, rw, wc, dinv_.p, nV_,
                              xcg<double>(), r_.p, z_.p, env_cg_.p, acc_.p);
  } else {
    if (n_warm) {
      put(warm_mask_, warm);
      if (decide) {
        launch_residual64_multi<T, T>(stream_, dt_a_.view, sched_.spmv_lanes, nEnv_, warm_mask_.p, d_row_ptr_.p,
                                      d_col_.p, vals_.p, nB_, xcg<T>(), v, w, res_.p, rref_.p,
                                      rref2_.p, rref3_.p, acc_.p);
        launch_ws_decide_multi(stream_, nEnv_, warm_mask_.p, acc_.p, warm_mask_.p, ws_ratio_.p, ws_ratio_hist_.p,
                               ws_alpha_.p,
                               ws_beta_.p, ws_gamma_.p);
      } else {
        launch_residual64<T, T>(stream_, dt_a_.view, sched_.spmv_lanes, nEnv_, warm_mask_.p, d_row_ptr_.p, d_col_.p,
                                vals_.p, nB_, xcg<T>(), res_.p, rref_.p, nullptr);
      }
    }
    T* accf = reinterpret_cast<T*>(acc_.p);
    launch_cg_reset<T>(stream_, nEnv_, solve_mask_.p, env_cg_.p, accf);
    launch_cg_init<T, T>(stream_, dt_v_.view, cfg, res_.p, wm, rref_.p, wa, v, rv, wb, w, rw, wc, dinv_.p, nV_,
                         xcg<T>(), r_.p, z_.p, env_cg_.p, accf);
  }
  tail_used_ = false;
  if (env_cg_applies()) {
    // the whole solve as one launch: each workgroup iterates whole
    // environments from the queue (largest first) to their stop decisions
    const int grid = std::max(1, std::min(active, std::max(1, num_cus_) * sched_.env_cg));
    launch_fill<int32_t>(stream_, 1, 0, env_next_.p);
    const EnvQueue eq{d_env_order_.p, nEnv_, env_max_rows_, env_next_.p, d_env_v0_.p};
    const KernelEvents* ev = timing_ok ? env_timer_events() : nullptr;
    if (precision_ != SDX_PREC_MODE_FP32)
      launch_cg_env<T, double>(stream_, eq, grid, size_t(env_lds_limit_), cfg, sched_.spmv_lanes, d_row_ptr_.p, d_col_.p, d_col16_.p,
                               vals_.p, nB_, dinv_.p, nV_, res_.p, z_.p, xcg<double>(), r_.p, p_.p, sv_.p, rref_.p,
                               env_cg_.p, acc_.p, num_active_.p, ev);
    else
      launch_cg_env<T, T>(stream_, eq, grid, size_t(env_lds_limit_), cfg, sched_.spmv_lanes, d_row_ptr_.p, d_col_.p, d_col16_.p, vals_.p,
                          nB_, dinv_.p, nV_, res_.p, z_.p, xcg<T>(), r_.p, p_.p, sv_.p, rref_.p, env_cg_.p,
                          reinterpret_cast<T*>(acc_.p), num_active_.p, ev);
    if (cfg.xd)  // the increments since the last refresh, into x
      launch_cg_fold<T, double>(stream_, int64_t(3) * nV_, xcg<double>(), static_cast<T*>(cfg.xd));
    return true;
  }
  // Bursts of B iterations are issued kBurstAhead bursts ahead of the host's
  // check of the active count, so the device does not wait for the host
  // between bursts even when a burst of the last few environments takes less
  // time than the host needs to issue the next one; bursts issued after every
  // environment finished exit at once.  With work lists, burst m starts with
  // the compaction, which also writes the count left by burst m - 1 into
  // pinned host memory (slot m mod kBurstSlots) and records event m; without
  // lists the count is copied after burst m and event m follows the copy.
  // After issuing burst m the host reads the count left by burst
  // m - kBurstAhead.
  const int B = std::max(1, sched_.cg_burst);
  volatile int32_t* hna = pinned_active_;
  const bool lists = sched_.active_lists != 0;
  // kernel timing samples whole solves only: a continuation of a few
  // environments would not cover the rows its launches are counted for
  const bool timed0 = timing_ok && burst_timed(0);
  int q = 0;  // compactions issued: list parity q & 1
  if (max_env_a_ == 0)
    for (int e = 0; e < nEnv_; ++e)
      max_env_a_ = std::max(max_env_a_, tiles_a_.first[size_t(e) + 1] - tiles_a_.first[size_t(e)]);
  // Grouped bursts: the first group_bursts_ bursts run group after group, so
  // that one group's matrix and vectors stay in the last-level cache across
  // the group's iterations.  They are queued without host checks.
  const int groups = lists ? cg_groups(sched_) : 1;
  group_bursts_ = 0;
  if (groups > 1 && sched_.group_iters > 0) {
    group_bursts_ = std::min((sched_.group_iters + B - 1) / B, cfg.max_iter / B + 1);
    if (group_plan_n_ != groups) plan_groups(groups);
    const int ng = int(group_a_.size()) - 1;
    // every group's compactions and bursts, in order (eager, or into a capture)
    auto issue_grouped = [&](bool timed_ok) {
      int qq = 0;
      for (int g = 0; g < ng; ++g)
        for (int m = 0; m < group_bursts_; ++m, ++qq) {
          const bool timed = m == 0 && timed_ok;
          launch_compact_tiles(stream_, dt_a_.view, dt_v_.view, group_a_[size_t(g)], group_a_[size_t(g) + 1],
                               group_v_[size_t(g)], group_v_[size_t(g) + 1], env_cg_.p, list_a_.p, list_v_.p,
                               list_count_.p, qq & 1, num_active_.p, nullptr);
          issue_cg(cfg, m, qq & 1, m * B, B, timed, 0);
        }
    };
    if (sched_.use_graphs && !timed0) {
      // the whole grouped phase as one graph: no launch boundary between bursts
      graphs_for(cfg, B);
      if (grouped_graph_ && (grouped_graph_groups_ != ng || grouped_graph_bursts_ != group_bursts_)) {
        (void)hipGraphExecDestroy(grouped_graph_);
        grouped_graph_ = nullptr;
      }
      if (!grouped_graph_) {
        hipGraph_t gr = nullptr;
        hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        issue_grouped(false);
        hip_check(hipStreamEndCapture(stream_, &gr), "hipStreamEndCapture");
        const hipError_t err = hipGraphInstantiate(&grouped_graph_, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (err != hipSuccess) grouped_graph_ = nullptr;
        hip_check(err, "hipGraphInstantiate");
        grouped_graph_groups_ = ng;
        grouped_graph_bursts_ = group_bursts_;
      }
      hip_check(hipGraphLaunch(grouped_graph_, stream_), "hipGraphLaunch");
    } else {
      issue_grouped(timed0);
    }
    q = ng * group_bursts_;
  }
  // grid class of the batch-wide bursts from the last published active
  // count: at most that many environments, each with at most max_env_a_
  // matvec tiles, are left
  int known_active = active;
  for (int m = group_bursts_;; ++m) {
    const bool issue = int64_t(m) * B <= int64_t(cfg.max_iter);
    const int slot = m % kBurstSlots;
    if (issue) {
      const bool timed = m == 0 && timed0;
      int cls = 0;
      if (lists) {
        const int64_t bound = int64_t(known_active) * max_env_a_;
        while (cls + 1 < kGridClasses && bound <= int64_t(kPersistentGridA >> (2 * (cls + 1)))) ++cls;
      }
      // the tail: few enough tiles for the smallest grid, one cooperative
      // launch per burst (always eager: not captured into graphs)
      const bool tail = lists && sched_.fused_tail && cls == kGridClasses - 1 && !timed;
      if (tail && !tail_used_) {
        launch_fill<int32_t>(stream_, 1, 0, tail_abort_.p);
        tail_used_ = true;
      }
      if (lists) {
        launch_compact_tiles(stream_, dt_a_.view, dt_v_.view, 0, dt_a_.view.n, 0, dt_v_.view.n, env_cg_.p, list_a_.p,
                             list_v_.p, list_count_.p, q & 1, num_active_.p, pinned_active_dev_ + slot,
                             tail ? tail_bar_.p : nullptr);
        hip_check(hipEventRecord(burst_done_[slot], stream_), "hipEventRecord");
      }
      if (tail)
        issue_cg_tail(cfg, q & 1, m * B, B);
      else if (sched_.use_graphs && !timed)
        hip_check(hipGraphLaunch(cg_graph(cfg, m, q & 1, B, cls), stream_), "hipGraphLaunch");
      else
        issue_cg(cfg, m, q & 1, m * B, B, timed, cls);
      ++q;
      if (!lists) {
        hip_check(hipMemcpyAsync(pinned_active_ + slot, num_active_.p, sizeof(int32_t), hipMemcpyDeviceToHost,
                                 stream_),
                  "hipMemcpyAsync");
        hip_check(hipEventRecord(burst_done_[slot], stream_), "hipEventRecord");
      }
    }
    if (!issue) break;  // every burst up to max_iter is queued
    const int j = m - kBurstAhead;  // check the count left by burst j
    // the compaction of burst j + 1 (a batch-wide one, which publishes the
    // count), or the copy after burst j
    if (j >= 0 && (!lists || j + 1 >= group_bursts_)) {
      const int s = (lists ? j + 1 : j) % kBurstSlots;
      hip_check(hipEventSynchronize(burst_done_[s]), "hipEventSynchronize");
      if (hna[s] <= 0) break;
      known_active = std::min(known_active, int(hna[s]));
    }
  }
  if (cfg.xd)  // the increments since the last refresh, into x
    launch_cg_fold<T, double>(stream_, int64_t(3) * nV_, xcg<double>(), static_cast<T*>(cfg.xd));
  return true;
}

// FP64 ||b - A x||^2 and ||b||^2 of every environment with mask[e] into
// norms2_, queued on the stream.
template <class T>
void World<T>::issue_true_residual(const std::vector<int32_t>& mask) {
  put(solve_mask_, mask);
  launch_fill<double>(stream_, int64_t(nEnv_) * kSlots * kSlotStride, 0.0, acc_.p);
  if (precision_ != SDX_PREC_MODE_FP32)
    launch_residual64<T, double>(stream_, dt_a_.view, sched_.spmv_lanes, nEnv_, solve_mask_.p, d_row_ptr_.p,
                                 d_col_.p, vals_.p, nB_, xcg<double>(), res_.p, nullptr, acc_.p);
  else
    launch_residual64<T, T>(stream_, dt_a_.view, sched_.spmv_lanes, nEnv_, solve_mask_.p, d_row_ptr_.p, d_col_.p,
                            vals_.p, nB_, xcg<T>(), res_.p, nullptr, acc_.p);
  launch_acc_finalize<double>(stream_, nEnv_, 2, acc_.p, norms2_.p);
}

// ||b - A x|| / ||b|| from the downloaded norms2_.
template <class T>
std::vector<double> World<T>::true_rel_from(const std::vector<int32_t>& mask, const std::vector<double>& q) const {
  std::vector<double> out(size_t(nEnv_), 0.0);
  for (int e = 0; e < nEnv_; ++e)
    if (mask[size_t(e)] && q[size_t(2 * e + 1)] > 0) out[size_t(e)] = std::sqrt(q[size_t(2 * e)] / q[size_t(2 * e + 1)]);
  return out;
}

// Block-Jacobi PCG on the assembled matrix for the environments with
// solve[e] != 0, from x0 = 0 (or dx_ where warm[e] is set).  The iterate is
// kept in the accumulator type (xcg_: FP64 in mixed precision); the solution
// is rounded into dx_ at the end.
// In mixed precision with the ResidualL2 norm, a solve that converged on the
// recursive residual is confirmed against the FP64 true residual and, if
// that misses the tolerance, continued from its iterate (at most twice).
// true_rel receives the FP64 ||b - A x|| / ||b|| of every solved environment.
template <class T>
std::vector<EnvCG> World<T>::cg_solve(const sdx_krylov_params& kp, const std::vector<int32_t>& solve,
                                      std::vector<double>* true_rel, const std::vector<int32_t>* warm_in,
                                      WarmStats* learned, const std::vector<double>* env_tol) {
  CgConfig cfg{};
  cfg.n_env = nEnv_;
  cfg.n_rows = nV_;
  cfg.norm_type = kp.norm_type;
  cfg.prec = kp.preconditioner;
  cfg.max_iter = kp.max_iter;
  cfg.refresh_every = precision_ == SDX_PREC_MODE_MIXED ? kp.refresh_every : 0;
  cfg.abort_if_not_spd = kp.abort_if_not_spd;
  cfg.abs_tol = kp.abs_tol;
  cfg.rel_tol = kp.rel_tol;
  cfg.rel_div_tol = kp.rel_div_tol;
  if (env_tol) {  // the values live on the device: the captured graphs stay valid across Newton iterations
    put(env_tol_, *env_tol);
    cfg.env_rel_tol = env_tol_.p;
  }
  auto rel_tol_of = [&](int e) { return env_tol ? (*env_tol)[size_t(e)] : kp.rel_tol; };
  // mixed precision: FP32 increments of the FP64 iterate between refreshes;
  // without refreshes every update adds into the FP64 iterate itself
  cfg.xd = (precision_ == SDX_PREC_MODE_MIXED && std::is_same<T, float>::value && cfg.refresh_every > 0)
               ? static_cast<void*>(xd_.p)
               : nullptr;
  std::vector<int32_t> warm = warm_in ? *warm_in : std::vector<int32_t>(size_t(nEnv_), 0);
  std::vector<int32_t> todo = solve;
  const int64_t n3 = int64_t(3) * nV_;
  const bool dd = precision_ != SDX_PREC_MODE_FP32;
  if (warm_in) {  // initial guesses arrive in dx_
    if (dd)
      launch_convert<T, double>(stream_, n3, dx_.p, xcg<double>());
    else
      launch_convert<T, T>(stream_, n3, dx_.p, xcg<T>());
  }
  // the solve, the true residuals and every download, then one wait
  std::vector<EnvCG> h;
  const bool ran = cg_run(cfg, todo, warm, learned != nullptr, h, true);
  issue_true_residual(solve);
  std::vector<double> q(size_t(2 * nEnv_));
  if (ran) fetch(env_cg_, h);
  if (learned) {
    learned->accepted.assign(size_t(nEnv_), 0);
    learned->ratio.assign(size_t(nEnv_), 0.0);
    fetch(warm_mask_, learned->accepted);
    learned->ratio_hist.assign(size_t(nEnv_), 0.0);
    fetch(ws_ratio_, learned->ratio);
    fetch(ws_ratio_hist_, learned->ratio_hist);
  }
  fetch(norms2_, q);
  std::vector<int32_t> tail_abort(1, 0);
  if (tail_used_) fetch(tail_abort_, tail_abort);
  fetch_all();
  if (tail_abort[0]) throw std::runtime_error("CG tail burst: a grid barrier wait ran out; the solve is invalid");
  std::vector<double> tr = true_rel_from(solve, q);
  const bool confirm = precision_ == SDX_PREC_MODE_MIXED && kp.norm_type == SDX_NORM_RESIDUAL_L2;
  for (int round = 0; confirm && round < 2; ++round) {
    int n = 0;
    for (int e = 0; e < nEnv_; ++e) {
      const EnvCG& c = h[size_t(e)];
      const bool redo = todo[size_t(e)] && c.status == kCgConverged && tr[size_t(e)] > rel_tol_of(e) &&
                        tr[size_t(e)] * std::sqrt(c.start) > kp.abs_tol && c.iter < kp.max_iter;
      todo[size_t(e)] = redo ? 1 : 0;
      warm[size_t(e)] = todo[size_t(e)];
      n += redo ? 1 : 0;
    }
    if (n == 0) break;
    CgConfig c2 = cfg;
    std::vector<EnvCG> h2;
    const bool ran2 = cg_run(c2, todo, warm, false, h2, false);
    issue_true_residual(todo);
    std::vector<double> q2(size_t(2 * nEnv_));
    if (ran2) fetch(env_cg_, h2);
    fetch(norms2_, q2);
    std::vector<int32_t> tail_abort2(1, 0);
    if (tail_used_) fetch(tail_abort_, tail_abort2);
    fetch_all();
    if (tail_abort2[0]) throw std::runtime_error("CG tail burst: a grid barrier wait ran out; the solve is invalid");
    std::vector<double> tr2 = true_rel_from(todo, q2);
    for (int e = 0; e < nEnv_; ++e) {
      if (!todo[size_t(e)]) continue;
      const int prev = h[size_t(e)].iter;
      h[size_t(e)] = h2[size_t(e)];
      h[size_t(e)].iter += prev;
      tr[size_t(e)] = tr2[size_t(e)];
    }
  }
  if (dd)  // the Newton update takes the solution in the state type
    launch_convert<double, T>(stream_, n3, xcg<double>(), dx_.p);
  else
    launch_convert<T, T>(stream_, n3, xcg<T>(), dx_.p);
  flush_timers();
  if (true_rel) *true_rel = tr;
  if (cfg.refresh_every > 0)
    for (auto& e : h) e.refreshes = e.iter / cfg.refresh_every;
  return h;
}

// Features of the learned warm start for every vertex, from the assembled
// residual and matrix and the previous Newton step.
template <class T>
void World<T>::ws_features() {
  launch_fill<double>(stream_, int64_t(nEnv_) * kSlots * kSlotStride, 0.0, acc_.p);
  launch_ws_scaled_residual<T>(stream_, mesh_, dt_v_.view, res_.p, vals_.p, hist1(), ws_y_.p, acc_.p);
  launch_ws_env_scale(stream_, nEnv_, acc_.p, ws_scale_.p);
  launch_ws_features<T>(stream_, mesh_, dt_v_.view, vals_.p, ws_y_.p, hist1(), ws_scale_.p, ws_F_.p);
}

template <class T>
void World<T>::set_ws_params(const float* params) {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  if (!params) {
    ws_on_ = false;
    return;
  }
  if (!ws_params_.p) {
    ws_params_.alloc(kMlpParams);
    ws_wpack_.alloc(kBf16W1 + 2 * kBf16W2 + kBf16W4);
    ws_bpack_.alloc(3 * kMlpHid + 16);
  }
  std::vector<float> h(params, params + kMlpParams);
  ws_params_.upload(h);
  uint16_t* w = ws_wpack_.p;
  float* b = ws_bpack_.p;
  launch_mlp_pack(stream_, ws_params_.p, w, w + kBf16W1, w + kBf16W1 + kBf16W2, w + kBf16W1 + 2 * kBf16W2, b,
                  b + kMlpHid, b + 2 * kMlpHid, b + 3 * kMlpHid);
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  ws_w_ = MlpBf16{w, w + kBf16W1, w + kBf16W1 + kBf16W2, w + kBf16W1 + 2 * kBf16W2, b, b + kMlpHid, b + 2 * kMlpHid,
                  b + 3 * kMlpHid};
  ws_on_ = true;
  wg_ = WarmGate{};
}

template <class T>
void World<T>::collect(int64_t capacity, int per_env) {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  coll_n_ = 0;
  coll_salt_ = 0;
  if (capacity <= 0) {
    coll_x_.free();
    coll_t_.free();
    coll_cap_ = 0;
    coll_per_env_ = 0;
    return;
  }
  coll_x_.alloc(size_t(capacity) * kMlpInPad);
  coll_t_.alloc(size_t(capacity) * kMlpOut);
  coll_cap_ = capacity;
  coll_per_env_ = per_env;
}

template <class T>
int64_t World<T>::collected(float* x, float* t, int64_t cap_rows) {
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  const int64_t n = std::min(coll_n_, cap_rows);
  if (x && n > 0)
    hip_check(hipMemcpy(x, coll_x_.p, sizeof(float) * size_t(n) * kMlpInPad, hipMemcpyDeviceToHost), "hipMemcpy");
  if (t && n > 0)
    hip_check(hipMemcpy(t, coll_t_.p, sizeof(float) * size_t(n) * kMlpOut, hipMemcpyDeviceToHost), "hipMemcpy");
  return coll_n_;
}

template <class T>
void World<T>::collected_device(const float** x, const float** t, int64_t* n) {
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  *x = coll_x_.p;
  *t = coll_t_.p;
  *n = coll_n_;
}

// Forcing term of the linear solve of Newton iteration k
// (NewtonSolver::SetLinearSolverRelativeTolerance, newton_solver.cpp:426-501):
// Eisenstat-Walker choices 1 and 2 from 0.01, their safeguard only above
// 0.01, clamped to [100 eps(T), 0.01]; choice 3 max(min(1/(2 + k),
// sqrt(|r|)), relTol); Constant keeps the previous value.
template <class T>
static double forcing_term(int strategy, int k, double res, double prev_res, double lin_res, double prev_eta,
                           double rel_tol) {
  const double kMaxEta = 0.01, kSafeguardEtaTh = 0.01;
  const double kMinEta = 100.0 * double(std::numeric_limits<T>::epsilon());
  const double alpha = 0.5 + 0.5 * std::sqrt(5.0);
  auto clamp = [&](double eta) { return std::min(std::max(eta, kMinEta), kMaxEta); };
  switch (strategy) {
    case SDX_LINTOL_EW1: {
      if (k == 0) return kMaxEta;
      const double safeguard = std::pow(prev_eta, alpha);
      double eta = std::fabs(res - lin_res) / prev_res;
      if (safeguard > kSafeguardEtaTh) eta = std::max(eta, safeguard);
      return clamp(eta);
    }
    case SDX_LINTOL_EW2: {
      if (k == 0) return kMaxEta;
      const double safeguard = 0.9 * std::pow(prev_eta, alpha);
      double eta = 0.9 * std::pow(res / prev_res, alpha);
      if (safeguard > kSafeguardEtaTh) eta = std::max(eta, safeguard);
      return clamp(eta);
    }
    case SDX_LINTOL_EW3:
      return std::max(std::min(1.0 / (2.0 + double(k)), std::sqrt(res)), rel_tol);
    default:
      return prev_eta;
  }
}

// One backward-Euler step of every environment: NewtonSolver<T>::Solve
// (newton_solver.cpp:603-876) run in lockstep over the environments, each
// with its own stop decisions, forcing terms, tries and line search.
//
// Tries (:736-860), with the default fitted-Hessian settings (joint friction
// off: the first try of an iteration is never its last): a try whose linear
// solve does not converge, or whose line search finds no improvement, is
// undone and retried once with the PSD-projected DResidual; the retry keeps
// its last line-search iterate or, with the gradient-descent fallback,
// searches along -res from where the iteration started.  Where the first try
// was already projected (PSD mode Always), the retry would repeat the same
// solve and search exactly: it is not run again, its iterations are counted
// and its outcome taken from the first.
template <class T>
void World<T>::step(double dt, const sdx_solver_params& sp, sdx_step_stats* stats, int n_stats) {
  if (!(dt > 0)) throw std::invalid_argument("dt must be positive");
  finalize();
  if (sp.schedule == -1 && !tuned_) autotune(dt, sp, 2);
  if (colliders_dirty_) upload_colliders();
  const int nE = nEnv_;
  const size_t n3 = size_t(3) * size_t(nV_);
  // EntityIncrementStep / PreStage: stage start = previous step.
  launch_copy<T>(stream_, int64_t(n3), u_.p, u0_.p);
  launch_copy<T>(stream_, int64_t(n3), v_.p, v0_.p);
  launch_apply_pins<T>(stream_, int64_t(n3), d_pinned_.p, d_pin_value_.p, u_.p);

  const int ls_type = sp.line_search_type;
  const bool need_obj = ls_type != SDX_LS_NONE && ls_type != SDX_LS_RESIDUAL_NORM;
  const bool gd = sp.gradient_descent_fallback != 0;
  const bool adaptive_tol = sp.linear_tolerance_strategy != SDX_LINTOL_CONSTANT;
  std::vector<EnvState> S(static_cast<size_t>(nE));
  for (EnvState& s : S) {
    s.psd = sp.psd_proj_mode;
    s.eta = sp.linear.rel_tol;
  }
  std::vector<int32_t> mask(size_t(nE), 1);
  std::vector<uint8_t> proj(size_t(nE), 0);
  // |dx| is needed by the RelStepNorm test of the next Newton iteration only
  const bool need_dx = sp.newton_max_iter > 1 && sp.rel_step_tol > 0;
  bool first_iter = true;
  while (true) {
    bool h_valid = false;
    for (int e = 0; e < nE; ++e) mask[size_t(e)] = S[size_t(e)].active ? 1 : 0;
    if (first_iter) {
      // residual + dresidual (newton_solver.cpp:667-686); every environment
      // starts in the same PSD mode
      set_env_mask(mask);
      assemble(dt, sp, true, sp.psd_proj_mode == SDX_PSD_ALWAYS ? 1 : 0);
      h_valid = true;
      auto sq = read_env_sq();
      for (int e = 0; e < nE; ++e) {
        S[size_t(e)].res = std::sqrt(sq[size_t(e)]);
        S[size_t(e)].res_known = true;
      }
      first_iter = false;
    }
    if (need_obj) {  // the merit at the current iterate
      set_env_mask(mask);
      queue_energy(dt, sp);
      std::vector<double> ob(static_cast<size_t>(nE));
      fetch(env_obj_, ob);
      fetch_all();
      for (int e = 0; e < nE; ++e)
        if (S[size_t(e)].active) S[size_t(e)].merit = ob[size_t(e)];
    }
    // stop criteria (newton_solver.cpp:304-314): ResNorm, MaxIters, RelStepNorm
    bool need_norms = false;
    for (int e = 0; e < nE; ++e)
      if (S[size_t(e)].active && S[size_t(e)].iters > 0 && S[size_t(e)].have_dx && sp.rel_step_tol > 0)
        need_norms = true;
    std::vector<double> n2;
    if (need_norms) {
      launch_fill<double>(stream_, int64_t(nEnv_) * kSlots * kSlotStride, 0.0, acc_.p);
      launch_env_norms<T>(stream_, dt_v_.view, nEnv_, u_.p, dx_.p, acc_.p);
      launch_acc_finalize<double>(stream_, nEnv_, 2, acc_.p, norms2_.p);
      n2.resize(size_t(2 * nE));
      get(norms2_, n2);
    }
    int n_active = 0;
    for (int e = 0; e < nE; ++e) {
      EnvState& s = S[size_t(e)];
      if (!s.active) continue;
      if (s.iters == 0) s.res0 = s.res;
      const bool absConv = s.res <= sp.newton_abs_tol;
      const bool relConv = s.iters > 0 && s.res <= sp.newton_rel_tol * s.res0;
      if (absConv || relConv) {
        s.conv = SDX_CONV_CONVERGED;
        s.active = false;
      } else if (s.iters >= sp.newton_max_iter) {
        s.conv = SDX_CONV_STOPPED;
        s.active = false;
      } else if (need_norms && s.iters > 0 && s.have_dx &&
                 s.dx_norm <= double(T(sp.rel_step_tol)) * std::sqrt(n2[size_t(2 * e)])) {
        s.conv = SDX_CONV_STOPPED;
        s.active = false;
      }
      if (s.active) ++n_active;
    }
    if (n_active == 0) break;
    for (int e = 0; e < nE; ++e) mask[size_t(e)] = S[size_t(e)].active ? 1 : 0;
    int newton_it = 0;  // every active environment is in the same Newton iteration
    std::vector<double> env_tol(size_t(nE), sp.linear.rel_tol);
    for (int e = 0; e < nE; ++e) {
      EnvState& s = S[size_t(e)];
      if (!s.active) continue;
      s.eta = forcing_term<T>(sp.linear_tolerance_strategy, s.iters, s.res, s.prev_res, s.lin_res, s.eta,
                              sp.linear.rel_tol);
      env_tol[size_t(e)] = s.eta;
      s.prev_res = s.res;
      s.res_start = s.res;
      s.merit0 = s.merit;
      newton_it = s.iters++;
    }
    hist_k_ = newton_it;  // the warm start draws on the steps of this Newton iteration
    launch_copy<T>(stream_, int64_t(n3), u_.p, sol0_.p);
    if (gd) launch_copy<T>(stream_, int64_t(n3), res_.p, res0_.p);
    for (int e = 0; e < nE; ++e) proj[size_t(e)] = S[size_t(e)].psd == SDX_PSD_ALWAYS ? 1 : 0;
    std::vector<int32_t> trying(mask);
    std::vector<uint8_t> retried(size_t(nE), 0);
    bool first_try = true;
    while (true) {
      int n_try = 0;
      for (int e = 0; e < nE; ++e) n_try += trying[size_t(e)] ? 1 : 0;
      if (n_try == 0) break;
      // the DResidual at the iterate the iteration started from (assembled with the residual on
      // the first iteration)
      if (!h_valid) assemble_h(dt, sp, trying, proj);
      h_valid = false;
      set_env_mask(trying);
      // linear solve, solution in dx_: from zero, or on the first try from the learned warm start
      // where its guess lowers the residual.  A Newton iteration whose guesses most environments
      // rejected in its last warm-started solve starts from zero for the next kWarmSkipSteps steps
      // (the network, features and guess residuals cost more than the few kept guesses save)
      const int hk = std::min(newton_it, kNewtonHist - 1);
      // the history-only start is the learned one with a zero network guess, which the
      // least-residual combination drops (its Gram rows vanish)
      const bool history_only = sp.warm_start == SDX_WARM_START_HISTORY;
      bool learned = first_try && ((sp.warm_start == SDX_WARM_START_LEARNED && ws_on_) || history_only);
      if (learned && wg_.skip[hk] > 0) {
        --wg_.skip[hk];
        learned = false;
      }
      // the network runs unless its gate holds this index on the previous steps alone
      bool network = learned && !history_only;
      if (network && wg_.net_skip[hk] > 0) {
        --wg_.net_skip[hk];
        network = false;
      }
      const bool collecting = first_try && coll_per_env_ > 0 && coll_n_ + int64_t(nE) * coll_per_env_ <= coll_cap_;
      if (network || collecting) ws_features();
      if (learned && !network)
        launch_fill<T>(stream_, int64_t(n3), T(0), dx_.p);
      else if (learned)
        launch_mlp_infer<T>(stream_, ws_w_, ws_F_.p, nV_, d_row_env_.p, ws_scale_.p, d_pinned_.p, nullptr, dx_.p);
      std::vector<double> true_rel;
      WarmStats ws;
      auto cg = cg_solve(sp.linear, trying, &true_rel, learned ? &trying : nullptr, learned ? &ws : nullptr,
                         adaptive_tol ? &env_tol : nullptr);
      if (collecting) {  // (features, Newton step / scale) pairs of this solve
        launch_ws_collect<T>(stream_, nE, d_env_v0_.p, env_mask_.p, coll_per_env_, coll_salt_++, ws_F_.p, dx_.p,
                             ws_scale_.p, coll_n_, coll_x_.p, coll_t_.p);
        coll_n_ += int64_t(nE) * coll_per_env_;
      }
      if (first_try) {
        launch_env_copy<T>(stream_, nV_, d_row_env_.p, env_mask_.p, hist1(), hist2());
        launch_env_copy<T>(stream_, nV_, d_row_env_.p, env_mask_.p, dx_.p, hist1());
      }
      if (learned) {
        int tried = 0, kept = 0;
        double log_gain = 0.0;  // sum of log(ratio / ratio_hist): what the network's guess adds
        for (int e = 0; e < nE; ++e)
          if (trying[size_t(e)]) {
            ++tried;
            kept += ws.accepted[size_t(e)] ? 1 : 0;
            if (network)
              log_gain += std::log(std::max(ws.ratio[size_t(e)], 1e-300) / std::max(ws.ratio_hist[size_t(e)], 1e-300));
          }
        // the first two warm-started solves of an iteration index run before
        // its history of previous Newton steps is filled, and never skip
        if (++wg_.seen[hk] > 2) {
          if (double(kept) < kWarmMinKept * double(tried)) wg_.skip[hk] = kWarmSkipSteps;
          if (network && tried > 0 && log_gain > double(tried) * std::log(kNetMinGain))
            wg_.net_skip[hk] = kWarmSkipSteps;
        }
      }
      std::vector<uint8_t> lin_ok(size_t(nE), 0);
      for (int e = 0; e < nE; ++e) {
        EnvState& s = S[size_t(e)];
        if (!trying[size_t(e)]) continue;
        const EnvCG& c = cg[size_t(e)];
        if (learned) {
          s.ws_tried++;
          s.ws_accepted += ws.accepted[size_t(e)] ? 1 : 0;
          s.ws_ratio = ws.ratio[size_t(e)];
          s.ws_ratio_hist = ws.ratio_hist[size_t(e)];
          s.ws_network = network ? 1 : 0;
        }
        s.lin_iters += c.iter;
        s.lin_conv = c.status == kCgConverged ? SDX_CONV_CONVERGED
                     : c.status == kCgDiverged ? SDX_CONV_DIVERGED
                                               : SDX_CONV_STOPPED;
        s.lin_rel = c.start > 0 ? std::sqrt(c.latest / c.start) : 0.0;
        s.lin_res = std::sqrt(std::max(c.latest, 0.0));
        s.lin_true_rel = true_rel[size_t(e)];
        s.refreshes += c.refreshes;
        s.have_dx = true;
        lin_ok[size_t(e)] = s.lin_conv == SDX_CONV_CONVERGED;
      }
      if (need_dx) {
        launch_fill<double>(stream_, int64_t(nEnv_) * kSlots * kSlotStride, 0.0, acc_.p);
        launch_env_norms<T>(stream_, dt_v_.view, nEnv_, dx_.p, dx_.p, acc_.p);
        launch_acc_finalize<double>(stream_, nEnv_, 2, acc_.p, norms2_.p);
        std::vector<double> m2(size_t(2 * nE));
        get(norms2_, m2);
        for (int e = 0; e < nE; ++e)
          if (trying[size_t(e)]) S[size_t(e)].dx_norm = std::sqrt(m2[size_t(2 * e)]);
      }
      // who steps: a converged solve, a last try, or a first try already projected (its retry
      // would be the same solve); a projected first try whose solve failed counts that solve twice
      std::vector<int32_t> take(size_t(nE), 0), keep_final(size_t(nE), 0);
      for (int e = 0; e < nE; ++e) {
        if (!trying[size_t(e)]) continue;
        const bool last = retried[size_t(e)] != 0;
        const bool same_retry = !last && proj[size_t(e)];
        if (lin_ok[size_t(e)] || last || same_retry) {
          take[size_t(e)] = 1;
          keep_final[size_t(e)] = (last || same_retry) && !gd ? 1 : 0;
          if (same_retry && !lin_ok[size_t(e)]) S[size_t(e)].lin_iters += cg[size_t(e)].iter;
        }
      }
      std::vector<uint8_t> improved;
      std::vector<int> ls_it;
      line_search(dt, sp, take, keep_final, S, improved, ls_it);
      std::vector<int32_t> next(size_t(nE), 0), gd_set(size_t(nE), 0);
      bool any_gd = false;
      for (int e = 0; e < nE; ++e) {
        if (!trying[size_t(e)]) continue;
        EnvState& s = S[size_t(e)];
        const bool last = retried[size_t(e)] != 0;
        const bool same_retry = !last && proj[size_t(e)];
        if (take[size_t(e)]) {
          s.ls_iters += ls_it[size_t(e)];
          if (improved[size_t(e)]) continue;
          if (same_retry && lin_ok[size_t(e)]) {  // the identical retry: same solve, same search
            s.lin_iters += cg[size_t(e)].iter;
            s.ls_iters += ls_it[size_t(e)];
          }
          if (last || same_retry) {
            if (gd) {
              gd_set[size_t(e)] = 1;
              any_gd = true;
            }
            continue;
          }
        }
        // retry with the projected DResidual from where the iteration started
        next[size_t(e)] = 1;
        retried[size_t(e)] = 1;
        proj[size_t(e)] = 1;
        if (s.psd == SDX_PSD_IF_FAIL_ALWAYS) s.psd = SDX_PSD_ALWAYS;
      }
      if (any_gd) {  // gradient-descent fallback along -res from the iteration's start
        put(ls_mask_, gd_set);
        launch_env_copy<T>(stream_, nV_, d_row_env_.p, ls_mask_.p, res0_.p, dx_.p);
        launch_env_copy<T>(stream_, nV_, d_row_env_.p, ls_mask_.p, res0_.p, res_.p);
        std::vector<int32_t> all_final(size_t(nE), 1);
        line_search(dt, sp, gd_set, all_final, S, improved, ls_it);
        for (int e = 0; e < nE; ++e)
          if (gd_set[size_t(e)]) S[size_t(e)].ls_iters += ls_it[size_t(e)];
      }
      trying = next;
      first_try = false;
    }
    // the residual at the new iterate where the line search did not leave it
    // (explosion control, newton_solver.cpp:856-874, and the next iteration)
    std::vector<int32_t> need_res(size_t(nE), 0);
    bool any_res = false;
    for (int e = 0; e < nE; ++e)
      if (S[size_t(e)].active && !S[size_t(e)].res_known) {
        need_res[size_t(e)] = 1;
        any_res = true;
      }
    if (any_res) {
      set_env_mask(need_res);
      assemble(dt, sp, false, 0);
      auto sq = read_env_sq();
      for (int e = 0; e < nE; ++e)
        if (need_res[size_t(e)]) {
          S[size_t(e)].res = std::sqrt(sq[size_t(e)]);
          S[size_t(e)].res_known = true;
        }
    }
    if (sp.explosion_control)
      for (int e = 0; e < nE; ++e) {
        EnvState& s = S[size_t(e)];
        if (s.active && (s.res > sp.rel_div_tol * s.res0 || s.res > sp.abs_div_tol || !std::isfinite(s.res))) {
          s.conv = SDX_CONV_DIVERGED;
          s.active = false;
        }
      }
  }
  // contact query at the step's result, before PostStage may reset diverged environments
  if (contact_query_) {
    const size_t ne = coll_entry_id_.size();
    if (ne) {
      if (wrench_.n < ne * kWrench) wrench_.alloc(ne * kWrench);
      if (wrench_dmin_.n < ne) wrench_dmin_.alloc(ne);
      launch_fill<double>(stream_, int64_t(ne * kWrench), 0.0, wrench_.p);
      launch_fill<long long>(stream_, int64_t(ne), std::numeric_limits<long long>::max(), wrench_dmin_.p);
      set_env_mask(std::vector<int32_t>(size_t(nE), 1));
      launch_contact_wrench<T>(stream_, mesh_, fields_, u_.p, u0_.p, dt, wrench_.p, wrench_dmin_.p);
      wrench_h_.assign(ne * kWrench, 0.0);
      wrench_dmin_h_.assign(ne, 0);
      fetch(wrench_, wrench_h_);
      fetch(wrench_dmin_, wrench_dmin_h_);
      fetch_all();
    }
    contact_ready_ = true;
  }
  // PostStage: v = (u - u0) / dt, reset on divergence (mochi_soft.cpp:135-163)
  std::vector<int32_t> reset(size_t(nE), 0);
  for (int e = 0; e < nE; ++e) reset[size_t(e)] = S[size_t(e)].conv == SDX_CONV_DIVERGED ? 1 : 0;
  put(reset_, reset);
  launch_post_step<T>(stream_, nV_, d_row_env_.p, reset_.p, dt, u0_.p, u_.p, v_.p, hist_.p, 2 * kNewtonHist);
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  // the colliders' targets of this step are the next step's stage start
  for (HostCollider& c : colliders_)
    if (c.p0[0] != c.p[0] || c.p0[1] != c.p[1] || c.p0[2] != c.p[2]) {
      for (int d = 0; d < 3; ++d) c.p0[d] = c.p[d];
      colliders_dirty_ = true;
    }
  flush_timers();
  if (stats) {
    for (int e = 0; e < std::min(nE, n_stats); ++e) {
      const EnvState& s = S[size_t(e)];
      sdx_step_stats& o = stats[e];
      std::memset(&o, 0, sizeof(o));
      o.newton_iters = s.iters;
      o.linear_iters = s.lin_iters;
      o.line_search_iters = s.ls_iters;
      o.convergence = s.conv;
      o.linear_convergence = s.lin_conv;
      o.fp64_refreshes = s.refreshes;
      o.res_norm0 = s.res0;
      o.res_norm = s.res;
      o.linear_rel_res = s.lin_rel;
      o.linear_true_rel_res = s.lin_true_rel;
      o.warm_start_tried = s.ws_tried;
      o.warm_start_accepted = s.ws_accepted;
      o.warm_start_rel_res = s.ws_ratio;
      o.warm_start_history_rel_res = s.ws_ratio_hist;
      o.warm_start_network = s.ws_network;
    }
  }
}

// Line search (line_search.h:76-390) along -dx_ from sol0_ (= u_ of these
// environments on entry, res_ the residual there) for the environments in
// take, all rules in lockstep: each trial moves the searching environments
// to u = sol0 - scale dx and evaluates what their rules read (residual norm,
// objective, res.dx) in one pass.  improved[e] tells whether a trial was
// accepted, ls_it[e] the trials rejected before it (max_iter when none).
// Without improvement an environment keeps its last trial if keep_final[e],
// else returns to sol0.  S[e].res, res_known and merit follow u.
template <class T>
void World<T>::line_search(double dt, const sdx_solver_params& sp, const std::vector<int32_t>& take,
                           const std::vector<int32_t>& keep_final, std::vector<EnvState>& S,
                           std::vector<uint8_t>& improved, std::vector<int>& ls_it) {
  const int nE = nEnv_;
  improved.assign(size_t(nE), 0);
  ls_it.assign(size_t(nE), 0);
  int n = 0;
  for (int e = 0; e < nE; ++e) n += take[size_t(e)] ? 1 : 0;
  if (n == 0) return;
  const int type = sp.line_search_type;
  if (type == SDX_LS_NONE) {  // (:76-87) take the step; the trial count is left as it was
    put(ls_scale_, std::vector<double>(size_t(nE), 1.0));
    put(ls_mask_, take);
    launch_ls_update<T>(stream_, nV_, d_row_env_.p, ls_mask_.p, ls_scale_.p, sol0_.p, dx_.p, u_.p);
    for (int e = 0; e < nE; ++e)
      if (take[size_t(e)]) {
        improved[size_t(e)] = 1;
        S[size_t(e)].res_known = false;
      }
    return;
  }
  // each environment's rule, and its slope res.inc at the start (inc = -dx)
  std::vector<int> rule(size_t(nE), type);
  std::vector<double> slope(size_t(nE), 0.0);
  if (type == SDX_LS_ARMIJO || type == SDX_LS_WOLFE_WEAK || type == SDX_LS_WOLFE_STRONG ||
      type == SDX_LS_ARMIJO_OR_RESIDUAL_NORM) {
    queue_env_dot(res_.p, dx_.p);
    std::vector<double> d(static_cast<size_t>(nE));
    fetch(env_dot_, d);
    fetch_all();
    for (int e = 0; e < nE; ++e) {
      slope[size_t(e)] = -d[size_t(e)];
      // not a descent direction: the residual-norm search (:343-348)
      if (type == SDX_LS_ARMIJO_OR_RESIDUAL_NORM && slope[size_t(e)] > 0) rule[size_t(e)] = SDX_LS_RESIDUAL_NORM;
    }
  }
  const double alpha = sp.line_search_alpha, w1 = sp.line_search_wolfe1, w2 = sp.line_search_wolfe2;
  const double inf = std::numeric_limits<double>::infinity();
  std::vector<double> scale(size_t(nE), 1.0), lo(size_t(nE), 0.0), hi(size_t(nE), inf);
  std::vector<int32_t> searching(take);
  std::vector<uint8_t> res_at_last(size_t(nE), 0);
  const int max_it = sp.line_search_max_iter;
  for (int it = 0; it < max_it; ++it) {
    bool want_res = false, want_obj = false, want_dot = false, any = false;
    for (int e = 0; e < nE; ++e) {
      if (!searching[size_t(e)]) continue;
      any = true;
      const int r = rule[size_t(e)];
      const bool wolfe = r == SDX_LS_WOLFE_WEAK || r == SDX_LS_WOLFE_STRONG;
      want_res |= r == SDX_LS_RESIDUAL_NORM || wolfe || r == SDX_LS_ARMIJO_OR_RESIDUAL_NORM;
      want_obj |= r != SDX_LS_RESIDUAL_NORM;
      want_dot |= wolfe;
    }
    if (!any) break;
    put(ls_scale_, scale);
    put(ls_mask_, searching);
    launch_ls_update<T>(stream_, nV_, d_row_env_.p, ls_mask_.p, ls_scale_.p, sol0_.p, dx_.p, u_.p);
    set_env_mask(searching);
    std::vector<double> sq(static_cast<size_t>(nE)), ob(static_cast<size_t>(nE)), dd(static_cast<size_t>(nE));
    if (want_res) {
      assemble(dt, sp, false, 0);
      fetch(env_sq_, sq);
    }
    if (want_obj) {
      queue_energy(dt, sp);
      fetch(env_obj_, ob);
    }
    if (want_dot) {
      queue_env_dot(res_.p, dx_.p);
      fetch(env_dot_, dd);
    }
    fetch_all();
    flush_timers();
    for (int e = 0; e < nE; ++e) {
      if (!searching[size_t(e)]) continue;
      EnvState& s = S[size_t(e)];
      const int r = rule[size_t(e)];
      const double a = scale[size_t(e)], m = ob[size_t(e)];
      if (want_res) s.res = std::sqrt(sq[size_t(e)]);
      res_at_last[size_t(e)] = want_res ? 1 : 0;
      bool ok = false, backtrack = true;
      switch (r) {
        case SDX_LS_SIMPLE:  // (:89-124)
          ok = m <= s.merit0 + std::fabs(s.merit0 * sp.line_search_max_rel_increase);
          break;
        case SDX_LS_ARMIJO:  // (:126-167)
          ok = m <= s.merit0 + a * w1 * slope[size_t(e)];
          break;
        case SDX_LS_WOLFE_WEAK:
        case SDX_LS_WOLFE_STRONG: {  // (:169-289)
          backtrack = false;
          const double g = -a * dd[size_t(e)];  // res(x).inc with inc = -a dx
          const double t1 = w1 * slope[size_t(e)], t2 = w2 * slope[size_t(e)];
          const bool strong = r == SDX_LS_WOLFE_STRONG;
          const bool bad = !std::isfinite(g) || !std::isfinite(m) || m > s.merit0 + a * t1;
          const bool shrink = bad || (strong && g > std::fabs(t2 * a));
          const bool grow = !shrink && (strong ? g < -std::fabs(t2 * a) : g < t2 * a);
          if (shrink) {
            hi[size_t(e)] = a;
            scale[size_t(e)] = lo[size_t(e)] + alpha * (a - lo[size_t(e)]);
          } else if (grow) {
            lo[size_t(e)] = a;
            scale[size_t(e)] = std::isfinite(hi[size_t(e)]) ? 0.5 * (a + hi[size_t(e)]) : a / alpha;
          } else {
            ok = true;
          }
          break;
        }
        case SDX_LS_ARMIJO_OR_RESIDUAL_NORM:  // (:338-390)
          ok = m <= s.merit0 + a * w1 * slope[size_t(e)] || s.res < s.res_start;
          break;
        default:  // residual norm (:291-336)
          ok = s.res <= s.res_start;
          break;
      }
      if (ok) {
        improved[size_t(e)] = 1;
        ls_it[size_t(e)] = it;
        searching[size_t(e)] = 0;
        if (r != SDX_LS_RESIDUAL_NORM) s.merit = m;
      } else {
        if (backtrack) scale[size_t(e)] *= alpha;
        ls_it[size_t(e)] = it + 1;
      }
    }
  }
  std::vector<int32_t> back(size_t(nE), 0);
  bool any_back = false;
  for (int e = 0; e < nE; ++e) {
    if (!take[size_t(e)]) continue;
    EnvState& s = S[size_t(e)];
    s.res_known = res_at_last[size_t(e)] != 0;
    if (!improved[size_t(e)] && !keep_final[size_t(e)]) {
      back[size_t(e)] = 1;
      any_back = true;
      s.res = s.res_start;
      s.res_known = false;
      s.merit = s.merit0;
    }
  }
  if (any_back) {
    put(ls_mask_, back);
    launch_env_copy<T>(stream_, nV_, d_row_env_.p, ls_mask_.p, sol0_.p, u_.p);
  }
}

template <class T>
void World<T>::get_state(int a, int which, double* out) {
  HostActor& A = actor(a);
  if (!finalized_) {
    const std::vector<double>& src = which == 1 ? A.v : A.u;
    for (int k = 0; k < 3 * A.nV; ++k) out[k] = src[size_t(k)] + (which == 2 ? A.X[size_t(k)] : 0.0);
    return;
  }
  std::vector<T> h(size_t(3 * A.nV));
  const T* src = (which == 1 ? v_.p : u_.p) + size_t(3) * size_t(env_v0_[size_t(a)]);
  hip_check(hipMemcpyAsync(h.data(), src, h.size() * sizeof(T), hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  for (int k = 0; k < 3 * A.nV; ++k) out[k] = double(h[size_t(k)]) + (which == 2 ? A.X[size_t(k)] : 0.0);
}

template <class T>
void World<T>::set_state(int a, int which, const double* in) {
  HostActor& A = actor(a);
  std::vector<double>& dst = which == 1 ? A.v : A.u;
  for (int k = 0; k < 3 * A.nV; ++k) dst[size_t(k)] = in[k];
  if (!finalized_) return;
  std::vector<T> h(size_t(3 * A.nV));
  for (int k = 0; k < 3 * A.nV; ++k) h[size_t(k)] = T(in[k]);
  T* d = (which == 1 ? v_.p : u_.p) + size_t(3) * size_t(env_v0_[size_t(a)]);
  hip_check(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream_), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

template <class T>
void World<T>::actor_bsr_size(int a, int32_t* nrows, int32_t* nnzb) {
  finalize();
  actor(a);
  const int v0 = env_v0_[size_t(a)], v1 = env_v0_[size_t(a) + 1];
  if (nrows) *nrows = v1 - v0;
  if (nnzb) *nnzb = h_row_ptr_[size_t(v1)] - h_row_ptr_[size_t(v0)];
}

template <class T>
void World<T>::assemble_actor(int a, double dt, const sdx_solver_params& sp, double* res, int32_t* ptr, int32_t* col,
                              double* vals) {
  finalize();
  actor(a);
  if (colliders_dirty_) upload_colliders();
  // stage start u0 = 0, v0 = 0 (inertia acts on u itself), as the oracle's assemble entry
  hip_check(hipMemsetAsync(u0_.p, 0, u0_.bytes(), stream_), "hipMemsetAsync");
  hip_check(hipMemsetAsync(v0_.p, 0, v0_.bytes(), stream_), "hipMemsetAsync");
  std::vector<int32_t> mask(size_t(nEnv_), 0);
  mask[size_t(a)] = 1;
  set_env_mask(mask);
  assemble(dt, sp, true, sp.psd_proj_mode == SDX_PSD_ALWAYS ? 1 : 0);
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  const int v0 = env_v0_[size_t(a)], v1 = env_v0_[size_t(a) + 1];
  const int b0 = h_row_ptr_[size_t(v0)], b1 = h_row_ptr_[size_t(v1)];
  if (res) {
    std::vector<T> h(size_t(3 * (v1 - v0)));
    hip_check(hipMemcpy(h.data(), res_.p + size_t(3) * size_t(v0), h.size() * sizeof(T), hipMemcpyDeviceToHost),
              "hipMemcpy");
    for (size_t k = 0; k < h.size(); ++k) res[k] = double(h[k]);
  }
  if (ptr)
    for (int i = v0; i <= v1; ++i) ptr[i - v0] = h_row_ptr_[size_t(i)] - b0;
  if (col)
    for (int k = b0; k < b1; ++k) col[k - b0] = h_col_[size_t(k)] - v0;
  if (vals) {
    std::vector<T> h(size_t(b1 - b0));
    for (int c = 0; c < 9; ++c) {
      hip_check(hipMemcpy(h.data(), vals_.p + size_t(c) * size_t(nB_) + size_t(b0), h.size() * sizeof(T),
                          hipMemcpyDeviceToHost),
                "hipMemcpy");
      for (size_t k = 0; k < h.size(); ++k) vals[9 * k + size_t(c)] = double(h[k]);
    }
  }
}

// Objective of one actor at its current displacement with stage start
// u0 = v0 = 0, as the assembly entry above.
template <class T>
double World<T>::objective_actor(int a, double dt, const sdx_solver_params& sp) {
  finalize();
  actor(a);
  if (colliders_dirty_) upload_colliders();
  hip_check(hipMemsetAsync(u0_.p, 0, u0_.bytes(), stream_), "hipMemsetAsync");
  hip_check(hipMemsetAsync(v0_.p, 0, v0_.bytes(), stream_), "hipMemsetAsync");
  std::vector<int32_t> mask(size_t(nEnv_), 0);
  mask[size_t(a)] = 1;
  set_env_mask(mask);
  queue_energy(dt, sp);
  std::vector<double> ob(static_cast<size_t>(nEnv_));
  fetch(env_obj_, ob);
  fetch_all();
  return ob[size_t(a)];
}

// AddBoundaryConditionDofsWorld (mochi_actor.h:187-190) and ClearBoundaryConditions
// (:211): a changed set of pinned DoFs rebuilds the gather items at the next step;
// new values for the same set only replace the targets.
template <class T>
void World<T>::set_boundary_conditions(int a, const int32_t* dofs, const double* values, int n, bool clear) {
  HostActor& A = actor(a);
  const int nd = 3 * A.nV;
  for (int i = 0; i < n; ++i) {
    if (dofs[i] < 0 || dofs[i] >= nd) throw std::invalid_argument("boundary-condition DoF out of range");
    if (!std::isfinite(values[i])) throw std::invalid_argument("boundary-condition values must be finite");
  }
  std::vector<uint8_t> pinned = clear ? A.perm : A.pinned;
  std::vector<double> pv = A.pin_value;
  if (clear)
    for (int k = 0; k < nd; ++k) pv[size_t(k)] = A.perm[size_t(k)] ? A.perm_value[size_t(k)] : 0.0;
  for (int i = 0; i < n; ++i) {
    pinned[size_t(dofs[i])] = 1;
    pv[size_t(dofs[i])] = values[i] - A.X[size_t(dofs[i])];
  }
  const bool same = pinned == A.pinned;
  if (!same) {
    invalidate();  // device state back to the host mirrors first
    A.pinned = pinned;
    A.pin_value = pv;
    return;
  }
  A.pin_value = pv;
  if (!finalized_) return;
  std::vector<T> h(pv.size());
  for (size_t k = 0; k < pv.size(); ++k) h[k] = T(pv[k]);
  T* d = d_pin_value_.p + size_t(3) * size_t(env_v0_[size_t(a)]);
  hip_check(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream_), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

template <class T>
void World<T>::set_contact_query(bool on) {
  contact_query_ = on;
  contact_ready_ = false;
}

template <class T>
void World<T>::contact_wrench(int a, int c, double out[7]) {
  actor(a);
  if (c < 0 || c >= int(colliders_.size())) throw std::out_of_range("no such collider");
  if (!contact_query_) throw std::logic_error("the contact query is not registered");
  if (!contact_ready_ || !finalized_)
    throw std::logic_error("contact query results are available after the next simulation step");
  for (int k = coll_first_h_[size_t(a)]; k < coll_first_h_[size_t(a) + 1]; ++k) {
    if (coll_entry_id_[size_t(k)] != c) continue;
    for (int j = 0; j < kWrench; ++j) out[j] = wrench_h_[size_t(k) * kWrench + size_t(j)];
    long long b = wrench_dmin_h_[size_t(k)];
    if (b == std::numeric_limits<long long>::max()) {
      out[6] = std::numeric_limits<double>::infinity();
    } else {
      if (b < 0) b ^= 0x7fffffffffffffffLL;
      std::memcpy(&out[6], &b, sizeof(double));
    }
    return;
  }
  throw std::invalid_argument("the collider does not act on this actor");
}

template <class T>
void World<T>::set_schedule(const sdx_schedule& sc) {
  const bool lanes_changed = sc.spmv_lanes != sched_.spmv_lanes;
  sched_ = sc;
  drop_graphs();
  group_plan_n_ = 0;  // the groups' tile ranges follow the tiles
  max_env_a_ = 0;
  if (finalized_ && lanes_changed) {
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    BuildTiles(env_v0_, nEnv_, kSpmvThreads / sched_.spmv_lanes, tiles_a_);
    upload_tiles(tiles_a_, dt_a_);
    alloc_lists();
  }
  if (finalized_) plan_assembly(asm_groups(sched_));
}

// Coordinate search over the schedule: each candidate runs `trials` timed
// steps (after one untimed step that captures its graphs) from the same saved
// state; the state is restored afterwards.
template <class T>
double World<T>::autotune(double dt, const sdx_solver_params& sp_in, int trials) {
  finalize();
  if (colliders_dirty_) upload_colliders();
  sdx_solver_params sp = sp_in;
  sp.schedule = 0;
  const int64_t n3 = int64_t(3) * nV_;
  std::vector<DBufBase*> reg;
  DBuf<T> su(reg, &stream_), sv(reg, &stream_), sh(reg, &stream_);
  const int64_t nh = int64_t(2 * kNewtonHist) * n3;
  su.alloc(size_t(n3));
  sv.alloc(size_t(n3));
  sh.alloc(size_t(nh));
  launch_copy<T>(stream_, n3, u_.p, su.p);
  launch_copy<T>(stream_, n3, v_.p, sv.p);
  launch_copy<T>(stream_, nh, hist_.p, sh.p);
  const int64_t coll_n0 = coll_n_;
  const uint32_t coll_salt0 = coll_salt_;
  const WarmGate wg0 = wg_;
  const std::vector<HostCollider> colliders0 = colliders_;  // a trial step advances their stage start
  auto restore = [&]() {
    colliders_ = colliders0;
    colliders_dirty_ = true;
    wg_ = wg0;
    launch_copy<T>(stream_, n3, su.p, u_.p);
    launch_copy<T>(stream_, n3, sv.p, v_.p);
    launch_copy<T>(stream_, nh, sh.p, hist_.p);
    coll_n_ = coll_n0;  // trial steps add no training samples
    coll_salt_ = coll_salt0;
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  };
  auto measure = [&](const sdx_schedule& c) {
    set_schedule(c);
    double best = std::numeric_limits<double>::infinity();
    for (int t = 0; t <= trials; ++t) {
      restore();
      const auto t0 = std::chrono::steady_clock::now();
      step(dt, sp, nullptr, 0);
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (t > 0) best = std::min(best, ms);
    }
    return best;
  };
  sdx_schedule cur = sched_;
  double cur_ms = measure(cur);
  auto consider = [&](sdx_schedule c) {
    const double ms = measure(c);
    if (ms < 0.98 * cur_ms) {
      cur = c;
      cur_ms = ms;
    }
  };
  for (int L : {4, 8, 16})
    if (L != cur.spmv_lanes) {
      sdx_schedule c = cur;
      c.spmv_lanes = L;
      consider(c);
    }
  // the whole solve as one launch per solve, where every environment fits
  // a workgroup's LDS; then its lanes per row, and the burst schedule only
  // where the solve still runs as bursts
  if (env_cg_fits()) {
    for (int ec : {0, 1, 2})
      if (ec != cur.env_cg) {
        sdx_schedule c = cur;
        c.env_cg = ec;
        consider(c);
      }
    if (cur.env_cg > 0)
      for (int L : {4, 8, 16})
        if (L != cur.spmv_lanes) {
          sdx_schedule c = cur;
          c.spmv_lanes = L;
          consider(c);
        }
  }
  const bool bursts = !(cur.env_cg > 0 && env_cg_fits());
  if (bursts) {
    sdx_schedule c = cur;
    c.active_lists = 1 - cur.active_lists;
    consider(c);
  }
  for (int eb : {64, 128, 256})
    if (eb != cur.element_block) {
      sdx_schedule c = cur;
      c.element_block = eb;
      consider(c);
    }
  for (int g : {1, 0})
    for (int B : {4, 8, 16, 32})
      if (bursts && (g != cur.use_graphs || B != cur.cg_burst)) {
        sdx_schedule c = cur;
        c.use_graphs = g;
        c.cg_burst = B;
        consider(c);
      }
  if (bursts && cur.active_lists && nEnv_ > 1) {
    const sdx_schedule base = cur;
    for (int ng : {1, 2, 3, 4, 6, 8})
      if (ng <= nEnv_ && ng != cg_groups(base)) {
        sdx_schedule c = base;
        c.cg_groups = ng;
        consider(c);
      }
    {
      sdx_schedule c = cur;
      c.fused_tail = 1 - cur.fused_tail;
      consider(c);
    }
    if (cg_groups(cur) > 1) {
      const sdx_schedule chosen = cur;
      for (int gi : {8, 24, 32})
        if (gi != chosen.group_iters) {
          sdx_schedule c = chosen;
          c.group_iters = gi;
          consider(c);
        }
    }
  }
  if (nEnv_ > 1 && !cur.fused_asm) {  // assembly groups: one (every record kept), the automatic count, half and twice it
    const int a = asm_groups(cur);
    std::vector<int> cand = {1, std::max(1, a / 2), asm_groups(sdx_schedule{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}), 2 * a};
    for (int ag : cand)
      if (ag <= nEnv_ && ag != asm_groups(cur)) {
        sdx_schedule c = cur;
        c.asm_groups = ag;
        consider(c);
      }
  }
  set_schedule(cur);
  restore();
  tuned_ = true;
  return cur_ms;
}

template <class T>
void World<T>::info(int64_t out[8]) {
  finalize();
  std::memset(out, 0, sizeof(int64_t) * 8);
  out[0] = nV_;
  out[1] = nT_;
  out[2] = nB_;
  out[3] = nF_;
  out[4] = nEnv_;
  out[5] = int64_t(tiles_a_.env.size());
  out[6] = int64_t(tiles_v_.env.size());
  const int groups = sched_.active_lists ? cg_groups(sched_) : 1;
  if (groups > 1 && sched_.group_iters > 0) {
    if (group_plan_n_ != groups) plan_groups(groups);
    out[7] = int64_t(group_a_.size()) - 1;
  } else {
    out[7] = 1;
  }
}

// ---- state snapshots ----------------------------------------------------


template <class T>
typename World<T>::Snapshot& World<T>::snapshot(int32_t handle) {
  auto it = snaps_.find(handle);
  if (it == snaps_.end()) throw std::out_of_range("no such state handle");
  return *it->second;
}

template <class T>
std::unique_ptr<typename World<T>::Snapshot> World<T>::take_snapshot() {
  finalize();
  const int64_t n3 = int64_t(3) * nV_, nh = int64_t(2 * kNewtonHist) * n3;
  auto s = std::make_unique<Snapshot>(&stream_);
  s->u.alloc(size_t(n3));
  s->v.alloc(size_t(n3));
  s->hist.alloc(size_t(nh));
  launch_copy<T>(stream_, n3, u_.p, s->u.p);
  launch_copy<T>(stream_, n3, v_.p, s->v.p);
  launch_copy<T>(stream_, nh, hist_.p, s->hist.p);
  s->coll_p.resize(6 * colliders_.size());
  for (size_t c = 0; c < colliders_.size(); ++c)
    for (int d = 0; d < 3; ++d) {
      s->coll_p[6 * c + size_t(d)] = colliders_[c].p[d];
      s->coll_p[6 * c + 3 + size_t(d)] = colliders_[c].p0[d];
    }
  s->wg = wg_;
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  return s;
}

template <class T>
void World<T>::put_snapshot(const Snapshot& s) {
  finalize();
  const int64_t n3 = int64_t(3) * nV_, nh = int64_t(2 * kNewtonHist) * n3;
  if (s.u.n != size_t(n3) || s.hist.n != size_t(nh) || s.coll_p.size() != 6 * colliders_.size())
    throw std::runtime_error("state was captured from a world of another size");
  launch_copy<T>(stream_, n3, s.u.p, u_.p);
  launch_copy<T>(stream_, n3, s.v.p, v_.p);
  launch_copy<T>(stream_, nh, s.hist.p, hist_.p);
  for (size_t c = 0; c < colliders_.size(); ++c)
    for (int d = 0; d < 3; ++d) {
      colliders_[c].p[d] = s.coll_p[6 * c + size_t(d)];
      colliders_[c].p0[d] = s.coll_p[6 * c + 3 + size_t(d)];
    }
  colliders_dirty_ = true;
  wg_ = s.wg;
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

template <class T>
int32_t World<T>::capture_state() {
  auto s = take_snapshot();
  if (next_snap_ == std::numeric_limits<int32_t>::max()) throw std::runtime_error("state handles exhausted");
  const int32_t h = next_snap_++;
  snaps_[h] = std::move(s);
  return h;
}

template <class T>
void World<T>::restore_state(int32_t handle, bool release) {
  put_snapshot(snapshot(handle));
  if (release) snaps_.erase(handle);
}

template <class T>
void World<T>::release_state(int32_t handle) {
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  if (handle == 0) {
    snaps_.clear();
    return;
  }
  if (snaps_.erase(handle) == 0) throw std::out_of_range("no such state handle");
}

// Equal values (NaN equal to NaN) of every captured field.
template <class T>
bool World<T>::equal_states(int32_t a, int32_t b) {
  const Snapshot& A = snapshot(a);
  const Snapshot& B = snapshot(b);
  if (A.u.n != B.u.n || A.hist.n != B.hist.n || A.coll_p != B.coll_p) return false;
  if (std::memcmp(&A.wg, &B.wg, sizeof(WarmGate)) != 0) return false;
  auto same = [](const DBuf<T>& x, const DBuf<T>& y) {
    std::vector<T> hx(x.n), hy(y.n);
    x.download(hx);
    y.download(hy);
    for (size_t i = 0; i < hx.size(); ++i)
      if (!(hx[i] == hy[i] || (hx[i] != hx[i] && hy[i] != hy[i]))) return false;
    return true;
  };
  return same(A.u, B.u) && same(A.v, B.v) && same(A.hist, B.hist);
}

template <class T>
size_t World<T>::state_size() {
  finalize();
  // header, the gate block, three actor blocks (a length per actor), the
  // collider block and the footer
  const size_t n3 = size_t(3) * size_t(nV_), na = actors_.size(), nc = colliders_.size();
  size_t n = 24 + 24 + size_t(capture::kGateCounters) * 4 + 4;
  if (na) n += 3 * (24 + 8 * na) + (2 + capture::kHistPlanes) * n3 * sizeof(T);
  if (nc) n += 24 + 6 * nc * sizeof(double);
  return n;
}

namespace {
// FNV-1a over what decides how a state steps: every actor's counts, rest
// positions, tets, material (homogeneous parameters and per-element fields)
// and permanently pinned DoFs with their values (clearable boundary
// conditions are targets that change from step to step); and every
// collider's kind, attachment, normal and radius (collider positions are part
// of the state itself).
class TopologyHash {
 public:
  void bytes(const void* p, size_t n) { h_ = capture::fnv1a(p, n, h_); }
  template <class V>
  void vec(const std::vector<V>& v) {
    const uint64_t n = v.size();
    bytes(&n, sizeof(n));
    if (n) bytes(v.data(), v.size() * sizeof(V));
  }
  template <class V>
  void pod(const V& v) {
    bytes(&v, sizeof(V));
  }
  uint64_t value() const { return h_; }

 private:
  uint64_t h_ = 1469598103934665603ull;
};

uint64_t actor_hash(const std::vector<HostActor>& actors) {
  TopologyHash h;
  for (const HostActor& a : actors) {
    h.pod(a.nV);
    h.pod(a.nT);
    h.vec(a.X);
    h.vec(a.tets);
    h.pod(a.params);
    h.vec(a.E);
    h.vec(a.nu);
    h.vec(a.perm);
    h.vec(a.perm_value);
  }
  return h.value();
}

uint64_t collider_hash(const std::vector<HostCollider>& colliders) {
  TopologyHash h;
  for (const HostCollider& c : colliders) {
    h.pod(c.kind);
    h.pod(c.actor);
    h.bytes(c.n, sizeof(c.n));
    h.pod(c.r);
  }
  return h.value();
}
}  // namespace

// The byte form: the container of include/sdx_fem.h's state formats, built
// from one download of each device array.
template <class T>
void World<T>::capture_bytes(uint8_t* out, size_t cap) {
  const size_t need = state_size();
  if (cap < need) throw std::invalid_argument("state buffer too small (sdx_state_size gives the size)");
  static_assert(2 * kNewtonHist == capture::kHistPlanes, "the state holds two previous steps per Newton iteration");
  static_assert(3 * kNewtonHist == capture::kGateCounters, "the state holds three gate counters per Newton iteration");
  capture::StateParts sp;
  sp.real_bytes = uint32_t(sizeof(T));
  sp.actor_hash = actor_hash(actors_);
  sp.collider_hash = collider_hash(colliders_);
  const size_t n3 = size_t(3) * size_t(nV_);
  for (size_t a = 0; a < actors_.size(); ++a) sp.actor_dofs.push_back(3 * int64_t(actors_[a].nV));
  // u and v straight into the state's arrays; the history planes [plane][3 nV]
  // regrouped per actor [plane][3 nV_a]
  sp.u.resize(n3 * sizeof(T));
  u_.download_raw(sp.u.data(), n3);
  sp.v.resize(n3 * sizeof(T));
  v_.download_raw(sp.v.data(), n3);
  std::vector<T> hh(size_t(capture::kHistPlanes) * n3);
  hist_.download_raw(hh.data(), hh.size());
  sp.hist.resize(hh.size() * sizeof(T));
  size_t o = 0;
  for (size_t a = 0; a < actors_.size(); ++a) {
    const size_t d = size_t(3 * actors_[a].nV), v0 = size_t(3 * env_v0_[a]);
    for (int p = 0; p < capture::kHistPlanes; ++p, o += d)
      std::memcpy(sp.hist.data() + o * sizeof(T), hh.data() + size_t(p) * n3 + v0, d * sizeof(T));
  }
  for (const HostCollider& c : colliders_) {
    sp.colliders.insert(sp.colliders.end(), c.p, c.p + 3);
    sp.colliders.insert(sp.colliders.end(), c.p0, c.p0 + 3);
  }
  std::memcpy(sp.gate, wg_.skip, sizeof(wg_.skip));
  std::memcpy(sp.gate + kNewtonHist, wg_.seen, sizeof(wg_.seen));
  std::memcpy(sp.gate + 2 * kNewtonHist, wg_.net_skip, sizeof(wg_.net_skip));
  if (capture::encoded_size(sp) != need) throw std::logic_error("state size does not match its encoding");
  capture::encode_into(sp, out);
}

// Restore the components of a byte form that `excluded` (SDX_STATE_* bits)
// does not name (RestorePartialState, mochi_capture.cpp:522-566); every
// component is checked against this world first, so a rejected state leaves
// the world as it was.
template <class T>
void World<T>::restore_bytes(const uint8_t* in, size_t n, uint32_t excluded) {
  finalize();
  uint32_t present = 0;
  const capture::StateParts sp = capture::decode(in, n, &present);
  const size_t n3 = size_t(3) * size_t(nV_);
  const bool has_actors = !actors_.empty();
  std::vector<int64_t> dofs;
  for (const HostActor& a : actors_) dofs.push_back(3 * int64_t(a.nV));
  const uint32_t actor_kinds = capture::kDisplacements | capture::kVelocities;
  const bool actors_ok = has_actors ? ((present & actor_kinds) == actor_kinds && sp.real_bytes == sizeof(T) &&
                                       sp.actor_dofs == dofs && sp.actor_hash == actor_hash(actors_))
                                    : (present & actor_kinds) == 0;
  const bool colliders_ok = colliders_.empty() ? (present & capture::kColliderTargets) == 0
                                               : ((present & capture::kColliderTargets) &&
                                                  sp.colliders.size() == 6 * colliders_.size() &&
                                                  sp.collider_hash == collider_hash(colliders_));
  if (!actors_ok || !colliders_ok)
    throw std::runtime_error("state was captured from a world of another precision or topology");
  if (!(excluded & capture::kColliderTargets))
    for (size_t c = 0; c < colliders_.size(); ++c) {
      std::memcpy(colliders_[c].p, sp.colliders.data() + 6 * c, 3 * sizeof(double));
      std::memcpy(colliders_[c].p0, sp.colliders.data() + 6 * c + 3, 3 * sizeof(double));
      colliders_dirty_ = true;
    }
  if (has_actors) {
    // one-off transfers: plain uploads, so the pinned staging ring keeps its size
    if (!(excluded & capture::kDisplacements)) u_.upload_raw(sp.u.data(), n3);
    if (!(excluded & capture::kVelocities)) v_.upload_raw(sp.v.data(), n3);
    if (!(excluded & capture::kWarmStart)) {
      std::vector<T> hh(size_t(capture::kHistPlanes) * n3);
      size_t o = 0;
      for (size_t a = 0; a < actors_.size(); ++a) {
        const size_t d = size_t(3 * actors_[a].nV), v0 = size_t(3 * env_v0_[a]);
        for (int p = 0; p < capture::kHistPlanes; ++p, o += d)
          std::memcpy(hh.data() + size_t(p) * n3 + v0, sp.hist.data() + o * sizeof(T), d * sizeof(T));
      }
      hist_.upload_raw(hh.data(), hh.size());
    }
  }
  if (!(excluded & capture::kWarmStart)) {
    std::memcpy(wg_.skip, sp.gate, sizeof(wg_.skip));
    std::memcpy(wg_.seen, sp.gate + kNewtonHist, sizeof(wg_.seen));
    std::memcpy(wg_.net_skip, sp.gate + 2 * kNewtonHist, sizeof(wg_.net_skip));
  }
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

template class World<float>;
template class World<double>;

}  // namespace sdx
