// Launch interface of the HIP kernels of the soft-body FEM step.
// Host code (sdx_world.cpp, sdx_pcg.cpp) only calls these wrappers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sdx_layout.h"

namespace sdx {

// Per-environment CG state, one per island, in device memory.
enum : int32_t { kCgActive = 0, kCgConverged = 1, kCgStopped = 2, kCgDiverged = 3, kCgIdle = 4 };

struct EnvCG {
  double start;     // stop-test scaling value (squared norm of b in the chosen norm)
  double latest;    // squared stop-test norm when the solve ended
  double rz;        // r.z when the solve ended
  double gamma[2];  // r_k.z_k of iteration k in slot k % 2
  double alpha[2];
  double mu[2];     // p_k.A p_k
  int32_t iter;     // completed CG iterations
  int32_t status;   // kCg*
  int32_t refreshes;
  int32_t pad;
};

struct CgConfig {
  int32_t n_env;          // environments in the world (accumulator layout)
  int32_t norm_type;      // SDX_NORM_*
  int32_t prec;           // SDX_PREC_*
  int32_t max_iter;
  int32_t refresh_every;  // 0: never
  int32_t abort_if_not_spd;
  int32_t n_rows;         // block rows of the matrix (size of z)
  int32_t pad1;
  double abs_tol, rel_tol, rel_div_tol;
  // mixed precision (T = float, accumulator double): the iterate is x + xd,
  // xd [3 n_rows] T holding the increments since the last refresh iteration,
  // which folds them into x (x is read and written in FP64 only then); NULL:
  // every update adds into x
  void* xd;
  // per-environment relative tolerance [n_env] (the Newton forcing terms);
  // NULL: rel_tol for every environment
  const double* env_rel_tol;
};

// Iteration k of the CG replaces the recursive residual by the FP64 true
// residual (formed in the matvec of iteration k) every refresh_every iterations.
__host__ __device__ inline bool cg_refresh(const CgConfig& cfg, int k) {
  return cfg.refresh_every > 0 && ((k + 1) % cfg.refresh_every) == 0;
}

// Work list of tiles (of environments still active) for the persistent CG
// kernels; list == nullptr: one tile per workgroup over all tiles.
struct WorkList {
  const int32_t* list;
  const int32_t* count;
};

// Event pair recorded around one kernel dispatch.
struct KernelEvents {
  hipEvent_t start, stop;
};

// Tiles: contiguous row ranges that never cross an environment.
struct Tiles {
  const int32_t* env;    // [n] environment of the tile
  const int32_t* row0;   // [n] first row
  const int32_t* row1;   // [n] one past the last row
  const int32_t* first;  // [nEnv+1] tile range of each environment
  int32_t n;
};

constexpr int kSpmvLanes = 16;    // lanes per block row of the plain matvec (CG matvec: 4, 8 or 16)
constexpr int kSpmvThreads = 256;
constexpr int kSpmvChunk = 16;    // blocks of a row whose loads the CG matvec issues together
constexpr int kSpmvRows = kSpmvThreads / kSpmvLanes;  // 16 rows per plain matvec tile
constexpr int kVecThreads = 256;  // rows per vector tile
// Cross-block accumulators: per environment and set, kSlots lines of
// kSlotStride values (one 128-byte line per slot for doubles).
constexpr int kSlots = 16;
constexpr int kSlotStride = 16;
constexpr int kCgSets = 6;

// Whole-solve CG (k_cg_env): one workgroup of kEnvThreads runs every iteration
// of one environment at a time, taking environments from a queue.
constexpr int kEnvThreads = 1024;
constexpr int kEnvWaves = kEnvThreads / 64;
constexpr int kEnvLdsMax = 160 * 1024;  // LDS bytes of a CU: no workgroup takes more
struct EnvQueue {
  const int32_t* order;  // environments in the order workgroups take them (largest first)
  int32_t n;             // entries of order
  int32_t max_rows;      // LDS rows per workgroup: at least every queued environment's rows
  int32_t* next;         // queue position, zero before the launch
  const int32_t* v0;     // [nEnv + 1] first row of each environment
};
// LDS of a workgroup: z [max_rows] (as Z4, or as 3 values with the CG
// vectors in LDS), then `planes` vectors of 3 max_rows values (w; with the CG
// vectors in LDS also r, p, s and the iterate's increments), then the two
// reduction areas (kEnvWaves partials of 1 and of 4 sums).
constexpr int kEnvPlanes = 1, kEnvPlanesVec = 5;
template <class T>
__host__ __device__ inline size_t env_lds_plane_offset(int max_rows, int plane, bool vec) {
  return size_t(max_rows) * (vec ? 3 : 4) * sizeof(T) + size_t(plane) * size_t(max_rows) * 3 * sizeof(T);
}
template <class T>
__host__ __device__ inline size_t env_lds_red_offset(int max_rows, bool vec) {
  return (env_lds_plane_offset<T>(max_rows, vec ? kEnvPlanesVec : kEnvPlanes, vec) + 15) & ~size_t(15);
}
template <class T, class A>
__host__ __device__ inline size_t env_lds_bytes(int max_rows, bool vec) {
  return env_lds_red_offset<T>(max_rows, vec) + size_t(kEnvWaves) * 5 * sizeof(A);
}

// Geometry planes of a tet, stored SoA with stride nT.
// kGBeta: stiffness damping coefficient beta [s]; kGBetaGeo: 1 where its tangent
// takes the geometric block.
enum : int {
  kGdN = 0,
  kGVol = 9,
  kGMu = 10,
  kGLambda = 11,
  kGMass = 12,
  kGAlpha = 13,
  kGBeta = 14,
  kGBetaGeo = 15,
  kGeomPlanes = 16
};

struct MeshView {
  int32_t nV, nT, nF, nB, nEnv;
  const int32_t* tets;      // [4 nT]
  const int32_t* tet_env;   // [nT]
  const int32_t* tris;      // [3 nF]
  const int32_t* tri_env;   // [nF]
  const int32_t* row_env;   // [nV]
  const int32_t* row_ptr;   // [nV+1]
  const int32_t* col;       // [nB]
  const int32_t* slot_row;  // [nB]
  const int32_t* diag_slot; // [nV]
  const int4* slot_desc;    // [nB] per BSR slot: first and end record, row | pin bits << kItemRowBits,
                            // environment | kind << 30 (kSlotUpper / kSlotLower / kSlotDiag)
  const int32_t* eptr;      // [nB+1] off-diagonal records of each upper slot (pair-sorted)
  int32_t nP;               // element patches
  const int32_t* patch_t0;  // [nP+1] tets of patch p: [patch_t0[p], patch_t0[p+1])
  const int32_t* patch_meta_ptr; // [nP+1] description of patch p: words [ptr[p], ptr[p+1]) of patch_meta
  const int32_t* patch_meta_mid; // [nP] first word of its pair part
  const uint32_t* patch_meta;    // patch descriptions (layout at kPatchMetaWords)
  const int32_t* pos_vc;    // [3 nF] vertex-sorted record of contact contribution (triangle, a)
  const int32_t* pos_ec;    // [3 nF] pair-sorted record of pair k of a boundary triangle
  uint8_t* coll_flags;      // [nF] bit 0 / bit 1: the triangle's residual / Hessian contact records
                            // hold nonzero values (or were never written)
  const int32_t* vtx_ptr;   // [nV+1] records (element and contact) of each vertex
  const uint8_t* pinned;    // [3 nV]
  const int32_t* coll_first; // [nEnv+1] into coll
  const double* coll;        // kCollStride doubles per collider: kind, p(3), n(3), r, p0(3), 0
  const int32_t* env_mask;   // [nEnv] 1: assemble/solve this environment
};

struct ContactParamsDev {
  double k, h, thr, max_align;
};

// Doubles per collider: kind, position p (the step's target), plane normal n,
// sphere radius r, stage-start position p0, padding.
constexpr int kCollStride = 12;
// Contact parameters per environment: penalty k, smoothing half-width h,
// threshold, max normal alignment, Coulomb mu, viscous friction, friction
// falloff velocity, normal damping.
constexpr int kContactParams = 8;

template <class T>
struct FieldPtrs {
  const T* X;   // rest positions [3nV]
  const T* geom;  // [kGeomPlanes * nT]
  const double* contact;  // [kContactParams nEnv]
};

// ---- assembly ----
// Element and contact residuals and Hessians are written to their
// destination order (records).  The tets of an environment form patches of
// up to kPatchTets consecutive tets; a patch sums its tets' contributions to
// each of its vertices and vertex pairs in LDS and writes one record per
// local vertex (residual in Rv [.][3], diagonal block in Dv [.][6]: 00 11 22
// 01 02 12) and one per local pair (off-diagonal block in Ev [.][9],
// row-major for the (min, max) vertex pair).  Boundary triangle f writes its
// own records at pos_vc[3 f + a] and pos_ec[3 f + k].  Every vertex / upper
// slot owns one contiguous run of records (one per patch that touches it,
// then the contact records).  Each array is padded by one 16-byte vector for
// the gathers' vector loads.
// (patch descriptions and the local pair order: sdx_layout.h)
constexpr int kItemRowBits = 26;  // rows of the BSR slots below 2^26
// Kinds of a BSR slot for the gather: an upper slot (i < j) sums its pair's
// records, a lower slot their transposes, a diagonal slot its vertex's
// diagonal-block records.
enum : int32_t { kSlotUpper = 0, kSlotLower = 1, kSlotDiag = 2 };
//
// An assembly runs over a contiguous range of environments (an assembly
// group): its tets [t0, t1) in patches [p0, p1), boundary triangles [f0, f1),
// BSR slots [s0, s1) and vector tiles [tile0, tile1).  Record indices stay global; the record pointers a
// launch takes are the group's buffers shifted by the group's first record,
// so a group writes and reads only its own records.  contact_always: every
// boundary triangle writes its contact records (zeros out of contact),
// because the records of the previous assembly of these triangles may have
// been overwritten by another group; 0: a triangle skips writing records
// that are known to hold zeros (its flag bits).
struct AsmRange {
  int32_t t0, t1, f0, f1, s0, s1, tile0, tile1;
  int32_t contact_always;
  int32_t p0, p1;
};
template <class T>
void launch_element(hipStream_t s, const MeshView& m, const AsmRange& r, const FieldPtrs<T>& f, const T* u,
                    const T* u0, const T* v0, double dt, const double grav[3], int wantH, int project, T* Rv, T* Dv,
                    T* Ev, int block = 128, const KernelEvents* ev = nullptr, bool damp = false,
                    const int32_t* plist = nullptr);
// u0: the stage-start displacement (relative sliding of the samples).
template <class T>
void launch_contact(hipStream_t s, const MeshView& m, const AsmRange& r, const FieldPtrs<T>& f, const T* u,
                    const T* u0, T dt, int wantH, T* Rv, T* Dv, T* Ev);
// BSR values (and the preconditioner blocks) from the records, over the
// slots of the range.
template <class T>
void launch_gather_bsr(hipStream_t s, const MeshView& m, const AsmRange& r, const T* Dv, const T* Ev, int prec,
                       T* vals, T* dinv, const KernelEvents* ev = nullptr);
// Fused assembly over row blocks (sdx_rowblocks.h): a workgroup per block
// computes every tet and boundary triangle touching its rows, sums their
// contributions to those rows in LDS and writes the rows' BSR values (with
// the Dirichlet rows and columns), the preconditioner blocks, the residual
// (Dirichlet rows zeroed) and the environment's ||res||^2 into acc (set 0,
// zeroed by the caller).  Environments with env_mask 0 are skipped.
struct RowBlockView {
  int32_t n;            // blocks
  int32_t max_slots;    // BSR slots of the largest block (its LDS)
  const int32_t* row0;  // [n+1] rows of block b: [row0[b], row0[b+1]) of rows
  const int32_t* tet0;  // [n+1] tets of block b in tets
  const int32_t* tri0;  // [n+1] boundary triangles of block b in tris
  const int32_t* env;   // [n]
  const int32_t* rows;  // global vertex of each block row (ascending within a block)
  const int32_t* rbase; // first LDS slot of each block row
  const int2* tets;     // (global tet, block-local row of each vertex: 8 bits each, 0xFF: another block)
  const int2* tris;     // (global triangle, block-local rows of its 3 vertices)
};
// LDS bytes of the fused assembly's workgroup for blocks of at most max_slots
// slots: values [slots][9], residual [rows][3], columns [slots], per row its
// first global slot, LDS base (plus the end) and vertex, the row of each slot.
constexpr int kRowBlockRowsMax = 64;
template <class T>
inline size_t rows_lds_bytes_host(int max_slots) {
  return size_t(max_slots) * 9 * sizeof(T) + size_t(kRowBlockRowsMax) * 3 * sizeof(T) + size_t(max_slots) * 4 +
         size_t(3 * kRowBlockRowsMax + 1) * 4 + size_t(max_slots);
}
template <class T, class A>
void launch_assemble_rows(hipStream_t s, const MeshView& m, const RowBlockView& rb, const FieldPtrs<T>& f,
                          const T* u, const T* u0, const T* v0, double dt, const double grav[3], int project,
                          int prec, T* vals, T* dinv, T* res, A* acc, bool damp = false,
                          const KernelEvents* ev = nullptr);
// acc: set 0 of the accumulators (zeroed by the caller); per-env ||res||^2.
template <class T, class A>
void launch_gather_res(hipStream_t s, const MeshView& m, const AsmRange& r, const Tiles& tv, const T* Rv, T* res,
                       A* acc);
template <class A>
void launch_fill(hipStream_t s, int64_t n, A v, A* d);
// out[e * nv + k] = sum over slots of set 0, k < nv.
template <class A>
void launch_acc_finalize(hipStream_t s, int nEnv, int nv, const A* acc, double* out);

// ---- CG ----
template <class A>
void launch_cg_reset(hipStream_t s, int nEnv, const int32_t* solve, EnvCG* env, A* acc);
// Vectors of the CG are [3 nV] except z, stored [4 nV] (padded per vertex).
// r0, z0, scaling and R_0.  warm (may be NULL): environments that keep x as
// the initial guess and take r0 from r64 = b - A x (launch_residual64);
// alpha (may be NULL): per-environment factor applied to the kept guess;
// v, rv, beta, w, rw, gamma (all NULL or none): two more guesses with their
// residuals (launch_residual64_multi) and factors, x0 = alpha x + beta v +
// gamma w.
template <class T, class A>
void launch_cg_init(hipStream_t s, const Tiles& tv, const CgConfig& cfg, const T* b, const int32_t* warm,
                    const double* r64, const double* alpha, const T* v, const double* rv, const double* beta,
                    const T* w, const double* rw, const double* gamma, const T* dinv, int32_t nV, A* x, T* r, T* z,
                    EnvCG* env, A* acc);
// CG iteration k: w = A z (+ FP64 true residual on refresh iterations), with
// `lanes` (4, 8, 16) lanes per block row and kSpmvThreads / lanes rows per
// tile; over all tiles, or persistent (`grid` >= 8 workgroups) over a list.
// col16 (may be NULL): the columns as signed 16-bit offsets from their rows,
// read instead of col.
template <class T, class A>
void launch_cg_spmv(hipStream_t s, const Tiles& ta, const WorkList& wl, int grid, const CgConfig& cfg, int k,
                    int lanes, const int32_t* row_ptr, const int32_t* col, const T* vals, int32_t nB, const T* z,
                    const A* x, const T* b, T* w, double* rref, EnvCG* env, A* acc, int32_t* num_active,
                    const KernelEvents* ev = nullptr, const int16_t* col16 = nullptr);
// CG iteration k: p, s = A p, x, r, z updates.  The iterate x is kept in the
// accumulator type A (FP64 in mixed precision; with cfg.xd, x + xd).
template <class T, class A>
void launch_cg_update(hipStream_t s, const Tiles& tv, const WorkList& wl, int grid, const CgConfig& cfg, int k,
                      const T* dinv, int32_t nV, const T* w, const double* rref, A* x, T* r, T* z, T* p, T* sv,
                      EnvCG* env, A* acc, int32_t* num_active, const KernelEvents* ev = nullptr);
// LDS bytes a workgroup may take on the current device (its reported limit,
// at most kEnvLdsMax).
int env_lds_limit();

// The whole CG of every active environment of q in one launch of `grid`
// workgroups (k_cg_env): the iterations of k_cg_spmv + k_cg_update from the
// state k_cg_init left, each environment's EnvCG record at the end.  r, p, s
// and the iterate's increments live in LDS too where q.max_rows allows it
// within lds_limit (r, p and s are then not written back, the next solve
// starting from k_cg_init; the increments are folded into x at the end).  Throws if
// q.max_rows needs more than lds_limit bytes of LDS even without them.
template <class T, class A>
void launch_cg_env(hipStream_t s, const EnvQueue& q, int grid, size_t lds_limit, const CgConfig& cfg, int lanes,
                   const int32_t* row_ptr, const int32_t* col, const int16_t* col16, const T* vals, int32_t nB,
                   const T* dinv, int32_t nV, const T* b, const T* z, A* x, T* r, T* p, T* sv, double* rref,
                   EnvCG* env, const A* acc, int32_t* num_active, const KernelEvents* ev);
// Work lists of the tiles of active environments among tiles [a0, a1) of ta
// and [v0, v1) of tv (slot `par` of counts, which must start at zero; slot
// par ^ 1 is cleared for the next burst).
// publish (may be NULL; device-visible pinned host memory): *publish =
// *num_active, the active count left by the previous kernels.
void launch_compact_tiles(hipStream_t s, const Tiles& ta, const Tiles& tv, int a0, int a1, int v0, int v1,
                          const EnvCG* env, int32_t* la, int32_t* lv, int32_t* counts, int par,
                          const int32_t* num_active, int32_t* publish, uint32_t* bar = nullptr);
// CG iterations [k0, k0 + n) (matvec and update of each) in one cooperative
// launch of `grid` workgroups over the work lists wa / wv, with grid-wide
// barriers between the phases; bar (zeroed by the compaction before the
// launch) counts barrier arrivals, and *abort is set if a barrier wait ran
// out (the solve is then invalid).
template <class T, class A>
void launch_cg_tail(hipStream_t s, const Tiles& ta, const Tiles& tv, const WorkList& wa, const WorkList& wv, int grid,
                    const CgConfig& cfg, int k0, int n, int lanes, const int32_t* row_ptr, const int32_t* col,
                    const int16_t* col16, const T* vals, int32_t nB, const T* dinv, int32_t nV, A* x, const T* b,
                    T* w, double* rref, T* r, T* z, T* p, T* sv, EnvCG* env, A* acc, int32_t* num_active,
                    uint32_t* bar, int32_t* abort);
// FP64 residual b - A x of the environments with mask[e]: into r64 when
// given; |b - A x|^2, |b|^2, b.Ax, |Ax|^2 into set 0 of acc when given
// (zeroed by the caller).  Matvec tiles of kSpmvThreads / lanes rows.
template <class T, class X>
void launch_residual64(hipStream_t s, const Tiles& ta, int lanes, int nEnv, const int32_t* mask,
                       const int32_t* row_ptr, const int32_t* col, const T* vals, int32_t nB, const X* x,
                       const T* b, double* r64, double* acc);
// The FP64 residuals ru = b - A u, rv = b - A v, rw = b - A w of the
// environments with mask[e], and into set 0 of acc (zeroed by the caller):
// |b|^2, b.Au, b.Av, b.Aw, |Au|^2, |Av|^2, |Aw|^2, Au.Av, Au.Aw, Av.Aw.
template <class T, class X>
void launch_residual64_multi(hipStream_t s, const Tiles& ta, int lanes, int nEnv, const int32_t* mask,
                             const int32_t* row_ptr, const int32_t* col, const T* vals, int32_t nB, const X* u,
                             const T* v, const T* w, const T* b, double* ru, double* rv, double* rw, double* acc);
// x[i] += xd[i], xd[i] = 0 over n values: the iterate of a mixed-precision CG
// (CgConfig.xd) made whole in x.
template <class T, class A>
void launch_cg_fold(hipStream_t s, int64_t n, A* x, T* xd);
// dst[i] = D(src[i])
template <class S, class D>
void launch_convert(hipStream_t s, int64_t n, const S* src, D* dst);
// Diagonal-block inverse of a BSR matrix (L1 entry; assembly fuses it).
template <class T>
void launch_block_inverse(hipStream_t s, int32_t nV, const int32_t* diag_slot, const T* vals, int32_t nB,
                          int prec, T* dinv);

// ---- vector helpers ----
template <class T>
void launch_copy(hipStream_t s, int64_t n, const T* src, T* dst);
template <class T>
void launch_apply_pins(hipStream_t s, int64_t n, const uint8_t* pinned, const T* value, T* u);
// x = sol0 - scale[env] * dx on rows whose env has ls_mask set.
template <class T>
void launch_ls_update(hipStream_t s, int32_t nV, const int32_t* row_env, const int32_t* ls_mask,
                      const double* scale, const T* sol0, const T* dx, T* x);
// v = (u - u0) / dt; environments with reset[env] get u = v = 0 and their
// rows of the n_hist [3 nV] vectors at hist zeroed.
template <class T>
void launch_post_step(hipStream_t s, int32_t nV, const int32_t* row_env, const int32_t* reset, double dt,
                      const T* u0, T* u, T* v, T* hist = nullptr, int n_hist = 0);
// dst = src on the rows of environments with mask[env].
template <class T>
void launch_env_copy(hipStream_t s, int32_t nV, const int32_t* row_env, const int32_t* mask, const T* src, T* dst);
// Per-environment |a|^2, |b|^2 added into set 0 of acc (zeroed by the caller).
template <class T>
void launch_env_norms(hipStream_t s, const Tiles& tv, int nEnv, const T* a, const T* b, double* acc);
// Per-environment a.b added into set 0 of acc (zeroed by the caller).
template <class T>
void launch_env_dot(hipStream_t s, const Tiles& tv, int nEnv, const T* a, const T* b, double* acc);
// Contact wrench of every collider (kWrench doubles each in out, zeroed by the
// caller: force on the soft samples, torque of the reaction on the collider
// about its position) and the least signed distance of its samples in contact
// in dmin (order-preserving int64 of the double, set to INT64_MAX by the caller),
// over the environments with env_mask set, at displacement u, stage start u0.
constexpr int kWrench = 6;
template <class T>
void launch_contact_wrench(hipStream_t s, const MeshView& m, const FieldPtrs<T>& f, const T* u, const T* u0, double dt,
                           double* out, long long* dmin);
// Objective (merit) of the environments with env_mask set at displacement u,
// stage start u0, v0: stress, gravity, inertia, mass damping and contact
// energies, per environment in FP64 into set 0 of acc (zeroed by the caller).
template <class T>
void launch_energy(hipStream_t s, const MeshView& m, const FieldPtrs<T>& f, const T* u, const T* u0, const T* v0,
                   double dt, const double grav[3], double* acc);

}  // namespace sdx
