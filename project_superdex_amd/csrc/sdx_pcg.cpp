// L1 entries: block-Jacobi PCG on a caller-supplied 3x3-block BSR matrix.
// Replace LinearSolver<T>::CudaIterativeSolve<3>
// (mochi_core/include/mochi_core/solvers/linear_solver.h:615-662): the matrix
// and right-hand side are copied to the device once per call, the
// preconditioner is the closed-form inverse of each diagonal block, and the
// CG runs with the same fused kernels as the FEM step.  The reference binds
// BlockSparseMatrixView<real const, 3> with real = float by default
// (mochi_core/include/mochi_core/mochi_config.h:106-116), so the matrix, b and
// x are taken as double (sdx_pcg_bsr3_solve) or float (sdx_pcg_bsr3_solve_f32)
// arrays; the arithmetic of the solve is params->precision_mode either way.
#include "sdx_pcg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sdx_world.h"

namespace sdx {

sdx_status pcg_check_args(int32_t device, int32_t n_block_rows, int32_t nnzb, const int32_t* ptr, const int32_t* col,
                          const void* values, int32_t layout, const void* b, const void* x,
                          const sdx_krylov_params* params) {
  if (!ptr || !col || !values || !b || !x || !params) return set_error(SDX_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n_block_rows <= 0 || nnzb <= 0) return set_error(SDX_ERR_INVALID_ARGUMENT, "empty matrix");
  if (layout != SDX_BSR_BLOCK_ROWMAJOR && layout != SDX_BSR_REF_ROW_STRIP)
    return set_error(SDX_ERR_INVALID_ARGUMENT, "unknown BSR layout");
  if (ptr[0] != 0 || ptr[n_block_rows] != nnzb)
    return set_error(SDX_ERR_INVALID_ARGUMENT, "row pointer does not span the blocks");
  for (int i = 0; i < n_block_rows; ++i)
    if (ptr[i + 1] < ptr[i]) return set_error(SDX_ERR_INVALID_ARGUMENT, "row pointer decreases");
  for (int k = 0; k < nnzb; ++k)
    if (col[k] < 0 || col[k] >= n_block_rows)
      return set_error(SDX_ERR_INVALID_ARGUMENT, "block column index out of range");
  if (params->max_iter <= 0) return set_error(SDX_ERR_INVALID_ARGUMENT, "max_iter must be positive");
  if (params->preconditioner < SDX_PREC_NONE || params->preconditioner > SDX_PREC_BLOCK_JACOBI)
    return set_error(SDX_ERR_INVALID_ARGUMENT, "unsupported preconditioner");
  if (params->norm_type < 0 || params->norm_type > 2) return set_error(SDX_ERR_INVALID_ARGUMENT, "unsupported norm");
  if (params->precision_mode < SDX_PREC_MODE_FP32 || params->precision_mode > SDX_PREC_MODE_FP64)
    return set_error(SDX_ERR_INVALID_ARGUMENT, "unknown precision mode");
  if (params->preconditioner != SDX_PREC_NONE)
    for (int i = 0; i < n_block_rows; ++i) {
      bool has = false;
      for (int k = ptr[i]; k < ptr[i + 1] && !has; ++k) has = col[k] == i;
      if (!has) return set_error(SDX_ERR_INVALID_ARGUMENT, "block row without a diagonal block");
    }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return set_error(SDX_ERR_DEVICE, "no such HIP device");
  return SDX_OK;
}

template <class T, class A>
void pcg_run(hipStream_t s, const PcgSystem<T, A>& d, const sdx_krylov_params& kp, sdx_linear_status* out) {
  CgConfig cfg{};
  cfg.n_env = 1;
  cfg.n_rows = d.nbr;
  cfg.norm_type = kp.norm_type;
  cfg.prec = kp.preconditioner;
  cfg.max_iter = kp.max_iter;
  cfg.refresh_every = kp.precision_mode == SDX_PREC_MODE_MIXED ? kp.refresh_every : 0;
  cfg.abort_if_not_spd = kp.abort_if_not_spd;
  cfg.abs_tol = kp.abs_tol;
  cfg.rel_tol = kp.rel_tol;
  cfg.rel_div_tol = kp.rel_div_tol;
  cfg.xd = nullptr;  // every update adds into the iterate in the accumulator type
  launch_fill<int32_t>(s, 1, 1, d.num_active);
  launch_fill<int32_t>(s, 1, 1, d.solve);
  launch_block_inverse<T>(s, d.nbr, d.diag, d.vals, d.nnzb, kp.preconditioner, d.dinv);
  const bool warm = !kp.initial_guess_zero;
  if (warm)  // r0 = b - A x0 in FP64
    launch_residual64<T, A>(s, d.ta, kSpmvLanes, 1, d.solve, d.ptr, d.col, d.vals, d.nnzb, d.x, d.b, d.rref, nullptr);
  launch_cg_reset<A>(s, 1, d.solve, d.env, d.acc);
  launch_cg_init<T, A>(s, d.tv, cfg, d.b, warm ? d.solve : nullptr, d.rref, /*alpha*/ nullptr, /*v*/ nullptr,
                       /*rv*/ nullptr, /*beta*/ nullptr, /*w*/ nullptr, /*rw*/ nullptr, /*gamma*/ nullptr, d.dinv,
                       d.nbr, d.x, d.r, d.z, d.env, d.acc);
  const WorkList none{nullptr, nullptr};
  int j = 0;
  int32_t na = 1;
  while (true) {
    for (int k = 0; k < 8; ++k, ++j) {
      launch_cg_spmv<T, A>(s, d.ta, none, 0, cfg, j, kSpmvLanes, d.ptr, d.col, d.vals, d.nnzb, d.z, d.x, d.b, d.w,
                           d.rref, d.env, d.acc, d.num_active);
      launch_cg_update<T, A>(s, d.tv, none, 0, cfg, j, d.dinv, d.nbr, d.w, d.rref, d.x, d.r, d.z, d.p, d.sv, d.env,
                             d.acc, d.num_active);
    }
    hip_check(hipMemcpyAsync(&na, d.num_active, sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (na <= 0 || j > kp.max_iter + 1) break;
  }
  EnvCG e;
  hip_check(hipMemcpyAsync(&e, d.env, sizeof(EnvCG), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (out) {
    out->num_iter_done = e.iter;
    out->convergence = e.status == kCgConverged ? SDX_CONV_CONVERGED
                       : e.status == kCgDiverged ? SDX_CONV_DIVERGED
                                                 : SDX_CONV_STOPPED;
    out->residual_norm = std::sqrt(std::max(e.latest, 0.0));
    out->relative_residual_norm = e.start > 0 ? std::sqrt(std::max(e.latest, 0.0) / e.start) : 0.0;
  }
}

template void pcg_run<float, float>(hipStream_t, const PcgSystem<float, float>&, const sdx_krylov_params&,
                                    sdx_linear_status*);
template void pcg_run<float, double>(hipStream_t, const PcgSystem<float, double>&, const sdx_krylov_params&,
                                     sdx_linear_status*);
template void pcg_run<double, double>(hipStream_t, const PcgSystem<double, double>&, const sdx_krylov_params&,
                                      sdx_linear_status*);

namespace {

template <class U>
struct Tmp {
  U* p = nullptr;
  explicit Tmp(size_t n) {
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(U));
    if (e == hipErrorOutOfMemory) throw std::bad_alloc();
    hip_check(e, "hipMalloc");
  }
  ~Tmp() {
    if (p) (void)hipFree(p);
  }
  Tmp(const Tmp&) = delete;
  Tmp& operator=(const Tmp&) = delete;
};

// Asynchronous upload; h must stay alive until the stream has run the copy
// (every caller synchronises before its host vectors go out of scope).
template <class U>
void put(Tmp<U>& d, const U* h, size_t n, hipStream_t s) {
  if (n) hip_check(hipMemcpyAsync(d.p, h, n * sizeof(U), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
}
template <class U>
void put(Tmp<U>& d, const std::vector<U>& h, hipStream_t s) {
  put(d, h.data(), h.size(), s);
}

void make_tiles(int n, int rows, std::vector<int32_t>& env, std::vector<int32_t>& r0, std::vector<int32_t>& r1,
                std::vector<int32_t>& first) {
  for (int r = 0; r < n; r += rows) {
    env.push_back(0);
    r0.push_back(r);
    r1.push_back(std::min(r + rows, n));
  }
  first = {0, int32_t(env.size())};
}

// V: the caller's value type (double or float); T, A: the solve's arithmetic.
template <class T, class A, class V>
sdx_status solve(int32_t device, int32_t nbr, int32_t nnzb, const int32_t* ptr, const int32_t* col, const V* values,
                 int32_t layout, const V* b, V* x, const sdx_krylov_params& kp, sdx_linear_status* out) {
  hip_check(hipSetDevice(device), "hipSetDevice");
  hipStream_t s;
  hip_check(hipStreamCreate(&s), "hipStreamCreate");
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);  // no copy from a host vector below is pending past this scope
      (void)hipStreamDestroy(s);
    }
  } guard{s};
  // SoA value planes
  std::vector<T> vals(size_t(9) * size_t(nnzb));
  std::vector<int32_t> diag(size_t(nbr), -1);
  for (int i = 0; i < nbr; ++i) {
    const int k0 = ptr[i], nr = ptr[i + 1] - ptr[i];
    for (int kk = 0; kk < nr; ++kk) {
      const int k = k0 + kk;
      if (col[k] == i) diag[size_t(i)] = k;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          const V v = layout == SDX_BSR_BLOCK_ROWMAJOR ? values[size_t(9) * size_t(k) + size_t(3 * r + c)]
                                                       : values[size_t(9) * size_t(k0) + size_t(r) * size_t(3 * nr) +
                                                                size_t(3 * kk + c)];
          vals[size_t(3 * r + c) * size_t(nnzb) + size_t(k)] = T(v);
        }
    }
  }
  for (int i = 0; i < nbr; ++i)  // read but unused without a preconditioner: any slot in range
    if (diag[size_t(i)] < 0) diag[size_t(i)] = 0;
  std::vector<int32_t> te, t0, t1, tf, ae, a0, a1, af;
  make_tiles(nbr, kVecThreads, te, t0, t1, tf);
  make_tiles(nbr, kSpmvRows, ae, a0, a1, af);
  const size_t n3 = size_t(3) * size_t(nbr);
  std::vector<T> hb(n3);
  std::vector<A> hx(n3);  // the iterate is kept in the accumulator type
  for (size_t k = 0; k < n3; ++k) {
    hb[k] = T(b[k]);
    hx[k] = A(x[k]);
  }
  Tmp<int32_t> dptr(size_t(nbr) + 1), dcol(static_cast<size_t>(nnzb)), ddiag(static_cast<size_t>(nbr));
  Tmp<T> dvals(vals.size()), ddinv(size_t(6) * size_t(nbr)), db(n3), dr(n3), dz(size_t(4) * size_t(nbr)), dp(n3),
      dsv(n3), dw(n3);
  Tmp<A> dx(n3);
  Tmp<double> drref(n3);
  Tmp<int32_t> dte(te.size()), dt0(t0.size()), dt1(t1.size()), dtf(2), dae(ae.size()), da0(a0.size()),
      da1(a1.size()), daf(2);
  Tmp<EnvCG> denv(1);
  Tmp<A> dacc(size_t(kCgSets) * kSlots * kSlotStride);
  Tmp<int32_t> dna(1), dsolve(1);
  put(dptr, ptr, size_t(nbr) + 1, s);
  put(dcol, col, size_t(nnzb), s);
  put(ddiag, diag, s);
  put(dvals, vals, s);
  put(db, hb, s);
  put(dx, hx, s);
  put(dte, te, s);
  put(dt0, t0, s);
  put(dt1, t1, s);
  put(dtf, tf, s);
  put(dae, ae, s);
  put(da0, a0, s);
  put(da1, a1, s);
  put(daf, af, s);
  PcgSystem<T, A> d;
  d.nbr = nbr;
  d.nnzb = nnzb;
  d.ptr = dptr.p;
  d.col = dcol.p;
  d.diag = ddiag.p;
  d.vals = dvals.p;
  d.b = db.p;
  d.x = dx.p;
  d.dinv = ddinv.p;
  d.r = dr.p;
  d.z = dz.p;
  d.p = dp.p;
  d.sv = dsv.p;
  d.w = dw.p;
  d.rref = drref.p;
  d.tv = Tiles{dte.p, dt0.p, dt1.p, dtf.p, int32_t(te.size())};
  d.ta = Tiles{dae.p, da0.p, da1.p, daf.p, int32_t(ae.size())};
  d.env = denv.p;
  d.acc = dacc.p;
  d.num_active = dna.p;
  d.solve = dsolve.p;
  pcg_run<T, A>(s, d, kp, out);
  hip_check(hipMemcpyAsync(hx.data(), dx.p, n3 * sizeof(A), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  for (size_t k = 0; k < n3; ++k) x[k] = V(hx[k]);
  return SDX_OK;
}

template <class V>
sdx_status solve_any(int32_t device, int32_t n_block_rows, int32_t nnzb, const int32_t* ptr, const int32_t* col,
                     const V* values, int32_t layout, const V* b, V* x, const sdx_krylov_params* params,
                     sdx_linear_status* out) {
  const sdx_status st = pcg_check_args(device, n_block_rows, nnzb, ptr, col, values, layout, b, x, params);
  if (st != SDX_OK) return st;
  return guarded([&]() -> sdx_status {
    switch (params->precision_mode) {
      case SDX_PREC_MODE_FP32:
        return solve<float, float, V>(device, n_block_rows, nnzb, ptr, col, values, layout, b, x, *params, out);
      case SDX_PREC_MODE_MIXED:
        return solve<float, double, V>(device, n_block_rows, nnzb, ptr, col, values, layout, b, x, *params, out);
      default:
        return solve<double, double, V>(device, n_block_rows, nnzb, ptr, col, values, layout, b, x, *params, out);
    }
  });
}

}  // namespace
}  // namespace sdx

extern "C" sdx_status sdx_pcg_bsr3_solve(int32_t device, int32_t n_block_rows, int32_t nnzb,
                                         const int32_t* block_row_ptr, const int32_t* block_col_idx,
                                         const double* values, int32_t layout, const double* b, double* x,
                                         const sdx_krylov_params* params, sdx_linear_status* out) {
  return sdx::solve_any<double>(device, n_block_rows, nnzb, block_row_ptr, block_col_idx, values, layout, b, x,
                                params, out);
}

extern "C" sdx_status sdx_pcg_bsr3_solve_f32(int32_t device, int32_t n_block_rows, int32_t nnzb,
                                             const int32_t* block_row_ptr, const int32_t* block_col_idx,
                                             const float* values, int32_t layout, const float* b, float* x,
                                             const sdx_krylov_params* params, sdx_linear_status* out) {
  return sdx::solve_any<float>(device, n_block_rows, nnzb, block_row_ptr, block_col_idx, values, layout, b, x,
                               params, out);
}
