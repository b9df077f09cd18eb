// L1 entry: block-Jacobi PCG on a caller-supplied 3x3-block BSR matrix.
// Replaces LinearSolver<T>::CudaIterativeSolve<3>
// (mochi_core/include/mochi_core/solvers/linear_solver.h:615-662): the matrix
// and right-hand side are copied to the device once, the preconditioner is
// the closed-form inverse of each diagonal block, and the CG runs with the
// same fused kernels as the FEM step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/sdx_fem.h"
#include "sdx_kernels.h"
#include "sdx_world.h"

namespace sdx {
namespace {

template <class T>
struct Tmp {
  T* p = nullptr;
  explicit Tmp(size_t n) { hip_check(hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc"); }
  ~Tmp() {
    if (p) (void)hipFree(p);
  }
  Tmp(const Tmp&) = delete;
  Tmp& operator=(const Tmp&) = delete;
};

template <class T>
void put(Tmp<T>& d, const std::vector<T>& h, hipStream_t s) {
  if (!h.empty()) hip_check(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
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

template <class T, class A>
sdx_status solve(int32_t device, int32_t nbr, int32_t nnzb, const int32_t* ptr, const int32_t* col, const double* values,
                 int32_t layout, const double* b, double* x, const sdx_krylov_params& kp, sdx_linear_status* out) {
  hip_check(hipSetDevice(device), "hipSetDevice");
  hipStream_t s;
  hip_check(hipStreamCreate(&s), "hipStreamCreate");
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
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
          const double v = layout == SDX_BSR_BLOCK_ROWMAJOR ? values[size_t(9) * size_t(k) + size_t(3 * r + c)]
                                                            : values[size_t(9) * size_t(k0) + size_t(r) * size_t(3 * nr) +
                                                                     size_t(3 * kk + c)];
          vals[size_t(3 * r + c) * size_t(nnzb) + size_t(k)] = T(v);
        }
    }
    if (kp.preconditioner != SDX_PREC_NONE && diag[size_t(i)] < 0)
      return set_error(SDX_ERR_INVALID_ARGUMENT, "block row without a diagonal block");
  }
  for (int i = 0; i < nbr; ++i)
    if (diag[size_t(i)] < 0) diag[size_t(i)] = ptr[i];  // unused without a preconditioner
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
  Tmp<double> dacc(size_t(kCgSets) * kSlots * kSlotStride);
  Tmp<int32_t> dna(1), dsolve(1);
  put(dptr, std::vector<int32_t>(ptr, ptr + nbr + 1), s);
  put(dcol, std::vector<int32_t>(col, col + nnzb), s);
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
  put(dna, std::vector<int32_t>(1, 1), s);
  put(dsolve, std::vector<int32_t>(1, 1), s);
  Tiles tv{dte.p, dt0.p, dt1.p, dtf.p, int32_t(te.size())};
  Tiles ta{dae.p, da0.p, da1.p, daf.p, int32_t(ae.size())};
  CgConfig cfg{};
  cfg.n_env = 1;
  cfg.n_rows = nbr;
  cfg.norm_type = kp.norm_type;
  cfg.prec = kp.preconditioner;
  cfg.max_iter = kp.max_iter;
  cfg.refresh_every = kp.precision_mode == SDX_PREC_MODE_MIXED ? kp.refresh_every : 0;
  cfg.abort_if_not_spd = kp.abort_if_not_spd;
  cfg.abs_tol = kp.abs_tol;
  cfg.rel_tol = kp.rel_tol;
  cfg.rel_div_tol = kp.rel_div_tol;
  launch_block_inverse<T>(s, nbr, ddiag.p, dvals.p, nnzb, kp.preconditioner, ddinv.p);
  const bool warm = !kp.initial_guess_zero;
  if (warm)  // r0 = b - A x0 in FP64
    launch_residual64<T, A>(s, ta, kSpmvLanes, 1, dsolve.p, dptr.p, dcol.p, dvals.p, nnzb, dx.p, db.p, drref.p, nullptr);
  A* acc = reinterpret_cast<A*>(dacc.p);
  launch_cg_reset<A>(s, 1, dsolve.p, denv.p, acc);
  launch_cg_init<T, A>(s, tv, cfg, db.p, warm ? dsolve.p : nullptr, drref.p, /*alpha*/ nullptr, /*v*/ nullptr,
                       /*rv*/ nullptr, /*beta*/ nullptr, /*w*/ nullptr, /*rw*/ nullptr, /*gamma*/ nullptr, ddinv.p, nbr,
                       dx.p, dr.p, dz.p, denv.p, acc);
  int j = 0;
  int32_t na = 1;
  while (true) {
    for (int k = 0; k < 8; ++k, ++j) {
      launch_cg_spmv<T, A>(s, ta, WorkList{nullptr, nullptr}, 0, cfg, j, kSpmvLanes, dptr.p, dcol.p, dvals.p, nnzb, dz.p, dx.p, db.p, dw.p, drref.p, denv.p,
                           acc, dna.p);
      launch_cg_update<T, A>(s, tv, WorkList{nullptr, nullptr}, 0, cfg, j, ddinv.p, nbr, dw.p, drref.p, dx.p, dr.p, dz.p, dp.p, dsv.p, denv.p, acc,
                             dna.p);
    }
    hip_check(hipMemcpyAsync(&na, dna.p, sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (na <= 0 || j > kp.max_iter + 1) break;
  }
  EnvCG e;
  hip_check(hipMemcpyAsync(&e, denv.p, sizeof(EnvCG), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipMemcpyAsync(hx.data(), dx.p, n3 * sizeof(A), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  for (size_t k = 0; k < n3; ++k) x[k] = double(hx[k]);
  if (out) {
    out->num_iter_done = e.iter;
    out->convergence = e.status == kCgConverged ? SDX_CONV_CONVERGED
                       : e.status == kCgDiverged ? SDX_CONV_DIVERGED
                                                 : SDX_CONV_STOPPED;
    out->residual_norm = std::sqrt(std::max(e.latest, 0.0));
    out->relative_residual_norm = e.start > 0 ? std::sqrt(std::max(e.latest, 0.0) / e.start) : 0.0;
  }
  return SDX_OK;
}

}  // namespace
}  // namespace sdx

extern "C" sdx_status sdx_pcg_bsr3_solve(int32_t device, int32_t n_block_rows, int32_t nnzb,
                                         const int32_t* block_row_ptr, const int32_t* block_col_idx,
                                         const double* values, int32_t layout, const double* b, double* x,
                                         const sdx_krylov_params* params, sdx_linear_status* out) {
  using sdx::set_error;
  if (!block_row_ptr || !block_col_idx || !values || !b || !x || !params)
    return set_error(SDX_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n_block_rows <= 0 || nnzb <= 0) return set_error(SDX_ERR_INVALID_ARGUMENT, "empty matrix");
  if (layout != SDX_BSR_BLOCK_ROWMAJOR && layout != SDX_BSR_REF_ROW_STRIP)
    return set_error(SDX_ERR_INVALID_ARGUMENT, "unknown BSR layout");
  if (block_row_ptr[0] != 0 || block_row_ptr[n_block_rows] != nnzb)
    return set_error(SDX_ERR_INVALID_ARGUMENT, "row pointer does not span the blocks");
  for (int i = 0; i < n_block_rows; ++i)
    if (block_row_ptr[i + 1] < block_row_ptr[i]) return set_error(SDX_ERR_INVALID_ARGUMENT, "row pointer decreases");
  for (int k = 0; k < nnzb; ++k)
    if (block_col_idx[k] < 0 || block_col_idx[k] >= n_block_rows)
      return set_error(SDX_ERR_INVALID_ARGUMENT, "block column index out of range");
  if (params->max_iter <= 0) return set_error(SDX_ERR_INVALID_ARGUMENT, "max_iter must be positive");
  if (params->preconditioner < SDX_PREC_NONE || params->preconditioner > SDX_PREC_BLOCK_JACOBI)
    return set_error(SDX_ERR_INVALID_ARGUMENT, "unsupported preconditioner");
  if (params->norm_type < 0 || params->norm_type > 2) return set_error(SDX_ERR_INVALID_ARGUMENT, "unsupported norm");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return set_error(SDX_ERR_DEVICE, "no such HIP device");
  try {
    switch (params->precision_mode) {
      case SDX_PREC_MODE_FP32:
        return sdx::solve<float, float>(device, n_block_rows, nnzb, block_row_ptr, block_col_idx, values, layout, b, x,
                                        *params, out);
      case SDX_PREC_MODE_MIXED:
        return sdx::solve<float, double>(device, n_block_rows, nnzb, block_row_ptr, block_col_idx, values, layout, b,
                                         x, *params, out);
      case SDX_PREC_MODE_FP64:
        return sdx::solve<double, double>(device, n_block_rows, nnzb, block_row_ptr, block_col_idx, values, layout, b,
                                          x, *params, out);
      default:
        return set_error(SDX_ERR_INVALID_ARGUMENT, "unknown precision mode");
    }
  } catch (const sdx::DeviceError& e) {
    return set_error(SDX_ERR_DEVICE, e.what());
  } catch (const std::bad_alloc&) {
    return set_error(SDX_ERR_OUT_OF_MEMORY, "out of memory");
  }
}
