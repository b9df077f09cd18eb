// Device side of the L1 linear solve shared by sdx_pcg_bsr3_solve(_f32) and
// its allocation-free variant: the block-Jacobi PCG of one system whose
// matrix, right-hand side and work vectors are already in device memory.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/sdx_fem.h"
#include "sdx_kernels.h"

namespace sdx {

// One system of nbr block rows and nnzb 3x3 blocks in device memory, values
// as 9 SoA planes (vals[c * nnzb + k]) of T, the iterate x in the
// accumulator type A.  The tiles cover the rows of environment 0: tv in
// kVecThreads-row vector tiles, ta in kSpmvRows-row matvec tiles.
template <class T, class A>
struct PcgSystem {
  int32_t nbr = 0, nnzb = 0;
  const int32_t* ptr = nullptr;   // [nbr + 1]
  const int32_t* col = nullptr;   // [nnzb]
  const int32_t* diag = nullptr;  // [nbr] slot of each row's diagonal block
  const T* vals = nullptr;        // [9 nnzb]
  const T* b = nullptr;           // [3 nbr]
  A* x = nullptr;                 // [3 nbr] initial guess in, solution out
  T* dinv = nullptr;              // [6 nbr]
  T *r = nullptr, *z = nullptr, *p = nullptr, *sv = nullptr, *w = nullptr;  // z: [4 nbr], others [3 nbr]
  double* rref = nullptr;         // [3 nbr]
  Tiles tv{}, ta{};
  EnvCG* env = nullptr;           // [1]
  A* acc = nullptr;               // [kCgSets * kSlots * kSlotStride]
  int32_t* num_active = nullptr;  // [1]
  int32_t* solve = nullptr;       // [1]
};

// Runs the PCG of `kp` on the stream (x is the initial guess unless
// kp.initial_guess_zero) and fills `out`; returns after the solution is in x.
// Every device value the solve reads before writing it is set here, so the
// work buffers may hold anything when it starts.
template <class T, class A>
void pcg_run(hipStream_t s, const PcgSystem<T, A>& d, const sdx_krylov_params& kp, sdx_linear_status* out);

// Argument checks shared by the L1 entry points (host-side arrays).
sdx_status pcg_check_args(int32_t device, int32_t n_block_rows, int32_t nnzb, const int32_t* ptr, const int32_t* col,
                          const void* values, int32_t layout, const void* b, const void* x,
                          const sdx_krylov_params* params);

}  // namespace sdx
