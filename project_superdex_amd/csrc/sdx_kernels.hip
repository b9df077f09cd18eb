// HIP kernels of the implicit soft-body FEM step for gfx950 (CDNA4).
//
// Data layout in HBM (V vertices, T tets, F boundary triangles, B BSR blocks):
//   DoF vectors ............ [3V] node-major xyz, FP32 (FP64 in fp64 mode);
//                            z padded to [4V] for one-load neighbour gathers
//   tet geometry ........... 14 SoA planes of T reals (coalesced per tet)
//   records ................ element and contact contributions in destination
//                            order: residual [4T + 3F][3] and diagonal blocks
//                            [4T + 3F][6] vertex-sorted, off-diagonal blocks
//                            [6T + 3F][9] sorted by (min, max) vertex pair
//   BSR values ............. 9 SoA planes of B reals (vals[c*B + slot])
//   block-Jacobi inverse ... 6 SoA planes of V reals (symmetric block)
// Assembly is atomic-free: the element and contact kernels write every
// contribution to its record, and every BSR slot / vertex sums one contiguous
// run of records in a fixed order, so results are deterministic.
//
// CG (pcg.h:70-238, Polak-Ribiere) takes two kernels per iteration:
//   cg_spmv:   w = A z (the only matvec), partial z.w
//   cg_update: p = z + beta p, s = w + beta s (= A p), x += alpha p,
//              r -= alpha s, z = D^-1 r, partial r.r, r.z, z.z, r.z_old
// Per-environment scalars are summed by spread atomic accumulators; every
// block of the next kernel derives alpha, beta and the stop test from them.
// Tiles never straddle environments, so hundreds of uneven environments are
// solved by the same launches.
#include <hip/hip_ext.h>

#include <stdexcept>
#include <string>

#include "sdx_device.h"
#include "sdx_kernels.h"
#include "sdx_math.h"
#include "sdx_rowblocks.h"

namespace sdx {

// ------------------------------------------------------------------ element

// Element patches, one wave each: lane i computes tet i of the patch (F, Smith
// P, 12-vector residual with stress, gravity, inertia and damping, 12 x 12
// Hessian with the consistent mass) and stages its contributions in the
// wave's LDS; the wave then sums them per local vertex and local vertex pair
// in the fixed order of the patch's lists and writes one record per local
// vertex (residual, diagonal block) and per local pair (off-diagonal block in
// (min, max) orientation).  Two phases share one stage region of kStageE
// values per lane: first the residual and the diagonal blocks (kStageV per
// tet), then the off-diagonal blocks, which wait in registers meanwhile.  The
// patch description is copied into LDS with 16-byte loads issued before the
// element's own loads.  Every lane of a workgroup passes every barrier (the
// control flow around them depends only on wantH, uniform over the launch).
constexpr int kStageV = 37;  // residual 12 + diagonal blocks 24, odd lane stride
constexpr int kStageE = 55;  // off-diagonal blocks 54, odd lane stride
constexpr int kMetaVecs = (kPatchMetaWords / 4 + 63) / 64;  // 16-byte loads per lane
static_assert(kPatchTets == 64, "one tet per lane of a wave");

extern __shared__ __align__(16) unsigned char sdx_dyn_lds[];

template <class T>
constexpr size_t patch_lds_bytes() {
  return size_t(kStageE) * 64 * sizeof(T) + size_t(kPatchMetaWords) * 4;
}

// The inputs of one tet: shape gradients, volume, material, mass and
// damping, its displacement u, u - predicted target (inertia) and u - u0
// (damping).  Loads issued in this order: tet geometry planes, then the
// vertex data.
template <class T>
struct TetIn {
  V3<T> dN[4];
  T vol, mu, lam, mass, alpha;
  T ue[12], w[12], wd[12];
};

template <class T>
__device__ __forceinline__ void load_tet(const FieldPtrs<T>& f, int nT, int t, const int nd[4],
                                         const T* __restrict__ u, const T* __restrict__ u0,
                                         const T* __restrict__ v0, T dt, TetIn<T>& in) {
  const T* g = f.geom;
  in.dN[1] = mk3(g[0 * nT + t], g[1 * nT + t], g[2 * nT + t]);
  in.dN[2] = mk3(g[3 * nT + t], g[4 * nT + t], g[5 * nT + t]);
  in.dN[3] = mk3(g[6 * nT + t], g[7 * nT + t], g[8 * nT + t]);
  in.dN[0] = scl3(T(-1), add3(add3(in.dN[1], in.dN[2]), in.dN[3]));
  in.vol = g[kGVol * nT + t];
  in.mu = g[kGMu * nT + t];
  in.lam = g[kGLambda * nT + t];
  in.mass = g[kGMass * nT + t];
  in.alpha = g[kGAlpha * nT + t];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T uu = u[3 * nd[a] + d], s0 = u0[3 * nd[a] + d];
      in.ue[3 * a + d] = uu;
      in.w[3 * a + d] = uu - (s0 + dt * v0[3 * nd[a] + d]);  // u - predicted target (batched_element_utils.h:90-120)
      in.wd[3 * a + d] = uu - s0;
    }
}

// One tet's contributions (fem_stress.h:107-262, fem_inertia.h:55-250,
// fem_gravity.h:44-125, fem_stress_damping.h): on_res(res[12]) receives the
// residual of its 4 vertices (stress, inertia of the consistent mass,
// gravity, damping), formed before the Hessian; emit(fa, fb, q) each Hessian
// block with the consistent-mass term, q = 6 symmetric entries (00 11 22 01
// 02 12) for fa == fb, else 9 row-major entries of block (fa, fb), fa < fb.
template <class T, bool kDamp, class OnRes, class Emit>
__device__ __forceinline__ void tet_contrib(const FieldPtrs<T>& f, int nT, int t, const TetIn<T>& in, T dt, T gx,
                                            T gy, T gz, bool wantH, bool project, OnRes&& on_res, Emit&& emit) {
  const T* g = f.geom;
  // stiffness damping (fem_stress_damping.h), kappa = beta / dt (mochi_soft.h:259-292)
  Viscous<T> vis;
  bool damp = false;
  if constexpr (kDamp) {
    const T beta = g[kGBeta * nT + t];
    damp = beta > T(0);
    if (damp) {
      T us[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) us[k] = in.ue[k] - in.wd[k];
      viscous_setup<T>(in.dN, in.ue, us, beta / dt, in.mu, in.lam, g[kGBetaGeo * nT + t] != T(0), project, vis);
    }
  }
  // consistent mass M_ab = m/20 (1 + delta_ab) (Q4 integrates P1 x P1 exactly)
  const T idt2 = T(1) / (dt * dt);
  const T mdamp = in.alpha > T(0) ? in.alpha / dt : T(0);
  const T m20 = in.mass * T(0.05);
  const T ms = (idt2 + mdamp) * m20;
  auto block = [&](int fa, int fb, const T* q) {
    if (fa == fb) {
      T dq[6];
#pragma unroll
      for (int r = 0; r < 6; ++r) dq[r] = q[r];
      if (kDamp && damp) viscous_block<T>(vis, in.dN[fa], in.dN[fb], true, in.vol, dq);
      dq[0] += T(2) * ms;
      dq[1] += T(2) * ms;
      dq[2] += T(2) * ms;
      emit(fa, fb, dq);
      return;
    }
    T o[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) o[r] = q[r];
    if (kDamp && damp) viscous_block<T>(vis, in.dN[fa], in.dN[fb], false, in.vol, o);
    o[0] += ms;
    o[4] += ms;
    o[8] += ms;
    emit(fa, fb, o);
  };
  // the residual of each vertex: stress, inertia of the consistent mass and
  // gravity (fem_gravity.h:108-118: res_f -= rho g vol / 4)
  auto res_of = [&](const T* res_in) {
    T res[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) res[k] = res_in[k];
    if (kDamp && damp) viscous_res<T>(vis, in.dN, in.vol, res);
    const T gq = T(0.25) * in.mass;
    const T* w = in.w;
    const T* wd = in.wd;
    const T sw[3] = {w[0] + w[3] + w[6] + w[9], w[1] + w[4] + w[7] + w[10], w[2] + w[5] + w[8] + w[11]};
    const T sd[3] = {wd[0] + wd[3] + wd[6] + wd[9], wd[1] + wd[4] + wd[7] + wd[10], wd[2] + wd[5] + wd[8] + wd[11]};
    const T gv[3] = {gx, gy, gz};
    T out[12];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        T r = res[3 * a + d] - gv[d] * gq + idt2 * m20 * (w[3 * a + d] + sw[d]);
        if (mdamp > T(0)) r += mdamp * m20 * (wd[3 * a + d] + sd[d]);
        out[3 * a + d] = r;
      }
    on_res(out);
  };
  smith_element<T>(in.dN, in.ue, in.vol, in.mu, in.lam, wantH, project, res_of, block);
}

template <class T, bool kDamp>
__global__ void __launch_bounds__(256) k_element(MeshView m, AsmRange ar, FieldPtrs<T> f, const T* __restrict__ u,
                                                 const T* __restrict__ u0, const T* __restrict__ v0, T dt,
                                                 T gx, T gy, T gz, int wantH, int project, T* __restrict__ Rv,
                                                 T* __restrict__ Dv, T* __restrict__ Ev,
                                                 const int32_t* __restrict__ plist) {
  const int wpb = int(blockDim.x) / 64, wave = int(threadIdx.x) / 64, lane = int(threadIdx.x) % 64;
  // patches [p0, p1), or entries [p0, p1) of the patch list plist
  const int q = ar.p0 + xcd_tile(blockIdx.x, gridDim.x) * wpb + wave;
  unsigned char* lds_w = sdx_dyn_lds + size_t(wave) * patch_lds_bytes<T>();
  T* stage = reinterpret_cast<T*>(lds_w);
  uint32_t* meta = reinterpret_cast<uint32_t*>(lds_w + size_t(kStageE) * 64 * sizeof(T));
  // a wave past the range loads the first patch's data and computes nothing
  const bool havep = q < ar.p1;
  const int qc = havep ? q : ar.p0;
  const int pc = plist ? plist[qc] : qc;
  // the whole description, or its vertex part for a residual
  const int mw0 = m.patch_meta_ptr[pc], nv4 = ((wantH ? m.patch_meta_ptr[pc + 1] : m.patch_meta_mid[pc]) - mw0) / 4;
  uint4 mv[kMetaVecs];
  const uint4* msrc = reinterpret_cast<const uint4*>(m.patch_meta + mw0);
#pragma unroll
  for (int k = 0; k < kMetaVecs; ++k) {
    const int i = lane + 64 * k;
    mv[k] = msrc[i < nv4 ? i : 0];
  }
  const int t0 = m.patch_t0[pc], t1 = m.patch_t0[pc + 1];
  const int env = m.tet_env[t0];
  const int nT = m.nT;
  const int t = t0 + lane < t1 ? t0 + lane : t0;
  const int4 tv = reinterpret_cast<const int4*>(m.tets)[t];
  const int nd[4] = {tv.x, tv.y, tv.z, tv.w};
  TetIn<T> in;
  load_tet<T>(f, nT, t, nd, u, u0, v0, dt, in);
  uint4* mdst = reinterpret_cast<uint4*>(meta);
#pragma unroll
  for (int k = 0; k < kMetaVecs; ++k) {
    const int i = lane + 64 * k;
    if (i < nv4) mdst[i] = mv[k];
  }
  const bool on = havep && m.env_mask[env] != 0;  // uniform over the wave
  const bool mine = on && t0 + lane < t1;
  T off[54];  // the off-diagonal blocks, staged in the second phase
  if (mine) {
    T* sv = stage + lane * kStageV;
    // diagonal blocks to the stage, off-diagonal blocks (pair index k of
    // (kPairA, kPairB)) to registers
    auto emit = [&](int fa, int fb, const T* q) {
      if (fa == fb) {
        T* d = sv + 12 + 6 * fa;
#pragma unroll
        for (int r = 0; r < 6; ++r) d[r] = q[r];
        return;
      }
      const int pk = fa == 0 ? fb - 1 : (fa == 1 ? fb + 1 : 5);
#pragma unroll
      for (int r = 0; r < 9; ++r) off[9 * pk + r] = q[r];
    };
    // the residual, staged before the Hessian is formed
    auto on_res = [&](const T* res) {
#pragma unroll
      for (int k = 0; k < 12; ++k) sv[k] = res[k];
    };
    tet_contrib<T, kDamp>(f, nT, t, in, dt, gx, gy, gz, wantH != 0, project != 0, on_res, emit);
  }
  __syncthreads();
  const uint32_t hd = meta[0];
  const int nLV = int(hd & 1023u), nLE = int((hd >> 10) & 1023u);
  const uint32_t* rec = meta + 1;
  const uint16_t* ls = reinterpret_cast<const uint16_t*>(meta + 1 + nLV);
  const uint16_t* ent = ls + nLV + 1;
  // phase 1: residual and diagonal block of every local vertex
  if (on)
    for (int j = lane; j < nLV; j += 64) {
      T r0 = T(0), r1 = T(0), r2 = T(0), d[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
      const int q1 = ls[j + 1];
      for (int q = ls[j]; q < q1; ++q) {
        const int c = ent[q];
        const T* sp = stage + (c >> 2) * kStageV;
        const int a = c & 3;
        r0 += sp[3 * a];
        r1 += sp[3 * a + 1];
        r2 += sp[3 * a + 2];
        if (wantH)
#pragma unroll
          for (int k = 0; k < 6; ++k) d[k] += sp[12 + 6 * a + k];
      }
      const size_t rp = rec[j];
      Rv[3 * rp] = r0;
      Rv[3 * rp + 1] = r1;
      Rv[3 * rp + 2] = r2;
      if (wantH)
#pragma unroll
        for (int k = 0; k < 6; ++k) Dv[6 * rp + size_t(k)] = d[k];
    }
  if (!wantH) return;
  __syncthreads();
  if (mine) {
    T* sv = stage + lane * kStageE;
#pragma unroll
    for (int k = 0; k < 54; ++k) sv[k] = off[k];
  }
  __syncthreads();
  // phase 2: off-diagonal block of every local pair, (min, max) orientation
  const uint32_t* prec = meta + PatchVertexWords(nLV, int(hd >> 20));
  const uint16_t* pls = reinterpret_cast<const uint16_t*>(prec + nLE);
  const uint16_t* pent = pls + nLE + 1;
  if (on)
    for (int j = lane; j < nLE; j += 64) {
      T o[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) o[k] = T(0);
      const int q1 = pls[j + 1];
      for (int q = pls[j]; q < q1; ++q) {
        const int c = pent[q];
        const int id = c & 0x7fff;
        const T* sp = stage + (id / 6) * kStageE + 9 * (id % 6);
        if (c >> 15) {
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) o[3 * r + cc] += sp[3 * cc + r];
        } else {
#pragma unroll
          for (int k = 0; k < 9; ++k) o[k] += sp[k];
        }
      }
      const size_t rp = prec[j];
#pragma unroll
      for (int k = 0; k < 9; ++k) Ev[9 * rp + size_t(k)] = o[k];
    }
}

// ------------------------------------------------------------------ contact

// Symmetric 3x3 block (a, b), a <= b, of a triangle's 6 in hc[36].
__host__ __device__ constexpr int hc_block(int a, int b) { return a == 0 ? b : (a == 1 ? 2 + b : 5); }
// Entry (r, c) of a symmetric 3x3 block stored as 00 11 22 01 02 12.
__device__ __forceinline__ int sym_entry(int r, int c) { return r == c ? r : 2 + r + c; }

// One trace-Q3 sample x of a boundary triangle (face normal fnv, displacement
// dxq = x - x0 over the step) against one analytic collider cd: the penalty
// (contact_utils.h:498-570) and its dissipation (contact_utils.h:630-935):
// C1-regularised Coulomb friction (activations.h:131-160), viscous friction
// and normal damping, each proportional to the penalty force at the sample's
// estimated stage-start distance d - n.pRel and faded by the alignment of the
// collider and face normals, with pRel = (x - x0) - (P - P0) the sample's
// displacement relative to the collider, and the fitted friction Hessian (the
// tangent-plane eigenvalue -mu fN df/|pT| in every tangent direction).
// Returns whether the sample touches; F is the force on the sample, its
// derivative cI I + cN n n^T (n = gr), and with want_e the energy density
// 1/2 k p^2 + fN (mu f(|pT|) + 1/2 cv/dt |pT|^2 + 1/2 cn/dt pn^2).
template <class T>
struct ContactSample {
  V3<T> F, gr;
  T cI, cN, dist;
  double e;
};

template <class T>
__device__ __forceinline__ bool contact_sample(const V3<T>& x, const V3<T>& dxq, const V3<T>& fnv,
                                               const double* __restrict__ cd, const double* __restrict__ cp, T dt,
                                               bool want_e, ContactSample<T>& out) {
  const T kp = T(cp[0]), hp = T(cp[1]), thr = T(cp[2]), maxAlign = T(cp[3]);
  const T mu = T(cp[4]), cvd = T(cp[5]) / dt, falloff = T(cp[6]) * dt, cnd = T(cp[7]) / dt;
  const int kind = (int)cd[0];
  const V3<T> P = mk3(T(cd[1]), T(cd[2]), T(cd[3]));
  T dist;
  V3<T> gr;
  if (kind == 0) {
    gr = mk3(T(cd[4]), T(cd[5]), T(cd[6]));
    dist = dot3(gr, sub3(x, P));
  } else {
    const V3<T> dv = sub3(x, P);
    const T l = dsqrt(dot3(dv, dv));
    gr = scl3(T(1) / l, dv);
    dist = l - T(cd[7]);
  }
  if (!(dist <= thr)) return false;
  const T align = dot3(gr, fnv);
  if (!(align <= maxAlign)) return false;
  T pen, dpen, ddpen;
  poly_relu<T>(-dist, hp, hp - thr, pen, dpen, ddpen);
  out.gr = gr;
  out.dist = dist;
  out.F = scl3(kp * pen * dpen, gr);
  out.cI = T(0);
  out.cN = -kp * (dpen * dpen + pen * ddpen);
  out.e = want_e ? 0.5 * double(kp) * double(pen) * double(pen) : 0.0;
  if (mu > T(0) || cvd > T(0) || cnd > T(0)) {
    const T fade = maxAlign > T(-1) ? (maxAlign - align) / (maxAlign + T(1)) : T(1);
    const V3<T> dP = mk3(T(cd[1] - cd[8]), T(cd[2] - cd[9]), T(cd[3] - cd[10]));  // P - P0
    const V3<T> pRel = sub3(dxq, dP);
    const T pn = dot3(gr, pRel);
    T p0, dp0, ddp0;
    poly_relu<T>(-(dist - pn), hp, hp - thr, p0, dp0, ddp0);
    const T fN = fade * kp * p0 * dp0;
    const V3<T> pT = sub3(pRel, scl3(pn, gr));
    const T pTn = dsqrt(dot3(pT, pT));
    T eigP = T(0), eigN = T(0);
    double ef = 0.0;
    if (mu > T(0)) {
      // IPCstepC1(|pT|, falloff): df / |pT| and the antiderivative f
      T dfx, fs;
      if (pTn >= falloff) {
        dfx = pTn == T(0) ? T(0) : T(1) / pTn;
        fs = pTn - falloff / T(3);
      } else {
        const T xt = pTn / falloff;
        dfx = (T(2) - xt) / falloff;
        fs = pTn * xt * (T(1) - xt / T(3));
      }
      const T tmp = mu * fN * dfx;
      out.F = sub3(out.F, scl3(tmp, pT));
      eigP = -tmp;
      ef += double(mu) * double(fs);
    }
    if (cvd > T(0)) {
      const T tmp = cvd * fN;
      out.F = sub3(out.F, scl3(tmp, pT));
      eigP -= tmp;
      ef += 0.5 * double(cvd) * double(pTn) * double(pTn);
    }
    if (cnd > T(0)) {
      const T tmp = cnd * fN;
      out.F = sub3(out.F, scl3(tmp * pn, gr));
      eigN = -tmp;
      ef += 0.5 * double(cnd) * double(pn) * double(pn);
    }
    out.cI = eigP;
    out.cN += eigN - eigP;
    if (want_e) out.e += double(fN) * ef;
  }
  return true;
}

// Trace-Q3 sample q of a triangle: barycentric weights (2/3 on vertex (q + 1) % 3).
template <class T>
__device__ __forceinline__ void tri_q3(int q, T N[3]) {
  const T kSix = T(1.0 / 6.0), kTwoThirds = T(2.0 / 3.0);
  N[0] = q == 2 ? kTwoThirds : kSix;
  N[1] = q == 0 ? kTwoThirds : kSix;
  N[2] = q == 1 ? kTwoThirds : kSix;
}

// Penalty contact and its dissipation (contact_sample) at the 3 trace-Q3
// samples of a boundary triangle against the analytic colliders of its
// environment (fem_traction.h:58-170): each sample adds -F N_i w to the
// residual of vertex i and -(cN n n^T + cI I) N_i N_j w to block (i, j).
// Outputs per triangle, written to their records (zeros when nothing
// touches): 3 residual rows, 3 diagonal blocks and the 3 blocks of its vertex
// pairs (symmetric, so no transposition).
// One boundary triangle tr of environment env against the environment's
// colliders: rc[9] its 3 vertices' residuals, with wantH hc[36] its 6 vertex
// pair blocks (hc_block order, each symmetric, 00 11 22 01 02 12).  Returns
// whether some sample touches (else rc and hc stay zero).
template <class T>
__device__ __forceinline__ bool contact_tri(const MeshView& m, const FieldPtrs<T>& f, const T* __restrict__ u,
                                            const T* __restrict__ u0, T dt, int tr, int env, bool wantH, T rc[9],
                                            T hc[36]) {
  bool touched = false;  // some quadrature point is in contact
#pragma unroll
  for (int k = 0; k < 9; ++k) rc[k] = T(0);
  if (wantH)
#pragma unroll
    for (int k = 0; k < 36; ++k) hc[k] = T(0);
  const int c0 = m.coll_first[env], c1 = m.coll_first[env + 1];
  if (c1 > c0) {
    const int nd[3] = {m.tris[3 * tr], m.tris[3 * tr + 1], m.tris[3 * tr + 2]};
    V3<T> xs[3], Xr[3], dxs[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      Xr[i] = mk3(f.X[3 * nd[i]], f.X[3 * nd[i] + 1], f.X[3 * nd[i] + 2]);
      const V3<T> ui = mk3(u[3 * nd[i]], u[3 * nd[i] + 1], u[3 * nd[i] + 2]);
      xs[i] = add3(Xr[i], ui);
      dxs[i] = sub3(ui, mk3(u0[3 * nd[i]], u0[3 * nd[i] + 1], u0[3 * nd[i] + 2]));  // x - x0
    }
    // rest area -> quad weight A0/3 ; current outward normal for the alignment mask
    const V3<T> n0 = cross3(sub3(Xr[1], Xr[0]), sub3(Xr[2], Xr[0]));
    const T w = T(0.5) * dsqrt(dot3(n0, n0)) / T(3);
    V3<T> fnv = cross3(sub3(xs[1], xs[0]), sub3(xs[2], xs[0]));
    fnv = scl3(T(1) / dsqrt(dot3(fnv, fnv)), fnv);
    const double* cp = f.contact + kContactParams * env;
    for (int c = c0; c < c1; ++c) {
      const double* cd = m.coll + kCollStride * c;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        T N[3];
        tri_q3<T>(q, N);
        const V3<T> x = add3(add3(scl3(N[0], xs[0]), scl3(N[1], xs[1])), scl3(N[2], xs[2]));
        const V3<T> dxq = add3(add3(scl3(N[0], dxs[0]), scl3(N[1], dxs[1])), scl3(N[2], dxs[2]));
        ContactSample<T> cs;
        if (!contact_sample<T>(x, dxq, fnv, cd, cp, dt, false, cs)) continue;
        touched = true;
        const T fv[3] = {cs.F.x, cs.F.y, cs.F.z};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int d = 0; d < 3; ++d) rc[3 * i + d] -= fv[d] * (N[i] * w);
        if (wantH) {
          const T gv[3] = {cs.gr.x, cs.gr.y, cs.gr.z};
          // -(cN n n^T + cI I) in 00 11 22 01 02 12 order
          const T gg[6] = {-(cs.cN * gv[0] * gv[0] + cs.cI), -(cs.cN * gv[1] * gv[1] + cs.cI),
                           -(cs.cN * gv[2] * gv[2] + cs.cI), -cs.cN * gv[0] * gv[1], -cs.cN * gv[0] * gv[2],
                           -cs.cN * gv[1] * gv[2]};
#pragma unroll
          for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = a; b < 3; ++b) {
              const T s = N[a] * N[b] * w;
              T* blk = hc + 6 * hc_block(a, b);
#pragma unroll
              for (int k = 0; k < 6; ++k) blk[k] += s * gg[k];
            }
        }
      }
    }
  }
  return touched;
}

template <class T>
__global__ void __launch_bounds__(256) k_contact(MeshView m, AsmRange ar, FieldPtrs<T> f, const T* __restrict__ u,
                                                 const T* __restrict__ u0, T dt, int wantH, T* __restrict__ Rv,
                                                 T* __restrict__ Dv, T* __restrict__ Ev) {
  const int tr = ar.f0 + int(blockIdx.x * blockDim.x + threadIdx.x);
  if (tr >= ar.f1) return;
  const int env = m.tri_env[tr];
  if (!m.env_mask[env]) return;
  T rc[9];
  T hc[36];
  const bool touched = contact_tri<T>(m, f, u, u0, dt, tr, env, wantH != 0, rc, hc);
  // A triangle out of contact has zero records: they are written only if
  // it is in contact now or its records still hold values from an earlier
  // write (flag bits), so most triangles write nothing; with contact_always
  // (records shared between assembly groups) every triangle writes them and
  // the flags stay set.
  const uint8_t fl = ar.contact_always ? uint8_t(3) : m.coll_flags[tr];
  uint8_t nf = uint8_t((fl & 2) | (touched ? 1 : 0));
  if (touched || (fl & 1)) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      T* re = Rv + (size_t)m.pos_vc[3 * tr + a] * 3;
      re[0] = rc[3 * a];
      re[1] = rc[3 * a + 1];
      re[2] = rc[3 * a + 2];
    }
  }
  if (wantH) {
    nf = uint8_t((nf & 1) | (touched ? 2 : 0));
    if (touched || (fl & 2)) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        T* d = Dv + (size_t)m.pos_vc[3 * tr + a] * 6;
#pragma unroll
        for (int k = 0; k < 6; ++k) d[k] = hc[6 * hc_block(a, a) + k];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T* q = hc + 6 * hc_block(kTriPairA[k], kTriPairB[k]);
        T* d = Ev + (size_t)m.pos_ec[3 * tr + k] * 9;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) d[3 * r + c] = q[sym_entry(r, c)];
      }
    }
  }
  if (nf != fl && !ar.contact_always) m.coll_flags[tr] = nf;
}

// ------------------------------------------------------------------ gathers

template <class T>
__device__ __forceinline__ void inv3(const T a[9], T o[9]) {
  const T c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
  const T c10 = a[2] * a[7] - a[1] * a[8], c11 = a[0] * a[8] - a[2] * a[6], c12 = a[1] * a[6] - a[0] * a[7];
  const T c20 = a[1] * a[5] - a[2] * a[4], c21 = a[2] * a[3] - a[0] * a[5], c22 = a[0] * a[4] - a[1] * a[3];
  const T id = T(1) / (a[0] * c00 + a[1] * c01 + a[2] * c02);
  o[0] = c00 * id;
  o[1] = c10 * id;
  o[2] = c20 * id;
  o[3] = c01 * id;
  o[4] = c11 * id;
  o[5] = c21 * id;
  o[6] = c02 * id;
  o[7] = c12 * id;
  o[8] = c22 * id;
}

// The preconditioner block of diagonal block b (row-major) for row `row`:
// block Jacobi (prec 2) its inverse, Jacobi (1) the reciprocal diagonal,
// none (0) the identity; 6 planes of nV, symmetric storage 00 11 22 01 02 12.
template <class T>
__device__ __forceinline__ void store_precond(const T b[9], int prec, T* __restrict__ dinv, int nV, int row) {
  T o[9];
  if (prec == 2) {
    inv3(b, o);
  } else {
#pragma unroll
    for (int q = 0; q < 9; ++q) o[q] = T(0);
    if (prec == 1) {
      o[0] = T(1) / b[0];
      o[4] = T(1) / b[4];
      o[8] = T(1) / b[8];
    } else {
      o[0] = o[4] = o[8] = T(1);
    }
  }
  const T sym[6] = {o[0], o[4], o[8], o[1], o[2], o[5]};
#pragma unroll
  for (int q = 0; q < 6; ++q) dinv[(size_t)q * nV + row] = sym[q];
}

// Every BSR slot: sum its element and contact contributions, apply Dirichlet
// rows/columns (mochi_solve.cpp:337-396), write the 9 SoA planes, and on
// diagonal slots the preconditioner block.
// 16-byte vector of T for staging copies: a native vector type, so an
// unrolled array of them stays in registers (HIP's float4 struct does not).
typedef float sdx_f4 __attribute__((ext_vector_type(4)));
typedef float sdx_f2 __attribute__((ext_vector_type(2)));
typedef double sdx_d2 __attribute__((ext_vector_type(2)));
template <class T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = sdx_f4;
};
template <>
struct Vec16<double> {
  using type = sdx_d2;
};

// LDS staging chunk of the gathers: 16 KB, four 16-byte loads per thread of
// a 256-thread workgroup.
template <class T>
constexpr int kStage = 16384 / int(sizeof(T));

// The workgroup copies src[c0, c1) (at most CH values, c1 > c0) into lds
// with 16-byte loads: stage_load issues a thread's loads into registers,
// stage_store writes them to LDS.  The loads are unconditional (a thread
// past the end reloads the last vector), so no branch separates them and a
// thread's loads are all in flight before the first wait; c0 is a multiple
// of the vector width and src is padded by one vector, so the last vector
// may run past c1 but never past the allocation.
template <class T, int NT, int CH>
struct Stage {
  using V = typename Vec16<T>::type;
  static constexpr int kW = 16 / int(sizeof(T));
  static constexpr int kU = CH / (kW * NT);  // vectors per thread
  static_assert(kU * kW * NT == CH, "chunk must be a whole number of vectors per thread");
  V v[kU];
  int nv;
  __device__ __forceinline__ void load(const T* __restrict__ src, int64_t c0, int64_t c1) {
    const V* s = reinterpret_cast<const V*>(src + c0);
    nv = int((c1 - c0 + kW - 1) / kW);
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int k = int(threadIdx.x) + u * NT;
      v[u] = s[k < nv ? k : nv - 1];
    }
  }
  __device__ __forceinline__ void store(T* lds) const {
    V* d = reinterpret_cast<V*>(lds);
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int k = int(threadIdx.x) + u * NT;
      if (k < nv) d[k] = v[u];
    }
  }
};

template <class T, int NT, int CH>
__device__ __forceinline__ void stage_chunk(const T* __restrict__ src, int64_t c0, int64_t c1, T* lds) {
  Stage<T, NT, CH> st;
  st.load(src, c0, c1);
  st.store(lds);
}

// Records k of [k0, k1) whose component e (of W per record) lies in the
// chunk [c0, c1): sum of lds[k W + e - c0].
template <class T, int W>
__device__ __forceinline__ T sum_run(const T* lds, int k0, int k1, int e, int64_t c0, int64_t c1) {
  const int64_t lo = c0 > e ? (c0 - e + W - 1) / W : 0;
  const int64_t hi = c1 > e ? (c1 - e + W - 1) / W : 0;
  const int a = int(lo > k0 ? lo : int64_t(k0)), b = int(hi < k1 ? hi : int64_t(k1));
  const T* p = lds + (int64_t(a) * W + e - c0);
  T s0 = T(0), s1 = T(0);
  int k = a;
  for (; k + 1 < b; k += 2, p += 2 * W) {
    s0 += p[0];
    s1 += p[W];
  }
  if (k < b) s0 += p[0];
  return s0 + s1;
}

// BSR gather: one lane per BSR slot, in slot order, so that the lanes of a
// wave write consecutive slots of each value plane.  A slot sums its records
// straight from global memory: a diagonal slot its vertex's diagonal-block
// records (00 11 22 01 02 12), an upper slot (i < j) its pair's off-diagonal
// records, a lower slot the same records transposed (each pair's records are
// read twice, every value written once, coalesced).  Then the Dirichlet
// rows / columns, and on a diagonal slot the preconditioner block.
constexpr int kGatherThreads = 256;

template <class T>
__global__ void __launch_bounds__(kGatherThreads) k_gather_bsr(MeshView m, AsmRange ar, const T* __restrict__ Dv,
                                                               const T* __restrict__ Ev, int prec,
                                                               T* __restrict__ vals, T* __restrict__ dinv) {
  const int k = ar.s0 + xcd_tile(blockIdx.x, gridDim.x) * kGatherThreads + int(threadIdx.x);
  if (k >= ar.s1) return;
  const int4 d = m.slot_desc[k];  // first and end record, row | pins, environment | kind
  if (!m.env_mask[d.w & 0x3fffffff]) return;
  const int kind = int(uint32_t(d.w) >> 30);
  T acc[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) acc[e] = T(0);
  if (kind == kSlotDiag) {
    for (int r = d.x; r < d.y; ++r) {
      const T* p = Dv + size_t(r) * 6;
#pragma unroll
      for (int e = 0; e < 6; ++e) acc[e] += p[e];
    }
  } else {
    for (int r = d.x; r < d.y; ++r) {
      const T* p = Ev + size_t(r) * 9;
#pragma unroll
      for (int e = 0; e < 9; ++e) acc[e] += p[e];
    }
  }
  // the block, row-major
  T b[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      b[3 * r + c] = kind == kSlotDiag ? acc[sym_entry(r, c)] : (kind == kSlotLower ? acc[3 * c + r] : acc[3 * r + c]);
  // Dirichlet rows / columns (mochi_solve.cpp:337-396): row pins in bits 0-2, column pins in bits 3-5
  const int pins = d.z >> kItemRowBits;
  const bool diag = kind == kSlotDiag;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (((pins >> r) | (pins >> (3 + c))) & 1) b[3 * r + c] = (diag && r == c) ? T(1) : T(0);
  const int nB = m.nB;
#pragma unroll
  for (int q = 0; q < 9; ++q) vals[(size_t)q * nB + k] = b[q];
  if (!diag) return;
  const int row = d.z & ((1 << kItemRowBits) - 1);
  store_precond<T>(b, prec, dinv, m.nV, row);
}

template <class T>
__global__ void __launch_bounds__(256) k_block_inverse(int nV, const int32_t* __restrict__ diag_slot,
                                                       const T* __restrict__ vals, int nB, int prec,
                                                       T* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  const int s = diag_slot[i];
  T a[9], o[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) a[k] = vals[(size_t)k * nB + s];
  if (prec == 2) {
    inv3(a, o);
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] = T(0);
    if (prec == 1) {
      o[0] = T(1) / a[0];
      o[4] = T(1) / a[4];
      o[8] = T(1) / a[8];
    } else {
      o[0] = o[4] = o[8] = T(1);
    }
  }
  // symmetric storage, 6 planes 00 11 22 01 02 12 (upper triangle of the inverse)
  const T sym[6] = {o[0], o[4], o[8], o[1], o[2], o[5]};
#pragma unroll
  for (int k = 0; k < 6; ++k) dinv[(size_t)k * nV + i] = sym[k];
}

// ------------------------------------------------------------- row blocks

// Fused assembly: one workgroup per row block (sdx_rowblocks.h).  The block's
// rows, their columns and LDS bases come first; every thread then computes
// tets of the block (and boundary triangles) and adds each contribution to a
// row of the block into LDS (float atomics, so the summation order varies
// from run to run); last, the block writes its rows' BSR values with the
// Dirichlet rows and columns (mochi_solve.cpp:337-396), the preconditioner
// blocks and the residual.  LDS: values [slots][9], residual [rows][3],
// columns [slots], per row its first global slot, LDS base (plus the end)
// and vertex, and the row of every slot (rows_lds_bytes_host).
constexpr int kRowBlockThreads = 256;
static_assert(kBlockRows <= kRowBlockRowsMax && kBlockRows <= kRowBlockThreads,
              "a row block's rows fit its LDS row arrays and one thread each");

// LDS index of column c in block row r (its columns lie in [base[r], base[r+1]), sorted).
__device__ __forceinline__ int row_slot(const int* __restrict__ lcol, const int* __restrict__ lbase, int r, int c) {
  int lo = lbase[r], hi = lbase[r + 1] - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (lcol[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Add a 3x3 block (row-major q, or its transpose) into LDS slot s.
template <class T>
__device__ __forceinline__ void lds_add_block(T* __restrict__ lv, int s, const T* q, bool transpose) {
  T* d = lv + 9 * s;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) atomicAdd(d + 3 * r + c, transpose ? q[3 * c + r] : q[3 * r + c]);
}
// The same for a symmetric block in 00 11 22 01 02 12 order.
template <class T>
__device__ __forceinline__ void lds_add_sym(T* __restrict__ lv, int s, const T* q) {
  T* d = lv + 9 * s;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) atomicAdd(d + 3 * r + c, q[sym_entry(r, c)]);
}

template <class T, class A, bool kDamp>
__global__ void __launch_bounds__(kRowBlockThreads) k_assemble_rows(MeshView m, RowBlockView rb, FieldPtrs<T> f,
                                                                    const T* __restrict__ u,
                                                                    const T* __restrict__ u0,
                                                                    const T* __restrict__ v0, T dt, T gx, T gy,
                                                                    T gz, int project, int prec,
                                                                    T* __restrict__ vals, T* __restrict__ dinv,
                                                                    T* __restrict__ res, A* acc) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int env = rb.env[b];
  if (!m.env_mask[env]) return;  // uniform over the workgroup
  const int ms = rb.max_slots;
  T* lv = reinterpret_cast<T*>(sdx_dyn_lds);
  T* lres = lv + size_t(9) * ms;
  int* lcol = reinterpret_cast<int*>(lres + 3 * kRowBlockRowsMax);
  int* lrow_g = lcol + ms;
  int* lbase = lrow_g + kRowBlockRowsMax;  // kRowBlockRowsMax + 1 entries
  int* lrow_v = lbase + kRowBlockRowsMax + 1;
  uint8_t* lslot_row = reinterpret_cast<uint8_t*>(lrow_v + kRowBlockRowsMax);
  const int tid = int(threadIdx.x);
  const int r0 = rb.row0[b], nr = rb.row0[b + 1] - r0;
  const int e0 = rb.tet0[b], ne = rb.tet0[b + 1] - e0;
  const int c0 = rb.tri0[b], nc = rb.tri0[b + 1] - c0;
  // rows: first global slot, LDS base, vertex; columns and slot rows
  if (tid < nr) {
    const int v = rb.rows[r0 + tid];
    const int s0 = m.row_ptr[v], len = m.row_ptr[v + 1] - s0;
    const int base = rb.rbase[r0 + tid];
    lrow_g[tid] = s0;
    lbase[tid] = base;
    lrow_v[tid] = v;
    if (tid == nr - 1) lbase[nr] = base + len;
    for (int k = 0; k < len; ++k) {
      lcol[base + k] = m.col[s0 + k];
      lslot_row[base + k] = uint8_t(tid);
    }
  }
  __syncthreads();
  const int nslots = lbase[nr];
  for (int j = tid; j < 9 * nslots; j += kRowBlockThreads) lv[j] = T(0);
  for (int j = tid; j < 3 * nr; j += kRowBlockThreads) lres[j] = T(0);
  __syncthreads();
  // tets
  const int nT = m.nT;
  for (int k = tid; k < ne; k += kRowBlockThreads) {
    const int2 te = rb.tets[e0 + k];
    const int t = te.x;
    const uint32_t lr = uint32_t(te.y);
    const int4 tv = reinterpret_cast<const int4*>(m.tets)[t];
    const int nd[4] = {tv.x, tv.y, tv.z, tv.w};
    TetIn<T> in;
    load_tet<T>(f, nT, t, nd, u, u0, v0, dt, in);
    auto own = [&](int a) { return int((lr >> (8 * a)) & kNotInBlock); };
    auto on_res = [&](const T* r) {
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int la = own(a);
        if (la == int(kNotInBlock)) continue;
#pragma unroll
        for (int d = 0; d < 3; ++d) atomicAdd(lres + 3 * la + d, r[3 * a + d]);
      }
    };
    auto emit = [&](int fa, int fb, const T* q) {
      const int la = own(fa);
      if (fa == fb) {
        if (la != int(kNotInBlock)) lds_add_sym<T>(lv, row_slot(lcol, lbase, la, nd[fa]), q);
        return;
      }
      if (la != int(kNotInBlock)) lds_add_block<T>(lv, row_slot(lcol, lbase, la, nd[fb]), q, false);
      const int lb = own(fb);
      if (lb != int(kNotInBlock)) lds_add_block<T>(lv, row_slot(lcol, lbase, lb, nd[fa]), q, true);
    };
    tet_contrib<T, kDamp>(f, nT, t, in, dt, gx, gy, gz, true, project != 0, on_res, emit);
  }
  // boundary triangles in contact
  for (int k = tid; k < nc; k += kRowBlockThreads) {
    const int2 fe = rb.tris[c0 + k];
    const int tr = fe.x;
    const uint32_t lr = uint32_t(fe.y);
    T rc[9], hc[36];
    if (!contact_tri<T>(m, f, u, u0, dt, tr, env, true, rc, hc)) continue;
    const int nd[3] = {m.tris[3 * tr], m.tris[3 * tr + 1], m.tris[3 * tr + 2]};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int li = int((lr >> (8 * i)) & kNotInBlock);
      if (li == int(kNotInBlock)) continue;
#pragma unroll
      for (int d = 0; d < 3; ++d) atomicAdd(lres + 3 * li + d, rc[3 * i + d]);
#pragma unroll
      for (int j = 0; j < 3; ++j)
        lds_add_sym<T>(lv, row_slot(lcol, lbase, li, nd[j]), hc + 6 * hc_block(i < j ? i : j, i < j ? j : i));
    }
  }
  __syncthreads();
  // the rows' BSR values, Dirichlet rows and columns, preconditioner blocks
  const int nB = m.nB;
  for (int j = tid; j < nslots; j += kRowBlockThreads) {
    const int r = lslot_row[j];
    const int v = lrow_v[r], c = lcol[j];
    const int gs = lrow_g[r] + (j - lbase[r]);
    T q[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) q[e] = lv[9 * j + e];
    const bool diag = c == v;
    const uint8_t* pv = m.pinned + 3 * v;
    const uint8_t* pc = m.pinned + 3 * c;
    const bool rp[3] = {pv[0] != 0, pv[1] != 0, pv[2] != 0}, cpn[3] = {pc[0] != 0, pc[1] != 0, pc[2] != 0};
#pragma unroll
    for (int rr = 0; rr < 3; ++rr)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        if (rp[rr] || cpn[cc]) q[3 * rr + cc] = (diag && rr == cc) ? T(1) : T(0);
#pragma unroll
    for (int e = 0; e < 9; ++e) vals[(size_t)e * nB + gs] = q[e];
    if (diag) store_precond<T>(q, prec, dinv, m.nV, v);
  }
  // the residual, Dirichlet rows zeroed, and ||res||^2
  A part[1] = {A(0)};
  if (tid < nr) {
    const int v = lrow_v[tid];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T x = m.pinned[3 * v + d] ? T(0) : lres[3 * tid + d];
      res[3 * v + d] = x;
      part[0] += A(x) * A(x);
    }
  }
  acc_add<A, 1>(part, acc_ptr(acc, 0, m.nEnv, env), b);
}

// One thread per vertex: residual = sum of its element and contact records,
// Dirichlet rows zeroed; per-environment ||res||^2 added into acc (set 0,
// zeroed by the caller).  The residual records of a tile's vertices are
// contiguous (vertex-sorted), so the workgroup stages them through LDS with
// 16-byte loads and each thread sums its run from LDS.
template <class T, class A>
__global__ void __launch_bounds__(kVecThreads) k_gather_res(MeshView m, AsmRange ar, Tiles tv,
                                                            const T* __restrict__ Rv, T* __restrict__ res, A* acc) {
  __shared__ typename Vec16<T>::type lds_v[kStage<T> * sizeof(T) / 16];
  T* lds = reinterpret_cast<T*>(lds_v);
  constexpr int kW = 16 / int(sizeof(T));
  const int tile = ar.tile0 + xcd_tile(blockIdx.x, gridDim.x);
  const int env = tv.env[tile];
  const int row0 = tv.row0[tile], row1 = tv.row1[tile];
  const int msk = m.env_mask[env];
  // the first chunk's records and the thread's record range are loaded
  // before the (uniform) environment mask is tested
  // (a tile of vertices in no element or triangle has no records)
  const int64_t start = (int64_t(m.vtx_ptr[row0]) * 3) & ~int64_t(kW - 1), end = int64_t(m.vtx_ptr[row1]) * 3;
  const bool any = start < end;
  Stage<T, kVecThreads, kStage<T>> st;
  int64_t c1 = start + kStage<T> < end ? start + kStage<T> : end;
  if (any) st.load(Rv, start, c1);
  const int i = row0 + int(threadIdx.x);
  const bool mine = i < row1;
  const int ic = mine ? i : row0;
  const int k0 = m.vtx_ptr[ic], k1 = m.vtx_ptr[ic + 1];
  if (!msk) return;  // uniform over the tile
  if (any) st.store(lds);
  __syncthreads();
  T r0 = T(0), r1 = T(0), r2 = T(0);
  for (int64_t c0 = start; any;) {
    if (mine) {
      r0 += sum_run<T, 3>(lds, k0, k1, 0, c0, c1);
      r1 += sum_run<T, 3>(lds, k0, k1, 1, c0, c1);
      r2 += sum_run<T, 3>(lds, k0, k1, 2, c0, c1);
    }
    c0 += kStage<T>;
    if (c0 >= end) break;  // uniform over the workgroup
    __syncthreads();
    c1 = c0 + kStage<T> < end ? c0 + kStage<T> : end;
    stage_chunk<T, kVecThreads, kStage<T>>(Rv, c0, c1, lds);
    __syncthreads();
  }
  A part[1] = {A(0)};
  if (mine) {
    if (m.pinned[3 * i]) r0 = T(0);
    if (m.pinned[3 * i + 1]) r1 = T(0);
    if (m.pinned[3 * i + 2]) r2 = T(0);
    res[3 * i] = r0;
    res[3 * i + 1] = r1;
    res[3 * i + 2] = r2;
    part[0] = A(r0) * A(r0) + A(r1) * A(r1) + A(r2) * A(r2);
  }
  acc_add<A, 1>(part, acc_ptr(acc, 0, m.nEnv, env), tile - tv.first[env]);
}

// ------------------------------------------------------------------ CG
//
// Single-gather Polak-Ribiere PCG.  With w = A z and s = A p kept by the
// recurrence s_k = w_k + beta_k s_{k-1}, iteration k needs only one matvec
// (w_k = A z_k) and one reduction point:
//   gamma_k = r_k.z_k, rho_k = r_k.z_{k-1}, delta_k = z_k.w_k
//   beta_k  = (gamma_k - [k >= 2] rho_k) / gamma_{k-1}          (pcg.h:205-219)
//   mu_k    = p_k.s_k = delta_k + 2 beta_k (rho_k - gamma_k) / alpha_{k-1}
//             + beta_k^2 mu_{k-1}                (exact for s = A p, symmetric A, M)
//   alpha_k = gamma_k / mu_k
// which is the reference iteration rearranged: p_k, x, r, z are the same
// vectors in exact arithmetic.
//
// Accumulator sets per environment (kCgSets of them):
//   0      scaling of the stop test from b: b.b, b.M^-1 b, |M^-1 b|^2
//   1, 2   R_k in set 1 + k % 2: r_k.r_k, r_k.z_k, z_k.z_k, r_k.z_{k-1}
//   3, 4   D_k in set 3 + k % 2: z_k.w_k
// cg_init writes R_0; cg_spmv(k) writes D_k and zeroes R_{k+1}; cg_update(k)
// writes R_{k+1} and zeroes D_{k+1}.  Both kernels of iteration k re-derive
// the same stop decision from R_k, so finished environments drop out of the
// matvec as well as the update.

constexpr int kSetScale = 0, kSetR = 1, kSetD = 3;

// The preconditioned residual z is stored padded to 4 components per vertex
// so that the matvec gathers a neighbour's z with one vector load.
template <class T>
struct alignas(4 * sizeof(T)) Z4 {
  T v[4];
};

// Raw buffer resources (gfx9 descriptor word 3 = 0x00020000: 32-bit data,
// no swizzle): a load names its array by a scalar descriptor and its element
// by a 32-bit byte offset in one register, where a global load needs a
// 64-bit address per element.  Loads past `bytes` return zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), short(0), int(bytes), 0x00020000);
}
template <class T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ float buf_ld<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <>
__device__ __forceinline__ double buf_ld<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <>
__device__ __forceinline__ int buf_ld<int>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return int(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
// The first three components of Z4 entry j.
__device__ __forceinline__ void buf_ld_z(__amdgpu_buffer_rsrc_t r, int j, float& a, float& b, float& c) {
  const sdx_f4 q = __builtin_bit_cast(sdx_f4, __builtin_amdgcn_raw_buffer_load_b128(r, uint32_t(j) * 16u, 0, 0));
  a = q.x;
  b = q.y;
  c = q.z;
}
__device__ __forceinline__ void buf_ld_z(__amdgpu_buffer_rsrc_t r, int j, double& a, double& b, double& c) {
  const sdx_d2 q = __builtin_bit_cast(sdx_d2, __builtin_amdgcn_raw_buffer_load_b128(r, uint32_t(j) * 32u, 0, 0));
  a = q.x;
  b = q.y;
  c = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, uint32_t(j) * 32u + 16u, 0, 0));
}

// U consecutive floats at byte offset off of a buffer in one load (U = 4, 2
// or 1; the offset need only be 4-byte aligned).
template <int U>
__device__ __forceinline__ void load_run(__amdgpu_buffer_rsrc_t r, uint32_t off, float (&v)[U]) {
  if constexpr (U == 4) {
    const sdx_f4 q = __builtin_bit_cast(sdx_f4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
  } else if constexpr (U == 2) {
    const sdx_f2 q = __builtin_bit_cast(sdx_f2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    v[0] = q.x;
    v[1] = q.y;
  } else {
    v[0] = buf_ld<float>(r, off);
  }
}

// Descriptors of the matvec operands: the 9 value planes, the column
// indices and z (each below 4 GB; checked by the launcher).
struct SpmvBufs {
  __amdgpu_buffer_rsrc_t v[9];
  __amdgpu_buffer_rsrc_t col, z;
};
template <class T>
__device__ __forceinline__ SpmvBufs spmv_bufs(const void* col, uint32_t col_bytes, const T* vals, int nB, const T* z,
                                              int nV) {
  SpmvBufs b;
#pragma unroll
  for (int c = 0; c < 9; ++c) b.v[c] = make_rsrc(vals + (size_t)c * nB, uint32_t(nB) * uint32_t(sizeof(T)));
  b.col = make_rsrc(col, col_bytes);
  b.z = make_rsrc(z, uint32_t(nV) * uint32_t(4 * sizeof(T)));
  return b;
}
template <class T>
__device__ __forceinline__ void store_z(T* z, int i, T a, T b, T c) {
  Z4<T> q;
  q.v[0] = a;
  q.v[1] = b;
  q.v[2] = c;
  q.v[3] = T(0);
  reinterpret_cast<Z4<T>*>(z)[i] = q;
}

// z = D^-1 r with the inverse of the (symmetric) diagonal block stored as 6
// planes 00 11 22 01 02 12; Jacobi keeps the reciprocal diagonal in planes 0-2.
template <class T>
__device__ __forceinline__ void apply_dinv(const T* __restrict__ dinv, int nV, int i, int prec, T r0, T r1, T r2,
                                           T& z0, T& z1, T& z2) {
  const T* d = dinv + i;
  const size_t s = nV;
  if (prec == 2) {
    const T d00 = d[0], d11 = d[s], d22 = d[2 * s], d01 = d[3 * s], d02 = d[4 * s], d12 = d[5 * s];
    z0 = d00 * r0 + d01 * r1 + d02 * r2;
    z1 = d01 * r0 + d11 * r1 + d12 * r2;
    z2 = d02 * r0 + d12 * r1 + d22 * r2;
  } else if (prec == 1) {
    z0 = d[0] * r0;
    z1 = d[s] * r1;
    z2 = d[2 * s] * r2;
  } else {
    z0 = r0;
    z1 = r1;
    z2 = r2;
  }
}

// Per-environment CG state that other workgroups of the same launch write
// (k_cg_tail runs both CG phases in one launch): read with vector loads at
// device scope, never through the scalar cache.
__device__ __forceinline__ int env_status(const EnvCG* env, int e) {
  return __hip_atomic_load(&env[e].status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double env_ld(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// StopTest (stopping_criterion.h:53-98): scaling from b, status of `latest`.
// Returns kCgActive, kCgConverged or kCgDiverged.
__device__ __forceinline__ int stop_test(const CgConfig& cfg, int e, double start, double latest) {
  const double div = (start * cfg.rel_div_tol) * cfg.rel_div_tol;
  const double rel = cfg.env_rel_tol ? cfg.env_rel_tol[e] : cfg.rel_tol;
  const double c1 = cfg.abs_tol * cfg.abs_tol, c2 = (start * rel) * rel;
  const double conv = c1 > c2 ? c1 : c2;
  if (!(latest == latest) || isinf(latest) || latest > div || latest < -conv) return kCgDiverged;
  if ((latest < 0.0 ? 0.0 : latest) <= conv) return kCgConverged;
  return kCgActive;
}

// Status of an environment before update k (pcg.h:186-222): stop test on
// r_k, a vanishing r.z, then the iteration budget.
__device__ __forceinline__ int cg_decide(const CgConfig& cfg, int e, int k, double start, double latest,
                                         double gamma) {
  int st = stop_test(cfg, e, start, latest);
  if (st == kCgActive && k >= 1 && gamma == 0.0) st = kCgDiverged;
  if (st == kCgActive && k >= cfg.max_iter) st = kCgStopped;
  return st;
}

// One block per environment: zero its accumulator sets and (re)start its CG state.
template <class A>
__global__ void k_cg_reset(int nEnv, const int32_t* __restrict__ solve, EnvCG* env, A* acc) {
  const int e = blockIdx.x;
  for (int set = 0; set < kCgSets; ++set) acc_zero(acc_ptr(acc, set, nEnv, e));
  if (threadIdx.x == 0) {
    EnvCG E = {};
    E.status = solve[e] ? kCgActive : kCgIdle;
    env[e] = E;
  }
}

// r0 = b - A x0 or b, z0 = M^-1 r0; scaling from b (SetScaling); R_0.
// Environments with warm[e] (warm may be NULL: none) keep x (scaled by
// alpha[e] when alpha is given) and take r0 from r64 = b - A x0 formed in
// FP64 (k_residual64); the others start from x = 0.
template <class T, class A>
__global__ void __launch_bounds__(kVecThreads) k_cg_init(Tiles tv, CgConfig cfg, const T* __restrict__ b,
                                                         const int32_t* __restrict__ warm,
                                                         const double* __restrict__ r64,
                                                         const double* __restrict__ alpha,
                                                         const T* __restrict__ v, const double* __restrict__ rv,
                                                         const double* __restrict__ beta, const T* __restrict__ w2,
                                                         const double* __restrict__ rw,
                                                         const double* __restrict__ gamma, const T* __restrict__ dinv,
                                                         int nV, A* __restrict__ x, T* __restrict__ r,
                                                         T* __restrict__ z, EnvCG* env, A* acc) {
  const int tile = blockIdx.x;
  const int e = tv.env[tile];
  if (env[e].status != kCgActive) return;
  const bool w = warm && warm[e];
  const int i = tv.row0[tile] + threadIdx.x;
  A sc[3] = {A(0), A(0), A(0)};        // b.b, b.Mb, Mb.Mb
  A rs[4] = {A(0), A(0), A(0), A(0)};  // r.r, r.z, z.z, 0
  if (i < tv.row1[tile]) {
    const T b0 = b[3 * i], b1 = b[3 * i + 1], b2 = b[3 * i + 2];
    T r0 = b0, r1 = b1, r2 = b2;
    if (w) {
      if (alpha && v) {  // x0 <- a x0 + c v + d w: r0 = a (b - A x0) + c (b - A v) + d (b - A w) + (1 - a - c - d) b
        const double a = alpha[e], c = beta[e], d = gamma[e], o = 1.0 - a - c - d;
        r0 = T(a * r64[3 * i] + c * rv[3 * i] + d * rw[3 * i] + o * double(b0));
        r1 = T(a * r64[3 * i + 1] + c * rv[3 * i + 1] + d * rw[3 * i + 1] + o * double(b1));
        r2 = T(a * r64[3 * i + 2] + c * rv[3 * i + 2] + d * rw[3 * i + 2] + o * double(b2));
        x[3 * i] = A(a * double(x[3 * i]) + c * double(v[3 * i]) + d * double(w2[3 * i]));
        x[3 * i + 1] = A(a * double(x[3 * i + 1]) + c * double(v[3 * i + 1]) + d * double(w2[3 * i + 1]));
        x[3 * i + 2] = A(a * double(x[3 * i + 2]) + c * double(v[3 * i + 2]) + d * double(w2[3 * i + 2]));
      } else if (alpha) {  // x0 <- a x0, so r0 = b - a A x0 = a r64 + (1 - a) b
        const double a = alpha[e];
        r0 = T(a * r64[3 * i] + (1.0 - a) * double(b0));
        r1 = T(a * r64[3 * i + 1] + (1.0 - a) * double(b1));
        r2 = T(a * r64[3 * i + 2] + (1.0 - a) * double(b2));
        x[3 * i] = A(a) * x[3 * i];
        x[3 * i + 1] = A(a) * x[3 * i + 1];
        x[3 * i + 2] = A(a) * x[3 * i + 2];
      } else {
        r0 = T(r64[3 * i]);
        r1 = T(r64[3 * i + 1]);
        r2 = T(r64[3 * i + 2]);
      }
    } else {
      x[3 * i] = A(0);
      x[3 * i + 1] = A(0);
      x[3 * i + 2] = A(0);
    }
    if (cfg.xd) {
      T* xd = static_cast<T*>(cfg.xd);
      xd[3 * i] = T(0);
      xd[3 * i + 1] = T(0);
      xd[3 * i + 2] = T(0);
    }
    T zb0, zb1, zb2, z0, z1, z2;
    apply_dinv(dinv, nV, i, cfg.prec, b0, b1, b2, zb0, zb1, zb2);
    apply_dinv(dinv, nV, i, cfg.prec, r0, r1, r2, z0, z1, z2);
    r[3 * i] = r0;
    r[3 * i + 1] = r1;
    r[3 * i + 2] = r2;
    store_z(z, i, z0, z1, z2);
    sc[0] = A(b0) * A(b0) + A(b1) * A(b1) + A(b2) * A(b2);
    sc[1] = A(b0) * A(zb0) + A(b1) * A(zb1) + A(b2) * A(zb2);
    sc[2] = A(zb0) * A(zb0) + A(zb1) * A(zb1) + A(zb2) * A(zb2);
    rs[0] = A(r0) * A(r0) + A(r1) * A(r1) + A(r2) * A(r2);
    rs[1] = A(r0) * A(z0) + A(r1) * A(z1) + A(r2) * A(z2);
    rs[2] = A(z0) * A(z0) + A(z1) * A(z1) + A(z2) * A(z2);
  }
  const int te = tile - tv.first[e];
  acc_add<A, 3>(sc, acc_ptr(acc, kSetScale, cfg.n_env, e), te);
  acc_add<A, 4>(rs, acc_ptr(acc, kSetR, cfg.n_env, e), te);
}

// Finish an accumulator read started with one slot load per lane.
__device__ __forceinline__ double slot_reduce(double v) {
  static_assert(kSlots == 16, "one accumulator slot per lane of a row of 16");
  return group_sum<kSlots>(v);
}


// CG iteration k, matvec half.  Every block re-derives the stop decision on
// r_k from R_k; the first tile of a finished environment records it.  Then
// w_k = A z_k is stored and z_k.w_k goes into D_k.  On refresh iterations the
// true residual rref = b - A x_k is also formed in FP64 (residual replacement).
template <class T, class A, bool kRefresh, int L, bool kC16>
__device__ __forceinline__ void cg_spmv_tile(int tile, const Tiles& ta, const CgConfig& cfg, int k,
                                             const int32_t* __restrict__ row_ptr, const SpmvBufs& bf,
                                             const T* __restrict__ z, const A* __restrict__ x,
                                             const T* __restrict__ b, T* __restrict__ w,
                                             double* __restrict__ rref, EnvCG* env, A* acc, int32_t* num_active) {
  const int e = ta.env[tile];
  // a work list is compacted only at burst starts: the tiles of an environment
  // that finished inside the burst return before loading their rows
  if (env_status(env, e) != kCgActive) return;
  const int nEnv = cfg.n_env;
  // raw accumulator slots (one per lane of each 16-lane group), issued first
  const int sl = (threadIdx.x & (kSlots - 1)) * kSlotStride;
  const int nsel = cfg.norm_type == 0 ? 0 : (cfg.norm_type == 1 ? 2 : 1);
  const A* s_r = acc_ptr(acc, kSetR + (k & 1), nEnv, e) + sl;
  double w_start = double(acc_ptr(acc, kSetScale, nEnv, e)[sl + nsel]);
  double w_latest = double(s_r[nsel]);
  double w_gamma = double(s_r[1]);
  const int i = ta.row0[tile] + (threadIdx.x / L);
  const int lane = threadIdx.x % L;
  const bool valid = i < ta.row1[tile];
  T a0 = T(0), a1 = T(0), a2 = T(0);
  double q0 = 0.0, q1 = 0.0, q2 = 0.0;
  T zi0 = T(0), zi1 = T(0), zi2 = T(0);
  if (valid) {
    const int k0 = row_ptr[i], k1 = row_ptr[i + 1];
    constexpr int U = kSpmvChunk / L > 0 ? kSpmvChunk / L : 1;
    // one block: w_i += A_ij z_j (one 16-byte gather of z_j), and on refresh
    // iterations the FP64 product with the iterate x_j + xd_j
    auto block = [&](T v0, T v1, T v2, T v3, T v4, T v5, T v6, T v7, T v8, int j) {
      T zj0, zj1, zj2;
      buf_ld_z(bf.z, j, zj0, zj1, zj2);  // FP64: 24 of 32 bytes
      a0 += v0 * zj0 + v1 * zj1 + v2 * zj2;
      a1 += v3 * zj0 + v4 * zj1 + v5 * zj2;
      a2 += v6 * zj0 + v7 * zj1 + v8 * zj2;
      if (kRefresh) {
        double x0 = x[3 * j], x1 = x[3 * j + 1], x2 = x[3 * j + 2];
        if (cfg.xd) {  // the increments not yet folded into x
          const T* xd = static_cast<const T*>(cfg.xd);
          x0 += double(xd[3 * j]);
          x1 += double(xd[3 * j + 1]);
          x2 += double(xd[3 * j + 2]);
        }
        q0 += double(v0) * x0 + double(v1) * x1 + double(v2) * x2;
        q1 += double(v3) * x0 + double(v4) * x1 + double(v5) * x2;
        q2 += double(v6) * x0 + double(v7) * x1 + double(v8) * x2;
      }
    };
    if constexpr (sizeof(T) == 4) {
      // runs of U L blocks: the lane takes U consecutive blocks of each run,
      // one U-wide load per value plane; its slots past the row's end weigh
      // zero and name the row itself (loads past the end of a buffer return
      // zero)
      for (int kb = k0 + lane * U; kb < k1; kb += U * L) {
        float m[9][U];
#pragma unroll
        for (int c = 0; c < 9; ++c) load_run<U>(bf.v[c], uint32_t(kb) * 4u, m[c]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const bool in = kb + u < k1;
          const T wu = in ? T(1) : T(0);
          int j;
          if (kC16)
            j = i + int(short(__builtin_amdgcn_raw_buffer_load_b16(bf.col, uint32_t(kb + u) * 2u, 0, 0)));
          else
            j = buf_ld<int>(bf.col, uint32_t(kb + u) * 4u);
          block(m[0][u] * wu, m[1][u] * wu, m[2][u] * wu, m[3][u] * wu, m[4][u] * wu, m[5][u] * wu, m[6][u] * wu,
                m[7][u] * wu, m[8][u] * wu, in ? j : i);
        }
      }
    } else {
      // FP64 values: a lane's blocks strided by L, indices past the row end
      // clamped to its last block (a row holds its diagonal) and weighted
      // by zero, so no load waits on a branch
      for (int kb = k0 + lane; kb < k1; kb += U * L) {
        int kk[U], jc[U];
        T wv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int kr = kb + u * L;
          kk[u] = kr < k1 ? kr : k1 - 1;
          wv[u] = kr < k1 ? T(1) : T(0);
          if (kC16)
            jc[u] = i + int(short(__builtin_amdgcn_raw_buffer_load_b16(bf.col, uint32_t(kk[u]) * 2u, 0, 0)));
          else
            jc[u] = buf_ld<int>(bf.col, uint32_t(kk[u]) * 4u);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t o = uint32_t(kk[u]) * uint32_t(sizeof(T));
          block(buf_ld<T>(bf.v[0], o) * wv[u], buf_ld<T>(bf.v[1], o) * wv[u], buf_ld<T>(bf.v[2], o) * wv[u],
                buf_ld<T>(bf.v[3], o) * wv[u], buf_ld<T>(bf.v[4], o) * wv[u], buf_ld<T>(bf.v[5], o) * wv[u],
                buf_ld<T>(bf.v[6], o) * wv[u], buf_ld<T>(bf.v[7], o) * wv[u], buf_ld<T>(bf.v[8], o) * wv[u], jc[u]);
        }
      }
    }
    if (lane == 0) {
      zi0 = z[4 * i];
      zi1 = z[4 * i + 1];
      zi2 = z[4 * i + 2];
    }
  }
  const double start = slot_reduce(w_start);
  const double latest = slot_reduce(w_latest);
  const double gamma = slot_reduce(w_gamma);
  const int st = cg_decide(cfg, e, k, start, latest, gamma);
  const bool lead = tile == ta.first[e];
  if (st != kCgActive) {
    if (lead && threadIdx.x == 0) {
      EnvCG& E = env[e];
      E.status = st;
      E.iter = k;
      E.latest = latest < 0.0 ? 0.0 : latest;
      E.start = start;
      E.rz = gamma;
      atomicSub(num_active, 1);
    }
    return;
  }
  if (lead) acc_zero(acc_ptr(acc, kSetR + ((k + 1) & 1), nEnv, e));
  a0 = group_sum<L>(a0);
  a1 = group_sum<L>(a1);
  a2 = group_sum<L>(a2);
  if (kRefresh) {
    q0 = group_sum<L>(q0);
    q1 = group_sum<L>(q1);
    q2 = group_sum<L>(q2);
  }
  A part[1] = {A(0)};
  if (valid && lane == 0) {
    w[3 * i] = a0;
    w[3 * i + 1] = a1;
    w[3 * i + 2] = a2;
    part[0] = A(zi0) * A(a0) + A(zi1) * A(a1) + A(zi2) * A(a2);
    if (kRefresh) {
      rref[3 * i] = double(b[3 * i]) - q0;
      rref[3 * i + 1] = double(b[3 * i + 1]) - q1;
      rref[3 * i + 2] = double(b[3 * i + 2]) - q2;
    }
  }
  acc_add<A, 1>(part, acc_ptr(acc, kSetD + (k & 1), nEnv, e), tile - ta.first[e]);
}

// CG iteration k, update half: beta_k, mu_k, alpha_k from R_k, D_k and the
// previous iteration's scalars; p_k = z_k + beta p, s_k = w_k + beta s,
// x += alpha p_k, r -= alpha s_k (r = rref - alpha s_k on refresh
// iterations), z = M^-1 r; R_{k+1} partials.  Row loads are issued before
// the scalars are known.
template <class T, class A, bool kRefresh>
__device__ __forceinline__ void cg_update_tile(int tile, const Tiles& tv, const CgConfig& cfg, int k,
                                               const T* __restrict__ dinv, int nV, const T* __restrict__ w,
                                               const double* __restrict__ rref, A* __restrict__ x,
                                               T* __restrict__ r, T* __restrict__ z, T* __restrict__ p,
                                               T* __restrict__ s, EnvCG* env, A* acc, int32_t* num_active) {
  const int e = tv.env[tile];
  if (env_status(env, e) != kCgActive) return;  // finished inside the burst: load nothing
  const int nEnv = cfg.n_env;
  const int sl = (threadIdx.x & (kSlots - 1)) * kSlotStride;
  const int nsel = cfg.norm_type == 0 ? 0 : (cfg.norm_type == 1 ? 2 : 1);
  const A* s_r = acc_ptr(acc, kSetR + (k & 1), nEnv, e) + sl;
  double w_start = double(acc_ptr(acc, kSetScale, nEnv, e)[sl + nsel]);
  double w_latest = double(s_r[nsel]);
  double w_gamma = double(s_r[1]);
  double w_rho = double(s_r[3]);
  double w_delta = double(acc_ptr(acc, kSetD + (k & 1), nEnv, e)[sl]);
  const int pk = (k + 1) & 1;  // slot of the previous iteration's scalars
  const double g_prev = env_ld(&env[e].gamma[pk]), a_prev = env_ld(&env[e].alpha[pk]), m_prev = env_ld(&env[e].mu[pk]);
  const int i = tv.row0[tile] + threadIdx.x;
  const bool valid = i < tv.row1[tile];
  A x0 = A(0), x1 = A(0), x2 = A(0);
  T z0 = T(0), z1 = T(0), z2 = T(0), w0 = T(0), w1 = T(0), w2 = T(0);
  T p0 = T(0), p1 = T(0), p2 = T(0), s0 = T(0), s1 = T(0), s2 = T(0);
  double ro0 = 0.0, ro1 = 0.0, ro2 = 0.0;
  T* xd = static_cast<T*>(cfg.xd);
  T d0 = T(0), d1 = T(0), d2 = T(0);
  if (valid) {
    if (xd) {  // increments since the last refresh; x itself only on refresh iterations
      d0 = xd[3 * i];
      d1 = xd[3 * i + 1];
      d2 = xd[3 * i + 2];
    }
    if (!xd || kRefresh) {
      x0 = x[3 * i];
      x1 = x[3 * i + 1];
      x2 = x[3 * i + 2];
    }
    z0 = z[4 * i];
    z1 = z[4 * i + 1];
    z2 = z[4 * i + 2];
    w0 = w[3 * i];
    w1 = w[3 * i + 1];
    w2 = w[3 * i + 2];
    if (k > 0) {
      p0 = p[3 * i];
      p1 = p[3 * i + 1];
      p2 = p[3 * i + 2];
      s0 = s[3 * i];
      s1 = s[3 * i + 1];
      s2 = s[3 * i + 2];
    }
    if (kRefresh) {
      ro0 = rref[3 * i];
      ro1 = rref[3 * i + 1];
      ro2 = rref[3 * i + 2];
    } else {
      ro0 = r[3 * i];
      ro1 = r[3 * i + 1];
      ro2 = r[3 * i + 2];
    }
  }
  const double start = slot_reduce(w_start);
  const double latest = slot_reduce(w_latest);
  const double gamma = slot_reduce(w_gamma);
  if (cg_decide(cfg, e, k, start, latest, gamma) != kCgActive) return;  // recorded by cg_spmv(k)
  const double rho = slot_reduce(w_rho);
  const double delta = slot_reduce(w_delta);
  double beta = 0.0, mu = delta;
  if (k >= 1) {
    beta = (gamma - (k >= 2 ? rho : 0.0)) / g_prev;
    mu = delta + 2.0 * beta * (rho - gamma) / a_prev + beta * beta * m_prev;
  }
  const bool lead = tile == tv.first[e];
  if (mu <= 0.0 && cfg.abort_if_not_spd) {
    // pcg.h:152-169: not SPD, abort with iteration k + 1
    if (lead && threadIdx.x == 0) {
      EnvCG& E = env[e];
      E.status = kCgDiverged;
      E.iter = k + 1;
      E.latest = latest < 0.0 ? 0.0 : latest;
      E.start = start;
      E.rz = gamma;
      atomicSub(num_active, 1);
    }
    return;
  }
  const double alpha = gamma / mu;
  if (lead) {
    acc_zero(acc_ptr(acc, kSetD + ((k + 1) & 1), nEnv, e));
    if (threadIdx.x == 0) {
      EnvCG& E = env[e];
      E.gamma[k & 1] = gamma;
      E.alpha[k & 1] = alpha;
      E.mu[k & 1] = mu;
    }
  }
  const T al = T(alpha), bt = T(beta);
  A part[4] = {A(0), A(0), A(0), A(0)};  // r.r, r.z, z.z, r.z_old
  if (valid) {
    p0 = z0 + bt * p0;
    p1 = z1 + bt * p1;
    p2 = z2 + bt * p2;
    s0 = w0 + bt * s0;
    s1 = w1 + bt * s1;
    s2 = w2 + bt * s2;
    p[3 * i] = p0;
    p[3 * i + 1] = p1;
    p[3 * i + 2] = p2;
    s[3 * i] = s0;
    s[3 * i + 1] = s1;
    s[3 * i + 2] = s2;
    // the iterate accumulates in the dot-product type (FP64 in mixed
    // precision); with xd, in FP32 increments that refresh iterations fold in
    if (xd && !kRefresh) {
      xd[3 * i] = d0 + al * p0;
      xd[3 * i + 1] = d1 + al * p1;
      xd[3 * i + 2] = d2 + al * p2;
    } else {
      x[3 * i] = x0 + A(d0) + A(alpha) * A(p0);
      x[3 * i + 1] = x1 + A(d1) + A(alpha) * A(p1);
      x[3 * i + 2] = x2 + A(d2) + A(alpha) * A(p2);
      if (xd) {
        xd[3 * i] = T(0);
        xd[3 * i + 1] = T(0);
        xd[3 * i + 2] = T(0);
      }
    }
    T r0, r1, r2;
    if (kRefresh) {
      r0 = T(ro0 - alpha * double(s0));
      r1 = T(ro1 - alpha * double(s1));
      r2 = T(ro2 - alpha * double(s2));
    } else {
      r0 = T(ro0) - al * s0;
      r1 = T(ro1) - al * s1;
      r2 = T(ro2) - al * s2;
    }
    r[3 * i] = r0;
    r[3 * i + 1] = r1;
    r[3 * i + 2] = r2;
    T n0, n1, n2;
    apply_dinv(dinv, nV, i, cfg.prec, r0, r1, r2, n0, n1, n2);
    store_z(z, i, n0, n1, n2);
    if (kRefresh) {
      // the mu recurrence of the next iteration needs z_{k+1}.(r_k this update
      // started from) = r_{k+1}.M^-1 rref: replace z_k by M^-1 rref
      apply_dinv(dinv, nV, i, cfg.prec, T(ro0), T(ro1), T(ro2), z0, z1, z2);
    }
    part[0] = A(r0) * A(r0) + A(r1) * A(r1) + A(r2) * A(r2);
    part[1] = A(r0) * A(n0) + A(r1) * A(n1) + A(r2) * A(n2);
    part[2] = A(n0) * A(n0) + A(n1) * A(n1) + A(n2) * A(n2);
    part[3] = A(r0) * A(z0) + A(r1) * A(z1) + A(r2) * A(z2);
  }
  acc_add<A, 4>(part, acc_ptr(acc, kSetR + ((k + 1) & 1), nEnv, e), tile - tv.first[e]);
}

// A CG kernel runs one tile per workgroup (list == NULL: tiles in XCD-
// contiguous order), or is persistent over a work list of the tiles of the
// environments still active: the list is split into 8 contiguous chunks, one
// per XCD, and each workgroup strides over its XCD's chunk.
template <class F>
__device__ __forceinline__ void for_each_tile(const WorkList& wl, int n_tiles, F&& f) {
  if (!wl.list) {
    f(xcd_tile(blockIdx.x, n_tiles));
    return;
  }
  constexpr int kXcd = 8;
  const int n = *wl.count;
  const int xc = blockIdx.x % kXcd, i = blockIdx.x / kXcd;
  const int nb = (int(gridDim.x) - xc + kXcd - 1) / kXcd;  // workgroups on this XCD (grid >= 8)
  const int q1 = int((int64_t)n * (xc + 1) / kXcd);
  for (int q = int((int64_t)n * xc / kXcd) + i; q < q1; q += nb) {
    f(wl.list[q]);
    __syncthreads();  // the tile's block reduction reuses LDS
  }
}

// kC16: col holds each block's column as a signed 16-bit offset from its row.
template <class T, class A, bool kRefresh, int L, bool kC16>
__global__ void __launch_bounds__(kSpmvThreads) k_cg_spmv(Tiles ta, WorkList wl, CgConfig cfg, int k,
                                                          const int32_t* __restrict__ row_ptr,
                                                          const void* __restrict__ col, const T* __restrict__ vals,
                                                          int nB, const T* __restrict__ z, const A* __restrict__ x,
                                                          const T* __restrict__ b, T* __restrict__ w,
                                                          double* __restrict__ rref, EnvCG* env, A* acc,
                                                          int32_t* num_active) {
  const SpmvBufs bf = spmv_bufs<T>(col, uint32_t(nB) * (kC16 ? 2u : 4u), vals, nB, z, cfg.n_rows);
  for_each_tile(wl, wl.list ? 0 : int(gridDim.x), [&](int tile) {
    cg_spmv_tile<T, A, kRefresh, L, kC16>(tile, ta, cfg, k, row_ptr, bf, z, x, b, w, rref, env, acc, num_active);
  });
}

template <class T, class A, bool kRefresh>
__global__ void __launch_bounds__(kVecThreads) k_cg_update(Tiles tv, WorkList wl, CgConfig cfg, int k,
                                                           const T* __restrict__ dinv, int nV,
                                                           const T* __restrict__ w, const double* __restrict__ rref,
                                                           A* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                           T* __restrict__ p, T* __restrict__ s, EnvCG* env, A* acc,
                                                           int32_t* num_active) {
  if (!wl.list) {  // vector tiles need no XCD placement
    cg_update_tile<T, A, kRefresh>(blockIdx.x, tv, cfg, k, dinv, nV, w, rref, x, r, z, p, s, env, acc, num_active);
    return;
  }
  for_each_tile(wl, 0, [&](int tile) {
    cg_update_tile<T, A, kRefresh>(tile, tv, cfg, k, dinv, nV, w, rref, x, r, z, p, s, env, acc, num_active);
  });
}

// ------------------------------------------------------------ whole-solve CG
//
// Environments never exchange data, so one workgroup can run every iteration
// of one environment's PCG by itself (schedule env_cg): workgroups take
// environments from a queue, largest first, and iterate each to its stop
// decision with the arithmetic of cg_spmv_tile and cg_update_tile.  z and w
// live in LDS, an iteration's sums are workgroup reductions in a fixed order,
// and nothing waits on another workgroup or on a kernel boundary.  The rows of
// an environment's matrix are re-read every iteration: with one or two
// workgroups per CU, the matrices in flight stay within the last-level cache.

// Sum N values over the workgroup, in every thread, in a fixed order: wave
// sums, then the waves' partials (LDS `red`, kEnvWaves * N) in wave order.
template <int N, class A>
__device__ __forceinline__ void env_block_sum(A (&v)[N], A* red) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < N; ++c) v[c] = wave_sum(v[c]);
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < N; ++c) red[wave * N + c] = v[c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < N; ++c) {
    A t = red[c];
    for (int w = 1; w < kEnvWaves; ++w) t += red[w * N + c];
    v[c] = t;
  }
}

// One environment's CG from the state k_cg_init left (r, z, x, the scale and
// R_0 accumulators) to its stop decision; the same iterations, scalars and
// stop tests as k_cg_spmv + k_cg_update, and the same EnvCG record at the end.
// kVec: r, p and s also live in LDS (rl, pl, sl), r loaded from r at the start.
template <class T, class A, int L, bool kC16, bool kVec>
__device__ __forceinline__ void cg_env_solve(int e, int v0, int nr, const CgConfig& cfg,
                                             const int32_t* __restrict__ row_ptr, const SpmvBufs& bf,
                                             const T* __restrict__ dinv, int nV, const T* __restrict__ b,
                                             const T* __restrict__ zin, A* __restrict__ x, T* __restrict__ r,
                                             T* __restrict__ p, T* __restrict__ s, double* __restrict__ rref,
                                             EnvCG* env, const A* __restrict__ acc, int32_t* num_active, Z4<T>* zl,
                                             T* z3, T* wl, T* rl, T* pl, T* sl, T* xl, A* red) {
  const int nsel = cfg.norm_type == 0 ? 0 : (cfg.norm_type == 1 ? 2 : 1);
  T* xd = static_cast<T*>(cfg.xd);
  // z of local row k: Z4 entries, or 3 values per row with the CG vectors in LDS
  auto zget = [&](int k) {
    if (kVec) {
      Z4<T> q;
      q.v[0] = z3[3 * k];
      q.v[1] = z3[3 * k + 1];
      q.v[2] = z3[3 * k + 2];
      q.v[3] = T(0);
      return q;
    }
    return zl[k];
  };
  for (int il = threadIdx.x; il < nr; il += kEnvThreads) {
    const Z4<T> q = reinterpret_cast<const Z4<T>*>(zin)[v0 + il];
    if (kVec) {
      z3[3 * il] = q.v[0];
      z3[3 * il + 1] = q.v[1];
      z3[3 * il + 2] = q.v[2];
      xl[3 * il] = T(0);
      xl[3 * il + 1] = T(0);
      xl[3 * il + 2] = T(0);
    } else {
      zl[il] = q;
    }
    if (kVec) {
      rl[3 * il] = r[3 * (v0 + il)];
      rl[3 * il + 1] = r[3 * (v0 + il) + 1];
      rl[3 * il + 2] = r[3 * (v0 + il) + 2];
    }
  }
  // where this thread's rows of r, p and s live: LDS (local row) or global (row)
  T* const rv = kVec ? rl : r;
  T* const pv = kVec ? pl : p;
  T* const sv = kVec ? sl : s;
  // the scale and R_0 of k_cg_init, summed over the accumulator slots in slot order
  double start = 0.0, R[4] = {0.0, 0.0, 0.0, 0.0};
  {
    const A* sc = acc_ptr(acc, kSetScale, cfg.n_env, e);
    const A* rs = acc_ptr(acc, kSetR, cfg.n_env, e);
    A st = A(0), r4[4] = {A(0), A(0), A(0), A(0)};
    for (int sl = 0; sl < kSlots; ++sl) {
      st += sc[sl * kSlotStride + nsel];
#pragma unroll
      for (int c = 0; c < 4; ++c) r4[c] += rs[sl * kSlotStride + c];
    }
    start = double(st);
#pragma unroll
    for (int c = 0; c < 4; ++c) R[c] = double(r4[c]);
  }
  // the iterate's increments of row i (local row il): LDS with the CG vectors there
  T* const xdv = kVec ? xl : xd;
  const bool has_xd = xd != nullptr;
  // at the end of the solve, with the increments in LDS: this thread's rows (the
  // update's) fold theirs into x
  auto fold = [&]() {
    if (kVec && has_xd)
      for (int il = threadIdx.x; il < nr; il += kEnvThreads) {
        const int i = v0 + il;
        x[3 * i] += A(xl[3 * il]);
        x[3 * i + 1] += A(xl[3 * il + 1]);
        x[3 * i + 2] += A(xl[3 * il + 2]);
      }
  };
  const int sub = threadIdx.x / L, lane = threadIdx.x % L;
  double g_prev = 0.0, a_prev = 0.0, m_prev = 0.0;
  __syncthreads();  // z in LDS
  for (int k = 0;; ++k) {
    const double latest = R[nsel], gamma = R[1], rho = R[3];
    const int st = cg_decide(cfg, e, k, start, latest, gamma);
    if (st != kCgActive) {
      fold();
      if (threadIdx.x == 0) {
        EnvCG& E = env[e];
        E.status = st;
        E.iter = k;
        E.latest = latest < 0.0 ? 0.0 : latest;
        E.start = start;
        E.rz = gamma;
        atomicSub(num_active, 1);
      }
      return;
    }
    const bool refresh = cg_refresh(cfg, k);
    // matvec: w = A z into LDS, z.w; on refresh iterations also the FP64
    // true residual rref = b - A (x + xd) of the iterate
    A dz[1] = {A(0)};
    for (int base = 0; base < nr; base += kEnvThreads / L) {
      const int il = base + sub;
      const bool valid = il < nr;
      const int i = v0 + il;
      T a0 = T(0), a1 = T(0), a2 = T(0);
      double q0 = 0.0, q1 = 0.0, q2 = 0.0;
      if (valid) {
        const int k0 = row_ptr[i], k1 = row_ptr[i + 1];
        constexpr int U = kSpmvChunk / L > 0 ? kSpmvChunk / L : 1;
        // one block of the row: w_i += A_ij z_j (z from LDS), and on refresh
        // iterations the FP64 product with the iterate x_j + xd_j
        auto block = [&](T m0, T m1, T m2, T m3, T m4, T m5, T m6, T m7, T m8, int j) {
          const Z4<T> zj = zget(j - v0);
          a0 += m0 * zj.v[0] + m1 * zj.v[1] + m2 * zj.v[2];
          a1 += m3 * zj.v[0] + m4 * zj.v[1] + m5 * zj.v[2];
          a2 += m6 * zj.v[0] + m7 * zj.v[1] + m8 * zj.v[2];
          if (refresh) {
            double x0 = x[3 * j], x1 = x[3 * j + 1], x2 = x[3 * j + 2];
            if (has_xd) {
              const int jv = kVec ? j - v0 : j;
              x0 += double(xdv[3 * jv]);
              x1 += double(xdv[3 * jv + 1]);
              x2 += double(xdv[3 * jv + 2]);
            }
            q0 += double(m0) * x0 + double(m1) * x1 + double(m2) * x2;
            q1 += double(m3) * x0 + double(m4) * x1 + double(m5) * x2;
            q2 += double(m6) * x0 + double(m7) * x1 + double(m8) * x2;
          }
        };
        if constexpr (sizeof(T) == 4) {
          // runs of U L blocks: the lane takes U consecutive blocks of each
          // run, one U-wide load per value plane; its slots past the row's
          // end weigh zero and name the row itself (loads past the end of a
          // buffer return zero)
          for (int kb = k0 + lane * U; kb < k1; kb += U * L) {
            float m[9][U];
#pragma unroll
            for (int c = 0; c < 9; ++c) load_run<U>(bf.v[c], uint32_t(kb) * 4u, m[c]);
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const bool in = kb + u < k1;
              const T wu = in ? T(1) : T(0);
              int j;
              if (kC16)
                j = i + int(short(__builtin_amdgcn_raw_buffer_load_b16(bf.col, uint32_t(kb + u) * 2u, 0, 0)));
              else
                j = buf_ld<int>(bf.col, uint32_t(kb + u) * 4u);
              block(m[0][u] * wu, m[1][u] * wu, m[2][u] * wu, m[3][u] * wu, m[4][u] * wu, m[5][u] * wu,
                    m[6][u] * wu, m[7][u] * wu, m[8][u] * wu, in ? j : i);
            }
          }
        } else {
          // FP64 values: a lane's blocks strided by L, clamped to the row's
          // last block (weighted zero) past its end
          for (int kb = k0 + lane; kb < k1; kb += U * L) {
            int kk[U], jc[U];
            T wv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const int kr = kb + u * L;
              kk[u] = kr < k1 ? kr : k1 - 1;
              wv[u] = kr < k1 ? T(1) : T(0);
              if (kC16)
                jc[u] = i + int(short(__builtin_amdgcn_raw_buffer_load_b16(bf.col, uint32_t(kk[u]) * 2u, 0, 0)));
              else
                jc[u] = buf_ld<int>(bf.col, uint32_t(kk[u]) * 4u);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const uint32_t o = uint32_t(kk[u]) * uint32_t(sizeof(T));
              block(buf_ld<T>(bf.v[0], o) * wv[u], buf_ld<T>(bf.v[1], o) * wv[u], buf_ld<T>(bf.v[2], o) * wv[u],
                    buf_ld<T>(bf.v[3], o) * wv[u], buf_ld<T>(bf.v[4], o) * wv[u], buf_ld<T>(bf.v[5], o) * wv[u],
                    buf_ld<T>(bf.v[6], o) * wv[u], buf_ld<T>(bf.v[7], o) * wv[u], buf_ld<T>(bf.v[8], o) * wv[u],
                    jc[u]);
            }
          }
        }
      }
      a0 = group_sum<L>(a0);
      a1 = group_sum<L>(a1);
      a2 = group_sum<L>(a2);
      if (refresh) {
        q0 = group_sum<L>(q0);
        q1 = group_sum<L>(q1);
        q2 = group_sum<L>(q2);
      }
      if (valid && lane == 0) {
        wl[3 * il] = a0;
        wl[3 * il + 1] = a1;
        wl[3 * il + 2] = a2;
        const Z4<T> zi = zget(il);
        dz[0] += A(zi.v[0]) * A(a0) + A(zi.v[1]) * A(a1) + A(zi.v[2]) * A(a2);
        if (refresh) {
          rref[3 * i] = double(b[3 * i]) - q0;
          rref[3 * i + 1] = double(b[3 * i + 1]) - q1;
          rref[3 * i + 2] = double(b[3 * i + 2]) - q2;
        }
      }
    }
    env_block_sum<1>(dz, red);  // also orders the w and rref writes before the update
    const double delta = double(dz[0]);
    double beta = 0.0, mu = delta;
    if (k >= 1) {
      beta = (gamma - (k >= 2 ? rho : 0.0)) / g_prev;
      mu = delta + 2.0 * beta * (rho - gamma) / a_prev + beta * beta * m_prev;
    }
    if (mu <= 0.0 && cfg.abort_if_not_spd) {  // pcg.h:152-169: not SPD, abort with iteration k + 1
      fold();
      if (threadIdx.x == 0) {
        EnvCG& E = env[e];
        E.status = kCgDiverged;
        E.iter = k + 1;
        E.latest = latest < 0.0 ? 0.0 : latest;
        E.start = start;
        E.rz = gamma;
        atomicSub(num_active, 1);
      }
      return;
    }
    const double alpha = gamma / mu;
    const T al = T(alpha), bt = T(beta);
    // update: p = z + beta p, s = w + beta s, the iterate, r, z = M^-1 r (LDS)
    A part[4] = {A(0), A(0), A(0), A(0)};  // r.r, r.z, z.z, r.z_old
    for (int il = threadIdx.x; il < nr; il += kEnvThreads) {
      const int i = v0 + il;
      const int iv = kVec ? il : i;  // row in rv, pv, sv
      T d0 = T(0), d1 = T(0), d2 = T(0);
      A x0 = A(0), x1 = A(0), x2 = A(0);
      if (has_xd) {
        d0 = xdv[3 * iv];
        d1 = xdv[3 * iv + 1];
        d2 = xdv[3 * iv + 2];
      }
      if (!has_xd || refresh) {
        x0 = x[3 * i];
        x1 = x[3 * i + 1];
        x2 = x[3 * i + 2];
      }
      T p0 = T(0), p1 = T(0), p2 = T(0), s0 = T(0), s1 = T(0), s2 = T(0);
      if (k > 0) {
        p0 = pv[3 * iv];
        p1 = pv[3 * iv + 1];
        p2 = pv[3 * iv + 2];
        s0 = sv[3 * iv];
        s1 = sv[3 * iv + 1];
        s2 = sv[3 * iv + 2];
      }
      double ro0, ro1, ro2;
      if (refresh) {
        ro0 = rref[3 * i];
        ro1 = rref[3 * i + 1];
        ro2 = rref[3 * i + 2];
      } else {
        ro0 = rv[3 * iv];
        ro1 = rv[3 * iv + 1];
        ro2 = rv[3 * iv + 2];
      }
      Z4<T> zo = zget(il);
      const T w0 = wl[3 * il], w1 = wl[3 * il + 1], w2 = wl[3 * il + 2];
      p0 = zo.v[0] + bt * p0;
      p1 = zo.v[1] + bt * p1;
      p2 = zo.v[2] + bt * p2;
      s0 = w0 + bt * s0;
      s1 = w1 + bt * s1;
      s2 = w2 + bt * s2;
      pv[3 * iv] = p0;
      pv[3 * iv + 1] = p1;
      pv[3 * iv + 2] = p2;
      sv[3 * iv] = s0;
      sv[3 * iv + 1] = s1;
      sv[3 * iv + 2] = s2;
      if (has_xd && !refresh) {
        xdv[3 * iv] = d0 + al * p0;
        xdv[3 * iv + 1] = d1 + al * p1;
        xdv[3 * iv + 2] = d2 + al * p2;
      } else {
        x[3 * i] = x0 + A(d0) + A(alpha) * A(p0);
        x[3 * i + 1] = x1 + A(d1) + A(alpha) * A(p1);
        x[3 * i + 2] = x2 + A(d2) + A(alpha) * A(p2);
        if (has_xd) {
          xdv[3 * iv] = T(0);
          xdv[3 * iv + 1] = T(0);
          xdv[3 * iv + 2] = T(0);
        }
      }
      T r0, r1, r2;
      if (refresh) {
        r0 = T(ro0 - alpha * double(s0));
        r1 = T(ro1 - alpha * double(s1));
        r2 = T(ro2 - alpha * double(s2));
      } else {
        r0 = T(ro0) - al * s0;
        r1 = T(ro1) - al * s1;
        r2 = T(ro2) - al * s2;
      }
      rv[3 * iv] = r0;
      rv[3 * iv + 1] = r1;
      rv[3 * iv + 2] = r2;
      T n0, n1, n2;
      apply_dinv(dinv, nV, i, cfg.prec, r0, r1, r2, n0, n1, n2);
      Z4<T> zn;
      zn.v[0] = n0;
      zn.v[1] = n1;
      zn.v[2] = n2;
      zn.v[3] = T(0);
      if (kVec) {
        z3[3 * il] = n0;
        z3[3 * il + 1] = n1;
        z3[3 * il + 2] = n2;
      } else {
        zl[il] = zn;
      }
      if (refresh)  // z_k replaced by M^-1 rref for the next mu recurrence
        apply_dinv(dinv, nV, i, cfg.prec, T(ro0), T(ro1), T(ro2), zo.v[0], zo.v[1], zo.v[2]);
      part[0] += A(r0) * A(r0) + A(r1) * A(r1) + A(r2) * A(r2);
      part[1] += A(r0) * A(n0) + A(r1) * A(n1) + A(r2) * A(n2);
      part[2] += A(n0) * A(n0) + A(n1) * A(n1) + A(n2) * A(n2);
      part[3] += A(r0) * A(zo.v[0]) + A(r1) * A(zo.v[1]) + A(r2) * A(zo.v[2]);
    }
    env_block_sum<4>(part, red + kEnvWaves);  // also orders the LDS z writes before the next matvec
#pragma unroll
    for (int c = 0; c < 4; ++c) R[c] = double(part[c]);
    g_prev = gamma;
    a_prev = alpha;
    m_prev = mu;
  }
}

template <class T, class A, int L, bool kC16, bool kVec>
__global__ void __launch_bounds__(kEnvThreads) k_cg_env(EnvQueue q, CgConfig cfg, const int32_t* __restrict__ row_ptr,
                                                        const void* __restrict__ col, const T* __restrict__ vals,
                                                        int nB, const T* __restrict__ dinv, int nV,
                                                        const T* __restrict__ b, const T* __restrict__ zin,
                                                        A* __restrict__ x, T* __restrict__ r, T* __restrict__ p,
                                                        T* __restrict__ s, double* __restrict__ rref, EnvCG* env,
                                                        const A* __restrict__ acc, int32_t* num_active) {
  extern __shared__ __align__(16) unsigned char env_lds[];
  Z4<T>* zl = kVec ? nullptr : reinterpret_cast<Z4<T>*>(env_lds);
  T* z3 = kVec ? reinterpret_cast<T*>(env_lds) : nullptr;
  auto plane = [&](int k) { return reinterpret_cast<T*>(env_lds + env_lds_plane_offset<T>(q.max_rows, k, kVec)); };
  T* wl = plane(0);
  T* rl = kVec ? plane(1) : nullptr;
  T* pl = kVec ? plane(2) : nullptr;
  T* sl = kVec ? plane(3) : nullptr;
  T* xl = kVec ? plane(4) : nullptr;
  A* red = reinterpret_cast<A*>(env_lds + env_lds_red_offset<T>(q.max_rows, kVec));
  __shared__ int s_next;
  const SpmvBufs bf = spmv_bufs<T>(col, uint32_t(nB) * (kC16 ? 2u : 4u), vals, nB, zin, nV);
  for (;;) {
    if (threadIdx.x == 0) s_next = atomicAdd(q.next, 1);
    __syncthreads();
    const int qi = s_next;
    __syncthreads();  // every thread has read s_next before thread 0 takes the next one
    if (qi >= q.n) return;
    const int e = q.order[qi];
    if (env[e].status != kCgActive) continue;
    const int v0 = q.v0[e], nr = q.v0[e + 1] - v0;
    if (nr > q.max_rows) {  // not reached: the launcher sized max_rows for every environment
      if (threadIdx.x == 0) {
        env[e].status = kCgDiverged;
        atomicSub(num_active, 1);
      }
      continue;
    }
    cg_env_solve<T, A, L, kC16, kVec>(e, v0, nr, cfg, row_ptr, bf, dinv, nV, b, zin, x, r, p, s, rref, env, acc,
                                      num_active, zl, z3, wl, rl, pl, sl, xl, red);
    __syncthreads();  // the next environment reuses the LDS
  }
}

// Grid-wide barrier of a cooperative launch (every workgroup resident):
// arrivals are counted on *bar, zeroed before the launch, and barrier j of
// the launch waits for j * gridDim.x of them.  The wait is bounded: past
// kBarrierSpins polls, or once another workgroup gave up, it sets *abort and
// returns false, and the caller leaves the kernel (the host then reports the
// solve as failed).  The fences before arriving and after leaving make every
// workgroup's writes of the phase visible to every other one.
constexpr uint32_t kBarrierSpins = 4000000u;
__device__ __forceinline__ bool grid_barrier(uint32_t* bar, uint32_t target, int32_t* abort) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    atomicAdd(bar, 1u);
    int ok = 1;
    for (uint32_t spins = 0; __hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target; ++spins) {
      if (spins > kBarrierSpins || __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        atomicExch(abort, 1);
        ok = 0;
        break;
      }
    }
    __threadfence();
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// A whole CG burst, iterations [k0, k0 + n), in one cooperative launch of
// a few workgroups: the matvec and update phases of every iteration stride
// over the work lists as k_cg_spmv and k_cg_update do, separated by grid
// barriers instead of kernel boundaries.  For the tail of a batched solve,
// when a few environments are left and a launch costs more than its work.
template <class T, class A, int L, bool kC16>
__global__ void __launch_bounds__(kSpmvThreads) k_cg_tail(Tiles ta, Tiles tv, WorkList wa, WorkList wv, CgConfig cfg,
                                                          int k0, int n, const int32_t* __restrict__ row_ptr,
                                                          const void* __restrict__ col, const T* __restrict__ vals,
                                                          int nB, const T* __restrict__ dinv, int nV, A* x,
                                                          const T* __restrict__ b, T* w, double* rref, T* r, T* z,
                                                          T* p, T* sv, EnvCG* env, A* acc, int32_t* num_active,
                                                          uint32_t* bar, int32_t* abort) {
  static_assert(kSpmvThreads == kVecThreads, "one workgroup size for both phases");
  const SpmvBufs bf = spmv_bufs<T>(col, uint32_t(nB) * (kC16 ? 2u : 4u), vals, nB, z, cfg.n_rows);
  uint32_t j = 0;
  for (int k = k0; k < k0 + n; ++k) {
    const bool refresh = cg_refresh(cfg, k);
    for_each_tile(wa, 0, [&](int tile) {
      if (refresh)
        cg_spmv_tile<T, A, true, L, kC16>(tile, ta, cfg, k, row_ptr, bf, z, x, b, w, rref, env, acc, num_active);
      else
        cg_spmv_tile<T, A, false, L, kC16>(tile, ta, cfg, k, row_ptr, bf, z, x, b, w, rref, env, acc, num_active);
    });
    if (!grid_barrier(bar, ++j * gridDim.x, abort)) return;
    for_each_tile(wv, 0, [&](int tile) {
      if (refresh)
        cg_update_tile<T, A, true>(tile, tv, cfg, k, dinv, nV, w, rref, x, r, z, p, sv, env, acc, num_active);
      else
        cg_update_tile<T, A, false>(tile, tv, cfg, k, dinv, nV, w, rref, x, r, z, p, sv, env, acc, num_active);
    });
    if (!grid_barrier(bar, ++j * gridDim.x, abort)) return;
  }
}

// Work lists of the matvec and vector tiles whose environment is still
// active, for the bursts' persistent CG kernels.  Blocks [0, nba) scan the
// matvec tiles, the rest the vector tiles; each reserves a contiguous run of
// its list with one atomic (list order follows block arrival).  counts:
// {matvec list, matvec list (other parity), vector list, vector list (other
// parity)}; this launch fills slot `par` and clears slot par ^ 1 for the next
// burst.
__global__ void __launch_bounds__(256) k_compact_tiles(Tiles ta, Tiles tv, int a0, int na, int v0, int nv,
                                                      const EnvCG* __restrict__ env, int32_t* __restrict__ la,
                                                      int32_t* __restrict__ lv, int32_t* counts, int par,
                                                      const int32_t* num_active, int32_t* publish,
                                                      uint32_t* bar) {
  // tiles [a0, a0 + na) of ta and [v0, v0 + nv) of tv: one environment group
  // (or all of them)
  const int nba = (na + 255) / 256;
  const bool isA = int(blockIdx.x) < nba;
  const Tiles& t = isA ? ta : tv;
  const int n = isA ? na : nv;
  const int i = (isA ? int(blockIdx.x) : int(blockIdx.x) - nba) * 256 + int(threadIdx.x);
  const int tile = (isA ? a0 : v0) + i;
  int32_t* list = isA ? la : lv;
  int32_t* cnt = counts + (isA ? 0 : 2) + par;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    counts[par ^ 1] = 0;
    counts[2 + (par ^ 1)] = 0;
    if (publish) *publish = *num_active;  // pinned host memory, read after this kernel's event
    if (bar) *bar = 0u;                     // the barrier counter of a k_cg_tail burst
  }
  const bool keep = i < n && env[t.env[tile]].status == kCgActive;
  const unsigned long long m = __ballot(keep);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ int ws[4];
  __shared__ int base;
  if (lane == 0) ws[wv] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int tot = ws[0] + ws[1] + ws[2] + ws[3];
    base = tot ? atomicAdd(cnt, tot) : 0;
  }
  __syncthreads();
  int off = base;
  for (int q = 0; q < wv; ++q) off += ws[q];
  if (keep) list[off + __popcll(m & ((1ull << lane) - 1ull))] = tile;
}

// The blocks k0..k1-1 of one row as lane `lane` of the row's L lanes takes
// them: f(a, j) with the block's 9 values in FP64 and its column j.  FP32
// values: runs of U L blocks (U = kSpmvChunk / L), the lane's U consecutive
// blocks of a run read with one U-wide load per value plane and one for the
// columns (bf: descriptors of the planes and the 32-bit columns); FP64
// values: one load per value, blocks strided by L.
template <class T, int L, class F>
__device__ __forceinline__ void row_blocks(const SpmvBufs& bf, const int32_t* __restrict__ col,
                                           const T* __restrict__ vals, int nB, int k0, int k1, int lane, F&& f) {
  if constexpr (sizeof(T) == 4) {
    constexpr int U = kSpmvChunk / L > 0 ? kSpmvChunk / L : 1;
    for (int kb = k0 + lane * U; kb < k1; kb += U * L) {
      float m[9][U], cj[U];
#pragma unroll
      for (int c = 0; c < 9; ++c) load_run<U>(bf.v[c], uint32_t(kb) * 4u, m[c]);
      load_run<U>(bf.col, uint32_t(kb) * 4u, cj);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (kb + u < k1) {
          double a[9];
#pragma unroll
          for (int c = 0; c < 9; ++c) a[c] = double(m[c][u]);
          f(a, __float_as_int(cj[u]));
        }
    }
  } else {
    for (int kk = k0 + lane; kk < k1; kk += L) {
      double a[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) a[c] = double(vals[(size_t)c * nB + kk]);
      f(a, col[kk]);
    }
  }
}

// FP64 residual b - A x of the environments with mask[e]: stored in r64
// when given, and |b - A x|^2, |b|^2, b.Ax, |Ax|^2 added into set 0 of acc
// when given (zeroed by the caller).
template <class T, class X, int L>
__global__ void __launch_bounds__(kSpmvThreads) k_residual64(Tiles ta, int nEnv, const int32_t* __restrict__ mask,
                                                             const int32_t* __restrict__ row_ptr,
                                                             const int32_t* __restrict__ col,
                                                             const T* __restrict__ vals, int nB,
                                                             const X* __restrict__ x, const T* __restrict__ b,
                                                             double* __restrict__ r64, double* acc) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int e = ta.env[tile];
  if (!mask[e]) return;
  const int i = ta.row0[tile] + (threadIdx.x / L);
  const int lane = threadIdx.x % L;
  const bool valid = i < ta.row1[tile];
  double q0 = 0.0, q1 = 0.0, q2 = 0.0;
  if (valid) {
    const SpmvBufs bf = spmv_bufs<T>(col, uint32_t(nB) * 4u, vals, nB, vals, 0);
    row_blocks<T, L>(bf, col, vals, nB, row_ptr[i], row_ptr[i + 1], lane, [&](const double (&a)[9], int jc) {
      const double x0 = x[3 * jc], x1 = x[3 * jc + 1], x2 = x[3 * jc + 2];
      q0 += a[0] * x0 + a[1] * x1 + a[2] * x2;
      q1 += a[3] * x0 + a[4] * x1 + a[5] * x2;
      q2 += a[6] * x0 + a[7] * x1 + a[8] * x2;
    });
  }
  q0 = group_sum<L>(q0);
  q1 = group_sum<L>(q1);
  q2 = group_sum<L>(q2);
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  if (valid && lane == 0) {
    const double b0 = b[3 * i], b1 = b[3 * i + 1], b2 = b[3 * i + 2];
    const double d0 = b0 - q0, d1 = b1 - q1, d2 = b2 - q2;
    if (r64) {
      r64[3 * i] = d0;
      r64[3 * i + 1] = d1;
      r64[3 * i + 2] = d2;
    }
    part[0] = d0 * d0 + d1 * d1 + d2 * d2;
    part[1] = b0 * b0 + b1 * b1 + b2 * b2;
    part[2] = b0 * q0 + b1 * q1 + b2 * q2;  // b.Ax and |Ax|^2: the best multiple of x
    part[3] = q0 * q0 + q1 * q1 + q2 * q2;
  }
  if (acc) acc_add<double, 4>(part, acc_ptr(acc, 0, nEnv, e), tile - ta.first[e]);
}

// The FP64 residuals b - A u, b - A v, b - A w of three initial guesses in
// one pass over the matrix, for the environments with mask[e]: into ru, rv,
// rw, and the ten sums |b|^2, b.Au, b.Av, b.Aw, |Au|^2, |Av|^2, |Aw|^2,
// Au.Av, Au.Aw, Av.Aw into set 0 of acc (zeroed by the caller): the normal
// equations of the combination with the least residual.
template <class T, class X, int L>
__global__ void __launch_bounds__(kSpmvThreads) k_residual64_multi(Tiles ta, int nEnv, const int32_t* __restrict__ mask,
                                                                   const int32_t* __restrict__ row_ptr,
                                                                   const int32_t* __restrict__ col,
                                                                   const T* __restrict__ vals, int nB,
                                                                   const X* __restrict__ u, const T* __restrict__ v,
                                                                   const T* __restrict__ w, const T* __restrict__ b,
                                                                   double* __restrict__ ru, double* __restrict__ rv,
                                                                   double* __restrict__ rw, double* acc) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int e = ta.env[tile];
  if (!mask[e]) return;
  const int i = ta.row0[tile] + (threadIdx.x / L);
  const int lane = threadIdx.x % L;
  const bool valid = i < ta.row1[tile];
  double p[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};  // A u, A v, A w rows
  if (valid) {
    const SpmvBufs bf = spmv_bufs<T>(col, uint32_t(nB) * 4u, vals, nB, vals, 0);
    row_blocks<T, L>(bf, col, vals, nB, row_ptr[i], row_ptr[i + 1], lane, [&](const double (&a)[9], int jc) {
      const double g[3][3] = {{double(u[3 * jc]), double(u[3 * jc + 1]), double(u[3 * jc + 2])},
                              {double(v[3 * jc]), double(v[3 * jc + 1]), double(v[3 * jc + 2])},
                              {double(w[3 * jc]), double(w[3 * jc + 1]), double(w[3 * jc + 2])}};
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int r = 0; r < 3; ++r) p[q][r] += a[3 * r] * g[q][0] + a[3 * r + 1] * g[q][1] + a[3 * r + 2] * g[q][2];
    });
  }
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int r = 0; r < 3; ++r) p[q][r] = group_sum<L>(p[q][r]);
  double part[10] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (valid && lane == 0) {
    const double bb[3] = {double(b[3 * i]), double(b[3 * i + 1]), double(b[3 * i + 2])};
    double* out[3] = {ru, rv, rw};
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int r = 0; r < 3; ++r) out[q][3 * i + r] = bb[r] - p[q][r];
    auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    part[0] = dot(bb, bb);
    part[1] = dot(bb, p[0]);
    part[2] = dot(bb, p[1]);
    part[3] = dot(bb, p[2]);
    part[4] = dot(p[0], p[0]);
    part[5] = dot(p[1], p[1]);
    part[6] = dot(p[2], p[2]);
    part[7] = dot(p[0], p[1]);
    part[8] = dot(p[0], p[2]);
    part[9] = dot(p[1], p[2]);
  }
  acc_add<double, 10>(part, acc_ptr(acc, 0, nEnv, e), tile - ta.first[e]);
}

// ------------------------------------------------------------------ vectors

template <class T>
__global__ void k_copy(int64_t n, const T* __restrict__ s, T* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

template <class T>
__global__ void k_apply_pins(int64_t n, const uint8_t* __restrict__ pin, const T* __restrict__ val, T* __restrict__ u) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && pin[i]) u[i] = val[i];
}

template <class T>
__global__ void k_ls_update(int nV, const int32_t* __restrict__ row_env, const int32_t* __restrict__ mask,
                            const double* __restrict__ scale, const T* __restrict__ sol0, const T* __restrict__ dx,
                            T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  const int e = row_env[i];
  if (!mask[e]) return;
  const T s = T(scale[e]);
#pragma unroll
  for (int d = 0; d < 3; ++d) x[3 * i + d] = sol0[3 * i + d] - s * dx[3 * i + d];
}

template <class T>
__global__ void k_post_step(int nV, const int32_t* __restrict__ row_env, const int32_t* __restrict__ reset, T idt,
                            const T* __restrict__ u0, T* __restrict__ u, T* __restrict__ v, T* __restrict__ hist,
                            int n_hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  const bool rs = reset[row_env[i]] != 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const T uu = u[3 * i + d];
    v[3 * i + d] = rs ? T(0) : (uu - u0[3 * i + d]) * idt;
    if (rs) {
      u[3 * i + d] = T(0);
      for (int h = 0; h < n_hist; ++h) hist[(size_t)h * 3 * nV + 3 * i + d] = T(0);
    }
  }
}

template <class T>
__global__ void k_env_copy(int nV, const int32_t* __restrict__ row_env, const int32_t* __restrict__ mask,
                           const T* __restrict__ src, T* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV || !mask[row_env[i]]) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) dst[3 * i + d] = src[3 * i + d];
}

template <class T>
__global__ void __launch_bounds__(kVecThreads) k_env_norms(Tiles tv, int nEnv, const T* __restrict__ a,
                                                           const T* __restrict__ b, double* acc) {
  const int tile = blockIdx.x;
  const int e = tv.env[tile];
  const int i = tv.row0[tile] + threadIdx.x;
  double part[2] = {0.0, 0.0};
  if (i < tv.row1[tile]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double av = a[3 * i + d], bv = b[3 * i + d];
      part[0] += av * av;
      part[1] += bv * bv;
    }
  }
  acc_add<double, 2>(part, acc_ptr(acc, 0, nEnv, e), tile - tv.first[e]);
}

// Per-environment a.b (FP64) added into set 0 of acc (zeroed by the caller).
template <class T>
__global__ void __launch_bounds__(kVecThreads) k_env_dot(Tiles tv, int nEnv, const T* __restrict__ a,
                                                         const T* __restrict__ b, double* acc) {
  const int tile = blockIdx.x;
  const int e = tv.env[tile];
  const int i = tv.row0[tile] + threadIdx.x;
  double part[1] = {0.0};
  if (i < tv.row1[tile])
#pragma unroll
    for (int d = 0; d < 3; ++d) part[0] += double(a[3 * i + d]) * double(b[3 * i + d]);
  acc_add<double, 1>(part, acc_ptr(acc, 0, nEnv, e), tile - tv.first[e]);
}

// v of every lane added into environment env's accumulator of set 0: with
// one atomic per wave when every lane of the wave holds the same environment
// (lanes past the end hold v = 0 and lane 0's environment), else per lane.
// Valid lanes are a prefix of the wave.
__device__ __forceinline__ void env_add_wave(double* acc, int nEnv, int env, bool valid, double v) {
  const int lane = int(threadIdx.x) & 63;
  const int e0 = __shfl(env, 0, 64);
  if (!valid) {
    env = e0;
    v = 0.0;
  }
  if (e0 < 0) return;  // no valid lane in this wave
  const int slot = (int(blockIdx.x) * int(blockDim.x >> 6) + int(threadIdx.x >> 6)) & (kSlots - 1);
  if (__all(env == e0)) {
    v = wave_sum(v);
    if (lane == 0) atomicAdd(acc_ptr(acc, 0, nEnv, e0) + slot * kSlotStride, v);
  } else if (v != 0.0) {
    atomicAdd(acc_ptr(acc, 0, nEnv, env) + slot * kSlotStride, v);
  }
}

// Objective of a backward-Euler step (the merit of the objective-based line
// searches, newton_solver.cpp:664-686; line_search.h:125-290, 338-390) of the
// environments with env_mask[e], summed per environment in FP64 into set 0 of
// acc (zeroed by the caller).  Per tet: the stress energy vol psi(F)
// (fem_stress.h), gravity -rho vol g.(X_c + u_c) at the centroid (the Q1
// rule of fem_gravity.h:93-111), inertia 1/(2 dt^2) sum_ab m_ab
// (u - x~)_a.(u - x~)_b with x~ = u0 + dt v0, and mass damping alpha/(2 dt)
// sum_ab m_ab (u - u0)_a.(u - u0)_b, with the consistent mass
// m_ab = rho vol / 20 (1 + delta_ab) (fem_inertia.h:55-130, exact for the Q4
// rule).  Per boundary-triangle sample: w times the contact energy density of
// contact_sample (contact_utils.h:560, :935; fem_traction.h:130-135).
// Blocks [0, n_tet_blocks) take tets, the rest triangles.
template <class T>
__global__ void __launch_bounds__(256) k_energy(MeshView m, FieldPtrs<T> f, const T* __restrict__ u,
                                                const T* __restrict__ u0, const T* __restrict__ v0, T dt, double gx,
                                                double gy, double gz, int n_tet_blocks, double* acc) {
  if (int(blockIdx.x) < n_tet_blocks) {
    const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
    const bool in = t < m.nT;
    const int env = in ? m.tet_env[t] : -1;
    const bool valid = in && m.env_mask[env];
    double e = 0.0;
    if (valid) {
      const int nT = m.nT;
      const int4 tv = reinterpret_cast<const int4*>(m.tets)[t];
      const int nd[4] = {tv.x, tv.y, tv.z, tv.w};
      const T* g = f.geom;
      V3<T> dN[4];
      dN[1] = mk3(g[0 * nT + t], g[1 * nT + t], g[2 * nT + t]);
      dN[2] = mk3(g[3 * nT + t], g[4 * nT + t], g[5 * nT + t]);
      dN[3] = mk3(g[6 * nT + t], g[7 * nT + t], g[8 * nT + t]);
      dN[0] = scl3(T(-1), add3(add3(dN[1], dN[2]), dN[3]));
      const double vol = double(g[kGVol * nT + t]), mu = double(g[kGMu * nT + t]);
      const double lam = double(g[kGLambda * nT + t]), mass = double(g[kGMass * nT + t]);
      const double alpha = double(g[kGAlpha * nT + t]);
      M33<T> F;
#pragma unroll
      for (int k = 0; k < 9; ++k) F.m[k] = (k % 4 == 0) ? T(1) : T(0);
      double s1 = 0.0, sd1 = 0.0, sum[3] = {0.0, 0.0, 0.0}, sumd[3] = {0.0, 0.0, 0.0}, xc[3] = {0.0, 0.0, 0.0};
      const double dtd = double(dt);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const T d[3] = {dN[a].x, dN[a].y, dN[a].z};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const T ua = u[3 * nd[a] + r];
#pragma unroll
          for (int c = 0; c < 3; ++c) F.m[r * 3 + c] += ua * d[c];
          const double us = double(u0[3 * nd[a] + r]);
          const double del = double(ua) - (us + dtd * double(v0[3 * nd[a] + r]));  // u - x~
          const double dd = double(ua) - us;                                          // u - u0
          s1 += del * del;
          sd1 += dd * dd;
          sum[r] += del;
          sumd[r] += dd;
          xc[r] += 0.25 * (double(f.X[3 * nd[a] + r]) + double(ua));
        }
      }
      const double m20 = mass / 20.0;
      e = vol * smith_energy<T>(F, mu, lam);
      const T beta = g[kGBeta * nT + t];
      if (beta > T(0)) {  // stiffness damping energy vol dE:S/2
        T ue[12], us[12];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            ue[3 * a + r] = u[3 * nd[a] + r];
            us[3 * a + r] = u0[3 * nd[a] + r];
          }
        Viscous<T> vis;
        double psi = 0.0;
        viscous_setup<T>(dN, ue, us, beta / dt, T(mu), T(lam), false, false, vis, &psi);
        e += vol * psi;
      }
      e -= mass * (gx * xc[0] + gy * xc[1] + gz * xc[2]);
      e += 0.5 / (dtd * dtd) * m20 * (s1 + sum[0] * sum[0] + sum[1] * sum[1] + sum[2] * sum[2]);
      if (alpha > 0.0)
        e += 0.5 * alpha / dtd * m20 * (sd1 + sumd[0] * sumd[0] + sumd[1] * sumd[1] + sumd[2] * sumd[2]);
    }
    env_add_wave(acc, m.nEnv, env, valid, e);
    return;
  }
  const int tr = int((blockIdx.x - unsigned(n_tet_blocks)) * blockDim.x + threadIdx.x);
  const bool in = tr < m.nF;
  const int env = in ? m.tri_env[tr] : -1;
  const bool valid = in && m.env_mask[env];
  double e = 0.0;
  if (valid) {
    const int c0 = m.coll_first[env], c1 = m.coll_first[env + 1];
    if (c1 > c0) {
      const int nd[3] = {m.tris[3 * tr], m.tris[3 * tr + 1], m.tris[3 * tr + 2]};
      V3<T> xs[3], Xr[3], dxs[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        Xr[i] = mk3(f.X[3 * nd[i]], f.X[3 * nd[i] + 1], f.X[3 * nd[i] + 2]);
        const V3<T> ui = mk3(u[3 * nd[i]], u[3 * nd[i] + 1], u[3 * nd[i] + 2]);
        xs[i] = add3(Xr[i], ui);
        dxs[i] = sub3(ui, mk3(u0[3 * nd[i]], u0[3 * nd[i] + 1], u0[3 * nd[i] + 2]));
      }
      const V3<T> n0 = cross3(sub3(Xr[1], Xr[0]), sub3(Xr[2], Xr[0]));
      const double w = 0.5 * double(dsqrt(dot3(n0, n0))) / 3.0;
      V3<T> fnv = cross3(sub3(xs[1], xs[0]), sub3(xs[2], xs[0]));
      fnv = scl3(T(1) / dsqrt(dot3(fnv, fnv)), fnv);
      const double* cp = f.contact + kContactParams * env;
      for (int c = c0; c < c1; ++c) {
        const double* cd = m.coll + kCollStride * c;
        for (int q = 0; q < 3; ++q) {
          T N[3];
          tri_q3<T>(q, N);
          const V3<T> x = add3(add3(scl3(N[0], xs[0]), scl3(N[1], xs[1])), scl3(N[2], xs[2]));
          const V3<T> dxq = add3(add3(scl3(N[0], dxs[0]), scl3(N[1], dxs[1])), scl3(N[2], dxs[2]));
          ContactSample<T> cs;
          if (contact_sample<T>(x, dxq, fnv, cd, cp, dt, true, cs)) e += w * cs.e;
        }
      }
    }
  }
  env_add_wave(acc, m.nEnv, env, valid, e);
}

// Order-preserving map of a double onto int64 (for atomicMin on distances).
__device__ __forceinline__ long long ordered_bits(double x) {
  const long long b = __double_as_longlong(x);
  return b >= 0 ? b : (b ^ 0x7fffffffffffffffLL);
}

// Contact wrench of every collider at displacement u with stage start u0,
// over the boundary-triangle samples of the environments with env_mask[e]:
// out[kWrench c + 0..2] += the contact force on the samples from collider c
// (penalty and dissipation, weighted), [3..5] += the torque of the reaction
// on the collider about its position p, and dmin[c] the least signed
// distance of the samples in contact (those within the detection threshold;
// ordered_bits encoding).  One thread per triangle.
template <class T>
__global__ void __launch_bounds__(256) k_contact_wrench(MeshView m, FieldPtrs<T> f, const T* __restrict__ u,
                                                        const T* __restrict__ u0, T dt, double* out,
                                                        long long* dmin) {
  const int tr = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (tr >= m.nF) return;
  const int env = m.tri_env[tr];
  if (!m.env_mask[env]) return;
  const int c0 = m.coll_first[env], c1 = m.coll_first[env + 1];
  if (c1 <= c0) return;
  const int nd[3] = {m.tris[3 * tr], m.tris[3 * tr + 1], m.tris[3 * tr + 2]};
  V3<T> xs[3], Xr[3], dxs[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Xr[i] = mk3(f.X[3 * nd[i]], f.X[3 * nd[i] + 1], f.X[3 * nd[i] + 2]);
    const V3<T> ui = mk3(u[3 * nd[i]], u[3 * nd[i] + 1], u[3 * nd[i] + 2]);
    xs[i] = add3(Xr[i], ui);
    dxs[i] = sub3(ui, mk3(u0[3 * nd[i]], u0[3 * nd[i] + 1], u0[3 * nd[i] + 2]));
  }
  const V3<T> n0 = cross3(sub3(Xr[1], Xr[0]), sub3(Xr[2], Xr[0]));
  const T w = T(0.5) * dsqrt(dot3(n0, n0)) / T(3);
  V3<T> fnv = cross3(sub3(xs[1], xs[0]), sub3(xs[2], xs[0]));
  fnv = scl3(T(1) / dsqrt(dot3(fnv, fnv)), fnv);
  const double* cp = f.contact + kContactParams * env;
  for (int c = c0; c < c1; ++c) {
    const double* cd = m.coll + kCollStride * c;
    const V3<T> P = mk3(T(cd[1]), T(cd[2]), T(cd[3]));
    double F[3] = {0, 0, 0}, Tq[3] = {0, 0, 0}, dm = 0;
    bool any = false;
    for (int q = 0; q < 3; ++q) {
      T N[3];
      tri_q3<T>(q, N);
      const V3<T> x = add3(add3(scl3(N[0], xs[0]), scl3(N[1], xs[1])), scl3(N[2], xs[2]));
      const V3<T> dxq = add3(add3(scl3(N[0], dxs[0]), scl3(N[1], dxs[1])), scl3(N[2], dxs[2]));
      ContactSample<T> cs;
      if (!contact_sample<T>(x, dxq, fnv, cd, cp, dt, false, cs)) continue;
      const V3<T> Fw = scl3(w, cs.F);
      const V3<T> tq = cross3(sub3(x, P), scl3(T(-1), Fw));
      F[0] += double(Fw.x);
      F[1] += double(Fw.y);
      F[2] += double(Fw.z);
      Tq[0] += double(tq.x);
      Tq[1] += double(tq.y);
      Tq[2] += double(tq.z);
      dm = any ? fmin(dm, double(cs.dist)) : double(cs.dist);
      any = true;
    }
    if (!any) continue;
    double* o = out + kWrench * c;
    for (int k = 0; k < 3; ++k) {
      atomicAdd(o + k, F[k]);
      atomicAdd(o + 3 + k, Tq[k]);
    }
    atomicMin(dmin + c, ordered_bits(dm));
  }
}

// Sum the kSlots lines of set 0 of each environment: out[e * nv + k].
template <class A>
__global__ void k_acc_finalize(int nEnv, int nv, const A* __restrict__ acc, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nEnv) return;
  const A* src = acc + (size_t)e * kSlots * kSlotStride;
  for (int k = 0; k < nv; ++k) {
    double s = 0.0;
    for (int sl = 0; sl < kSlots; ++sl) s += double(src[sl * kSlotStride + k]);
    out[(size_t)e * nv + k] = s;
  }
}

template <class A>
__global__ void k_fill(int64_t n, A v, A* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = v;
}

// ------------------------------------------------------------------ launchers

static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Plain launch, or an event-bracketed one (hipExtLaunchKernelGGL records the
// events at the kernel's own dispatch) when ev is given.
#define SDX_LAUNCH(ev, kern, grid, block, s, ...)                                                        \
  do {                                                                                                 \
    if (ev)                                                                                            \
      hipExtLaunchKernelGGL(kern, grid, block, 0, s, (ev)->start, (ev)->stop, 0, __VA_ARGS__);         \
    else                                                                                               \
      hipLaunchKernelGGL(kern, grid, block, 0, s, __VA_ARGS__);                                        \
  } while (0)

// One wave per patch, block / 64 patches per workgroup (fewer where their
// LDS would pass 64 KB).
template <class T>
void launch_element(hipStream_t s, const MeshView& m, const AsmRange& r, const FieldPtrs<T>& f, const T* u,
                    const T* u0, const T* v0, double dt, const double grav[3], int wantH, int project, T* Rv, T* Dv,
                    T* Ev, int block, const KernelEvents* ev, bool damp, const int32_t* plist) {
  if (r.p1 <= r.p0) return;
  int wpb = block / 64;
  while (wpb > 1 && size_t(wpb) * patch_lds_bytes<T>() > size_t(65536)) --wpb;
  if (wpb < 1) wpb = 1;
  const size_t lds = size_t(wpb) * patch_lds_bytes<T>();
  const dim3 grid(blocks_for(r.p1 - r.p0, wpb)), blk(unsigned(64 * wpb));
  auto go = [&](auto kern) {
    if (ev)
      hipExtLaunchKernelGGL(kern, grid, blk, lds, s, ev->start, ev->stop, 0, m, r, f, u, u0, v0, T(dt), T(grav[0]),
                            T(grav[1]), T(grav[2]), wantH, project, Rv, Dv, Ev, plist);
    else
      hipLaunchKernelGGL(kern, grid, blk, lds, s, m, r, f, u, u0, v0, T(dt), T(grav[0]), T(grav[1]), T(grav[2]),
                         wantH, project, Rv, Dv, Ev, plist);
  };
  if (damp)
    go(k_element<T, true>);
  else
    go(k_element<T, false>);
}

template <class T>
void launch_contact(hipStream_t s, const MeshView& m, const AsmRange& r, const FieldPtrs<T>& f, const T* u,
                    const T* u0, T dt, int wantH, T* Rv, T* Dv, T* Ev) {
  if (r.f1 <= r.f0) return;
  hipLaunchKernelGGL(k_contact<T>, dim3(blocks_for(r.f1 - r.f0, 256)), dim3(256), 0, s, m, r, f, u, u0, dt, wantH,
                     Rv, Dv, Ev);
}

template <class T>
void launch_gather_bsr(hipStream_t s, const MeshView& m, const AsmRange& r, const T* Dv, const T* Ev, int prec,
                       T* vals, T* dinv, const KernelEvents* ev) {
  if (r.s1 <= r.s0) return;
  SDX_LAUNCH(ev, k_gather_bsr<T>, dim3(blocks_for(r.s1 - r.s0, kGatherThreads)), dim3(kGatherThreads), s, m, r, Dv,
             Ev, prec, vals, dinv);
}

template <class T, class A>
void launch_assemble_rows(hipStream_t s, const MeshView& m, const RowBlockView& rb, const FieldPtrs<T>& f,
                          const T* u, const T* u0, const T* v0, double dt, const double grav[3], int project,
                          int prec, T* vals, T* dinv, T* res, A* acc, bool damp, const KernelEvents* ev) {
  if (rb.n <= 0) return;
  const size_t lds = rows_lds_bytes_host<T>(rb.max_slots);
  if (lds > size_t(65536)) throw std::invalid_argument("row blocks too large for the fused assembly's LDS");
  auto go = [&](auto kern) {
    if (ev)
      hipExtLaunchKernelGGL(kern, dim3(unsigned(rb.n)), dim3(kRowBlockThreads), lds, s, ev->start, ev->stop, 0, m,
                            rb, f, u, u0, v0, T(dt), T(grav[0]), T(grav[1]), T(grav[2]), project, prec, vals, dinv,
                            res, acc);
    else
      hipLaunchKernelGGL(kern, dim3(unsigned(rb.n)), dim3(kRowBlockThreads), lds, s, m, rb, f, u, u0, v0, T(dt),
                         T(grav[0]), T(grav[1]), T(grav[2]), project, prec, vals, dinv, res, acc);
  };
  if (damp)
    go(k_assemble_rows<T, A, true>);
  else
    go(k_assemble_rows<T, A, false>);
}

template <class T, class A>
void launch_gather_res(hipStream_t s, const MeshView& m, const AsmRange& r, const Tiles& tv, const T* Rv, T* res,
                       A* acc) {
  if (r.tile1 <= r.tile0) return;
  hipLaunchKernelGGL((k_gather_res<T, A>), dim3(unsigned(r.tile1 - r.tile0)), dim3(kVecThreads), 0, s, m, r, tv, Rv,
                     res, acc);
}

template <class T>
void launch_contact_wrench(hipStream_t s, const MeshView& m, const FieldPtrs<T>& f, const T* u, const T* u0, double dt,
                           double* out, long long* dmin) {
  if (m.nF == 0) return;
  hipLaunchKernelGGL(k_contact_wrench<T>, dim3(blocks_for(m.nF, 256)), dim3(256), 0, s, m, f, u, u0, T(dt), out,
                     dmin);
}

template <class A>
void launch_fill(hipStream_t s, int64_t n, A v, A* d) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill<A>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, v, d);
}

template <class A>
void launch_acc_finalize(hipStream_t s, int nEnv, int nv, const A* acc, double* out) {
  if (nEnv == 0) return;
  hipLaunchKernelGGL(k_acc_finalize<A>, dim3(blocks_for(nEnv, 64)), dim3(64), 0, s, nEnv, nv, acc, out);
}

template <class A>
void launch_cg_reset(hipStream_t s, int nEnv, const int32_t* solve, EnvCG* env, A* acc) {
  if (nEnv == 0) return;
  hipLaunchKernelGGL(k_cg_reset<A>, dim3(nEnv), dim3(256), 0, s, nEnv, solve, env, acc);
}

template <class T, class A>
void launch_cg_init(hipStream_t s, const Tiles& tv, const CgConfig& cfg, const T* b, const int32_t* warm,
                    const double* r64, const double* alpha, const T* v, const double* rv, const double* beta,
                    const T* w, const double* rw, const double* gamma, const T* dinv, int32_t nV, A* x, T* r, T* z,
                    EnvCG* env, A* acc) {
  if (tv.n == 0) return;
  hipLaunchKernelGGL((k_cg_init<T, A>), dim3(tv.n), dim3(kVecThreads), 0, s, tv, cfg, b, warm, r64, alpha, v, rv,
                     beta, w, rw, gamma, dinv, nV, x, r, z, env, acc);
}

template <class T, class A>
void launch_cg_spmv(hipStream_t s, const Tiles& ta, const WorkList& wl, int grid, const CgConfig& cfg, int k,
                    int lanes, const int32_t* row_ptr, const int32_t* col, const T* vals, int32_t nB, const T* z,
                    const A* x, const T* b, T* w, double* rref, EnvCG* env, A* acc, int32_t* num_active,
                    const KernelEvents* ev, const int16_t* col16) {
  if (ta.n == 0) return;
  if (uint64_t(nB) * sizeof(T) >= (uint64_t(1) << 32) || uint64_t(cfg.n_rows) * 4 * sizeof(T) >= (uint64_t(1) << 32))
    throw std::invalid_argument("matrix too large for 32-bit buffer offsets of the CG matvec");
  const int g = wl.list ? grid : ta.n;
  const bool refresh = cg_refresh(cfg, k);
  const void* cp = col16 ? static_cast<const void*>(col16) : static_cast<const void*>(col);
#define SDX_SPMV(R, L)                                                                                                \
  do {                                                                                                                \
    if (col16)                                                                                                        \
      SDX_LAUNCH(ev, (k_cg_spmv<T, A, R, L, true>), dim3(g), dim3(kSpmvThreads), s, ta, wl, cfg, k, row_ptr, cp, vals, \
                 nB, z, x, b, w, rref, env, acc, num_active);                                                         \
    else                                                                                                              \
      SDX_LAUNCH(ev, (k_cg_spmv<T, A, R, L, false>), dim3(g), dim3(kSpmvThreads), s, ta, wl, cfg, k, row_ptr, cp,      \
                 vals, nB, z, x, b, w, rref, env, acc, num_active);                                                   \
  } while (0)
  switch (lanes) {
    case 4:
      if (refresh) SDX_SPMV(true, 4); else SDX_SPMV(false, 4);
      break;
    case 8:
      if (refresh) SDX_SPMV(true, 8); else SDX_SPMV(false, 8);
      break;
    default:
      if (refresh) SDX_SPMV(true, 16); else SDX_SPMV(false, 16);
      break;
  }
#undef SDX_SPMV
}

template <class T, class A>
void launch_cg_update(hipStream_t s, const Tiles& tv, const WorkList& wl, int grid, const CgConfig& cfg, int k,
                      const T* dinv, int32_t nV, const T* w, const double* rref, A* x, T* r, T* z, T* p, T* sv,
                      EnvCG* env, A* acc, int32_t* num_active, const KernelEvents* ev) {
  if (tv.n == 0) return;
  const int g = wl.list ? grid : tv.n;
  if (cg_refresh(cfg, k))
    SDX_LAUNCH(ev, (k_cg_update<T, A, true>), dim3(g), dim3(kVecThreads), s, tv, wl, cfg, k, dinv, nV, w, rref, x,
               r, z, p, sv, env, acc, num_active);
  else
    SDX_LAUNCH(ev, (k_cg_update<T, A, false>), dim3(g), dim3(kVecThreads), s, tv, wl, cfg, k, dinv, nV, w, rref, x,
               r, z, p, sv, env, acc, num_active);
}

int env_lds_limit() {
  int dev = 0, lim = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lim <= 0)
    return 64 * 1024;
  return std::min(lim, kEnvLdsMax);
}

template <class T, class A>
void launch_cg_env(hipStream_t s, const EnvQueue& q, int grid, size_t lds_limit, const CgConfig& cfg, int lanes,
                   const int32_t* row_ptr, const int32_t* col, const int16_t* col16, const T* vals, int32_t nB,
                   const T* dinv, int32_t nV, const T* b, const T* z, A* x, T* r, T* p, T* sv, double* rref,
                   EnvCG* env, const A* acc, int32_t* num_active, const KernelEvents* ev) {
  if (q.n <= 0 || grid <= 0) return;
  if (q.max_rows <= 0 || !q.order || !q.next || !q.v0)
    throw std::invalid_argument("cg_env: empty queue description");
  const size_t lim = std::min(lds_limit, size_t(kEnvLdsMax));
  const bool vec = env_lds_bytes<T, A>(q.max_rows, true) <= lim;  // r, p, s and increments in LDS too
  const size_t lds = env_lds_bytes<T, A>(q.max_rows, vec);
  if (lds > lim) throw std::invalid_argument("cg_env: an environment's rows need more LDS than a workgroup may take");
  if (uint64_t(nB) * sizeof(T) >= (uint64_t(1) << 32) || uint64_t(cfg.n_rows) * 4 * sizeof(T) >= (uint64_t(1) << 32))
    throw std::invalid_argument("matrix too large for 32-bit buffer offsets of the CG matvec");
  const void* cp = col16 ? static_cast<const void*>(col16) : static_cast<const void*>(col);
  // above 64 KB a kernel's dynamic LDS must be allowed for it first
#define SDX_ENV1(L, C, V)                                                                                       \
  do {                                                                                                          \
    if (lds > size_t(64 * 1024)) {                                                                              \
      const hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_cg_env<T, A, L, C, V>),        \
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));           \
      if (err != hipSuccess) throw std::runtime_error("cg_env: hipFuncSetAttribute refused the workgroup's LDS");  \
    }                                                                                                           \
    if (ev)                                                                                                     \
      hipExtLaunchKernelGGL((k_cg_env<T, A, L, C, V>), dim3(grid), dim3(kEnvThreads), lds, s, ev->start,         \
                            ev->stop, 0, q, cfg, row_ptr, cp, vals, nB, dinv, nV, b, z, x, r, p, sv, rref, env,  \
                            acc, num_active);                                                                   \
    else                                                                                                        \
      hipLaunchKernelGGL((k_cg_env<T, A, L, C, V>), dim3(grid), dim3(kEnvThreads), lds, s, q, cfg, row_ptr, cp,  \
                         vals, nB, dinv, nV, b, z, x, r, p, sv, rref, env, acc, num_active);                    \
  } while (0)
#define SDX_ENV(L, C)             \
  do {                            \
    if (vec)                      \
      SDX_ENV1(L, C, true);       \
    else                          \
      SDX_ENV1(L, C, false);      \
  } while (0)
  switch (lanes) {
    case 4:
      if (col16) SDX_ENV(4, true); else SDX_ENV(4, false);
      break;
    case 8:
      if (col16) SDX_ENV(8, true); else SDX_ENV(8, false);
      break;
    default:
      if (col16) SDX_ENV(16, true); else SDX_ENV(16, false);
      break;
  }
#undef SDX_ENV
#undef SDX_ENV1
}

void launch_compact_tiles(hipStream_t s, const Tiles& ta, const Tiles& tv, int a0, int a1, int v0, int v1,
                          const EnvCG* env, int32_t* la, int32_t* lv, int32_t* counts, int par,
                          const int32_t* num_active, int32_t* publish, uint32_t* bar) {
  if (a0 < 0 || a1 < a0 || a1 > ta.n || v0 < 0 || v1 < v0 || v1 > tv.n)
    throw std::runtime_error("compact_tiles: tile range outside the tiles");
  const int na = a1 - a0, nv = v1 - v0;
  // at least one workgroup: it clears the other parity's counts and publishes
  const int nb = std::max(1, (na + 255) / 256 + (nv + 255) / 256);
  hipLaunchKernelGGL(k_compact_tiles, dim3(nb), dim3(256), 0, s, ta, tv, a0, na, v0, nv, env, la, lv, counts, par,
                     num_active, publish, bar);
}

template <class T, class A>
void launch_cg_tail(hipStream_t s, const Tiles& ta, const Tiles& tv, const WorkList& wa, const WorkList& wv, int grid,
                    const CgConfig& cfg, int k0, int n, int lanes, const int32_t* row_ptr, const int32_t* col,
                    const int16_t* col16, const T* vals, int32_t nB, const T* dinv, int32_t nV, A* x, const T* b,
                    T* w, double* rref, T* r, T* z, T* p, T* sv, EnvCG* env, A* acc, int32_t* num_active,
                    uint32_t* bar, int32_t* abort) {
  if (ta.n == 0 || tv.n == 0 || n <= 0) return;
  if (!wa.list || !wv.list || grid < 8) throw std::invalid_argument("cg_tail: work lists and >= 8 workgroups required");
  if (uint64_t(nB) * sizeof(T) >= (uint64_t(1) << 32) || uint64_t(cfg.n_rows) * 4 * sizeof(T) >= (uint64_t(1) << 32))
    throw std::invalid_argument("matrix too large for 32-bit buffer offsets of the CG matvec");
  const void* cp = col16 ? static_cast<const void*>(col16) : static_cast<const void*>(col);
  const void* kern = nullptr;
#define SDX_TAIL(L)                                                                                     \
  kern = col16 ? reinterpret_cast<const void*>(&k_cg_tail<T, A, L, true>)                              \
               : reinterpret_cast<const void*>(&k_cg_tail<T, A, L, false>)
  switch (lanes) {
    case 4: SDX_TAIL(4); break;
    case 8: SDX_TAIL(8); break;
    default: SDX_TAIL(16); break;
  }
#undef SDX_TAIL
  Tiles a = ta, v = tv;
  WorkList la = wa, lv = wv;
  CgConfig c = cfg;
  // one pointer per kernel parameter, in the kernel's order
  void* args[] = {&a, &v,    &la,   &lv, &c, &k0, &n,  &row_ptr, &cp,  &vals,       &nB,  &dinv, &nV,
                  &x, &b,    &w,    &rref, &r, &z,  &p, &sv,      &env, &acc, &num_active, &bar, &abort};
  const hipError_t e = hipLaunchCooperativeKernel(kern, dim3(unsigned(grid)), dim3(kSpmvThreads), args, 0, s);
  if (e != hipSuccess) throw std::runtime_error(std::string("hipLaunchCooperativeKernel: ") + hipGetErrorString(e));
}

template <class T, class X>
void launch_residual64(hipStream_t s, const Tiles& ta, int lanes, int nEnv, const int32_t* mask,
                       const int32_t* row_ptr, const int32_t* col, const T* vals, int32_t nB, const X* x,
                       const T* b, double* r64, double* acc) {
  if (ta.n == 0) return;
  if (uint64_t(nB) * sizeof(T) >= (uint64_t(1) << 32))
    throw std::invalid_argument("matrix too large for 32-bit buffer offsets of the FP64 residual");
  auto kern = lanes == 4   ? k_residual64<T, X, 4>
              : lanes == 8 ? k_residual64<T, X, 8>
                           : k_residual64<T, X, 16>;
  hipLaunchKernelGGL(kern, dim3(ta.n), dim3(kSpmvThreads), 0, s, ta, nEnv, mask, row_ptr, col, vals, nB, x, b, r64,
                     acc);
}

template <class T, class X>
void launch_residual64_multi(hipStream_t s, const Tiles& ta, int lanes, int nEnv, const int32_t* mask,
                             const int32_t* row_ptr, const int32_t* col, const T* vals, int32_t nB, const X* u,
                             const T* v, const T* w, const T* b, double* ru, double* rv, double* rw, double* acc) {
  if (ta.n == 0) return;
  if (uint64_t(nB) * sizeof(T) >= (uint64_t(1) << 32))
    throw std::invalid_argument("matrix too large for 32-bit buffer offsets of the FP64 residual");
  auto kern = lanes == 4   ? k_residual64_multi<T, X, 4>
              : lanes == 8 ? k_residual64_multi<T, X, 8>
                           : k_residual64_multi<T, X, 16>;
  hipLaunchKernelGGL(kern, dim3(ta.n), dim3(kSpmvThreads), 0, s, ta, nEnv, mask, row_ptr, col, vals, nB, u, v, w, b,
                     ru, rv, rw, acc);
}

template <class S, class D>
__global__ void k_convert(int64_t n, const S* __restrict__ s, D* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = D(s[i]);
}

template <class S, class D>
void launch_convert(hipStream_t s, int64_t n, const S* src, D* dst) {
  if (n == 0) return;
  hipLaunchKernelGGL((k_convert<S, D>), dim3(blocks_for(n, 256)), dim3(256), 0, s, n, src, dst);
}

template <class T, class A>
__global__ void k_cg_fold(int64_t n, A* __restrict__ x, T* __restrict__ xd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] += A(xd[i]);
  xd[i] = T(0);
}

template <class T, class A>
void launch_cg_fold(hipStream_t s, int64_t n, A* x, T* xd) {
  if (n == 0) return;
  hipLaunchKernelGGL((k_cg_fold<T, A>), dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, xd);
}

template <class T>
void launch_block_inverse(hipStream_t s, int32_t nV, const int32_t* diag_slot, const T* vals, int32_t nB, int prec,
                          T* dinv) {
  if (nV == 0) return;
  hipLaunchKernelGGL(k_block_inverse<T>, dim3(blocks_for(nV, 256)), dim3(256), 0, s, nV, diag_slot, vals, nB, prec,
                     dinv);
}

template <class T>
void launch_copy(hipStream_t s, int64_t n, const T* src, T* dst) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_copy<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, src, dst);
}

template <class T>
void launch_apply_pins(hipStream_t s, int64_t n, const uint8_t* pinned, const T* value, T* u) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_apply_pins<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, pinned, value, u);
}

template <class T>
void launch_ls_update(hipStream_t s, int32_t nV, const int32_t* row_env, const int32_t* ls_mask, const double* scale,
                      const T* sol0, const T* dx, T* x) {
  if (nV == 0) return;
  hipLaunchKernelGGL(k_ls_update<T>, dim3(blocks_for(nV, 256)), dim3(256), 0, s, nV, row_env, ls_mask, scale, sol0,
                     dx, x);
}

template <class T>
void launch_post_step(hipStream_t s, int32_t nV, const int32_t* row_env, const int32_t* reset, double dt,
                      const T* u0, T* u, T* v, T* hist, int n_hist) {
  if (nV == 0) return;
  hipLaunchKernelGGL(k_post_step<T>, dim3(blocks_for(nV, 256)), dim3(256), 0, s, nV, row_env, reset,
                     T(1.0 / dt), u0, u, v, hist, n_hist);
}

template <class T>
void launch_env_copy(hipStream_t s, int32_t nV, const int32_t* row_env, const int32_t* mask, const T* src, T* dst) {
  if (nV == 0) return;
  hipLaunchKernelGGL(k_env_copy<T>, dim3(blocks_for(nV, 256)), dim3(256), 0, s, nV, row_env, mask, src, dst);
}

template <class T>
void launch_env_norms(hipStream_t s, const Tiles& tv, int nEnv, const T* a, const T* b, double* acc) {
  if (tv.n == 0) return;
  hipLaunchKernelGGL(k_env_norms<T>, dim3(tv.n), dim3(kVecThreads), 0, s, tv, nEnv, a, b, acc);
}

template <class T>
void launch_env_dot(hipStream_t s, const Tiles& tv, int nEnv, const T* a, const T* b, double* acc) {
  if (tv.n == 0) return;
  hipLaunchKernelGGL(k_env_dot<T>, dim3(tv.n), dim3(kVecThreads), 0, s, tv, nEnv, a, b, acc);
}

template <class T>
void launch_energy(hipStream_t s, const MeshView& m, const FieldPtrs<T>& f, const T* u, const T* u0, const T* v0,
                   double dt, const double grav[3], double* acc) {
  const int nbt = int(blocks_for(m.nT, 256)), nbf = int(blocks_for(m.nF, 256));
  if (nbt + nbf == 0) return;
  hipLaunchKernelGGL(k_energy<T>, dim3(unsigned(nbt + nbf)), dim3(256), 0, s, m, f, u, u0, v0, T(dt), grav[0],
                     grav[1], grav[2], nbt, acc);
}

#define SDX_INST_T(T)                                                                                            \
  template void launch_element<T>(hipStream_t, const MeshView&, const AsmRange&, const FieldPtrs<T>&, const T*, \
                                  const T*, const T*, double, const double*, int, int, T*, T*, T*, int,         \
                                  const KernelEvents*, bool, const int32_t*);                                   \
  template void launch_contact<T>(hipStream_t, const MeshView&, const AsmRange&, const FieldPtrs<T>&, const T*, \
                                  const T*, T, int, T*, T*, T*);                                                \
  template void launch_gather_bsr<T>(hipStream_t, const MeshView&, const AsmRange&, const T*, const T*, int, T*, \
                                     T*, const KernelEvents*);                                                  \
  template void launch_block_inverse<T>(hipStream_t, int32_t, const int32_t*, const T*, int32_t, int, T*);     \
  template void launch_copy<T>(hipStream_t, int64_t, const T*, T*);                                             \
  template void launch_apply_pins<T>(hipStream_t, int64_t, const uint8_t*, const T*, T*);                       \
  template void launch_ls_update<T>(hipStream_t, int32_t, const int32_t*, const int32_t*, const double*,        \
                                    const T*, const T*, T*);                                                    \
  template void launch_post_step<T>(hipStream_t, int32_t, const int32_t*, const int32_t*, double, const T*, T*, \
                                    T*, T*, int);                                                               \
  template void launch_env_copy<T>(hipStream_t, int32_t, const int32_t*, const int32_t*, const T*, T*);        \
  template void launch_env_norms<T>(hipStream_t, const Tiles&, int, const T*, const T*, double*);              \
  template void launch_env_dot<T>(hipStream_t, const Tiles&, int, const T*, const T*, double*);                \
  template void launch_contact_wrench<T>(hipStream_t, const MeshView&, const FieldPtrs<T>&, const T*, const T*, \
                                         double, double*, long long*);                                         \
  template void launch_energy<T>(hipStream_t, const MeshView&, const FieldPtrs<T>&, const T*, const T*,         \
                                 const T*, double, const double*, double*);                                     \
  template void launch_fill<T>(hipStream_t, int64_t, T, T*);                                                    \
  template void launch_acc_finalize<T>(hipStream_t, int, int, const T*, double*);                               \
  template void launch_cg_reset<T>(hipStream_t, int, const int32_t*, EnvCG*, T*);

#define SDX_INST_TA(T, A)                                                                                         \
  template void launch_assemble_rows<T, A>(hipStream_t, const MeshView&, const RowBlockView&, const FieldPtrs<T>&, \
                                           const T*, const T*, const T*, double, const double*, int, int, T*, T*, \
                                           T*, A*, bool, const KernelEvents*);                                    \
  template void launch_gather_res<T, A>(hipStream_t, const MeshView&, const AsmRange&, const Tiles&, const T*,  \
                                        T*, A*);                                                                \
  template void launch_cg_init<T, A>(hipStream_t, const Tiles&, const CgConfig&, const T*, const int32_t*,       \
                                     const double*, const double*, const T*, const double*, const double*,      \
                                     const T*, const double*, const double*, const T*, int32_t, A*, T*, T*,     \
                                     EnvCG*, A*);                                                               \
  template void launch_residual64_multi<T, A>(hipStream_t, const Tiles&, int, int, const int32_t*,             \
                                              const int32_t*, const int32_t*, const T*, int32_t, const A*,      \
                                              const T*, const T*, const T*, double*, double*, double*,          \
                                              double*);                                                         \
  template void launch_residual64<T, A>(hipStream_t, const Tiles&, int, int, const int32_t*, const int32_t*,    \
                                        const int32_t*, const T*, int32_t, const A*, const T*, double*,         \
                                        double*);                                                               \
  template void launch_cg_spmv<T, A>(hipStream_t, const Tiles&, const WorkList&, int, const CgConfig&, int, int, \
                                     const int32_t*, const int32_t*, const T*, int32_t, const T*, const A*,      \
                                     const T*, T*, double*, EnvCG*, A*, int32_t*, const KernelEvents*,           \
                                     const int16_t*);                                                            \
  template void launch_cg_update<T, A>(hipStream_t, const Tiles&, const WorkList&, int, const CgConfig&, int,    \
                                       const T*, int32_t, const T*, const double*, A*, T*, T*, T*, T*, EnvCG*,   \
                                       A*, int32_t*, const KernelEvents*);                                       \
  template void launch_cg_env<T, A>(hipStream_t, const EnvQueue&, int, size_t, const CgConfig&, int,            \
                                    const int32_t*,                                                               \
                                    const int32_t*, const int16_t*, const T*, int32_t, const T*, int32_t, const T*, \
                                    const T*, A*, T*, T*, T*, double*, EnvCG*, const A*, int32_t*,                \
                                    const KernelEvents*);                                                         \
  template void launch_cg_tail<T, A>(hipStream_t, const Tiles&, const Tiles&, const WorkList&, const WorkList&,  \
                                     int, const CgConfig&, int, int, int, const int32_t*, const int32_t*,         \
                                     const int16_t*, const T*, int32_t, const T*, int32_t, A*, const T*, T*,      \
                                     double*, T*, T*, T*, T*, EnvCG*, A*, int32_t*, uint32_t*, int32_t*);

SDX_INST_T(float)
SDX_INST_T(double)
SDX_INST_TA(float, float)
SDX_INST_TA(float, double)
SDX_INST_TA(double, double)
template void launch_convert<float, float>(hipStream_t, int64_t, const float*, float*);
template void launch_fill<int32_t>(hipStream_t, int64_t, int32_t, int32_t*);
template void launch_fill<long long>(hipStream_t, int64_t, long long, long long*);
template void launch_convert<float, double>(hipStream_t, int64_t, const float*, double*);
template void launch_convert<double, float>(hipStream_t, int64_t, const double*, float*);
template void launch_convert<double, double>(hipStream_t, int64_t, const double*, double*);
template void launch_cg_fold<float, double>(hipStream_t, int64_t, double*, float*);
template void launch_cg_fold<double, double>(hipStream_t, int64_t, double*, double*);

}  // namespace sdx
