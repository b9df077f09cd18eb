// Device math of the soft-body element: 3x3 algebra, a robust symmetric 3x3
// eigen-solver, the rotation-variant SVD, the Smith et al. 2018 stable
// Neo-Hookean response and its PSD-projected element Hessian, and the
// PolyReLU contact penalty.
//
// Reference behaviour followed (paths under superdex_physics/libraries/mochi/):
//   mochi_core/src/materials/batched_smith_neo_hookean.cpp:29-269
//   mochi_core/include/mochi_core/element_operations/fem_stress.h:148-259
//   mochi_core/include/mochi_core/utils/activations.h:113-128
#pragma once

#include <hip/hip_runtime.h>

namespace sdx {

template <class T>
struct V3 {
  T x, y, z;
};

template <class T>
__device__ __forceinline__ V3<T> mk3(T a, T b, T c) {
  return V3<T>{a, b, c};
}
template <class T>
__device__ __forceinline__ T dot3(V3<T> a, V3<T> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
__device__ __forceinline__ V3<T> cross3(V3<T> a, V3<T> b) {
  return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
__device__ __forceinline__ V3<T> add3(V3<T> a, V3<T> b) {
  return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
__device__ __forceinline__ V3<T> sub3(V3<T> a, V3<T> b) {
  return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
__device__ __forceinline__ V3<T> scl3(T s, V3<T> a) {
  return V3<T>{s * a.x, s * a.y, s * a.z};
}

// Row-major 3x3 matrix.
template <class T>
struct M33 {
  T m[9];
};

template <class T>
__device__ __forceinline__ V3<T> mul(const M33<T>& A, V3<T> v) {
  return V3<T>{A.m[0] * v.x + A.m[1] * v.y + A.m[2] * v.z, A.m[3] * v.x + A.m[4] * v.y + A.m[5] * v.z,
               A.m[6] * v.x + A.m[7] * v.y + A.m[8] * v.z};
}
template <class T>
__device__ __forceinline__ V3<T> mulT(const M33<T>& A, V3<T> v) {
  return V3<T>{A.m[0] * v.x + A.m[3] * v.y + A.m[6] * v.z, A.m[1] * v.x + A.m[4] * v.y + A.m[7] * v.z,
               A.m[2] * v.x + A.m[5] * v.y + A.m[8] * v.z};
}
template <class T>
__device__ __forceinline__ T det33(const M33<T>& F) {
  return F.m[0] * (F.m[4] * F.m[8] - F.m[5] * F.m[7]) - F.m[1] * (F.m[3] * F.m[8] - F.m[5] * F.m[6]) +
         F.m[2] * (F.m[3] * F.m[7] - F.m[4] * F.m[6]);
}
// cof[i][j] = dJ/dF_ij
template <class T>
__device__ __forceinline__ M33<T> cofactor(const M33<T>& F) {
  M33<T> c;
  c.m[0] = F.m[4] * F.m[8] - F.m[5] * F.m[7];
  c.m[1] = F.m[5] * F.m[6] - F.m[3] * F.m[8];
  c.m[2] = F.m[3] * F.m[7] - F.m[4] * F.m[6];
  c.m[3] = F.m[2] * F.m[7] - F.m[1] * F.m[8];
  c.m[4] = F.m[0] * F.m[8] - F.m[2] * F.m[6];
  c.m[5] = F.m[1] * F.m[6] - F.m[0] * F.m[7];
  c.m[6] = F.m[1] * F.m[5] - F.m[2] * F.m[4];
  c.m[7] = F.m[2] * F.m[3] - F.m[0] * F.m[5];
  c.m[8] = F.m[0] * F.m[4] - F.m[1] * F.m[3];
  return c;
}

template <class T>
__device__ __forceinline__ T dsqrt(T x) {
  return sqrt(x);
}

// Unit vector orthogonal to unit v.
template <class T>
__device__ __forceinline__ V3<T> any_orthogonal(V3<T> v) {
  V3<T> a = (fabs(v.x) < T(0.9)) ? mk3<T>(1, 0, 0) : mk3<T>(0, 1, 0);
  V3<T> u = cross3(v, a);
  T n = dsqrt(dot3(u, u));
  return scl3(T(1) / n, u);
}

// Symmetric 3x3 eigen-decomposition, eigenvalues descending.  Eigenvalues by
// the trigonometric (Cardano) formula; the eigenvector of the most isolated
// eigenvalue from the largest cross product of two rows of (S - lambda I);
// the remaining pair by a 2x2 Jacobi rotation in the orthogonal plane, which
// stays well defined when the pair is degenerate.
// S = {s00, s11, s22, s01, s02, s12}; V columns are the eigenvectors.
template <class T>
__device__ void sym_eig3(const T S[6], T lam[3], V3<T> vec[3]) {
  const T s00 = S[0], s11 = S[1], s22 = S[2], s01 = S[3], s02 = S[4], s12 = S[5];
  const T p1 = s01 * s01 + s02 * s02 + s12 * s12;
  const T q = (s00 + s11 + s22) / T(3);
  const T d0 = s00 - q, d1 = s11 - q, d2 = s22 - q;
  const T p2 = d0 * d0 + d1 * d1 + d2 * d2 + T(2) * p1;
  const T p = dsqrt(p2 / T(6));
  const T scale = fabs(q) > T(1) ? fabs(q) : T(1);
  if (!(p > T(64) * (sizeof(T) == 4 ? T(1.1920929e-7) : T(2.220446049250313e-16)) * scale)) {
    lam[0] = lam[1] = lam[2] = q;
    vec[0] = mk3<T>(1, 0, 0);
    vec[1] = mk3<T>(0, 1, 0);
    vec[2] = mk3<T>(0, 0, 1);
    return;
  }
  const T ip = T(1) / p;
  const T b00 = d0 * ip, b11 = d1 * ip, b22 = d2 * ip, b01 = s01 * ip, b02 = s02 * ip, b12 = s12 * ip;
  T r = T(0.5) * (b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) + b02 * (b01 * b12 - b11 * b02));
  r = r < T(-1) ? T(-1) : (r > T(1) ? T(1) : r);
  const T phi = acos(r) / T(3);
  const T l0 = q + T(2) * p * cos(phi);
  const T l2 = q + T(2) * p * cos(phi + T(2.0943951023931954923));
  const T l1 = T(3) * q - l0 - l2;
  const int d = ((l0 - l1) > (l1 - l2)) ? 0 : 2;
  const T ld = d == 0 ? l0 : l2;
  const V3<T> r0 = mk3(s00 - ld, s01, s02), r1 = mk3(s01, s11 - ld, s12), r2 = mk3(s02, s12, s22 - ld);
  V3<T> c0 = cross3(r0, r1), c1 = cross3(r0, r2), c2 = cross3(r1, r2);
  T n0 = dot3(c0, c0), n1 = dot3(c1, c1), n2 = dot3(c2, c2);
  V3<T> vd = c0;
  T nd = n0;
  if (n1 > nd) {
    vd = c1;
    nd = n1;
  }
  if (n2 > nd) {
    vd = c2;
    nd = n2;
  }
  vd = scl3(T(1) / dsqrt(nd), vd);
  const V3<T> u = any_orthogonal(vd);
  const V3<T> w = cross3(vd, u);
  auto quad = [&](V3<T> a, V3<T> b) {
    V3<T> Sb = mk3(s00 * b.x + s01 * b.y + s02 * b.z, s01 * b.x + s11 * b.y + s12 * b.z,
                   s02 * b.x + s12 * b.y + s22 * b.z);
    return dot3(a, Sb);
  };
  const T a11 = quad(u, u), a12 = quad(u, w), a22 = quad(w, w);
  // Jacobi rotation diagonalising [[a11, a12], [a12, a22]]
  T c = T(1), s = T(0);
  if (a12 != T(0)) {
    const T tau = (a22 - a11) / (T(2) * a12);
    const T t = (tau >= T(0) ? T(1) : T(-1)) / (fabs(tau) + dsqrt(T(1) + tau * tau));
    c = T(1) / dsqrt(T(1) + t * t);
    s = t * c;
  }
  V3<T> e1 = sub3(scl3(c, u), scl3(s, w));
  V3<T> e2 = add3(scl3(s, u), scl3(c, w));
  T m1 = quad(e1, e1), m2 = quad(e2, e2);
  if (m1 < m2) {
    V3<T> te = e1;
    e1 = e2;
    e2 = te;
    T tm = m1;
    m1 = m2;
    m2 = tm;
  }
  if (d == 0) {
    lam[0] = l0;
    lam[1] = m1;
    lam[2] = m2;
    vec[0] = vd;
    vec[1] = e1;
    vec[2] = e2;
  } else {
    lam[0] = m1;
    lam[1] = m2;
    lam[2] = l2;
    vec[0] = e1;
    vec[1] = e2;
    vec[2] = vd;
  }
}

// Rotation-variant SVD F = U diag(sigma) V^T with det U = det V = +1 and
// sigma[2] < 0 iff det F < 0 (the convention of the reference's
// BatchedRotationVariantSvd3x3).  V from the eigenvectors of F^T F; U by
// Gram-Schmidt on F V.
template <class T>
__device__ void svd_rv(const M33<T>& F, V3<T> U[3], T sigma[3], V3<T> V[3]) {
  T S[6];
  S[0] = F.m[0] * F.m[0] + F.m[3] * F.m[3] + F.m[6] * F.m[6];
  S[1] = F.m[1] * F.m[1] + F.m[4] * F.m[4] + F.m[7] * F.m[7];
  S[2] = F.m[2] * F.m[2] + F.m[5] * F.m[5] + F.m[8] * F.m[8];
  S[3] = F.m[0] * F.m[1] + F.m[3] * F.m[4] + F.m[6] * F.m[7];
  S[4] = F.m[0] * F.m[2] + F.m[3] * F.m[5] + F.m[6] * F.m[8];
  S[5] = F.m[1] * F.m[2] + F.m[4] * F.m[5] + F.m[7] * F.m[8];
  T lam[3];
  sym_eig3(S, lam, V);
  if (dot3(cross3(V[0], V[1]), V[2]) < T(0)) V[2] = scl3(T(-1), V[2]);
  const V3<T> f0 = mul(F, V[0]);
  const T s0 = dsqrt(dot3(f0, f0));
  const T tiny = sizeof(T) == 4 ? T(1e-20) : T(1e-150);
  U[0] = s0 > tiny ? scl3(T(1) / s0, f0) : mk3<T>(1, 0, 0);
  const V3<T> f1 = mul(F, V[1]);
  const V3<T> f1p = sub3(f1, scl3(dot3(U[0], f1), U[0]));
  const T s1n = dsqrt(dot3(f1p, f1p));
  U[1] = s1n > tiny * (s0 > T(1) ? s0 : T(1)) ? scl3(T(1) / s1n, f1p) : any_orthogonal(U[0]);
  U[2] = cross3(U[0], U[1]);
  sigma[0] = s0;
  sigma[1] = dot3(U[1], f1);
  sigma[2] = dot3(U[2], mul(F, V[2]));
}

// PolyReLU (activations.h:113-128): C2 ramp of half-width t centred at shift.
template <class T>
__device__ __forceinline__ void poly_relu(T x, T t, T shift, T& f, T& df, T& ddf) {
  x -= shift;
  const T xt = x / t, x2t2 = xt * xt, x3t3 = x2t2 * xt;
  const bool mask = (t <= T(0)) || (fabs(x) > t);
  const T fv = mask ? x : T(3.0 / 16.0) * t + (T(0.5) + T(3.0 / 8.0) * xt - T(1.0 / 16.0) * x3t3) * x;
  f = fv > T(0) ? fv : T(0);
  T dfv = mask ? (x < T(0) ? T(0) : T(1)) : T(0.5) + T(0.75) * xt - T(0.25) * x3t3;
  df = dfv < T(0) ? T(0) : (dfv > T(1) ? T(1) : dfv);
  const T ddv = mask ? T(0) : T(0.75) * (T(1) - x2t2) / t;
  ddf = ddv > T(0) ? ddv : T(0);
}

// Offsets of the 10 upper blocks (a <= b) of a packed symmetric 12x12 element
// Hessian: diagonal blocks keep their 6 upper entries {00,11,22,01,02,12},
// off-diagonal blocks all 9 row-major.  Total 78, padded to 80.
__host__ __device__ constexpr int he_block_offset(int a, int b) {
  // order (0,0)(0,1)(0,2)(0,3)(1,1)(1,2)(1,3)(2,2)(2,3)(3,3)
  return a == 0 ? (b == 0 ? 0 : 6 + 9 * (b - 1))
       : a == 1 ? (b == 1 ? 33 : 39 + 9 * (b - 2))
       : a == 2 ? (b == 2 ? 57 : 63)
                : 72;
}

// Rank-1 update of the packed Hessian by eigenpair MODE of Smith's
// eigensystem (0-2 scaling, 3-8 twist / flip about axis (MODE - 3) / 2), as a
// template so every array index is a compile-time constant.
template <int MODE, class T>
__device__ __forceinline__ void smith_mode(const T la[3], const V3<T> ae[3], const V3<T> w[4], const V3<T> U[3],
                                           const T s[3], T c0, T c1, T vol, T eps, T H[78]) {
  T lk;
  V3<T> y[4];
  if (MODE < 3) {
    // scaling: Q = U diag(ae) V^T  ->  y_f = U (ae .* w_f)
    lk = la[MODE < 3 ? MODE : 0];
    const V3<T> e = ae[MODE < 3 ? MODE : 0];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const T t0 = e.x * w[f].x, t1 = e.y * w[f].y, t2 = e.z * w[f].z;
      y[f] = add3(add3(scl3(t0, U[0]), scl3(t1, U[1])), scl3(t2, U[2]));
    }
  } else {
    // twist / flip in the plane (j, k) orthogonal to axis i, oriented j < k
    constexpr int i = MODE < 3 ? 0 : (MODE - 3) >> 1;
    constexpr bool twist = MODE < 3 ? false : ((MODE - 3) & 1) == 0;
    constexpr int j = (i + 1) % 3, k = (i + 2) % 3;
    constexpr int jj = j < k ? j : k, kk = j < k ? k : j;
    lk = twist ? (c0 + c1 * s[i]) : (c0 - c1 * s[i]);
    const T sg = twist ? T(-1) : T(1);
    const T isq2 = T(0.70710678118654752440);
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const T wv[3] = {w[f].x, w[f].y, w[f].z};
      // M = e_jj e_kk^T + sg e_kk e_jj^T  ->  M w = e_jj w_kk + sg e_kk w_jj
      y[f] = scl3(isq2, add3(scl3(wv[kk], U[jj]), scl3(sg * wv[jj], U[kk])));
    }
  }
  const T lam = (lk > eps ? lk : eps) * vol;
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int g = f; g < 4; ++g) {
      T* blk = H + he_block_offset(f, g);
      const T yf[3] = {y[f].x, y[f].y, y[f].z}, yg[3] = {y[g].x, y[g].y, y[g].z};
      if (f == g) {
        blk[0] += lam * yf[0] * yf[0];
        blk[1] += lam * yf[1] * yf[1];
        blk[2] += lam * yf[2] * yf[2];
        blk[3] += lam * yf[0] * yf[1];
        blk[4] += lam * yf[0] * yf[2];
        blk[5] += lam * yf[1] * yf[2];
      } else {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) blk[r * 3 + c] += lam * yf[r] * yg[c];
      }
    }
}

// Smith stable Neo-Hookean energy density (batched_smith_neo_hookean.cpp:51-68)
//   psi = mu^/2 (Ic - 3) + lambda^/2 (J - alpha)^2 - mu^/2 ln(Ic + 1),
//   mu^ = 4/3 mu, lambda^ = lambda + 5/6 mu, alpha = 1 + 3/4 mu^/lambda^,
// evaluated in FP64 from F.
template <class T>
__device__ __forceinline__ double smith_energy(const M33<T>& F, double mu, double lambda) {
  double f[9], Ic = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    f[i] = double(F.m[i]);
    Ic += f[i] * f[i];
  }
  const double J = f[0] * (f[4] * f[8] - f[5] * f[7]) - f[1] * (f[3] * f[8] - f[5] * f[6]) +
                   f[2] * (f[3] * f[7] - f[4] * f[6]);
  const double muHat = mu * (4.0 / 3.0), lambdaHat = lambda + mu * (5.0 / 6.0);
  const double alpha = 1.0 + 0.75 * muHat / lambdaHat;
  const double Jma = J - alpha;
  return 0.5 * muHat * (Ic - 3.0) + 0.5 * lambdaHat * Jma * Jma - 0.5 * muHat * log(Ic + 1.0);
}

// Stiffness (viscous) damping of one tet (fem_stress_damping.h:102-501),
// isotropic rest stiffness C0 = Lame(lambda, mu) of the Smith material:
//   dE = E(F) - E(Fs), S = kappa (lambda tr(dE) I + 2 mu dE), kappa = beta/dt
//   residual  vol F S dN_f
//   tangent   vol kappa [lambda a_f a_g^T + mu ((F F^T)(dN_f.dN_g) + a_g a_f^T)], a = F dN,
//             plus (geometric term on) vol (dN_f . Sg dN_g) I, Sg = S projected to
//             PSD (eigenvalues clamped at eps) where it is not positive definite and
//             the tangent is projected.
// Symmetric 3x3 values in the order 00 11 22 01 02 12.
template <class T>
struct Viscous {
  M33<T> F;
  T S[6], Sg[6], FFt[6];
  T kl, km;
};

// psi (if not NULL) receives the energy density dE:S/2 in FP64.
template <class T>
__device__ __forceinline__ void viscous_setup(const V3<T> dN[4], const T u[12], const T us[12], T kappa, T mu,
                                              T lambda, bool geo, bool project, Viscous<T>& v,
                                              double* psi = nullptr) {
  M33<T> Fs;
#pragma unroll
  for (int k = 0; k < 9; ++k) v.F.m[k] = Fs.m[k] = (k % 4 == 0) ? T(1) : T(0);
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const T d[3] = {dN[f].x, dN[f].y, dN[f].z};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        v.F.m[r * 3 + c] += u[f * 3 + r] * d[c];
        Fs.m[r * 3 + c] += us[f * 3 + r] * d[c];
      }
  }
  const int ii[6] = {0, 1, 2, 0, 0, 1}, jj[6] = {0, 1, 2, 1, 2, 2};
  T dE[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    T c = T(0);
#pragma unroll
    for (int p = 0; p < 3; ++p)
      c += v.F.m[p * 3 + ii[k]] * v.F.m[p * 3 + jj[k]] - Fs.m[p * 3 + ii[k]] * Fs.m[p * 3 + jj[k]];
    dE[k] = T(0.5) * c;
    v.FFt[k] = v.F.m[ii[k] * 3] * v.F.m[jj[k] * 3] + v.F.m[ii[k] * 3 + 1] * v.F.m[jj[k] * 3 + 1] +
               v.F.m[ii[k] * 3 + 2] * v.F.m[jj[k] * 3 + 2];
  }
  const T tr = dE[0] + dE[1] + dE[2];
  v.kl = kappa * lambda;
  v.km = kappa * mu;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    v.S[k] = T(2) * v.km * dE[k] + (k < 3 ? v.kl * tr : T(0));
    v.Sg[k] = geo ? v.S[k] : T(0);
  }
  if (psi) {
    double e = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) e += (k < 3 ? 0.5 : 1.0) * double(v.S[k]) * double(dE[k]);
    *psi = e;
  }
  if (geo && project) {
    const T* q = v.S;
    const T m2 = q[0] * q[1] - q[3] * q[3];
    const T m3 = q[0] * (q[1] * q[2] - q[5] * q[5]) - q[3] * (q[3] * q[2] - q[5] * q[4]) +
                 q[4] * (q[3] * q[5] - q[1] * q[4]);
    if (!(q[0] > T(0) && m2 > T(0) && m3 > T(0))) {
      T la[3];
      V3<T> ev[3];
      sym_eig3(v.S, la, ev);
      const T eps = sizeof(T) == 4 ? T(1.1920928955078125e-07) : T(2.220446049250313e-16);
#pragma unroll
      for (int k = 0; k < 3; ++k) la[k] = la[k] > eps ? la[k] : eps;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const V3<T> a = ev[0], b = ev[1], c = ev[2];
        const T ai[3] = {a.x, a.y, a.z}, bi[3] = {b.x, b.y, b.z}, ci[3] = {c.x, c.y, c.z};
        v.Sg[k] = la[0] * ai[ii[k]] * ai[jj[k]] + la[1] * bi[ii[k]] * bi[jj[k]] + la[2] * ci[ii[k]] * ci[jj[k]];
      }
    }
  }
}

// Residual res_f += vol F S dN_f.
template <class T>
__device__ __forceinline__ void viscous_res(const Viscous<T>& v, const V3<T> dN[4], T vol, T res[12]) {
  const T* S = v.S;
  const M33<T> Sm = {{S[0], S[3], S[4], S[3], S[1], S[5], S[4], S[5], S[2]}};
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const V3<T> sd = mul(Sm, dN[f]);
    const V3<T> p = mul(v.F, sd);
    res[3 * f + 0] += vol * p.x;
    res[3 * f + 1] += vol * p.y;
    res[3 * f + 2] += vol * p.z;
  }
}

// Block (f, g) of the viscous tangent added to blk: f == g the 6 entries
// 00 11 22 01 02 12, else 9 row-major.
template <class T>
__device__ __forceinline__ void viscous_block(const Viscous<T>& v, const V3<T>& dNf, const V3<T>& dNg, bool diag,
                                              T vol, T* blk) {
  const V3<T> af = mul(v.F, dNf), ag = mul(v.F, dNg);
  const T dd = dot3(dNf, dNg);
  const T* Sg = v.Sg;
  const M33<T> Sm = {{Sg[0], Sg[3], Sg[4], Sg[3], Sg[1], Sg[5], Sg[4], Sg[5], Sg[2]}};
  const T gs = dot3(dNf, mul(Sm, dNg));
  const T a1[3] = {af.x, af.y, af.z}, a2[3] = {ag.x, ag.y, ag.z};
  const T* Q = v.FFt;
  const T FF[9] = {Q[0], Q[3], Q[4], Q[3], Q[1], Q[5], Q[4], Q[5], Q[2]};
  auto entry = [&](int i, int j) {
    T x = v.kl * a1[i] * a2[j] + v.km * (FF[i * 3 + j] * dd + a2[i] * a1[j]);
    if (i == j) x += gs;
    return vol * x;
  };
  if (diag) {
    blk[0] += entry(0, 0);
    blk[1] += entry(1, 1);
    blk[2] += entry(2, 2);
    blk[3] += entry(0, 1);
    blk[4] += entry(0, 2);
    blk[5] += entry(1, 2);
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) blk[i * 3 + j] += entry(i, j);
  }
}

// Element kernel core: stress residual (12) and, if wantH, the 10 upper 3x3
// blocks of the Hessian of one tet, both scaled by the Q1 weight `vol`; each
// block goes to emit(f, g, blk) as soon as it is formed (f == g: 6 entries
// 00 11 22 01 02 12, else 9 row-major), so the 78 values are never live
// together on the direct path.
//   F = I + sum_f u_f (x) dN_f
//   res_f = vol * P dN_f
//   direct tangent:   H_fg = vol [c3 a_f a_g^T + lh b_f b_g^T + c2 (dN_f.dN_g) I
//                              + lh (J - alpha) skew(F (dN_f x dN_g))]
//     with a_f = F dN_f, b_f = cof(F) dN_f, skew(v)_rc = eps_rcm v_m
//   projected tangent: H_fg = vol sum_k max(l_k, eps) (Q_k dN_f)(Q_k dN_g)^T
//     over Smith's 9 eigenpairs (3 scaling, 3 twist, 3 flip).
template <class T, class OnRes, class Emit>
__device__ __forceinline__ void smith_element(const V3<T> dN[4], const T u[12], T vol, T mu, T lambda, bool wantH,
                                              bool projectPsd, OnRes&& on_res, Emit&& emit) {
  M33<T> F;
  F.m[0] = T(1);
  F.m[1] = T(0);
  F.m[2] = T(0);
  F.m[3] = T(0);
  F.m[4] = T(1);
  F.m[5] = T(0);
  F.m[6] = T(0);
  F.m[7] = T(0);
  F.m[8] = T(1);
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const T d[3] = {dN[f].x, dN[f].y, dN[f].z};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) F.m[r * 3 + c] += u[f * 3 + r] * d[c];
  }
  const T one = T(1);
  const T muHat = mu * T(4.0 / 3.0);
  const T lambdaHat = lambda + mu * T(5.0 / 6.0);
  const T alpha = one + T(0.75) * muHat / lambdaHat;
  T Ic = T(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) Ic += F.m[i] * F.m[i];
  const T IcPlus1 = Ic + one;
  const T IcPlus1Inv = one / IcPlus1;
  const T J = det33(F);
  const T Jma = J - alpha;
  const M33<T> cof = cofactor(F);
  const T c0 = muHat * (one - IcPlus1Inv);
  const T c1 = lambdaHat * Jma;
  // res_f = vol * (c0 F + c1 cof) dN_f, handed to on_res before the Hessian
  // so that nothing of the residual stays live through it
  {
    T res[12];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const V3<T> a = mul(F, dN[f]);
      const V3<T> b = mul(cof, dN[f]);
      res[3 * f + 0] = vol * (c0 * a.x + c1 * b.x);
      res[3 * f + 1] = vol * (c0 * a.y + c1 * b.y);
      res[3 * f + 2] = vol * (c0 * a.z + c1 * b.z);
    }
    on_res(res);
  }
  if (!wantH) return;
  // Conservative indefiniteness test (Ic bounds max sigma^2); a tangent that
  // passes it is PSD, so the direct form equals its projection.
  const bool indefinite = projectPsd && ((c1 * c1 * Ic > c0 * c0) || (J < T(0)));
  if (!indefinite) {
    const T c2 = c0;
    const T c3 = T(2) * muHat * IcPlus1Inv * IcPlus1Inv;
    V3<T> a[4], b[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      a[f] = mul(F, dN[f]);
      b[f] = mul(cof, dN[f]);
    }
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int g = f; g < 4; ++g) {
        const T av[3] = {a[f].x, a[f].y, a[f].z}, ag[3] = {a[g].x, a[g].y, a[g].z};
        const T bv[3] = {b[f].x, b[f].y, b[f].z}, bg[3] = {b[g].x, b[g].y, b[g].z};
        const T dd = dot3(dN[f], dN[g]);
        const V3<T> sk = mul(F, cross3(dN[f], dN[g]));
        const T skv[3] = {sk.x, sk.y, sk.z};
        T blk[9];
        if (f == g) {
          // entries 00,11,22,01,02,12 ; skew term vanishes on the diagonal block
          const int rr[6] = {0, 1, 2, 0, 0, 1}, cc[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            T v = c3 * av[rr[k]] * ag[cc[k]] + lambdaHat * bv[rr[k]] * bg[cc[k]];
            if (rr[k] == cc[k]) v += c2 * dd;
            blk[k] = vol * v;
          }
        } else {
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              T v = c3 * av[r] * ag[c] + lambdaHat * bv[r] * bg[c];
              if (r == c) v += c2 * dd;
              // skew(v)_rc = eps_rcm v_m
              const int m = 3 - r - c;
              if (r != c) {
                const T e = ((r + 1) % 3 == c) ? T(1) : T(-1);
                v += c1 * e * skv[m];
              }
              blk[r * 3 + c] = vol * v;
            }
        }
        emit(f, g, blk);
      }
    return;
  }
  // Eigensystem path (Smith et al. 2018), clamped at eps (material_params_utils.h:66).
  V3<T> U[3], Vv[3];
  T s[3];
  svd_rv(F, U, s, Vv);
  const T eps = sizeof(T) == 4 ? T(1.1920928955078125e-07) : T(2.220446049250313e-16);
  const T i2 = IcPlus1Inv * IcPlus1Inv;
  const T mu2 = T(2) * muHat;
  const T J2 = (T(2) * J - alpha) * lambdaHat;
  const T Aq[6] = {(mu2 * s[0] * s[0] - muHat * IcPlus1) * i2 + lambdaHat * s[1] * s[1] * s[2] * s[2] + muHat,
                   (mu2 * s[1] * s[1] - muHat * IcPlus1) * i2 + lambdaHat * s[0] * s[0] * s[2] * s[2] + muHat,
                   (mu2 * s[2] * s[2] - muHat * IcPlus1) * i2 + lambdaHat * s[0] * s[0] * s[1] * s[1] + muHat,
                   J2 * s[2] + mu2 * s[0] * s[1] * i2,
                   J2 * s[1] + mu2 * s[0] * s[2] * i2,
                   J2 * s[0] + mu2 * s[1] * s[2] * i2};
  T la[3];
  V3<T> ae[3];
  sym_eig3(Aq, la, ae);
  // w_f = V^T dN_f
  V3<T> w[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) w[f] = mk3(dot3(Vv[0], dN[f]), dot3(Vv[1], dN[f]), dot3(Vv[2], dN[f]));
  T H[78];
#pragma unroll
  for (int k = 0; k < 78; ++k) H[k] = T(0);
  smith_mode<0>(la, ae, w, U, s, c0, c1, vol, eps, H);
  smith_mode<1>(la, ae, w, U, s, c0, c1, vol, eps, H);
  smith_mode<2>(la, ae, w, U, s, c0, c1, vol, eps, H);
  smith_mode<3>(la, ae, w, U, s, c0, c1, vol, eps, H);
  smith_mode<4>(la, ae, w, U, s, c0, c1, vol, eps, H);
  smith_mode<5>(la, ae, w, U, s, c0, c1, vol, eps, H);
  smith_mode<6>(la, ae, w, U, s, c0, c1, vol, eps, H);
  smith_mode<7>(la, ae, w, U, s, c0, c1, vol, eps, H);
  smith_mode<8>(la, ae, w, U, s, c0, c1, vol, eps, H);
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int g = f; g < 4; ++g) emit(f, g, H + he_block_offset(f, g));
}

}  // namespace sdx
