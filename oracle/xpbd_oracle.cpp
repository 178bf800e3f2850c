// CPU restatement of the reference XPBD Neo-Hookean solver (TEST INFRASTRUCTURE).
//
// This file is the correctness oracle and the timed CPU baseline.  It is never
// part of the product path: only tests, the smoke check and the benchmark's
// CPU-baseline leg load it.  It restates, kernel by kernel and in the same
// order, the reference solver written in Taichi:
//
//   xpbd_gpu/neohookean_solver.py   (RaymondMcGuire/rdc-xpbd-taichi)
//   xpbd_gpu/xpbd_base.py
//   xpbd_gpu/sdf.py
//   xpbd_gpu/constants.py
//
// Each function names the reference lines it follows.  The template parameter
// R is the arithmetic type: float restates the reference's f32 fields, double
// is the FP64 run used for the error-vs-FP64 figures.
//
// Scatter semantics.  The reference accumulates per-vertex corrections with
// float atomics (neohookean_solver.py:443-451), whose order is unspecified.
// This oracle fixes one order: ascending tet index.  Per-tet corner values are
// computed (optionally with OpenMP threads) into a buffer and summed per
// vertex through a vertex->(tet,corner) adjacency list sorted by tet, which is
// bit-identical to a sequential scatter loop over tets, for any thread count.
// Scalar sums (residual, mass, errors) are likewise summed sequentially in tet
// order.  Host-side scalars that the reference keeps in Python doubles (rho,
// omega, time) are doubles here for both R.
//
// Built with -ffp-contract=off so that no fused multiply-add changes the op
// order written below (liboracle_xpbd.so).  The same source is also built with
// contraction on (-ffp-contract=fast -mfma, liboracle_xpbd_fma.so): that
// flavour, and the run-time variants below, exist only to measure how far an
// f32 run moves when its arithmetic changes in ways the reference leaves
// open, so that the tolerance gates can be set at that level.
//
// Variants (Params.variant, bit flags; 0 is the restatement itself):
//   kVarOrder  every scatter and every sum is taken in a seeded random order
//              (the reference's float atomics fix no order)
//   kVarUlp    1/x and sqrt are rounded faithfully (exact when representable,
//              otherwise either neighbouring float, by a seeded choice) where
//              a GPU would use its native 1-ulp instructions (rest-shape
//              inverse, Jacobi average, residual), and x / c is formed as
//              x * (1/c) there
//   kVarFold   the compliance is formed as (6 / (lame dt^2)) * (1 / det Dm)
//              instead of 1 / (lame V0) / dt^2
//
// Pins.  Every function is checked against closed forms, brute force or
// library routines by tests/test_oracle_pins.py (Lame constants, rest shape,
// lumped mass and pinning, free fall, each collider, gradients against finite
// differences, the 2x2 solve and its fallback, the Jacobi average, the
// residual at rest, the omega schedule, f32 against f64) and its outputs are
// frozen in tests/golden.  Parity unpinned: the operation order inside Taichi's
// Matrix.inverse()/determinant() (restated here by their cofactor formulas)
// and the order of the reference's float-atomic scatter (fixed here to
// ascending tet index); both only change round-off.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kEpsilon = 1e-12;  // xpbd_gpu/constants.py:3

constexpr int kVarOrder = 1;
constexpr int kVarUlp = 2;
constexpr int kVarFold = 4;

// A seeded choice of -1, 0 or +1 ulp for the value at (a, b).
inline int ulp_step(uint64_t seed, uint64_t a, uint64_t b) {
  uint64_t h = seed * 0x9E3779B97F4A7C15ull ^ (a + 0x632BE59BD9B4E019ull) * 0xBF58476D1CE4E5B9ull ^
               (b + 0x94D049BB133111EBull) * 0x94D049BB133111EBull;
  h ^= h >> 31;
  h *= 0xD6E8FEB86659FD93ull;
  h ^= h >> 32;
  return int(h % 3) - 1;
}

// A faithfully rounded f32 result for the exact value `exact` (computed in
// double from f32 operands): the value itself when it is representable,
// otherwise one of the two floats around it -- the nearest for `step` 1 or
// 2, the other one for `step` 3 (ulp_step() + 2 picks each with probability
// 1/3).  This is the accuracy of 1-ulp instructions: exact results stay
// exact.  (Rounding the double to float is the correctly rounded f32 result:
// double carries more than 2 x 24 + 2 bits.)
inline float faithful_f32(double exact, int step) {
  const float f = float(exact);
  if (double(f) == exact || step < 3) return f;
  return std::nextafter(f, double(f) < exact ? INFINITY : -INFINITY);
}

// 1/x and sqrt(x) as the kVarUlp variant forms them (f32 only; the f64 run
// and step 0 give the correctly rounded result, by the ordinary operation).
template <typename R>
inline R recip(R x, int step) {
  if (sizeof(R) != sizeof(float) || step == 0) return R(1) / x;
  return R(faithful_f32(1.0 / double(x), step));
}
template <typename R>
inline R root(R x, int step) {
  if (sizeof(R) != sizeof(float) || step == 0) return std::sqrt(x);
  return R(faithful_f32(std::sqrt(double(x)), step));
}

template <typename R>
struct M3 {  // row-major 3x3
  R m[3][3];
};

template <typename R>
inline void cross3(const R a[3], const R b[3], R out[3]) {
  // ti.Vector.cross: (a1 b2 - a2 b1, a2 b0 - a0 b2, a0 b1 - a1 b0)
  out[0] = a[1] * b[2] - a[2] * b[1];
  out[1] = a[2] * b[0] - a[0] * b[2];
  out[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename R>
inline R dot3(const R a[3], const R b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// 3x3 determinant by cofactor expansion along the first column (the form
// Taichi's Matrix.determinant uses for n == 3).
template <typename R>
inline R det3(const M3<R>& a) {
  const auto& m = a.m;
  return m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
         m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
         m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
}

// 3x3 inverse as adjugate / determinant (Matrix.inverse for n == 3); a
// non-zero `step` forms the reciprocal of the determinant faithfully
// rounded (kVarUlp).
template <typename R>
inline M3<R> inv3(const M3<R>& a, int step = 0) {
  const auto& m = a.m;
  const R inv_det = recip(det3(a), step);
  M3<R> r;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
      const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      r.m[j][i] = inv_det * (m[i1][j1] * m[i2][j2] - m[i2][j1] * m[i1][j2]);
    }
  }
  return r;
}

// A @ B
template <typename R>
inline M3<R> mul(const M3<R>& a, const M3<R>& b) {
  M3<R> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
  return r;
}

// A @ B^T
template <typename R>
inline M3<R> mul_bt(const M3<R>& a, const M3<R>& b) {
  M3<R> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.m[i][j] = a.m[i][0] * b.m[j][0] + a.m[i][1] * b.m[j][1] + a.m[i][2] * b.m[j][2];
  return r;
}

// Matrix.norm_sqr(): sum of squared entries, row-major.
template <typename R>
inline R norm_sqr(const M3<R>& a) {
  R s = R(0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s += a.m[i][j] * a.m[i][j];
  return s;
}

// F = Ds @ inv_Dm with Ds = cols(p1-p0, p2-p0, p3-p0)
// (neohookean_solver.py:371-373).
template <typename R>
inline M3<R> deformation_gradient(const R* p0, const R* p1, const R* p2, const R* p3,
                                  const M3<R>& inv_dm) {
  M3<R> ds;
  for (int r = 0; r < 3; ++r) {
    ds.m[r][0] = p1[r] - p0[r];
    ds.m[r][1] = p2[r] - p0[r];
    ds.m[r][2] = p3[r] - p0[r];
  }
  return mul(ds, inv_dm);
}

// Gradients of C w.r.t. the 4 vertices from dC/dF: columns of dC/dF @ inv_Dm^T
// for vertices 1..3, and g0 = -(g1+g2+g3) (neohookean_solver.py:382-404).
template <typename R>
inline void vertex_grads(const M3<R>& dcdf, const M3<R>& inv_dm, R g[4][3]) {
  const M3<R> g123 = mul_bt(dcdf, inv_dm);
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) g[c + 1][r] = g123.m[r][c];
  for (int r = 0; r < 3; ++r) g[0][r] = -(g[1][r] + g[2][r] + g[3][r]);
}

// dC_H/dF = cols(F1 x F2, F2 x F0, F0 x F1) (neohookean_solver.py:391-398).
template <typename R>
inline M3<R> ddet_dF(const M3<R>& f) {
  R c0[3], c1[3], c2[3], x[3];
  for (int r = 0; r < 3; ++r) {
    c0[r] = f.m[r][0];
    c1[r] = f.m[r][1];
    c2[r] = f.m[r][2];
  }
  M3<R> g;
  cross3(c1, c2, x);
  for (int r = 0; r < 3; ++r) g.m[r][0] = x[r];
  cross3(c2, c0, x);
  for (int r = 0; r < 3; ++r) g.m[r][1] = x[r];
  cross3(c0, c1, x);
  for (int r = 0; r < 3; ++r) g.m[r][2] = x[r];
  return g;
}

// solve2x2_pivot_lu (neohookean_solver.py:331-356).  Returns ok.
template <typename R>
inline int solve2x2_pivot_lu(R a00, R a01, R a10, R a11, R b0, R b1, R eps, R* x0, R* x1) {
  int ok = 1;
  if (std::fabs(a00) < std::fabs(a10)) {
    std::swap(a00, a10);
    std::swap(a01, a11);
    std::swap(b0, b1);
  }
  if (std::fabs(a00) < eps) ok = 0;
  *x0 = R(0);
  *x1 = R(0);
  if (ok == 1) {
    const R m = a10 / a00;
    const R a11_ = a11 - m * a01;
    const R b1_ = b1 - m * b0;
    if (std::fabs(a11_) < eps) {
      ok = 0;
    } else {
      *x1 = b1_ / a11_;
      *x0 = (b0 - a01 * *x1) / a00;
    }
  }
  return ok;
}

// The compliances alpha~_D = 1 / (lame_D V0) / dt^2 and alpha~_H (with
// lame_D = mu = lame2, lame_H = lambda = lame1), neohookean_solver.py:413-419
// (also :178, :238, :318-319); kVarFold forms them as
// (6 / (lame dt^2)) * (1 / det Dm) instead.
template <typename R>
inline void compliance(R lame1, R lame2, R V0, R inv_det, R dt, int variant, R* aD, R* aH) {
  if (variant & kVarFold) {
    *aD = R(6.0 / (double(lame2) * double(dt) * double(dt))) * inv_det;
    *aH = R(6.0 / (double(lame1) * double(dt) * double(dt))) * inv_det;
  } else {
    const R alpha_D = R(1) / (lame2 * V0);
    const R alpha_H = R(1) / (lame1 * V0);
    *aD = alpha_D / (dt * dt);
    *aH = alpha_H / (dt * dt);
  }
}

// Block projection of one tet, solve_neohookean_block
// (neohookean_solver.py:358-451): C_D = |F|^2 - 3, C_H = det F - 1, their
// vertex gradients, the 2x2 system with the compliances on its diagonal,
// the pivoted LU solve and its diagonal fallback (:423-432), the multiplier
// update and the four corner corrections w_i (g_Di dlD + g_Hi dlH).
// grads (optional, 24): g_D then g_H per vertex.  aux (optional, 8): C_D,
// C_H, A11, A12, A22, dlD, dlH, ok.
template <typename R>
inline void block_projection(const R* const p[4], const R w[4], const M3<R>& idm, R aD, R aH,
                             R& lam_d, R& lam_h, R* corner, R* grads, R* aux) {
  const M3<R> F = deformation_gradient(p[0], p[1], p[2], p[3], idm);
  const R C_D = norm_sqr(F) - R(3);
  const R C_H = det3(F) - R(1);
  M3<R> gfd;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) gfd.m[i][j] = F.m[i][j] * R(2);
  R gD[4][3], gH[4][3];
  vertex_grads(gfd, idm, gD);
  vertex_grads(ddet_dF(F), idm, gH);
  R A11 = w[0] * dot3(gD[0], gD[0]) + w[1] * dot3(gD[1], gD[1]) +
          w[2] * dot3(gD[2], gD[2]) + w[3] * dot3(gD[3], gD[3]);
  R A22 = w[0] * dot3(gH[0], gH[0]) + w[1] * dot3(gH[1], gH[1]) +
          w[2] * dot3(gH[2], gH[2]) + w[3] * dot3(gH[3], gH[3]);
  const R A12 = w[0] * dot3(gD[0], gH[0]) + w[1] * dot3(gD[1], gH[1]) +
                w[2] * dot3(gD[2], gH[2]) + w[3] * dot3(gD[3], gH[3]);
  A11 += aD;
  A22 += aH;
  const R b1 = -(C_D + aD * lam_d);
  const R b2 = -(C_H + aH * lam_h);
  const R eps = R(kEpsilon);
  R dlD, dlH;
  const int ok = solve2x2_pivot_lu(A11, A12, A12, A22, b1, b2, eps, &dlD, &dlH);
  if (ok == 0) {  // neohookean_solver.py:425-432
    dlD = R(0);
    dlH = R(0);
    if (std::fabs(A11) > eps) dlD = b1 / A11;
    if (std::fabs(A22) > eps) dlH = b2 / A22;
  }
  lam_d += dlD;
  lam_h += dlH;
  for (int k = 0; k < 4; ++k)
    for (int r = 0; r < 3; ++r) corner[3 * k + r] = w[k] * (gD[k][r] * dlD + gH[k][r] * dlH);
  if (grads)
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r) {
        grads[3 * k + r] = gD[k][r];
        grads[12 + 3 * k + r] = gH[k][r];
      }
  if (aux) {
    aux[0] = C_D;
    aux[1] = C_H;
    aux[2] = A11;
    aux[3] = A12;
    aux[4] = A22;
    aux[5] = dlD;
    aux[6] = dlH;
    aux[7] = R(ok);
  }
}

// One scalar constraint of split mode on one tet: solve_hydrostatic
// (neohookean_solver.py:196-254) when hydro, solve_deviatoric (:144-194)
// otherwise.  C and w_sum = sum_i w_i |g_i|^2 go to info[0..1] (optional).
// When w_sum > EPSILON: dlambda = -(C + alpha~ lambda) / (w_sum + alpha~),
// times the relaxation, lambda += dlambda, corner i = dlambda w_i g_i, and 1
// (the cons_num increment) is returned; otherwise nothing changes and 0.
template <typename R>
inline int split_projection(const R* const p[4], const R w[4], const M3<R>& idm, bool hydro,
                            R alpha_tilde, R relax, R& lam, R* corner, R* info) {
  const M3<R> F = deformation_gradient(p[0], p[1], p[2], p[3], idm);
  R C;
  M3<R> gf;
  if (hydro) {
    C = det3(F) - R(1);
    gf = ddet_dF(F);
  } else {
    C = norm_sqr(F) - R(3);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) gf.m[i][j] = R(2) * F.m[i][j];
  }
  R g[4][3];
  vertex_grads(gf, idm, g);
  R w_sum = w[0] * dot3(g[0], g[0]);
  w_sum += w[1] * dot3(g[1], g[1]);
  w_sum += w[2] * dot3(g[2], g[2]);
  w_sum += w[3] * dot3(g[3], g[3]);
  if (info) {
    info[0] = C;
    info[1] = w_sum;
  }
  if (!(w_sum > R(kEpsilon))) return 0;
  R d_lambda = -(C + alpha_tilde * lam) / (w_sum + alpha_tilde);
  d_lambda *= relax;
  lam += d_lambda;
  for (int k = 0; k < 4; ++k)
    for (int r = 0; r < 3; ++r) corner[3 * k + r] = d_lambda * w[k] * g[k][r];
  return 1;
}

// compute_xpbd_residual for one tet (neohookean_solver.py:298-329):
// r = sqrt((C_D + alpha~_D lambda_D)^2 + (C_H + alpha~_H lambda_H)^2); C_D
// and C_H (the monitor kernels' quantities, :256-296) go to *C_D / *C_H.
// A non-zero `step` forms the square root faithfully rounded (kVarUlp).
template <typename R>
inline R residual_value(const R* const x[4], const M3<R>& idm, R aD, R aH, R lam_d, R lam_h,
                        R* C_D, R* C_H, int step = 0) {
  const M3<R> F = deformation_gradient(x[0], x[1], x[2], x[3], idm);
  *C_D = norm_sqr(F) - R(3);
  *C_H = det3(F) - R(1);
  const R rD = *C_D + aD * lam_d;
  const R rH = *C_H + aH * lam_h;
  return root(rD * rD + rH * rH, step);
}

// _chebyshev_update (neohookean_solver.py:131-137) on n scalars: with x the
// Jacobi iterate and curr the iterate it started from, x_new = gamma (x -
// curr) + curr, then x = prev + omega (x_new - prev).
template <typename R>
inline void chebyshev_blend_range(R* x, const R* curr, const R* prev, size_t lo, size_t hi,
                                  R omega, R gamma) {
  for (size_t i = lo; i < hi; ++i) {
    const R x_new = gamma * (x[i] - curr[i]) + curr[i];
    x[i] = prev[i] + omega * (x_new - prev[i]);
  }
}

struct SdfDesc {
  int kind;        // 0 plane, 1 sphere, 2 cube, 3 hang
  int fixed;
  double p[4];     // plane: point xyz | sphere: center xyz, radius | cube: center | hang: p[3] radius
  double dir[3];   // plane: unit normal (normalized in double at construction, sdf.py:66) |
                   // cube: edge lengths
  double vel[3];   // the obstacle's velocity, SdfModel.vel (sdf.py:9)
};

// One collider as the reference's f32 fields hold it (sdf.py:6-124).  Taichi's
// norm(eps) is sqrt(|q|^2 + eps), normalized(eps) is q / (|q| + eps).
template <typename R>
struct Sdf {
  int kind, fixed;
  R p[3], r, d[3], vel[3];
  std::vector<R> pts;  // hang: sphere centres, xyz per sphere
  explicit Sdf(const SdfDesc& s) : kind(s.kind), fixed(s.fixed) {
    for (int i = 0; i < 3; ++i) {
      p[i] = R(s.p[i]);
      d[i] = kind == 2 ? R(s.dir[i] * 0.5) : R(s.dir[i]);  // cube: half_size = size * 0.5
      vel[i] = R(s.vel[i]);
    }
    r = R(s.p[3]);
    if (kind == 3) fixed = 1;  // HangSdfModel(fixed=True), sdf.py:31
  }
  R dist(const R* x) const {
    if (kind == 3) {  // HangSdfModel.dist (sdf.py:38-43)
      R dist = R(1e5);
      for (size_t i = 0; i + 2 < pts.size(); i += 3) {
        const R q[3] = {x[0] - pts[i], x[1] - pts[i + 1], x[2] - pts[i + 2]};
        const R di = std::sqrt(dot3(q, q) + R(kEpsilon)) - r;
        dist = std::min(di, dist);
      }
      return dist;
    }
    R q[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
    if (kind == 0) return dot3(q, d);                   // PlaneSdf.dist (sdf.py:69-70)
    if (kind == 1) return std::sqrt(dot3(q, q)) - r;    // SphereSdf.dist (sdf.py:87-88)
    R e[3], o[3];                                        // CubeSdf.dist (sdf.py:107-110)
    for (int k = 0; k < 3; ++k) {
      e[k] = std::fabs(q[k]) - d[k];
      o[k] = std::max(e[k], R(0));
    }
    return std::sqrt(dot3(o, o)) + std::min(std::max(e[0], std::max(e[1], e[2])), R(0));
  }
  void normal(const R* x, R* n) const {
    n[0] = n[1] = n[2] = R(0);
    if (kind == 0) {  // PlaneSdf.normal (sdf.py:72-74)
      for (int i = 0; i < 3; ++i) n[i] = d[i];
      return;
    }
    if (kind == 3) {  // HangSdfModel.normal (sdf.py:45-53): towards sphere 0
      R dist = R(1e5);
      bool hit = false;
      for (size_t i = 0; i + 2 < pts.size(); i += 3) {
        const R q[3] = {x[0] - pts[i], x[1] - pts[i + 1], x[2] - pts[i + 2]};
        const R di = std::sqrt(dot3(q, q) + R(kEpsilon)) - r;
        if (di < dist) {
          dist = di;
          hit = true;
        }
      }
      if (hit) {
        const R q[3] = {x[0] - pts[0], x[1] - pts[1], x[2] - pts[2]};
        const R inv_len = R(1) / (std::sqrt(dot3(q, q)) + R(kEpsilon));
        for (int i = 0; i < 3; ++i) n[i] = inv_len * q[i];
      }
      return;
    }
    R q[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
    if (kind == 1) {  // SphereSdf.normal = normalized(EPSILON) (sdf.py:90-92)
      const R inv_len = R(1) / (std::sqrt(dot3(q, q)) + R(kEpsilon));
      for (int i = 0; i < 3; ++i) n[i] = inv_len * q[i];
      return;
    }
    int axis = 0;  // CubeSdf.normal (sdf.py:112-124)
    if (std::fabs(q[1]) > std::fabs(q[0])) axis = 1;
    if (std::fabs(q[2]) > std::fabs(q[axis])) axis = 2;
    n[axis] = q[axis] > R(0) ? R(1) : R(-1);
  }
  // SdfModel.check (sdf.py:13-27): returns inside, fills n.  diff_vel is the
  // obstacle's velocity minus the vertex's (zero unless the caller sets it).
  bool check(const R* x, const R* v, R* n) const {
    const R phi = dist(x);
    bool inside = false;
    n[0] = n[1] = n[2] = R(0);
    if (phi < R(0)) {
      normal(x, n);
      const R dv[3] = {vel[0] - v[0], vel[1] - v[1], vel[2] - v[2]};
      const R dotnv = dot3(n, dv);
      if (dotnv > R(0) || fixed) inside = true;
    }
    return inside;
  }
};

struct Params {
  double young, poisson, scale, offset[3];
  double frame_dt, dt;
  int substeps, iters;
  double hydro_relax, dev_relax;
  int cheb_enable, cheb_warmup;
  double cheb_gamma;
  int block_enable, monitor;
  double damping, gravity[3], total_mass;
  int variant;    // kVar* flags; 0 for the restatement itself
  uint64_t seed;  // of the variant's random orders and ulp steps
};

struct IterLog {
  int frame, substep, iter;
  double r_hat, rho_hat, omega, hydro_error, dev_error;
};

template <typename R>
class Solver {
 public:
  Solver(int nv, int nt, const float* rest, const int* tets, const Params& prm,
         const SdfDesc* sdfs, int n_sdf, int threads)
      : nv_(nv), nt_(nt), prm_(prm), threads_(threads) {
    tets_.assign(tets, tets + 4 * size_t(nt));
    for (int i = 0; i < n_sdf; ++i) sdfs_.emplace_back(sdfs[i]);
    x_.resize(3 * size_t(nv));
    for (size_t i = 0; i < 3 * size_t(nv); ++i) x_[i] = R(rest[i]);
    pred_.assign(3 * size_t(nv), R(0));
    v_.assign(3 * size_t(nv), R(0));
    m_.assign(nv, R(0));
    inv_m_.assign(nv, R(0));
    dx_.assign(3 * size_t(nv), R(0));
    cons_.assign(nv, 0u);
    prev_.assign(3 * size_t(nv), R(0));
    curr_.assign(3 * size_t(nv), R(0));
    V0_.assign(nt, R(0));
    inv_det_.assign(nt, R(0));
    inv_dm_.resize(nt);
    lam_h_.assign(nt, R(0));
    lam_d_.assign(nt, R(0));
    corner_a_.assign(12 * size_t(nt), R(0));
    corner_b_.assign(12 * size_t(nt), R(0));
    corner_ok_.assign(2 * size_t(nt), 0);
    per_tet_.assign(nt, R(0));
    per_tet2_.assign(nt, R(0));
    per_tet3_.assign(nt, R(0));
    variant_ = prm.variant;
    seed_ = prm.seed;
    sum_order_.resize(nt);
    for (int t = 0; t < nt; ++t) sum_order_[t] = t;
    if (variant_ & kVarOrder) {
      std::mt19937_64 rng(seed_ * 2654435761ull + 17);
      std::shuffle(sum_order_.begin(), sum_order_.end(), rng);
    }
    build_adjacency();
    set_lame();
    gravity_[0] = R(prm_.gravity[0]);
    gravity_[1] = R(prm_.gravity[1]);
    gravity_[2] = R(prm_.gravity[2]);
    initialize(R(prm_.scale), prm_.offset);
    total_mass_ = prm_.total_mass;
    compute_lumped_mass();
  }

  // ---- state access -------------------------------------------------------
  void get_x(double* out) const {
    for (size_t i = 0; i < x_.size(); ++i) out[i] = double(x_[i]);
  }
  void set_x(const double* in) {
    for (size_t i = 0; i < x_.size(); ++i) x_[i] = R(in[i]);
  }
  void get_v(double* out) const {
    for (size_t i = 0; i < v_.size(); ++i) out[i] = double(v_[i]);
  }
  void set_v(const double* in) {
    for (size_t i = 0; i < v_.size(); ++i) v_[i] = R(in[i]);
  }
  void get_pred(double* out) const {
    for (size_t i = 0; i < pred_.size(); ++i) out[i] = double(pred_[i]);
  }
  void get_mass(double* m, double* inv_m) const {
    for (int i = 0; i < nv_; ++i) {
      m[i] = double(m_[i]);
      inv_m[i] = double(inv_m_[i]);
    }
  }
  void get_tet_state(double* v0, double* inv_dm, double* lam_d, double* lam_h) const {
    for (int t = 0; t < nt_; ++t) {
      if (v0) v0[t] = double(V0_[t]);
      if (inv_dm)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) inv_dm[9 * size_t(t) + 3 * i + j] = double(inv_dm_[t].m[i][j]);
      if (lam_d) lam_d[t] = double(lam_d_[t]);
      if (lam_h) lam_h[t] = double(lam_h_[t]);
    }
  }
  void set_gravity(const double* g) {
    for (int i = 0; i < 3; ++i) gravity_[i] = R(g[i]);
  }
  void set_total_mass(double tm) { total_mass_ = tm; }
  double rho_hat() const { return rho_hat_; }
  int sdf_eval(int i, const double* x, const double* v, double* out) const {
    if (i < 0 || i >= int(sdfs_.size())) return -1;
    const R xr[3] = {R(x[0]), R(x[1]), R(x[2])}, vr[3] = {R(v[0]), R(v[1]), R(v[2])};
    R n[3];
    out[0] = double(sdfs_[i].dist(xr));
    sdfs_[i].normal(xr, n);
    for (int k = 0; k < 3; ++k) out[1 + k] = double(n[k]);
    return sdfs_[i].check(xr, vr, n) ? 1 : 0;
  }
  void set_sdf_points(int i, const double* xyz, int n) {
    if (i < 0 || i >= int(sdfs_.size())) return;
    sdfs_[i].pts.assign(3 * size_t(n), R(0));
    for (size_t k = 0; k < 3 * size_t(n); ++k) sdfs_[i].pts[k] = R(float(xyz[k]));
  }
  void set_rho_hat(double v) { rho_hat_ = v; }
  void set_pred(const double* in) {
    for (size_t i = 0; i < pred_.size(); ++i) pred_[i] = R(in[i]);
  }
  void set_lambda(const double* ld, const double* lh) {
    for (int t = 0; t < nt_; ++t) {
      lam_d_[t] = R(ld[t]);
      lam_h_[t] = R(lh[t]);
    }
  }
  // Residual sums on the current pred_ and multipliers with time step dt:
  // out = sum r, sum |C_H|, sum |C_D| (the quantities of
  // neohookean_solver.py:298-329 and :256-296).
  void residual_now(double dt, double* out) {
    R hs, ds;
    const R rs = residual(R(dt), &hs, &ds);
    out[0] = double(rs);
    out[1] = double(hs);
    out[2] = double(ds);
  }
  // Per-iteration trace of the Chebyshev sequence (first substep after it is
  // enabled): per iteration the iterate it started from (curr), the Jacobi
  // iterate, the iterate after the blend, and omega.
  void set_trace(int on) {
    trace_on_ = on != 0;
    trace_.clear();
  }
  const std::vector<double>& trace() const { return trace_; }
  double time() const { return time_; }
  int frame() const { return frame_; }
  const std::vector<IterLog>& log() const { return log_; }
  void clear_log() { log_.clear(); }

  // ---- setup kernels --------------------------------------------------------

  // calc_first_lame / calc_second_lame (neohookean_solver.py:7-12), evaluated
  // in Python doubles and frozen into the kernels as f32 constants.
  void set_lame() {
    const double E = prm_.young, nu = prm_.poisson;
    const double mu = E / (2.0 * (1.0 + nu));
    const double la = 2.0 * nu / (1.0 - 2.0 * nu) * mu;
    lame1_ = R(la);
    lame2_ = R(mu);
  }

  // initialize (neohookean_solver.py:97-107) with calc_tet_volume
  // (xpbd_base.py:6-8).
  void initialize(R scale, const double* offset) {
    const R off[3] = {R(offset[0]), R(offset[1]), R(offset[2])};
    for (int v = 0; v < nv_; ++v)
      for (int k = 0; k < 3; ++k) x_[3 * v + k] = x_[3 * v + k] * scale + off[k];
    for (int t = 0; t < nt_; ++t) {
      const R* p[4];
      for (int k = 0; k < 4; ++k) p[k] = &x_[3 * size_t(tets_[4 * size_t(t) + k])];
      V0_[t] = tet_volume(p[0], p[1], p[2], p[3]);
      M3<R> dm;
      for (int r = 0; r < 3; ++r) {
        dm.m[r][0] = p[1][r] - p[0][r];
        dm.m[r][1] = p[2][r] - p[0][r];
        dm.m[r][2] = p[3][r] - p[0][r];
      }
      const int step = (variant_ & kVarUlp) ? ulp_step(seed_, uint64_t(t), 0x1d) + 2 : 0;
      inv_dm_[t] = inv3(dm, step);
      inv_det_[t] = recip(det3(dm), step);
    }
  }

  static R tet_volume(const R* p0, const R* p1, const R* p2, const R* p3) {
    const R a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const R b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const R c[3] = {p3[0] - p0[0], p3[1] - p0[1], p3[2] - p0[2]};
    R ab[3];
    cross3(a, b, ab);
    return dot3(ab, c) / R(6);
  }

  // _compute_lumped_mass_kernel + _apply_fixed_vertices_for_sdf
  // (xpbd_base.py:89-137).
  void compute_lumped_mass() {
    for (int t = 0; t < nt_; ++t) {
      const R* p[4];
      for (int k = 0; k < 4; ++k) p[k] = &x_[3 * size_t(tets_[4 * size_t(t) + k])];
      per_tet_[t] = tet_volume(p[0], p[1], p[2], p[3]);
    }
    R total_volume = R(0);
    for (int v = 0; v < nv_; ++v) m_[v] = R(0);
    for (int i = 0; i < nt_; ++i) {
      const int t = sum_order_[i];
      const R quarter = per_tet_[t] * R(0.25);
      for (int k = 0; k < 4; ++k) m_[tets_[4 * size_t(t) + k]] += quarter;
      total_volume += per_tet_[t];
    }
    const R mass_scale = R(total_mass_) / total_volume;
    for (int v = 0; v < nv_; ++v) {
      m_[v] *= mass_scale;
      inv_m_[v] = m_[v] > R(0) ? R(1) / m_[v] : R(0);
    }
    for (const auto& s : sdfs_) {
      for (int v = 0; v < nv_; ++v) {
        if (inv_m_[v] > R(0)) {
          R n[3];
          const bool inside = s.check(&x_[3 * size_t(v)], &v_[3 * size_t(v)], n);
          if (s.fixed && inside) inv_m_[v] = R(0);
        }
      }
    }
  }

  // ---- per-substep kernels -------------------------------------------------

  // apply_external_forces (xpbd_base.py:139-146)
  void apply_external_forces(R dt) {
#pragma omp parallel for num_threads(threads_) schedule(static)
    for (int v = 0; v < nv_; ++v) {
      if (inv_m_[v] > R(0))
        for (int k = 0; k < 3; ++k) v_[3 * v + k] += gravity_[k] * dt;
      for (int k = 0; k < 3; ++k) pred_[3 * v + k] = x_[3 * v + k] + v_[3 * v + k] * dt;
    }
  }

  // _solve_sdf_collision_one / solve_sdf_collision (xpbd_base.py:162-179)
  void solve_sdf_collision() {
    for (const auto& s : sdfs_) {
#pragma omp parallel for num_threads(threads_) schedule(static)
      for (int v = 0; v < nv_; ++v) {
        if (inv_m_[v] > R(0)) {
          R n[3];
          R* x = &pred_[3 * size_t(v)];
          if (s.check(x, &v_[3 * size_t(v)], n)) {
            const R depth = -s.dist(x);
            if (depth > R(0))
              for (int k = 0; k < 3; ++k) x[k] += n[k] * depth;
          }
        }
      }
    }
  }

  // advance (xpbd_base.py:148-160).  ti.exp of an f32 argument.
  void advance(R dt, R damping) {
    const R decay = std::exp(-dt * damping);
#pragma omp parallel for num_threads(threads_) schedule(static)
    for (int v = 0; v < nv_; ++v) {
      if (inv_m_[v] <= R(0)) {
        for (int k = 0; k < 3; ++k) pred_[3 * v + k] = x_[3 * v + k];
      } else {
        for (int k = 0; k < 3; ++k) {
          const R raw_v = (pred_[3 * v + k] - x_[3 * v + k]) / dt;
          v_[3 * v + k] = raw_v * decay;
          x_[3 * v + k] = pred_[3 * v + k];
        }
      }
    }
  }

  // ---- Jacobi iteration kernels ----------------------------------------------

  // Per-tet block projection of solve_neohookean_block
  // (neohookean_solver.py:358-451): the 4 corner corrections go into `corner`
  // (12 values) and the multipliers are updated.
  void project_block_tet(int t, R dt, R* corner) {
    const int* tv = &tets_[4 * size_t(t)];
    const R* p[4];
    R w[4];
    for (int k = 0; k < 4; ++k) {
      p[k] = &pred_[3 * size_t(tv[k])];
      w[k] = inv_m_[tv[k]];
    }
    R aD, aH;
    compliance(lame1_, lame2_, V0_[t], inv_det_[t], dt, variant_, &aD, &aH);
    block_projection(p, w, inv_dm_[t], aD, aH, lam_d_[t], lam_h_[t], corner, (R*)nullptr,
                     (R*)nullptr);
  }

  // One scalar constraint of split mode: solve_hydrostatic
  // (neohookean_solver.py:196-254) when hydro, solve_deviatoric (:144-194)
  // otherwise.  Returns whether w_sum > EPSILON (the cons_num increment).
  int project_split_tet(int t, R dt, bool hydro, R* corner) {
    const int* tv = &tets_[4 * size_t(t)];
    const R* p[4];
    R w[4];
    for (int k = 0; k < 4; ++k) {
      p[k] = &pred_[3 * size_t(tv[k])];
      w[k] = inv_m_[tv[k]];
    }
    R aD, aH;
    compliance(lame1_, lame2_, V0_[t], inv_det_[t], dt, variant_, &aD, &aH);
    return split_projection(p, w, inv_dm_[t], hydro, hydro ? aH : aD,
                            R(hydro ? prm_.hydro_relax : prm_.dev_relax),
                            hydro ? lam_h_[t] : lam_d_[t], corner, (R*)nullptr);
  }

  // pre_solve + projection + post_solve (neohookean_solver.py:109-119, 482-490).
  void jacobi_iteration(R dt) {
    if (prm_.block_enable) {
#pragma omp parallel for num_threads(threads_) schedule(static)
      for (int t = 0; t < nt_; ++t) project_block_tet(t, dt, &corner_a_[12 * size_t(t)]);
    } else {
      // solve_hydrostatic runs over all tets before solve_deviatoric.
#pragma omp parallel for num_threads(threads_) schedule(static)
      for (int t = 0; t < nt_; ++t) {
        corner_ok_[2 * size_t(t)] = project_split_tet(t, dt, true, &corner_a_[12 * size_t(t)]);
        corner_ok_[2 * size_t(t) + 1] = project_split_tet(t, dt, false, &corner_b_[12 * size_t(t)]);
      }
    }
    // delta_x / cons_num accumulated in ascending tet order, then post_solve.
#pragma omp parallel for num_threads(threads_) schedule(static)
    for (int v = 0; v < nv_; ++v) {
      R d[3] = {R(0), R(0), R(0)};
      uint32_t cnt = 0;
      const int b = adj_off_[v], e = adj_off_[v + 1];
      if (prm_.block_enable) {
        for (int a = b; a < e; ++a) {
          const R* c = &corner_a_[3 * size_t(adj_[a])];
          d[0] += c[0];
          d[1] += c[1];
          d[2] += c[2];
          ++cnt;
        }
      } else {
        for (int pass = 0; pass < 2; ++pass) {
          const std::vector<R>& buf = pass == 0 ? corner_a_ : corner_b_;
          for (int a = b; a < e; ++a) {
            const int slot = adj_[a];
            if (!corner_ok_[2 * size_t(slot / 4) + pass]) continue;
            const R* c = &buf[3 * size_t(slot)];
            d[0] += c[0];
            d[1] += c[1];
            d[2] += c[2];
            ++cnt;
          }
        }
      }
      for (int k = 0; k < 3; ++k) dx_[3 * v + k] = d[k];
      cons_[v] = cnt;
      if (cnt > 0) {
        if (variant_ & kVarUlp) {
          const R ic = recip(R(cnt), ulp_step(seed_, uint64_t(v), 2 * iter_count_ + 1) + 2);
          for (int k = 0; k < 3; ++k) pred_[3 * v + k] += d[k] * ic;
        } else {
          for (int k = 0; k < 3; ++k) pred_[3 * v + k] += d[k] / R(cnt);
        }
      }
    }
    ++iter_count_;
  }

  // compute_xpbd_residual (neohookean_solver.py:298-329), plus the monitor
  // kernels compute_hydrostatic_error / compute_deviatoric_error (:256-296)
  // evaluated on the same positions.  Returns sum r (in R) and fills the sums
  // of |C_H| and |C_D|.
  R residual(R dt, R* hydro_sum, R* dev_sum) {
#pragma omp parallel for num_threads(threads_) schedule(static)
    for (int t = 0; t < nt_; ++t) {
      const int* tv = &tets_[4 * size_t(t)];
      const R* x[4] = {&pred_[3 * size_t(tv[0])], &pred_[3 * size_t(tv[1])],
                       &pred_[3 * size_t(tv[2])], &pred_[3 * size_t(tv[3])]};
      R aD, aH, C_D, C_H;
      compliance(lame1_, lame2_, V0_[t], inv_det_[t], dt, variant_, &aD, &aH);
      const int step =
          (variant_ & kVarUlp) ? ulp_step(seed_, uint64_t(t), 2 * iter_count_ + 2) + 2 : 0;
      per_tet_[t] = residual_value(x, inv_dm_[t], aD, aH, lam_d_[t], lam_h_[t], &C_D, &C_H, step);
      per_tet2_[t] = C_H;
      per_tet3_[t] = C_D;
    }
    R s = R(0), sh = R(0), sd = R(0);
    for (int i = 0; i < nt_; ++i) {
      const int t = sum_order_[i];
      s += per_tet_[t];
      sh += std::fabs(per_tet2_[t]);
      sd += std::fabs(per_tet3_[t]);
    }
    *hydro_sum = sh;
    *dev_sum = sd;
    return s;
  }

  // solve_step (neohookean_solver.py:453-560)
  void solve_step(double sub_dt) {
    const R dt = R(sub_dt);
    std::fill(lam_h_.begin(), lam_h_.end(), R(0));
    std::fill(lam_d_.begin(), lam_d_.end(), R(0));
    bool have_prev_r = false;
    double prev_r_hat = 0.0;
    solve_sdf_collision();
    const bool use_cheb = prm_.cheb_enable != 0;
    double omega = 1.0;
    if (use_cheb) prev_ = pred_;  // _save_curr_to_prev
    const R gamma = R(prm_.cheb_gamma);
    const bool tracing = trace_on_ && trace_.empty();
    for (int it = 0; it < prm_.iters; ++it) {
      if (use_cheb) curr_ = pred_;  // _save_curr_to_savebuf
      if (tracing) trace_.insert(trace_.end(), pred_.begin(), pred_.end());
      jacobi_iteration(dt);
      if (tracing) trace_.insert(trace_.end(), pred_.begin(), pred_.end());
      double r_hat = 0.0;
      R hs = R(0), ds = R(0);
      const bool need_res = use_cheb || prm_.monitor;
      R rs = R(0);
      if (need_res) rs = residual(dt, &hs, &ds);
      if (use_cheb) {
        // neohookean_solver.py:494-504
        const int cnt = nt_;
        r_hat = cnt > 0 ? double(rs) / double(cnt) : 0.0;
        if (have_prev_r && prev_r_hat > kEpsilon) {
          double rho_inst = r_hat / (prev_r_hat + kEpsilon);
          rho_inst = std::max(kRhoMin, std::min(kRhoMax, rho_inst));
          rho_hat_ = (1.0 - kRhoBeta) * rho_hat_ + kRhoBeta * rho_inst;
        }
        prev_r_hat = r_hat;
        have_prev_r = true;
        // neohookean_solver.py:537-546
        const double rho = rho_hat_;
        const double rho2 = rho * rho;
        if (it < prm_.cheb_warmup)
          omega = 1.0;
        else if (it == prm_.cheb_warmup)
          omega = 2.0 / (2.0 - rho2);
        else
          omega = 4.0 / (4.0 - rho2 * omega);
        chebyshev_update(R(omega), gamma);
        prev_ = curr_;  // _rotate_prev_with_save
      }
      if (tracing) {
        trace_.insert(trace_.end(), pred_.begin(), pred_.end());
        trace_.push_back(use_cheb ? omega : 1.0);
      }
      IterLog rec;
      rec.frame = frame_;
      rec.substep = substep_;
      rec.iter = it;
      rec.r_hat = r_hat;
      rec.rho_hat = rho_hat_;
      rec.omega = use_cheb ? omega : 1.0;
      const double nc = nt_ > 0 ? double(nt_) : 1.0;
      rec.hydro_error = prm_.monitor && nt_ > 0 ? double(hs) / nc : 0.0;
      rec.dev_error = prm_.monitor && nt_ > 0 ? double(ds) / nc : 0.0;
      if (log_enabled_) log_.push_back(rec);
    }
  }

  // _chebyshev_update (neohookean_solver.py:131-137)
  void chebyshev_update(R omega, R gamma) {
    const long n = long(pred_.size());
#pragma omp parallel for num_threads(threads_) schedule(static)
    for (long i = 0; i < n; ++i)
      chebyshev_blend_range(pred_.data(), curr_.data(), prev_.data(), size_t(i), size_t(i) + 1,
                            omega, gamma);
  }

  // solve (neohookean_solver.py:562-587)
  void solve_frame() {
    double frame_time_left = prm_.frame_dt;
    substep_ = 0;
    while (frame_time_left > 0.0) {
      const double dt0 = std::min(prm_.dt, frame_time_left);
      frame_time_left -= dt0;
      const double sub_dt = dt0 / prm_.substeps;
      for (int s = 0; s < prm_.substeps; ++s) {
        substep_ = s;
        substep(sub_dt);
      }
      time_ += dt0;
    }
    frame_ += 1;
  }

  // apply_external_forces -> solve_step -> advance (neohookean_solver.py:577-579)
  void substep(double sub_dt) {
    apply_external_forces(R(sub_dt));
    solve_step(sub_dt);
    advance(R(sub_dt), R(prm_.damping));
  }

  bool log_enabled_ = true;

 private:
  static constexpr double kRhoMin = 0.00001, kRhoMax = 0.99999, kRhoBeta = 0.2;

  void build_adjacency() {
    adj_off_.assign(nv_ + 1, 0);
    for (size_t i = 0; i < tets_.size(); ++i) adj_off_[tets_[i] + 1]++;
    for (int v = 0; v < nv_; ++v) adj_off_[v + 1] += adj_off_[v];
    adj_.resize(tets_.size());
    std::vector<int> fill(adj_off_.begin(), adj_off_.end() - 1);
    for (int t = 0; t < nt_; ++t)  // ascending t -> each list sorted by tet
      for (int k = 0; k < 4; ++k) adj_[fill[tets_[4 * size_t(t) + k]]++] = 4 * t + k;
    if (variant_ & kVarOrder) {  // each vertex's corrections in a random order
      std::mt19937_64 rng(seed_ * 0x9E3779B97F4A7C15ull + 5);
      for (int v = 0; v < nv_; ++v)
        std::shuffle(adj_.begin() + adj_off_[v], adj_.begin() + adj_off_[v + 1], rng);
    }
  }

  int nv_, nt_;
  Params prm_;
  int threads_;
  std::vector<int> tets_;
  std::vector<Sdf<R>> sdfs_;
  std::vector<R> x_, pred_, v_, m_, inv_m_, dx_, prev_, curr_;
  std::vector<uint32_t> cons_;
  std::vector<R> V0_, inv_det_, lam_h_, lam_d_;
  std::vector<M3<R>> inv_dm_;
  int variant_ = 0;
  uint64_t seed_ = 0;
  uint64_t iter_count_ = 0;     // Jacobi iterations run (keys the ulp steps)
  std::vector<int> sum_order_;  // order of the tet sums (a permutation under kVarOrder)
  bool trace_on_ = false;
  std::vector<double> trace_;
  std::vector<R> corner_a_, corner_b_, per_tet_, per_tet2_, per_tet3_;
  std::vector<int> corner_ok_;
  std::vector<int> adj_off_, adj_;
  R lame1_, lame2_;
  R gravity_[3];
  double total_mass_ = 1.0;
  double rho_hat_ = 0.999;  // neohookean_solver.py:65, persists across substeps
  double time_ = 0.0;
  int frame_ = 0, substep_ = 0;
  std::vector<IterLog> log_;
};

struct Handle {
  int precision;  // 32 or 64
  Solver<float>* s32 = nullptr;
  Solver<double>* s64 = nullptr;
};

template <typename F>
auto dispatch(Handle* h, F&& f) {
  if (h->precision == 32) return f(*h->s32);
  return f(*h->s64);
}

Params params_from(const double* p) {
  Params q;
  q.young = p[0];
  q.poisson = p[1];
  q.scale = p[2];
  q.offset[0] = p[3];
  q.offset[1] = p[4];
  q.offset[2] = p[5];
  q.frame_dt = p[6];
  q.dt = p[7];
  q.substeps = int(p[8]);
  q.iters = int(p[9]);
  q.hydro_relax = p[10];
  q.dev_relax = p[11];
  q.cheb_enable = int(p[12]);
  q.cheb_warmup = int(p[13]);
  q.cheb_gamma = p[14];
  q.block_enable = int(p[15]);
  q.monitor = int(p[16]);
  q.damping = p[17];
  q.gravity[0] = p[18];
  q.gravity[1] = p[19];
  q.gravity[2] = p[20];
  q.total_mass = p[21];
  q.variant = int(p[22]);
  q.seed = uint64_t(p[23]);
  return q;
}

}  // namespace

extern "C" {

// params: 24 doubles, in the order read by params_from().
// sdf_kind[n] (0 plane, 1 sphere, 2 cube, 3 hang), sdf_fixed[n], sdf_p[4n], sdf_dir[3n],
// sdf_vel[3n]; a hang collider's centres follow with oracle_set_sdf_points.
void* oracle_create(int precision, int nv, int nt, const float* rest, const int* tets,
                    const double* params, int n_sdf, const int* sdf_kind, const int* sdf_fixed,
                    const double* sdf_p, const double* sdf_dir, const double* sdf_vel,
                    int threads) {
  std::vector<SdfDesc> sd(n_sdf);
  for (int i = 0; i < n_sdf; ++i) {
    sd[i].kind = sdf_kind[i];
    sd[i].fixed = sdf_fixed[i];
    for (int k = 0; k < 4; ++k) sd[i].p[k] = sdf_p[4 * i + k];
    for (int k = 0; k < 3; ++k) sd[i].dir[k] = sdf_dir[3 * i + k];
    for (int k = 0; k < 3; ++k) sd[i].vel[k] = sdf_vel[3 * i + k];
  }
  const Params prm = params_from(params);
  Handle* h = new Handle;
  h->precision = precision;
  if (threads < 1) threads = 1;
  if (precision == 32)
    h->s32 = new Solver<float>(nv, nt, rest, tets, prm, sd.data(), n_sdf, threads);
  else
    h->s64 = new Solver<double>(nv, nt, rest, tets, prm, sd.data(), n_sdf, threads);
  return h;
}

void oracle_destroy(void* hp) {
  Handle* h = static_cast<Handle*>(hp);
  delete h->s32;
  delete h->s64;
  delete h;
}

void oracle_frame(void* hp, int n) {
  Handle* h = static_cast<Handle*>(hp);
  for (int i = 0; i < n; ++i) dispatch(h, [](auto& s) { s.solve_frame(); return 0; });
}

void oracle_substep(void* hp, double sub_dt) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.substep(sub_dt); return 0; });
}

void oracle_get_x(void* hp, double* out) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.get_x(out); return 0; });
}
void oracle_set_x(void* hp, const double* in) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.set_x(in); return 0; });
}
void oracle_get_v(void* hp, double* out) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.get_v(out); return 0; });
}
void oracle_set_v(void* hp, const double* in) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.set_v(in); return 0; });
}
void oracle_get_pred(void* hp, double* out) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.get_pred(out); return 0; });
}
void oracle_get_mass(void* hp, double* m, double* inv_m) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.get_mass(m, inv_m); return 0; });
}
void oracle_get_tet_state(void* hp, double* v0, double* inv_dm, double* lam_d, double* lam_h) {
  dispatch(static_cast<Handle*>(hp),
           [&](auto& s) { s.get_tet_state(v0, inv_dm, lam_d, lam_h); return 0; });
}
void oracle_set_gravity(void* hp, const double* g) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.set_gravity(g); return 0; });
}
void oracle_compute_lumped_mass(void* hp, double total_mass) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) {
    s.set_total_mass(total_mass);
    s.compute_lumped_mass();
    return 0;
  });
}
double oracle_rho_hat(void* hp) {
  return dispatch(static_cast<Handle*>(hp), [](auto& s) { return s.rho_hat(); });
}
// dist (out[0]) and normal (out[1..3]) of collider i at x, and whether a
// vertex with velocity v counts as inside (returned)
int oracle_sdf_eval(void* hp, int i, const double* x, const double* v, double* out) {
  return dispatch(static_cast<Handle*>(hp), [&](auto& s) { return s.sdf_eval(i, x, v, out); });
}
// sphere centres (xyz per sphere) of hang collider i, as f32 values
void oracle_set_sdf_points(void* hp, int i, const double* xyz, int n) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.set_sdf_points(i, xyz, n); return 0; });
}
// restores a saved rho_hat (a run continued from a stored state)
void oracle_set_rho_hat(void* hp, double v) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.set_rho_hat(v); return 0; });
}
void oracle_set_log(void* hp, int enabled) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.log_enabled_ = enabled != 0; return 0; });
}
int oracle_log_size(void* hp) {
  return dispatch(static_cast<Handle*>(hp), [](auto& s) { return int(s.log().size()); });
}
// rows of 8 doubles: frame, substep, iter, r_hat, rho_hat, omega, hydro_error, dev_error
void oracle_get_log(void* hp, double* out) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) {
    size_t i = 0;
    for (const auto& r : s.log()) {
      out[8 * i + 0] = r.frame;
      out[8 * i + 1] = r.substep;
      out[8 * i + 2] = r.iter;
      out[8 * i + 3] = r.r_hat;
      out[8 * i + 4] = r.rho_hat;
      out[8 * i + 5] = r.omega;
      out[8 * i + 6] = r.hydro_error;
      out[8 * i + 7] = r.dev_error;
      ++i;
    }
    return 0;
  });
}
void oracle_clear_log(void* hp) {
  dispatch(static_cast<Handle*>(hp), [](auto& s) { s.clear_log(); return 0; });
}

// ---- state access for the pins ------------------------------------------------

void oracle_set_pred(void* hp, const double* in) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.set_pred(in); return 0; });
}
void oracle_set_lambda(void* hp, const double* lam_d, const double* lam_h) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.set_lambda(lam_d, lam_h); return 0; });
}
// sum r, sum |C_H|, sum |C_D| over the tets at the current pred_x and
// multipliers, with time step dt
void oracle_residual_now(void* hp, double dt, double* out) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.residual_now(dt, out); return 0; });
}
// Trace of the first substep's Chebyshev sequence after enabling: per
// iteration curr (3V), Jacobi iterate (3V), blended iterate (3V), omega.
void oracle_set_trace(void* hp, int on) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) { s.set_trace(on); return 0; });
}
int oracle_trace_size(void* hp) {
  return dispatch(static_cast<Handle*>(hp), [](auto& s) { return int(s.trace().size()); });
}
void oracle_get_trace(void* hp, double* out) {
  dispatch(static_cast<Handle*>(hp), [&](auto& s) {
    std::copy(s.trace().begin(), s.trace().end(), out);
    return 0;
  });
}

// ---- single-tet entry points (for closed-form and finite-difference pins) ----
// They call the same per-tet functions as the solver above.

}  // extern "C"

namespace {
template <typename R>
M3<R> load_m3(const double* a) {
  M3<R> m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.m[i][j] = R(a[3 * i + j]);
  return m;
}

template <typename R>
void block_tet(const double* p, const double* w, const double* inv_dm, double v0, double lame1,
               double lame2, double dt, double* lam, double* corner, double* grads, double* aux) {
  R pr[12], wr[4], c[12], g[24], a[8];
  for (int i = 0; i < 12; ++i) pr[i] = R(p[i]);
  for (int i = 0; i < 4; ++i) wr[i] = R(w[i]);
  const R* pp[4] = {&pr[0], &pr[3], &pr[6], &pr[9]};
  const M3<R> idm = load_m3<R>(inv_dm);
  R aD, aH;
  compliance(R(lame1), R(lame2), R(v0), R(0), R(dt), 0, &aD, &aH);
  R ld = R(lam[0]), lh = R(lam[1]);
  block_projection(pp, wr, idm, aD, aH, ld, lh, c, g, a);
  lam[0] = double(ld);
  lam[1] = double(lh);
  for (int i = 0; i < 12; ++i) corner[i] = double(c[i]);
  if (grads)
    for (int i = 0; i < 24; ++i) grads[i] = double(g[i]);
  if (aux)
    for (int i = 0; i < 8; ++i) aux[i] = double(a[i]);
}
}  // namespace

extern "C" {

// Block projection of one tet in isolation (precision 32 or 64).  p: 4x3
// positions, w: 4 inverse masses, inv_dm: 3x3 row-major, lam: [lambda_D,
// lambda_H] in/out, corner: 4x3 corrections, grads (24, optional): g_D then
// g_H per vertex; aux (8): C_D, C_H, A11, A12, A22, dlamD, dlamH, ok.
void oracle_block_tet(int precision, const double* p, const double* w, const double* inv_dm,
                      double v0, double lame1, double lame2, double dt, double* lam,
                      double* corner, double* grads, double* aux) {
  if (precision == 32)
    block_tet<float>(p, w, inv_dm, v0, lame1, lame2, dt, lam, corner, grads, aux);
  else
    block_tet<double>(p, w, inv_dm, v0, lame1, lame2, dt, lam, corner, grads, aux);
}

void oracle_block_tet64(const double* p, const double* w, const double* inv_dm, double v0,
                        double lame1, double lame2, double dt, double* lam, double* corner,
                        double* grads, double* aux) {
  block_tet<double>(p, w, inv_dm, v0, lame1, lame2, dt, lam, corner, grads, aux);
}

// One split-mode constraint of one tet in float64 (hydro != 0: C_H, else
// C_D) with compliance 1 / (lame V0) / dt^2 and relaxation `relax`.  lam
// in/out; corner: 4x3; info (2): C, w_sum.  Returns the cons_num increment.
int oracle_split_tet64(const double* p, const double* w, const double* inv_dm, double v0,
                       double lame, double dt, double relax, int hydro, double* lam,
                       double* corner, double* info) {
  const double* pp[4] = {&p[0], &p[3], &p[6], &p[9]};
  const M3<double> idm = load_m3<double>(inv_dm);
  double aD, aH;
  compliance(lame, lame, v0, 0.0, dt, 0, &aD, &aH);
  std::fill(corner, corner + 12, 0.0);
  return split_projection(pp, w, idm, hydro != 0, hydro ? aH : aD, relax, *lam, corner, info);
}

// Residual of one tet in float64: out (3) = r, C_D, C_H.
void oracle_residual_tet64(const double* p, const double* inv_dm, double v0, double lame1,
                           double lame2, double dt, double lam_d, double lam_h, double* out) {
  const double* pp[4] = {&p[0], &p[3], &p[6], &p[9]};
  const M3<double> idm = load_m3<double>(inv_dm);
  double aD, aH;
  compliance(lame1, lame2, v0, 0.0, dt, 0, &aD, &aH);
  out[0] = residual_value(pp, idm, aD, aH, lam_d, lam_h, &out[1], &out[2]);
}

// The Chebyshev blend of n values in float64 (_chebyshev_update).
void oracle_cheb_blend64(double* x, const double* curr, const double* prev, int n, double omega,
                         double gamma) {
  chebyshev_blend_range(x, curr, prev, 0, size_t(n), omega, gamma);
}

int oracle_solve2x2_64(double a00, double a01, double a10, double a11, double b0, double b1,
                       double eps, double* x) {
  return solve2x2_pivot_lu(a00, a01, a10, a11, b0, b1, eps, &x[0], &x[1]);
}

// inverse and determinant of a row-major 3x3 as the oracle evaluates them
void oracle_inv3_64(const double* a, double* out, double* det) {
  M3<double> m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.m[i][j] = a[3 * i + j];
  const M3<double> r = inv3(m);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = r.m[i][j];
  *det = det3(m);
}

}  // extern "C"
