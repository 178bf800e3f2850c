// Device-side data structures and per-tet math shared by the gfx950 kernels.
//
// Layout in HBM (one scene):
//   vertex streams, float4 per vertex (16 B, one 64 B line holds 4 vertices):
//     pos   pred_x.xyz | inv_m          gathered by every tet, so the inverse mass
//                                       rides in the same 16 B record
//     xs    x.xyz      | inv_m
//     vel   v.xyz      | 0
//     prev  Chebyshev y_prev
//     xj    Jacobi-averaged positions of the current iteration
//     dacc  delta_x.xyz | cons_num      (float atomics)
//   tet streams (SoA over tets, coalesced per wavefront):
//     tet   int4 vertex ids
//     dm0   float4 (m00 m01 m02 m10) of inv_Dm, row-major
//     dm1   float4 (m11 m12 m20 m21)
//     dm2   float2 (m22, V0)
//     lam   float2 (lambda_D, lambda_H)
//   deterministic mode: corner float4[4T] per tet-corner corrections and a
//   vertex -> corner CSR sorted by tet.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/xpbd_abi.h"

namespace xpbd {

// Pointers held in the device-side scene table point into global memory.  The
// kernels read them from that table, so the compiler cannot infer their
// address space and would emit flat loads, which count against both the
// vector-memory and the LDS wait counters: every LDS wait would then also wait
// for the global loads in flight.  The device pass therefore types them as
// address space 1 (global); the host pass sees plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define XPBD_GLOBAL_AS __attribute__((address_space(1)))
#else
#define XPBD_GLOBAL_AS
#endif
template <class T>
using gptr = XPBD_GLOBAL_AS T*;

struct SdfDev {
  int kind;   // XPBD_SDF_PLANE / SPHERE / CUBE / HANG
  int fixed;
  float p[3];  // plane point | sphere or cube centre
  float r;     // sphere radius | hang sphere radius
  float d[3];  // plane unit normal | cube half size
  int n_pts;   // hang: number of spheres
  gptr<const float4> pts;  // hang: sphere centres (xyz)
  float vel[3];  // the obstacle's velocity (SdfModel.vel, sdf.py:9)
  int pad;
};

// Per-scene scalar state that persists across launches (rho_hat persists over
// the solver's life, neohookean_solver.py:65; the rest is per substep).
struct SceneCtrl {
  double rho_hat;
  double prev_r_hat;
  double omega;
  double r_hat;
  int have_prev;
  int status;
  int pad[2];
};

struct SceneDev {
  // vertex streams (n_verts)
  gptr<float4> pos;
  gptr<float4> xs;
  gptr<float4> vel;
  gptr<float4> prev;
  gptr<float4> xj;
  gptr<float4> dacc;
  gptr<const float4> rest;  // rest positions (after scale/offset), w unused
  gptr<const float> valence;
  // flat layout: one thread per tet
  gptr<const int> adj_off;
  gptr<const int> adj;
  gptr<float4> corner;
  gptr<const int4> tet;
  gptr<const float4> dm0;
  gptr<const float4> dm1;
  gptr<const float2> dm2;
  // patch layout: one workgroup per patch of <= 512 tets / <= 256 vertices
  gptr<const int> p_tet_begin;   // n_patches + 1
  gptr<const int> p_vert_begin;  // n_patches + 1: range of the patch's vertex records
  gptr<const int> p_vert_ids;    // per patch vertex: global vertex id
  gptr<const int2> p_vert_sum;   // per patch vertex: (corner base | corner count << 16 |
                            //   owner << 31, slot); the owner patch writes the vertex
  gptr<float4> ybuf[3];          // rotating Chebyshev iterates y^k (ybuf[0] is pos)
  gptr<const uint2> tet_local;   // per tet: 4 x u8 local vertex ids, 4 x u8 corner ranks
  gptr<const int> vslot_off;     // n_verts + 1: each vertex's slots, contiguous, ascending patch
  gptr<float4> slot_acc;         // per slot: partial delta_x (xyz) and constraint count (w)
  gptr<float2> lam;
  gptr<double> partial;  // per tet workgroup / patch: sum r, sum |C_H|, sum |C_D| (SoA,
                         // component c at [c * n_partials])
  gptr<SceneCtrl> ctrl;   // two slots, see the controller update in the kernels
  gptr<double> log;      // 6 doubles per iteration: iter, r_hat, rho_hat, omega, hydro, dev
  // persistent substep: scene-barrier arrivals, workgroups done, error word
  gptr<unsigned> sync;
  // persistent substep: residual partials posted by value, kMailboxWords per
  // residual group (2 iterations' parity x 3 components, SoA over groups)
  gptr<double> mailbox;
  int n_verts;
  int n_tets;
  int n_partials;
  int tet_block;    // flat layout: threads per tet workgroup
  int n_patches;
  int layout;       // 0 flat, 1 patch
  int fused;        // patch layout with Chebyshev: blend and average folded into the tet passes
  // parameters as the reference kernels see them (f32 constants / arguments)
  float lame1, lame2;
  float dt;
  float gamma;
  float hydro_relax, dev_relax;
  float decay;      // exp(-dt * damping), evaluated once on the host in f32
  // compliance scales: alpha~ = 1 / (lame V0 dt^2) = alpha_k / det(Dm), with
  // alpha_k = 6 / (lame dt^2) (deviatoric: lame2, hydrostatic: lame1)
  float alpha_kd, alpha_kh;
  float g[3];
  int block_mode;   // block_neohookean_enable
  int cheb;         // cheb_enable
  int monitor;      // monitor_convergence
  int csr;          // deterministic accumulation
  int warmup;
  int iters;
  int n_sdf;
  int pad;
  SdfDev sdf[XPBD_MAX_SDF];
};

// Work map entry: which scene a workgroup belongs to and its first element.
struct WorkItem {
  int scene;
  int first;
};

// One patch of one scene, as walked by the persistent patch kernels.
struct __align__(32) PatchItem {
  int scene;
  int patch;
  int tet_begin;
  int vert_begin;
  int n_tets;
  int n_verts;
  int pad[2];
};

constexpr float kEps = 1e-12f;  // xpbd_gpu/constants.py:3 as an f32 constant

// Patch shape, shared by the host's patch builder and the patch kernels: a
// patch holds at most kPatchTets consecutive tets (kPatchTetsPerThread per
// thread of a kPatchThreads workgroup) touching at most kPatchVerts vertices
// (one per thread; local ids and corner ranks fit in a byte).
constexpr int kPatchThreads = 256;
constexpr int kPatchTetsPerThread = 2;
constexpr int kPatchTets = kPatchThreads * kPatchTetsPerThread;
constexpr int kPatchVerts = 256;

// Residual partial sums of the patch layout: one per group of kPartialGroup
// consecutive patches of a scene (scene-local patch index / kPartialGroup), so
// a scene's sums do not depend on how work is split or which scenes share a
// batch.  The residual kernel's work ranges start at group boundaries.
constexpr int kPartialGroup = 2;

// Persistent substep kernel: one workgroup of kPersistThreads threads per
// residual group of kPartialGroup patches (one patch per half).
constexpr int kPersistThreads = kPatchThreads * kPartialGroup;

// The persistent substep's residual mailbox: per residual group, 2
// iterations' parity x 3 components, each a double posted by value.  A word
// holding kMailboxEmpty has not been posted: no double widened from f32 or
// summed from such doubles has this bit pattern (a NaN with all 52 payload
// bits set; widened f32 NaNs keep the low 29 bits zero, and arithmetic
// propagates payloads without setting bits).
constexpr int kMailboxWords = 6;
constexpr unsigned long long kMailboxEmpty = ~0ull;

// One workgroup of the batched persistent substep: the member scene, its
// residual group in that scene, and where the scene's patch items start in
// the batch's concatenated item list.
struct PersistGroup {
  int scene;
  int group;
  int item0;
  int pad;
};

// Reciprocal and square root on the native gfx950 instructions (v_rcp_f32,
// v_sqrt_f32: about 1 ulp) instead of the correctly rounded sequences (a
// dozen instructions each), where the result is not amplified (rest-shape
// inverse, Jacobi average, residual).  The 2x2 block solve divides exactly.
__device__ __forceinline__ float frcp(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }

__device__ __forceinline__ float3 f3(const float4& a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) {
  return make_float3(a.x + b.x, a.y + b.y, a.z + b.z);
}
__device__ __forceinline__ float3 operator-(float3 a, float3 b) {
  return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ float3 operator*(float s, float3 a) {
  return make_float3(s * a.x, s * a.y, s * a.z);
}
__device__ __forceinline__ float3 operator*(float3 a, float s) {
  return make_float3(a.x * s, a.y * s, a.z * s);
}
__device__ __forceinline__ float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Row-major 3x3 held in registers.
struct Mat3 {
  float m[3][3];
};

__device__ __forceinline__ Mat3 load_inv_dm(const float4& a, const float4& b, const float2& c) {
  Mat3 r;
  r.m[0][0] = a.x; r.m[0][1] = a.y; r.m[0][2] = a.z;
  r.m[1][0] = a.w; r.m[1][1] = b.x; r.m[1][2] = b.y;
  r.m[2][0] = b.z; r.m[2][1] = b.w; r.m[2][2] = c.x;
  return r;
}

// F = Ds @ inv_Dm, Ds = cols(p1-p0, p2-p0, p3-p0)  (neohookean_solver.py:371-373)
__device__ __forceinline__ Mat3 deformation_gradient(float3 p0, float3 p1, float3 p2, float3 p3,
                                                     const Mat3& idm) {
  const float3 e1 = p1 - p0, e2 = p2 - p0, e3 = p3 - p0;
  const float ds[3][3] = {{e1.x, e2.x, e3.x}, {e1.y, e2.y, e3.y}, {e1.z, e2.z, e3.z}};
  Mat3 f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      f.m[i][j] = ds[i][0] * idm.m[0][j] + ds[i][1] * idm.m[1][j] + ds[i][2] * idm.m[2][j];
  return f;
}

__device__ __forceinline__ float norm_sqr(const Mat3& a) {
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) s += a.m[i][j] * a.m[i][j];
  return s;
}

// cofactor expansion along the first column
__device__ __forceinline__ float det3(const Mat3& a) {
  const auto& m = a.m;
  return m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
         m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
         m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
}

// Vertex gradients g1..g3 = columns of dCdF @ inv_Dm^T, g0 = -(g1+g2+g3)
// (neohookean_solver.py:382-404).
__device__ __forceinline__ void vertex_grads(const Mat3& dcdf, const Mat3& idm, float3 g[4]) {
  float c[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      c[i][j] = dcdf.m[i][0] * idm.m[j][0] + dcdf.m[i][1] * idm.m[j][1] + dcdf.m[i][2] * idm.m[j][2];
  g[1] = make_float3(c[0][0], c[1][0], c[2][0]);
  g[2] = make_float3(c[0][1], c[1][1], c[2][1]);
  g[3] = make_float3(c[0][2], c[1][2], c[2][2]);
  g[0] = make_float3(-(g[1].x + g[2].x + g[3].x), -(g[1].y + g[2].y + g[3].y),
                     -(g[1].z + g[2].z + g[3].z));
}

// dC_H/dF = cols(F1 x F2, F2 x F0, F0 x F1)  (neohookean_solver.py:391-398)
__device__ __forceinline__ Mat3 ddet_dF(const Mat3& f) {
  const float3 c0 = make_float3(f.m[0][0], f.m[1][0], f.m[2][0]);
  const float3 c1 = make_float3(f.m[0][1], f.m[1][1], f.m[2][1]);
  const float3 c2 = make_float3(f.m[0][2], f.m[1][2], f.m[2][2]);
  const float3 x0 = cross(c1, c2), x1 = cross(c2, c0), x2 = cross(c0, c1);
  Mat3 g;
  g.m[0][0] = x0.x; g.m[1][0] = x0.y; g.m[2][0] = x0.z;
  g.m[0][1] = x1.x; g.m[1][1] = x1.y; g.m[2][1] = x1.z;
  g.m[0][2] = x2.x; g.m[1][2] = x2.y; g.m[2][2] = x2.z;
  return g;
}

// solve2x2_pivot_lu (neohookean_solver.py:331-356) with the diagonal fallback
// of :425-432, in the reference's operation sequence with correctly rounded
// divisions and no fused multiply-add.  Near F = I the gradients of C_D and
// C_H are parallel (grad C_D = 2 grad C_H at rest), so for stiff materials the
// system is nearly singular and `a11 - m * a01` cancels to the compliance
// terms: the result then depends on the exact rounding of these steps, and a
// fast reciprocal or a contracted multiply-add changes it by O(1).
__device__ __forceinline__ void solve2x2_with_fallback(float A11, float A12, float A22, float b1,
                                                       float b2, float& x0, float& x1) {
#pragma clang fp contract(off)
  float a00 = A11, a01 = A12, a10 = A12, a11 = A22, c0 = b1, c1 = b2;
  int ok = 1;
  if (fabsf(a00) < fabsf(a10)) {
    float t;
    t = a00; a00 = a10; a10 = t;
    t = a01; a01 = a11; a11 = t;
    t = c0; c0 = c1; c1 = t;
  }
  if (fabsf(a00) < kEps) ok = 0;
  x0 = 0.f;
  x1 = 0.f;
  if (ok) {
    const float m = a10 / a00;
    const float a11_ = a11 - m * a01;
    const float c1_ = c1 - m * c0;
    if (fabsf(a11_) < kEps) {
      ok = 0;
    } else {
      x1 = c1_ / a11_;
      x0 = (c0 - a01 * x1) / a00;
    }
  }
  if (!ok) {
    x0 = 0.f;
    x1 = 0.f;
    if (fabsf(A11) > kEps) x0 = b1 / A11;
    if (fabsf(A22) > kEps) x1 = b2 / A22;
  }
}

// SdfModel.check + dist + normal (sdf.py:13-124) in scalar f32 arithmetic,
// shared by the kernels and the host's fixed-vertex pinning.  Taichi's
// norm(eps) is sqrt(|q|^2 + eps) and normalized(eps) is q / (|q| + eps).
__host__ __device__ inline float sdf_dist(const SdfDev& s, const float x[3]) {
  if (s.kind == XPBD_SDF_HANG) {  // HangSdfModel.dist (sdf.py:38-43)
    float dist = 1e5f;
    for (int i = 0; i < s.n_pts; ++i) {
      const float4 c = s.pts[i];
      const float q0 = x[0] - c.x, q1 = x[1] - c.y, q2 = x[2] - c.z;
      const float di = sqrtf(q0 * q0 + q1 * q1 + q2 * q2 + kEps) - s.r;
      dist = di < dist ? di : dist;
    }
    return dist;
  }
  const float q0 = x[0] - s.p[0], q1 = x[1] - s.p[1], q2 = x[2] - s.p[2];
  if (s.kind == XPBD_SDF_PLANE) return q0 * s.d[0] + q1 * s.d[1] + q2 * s.d[2];
  if (s.kind == XPBD_SDF_SPHERE) return sqrtf(q0 * q0 + q1 * q1 + q2 * q2) - s.r;
  // CubeSdf.dist (sdf.py:107-110): |max(d, 0)| + min(max(d.x, d.y, d.z), 0)
  const float e0 = fabsf(q0) - s.d[0], e1 = fabsf(q1) - s.d[1], e2 = fabsf(q2) - s.d[2];
  const float o0 = fmaxf(e0, 0.f), o1 = fmaxf(e1, 0.f), o2 = fmaxf(e2, 0.f);
  return sqrtf(o0 * o0 + o1 * o1 + o2 * o2) + fminf(fmaxf(e0, fmaxf(e1, e2)), 0.f);
}

__host__ __device__ inline void sdf_normal(const SdfDev& s, const float x[3], float n[3]) {
  if (s.kind == XPBD_SDF_PLANE) {
    n[0] = s.d[0];
    n[1] = s.d[1];
    n[2] = s.d[2];
    return;
  }
  if (s.kind == XPBD_SDF_HANG) {
    // HangSdfModel.normal (sdf.py:45-53): the nearest sphere is searched, but
    // the normal is taken from sphere 0 (the reference's behaviour, kept)
    n[0] = n[1] = n[2] = 0.f;
    if (s.n_pts <= 0) return;
    float dist = 1e5f;
    bool hit = false;
    for (int i = 0; i < s.n_pts; ++i) {
      const float4 c = s.pts[i];
      const float q0 = x[0] - c.x, q1 = x[1] - c.y, q2 = x[2] - c.z;
      const float di = sqrtf(q0 * q0 + q1 * q1 + q2 * q2 + kEps) - s.r;
      if (di < dist) {
        dist = di;
        hit = true;
      }
    }
    if (hit) {
      const float4 c = s.pts[0];
      const float q0 = x[0] - c.x, q1 = x[1] - c.y, q2 = x[2] - c.z;
      const float inv = 1.f / (sqrtf(q0 * q0 + q1 * q1 + q2 * q2) + kEps);
      n[0] = inv * q0;
      n[1] = inv * q1;
      n[2] = inv * q2;
    }
    return;
  }
  const float q[3] = {x[0] - s.p[0], x[1] - s.p[1], x[2] - s.p[2]};
  if (s.kind == XPBD_SDF_SPHERE) {
    const float inv = 1.f / (sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) + kEps);
    n[0] = inv * q[0];
    n[1] = inv * q[1];
    n[2] = inv * q[2];
    return;
  }
  // CubeSdf.normal (sdf.py:112-124): the axis of the largest |x - centre|
  int axis = 0;
  if (fabsf(q[1]) > fabsf(q[0])) axis = 1;
  if (fabsf(q[2]) > fabsf(q[axis])) axis = 2;
  n[0] = n[1] = n[2] = 0.f;
  n[axis] = q[axis] > 0.f ? 1.f : -1.f;
}

// SdfModel.check (sdf.py:13-27): inside when phi < 0 and (n . (vel_sdf - v) > 0
// or fixed), with vel_sdf the obstacle's velocity.
__host__ __device__ inline bool sdf_inside(const SdfDev& s, const float x[3], const float v[3],
                                           float n[3]) {
  n[0] = n[1] = n[2] = 0.f;
  if (!(sdf_dist(s, x) < 0.f)) return false;
  sdf_normal(s, x, n);
  const float dv0 = s.vel[0] - v[0], dv1 = s.vel[1] - v[1], dv2 = s.vel[2] - v[2];
  const float dotnv = n[0] * dv0 + n[1] * dv1 + n[2] * dv2;
  return dotnv > 0.f || s.fixed;
}

// One collider projection of one vertex (xpbd_base.py:162-174).
__device__ __forceinline__ float3 sdf_project(const SdfDev& s, float3 x, float3 v) {
  const float xa[3] = {x.x, x.y, x.z}, va[3] = {v.x, v.y, v.z};
  float n[3];
  if (sdf_inside(s, xa, va, n)) {
    const float depth = -sdf_dist(s, xa);
    if (depth > 0.f) x = x + make_float3(n[0], n[1], n[2]) * depth;
  }
  return x;
}

// inv_Dm and 1/det(Dm) of a tet from its rest positions (initialize,
// neohookean_solver.py:97-107; V0 = det(Dm) / 6 as calc_tet_volume,
// xpbd_base.py:6-8), recomputed where used.  With Dm = cols(a, b, c) the
// rows of inv_Dm are (b x c, c x a, a x b) / det, det = a . (b x c).
__device__ __forceinline__ void rest_shape(float3 p0, float3 p1, float3 p2, float3 p3, Mat3& inv,
                                           float& inv_det) {
  const float3 a = p1 - p0, b = p2 - p0, c = p3 - p0;
  const float3 bc = cross(b, c), ca = cross(c, a), ab = cross(a, b);
  inv_det = frcp(dot(a, bc));
  inv.m[0][0] = bc.x * inv_det; inv.m[0][1] = bc.y * inv_det; inv.m[0][2] = bc.z * inv_det;
  inv.m[1][0] = ca.x * inv_det; inv.m[1][1] = ca.y * inv_det; inv.m[1][2] = ca.z * inv_det;
  inv.m[2][0] = ab.x * inv_det; inv.m[2][1] = ab.y * inv_det; inv.m[2][2] = ab.z * inv_det;
}

}  // namespace xpbd
