// gfx950 kernels of the XPBD Neo-Hookean substep.
//
// Every kernel walks a work map: workgroup b handles item map[b] of scene
// map[b].scene (a range of vertices or tets, or one patch), whose buffers and
// parameters come from a device-resident scene table.  One solver is a table
// of one scene; the batched form packs many ragged scenes into the same
// launches.
//
// Two tet layouts share the per-tet math:
//  flat   one thread per tet; inv_Dm/V0 streamed from HBM; corrections go to
//         per-vertex float atomics or to per-corner slots gathered per vertex.
//  patch  one workgroup per patch (<= 512 bisection-ordered tets, two per
//         thread, <= 256 vertices); the patch's vertices are staged in LDS,
//         inv_Dm/V0 are recomputed from the staged rest positions, corrections
//         accumulate in LDS and leave the workgroup as one partial per
//         (patch, vertex) slot, so no global atomics are issued.
#include "xpbd_device.h"
#include "xpbd_kernels.h"

#include <hip/hip_ext.h>

#include <mutex>

namespace xpbd {

namespace {

__device__ __forceinline__ void atomic_add3(float4* dst, float3 v) {
  unsafeAtomicAdd(&dst->x, v.x);
  unsafeAtomicAdd(&dst->y, v.y);
  unsafeAtomicAdd(&dst->z, v.z);
}

// Sum over the 64 lanes of a wavefront (lane-exchange butterfly); every lane
// receives the total, in the same order on every run.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

struct TetProjection {
  float3 dp[4];
  float cnt[4];
};

// The per-tet constraint projection of one Jacobi iteration.
//  block mode: solve_neohookean_block (neohookean_solver.py:358-451)
//  split mode: solve_hydrostatic (:196-254) then solve_deviatoric (:144-194)
//              on the same positions.
__device__ __forceinline__ void project_tet(const SceneDev& S, const float4 q[4], const Mat3& idm,
                                            float inv_det, float2& lam, TetProjection& out) {
  const float w[4] = {q[0].w, q[1].w, q[2].w, q[3].w};
  const Mat3 F = deformation_gradient(f3(q[0]), f3(q[1]), f3(q[2]), f3(q[3]), idm);
  // alpha~ = 1 / (lame V0 dt^2) = alpha_k / det(Dm)
  const float aD = S.alpha_kd * inv_det;
  const float aH = S.alpha_kh * inv_det;
  if (S.block_mode) {
    const float C_D = norm_sqr(F) - 3.f;
    // det F by expansion along the first column, from the cofactor matrix
    // the gradient needs anyway (the same products as det3)
    const Mat3 H = ddet_dF(F);
    const float C_H = F.m[0][0] * H.m[0][0] + F.m[1][0] * H.m[1][0] + F.m[2][0] * H.m[2][0] - 1.f;
    // grad C_D = 2 F inv_Dm^T: the gradients gF of F inv_Dm^T are formed and
    // the factor 2 is applied to the scalars below (exact in floating point)
    float3 gF[4], gH[4];
    vertex_grads(F, idm, gF);
    vertex_grads(H, idm, gH);
    const float sFF = w[0] * dot(gF[0], gF[0]) + w[1] * dot(gF[1], gF[1]) +
                      w[2] * dot(gF[2], gF[2]) + w[3] * dot(gF[3], gF[3]);
    const float sHH = w[0] * dot(gH[0], gH[0]) + w[1] * dot(gH[1], gH[1]) +
                      w[2] * dot(gH[2], gH[2]) + w[3] * dot(gH[3], gH[3]);
    const float sFH = w[0] * dot(gF[0], gH[0]) + w[1] * dot(gF[1], gH[1]) +
                      w[2] * dot(gF[2], gH[2]) + w[3] * dot(gF[3], gH[3]);
    const float A11 = 4.f * sFF + aD;
    const float A22 = sHH + aH;
    const float A12 = 2.f * sFH;
    const float b1 = -(C_D + aD * lam.x);
    const float b2 = -(C_H + aH * lam.y);
    float dlD, dlH;
    solve2x2_with_fallback(A11, A12, A22, b1, b2, dlD, dlH);
    lam.x += dlD;
    lam.y += dlH;
    const float dlD2 = 2.f * dlD;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      out.dp[k] = gF[k] * (w[k] * dlD2) + gH[k] * (w[k] * dlH);
      out.cnt[k] = 1.f;
    }
  } else {
    float3 gH[4], gD[4];
    vertex_grads(ddet_dF(F), idm, gH);
    const float C_H = det3(F) - 1.f;
    float wsH = w[0] * dot(gH[0], gH[0]);
    wsH += w[1] * dot(gH[1], gH[1]);
    wsH += w[2] * dot(gH[2], gH[2]);
    wsH += w[3] * dot(gH[3], gH[3]);
    float dlH = 0.f;
    const int okH = wsH > kEps;
    if (okH) {
      dlH = -(C_H + aH * lam.y) / (wsH + aH);
      dlH *= S.hydro_relax;
      lam.y += dlH;
    }
    Mat3 gfd;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) gfd.m[i][j] = 2.f * F.m[i][j];
    vertex_grads(gfd, idm, gD);
    const float C_D = norm_sqr(F) - 3.f;
    float wsD = w[0] * dot(gD[0], gD[0]);
    wsD += w[1] * dot(gD[1], gD[1]);
    wsD += w[2] * dot(gD[2], gD[2]);
    wsD += w[3] * dot(gD[3], gD[3]);
    float dlD = 0.f;
    const int okD = wsD > kEps;
    if (okD) {
      dlD = -(C_D + aD * lam.x) / (wsD + aD);
      dlD *= S.dev_relax;
      lam.x += dlD;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float3 a = make_float3(0.f, 0.f, 0.f);
      if (okH) a = a + gH[k] * (dlH * w[k]);
      if (okD) a = a + gD[k] * (dlD * w[k]);
      out.dp[k] = a;
      out.cnt[k] = float(okH + okD);
    }
  }
}

// compute_xpbd_residual (neohookean_solver.py:298-329) for one tet, plus the
// |C_H| and |C_D| of the monitor kernels (:256-296).
__device__ __forceinline__ void residual_tet(const SceneDev& S, float3 x0, float3 x1, float3 x2,
                                             float3 x3, const Mat3& idm, float inv_det,
                                             float2 lam, float& r, float& ah, float& ad) {
  const Mat3 F = deformation_gradient(x0, x1, x2, x3, idm);
  const float C_D = norm_sqr(F) - 3.f;
  const float C_H = det3(F) - 1.f;
  r = 0.f;
  if (S.cheb) {
    const float aD = S.alpha_kd * inv_det;
    const float aH = S.alpha_kh * inv_det;
    const float rD = C_D + aD * lam.x;
    const float rH = C_H + aH * lam.y;
    r = fsqrt(rD * rD + rH * rH);
  }
  ah = fabsf(C_H);
  ad = fabsf(C_D);
}

// Workgroup sum of (r, |C_H|, |C_D|) in a fixed order: f32 wave sums, then
// the waves in order in double; thread 0 gets the totals.  |C_H| and |C_D| are
// only summed when `all3` (convergence monitor on; uniform).  The caller puts
// a workgroup barrier between two calls.
__device__ __forceinline__ void block_sums(float r, float ah, float ad, bool all3, double tot[3]) {
  __shared__ float red[3][16];
  r = wave_sum(r);
  if (all3) {
    ah = wave_sum(ah);
    ad = wave_sum(ad);
  } else {
    ah = ad = 0.f;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  if (lane == 0) {
    red[0][wave] = r;
    red[1][wave] = ah;
    red[2][wave] = ad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < nwave; ++i) {
      s0 += red[0][i];
      s1 += red[1][i];
      s2 += red[2][i];
    }
    tot[0] = s0;
    tot[1] = s1;
    tot[2] = s2;
  }
}

// Data that the workgroups of one persistent launch hand to each other (slot
// partials across a scene barrier, residual partials through the mailbox) is
// stored and loaded as agent-scope atomics: global stores and loads with the
// sc1 bit, performed at the device's point of coherence, so they need no L2
// write-back before the barrier and no L2 invalidate after it.  A float4 goes
// as two 8-byte words.
__device__ __forceinline__ void xstore4(gptr<float4> p, float4 v) {
  gptr<unsigned long long> q = (gptr<unsigned long long>)p;
  __hip_atomic_store(q, __builtin_bit_cast(unsigned long long, make_float2(v.x, v.y)),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, __builtin_bit_cast(unsigned long long, make_float2(v.z, v.w)),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float4 xload4(gptr<const float4> p) {
  gptr<const unsigned long long> q = (gptr<const unsigned long long>)p;
  const float2 a = __builtin_bit_cast(
      float2, __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  const float2 b = __builtin_bit_cast(
      float2, __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  return make_float4(a.x, a.y, b.x, b.y);
}
__device__ __forceinline__ void xstore_d(gptr<double> p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// s_waitcnt immediate that waits for vector memory only (gfx9 encoding:
// vmcnt 0 in bits 3:0 and 15:14, expcnt 7 in bits 6:4, lgkmcnt 15 in 11:8).
constexpr unsigned kWaitVmcnt0 = 0x0F70;

// Stores partial idx (component c at partial[c * n_partials + idx]).
__device__ __forceinline__ void store_partial(const SceneDev& S, int idx, const double tot[3]) {
  const size_t n = size_t(S.n_partials);
  S.partial[idx] = tot[0];
  S.partial[n + idx] = tot[1];
  S.partial[2 * n + idx] = tot[2];
}

// Posts residual group `group`'s partials of iteration `it` to the scene's
// mailbox (parity it & 1), each component an agent-scope atomic 8-byte store:
// the value itself tells a reader that it has arrived.
__device__ __forceinline__ void post_partials(const SceneDev& S, int it, int group,
                                              const double tot[3]) {
  const size_t G = size_t(S.n_partials);
  gptr<double> box = S.mailbox + size_t(it & 1) * 3 * G + group;
  xstore_d(box, tot[0]);
  xstore_d(box + G, tot[1]);
  xstore_d(box + 2 * G, tot[2]);
}

// delta_x (xyz) and the constraint count (w) of vertex v summed over its slots
// [b, e) in ascending patch order.  kX: the slots were written by other
// workgroups of the same launch (exchange loads).
template <bool kX = false>
__device__ __forceinline__ float4 slot_sum(const SceneDev& S, int b, int e) {
  auto ld = [&](int i) -> float4 {
    if constexpr (kX)
      return xload4(S.slot_acc + i);
    else
      return S.slot_acc[i];
  };
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  int a = b;
  // four slot loads in flight, added in the same order as one at a time
  for (; a + 4 <= e; a += 4) {
    const float4 c0 = ld(a), c1 = ld(a + 1), c2 = ld(a + 2), c3 = ld(a + 3);
    acc.x += c0.x; acc.y += c0.y; acc.z += c0.z; acc.w += c0.w;
    acc.x += c1.x; acc.y += c1.y; acc.z += c1.z; acc.w += c1.w;
    acc.x += c2.x; acc.y += c2.y; acc.z += c2.z; acc.w += c2.w;
    acc.x += c3.x; acc.y += c3.y; acc.z += c3.z; acc.w += c3.w;
  }
  for (; a < e; ++a) {
    const float4 c = ld(a);
    acc.x += c.x;
    acc.y += c.y;
    acc.z += c.z;
    acc.w += c.w;
  }
  return acc;
}

// x_J = y + delta_x / cons_num (post_solve, neohookean_solver.py:115-119);
// cons_num is the valence in block mode, the summed count in split mode.
__device__ __forceinline__ float4 jacobi_from(const SceneDev& S, float4 acc, float valence,
                                              float4 p) {
  const float cnt = S.block_mode ? valence : acc.w;
  float3 x = f3(p);
  if (cnt > 0.f) {
    const float ic = frcp(cnt);
    x = x + make_float3(acc.x * ic, acc.y * ic, acc.z * ic);
  }
  return make_float4(x.x, x.y, x.z, p.w);
}

// x_J of vertex v with delta_x summed from its patch slots.  In block mode
// with deterministic accumulation every corner adds a count of exactly 1, so
// the slots' summed count is the valence (the same float) and the valence is
// not read; LDS atomics in block mode leave the count at 0.
__device__ __forceinline__ float4 jacobi_average_slots(const SceneDev& S, int v, int it) {
  const float4 acc = slot_sum(S, S.vslot_off[v], S.vslot_off[v + 1]);
  const float4 p = S.fused ? S.ybuf[it % 3][v] : S.pos[v];
  return jacobi_from(S, acc, S.block_mode && !S.csr ? S.valence[v] : acc.w, p);
}

// apply_external_forces (xpbd_base.py:139-146) and the collider projections
// of solve_sdf_collision (xpbd_base.py:162-179) for one vertex: from x (xs),
// v and inv_m (w of pos) to the predicted position; vel is updated.
__device__ __forceinline__ float3 predict_vertex(const SceneDev& S, float3 xs, float3& vel,
                                                 float w, bool external, bool collide,
                                                 float3 x) {
  if (external) {
    const float dt = S.dt;
    if (w > 0.f) vel = vel + make_float3(S.g[0], S.g[1], S.g[2]) * dt;
    x = xs + vel * dt;
  }
  if (collide && w > 0.f) {
    for (int s = 0; s < S.n_sdf; ++s) x = sdf_project(S.sdf[s], x, vel);
  }
  return x;
}

// advance (xpbd_base.py:148-160) of one vertex with final position p and start
// position x: pinned vertices go back to x; free ones get v = (p - x) / dt *
// exp(-dt damping) and x = p.  Writes pos (pinned) or vel and xs (free).
__device__ __forceinline__ void advance_vertex(const SceneDev& S, int v, float4 p, float4 x) {
  if (p.w <= 0.f) {
    S.pos[v] = make_float4(x.x, x.y, x.z, p.w);
  } else {
    const float dt = S.dt;
    const float3 raw = make_float3((p.x - x.x) / dt, (p.y - x.y) / dt, (p.z - x.z) / dt);
    const float3 vel = raw * S.decay;
    S.vel[v] = make_float4(vel.x, vel.y, vel.z, 0.f);
    S.xs[v] = make_float4(p.x, p.y, p.z, x.w);
  }
}

}  // namespace

// apply_external_forces (xpbd_base.py:139-146), then, when `collide`, the
// SDF projections of solve_sdf_collision (xpbd_base.py:162-179) and, when
// `save_prev`, _save_curr_to_prev (neohookean_solver.py:121-124).
__global__ void k_predict(const SceneDev* __restrict__ scenes, const WorkItem* __restrict__ map,
                          int external, int collide, int save_prev) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  const int v = wi.first + threadIdx.x;
  if (v >= S.n_verts) return;
  float4 p = S.pos[v];
  float3 vel = f3(S.vel[v]);
  const float w = p.w;
  const float3 x = predict_vertex(S, external ? f3(S.xs[v]) : f3(p), vel, w, external != 0,
                                  collide != 0, f3(p));
  if (external) S.vel[v] = make_float4(vel.x, vel.y, vel.z, 0.f);
  p = make_float4(x.x, x.y, x.z, w);
  S.pos[v] = p;
  if (save_prev) S.prev[v] = p;
}

// ---- flat layout ---------------------------------------------------------------

// One thread per tet.  `first` marks the first iteration of a substep: the
// multipliers start at 0 (neohookean_solver.py:458-459).
__global__ void k_project(const SceneDev* __restrict__ scenes, const WorkItem* __restrict__ map,
                          int first) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  const int t = wi.first + threadIdx.x;
  if (t >= S.n_tets) return;
  const int4 id = S.tet[t];
  const float4 q[4] = {S.pos[id.x], S.pos[id.y], S.pos[id.z], S.pos[id.w]};
  const float2 d2 = S.dm2[t];
  const Mat3 idm = load_inv_dm(S.dm0[t], S.dm1[t], d2);
  float2 lam = first ? make_float2(0.f, 0.f) : S.lam[t];
  TetProjection P;
  project_tet(S, q, idm, frcp(6.f * d2.y), lam, P);
  S.lam[t] = lam;
  if (S.csr) {
    float4* c = S.corner + 4 * size_t(t);
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = make_float4(P.dp[k].x, P.dp[k].y, P.dp[k].z, P.cnt[k]);
  } else {
    const int ids[4] = {id.x, id.y, id.z, id.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      atomic_add3(S.dacc + ids[k], P.dp[k]);
      if (!S.block_mode && P.cnt[k] != 0.f) unsafeAtomicAdd(&S.dacc[ids[k]].w, P.cnt[k]);
    }
  }
}

// pre_solve / post_solve (neohookean_solver.py:109-119): x_J = y + delta_x /
// cons_num.  With Chebyshev on, x_J goes to xj (y stays in pos for the blend);
// otherwise pos is updated in place.  Atomic accumulators are cleared here.
__global__ void k_average(const SceneDev* __restrict__ scenes, const WorkItem* __restrict__ map) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  const int v = wi.first + threadIdx.x;
  if (v >= S.n_verts) return;
  float3 d;
  float cnt;
  if (S.csr) {
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int b = S.adj_off[v], e = S.adj_off[v + 1];
    for (int a = b; a < e; ++a) {
      const float4 c = S.corner[S.adj[a]];
      acc.x += c.x;
      acc.y += c.y;
      acc.z += c.z;
      acc.w += c.w;
    }
    d = f3(acc);
    cnt = S.block_mode ? S.valence[v] : acc.w;
  } else {
    const float4 acc = S.dacc[v];
    d = f3(acc);
    cnt = S.block_mode ? S.valence[v] : acc.w;
    S.dacc[v] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  const float4 p = S.pos[v];
  float3 x = f3(p);
  if (cnt > 0.f) x = x + make_float3(d.x / cnt, d.y / cnt, d.z / cnt);
  const float4 out = make_float4(x.x, x.y, x.z, p.w);
  if (S.cheb)
    S.xj[v] = out;
  else
    S.pos[v] = out;
}

// Residual per tet on x_J; one partial sum per workgroup.
__global__ void k_residual(const SceneDev* __restrict__ scenes, const WorkItem* __restrict__ map) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  const int t = wi.first + threadIdx.x;
  const float4* X = S.cheb ? S.xj : S.pos;
  float r = 0.f, ah = 0.f, ad = 0.f;
  if (t < S.n_tets) {
    const int4 id = S.tet[t];
    const float2 d2 = S.dm2[t];
    const Mat3 idm = load_inv_dm(S.dm0[t], S.dm1[t], d2);
    residual_tet(S, f3(X[id.x]), f3(X[id.y]), f3(X[id.z]), f3(X[id.w]), idm, frcp(6.f * d2.y),
                 S.lam[t], r,
                 ah, ad);
  }
  double tot[3];
  block_sums(r, ah, ad, S.monitor != 0, tot);
  if (threadIdx.x == 0) store_partial(S, wi.first / S.tet_block, tot);
}

// ---- patch layout ----------------------------------------------------------------

// The Chebyshev blend of one vertex (_chebyshev_update + _rotate_prev_with_save,
// neohookean_solver.py:131-142): y^{k+1} = prev + omega (gamma (x_J - y) + y - prev).
__device__ __forceinline__ float4 cheb_blend(const SceneDev& S, float4 y4, float4 xj4, float4 pr4,
                                             float omega) {
  const float gamma = S.gamma;
  const float3 y = f3(y4), xj = f3(xj4), pr = f3(pr4);
  const float3 x_new = gamma * (xj - y) + y;
  const float3 nw = pr + omega * (x_new - pr);
  return make_float4(nw.x, nw.y, nw.z, y4.w);
}

// The inputs of one patch vertex's staged iterate y^it: in the fused sequence
// y^(it-1) (ybuf[(it-1)%3]), prev = y^(it-2) (or y^0 when it == 1) and
// x_J^(it-1) (xj), otherwise pos; and the rest position.  They are loaded one
// patch ahead and blended at staging, so the loads stay in flight while the
// previous patch is computed.
struct StageIn {
  float4 ya, yb, xa, r;
};

__device__ __forceinline__ void load_stage(const SceneDev& S, int g, int it, bool blend,
                                           StageIn& in) {
  if (blend) {
    in.ya = S.ybuf[(it - 1) % 3][g];
    // selecting the buffer (not the value) keeps both loads global
    in.yb = S.ybuf[it == 1 ? 0 : (it - 2) % 3][g];
    in.xa = S.xj[g];
  } else {
    in.ya = S.pos[g];
  }
  in.r = S.rest[g];
}

// y^it of a staged vertex: y^it = blend(y^(it-1), x_J^(it-1), prev, omega^(it-1))
// when blending, else the loaded position.  The vertex's owner patch stores it
// in ybuf[it%3]; the other patches holding the vertex compute the same bits.
__device__ __forceinline__ float4 blend_stage(const SceneDev& S, const StageIn& in, int g, int it,
                                              bool blend, bool owner, float omega) {
  if (!blend) return in.ya;
  const float4 y = cheb_blend(S, in.ya, in.xa, in.yb, omega);
  if (owner) S.ybuf[it % 3][g] = y;
  return y;
}

// ---- per-scene scalar control ----------------------------------------------------

// Workgroup size of every scene reduction, so that the standalone reduce and
// the one folded into the projection sum in the same order.
constexpr int kReduceThreads = kPatchThreads;

// LDS doubles a scene reduction needs as scratch.
constexpr int kReduceScratch = 3 * (kReduceThreads / 64);
// float4 entries holding a reduction's scratch plus the evaluated omega
// (kReduceScratch + 1 doubles).
constexpr int kOmegaScratchF4 = (8 * (kReduceScratch + 1) + 15) / 16;

// Fixed-order sums of a scene's partials (component 0, and 1-2 when `all3`)
// by one workgroup of kReduceThreads threads: each thread adds its strided
// entries in ascending order, then waves and workgroup combine in a fixed
// tree.  Every workgroup that calls it gets the same bits; thread 0 gets the
// totals.  `red` is LDS scratch of kReduceScratch doubles.  Ends with a
// workgroup barrier.
__device__ __forceinline__ void scene_sums(const SceneDev& S, bool all3, double tot[3],
                                           double* red) {
  constexpr int kW = kReduceThreads / 64;
  const int n = S.n_partials;
  const int tid = threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c > 0 && !all3) break;
    gptr<const double> p = S.partial + size_t(c) * size_t(n);
    double sum = 0.0;
    // eight loads in flight per step, all issued before the first add (the
    // indices are clamped, not branched on), then added in ascending order
    for (int b = tid; b < n; b += 8 * kReduceThreads) {
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = p[min(b + k * kReduceThreads, n - 1)];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (b + k * kReduceThreads < n) sum += v[k];
    }
    acc[c] = wave_sum_d(sum);
  }
  const int lane = tid & 63, wave = tid >> 6;
  if (lane == 0) {
    red[wave] = acc[0];
    red[kW + wave] = acc[1];
    red[2 * kW + wave] = acc[2];
  }
  __syncthreads();
  if (tid == 0) {
    tot[0] = tot[1] = tot[2] = 0.0;
    for (int w = 0; w < kW; ++w) {
      tot[0] += red[w];
      tot[1] += red[kW + w];
      tot[2] += red[2 * kW + w];
    }
  }
  __syncthreads();
}

// The residual-driven rho estimate and Chebyshev omega after iteration it
// (neohookean_solver.py:461-466, 492-504, 537-546), in double, from the
// controller state after iteration it-1 and the scene's residual sums; also
// forms the iteration's log row (iter, r_hat, rho_hat, omega, hydro, dev).
__device__ __forceinline__ SceneCtrl ctrl_update(const SceneDev& S, SceneCtrl c, int it,
                                                 const double tot[3], double L[6]) {
  if (it == 0) {  // prev_r_hat = None and omega = 1.0 per substep (:461, :466)
    c.have_prev = 0;
    c.omega = 1.0;
  }
  const double cnt = double(S.n_tets);
  double r_hat = 0.0;
  double omega = 1.0;
  double rho = c.rho_hat;
  if (S.cheb) {
    r_hat = S.n_tets > 0 ? tot[0] / cnt : 0.0;
    if (c.have_prev && c.prev_r_hat > 1e-12) {
      double rho_inst = r_hat / (c.prev_r_hat + 1e-12);
      rho_inst = fmax(0.00001, fmin(0.99999, rho_inst));
      rho = (1.0 - 0.2) * rho + 0.2 * rho_inst;
      c.rho_hat = rho;
    }
    c.prev_r_hat = r_hat;
    c.have_prev = 1;
    const double rho2 = rho * rho;
    if (it < S.warmup)
      omega = 1.0;
    else if (it == S.warmup)
      omega = 2.0 / (2.0 - rho2);
    else
      omega = 4.0 / (4.0 - rho2 * c.omega);
    c.omega = omega;
    c.r_hat = r_hat;
  }
  L[0] = it;
  L[1] = r_hat;
  L[2] = rho;
  L[3] = omega;
  L[4] = S.monitor && S.n_tets > 0 ? tot[1] / cnt : 0.0;
  L[5] = S.monitor && S.n_tets > 0 ? tot[2] / cnt : 0.0;
  return c;
}

__device__ __forceinline__ void store_log_row(const SceneDev& S, int it, const double L[6]) {
  gptr<double> dst = S.log + 6 * size_t(it);
#pragma unroll
  for (int k = 0; k < 6; ++k) dst[k] = L[k];
}

// The controller state is double-buffered: the update after iteration k reads
// ctrl[(k + 1) & 1] and writes ctrl[k & 1].  In the fused sequence every
// projection workgroup of iteration it evaluates the update after it-1 for
// its scenes itself (identical bits in all of them), so the separate reduce
// launch is needed only after the last iteration; the workgroup holding the
// scene's patch 0 stores the new state and the log row, into the slot no
// workgroup of this launch reads.  Returns omega^(it-1) to every thread.
// `red` is LDS scratch of kReduceScratch + 1 doubles; it is read after the
// last barrier, so the caller separates its next writes there by a barrier.
__device__ __forceinline__ float scene_omega(const SceneDev& S, int it, bool writer,
                                             double* red) {
  double tot[3];
  scene_sums(S, S.monitor != 0, tot, red);
  float* s_omega = reinterpret_cast<float*>(red + kReduceScratch);
  if (threadIdx.x == 0) {
    double L[6];
    const SceneCtrl c = ctrl_update(S, S.ctrl[it & 1], it - 1, tot, L);
    *s_omega = float(c.omega);
    if (writer) {
      S.ctrl[(it - 1) & 1] = c;
      store_log_row(S, it - 1, L);
    }
  }
  __syncthreads();
  return *s_omega;
}

// Out-of-line form for the projection's scene changes inside its loop (batches
// only), which keeps the loop's register use down.
__device__ __noinline__ float fused_omega(const SceneDev& S, int it, bool writer, double* red) {
  return scene_omega(S, it, writer, red);
}

// Tet data of one patch held by a thread: tet tid + j * kPatchThreads for j
// in [0, kPatchTetsPerThread) (local ids and corner ranks, multipliers).
struct PatchTets {
  uint2 loc[kPatchTetsPerThread];
  float2 lam[kPatchTetsPerThread];
};

// `lt` is the thread's index among the kPatchThreads threads working on h.
__device__ __forceinline__ void load_patch_tets(const SceneDev& S, const PatchItem& h, bool lam,
                                                PatchTets& d, int lt = threadIdx.x) {
#pragma unroll
  for (int j = 0; j < kPatchTetsPerThread; ++j) {
    const int t = lt + j * kPatchThreads;
    d.loc[j] = make_uint2(0u, 0u);
    d.lam[j] = make_float2(0.f, 0.f);
    if (t < h.n_tets) {
      d.loc[j] = S.tet_local[h.tet_begin + t];
      if (lam) d.lam[j] = S.lam[h.tet_begin + t];
    }
  }
}

// Persistent patch projection.  Workgroup b walks the patch items
// [ranges[b].x, ranges[b].y): it stages a patch's vertices and rest positions
// in LDS, projects every tet, accumulates the corrections per patch vertex in
// LDS, and writes one partial per (patch, vertex) slot.  The loads are
// software-pipelined: while patch i is computed, the vertex inputs and tet
// data of patch i+1 and the vertex ids of patch i+2 are in flight.
// Deterministic mode stores each tet corner at its vertex's rank position in
// an LDS array sorted by vertex (ranks follow tet order) and sums each
// vertex's run in that order; atomic mode adds into per-vertex LDS cells.
// kMultiScene: the patch ranges may cross scenes (batches), so omega is
// re-evaluated on a scene change.
template <bool kMultiScene>
__global__ void __launch_bounds__(kPatchThreads)
    k_patch_project(const SceneDev* __restrict__ scenes, const PatchItem* __restrict__ items,
                    const int2* __restrict__ ranges, int it) {
  const int2 rg = ranges[blockIdx.x];
  if (rg.x >= rg.y) return;
  const bool first = it == 0;
  __shared__ float4 s_y[kPatchVerts];
  // rest positions; w holds the start of the vertex's run in s_sorted
  __shared__ float4 s_rest[kPatchVerts];
  // corner corrections sorted by vertex (deterministic) or per-vertex LDS
  // atomics (the first kPatchVerts entries); the omega evaluation's scratch
  // sits in the last entries, which the atomic accumulators never reach and
  // which the deterministic corner writes only fill after the workgroup
  // barrier that follows every read of the evaluated omega
  __shared__ float4 s_sorted[4 * kPatchTets];
  float4* s_acc = s_sorted;
  static_assert(kPatchVerts + kOmegaScratchF4 <= 4 * kPatchTets, "omega scratch overlaps s_acc");
  double* red = reinterpret_cast<double*>(s_sorted + 4 * kPatchTets - kOmegaScratchF4);
  const int tid = threadIdx.x;
  const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);

  // stage 0 (current patch): vertex ids and inputs, tet data
  PatchItem h0 = items[rg.x];
  const SceneDev& S0 = scenes[h0.scene];
  bool blend0 = S0.fused && !first;
  int g0 = 0;
  int2 vs0 = make_int2(0, 0);
  StageIn in0{zero, zero, zero, zero};
  PatchTets t0;
  if (tid < h0.n_verts) {
    g0 = S0.p_vert_ids[h0.vert_begin + tid];
    vs0 = S0.p_vert_sum[h0.vert_begin + tid];
    load_stage(S0, g0, it, blend0, in0);
  }
  load_patch_tets(S0, h0, !first, t0);
  // stage 1 (next patch): vertex ids only
  bool has1 = rg.x + 1 < rg.y;
  PatchItem h1 = h0;
  int g1 = 0;
  int2 vs1 = make_int2(0, 0);
  if (has1) {
    h1 = items[rg.x + 1];
    const SceneDev& S = scenes[h1.scene];
    if (tid < h1.n_verts) {
      g1 = S.p_vert_ids[h1.vert_begin + tid];
      vs1 = S.p_vert_sum[h1.vert_begin + tid];
    }
  }
  // omega^(it-1) of the scene of the patch staged next, evaluated while the
  // first patch's loads are in flight (the staging below writes no LDS the
  // evaluation reads, and a barrier follows it)
  float omega = 1.f;
  int omega_scene = -1;
  if (blend0) {
    omega = scene_omega(S0, it, h0.patch == 0, red);
    omega_scene = h0.scene;
  }
  for (int i = rg.x; i < rg.y; ++i) {
    const SceneDev& S = scenes[h0.scene];
    const bool csr = S.csr != 0;
    if (tid < h0.n_verts) {
      s_y[tid] = blend_stage(S, in0, g0, it, blend0, vs0.x < 0, omega);
      s_rest[tid] = make_float4(in0.r.x, in0.r.y, in0.r.z, __int_as_float(vs0.x & 0xffff));
      if (!csr) s_acc[tid] = zero;
    }
    __syncthreads();
    // in flight during the compute below: inputs of patch i+1, ids of patch i+2
    StageIn in1{zero, zero, zero, zero};
    bool blend1 = false;
    PatchTets t1;
    if (has1) {
      const SceneDev& S1 = scenes[h1.scene];
      blend1 = S1.fused && !first;
      if (kMultiScene && blend1 && h1.scene != omega_scene) {  // uniform over the workgroup
        omega = fused_omega(S1, it, h1.patch == 0, red);
        omega_scene = h1.scene;
        __syncthreads();  // omega is read before s_sorted is written again
      }
      if (tid < h1.n_verts) load_stage(S1, g1, it, blend1, in1);
      load_patch_tets(S1, h1, !first, t1);
    }
    const bool has2 = i + 2 < rg.y;
    PatchItem h2 = h1;
    int g2 = 0;
    int2 vs2 = make_int2(0, 0);
    if (has2) {
      h2 = items[i + 2];
      const SceneDev& S2 = scenes[h2.scene];
      if (tid < h2.n_verts) {
        g2 = S2.p_vert_ids[h2.vert_begin + tid];
        vs2 = S2.p_vert_sum[h2.vert_begin + tid];
      }
    }
#pragma unroll
    for (int j = 0; j < kPatchTetsPerThread; ++j) {
      const int t = tid + j * kPatchThreads;
      if (t < h0.n_tets) {
        const uint2 loc = t0.loc[j];
        const int l[4] = {int(loc.x & 0xffu), int((loc.x >> 8) & 0xffu),
                          int((loc.x >> 16) & 0xffu), int(loc.x >> 24)};
        const float4 q[4] = {s_y[l[0]], s_y[l[1]], s_y[l[2]], s_y[l[3]]};
        const float4 rs[4] = {s_rest[l[0]], s_rest[l[1]], s_rest[l[2]], s_rest[l[3]]};
        Mat3 idm;
        float inv_det;
        rest_shape(f3(rs[0]), f3(rs[1]), f3(rs[2]), f3(rs[3]), idm, inv_det);
        float2 lam = t0.lam[j];
        TetProjection P;
        project_tet(S, q, idm, inv_det, lam, P);
        S.lam[h0.tet_begin + t] = lam;
        if (csr) {
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int rank = int((loc.y >> (8 * k)) & 0xffu);
            s_sorted[__float_as_int(rs[k].w) + rank] =
                make_float4(P.dp[k].x, P.dp[k].y, P.dp[k].z, P.cnt[k]);
          }
        } else {
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            atomicAdd(&s_acc[l[k]].x, P.dp[k].x);
            atomicAdd(&s_acc[l[k]].y, P.dp[k].y);
            atomicAdd(&s_acc[l[k]].z, P.dp[k].z);
            if (!S.block_mode) atomicAdd(&s_acc[l[k]].w, P.cnt[k]);
          }
        }
      }
    }
    __syncthreads();
    if (tid < h0.n_verts) {
      float4 a;
      if (csr) {
        a = zero;
        const int cb = vs0.x & 0xffff, ce = cb + ((vs0.x >> 16) & 0x7fff);
        int c = cb;
        // four reads in flight, added in the same order as one at a time
        for (; c + 4 <= ce; c += 4) {
          const float4 v0 = s_sorted[c], v1 = s_sorted[c + 1], v2 = s_sorted[c + 2],
                       v3 = s_sorted[c + 3];
          a.x += v0.x; a.y += v0.y; a.z += v0.z; a.w += v0.w;
          a.x += v1.x; a.y += v1.y; a.z += v1.z; a.w += v1.w;
          a.x += v2.x; a.y += v2.y; a.z += v2.z; a.w += v2.w;
          a.x += v3.x; a.y += v3.y; a.z += v3.z; a.w += v3.w;
        }
        for (; c < ce; ++c) {
          const float4 v = s_sorted[c];
          a.x += v.x;
          a.y += v.y;
          a.z += v.z;
          a.w += v.w;
        }
      } else {
        a = s_acc[tid];
      }
      S.slot_acc[vs0.y] = a;
    }
    __syncthreads();  // the LDS arrays are refilled by the next patch
    h0 = h1;
    g0 = g1;
    vs0 = vs1;
    in0 = in1;
    blend0 = blend1;
    t0 = t1;
    h1 = h2;
    g1 = g2;
    vs1 = vs2;
    has1 = has2;
  }
}

// Jacobi average from the slots (cheb: into xj; otherwise pos in place).
__global__ void k_patch_average(const SceneDev* __restrict__ scenes,
                                const WorkItem* __restrict__ map, int it) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  const int v = wi.first + threadIdx.x;
  if (v >= S.n_verts) return;
  const float4 out = jacobi_average_slots(S, v, it);
  if (S.cheb)
    S.xj[v] = out;
  else
    S.pos[v] = out;
}

// x_J of vertex g as staged by the residual: the Jacobi average in xj
// (Chebyshev on) or pos (off), written by k_patch_average.
__device__ __forceinline__ float4 stage_jacobi(const SceneDev& S, int g) {
  return S.cheb ? S.xj[g] : S.pos[g];
}

// Persistent patch residual, pipelined like k_patch_project: x_J (Chebyshev
// on) or pos (off) and the rest positions of a patch are staged in LDS; the
// sums (r, |C_H|, |C_D|) of each group of kPartialGroup patches form one
// partial of the scene: each lane adds its tets over the group's patches and
// the workgroup reduces once per group, in a fixed order.  Each range starts
// at a group boundary.
__global__ void __launch_bounds__(kPatchThreads)
    k_patch_residual(const SceneDev* __restrict__ scenes, const PatchItem* __restrict__ items,
                     const int2* __restrict__ ranges) {
  const int2 rg = ranges[blockIdx.x];
  if (rg.x >= rg.y) return;
  __shared__ float4 s_x[kPatchVerts];
  __shared__ float4 s_rest[kPatchVerts];
  const int tid = threadIdx.x;
  PatchItem h0 = items[rg.x];
  float4 x0 = make_float4(0.f, 0.f, 0.f, 0.f), r0 = x0;
  PatchTets t0;
  {
    const SceneDev& S = scenes[h0.scene];
    if (tid < h0.n_verts) {
      const int g = S.p_vert_ids[h0.vert_begin + tid];
      x0 = stage_jacobi(S, g);
      r0 = S.rest[g];
    }
    load_patch_tets(S, h0, true, t0);
  }
  bool has1 = rg.x + 1 < rg.y;
  PatchItem h1 = h0;
  int g1 = 0;
  if (has1) {
    h1 = items[rg.x + 1];
    if (tid < h1.n_verts) g1 = scenes[h1.scene].p_vert_ids[h1.vert_begin + tid];
  }
  // each lane's running sums over the current patch group, reduced over the
  // workgroup at the group's last patch (the range starts at a group boundary)
  float acc_r = 0.f, acc_h = 0.f, acc_d = 0.f;
  for (int i = rg.x; i < rg.y; ++i) {
    const SceneDev& S = scenes[h0.scene];
    if (tid < h0.n_verts) {
      s_x[tid] = x0;
      s_rest[tid] = r0;
    }
    __syncthreads();
    float4 x1 = make_float4(0.f, 0.f, 0.f, 0.f), r1 = x1;
    PatchTets t1;
    if (has1) {
      const SceneDev& S1 = scenes[h1.scene];
      if (tid < h1.n_verts) {
        x1 = stage_jacobi(S1, g1);
        r1 = S1.rest[g1];
      }
      load_patch_tets(S1, h1, true, t1);
    }
    const bool has2 = i + 2 < rg.y;
    PatchItem h2 = h1;
    int g2 = 0;
    if (has2) {
      h2 = items[i + 2];
      if (tid < h2.n_verts) g2 = scenes[h2.scene].p_vert_ids[h2.vert_begin + tid];
    }
#pragma unroll
    for (int j = 0; j < kPatchTetsPerThread; ++j) {
      if (tid + j * kPatchThreads < h0.n_tets) {
        const uint2 loc = t0.loc[j];
        const int l[4] = {int(loc.x & 0xffu), int((loc.x >> 8) & 0xffu),
                          int((loc.x >> 16) & 0xffu), int(loc.x >> 24)};
        Mat3 idm;
        float inv_det;
        rest_shape(f3(s_rest[l[0]]), f3(s_rest[l[1]]), f3(s_rest[l[2]]), f3(s_rest[l[3]]), idm,
                   inv_det);
        float r, ah, ad;
        residual_tet(S, f3(s_x[l[0]]), f3(s_x[l[1]]), f3(s_x[l[2]]), f3(s_x[l[3]]), idm, inv_det,
                     t0.lam[j], r, ah, ad);
        acc_r += r;
        acc_h += ah;
        acc_d += ad;
      }
    }
    if ((h0.patch + 1) % kPartialGroup == 0 || h0.patch + 1 == S.n_patches) {  // uniform
      double tot[3];
      block_sums(acc_r, acc_h, acc_d, S.monitor != 0, tot);
      if (tid == 0) store_partial(S, h0.patch / kPartialGroup, tot);
      acc_r = acc_h = acc_d = 0.f;
    }
    __syncthreads();
    h0 = h1;
    x0 = x1;
    r0 = r1;
    t0 = t1;
    h1 = h2;
    g1 = g2;
    has1 = has2;
  }
}

// One workgroup of kReduceThreads per scene: the controller update after
// iteration it from the scene's partials, stored into both state slots, and
// the iteration's log row.
__global__ void __launch_bounds__(kReduceThreads) k_reduce(const SceneDev* __restrict__ scenes,
                                                           int it) {
  __shared__ double red[kReduceScratch];
  const SceneDev& S = scenes[blockIdx.x];
  if (it >= S.iters) return;
  double tot[3];
  scene_sums(S, true, tot, red);
  if (threadIdx.x != 0) return;
  double L[6];
  const SceneCtrl c = ctrl_update(S, S.ctrl[(it + 1) & 1], it, tot, L);
  S.ctrl[0] = c;
  S.ctrl[1] = c;
  store_log_row(S, it, L);
}

// _chebyshev_update + _rotate_prev_with_save (neohookean_solver.py:131-142).
__global__ void k_blend(const SceneDev* __restrict__ scenes, const WorkItem* __restrict__ map) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  const int v = wi.first + threadIdx.x;
  if (v >= S.n_verts) return;
  const float omega = float(S.ctrl[0].omega);  // both slots hold the update of this iteration
  const float gamma = S.gamma;
  const float4 y4 = S.pos[v];
  const float3 y = f3(y4), xj = f3(S.xj[v]), pr = f3(S.prev[v]);
  const float3 x_new = gamma * (xj - y) + y;
  const float3 nw = pr + omega * (x_new - pr);
  S.prev[v] = make_float4(y.x, y.y, y.z, 0.f);
  S.pos[v] = make_float4(nw.x, nw.y, nw.z, y4.w);
}

// Blend of the last iteration of the fused sequence: y^n_iter into pos.
__global__ void k_final_blend(const SceneDev* __restrict__ scenes,
                              const WorkItem* __restrict__ map, int n_iter) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  const int v = wi.first + threadIdx.x;
  if (v >= S.n_verts || n_iter <= 0 || !S.fused) return;
  const int k = n_iter - 1;
  const float4 y = S.ybuf[k % 3][v];
  const float4 pr = S.ybuf[k == 0 ? 0 : (k - 1) % 3][v];  // y itself at k == 0
  // the last iteration's update comes from the standalone reduce (both slots)
  S.pos[v] = cheb_blend(S, y, S.xj[v], pr, float(S.ctrl[0].omega));
}

// advance (xpbd_base.py:148-160).
__global__ void k_advance(const SceneDev* __restrict__ scenes, const WorkItem* __restrict__ map) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  const int v = wi.first + threadIdx.x;
  if (v >= S.n_verts) return;
  advance_vertex(S, v, S.pos[v], S.xs[v]);
}

// ---- persistent substep (small scenes) -------------------------------------------

// A scene-wide barrier over the persistent substep's workgroups of one scene.
// Everything the workgroups hand to each other across it is written and read
// with agent-scope atomics (xstore4, xload4), so the
// barrier orders only their completion: every wave first waits until its own
// stores have completed (vmcnt 0), the workgroup barrier then holds thread 0
// until all waves have, and thread 0 counts the workgroup in on the scene's
// arrival word and waits, without blocking forever, until `target` arrivals
// are counted.  Loads issued after the closing workgroup barrier see every
// counted workgroup's stores.  No L2 write-back or invalidate is issued (an
// agent-scope release and acquire would be one of each, per workgroup, on
// the XCD that runs it).  The wait ends early when any workgroup has set the
// error word, and after `timeout` ticks of the 100 MHz wall clock it sets the
// error word itself (the workgroups are not all resident).  Returns false on
// either: the caller then leaves its loops.
__device__ __forceinline__ bool scene_barrier(gptr<unsigned> sync, gptr<unsigned> err,
                                              unsigned target, unsigned long long timeout,
                                              int* s_ok) {
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __hip_atomic_fetch_add(&sync[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long t0 = wall_clock64();
    while (__hip_atomic_load(&sync[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
        ok = 0;
        break;
      }
      if (wall_clock64() - t0 > timeout) {
        __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    *s_ok = ok;
  }
  __syncthreads();
  return *s_ok != 0;
}

// One wave's share of a mailbox wait: lane l polls residual group w0 + l's
// words of iteration `it` (components 0, and 1-2 when `all3`) with
// agent-scope atomic loads until none of the wave's is empty, so the values
// arrive with the wait, then sums them over the wave as scene_sums does (each
// lane's entry added to 0.0, then the lane-exchange butterfly): every lane
// gets the wave's sums.  Lanes without an entry add nothing.  The wave gives
// up when the error word is set, or sets it after `timeout` wall-clock ticks;
// returns 0 then (uniform over the wave), else 1.
__device__ __forceinline__ int mailbox_wave(const SceneDev& S, int it, bool all3, int w0,
                                            double acc[3], gptr<unsigned> err,
                                            unsigned long long timeout) {
  const int G = S.n_partials;
  const int e = w0 + int(threadIdx.x & 63);
  gptr<const unsigned long long> box =
      (gptr<const unsigned long long>)(S.mailbox + size_t(it & 1) * 3 * size_t(G));
  const int nc = all3 ? 3 : 1;
  unsigned long long u[3] = {0ull, 0ull, 0ull};
  int ok = 1;
  const unsigned long long t0 = wall_clock64();
  while (true) {
    bool have = true;
    if (e < G) {
      for (int c = 0; c < nc; ++c) {
        u[c] = __hip_atomic_load(box + size_t(c) * G + e, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
        have = have && u[c] != kMailboxEmpty;
      }
    }
    if (__all(have)) break;
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
      ok = 0;
      break;
    }
    if (wall_clock64() - t0 > timeout) {
      if ((threadIdx.x & 63) == 0)
        __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ok = 0;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    acc[c] = 0.0;
    if (c > 0 && !all3) continue;
    double sum = 0.0;
    if (e < G) sum += __builtin_bit_cast(double, u[c]);
    acc[c] = wave_sum_d(sum);
  }
  return ok;
}

// The sums of every residual group's partials of iteration `it` (component 0,
// and 1-2 when `all3`), waiting for each to be posted in the mailbox: wave w
// of the first kReduceThreads threads takes groups 64w..64w+63
// (mailbox_wave), then thread 0 adds the waves' sums in order, so the sums
// are those of scene_sums (thread t holding entry t), the same bits.  Returns
// false to every thread when a wave gave up.  Thread 0 gets the totals.  Ends
// with a workgroup barrier.
__device__ __forceinline__ bool mailbox_sums(const SceneDev& S, int it, bool all3, double tot[3],
                                             double* red, gptr<unsigned> err,
                                             unsigned long long timeout, int* s_wok) {
  constexpr int kW = kReduceThreads / 64;
  const int G = S.n_partials;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < kReduceThreads) {
    double acc[3] = {0.0, 0.0, 0.0};
    int ok = 1;
    if (wave * 64 < G) ok = mailbox_wave(S, it, all3, wave * 64, acc, err, timeout);
    if (lane == 0) {
      s_wok[wave] = ok;
      red[wave] = acc[0];
      red[kW + wave] = acc[1];
      red[2 * kW + wave] = acc[2];
    }
  }
  __syncthreads();
  bool all_ok = true;
#pragma unroll
  for (int w = 0; w < kW; ++w) all_ok = all_ok && s_wok[w] != 0;
  if (tid == 0) {
    tot[0] = tot[1] = tot[2] = 0.0;
    for (int w = 0; w < kW; ++w) {
      tot[0] += red[w];
      tot[1] += red[kW + w];
      tot[2] += red[2 * kW + w];
    }
  }
  __syncthreads();
  return all_ok;
}

// One whole substep of one scene (predict, every Jacobi iteration with its
// rho/omega update and Chebyshev blend, the final blend and advance) in one
// launch, for scenes of few patches, whose phased launches are bound by
// latency.  Workgroup g holds residual group g: patches 2g and 2g + 1, one per
// half of its kPersistThreads threads, each half working like a workgroup of
// the phased kernels.  The iterate, its predecessor, the Jacobi iterate and
// the rest positions of the group's vertices stay in LDS, and the
// multipliers in registers, for the whole substep; each patch predicts and
// blends its own copies of its vertices (the same bits in every patch that
// holds a vertex, as in the phased kernels).  Per iteration, one scene
// barrier after the slot partials are written, and one wait for the residual
// partials, which arrive by value in the scene's mailbox.  Every sum is taken
// in the phased kernels' order, so the results are bit-identical to them in
// deterministic mode.  With Chebyshev off the next iterate is x_J itself,
// and the residual pass and its wait run only for the convergence monitor.
// Requires the patch layout without isolated vertices, at least one
// iteration, and all of the launch's workgroups resident at once.
//
// kBatch = false: one scene (scenes[0]), workgroup g = residual group g, and
// the scene's own error word (sync[2]).  kBatch = true: the members of a
// batch side by side, workgroup b doing groups[b]; each scene's workgroups
// meet at that scene's barrier words only, and one error word serves the
// launch.  A member's results are the same bits as alone.
template <bool kBatch>
__global__ void __launch_bounds__(kPersistThreads)
    k_substep_persistent(const SceneDev* __restrict__ scenes, const PatchItem* __restrict__ items,
                         const PersistGroup* __restrict__ groups, unsigned* err_word,
                         unsigned long long timeout, int xcd_stride, int xcd) {
  int scene = 0, group = blockIdx.x;
  if constexpr (kBatch) {
    const PersistGroup d = groups[blockIdx.x];
    if (d.scene < 0) return;  // a hole in the per-XCD layout
    scene = d.scene;
    group = d.group;
    items += d.item0;
  } else {
    // Workgroups are dispatched to the XCDs round-robin by block index: with
    // xcd_stride = the XCD count, the scene's group g is block
    // g * xcd_stride + xcd, so all of its workgroups share one L2.  The other
    // blocks of the grid (xcd_stride * groups) leave at once.
    if (int(blockIdx.x) % xcd_stride != xcd) return;
    group = int(blockIdx.x) / xcd_stride;
  }
  const SceneDev& S = scenes[scene];
  gptr<unsigned> err;
  if constexpr (kBatch)
    err = (gptr<unsigned>)err_word;
  else
    err = S.sync + 2;
  const int tid = threadIdx.x;
  const int half = tid / kPatchThreads, lt = tid % kPatchThreads;
  const int p = group * kPartialGroup + half;
  const bool active = p < S.n_patches;
  const bool csr = S.csr != 0;
  const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
  __shared__ float4 s_y[kPartialGroup][kPatchVerts];     // y^k of the patch's vertices
  __shared__ float4 s_prev[kPartialGroup][kPatchVerts];  // y^(k-1) (y^0 at k = 0, 1)
  __shared__ float4 s_x[kPartialGroup][kPatchVerts];     // x_J^k
  __shared__ float4 s_rest[kPartialGroup][kPatchVerts];  // rest position | corner run base
  __shared__ float4 s_sorted[kPartialGroup][4 * kPatchTets];
  __shared__ double s_red[kReduceScratch];
  __shared__ float s_omega;
  __shared__ int s_ok;
  __shared__ int s_wok[kReduceThreads / 64];
  __shared__ SceneCtrl s_ctrl;
  gptr<unsigned> sync = S.sync;
  float4* my_sorted = s_sorted[half];

  // the patch's vertices: ids, LDS corner runs, slot ranges; the thread's tets
  PatchItem h;
  h.n_tets = 0;
  h.n_verts = 0;
  if (active) h = items[p];
  int g = 0, sb = 0, se = 0;
  int2 vs = make_int2(0, 0);
  float val = 0.f;
  if (lt < h.n_verts) {
    g = S.p_vert_ids[h.vert_begin + lt];
    vs = S.p_vert_sum[h.vert_begin + lt];
    sb = S.vslot_off[g];
    se = S.vslot_off[g + 1];
    val = S.valence[g];
  }
  PatchTets t;
  load_patch_tets(S, h, false, t, lt);  // the multipliers start at 0 (:458-459)
  // predict + collide (each patch for its own copy of its vertices)
  float4 xs0 = zero;
  if (lt < h.n_verts) {
    const float4 pw = S.pos[g];
    xs0 = S.xs[g];
    float3 vel = f3(S.vel[g]);
    const float3 x = predict_vertex(S, f3(xs0), vel, pw.w, true, true, f3(pw));
    const float4 y0 = make_float4(x.x, x.y, x.z, pw.w);
    s_y[half][lt] = y0;
    s_prev[half][lt] = y0;
    const float4 r = S.rest[g];
    s_rest[half][lt] = make_float4(r.x, r.y, r.z, __int_as_float(vs.x & 0xffff));
  }
  if (tid == 0) s_ctrl = S.ctrl[0];  // between substeps both slots hold the same state
  __syncthreads();
  // rest shape of the thread's tets, kept in registers for the substep
  Mat3 idm[kPatchTetsPerThread];
  float inv_det[kPatchTetsPerThread];
#pragma unroll
  for (int j = 0; j < kPatchTetsPerThread; ++j) {
    const uint2 loc = t.loc[j];
    const int l[4] = {int(loc.x & 0xffu), int((loc.x >> 8) & 0xffu), int((loc.x >> 16) & 0xffu),
                      int(loc.x >> 24)};
    rest_shape(f3(s_rest[half][l[0]]), f3(s_rest[half][l[1]]), f3(s_rest[half][l[2]]),
               f3(s_rest[half][l[3]]), idm[j], inv_det[j]);
  }
  const unsigned G = unsigned(S.n_partials);
  unsigned barriers = 0;
  bool ok = true;
  const bool writer = group == 0;
  // Chebyshev on: the blend folded into the iteration (the fused sequence);
  // off: the next iterate is x_J itself (post_solve writes pred_x).  The
  // residual pass runs for Chebyshev or the convergence monitor, as in the
  // phased kernels.
  const bool fused = S.cheb != 0;
  const bool need_res = S.cheb || S.monitor;
  for (int it = 0; it < S.iters && ok; ++it) {
    if (it > 0) {
      if (need_res) {
        // rho/omega update (and the log row) after iteration it - 1 from
        // every group's partial, as each arrives in the mailbox
        auto update = [&](const double tot[3]) {
          double L[6];
          const SceneCtrl c = ctrl_update(S, s_ctrl, it - 1, tot, L);
          s_ctrl = c;
          s_omega = float(c.omega);
          if (writer) {
            S.ctrl[(it - 1) & 1] = c;
            store_log_row(S, it - 1, L);
          }
        };
        if (G <= 64) {
          // One wave holds every group, and the other waves of mailbox_sums
          // would add exact zeros: wave 0 waits and sums alone, thread 0
          // adds the waves' sums in the same order and updates the
          // controller, and one workgroup barrier follows instead of three.
          if (tid < 64) {
            double acc[3];
            const int ok0 = mailbox_wave(S, it - 1, S.monitor != 0, 0, acc, err, timeout);
            if (tid == 0) {
              s_ok = ok0;
              if (ok0) {
                double tot[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                  tot[c] = 0.0;
                  for (int w = 0; w < kReduceThreads / 64; ++w) tot[c] += w == 0 ? acc[c] : 0.0;
                }
                update(tot);
              }
            }
          }
          __syncthreads();
          ok = s_ok != 0;
          if (!ok) break;
        } else {
          double tot[3];
          ok = mailbox_sums(S, it - 1, S.monitor != 0, tot, s_red, err, timeout, s_wok);
          if (!ok) break;
          if (tid == 0) update(tot);
          __syncthreads();
        }
      }
      if (lt < h.n_verts) {
        if (fused) {  // y^it from y^(it-1), x_J^(it-1) and y^(it-2)
          const float4 y = s_y[half][lt];
          s_y[half][lt] = cheb_blend(S, y, s_x[half][lt], s_prev[half][lt], s_omega);
          s_prev[half][lt] = y;
        } else {
          s_y[half][lt] = s_x[half][lt];
        }
      }
      __syncthreads();
    }
    // projection: corner corrections into LDS, summed per vertex into its slot
    if (!csr && lt < h.n_verts) my_sorted[lt] = zero;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPatchTetsPerThread; ++j) {
      const int tt = lt + j * kPatchThreads;
      if (tt < h.n_tets) {
        const uint2 loc = t.loc[j];
        const int l[4] = {int(loc.x & 0xffu), int((loc.x >> 8) & 0xffu),
                          int((loc.x >> 16) & 0xffu), int(loc.x >> 24)};
        const float4 q[4] = {s_y[half][l[0]], s_y[half][l[1]], s_y[half][l[2]], s_y[half][l[3]]};
        TetProjection P;
        project_tet(S, q, idm[j], inv_det[j], t.lam[j], P);
        if (csr) {
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int rank = int((loc.y >> (8 * k)) & 0xffu);
            my_sorted[__float_as_int(s_rest[half][l[k]].w) + rank] =
                make_float4(P.dp[k].x, P.dp[k].y, P.dp[k].z, P.cnt[k]);
          }
        } else {
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            atomicAdd(&my_sorted[l[k]].x, P.dp[k].x);
            atomicAdd(&my_sorted[l[k]].y, P.dp[k].y);
            atomicAdd(&my_sorted[l[k]].z, P.dp[k].z);
            if (!S.block_mode) atomicAdd(&my_sorted[l[k]].w, P.cnt[k]);
          }
        }
      }
    }
    __syncthreads();
    if (lt < h.n_verts) {
      float4 a;
      if (csr) {
        a = zero;
        const int cb = vs.x & 0xffff, ce = cb + ((vs.x >> 16) & 0x7fff);
        int c = cb;
        for (; c + 4 <= ce; c += 4) {
          const float4 v0 = my_sorted[c], v1 = my_sorted[c + 1], v2 = my_sorted[c + 2],
                       v3 = my_sorted[c + 3];
          a.x += v0.x; a.y += v0.y; a.z += v0.z; a.w += v0.w;
          a.x += v1.x; a.y += v1.y; a.z += v1.z; a.w += v1.w;
          a.x += v2.x; a.y += v2.y; a.z += v2.z; a.w += v2.w;
          a.x += v3.x; a.y += v3.y; a.z += v3.z; a.w += v3.w;
        }
        for (; c < ce; ++c) {
          const float4 v = my_sorted[c];
          a.x += v.x;
          a.y += v.y;
          a.z += v.z;
          a.w += v.w;
        }
      } else {
        a = my_sorted[lt];
      }
      xstore4(S.slot_acc + vs.y, a);
    }
    ok = scene_barrier(sync, err, ++barriers * G, timeout, &s_ok);
    if (!ok) break;
    // Every group has taken the partials of iteration it - 1 before arriving
    // here, so this group's words of that parity are emptied for it + 1.
    if (need_res && tid < 3)
      __hip_atomic_store((gptr<unsigned long long>)S.mailbox +
                             (size_t((it + 1) & 1) * 3 + tid) * G + group,
                         kMailboxEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // Jacobi average of the patch's vertices from every slot of each
    if (lt < h.n_verts) s_x[half][lt] = jacobi_from(S, slot_sum<true>(S, sb, se), val, s_y[half][lt]);
    __syncthreads();
    if (need_res) {
      // residual at x_J: lane sums over the group's two patches in order
      float rv[kPatchTetsPerThread], hv[kPatchTetsPerThread], dv[kPatchTetsPerThread];
#pragma unroll
      for (int j = 0; j < kPatchTetsPerThread; ++j) {
        rv[j] = hv[j] = dv[j] = 0.f;
        if (lt + j * kPatchThreads < h.n_tets) {
          const uint2 loc = t.loc[j];
          const int l[4] = {int(loc.x & 0xffu), int((loc.x >> 8) & 0xffu),
                            int((loc.x >> 16) & 0xffu), int(loc.x >> 24)};
          residual_tet(S, f3(s_x[half][l[0]]), f3(s_x[half][l[1]]), f3(s_x[half][l[2]]),
                       f3(s_x[half][l[3]]), idm[j], inv_det[j], t.lam[j], rv[j], hv[j], dv[j]);
        }
      }
      // the second patch's terms go to LDS (the corner array is free now)
      float* xch = reinterpret_cast<float*>(s_sorted[0]);
      if (half == 1) {
#pragma unroll
        for (int j = 0; j < kPatchTetsPerThread; ++j) {
          xch[(3 * j + 0) * kPatchThreads + lt] = rv[j];
          xch[(3 * j + 1) * kPatchThreads + lt] = hv[j];
          xch[(3 * j + 2) * kPatchThreads + lt] = dv[j];
        }
      }
      __syncthreads();
      float acc_r = 0.f, acc_h = 0.f, acc_d = 0.f;
      if (half == 0) {
#pragma unroll
        for (int j = 0; j < kPatchTetsPerThread; ++j) {
          acc_r += rv[j];
          acc_h += hv[j];
          acc_d += dv[j];
        }
#pragma unroll
        for (int j = 0; j < kPatchTetsPerThread; ++j) {
          acc_r += xch[(3 * j + 0) * kPatchThreads + lt];
          acc_h += xch[(3 * j + 1) * kPatchThreads + lt];
          acc_d += xch[(3 * j + 2) * kPatchThreads + lt];
        }
      }
      double tot[3];
      block_sums(acc_r, acc_h, acc_d, S.monitor != 0, tot);
      if (tid == 0) post_partials(S, it, group, tot);
    }
  }
  double tot[3];
  if (ok && need_res) ok = mailbox_sums(S, S.iters - 1, true, tot, s_red, err, timeout, s_wok);
  if (ok) {
    // the update after the last iteration (k_reduce), the final blend
    // (k_final_blend) or x_J itself, and advance (k_advance)
    if (tid == 0 && need_res) {
      double L[6];
      const SceneCtrl c = ctrl_update(S, s_ctrl, S.iters - 1, tot, L);
      s_omega = float(c.omega);
      if (writer) {
        S.ctrl[0] = c;
        S.ctrl[1] = c;
        store_log_row(S, S.iters - 1, L);
      }
    }
    __syncthreads();
    if (lt < h.n_verts && vs.x < 0) {  // the vertex's owner patch writes it
      const float4 pf =
          fused ? cheb_blend(S, s_y[half][lt], s_x[half][lt], s_prev[half][lt], s_omega)
                : s_x[half][lt];
      S.pos[g] = pf;
      if (fused) S.xj[g] = s_x[half][lt];
      advance_vertex(S, g, pf, xs0);
    }
#pragma unroll
    for (int j = 0; j < kPatchTetsPerThread; ++j)
      if (lt + j * kPatchThreads < h.n_tets) S.lam[h.tet_begin + lt + j * kPatchThreads] = t.lam[j];
  }
  // the last workgroup out (every other has finished with the barrier words
  // and the mailbox) resets them for the next launch
  __syncthreads();
  if (tid == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned done =
        __hip_atomic_fetch_add(&sync[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_ok = done == G - 1;
  }
  __syncthreads();
  if (s_ok) {
    gptr<unsigned long long> box = (gptr<unsigned long long>)S.mailbox;
    for (int i = tid; i < kMailboxWords * int(G); i += kPersistThreads)
      __hip_atomic_store(box + i, kMailboxEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (tid == 0) {
      __hip_atomic_store(&sync[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&sync[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Writes one scene table entry from the launch's arguments.  Being a launch
// on the scene's stream, it takes effect after the substeps queued before it
// and before those queued after, with no wait on the host.
__global__ void k_store_scene(SceneDev* __restrict__ dst, const SceneDev src) {
  if (threadIdx.x == 0) *dst = src;
}

// Copies between a caller-order (V,3) float array and an internal float4
// stream: dst[old_to_new[i]] <- src[i] (scatter) or dst[i] <- src[old_to_new[i]].
__global__ void k_scatter_xyz(float4* __restrict__ dst, const float* __restrict__ src,
                              const int* __restrict__ old_to_new, int n, int keep_w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = old_to_new[i];
  const float w = keep_w ? dst[j].w : 0.f;
  dst[j] = make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], w);
}

__global__ void k_gather_xyz(float* __restrict__ dst, const float4* __restrict__ src,
                             const int* __restrict__ old_to_new, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 a = src[old_to_new[i]];
  dst[3 * i] = a.x;
  dst[3 * i + 1] = a.y;
  dst[3 * i + 2] = a.z;
}

// Sets *flag when any position is not finite.
__global__ void k_check_finite(const float4* __restrict__ pos, int n, int* flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 a = pos[i];
  if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(a.z)) atomicOr(flag, 1);
}

// ---- the no-allocation work-list array ------------------------------------------

// Work lists and scene table of a batch re-targeted without allocating
// (xpbd_batch_set_members_noalloc): a fixed array in device memory, one copy
// per GPU, held until the process ends.  Laid out as the scene table
// (XPBD_NOALLOC_MAX_SCENES entries) followed by XPBD_NOALLOC_LIST_BYTES of
// work lists.
constexpr size_t kNoallocTableBytes =
    (sizeof(SceneDev) * XPBD_NOALLOC_MAX_SCENES + 255) / 256 * 256;
constexpr size_t kNoallocBytes = kNoallocTableBytes + XPBD_NOALLOC_LIST_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, void** table, void** lists) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    // once per device, with that device current: the array's address on this
    // GPU (the host-side name of the array is not a device address)
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *table = g_scratch_ptr[device];
  *lists = static_cast<unsigned char*>(g_scratch_ptr[device]) + kNoallocTableBytes;
  return hipSuccess;
}

// ---- launch wrappers ---------------------------------------------------------

hipError_t launch_predict(const SceneDev* scenes, const WorkItem* vmap, int nblocks, int bs,
                          int external, int collide, int save_prev, hipStream_t st) {
  k_predict<<<nblocks, bs, 0, st>>>(scenes, vmap, external, collide, save_prev);
  return hipGetLastError();
}
hipError_t launch_project(const SceneDev* scenes, const WorkItem* tmap, int nblocks, int bs,
                          int first, hipStream_t st) {
  k_project<<<nblocks, bs, 0, st>>>(scenes, tmap, first);
  return hipGetLastError();
}
hipError_t launch_average(const SceneDev* scenes, const WorkItem* vmap, int nblocks, int bs,
                          hipStream_t st) {
  k_average<<<nblocks, bs, 0, st>>>(scenes, vmap);
  return hipGetLastError();
}
hipError_t launch_residual(const SceneDev* scenes, const WorkItem* tmap, int nblocks, int bs,
                           hipStream_t st) {
  k_residual<<<nblocks, bs, 0, st>>>(scenes, tmap);
  return hipGetLastError();
}
hipError_t launch_patch_project(const SceneDev* scenes, const PatchItem* items,
                                const int2* ranges, int nblocks, int it, bool multi_scene,
                                hipStream_t st, hipEvent_t start, hipEvent_t stop) {
  if (nblocks <= 0) return hipSuccess;
  if (start || stop) {
    // events recorded by the dispatch itself: the kernel's own start and end
    void* args[] = {&scenes, &items, &ranges, &it};
    const void* fn = multi_scene ? reinterpret_cast<const void*>(&k_patch_project<true>)
                                 : reinterpret_cast<const void*>(&k_patch_project<false>);
    return hipExtLaunchKernel(fn, dim3(nblocks), dim3(kPatchThreads), args, 0, st, start, stop,
                              0);
  }
  if (multi_scene)
    k_patch_project<true><<<nblocks, kPatchThreads, 0, st>>>(scenes, items, ranges, it);
  else
    k_patch_project<false><<<nblocks, kPatchThreads, 0, st>>>(scenes, items, ranges, it);
  return hipGetLastError();
}
hipError_t launch_patch_average(const SceneDev* scenes, const WorkItem* vmap, int nblocks,
                                int bs, int it, hipStream_t st) {
  k_patch_average<<<nblocks, bs, 0, st>>>(scenes, vmap, it);
  return hipGetLastError();
}
hipError_t launch_patch_residual(const SceneDev* scenes, const PatchItem* items,
                                 const int2* ranges, int nblocks, hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  k_patch_residual<<<nblocks, kPatchThreads, 0, st>>>(scenes, items, ranges);
  return hipGetLastError();
}
hipError_t patch_kernel_occupancy(int* project_blocks_per_cu, int* residual_blocks_per_cu) {
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      project_blocks_per_cu, k_patch_project<true>, kPatchThreads, 0);
  if (e != hipSuccess) return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(residual_blocks_per_cu, k_patch_residual,
                                                      kPatchThreads, 0);
}
hipError_t launch_reduce(const SceneDev* scenes, int n_scenes, int it, hipStream_t st) {
  k_reduce<<<n_scenes, kReduceThreads, 0, st>>>(scenes, it);
  return hipGetLastError();
}
hipError_t launch_blend(const SceneDev* scenes, const WorkItem* vmap, int nblocks, int bs,
                        hipStream_t st) {
  k_blend<<<nblocks, bs, 0, st>>>(scenes, vmap);
  return hipGetLastError();
}
hipError_t launch_final_blend(const SceneDev* scenes, const WorkItem* vmap, int nblocks, int bs,
                              int n_iter, hipStream_t st) {
  k_final_blend<<<nblocks, bs, 0, st>>>(scenes, vmap, n_iter);
  return hipGetLastError();
}
hipError_t launch_substep_persistent(const SceneDev* scenes, const PatchItem* items, int n_groups,
                                    unsigned long long timeout_ticks, int xcd_stride, int xcd,
                                    hipStream_t st) {
  if (n_groups <= 0 || xcd_stride < 1 || xcd < 0 || xcd >= xcd_stride)
    return hipErrorInvalidValue;
  k_substep_persistent<false><<<n_groups * xcd_stride, kPersistThreads, 0, st>>>(
      scenes, items, nullptr, nullptr, timeout_ticks, xcd_stride, xcd);
  return hipGetLastError();
}
hipError_t launch_batch_substep_persistent(const SceneDev* scenes, const PatchItem* items,
                                          const PersistGroup* groups, int n_groups,
                                          unsigned* err, unsigned long long timeout_ticks,
                                          hipStream_t st) {
  if (n_groups <= 0 || !groups || !err) return hipErrorInvalidValue;
  k_substep_persistent<true><<<n_groups, kPersistThreads, 0, st>>>(scenes, items, groups, err,
                                                                   timeout_ticks, 1, 0);
  return hipGetLastError();
}
hipError_t persistent_occupancy(int* blocks_per_cu) {
  int a = 0, b = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_substep_persistent<false>,
                                                              kPersistThreads, 0);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_substep_persistent<true>,
                                                   kPersistThreads, 0);
  *blocks_per_cu = a < b ? a : b;
  return e;
}
hipError_t launch_store_scene(SceneDev* dst, const SceneDev& src, hipStream_t st) {
  k_store_scene<<<1, 64, 0, st>>>(dst, src);
  return hipGetLastError();
}
hipError_t launch_advance(const SceneDev* scenes, const WorkItem* vmap, int nblocks, int bs,
                          hipStream_t st) {
  k_advance<<<nblocks, bs, 0, st>>>(scenes, vmap);
  return hipGetLastError();
}
hipError_t launch_scatter_xyz(float4* dst, const float* src, const int* old_to_new, int n,
                              int keep_w, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_scatter_xyz<<<(n + 255) / 256, 256, 0, st>>>(dst, src, old_to_new, n, keep_w);
  return hipGetLastError();
}
hipError_t launch_gather_xyz(float* dst, const float4* src, const int* old_to_new, int n,
                             hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_gather_xyz<<<(n + 255) / 256, 256, 0, st>>>(dst, src, old_to_new, n);
  return hipGetLastError();
}
hipError_t launch_check_finite(const float4* pos, int n, int* flag, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_check_finite<<<(n + 255) / 256, 256, 0, st>>>(pos, n, flag);
  return hipGetLastError();
}

}  // namespace xpbd
