// gfx950 kernels of the XPBD Neo-Hookean substep (one launch per phase).
//
// Every kernel walks a work map: workgroup b handles elements
// [map[b].first, map[b].first + blockDim.x) of scene map[b].scene, whose
// buffers and parameters are read from a device-resident scene table.  One
// solver is a table of one scene; the batched form packs many ragged scenes
// into the same launches.
#include "xpbd_device.h"
#include "xpbd_kernels.h"

namespace xpbd {

namespace {

__device__ __forceinline__ void atomic_add3(float4* dst, float3 v) {
  unsafeAtomicAdd(&dst->x, v.x);
  unsafeAtomicAdd(&dst->y, v.y);
  unsafeAtomicAdd(&dst->z, v.z);
}

// Sum of a float over the 64 lanes of a wavefront (butterfly through DPP /
// ds_swizzle lane exchanges); every lane receives the total.
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
  float3 x = f3(p);
  float3 vel = f3(S.vel[v]);
  const float w = p.w;
  if (external) {
    const float dt = S.dt;
    if (w > 0.f) vel = vel + make_float3(S.g[0], S.g[1], S.g[2]) * dt;
    x = f3(S.xs[v]) + vel * dt;
    S.vel[v] = make_float4(vel.x, vel.y, vel.z, 0.f);
  }
  if (collide && w > 0.f) {
    for (int s = 0; s < S.n_sdf; ++s) x = sdf_project(S.sdf[s], x, vel);
  }
  p = make_float4(x.x, x.y, x.z, w);
  S.pos[v] = p;
  if (save_prev) S.prev[v] = p;
}

// Per-tet projection of one Jacobi iteration.
//  block mode: solve_neohookean_block (neohookean_solver.py:358-451)
//  split mode: solve_hydrostatic (:196-254) then solve_deviatoric (:144-194),
//              evaluated on the same positions and summed per vertex.
// `first` marks the first iteration of a substep: the multipliers start at 0
// (neohookean_solver.py:458-459) and the substep's Chebyshev control resets
// (:461, :466).
__global__ void k_project(const SceneDev* __restrict__ scenes, const WorkItem* __restrict__ map,
                          int first) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  if (first && wi.first == 0 && threadIdx.x == 0) {
    S.ctrl->have_prev = 0;
    S.ctrl->omega = 1.0;
  }
  const int t = wi.first + threadIdx.x;
  if (t >= S.n_tets) return;
  const int4 id = S.tet[t];
  const float4 q0 = S.pos[id.x], q1 = S.pos[id.y], q2 = S.pos[id.z], q3 = S.pos[id.w];
  const float2 d2 = S.dm2[t];
  const Mat3 idm = load_inv_dm(S.dm0[t], S.dm1[t], d2);
  const float V0 = d2.y;
  float2 lam = first ? make_float2(0.f, 0.f) : S.lam[t];
  const float w[4] = {q0.w, q1.w, q2.w, q3.w};
  const Mat3 F = deformation_gradient(f3(q0), f3(q1), f3(q2), f3(q3), idm);
  const float dt = S.dt;
  float3 dp[4];
  float cnt[4] = {1.f, 1.f, 1.f, 1.f};
  if (S.block_mode) {
    const float C_D = norm_sqr(F) - 3.f;
    const float C_H = det3(F) - 1.f;
    Mat3 gfd;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) gfd.m[i][j] = F.m[i][j] * 2.f;
    float3 gD[4], gH[4];
    vertex_grads(gfd, idm, gD);
    vertex_grads(ddet_dF(F), idm, gH);
    float A11 = w[0] * dot(gD[0], gD[0]) + w[1] * dot(gD[1], gD[1]) + w[2] * dot(gD[2], gD[2]) +
                w[3] * dot(gD[3], gD[3]);
    float A22 = w[0] * dot(gH[0], gH[0]) + w[1] * dot(gH[1], gH[1]) + w[2] * dot(gH[2], gH[2]) +
                w[3] * dot(gH[3], gH[3]);
    const float A12 = w[0] * dot(gD[0], gH[0]) + w[1] * dot(gD[1], gH[1]) +
                      w[2] * dot(gD[2], gH[2]) + w[3] * dot(gD[3], gH[3]);
    const float aD = (1.f / (S.lame2 * V0)) / (dt * dt);
    const float aH = (1.f / (S.lame1 * V0)) / (dt * dt);
    A11 += aD;
    A22 += aH;
    const float b1 = -(C_D + aD * lam.x);
    const float b2 = -(C_H + aH * lam.y);
    float dlD, dlH;
    solve2x2_with_fallback(A11, A12, A22, b1, b2, dlD, dlH);
    lam.x += dlD;
    lam.y += dlH;
#pragma unroll
    for (int k = 0; k < 4; ++k) dp[k] = w[k] * (gD[k] * dlD + gH[k] * dlH);
  } else {
    // hydrostatic constraint C_H = det F - 1
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
      const float at = (1.f / (S.lame1 * V0)) / (dt * dt);
      dlH = -(C_H + at * lam.y) / (wsH + at);
      dlH *= S.hydro_relax;
      lam.y += dlH;
    }
    // deviatoric constraint C_D = |F|^2 - 3
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
      const float at = (1.f / (S.lame2 * V0)) / (dt * dt);
      dlD = -(C_D + at * lam.x) / (wsD + at);
      dlD *= S.dev_relax;
      lam.x += dlD;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float3 a = make_float3(0.f, 0.f, 0.f);
      if (okH) a = a + gH[k] * (dlH * w[k]);
      if (okD) a = a + gD[k] * (dlD * w[k]);
      dp[k] = a;
      cnt[k] = float(okH + okD);
    }
  }
  S.lam[t] = lam;
  if (S.csr) {
    float4* c = S.corner + 4 * size_t(t);
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = make_float4(dp[k].x, dp[k].y, dp[k].z, cnt[k]);
  } else {
    const int ids[4] = {id.x, id.y, id.z, id.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      atomic_add3(S.dacc + ids[k], dp[k]);
      if (!S.block_mode && cnt[k] != 0.f) unsafeAtomicAdd(&S.dacc[ids[k]].w, cnt[k]);
    }
  }
}

// pre_solve / post_solve (neohookean_solver.py:109-119): x_J = y + delta_x /
// cons_num.  With Chebyshev on, x_J goes to xj (y stays in pos for the blend);
// otherwise pos is updated in place.  Atomic accumulators are cleared here
// for the next iteration.
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
    cnt = acc.w;
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

// compute_xpbd_residual (neohookean_solver.py:298-329) and the monitor sums
// of compute_hydrostatic_error / compute_deviatoric_error (:256-296) on x_J;
// the last workgroup of each scene then performs the host-side rho/omega
// update of solve_step (:492-504, :537-546) in double on the device.
__global__ void k_residual(const SceneDev* __restrict__ scenes, const WorkItem* __restrict__ map,
                           int it) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  const int t = wi.first + threadIdx.x;
  const float4* X = S.cheb ? S.xj : S.pos;
  float r = 0.f, ah = 0.f, ad = 0.f;
  if (t < S.n_tets) {
    const int4 id = S.tet[t];
    const float2 d2 = S.dm2[t];
    const Mat3 idm = load_inv_dm(S.dm0[t], S.dm1[t], d2);
    const Mat3 F = deformation_gradient(f3(X[id.x]), f3(X[id.y]), f3(X[id.z]), f3(X[id.w]), idm);
    const float C_D = norm_sqr(F) - 3.f;
    const float C_H = det3(F) - 1.f;
    if (S.cheb) {
      const float2 lam = S.lam[t];
      const float dt = S.dt;
      const float aD = (1.f / (S.lame2 * d2.y)) / (dt * dt);
      const float aH = (1.f / (S.lame1 * d2.y)) / (dt * dt);
      const float rD = C_D + aD * lam.x;
      const float rH = C_H + aH * lam.y;
      r = sqrtf(rD * rD + rH * rH);
    }
    ah = fabsf(C_H);
    ad = fabsf(C_D);
  }
  // workgroup sum: wavefront butterfly, then one value per wave through LDS
  __shared__ float red[3][16];
  __shared__ int last;
  r = wave_sum(r);
  ah = wave_sum(ah);
  ad = wave_sum(ad);
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
    double* part = S.partial + 3 * size_t(wi.first / S.tet_block);
    part[0] = s0;
    part[1] = s1;
    part[2] = s2;
    const unsigned done =
        __hip_atomic_fetch_add(S.ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (done == unsigned(S.n_tet_blocks - 1));
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // Deterministic fixed-order sum of this scene's partials: strided per
  // thread, butterfly per wave, then waves in order.
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int b = threadIdx.x; b < S.n_tet_blocks; b += blockDim.x) {
    const double* part = S.partial + 3 * size_t(b);
    s0 += part[0];
    s1 += part[1];
    s2 += part[2];
  }
  s0 = wave_sum_d(s0);
  s1 = wave_sum_d(s1);
  s2 = wave_sum_d(s2);
  __shared__ double dred[3][16];
  if (lane == 0) {
    dred[0][wave] = s0;
    dred[1][wave] = s1;
    dred[2][wave] = s2;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double tot[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < nwave; ++i) {
    tot[0] += dred[0][i];
    tot[1] += dred[1][i];
    tot[2] += dred[2][i];
  }
  SceneCtrl* c = S.ctrl;
  const double cnt = double(S.n_tets);
  double r_hat = 0.0;
  double omega = 1.0;
  double rho = c->rho_hat;
  if (S.cheb) {
    r_hat = S.n_tets > 0 ? tot[0] / cnt : 0.0;
    if (c->have_prev && c->prev_r_hat > 1e-12) {
      double rho_inst = r_hat / (c->prev_r_hat + 1e-12);
      rho_inst = fmax(0.00001, fmin(0.99999, rho_inst));
      rho = (1.0 - 0.2) * rho + 0.2 * rho_inst;
      c->rho_hat = rho;
    }
    c->prev_r_hat = r_hat;
    c->have_prev = 1;
    const double rho2 = rho * rho;
    if (it < S.warmup)
      omega = 1.0;
    else if (it == S.warmup)
      omega = 2.0 / (2.0 - rho2);
    else
      omega = 4.0 / (4.0 - rho2 * c->omega);
    c->omega = omega;
    c->r_hat = r_hat;
  }
  double* L = S.log + 6 * size_t(it);
  L[0] = it;
  L[1] = r_hat;
  L[2] = rho;
  L[3] = omega;
  L[4] = S.monitor && S.n_tets > 0 ? tot[1] / cnt : 0.0;
  L[5] = S.monitor && S.n_tets > 0 ? tot[2] / cnt : 0.0;
  __hip_atomic_store(S.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// _chebyshev_update + _rotate_prev_with_save (neohookean_solver.py:131-142).
__global__ void k_blend(const SceneDev* __restrict__ scenes, const WorkItem* __restrict__ map) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  const int v = wi.first + threadIdx.x;
  if (v >= S.n_verts) return;
  const float omega = float(S.ctrl->omega);
  const float gamma = S.gamma;
  const float4 y4 = S.pos[v];
  const float3 y = f3(y4), xj = f3(S.xj[v]), pr = f3(S.prev[v]);
  const float3 x_new = gamma * (xj - y) + y;
  const float3 nw = pr + omega * (x_new - pr);
  S.prev[v] = make_float4(y.x, y.y, y.z, 0.f);
  S.pos[v] = make_float4(nw.x, nw.y, nw.z, y4.w);
}

// advance (xpbd_base.py:148-160).
__global__ void k_advance(const SceneDev* __restrict__ scenes, const WorkItem* __restrict__ map) {
  const WorkItem wi = map[blockIdx.x];
  const SceneDev& S = scenes[wi.scene];
  const int v = wi.first + threadIdx.x;
  if (v >= S.n_verts) return;
  const float4 p = S.pos[v];
  const float4 x = S.xs[v];
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

// Copies between a caller-order (V,3) float array and an internal float4
// stream: dst[perm[i]] <- src[i] (scatter) or dst[i] <- src[perm[i]] (gather).
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

// Returns in *flag a nonzero value when any position is not finite.
__global__ void k_check_finite(const float4* __restrict__ pos, int n, int* flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 a = pos[i];
  if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(a.z)) atomicOr(flag, 1);
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
                           int it, hipStream_t st) {
  k_residual<<<nblocks, bs, 0, st>>>(scenes, tmap, it);
  return hipGetLastError();
}
hipError_t launch_blend(const SceneDev* scenes, const WorkItem* vmap, int nblocks, int bs,
                        hipStream_t st) {
  k_blend<<<nblocks, bs, 0, st>>>(scenes, vmap);
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
