// Host side of libxpbd_hip.so: the C ABI of include/xpbd_abi.h, mesh set-up
// (reordering, rest-shape data, lumped masses, adjacency), and the launch
// sequence of a substep (one kernel per phase, optionally captured in a
// hipGraph).
//
// Set-up arithmetic restates the reference's f32 initialisation kernels
// (neohookean_solver.py:97-107, xpbd_base.py:89-137) on the host.  It is
// compiled without floating-point contraction so that each f32 operation is
// rounded as written.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/xpbd_abi.h"
#include "xpbd_device.h"
#include "xpbd_kernels.h"

using xpbd::SceneCtrl;
using xpbd::SceneDev;
using xpbd::SdfDev;
using xpbd::WorkItem;
using xpbd::PatchItem;
using xpbd::PersistGroup;
using xpbd::kPatchTets;
using xpbd::kPatchVerts;

namespace {

thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return set_error(e_ == hipErrorOutOfMemory ? XPBD_ENOMEM : XPBD_EHIP, "%s: %s (%s:%d)", \
                       #expr, hipGetErrorString(e_), __FILE__, __LINE__);                \
  } while (0)

#define TRY(expr)               \
  do {                          \
    int r_ = (expr);            \
    if (r_ != XPBD_OK) return r_; \
  } while (0)

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;    // elements in use
  size_t cap = 0;  // elements allocated
  int alloc(size_t count) {
    release();
    n = cap = count;
    if (count == 0) return XPBD_OK;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    return XPBD_OK;
  }
  int upload(const T* src, size_t count) {
    TRY(alloc(count));
    if (count) HIP_TRY(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
    return XPBD_OK;
  }
  // upload into the existing allocation when it is large enough
  int upload_reuse(const T* src, size_t count) {
    if (count > cap || !p) {
      TRY(alloc(count));
    } else {
      n = count;
    }
    if (count) HIP_TRY(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
    return XPBD_OK;
  }
  int zero() {
    if (n) HIP_TRY(hipMemset(p, 0, n * sizeof(T)));
    return XPBD_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = cap = 0;
  }
  ~DevBuf() { release(); }
};

// 10-bit-per-axis Morton code of a point inside [lo, hi].
uint32_t morton3(const float* x, const float* lo, const float* ext) {
  uint32_t code = 0;
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    float u = ext[k] > 0.f ? (x[k] - lo[k]) / ext[k] : 0.f;
    u = std::min(std::max(u, 0.f), 1.f);
    q[k] = std::min<uint32_t>(1023u, uint32_t(u * 1023.f));
  }
  for (int b = 9; b >= 0; --b)
    for (int k = 0; k < 3; ++k) code = (code << 1) | ((q[k] >> b) & 1u);
  return code;
}

float tet_volume_f32(const float* p0, const float* p1, const float* p2, const float* p3) {
  // calc_tet_volume (xpbd_base.py:6-8): (p1-p0) x (p2-p0) . (p3-p0) / 6
  const float a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
  const float b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
  const float c[3] = {p3[0] - p0[0], p3[1] - p0[1], p3[2] - p0[2]};
  const float ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                       a[0] * b[1] - a[1] * b[0]};
  return (ab[0] * c[0] + ab[1] * c[1] + ab[2] * c[2]) / 6.f;
}

// Matrix.inverse for 3x3: adjugate times 1/det, det by first-column expansion.
void inv3_f32(const float m[3][3], float out[3][3]) {
  const float det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                    m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                    m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
  const float inv_det = 1.f / det;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      out[j][i] = inv_det * (m[i1][j1] * m[i2][j2] - m[i2][j1] * m[i1][j2]);
    }
}

// Device form of a collider (the parameters as the reference's f32 fields
// hold them); a hang collider's sphere centres at `pts`.
SdfDev to_dev(const xpbd_sdf& s, const float4* pts) {
  SdfDev d;
  std::memset(&d, 0, sizeof(d));
  d.kind = s.kind;
  d.fixed = s.fixed;
  for (int k = 0; k < 3; ++k) d.p[k] = float(s.p[k]);
  for (int k = 0; k < 3; ++k) d.vel[k] = float(s.vel[k]);
  d.r = float(s.p[3]);
  if (s.kind == XPBD_SDF_PLANE) {
    // PlaneSdf normalizes dir at construction in Python (sdf.py:66).
    const double n = std::sqrt(s.dir[0] * s.dir[0] + s.dir[1] * s.dir[1] + s.dir[2] * s.dir[2]);
    for (int k = 0; k < 3; ++k) d.d[k] = n > 0 ? float(s.dir[k] / n) : 0.f;
  } else if (s.kind == XPBD_SDF_CUBE) {
    for (int k = 0; k < 3; ++k) d.d[k] = float(s.dir[k] * 0.5);  // half_size (sdf.py:104)
  } else if (s.kind == XPBD_SDF_HANG) {
    d.fixed = 1;  // HangSdfModel is always fixed (sdf.py:31)
    d.n_pts = s.n_points;
    d.pts = pts;
  }
  return d;
}

int validate_sdfs(const xpbd_sdf* sdfs, int n) {
  if (n < 1) return set_error(XPBD_EINVAL, "At least one SDF object must be provided.");
  if (n > XPBD_MAX_SDF) return set_error(XPBD_EINVAL, "at most %d SDFs are supported", XPBD_MAX_SDF);
  if (!sdfs) return set_error(XPBD_EINVAL, "sdfs is NULL");
  for (int i = 0; i < n; ++i) {
    const int k = sdfs[i].kind;
    if (k != XPBD_SDF_PLANE && k != XPBD_SDF_SPHERE && k != XPBD_SDF_CUBE && k != XPBD_SDF_HANG)
      return set_error(XPBD_EINVAL, "sdf %d: unknown kind %d", i, k);
    if (k == XPBD_SDF_HANG &&
        (!sdfs[i].points || sdfs[i].n_points < 1 || sdfs[i].n_points > XPBD_MAX_HANG_POINTS))
      return set_error(XPBD_EINVAL, "sdf %d: a hang collider needs 1..%d points", i,
                       XPBD_MAX_HANG_POINTS);
  }
  return XPBD_OK;
}

// Patches: consecutive tets grouped into workgroup-sized pieces of at most
// kPatchTets tets touching at most kPatchVerts vertices (local ids and corner
// ranks fit in a byte).  Each (patch, vertex) pair is a slot;
// a vertex's slots are listed in ascending patch order.

// LDS slot order of a patch's vertices.  The patch kernels gather staged
// vertices (16 B each) by local id, corner k of every lane's tet at once;
// lanes that read different slots in the same group of four banks conflict.
// The slots are coloured by slot % kLdsPassLanes, modelling groups of
// kLdsPassLanes consecutive lanes as served together (8 measured fewer
// conflicts than 16 on gfx950).  Greedy colouring, most constrained vertex
// first: each vertex takes the colour, among those with free slots, shared by
// the fewest vertices it is gathered with in the same lane group; ties go to
// the lower colour.  Only the slot of each vertex changes, not the order of
// any sum, so results are bit-identical.  Returns the vertices in slot order.
constexpr int kLdsPassLanes = 8;

// The default wait of one scene barrier of the persistent substep before it
// gives up: 2 s of the 100 MHz wall clock.
constexpr unsigned long long kPersistTimeoutTicks = 200000000ull;
constexpr unsigned long long kWallClockTicksPerUs = 100ull;

std::vector<int> lds_slot_order(const std::vector<int>& verts, const std::vector<int>& tets,
                                int t0, int t1, std::vector<int>& local) {
  const int n = int(verts.size());
  for (int i = 0; i < n; ++i) local[verts[i]] = i;  // dense index while colouring
  const int n_groups = (t1 - t0 + kLdsPassLanes - 1) / kLdsPassLanes;
  // passes[k * n_groups + g]: the vertices corner k of pass g gathers
  std::vector<std::vector<int>> passes(size_t(4) * n_groups);
  std::vector<std::vector<int>> of_vert(n);
  for (int t = t0; t < t1; ++t) {
    const int g = (t - t0) / kLdsPassLanes;
    for (int k = 0; k < 4; ++k) {
      const int v = local[tets[4 * size_t(t) + k]];
      const int c = k * n_groups + g;
      passes[c].push_back(v);
      if (of_vert[v].empty() || of_vert[v].back() != c) of_vert[v].push_back(c);
    }
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return of_vert[a].size() > of_vert[b].size();
  });
  const int B = kLdsPassLanes;
  std::vector<int> cap(B, 0), color(n, -1);
  for (int s = 0; s < n; ++s) cap[s % B]++;
  std::vector<int> cost(B);
  for (int v : order) {
    std::fill(cost.begin(), cost.end(), 0);
    for (int c : of_vert[v])
      for (int u : passes[c])
        if (u != v && color[u] >= 0) cost[color[u]]++;
    int best = -1;
    for (int b = 0; b < B; ++b)
      if (cap[b] > 0 && (best < 0 || cost[b] < cost[best])) best = b;
    color[v] = best;
    cap[best]--;
  }
  // slot s holds a vertex of colour s % B; within a colour, ascending index
  std::vector<int> next(B);
  for (int b = 0; b < B; ++b) next[b] = b;
  std::vector<int> out(n);
  for (int i = 0; i < n; ++i) {
    const int s = next[color[i]];
    next[color[i]] += B;
    out[s] = verts[i];
  }
  return out;
}

struct Patches {
  std::vector<int> tet_begin, vert_begin, vert_ids, vslot_off;
  std::vector<int2> vert_sum;     // (corner base | corner count << 16, slot)
  std::vector<uint2> tet_local;   // (4 x u8 local ids, 4 x u8 corner ranks)
};

void build_patches(const std::vector<int>& tets, int nv, int nt, Patches& P,
                   const std::vector<int>* cuts = nullptr) {
  P.tet_begin.assign(1, 0);
  P.vert_begin.assign(1, 0);
  P.vert_ids.clear();
  P.vert_sum.clear();
  P.tet_local.assign(nt, make_uint2(0u, 0u));
  std::vector<int> stamp(nv, -1), local(nv, 0), lcount(nv, 0);
  std::vector<int> cur;
  std::vector<int> patch_of_slot;
  int start = 0;
  auto close = [&](int end) {
    std::sort(cur.begin(), cur.end());
    cur = lds_slot_order(cur, tets, start, end, local);
    for (size_t i = 0; i < cur.size(); ++i) local[cur[i]] = int(i);
    std::vector<int> cnt(cur.size(), 0);
    for (int t = start; t < end; ++t) {
      unsigned ids = 0, ranks = 0;
      for (int k = 0; k < 4; ++k) {
        const int l = local[tets[4 * size_t(t) + k]];
        ids |= unsigned(l) << (8 * k);
        ranks |= unsigned(cnt[l]++) << (8 * k);  // corners of a vertex in tet order
      }
      P.tet_local[t] = make_uint2(ids, ranks);
    }
    int base = 0;
    for (size_t i = 0; i < cur.size(); ++i) {
      P.vert_ids.push_back(cur[i]);
      P.vert_sum.push_back(make_int2(base | (cnt[i] << 16), 0));
      patch_of_slot.push_back(int(P.tet_begin.size()) - 1);
      base += cnt[i];
    }
    P.tet_begin.push_back(end);
    P.vert_begin.push_back(int(P.vert_ids.size()));
    cur.clear();
    start = end;
  };
  size_t next_cut = 0;
  for (int t = 0; t < nt; ++t) {
    const int p = int(P.tet_begin.size()) - 1;
    if (cuts && next_cut < cuts->size() && t == (*cuts)[next_cut] && t > start) {
      close(t);
      ++next_cut;
    }
    int fresh = 0, crowded = 0;
    for (int k = 0; k < 4; ++k) {
      const int g = tets[4 * size_t(t) + k];
      fresh += stamp[g] != p;
      crowded |= stamp[g] == p && lcount[g] >= 255;  // a rank must fit in a byte
    }
    if (t - start == kPatchTets || int(cur.size()) + fresh > kPatchVerts || crowded) close(t);
    const int q = int(P.tet_begin.size()) - 1;
    for (int k = 0; k < 4; ++k) {
      const int g = tets[4 * size_t(t) + k];
      if (stamp[g] != q) {
        stamp[g] = q;
        lcount[g] = 0;
        cur.push_back(g);
      }
      lcount[g]++;
    }
  }
  if (nt > start) close(nt);
  // slots of a vertex are contiguous and in ascending patch order
  P.vslot_off.assign(nv + 1, 0);
  for (int g : P.vert_ids) P.vslot_off[g + 1]++;
  for (int v = 0; v < nv; ++v) P.vslot_off[v + 1] += P.vslot_off[v];
  std::vector<int> fill(P.vslot_off.begin(), P.vslot_off.end() - 1);
  for (size_t i = 0; i < P.vert_ids.size(); ++i) {
    const int g = P.vert_ids[i];
    P.vert_sum[i].y = fill[g]++;
    // the patch holding a vertex's first slot owns the vertex's writes
    if (P.vert_sum[i].y == P.vslot_off[g]) P.vert_sum[i].x |= int(0x80000000u);
  }
}

// Recursive coordinate bisection of the tets by centroid: the tet set is cut
// at a multiple of kPatchTets along the longest axis of its centroid box
// until every piece has at most kPatchTets tets and kPatchVerts vertices (and
// no vertex in more than 255 of its tets).  The pieces are compact blobs, so
// a patch stages few vertices per tet; `order` lists the tets piece by piece
// and `cuts` the end of each piece.
void rcb_order(const std::vector<float>& cent, const int* tets, int nt, int nv,
               std::vector<int>& order, std::vector<int>& cuts) {
  struct Rec {
    float c[3];
    int id;
  };
  // records are partitioned in place, so every pass streams through memory
  std::vector<Rec> rec(nt);
  for (int t = 0; t < nt; ++t) {
    rec[t] = {{cent[3 * size_t(t)], cent[3 * size_t(t) + 1], cent[3 * size_t(t) + 2]}, t};
  }
  cuts.clear();
  std::vector<int> stamp(nv, -1), lcount(nv, 0);
  int stamp_id = 0;
  auto leaf_ok = [&](int lo, int hi) {
    ++stamp_id;
    int nverts = 0;
    for (int i = lo; i < hi; ++i)
      for (int k = 0; k < 4; ++k) {
        const int g = tets[4 * size_t(rec[i].id) + k];
        if (stamp[g] != stamp_id) {
          stamp[g] = stamp_id;
          lcount[g] = 0;
          ++nverts;
        }
        if (++lcount[g] > 255) return false;
      }
    return nverts <= kPatchVerts;
  };
  std::vector<int2> stack;
  if (nt > 0) stack.push_back(make_int2(0, nt));
  while (!stack.empty()) {
    const int2 r = stack.back();
    stack.pop_back();
    const int lo = r.x, hi = r.y, n = hi - lo;
    if (n <= kPatchTets && leaf_ok(lo, hi)) {
      cuts.push_back(hi);
      continue;
    }
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = lo; i < hi; ++i)
      for (int k = 0; k < 3; ++k) {
        mn[k] = std::min(mn[k], rec[i].c[k]);
        mx[k] = std::max(mx[k], rec[i].c[k]);
      }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
      if (mx[k] - mn[k] > mx[axis] - mn[axis]) axis = k;
    int m;
    if (n > kPatchTets) {
      const int pieces = (n + kPatchTets - 1) / kPatchTets;
      m = lo + (pieces / 2) * kPatchTets;
    } else {
      m = lo + n / 2;
    }
    std::nth_element(rec.begin() + lo, rec.begin() + m, rec.begin() + hi,
                     [axis](const Rec& a, const Rec& b) { return a.c[axis] < b.c[axis]; });
    // depth first, low half first: push high then low
    stack.push_back(make_int2(m, hi));
    stack.push_back(make_int2(lo, m));
  }
  order.resize(nt);
  for (int t = 0; t < nt; ++t) order[t] = rec[t].id;
}

}  // namespace

struct TimedSpan {
  hipEvent_t begin = nullptr, end = nullptr;
  int launches = 0;
};

struct xpbd_solver {
  int device = 0;
  hipStream_t stream = nullptr;
  xpbd_desc prm{};                 // scalar parameters (mesh pointers cleared)
  std::vector<xpbd_sdf> sdfs;                    // points copied into sdf_pts_*
  std::vector<std::vector<float4>> sdf_pts_host;  // hang colliders' sphere centres
  DevBuf<float4> sdf_pts_dev[XPBD_MAX_SDF];
  int nv = 0, nt = 0;
  int bs_t = 256, bs_v = 256;
  std::vector<int> v_new2old, v_old2new, t_new2old;
  std::vector<int> tets_orig;      // caller order, for the visual indices
  std::vector<int> tets_new;       // internal order, internal vertex ids
  std::vector<float> m_host, inv_m_host;  // internal order
  int n_degenerate = 0;     // tets of volume <= EPSILON at the last mass computation
  int first_degenerate = -1;  // the first of them in the caller's tet order

  DevBuf<float4> pos, xs, vel, prev, xj, dacc, corner, rest, slot_acc, ybuf1, ybuf2;
  bool has_isolated = false;
  DevBuf<int> p_tet_begin, p_vert_begin, p_vert_ids, vslot_off;
  DevBuf<int2> p_vert_sum;
  DevBuf<uint2> tet_local;
  DevBuf<xpbd::PatchItem> pitems;
  std::vector<PatchItem> pitems_host;
  DevBuf<int2> prange_proj, prange_res;
  int grid_proj = 0, grid_res = 0;
  // persistent substep: barrier words (arrivals, workgroups done, error), the
  // most workgroups of its kernel that can be resident at once, and whether it
  // has run since the error word was last read
  DevBuf<unsigned> sync;
  // persistent substep: each residual group's partials of the last two
  // iterations, posted by value (empty words hold kMailboxEmpty)
  DevBuf<double> mailbox;
  int persist_capacity = 0;
  // workgroups of that kernel resident at once on one XCD, and the XCD this
  // solver's persistent substep runs on when the scene fits there
  int persist_xcd_capacity = 0;
  int home_xcd = 0;
  bool persist_pending = false;
  unsigned long long persist_timeout = kPersistTimeoutTicks;
  int n_patches = 0;
  int layout = XPBD_LAYOUT_PATCH;
  DevBuf<float> valence;
  DevBuf<int> adj_off, adj, d_old2new;
  DevBuf<int4> tet;
  DevBuf<float4> dm0, dm1;
  DevBuf<float2> dm2, lam;
  DevBuf<double> partial, logbuf;
  DevBuf<SceneCtrl> ctrl;
  DevBuf<SceneDev> table;
  DevBuf<WorkItem> vmap, tmap;
  DevBuf<float> scratch_xyz;
  DevBuf<int> flag;
  int nvb = 0, ntb = 0;

  SceneDev host_scene{};
  bool table_valid = false;

  int frame = 0, substep = 0;
  int sub_pos = 0;  // substeps of the current frame already run by a batch
  double time = 0.0;
  bool logging = false;
  std::vector<double> log_rows;
  double* log_host = nullptr;     // pinned staging for one substep's rows

  int path = XPBD_PATH_GRAPH;
  hipGraphExec_t gexec = nullptr;
  int graph_key = -1;

  // kernel timing: 0 off; 1 times each projection launch with an event pair
  // (recorded by the dispatch itself in the patch layout); r > 1 times r
  // back-to-back replays of each projection (lambda saved before and restored
  // after, so the solve itself is unchanged)
  int timing = 0;
  DevBuf<float2> lam_save;
  std::vector<TimedSpan> ev_pending, ev_free;
  double timed_ms = 0.0;
  int64_t timed_launches = 0;

  int64_t device_bytes() const {
    int64_t b = 0;
    auto add = [&](const auto& buf) { b += int64_t(buf.n) * int64_t(sizeof(*buf.p)); };
    add(pos); add(xs); add(vel); add(prev); add(xj); add(dacc); add(corner); add(rest);
    add(slot_acc); add(ybuf1); add(ybuf2); add(p_tet_begin); add(p_vert_begin); add(p_vert_ids);
    add(vslot_off); add(p_vert_sum); add(tet_local); add(pitems); add(prange_proj);
    add(prange_res); add(valence); add(adj_off); add(adj); add(d_old2new); add(tet); add(dm0);
    add(dm1); add(dm2); add(lam); add(partial); add(logbuf); add(ctrl); add(table); add(vmap);
    add(tmap); add(scratch_xyz); add(flag); add(lam_save); add(sync); add(mailbox);
    return b;
  }

  ~xpbd_solver() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (auto& e : ev_pending) {
      (void)hipEventDestroy(e.begin);
      (void)hipEventDestroy(e.end);
    }
    for (auto& e : ev_free) {
      (void)hipEventDestroy(e.begin);
      (void)hipEventDestroy(e.end);
    }
    if (log_host) (void)hipHostFree(log_host);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Waits for the handle's stream.  When the persistent substep has run since
// the last wait, its error word is read: a set word (its workgroups were not
// all resident, so a scene barrier gave up) is cleared and reported.
int stream_sync(xpbd_solver* s) {
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (!s->persist_pending) return XPBD_OK;
  s->persist_pending = false;
  unsigned err = 0;
  HIP_TRY(hipMemcpy(&err, s->sync.p + 2, sizeof(err), hipMemcpyDeviceToHost));
  if (err) {
    const unsigned zero = 0;
    HIP_TRY(hipMemcpy(s->sync.p + 2, &zero, sizeof(zero), hipMemcpyHostToDevice));
    return set_error(XPBD_EHIP, "persistent substep: a scene barrier timed out (its workgroups "
                                "were not all resident); the substep's results are invalid");
  }
  return XPBD_OK;
}

int validate_params(const xpbd_desc* d) {
  if (d->substeps_num < 1) return set_error(XPBD_EINVAL, "substeps_num must be >= 1");
  if (d->constraints_iter < 0) return set_error(XPBD_EINVAL, "constraints_iter must be >= 0");
  if (d->block_size < 64 || d->block_size > 1024 || d->block_size % 64)
    return set_error(XPBD_EINVAL, "block_size must be a multiple of 64 in [64, 1024]");
  if (d->accumulate_mode != XPBD_ACCUM_ATOMIC && d->accumulate_mode != XPBD_ACCUM_CSR)
    return set_error(XPBD_EINVAL, "unknown accumulate_mode %d", d->accumulate_mode);
  if (!(d->frame_dt > 0) || !(d->dt > 0)) return set_error(XPBD_EINVAL, "frame_dt and dt must be > 0");
  if (d->layout != XPBD_LAYOUT_FLAT && d->layout != XPBD_LAYOUT_PATCH)
    return set_error(XPBD_EINVAL, "unknown layout %d", d->layout);
  if (d->reorder != XPBD_ORDER_NONE && d->reorder != XPBD_ORDER_BISECT &&
      d->reorder != XPBD_ORDER_MORTON)
    return set_error(XPBD_EINVAL, "unknown order %d", d->reorder);
  return XPBD_OK;
}

// Parameters as the reference kernels see them: Python doubles frozen into
// f32 constants (neohookean_solver.py:7-12, 43-44).
void fill_scene_params(xpbd_solver* s, SceneDev& S, double dt, double damping) {
  const xpbd_desc& d = s->prm;
  const double E = d.young_modulus, nu = d.poisson_ratio;
  const double mu = E / (2.0 * (1.0 + nu));
  const double la = 2.0 * nu / (1.0 - 2.0 * nu) * mu;
  S.lame1 = float(la);
  S.lame2 = float(mu);
  S.dt = float(dt);
  // alpha~ = 1 / (lame V0 dt^2) = alpha_k / det(Dm) with V0 = det(Dm) / 6
  S.alpha_kd = float(6.0 / (double(S.lame2) * double(S.dt) * double(S.dt)));
  S.alpha_kh = float(6.0 / (double(S.lame1) * double(S.dt) * double(S.dt)));
  S.gamma = float(d.cheb_gamma);
  S.hydro_relax = float(d.hydrostatic_relaxation);
  S.dev_relax = float(d.deviatoric_relaxation);
  S.decay = std::exp(-S.dt * float(damping));  // ti.exp(-dt * damping) in f32
  for (int k = 0; k < 3; ++k) S.g[k] = float(d.gravity[k]);
  S.block_mode = d.block_neohookean_enable ? 1 : 0;
  S.cheb = d.cheb_enable ? 1 : 0;
  S.monitor = d.monitor_convergence ? 1 : 0;
  S.csr = d.accumulate_mode == XPBD_ACCUM_CSR ? 1 : 0;
  S.fused = (s->layout == XPBD_LAYOUT_PATCH && S.cheb && !s->has_isolated) ? 1 : 0;
  S.warmup = d.cheb_warmup;
  S.iters = d.constraints_iter;
  S.n_sdf = int(s->sdfs.size());
  for (int i = 0; i < S.n_sdf; ++i) S.sdf[i] = to_dev(s->sdfs[i], s->sdf_pts_dev[i].p);
}

int build_scene(xpbd_solver* s) {
  SceneDev& S = s->host_scene;
  std::memset(&S, 0, sizeof(S));
  S.pos = s->pos.p;
  S.xs = s->xs.p;
  S.vel = s->vel.p;
  S.prev = s->prev.p;
  S.xj = s->xj.p;
  S.dacc = s->dacc.p;
  S.valence = s->valence.p;
  S.adj_off = s->adj_off.p;
  S.adj = s->adj.p;
  S.corner = s->corner.p;
  S.tet = s->tet.p;
  S.dm0 = s->dm0.p;
  S.dm1 = s->dm1.p;
  S.dm2 = s->dm2.p;
  S.rest = s->rest.p;
  S.p_tet_begin = s->p_tet_begin.p;
  S.p_vert_begin = s->p_vert_begin.p;
  S.p_vert_ids = s->p_vert_ids.p;
  S.p_vert_sum = s->p_vert_sum.p;
  S.ybuf[0] = s->pos.p;
  S.ybuf[1] = s->ybuf1.p;
  S.ybuf[2] = s->ybuf2.p;
  S.tet_local = s->tet_local.p;
  S.vslot_off = s->vslot_off.p;
  S.slot_acc = s->slot_acc.p;
  S.lam = s->lam.p;
  S.partial = s->partial.p;
  S.ctrl = s->ctrl.p;
  S.log = s->logbuf.p;
  S.sync = s->sync.p;
  S.mailbox = s->mailbox.p;
  S.n_verts = s->nv;
  S.n_tets = s->nt;
  S.layout = s->layout;
  S.n_patches = s->n_patches;
  S.n_partials = s->layout == XPBD_LAYOUT_PATCH
                     ? (s->n_patches + xpbd::kPartialGroup - 1) / xpbd::kPartialGroup
                     : s->ntb;
  S.tet_block = s->bs_t;
  fill_scene_params(s, S, s->prm.dt / s->prm.substeps_num, s->prm.damping);
  s->table_valid = false;
  return XPBD_OK;
}

// Uploads the scene table if (dt, damping) or any parameter changed.
// Brings the device scene table up to date with the parameters of the next
// substep.  The store is a launch on the solver's stream, so substeps already
// queued keep the parameters they were queued with and the host does not wait:
// a frame whose last time chunk is shorter (frame_dt not a multiple of dt)
// changes sub_dt twice per frame without a round trip.
int sync_table(xpbd_solver* s, double dt, double damping) {
  SceneDev S = s->host_scene;
  fill_scene_params(s, S, dt, damping);
  if (s->table_valid && std::memcmp(&S, &s->host_scene, sizeof(S)) == 0) return XPBD_OK;
  s->host_scene = S;
  HIP_TRY(xpbd::launch_store_scene(s->table.p, s->host_scene, s->stream));
  s->table_valid = true;
  return XPBD_OK;
}

// Workgroups are dispatched round-robin over the accelerator dies (XCDs), each
// with its own L2: workgroup b runs on XCD b % n_xcd.  This returns, for each
// workgroup, the work item it should take so that every XCD walks one
// contiguous range of the (Morton-ordered) items, and neighbouring patches and
// their vertices meet in the same L2.
std::vector<int> xcd_order(int n, int n_xcd) {
  std::vector<int> order(n);
  if (n_xcd <= 1) {
    std::iota(order.begin(), order.end(), 0);
    return order;
  }
  std::vector<int> start(n_xcd + 1, 0);
  for (int x = 0; x < n_xcd; ++x) start[x + 1] = start[x] + (n > x ? (n - x - 1) / n_xcd + 1 : 0);
  for (int b = 0; b < n; ++b) order[b] = start[b % n_xcd] + b / n_xcd;
  return order;
}

int xcd_count() {
  const char* e = std::getenv("XPBD_XCDS");
  const int v = e ? std::atoi(e) : 8;  // MI355X: 8 XCDs of 32 CUs
  return v >= 1 ? v : 1;
}

// Splits n_items work items over n_blocks persistent workgroups: each XCD gets
// a contiguous share proportional to its number of workgroups, and each
// workgroup a contiguous run within its XCD's share.  Every item is covered
// exactly once (checked).
int xcd_ranges(int n_items, int n_blocks, int n_xcd, std::vector<int2>& ranges) {
  ranges.assign(n_blocks, make_int2(0, 0));
  if (n_blocks <= 0) return XPBD_OK;
  n_xcd = std::max(1, std::min(n_xcd, n_blocks));
  std::vector<int> nb(n_xcd, 0);
  for (int b = 0; b < n_blocks; ++b) nb[b % n_xcd]++;
  long long before = 0;
  for (int x = 0; x < n_xcd; ++x) {
    const int is = int(before * n_items / n_blocks);
    before += nb[x];
    const int ie = int(before * n_items / n_blocks);
    for (int j = 0; j < nb[x]; ++j) {
      const int b = x + j * n_xcd;
      ranges[b] = make_int2(is + int((long long)j * (ie - is) / nb[x]),
                            is + int((long long)(j + 1) * (ie - is) / nb[x]));
    }
  }
  std::vector<char> seen(n_items, 0);
  for (const int2& r : ranges)
    for (int i = r.x; i < r.y; ++i) {
      if (i < 0 || i >= n_items || seen[i]) return set_error(XPBD_ESTATE, "work split overlaps");
      seen[i] = 1;
    }
  for (char c : seen)
    if (!c) return set_error(XPBD_ESTATE, "work split misses an item");
  return XPBD_OK;
}

// Residual work ranges: as xcd_ranges, over units that each hold one group of
// kPartialGroup consecutive patches of one scene, so every range starts at a
// group boundary and each group's partial is summed by one workgroup in order.
int grouped_ranges(const std::vector<PatchItem>& items, int n_blocks, int n_xcd,
                   std::vector<int2>& ranges, int* n_used) {
  std::vector<int> starts;
  for (size_t i = 0; i < items.size(); ++i)
    if (i == 0 || items[i].scene != items[i - 1].scene || items[i].patch % xpbd::kPartialGroup == 0)
      starts.push_back(int(i));
  const int n_units = int(starts.size());
  const int nb = std::max(1, std::min(n_blocks, n_units));
  std::vector<int2> ur;
  TRY(xcd_ranges(n_units, nb, n_xcd, ur));
  ranges.resize(ur.size());
  const int n_items = int(items.size());
  for (size_t b = 0; b < ur.size(); ++b) {
    const int lo = ur[b].x < n_units ? starts[ur[b].x] : n_items;
    const int hi = ur[b].y < n_units ? starts[ur[b].y] : n_items;
    ranges[b] = make_int2(lo, hi);
  }
  *n_used = nb;
  return XPBD_OK;
}

int build_maps(xpbd_solver* s) {
  std::vector<WorkItem> vm, tm;
  const int nx = xcd_count();
  const int nvb = (s->nv + s->bs_v - 1) / s->bs_v, ntb = (s->nt + s->bs_t - 1) / s->bs_t;
  for (int b : xcd_order(nvb, nx)) vm.push_back({0, b * s->bs_v});
  for (int b : xcd_order(ntb, nx)) tm.push_back({0, b * s->bs_t});
  s->nvb = int(vm.size());
  s->ntb = int(tm.size());
  TRY(s->vmap.upload(vm.data(), vm.size()));
  TRY(s->tmap.upload(tm.data(), tm.size()));
  return XPBD_OK;
}

// _compute_lumped_mass_kernel + _apply_fixed_vertices_for_sdf
// (xpbd_base.py:89-137) on host copies of x and v (internal order); the
// atomic accumulation order of the reference is unspecified, this sums in
// ascending internal tet order.
void lumped_mass_host(xpbd_solver* s, const std::vector<float>& x, const std::vector<float>& v,
                      double total_mass) {
  s->m_host.assign(s->nv, 0.f);
  s->inv_m_host.assign(s->nv, 0.f);
  float total_volume = 0.f;
  s->n_degenerate = 0;
  s->first_degenerate = -1;
  for (int t = 0; t < s->nt; ++t) {
    const int* tv = &s->tets_new[4 * size_t(t)];
    const float vol = tet_volume_f32(&x[3 * size_t(tv[0])], &x[3 * size_t(tv[1])],
                                     &x[3 * size_t(tv[2])], &x[3 * size_t(tv[3])]);
    if (vol <= xpbd::kEps) {  // the reference's warning condition (xpbd_base.py:104-106)
      const int caller = s->t_new2old[t];
      s->n_degenerate++;
      if (s->first_degenerate < 0 || caller < s->first_degenerate) s->first_degenerate = caller;
    }
    const float quarter = vol * 0.25f;
    for (int k = 0; k < 4; ++k) s->m_host[tv[k]] += quarter;
    total_volume += vol;
  }
  const float mass_scale = float(total_mass) / total_volume;
  for (int i = 0; i < s->nv; ++i) {
    s->m_host[i] *= mass_scale;
    s->inv_m_host[i] = s->m_host[i] > 0.f ? 1.f / s->m_host[i] : 0.f;
  }
  for (size_t j = 0; j < s->sdfs.size(); ++j) {
    const SdfDev d = to_dev(s->sdfs[j], s->sdf_pts_host[j].data());
    float n[3];
    for (int i = 0; i < s->nv; ++i)
      if (s->inv_m_host[i] > 0.f && d.fixed &&
          xpbd::sdf_inside(d, &x[3 * size_t(i)], &v[3 * size_t(i)], n))
        s->inv_m_host[i] = 0.f;
  }
}

// Keeps copies of the colliders; a hang collider's points are copied to the
// host and to a device buffer of this solver.
int store_sdfs(xpbd_solver* s, const xpbd_sdf* sdfs, int n) {
  TRY(validate_sdfs(sdfs, n));
  s->sdfs.assign(sdfs, sdfs + n);
  s->sdf_pts_host.assign(n, {});
  for (int i = 0; i < n; ++i) {
    xpbd_sdf& c = s->sdfs[i];
    if (c.kind == XPBD_SDF_HANG) {
      std::vector<float4>& h = s->sdf_pts_host[i];
      h.resize(c.n_points);
      for (int k = 0; k < c.n_points; ++k)
        h[k] = make_float4(c.points[3 * k], c.points[3 * k + 1], c.points[3 * k + 2], 0.f);
      TRY(s->sdf_pts_dev[i].upload(h.data(), h.size()));
    } else {
      s->sdf_pts_dev[i].release();
      c.n_points = 0;
    }
    c.points = nullptr;  // the caller's buffer is not kept
  }
  s->prm.n_sdf = n;
  return XPBD_OK;
}

int download_vec4_xyz(xpbd_solver* s, const DevBuf<float4>& b, std::vector<float>& out,
                      std::vector<float>* w = nullptr) {
  std::vector<float4> tmp(s->nv);
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (s->nv) HIP_TRY(hipMemcpy(tmp.data(), b.p, sizeof(float4) * s->nv, hipMemcpyDeviceToHost));
  out.resize(3 * size_t(s->nv));
  if (w) w->resize(s->nv);
  for (int i = 0; i < s->nv; ++i) {
    out[3 * i] = tmp[i].x;
    out[3 * i + 1] = tmp[i].y;
    out[3 * i + 2] = tmp[i].z;
    if (w) (*w)[i] = tmp[i].w;
  }
  return XPBD_OK;
}

int upload_inv_m(xpbd_solver* s) {
  // inv_m lives in the w lane of pos and xs
  std::vector<float4> tmp(s->nv);
  for (auto* b : {&s->pos, &s->xs}) {
    HIP_TRY(hipMemcpy(tmp.data(), b->p, sizeof(float4) * s->nv, hipMemcpyDeviceToHost));
    for (int i = 0; i < s->nv; ++i) tmp[i].w = s->inv_m_host[i];
    HIP_TRY(hipMemcpy(b->p, tmp.data(), sizeof(float4) * s->nv, hipMemcpyHostToDevice));
  }
  return XPBD_OK;
}

int events_acquire(xpbd_solver* s, TimedSpan& ev) {
  if (!s->ev_free.empty()) {
    ev = s->ev_free.back();
    s->ev_free.pop_back();
  } else {
    HIP_TRY(hipEventCreate(&ev.begin));
    HIP_TRY(hipEventCreate(&ev.end));
  }
  return XPBD_OK;
}

int events_begin(xpbd_solver* s, TimedSpan& ev) {
  TRY(events_acquire(s, ev));
  HIP_TRY(hipEventRecord(ev.begin, s->stream));
  return XPBD_OK;
}

int events_end(xpbd_solver* s, TimedSpan& ev, int launches) {
  HIP_TRY(hipEventRecord(ev.end, s->stream));
  ev.launches = launches;
  s->ev_pending.push_back(ev);
  return XPBD_OK;
}

int events_collect(xpbd_solver* s) {
  for (auto& ev : s->ev_pending) {
    HIP_TRY(hipEventSynchronize(ev.end));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ev.begin, ev.end));
    s->timed_ms += ms;
    s->timed_launches += ev.launches;
    s->ev_free.push_back(ev);
  }
  s->ev_pending.clear();
  return XPBD_OK;
}

// The Jacobi iterations of solve_step (neohookean_solver.py:474-549).
//  fused (patch layout, Chebyshev on): per iteration a projection (which
//    first evaluates the previous iteration's rho/omega update and forms its
//    Chebyshev blend), the Jacobi average and a residual; after the last
//    iteration the standalone rho/omega update and a vertex pass forming the
//    final blend.
//  otherwise: projection, Jacobi average, [residual, rho/omega], [blend].
int enqueue_iterations(xpbd_solver* s, bool capturing) {
  const SceneDev& S = s->host_scene;
  const SceneDev* T = s->table.p;
  const bool patch = s->layout == XPBD_LAYOUT_PATCH;
  auto project = [&](int it) -> int {
    if (patch)
      HIP_TRY(xpbd::launch_patch_project(T, s->pitems.p, s->prange_proj.p, s->grid_proj, it,
                                         false, s->stream));
    else
      HIP_TRY(xpbd::launch_project(T, s->tmap.p, s->ntb, s->bs_t, it == 0, s->stream));
    return XPBD_OK;
  };
  // one patch projection timed by events that its own dispatch records (the
  // kernel's start and end, without the gaps between launches)
  auto timed_project = [&](int it) -> int {
    TimedSpan ev;
    TRY(events_acquire(s, ev));
    HIP_TRY(xpbd::launch_patch_project(T, s->pitems.p, s->prange_proj.p, s->grid_proj, it,
                                       false, s->stream, ev.begin, ev.end));
    ev.launches = 1;
    s->ev_pending.push_back(ev);
    return XPBD_OK;
  };
  const size_t lam_bytes = sizeof(float2) * s->lam.n;
  for (int it = 0; it < S.iters; ++it) {
    TimedSpan ev;
    const int reps = capturing ? 0 : s->timing;
    if (reps > 1) {
      // The projection only writes lambda, the slot partials and the
      // iterate buffer of this iteration, from inputs it does not write, so
      // replays differ from the real launch only in the lambda they leave.
      HIP_TRY(hipMemcpyAsync(s->lam_save.p, S.lam, lam_bytes, hipMemcpyDeviceToDevice,
                             s->stream));
      if (patch) {
        for (int r = 0; r < reps; ++r) TRY(timed_project(it));
      } else {
        TRY(events_begin(s, ev));
        for (int r = 0; r < reps; ++r) TRY(project(it));
        TRY(events_end(s, ev, reps));
      }
      HIP_TRY(hipMemcpyAsync(S.lam, s->lam_save.p, lam_bytes, hipMemcpyDeviceToDevice,
                             s->stream));
    }
    if (reps == 1 && patch) {
      TRY(timed_project(it));
    } else {
      if (reps == 1) TRY(events_begin(s, ev));
      TRY(project(it));
      if (reps == 1) TRY(events_end(s, ev, 1));
    }
    if (patch)
      HIP_TRY(xpbd::launch_patch_average(T, s->vmap.p, s->nvb, s->bs_v, it, s->stream));
    else
      HIP_TRY(xpbd::launch_average(T, s->vmap.p, s->nvb, s->bs_v, s->stream));
    if (S.cheb || S.monitor) {
      if (patch)
        HIP_TRY(xpbd::launch_patch_residual(T, s->pitems.p, s->prange_res.p, s->grid_res,
                                            s->stream));
      else
        HIP_TRY(xpbd::launch_residual(T, s->tmap.p, s->ntb, s->bs_t, s->stream));
      // fused: the next projection evaluates this iteration's controller update
      if (!S.fused || it == S.iters - 1) HIP_TRY(xpbd::launch_reduce(T, 1, it, s->stream));
    }
    if (S.cheb && !S.fused) HIP_TRY(xpbd::launch_blend(T, s->vmap.p, s->nvb, s->bs_v, s->stream));
  }
  if (S.fused && S.iters > 0)
    HIP_TRY(xpbd::launch_final_blend(T, s->vmap.p, s->nvb, s->bs_v, S.iters, s->stream));
  return XPBD_OK;
}

int enqueue_solve_step(xpbd_solver* s, bool capturing) {
  // lambda reset happens in the first projection; collision then save prev
  HIP_TRY(xpbd::launch_predict(s->table.p, s->vmap.p, s->nvb, s->bs_v, 0, 1,
                               s->host_scene.cheb, s->stream));
  return enqueue_iterations(s, capturing);
}

int enqueue_substep(xpbd_solver* s, bool capturing) {
  HIP_TRY(xpbd::launch_predict(s->table.p, s->vmap.p, s->nvb, s->bs_v, 1, 1,
                               s->host_scene.cheb, s->stream));
  TRY(enqueue_iterations(s, capturing));
  HIP_TRY(xpbd::launch_advance(s->table.p, s->vmap.p, s->nvb, s->bs_v, s->stream));
  return XPBD_OK;
}

int graph_key_of(const SceneDev& S) {
  return (S.iters << 6) | (S.fused << 5) | (S.cheb << 0) | (S.monitor << 1) | (S.block_mode << 2) | (S.csr << 3) |
         (S.layout << 4);
}

int fetch_log(xpbd_solver* s) {
  if (!s->logging) return XPBD_OK;
  const int iters = s->host_scene.iters;
  if (iters == 0) return XPBD_OK;
  HIP_TRY(hipMemcpyAsync(s->log_host, s->logbuf.p, sizeof(double) * 6 * iters,
                         hipMemcpyDeviceToHost, s->stream));
  TRY(stream_sync(s));
  for (int it = 0; it < iters; ++it) {
    const double* r = s->log_host + 6 * it;
    const double row[8] = {double(s->frame), double(s->substep), r[0], r[1], r[2], r[3], r[4], r[5]};
    s->log_rows.insert(s->log_rows.end(), row, row + 8);
  }
  return XPBD_OK;
}

// The largest scene the persistent substep is used for.
constexpr int kPersistMaxGroups = 256;  // 512 patches, one workgroup per CU
// (its residual mailbox is read by one thread per group of a 256-thread half)
static_assert(kPersistMaxGroups <= xpbd::kPatchThreads, "mailbox readers");
// XPBD_PATH_AUTO takes the persistent substep up to this many groups.  Per
// iteration on resting boxes (persistent against the graph of phased
// kernels): 24 groups 10.9-11.2 against 16.0 us, 47 groups 13.3 against 16.1,
// 81 groups 15.0 against 16.6, 116 groups 16.2 against 16.6, 129 groups 16.2
// against 17.1, 159 groups 17.6 against 17.6; the 32^3 twisted grid (192
// groups) 17.1 against 17.0.
constexpr int kAutoPersistGroups = 144;
// A batch takes the persistent substep automatically up to this many
// workgroups in all (every member within kAutoPersistGroups).  With the scene
// barriers free of L2 write-backs and invalidates, a batch of small members
// stays near the one-member cost as it grows: members of 17 groups run an
// iteration in 10.9 us at 34 groups in all, 11.2 us at 136 and 12.2 us at
// 238, against 14.8-17.3 us for the phased launches.
constexpr int kAutoBatchGroups = 256;

// Whether the persistent substep kernel can run this scene: the patch layout
// without isolated vertices (Chebyshev on or off), at least one iteration,
// and few enough groups to be resident together.
bool persistent_ok(const xpbd_solver* s) {
  const SceneDev& S = s->host_scene;
  const int groups = S.n_partials;
  return s->layout == XPBD_LAYOUT_PATCH && !s->has_isolated && S.iters >= 1 && groups >= 1 &&
         groups <= kPersistMaxGroups && groups <= s->persist_capacity;
}

int run_substep(xpbd_solver* s, double sub_dt) {
  TRY(sync_table(s, sub_dt, s->prm.damping));
  const bool persist = (s->path == XPBD_PATH_PERSISTENT ||
                        (s->path == XPBD_PATH_AUTO &&
                         s->host_scene.n_partials <= kAutoPersistGroups)) &&
                       !s->timing && persistent_ok(s);
  if (persist) {
    // all of the scene's workgroups on one XCD when they fit there: its
    // barriers then meet in one L2
    const int groups = s->host_scene.n_partials, nx = xcd_count();
    const bool one_xcd = nx > 1 && groups <= s->persist_xcd_capacity;
    HIP_TRY(xpbd::launch_substep_persistent(s->table.p, s->pitems.p, groups, s->persist_timeout,
                                            one_xcd ? nx : 1, one_xcd ? s->home_xcd % nx : 0,
                                            s->stream));
    s->persist_pending = true;
    return fetch_log(s);
  }
  if (s->path != XPBD_PATH_KERNELS && !s->timing) {
    const int key = graph_key_of(s->host_scene);
    if (!s->gexec || key != s->graph_key) {
      if (s->gexec) {
        HIP_TRY(hipGraphExecDestroy(s->gexec));
        s->gexec = nullptr;
      }
      hipGraph_t g;
      HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_substep(s, true);
      const hipError_t ec = hipStreamEndCapture(s->stream, &g);
      if (rc != XPBD_OK) return rc;
      HIP_TRY(ec);
      const hipError_t ei = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIP_TRY(ei);
      s->graph_key = key;
    }
    HIP_TRY(hipGraphLaunch(s->gexec, s->stream));
  } else {
    TRY(enqueue_substep(s, false));
  }
  return fetch_log(s);
}

int set_positions_common(xpbd_solver* s, const float* src_dev) {
  HIP_TRY(xpbd::launch_scatter_xyz(s->xs.p, src_dev, s->d_old2new.p, s->nv, 1, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return XPBD_OK;
}

}  // namespace

extern "C" {

int xpbd_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return XPBD_EINVAL;
  std::snprintf(buf, len, "%s", g_last_error.c_str());
  return XPBD_OK;
}

int xpbd_abi_info(int32_t* version, size_t* desc_size, size_t* sdf_size) {
  if (version) *version = XPBD_ABI_VERSION;
  if (desc_size) *desc_size = sizeof(xpbd_desc);
  if (sdf_size) *sdf_size = sizeof(xpbd_sdf);
  return XPBD_OK;
}

int xpbd_desc_init(xpbd_desc* d) {
  if (!d) return set_error(XPBD_EINVAL, "desc is NULL");
  std::memset(d, 0, sizeof(*d));
  d->young_modulus = 200000.0;
  d->poisson_ratio = 0.495;
  d->scale = 1.0;
  d->frame_dt = 1e-2;
  d->dt = 1e-2;
  d->substeps_num = 20;
  d->constraints_iter = 5;
  d->hydrostatic_relaxation = 1.2;
  d->deviatoric_relaxation = 1.2;
  d->cheb_enable = 1;
  d->cheb_warmup = 5;
  d->cheb_gamma = 0.6666;
  d->block_neohookean_enable = 0;
  d->monitor_convergence = 0;
  d->damping = 5.0;
  d->gravity[1] = -9.8;
  d->total_mass = 1.0;
  d->accumulate_mode = XPBD_ACCUM_CSR;
  d->reorder = XPBD_ORDER_BISECT;
  d->block_size = 256;
  d->device = 0;
  d->layout = XPBD_LAYOUT_PATCH;
  return XPBD_OK;
}

int xpbd_create(const xpbd_desc* d, xpbd_solver** out) {
  if (!d || !out) return set_error(XPBD_EINVAL, "NULL argument");
  *out = nullptr;
  if (d->n_verts <= 0 || d->n_tets <= 0 || !d->rest_positions || !d->tets)
    return set_error(XPBD_EINVAL, "mesh must have vertices and tets");
  TRY(validate_params(d));
  TRY(validate_sdfs(d->sdfs, d->n_sdf));
  const int nv = d->n_verts, nt = d->n_tets;
  for (size_t i = 0; i < 4 * size_t(nt); ++i)
    if (d->tets[i] < 0 || d->tets[i] >= nv)
      return set_error(XPBD_EINVAL, "tet %zu references vertex %d outside [0, %d)", i / 4,
                       d->tets[i], nv);
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (d->device < 0 || d->device >= ndev)
    return set_error(XPBD_EINVAL, "device %d not present (%d visible)", d->device, ndev);
  HIP_TRY(hipSetDevice(d->device));

  auto s = std::make_unique<xpbd_solver>();
  s->device = d->device;
  s->prm = *d;
  s->prm.rest_positions = nullptr;
  s->prm.tets = nullptr;
  s->prm.sdfs = nullptr;
  TRY(store_sdfs(s.get(), d->sdfs, d->n_sdf));
  s->nv = nv;
  s->nt = nt;
  s->bs_t = d->block_size;
  s->bs_v = 256;
  HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));

  // initialize (neohookean_solver.py:99-100): x = x * scale + offset, in f32
  const float scale = float(d->scale);
  const float off[3] = {float(d->offset[0]), float(d->offset[1]), float(d->offset[2])};
  std::vector<float> x0(3 * size_t(nv));
  for (int i = 0; i < nv; ++i)
    for (int k = 0; k < 3; ++k) x0[3 * size_t(i) + k] = d->rest_positions[3 * size_t(i) + k] * scale + off[k];

  // internal ordering
  s->v_new2old.resize(nv);
  s->t_new2old.resize(nt);
  std::iota(s->v_new2old.begin(), s->v_new2old.end(), 0);
  std::iota(s->t_new2old.begin(), s->t_new2old.end(), 0);
  if (d->reorder) {
    float lo[3] = {x0[0], x0[1], x0[2]}, hi[3] = {x0[0], x0[1], x0[2]}, ext[3];
    for (int i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], x0[3 * size_t(i) + k]);
        hi[k] = std::max(hi[k], x0[3 * size_t(i) + k]);
      }
    for (int k = 0; k < 3; ++k) ext[k] = hi[k] - lo[k];
    std::vector<uint32_t> vc(nv), tc(nt);
    for (int i = 0; i < nv; ++i) vc[i] = morton3(&x0[3 * size_t(i)], lo, ext);
    std::stable_sort(s->v_new2old.begin(), s->v_new2old.end(),
                     [&](int a, int b) { return vc[a] < vc[b]; });
    for (int t = 0; t < nt; ++t) {
      float c[3] = {0.f, 0.f, 0.f};
      for (int k = 0; k < 4; ++k)
        for (int j = 0; j < 3; ++j) c[j] += 0.25f * x0[3 * size_t(d->tets[4 * size_t(t) + k]) + j];
      tc[t] = morton3(c, lo, ext);
    }
    std::stable_sort(s->t_new2old.begin(), s->t_new2old.end(),
                     [&](int a, int b) { return tc[a] < tc[b]; });
  }
  std::vector<int> patch_cuts;
  if (d->reorder == XPBD_ORDER_BISECT && d->layout == XPBD_LAYOUT_PATCH) {
    // compact patches by coordinate bisection; vertices numbered by their first
    // appearance in patch order, so a patch's vertices are nearly contiguous
    std::vector<float> cent(3 * size_t(nt));
    for (int t = 0; t < nt; ++t)
      for (int j = 0; j < 3; ++j) {
        float c = 0.f;
        for (int k = 0; k < 4; ++k) c += 0.25f * x0[3 * size_t(d->tets[4 * size_t(t) + k]) + j];
        cent[3 * size_t(t) + j] = c;
      }
    rcb_order(cent, d->tets, nt, nv, s->t_new2old, patch_cuts);
    std::vector<int> seen(nv, 0);
    s->v_new2old.clear();
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < 4; ++k) {
        const int g = d->tets[4 * size_t(s->t_new2old[t]) + k];
        if (!seen[g]) {
          seen[g] = 1;
          s->v_new2old.push_back(g);
        }
      }
    for (int g = 0; g < nv; ++g)
      if (!seen[g]) s->v_new2old.push_back(g);
  }
  s->v_old2new.resize(nv);
  for (int i = 0; i < nv; ++i) s->v_old2new[s->v_new2old[i]] = i;
  s->tets_orig.assign(d->tets, d->tets + 4 * size_t(nt));
  s->tets_new.resize(4 * size_t(nt));
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < 4; ++k)
      s->tets_new[4 * size_t(t) + k] = s->v_old2new[d->tets[4 * size_t(s->t_new2old[t]) + k]];
  std::vector<float> x(3 * size_t(nv));
  for (int i = 0; i < nv; ++i)
    for (int k = 0; k < 3; ++k) x[3 * size_t(i) + k] = x0[3 * size_t(s->v_new2old[i]) + k];

  // initialize (neohookean_solver.py:102-107): V0 and inv_Dm per tet
  std::vector<float4> h_dm0(nt), h_dm1(nt);
  std::vector<float2> h_dm2(nt);
  std::vector<int4> h_tet(nt);
  for (int t = 0; t < nt; ++t) {
    const int* tv = &s->tets_new[4 * size_t(t)];
    const float* p[4];
    for (int k = 0; k < 4; ++k) p[k] = &x[3 * size_t(tv[k])];
    const float V0 = tet_volume_f32(p[0], p[1], p[2], p[3]);
    float dm[3][3], inv[3][3];
    for (int r = 0; r < 3; ++r) {
      dm[r][0] = p[1][r] - p[0][r];
      dm[r][1] = p[2][r] - p[0][r];
      dm[r][2] = p[3][r] - p[0][r];
    }
    inv3_f32(dm, inv);
    h_dm0[t] = make_float4(inv[0][0], inv[0][1], inv[0][2], inv[1][0]);
    h_dm1[t] = make_float4(inv[1][1], inv[1][2], inv[2][0], inv[2][1]);
    h_dm2[t] = make_float2(inv[2][2], V0);
    h_tet[t] = make_int4(tv[0], tv[1], tv[2], tv[3]);
  }

  // compute_lumped_mass (xpbd_base.py:133-137) with zero velocities
  std::vector<float> v0(3 * size_t(nv), 0.f);
  lumped_mass_host(s.get(), x, v0, d->total_mass);

  // valence (cons_num in block mode) and the vertex -> corner CSR
  std::vector<float> val(nv, 0.f);
  std::vector<int> off_v(nv + 1, 0);
  for (int i = 0; i < 4 * nt; ++i) off_v[s->tets_new[i] + 1]++;
  for (int i = 0; i < nv; ++i) {
    val[i] = float(off_v[i + 1]);
    off_v[i + 1] += off_v[i];
  }

  std::vector<float4> h_pos(nv), h_zero(nv, make_float4(0.f, 0.f, 0.f, 0.f));
  for (int i = 0; i < nv; ++i)
    h_pos[i] = make_float4(x[3 * size_t(i)], x[3 * size_t(i) + 1], x[3 * size_t(i) + 2], s->inv_m_host[i]);

  TRY(s->pos.upload(h_pos.data(), nv));
  TRY(s->xs.upload(h_pos.data(), nv));
  TRY(s->vel.upload(h_zero.data(), nv));
  TRY(s->prev.upload(h_zero.data(), nv));
  TRY(s->xj.upload(h_zero.data(), nv));
  TRY(s->dacc.upload(h_zero.data(), nv));
  TRY(s->valence.upload(val.data(), nv));
  {
    std::vector<float4> h_rest(nv);
    for (int i = 0; i < nv; ++i)
      h_rest[i] = make_float4(x[3 * size_t(i)], x[3 * size_t(i) + 1], x[3 * size_t(i) + 2], 0.f);
    TRY(s->rest.upload(h_rest.data(), nv));
  }
  s->layout = d->layout;
  if (d->layout == XPBD_LAYOUT_FLAT) {
    TRY(s->dm0.upload(h_dm0.data(), nt));
    TRY(s->dm1.upload(h_dm1.data(), nt));
    TRY(s->dm2.upload(h_dm2.data(), nt));
  }
  if (d->layout == XPBD_LAYOUT_FLAT) {
    TRY(s->tet.upload(h_tet.data(), nt));
  } else {
    Patches P;
    build_patches(s->tets_new, nv, nt, P, patch_cuts.empty() ? nullptr : &patch_cuts);
    s->n_patches = int(P.tet_begin.size()) - 1;
    TRY(s->p_tet_begin.upload(P.tet_begin.data(), P.tet_begin.size()));
    TRY(s->p_vert_begin.upload(P.vert_begin.data(), P.vert_begin.size()));
    TRY(s->p_vert_ids.upload(P.vert_ids.data(), P.vert_ids.size()));
    TRY(s->p_vert_sum.upload(P.vert_sum.data(), P.vert_sum.size()));
    TRY(s->tet_local.upload(P.tet_local.data(), P.tet_local.size()));
    TRY(s->vslot_off.upload(P.vslot_off.data(), P.vslot_off.size()));
    TRY(s->slot_acc.alloc(P.vert_ids.size()));
    TRY(s->slot_acc.zero());
    TRY(s->ybuf1.alloc(nv));
    TRY(s->ybuf2.alloc(nv));
    for (int i = 0; i < nv; ++i)
      if (val[i] == 0.f) s->has_isolated = true;
    std::vector<PatchItem> items(s->n_patches);
    for (int p = 0; p < s->n_patches; ++p) {
      PatchItem& it = items[p];
      std::memset(&it, 0, sizeof(it));
      it.scene = 0;
      it.patch = p;
      it.tet_begin = P.tet_begin[p];
      it.vert_begin = P.vert_begin[p];
      it.n_tets = P.tet_begin[p + 1] - P.tet_begin[p];
      it.n_verts = P.vert_begin[p + 1] - P.vert_begin[p];
    }
    TRY(s->pitems.upload(items.data(), items.size()));
    s->pitems_host = items;
    int occ_p = 0, occ_r = 0, n_cu = 0;
    HIP_TRY(xpbd::patch_kernel_occupancy(&occ_p, &occ_r));
    HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, d->device));
    s->grid_proj = std::max(1, std::min(s->n_patches, n_cu * std::max(1, occ_p)));
    s->grid_res = std::max(1, std::min(s->n_patches, n_cu * std::max(1, occ_r)));
    std::vector<int2> rp, rr;
    TRY(xcd_ranges(s->n_patches, s->grid_proj, xcd_count(), rp));
    TRY(grouped_ranges(items, s->grid_res, xcd_count(), rr, &s->grid_res));
    TRY(s->prange_proj.upload(rp.data(), rp.size()));
    TRY(s->prange_res.upload(rr.data(), rr.size()));
    int occ_s = 0;
    HIP_TRY(xpbd::persistent_occupancy(&occ_s));
    s->persist_capacity = n_cu * occ_s;
    s->persist_xcd_capacity = n_cu / xcd_count() * occ_s;
    static std::atomic<unsigned> next_xcd{0};
    s->home_xcd = int(next_xcd.fetch_add(1u) % unsigned(xcd_count()));
  }
  TRY(s->sync.alloc(4));
  TRY(s->sync.zero());
  if (s->layout == XPBD_LAYOUT_PATCH) {
    const size_t groups = size_t(s->n_patches + xpbd::kPartialGroup - 1) / xpbd::kPartialGroup;
    if (groups <= size_t(kPersistMaxGroups)) {  // scenes the persistent substep can run
      TRY(s->mailbox.alloc(xpbd::kMailboxWords * groups));
      HIP_TRY(hipMemset(s->mailbox.p, 0xff, s->mailbox.n * sizeof(double)));  // all empty
    }
  }
  TRY(s->lam.alloc(nt));
  TRY(s->lam.zero());
  TRY(s->d_old2new.upload(s->v_old2new.data(), nv));
  TRY(s->scratch_xyz.alloc(3 * size_t(nv)));
  TRY(s->flag.alloc(1));
  if (d->accumulate_mode == XPBD_ACCUM_CSR && d->layout == XPBD_LAYOUT_FLAT) {
    std::vector<int> adj(4 * size_t(nt));
    std::vector<int> fill(off_v.begin(), off_v.end() - 1);
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < 4; ++k) adj[fill[s->tets_new[4 * size_t(t) + k]]++] = 4 * t + k;
    TRY(s->adj_off.upload(off_v.data(), nv + 1));
    TRY(s->adj.upload(adj.data(), adj.size()));
    TRY(s->corner.alloc(4 * size_t(nt)));
    TRY(s->corner.zero());
  }
  TRY(build_maps(s.get()));
  TRY(s->partial.alloc(3 * size_t(std::max(s->ntb, s->n_patches))));
  TRY(s->partial.zero());
  SceneCtrl c0;
  std::memset(&c0, 0, sizeof(c0));
  c0.rho_hat = 0.999;  // neohookean_solver.py:65
  c0.omega = 1.0;
  const SceneCtrl c2[2] = {c0, c0};  // two slots (double-buffered controller state)
  TRY(s->ctrl.upload(c2, 2));
  TRY(s->logbuf.alloc(6 * size_t(std::max(1, d->constraints_iter))));
  TRY(s->logbuf.zero());
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->log_host),
                        sizeof(double) * 6 * std::max(1, d->constraints_iter), 0));
  TRY(s->table.alloc(1));
  TRY(build_scene(s.get()));
  TRY(sync_table(s.get(), s->prm.dt / s->prm.substeps_num, s->prm.damping));
  HIP_TRY(hipDeviceSynchronize());
  *out = s.release();
  return XPBD_OK;
}

int xpbd_destroy(xpbd_solver* s) {
  if (!s) return XPBD_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  delete s;
  return XPBD_OK;
}

#define CHECK_HANDLE(s)                                               \
  do {                                                                \
    if (!(s)) return set_error(XPBD_EINVAL, "solver handle is NULL"); \
    HIP_TRY(hipSetDevice((s)->device));                               \
  } while (0)

int xpbd_update_params(xpbd_solver* s, const xpbd_desc* d) {
  CHECK_HANDLE(s);
  if (!d) return set_error(XPBD_EINVAL, "desc is NULL");
  TRY(validate_params(d));
  if ((d->accumulate_mode == XPBD_ACCUM_CSR) != (s->prm.accumulate_mode == XPBD_ACCUM_CSR))
    return set_error(XPBD_EINVAL, "accumulate_mode is fixed at creation");
  if (d->block_size != s->prm.block_size)
    return set_error(XPBD_EINVAL, "block_size is fixed at creation");
  const int old_iters = s->prm.constraints_iter;
  xpbd_desc keep = s->prm;
  s->prm = *d;
  s->prm.rest_positions = nullptr;
  s->prm.tets = nullptr;
  s->prm.sdfs = nullptr;
  s->prm.n_sdf = keep.n_sdf;
  s->prm.reorder = keep.reorder;
  s->prm.device = keep.device;
  if (d->constraints_iter > old_iters) {
    TRY(stream_sync(s));
    TRY(s->logbuf.alloc(6 * size_t(d->constraints_iter)));
    TRY(s->logbuf.zero());
    if (s->log_host) (void)hipHostFree(s->log_host);
    s->log_host = nullptr;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->log_host),
                          sizeof(double) * 6 * d->constraints_iter, 0));
    s->host_scene.log = s->logbuf.p;
  }
  s->table_valid = false;
  return sync_table(s, s->prm.dt / s->prm.substeps_num, s->prm.damping);
}

int xpbd_set_sdfs(xpbd_solver* s, const xpbd_sdf* sdfs, int32_t n) {
  CHECK_HANDLE(s);
  TRY(store_sdfs(s, sdfs, n));
  s->table_valid = false;
  return sync_table(s, double(s->host_scene.dt), s->prm.damping);
}

int xpbd_set_gravity(xpbd_solver* s, const double g[3]) {
  CHECK_HANDLE(s);
  if (!g) return set_error(XPBD_EINVAL, "gravity is NULL");
  for (int k = 0; k < 3; ++k) s->prm.gravity[k] = g[k];
  s->table_valid = false;
  return XPBD_OK;  // uploaded before the next launch
}

int xpbd_compute_lumped_mass(xpbd_solver* s, double total_mass) {
  CHECK_HANDLE(s);
  std::vector<float> x, v;
  TRY(download_vec4_xyz(s, s->xs, x));
  TRY(download_vec4_xyz(s, s->vel, v));
  s->prm.total_mass = total_mass;
  lumped_mass_host(s, x, v, total_mass);
  return upload_inv_m(s);
}

int xpbd_get_positions(xpbd_solver* s, float* xyz) {
  CHECK_HANDLE(s);
  if (!xyz) return set_error(XPBD_EINVAL, "NULL buffer");
  HIP_TRY(xpbd::launch_gather_xyz(s->scratch_xyz.p, s->xs.p, s->d_old2new.p, s->nv, s->stream));
  HIP_TRY(hipMemcpyAsync(xyz, s->scratch_xyz.p, sizeof(float) * 3 * s->nv, hipMemcpyDeviceToHost,
                         s->stream));
  TRY(stream_sync(s));
  return XPBD_OK;
}

int xpbd_get_positions_device(xpbd_solver* s, void* dev_xyz) {
  CHECK_HANDLE(s);
  if (!dev_xyz) return set_error(XPBD_EINVAL, "NULL buffer");
  HIP_TRY(xpbd::launch_gather_xyz(static_cast<float*>(dev_xyz), s->xs.p, s->d_old2new.p, s->nv,
                                  s->stream));
  TRY(stream_sync(s));
  return XPBD_OK;
}

int xpbd_set_positions(xpbd_solver* s, const float* xyz) {
  CHECK_HANDLE(s);
  if (!xyz) return set_error(XPBD_EINVAL, "NULL buffer");
  TRY(stream_sync(s));
  HIP_TRY(hipMemcpy(s->scratch_xyz.p, xyz, sizeof(float) * 3 * s->nv, hipMemcpyHostToDevice));
  return set_positions_common(s, s->scratch_xyz.p);
}

int xpbd_set_positions_device(xpbd_solver* s, const void* dev_xyz) {
  CHECK_HANDLE(s);
  if (!dev_xyz) return set_error(XPBD_EINVAL, "NULL buffer");
  return set_positions_common(s, static_cast<const float*>(dev_xyz));
}

int xpbd_get_velocities(xpbd_solver* s, float* xyz) {
  CHECK_HANDLE(s);
  if (!xyz) return set_error(XPBD_EINVAL, "NULL buffer");
  HIP_TRY(xpbd::launch_gather_xyz(s->scratch_xyz.p, s->vel.p, s->d_old2new.p, s->nv, s->stream));
  HIP_TRY(hipMemcpyAsync(xyz, s->scratch_xyz.p, sizeof(float) * 3 * s->nv, hipMemcpyDeviceToHost,
                         s->stream));
  TRY(stream_sync(s));
  return XPBD_OK;
}

int xpbd_set_velocities(xpbd_solver* s, const float* xyz) {
  CHECK_HANDLE(s);
  if (!xyz) return set_error(XPBD_EINVAL, "NULL buffer");
  TRY(stream_sync(s));
  HIP_TRY(hipMemcpy(s->scratch_xyz.p, xyz, sizeof(float) * 3 * s->nv, hipMemcpyHostToDevice));
  HIP_TRY(xpbd::launch_scatter_xyz(s->vel.p, s->scratch_xyz.p, s->d_old2new.p, s->nv, 0, s->stream));
  TRY(stream_sync(s));
  return XPBD_OK;
}

int xpbd_get_predicted(xpbd_solver* s, float* xyz) {
  CHECK_HANDLE(s);
  if (!xyz) return set_error(XPBD_EINVAL, "NULL buffer");
  HIP_TRY(xpbd::launch_gather_xyz(s->scratch_xyz.p, s->pos.p, s->d_old2new.p, s->nv, s->stream));
  HIP_TRY(hipMemcpyAsync(xyz, s->scratch_xyz.p, sizeof(float) * 3 * s->nv, hipMemcpyDeviceToHost,
                         s->stream));
  TRY(stream_sync(s));
  return XPBD_OK;
}

int xpbd_get_masses(xpbd_solver* s, float* m, float* inv_m) {
  CHECK_HANDLE(s);
  for (int i = 0; i < s->nv; ++i) {
    const int j = s->v_old2new[i];
    if (m) m[i] = s->m_host[j];
    if (inv_m) inv_m[i] = s->inv_m_host[j];
  }
  return XPBD_OK;
}

int xpbd_get_degenerate_tets(xpbd_solver* s, int32_t* count, int32_t* first) {
  if (!s) return set_error(XPBD_EINVAL, "solver handle is NULL");
  if (count) *count = s->n_degenerate;
  if (first) *first = s->first_degenerate;
  return XPBD_OK;
}

int xpbd_get_multipliers(xpbd_solver* s, float* lambda_d, float* lambda_h) {
  CHECK_HANDLE(s);
  std::vector<float2> l(s->nt);
  TRY(stream_sync(s));
  HIP_TRY(hipMemcpy(l.data(), s->lam.p, sizeof(float2) * s->nt, hipMemcpyDeviceToHost));
  for (int t = 0; t < s->nt; ++t) {
    const int o = s->t_new2old[t];
    if (lambda_d) lambda_d[o] = l[t].x;
    if (lambda_h) lambda_h[o] = l[t].y;
  }
  return XPBD_OK;
}

int xpbd_get_visual_indices(xpbd_solver* s, uint32_t* out) {
  if (!s || !out) return set_error(XPBD_EINVAL, "NULL argument");
  // initialize_visual_indices (xpbd_base.py:78-87)
  static const int ind[4][3] = {{0, 2, 1}, {0, 3, 2}, {0, 1, 3}, {1, 2, 3}};
  for (int c = 0; c < s->nt; ++c)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 3; ++j)
        out[(size_t(c) * 4 + i) * 3 + j] = uint32_t(s->tets_orig[4 * size_t(c) + ind[i][j]]);
  return XPBD_OK;
}

int xpbd_apply_external_forces(xpbd_solver* s, double dt) {
  CHECK_HANDLE(s);
  TRY(sync_table(s, dt, s->prm.damping));
  HIP_TRY(xpbd::launch_predict(s->table.p, s->vmap.p, s->nvb, s->bs_v, 1, 0, 0, s->stream));
  return XPBD_OK;
}

int xpbd_solve_sdf_collision(xpbd_solver* s, double dt) {
  CHECK_HANDLE(s);
  TRY(sync_table(s, dt, s->prm.damping));
  HIP_TRY(xpbd::launch_predict(s->table.p, s->vmap.p, s->nvb, s->bs_v, 0, 1, 0, s->stream));
  return XPBD_OK;
}

int xpbd_solve_step(xpbd_solver* s, double dt) {
  CHECK_HANDLE(s);
  TRY(sync_table(s, dt, s->prm.damping));
  TRY(enqueue_solve_step(s, false));
  return fetch_log(s);
}

int xpbd_advance(xpbd_solver* s, double dt, double damping) {
  CHECK_HANDLE(s);
  TRY(sync_table(s, dt, damping));
  HIP_TRY(xpbd::launch_advance(s->table.p, s->vmap.p, s->nvb, s->bs_v, s->stream));
  return XPBD_OK;
}

int xpbd_substep(xpbd_solver* s, double sub_dt) {
  CHECK_HANDLE(s);
  return run_substep(s, sub_dt);
}

int xpbd_frame(xpbd_solver* s, int32_t n_frames) {
  CHECK_HANDLE(s);
  for (int f = 0; f < n_frames; ++f) {
    // solve (neohookean_solver.py:562-587)
    double frame_time_left = s->prm.frame_dt;
    s->substep = 0;
    while (frame_time_left > 0.0) {
      const double dt0 = std::min(s->prm.dt, frame_time_left);
      frame_time_left -= dt0;
      const double sub_dt = dt0 / s->prm.substeps_num;
      for (int k = 0; k < s->prm.substeps_num; ++k) {
        s->substep = k;
        TRY(run_substep(s, sub_dt));
      }
      s->time += dt0;
    }
    s->frame += 1;
  }
  return XPBD_OK;
}

int xpbd_synchronize(xpbd_solver* s) {
  CHECK_HANDLE(s);
  TRY(stream_sync(s));
  return XPBD_OK;
}

int xpbd_get_stream(xpbd_solver* s, void** stream) {
  if (!s || !stream) return set_error(XPBD_EINVAL, "NULL argument");
  *stream = static_cast<void*>(s->stream);
  return XPBD_OK;
}

int xpbd_set_logging(xpbd_solver* s, int32_t enabled) {
  if (!s) return set_error(XPBD_EINVAL, "solver handle is NULL");
  s->logging = enabled != 0;
  return XPBD_OK;
}

int xpbd_log_size(xpbd_solver* s, int64_t* rows) {
  if (!s || !rows) return set_error(XPBD_EINVAL, "NULL argument");
  *rows = int64_t(s->log_rows.size() / 8);
  return XPBD_OK;
}

int xpbd_get_log(xpbd_solver* s, double* rows, int64_t max_rows) {
  if (!s || !rows) return set_error(XPBD_EINVAL, "NULL argument");
  const size_t n = std::min<size_t>(size_t(std::max<int64_t>(max_rows, 0)), s->log_rows.size() / 8);
  std::memcpy(rows, s->log_rows.data(), n * 8 * sizeof(double));
  return XPBD_OK;
}

int xpbd_clear_log(xpbd_solver* s) {
  if (!s) return set_error(XPBD_EINVAL, "solver handle is NULL");
  s->log_rows.clear();
  return XPBD_OK;
}

int xpbd_get_state(xpbd_solver* s, int32_t* frame, int32_t* substep, double* time,
                   double* rho_hat) {
  CHECK_HANDLE(s);
  if (frame) *frame = s->frame;
  if (substep) *substep = s->substep;
  if (time) *time = s->time;
  if (rho_hat) {
    SceneCtrl c;
    TRY(stream_sync(s));
    // between substeps both controller slots hold the same state
    HIP_TRY(hipMemcpy(&c, s->ctrl.p, sizeof(c), hipMemcpyDeviceToHost));
    *rho_hat = c.rho_hat;
  }
  return XPBD_OK;
}

int xpbd_set_state(xpbd_solver* s, int32_t frame, double time, double rho_hat) {
  CHECK_HANDLE(s);
  SceneCtrl c;
  TRY(stream_sync(s));
  HIP_TRY(hipMemcpy(&c, s->ctrl.p, sizeof(c), hipMemcpyDeviceToHost));
  if (!std::isnan(rho_hat)) {
    c.rho_hat = rho_hat;
    const SceneCtrl c2[2] = {c, c};
    HIP_TRY(hipMemcpy(s->ctrl.p, c2, sizeof(c2), hipMemcpyHostToDevice));
  }
  s->frame = frame;
  s->time = time;
  s->substep = 0;
  s->sub_pos = 0;  // a restored state starts at a frame boundary
  return XPBD_OK;
}

int xpbd_get_sizes(xpbd_solver* s, int32_t* n_verts, int32_t* n_tets) {
  if (!s) return set_error(XPBD_EINVAL, "solver handle is NULL");
  if (n_verts) *n_verts = s->nv;
  if (n_tets) *n_tets = s->nt;
  return XPBD_OK;
}

int xpbd_set_barrier_timeout(xpbd_solver* s, int64_t microseconds) {
  if (!s) return set_error(XPBD_EINVAL, "solver handle is NULL");
  if (microseconds < 0 || microseconds > int64_t(1e9))
    return set_error(XPBD_EINVAL, "barrier timeout must be in [0, 1e9] microseconds");
  s->persist_timeout = (unsigned long long)microseconds * kWallClockTicksPerUs;
  return XPBD_OK;
}

int xpbd_set_exec_path(xpbd_solver* s, int32_t path) {
  if (!s) return set_error(XPBD_EINVAL, "solver handle is NULL");
  if (path != XPBD_PATH_KERNELS && path != XPBD_PATH_GRAPH && path != XPBD_PATH_PERSISTENT &&
      path != XPBD_PATH_AUTO)
    return set_error(XPBD_EINVAL, "unknown exec path %d", path);
  s->path = path;
  return XPBD_OK;
}

int xpbd_set_kernel_timing(xpbd_solver* s, int32_t enabled) {
  CHECK_HANDLE(s);
  if (enabled < 0 || enabled > 1000)
    return set_error(XPBD_EINVAL, "kernel timing mode %d outside [0, 1000]", enabled);
  TRY(events_collect(s));
  if (enabled > 1 && s->lam_save.n < s->lam.n) TRY(s->lam_save.alloc(s->lam.n));
  s->timing = enabled;
  s->timed_ms = 0.0;
  s->timed_launches = 0;
  return XPBD_OK;
}

int xpbd_get_kernel_time(xpbd_solver* s, double* total_ms, int64_t* launches) {
  CHECK_HANDLE(s);
  TRY(events_collect(s));
  if (total_ms) *total_ms = s->timed_ms;
  if (launches) *launches = s->timed_launches;
  return XPBD_OK;
}

int xpbd_patch_stats(const float* positions, const int32_t* tets, int32_t n_tets, int32_t n_verts,
                     int32_t* n_patches, int64_t* n_slots, int32_t* max_patch_verts,
                     int32_t* max_patch_tets) {
  if (!tets || n_tets < 0 || n_verts <= 0) return set_error(XPBD_EINVAL, "bad mesh");
  for (size_t i = 0; i < 4 * size_t(n_tets); ++i)
    if (tets[i] < 0 || tets[i] >= n_verts) return set_error(XPBD_EINVAL, "vertex id out of range");
  std::vector<int> t(tets, tets + 4 * size_t(n_tets));
  std::vector<int> cuts;
  if (positions) {  // the ordering the solver uses: bisection + first-appearance numbering
    std::vector<float> cent(3 * size_t(n_tets));
    for (int e = 0; e < n_tets; ++e)
      for (int j = 0; j < 3; ++j) {
        float c = 0.f;
        for (int k = 0; k < 4; ++k) c += 0.25f * positions[3 * size_t(tets[4 * size_t(e) + k]) + j];
        cent[3 * size_t(e) + j] = c;
      }
    std::vector<int> order;
    rcb_order(cent, tets, n_tets, n_verts, order, cuts);
    std::vector<int> newid(n_verts, -1);
    int next = 0;
    for (int e = 0; e < n_tets; ++e)
      for (int k = 0; k < 4; ++k) {
        const int g = tets[4 * size_t(order[e]) + k];
        if (newid[g] < 0) newid[g] = next++;
      }
    for (int g = 0; g < n_verts; ++g)
      if (newid[g] < 0) newid[g] = next++;
    for (int e = 0; e < n_tets; ++e)
      for (int k = 0; k < 4; ++k) t[4 * size_t(e) + k] = newid[tets[4 * size_t(order[e]) + k]];
  }
  Patches P;
  build_patches(t, n_verts, n_tets, P, cuts.empty() ? nullptr : &cuts);
  const int np = int(P.tet_begin.size()) - 1;
  int mv = 0, mt = 0;
  for (int p = 0; p < np; ++p) {
    mv = std::max(mv, P.vert_begin[p + 1] - P.vert_begin[p]);
    mt = std::max(mt, P.tet_begin[p + 1] - P.tet_begin[p]);
  }
  // every tet's local ids must map back to its own vertices
  for (int p = 0; p < np; ++p)
    for (int tt = P.tet_begin[p]; tt < P.tet_begin[p + 1]; ++tt)
      for (int k = 0; k < 4; ++k) {
        const int l = int((P.tet_local[tt].x >> (8 * k)) & 0xffu);
        if (P.vert_begin[p] + l >= P.vert_begin[p + 1] ||
            P.vert_ids[P.vert_begin[p] + l] != t[4 * size_t(tt) + k])
          return set_error(XPBD_ESTATE, "patch %d tet %d corner %d maps to the wrong vertex", p, tt, k);
      }
  // corner positions base(l) + rank cover [0, 4 * tets) once per patch, and
  // each vertex's slot lies in its own contiguous range, once
  std::vector<char> seen_slot(P.vert_ids.size(), 0);
  for (int p = 0; p < np; ++p) {
    const int tb = P.tet_begin[p], ntp = P.tet_begin[p + 1] - tb, vb = P.vert_begin[p];
    std::vector<char> seen(4 * size_t(ntp), 0);
    for (int tt = tb; tt < tb + ntp; ++tt)
      for (int k = 0; k < 4; ++k) {
        const int l = int((P.tet_local[tt].x >> (8 * k)) & 0xffu);
        const int rank = int((P.tet_local[tt].y >> (8 * k)) & 0xffu);
        const int base = P.vert_sum[vb + l].x & 0xffff, cnt = (P.vert_sum[vb + l].x >> 16) & 0x7fff;
        const int pos = base + rank;
        if (rank >= cnt || pos >= 4 * ntp || seen[pos])
          return set_error(XPBD_ESTATE, "patch %d: corner position %d invalid or reused", p, pos);
        seen[pos] = 1;
      }
    for (int i = vb; i < P.vert_begin[p + 1]; ++i) {
      const int g = P.vert_ids[i], dst = P.vert_sum[i].y;
      if (dst < P.vslot_off[g] || dst >= P.vslot_off[g + 1] || seen_slot[dst])
        return set_error(XPBD_ESTATE, "slot %d of vertex %d out of its range or reused", dst, g);
      if ((P.vert_sum[i].x < 0) != (dst == P.vslot_off[g]))
        return set_error(XPBD_ESTATE, "vertex %d: owner flag not on its first slot", g);
      seen_slot[dst] = 1;
    }
  }
  if (n_patches) *n_patches = np;
  if (n_slots) *n_slots = int64_t(P.vert_ids.size());
  if (max_patch_verts) *max_patch_verts = mv;
  if (max_patch_tets) *max_patch_tets = mt;
  return XPBD_OK;
}

int xpbd_check_finite(xpbd_solver* s) {
  CHECK_HANDLE(s);
  int zero = 0, hit = 0;
  HIP_TRY(hipMemcpyAsync(s->flag.p, &zero, sizeof(int), hipMemcpyHostToDevice, s->stream));
  HIP_TRY(xpbd::launch_check_finite(s->xs.p, s->nv, s->flag.p, s->stream));
  HIP_TRY(xpbd::launch_check_finite(s->pos.p, s->nv, s->flag.p, s->stream));
  HIP_TRY(hipMemcpyAsync(&hit, s->flag.p, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  TRY(stream_sync(s));
  if (hit) return set_error(XPBD_ENUMERIC, "non-finite vertex position in the solver state");
  return XPBD_OK;
}

}  // extern "C"

// ---- batched form ------------------------------------------------------------

struct xpbd_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<xpbd_solver*> scenes;
  xpbd_batch_opts opts{};
  DevBuf<SceneDev> table;
  std::vector<SceneDev> table_host;
  DevBuf<WorkItem> vmap;
  DevBuf<PatchItem> pitems;
  DevBuf<int2> prange_proj, prange_res;
  int nvb = 0, n_items = 0, grid_proj = 0, grid_res = 0;
  int max_scenes_per_range = 0;
  int64_t n_verts = 0, n_tets = 0;
  SceneDev proto{};  // flags and iteration count shared by the members
  hipGraphExec_t gexec = nullptr;
  int graph_substeps = 0;
  // where the scene table and work lists live: the batch's own buffers, or
  // (noalloc) the per-GPU array of xpbd_batch_set_members_noalloc
  bool noalloc = false;
  SceneDev* table_d = nullptr;
  WorkItem* vmap_d = nullptr;
  PatchItem* items_d = nullptr;
  int2* rp_d = nullptr;
  int2* rr_d = nullptr;
  // persistent substep: one workgroup per residual group of every member, the
  // launch's error word, and whether it has run since the word was last read
  DevBuf<PersistGroup> pgroups;
  PersistGroup* pgroups_d = nullptr;
  int n_pgroups = 0;       // entries of the groups list: the launch's grid
  int n_pgroups_live = 0;  // of which workgroups with work (the rest are holes)
  DevBuf<unsigned> perr;
  int persist_capacity = 0;
  int persist_xcd_capacity = 0;
  bool persist_pending = false;
  unsigned long long persist_timeout = kPersistTimeoutTicks;

  ~xpbd_batch() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// The batch whose lists each GPU's no-allocation array holds.
constexpr int kMaxNoallocDevices = 64;
std::mutex g_noalloc_mutex;
xpbd_batch* g_noalloc_owner[kMaxNoallocDevices] = {};

void noalloc_release(xpbd_batch* b) {
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (b->device >= 0 && b->device < kMaxNoallocDevices && g_noalloc_owner[b->device] == b)
    g_noalloc_owner[b->device] = nullptr;
}

// Members of one batch share what the launch sequence depends on: the
// iteration count, time step, substep count and mode flags (checked on the
// members' current parameters).
int batch_uniform_params(xpbd_solver* const* scenes, int n) {
  const xpbd_solver* s0 = scenes[0];
  for (int i = 1; i < n; ++i) {
    const xpbd_solver* s = scenes[i];
    const xpbd_desc &a = s->prm, &r = s0->prm;
    if (a.constraints_iter != r.constraints_iter || a.substeps_num != r.substeps_num ||
        a.frame_dt != r.frame_dt || a.dt != r.dt || a.cheb_enable != r.cheb_enable ||
        a.block_neohookean_enable != r.block_neohookean_enable ||
        a.monitor_convergence != r.monitor_convergence || a.accumulate_mode != r.accumulate_mode ||
        s->has_isolated != s0->has_isolated)
      return set_error(XPBD_EINVAL, "scene %d: iteration count, time step or mode flags differ", i);
  }
  return XPBD_OK;
}

int batch_sync_table(xpbd_batch* b) {
  // a member's parameters may have changed since the batch was built (for
  // example by xpbd_update_params): the members must still agree, and a
  // change of the shared sequence drops the captured graph
  if (batch_uniform_params(b->scenes.data(), int(b->scenes.size())) != XPBD_OK)
    return set_error(XPBD_ESTATE, "batch members no longer share the iteration count, time "
                                  "step or mode flags: re-target the batch");
  bool changed = b->table_host.size() != b->scenes.size();
  b->table_host.resize(b->scenes.size());
  for (size_t i = 0; i < b->scenes.size(); ++i) {
    xpbd_solver* s = b->scenes[i];
    SceneDev S = s->host_scene;  // the member's buffers, with its current parameters
    fill_scene_params(s, S, s->prm.dt / s->prm.substeps_num, s->prm.damping);
    if (changed || std::memcmp(&S, &b->table_host[i], sizeof(S)) != 0) {
      b->table_host[i] = S;
      changed = true;
    }
  }
  if (!changed) return XPBD_OK;
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipMemcpy(b->table_d, b->table_host.data(), sizeof(SceneDev) * b->table_host.size(),
                    hipMemcpyHostToDevice));
  const SceneDev& p = b->table_host[0];
  if (p.iters != b->proto.iters || p.fused != b->proto.fused || p.cheb != b->proto.cheb ||
      p.monitor != b->proto.monitor) {
    if (b->gexec) {
      HIP_TRY(hipGraphExecDestroy(b->gexec));
      b->gexec = nullptr;
      b->graph_substeps = 0;
    }
  }
  b->proto = p;
  return XPBD_OK;
}

// Members of one batch: distinct, on the batch's device, patch layout, and
// the same iteration count, time step and mode flags.  Their own streams are
// drained, since from now on they run on the batch's stream.
int batch_check_members(xpbd_solver* const* scenes, int n, int device) {
  xpbd_solver* s0 = scenes[0];
  if (!s0) return set_error(XPBD_EINVAL, "NULL scene");
  for (int i = 0; i < n; ++i) {
    xpbd_solver* s = scenes[i];
    if (!s) return set_error(XPBD_EINVAL, "NULL scene %d", i);
    for (int j = 0; j < i; ++j)
      if (scenes[j] == s) return set_error(XPBD_EINVAL, "scene %d appears twice", i);
    if (s->device != device) return set_error(XPBD_EINVAL, "scene %d on another device", i);
    if (s->layout != XPBD_LAYOUT_PATCH)
      return set_error(XPBD_EINVAL, "scene %d: batches need the patch layout", i);
  }
  TRY(batch_uniform_params(scenes, n));
  for (int i = 0; i < n; ++i) HIP_TRY(hipStreamSynchronize(scenes[i]->stream));
  return XPBD_OK;
}

// Places each member's persistent-substep workgroups on one XCD (dispatch is
// round-robin over the n_xcd XCDs by block index, so entry slot * n_xcd + x
// runs on XCD x): members in order, each to the XCD with the fewest
// workgroups so far, so that a member's scene barriers meet in one L2.
// Unused entries are holes (scene -1) that leave at once.  The list is left
// as it is (contiguous, members spread over the XCDs) when some member would
// not fit on one XCD beside those placed before it (`cap` workgroups
// resident per XCD).
int persist_xcd_layout(std::vector<PersistGroup>& pg, int n_xcd, int cap) {
  if (pg.empty() || n_xcd <= 1 || cap < 1) return XPBD_OK;
  std::vector<int> load(n_xcd, 0);
  std::vector<int2> at(pg.size());  // (xcd, slot) of each entry
  size_t i = 0;
  while (i < pg.size()) {
    size_t e = i;
    while (e < pg.size() && pg[e].scene == pg[i].scene) ++e;
    const int n = int(e - i);
    const int x = int(std::min_element(load.begin(), load.end()) - load.begin());
    if (load[x] + n > cap) return XPBD_OK;
    for (int k = 0; k < n; ++k) at[i + k] = make_int2(x, load[x] + k);
    load[x] += n;
    i = e;
  }
  const int rows = *std::max_element(load.begin(), load.end());
  std::vector<PersistGroup> out(size_t(rows) * n_xcd, PersistGroup{-1, 0, 0, 0});
  for (size_t k = 0; k < pg.size(); ++k) out[size_t(at[k].y) * n_xcd + at[k].x] = pg[k];
  pg.swap(out);
  return XPBD_OK;
}

// The batch's concatenated work lists (vertex blocks and patch items in member
// order), its persistent grids and ranges, and its scene table, for the given
// members.  They go to the batch's own device buffers (reused when large
// enough from an earlier membership) or, with `noalloc`, to the per-GPU
// no-allocation array; a membership that does not fit there is refused
// before the batch changes.
int batch_build(xpbd_batch* b, xpbd_solver* const* scenes, int n_scenes, bool noalloc) {
  const xpbd_batch_opts& o = b->opts;
  std::vector<WorkItem> vm;
  std::vector<PatchItem> items;
  std::vector<PersistGroup> pg;
  int64_t n_verts = 0, n_tets = 0;
  for (int i = 0; i < n_scenes; ++i) {
    xpbd_solver* s = scenes[i];
    for (int v = 0; v < s->nv; v += o.vertex_block) vm.push_back({i, v});
    if (s->layout == XPBD_LAYOUT_PATCH)
      for (int g = 0; g < s->host_scene.n_partials; ++g) pg.push_back({i, g, int(items.size()), 0});
    for (PatchItem it : s->pitems_host) {
      it.scene = i;
      items.push_back(it);
    }
    n_verts += s->nv;
    n_tets += s->nt;
  }
  const int nx = o.xcd_split ? xcd_count() : 1;
  const int n_pg_live = int(pg.size());
  if (nx > 1) TRY(persist_xcd_layout(pg, nx, b->persist_xcd_capacity));
  std::vector<WorkItem> vmo(vm.size());
  const std::vector<int> order = xcd_order(int(vm.size()), nx);
  for (size_t k = 0; k < vm.size(); ++k) vmo[k] = vm[order[k]];
  const int n_items = int(items.size());
  int occ_p = 0, occ_r = 0, n_cu = 0;
  HIP_TRY(xpbd::patch_kernel_occupancy(&occ_p, &occ_r));
  HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, b->device));
  const long long full_p = (long long)n_cu * std::max(1, occ_p);
  const long long full_r = (long long)n_cu * std::max(1, occ_r);
  const int grid_proj = int(std::max(1LL, std::min<long long>(n_items, full_p * o.grid_pct / 100)));
  int grid_res = int(std::max(1LL, std::min<long long>(n_items, full_r * o.grid_pct / 100)));
  std::vector<int2> rp, rr;
  TRY(xcd_ranges(n_items, grid_proj, nx, rp));
  TRY(grouped_ranges(items, grid_res, nx, rr, &grid_res));
  if (noalloc) {
    auto a256 = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t bytes = a256(sizeof(WorkItem) * vmo.size()) + a256(sizeof(PatchItem) * items.size()) +
                         a256(sizeof(int2) * rp.size()) + a256(sizeof(int2) * rr.size()) +
                         a256(sizeof(PersistGroup) * pg.size());
    if (n_scenes > XPBD_NOALLOC_MAX_SCENES)
      return set_error(XPBD_EINVAL, "%d members: a batch re-targeted without allocation holds "
                       "at most %d", n_scenes, XPBD_NOALLOC_MAX_SCENES);
    if (bytes > size_t(XPBD_NOALLOC_LIST_BYTES))
      return set_error(XPBD_EINVAL, "the members' work lists need %zu bytes; a batch re-targeted "
                       "without allocation holds at most %d", bytes, XPBD_NOALLOC_LIST_BYTES);
    void *table = nullptr, *lists = nullptr;
    HIP_TRY(xpbd::noalloc_scratch(b->device, &table, &lists));
    unsigned char* at = static_cast<unsigned char*>(lists);
    auto place = [&](auto*& dst, const auto& v) -> int {
      using T = typename std::decay_t<decltype(v)>::value_type;
      dst = reinterpret_cast<std::remove_reference_t<decltype(dst)>>(at);
      if (!v.empty()) HIP_TRY(hipMemcpy(at, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
      at += a256(sizeof(T) * v.size());
      return XPBD_OK;
    };
    {
      std::lock_guard<std::mutex> lock(g_noalloc_mutex);
      g_noalloc_owner[b->device] = b;  // a batch that held the array loses its lists
    }
    TRY(place(b->vmap_d, vmo));
    TRY(place(b->items_d, items));
    TRY(place(b->rp_d, rp));
    TRY(place(b->rr_d, rr));
    TRY(place(b->pgroups_d, pg));
    b->table_d = static_cast<SceneDev*>(table);
    b->noalloc = true;
  } else {
    noalloc_release(b);
    TRY(b->vmap.upload_reuse(vmo.data(), vmo.size()));
    TRY(b->pitems.upload_reuse(items.data(), items.size()));
    TRY(b->prange_proj.upload_reuse(rp.data(), rp.size()));
    TRY(b->prange_res.upload_reuse(rr.data(), rr.size()));
    TRY(b->pgroups.upload_reuse(pg.data(), pg.size()));
    if (b->table.cap < size_t(n_scenes)) TRY(b->table.alloc(n_scenes));
    b->table.n = size_t(n_scenes);
    b->vmap_d = b->vmap.p;
    b->items_d = b->pitems.p;
    b->rp_d = b->prange_proj.p;
    b->rr_d = b->prange_res.p;
    b->pgroups_d = b->pgroups.p;
    b->table_d = b->table.p;
    b->noalloc = false;
  }
  b->n_pgroups = int(pg.size());
  b->n_pgroups_live = n_pg_live;
  b->scenes.assign(scenes, scenes + n_scenes);
  b->n_verts = n_verts;
  b->n_tets = n_tets;
  b->nvb = int(vmo.size());
  b->n_items = n_items;
  b->grid_proj = grid_proj;
  b->grid_res = grid_res;
  b->max_scenes_per_range = 0;
  for (const int2& r : rp) {
    int k = r.y > r.x ? 1 : 0;
    for (int i = r.x + 1; i < r.y; ++i) k += items[i].scene != items[i - 1].scene;
    b->max_scenes_per_range = std::max(b->max_scenes_per_range, k);
  }
  b->table_host.clear();  // forces the upload
  b->proto = b->scenes[0]->host_scene;
  TRY(batch_sync_table(b));
  return XPBD_OK;
}

// One substep of every member: predict + collide + save, the Jacobi
// iterations, the final blend and advance (the fused patch sequence of
// enqueue_iterations, over the concatenated work list).
int batch_enqueue_substep(xpbd_batch* b) {
  const SceneDev& P = b->proto;
  const SceneDev* T = b->table_d;
  const int bs = b->opts.vertex_block;
  const int ns = int(b->scenes.size());
  HIP_TRY(xpbd::launch_predict(T, b->vmap_d, b->nvb, bs, 1, 1, P.cheb, b->stream));
  for (int it = 0; it < P.iters; ++it) {
    HIP_TRY(xpbd::launch_patch_project(T, b->items_d, b->rp_d, b->grid_proj, it, ns > 1,
                                       b->stream));
    HIP_TRY(xpbd::launch_patch_average(T, b->vmap_d, b->nvb, bs, it, b->stream));
    if (P.cheb || P.monitor) {
      HIP_TRY(xpbd::launch_patch_residual(T, b->items_d, b->rr_d, b->grid_res, b->stream));
      if (!P.fused || it == P.iters - 1) HIP_TRY(xpbd::launch_reduce(T, ns, it, b->stream));
    }
    if (P.cheb && !P.fused) HIP_TRY(xpbd::launch_blend(T, b->vmap_d, b->nvb, bs, b->stream));
  }
  if (P.fused && P.iters > 0)
    HIP_TRY(xpbd::launch_final_blend(T, b->vmap_d, b->nvb, bs, P.iters, b->stream));
  HIP_TRY(xpbd::launch_advance(T, b->vmap_d, b->nvb, bs, b->stream));
  return XPBD_OK;
}

// Per-member frame accounting: members may join a batch mid-frame.
void batch_count_substep(xpbd_batch* b) {
  for (xpbd_solver* s : b->scenes) {
    s->substep = s->sub_pos;
    if (++s->sub_pos == s->prm.substeps_num) {
      s->sub_pos = 0;
      s->time += s->prm.dt;
      s->frame += 1;
    }
  }
}

// Whether the batch's next substeps take the persistent substep kernel: every
// member could take it alone, and all their workgroups fit on the device at
// once.  opts.persistent 1 (automatic) also asks, as XPBD_PATH_AUTO does for one
// scene, that no member's barrier spans more than kAutoPersistGroups
// workgroups, and that the launch stays within kAutoBatchGroups.
bool batch_persistent(const xpbd_batch* b) {
  if (b->opts.persistent == 0 || b->n_pgroups_live < 1 ||
      b->n_pgroups_live > b->persist_capacity)
    return false;
  for (const xpbd_solver* s : b->scenes) {
    if (!persistent_ok(s)) return false;
    if (b->opts.persistent == 1 && s->host_scene.n_partials > kAutoPersistGroups) return false;
  }
  return b->opts.persistent == 2 || b->n_pgroups_live <= kAutoBatchGroups;
}

bool batch_at_frame_start(const xpbd_batch* b) {
  for (const xpbd_solver* s : b->scenes)
    if (s->sub_pos != 0) return false;
  return true;
}

int batch_run(xpbd_batch* b, int n_substeps) {
  if (b->noalloc) {
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_noalloc_owner[b->device] != b)
      return set_error(XPBD_ESTATE, "another batch's lists replaced this batch's in the "
                                    "no-allocation array of device %d: re-target it", b->device);
  }
  TRY(batch_sync_table(b));
  if (batch_persistent(b)) {
    for (int k = 0; k < n_substeps; ++k) {
      HIP_TRY(xpbd::launch_batch_substep_persistent(b->table_d, b->items_d, b->pgroups_d,
                                                    b->n_pgroups, b->perr.p, b->persist_timeout,
                                                    b->stream));
      b->persist_pending = true;
      batch_count_substep(b);
    }
    return XPBD_OK;
  }
  if (b->opts.persistent == 2)
    return set_error(XPBD_EINVAL, "the persistent substep does not apply to this batch (patch "
                                  "layout without isolated vertices, at least one iteration, at "
                                  "most %d groups per member and %d in all)", kPersistMaxGroups,
                     b->persist_capacity);
  if (!b->opts.graph) {
    for (int k = 0; k < n_substeps; ++k) {
      TRY(batch_enqueue_substep(b));
      batch_count_substep(b);
    }
    return XPBD_OK;
  }
  const int per = b->opts.graph_substeps;
  int done = 0;
  while (done < n_substeps) {
    const int chunk = (per > 1 && n_substeps - done >= per && batch_at_frame_start(b)) ? per : 1;
    if (!b->gexec || b->graph_substeps != chunk) {
      if (b->gexec) {
        HIP_TRY(hipGraphExecDestroy(b->gexec));
        b->gexec = nullptr;
      }
      hipGraph_t g;
      HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = XPBD_OK;
      for (int k = 0; k < chunk && rc == XPBD_OK; ++k) rc = batch_enqueue_substep(b);
      const hipError_t ec = hipStreamEndCapture(b->stream, &g);
      if (rc != XPBD_OK) return rc;
      HIP_TRY(ec);
      const hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIP_TRY(ei);
      b->graph_substeps = chunk;
    }
    HIP_TRY(hipGraphLaunch(b->gexec, b->stream));
    for (int k = 0; k < chunk; ++k) batch_count_substep(b);
    done += chunk;
  }
  return XPBD_OK;
}

}  // namespace

extern "C" {

int xpbd_get_device_bytes(xpbd_solver* s, int64_t* bytes) {
  if (!s || !bytes) return set_error(XPBD_EINVAL, "NULL argument");
  *bytes = s->device_bytes();
  return XPBD_OK;
}

int xpbd_batch_opts_init(xpbd_batch_opts* o) {
  if (!o) return set_error(XPBD_EINVAL, "opts is NULL");
  o->graph = 1;
  o->graph_substeps = 1;
  o->vertex_block = 256;
  o->grid_pct = 100;
  o->xcd_split = 1;
  o->persistent = 1;
  return XPBD_OK;
}

int xpbd_batch_create(xpbd_solver* const* scenes, int32_t n_scenes, const xpbd_batch_opts* opts,
                      xpbd_batch** out) {
  if (!scenes || n_scenes <= 0 || !out) return set_error(XPBD_EINVAL, "batch needs scenes");
  *out = nullptr;
  auto b = std::make_unique<xpbd_batch>();
  if (opts)
    b->opts = *opts;
  else
    xpbd_batch_opts_init(&b->opts);
  const xpbd_batch_opts& o = b->opts;
  if (o.vertex_block < 64 || o.vertex_block > 1024 || o.vertex_block % 64)
    return set_error(XPBD_EINVAL, "vertex_block must be a multiple of 64 in [64, 1024]");
  if (o.grid_pct < 1 || o.grid_pct > 100) return set_error(XPBD_EINVAL, "grid_pct in [1, 100]");
  if (o.persistent < 0 || o.persistent > 2) return set_error(XPBD_EINVAL, "persistent in 0..2");
  if (!scenes[0]) return set_error(XPBD_EINVAL, "NULL scene");
  b->device = scenes[0]->device;
  HIP_TRY(hipSetDevice(b->device));
  TRY(batch_check_members(scenes, n_scenes, b->device));
  HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  {
    int occ = 0, n_cu = 0;
    HIP_TRY(xpbd::persistent_occupancy(&occ));
    HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, b->device));
    b->persist_capacity = n_cu * occ;
    b->persist_xcd_capacity = n_cu / xcd_count() * occ;
  }
  TRY(b->perr.alloc(1));
  TRY(b->perr.zero());
  TRY(batch_build(b.get(), scenes, n_scenes, false));
  *out = b.release();
  return XPBD_OK;
}

int xpbd_batch_set_members(xpbd_batch* b, xpbd_solver* const* scenes, int32_t n_scenes) {
  if (!b) return set_error(XPBD_EINVAL, "batch is NULL");
  if (!scenes || n_scenes <= 0) return set_error(XPBD_EINVAL, "batch needs scenes");
  HIP_TRY(hipSetDevice(b->device));
  TRY(batch_check_members(scenes, n_scenes, b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));  // the lists below are in use until then
  if (b->gexec) {
    HIP_TRY(hipGraphExecDestroy(b->gexec));
    b->gexec = nullptr;
    b->graph_substeps = 0;
  }
  return batch_build(b, scenes, n_scenes, false);
}

int xpbd_batch_set_members_noalloc(xpbd_batch* b, xpbd_solver* const* scenes, int32_t n_scenes) {
  if (!b) return set_error(XPBD_EINVAL, "batch is NULL");
  if (!scenes || n_scenes <= 0) return set_error(XPBD_EINVAL, "batch needs scenes");
  if (b->opts.graph)
    return set_error(XPBD_EINVAL, "a batch re-targeted without allocation must not replay "
                                  "graphs (opts.graph = 0)");
  if (b->device < 0 || b->device >= kMaxNoallocDevices)
    return set_error(XPBD_EINVAL, "device %d out of range", b->device);
  HIP_TRY(hipSetDevice(b->device));
  TRY(batch_check_members(scenes, n_scenes, b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));  // the lists below are in use until then
  return batch_build(b, scenes, n_scenes, true);
}

int xpbd_batch_destroy(xpbd_batch* b) {
  if (!b) return XPBD_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  noalloc_release(b);
  delete b;
  return XPBD_OK;
}

int xpbd_batch_substep(xpbd_batch* b) {
  if (!b) return set_error(XPBD_EINVAL, "batch is NULL");
  HIP_TRY(hipSetDevice(b->device));
  return batch_run(b, 1);
}

int xpbd_batch_frame(xpbd_batch* b, int32_t n_frames) {
  if (!b) return set_error(XPBD_EINVAL, "batch is NULL");
  HIP_TRY(hipSetDevice(b->device));
  const xpbd_desc& d = b->scenes[0]->prm;
  // frame chunking of solve() (neohookean_solver.py:566-571); with dt ==
  // frame_dt a frame is exactly substeps_num substeps
  if (d.dt != d.frame_dt)
    return set_error(XPBD_EINVAL, "batch frames need dt == frame_dt; step substeps instead");
  if (!batch_at_frame_start(b))
    return set_error(XPBD_ESTATE, "a member's frame is partly stepped; finish it with substeps");
  return batch_run(b, n_frames * d.substeps_num);
}

int xpbd_batch_synchronize(xpbd_batch* b) {
  if (!b) return set_error(XPBD_EINVAL, "batch is NULL");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (!b->persist_pending) return XPBD_OK;
  b->persist_pending = false;
  unsigned err = 0;
  HIP_TRY(hipMemcpy(&err, b->perr.p, sizeof(err), hipMemcpyDeviceToHost));
  if (err) {
    TRY(b->perr.zero());
    return set_error(XPBD_EHIP, "persistent batch substep: a scene barrier timed out (its "
                                "workgroups were not all resident); the results are invalid");
  }
  return XPBD_OK;
}

int xpbd_batch_set_barrier_timeout(xpbd_batch* b, int64_t microseconds) {
  if (!b) return set_error(XPBD_EINVAL, "batch is NULL");
  if (microseconds < 0 || microseconds > int64_t(1e9))
    return set_error(XPBD_EINVAL, "barrier timeout must be in [0, 1e9] microseconds");
  b->persist_timeout = (unsigned long long)microseconds * kWallClockTicksPerUs;
  return XPBD_OK;
}

int xpbd_batch_path(xpbd_batch* b, int32_t* path) {
  if (!b || !path) return set_error(XPBD_EINVAL, "NULL argument");
  TRY(batch_sync_table(b));
  *path = batch_persistent(b) ? XPBD_PATH_PERSISTENT
                              : (b->opts.graph ? XPBD_PATH_GRAPH : XPBD_PATH_KERNELS);
  return XPBD_OK;
}

int xpbd_batch_get_stream(xpbd_batch* b, void** stream) {
  if (!b || !stream) return set_error(XPBD_EINVAL, "NULL argument");
  *stream = static_cast<void*>(b->stream);
  return XPBD_OK;
}

int xpbd_batch_grid(xpbd_batch* b, int32_t* grid_project, int32_t* grid_residual,
                    int32_t* max_scenes_per_range) {
  if (!b) return set_error(XPBD_EINVAL, "batch is NULL");
  if (grid_project) *grid_project = b->grid_proj;
  if (grid_residual) *grid_residual = b->grid_res;
  if (max_scenes_per_range) *max_scenes_per_range = b->max_scenes_per_range;
  return XPBD_OK;
}

int xpbd_batch_info(xpbd_batch* b, int32_t* n_scenes, int64_t* n_verts, int64_t* n_tets,
                    int32_t* n_patches) {
  if (!b) return set_error(XPBD_EINVAL, "batch is NULL");
  if (n_scenes) *n_scenes = int32_t(b->scenes.size());
  if (n_verts) *n_verts = b->n_verts;
  if (n_tets) *n_tets = b->n_tets;
  if (n_patches) *n_patches = b->n_items;
  return XPBD_OK;
}

}  // extern "C"
