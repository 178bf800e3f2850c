// Host-callable launch wrappers of the gfx950 kernels in xpbd_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

namespace xpbd {

struct SceneDev;
struct WorkItem;
struct PatchItem;

hipError_t launch_predict(const SceneDev* scenes, const WorkItem* vmap, int nblocks, int bs,
                          int external, int collide, int save_prev, hipStream_t st);
hipError_t launch_project(const SceneDev* scenes, const WorkItem* tmap, int nblocks, int bs,
                          int first, hipStream_t st);
hipError_t launch_average(const SceneDev* scenes, const WorkItem* vmap, int nblocks, int bs,
                          hipStream_t st);
hipError_t launch_residual(const SceneDev* scenes, const WorkItem* tmap, int nblocks, int bs,
                           hipStream_t st);
// multi_scene: the patch ranges may cross scenes (batches)
hipError_t launch_patch_project(const SceneDev* scenes, const PatchItem* items,
                                const int2* ranges, int nblocks, int it, bool multi_scene,
                                hipStream_t st, hipEvent_t start = nullptr,
                                hipEvent_t stop = nullptr);
hipError_t launch_patch_average(const SceneDev* scenes, const WorkItem* vmap, int nblocks,
                                int bs, int it, hipStream_t st);
hipError_t launch_patch_residual(const SceneDev* scenes, const PatchItem* items,
                                 const int2* ranges, int nblocks, hipStream_t st);
hipError_t patch_kernel_occupancy(int* project_blocks_per_cu, int* residual_blocks_per_cu);
hipError_t launch_final_blend(const SceneDev* scenes, const WorkItem* vmap, int nblocks, int bs,
                              int n_iter, hipStream_t st);
hipError_t launch_reduce(const SceneDev* scenes, int n_scenes, int it, hipStream_t st);
hipError_t launch_blend(const SceneDev* scenes, const WorkItem* vmap, int nblocks, int bs,
                        hipStream_t st);
hipError_t launch_advance(const SceneDev* scenes, const WorkItem* vmap, int nblocks, int bs,
                          hipStream_t st);
// One substep of one scene in one launch (n_groups workgroups, all resident);
// a scene barrier gives up after timeout_ticks of the 100 MHz wall clock.
// xcd_stride > 1 places every workgroup on XCD `xcd` (grid n_groups *
// xcd_stride, round-robin dispatch; the other blocks leave at once).
hipError_t launch_substep_persistent(const SceneDev* scenes, const PatchItem* items, int n_groups,
                                    unsigned long long timeout_ticks, int xcd_stride, int xcd,
                                    hipStream_t st);
// One substep of every member of a batch in one launch: workgroup b does
// groups[b] (n_groups in all, holes marked by scene -1 leave at once; the
// others all resident); a barrier that gives up sets *err.
hipError_t launch_batch_substep_persistent(const SceneDev* scenes, const PatchItem* items,
                                          const PersistGroup* groups, int n_groups,
                                          unsigned* err, unsigned long long timeout_ticks,
                                          hipStream_t st);
// Stores *dst = src in stream order (no host wait).
hipError_t launch_store_scene(SceneDev* dst, const SceneDev& src, hipStream_t st);
// Workgroups of either persistent substep kernel resident per CU.
hipError_t persistent_occupancy(int* blocks_per_cu);
hipError_t launch_scatter_xyz(float4* dst, const float* src, const int* old_to_new, int n,
                              int keep_w, hipStream_t st);
hipError_t launch_gather_xyz(float* dst, const float4* src, const int* old_to_new, int n,
                             hipStream_t st);
hipError_t launch_check_finite(const float4* pos, int n, int* flag, hipStream_t st);
// Device addresses, on `device` (which must be current), of the no-allocation
// scene table (XPBD_NOALLOC_MAX_SCENES entries) and work-list area
// (XPBD_NOALLOC_LIST_BYTES).
hipError_t noalloc_scratch(int device, void** table, void** lists);

}  // namespace xpbd
