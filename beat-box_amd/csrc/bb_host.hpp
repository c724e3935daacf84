// bb_host.hpp — host-side internals shared by the C-ABI (bb_capi.cpp) and
// the serving loop (bb_server.cpp): the batch object, its device arrays, and
// the helpers that lay worlds out in them and upload their initial state.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <functional>
#include <vector>

#include "../../include/bb_capi.h"
#include "bb_dev.hpp"

namespace bb {
// One step of a lane: k_team first when `team`, then the per-world stages.
void launch_step(const Lane& l, hipStream_t st, bool team);
// One step of up to two lanes on one stream, stage by stage (both lanes'
// kernels of a stage back to back), with events ev[0..7] around the stages.
void launch_step_timed(const Lane* lanes, int nl, hipStream_t st, hipEvent_t* ev);
void launch_copy(const float4* src, float4* dst, size_t n, hipStream_t st);
void launch_sum_counts(const Dev& d, unsigned long long* out, hipStream_t st);
void launch_energy(const Dev& d, double* out, hipStream_t st);
void launch_instances(const Dev& d, uint32_t w, uint32_t n, float* dst, hipStream_t st);
// World w's `nodes` (= 2n-1) BVH nodes as 40-byte LBVHNode records into dst.
void launch_lbvh_export(const Dev& d, uint32_t w, uint32_t nodes, void* dst, hipStream_t st);
uint32_t team_capacity();
}  // namespace bb

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "beatbox: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return BB_ERR_DEVICE;                                                        \
    }                                                                              \
  } while (0)

namespace bb {
namespace host {

// Rows per world whose body state k_pgs may keep in LDS: 80 B each, within
// the 160 KB a gfx950 workgroup can allocate.
constexpr uint32_t kMaxPgsLdsRows = 2048;

struct HostWorld {
  bb_world_desc desc;
  std::vector<bb_body_desc> bodies;  // id order
  bool has_voxel = false;            // a body is a voxel shape (kModeVoxel)
  uint32_t pair_scale = 1;           // pair capacity multiplier, doubled on overflow
  bool no_team = false;              // stepped by one workgroup whatever its size
  bool skip_ws_once = false;
  bool drop_prev = false;
  uint32_t body_off = 0, act_off = 0, pair_off = 0, pair_cap = 0, n = 0, a = 0;
};

// Device allocations of a batch.  With count_only set nothing is allocated
// and `bytes` sums what would be (used to price a world for admission).
// With carve_base set nothing is allocated either: the arrays are cut, 256-byte
// aligned, from the caller's range [carve_base, carve_base + carve_cap), a
// request past its end yields a null pointer, `bytes` is the end of the last
// cut, and release() frees nothing.
struct Buffers {
  std::vector<void*> ptrs;
  std::vector<size_t> sizes;  // bytes of each allocation, in ptrs order
  uint64_t bytes = 0;
  bool count_only = false;
  char* carve_base = nullptr;
  size_t carve_cap = 0;
  template <class T>
  T* alloc(size_t count) {
    void* p = nullptr;
    size_t n = std::max<size_t>(count, 1) * sizeof(T);
    if (count_only) {
      bytes += n;
      return nullptr;
    }
    if (carve_base) {
      const size_t off = (bytes + 255u) & ~(size_t)255u;
      if (off <= carve_cap && n <= carve_cap - off) {
        p = carve_base + off;
        bytes = off + n;
      }
    } else if (hipMalloc(&p, n) != hipSuccess) {
      p = nullptr;
    } else {
      bytes += n;
    }
    ptrs.push_back(p);
    sizes.push_back(n);
    return (T*)p;
  }
  void release() {
    if (!carve_base)
      for (void* p : ptrs)
        if (p) (void)hipFree(p);
    ptrs.clear();
    sizes.clear();
    bytes = 0;
  }
};

// Host images of the initial state of a contiguous range of body rows (both
// parities get the same rows) and active slots.  Pointers may address
// pinned memory for asynchronous uploads.
struct Images {
  uint32_t* id;
  uint32_t* flags;
  uint32_t* entries;
  uint32_t* rank;
  uint32_t* active_ids;
  float4* pos;
  float4* rot;
  float4* vel;
  float4* omg;
  float4* pmin;
  float4* pmax;
  float4* mass;
  float4* ii0;
  float4* ii1;
  float4* ii2;
};
// Bytes of image data per body row and per active slot.
constexpr size_t kImageBodyBytes = 4 * 4 + 10 * 16;
constexpr size_t kImageActiveBytes = 4;
// Carves an Images set for nb rows and na active slots out of `mem`
// (kImageBodyBytes * nb + kImageActiveBytes * na bytes, 16-byte aligned).
Images carve_images(void* mem, uint32_t nb, uint32_t na);

}  // namespace host
}  // namespace bb

struct bb_batch {
  std::vector<bb::host::HostWorld> worlds;
  bb::host::Buffers buf;
  bool built = false;
  bool dirty = true;
  bool stepped = false;
  bb::Dev base{};  // pointers common to both parities
  bb::Rows rows[2];
  uint32_t* entries[2] = {nullptr, nullptr};
  uint2* pairs[2] = {nullptr, nullptr};
  uint32_t* npairs[2] = {nullptr, nullptr};
  uint32_t* pstart[2] = {nullptr, nullptr};
  bb::Manifolds man[2];
  int latest = 0;  // buffer set holding the newest state
  uint32_t NB = 0, NA = 0, NP = 0;
  hipStream_t stream = nullptr;
  uint32_t block = 0;
  uint32_t pgs_lds = 512;  // BB_PARAM_PGS_LDS_ROWS
  uint32_t team_bodies = 8192;  // BB_PARAM_TEAM_BODIES: worlds this large run in k_team
  // BB_PARAM_LANE_BODIES / _LANE_BLOCK: worlds this large (0: none) run the
  // per-world kernels with lane_block threads on a second stream
  uint32_t lane_bodies = 384;
  uint32_t lane_block = 512;
  uint32_t lane_lds = 1024;  // BB_PARAM_LANE_LDS_ROWS: LDS rows of k_pgs in that lane
  uint32_t team_spin = 0;    // BB_PARAM_TEAM_SPIN: spins of a team barrier wait (0: default)
  uint32_t pair_scale = 1;   // BB_PARAM_PAIR_SCALE: pair slots per world, times the default
  uint32_t fused = 0;        // BB_PARAM_FUSED: each lane's step as one k_fused launch
  uint32_t sort_blocks = 0;  // BB_PARAM_SORT_BLOCKS (0: kTeamSortBlocks)
  uint32_t avbd_lanes = 0;   // BB_PARAM_AVBD_LANES (0: 4)
  uint32_t pgs_prefetch = 0; // BB_PARAM_PGS_PREFETCH
  uint32_t walk_budget = 0;  // BB_PARAM_WALK_BUDGET (0: kWalkBudgetDefault)
  uint32_t team_blocks = 0;  // BB_PARAM_TEAM_BLOCKS (0: half the resident workgroups)
  uint32_t pgs_block = 0;    // BB_PARAM_PGS_BLOCK (0: the lane's workgroup size)
  uint32_t n_wide = 0;  // worlds in that lane: the first n_wide of the order table
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bb::TeamTable team_table{};
  uint64_t step_index = 0;
  std::vector<bb::WorldInfo> infos;
  hipGraphExec_t graph = nullptr;
  int graph_parity = -1;
  // bb_batch_build_split: the build stage of the slowest team, split by the
  // team's clock marks, summed over the last bb_batch_profile call's steps
  float build_ms[BB_BUILD_COUNT] = {};
  float avbd_ms[BB_AVBD_COUNT] = {};
  // voxel shape pools shared by the worlds (host copies; the device arrays
  // are allocated at their maximum sizes and refreshed when vox_dirty)
  std::vector<bb::VoxShape> vox_shapes;
  std::vector<bb_voxel_shape_info> vox_info;
  std::vector<float> vox_density;
  std::vector<uint32_t> vox_occ, vox_surf;
  bool vox_dirty = false;
  bb::VoxShape* d_vshape = nullptr;
  uint32_t* d_vocc = nullptr;
  uint32_t* d_vsurf = nullptr;
};

namespace bb {
namespace host {

uint32_t pair_capacity(uint32_t n);
bool has_device();
bool valid_desc(const bb_world_desc* d);
bool valid_body(const bb_body_desc& b);
// valid_body, or a voxel body whose shape id names one of `shapes` shapes.
bool valid_body_with_voxels(const bb_body_desc& b, size_t shapes);
// SceneManager::build_voxel_shape into the batch's host pools.
bb_status add_voxel_shape(bb_batch* b, const bb_voxel_shape_desc* d, bb_voxel_shape_info* out);
// Copies the host voxel pools to the device when they changed.
bb_status upload_voxels(bb_batch* b);
// Parity mode keeps the reference cap of 1,024 bodies per world.
bool over_reference_cap(const bb_world_desc& d, uint32_t n);
WorldInfo info_of(const HostWorld& hw);
// Allocates (or, with A.count_only, prices, or, with A.carve_base, cuts from
// the caller's range) every device array of a batch with W world slots, NB
// body rows, NA active slots and NP pair slots, and points b->base / rows /
// man at them.  Allocated arrays get their zero-start tables cleared; a carved
// layout is left for its owner to clear.
bb_status alloc_arrays(bb_batch* b, Buffers& A, uint32_t W, uint32_t NB, uint32_t NA, uint32_t NP);
// Device bytes of a batch of these capacities.
uint64_t device_bytes(uint32_t W, uint32_t NB, uint32_t NA, uint32_t NP);
// Writes world hw's initial rows into img at row r0 and active slot a0 of
// the image range (image index = device index - range start).
void fill_images(const HostWorld& hw, uint32_t r0, uint32_t a0, const Images& img,
                 const std::vector<bb_voxel_shape_info>* vox = nullptr);
// Copies an image range to the device: rows [body_lo, body_lo+nb) of both
// parities and of the property arrays, active slots [act_lo, act_lo+na).
bb_status upload_images(bb_batch* b, const Images& img, uint32_t body_lo, uint32_t nb, uint32_t act_lo,
                        uint32_t na, hipStream_t st);
// Clears the per-world device state a fresh world starts from: previous pair
// counts (both parities), pair-row offsets, sleep timers, velocities of the
// previous step, and the world's counters.
bb_status reset_world_state(bb_batch* b, uint32_t w, hipStream_t st);
uint32_t pgs_lds_rows(const bb_batch* b);
// Doubles the pair capacity of a built single-world batch, keeping every
// device array's contents (both parities, the manifolds included).
bb_status grow_pairs(bb_batch* b);
// Reads the batch's overflow word (set by any world whose step flagged an
// overflow since it was last cleared) after the batch stream has finished.
bb_status read_overflow_any(bb_batch* b, uint32_t* out);
// Clears the sticky overflow words of every world and the batch's word.
bb_status clear_overflow(bb_batch* b);
// Counters of one world (raw device slots, ST_COUNT of them) as the C-ABI
// struct.
void stats_from_raw(const uint32_t* raw, uint64_t step_index, bb_step_stats* out);
uint32_t pick_block(const bb_batch* b);
// The lanes of one step over the device view d: the wide lane (when any world
// is in it) and the rest.  Returns the lane count (1 or 2).
int lanes_of(const bb_batch* b, const Dev& d, Lane* out);
// Chooses the worlds that run as workgroup teams in k_team (at least
// team_bodies bodies, workgroups in proportion to size within what can be
// resident at once), sets their kModeTeam bit in b->infos and fills
// b->team_table.  Returns the workgroups of the k_team launch (0: none).
uint32_t assign_teams(bb_batch* b, uint32_t grid_cap);
// Workgroup-to-world order of the per-world kernels: descending body count
// (ties by index), team worlds last (their workgroups exit at once).
void world_order(const std::vector<WorldInfo>& infos, uint32_t* out);
Dev dev_for(const bb_batch* b, int cur);
// 1 when a world of the batch runs the per-island solver path.
uint32_t any_island(const bb_batch* b);
// One step of every world of a built batch on b->stream (flips the parity).
bb_status launch_one(bb_batch* b);
// Bodies of world w (n >= its body count) in `order`, after b->stream drains.
bb_status read_world(bb_batch* b, uint32_t w, bb_body_state* out, uint32_t n, bb_order order);
// Counters of world w's last step.
bb_status world_stats(bb_batch* b, uint32_t w, bb_step_stats* out);
// Checkpoint arrays of a world: what its next step reads from the device
// (rows of one parity in row order, the id -> row map, pair-row offsets,
// island slots, the pair list and the manifold slots that warm start it) and
// its counters.  world_state_bytes is their size for n rows and p pairs;
// world_state_arrays hands each array's device address (world slot w,
// parity `parity`) and size to fn, in blob order.
size_t world_state_bytes(uint32_t n, uint32_t p);
void world_state_arrays(bb_batch* b, uint32_t w, int parity, uint32_t n, uint32_t p,
                        const std::function<void(void*, size_t)>& fn);
// Host check of the indices in a world's checkpoint arrays (ids a
// permutation of [0, n), the id -> row map consistent with them, pair rows
// and manifold bodies below n, at most 8 contacts per manifold).
bool valid_blob_indices(const char* in, uint32_t n, uint32_t p);
// The voxel-shape section of a checkpoint: per shape its record (dims, cell
// size, density, word count) and occupancy words.  put_shapes writes b's
// (shapes_bytes of them); take_shapes adds a section's shapes to b as ids
// 1, 2, ... (null: refused); shapes_section_bytes checks one on the host
// (0: malformed; a section without shapes is also 0 bytes).
size_t shapes_bytes(const bb_batch* b);
char* put_shapes(const bb_batch* b, char* o);
const char* take_shapes(bb_batch* b, const char* in, const char* end, uint32_t nshapes, uint32_t shape_words);
size_t shapes_section_bytes(const char* in, const char* end, uint32_t nshapes, uint32_t shape_words);

}  // namespace host
}  // namespace bb
