// bb_noalloc.hip — bb_world_step_noalloc: a fresh world stepped and read back
// without any device allocation.
//
// bb_world_create + bb_world_add_bodies + bb_world_step + bb_world_read_bodies
// allocate the world's device arrays with hipMalloc.  This variant cuts the
// same arrays, in the same layout, from `scratch`, a static device array of
// this file, and queues everything (clearing, upload, the step kernels) on the
// caller's stream.  The step kernels are the ones bb_world_step launches, with
// the same per-world parameters, so the results are the same bit for bit; the
// only difference is that the pair capacity is fixed by what the array holds
// instead of growing on overflow.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <vector>

#include "bb_host.hpp"

using namespace bb;
using namespace bb::host;

namespace {

constexpr size_t kScratchBytes = BB_NOALLOC_SCRATCH_BYTES;
// Every array of a world stepped by bb_world_step_noalloc.  In host code the
// name stands for a host-side shadow, not device memory: host code uses the
// device address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock[kMaxDevices];          // one call at a time per device
char* g_scratch[kMaxDevices] = {};       // device address of `scratch`, per device

// The device address of `scratch` on device `dev` (current), looked up on the
// first call on that device.  Caller holds g_lock[dev].
bb_status scratch_on(int dev, char** out) {
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      std::fprintf(stderr, "beatbox: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      return BB_ERR_DEVICE;
    }
    g_scratch[dev] = static_cast<char*>(p);
  }
  *out = g_scratch[dev];
  return BB_OK;
}

bb_status step_noalloc(const bb_world_desc* desc, const bb_body_desc* bodies, uint32_t n, uint32_t nsteps,
                       bb_body_state* out, uint32_t out_n, bb_step_stats* stats, void* stream) {
  if (!valid_desc(desc) || !bodies || !out || n == 0 || out_n < n) return BB_ERR_INVALID;
  if (n > BB_NOALLOC_MAX_BODIES) {
    std::fprintf(stderr, "beatbox: bb_world_step_noalloc: %u bodies exceed its maximum of %u\n", n,
                 (unsigned)BB_NOALLOC_MAX_BODIES);
    return BB_ERR_CAPACITY;
  }
  uint32_t a = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (!valid_body(bodies[i])) return BB_ERR_INVALID;  // boxes and spheres only
    if (bodies[i].flags & BB_BODY_DYNAMIC) ++a;
  }
  if (desc->capacity_bodies && n > desc->capacity_bodies) return BB_ERR_CAPACITY;
  if (over_reference_cap(*desc, n)) return BB_ERR_CAPACITY;
  if (!has_device()) return BB_ERR_NO_DEVICE;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return BB_ERR_DEVICE;
  std::lock_guard<std::mutex> guard(g_lock[dev]);
  char* base = nullptr;
  bb_status rc = scratch_on(dev, &base);
  if (rc != BB_OK) return rc;
  const hipStream_t st = static_cast<hipStream_t>(stream);

  // a one-world batch on the caller's stream: one lane, no teams; a world
  // the size of bb_world_step's wide lane gets that lane's workgroup size and
  // LDS rows (settings that leave every result unchanged)
  bb_batch b;
  b.stream = st;
  b.team_bodies = 0;
  if (n >= b.lane_bodies) {
    b.block = b.lane_block;
    b.pgs_lds = b.lane_lds;
  }
  b.lane_bodies = 0;
  HostWorld hw;
  hw.desc = *desc;
  hw.bodies.assign(bodies, bodies + n);
  hw.n = n;
  hw.a = a;
  b.worlds.push_back(hw);
  // the largest pair capacity (the default times 16, 8, 4, 2 or 1) that fits
  rc = BB_ERR_CAPACITY;
  uint32_t cap = 0;
  for (uint32_t scale = 16; scale >= 1 && rc != BB_OK; scale /= 2) {
    cap = pair_capacity(n) * scale;
    b.buf = Buffers{};
    b.buf.carve_base = base;
    b.buf.carve_cap = kScratchBytes;
    rc = alloc_arrays(&b, b.buf, 1, n, a, cap);
  }
  if (rc != BB_OK) {
    std::fprintf(stderr, "beatbox: bb_world_step_noalloc: %u bodies do not fit %zu bytes of scratch\n", n,
                 kScratchBytes);
    return BB_ERR_CAPACITY;
  }
  HostWorld& w0 = b.worlds[0];
  w0.pair_cap = cap;
  b.NB = n;
  b.NA = a;
  b.NP = cap;
  // the array keeps the previous call's contents: clear all of this layout
  HIPCHK(hipMemsetAsync(base, 0, b.buf.bytes, st));
  std::vector<char> mem(kImageBodyBytes * (size_t)n + kImageActiveBytes * std::max(a, 1u) + 64);
  const Images img = carve_images((void*)(((uintptr_t)mem.data() + 15) & ~(uintptr_t)15), n, std::max(a, 1u));
  fill_images(w0, 0, 0, img, nullptr);
  rc = upload_images(&b, img, 0, n, 0, a, st);
  if (rc != BB_OK) return rc;
  // world table: the first step has no previous manifolds, the others do
  WorldInfo wi[2] = {info_of(w0), info_of(w0)};
  wi[0].mode |= kModeDropPrev;
  const uint32_t order0 = 0;
  HIPCHK(hipMemcpyAsync((void*)b.base.order, &order0, 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(b.base.worlds, &wi[0], sizeof(WorldInfo), hipMemcpyHostToDevice, st));
  b.infos.assign(1, wi[1]);
  b.latest = 1;  // the first step reads set 1 and writes set 0
  b.base.pgs_lcap = pgs_lds_rows(&b);
  b.base.any_voxel = 0;
  b.base.any_island = any_island(&b);
  b.base.team_grid = 0;
  b.built = true;
  b.dirty = false;
  b.vox_dirty = false;
  for (uint32_t s = 0; s < nsteps; ++s) {
    if (s == 1) HIPCHK(hipMemcpyAsync(b.base.worlds, &wi[1], sizeof(WorldInfo), hipMemcpyHostToDevice, st));
    rc = launch_one(&b);
    if (rc != BB_OK) {
      (void)hipStreamSynchronize(st);
      return rc;
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  rc = read_world(&b, 0, out, out_n, BB_ORDER_BY_ID);
  if (rc == BB_OK && stats) rc = world_stats(&b, 0, stats);
  if (rc != BB_OK) return rc;
  uint32_t sticky = 0;
  HIPCHK(hipMemcpy(&sticky, b.base.stats + ST_OVF_STICKY, 4, hipMemcpyDeviceToHost));
  return sticky ? BB_ERR_OVERFLOW : BB_OK;
}

}  // namespace

extern "C" bb_status bb_world_step_noalloc(const bb_world_desc* desc, const bb_body_desc* bodies, uint32_t n,
                                           uint32_t nsteps, bb_body_state* out, uint32_t out_n,
                                           bb_step_stats* stats, void* stream) {
  return step_noalloc(desc, bodies, n, nsteps, out, out_n, stats, stream);
}
