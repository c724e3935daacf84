// bb_server.cpp — serving loop over a capacity-allocated batch.
//
// The device arrays are allocated once for the world slots and the body,
// active and pair capacities that the HBM budget holds.  A world occupies one
// slot and one contiguous range in each arena (first fit); the kernels find
// it through its slot's WorldInfo, so admitting or retiring a world only
// rewrites its table entry and its ranges.  Empty slots have n = 0 and their
// workgroups return at once, so a step launch (one hipGraph per parity over a
// slot range) never changes shape.
//
// Pools: the slots are split into up to three pools by world size (small
// worlds, one workgroup each; mid-sized worlds, one larger workgroup each;
// large worlds, stepped by workgroup teams).  Each pool has its own slot
// range, queue, stream, graphs, parity and pipeline, and steps at its own
// cadence: a small world's step latency is its own pool's kernels' time, not
// that of the slowest resident world.  The arenas and the HBM budget are
// shared, and every pool touches only its own worlds' rows, so pools on
// different streams never write the same memory.
//
// Host/device pipelining, per pool: per pool step the host stages new
// worlds' images and table entries in a pinned chunk, enqueues their uploads,
// the step and the read-back of retiring worlds on the pool's stream, and
// records an event.  A chunk is reused only after its step's event has
// completed, and at most kAhead steps of a pool are in flight.  A retired
// world's ranges are handed out again at once: an upload into them by another
// pool waits for the read-back's event first (same pool: stream order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <unordered_map>
#include <thread>
#include <vector>

#include "bb_host.hpp"

using namespace bb;
using namespace bb::host;

namespace {

constexpr int kAhead = 2;         // server steps in flight
constexpr int kChunks = kAhead + 1;

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator of [0, cap) with coalescing frees.
struct RangeAlloc {
  std::map<uint32_t, uint32_t> free_;  // start -> length
  void reset(uint32_t cap) {
    free_.clear();
    if (cap) free_[0] = cap;
  }
  bool alloc(uint32_t n, uint32_t* out) {
    if (n == 0) {
      *out = 0;
      return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      const uint32_t start = it->first, len = it->second;
      free_.erase(it);
      if (len > n) free_[start + n] = len - n;
      *out = start;
      return true;
    }
    return false;
  }
  void release(uint32_t start, uint32_t n) {
    if (n == 0) return;
    auto it = free_.emplace(start, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
};

// Pinned staging chunk, grown on demand (only while its step is complete).
struct Chunk {
  char* mem = nullptr;
  size_t cap = 0, used = 0;
  bb_status reserve(size_t bytes) {
    if (bytes <= cap) return BB_OK;
    size_t n = cap ? cap : (size_t)1 << 20;
    while (n < bytes) n *= 2;
    if (mem) (void)hipHostFree(mem);
    mem = nullptr;
    cap = 0;
    if (hipHostMalloc((void**)&mem, n, hipHostMallocDefault) != hipSuccess) return BB_ERR_DEVICE;
    cap = n;
    return BB_OK;
  }
  char* take(size_t bytes) {
    used = (used + 15) & ~(size_t)15;
    char* p = mem + used;
    used += bytes;
    return p;
  }
};

struct Request {
  uint64_t ticket;
  bb_world_desc desc;
  std::vector<bb_body_desc> bodies;
  uint32_t steps;
  uint32_t active;
  uint64_t bytes;
  double submit_ms;
  // a request runs again from its first step when a step of it overflowed:
  // with twice the pair capacity (pairs past it were dropped), or by one
  // workgroup instead of a team (the team left a step unfinished)
  uint32_t pair_scale = 1;
  bool no_team = false;
  uint32_t reruns = 0;
  // a world carried over by a checkpoint: its state arrays (the order of
  // world_state_arrays), its pair count and the steps it had done; admission
  // puts them into the newest parity of its pool instead of a fresh start
  std::vector<char> state;
  uint32_t npairs = 0, done = 0;
};

// Pair-capacity doublings and team fallbacks a request may take before it
// is reported as failed.
constexpr uint32_t kMaxReruns = 17;

struct Slot {
  bool used = false;
  uint64_t ticket = 0;
  uint32_t steps = 0, done = 0;
  uint64_t bytes = 0;
  uint32_t body_off = 0, n = 0, act_off = 0, a = 0, pair_off = 0, pair_cap = 0;
  bool fresh = false;  // admitted for the step being assembled
};

struct Readback {  // a retiring world's rows, staged in a pinned chunk
  uint64_t ticket;
  uint32_t n;
  uint64_t step_index;
  uint32_t *id, *flags, *islot, *stats;
  float4 *pos, *rot, *vel, *omg;
};

struct StepRecord {
  hipEvent_t done = nullptr;   // after the step (and its read-backs)
  hipEvent_t admit = nullptr;  // before the step, after this step's uploads
  int chunk = 0;
  uint64_t carried = 0;        // world-steps of worlds resident before
  uint64_t admitted = 0;       // world-steps of worlds admitted for this step
  uint64_t bodies = 0;
  std::vector<Readback> readbacks;
};

struct Result {
  std::vector<bb_body_state> bodies;
  bb_step_stats stats;
  bool failed = false;
};

// A pool: the worlds whose body count lies in [lo, hi), in slots [w0, w1).
struct Pool {
  uint32_t lo = 0, hi = 0xFFFFFFFFu;
  uint32_t w0 = 0, w1 = 0;
  uint32_t block = 256;  // per-world workgroup size
  uint32_t lcap = 512;   // PGS LDS rows of its per-world kernels
  bool teams = false;    // the pool launches k_team
  hipStream_t stream = nullptr;
  std::deque<Request> queue;
  std::vector<uint32_t> free_slots;
  std::deque<StepRecord> inflight;
  Chunk chunks[kChunks];
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int latest = 1;  // parity holding the newest rows of this pool's worlds
  uint64_t step_counter = 0;
  bool table_dirty = false;
  hipEvent_t last_done = nullptr;
  bool have_prev_done = false;
  std::vector<std::pair<float, uint64_t>> step_samples;  // (ms, world-steps)
  uint64_t world_steps = 0;
  uint32_t running = 0;
};

}  // namespace

struct bb_server {
  bb_batch* b = nullptr;
  bb_server_desc desc{};
  uint32_t W = 0, NB = 0, NA = 0, NP = 0;
  uint64_t cW = 0, cB = 0, cA = 0, cP = 0;  // device bytes per slot/body/active/pair slot
  RangeAlloc rb, ra, rp;
  std::vector<Slot> slots;
  std::vector<Pool> pools;
  std::unordered_map<uint64_t, Result> results;
  std::unordered_map<uint64_t, uint32_t> running;  // ticket -> slot
  // requests on the device or in a read-back, kept to run again on overflow
  std::unordered_map<uint64_t, Request> origin;
  // ranges freed by a read-back still in flight on a pool's stream: another
  // pool that uploads into them waits for that event (hipStreamWaitEvent)
  std::vector<std::pair<uint32_t, hipEvent_t>> released;  // (pool, its newest step event)
  uint64_t reruns = 0;
  std::vector<hipEvent_t> spare_events;
  uint64_t next_ticket = 1, step_counter = 0;
  uint64_t bytes_in_use = 0, resident_bodies = 0;
  uint64_t submitted = 0, admitted_n = 0, retired_n = 0;
  uint64_t world_steps = 0, body_steps = 0;
  hipEvent_t first = nullptr;
  bool started = false;
  std::vector<double> queue_samples;
  double device_ms = 0.0;
};

namespace {

hipEvent_t take_event(bb_server* s) {
  if (!s->spare_events.empty()) {
    hipEvent_t e = s->spare_events.back();
    s->spare_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Pair slots of a request per unit of its pair scale and per default pair
// capacity: six when a body is a voxel shape (a voxel pair takes up to six
// slots, one per contact-face code), else one.
uint32_t pair_factor(const std::vector<bb_body_desc>& bodies) {
  for (const bb_body_desc& bd : bodies)
    if (BB_SHAPE_KIND(bd.shape) == BB_SHAPE_VOXEL) return 6u;
  return 1u;
}

uint64_t price(const bb_server* s, uint32_t n, uint32_t a, uint32_t pair_scale = 1) {
  return s->cW + s->cB * n + s->cA * a + s->cP * (uint64_t)pair_capacity(n) * pair_scale;
}

uint32_t pool_of(const bb_server* s, uint32_t n) {
  for (uint32_t p = 0; p < (uint32_t)s->pools.size(); ++p)
    if (n >= s->pools[p].lo && n < s->pools[p].hi) return p;
  return (uint32_t)s->pools.size() - 1;
}

void to_states(const Readback& r, Result& res) {
  res.bodies.resize(r.n);
  for (uint32_t row = 0; row < r.n; ++row) {
    const uint32_t id = r.id[row];
    if (id >= r.n) continue;
    bb_body_state& o = res.bodies[id];
    o.id = id;
    o.flags = r.flags[row];
    o.island_index = r.islot[row];
    o.row = row;
    o.position[0] = r.pos[row].x; o.position[1] = r.pos[row].y; o.position[2] = r.pos[row].z;
    o.rotation[0] = r.rot[row].x; o.rotation[1] = r.rot[row].y; o.rotation[2] = r.rot[row].z;
    o.rotation[3] = r.rot[row].w;
    o.velocity[0] = r.vel[row].x; o.velocity[1] = r.vel[row].y; o.velocity[2] = r.vel[row].z;
    o.omega[0] = r.omg[row].x; o.omega[1] = r.omg[row].y; o.omega[2] = r.omg[row].z;
  }
  stats_from_raw(r.stats, r.step_index, &res.stats);
}

// Completes the oldest in-flight step of pool p: latency samples, results.
bb_status finish_oldest(bb_server* s, uint32_t p) {
  Pool& P = s->pools[p];
  StepRecord& rec = P.inflight.front();
  HIPCHK(hipEventSynchronize(rec.done));
  float ms = 0.0f;
  if (P.have_prev_done) {
    HIPCHK(hipEventElapsedTime(&ms, P.last_done, rec.done));
    if (rec.carried) P.step_samples.emplace_back(ms, rec.carried);
  }
  if (rec.admitted) {
    float am = 0.0f;
    HIPCHK(hipEventElapsedTime(&am, rec.admit, rec.done));
    P.step_samples.emplace_back(am, rec.admitted);
  }
  float total = 0.0f;
  HIPCHK(hipEventElapsedTime(&total, s->first, rec.done));
  s->device_ms = std::max(s->device_ms, (double)total);
  for (const Readback& r : rec.readbacks) {
    auto oit = s->origin.find(r.ticket);
    const uint32_t ovf = r.stats[ST_OVF_STICKY];
    if (ovf && oit != s->origin.end() && oit->second.reruns < kMaxReruns) {
      // a step of this world dropped pairs or was left unfinished by its
      // team: run the request again from its first step, ahead of its pool's
      // queue, if it fits the budget and the pair arena that way
      const Request& o = oit->second;
      const uint32_t scale = (ovf & 1u) ? o.pair_scale * 2u : o.pair_scale;
      const uint32_t nb = (uint32_t)o.bodies.size(), pf = pair_factor(o.bodies);
      const uint64_t bytes = price(s, nb, o.active, scale * pf);
      if (bytes <= s->desc.hbm_budget_bytes && (uint64_t)pair_capacity(nb) * scale * pf <= s->NP) {
        Request q = std::move(oit->second);
        s->origin.erase(oit);
        q.state.clear();  // a rerun starts from the first step
        q.done = 0;
        q.npairs = 0;
        q.pair_scale = scale;
        if (ovf & 2u) q.no_team = true;
        ++q.reruns;
        ++s->reruns;
        q.bytes = bytes;
        P.queue.push_front(std::move(q));
        continue;
      }
    }
    Result& res = s->results[r.ticket];
    to_states(r, res);
    res.failed = ovf != 0;
    if (oit != s->origin.end()) s->origin.erase(oit);
  }
  // read-backs of this pool whose event is complete no longer hold anyone up
  for (size_t i = 0; i < s->released.size();)
    if (s->released[i].second == rec.done) s->released.erase(s->released.begin() + i);
    else ++i;
  if (P.have_prev_done) s->spare_events.push_back(P.last_done);
  P.last_done = rec.done;
  P.have_prev_done = true;
  s->spare_events.push_back(rec.admit);
  P.chunks[rec.chunk].used = 0;
  P.inflight.pop_front();
  return BB_OK;
}

// The step graphs of pool p (one per parity): k_team first when the pool has
// teams, then the per-world stage kernels over its slots in its order.
bb_status ensure_graphs(bb_server* s, uint32_t p) {
  Pool& P = s->pools[p];
  for (int par = 0; par < 2; ++par) {
    if (P.graph[par]) continue;
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(P.stream, hipStreamCaptureModeThreadLocal));
    Dev d = dev_for(s->b, par);
    d.order = d.order + P.w0;
    d.pgs_lcap = P.lcap;
    if (!P.teams) d.team_grid = 0;
    // with voxel shapes registered any team may hold voxel bodies
    d.team_vox = s->b->vox_shapes.empty() ? 0u : 1u;
    launch_step(Lane{d, P.block, P.w1 - P.w0}, P.stream, true);
    HIPCHK(hipStreamEndCapture(P.stream, &g));
    HIPCHK(hipGraphInstantiate(&P.graph[par], g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
  }
  return BB_OK;
}

// One step of pool p: admission from its queue, the step of its slots,
// read-back and release of its worlds that finish with it.
bb_status pool_step(bb_server* s, uint32_t p) {
  bb_batch* b = s->b;
  Pool& P = s->pools[p];
  while ((int)P.inflight.size() >= kAhead) {
    bb_status st = finish_oldest(s, p);
    if (st != BB_OK) return st;
  }
  const int chunk = (int)(P.step_counter % kChunks);
  Chunk& ch = P.chunks[chunk];
  // ---- admission: first come first served while the budget, the ranges,
  // a slot and the resident-body cap allow
  std::vector<uint32_t> admitted;
  size_t stage_bytes = 0;
  const double t_now = now_ms();
  while (!P.queue.empty() && !P.free_slots.empty()) {
    Request& q = P.queue.front();
    const uint32_t n = (uint32_t)q.bodies.size();
    if (s->bytes_in_use + q.bytes > s->desc.hbm_budget_bytes) break;
    if (s->desc.max_resident_bodies && s->resident_bodies > 0 &&
        s->resident_bodies + n > s->desc.max_resident_bodies)
      break;
    uint32_t bo, ao, po;
    if (!s->rb.alloc(n, &bo)) break;
    if (!s->ra.alloc(q.active, &ao)) {
      s->rb.release(bo, n);
      break;
    }
    const uint32_t pf = pair_factor(q.bodies);
    const uint32_t pcap = pair_capacity(n) * q.pair_scale * pf;
    if ((uint64_t)pcap > s->NP || !s->rp.alloc(pcap, &po)) {
      s->rb.release(bo, n);
      s->ra.release(ao, q.active);
      break;
    }
    const uint32_t w = P.free_slots.back();
    P.free_slots.pop_back();
    Slot& sl = s->slots[w];
    sl = Slot{};
    sl.used = true;
    sl.fresh = true;
    sl.ticket = q.ticket;
    sl.steps = q.steps;
    sl.done = q.done;
    sl.bytes = q.bytes;
    sl.body_off = bo;
    sl.n = n;
    sl.act_off = ao;
    sl.a = q.active;
    sl.pair_off = po;
    sl.pair_cap = pcap;
    HostWorld& hw = b->worlds[w];
    hw = HostWorld{};
    hw.desc = q.desc;
    hw.bodies = q.bodies;
    hw.pair_scale = q.pair_scale;
    hw.has_voxel = pf > 1u;
    hw.no_team = q.no_team;
    hw.n = n;
    hw.a = q.active;
    hw.body_off = bo;
    hw.act_off = ao;
    hw.pair_off = po;
    hw.pair_cap = sl.pair_cap;
    s->bytes_in_use += q.bytes;
    s->resident_bodies += n;
    s->running[q.ticket] = w;
    ++P.running;
    if (!q.reruns && q.state.empty()) {
      s->queue_samples.push_back(t_now - q.submit_ms);
      ++s->admitted_n;
    }
    const uint64_t tk = q.ticket;
    stage_bytes += kImageBodyBytes * std::max(n, 1u) + kImageActiveBytes * std::max(q.active, 1u) + 64;
    stage_bytes += q.state.size() + 64;
    admitted.push_back(w);
    s->origin[tk] = std::move(q);
    P.queue.pop_front();
  }
  // ---- read-back sizes of the worlds that finish with this step
  std::vector<uint32_t> retiring;
  for (uint32_t w = P.w0; w < P.w1; ++w) {
    const Slot& sl = s->slots[w];
    if (sl.used && sl.done + 1 == sl.steps) {
      retiring.push_back(w);
      stage_bytes += (size_t)sl.n * (4 * 3 + 16 * 4) + ST_COUNT * 4 + 128;
    }
  }
  stage_bytes += (sizeof(WorldInfo) + sizeof(uint32_t)) * (size_t)(P.w1 - P.w0) + sizeof(TeamTable) + 96;
  bb_status st = ch.reserve(stage_bytes + 4096);
  if (st != BB_OK) return st;
  ch.used = 0;
  hipStream_t stream = P.stream;
  // ---- ranges another pool freed this recently may still be read back on
  // that pool's stream: order the uploads after those read-backs
  if (!admitted.empty())
    for (const auto& r : s->released)
      if (r.first != p) HIPCHK(hipStreamWaitEvent(stream, r.second, 0));
  // ---- uploads of the admitted worlds and their table entries
  for (uint32_t w : admitted) {
    HostWorld& hw = b->worlds[w];
    const uint32_t n = hw.n, a = hw.a;
    void* mem = ch.take(kImageBodyBytes * std::max(n, 1u) + kImageActiveBytes * std::max(a, 1u));
    const Images img = carve_images(mem, std::max(n, 1u), std::max(a, 1u));
    fill_images(hw, 0, 0, img, &b->vox_info);
    st = upload_images(b, img, hw.body_off, n, hw.act_off, a, stream);
    if (st != BB_OK) return st;
    st = reset_world_state(b, w, stream);
    if (st != BB_OK) return st;
    Request& org = s->origin[s->slots[w].ticket];
    if (!org.state.empty()) {
      // a checkpointed world: its rows, pairs, manifolds and counters go
      // into the parity the next step reads (after the fresh image above,
      // in stream order), with its pair count, so it warm starts as it
      // would have in the server that wrote the checkpoint
      char* src = (char*)ch.take(org.state.size() + 4);
      std::memcpy(src, org.state.data(), org.state.size());
      std::memcpy(src + org.state.size(), &org.npairs, 4);
      hipError_t rc = hipSuccess;
      const char* at = src;
      world_state_arrays(b, w, P.latest, n, org.npairs, [&](void* dst, size_t bytes) {
        if (rc == hipSuccess && bytes) rc = hipMemcpyAsync(dst, at, bytes, hipMemcpyHostToDevice, stream);
        at += bytes;
      });
      if (rc == hipSuccess)
        rc = hipMemcpyAsync(b->npairs[P.latest] + w, src + org.state.size(), 4, hipMemcpyHostToDevice, stream);
      HIPCHK(rc);
      org.state.clear();
      org.state.shrink_to_fit();
    }
    WorldInfo wi = info_of(hw);
    wi.mode &= kModeVoxel;  // the cleared pair counts already give an empty previous step
    b->infos[w] = wi;
    hw.bodies.clear();
    hw.bodies.shrink_to_fit();
  }
  // ---- world table: after an admission or a retirement, choose the team
  // worlds again and upload this pool's part of the table (and the team
  // table, which only a pool with teams uses)
  if (!admitted.empty() || P.table_dirty) {
    if (P.teams && b->base.team_grid) {
      assign_teams(b, b->base.team_grid);
      TeamTable* tt = (TeamTable*)ch.take(sizeof(TeamTable));
      *tt = b->team_table;
      HIPCHK(hipMemcpyAsync(b->base.teams, tt, sizeof(TeamTable), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemsetAsync(b->base.team_ctl, 0, sizeof(TeamCtl) * kMaxTeams, stream));
    }
    const uint32_t nw = P.w1 - P.w0;
    WorldInfo* tab = (WorldInfo*)ch.take(sizeof(WorldInfo) * nw);
    std::memcpy(tab, b->infos.data() + P.w0, sizeof(WorldInfo) * nw);
    HIPCHK(hipMemcpyAsync(b->base.worlds + P.w0, tab, sizeof(WorldInfo) * nw, hipMemcpyHostToDevice, stream));
    uint32_t* ord = (uint32_t*)ch.take(sizeof(uint32_t) * nw);
    const std::vector<WorldInfo> part(b->infos.begin() + P.w0, b->infos.begin() + P.w1);
    world_order(part, ord);
    for (uint32_t i = 0; i < nw; ++i) ord[i] += P.w0;
    HIPCHK(hipMemcpyAsync((void*)(b->base.order + P.w0), ord, sizeof(uint32_t) * nw, hipMemcpyHostToDevice,
                          stream));
    P.table_dirty = false;
  }
  StepRecord rec;
  rec.chunk = chunk;
  rec.admit = take_event(s);
  rec.done = take_event(s);
  if (!rec.admit || !rec.done) return BB_ERR_DEVICE;
  HIPCHK(hipEventRecord(rec.admit, stream));
  if (!s->started) {
    HIPCHK(hipEventRecord(s->first, stream));
    s->started = true;
  }
  // ---- the step: every slot of the pool, one launch sequence
  st = ensure_graphs(s, p);
  if (st != BB_OK) return st;
  const int cur = P.latest ^ 1;
  HIPCHK(hipGraphLaunch(P.graph[cur], stream));
  P.latest = cur;
  // ---- per-world bookkeeping; read-back and release of finished worlds
  for (uint32_t w = P.w0; w < P.w1; ++w) {
    Slot& sl = s->slots[w];
    if (!sl.used) continue;
    if (sl.fresh) ++rec.admitted;
    else ++rec.carried;
    sl.fresh = false;
    ++sl.done;
    rec.bodies += sl.n;
  }
  const Rows& R = b->rows[cur];
  for (uint32_t w : retiring) {
    Slot& sl = s->slots[w];
    Readback r{};
    r.ticket = sl.ticket;
    r.n = sl.n;
    r.step_index = sl.steps - 1;
    const size_t n = sl.n, g = sl.body_off;
    r.id = (uint32_t*)ch.take(n * 4);
    r.flags = (uint32_t*)ch.take(n * 4);
    r.islot = (uint32_t*)ch.take(n * 4);
    r.pos = (float4*)ch.take(n * 16);
    r.rot = (float4*)ch.take(n * 16);
    r.vel = (float4*)ch.take(n * 16);
    r.omg = (float4*)ch.take(n * 16);
    r.stats = (uint32_t*)ch.take(ST_COUNT * 4);
    HIPCHK(hipMemcpyAsync(r.id, R.id + g, n * 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(r.flags, R.flags + g, n * 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(r.islot, b->base.islot + g, n * 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(r.pos, R.pos + g, n * 16, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(r.rot, R.rot + g, n * 16, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(r.vel, R.vel + g, n * 16, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(r.omg, R.omg + g, n * 16, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(r.stats, b->base.stats + (size_t)w * ST_COUNT, ST_COUNT * 4, hipMemcpyDeviceToHost,
                          stream));
    rec.readbacks.push_back(r);
    // the slot becomes empty for the next step; its ranges are free now
    // (later uploads into them are ordered after this read-back)
    b->infos[w] = WorldInfo{};
    P.table_dirty = true;
    s->rb.release(sl.body_off, sl.n);
    s->ra.release(sl.act_off, sl.a);
    s->rp.release(sl.pair_off, sl.pair_cap);
    s->bytes_in_use -= sl.bytes;
    s->resident_bodies -= sl.n;
    s->running.erase(sl.ticket);
    --P.running;
    b->worlds[w] = HostWorld{};
    sl = Slot{};
    P.free_slots.push_back(w);
    ++s->retired_n;
  }
  HIPCHK(hipEventRecord(rec.done, stream));
  if (!retiring.empty() && s->pools.size() > 1) s->released.emplace_back(p, rec.done);
  P.world_steps += rec.carried + rec.admitted;
  s->world_steps += rec.carried + rec.admitted;
  s->body_steps += rec.bodies;
  P.inflight.push_back(std::move(rec));
  ++P.step_counter;
  ++s->step_counter;
  return BB_OK;
}

bool pool_busy(const Pool& P) { return !P.queue.empty() || P.running > 0; }

bool idle(const bb_server* s) {
  for (const Pool& P : s->pools)
    if (pool_busy(P)) return false;
  return true;
}

bool any_inflight(const bb_server* s) {
  for (const Pool& P : s->pools)
    if (!P.inflight.empty()) return true;
  return false;
}

// Completes the in-flight steps that have finished, without blocking.
bb_status collect(bb_server* s) {
  for (uint32_t p = 0; p < (uint32_t)s->pools.size(); ++p) {
    Pool& P = s->pools[p];
    while (!P.inflight.empty() && hipEventQuery(P.inflight.front().done) == hipSuccess) {
      bb_status st = finish_oldest(s, p);
      if (st != BB_OK) return st;
    }
  }
  return BB_OK;
}

// One server round: every pool with work and room in its pipeline steps
// once; when no pool has room, waits until one has.
bb_status server_round(bb_server* s) {
  bb_status st = collect(s);
  if (st != BB_OK) return st;
  for (;;) {
    bool issued = false;
    for (uint32_t p = 0; p < (uint32_t)s->pools.size(); ++p) {
      Pool& P = s->pools[p];
      if (!pool_busy(P) || (int)P.inflight.size() >= kAhead) continue;
      st = pool_step(s, p);
      if (st != BB_OK) return st;
      issued = true;
    }
    if (issued) return BB_OK;
    if (s->pools.size() == 1) {
      Pool& P = s->pools[0];
      if (!pool_busy(P)) return BB_OK;
      st = finish_oldest(s, 0);  // one pool: wait for it, as a lockstep loop does
      if (st != BB_OK) return st;
      continue;
    }
    bool any = false;
    for (const Pool& P : s->pools) any = any || (pool_busy(P) && !P.inflight.empty());
    if (!any) return BB_OK;
    std::this_thread::yield();
    st = collect(s);
    if (st != BB_OK) return st;
  }
}

double percentile(std::vector<std::pair<float, uint64_t>> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  uint64_t total = 0;
  for (auto& x : v) total += x.second;
  const double target = q * (double)total;
  uint64_t acc = 0;
  for (auto& x : v) {
    acc += x.second;
    if ((double)acc >= target) return x.first;
  }
  return v.back().first;
}

}  // namespace

extern "C" {

bb_status bb_server_create(const bb_server_desc* desc, bb_server** out) {
  if (!desc || !out || desc->max_worlds == 0 || desc->hbm_budget_bytes == 0) return BB_ERR_INVALID;
  if (desc->block != 0 && desc->block != 128 && desc->block != 256 && desc->block != 512 && desc->block != 1024)
    return BB_ERR_INVALID;
  if (!has_device()) return BB_ERR_NO_DEVICE;
  bb_server* s = new (std::nothrow) bb_server();
  if (!s) return BB_ERR_CAPACITY;
  s->desc = *desc;
  const uint32_t team_bodies = desc->team_bodies == 0 ? 8192u : desc->team_bodies;  // 0xFFFFFFFF: never
  // pools by body count: [0, s), [s, 4s), [4s, team), [team, inf) with
  // s = split (the middle ones only below team, at most BB_SERVER_MAX_POOLS
  // pools); one pool without a split
  std::vector<std::pair<uint32_t, uint32_t>> bounds;
  const uint32_t split = desc->split_bodies;
  if (split == 0) {
    bounds.emplace_back(0u, 0xFFFFFFFFu);
  } else {
    std::vector<uint32_t> cuts{split};
    const bool teams = team_bodies != 0xFFFFFFFFu && split < team_bodies;
    for (uint64_t next = 4ull * split; teams && next < team_bodies && cuts.size() + 2 < BB_SERVER_MAX_POOLS;
         next *= 4)
      cuts.push_back((uint32_t)next);
    if (teams) cuts.push_back(team_bodies);
    uint32_t lo = 0;
    for (uint32_t c : cuts) {
      bounds.emplace_back(lo, c);
      lo = c;
    }
    bounds.emplace_back(lo, 0xFFFFFFFFu);
  }
  s->W = desc->max_worlds * (uint32_t)bounds.size();
  const uint64_t base = device_bytes(1, 1, 1, 1);
  s->cW = device_bytes(2, 1, 1, 1) - base;
  s->cB = device_bytes(1, 2, 1, 1) - base;
  s->cA = device_bytes(1, 1, 2, 1) - base;
  s->cP = device_bytes(1, 1, 1, 2) - base;
  uint64_t nb = desc->max_bodies;
  if (nb == 0) {
    const uint64_t fixed = s->W * (s->cW + 64 * s->cP);
    const uint64_t per = s->cB + s->cA + 8 * s->cP;
    nb = desc->hbm_budget_bytes > fixed ? (desc->hbm_budget_bytes - fixed) / per : 0;
  }
  if (nb == 0 || nb > 0x7FFFFFFFull) {
    delete s;
    return BB_ERR_BUDGET;
  }
  s->NB = (uint32_t)nb;
  s->NA = (uint32_t)nb;
  const uint64_t np = 8 * nb + 64ull * s->W;
  if (np > 0xFFFFFFFFull / 8) {
    delete s;
    return BB_ERR_BUDGET;
  }
  s->NP = (uint32_t)np;
  s->b = new (std::nothrow) bb_batch();
  if (!s->b) {
    delete s;
    return BB_ERR_CAPACITY;
  }
  bb_batch* b = s->b;
  b->block = desc->block ? desc->block : 256;
  b->team_bodies = desc->team_bodies == 0xFFFFFFFFu ? 0u : team_bodies;
  s->pools.resize(bounds.size());
  bb_status st = BB_OK;
  for (uint32_t p = 0; p < (uint32_t)bounds.size() && st == BB_OK; ++p) {
    Pool& P = s->pools[p];
    P.lo = bounds[p].first;
    P.hi = bounds[p].second;
    P.w0 = p * desc->max_worlds;
    P.w1 = P.w0 + desc->max_worlds;
    // the first pool (the only one without a split) runs the descriptor's
    // workgroup size; the larger worlds' pools run 512-thread workgroups
    // that keep 1,024 rows of body state in LDS
    P.block = p == 0 ? b->block : 512u;
    P.lcap = p == 0 ? b->pgs_lds : 1024u;
    P.teams = b->team_bodies != 0 && P.hi > b->team_bodies;
    for (uint32_t w = P.w1; w > P.w0; --w) P.free_slots.push_back(w - 1);
    if (hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking) != hipSuccess) st = BB_ERR_DEVICE;
  }
  b->stream = s->pools[0].stream;
  if (st == BB_OK) st = alloc_arrays(b, b->buf, s->W, s->NB, s->NA, s->NP);
  if (st == BB_OK) {
    b->NB = s->NB;
    b->NA = s->NA;
    b->NP = s->NP;
    b->worlds.resize(s->W);
    b->infos.assign(s->W, WorldInfo{});
    b->base.pgs_lcap = b->pgs_lds;
    // requests may run the per-island solver path; its kernels exit at once
    // for every other world
    b->base.any_island = 1;
    // the k_team launch keeps one shape (half of what can be resident) and
    // the team table in device memory says which workgroups step which world
    b->base.team_grid = b->team_bodies ? std::min(team_capacity() / 2, kMaxTeamBlocks) : 0;
    std::vector<uint32_t> order(s->W);
    for (uint32_t w = 0; w < s->W; ++w) order[w] = w;  // empty slots, in slot order
    if (hipMemcpy(b->base.worlds, b->infos.data(), s->W * sizeof(WorldInfo), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy((void*)b->base.order, order.data(), s->W * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipEventCreate(&s->first) != hipSuccess)
      st = BB_ERR_DEVICE;
    b->latest = 1;
    b->built = true;
    b->dirty = false;
  }
  if (st != BB_OK) {
    bb_server_destroy(s);
    return st;
  }
  s->rb.reset(s->NB);
  s->ra.reset(s->NA);
  s->rp.reset(s->NP);
  s->slots.resize(s->W);
  *out = s;
  return BB_OK;
}

// Voxel shapes for the server's requests: every pool's steps complete first,
// the shapes go to the device pools the batch allocated at creation, and the
// pools' step graphs are made again (now with the voxel narrow phase, and the
// k_team instance that has it).  add: one shape from a descriptor; the
// checkpoint path adds a blob's section.
static bb_status shapes_changed(bb_server* s) {
  for (uint32_t p = 0; p < (uint32_t)s->pools.size(); ++p)
    while (!s->pools[p].inflight.empty()) {
      bb_status st = finish_oldest(s, p);
      if (st != BB_OK) return st;
    }
  for (Pool& P : s->pools) {
    HIPCHK(hipStreamSynchronize(P.stream));
    for (hipGraphExec_t& g : P.graph)
      if (g) {
        HIPCHK(hipGraphExecDestroy(g));
        g = nullptr;
      }
  }
  bb_batch* b = s->b;
  bb_status st = upload_voxels(b);
  if (st != BB_OK) return st;
  b->base.any_voxel = 1;
  return BB_OK;
}

static bb_status adopt_shapes(bb_server* s, const char* in, const char* end, uint32_t nshapes, uint32_t shape_words) {
  if (!s->b->vox_shapes.empty() || !take_shapes(s->b, in, end, nshapes, shape_words)) return BB_ERR_INVALID;
  return shapes_changed(s);
}

bb_status bb_server_add_voxel_shape(bb_server* s, const bb_voxel_shape_desc* desc, bb_voxel_shape_info* out) {
  if (!s) return BB_ERR_INVALID;
  bb_status st = add_voxel_shape(s->b, desc, out);
  if (st != BB_OK) return st;
  return shapes_changed(s);
}

bb_status bb_server_submit(bb_server* s, const bb_world_desc* desc, const bb_body_desc* bodies, uint32_t n,
                           uint32_t steps, uint64_t* ticket) {
  if (!s || !valid_desc(desc) || n == 0 || !bodies || steps == 0) return BB_ERR_INVALID;
  for (uint32_t i = 0; i < n; ++i)
    if (!valid_body_with_voxels(bodies[i], s->b->vox_shapes.size())) return BB_ERR_INVALID;
  if (over_reference_cap(*desc, n)) return BB_ERR_CAPACITY;
  uint32_t a = 0;
  for (uint32_t i = 0; i < n; ++i)
    if (bodies[i].flags & BB_BODY_DYNAMIC) ++a;
  Request q;
  q.bodies.assign(bodies, bodies + n);
  const uint32_t pf = pair_factor(q.bodies);
  q.bytes = price(s, n, a, pf);
  if (q.bytes > s->desc.hbm_budget_bytes || n > s->NB || (uint64_t)pair_capacity(n) * pf > s->NP)
    return BB_ERR_BUDGET;
  q.ticket = s->next_ticket++;
  q.desc = *desc;
  q.steps = steps;
  q.active = a;
  q.submit_ms = now_ms();
  Pool& P = s->pools[pool_of(s, n)];
  P.queue.push_back(std::move(q));
  ++s->submitted;
  if (ticket) *ticket = P.queue.back().ticket;
  return BB_OK;
}

bb_status bb_server_step(bb_server* s, uint32_t nsteps) {
  if (!s) return BB_ERR_INVALID;
  for (uint32_t k = 0; k < nsteps; ++k) {
    bb_status st = server_round(s);
    if (st != BB_OK) return st;
  }
  return BB_OK;
}

bb_status bb_server_drain(bb_server* s) {
  if (!s) return BB_ERR_INVALID;
  // a read-back that completes may put its request back in a queue (a rerun
  // after an overflow), so drain until nothing is left anywhere
  while (!idle(s) || any_inflight(s)) {
    for (const Pool& P : s->pools)
      if (P.running == 0 && !P.queue.empty() && P.inflight.empty() && s->running.empty()) {
        // nothing resident yet the head of the queue does not fit: cannot
        // happen for a request accepted by submit, but never spin on it
        if (P.queue.front().bytes > s->desc.hbm_budget_bytes) return BB_ERR_BUDGET;
      }
    if (!idle(s)) {
      bb_status st = server_round(s);
      if (st != BB_OK) return st;
      continue;
    }
    for (uint32_t p = 0; p < (uint32_t)s->pools.size(); ++p)
      while (!s->pools[p].inflight.empty()) {
        bb_status st = finish_oldest(s, p);
        if (st != BB_OK) return st;
      }
  }
  return BB_OK;
}

bb_status bb_server_poll(bb_server* s, uint64_t ticket, uint32_t* state, bb_body_state* out, uint32_t n,
                         bb_step_stats* stats) {
  if (!s || !state) return BB_ERR_INVALID;
  // results of completed steps become visible without blocking
  bb_status cst = collect(s);
  if (cst != BB_OK) return cst;
  auto it = s->results.find(ticket);
  if (it == s->results.end()) {
    if (s->running.count(ticket)) {
      *state = BB_REQ_RUNNING;
      return BB_OK;
    }
    for (const Pool& P : s->pools) {
      for (const StepRecord& rec : P.inflight)
        for (const Readback& r : rec.readbacks)
          if (r.ticket == ticket) {
            *state = BB_REQ_RUNNING;
            return BB_OK;
          }
      for (const Request& q : P.queue)
        if (q.ticket == ticket) {
          *state = BB_REQ_QUEUED;
          return BB_OK;
        }
    }
    *state = BB_REQ_UNKNOWN;
    return BB_OK;
  }
  if (out && n < it->second.bodies.size()) return BB_ERR_INVALID;
  if (out) std::memcpy(out, it->second.bodies.data(), it->second.bodies.size() * sizeof(bb_body_state));
  if (stats) *stats = it->second.stats;
  *state = it->second.failed ? BB_REQ_FAILED : BB_REQ_DONE;
  s->results.erase(it);
  return BB_OK;
}

// ---- checkpoint of the serving loop.  Blob: a header, then one record per
// world (resident worlds first, in slot order, then the queues, pool by pool,
// in queue order), then one per result not yet polled.  A resident world's
// record carries its request (descriptor, body descriptors, steps, reruns)
// and its state arrays; a queued world's carries its request only.
struct CkptHeader {
  char magic[8];
  uint32_t version, worlds, results, nshapes;
  uint64_t next_ticket;
  uint32_t shape_words, pad;  // the voxel shapes' section follows the header
};
struct CkptWorld {
  uint64_t ticket;
  bb_world_desc desc;
  uint32_t n, active, steps, done;
  uint32_t pair_scale, no_team, reruns, npairs;
  uint32_t resident, pad;
};
struct CkptResult {
  uint64_t ticket;
  uint32_t n, failed;
  bb_step_stats stats;
};
constexpr char kCkptMagic[8] = {'B', 'B', 'S', 'R', 'V', 'C', 'K', '2'};
constexpr uint32_t kCkptVersion = 2;
// pair-capacity doublings a checkpointed request may carry
constexpr uint32_t kMaxPairScale = 1u << kMaxReruns;

bb_status bb_server_checkpoint(bb_server* s, void* dst, size_t cap, size_t* size) {
  if (!s || !size) return BB_ERR_INVALID;
  // the steps in flight complete first: their read-backs become results, or
  // go back to a queue to run again
  for (uint32_t p = 0; p < (uint32_t)s->pools.size(); ++p)
    while (!s->pools[p].inflight.empty()) {
      bb_status st = finish_oldest(s, p);
      if (st != BB_OK) return st;
    }
  for (const Pool& P : s->pools) HIPCHK(hipStreamSynchronize(P.stream));
  bb_batch* b = s->b;
  struct Res {
    uint32_t w, npairs;
    int parity;  // the newest parity of the world's pool
    const Request* q;
  };
  std::vector<Res> res;
  for (uint32_t p = 0; p < (uint32_t)s->pools.size(); ++p) {
    const Pool& P = s->pools[p];
    for (uint32_t w = P.w0; w < P.w1; ++w) {
      const Slot& sl = s->slots[w];
      if (!sl.used) continue;
      auto it = s->origin.find(sl.ticket);
      if (it == s->origin.end()) return BB_ERR_INVALID;
      uint32_t np = 0;
      HIPCHK(hipMemcpy(&np, b->npairs[P.latest] + w, 4, hipMemcpyDeviceToHost));
      res.push_back(Res{w, std::min(np, sl.pair_cap), P.latest, &it->second});
    }
  }
  size_t bytes = sizeof(CkptHeader) + shapes_bytes(b);
  uint32_t nworlds = (uint32_t)res.size();
  for (const Res& r : res)
    bytes += sizeof(CkptWorld) + r.q->bodies.size() * sizeof(bb_body_desc) +
             world_state_bytes(s->slots[r.w].n, r.npairs);
  for (const Pool& P : s->pools)
    for (const Request& q : P.queue) {
      bytes += sizeof(CkptWorld) + q.bodies.size() * sizeof(bb_body_desc) + q.state.size();
      ++nworlds;
    }
  for (const auto& kv : s->results) bytes += sizeof(CkptResult) + kv.second.bodies.size() * sizeof(bb_body_state);
  *size = bytes;
  if (!dst) return BB_OK;
  if (cap < bytes) return BB_ERR_CAPACITY;
  char* o = (char*)dst;
  auto put = [&](const void* src, size_t n) {
    if (n) std::memcpy(o, src, n);
    o += n;
  };
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, 8);
  h.version = kCkptVersion;
  h.worlds = nworlds;
  h.results = (uint32_t)s->results.size();
  h.next_ticket = s->next_ticket;
  h.nshapes = (uint32_t)b->vox_shapes.size();
  h.shape_words = (uint32_t)b->vox_occ.size();
  put(&h, sizeof(h));
  o = put_shapes(b, o);
  auto record = [&](const Request& q, uint32_t done, uint32_t npairs, bool resident) {
    CkptWorld r{};
    r.ticket = q.ticket;
    r.desc = q.desc;
    r.n = (uint32_t)q.bodies.size();
    r.active = q.active;
    r.steps = q.steps;
    r.done = done;
    r.pair_scale = q.pair_scale;
    r.no_team = q.no_team ? 1u : 0u;
    r.reruns = q.reruns;
    r.npairs = npairs;
    r.resident = resident ? 1u : 0u;
    put(&r, sizeof(r));
    put(q.bodies.data(), q.bodies.size() * sizeof(bb_body_desc));
  };
  for (const Res& r : res) {
    const Slot& sl = s->slots[r.w];
    record(*r.q, sl.done, r.npairs, true);
    hipError_t rc = hipSuccess;
    world_state_arrays(b, r.w, r.parity, sl.n, r.npairs, [&](void* src, size_t n) {
      if (rc == hipSuccess && n) rc = hipMemcpy(o, src, n, hipMemcpyDeviceToHost);
      o += n;
    });
    HIPCHK(rc);
  }
  for (const Pool& P : s->pools)
    for (const Request& q : P.queue) {
      // a request queued with a checkpointed state (restored, not yet
      // admitted) keeps it
      record(q, q.done, q.npairs, !q.state.empty());
      put(q.state.data(), q.state.size());
    }
  for (const auto& kv : s->results) {
    CkptResult r{};
    r.ticket = kv.first;
    r.n = (uint32_t)kv.second.bodies.size();
    r.failed = kv.second.failed ? 1u : 0u;
    r.stats = kv.second.stats;
    put(&r, sizeof(r));
    put(kv.second.bodies.data(), kv.second.bodies.size() * sizeof(bb_body_state));
  }
  return BB_OK;
}

bb_status bb_server_restore(const bb_server_desc* desc, const void* src, size_t size, bb_server** out) {
  if (!desc || !src || !out || size < sizeof(CkptHeader)) return BB_ERR_INVALID;
  *out = nullptr;
  const char* in = (const char*)src;
  const char* end = in + size;
  CkptHeader h;
  std::memcpy(&h, in, sizeof(h));
  in += sizeof(h);
  if (std::memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != kCkptVersion) return BB_ERR_INVALID;
  // the whole blob is checked on the host before a server is made
  const char* shapes = in;
  const size_t shapes_len = shapes_section_bytes(in, end, h.nshapes, h.shape_words);
  if (h.nshapes && !shapes_len) return BB_ERR_INVALID;
  in += shapes_len;
  struct Rec {
    CkptWorld r;
    const char* bodies;
    const char* state;
  };
  std::vector<Rec> recs;
  uint64_t max_ticket = 0;
  auto take = [&](size_t n) -> const char* {
    if ((size_t)(end - in) < n) return nullptr;
    const char* p = in;
    in += n;
    return p;
  };
  for (uint32_t k = 0; k < h.worlds; ++k) {
    const char* p = take(sizeof(CkptWorld));
    if (!p) return BB_ERR_INVALID;
    Rec rc{};
    std::memcpy(&rc.r, p, sizeof(CkptWorld));
    const CkptWorld& r = rc.r;
    if (!valid_desc(&r.desc) || r.n == 0 || r.steps == 0 || r.done >= r.steps || r.pair_scale == 0 ||
        r.pair_scale > kMaxPairScale || r.reruns > kMaxReruns || r.ticket == 0 || r.resident > 1 ||
        (!r.resident && (r.done || r.npairs)) || over_reference_cap(r.desc, r.n))
      return BB_ERR_INVALID;
    rc.bodies = take((size_t)r.n * sizeof(bb_body_desc));
    if (!rc.bodies) return BB_ERR_INVALID;
    uint32_t a = 0, pf = 1;
    for (uint32_t i = 0; i < r.n; ++i) {
      bb_body_desc bd;
      std::memcpy(&bd, rc.bodies + (size_t)i * sizeof(bb_body_desc), sizeof(bd));
      if (!valid_body_with_voxels(bd, h.nshapes)) return BB_ERR_INVALID;
      if (bd.flags & BB_BODY_DYNAMIC) ++a;
      if (BB_SHAPE_KIND(bd.shape) == BB_SHAPE_VOXEL) pf = 6;
    }
    if (a != r.active) return BB_ERR_INVALID;
    if ((uint64_t)r.npairs > (uint64_t)pair_capacity(r.n) * r.pair_scale * pf) return BB_ERR_INVALID;
    if (r.resident) {
      rc.state = take(world_state_bytes(r.n, r.npairs));
      if (!rc.state || !valid_blob_indices(rc.state, r.n, r.npairs)) return BB_ERR_INVALID;
    }
    max_ticket = std::max(max_ticket, r.ticket);
    recs.push_back(rc);
  }
  struct ResRec {
    CkptResult r;
    const char* bodies;
  };
  std::vector<ResRec> results;
  for (uint32_t k = 0; k < h.results; ++k) {
    const char* p = take(sizeof(CkptResult));
    if (!p) return BB_ERR_INVALID;
    ResRec rr{};
    std::memcpy(&rr.r, p, sizeof(CkptResult));
    if (rr.r.n > (1u << 30) || rr.r.failed > 1) return BB_ERR_INVALID;
    rr.bodies = take((size_t)rr.r.n * sizeof(bb_body_state));
    if (!rr.bodies) return BB_ERR_INVALID;
    max_ticket = std::max(max_ticket, rr.r.ticket);
    results.push_back(rr);
  }
  if (in != end) return BB_ERR_INVALID;
  bb_server* s = nullptr;
  bb_status st = bb_server_create(desc, &s);
  if (st != BB_OK) return st;
  if (h.nshapes) {
    st = adopt_shapes(s, shapes, shapes + shapes_len, h.nshapes, h.shape_words);
    if (st != BB_OK) {
      bb_server_destroy(s);
      return st;
    }
  }
  for (const Rec& rc : recs) {
    const CkptWorld& r = rc.r;
    Request q;
    q.ticket = r.ticket;
    q.desc = r.desc;
    q.bodies.resize(r.n);
    std::memcpy(q.bodies.data(), rc.bodies, (size_t)r.n * sizeof(bb_body_desc));
    q.steps = r.steps;
    q.active = r.active;
    q.pair_scale = r.pair_scale;
    q.no_team = r.no_team != 0;
    q.reruns = r.reruns;
    const uint32_t pf = pair_factor(q.bodies);
    q.bytes = price(s, r.n, r.active, r.pair_scale * pf);
    q.submit_ms = now_ms();
    if (q.bytes > s->desc.hbm_budget_bytes || r.n > s->NB ||
        (uint64_t)pair_capacity(r.n) * r.pair_scale * pf > s->NP) {
      bb_server_destroy(s);
      return BB_ERR_BUDGET;
    }
    if (r.resident) {
      q.state.assign(rc.state, rc.state + world_state_bytes(r.n, r.npairs));
      q.npairs = r.npairs;
      q.done = r.done;
    }
    s->pools[pool_of(s, r.n)].queue.push_back(std::move(q));
    ++s->submitted;
  }
  for (const ResRec& rr : results) {
    Result& res = s->results[rr.r.ticket];
    res.bodies.resize(rr.r.n);
    std::memcpy(res.bodies.data(), rr.bodies, (size_t)rr.r.n * sizeof(bb_body_state));
    res.stats = rr.r.stats;
    res.failed = rr.r.failed != 0;
  }
  s->next_ticket = std::max<uint64_t>(h.next_ticket, max_ticket + 1);
  *out = s;
  return BB_OK;
}

bb_status bb_server_report_get(bb_server* s, bb_server_report* out) {
  if (!s || !out) return BB_ERR_INVALID;
  std::memset(out, 0, sizeof(*out));
  out->submitted = s->submitted;
  out->admitted = s->admitted_n;
  out->retired = s->retired_n;
  for (const Pool& P : s->pools) out->queued += P.queue.size();
  out->resident_worlds = s->running.size();
  out->resident_bodies = s->resident_bodies;
  out->bytes_in_use = s->bytes_in_use;
  out->bytes_budget = s->desc.hbm_budget_bytes;
  out->bytes_allocated = s->b ? s->b->buf.bytes : 0;
  out->server_steps = s->step_counter;
  out->world_steps = s->world_steps;
  out->body_steps = s->body_steps;
  out->device_ms = s->device_ms;
  std::vector<std::pair<float, uint64_t>> all;
  for (const Pool& P : s->pools) all.insert(all.end(), P.step_samples.begin(), P.step_samples.end());
  out->step_ms_p50 = percentile(all, 0.50);
  out->step_ms_p99 = percentile(all, 0.99);
  std::vector<std::pair<float, uint64_t>> qs;
  qs.reserve(s->queue_samples.size());
  for (double v : s->queue_samples) qs.emplace_back((float)v, 1);
  out->queue_ms_p50 = percentile(qs, 0.50);
  out->queue_ms_p99 = percentile(qs, 0.99);
  out->reruns = s->reruns;
  out->pools = (uint32_t)std::min<size_t>(s->pools.size(), BB_SERVER_MAX_POOLS);
  for (uint32_t p = 0; p < out->pools; ++p) {
    const Pool& P = s->pools[p];
    out->pool_min_bodies[p] = P.lo;
    out->pool_world_steps[p] = P.world_steps;
    out->pool_step_ms_p50[p] = percentile(P.step_samples, 0.50);
    out->pool_step_ms_p99[p] = percentile(P.step_samples, 0.99);
  }
  return BB_OK;
}

void bb_server_destroy(bb_server* s) {
  if (!s) return;
  for (Pool& P : s->pools)
    if (P.stream) (void)hipStreamSynchronize(P.stream);
  for (Pool& P : s->pools) {
    for (auto& g : P.graph)
      if (g) (void)hipGraphExecDestroy(g);
    for (auto& rec : P.inflight) {
      if (rec.done) (void)hipEventDestroy(rec.done);
      if (rec.admit) (void)hipEventDestroy(rec.admit);
    }
    if (P.have_prev_done && P.last_done) (void)hipEventDestroy(P.last_done);
    for (auto& c : P.chunks)
      if (c.mem) (void)hipHostFree(c.mem);
    if (P.stream) (void)hipStreamDestroy(P.stream);
  }
  if (s->b) {
    s->b->buf.release();
    s->b->stream = nullptr;  // one of the pools' streams, destroyed above
    delete s->b;
  }
  for (hipEvent_t e : s->spare_events) (void)hipEventDestroy(e);
  if (s->first) (void)hipEventDestroy(s->first);
  delete s;
}

}  // extern "C"
