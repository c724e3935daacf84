// bb_server.cpp — serving loop over a capacity-allocated batch.
//
// The device arrays are allocated once for `max_worlds` slots and the body,
// active and pair capacities that the HBM budget holds.  A world occupies one
// slot and one contiguous range in each arena (first fit); the kernels find
// it through its slot's WorldInfo, so admitting or retiring a world only
// rewrites its table entry and its ranges.  Empty slots have n = 0 and their
// workgroups return at once, so the step launch (one hipGraph per parity over
// all slots) never changes shape.
//
// Host/device pipelining: per server step the host stages new worlds' images
// and table entries in a pinned chunk, enqueues their uploads, the step and
// the read-back of retiring worlds on one stream, and records an event.  A
// chunk is reused only after its step's event has completed, and at most
// kAhead steps are in flight.  A retired world's ranges are handed out again
// at once: the new owner's uploads are enqueued after the read-back, so
// stream order protects the old contents.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <unordered_map>
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

}  // namespace

struct bb_server {
  bb_batch* b = nullptr;
  bb_server_desc desc{};
  uint32_t W = 0, NB = 0, NA = 0, NP = 0;
  uint64_t cW = 0, cB = 0, cA = 0, cP = 0;  // device bytes per slot/body/active/pair slot
  RangeAlloc rb, ra, rp;
  std::vector<Slot> slots;
  std::vector<uint32_t> free_slots;
  std::deque<Request> queue;
  std::unordered_map<uint64_t, Result> results;
  std::unordered_map<uint64_t, uint32_t> running;  // ticket -> slot
  // requests on the device or in a read-back, kept to run again on overflow
  std::unordered_map<uint64_t, Request> origin;
  uint64_t reruns = 0;
  std::deque<StepRecord> inflight;
  std::vector<hipEvent_t> spare_events;
  Chunk chunks[kChunks];
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  uint64_t next_ticket = 1, step_counter = 0;
  uint64_t bytes_in_use = 0, resident_bodies = 0;
  uint64_t submitted = 0, admitted_n = 0, retired_n = 0;
  uint64_t world_steps = 0, body_steps = 0;
  bool table_dirty = false;  // a retirement changed the world table
  hipEvent_t first = nullptr, last_done = nullptr;
  bool have_prev_done = false;
  std::vector<std::pair<float, uint64_t>> step_samples;  // (ms, world-steps)
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

uint64_t price(const bb_server* s, uint32_t n, uint32_t a, uint32_t pair_scale = 1) {
  return s->cW + s->cB * n + s->cA * a + s->cP * (uint64_t)pair_capacity(n) * pair_scale;
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

// Completes the oldest in-flight step: latency samples, results.
bb_status finish_oldest(bb_server* s) {
  StepRecord& rec = s->inflight.front();
  HIPCHK(hipEventSynchronize(rec.done));
  float ms = 0.0f;
  if (s->have_prev_done) {
    HIPCHK(hipEventElapsedTime(&ms, s->last_done, rec.done));
    if (rec.carried) s->step_samples.emplace_back(ms, rec.carried);
  }
  if (rec.admitted) {
    float am = 0.0f;
    HIPCHK(hipEventElapsedTime(&am, rec.admit, rec.done));
    s->step_samples.emplace_back(am, rec.admitted);
  }
  float total = 0.0f;
  HIPCHK(hipEventElapsedTime(&total, s->first, rec.done));
  s->device_ms = total;
  for (const Readback& r : rec.readbacks) {
    auto oit = s->origin.find(r.ticket);
    const uint32_t ovf = r.stats[ST_OVF_STICKY];
    if (ovf && oit != s->origin.end() && oit->second.reruns < kMaxReruns) {
      // a step of this world dropped pairs or was left unfinished by its
      // team: run the request again from its first step, ahead of the queue,
      // if it fits the budget and the pair arena that way
      const Request& o = oit->second;
      const uint32_t scale = (ovf & 1u) ? o.pair_scale * 2u : o.pair_scale;
      const uint32_t nb = (uint32_t)o.bodies.size();
      const uint64_t bytes = price(s, nb, o.active, scale);
      if (bytes <= s->desc.hbm_budget_bytes && (uint64_t)pair_capacity(nb) * scale <= s->NP) {
        Request q = std::move(oit->second);
        s->origin.erase(oit);
        q.pair_scale = scale;
        if (ovf & 2u) q.no_team = true;
        ++q.reruns;
        ++s->reruns;
        q.bytes = bytes;
        s->queue.push_front(std::move(q));
        continue;
      }
    }
    Result& res = s->results[r.ticket];
    to_states(r, res);
    res.failed = ovf != 0;
    if (oit != s->origin.end()) s->origin.erase(oit);
  }
  if (s->have_prev_done) s->spare_events.push_back(s->last_done);
  s->last_done = rec.done;
  s->have_prev_done = true;
  s->spare_events.push_back(rec.admit);
  s->chunks[rec.chunk].used = 0;
  s->inflight.pop_front();
  return BB_OK;
}

bb_status ensure_graphs(bb_server* s) {
  for (int p = 0; p < 2; ++p) {
    if (s->graph[p]) continue;
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(s->b->stream, hipStreamCaptureModeThreadLocal));
    const Dev d = dev_for(s->b, p);
    launch_step(Lane{d, s->b->block, d.W}, s->b->stream, true);
    HIPCHK(hipStreamEndCapture(s->b->stream, &g));
    HIPCHK(hipGraphInstantiate(&s->graph[p], g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
  }
  return BB_OK;
}

bb_status server_step(bb_server* s) {
  bb_batch* b = s->b;
  while ((int)s->inflight.size() >= kAhead) {
    bb_status st = finish_oldest(s);
    if (st != BB_OK) return st;
  }
  const int chunk = (int)(s->step_counter % kChunks);
  Chunk& ch = s->chunks[chunk];
  // ---- admission: first come first served while the budget, the ranges,
  // a slot and the resident-body cap allow
  std::vector<uint32_t> admitted;
  size_t stage_bytes = 0;
  const double t_now = now_ms();
  while (!s->queue.empty() && !s->free_slots.empty()) {
    Request& q = s->queue.front();
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
    const uint32_t pcap = pair_capacity(n) * q.pair_scale;
    if ((uint64_t)pcap > s->NP || !s->rp.alloc(pcap, &po)) {
      s->rb.release(bo, n);
      s->ra.release(ao, q.active);
      break;
    }
    const uint32_t w = s->free_slots.back();
    s->free_slots.pop_back();
    Slot& sl = s->slots[w];
    sl = Slot{};
    sl.used = true;
    sl.fresh = true;
    sl.ticket = q.ticket;
    sl.steps = q.steps;
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
    if (!q.reruns) {
      s->queue_samples.push_back(t_now - q.submit_ms);
      ++s->admitted_n;
    }
    const uint64_t tk = q.ticket;
    stage_bytes += kImageBodyBytes * std::max(n, 1u) + kImageActiveBytes * std::max(q.active, 1u) + 64;
    admitted.push_back(w);
    s->origin[tk] = std::move(q);
    s->queue.pop_front();
  }
  // ---- read-back sizes of the worlds that finish with this step
  std::vector<uint32_t> retiring;
  for (uint32_t w = 0; w < s->W; ++w) {
    const Slot& sl = s->slots[w];
    if (sl.used && sl.done + 1 == sl.steps) {
      retiring.push_back(w);
      stage_bytes += (size_t)sl.n * (4 * 3 + 16 * 4) + ST_COUNT * 4 + 128;
    }
  }
  stage_bytes += (sizeof(WorldInfo) + sizeof(uint32_t)) * (size_t)s->W + sizeof(TeamTable) + 96;
  bb_status st = ch.reserve(stage_bytes + 4096);
  if (st != BB_OK) return st;
  ch.used = 0;
  // ---- uploads of the admitted worlds and their table entries
  hipStream_t stream = b->stream;
  for (uint32_t w : admitted) {
    HostWorld& hw = b->worlds[w];
    const uint32_t n = hw.n, a = hw.a;
    void* mem = ch.take(kImageBodyBytes * std::max(n, 1u) + kImageActiveBytes * std::max(a, 1u));
    const Images img = carve_images(mem, std::max(n, 1u), std::max(a, 1u));
    fill_images(hw, 0, 0, img);
    st = upload_images(b, img, hw.body_off, n, hw.act_off, a, stream);
    if (st != BB_OK) return st;
    st = reset_world_state(b, w, stream);
    if (st != BB_OK) return st;
    WorldInfo wi = info_of(hw);
    wi.mode = 0;  // the cleared pair counts already give an empty previous step
    b->infos[w] = wi;
    hw.bodies.clear();
    hw.bodies.shrink_to_fit();
  }
  // ---- world table: after an admission or a retirement, choose the team
  // worlds again and upload the whole table (and the team table)
  if (!admitted.empty() || s->table_dirty) {
    if (b->base.team_grid) {
      assign_teams(b, b->base.team_grid);
      TeamTable* tt = (TeamTable*)ch.take(sizeof(TeamTable));
      *tt = b->team_table;
      HIPCHK(hipMemcpyAsync(b->base.teams, tt, sizeof(TeamTable), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemsetAsync(b->base.team_ctl, 0, sizeof(TeamCtl) * kMaxTeams, stream));
    }
    WorldInfo* tab = (WorldInfo*)ch.take(sizeof(WorldInfo) * s->W);
    std::memcpy(tab, b->infos.data(), sizeof(WorldInfo) * s->W);
    HIPCHK(hipMemcpyAsync(b->base.worlds, tab, sizeof(WorldInfo) * s->W, hipMemcpyHostToDevice, stream));
    uint32_t* ord = (uint32_t*)ch.take(sizeof(uint32_t) * s->W);
    world_order(b->infos, ord);
    HIPCHK(hipMemcpyAsync((void*)b->base.order, ord, sizeof(uint32_t) * s->W, hipMemcpyHostToDevice, stream));
    s->table_dirty = false;
  }
  StepRecord rec;
  rec.chunk = chunk;
  rec.admit = take_event(s);
  rec.done = take_event(s);
  if (!rec.admit || !rec.done) return BB_ERR_DEVICE;
  HIPCHK(hipEventRecord(rec.admit, stream));
  if (s->step_counter == 0) HIPCHK(hipEventRecord(s->first, stream));
  // ---- the step: every slot, one launch sequence
  st = ensure_graphs(s);
  if (st != BB_OK) return st;
  const int cur = b->latest ^ 1;
  HIPCHK(hipGraphLaunch(s->graph[cur], stream));
  b->latest = cur;
  ++b->step_index;
  // ---- per-world bookkeeping; read-back and release of finished worlds
  for (uint32_t w = 0; w < s->W; ++w) {
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
    s->table_dirty = true;
    s->rb.release(sl.body_off, sl.n);
    s->ra.release(sl.act_off, sl.a);
    s->rp.release(sl.pair_off, sl.pair_cap);
    s->bytes_in_use -= sl.bytes;
    s->resident_bodies -= sl.n;
    s->running.erase(sl.ticket);
    b->worlds[w] = HostWorld{};
    sl = Slot{};
    s->free_slots.push_back(w);
    ++s->retired_n;
  }
  HIPCHK(hipEventRecord(rec.done, stream));
  s->world_steps += rec.carried + rec.admitted;
  s->body_steps += rec.bodies;
  s->inflight.push_back(std::move(rec));
  ++s->step_counter;
  return BB_OK;
}

bool idle(const bb_server* s) { return s->queue.empty() && s->running.empty(); }

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
  if (desc->block != 0 && desc->block != 256 && desc->block != 512 && desc->block != 1024) return BB_ERR_INVALID;
  if (!has_device()) return BB_ERR_NO_DEVICE;
  bb_server* s = new (std::nothrow) bb_server();
  if (!s) return BB_ERR_CAPACITY;
  s->desc = *desc;
  s->W = desc->max_worlds;
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
  if (desc->team_bodies) b->team_bodies = desc->team_bodies == 0xFFFFFFFFu ? 0u : desc->team_bodies;
  bb_status st = BB_OK;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) st = BB_ERR_DEVICE;
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
    world_order(b->infos, order.data());
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
  for (uint32_t w = s->W; w > 0; --w) s->free_slots.push_back(w - 1);
  *out = s;
  return BB_OK;
}

bb_status bb_server_submit(bb_server* s, const bb_world_desc* desc, const bb_body_desc* bodies, uint32_t n,
                           uint32_t steps, uint64_t* ticket) {
  if (!s || !valid_desc(desc) || n == 0 || !bodies || steps == 0) return BB_ERR_INVALID;
  for (uint32_t i = 0; i < n; ++i)
    if (!valid_body(bodies[i])) return BB_ERR_INVALID;
  if (over_reference_cap(*desc, n)) return BB_ERR_CAPACITY;
  uint32_t a = 0;
  for (uint32_t i = 0; i < n; ++i)
    if (bodies[i].flags & BB_BODY_DYNAMIC) ++a;
  Request q;
  q.bytes = price(s, n, a);
  if (q.bytes > s->desc.hbm_budget_bytes || n > s->NB || (uint64_t)pair_capacity(n) > s->NP) return BB_ERR_BUDGET;
  q.ticket = s->next_ticket++;
  q.desc = *desc;
  q.bodies.assign(bodies, bodies + n);
  q.steps = steps;
  q.active = a;
  q.submit_ms = now_ms();
  s->queue.push_back(std::move(q));
  ++s->submitted;
  if (ticket) *ticket = s->queue.back().ticket;
  return BB_OK;
}

bb_status bb_server_step(bb_server* s, uint32_t nsteps) {
  if (!s) return BB_ERR_INVALID;
  for (uint32_t k = 0; k < nsteps; ++k) {
    bb_status st = server_step(s);
    if (st != BB_OK) return st;
  }
  return BB_OK;
}

bb_status bb_server_drain(bb_server* s) {
  if (!s) return BB_ERR_INVALID;
  // a read-back that completes may put its request back in the queue (a
  // rerun after an overflow), so drain until nothing is left anywhere
  while (!idle(s) || !s->inflight.empty()) {
    while (!idle(s)) {
      if (s->running.empty() && !s->queue.empty() && s->free_slots.size() == s->W) {
        // nothing resident yet the head of the queue does not fit: cannot
        // happen for a request accepted by submit, but never spin on it
        const Request& q = s->queue.front();
        if (q.bytes > s->desc.hbm_budget_bytes) return BB_ERR_BUDGET;
      }
      bb_status st = server_step(s);
      if (st != BB_OK) return st;
    }
    while (!s->inflight.empty()) {
      bb_status st = finish_oldest(s);
      if (st != BB_OK) return st;
    }
  }
  return BB_OK;
}

bb_status bb_server_poll(bb_server* s, uint64_t ticket, uint32_t* state, bb_body_state* out, uint32_t n,
                         bb_step_stats* stats) {
  if (!s || !state) return BB_ERR_INVALID;
  // results of completed steps become visible without blocking
  while (!s->inflight.empty() && hipEventQuery(s->inflight.front().done) == hipSuccess) {
    bb_status st = finish_oldest(s);
    if (st != BB_OK) return st;
  }
  auto it = s->results.find(ticket);
  if (it == s->results.end()) {
    if (s->running.count(ticket)) {
      *state = BB_REQ_RUNNING;
      return BB_OK;
    }
    for (const StepRecord& rec : s->inflight)
      for (const Readback& r : rec.readbacks)
        if (r.ticket == ticket) {
          *state = BB_REQ_RUNNING;
          return BB_OK;
        }
    for (const Request& q : s->queue)
      if (q.ticket == ticket) {
        *state = BB_REQ_QUEUED;
        return BB_OK;
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

bb_status bb_server_report_get(bb_server* s, bb_server_report* out) {
  if (!s || !out) return BB_ERR_INVALID;
  std::memset(out, 0, sizeof(*out));
  out->submitted = s->submitted;
  out->admitted = s->admitted_n;
  out->retired = s->retired_n;
  out->queued = s->queue.size();
  out->resident_worlds = s->running.size();
  out->resident_bodies = s->resident_bodies;
  out->bytes_in_use = s->bytes_in_use;
  out->bytes_budget = s->desc.hbm_budget_bytes;
  out->bytes_allocated = s->b ? s->b->buf.bytes : 0;
  out->server_steps = s->step_counter;
  out->world_steps = s->world_steps;
  out->body_steps = s->body_steps;
  out->device_ms = s->device_ms;
  out->step_ms_p50 = percentile(s->step_samples, 0.50);
  out->step_ms_p99 = percentile(s->step_samples, 0.99);
  std::vector<std::pair<float, uint64_t>> qs;
  qs.reserve(s->queue_samples.size());
  for (double v : s->queue_samples) qs.emplace_back((float)v, 1);
  out->queue_ms_p50 = percentile(qs, 0.50);
  out->queue_ms_p99 = percentile(qs, 0.99);
  out->reruns = s->reruns;
  return BB_OK;
}

void bb_server_destroy(bb_server* s) {
  if (!s) return;
  if (s->b) {
    if (s->b->stream) (void)hipStreamSynchronize(s->b->stream);
    for (auto& g : s->graph)
      if (g) (void)hipGraphExecDestroy(g);
    s->b->buf.release();
    if (s->b->stream) (void)hipStreamDestroy(s->b->stream);
    delete s->b;
  }
  for (auto& rec : s->inflight) {
    if (rec.done) (void)hipEventDestroy(rec.done);
    if (rec.admit) (void)hipEventDestroy(rec.admit);
  }
  for (hipEvent_t e : s->spare_events) (void)hipEventDestroy(e);
  if (s->have_prev_done && s->last_done) (void)hipEventDestroy(s->last_done);
  if (s->first) (void)hipEventDestroy(s->first);
  for (auto& c : s->chunks)
    if (c.mem) (void)hipHostFree(c.mem);
  delete s;
}

}  // extern "C"
