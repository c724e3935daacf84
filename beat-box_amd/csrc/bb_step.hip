// bb_step.hip — the rigid-body step on gfx950, one workgroup per world.
//
// The reference step (RigidBodyManager::simulate, src/rigid_body_manager.cpp:
// 1606-1622) issues ~1,350-1,550 barrier-separated dispatches per step.  Here
// each stage is a kernel whose grid has one workgroup per world; inside a
// workgroup the stage's passes are separated by s_barrier instead of kernel
// boundaries, so a batch of thousands of worlds runs every stage in one
// launch and a single world pays a handful of launches per step.
//
// Stages (reference pass order, src/rigid_body_manager.cpp:1284-1427):
//   k_build   Morton keys, stable LSD radix sort, Karras LBVH, bounds, reorder
//   k_broad   per-row BVH traversal, pairs in (a, b) order
//   k_narrow  SAT + clipping (+ sphere closed forms), warm start, union-find
//   k_post    advect, islands, sleeping, edge colouring, colour buckets
//   k_pgs     PGS_SOFT colour sweeps + position integration
//   k_avbd_*  AVBD body colouring, prepare, warm start, primal/dual sweeps
//   k_finish  end-of-step counters
// Everything is deterministic: results do not depend on thread timing, and a
// world's results do not depend on the other worlds of the batch.
#include <hip/hip_runtime.h>

#include "../../include/bb_capi.h"
#include "bb_dev.hpp"
#include "bb_math.hpp"

namespace bb {

#define WAVE 64
// Lanes per body in the AVBD primal sweeps of the per-world kernels (the
// default of BB_PARAM_AVBD_LANES) and of workgroup teams.
static constexpr uint32_t kAvG = 4, kAvGTeam = 8;

// ============================================================ block helpers
template <int BLOCK>
struct ScanLds {
  uint32_t w[BLOCK / WAVE + 1];
};

// Exclusive block scan of one value per thread; returns the prefix and the
// block total.  All threads of the block must call it.
template <int BLOCK>
BB_DI uint32_t block_scan(uint32_t v, ScanLds<BLOCK>& s, uint32_t& total) {
  const uint32_t lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    uint32_t y = __shfl_up(x, o, WAVE);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == WAVE - 1) s.w[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t sum = 0;
    for (int k = 0; k < BLOCK / WAVE; ++k) {
      uint32_t t = s.w[k];
      s.w[k] = sum;
      sum += t;
    }
    s.w[BLOCK / WAVE] = sum;
  }
  __syncthreads();
  uint32_t r = s.w[wave] + x - v;
  total = s.w[BLOCK / WAVE];
  __syncthreads();
  return r;
}

BB_DI void stat_add(const Dev& d, uint32_t w, uint32_t slot, uint32_t v) {
  if (v) atomicAdd(&d.stats[w * ST_COUNT + slot], v);
}
BB_DI void stat_max(const Dev& d, uint32_t w, uint32_t slot, uint32_t v) {
  atomicMax(&d.stats[w * ST_COUNT + slot], v);
}

// ============================================================ body helpers
struct Shape {
  v3 mn, mx;
  float radius;
  uint32_t kind;
};
// pmax.w holds the body's shape word as a float: 0 box, 1 sphere, or
// BB_SHAPE_VOXEL | voxel shape id << 8.
BB_DI uint32_t shape_word(float w) { return (uint32_t)w; }
BB_DI Shape shape_of(const Dev& d, uint32_t gid) {
  float4 a = d.props.pmin[gid], b = d.props.pmax[gid];
  return Shape{ld3(a), ld3(b), a.w, shape_word(b.w) & 0xFFu};
}

// RigidBody::get_AABB_from_OBB (src/shaders/extensions.slang:195-210);
// spheres (extension) use centre +- radius.
BB_DI void body_aabb(const Shape& s, v3 pos, qt q, v3& mn, v3& mx) {
  if (s.kind == BB_SHAPE_SPHERE) {
    v3 r = V(s.radius, s.radius, s.radius);
    mn = pos - r;
    mx = pos + r;
    return;
  }
  mn = V(3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f);
  mx = V(-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f);
#pragma unroll
  for (uint32_t i = 0; i < 8; ++i) {
    v3 c;
    c.x = (i & 1) == 0 ? s.mn.x : s.mx.x;
    c.y = (i & 2) == 0 ? s.mn.y : s.mx.y;
    c.z = (i & 4) == 0 ? s.mn.z : s.mx.z;
    c = qrot(q, c) + pos;
    mn = vmin(mn, c);
    mx = vmax(mx, c);
  }
}

// World inverse inertia (src/shaders/extensions.slang:184-188): with the
// parity quirk Rt*(I*transpose(Rt)) where Rt = get_rotation_matrix() =
// transpose(R); otherwise R*(I*R^T).
BB_DI mat3 world_inv_inertia(const Dev& d, uint32_t gid, qt q, bool quirk) {
  mat3 I = mrows(ld3(d.props.ii0[gid]), ld3(d.props.ii1[gid]), ld3(d.props.ii2[gid]));
  mat3 R = qmat(q);
  if (quirk) {
    mat3 Rt = mT(R);
    return mmul(Rt, mmul(I, mT(Rt)));
  }
  return mmul(R, mmul(I, mT(R)));
}

// ================================================================== Morton
// MortonCodeGenerator (src/shaders/simulation/LBVH.slang:4-32)
BB_DI uint32_t expand10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
BB_DI uint32_t morton30(v3 p) {
  float x = fmn(fmx(p.x * 1024.0f, 0.0f), 1023.0f);
  float y = fmn(fmx(p.y * 1024.0f, 0.0f), 1023.0f);
  float z = fmn(fmx(p.z * 1024.0f, 0.0f), 1023.0f);
  return expand10((uint32_t)x) * 4 + expand10((uint32_t)y) * 2 + expand10((uint32_t)z);
}
// 63-bit extension: 21 bits per axis, x in the highest position.
BB_DI uint64_t expand21(uint64_t x) {
  x &= 0x1fffffull;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}
BB_DI uint64_t morton63(v3 p) {
  float x = fmn(fmx(p.x * 2097152.0f, 0.0f), 2097151.0f);
  float y = fmn(fmx(p.y * 2097152.0f, 0.0f), 2097151.0f);
  float z = fmn(fmx(p.z * 2097152.0f, 0.0f), 2097151.0f);
  return (expand21((uint64_t)x) << 2) | (expand21((uint64_t)y) << 1) | expand21((uint64_t)z);
}

// =========================================================== block sort
// Stable LSD radix sort of (key, index) with 8-bit digits.  Per pass: an LDS
// digit histogram, a block scan of the 256 bins, then a stable scatter in
// chunks of BLOCK elements: wave ballots over the 8 digit bits give each lane
// its peers (same digit) in the wave, per-wave digit counts in LDS give the
// offsets of lower waves.  Passes whose digit is constant are skipped (a
// stable pass over one digit value is the identity).
template <int BLOCK>
struct SortLds {
  uint32_t base[256];
  uint32_t wcnt[(BLOCK / WAVE) * 256];
  ScanLds<BLOCK> scan;
};
// The team sort's scatter takes kSortTile keys per thread per tile (sub-tile
// j holds keys [j*BLOCK, (j+1)*BLOCK) of the tile): per (sub-tile, wave,
// digit) counts, zero between tiles, and the offsets made from them.
static constexpr uint32_t kSortTile = 4;
template <int BLOCK>
struct TeamSortLds {
  SortLds<BLOCK> s;
  uint32_t cnt[kSortTile * (BLOCK / WAVE) * 256];
  uint32_t off[kSortTile * (BLOCK / WAVE) * 256];
};

template <int BLOCK>
BB_DI bool block_radix_sort(uint64_t* ka, uint32_t* ia, uint64_t* kb, uint32_t* ib, uint32_t n, int passes,
                            SortLds<BLOCK>& L) {
  const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
  for (uint32_t i = tid; i < (BLOCK / WAVE) * 256; i += BLOCK) L.wcnt[i] = 0;
  bool in_a = true;
  for (int pass = 0; pass < passes; ++pass) {
    const uint32_t shift = 8u * pass;
    uint64_t* kin = in_a ? ka : kb;
    uint32_t* iin = in_a ? ia : ib;
    uint64_t* kout = in_a ? kb : ka;
    uint32_t* iout = in_a ? ib : ia;
    for (uint32_t i = tid; i < 256; i += BLOCK) L.base[i] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n; i += BLOCK) atomicAdd(&L.base[(uint32_t)(kin[i] >> shift) & 255u], 1u);
    __syncthreads();
    const bool skip = n == 0 || L.base[(uint32_t)(kin[0] >> shift) & 255u] == n;
    __syncthreads();
    if (skip) continue;
    {
      // exclusive scan of the 256 digit counts (256 / BLOCK consecutive
      // digits per thread in a smaller block)
      constexpr uint32_t per = BLOCK >= 256 ? 1u : 256u / BLOCK;
      uint32_t v[per], sum = 0, total;
#pragma unroll
      for (uint32_t q = 0; q < per; ++q) {
        const uint32_t dg = tid * per + q;
        v[q] = dg < 256 ? L.base[dg] : 0u;
        sum += v[q];
      }
      uint32_t ex = block_scan<BLOCK>(sum, L.scan, total);
#pragma unroll
      for (uint32_t q = 0; q < per; ++q) {
        const uint32_t dg = tid * per + q;
        if (dg < 256) L.base[dg] = ex;
        ex += v[q];
      }
      __syncthreads();
    }
    for (uint32_t c0 = 0; c0 < n; c0 += BLOCK) {
      const uint32_t i = c0 + tid;
      const bool valid = i < n;
      uint64_t key = valid ? kin[i] : 0ull;
      uint32_t idx = valid ? iin[i] : 0u;
      uint32_t dg = (uint32_t)(key >> shift) & 255u;
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        uint64_t m = __ballot((dg >> b) & 1u);
        peers &= ((dg >> b) & 1u) ? m : ~m;
      }
      const uint32_t rank = __popcll(peers & ((1ull << lane) - 1ull));
      const uint32_t cnt = __popcll(peers);
      const bool leader = valid && lane == (uint32_t)(__ffsll((unsigned long long)peers) - 1);
      if (leader) L.wcnt[wave * 256 + dg] = cnt;
      __syncthreads();
      if (valid) {
        uint32_t off = L.base[dg] + rank;
        for (uint32_t w2 = 0; w2 < wave; ++w2) off += L.wcnt[w2 * 256 + dg];
        kout[off] = key;
        iout[off] = idx;
      }
      __syncthreads();
      if (leader) {
        atomicAdd(&L.base[dg], cnt);
        L.wcnt[wave * 256 + dg] = 0;
      }
      __syncthreads();
    }
    in_a = !in_a;
  }
  return in_a;  // true: sorted data in (ka, ia)
}

// ================================================================== teams
// A team of threads advances one world through the stage functions s_*.
// BlockTeam is one workgroup (the per-world stage kernels, many worlds per
// launch).  GroupTeam is a contiguous range of workgroups of k_team (large
// worlds); its barrier is an arrival counter plus generation in global
// memory, and its scan combines per-workgroup partials.  Team scalars of a
// stage live in LDS for a BlockTeam and in global memory for a GroupTeam; the
// stage code is the same.
template <int BLOCK>
struct BlockTeam {
  ScanLds<BLOCK>* scan_lds;
  SortLds<BLOCK>* sort_lds;
  // AVBD primal: kAvG lanes per body; empty cascade buckets skip their barrier
  static constexpr uint32_t kAvGroup = kAvG;
  static constexpr bool kSkipEmpty = true;
  static constexpr bool kGlobalScalars = false;  // stage scalars in LDS
  static constexpr int kBlock = BLOCK;
  BB_DI uint32_t rank() const { return threadIdx.x; }
  BB_DI uint32_t size() const { return BLOCK; }
  // this workgroup's index in the team, and the team's workgroups
  BB_DI uint32_t group() const { return 0; }
  BB_DI uint32_t groups() const { return 1; }
  BB_DI void sync() const { __syncthreads(); }
  BB_DI void fence() const { __threadfence_block(); }
  // every leaf range is this workgroup's (see GroupTeam::leaves_local)
  BB_DI bool leaves_local(uint32_t, uint32_t) const { return true; }
  BB_DI bool lead_group() const { return true; }
  BB_DI uint32_t scan(uint32_t v, uint32_t& total) const { return block_scan<BLOCK>(v, *scan_lds, total); }
  BB_DI bool radix_sort(uint64_t* ka, uint32_t* ia, uint64_t* kb, uint32_t* ib, uint32_t n, int passes) const {
    return block_radix_sort<BLOCK>(ka, ia, kb, ib, n, passes, *sort_lds);
  }
  // clock marks inside a stage are taken by workgroup teams only
  BB_DI void mark(uint32_t) const {}
  BB_DI void mark_lead(uint32_t) const {}
  BB_DI bool sorted_in_a(bool in_a) const { return in_a; }
};

// Spins before a team barrier wait gives up (about a second).  A wait that
// runs out sets TeamCtl::error.  Every workgroup that then reaches a team
// barrier, or is still waiting at one, leaves the step at that barrier
// (GroupTeam::abandon): no stage code runs on data the team did not finish,
// so no index computed from it is ever used.  The world's overflow word gets
// bit 1, and the host replays the step.
static constexpr uint32_t kTeamSpinLimit = 1u << 24;

template <int BLOCK>
struct GroupTeam {
  ScanLds<BLOCK>* scan_lds;
  TeamSortLds<BLOCK>* sort_lds;
  TeamCtl* ctl;
  uint32_t* hist;  // [nblocks * 256] this team's radix histograms / offsets
  uint32_t blk, nblocks;
  uint32_t* dead;  // LDS words: thread 0's verdict of the last barrier, then its
                   // barrier counts on the team's and the sub-team's counter
  uint32_t* stats; // the world's counters (ST_COUNT words)
  uint32_t* ovf_any;
  uint32_t spin;   // spins before a wait gives up
  unsigned long long* clk;  // this team's clock marks (profiling launches), or null
  uint32_t sort_blocks;     // the first sort_blocks workgroups run the radix sort
  // AVBD primal: kAvGTeam lanes per body (a pile's bodies have half a dozen
  // manifolds, so eight lanes take about one each); empty cascade buckets
  // skip their barrier (the bucket offsets are team scalars every workgroup
  // reads after the same barrier, so all of them skip the same buckets)
  static constexpr uint32_t kAvGroup = kAvGTeam;
  static constexpr bool kSkipEmpty = true;
  static constexpr bool kGlobalScalars = true;  // stage scalars in a global slot
  static constexpr int kBlock = BLOCK;
  BB_DI uint32_t rank() const { return blk * BLOCK + threadIdx.x; }
  BB_DI uint32_t size() const { return nblocks * BLOCK; }
  BB_DI uint32_t group() const { return blk; }
  BB_DI uint32_t groups() const { return nblocks; }
  BB_DI void fence() const { __threadfence(); }
  // Clock mark m of this step, called right after a team barrier: every
  // workgroup has reached that point when rank 0 stamps it.
  BB_DI void mark(uint32_t m) const {
    if (clk && rank() == 0) clk[m] = wall_clock64();
  }
  // Clock mark m once the lead workgroup (rank 0's) is past a point of its
  // own work (a workgroup barrier, not a team one; profiling launches only).
  BB_DI void mark_lead(uint32_t m) const {
    if (!clk) return;  // workgroup-uniform
    __syncthreads();
    if (rank() == 0) clk[m] = wall_clock64();
  }
  // With leaf i handled by rank i mod size(), a tree node over leaves
  // [first, last] inside one BLOCK-aligned run is touched by one workgroup
  // alone: a bottom-up pass orders its steps there with workgroup scope
  // instead of device-scope fences (an L2 writeback and invalidate each on
  // a multi-XCD chip).
  BB_DI bool leaves_local(uint32_t first, uint32_t last) const { return first / BLOCK == last / BLOCK; }
  // the workgroup that finishes the nodes above the runs
  BB_DI bool lead_group() const { return blk == 0; }
  BB_DI uint32_t err() const { return __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
  // Thread 0: this workgroup has left the step.  The last workgroup of the
  // team to leave resets the control block for the next launch (every other
  // member is past its last use of it).
  BB_DI void leave() const {
    __threadfence();
    if (atomicAdd(&ctl->exited, 1u) == nblocks - 1u) {
      __hip_atomic_store(&ctl->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctl->scount, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctl->error, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctl->exited, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __threadfence();
    }
  }
  // The team did not meet at a barrier: flag the world (overflow bit 1, also
  // in its sticky word and the batch's overflow word) and end every wave of
  // this workgroup here.  Called by all threads, workgroup-uniformly.
  BB_DI void abandon() const {
    if (threadIdx.x == 0) {
      atomicOr(&stats[ST_OVERFLOW], 2u);
      atomicOr(&stats[ST_OVF_STICKY], 2u);
      atomicOr(ovf_any, 1u);
      leave();
    }
    __builtin_amdgcn_endpgm();
  }
  // Barrier of `members` workgroups on the arrival counter *cnt, which only
  // grows during a launch (the last workgroup to leave resets it): this
  // workgroup's e-th barrier on it is met once it reads e * members.  *epoch
  // (LDS, thread 0's) counts this workgroup's barriers on the counter.  A
  // member that finds the team's error word set while it waits, or that waits
  // past the spin limit (and sets the word), leaves the step; one that met the
  // barrier sees a failure at its next one.
  BB_DI void barrier(uint32_t* cnt, uint32_t* epoch, uint32_t members) const {
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t target = (*epoch += 1u) * members;
      uint32_t bad = 0u;
      // release: the workgroup's writes before the arrival (L2 write-back)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      if (atomicAdd(cnt, 1u) + 1u != target) {
        uint32_t spins = 0;
        while ((int32_t)(__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
          if ((bad = err()) != 0u) break;
          if (++spins > spin) {
            atomicOr(&ctl->error, 1u);
            bad = 1u;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      // acquire: the other members' writes before their arrivals (L2 invalidate)
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      dead[0] = bad;
    }
    __syncthreads();
    if (dead[0]) abandon();
  }
  BB_DI void sync() const { barrier(&ctl->count, &dead[1], nblocks); }
  // Barrier of the sort's sub-team, the first `members` workgroups.
  BB_DI void sync_first(uint32_t members) const { barrier(&ctl->scount, &dead[2], members); }
  // Exclusive scan over the whole team (rank order); two team barriers.
  BB_DI uint32_t scan(uint32_t v, uint32_t& total) const { return scan_first(v, total, nblocks); }
  // The same over the first `members` workgroups (rank order), with their
  // barrier (the whole team's when members = nblocks).
  BB_DI uint32_t scan_first(uint32_t v, uint32_t& total, uint32_t members) const {
    auto bar = [&]() {
      if (members == nblocks) sync();
      else sync_first(members);
    };
    uint32_t btot;
    const uint32_t ex = block_scan<BLOCK>(v, *scan_lds, btot);
    if (threadIdx.x == 0) ctl->part[blk] = btot;
    bar();
    if (threadIdx.x < WAVE) {
      uint32_t pre = 0, all = 0;
      for (uint32_t b = threadIdx.x; b < members; b += WAVE) {
        const uint32_t x = ctl->part[b];
        all += x;
        if (b < blk) pre += x;
      }
      for (int o = 32; o > 0; o >>= 1) {
        pre += __shfl_down(pre, o, WAVE);
        all += __shfl_down(all, o, WAVE);
      }
      if (threadIdx.x == 0) {
        scan_lds->w[0] = pre;
        scan_lds->w[1] = all;
      }
    }
    __syncthreads();
    const uint32_t pre = scan_lds->w[0];
    total = scan_lds->w[1];
    bar();
    return pre + ex;
  }
  // Stable LSD radix sort over the sort's sub-team (the first sort_blocks
  // workgroups; the others go on to the barrier after the sort): per pass
  // each member histograms its contiguous chunk, one scan over the
  // (digit, workgroup) counts in digit-major order gives every member its
  // output offset per digit, and each member scatters its chunk with the
  // ballot ranking of the one-workgroup sort.  Passes whose digit is constant
  // are skipped.  Returns, in the members, the side the keys ended on, which
  // sorted_in_a hands to the whole team after the next team barrier.
  BB_DI bool radix_sort(uint64_t* ka, uint32_t* ia, uint64_t* kb, uint32_t* ib, uint32_t n, int passes) const {
    const uint32_t S = sort_blocks < 1u ? 1u : (sort_blocks < nblocks ? sort_blocks : nblocks);
    if (blk >= S) return true;  // not a member: the result comes through ctl->sorted
    SortLds<BLOCK>& L = sort_lds->s;
    TeamSortLds<BLOCK>& X = *sort_lds;
    constexpr uint32_t NW = BLOCK / WAVE, K = kSortTile;
    const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    const uint32_t chunk = (n + S - 1) / S;
    const uint32_t lo = blk * chunk < n ? blk * chunk : n;
    const uint32_t hi = lo + chunk < n ? lo + chunk : n;
    auto bar = [&]() {
      if (S == nblocks) sync();
      else sync_first(S);
    };
    for (uint32_t i = tid; i < K * NW * 256; i += BLOCK) X.cnt[i] = 0;
    bool in_a = true;
    for (int pass = 0; pass < passes; ++pass) {
      const uint32_t shift = 8u * pass;
      uint64_t* kin = in_a ? ka : kb;
      uint32_t* iin = in_a ? ia : ib;
      uint64_t* kout = in_a ? kb : ka;
      uint32_t* iout = in_a ? ib : ia;
      for (uint32_t i = tid; i < 256; i += BLOCK) L.base[i] = 0;
      if (blk == 0 && tid == 0) ctl->diff = 0;
      bar();
      const uint32_t d0 = n ? (uint32_t)(kin[0] >> shift) & 255u : 0u;
      bool differs = false;
      // histogram, K keys per thread in flight
      for (uint32_t c0 = lo; c0 < hi; c0 += K * BLOCK) {
        uint32_t dgs[K];
#pragma unroll
        for (uint32_t j = 0; j < K; ++j) {
          const uint32_t i = c0 + j * BLOCK + tid;
          dgs[j] = i < hi ? (uint32_t)(kin[i] >> shift) & 255u : 256u;
        }
#pragma unroll
        for (uint32_t j = 0; j < K; ++j)
          if (dgs[j] < 256u) {
            atomicAdd(&L.base[dgs[j]], 1u);
            differs |= dgs[j] != d0;
          }
      }
      if (differs) atomicOr(&ctl->diff, 1u);
      __syncthreads();
      for (uint32_t i = tid; i < 256; i += BLOCK) hist[i * S + blk] = L.base[i];
      bar();
      if (ctl->diff == 0u) continue;  // uniform over the members: read after their barrier
      // offsets: exclusive scan of hist in (digit, workgroup) order
      uint32_t carry = 0;
      for (uint32_t e0 = 0; e0 < 256u * S; e0 += S * BLOCK) {
        const uint32_t e = e0 + blk * BLOCK + tid;
        const uint32_t v = e < 256u * S ? hist[e] : 0u;
        uint32_t tot;
        const uint32_t ex = scan_first(v, tot, S);
        if (e < 256u * S) hist[e] = carry + ex;
        carry += tot;
      }
      bar();
      for (uint32_t i = tid; i < 256; i += BLOCK) L.base[i] = hist[i * S + blk];
      __syncthreads();
      // scatter, a tile of K * BLOCK keys at a time: every key's rank among
      // its wave's peers by ballots, the (sub-tile, wave, digit) counts, one
      // walk per digit over them in key order for the offsets, then the
      // writes (two workgroup barriers per tile)
      for (uint32_t c0 = lo; c0 < hi; c0 += K * BLOCK) {
        uint64_t key[K];
        uint32_t idx[K], dg[K], rk[K];
        bool valid[K];
#pragma unroll
        for (uint32_t j = 0; j < K; ++j) {
          const uint32_t i = c0 + j * BLOCK + tid;
          valid[j] = i < hi;
          key[j] = valid[j] ? kin[i] : 0ull;
          idx[j] = valid[j] ? iin[i] : 0u;
          dg[j] = (uint32_t)(key[j] >> shift) & 255u;
        }
#pragma unroll
        for (uint32_t j = 0; j < K; ++j) {
          uint64_t peers = __ballot(valid[j]);
#pragma unroll
          for (int b = 0; b < 8; ++b) {
            const uint64_t m = __ballot((dg[j] >> b) & 1u);
            peers &= ((dg[j] >> b) & 1u) ? m : ~m;
          }
          rk[j] = __popcll(peers & ((1ull << lane) - 1ull));
          if (valid[j] && lane == (uint32_t)(__ffsll((unsigned long long)peers) - 1))
            X.cnt[(j * NW + wave) * 256 + dg[j]] = __popcll(peers);
        }
        __syncthreads();
        for (uint32_t dd = tid; dd < 256; dd += BLOCK) {
          uint32_t run = L.base[dd];
          for (uint32_t s2 = 0; s2 < K * NW; ++s2) {
            const uint32_t c = X.cnt[s2 * 256 + dd];
            X.off[s2 * 256 + dd] = run;
            X.cnt[s2 * 256 + dd] = 0;
            run += c;
          }
          L.base[dd] = run;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < K; ++j)
          if (valid[j]) {
            const uint32_t o = X.off[(j * NW + wave) * 256 + dg[j]] + rk[j];
            if (o < n) {  // always true unless a barrier wait ran out
              kout[o] = key[j];
              iout[o] = idx[j];
            }
          }
      }
      bar();
      in_a = !in_a;
    }
    if (blk == 0 && tid == 0) __hip_atomic_store(&ctl->sorted, in_a ? 1u : 0u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
    return in_a;
  }
  // After the team barrier that follows radix_sort: the side the sorted keys
  // are on, the same in every workgroup.
  BB_DI bool sorted_in_a(bool) const {
    return __hip_atomic_load(&ctl->sorted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
  }
};

// =================================================================== LBVH
struct KeyView {
  const uint64_t* k;
  int n;
  int bits;
};
// LBVH::delta (src/shaders/simulation/LBVH.slang:159-170)
BB_DI int lb_delta(const KeyView& K, int i, uint64_t ci, int j) {
  if (j < 0 || j > K.n - 1) return -1;
  uint64_t cj = K.k[j];
  if (K.bits == 30) {
    if (ci == cj) return 32 + __clz((uint32_t)(i ^ j));
    return __clz((uint32_t)(ci ^ cj));
  }
  if (ci == cj) return 64 + __clz((uint32_t)(i ^ j));
  return __clzll(ci ^ cj);
}
// LBVH::determine_range / find_split (LBVH.slang:172-221)
BB_DI void lb_range(const KeyView& K, int index, int& first, int& last) {
  uint64_t code = K.k[index];
  int dl = lb_delta(K, index, code, index - 1);
  int dr = lb_delta(K, index, code, index + 1);
  int dd = (dr >= dl) ? 1 : -1;
  int dmin = dl < dr ? dl : dr;
  int lmax = 2;
  while (lb_delta(K, index, code, index + lmax * dd) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t > 0; t >>= 1)
    if (lb_delta(K, index, code, index + (l + t) * dd) > dmin) l += t;
  int j = index + l * dd;
  first = index < j ? index : j;
  last = index < j ? j : index;
}
BB_DI int lb_split(const KeyView& K, int first, int last) {
  uint64_t fc = K.k[first];
  int common = lb_delta(K, first, fc, last);
  int split = first, step = last - first;
  do {
    step = (step + 1) >> 1;
    int ns = split + step;
    if (ns < last) {
      if (lb_delta(K, first, fc, ns) > common) split = ns;
    }
  } while (step > 1);
  return split;
}

// ================================================================ k_build
struct BuildShared {
  uint32_t lo[3], hi[3];
  uint32_t nroots;   // bounds pass: nodes handed up from the 512-leaf runs
  uint32_t nroots2;  // and from the spans of kClimbRuns runs (teams)
};
// Leaf runs per span of the bounds pass's middle level (teams): the nodes
// above the runs but inside one span are climbed by one workgroup each, so
// the lead workgroup only climbs what lies above the spans.
static constexpr uint32_t kClimbRuns = 16;

template <class T>
BB_DI void s_build(const Dev& d, uint32_t w, const T& t, BuildShared& L) {
  const uint32_t tid = t.rank(), NT = t.size();
  const WorldInfo W = d.worlds[w];
  const uint32_t N = W.n, off = W.body_off;
  // per-step resets (an empty world slot keeps zero counters)
  if (tid < ST_COUNT && tid != ST_OVF_STICKY) d.stats[w * ST_COUNT + tid] = 0;
  if (N == 0) return;
  for (uint32_t k = tid; k < W.a; k += NT) d.link[W.act_off + k] = k;
  for (uint32_t i = tid; i < N; i += NT) d.deg[off + i] = 0;
  // ---- Morton keys from the previous rows (RB_sim.slang:75-111)
  uint64_t* ka = d.keys_a + off;
  uint64_t* kb = d.keys_b + off;
  uint32_t* ia = d.idx_a + off;
  uint32_t* ib = d.idx_b + off;
  if (W.morton_bits == 30) {
    for (uint32_t i = tid; i < N; i += NT) {
      v3 p = ld3(d.prev.pos[off + i]);
      v3 q = p - V(-64.0f, -64.0f, -64.0f);
      v3 ext = V(128.0f, 128.0f, 128.0f);
      ka[i] = morton30(V(q.x / ext.x, q.y / ext.y, q.z / ext.z));
      ia[i] = i;
    }
  } else {
    if (tid < 3) {
      L.lo[tid] = 0xFFFFFFFFu;
      L.hi[tid] = 0u;
    }
    t.sync();
    // per-thread, then per-wave extremes; one lane per wave merges them (a
    // team's L lives in global memory, where a million same-address atomics
    // would serialise)
    uint32_t mlo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mhi[3] = {0u, 0u, 0u};
    for (uint32_t i = tid; i < N; i += NT) {
      const float4 p = d.prev.pos[off + i];
      const uint32_t m[3] = {omap(p.x), omap(p.y), omap(p.z)};
      for (int k = 0; k < 3; ++k) {
        mlo[k] = m[k] < mlo[k] ? m[k] : mlo[k];
        mhi[k] = m[k] > mhi[k] ? m[k] : mhi[k];
      }
    }
    for (int o = 32; o > 0; o >>= 1)
      for (int k = 0; k < 3; ++k) {
        const uint32_t a = __shfl_down(mlo[k], o, 64), b = __shfl_down(mhi[k], o, 64);
        mlo[k] = a < mlo[k] ? a : mlo[k];
        mhi[k] = b > mhi[k] ? b : mhi[k];
      }
    if ((threadIdx.x & 63u) == 0u)
      for (int k = 0; k < 3; ++k) {
        atomicMin(&L.lo[k], mlo[k]);
        atomicMax(&L.hi[k], mhi[k]);
      }
    t.sync();
    v3 lo = V(ounmap(L.lo[0]), ounmap(L.lo[1]), ounmap(L.lo[2]));
    v3 hi = V(ounmap(L.hi[0]), ounmap(L.hi[1]), ounmap(L.hi[2]));
    v3 ext = hi - lo;
    for (uint32_t i = tid; i < N; i += NT) {
      v3 dd = ld3(d.prev.pos[off + i]) - lo;
      v3 p;
      p.x = ext.x > 0.0f ? dd.x / ext.x : 0.0f;
      p.y = ext.y > 0.0f ? dd.y / ext.y : 0.0f;
      p.z = ext.z > 0.0f ? dd.z / ext.z : 0.0f;
      ka[i] = morton63(p);
      ia[i] = i;
    }
  }
  t.sync();
  t.mark(kMarkKeys);
  // tap copy of the unsorted keys lives in keys_b until the sort overwrites it;
  // the host reads taps from a dedicated copy made by k_finish instead.
  // ---- stable LSD radix sort (LBVH.slang:34-152)
  const int passes = W.morton_bits == 30 ? 4 : 8;
  bool in_a = t.radix_sort(ka, ia, kb, ib, N, passes);
  t.sync();
  in_a = t.sorted_in_a(in_a);
  if (!in_a) {
    for (uint32_t i = tid; i < N; i += NT) {
      ka[i] = kb[i];
      ia[i] = ib[i];
    }
    t.sync();
  }
  t.mark(kMarkSort);
  // ---- LBVH hierarchy (RB_sim.slang:206-278)
  const uint32_t nb = 2 * off;
  KeyView K{ka, (int)N, W.morton_bits == 30 ? 30 : 63};
  const uint32_t leaf_off = N - 1;
  for (uint32_t i = tid; i < N; i += NT) {
    uint32_t src = off + ia[i];
    uint32_t id = d.prev.id[src];
    Shape s = shape_of(d, off + id);
    v3 mn, mx;
    body_aabb(s, ld3(d.prev.pos[src]), ldq(d.prev.rot[src]), mn, mx);
    d.nmin[nb + leaf_off + i] = st3(mn, __int_as_float(-1));
    d.nmax[nb + leaf_off + i] = st3(mx, __int_as_float((int)i));
  }
  for (uint32_t i = tid; i + 1 < N; i += NT) {
    int first, last;
    lb_range(K, (int)i, first, last);
    int split = lb_split(K, first, last);
    int ca = (split == first) ? (int)leaf_off + split : split;
    int cb = (split + 1 == last) ? (int)leaf_off + split + 1 : split + 1;
    d.nmin[nb + i] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(ca));
    d.nmax[nb + i] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(cb));
    d.parent[nb + ca] = i;
    d.parent[nb + cb] = i;
    d.visit[nb + i] = 0;
    kb[i] = (uint64_t)(uint32_t)first | ((uint64_t)(uint32_t)last << 32);  // leaf range (the sort is done with kb)
  }
  if (tid == 0) {
    L.nroots = 0;
    L.nroots2 = 0;
  }
  t.sync();
  t.mark(kMarkHierarchy);
  // ---- bounds, bottom-up: the second child to arrive merges (exact min/max
  // through the order-preserving map, same values as the reference's integer
  // atomics, RB_sim.slang:282-349)
  // Two phases: each workgroup climbs the nodes over its own runs of leaves
  // (workgroup-scope ordering, see leaves_local) and hands the first node
  // above them to the lead workgroup (ib, free after the sort); after a
  // barrier the lead workgroup climbs the rest.  A node's bounds are the same
  // whichever child arrives last.
  auto merge = [&](uint32_t p) -> bool {
    const float4 a0 = d.nmin[nb + p], a1 = d.nmax[nb + p];
    const int l = __float_as_int(a0.w), r = __float_as_int(a1.w);
    if (l < 0 || r < 0 || l >= (int)(2 * N - 1) || r >= (int)(2 * N - 1)) return false;
    const float4 lm = d.nmin[nb + l], lx = d.nmax[nb + l], rm = d.nmin[nb + r], rx = d.nmax[nb + r];
    d.nmin[nb + p] = make_float4(omin(lm.x, rm.x), omin(lm.y, rm.y), omin(lm.z, rm.z), a0.w);
    d.nmax[nb + p] = make_float4(omax(lx.x, rx.x), omax(lx.y, rx.y), omax(lx.z, rx.z), a1.w);
    return true;
  };
  if (N > 1) {
    for (uint32_t i = tid; i < N; i += NT) {
      uint32_t p = d.parent[nb + leaf_off + i];
      while (p + 1 < N) {  // p is an internal node index in [0, N-2]
        const uint64_t rg = kb[p];
        if (!t.leaves_local((uint32_t)rg, (uint32_t)(rg >> 32))) {
          const uint32_t k = atomicAdd(&L.nroots, 1u);
          if (k < N) ib[k] = p;  // at most one node per run, far fewer than N
          break;
        }
        __threadfence_block();
        if (atomicAdd(&d.visit[nb + p], 1u) == 0u) break;
        __threadfence_block();
        if (!merge(p) || p == 0) break;
        p = d.parent[nb + p];
      }
    }
    t.sync();
    t.mark(kMarkBounds);
    // the nodes handed up from the runs: a team climbs those inside a span of
    // kClimbRuns runs in the span's own workgroup (span s: workgroup s mod
    // team size; no other workgroup touches them in this phase, so
    // workgroup-scope ordering suffices again) and hands the first node above
    // each span on (pcount, free until the broad phase); after a team barrier
    // the lead workgroup climbs the rest.  A one-workgroup world hands
    // nothing up.
    const uint32_t* top = ib;
    uint32_t ntop = L.nroots < N ? L.nroots : N;
    if constexpr (T::kGlobalScalars) {
      const uint32_t span = kClimbRuns * (uint32_t)T::kBlock;
      uint32_t* ib2 = d.pcount + off;
      for (uint32_t k = threadIdx.x; k < ntop; k += blockDim.x) {
        uint32_t p = ib[k];
        if (((uint32_t)kb[p] / span) % t.groups() != t.group()) continue;
        while (p + 1 < N) {
          const uint64_t rg = kb[p];
          if ((uint32_t)rg / span != (uint32_t)(rg >> 32) / span) {
            const uint32_t k2 = atomicAdd(&L.nroots2, 1u);
            if (k2 < N) ib2[k2] = p;
            break;
          }
          __threadfence_block();
          if (atomicAdd(&d.visit[nb + p], 1u) == 0u) break;
          __threadfence_block();
          if (!merge(p) || p == 0) break;
          p = d.parent[nb + p];
        }
      }
      t.sync();
      top = ib2;
      ntop = L.nroots2 < N ? L.nroots2 : N;
    }
    if (t.lead_group()) {
      for (uint32_t k = threadIdx.x; k < ntop; k += blockDim.x) {
        uint32_t p = top[k];
        while (p + 1 < N) {
          __threadfence_block();
          if (atomicAdd(&d.visit[nb + p], 1u) == 0u) break;
          __threadfence_block();
          if (!merge(p) || p == 0) break;
          p = d.parent[nb + p];
        }
      }
      t.mark_lead(kMarkClimb);
    }
  }
  // ---- reorder: rows[i] = prev rows[sorted index] (RB_sim.slang:351-386)
  for (uint32_t i = tid; i < N; i += NT) {
    uint32_t src = off + ia[i], dst = off + i;
    uint32_t id = d.prev.id[src];
    d.cur.id[dst] = id;
    d.cur.flags[dst] = d.prev.flags[src] & ~(uint32_t)BB_BODY_COLLIDING;
    d.cur.sleep[dst] = d.prev.sleep[src];
    d.cur.pos[dst] = d.prev.pos[src];
    d.cur.rot[dst] = d.prev.rot[src];
    d.cur.vel[dst] = d.prev.vel[src];
    d.cur.omg[dst] = d.prev.omg[src];
    d.cur.pvel[dst] = d.prev.pvel[src];
    d.entries_cur[off + id] = i;
    // per-row copies of the id-indexed lookups the narrow phase and the
    // islands make per manifold (one gather per body here instead of two per
    // manifold there)
    d.row_prev[dst] = d.entries_prev[off + id];
    d.row_rank[dst] = d.props.active_rank[off + id];
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_build(Dev d) {
  __shared__ BuildShared L;
  __shared__ ScanLds<BLOCK> S;
  __shared__ SortLds<BLOCK> Q;
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{&S, &Q};
  s_build(d, w, t, L);
}

// Exclusive scan of value(i) over i in [0, n), in index order, by the whole
// team, with one team scan: each wave takes a contiguous run of 64-row
// chunks and sums it, the team scan of the wave sums orders the waves, and
// each wave walks its run again, handing sink(i, offset) every row (a scan
// per NT-row chunk costs a team scan and two barriers per chunk, which adds
// up over a million rows).  value must read what no sink call writes, except
// at its own index; returns the total to every thread.  All threads call it.
template <class T, class Val, class Sink>
BB_DI uint32_t scan_range(const T& t, uint32_t n, Val&& value, Sink&& sink) {
  const uint32_t lane = t.rank() & (WAVE - 1), wv = t.rank() / WAVE, nw = t.size() / WAVE;
  const uint32_t chunks = (n + WAVE - 1) / WAVE, per = (chunks + nw - 1) / nw;
  const uint32_t c0 = wv * per < chunks ? wv * per : chunks;
  const uint32_t c1 = c0 + per < chunks ? c0 + per : chunks;
  uint32_t sum = 0;
  for (uint32_t c = c0; c < c1; ++c) {
    const uint32_t i = c * WAVE + lane;
    if (i < n) sum += value(i);
  }
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_down(sum, o, WAVE);
  uint32_t total;
  const uint32_t ex = t.scan(lane == 0 ? sum : 0u, total);
  uint32_t carry = __shfl(ex, 0, WAVE);  // the sums of the waves before this one
  for (uint32_t c = c0; c < c1; ++c) {
    const uint32_t i = c * WAVE + lane;
    const uint32_t v = i < n ? value(i) : 0u;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, WAVE);
      if (lane >= (uint32_t)o) x += y;
    }
    if (i < n) sink(i, carry + x - v);
    carry += __shfl(x, WAVE - 1, WAVE);
  }
  return total;
}

// ================================================================ k_broad
BB_DI bool overlap(v3 amn, v3 amx, float4 bmn, float4 bmx) {  // Aabb::overlap, math.hpp:213-217
  return amn.x <= bmx.x && amx.x >= bmn.x && amn.y <= bmx.y && amx.y >= bmn.y && amn.z <= bmx.z &&
         amx.z >= bmn.z;
}

// LBVH::broad_phase (LBVH.slang:223-286): stack traversal, one pair (i, j)
// per overlapping leaf with row j > i.  Each node record carries its
// children's indices (nmin.w, nmax.w), so descending needs one dependent load
// per level and the stack holds child-index pairs.  walk() starts below a
// node whose children are (l, r); emit(k, j) receives the k-th partner j.
// Partners come out in ascending row order: leaf k holds row k, and a Karras
// node's left subtree covers the rows before its right subtree's.
//
// A left subtree whose rows all lie at or below i is skipped without loading
// it: the left child of a node with split s covers rows [first, s], and s
// follows from the right child's index (s + 1, or the leaf N - 1 + s + 1),
// so the pairs are those of the full traversal.
//
// `budget` bounds the nodes a thread expands; when it runs out, over_budget()
// decides: true hands the row over (walk returns kMAX), false walks on.
template <class Emit, class Over>
BB_DI uint32_t walk(const Dev& d, uint32_t nb, uint32_t N, uint32_t i, v3 qmn, v3 qmx, int l, int r,
                    uint32_t budget, Over&& over_budget, Emit&& emit) {
  // depth of a Karras tree is at most key bits + index bits (63 + 32 < 128)
  const int nn = (int)(2 * N - 1), leaf_off = (int)N - 1;
  int2 stack[128];
  uint32_t sp = 0, visits = 0;
  uint32_t cnt = 0;
  for (;;) {
    if (l < 0 || r < 0 || l >= nn || r >= nn) break;
    if (++visits > budget) {
      if (over_budget()) return kMAX;
      budget = kMAX;
    }
    // rows of the left subtree end at its split
    const bool lskip = (r >= leaf_off ? r - leaf_off : r) - 1 <= (int)i;
    const float4 rm = d.nmin[nb + r], rx = d.nmax[nb + r];
    bool ol = false, ll = false;
    uint32_t lp = 0;
    float4 lm = rm, lx = rx;
    if (!lskip) {
      lm = d.nmin[nb + l];
      lx = d.nmax[nb + l];
      ol = overlap(qmn, qmx, lm, lx);
      ll = __float_as_int(lm.w) < 0;
      lp = (uint32_t)__float_as_int(lx.w);
    }
    const bool orr = overlap(qmn, qmx, rm, rx);
    const bool rl = __float_as_int(rm.w) < 0;
    const uint32_t rp = (uint32_t)__float_as_int(rx.w);
    if (ol && ll && i < lp) emit(cnt++, lp);
    if (orr && rl && i < rp) emit(cnt++, rp);
    const bool tl = ol && !ll, tr = orr && !rl;
    if (tl) {
      if (tr && sp < 128) stack[sp++] = make_int2(__float_as_int(rm.w), __float_as_int(rx.w));
      l = __float_as_int(lm.w);
      r = __float_as_int(lx.w);
    } else if (tr) {
      l = __float_as_int(rm.w);
      r = __float_as_int(rx.w);
    } else {
      if (sp == 0) break;
      const int2 c = stack[--sp];
      l = c.x;
      r = c.y;
    }
  }
  return cnt;
}

// The whole traversal of row i from the root.
template <class Emit>
BB_DI uint32_t traverse(const Dev& d, uint32_t nb, uint32_t N, uint32_t i, v3 qmn, v3 qmx, Emit&& emit) {
  if (N < 2) return 0;
  const float4 r0m = d.nmin[nb], r0x = d.nmax[nb];
  return walk(d, nb, N, i, qmn, qmx, __float_as_int(r0m.w), __float_as_int(r0x.w), kMAX, [] { return false; },
              emit);
}

// A row whose walk runs past its budget (the ground under a pile overlaps
// thousands of bodies) is walked by a whole workgroup instead: the frontier
// below the root is widened level by level, in leaf order, until it holds a
// node per thread, then each thread walks below its frontier nodes.  The
// frontier's nodes cover disjoint runs of rows in ascending order, so an
// exclusive scan of the per-thread counts places every partner where the
// one-thread walk would have put it.
template <int BLOCK>
struct WalkLds {
  uint32_t fr[2][2 * BLOCK];  // frontier nodes (leaf index N - 1 + row for a partner)
};
static constexpr uint32_t kMaxDefer = 64;   // rows per world handed to a workgroup
static constexpr uint32_t kDeferBit = 0x80000000u;  // pcount mark of such a row

// Count (out == nullptr) or write the first `keep` partners of row i to
// out[0..keep): called by every thread of the workgroup; returns the count.
template <int BLOCK>
BB_DI uint32_t coop_walk(const Dev& d, uint32_t nb, uint32_t N, uint32_t i, v3 qmn, v3 qmx, WalkLds<BLOCK>& F,
                         ScanLds<BLOCK>& S, uint2* out, uint32_t keep) {
  const uint32_t tid = threadIdx.x;
  const int nn = (int)(2 * N - 1), leaf_off = (int)N - 1;
  if (N < 2) return 0;
  uint32_t cur = 0, nf = 1;
  if (tid == 0) F.fr[0][0] = 0u;  // the root (tested no more than walk tests it)
  __syncthreads();
  for (uint32_t round = 0; round < 128; ++round) {
    uint32_t n = 0, a = 0, b = 0, inner = 0;
    if (tid < nf) {
      const int x = (int)F.fr[cur][tid];
      if (x >= leaf_off) {
        a = (uint32_t)x;
        n = 1;
      } else {
        const int l = __float_as_int(d.nmin[nb + x].w), r = __float_as_int(d.nmax[nb + x].w);
        if (l >= 0 && r >= 0 && l < nn && r < nn) {
          const bool lskip = (r >= leaf_off ? r - leaf_off : r) - 1 <= (int)i;
          const float4 rm = d.nmin[nb + r], rx = d.nmax[nb + r];
          const bool orr = overlap(qmn, qmx, rm, rx);
          const bool rl = __float_as_int(rm.w) < 0;
          bool ol = false, ll = false;
          uint32_t lp = 0;
          if (!lskip) {
            const float4 lm = d.nmin[nb + l], lx = d.nmax[nb + l];
            ol = overlap(qmn, qmx, lm, lx);
            ll = __float_as_int(lm.w) < 0;
            lp = (uint32_t)__float_as_int(lx.w);
          }
          const uint32_t rp = (uint32_t)__float_as_int(rx.w);
          uint32_t o[2];
          if (ol && (!ll || i < lp)) {
            o[n++] = (uint32_t)l;
            inner += ll ? 0u : 1u;
          }
          if (orr && (!rl || i < rp)) {
            o[n++] = (uint32_t)r;
            inner += rl ? 0u : 1u;
          }
          a = o[0];
          b = o[1];
        }
      }
    }
    // one scan places the new frontier and counts its inner nodes (high
    // half; the frontier stays below 2 * BLOCK <= 2048 entries)
    uint32_t total;
    const uint32_t ex = block_scan<BLOCK>(n | inner << 16, S, total) & 0xFFFFu;  // the old frontier is read
    if (n > 0) F.fr[cur ^ 1u][ex] = a;
    if (n > 1) F.fr[cur ^ 1u][ex + 1] = b;
    __syncthreads();
    cur ^= 1u;
    nf = total & 0xFFFFu;
    if ((total >> 16) == 0u || nf >= (uint32_t)BLOCK) break;
  }
  // each thread walks below its frontier nodes (two each past BLOCK nodes)
  const uint32_t per = nf > (uint32_t)BLOCK ? 2u : 1u;
  auto below = [&](uint32_t e, auto&& emit) -> uint32_t {
    const int x = (int)F.fr[cur][e];
    if (x >= leaf_off) {
      emit(0u, (uint32_t)(x - leaf_off));
      return 1u;
    }
    return walk(d, nb, N, i, qmn, qmx, __float_as_int(d.nmin[nb + x].w), __float_as_int(d.nmax[nb + x].w), kMAX,
                [] { return false; }, emit);
  };
  uint32_t mine = 0;
  for (uint32_t q = 0; q < per; ++q) {
    const uint32_t e = tid * per + q;
    if (e < nf) mine += below(e, [](uint32_t, uint32_t) {});
  }
  uint32_t total;
  uint32_t at = block_scan<BLOCK>(mine, S, total);
  if (out) {
    for (uint32_t q = 0; q < per; ++q) {
      const uint32_t e = tid * per + q;
      if (e < nf) {
        const uint32_t base = at;
        at += below(e, [&](uint32_t k, uint32_t j) {
          if (base + k < keep) out[base + k] = make_uint2(i, j);
        });
      }
    }
  }
  __syncthreads();  // the frontier is free for the next row
  return total;
}

// Rows keep up to kPairStage partners in a staging area (the colouring
// scratch `pick`, whose 8n+64 slots per world hold 8 per row) during the
// counting traversal; only rows with more partners traverse a second time.
static constexpr uint32_t kPairStage = 8;

BB_DI void cas_u32(uint32_t& a, uint32_t& b) {
  const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  a = lo;
  b = hi;
}
// Optimal 19-comparator sorting network for 8 keys (Knuth, TAOCP vol. 3).
BB_DI void sort8(uint32_t (&k)[8]) {
  cas_u32(k[0], k[2]); cas_u32(k[1], k[3]); cas_u32(k[4], k[6]); cas_u32(k[5], k[7]);
  cas_u32(k[0], k[4]); cas_u32(k[1], k[5]); cas_u32(k[2], k[6]); cas_u32(k[3], k[7]);
  cas_u32(k[0], k[1]); cas_u32(k[2], k[3]); cas_u32(k[4], k[5]); cas_u32(k[6], k[7]);
  cas_u32(k[2], k[4]); cas_u32(k[3], k[5]);
  cas_u32(k[1], k[4]); cas_u32(k[3], k[6]);
  cas_u32(k[1], k[2]); cas_u32(k[3], k[4]); cas_u32(k[5], k[6]);
}

struct BroadShared {
  uint32_t carry;
  uint32_t pairs;   // voxel worlds: pairs, as opposed to pair slots
  uint32_t ndefer;  // rows handed to a workgroup (coop_walk), listed in defer
  uint32_t defer[kMaxDefer];
};

template <class T>
BB_DI void s_broad(const Dev& d, uint32_t w, const T& t, BroadShared& L, WalkLds<T::kBlock>& F) {
  const uint32_t tid = t.rank(), NT = t.size();
  const WorldInfo W = d.worlds[w];
  const uint32_t N = W.n, off = W.body_off, nb = 2 * off;
  uint32_t* stage = d.pick + W.pair_off;
  // voxel worlds: a pair with a voxel body takes six slots (one per contact-
  // face code), so the per-row offsets count slots, and the pair counter
  // counts pairs
  const bool vox = (W.mode & kModeVoxel) != 0;
  auto is_vox = [&](uint32_t row) {
    return (shape_word(d.props.pmax[off + d.cur.id[off + row]].w) & 0xFFu) == BB_SHAPE_VOXEL;
  };
  if (tid == 0) {
    L.carry = 0;
    L.pairs = 0;
    L.ndefer = 0;
  }
  t.sync();
  // partner counts (slots of voxel worlds parked in poff); a row whose walk
  // runs past the budget is listed for a workgroup and marked kDeferBit
  const uint32_t budget = vox ? kMAX : (d.walk_budget ? d.walk_budget : kWalkBudgetDefault);
  for (uint32_t i = tid; i < N; i += NT) {
    const float4 mn = d.nmin[nb + N - 1 + i], mx = d.nmax[nb + N - 1 + i];
    uint32_t* st = stage + i * kPairStage;
    const bool vi = vox && is_vox(i);
    uint32_t slots = 0, cnt = 0;
    if (N >= 2) {
      const float4 r0m = d.nmin[nb], r0x = d.nmax[nb];
      cnt = walk(
          d, nb, N, i, ld3(mn), ld3(mx), __float_as_int(r0m.w), __float_as_int(r0x.w), budget,
          [&]() {
            const uint32_t k = atomicAdd(&L.ndefer, 1u);
            if (k >= kMaxDefer) return false;
            L.defer[k] = i;
            return true;
          },
          [&](uint32_t k, uint32_t j) {
            if (k < kPairStage) st[k] = j;
            if (vox) slots += (vi || is_vox(j)) ? 6u : 1u;
          });
    }
    d.pcount[off + i] = cnt == kMAX ? kDeferBit : cnt;
    if (vox) d.poff[off + i] = slots;
  }
  t.sync();
  const uint32_t nd = L.ndefer < kMaxDefer ? L.ndefer : kMaxDefer;
  for (uint32_t k = t.group(); k < nd; k += t.groups()) {
    const uint32_t i = L.defer[k];
    const float4 mn = d.nmin[nb + N - 1 + i], mx = d.nmax[nb + N - 1 + i];
    const uint32_t cnt = coop_walk<T::kBlock>(d, nb, N, i, ld3(mn), ld3(mx), F, *t.scan_lds, nullptr, 0u);
    if (threadIdx.x == 0) d.pcount[off + i] = kDeferBit | cnt;
  }
  if (nd) t.sync();
  // per-row offsets of the pair slots (voxel worlds: slots parked in poff,
  // and the pairs counted apart)
  const uint32_t vox_pairs =
      vox ? scan_range(t, N, [&](uint32_t i) { return d.pcount[off + i] & ~kDeferBit; }, [](uint32_t, uint32_t) {})
          : 0u;
  const uint32_t P = scan_range(
      t, N, [&](uint32_t i) { return vox ? d.poff[off + i] : (d.pcount[off + i] & ~kDeferBit); },
      [&](uint32_t i, uint32_t ex) {
        d.poff[off + i] = ex;
        d.pstart_cur[off + i] = ex;
      });
  t.sync();
  const uint32_t cap = W.pair_cap;
  if (tid == 0) {
    d.npairs_cur[w] = P < cap ? P : cap;
    d.stats[w * ST_COUNT + ST_P] = vox ? vox_pairs : P;
    if (P > cap) d.stats[w * ST_COUNT + ST_OVERFLOW] = 1;
  }
  uint2* pairs = d.pairs_cur + W.pair_off;
  if (vox) {
    // partners in ascending row order, each written 1 or 6 times; slots past
    // the capacity are dropped (the overflow flag is set)
    for (uint32_t i = tid; i < N; i += NT) {
      const uint32_t o = d.poff[off + i];
      const uint32_t cnt = d.pcount[off + i];
      if (cnt == 0 || o >= cap) continue;
      const bool vi = is_vox(i);
      uint32_t* st = stage + i * kPairStage;
      uint32_t k8[8];
      if (cnt <= kPairStage) {
        static_for<8>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          k8[j] = (uint32_t)j < cnt ? st[j] : 0xFFFFFFFFu;
        });
        sort8(k8);
      }
      uint32_t at = o;
      auto put = [&](uint32_t j) {
        const uint32_t rep = (vi || is_vox(j)) ? 6u : 1u;
        for (uint32_t r = 0; r < rep; ++r, ++at)
          if (at < cap) pairs[at] = make_uint2(i, j);
      };
      if (cnt <= kPairStage) {
        static_for<8>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          if ((uint32_t)j < cnt) put(k8[j]);
        });
        continue;
      }
      // more partners than the staging area: traverse again into the slot
      // range (it holds at least cnt entries), sort there, then expand from
      // the back (an entry only moves to a higher or equal slot)
      const uint32_t room = cap - o;
      const uint32_t keep = cnt < room ? cnt : room;
      float4 mn = d.nmin[nb + N - 1 + i], mx = d.nmax[nb + N - 1 + i];
      traverse(d, nb, N, i, ld3(mn), ld3(mx), [&](uint32_t k, uint32_t j) {
        if (k < keep) pairs[o + k] = make_uint2(i, j);
      });
      for (uint32_t a = 1; a < keep; ++a) {
        uint2 x = pairs[o + a];
        uint32_t b2 = a;
        while (b2 > 0 && pairs[o + b2 - 1].y > x.y) {
          pairs[o + b2] = pairs[o + b2 - 1];
          --b2;
        }
        pairs[o + b2] = x;
      }
      uint32_t end = o;
      for (uint32_t a = 0; a < keep; ++a) end += (vi || is_vox(pairs[o + a].y)) ? 6u : 1u;
      for (uint32_t a = keep; a-- > 0;) {
        const uint32_t j = pairs[o + a].y;
        const uint32_t rep = (vi || is_vox(j)) ? 6u : 1u;
        end -= rep;
        for (uint32_t r = 0; r < rep; ++r)
          if (end + r < cap) pairs[end + r] = make_uint2(i, j);
      }
    }
    return;
  }
  for (uint32_t i = tid; i < N; i += NT) {
    const uint32_t o = d.poff[off + i];
    if (o >= cap) continue;
    const uint32_t room = cap - o;
    uint32_t cnt = d.pcount[off + i];
    if (cnt & kDeferBit) continue;  // written by its workgroup below
    const uint32_t keep = cnt < room ? cnt : room;  // first `room` in traversal order
    if (cnt <= kPairStage) {
      // canonical order: ascending partner row
      uint32_t k8[8];
      static_for<8>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        k8[j] = (uint32_t)j < keep ? stage[i * kPairStage + j] : 0xFFFFFFFFu;
      });
      sort8(k8);
      static_for<8>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if ((uint32_t)j < keep) pairs[o + j] = make_uint2(i, k8[j]);
      });
      continue;
    }
    // more partners than the staging area holds: traverse again, writing the
    // first `room` partners straight into the pair list, then sort them
    float4 mn = d.nmin[nb + N - 1 + i], mx = d.nmax[nb + N - 1 + i];
    traverse(d, nb, N, i, ld3(mn), ld3(mx), [&](uint32_t k, uint32_t j) {
      if (k < keep) pairs[o + k] = make_uint2(i, j);
    });
    for (uint32_t a = 1; a < keep; ++a) {
      uint2 x = pairs[o + a];
      uint32_t b2 = a;
      while (b2 > 0 && pairs[o + b2 - 1].y > x.y) {
        pairs[o + b2] = pairs[o + b2 - 1];
        --b2;
      }
      pairs[o + b2] = x;
    }
  }
  // the rows handed to a workgroup: the first `room` partners, in order
  for (uint32_t k = t.group(); k < nd; k += t.groups()) {
    const uint32_t i = L.defer[k];
    const uint32_t o = d.poff[off + i];
    if (o >= cap) continue;  // workgroup-uniform
    const uint32_t cnt = d.pcount[off + i] & ~kDeferBit, room = cap - o;
    const float4 mn = d.nmin[nb + N - 1 + i], mx = d.nmax[nb + N - 1 + i];
    (void)coop_walk<T::kBlock>(d, nb, N, i, ld3(mn), ld3(mx), F, *t.scan_lds, pairs + o, cnt < room ? cnt : room);
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_broad(Dev d) {
  __shared__ BroadShared L;
  __shared__ ScanLds<BLOCK> S;
  __shared__ WalkLds<BLOCK> F;
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{&S, nullptr};
  s_broad(d, w, t, L, F);
}

// ==================================================== narrow phase (SAT)
// OBBvsOBBCollision (src/shaders/simulation/collision_detection.slang:5-877),
// restated for one pair per thread.
struct Tx {
  mat3 R;
  v3 p;
};
BB_DI v3 tx_rot(const Tx& t, v3 v) { return mulvm(v, t.R); }
BB_DI v3 tx_rotT(const Tx& t, v3 v) { return mulvm(v, mT(t.R)); }
BB_DI v3 tx_tf(const Tx& t, v3 v) { return mulvm(v, t.R) + t.p; }
BB_DI v3 tx_tfT(const Tx& t, v3 v) { return mulvm(v - t.p, mT(t.R)); }

// Clip vertex: position and the four feature ids (in_ref, out_ref, in_inc,
// out_inc) packed one per byte.  Polygons live in fixed 8-entry register
// arrays; every index is a compile-time constant after unrolling.
struct CV {
  float x, y, z;
  uint32_t f;
};
BB_DI v3 cv_v(const CV& c) { return V(c.x, c.y, c.z); }
BB_DI void cv_set(CV& c, v3 v) { c.x = v.x; c.y = v.y; c.z = v.z; }
BB_DI uint32_t cv_fp(const CV& c, int k) { return (c.f >> (8 * k)) & 0xFFu; }

enum : uint32_t {
  PXPY = 0, PXNY = 1, PXPZ = 2, PXNZ = 3, NXPY = 4, NXNY = 5, NXPZ = 6, NXNZ = 7,
  PYPZ = 8, PYNZ = 9, NYPZ = 10, NYNZ = 11
};

BB_DI void set_inc(CV (&o)[8], int i, v3 v, uint32_t in, uint32_t out) {
  o[i] = CV{v.x, v.y, v.z, (in << 16) | (out << 24)};
}

// compute_incident_face (collision_detection.slang:124-264)
BB_DI void incident_face(const Tx& itx, v3 e, v3 n, CV (&o)[8]) {
  n = -tx_rotT(itx, n);
  v3 an = vabs(n);
  if (an.x > an.y && an.x > an.z) {
    if (n.x > 0.0f) {
      set_inc(o, 0, V(e.x, e.y, -e.z), PYNZ, PXNY);
      set_inc(o, 1, V(e.x, e.y, e.z), PXNY, PYPZ);
      set_inc(o, 2, V(e.x, -e.y, e.z), PYPZ, NXNZ);
      set_inc(o, 3, V(e.x, -e.y, -e.z), NXNZ, PYNZ);
    } else {
      set_inc(o, 0, V(-e.x, -e.y, e.z), NXNY, NYNZ);
      set_inc(o, 1, V(-e.x, e.y, e.z), NYNZ, PXNZ);
      set_inc(o, 2, V(-e.x, e.y, -e.z), PXNZ, NYPZ);
      set_inc(o, 3, V(-e.x, -e.y, -e.z), NYPZ, NXNY);
    }
  } else if (an.y > an.x && an.y > an.z) {
    if (n.y > 0.0f) {
      set_inc(o, 0, V(-e.x, e.y, e.z), PXNZ, PXPY);
      set_inc(o, 1, V(e.x, e.y, e.z), PXPY, PXNY);
      set_inc(o, 2, V(e.x, e.y, -e.z), PXNY, PXPZ);
      set_inc(o, 3, V(-e.x, e.y, -e.z), PXPZ, PXNZ);
    } else {
      set_inc(o, 0, V(e.x, -e.y, e.z), NXNZ, NXPY);
      set_inc(o, 1, V(-e.x, -e.y, e.z), NXPY, NXNY);
      set_inc(o, 2, V(-e.x, -e.y, -e.z), NXNY, NXPZ);
      set_inc(o, 3, V(e.x, -e.y, -e.z), NXPZ, NXNZ);
    }
  } else {
    if (n.z > 0.0f) {
      set_inc(o, 0, V(-e.x, e.y, e.z), PXPY, NYNZ);
      set_inc(o, 1, V(-e.x, -e.y, e.z), NYNZ, NXPY);
      set_inc(o, 2, V(e.x, -e.y, e.z), NXPY, PYPZ);
      set_inc(o, 3, V(e.x, e.y, e.z), PYPZ, PXPY);
    } else {
      set_inc(o, 0, V(e.x, -e.y, -e.z), PYNZ, NXPZ);
      set_inc(o, 1, V(-e.x, -e.y, -e.z), NXPZ, NYPZ);
      set_inc(o, 2, V(-e.x, e.y, -e.z), NYPZ, PXPZ);
      set_inc(o, 3, V(e.x, e.y, -e.z), PXPZ, PYNZ);
    }
  }
  static_for<4>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    cv_set(o[i], tx_tf(itx, cv_v(o[i])));
  });
}

// compute_reference_edges_and_basis (collision_detection.slang:266-336)
BB_DI void reference_edges(v3 er, const Tx& rtx, v3 n, int axis, uint32_t* ce, mat3& basis, v3& e) {
  n = tx_rotT(rtx, n);
  if (axis > 2) axis -= 3;
  v3 r0 = mrow(rtx.R, 0), r1 = mrow(rtx.R, 1), r2 = mrow(rtx.R, 2);
  if (axis == 0) {
    if (n.x > 0.0f) {
      ce[0] = PXNY; ce[1] = PYPZ; ce[2] = NXNZ; ce[3] = PYNZ;
      e = V(er.y, er.z, er.x); basis = mrows(r1, r2, r0);
    } else {
      ce[0] = NYNZ; ce[1] = PXNZ; ce[2] = NYPZ; ce[3] = NXNY;
      e = V(er.z, er.y, er.x); basis = mrows(r2, r1, -r0);
    }
  } else if (axis == 1) {
    if (n.y > 0.0f) {
      ce[0] = PXPY; ce[1] = PXNY; ce[2] = PXPZ; ce[3] = PXNZ;
      e = V(er.z, er.x, er.y); basis = mrows(r2, r0, r1);
    } else {
      ce[0] = NXPY; ce[1] = NXNY; ce[2] = NXPZ; ce[3] = NXNZ;
      e = V(er.z, er.x, er.y); basis = mrows(r2, -r0, -r1);
    }
  } else {
    if (n.z > 0.0f) {
      ce[0] = NYNZ; ce[1] = NXPY; ce[2] = PYPZ; ce[3] = PXPY;
      e = V(er.y, er.x, er.z); basis = mrows(-r1, r0, r2);
    } else {
      ce[0] = NXPZ; ce[1] = NYPZ; ce[2] = PXPZ; ce[3] = PYNZ;
      e = V(er.y, er.x, er.z); basis = mrows(-r1, -r0, -r2);
    }
  }
}

BB_DI bool clip_on(float v) { return (v < 0.005f) && (v > -0.005f); }

BB_DI CV cv_sel(bool c, const CV& x, const CV& y) {
  return CV{c ? x.x : y.x, c ? x.y : y.y, c ? x.z : y.z, c ? x.f : y.f};
}
// Writes x into o[k] (k < 8) with a select per entry: no runtime index and
// no conditional store, so the array stays in registers.
BB_DI void cv_put(CV (&o)[8], int k, const CV& x) {
  static_for<8>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    o[j] = cv_sel(j == k, x, o[j]);
  });
}
BB_DI CV cv_lerp(const CV& a, const CV& b, float t) {
  return CV{a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, 0u};
}

// ortographic_clip (collision_detection.slang:352-388) against the side
// plane sign * v[AX] = e; output capped at 8.
template <int MAXIN, int AX, int SG>
BB_DI int clip_pass(float e, uint32_t ce, const CV (&in)[8], int nin, CV (&out)[8]) {
  const float sign = SG > 0 ? 1.0f : -1.0f;
  CV a = in[0];
  static_for<MAXIN>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k > 0) a = cv_sel(k == nin - 1, in[k], a);
  });
  int nout = 0;
  static_for<MAXIN>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    if (i < nin) {
      const CV b = in[i];
      const float da = sign * (AX ? a.y : a.x) - e;
      const float db = sign * (AX ? b.y : b.x) - e;
      if ((da < 0.0f && db < 0.0f) || clip_on(da) || clip_on(db)) {
        if (nout < 8) cv_put(out, nout++, b);
      } else if (da < 0.0f && db >= 0.0f) {
        CV c = cv_lerp(a, b, da / (da - db));
        c.f = (b.f & ~0x0000FF00u) | (ce << 8);
        if (nout < 8) cv_put(out, nout++, c);
      } else if (da >= 0.0f && db < 0.0f) {
        CV c = cv_lerp(a, b, da / (da - db));
        c.f = (a.f & ~0x000000FFu) | ce;
        if (nout < 8) cv_put(out, nout++, c);
        if (nout < 8) cv_put(out, nout++, b);
      }
      a = b;
    }
  });
  return nout;
}

BB_DI float c_sign(float x) { return x < 0.0f ? -1.0f : 1.0f; }

// support_edge (collision_detection.slang:459-500)
BB_DI void support_edge(const Tx& tx, v3 e, v3 n, v3& ca, v3& cb) {
  n = tx_rotT(tx, n);
  v3 an = vabs(n);
  int axis;
  if (an.x <= an.y && an.x <= an.z) axis = 0;
  else if (an.y <= an.x && an.y <= an.z) axis = 1;
  else axis = 2;
  v3 v0 = V(c_sign(n.x) * e.x, c_sign(n.y) * e.y, c_sign(n.z) * e.z);
  v3 v1 = v0;
  if (axis == 0) { v0.x = e.x; v1.x = -e.x; }
  else if (axis == 1) { v0.y = e.y; v1.y = -e.y; }
  else { v0.z = e.z; v1.z = -e.z; }
  ca = tx_tf(tx, v0);
  cb = tx_tf(tx, v1);
}

// edges_contact (collision_detection.slang:502-553)
BB_DI void edges_contact(v3 PA, v3 QA, v3 PB, v3 QB, v3& CA, v3& CB) {
  v3 DA = QA - PA, DB = QB - PB, r = PA - PB;
  float a = dot(DA, DA), e = dot(DB, DB), f = dot(DB, r), c = dot(DA, r), b = dot(DA, DB);
  float denom = a * e - b * b;
  if (denom == 0.0f) {
    float t = dot(r, DA) / a;
    t = fcl(t, 0.0f, 1.0f);
    CA = PA + DA * t;
    v3 s = CA - PB;
    float u = dot(s, DB) / e;
    u = fcl(u, 0.0f, 1.0f);
    CB = PB + DB * u;
    return;
  }
  float TA = (b * f - c * e) / denom;
  TA = fcl(TA, 0.0f, 1.0f);
  float TB = (b * TA + f) / e;
  if (TB < 0.0f) {
    TB = 0.0f;
    TA = fcl(-c / a, 0.0f, 1.0f);
  } else if (TB > 1.0f) {
    TB = 1.0f;
    TA = fcl((b - c) / a, 0.0f, 1.0f);
  }
  CA = PA + DA * TA;
  CB = PB + DB * TB;
}

struct BodyView {
  v3 pos;
  qt rot;
  v3 mn, mx;
  float radius;
  uint32_t kind;
  uint32_t vid;  // voxel shape id (1-based) of a voxel body
};

// Detected contacts go straight into the output manifold slot: position +
// penetration and feature pair of contact k at base + k.
struct CSink {
  float4* cpos;
  uint4* cfp;
  uint32_t base;
  bool capped;  // box-box penetrations are floored at -0.25
  float deep;   // min(0, penetrations written)
  BB_DI void put(int k, v3 p, float pen, uint32_t f0, uint32_t f1, uint32_t f2, uint32_t f3) {
    if (capped) pen = fmx(pen, -0.25f);
    deep = fmn(deep, pen);
    cpos[base + k] = st3(p, pen);
    cfp[base + k] = make_uint4(f0, f1, f2, f3);
  }
};

// OBBvsOBBCollision::collision_detection (collision_detection.slang:568-877).
// Returns the contact count (0: no manifold); key and normal on output.
BB_DI int obb_obb(const BodyView& A, const BodyView& B, CSink& sk, int& key, v3& normal_out) {
  key = -1;
  Tx txA{mT(qmat(A.rot)), A.pos}, txB{mT(qmat(B.rot)), B.pos};
  v3 eA = (A.mx - A.mn) * 0.5f, eB = (B.mx - B.mn) * 0.5f;
  mat3 C = mmul(mT(txA.R), txB.R);
  mat3 aC = mabs(C);
  bool parallel = false;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (aC.m[i][j] > 0.9998f) parallel = true;
  v3 t = qrot_inv(A.rot, B.pos - A.pos);
  float a_max = -3.402823466e+38f, b_max = -3.402823466e+38f, e_max = -3.402823466e+38f;
  int a_axis = -1, b_axis = -1, e_axis = -1;
  v3 n_a = V(0, 0, 0), n_b = V(0, 0, 0), n_e = V(0, 0, 0);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float s = fab(cmp(t, i)) - (cmp(eA, i) + dot(mcol(aC, i), eB));
    if (s > 0.0f) return 0;
    if (s > a_max) { a_max = s; a_axis = i; n_a = mrow(txA.R, i); }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float proj = dot(t, mrow(C, i));
    float s = fab(proj) - (cmp(eB, i) + dot(mrow(aC, i), eA));
    if (s > 0.0f) return 0;
    if (s > b_max) { b_max = s; b_axis = 3 + i; n_b = mrow(txB.R, i); }
  }
  if (!parallel) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        float rA = cmp(eA, i1) * aC.m[j][i2] + cmp(eA, i2) * aC.m[j][i1];
        float rB = cmp(eB, j1) * aC.m[j2][i] + cmp(eB, j2) * aC.m[j1][i];
        float s = fab(cmp(t, i2) * C.m[j][i1] - cmp(t, i1) * C.m[j][i2]) - (rA + rB);
        v3 ax = V(0, 0, 0);
        setcmp(ax, i1, -C.m[j][i2]);
        setcmp(ax, i2, C.m[j][i1]);
        if (s > 0.0f) return 0;
        float l = 1.0f / vlen(ax);
        s *= l;
        if (s > e_max) { e_max = s; e_axis = 6 + i * 3 + j; n_e = ax * l; }
      }
    }
  }
  int axis;
  float s_max;
  v3 normal;
  float face_max = fmx(a_max, b_max);
  if (0.95f * e_max > face_max + 0.01f) { axis = e_axis; s_max = e_max; normal = n_e; }
  else if (0.95f * b_max > a_max + 0.01f) { axis = b_axis; s_max = b_max; normal = n_b; }
  else { axis = a_axis; s_max = a_max; normal = n_a; }
  if (dot(normal, B.pos - A.pos) < 0.0f) normal = -normal;
  if (axis == -1) return 0;
  int count;
  if (axis <= 5) {
    Tx rtx, itx;
    v3 er, ei;
    bool flip = false;
    if (axis <= 2) { rtx = txA; itx = txB; er = eA; ei = eB; }
    else { rtx = txB; itx = txA; er = eB; ei = eA; flip = true; normal = -normal; }
    CV inc[8], tmp[8];
    incident_face(itx, ei, normal, inc);
    uint32_t ce[4];
    mat3 basis;
    v3 e;
    reference_edges(er, rtx, normal, axis, ce, basis, e);
    Tx tx{basis, rtx.p};
    // clip (collision_detection.slang:390-452)
    static_for<4>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      cv_set(inc[i], tx_tfT(tx, cv_v(inc[i])));
    });
    int nout = clip_pass<4, 0, 1>(e.x, ce[0], inc, 4, tmp);
    if (nout == 0) return 0;
    int nin = clip_pass<8, 1, 1>(e.y, ce[1], tmp, nout, inc);
    if (nin == 0) return 0;
    nout = clip_pass<8, 0, -1>(e.x, ce[2], inc, nin, tmp);
    if (nout == 0) return 0;
    nin = clip_pass<8, 1, -1>(e.y, ce[3], tmp, nout, inc);
    if (nin == 0) return 0;
    int k = 0;
    static_for<8>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if (i < nin) {
        const float dz = inc[i].z - e.z;
        if (dz <= 0.0f) {
          const v3 pw = tx_tf(tx, cv_v(inc[i]));
          const uint32_t f0 = cv_fp(inc[i], 0), f1 = cv_fp(inc[i], 1), f2 = cv_fp(inc[i], 2), f3 = cv_fp(inc[i], 3);
          if (flip) sk.put(k, pw, dz, f2, f3, f0, f1);
          else sk.put(k, pw, dz, f0, f1, f2, f3);
          ++k;
        }
      }
    });
    if (k == 0) return 0;
    count = k;
    normal_out = flip ? -normal : normal;
  } else {
    normal = tx_rot(txA, normal);
    if (dot(normal, B.pos - A.pos) < 0.0f) normal = -normal;
    v3 PA, QA, PB, QB, CA, CB;
    support_edge(txA, eA, normal, PA, QA);
    support_edge(txB, eB, -normal, PB, QB);
    edges_contact(PA, QA, PB, QB, CA, CB);
    count = 1;
    normal_out = normal;
    sk.put(0, (CA + CB) * 0.5f, s_max, (uint32_t)axis, (uint32_t)axis, (uint32_t)axis, (uint32_t)axis);
  }
  key = axis;
  return count;
}

// Sphere contacts (extension; the reference has no spheres): sphere-sphere
// along the centre line, box-sphere through the clamped local closest point
// or, for a centre inside the box, the face of least penetration.
static constexpr int KEY_SPHERE_SPHERE = 20, KEY_BOX_SPHERE = 21;
BB_DI int sphere_sphere(const BodyView& A, const BodyView& B, CSink& sk, int& key, v3& normal_out) {
  key = KEY_SPHERE_SPHERE;
  v3 dd = B.pos - A.pos;
  float dist = vlen(dd);
  float pen = dist - (A.radius + B.radius);
  if (pen > 0.0f) return 0;
  v3 n = dist > 0.0f ? vdiv(dd, dist) : V(0.0f, 1.0f, 0.0f);
  v3 pa = A.pos + n * A.radius, pb = B.pos - n * B.radius;
  normal_out = n;
  const uint32_t k = (uint32_t)KEY_SPHERE_SPHERE;
  sk.put(0, (pa + pb) * 0.5f, pen, k, k, k, k);
  return 1;
}
BB_DI int box_sphere(const BodyView& A, const BodyView& B, CSink& sk, int& key, v3& normal_out) {
  key = KEY_BOX_SPHERE;
  const bool a_box = A.kind == BB_SHAPE_BOX;
  const BodyView& X = a_box ? A : B;
  const BodyView& S = a_box ? B : A;
  v3 e = (X.mx - X.mn) * 0.5f;
  v3 local = qrot_inv(X.rot, S.pos - X.pos);
  v3 q = V(fcl(local.x, -e.x, e.x), fcl(local.y, -e.y, e.y), fcl(local.z, -e.z, e.z));
  v3 dd = local - q;
  float dist2 = dot(dd, dd);
  v3 nl, surf;
  float pen;
  if (dist2 > 0.0f) {
    float dist = __builtin_sqrtf(dist2);
    pen = dist - S.radius;
    if (pen > 0.0f) return 0;
    nl = vdiv(dd, dist);
    surf = q;
  } else {
    float dx = e.x - fab(local.x), dy = e.y - fab(local.y), dz = e.z - fab(local.z);
    int ax = 0;
    float dmin = dx;
    if (dy < dmin) { dmin = dy; ax = 1; }
    if (dz < dmin) { dmin = dz; ax = 2; }
    nl = V(0, 0, 0);
    setcmp(nl, ax, cmp(local, ax) < 0.0f ? -1.0f : 1.0f);
    surf = local;
    setcmp(surf, ax, cmp(nl, ax) * cmp(e, ax));
    pen = -dmin - S.radius;
  }
  v3 n = qrot(X.rot, nl);
  v3 surf_w = qrot(X.rot, surf) + X.pos;
  v3 sph_w = S.pos - n * S.radius;
  normal_out = a_box ? n : -n;
  const uint32_t k = (uint32_t)KEY_BOX_SPHERE;
  sk.put(0, (surf_w + sph_w) * 0.5f, pen, k, k, k, k);
  return 1;
}

BB_DI BodyView body_view(const Dev& d, uint32_t off, uint32_t row) {
  uint32_t g = off + row;
  uint32_t id = d.cur.id[g];
  float4 a = d.props.pmin[off + id], b = d.props.pmax[off + id];
  const uint32_t word = shape_word(b.w);
  return BodyView{ld3(d.cur.pos[g]), ldq(d.cur.rot[g]), ld3(a), ld3(b), a.w, word & 0xFFu, word >> 8};
}
// ================================================== voxel narrow phase
// Voxel-bitmask narrow phase (src/shaders/simulation/collision_detection.slang:
// 1001-1512), one thread per (pair, contact-face code) slot: the surface
// cells of the body with the shorter surface list are sampled against the
// other body's occupancy (or its OBB: a box's bounds, a sphere's bounding
// cube) along axis code/2 of the other body's frame, sign + for odd codes;
// a four-way extreme reducer keeps at most four cells, emitted as at most
// four contacts.  The reducer's arrays are indexed only by unrolled loop
// counters so they stay in registers.
struct VoxRed {
  v3 pos[4];
  float pen[4];
  v3 nrm[4];
  uint32_t fa[4], fb[4];
  float score[4];
  uint32_t samples, interior;
};

// voxel_occ_bit (collision_detection.slang:1009-1018)
BB_DI bool vox_bit(const Dev& d, const VoxShape& s, int cx, int cy, int cz) {
  if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)s.dims[0] || cy >= (int)s.dims[1] || cz >= (int)s.dims[2])
    return false;
  const uint32_t bit = (uint32_t)cx + (uint32_t)cy * s.dims[0] + (uint32_t)cz * s.dims[0] * s.dims[1];
  return ((d.vocc[s.occ_off + bit / 32u] >> (bit % 32u)) & 1u) != 0u;
}

BB_DI v3 vox_origin(const VoxShape& s) { return V(s.ox, s.oy, s.oz); }

// VoxelManifoldReducer::add (collision_detection.slang:1063-1086): per
// extreme direction the higher score bucket, then the deeper depth bucket,
// then the lower cell id.
BB_DI void vox_add(VoxRed& r, v3 p, float pe, v3 n, uint32_t fa, uint32_t fb, float su, float sv, float quant) {
  const float sc4[4] = {su + sv, su - sv, -su + sv, -su - sv};
  const int nb = (int)__builtin_floorf(pe / quant);
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    const bool empty = r.pen[o] > 0.5e9f;
    const int nb_s = (int)__builtin_floorf(sc4[o] / quant);
    const int ob_s = empty ? -2147483640 : (int)__builtin_floorf(r.score[o] / quant);
    const int ob = empty ? 0 : (int)__builtin_floorf(r.pen[o] / quant);
    if (empty || nb_s > ob_s || (nb_s == ob_s && (nb < ob || (nb == ob && fa < r.fa[o])))) {
      r.pos[o] = p;
      r.pen[o] = pe;
      r.nrm[o] = n;
      r.fa[o] = fa;
      r.fb[o] = fb;
      r.score[o] = sc4[o];
    }
  }
  ++r.samples;
}

BB_DI int vox_ci(int x, int y, int z, uint32_t k) { return k == 0 ? x : (k == 1 ? y : z); }

// voxel_sample_body (collision_detection.slang:1096-1266)
BB_DI void vox_sample(const Dev& d, const BodyView& r1, const VoxShape& s1, const BodyView& r2,
                      const VoxShape& s2, bool s2_vox, VoxRed& red, uint32_t code) {
  const float vs1 = s1.vs;
  const v3 a0 = qrot_inv(r2.rot, qrot(r1.rot, V(1.0f, 0.0f, 0.0f)));
  const v3 a1 = qrot_inv(r2.rot, qrot(r1.rot, V(0.0f, 1.0f, 0.0f)));
  const v3 a2 = qrot_inv(r2.rot, qrot(r1.rot, V(0.0f, 0.0f, 1.0f)));
  const v3 sup = V((fab(a0.x) + fab(a1.x)) + fab(a2.x), (fab(a0.y) + fab(a1.y)) + fab(a2.y),
                   (fab(a0.z) + fab(a1.z)) + fab(a2.z)) *
                 (0.5f * vs1);
  const uint32_t axk = code / 2u;
  const float sgn_pass = (code & 1u) != 0u ? 1.0f : -1.0f;
  const int isg = (int)sgn_pass;
  const v3 o1 = vox_origin(s1), o2 = vox_origin(s2);
  // the pass normal, outward of r2, as each branch of the reference forms it
  const v3 n_pass = qrot(r2.rot, V(axk == 0u ? sgn_pass : 0.0f, axk == 1u ? sgn_pass : 0.0f,
                                   axk == 2u ? sgn_pass : 0.0f));
  const v3 n_pass_box = qrot(r2.rot, V(axk == 0u ? 1.0f : 0.0f, axk == 1u ? 1.0f : 0.0f, axk == 2u ? 1.0f : 0.0f) *
                                         sgn_pass);
  for (uint32_t i = 0; i < s1.surf_count; ++i) {
    const uint32_t e = d.vsurf[s1.surf_off + i];
    const v3 local = V(o1.x + ((float)(e & 0xFFu) + 0.5f) * vs1, o1.y + ((float)((e >> 8u) & 0xFFu) + 0.5f) * vs1,
                       o1.z + ((float)((e >> 16u) & 0xFFu) + 0.5f) * vs1);
    const v3 pw = r1.pos + qrot(r1.rot, local);
    float pe, su, sv;
    uint32_t fb;
    if (s2_vox) {
      const float vs2 = s2.vs;
      const v3 l2 = qrot_inv(r2.rot, pw - r2.pos);
      su = axk == 0u ? l2.y : l2.x;
      sv = axk == 2u ? l2.y : l2.z;
      const v3 g = vdiv(l2 - o2, vs2);
      const int cx = (int)__builtin_floorf(g.x), cy = (int)__builtin_floorf(g.y), cz = (int)__builtin_floorf(g.z);
      if (!vox_bit(d, s2, cx, cy, cz)) {
        // hover band: the occupied cell just behind along the pass axis
        const int ix = axk == 0u ? cx - isg : cx, iy = axk == 1u ? cy - isg : cy, iz = axk == 2u ? cz - isg : cz;
        if (!vox_bit(d, s2, ix, iy, iz)) continue;
        const float go_a = cmp(o2, (int)axk);
        const float ci_a = (float)vox_ci(ix, iy, iz, axk);
        const float dh = cmp(l2, (int)axk) - (go_a + (ci_a + 0.5f) * vs2);
        const float ovh = (0.5f * vs2 + cmp(sup, (int)axk)) - sgn_pass * dh;
        pe = -ovh;
        fb = (uint32_t)(ix & 0xFF) | ((uint32_t)(iy & 0xFF) << 8u) | ((uint32_t)(iz & 0xFF) << 16u);
      } else {
        // cell-face contact: the smallest outward-feasible exit bounds the
        // pass axis's dominance band
        const v3 cc_l = o2 + (V((float)cx, (float)cy, (float)cz) + V(0.5f, 0.5f, 0.5f)) * vs2;
        const v3 dl = l2 - cc_l;
        float best = 1e9f;
#pragma unroll
        for (uint32_t ax = 0; ax < 3u; ++ax) {
          const float di = cmp(dl, (int)ax);
          const int sg = di >= 0.0f ? 1 : -1;
          const int nx = ax == 0u ? cx + sg : cx, ny = ax == 1u ? cy + sg : cy, nz = ax == 2u ? cz + sg : cz;
          if (vox_bit(d, s2, nx, ny, nz)) continue;
          const float ov = (0.5f * vs2 + cmp(sup, (int)ax)) - fab(di);
          if (ov < best) best = ov;
        }
        if (best > 0.5e9f) {
          red.interior += 1u;
          continue;
        }
        const int px = axk == 0u ? cx + isg : cx, py = axk == 1u ? cy + isg : cy, pz = axk == 2u ? cz + isg : cz;
        if (vox_bit(d, s2, px, py, pz)) continue;
        const float ovp = (0.5f * vs2 + cmp(sup, (int)axk)) - sgn_pass * cmp(dl, (int)axk);
        if (ovp > best + 0.15f * vs2) continue;
        pe = -ovp;
        fb = (uint32_t)(cx & 0xFF) | ((uint32_t)(cy & 0xFF) << 8u) | ((uint32_t)(cz & 0xFF) << 16u);
      }
    } else {
      // an OBB partner: the cell against the box faces
      v3 l2 = qrot_inv(r2.rot, pw - r2.pos);
      const v3 half = (r2.mx - r2.mn) * 0.5f;
      const v3 center_off = (r2.mx + r2.mn) * 0.5f;
      l2 = l2 - center_off;
      su = axk == 0u ? l2.y : l2.x;
      sv = axk == 2u ? l2.y : l2.z;
      const v3 q = vabs(l2) - half;
      if (q.x > sup.x || q.y > sup.y || q.z > sup.z) continue;
      const v3 qs = q - sup;
      const float best_q = fmx(qs.x, fmx(qs.y, qs.z));
      const float q_pass = (sgn_pass * cmp(l2, (int)axk) - cmp(half, (int)axk)) - cmp(sup, (int)axk);
      if (q_pass < best_q - 0.15f * vs1) continue;
      pe = q_pass;
      fb = axk;
    }
    if (pe >= 0.02f) continue;
    pe = fmx(pe, s2_vox ? -0.5f * vs1 : -vs1);
    vox_add(red, pw, pe, s2_vox ? n_pass : n_pass_box, e & 0xFFFFFFu, fb, su, sv, 0.5f * vs1);
  }
}

// voxel_obb_prefilter (collision_detection.slang:1268-1294)
BB_DI bool vox_prefilter(const BodyView& a, const BodyView& b) {
  const v3 ha = (a.mx - a.mn) * 0.5f, hb = (b.mx - b.mn) * 0.5f;
  const v3 ca = a.pos + qrot(a.rot, (a.mx + a.mn) * 0.5f);
  const v3 cb = b.pos + qrot(b.rot, (b.mx + b.mn) * 0.5f);
  const v3 dd = cb - ca;
  const v3 ax[3] = {qrot(a.rot, V(1, 0, 0)), qrot(a.rot, V(0, 1, 0)), qrot(a.rot, V(0, 0, 1))};
  const v3 bx[3] = {qrot(b.rot, V(1, 0, 0)), qrot(b.rot, V(0, 1, 0)), qrot(b.rot, V(0, 0, 1))};
  const float haa[3] = {ha.x, ha.y, ha.z}, hbb[3] = {hb.x, hb.y, hb.z};
  bool sep = false;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float rb = (hbb[0] * fab(dot(ax[i], bx[0])) + hbb[1] * fab(dot(ax[i], bx[1]))) + hbb[2] * fab(dot(ax[i], bx[2]));
    const float ra = (haa[0] * fab(dot(bx[i], ax[0])) + haa[1] * fab(dot(bx[i], ax[1]))) + haa[2] * fab(dot(bx[i], ax[2]));
    sep = sep || fab(dot(dd, ax[i])) > (haa[i] + rb) + 0.01f || fab(dot(dd, bx[i])) > (hbb[i] + ra) + 0.01f;
  }
  return !sep;
}

// voxel_corner_supported (collision_detection.slang:1302-1322)
BB_DI bool vox_corner_ok(const Dev& d, v3 p, v3 n_sep, const BodyView& r2, const VoxShape& s2, bool s2_vox,
                         uint32_t axk) {
  if (s2_vox) {
    const float vs2 = s2.vs;
    const v3 l2 = qrot_inv(r2.rot, (p - n_sep * (0.3f * vs2)) - r2.pos);
    const v3 g = vdiv(l2 - vox_origin(s2), vs2);
    return vox_bit(d, s2, (int)__builtin_floorf(g.x), (int)__builtin_floorf(g.y), (int)__builtin_floorf(g.z));
  }
  const v3 half = (r2.mx - r2.mn) * 0.5f;
  const v3 co = (r2.mx + r2.mn) * 0.5f;
  const v3 l2 = qrot_inv(r2.rot, p - r2.pos) - co;
  const float lu = axk == 0u ? l2.y : l2.x, lv = axk == 2u ? l2.y : l2.z;
  const float hu = axk == 0u ? half.y : half.x, hv = axk == 2u ? half.y : half.z;
  return fab(lu) <= hu + 0.01f && fab(lv) <= hv + 0.01f;
}

// voxel_emit_manifold (collision_detection.slang:1324-1432): kept cells in
// slot order; one cell gives its four face corners, two cells two each,
// more their centres; unsupported corners retreat toward the cell centre.
BB_DI int vox_emit(const Dev& d, const VoxRed& red, bool flip, v3 n_sep, float vs, const BodyView& r2,
                   const VoxShape& s2, bool s2_vox, uint32_t axk, CSink& sk, v3& n_out) {
  if (red.samples == 0u) return 0;
  n_out = flip ? n_sep : -n_sep;
  uint32_t keepm = 0u;
  float align[4];
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    align[o] = dot(red.nrm[o], n_sep);
    bool ok = red.pen[o] < 0.02f && align[o] > 0.1f;
#pragma unroll
    for (int q = 0; q < o; ++q)
      if (((keepm >> q) & 1u) && red.fa[q] == red.fa[o]) ok = false;
    if (ok) keepm |= 1u << o;
  }
  const uint32_t kn = (uint32_t)__popc(keepm);
  const uint32_t reps = kn == 1u ? 4u : (kn == 2u ? 2u : 1u);
  const v3 t1 = vnorm(cross(n_sep, fab(n_sep.y) < 0.9f ? V(0.0f, 1.0f, 0.0f) : V(1.0f, 0.0f, 0.0f)));
  const v3 t2 = cross(n_sep, t1);
  const v3 da = ((t1 + t2) * 0.318f) * vs;
  const v3 db = ((t1 - t2) * 0.318f) * vs;
  int out_n = 0;
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    if (!((keepm >> o) & 1u)) continue;
    const uint32_t fa = red.fa[o];
    const uint32_t vpar = (fa ^ (fa >> 8u) ^ (fa >> 16u)) & 1u;
    v3 em[4];
    uint32_t en = 0u;
#pragma unroll
    for (uint32_t c = 0; c < 4u; ++c) {
      if (c >= reps) break;
      v3 p = red.pos[o];
      if (reps > 1u) {
        const v3 offv = ((c < 2u) == (vpar == 0u)) ? da : db;
        const v3 off = offv * ((c & 1u) == 0u ? 1.0f : -1.0f);
        p = red.pos[o] + off;
        if (!vox_corner_ok(d, p, n_sep, r2, s2, s2_vox, axk)) {
          p = red.pos[o] + off * 0.5f;
          if (!vox_corner_ok(d, p, n_sep, r2, s2, s2_vox, axk)) p = red.pos[o];
        }
      }
      bool dup = false;
#pragma unroll
      for (uint32_t q = 0; q < 4u; ++q)
        if (q < en) {
          const v3 dp = p - em[q];
          if (dot(dp, dp) < 1.0e-12f) dup = true;
        }
      if (dup) continue;
#pragma unroll
      for (uint32_t q = 0; q < 4u; ++q)
        if (q == en) em[q] = p;
      ++en;
      sk.put(out_n, p, red.pen[o] * align[o], flip ? 1u : 0u, red.fb[o], fa | (c << 24u), fa | (c << 24u));
      ++out_n;
    }
  }
  return out_n;
}

// voxel_axis_collision_detection (collision_detection.slang:1446-1512):
// the manifold of code `code`, key 90 + code.
BB_DI int vox_collide(const Dev& d, const BodyView& A, const BodyView& B, uint32_t code, CSink& sk, int& key,
                      v3& n) {
  key = (int)(90u + code);
  if (!vox_prefilter(A, B)) return 0;
  const bool a_vox = A.kind == BB_SHAPE_VOXEL, b_vox = B.kind == BB_SHAPE_VOXEL;
  const VoxShape sA = d.vshape[a_vox ? A.vid - 1u : 0u];
  const VoxShape sB = d.vshape[b_vox ? B.vid - 1u : 0u];
  const bool flip = !a_vox || (b_vox && sB.surf_count < sA.surf_count);
  const BodyView& r1 = flip ? B : A;
  const BodyView& r2 = flip ? A : B;
  const VoxShape s1 = flip ? sB : sA;
  const VoxShape s2 = flip ? sA : sB;
  const bool s2_vox = flip ? a_vox : b_vox;
  VoxRed red;
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    red.pen[o] = 1.0e9f;
    red.fa[o] = 0u;
    red.fb[o] = 0u;
    red.score[o] = -1.0e30f;
    red.pos[o] = V(0.0f, 0.0f, 0.0f);
    red.nrm[o] = V(0.0f, 0.0f, 0.0f);
  }
  red.samples = 0u;
  red.interior = 0u;
  vox_sample(d, r1, s1, r2, s2, s2_vox, red, code);
  if (red.samples == 0u) {
    if (code == 0u && red.interior > 0u) {
      // deep-interpenetration rescue: one centre-to-centre contact
      const v3 dd = r1.pos - r2.pos;
      const v3 n_res = dot(dd, dd) > 1e-8f ? vnorm(dd) : V(0.0f, 1.0f, 0.0f);
      n = flip ? n_res : -n_res;
      sk.put(0, (r1.pos + r2.pos) * 0.5f, -0.5f * s1.vs, 0xFFFFFFu, 0xFFFFFFu, 0xFFFFFFu, 0xFFFFFFu);
      return 1;
    }
    return 0;
  }
  const uint32_t ax = code / 2u;
  const float sgn = (code & 1u) != 0u ? 1.0f : -1.0f;
  const v3 n_world = qrot(r2.rot, V(ax == 0u ? sgn : 0.0f, ax == 1u ? sgn : 0.0f, ax == 2u ? sgn : 0.0f));
  return vox_emit(d, red, flip, n_world, s1.vs, r2, s2, s2_vox, ax, sk, n);
}



// IslandBuilder::link_bodies (src/shaders/simulation/islands.slang:153-213):
// lock-free union-find, links always point to the lower index, so every
// component's root is its minimum active index whatever the timing.
// A root as seen now: the links are read with device-scope atomic loads,
// not read-modify-writes (an atomicAdd of 0 per hop serialises every thread
// of a world on the popular root's address).  A stale root only makes the
// CAS in uf_link fail and walk again; the final roots (the minimum index of
// each set) do not depend on the interleaving.
BB_DI uint32_t uf_root(uint32_t* link, uint32_t x) {
  for (;;) {
    const uint32_t y = __hip_atomic_load(&link[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (y == x) return x;
    x = y;
  }
}
BB_DI void uf_link(uint32_t* link, uint32_t f, uint32_t s) {
  uint32_t fl = f, sl = s;
  for (;;) {
    fl = uf_root(link, fl);
    sl = uf_root(link, sl);
    if (fl != sl) {
      if (fl < sl) {
        if (atomicCAS(&link[sl], sl, fl) != sl) continue;
      } else {
        if (atomicCAS(&link[fl], fl, sl) != fl) continue;
      }
    }
    uint32_t lo = fl < sl ? fl : sl;
    atomicMin(&link[f], lo);
    atomicMin(&link[s], lo);
    break;
  }
}

// Reads the previous step's manifold of this pair and carries its impulses
// over to the new contacts already written at slot*8 (find_old_collision +
// Manifold::update_from_old, collision_detection.slang:914-963,
// src/shaders/extensions.slang:305-473).  Returns the bitmask of new
// contacts that inherited state and sets `error`.  Old and new contact
// positions and feature ids are held in registers (all indices are
// compile-time constants), so the matching loops issue no memory traffic.
BB_DI uint32_t warm_start(const Dev& d, uint32_t os, bool swapped, int key, int count, v3 n, uint32_t slot,
                          float deep_new, int32_t& error) {
  const int4 oh = d.man_prev.hdr[os];
  if (oh.w <= 0) return 0;
  const float sgn = swapped ? -1.0f : 1.0f;
  const bool cross_key = key != oh.z || swapped;
  const v3 on = ld3(d.man_prev.nrm[os]);
  const int ocount = oh.w;
  if (cross_key && !(dot(n, on * sgn) > 0.7f)) return 0;
  const uint32_t nb = slot * 8, ob = os * 8;
  float4 op[8], np[8];  // position + penetration
  uint2 of[8], nf[8];   // feature ids in_inc, out_inc
  static_for<8>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    float4 p4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint2 f2 = make_uint2(0u, 0u);
    if (j < ocount) {
      p4 = d.man_prev.cpos[ob + j];
      const uint4 f = d.man_prev.cfp[ob + j];
      f2 = make_uint2(f.z, f.w);
    }
    op[j] = p4;
    of[j] = f2;
    float4 q4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint2 g2 = make_uint2(0u, 0u);
    if (j < count) {
      q4 = d.man_cur.cpos[nb + j];
      const uint4 f = d.man_cur.cfp[nb + j];
      g2 = make_uint2(f.z, f.w);
    }
    np[j] = q4;
    nf[j] = g2;
  });
  // match map: byte i = old index j (pass 1), 0x80 | j (pass 2), 0xFF none
  uint64_t map = ~0ull;
  uint32_t used_old = 0, matched_new = 0, snap = 0;
  // pass 1: same feature pair within 0.1 m (same key, not swapped)
  if (!cross_key) {
    static_for<8>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if (i < count) {
        bool found = false;
        static_for<8>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          if (j < ocount && !found && !((used_old >> j) & 1u)) {
            const v3 opos = ld3(op[j]), npos = ld3(np[i]);
            const v3 dp = opos - npos;
            const bool feat = of[j].x == nf[i].x && of[j].y == nf[i].y;
            const bool close = dot(dp, dp) < 0.1f * 0.1f;
            if (feat && close) {
              found = true;
              used_old |= 1u << j;
              matched_new |= 1u << i;
              map = (map & ~(0xFFull << (8 * i))) | ((uint64_t)j << (8 * i));
              const v3 d2 = npos - opos;
              if (dot(d2, d2) < 0.01f) snap |= 1u << i;
            }
          }
        });
      }
    });
  }
  // pass 2: nearest remaining pairs within 0.35 m, one pair per round
  for (int round = 0; round < count; ++round) {
    float best = 0.1225f;
    int bi = -1, bj = -1;
    static_for<8>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if (i < count && !((matched_new >> i) & 1u)) {
        static_for<8>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          if (j < ocount && !((used_old >> j) & 1u)) {
            const v3 dp = ld3(np[i]) - ld3(op[j]);
            const float dd2 = dot(dp, dp);
            if (dd2 < best) { best = dd2; bi = i; bj = j; }
          }
        });
      }
    });
    if (bi < 0) break;
    map = (map & ~(0xFFull << (8 * bi))) | ((uint64_t)(0x80 | bj) << (8 * bi));
    matched_new |= 1u << bi;
    used_old |= 1u << bj;
  }
  // carry the matched contacts' state over
  static_for<8>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const uint32_t m = (uint32_t)(map >> (8 * i)) & 0xFFu;
    if (i < count && m != 0xFFu) {
      const bool p2 = (m & 0x80u) != 0;
      const uint32_t oc = ob + (m & 0x7Fu), c = nb + i;
      const float4 ol = d.man_prev.clam[oc], om = d.man_prev.cmas[oc];
      const float4 oa = d.man_prev.cana[oc], obb = d.man_prev.canb[oc];
      const uint32_t ost = __float_as_uint(obb.w);
      d.man_cur.clam[c] = make_float4(ol.x, p2 ? ol.y * sgn : ol.y, ol.z, ol.w);
      d.man_cur.cmas[c] = make_float4(0.0f, om.y, om.z, 0.0f);
      if (ost != 0) {
        const bool sw = p2 && swapped;
        d.man_cur.cana[c] = st3(sw ? ld3(obb) : ld3(oa), 0.0f);
        d.man_cur.canb[c] = st3(sw ? ld3(oa) : ld3(obb), __uint_as_float(ost));
      } else {
        d.man_cur.cana[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        d.man_cur.canb[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
      if (p2) d.man_cur.cfp[c] = d.man_prev.cfp[oc];
      if ((snap >> i) & 1u) {
        const float4 o4 = d.man_prev.cpos[oc];
        d.man_cur.cpos[c] = make_float4(o4.x, o4.y, o4.z, np[i].w);
      }
      const uint32_t ows = d.man_prev.cws[oc];
      d.man_cur.cws[c] = ows + 1u < 255u ? ows + 1u : 255u;
    }
  });
  float deep_old = 0.0f;
  static_for<8>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j < ocount) deep_old = fmn(deep_old, op[j].w);
  });
  error = (int32_t)fcl((deep_new - deep_old) * 10000.0f, -32000.0f, 32000.0f);
  return matched_new;
}

// Narrow phase, part 1: contact detection.  Writes the slot header {a, b,
// key, count} (count 0: no manifold), the normal, the body centres (the
// first one's w holds the deepest penetration, min(0, pen)) and the
// contacts' positions, penetrations and feature pairs.
// kVox: the instance for worlds with voxel bodies (k_narrow_vox); the other
// instance skips them, so the box and sphere kernels carry no voxel code.
template <bool kVox, class T>
BB_DI void s_narrow(const Dev& d, uint32_t w, const T& t) {
  const uint32_t tid = t.rank(), NT = t.size();
  const WorldInfo W = d.worlds[w];
  if (((W.mode & kModeVoxel) != 0) != kVox) return;
  const uint32_t off = W.body_off, P = d.npairs_cur[w];
  const uint2* pairs = d.pairs_cur + W.pair_off;
  for (uint32_t p = tid; p < P; p += NT) {
    const uint32_t slot = W.pair_off + p;
    const uint2 pr = pairs[p];
    const uint32_t ga = off + pr.x, gb = off + pr.y;
    const uint32_t fa = d.cur.flags[ga], fb = d.cur.flags[gb];
    const bool da = (fa & BB_BODY_DYNAMIC) != 0, db = (fb & BB_BODY_DYNAMIC) != 0;
    if (!da && !db) {
      d.man_cur.hdr[slot] = make_int4((int)pr.x, (int)pr.y, -1, 0);
      continue;
    }
    const BodyView A = body_view(d, off, pr.x), B = body_view(d, off, pr.y);
    CSink sk{d.man_cur.cpos, d.man_cur.cfp, slot * 8, false, 0.0f};
    int key, count;
    v3 n;
    if (kVox && (A.kind == BB_SHAPE_VOXEL || B.kind == BB_SHAPE_VOXEL)) {
      // the pair's contact-face code is its slot's rank among the pair's six
      uint32_t code = 0;
      while (code < 5u && code < p && pairs[p - code - 1].x == pr.x && pairs[p - code - 1].y == pr.y) ++code;
      count = vox_collide(d, A, B, code, sk, key, n);
    } else if (A.kind == BB_SHAPE_BOX && B.kind == BB_SHAPE_BOX) {
      sk.capped = true;
      count = obb_obb(A, B, sk, key, n);
    } else if (A.kind == BB_SHAPE_SPHERE && B.kind == BB_SHAPE_SPHERE) {
      count = sphere_sphere(A, B, sk, key, n);
    } else {
      count = box_sphere(A, B, sk, key, n);
    }
    if (count == 0) {
      d.man_cur.hdr[slot] = make_int4((int)pr.x, (int)pr.y, -1, 0);
      continue;
    }
    d.man_cur.hdr[slot] = make_int4((int)pr.x, (int)pr.y, key, count);
    d.man_cur.nrm[slot] = st3(n, 0.0f);
    d.man_cur.anca[slot] = st3(A.pos, sk.deep);  // w: deepest penetration, min(0, pen)
    d.man_cur.ancb[slot] = st3(B.pos, 0.0f);
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_narrow_vox(Dev d) {
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{nullptr, nullptr};
  s_narrow<true>(d, w, t);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_narrow(Dev d) {
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{nullptr, nullptr};
  s_narrow<false>(d, w, t);
}

// Narrow phase, part 2, per manifold: tangent basis, warm start from the
// previous step's manifold of the pair, fresh contacts' zero impulses,
// counters, contact degrees and the union-find link.
template <class T>
BB_DI void s_contacts(const Dev& d, uint32_t w, const T& t) {
  const uint32_t tid = t.rank(), NT = t.size();
  const WorldInfo W = d.worlds[w];
  const uint32_t off = W.body_off, P = d.npairs_cur[w];
  const bool warm = (W.flags & BB_SIM_WARM_STARTING) != 0 && (W.mode & kModeDropPrev) == 0;
  const uint32_t Pprev = d.npairs_prev[w];
  const uint2* ppairs = d.pairs_prev + W.pair_off;
  uint32_t fresh = 0, pen_mm = 0, ccount = 0, mcount = 0;
  for (uint32_t p = tid; p < P; p += NT) {
    const uint32_t slot = W.pair_off + p;
    const int4 h = d.man_cur.hdr[slot];
    const int count = h.w, key = h.z;
    if (count <= 0) continue;
    const uint32_t ga = off + (uint32_t)h.x, gb = off + (uint32_t)h.y;
    const v3 n = ld3(d.man_cur.nrm[slot]);
    // Manifold::compute_basis (src/shaders/extensions.slang:262-272)
    v3 t0, t1;
    if (fab(n.x) > fab(n.y)) t0 = V(n.z, 0.0f, -n.x);
    else t0 = V(0.0f, -n.z, n.y);
    t0 = vnorm(t0);
    t1 = cross(t0, n);
    const float deep = d.man_cur.anca[slot].w;
    int32_t error = 0x7FFFFFFF;
    uint32_t matched = 0;
    // warm start: the previous step's manifold of this pair, found through
    // the previous pair list's per-row offsets (lists ascend by partner)
    if (warm && Pprev > 0) {
      const uint32_t pa = d.row_prev[ga], pb = d.row_prev[gb];
      if (pa < W.prev_n && pb < W.prev_n) {
        const uint32_t lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
        uint32_t L0 = d.pstart_prev[off + lo];
        const uint32_t L1 = lo + 1 < W.prev_n ? d.pstart_prev[off + lo + 1] : Pprev;
        const uint32_t end = L1 < Pprev ? L1 : Pprev;
        // the first entry of row lo's list whose partner is not below hi: a
        // binary search (a ground's list holds hundreds of partners, and a
        // walk along it per pair was the stage's tail)
        for (uint32_t R = end; L0 < R;) {
          const uint32_t mid = L0 + (R - L0) / 2;
          if (ppairs[mid].y < hi) L0 = mid + 1;
          else R = mid;
        }
        if (key >= 90 && key <= 95) {
          // voxel manifolds (keys 90..95) match only the same ordered pair
          // under the same key (find_old_collision's exact_only,
          // collision_detection.slang:914-963): the old pair's slot of the
          // same contact-face code
          const uint32_t os = L0 + (uint32_t)(key - 90);
          if (pa == lo && os < end && ppairs[os].y == hi)
            matched = warm_start(d, W.pair_off + os, false, key, count, n, slot, deep, error);
        } else if (L0 < end && ppairs[L0].y == hi) {
          matched = warm_start(d, W.pair_off + L0, pa != lo, key, count, n, slot, deep, error);
        }
      }
    }
    // fresh contacts start from zero impulses
    for (int i = 0; i < count; ++i) {
      if (matched & (1u << i)) continue;
      const uint32_t c = slot * 8 + i;
      const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      d.man_cur.clam[c] = z;
      d.man_cur.cmas[c] = z;
      d.man_cur.cana[c] = z;
      d.man_cur.canb[c] = z;
      d.man_cur.cws[c] = 0u;
      ++fresh;
    }
    d.man_cur.nrm[slot] = st3(n, __int_as_float(error));
    d.man_cur.t0[slot] = st3(t0, 0.0f);
    d.man_cur.t1[slot] = st3(t1, 0.0f);
    if (deep < 0.0f) {
      const uint32_t mm = (uint32_t)fmn(-deep * 1000.0f, 1.0e9f);
      pen_mm = pen_mm > mm ? pen_mm : mm;
    }
    ccount += (uint32_t)count;
    ++mcount;
    atomicAdd(&d.deg[ga], 1u);
    atomicAdd(&d.deg[gb], 1u);
    const bool da = (d.cur.flags[ga] & BB_BODY_DYNAMIC) != 0, db = (d.cur.flags[gb] & BB_BODY_DYNAMIC) != 0;
    const uint32_t ra = da ? d.row_rank[ga] : kMAX;
    const uint32_t rb = db ? d.row_rank[gb] : kMAX;
    if (ra < W.a && rb < W.a) uf_link(d.link + W.act_off, ra, rb);
  }
  stat_add(d, w, ST_FRESH, fresh);
  stat_add(d, w, ST_C, ccount);
  stat_add(d, w, ST_M, mcount);
  if (pen_mm) stat_max(d, w, ST_PEN, pen_mm);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_contacts(Dev d) {
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{nullptr, nullptr};
  s_contacts(d, w, t);
}

// ================================================================= k_post
// RigidBody::advect / advance (src/shaders/extensions.slang:127-139); the
// damping factor exp(-dt*0.1) is evaluated once on the host.
BB_DI void advect_row(const Dev& d, uint32_t off, uint32_t i, const WorldInfo& W) {
  const uint32_t g = off + i;
  const uint32_t fl = d.cur.flags[g];
  if (!((fl & BB_BODY_DYNAMIC) && !(fl & BB_BODY_SLEEPING))) return;
  v3 v = ld3(d.cur.vel[g]), om = ld3(d.cur.omg[g]);
  if (fl & BB_BODY_GRAVITY) {
    const uint32_t id = d.cur.id[g];
    float inv_mass = d.props.mass[off + id].y;
    v3 impulse = V(0.0f, W.gravity, 0.0f) * W.dt;
    v3 lin = impulse * inv_mass;
    v = v + lin;
    v3 p = ld3(d.cur.pos[g]);
    v3 rr = p - p;
    v3 torque = cross(rr, impulse);
    mat3 I = mrows(ld3(d.props.ii0[off + id]), ld3(d.props.ii1[off + id]), ld3(d.props.ii2[off + id]));
    om = om + mulvm(torque, I);
  }
  v = v * W.damp;
  om = om * W.damp;
  d.cur.vel[g] = st3(v, 0.0f);
  d.cur.omg[g] = st3(om, 0.0f);
}

BB_DI uint32_t gc_hash(uint32_t x) {  // coloring.slang:114-122
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
BB_DI uint32_t gc_priority(uint32_t pid, uint32_t round) {  // coloring.slang:132-138
  uint32_t h = gc_hash(pid + round * 0x85ebca6bu) & 0xFFFu;
  if (h == 0u) h = 1u;
  return (h << 20) | (pid & 0xFFFFFu);
}

template <class T>
BB_DI void iw_refresh(const Dev& d, const WorldInfo& W, const T& t);
BB_DI bool pgs_solvable(const Dev& d, uint32_t off, int4 h);
BB_DI void pgs_compact(const Dev& d, const WorldInfo& W, uint32_t k, uint32_t slot);

// Per-row adjacency of a world in CSR form: the manifold slots of each
// dynamic row in ascending order at adj[2*pair_off + adj_off[row]], deg[row]
// entries (deg also counts a row's static partners' manifolds; static rows'
// lists are left unfilled).  `carry` is a team-shared scalar.
template <class T>
BB_DI void build_adj(const Dev& d, const WorldInfo& W, uint32_t P, const T& t, uint32_t& carry) {
  const uint32_t tid = t.rank(), NT = t.size();
  const uint32_t N = W.n, off = W.body_off;
  (void)carry;
  (void)scan_range(t, N, [&](uint32_t i) { return d.deg[off + i]; },
                   [&](uint32_t i, uint32_t ex) {
                     d.adj_off[off + i] = ex;
                     d.pcount[off + i] = 0;  // fill cursor
                   });
  t.sync();
  uint32_t* adj = d.adj + 2 * W.pair_off;
  for (uint32_t p = tid; p < P; p += NT) {
    const uint32_t slot = W.pair_off + p;
    int4 h = d.man_cur.hdr[slot];
    if (h.w <= 0) continue;
    const uint32_t ga = off + h.x, gb = off + h.y;
    if (d.cur.flags[ga] & BB_BODY_DYNAMIC) adj[d.adj_off[ga] + atomicAdd(&d.pcount[ga], 1u)] = slot;
    if (d.cur.flags[gb] & BB_BODY_DYNAMIC) adj[d.adj_off[gb] + atomicAdd(&d.pcount[gb], 1u)] = slot;
  }
  t.sync();
  for (uint32_t i = tid; i < N; i += NT) {
    const uint32_t g = off + i;
    if (!(d.cur.flags[g] & BB_BODY_DYNAMIC)) continue;
    uint32_t* l = adj + d.adj_off[g];
    const uint32_t cnt = d.deg[g];
    for (uint32_t a = 1; a < cnt; ++a) {
      uint32_t x = l[a], b = a;
      while (b > 0 && l[b - 1] > x) {
        l[b] = l[b - 1];
        --b;
      }
      l[b] = x;
    }
  }
  t.sync();
}

// Exact dependency levels of the per-island path (per-world kernels): one
// thread walks the slots in order with each row's next free level in LDS;
// the slots are staged kLvlChunk at a time so the walk touches LDS only.
// Levels are below the slot count, which fits 16 bits for worlds of at most
// kLvlRows rows whose slot count stays below 65,536.
static constexpr uint32_t kLvlRows = 4096, kLvlChunk = 256;
struct LevelLds {
  uint16_t lvl[kLvlRows];  // 1 + level of the last manifold on the row (0: none)
  uint32_t sa[kLvlChunk], sb[kLvlChunk];  // staged rows (kLvlSkip: not solved)
  uint16_t out[kLvlChunk];
  uint32_t top;
};
static constexpr uint32_t kLvlSkip = 0xFFFFFFFFu, kLvlStatic = 0x80000000u;

struct PostShared {
  uint32_t carry;
  uint32_t flag;
  uint32_t bcount[kBuckets];
  uint32_t bfill[kBuckets];
};

// Buckets of the solvable manifolds (at least one awake dynamic body; the
// others are skipped by every sweep, RB_sim.slang:1424) by bkey(p): a colour,
// or a capped dependency level (kColors and above: the overflow bucket;
// kMAX - 1: no manifold).  Order inside a bucket is free (no shared dynamic
// body); the overflow bucket is filled in ascending slot order for its
// serial solve.  Then the world inverse inertia and the solver rows.
template <class T, class K>
BB_DI void build_rows(const Dev& d, const WorldInfo& W, uint32_t w, const T& t, PostShared& L, K&& bkey) {
  const uint32_t tid = t.rank(), NT = t.size();
  const uint32_t off = W.body_off, P = d.npairs_cur[w];
  if (tid < kBuckets) L.bcount[tid] = 0;
  t.sync();
  for (uint32_t p = tid; p < P; p += NT) {
    uint32_t cc = bkey(p);
    if (cc == kMAX - 1) continue;
    int4 h = d.man_cur.hdr[W.pair_off + p];
    if (!pgs_solvable(d, off, h)) continue;
    atomicAdd(&L.bcount[cc < kColors ? cc : kColors], 1u);
  }
  t.sync();
  if (tid == 0) {
    uint32_t sum = 0;
    uint32_t* bo = d.bucket_off + w * (kBuckets + 1);
    for (uint32_t c2 = 0; c2 < kBuckets; ++c2) {
      bo[c2] = sum;
      L.bfill[c2] = sum;
      sum += L.bcount[c2];
    }
    bo[kBuckets] = sum;
  }
  t.sync();
  uint32_t* bucket = d.bucket + W.pair_off;
  for (uint32_t p = tid; p < P; p += NT) {
    uint32_t cc = bkey(p);
    if (cc >= kColors) continue;
    int4 h = d.man_cur.hdr[W.pair_off + p];
    if (!pgs_solvable(d, off, h)) continue;
    bucket[atomicAdd(&L.bfill[cc], 1u)] = W.pair_off + p;
  }
  if (L.bcount[kColors]) {
    if (tid == 0) L.carry = L.bfill[kColors];
    t.sync();
    for (uint32_t c0 = 0; c0 < P; c0 += NT) {
      uint32_t p = c0 + tid;
      uint32_t f = 0;
      const uint32_t cc = p < P ? bkey(p) : kMAX - 1;
      if (cc != kMAX - 1 && cc >= kColors)
        f = pgs_solvable(d, off, d.man_cur.hdr[W.pair_off + p]) ? 1u : 0u;
      uint32_t tot;
      uint32_t ex = t.scan(f, tot);
      if (f) bucket[L.carry + ex] = W.pair_off + p;
      t.sync();
      if (tid == 0) L.carry += tot;
      t.sync();
    }
  }
  // ---- world inverse inertia of every row, then the solver rows
  iw_refresh(d, W, t);
  t.sync();
  const uint32_t total = d.bucket_off[w * (kBuckets + 1) + kBuckets];
  for (uint32_t k = tid; k < total; k += NT) pgs_compact(d, W, W.pair_off + k, bucket[k]);
}

template <class T>
BB_DI void s_post(const Dev& d, uint32_t w, const T& t, PostShared& L) {
  const uint32_t tid = t.rank(), NT = t.size();
  const WorldInfo W = d.worlds[w];
  const uint32_t N = W.n, off = W.body_off, A = W.a;
  const uint32_t P = d.npairs_cur[w];
  const uint32_t flags = W.flags;
  const bool pgs = W.solver != BB_SOLVER_AVBD;
  const uint32_t* link = d.link + W.act_off;
  // ---- advect, PGS family only (RB_sim.slang:518-554).  With the parity
  // quirk the dispatch covers ceil(A/32)*32 threads indexing rows < N.
  if (pgs && (flags & BB_SIM_ADVECTION)) {
    uint32_t lim = N;
    if (W.quirks & BB_QUIRK_ACTIVE_DISPATCH_ROWS) {
      uint32_t r = ((A + 31u) / 32u) * 32u;
      lim = r < N ? r : N;
    }
    for (uint32_t i = tid; i < lim; i += NT) advect_row(d, off, i, W);
  }
  // ---- islands (islands.slang:215-302): root = min active index; slots in
  // ascending root order
  for (uint32_t k = tid; k < A; k += NT) {
    uint32_t x = k;
    while (link[x] != x) x = link[x];
    d.ascratch[W.act_off + k] = x;
  }
  t.sync();
  const uint32_t* root = d.ascratch + W.act_off;
  const uint32_t islands = scan_range(t, A, [&](uint32_t k) { return root[k] == k ? 1u : 0u; },
                                      [&](uint32_t k, uint32_t ex) {
                                        if (root[k] == k) d.poff[off + k] = ex;  // slot of root k
                                      });
  for (uint32_t i = tid; i < N; i += NT) {
    d.island[off + i] = kMAX;
    d.islot[off + i] = kMAX;
  }
  t.sync();
  for (uint32_t k = tid; k < A; k += NT) {
    uint32_t id = d.props.active_ids[W.act_off + k];
    uint32_t row = d.entries_cur[off + id];
    uint32_t root = d.ascratch[W.act_off + k];
    d.island[off + row] = root;
    d.islot[off + row] = d.poff[off + root];
  }
  // contact islands: islands owning at least one manifold
  for (uint32_t k = tid; k < A; k += NT) d.pcount[off + k] = 0;
  t.sync();
  for (uint32_t p = tid; p < P; p += NT) {
    int4 h = d.man_cur.hdr[W.pair_off + p];
    if (h.w <= 0) continue;
    uint32_t ia = d.row_rank[off + h.x];
    uint32_t ib = d.row_rank[off + h.y];
    uint32_t mn = ia < ib ? ia : ib;
    if (mn < A) d.pcount[off + d.ascratch[W.act_off + mn]] = 1;
  }
  t.sync();
  {
    uint32_t cnt = 0;
    for (uint32_t k = tid; k < A; k += NT) cnt += d.pcount[off + k];
    if (tid == 0) L.flag = 0;
    t.sync();
    atomicAdd(&L.flag, cnt);
    t.sync();
    if (tid == 0) {
      d.stats[w * ST_COUNT + ST_ISLANDS] = islands;
      d.stats[w * ST_COUNT + ST_CISLANDS] = L.flag;
    }
    t.sync();
  }
  // ---- sleeping (RB_sim.slang:556-737)
  const bool sleep_on = (flags & BB_SIM_SLEEPING_ENABLED) != 0;
  for (uint32_t i = tid; i < N; i += NT) {
    const uint32_t g = off + i;
    uint32_t fl = d.cur.flags[g];
    if (!(fl & BB_BODY_DYNAMIC)) continue;
    if (!sleep_on) {
      d.cur.sleep[g] = 0;
      d.cur.flags[g] = fl & ~(uint32_t)BB_BODY_SLEEPING;
      continue;
    }
    uint32_t st = d.cur.sleep[g];
    if (d.deg[g] == 0 && (fl & BB_BODY_SLEEPING)) {
      uint32_t t = st & 0x7FFFFFFFu;
      d.cur.sleep[g] = t >= 10u ? t - 10u : 0u;
      continue;
    }
    v3 v = ld3(d.cur.vel[g]), om = ld3(d.cur.omg[g]);
    float lin2 = dot(v, v), ang2 = dot(om, om);
    uint32_t timer = st & 0x7FFFFFFFu;
    bool calm = lin2 < 0.0144f && ang2 < 0.0225f;
    d.cur.sleep[g] = calm ? (timer + 1u < 30u ? timer + 1u : 30u) : (timer >= 4u ? timer - 4u : 0u);
  }
  t.sync();
  if (sleep_on) {
    for (uint32_t p = tid; p < P; p += NT) {
      const uint32_t slot = W.pair_off + p;
      int4 h = d.man_cur.hdr[slot];
      if (h.w <= 0) continue;
      const uint32_t ga = off + h.x, gb = off + h.y;
      bool adyn = (d.cur.flags[ga] & BB_BODY_DYNAMIC) != 0, bdyn = (d.cur.flags[gb] & BB_BODY_DYNAMIC) != 0;
      bool aloud = adyn && ((d.cur.sleep[ga] & 0x7FFFFFFFu) < 30u);
      bool bloud = bdyn && ((d.cur.sleep[gb] & 0x7FFFFFFFu) < 30u);
      int32_t err = __float_as_int(d.man_cur.nrm[slot].w);
      bool stalled = err != 0x7FFFFFFF && (err < 0 ? -err : err) < 10;
      // a contact deeper than 5 cm (the narrow phase keeps min(0, pen) in anca.w)
      const bool deep = !stalled && d.man_cur.anca[slot].w < -0.05f;
      if ((aloud || deep) && bdyn) atomicOr(&d.cur.sleep[gb], 0x80000000u);
      if ((bloud || deep) && adyn) atomicOr(&d.cur.sleep[ga], 0x80000000u);
    }
    t.sync();
    uint32_t sl_count = 0;
    for (uint32_t i = tid; i < N; i += NT) {
      const uint32_t g = off + i;
      uint32_t fl = d.cur.flags[g];
      if (!(fl & BB_BODY_DYNAMIC)) continue;
      uint32_t st = d.cur.sleep[g];
      bool vetoed = (st & 0x80000000u) != 0;
      bool quiet = (st & 0x7FFFFFFFu) >= 30u;
      bool sl = quiet && !vetoed;
      bool was = (fl & BB_BODY_SLEEPING) != 0;
      if (sl) {
        if (!was) {
          d.cur.vel[g] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          d.cur.omg[g] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          d.cur.flags[g] = fl | BB_BODY_SLEEPING;
        }
        ++sl_count;
      } else if (was) {
        d.cur.flags[g] = fl & ~(uint32_t)BB_BODY_SLEEPING;
      }
    }
    stat_add(d, w, ST_SLEEPING, sl_count);
  }
  t.sync();
  if (!pgs) return;
  // ---- edge colouring (coloring.slang:44-292).  Phase 2 commits at the
  // colour picked in phase 1; the owner cells touched in a round are cleared
  // by a third phase so the table is zero between rounds.  Only uncoloured
  // manifolds are visited: they sit in a pending list of records, {a row |
  // static << 31, b row | static << 31, pair hash, slot}, compacted after
  // every round.
  uint32_t* color = d.color + W.pair_off;
  uint32_t* pick = d.pick + W.pair_off;
  // pending records in two buffers, taken in turn (the solver rows' headers
  // and normals, both free until the rows are built)
  int4* const rec_a = d.sv_hdr + W.pair_off;
  int4* const rec_b = reinterpret_cast<int4*>(d.sv_n + W.pair_off);
  int4* rec = rec_a;
  for (uint32_t i = tid; i < N; i += NT) d.mask[off + i] = 0;
  const uint32_t npend0 = scan_range(
      t, P, [&](uint32_t p) { return d.man_cur.hdr[W.pair_off + p].w > 0 ? 1u : 0u; },
      [&](uint32_t p, uint32_t ex) {
        const int4 h = d.man_cur.hdr[W.pair_off + p];
        const bool f = h.w > 0;
        color[p] = f ? kMAX : kMAX - 1;
        if (!f) return;
        const uint32_t ga = off + h.x, gb = off + h.y;
        const uint32_t sa = (d.cur.flags[ga] & BB_BODY_DYNAMIC) ? 0u : 0x80000000u;
        const uint32_t sb = (d.cur.flags[gb] & BB_BODY_DYNAMIC) ? 0u : 0x80000000u;
        const uint32_t ida = d.cur.id[ga], idb = d.cur.id[gb];
        const uint32_t lo = ida < idb ? ida : idb, hi = ida < idb ? idb : ida;
        const uint32_t pid = gc_hash(lo * 0x9e3779b9u + hi * 0x85ebca6bu + (uint32_t)h.z * 0xc2b2ae35u);
        rec[ex] = make_int4((int)(ga | sa), (int)(gb | sb), (int)pid, (int)p);
      });
  if (tid == 0) L.carry = npend0;
  t.sync();
  for (uint32_t round = 1; round <= kColors; ++round) {
    const uint32_t npend = L.carry;
    if (npend == 0) break;
    if (tid == 0) L.flag = 0;
    t.sync();
    for (uint32_t i = tid; i < npend; i += NT) {
      const int4 r = rec[i];
      const uint32_t ga = (uint32_t)r.x & 0x7FFFFFFFu, gb = (uint32_t)r.y & 0x7FFFFFFFu;
      const bool sa = r.x < 0, sb = r.y < 0;
      const uint32_t avail = ~((sa ? 0u : d.mask[ga]) | (sb ? 0u : d.mask[gb]));
      if (avail == 0) {
        pick[i] = kColors;
        continue;
      }
      const uint32_t cc = (uint32_t)(__ffs(avail) - 1);
      pick[i] = cc;
      const uint32_t pr = gc_priority((uint32_t)r.z, round);
      if (!sa) atomicMax(&d.owner[ga * kColors + cc], pr);
      if (!sb) atomicMax(&d.owner[gb * kColors + cc], pr);
      L.flag = 1;
    }
    t.sync();
    const bool any = L.flag != 0;
    for (uint32_t i = tid; i < npend; i += NT) {
      const uint32_t cc = pick[i];
      if (cc >= kColors) continue;
      const int4 r = rec[i];
      const uint32_t ga = (uint32_t)r.x & 0x7FFFFFFFu, gb = (uint32_t)r.y & 0x7FFFFFFFu;
      const bool sa = r.x < 0, sb = r.y < 0;
      const uint32_t pr = gc_priority((uint32_t)r.z, round);
      const bool wa = sa || d.owner[ga * kColors + cc] == pr;
      const bool wb = sb || d.owner[gb * kColors + cc] == pr;
      if (wa && wb) {
        color[(uint32_t)r.w] = cc;
        if (!sa) atomicOr(&d.mask[ga], 1u << cc);
        if (!sb) atomicOr(&d.mask[gb], 1u << cc);
      }
    }
    t.sync();
    for (uint32_t i = tid; i < npend; i += NT) {
      const uint32_t cc = pick[i];
      if (cc >= kColors) continue;
      const int4 r = rec[i];
      d.owner[((uint32_t)r.x & 0x7FFFFFFFu) * kColors + cc] = 0;
      d.owner[((uint32_t)r.y & 0x7FFFFFFFu) * kColors + cc] = 0;
    }
    if (tid == 0) L.carry = 0;
    t.sync();
    if (!any) break;
    // keep the still uncoloured records, in order, in the other buffer (one
    // scan over the range; a second walk over one buffer would overwrite
    // records not yet read)
    const int4* cur = rec;
    int4* nxt = rec == rec_a ? rec_b : rec_a;
    const uint32_t kept = scan_range(t, npend, [&](uint32_t i) { return color[(uint32_t)cur[i].w] == kMAX ? 1u : 0u; },
                                     [&](uint32_t i, uint32_t ex) {
                                       const int4 r = cur[i];
                                       if (color[(uint32_t)r.w] == kMAX) nxt[ex] = r;
                                     });
    rec = nxt;
    if (tid == 0) L.carry = kept;
    t.sync();
  }
  // ---- validator (coloring.slang:214-292) and colour buckets
  if (tid < kBuckets) L.bcount[tid] = 0;
  t.sync();
  uint32_t viol = 0, used = 0;
  for (uint32_t p = tid; p < P; p += NT) {
    uint32_t cc = color[p];
    if (cc == kMAX - 1) continue;  // no manifold in this slot
    if (cc >= kColors) {
      atomicAdd(&L.bcount[kColors], 1u);
      continue;
    }
    atomicAdd(&L.bcount[cc], 1u);
    used = used > cc + 1 ? used : cc + 1;
    int4 h = d.man_cur.hdr[W.pair_off + p];
    const uint32_t ga = off + h.x, gb = off + h.y;
    if (d.cur.flags[ga] & BB_BODY_DYNAMIC)
      if (atomicAdd(&d.owner[ga * kColors + cc], 1u) >= 1u) ++viol;
    if (d.cur.flags[gb] & BB_BODY_DYNAMIC)
      if (atomicAdd(&d.owner[gb * kColors + cc], 1u) >= 1u) ++viol;
  }
  stat_add(d, w, ST_GC_VIOL, viol);
  if (used) stat_max(d, w, ST_GC_COUNT, used);
  t.sync();
  for (uint32_t p = tid; p < P; p += NT) {
    uint32_t cc = color[p];
    if (cc >= kColors) continue;
    int4 h = d.man_cur.hdr[W.pair_off + p];
    d.owner[(off + h.x) * kColors + cc] = 0;
    d.owner[(off + h.y) * kColors + cc] = 0;
  }
  t.sync();
  if (tid == 0) d.stats[w * ST_COUNT + ST_GC_OVF] = L.bcount[kColors];
  t.sync();
  // the per-island path (USE_GRAPH_COLORING off) builds its solver rows in
  // s_levels
  if ((flags & BB_SIM_USE_GRAPH_COLORING) == 0) return;
  build_rows(d, W, w, t, L, [&](uint32_t p) { return color[p]; });
}

// Solver rows of the per-island path: dependency levels, level buckets and
// the rows (after s_post; per-world kernels with LL in LDS, teams without).
template <class T>
BB_DI void s_levels(const Dev& d, uint32_t w, const T& t, PostShared& L, LevelLds* LL) {
  const uint32_t tid = t.rank(), NT = t.size();
  const WorldInfo W = d.worlds[w];
  if (W.solver == BB_SOLVER_AVBD || W.n == 0 || (W.flags & BB_SIM_USE_GRAPH_COLORING)) return;
  const uint32_t N = W.n, off = W.body_off, P = d.npairs_cur[w];
  uint32_t* color = d.color + W.pair_off;
  uint32_t* pick = d.pick + W.pair_off;
  // ---- per-island mode (USE_GRAPH_COLORING off, rigid_body_manager.cpp:
  // 1390-1399): the reference walks each contact island's manifolds serially
  // in packed order (= slot order).  That order is kept exactly by parallel
  // sweeps over dependency levels: a manifold's level is one more than the
  // level of the previous solvable manifold (lower slot) on either of its
  // dynamic rows, so the manifolds of one level share no dynamic row and any
  // two that share one are solved in slot order.  Manifolds no sweep solves
  // (no awake dynamic body) get no level.
  //  - Per-world kernels (LL, at most kLvlRows rows, fewer than 65,536
  //    slots): exact levels from one in-order walk, then a counting sort of
  //    the rows by level; every level is a sweep of its own (d.nlev[w]).
  //  - Teams and larger worlds: levels capped at kColors, the longest-path
  //    fixpoint over the rows' ascending slot lists (at most kColors + 1
  //    rounds, each settling one more step of every path); the capped
  //    manifolds form the overflow bucket, solved serially in ascending slot
  //    order after the levels (a manifold after an overflow one on a shared
  //    row is itself overflow).
  // The levels live in `pick`, free after the colouring.
  const bool exact = LL != nullptr && N <= kLvlRows && P < 65536u;
  uint32_t* key = pick;
  if (exact) {
    for (uint32_t i = tid; i < N; i += NT) LL->lvl[i] = 0;
    if (tid == 0) LL->top = 0;
    t.sync();
    for (uint32_t c0 = 0; c0 < P; c0 += kLvlChunk) {
      const uint32_t nc = P - c0 < kLvlChunk ? P - c0 : kLvlChunk;
      for (uint32_t q = tid; q < nc; q += NT) {
        const uint32_t p = c0 + q;
        uint32_t a = kLvlSkip, b = kLvlSkip;
        if (color[p] != kMAX - 1) {
          const int4 h = d.man_cur.hdr[W.pair_off + p];
          if (pgs_solvable(d, off, h)) {
            a = (uint32_t)h.x | ((d.cur.flags[off + h.x] & BB_BODY_DYNAMIC) ? 0u : kLvlStatic);
            b = (uint32_t)h.y | ((d.cur.flags[off + h.y] & BB_BODY_DYNAMIC) ? 0u : kLvlStatic);
          }
        }
        LL->sa[q] = a;
        LL->sb[q] = b;
      }
      t.sync();
      if (tid == 0) {
        uint32_t top = LL->top;
#pragma unroll 1
        for (uint32_t q = 0; q < nc; ++q) {
          const uint32_t a = LL->sa[q], b = LL->sb[q];
          if (a == kLvlSkip) {
            LL->out[q] = 0xFFFFu;
            continue;
          }
          const bool da = (a & kLvlStatic) == 0, db = (b & kLvlStatic) == 0;
          const uint32_t ra = a & ~kLvlStatic, rb = b & ~kLvlStatic;
          uint32_t lv = da ? LL->lvl[ra] : 0u;
          if (db) {
            const uint32_t lb = LL->lvl[rb];
            lv = lb > lv ? lb : lv;
          }
          LL->out[q] = (uint16_t)lv;
          if (da) LL->lvl[ra] = (uint16_t)(lv + 1u);
          if (db) LL->lvl[rb] = (uint16_t)(lv + 1u);
          top = lv + 1u > top ? lv + 1u : top;
        }
        LL->top = top;
      }
      t.sync();
      for (uint32_t q = tid; q < nc; q += NT) key[c0 + q] = LL->out[q] == 0xFFFFu ? kMAX - 1 : LL->out[q];
      t.sync();
    }
  } else {
    build_adj(d, W, P, t, L.carry);
    for (uint32_t p = tid; p < P; p += NT) key[p] = color[p] == kMAX - 1 ? kMAX - 1 : 0u;
    t.sync();
    const uint32_t* adj = d.adj + 2 * W.pair_off;
    uint32_t* lev = d.pick;  // indexed by absolute slot, as adj holds them
    for (uint32_t round = 0; round <= kColors + 1; ++round) {
      if (tid == 0) L.flag = 0;
      t.sync();
      for (uint32_t i = tid; i < N; i += NT) {
        const uint32_t g = off + i;
        const uint32_t cnt = d.deg[g];
        if (!(d.cur.flags[g] & BB_BODY_DYNAMIC) || cnt < 2) continue;
        const uint32_t* l = adj + d.adj_off[g];
        uint32_t lv = __hip_atomic_load(&lev[l[0]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (uint32_t k = 1; k < cnt; ++k) {
          const uint32_t cand = lv + 1 < kColors ? lv + 1 : kColors;
          const uint32_t old = atomicMax(&lev[l[k]], cand);
          if (old < cand) {
            L.flag = 1;
            lv = cand;
          } else {
            lv = old;
          }
        }
      }
      t.sync();
      const bool more = L.flag != 0;
      t.sync();
      if (!more) break;
    }
  }
  if (tid == 0) d.nlev[w] = 0u;
  if (exact) {
    // counting sort of the solved rows by level: level offsets in the
    // world's adjacency area (2 words per slot, free in PGS worlds), fill
    // cursors in the solver row headers (free until the rows are built)
    const uint32_t nlev = LL->top;
    uint32_t* lvoff = d.adj + 2 * W.pair_off;
    uint32_t* cursor = reinterpret_cast<uint32_t*>(d.sv_hdr + W.pair_off);
    uint32_t* bucket = d.bucket + W.pair_off;
    for (uint32_t l = tid; l <= nlev; l += NT) lvoff[l] = 0;
    t.sync();
    for (uint32_t p = tid; p < P; p += NT) {
      const uint32_t k = key[p];
      if (k != kMAX - 1) atomicAdd(&lvoff[k], 1u);
    }
    t.sync();
    if (tid == 0) L.carry = 0;
    t.sync();
    for (uint32_t c0 = 0; c0 <= nlev; c0 += NT) {
      const uint32_t l = c0 + tid;
      const uint32_t v = l <= nlev ? lvoff[l] : 0u;
      uint32_t tot;
      const uint32_t ex = t.scan(v, tot);
      if (l <= nlev) {
        lvoff[l] = L.carry + ex;
        cursor[l] = L.carry + ex;
      }
      t.sync();
      if (tid == 0) L.carry += tot;
      t.sync();
    }
    for (uint32_t p = tid; p < P; p += NT) {
      const uint32_t k = key[p];
      if (k != kMAX - 1) bucket[atomicAdd(&cursor[k], 1u)] = W.pair_off + p;
    }
    t.sync();
    const uint32_t total = lvoff[nlev];
    if (tid == 0) {
      uint32_t* bo = d.bucket_off + w * (kBuckets + 1);
      for (uint32_t c2 = 0; c2 < kBuckets; ++c2) bo[c2] = 0;  // no colour buckets, no overflow
      bo[kBuckets] = total;
      d.nlev[w] = nlev;
    }
    iw_refresh(d, W, t);
    t.sync();
    for (uint32_t k = tid; k < total; k += NT) pgs_compact(d, W, W.pair_off + k, bucket[k]);
    return;
  }
  build_rows(d, W, w, t, L, [&](uint32_t p) {
    return __hip_atomic_load(&key[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  });
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_post(Dev d) {
  __shared__ PostShared L;
  __shared__ ScanLds<BLOCK> S;
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{&S, nullptr};
  s_post(d, w, t, L);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_levels(Dev d) {
  __shared__ PostShared L;
  __shared__ ScanLds<BLOCK> S;
  __shared__ LevelLds LL;
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{&S, nullptr};
  s_levels(d, w, t, L, &LL);
}

// ================================================================== k_pgs
// PGS_SOFT::pre_step / solve (src/shaders/simulation/solvers.slang:142-279),
// colour by colour (RB_sim.slang:1408-1578).  After colouring, the solvable
// manifolds are compacted into colour-sorted solver rows: row k holds the
// manifold constants, and contact c of row k lives at c*NP + k, so one sweep
// over a colour streams contiguous memory.  The velocity-independent pre-step
// quantities (arms, effective masses, soft coefficients) are computed once
// at compaction; the warm-start impulses are then applied colour by colour.
// Statics are never written; rows of one colour share no dynamic body.
// Soft contact coefficients of PGS_SOFT::pre_step (solvers.slang:142-200)
// for contact hertz ch: zeta = 10, omega = 2*pi*ch.
struct Soft {
  float bias_factor, impulse_coef, mass_coef;
};
BB_DI Soft soft_coefs(float dt, float ch) {
  const float zeta = 10.0f;
  const float omega = 2.0f * 3.14159265359f * ch;
  const float cc = dt * omega * (2.0f * zeta + dt * omega);
  Soft s;
  s.bias_factor = omega / (2.0f * zeta + dt * omega);
  s.impulse_coef = 1.0f / (1.0f + cc);
  s.mass_coef = cc * s.impulse_coef;
  return s;
}

struct Pgs {
  bool accum, friction;
  bool plain;  // PGS (per-island path only), else PGS_SOFT
  float inv_dt;
  Soft soft[2];  // [0]: both bodies movable, [1]: one immovable (2x hertz)
};

// PGS::pre_step's Baumgarte bias (solvers.slang:28) from the contact's
// penetration: PENETRATION_FACTOR 0.01 m allowed, BIAS_FACTOR 0.2
// (shared.inl:82-83).
BB_DI float plain_bias(float pen, float inv_dt) {
  return -0.2f * inv_dt * fmn(pen + 0.01f, 0.0f);
}

// Manifold::compute_basis (src/shaders/extensions.slang:262-272)
BB_DI void contact_basis(v3 n, v3& t0, v3& t1) {
  if (fab(n.x) > fab(n.y)) t0 = V(n.z, 0.0f, -n.x);
  else t0 = V(0.0f, -n.z, n.y);
  t0 = vnorm(t0);
  t1 = cross(t0, n);
}

BB_DI mat3 iw_load(const Dev& d, uint32_t g) {
  return mrows(ld3(d.iw[g * 3 + 0]), ld3(d.iw[g * 3 + 1]), ld3(d.iw[g * 3 + 2]));
}

// Caches R*I^-1*R^T (or the quirk form) per row with inv_mass in row0.w.
template <class T>
BB_DI void iw_refresh(const Dev& d, const WorldInfo& W, const T& t) {
  const bool quirk = (W.quirks & BB_QUIRK_TRANSPOSED_WORLD_INERTIA) != 0;
  for (uint32_t i = t.rank(); i < W.n; i += t.size()) {
    const uint32_t g = W.body_off + i;
    const uint32_t id = d.cur.id[g];
    mat3 M = world_inv_inertia(d, W.body_off + id, ldq(d.cur.rot[g]), quirk);
    d.iw[g * 3 + 0] = st3(mrow(M, 0), d.props.mass[W.body_off + id].y);
    d.iw[g * 3 + 1] = st3(mrow(M, 1), 0.0f);
    d.iw[g * 3 + 2] = st3(mrow(M, 2), 0.0f);
  }
}

BB_DI bool pgs_solvable(const Dev& d, uint32_t off, int4 h) {
  const uint32_t fa = d.cur.flags[off + h.x], fb = d.cur.flags[off + h.y];
  const bool awa = (fa & BB_BODY_DYNAMIC) && !(fa & BB_BODY_SLEEPING);
  const bool awb = (fb & BB_BODY_DYNAMIC) && !(fb & BB_BODY_SLEEPING);
  return awa || awb;
}

// Solver row k from manifold slot: the velocity-independent quantities of
// PGS_SOFT::pre_step (arms, effective masses).  The row keeps the normal
// only: the tangents are recomputed from it exactly as Manifold::compute_basis
// made them, and the soft coefficients depend only on whether either body
// is immovable (contact hertz doubled, bit 1024 of the count word).
BB_DI void pgs_compact(const Dev& d, const WorldInfo& W, uint32_t k, uint32_t slot) {
  const uint32_t off = W.body_off, NP = d.np;
  const int4 h = d.man_cur.hdr[slot];
  const uint32_t ga = off + h.x, gb = off + h.y;
  const bool da = (d.cur.flags[ga] & BB_BODY_DYNAMIC) != 0, db = (d.cur.flags[gb] & BB_BODY_DYNAMIC) != 0;
  const float4 ia0 = d.iw[ga * 3], ib0 = d.iw[gb * 3];
  const float ima = ia0.w, imb = ib0.w;
  const mat3 Ia = iw_load(d, ga), Ib = iw_load(d, gb);
  const float fa = d.props.mass[off + d.cur.id[ga]].z, fb = d.props.mass[off + d.cur.id[gb]].z;
  const v3 n = ld3(d.man_cur.nrm[slot]);
  const v3 t0 = ld3(d.man_cur.t0[slot]), t1 = ld3(d.man_cur.t1[slot]);
  const v3 pa = ld3(d.man_cur.anca[slot]), pb = ld3(d.man_cur.ancb[slot]);
  const bool soft_static = ima == 0.0f || imb == 0.0f;
  for (int i = 0; i < h.w; ++i) {
    const float4 cp = d.man_cur.cpos[slot * 8 + i];
    const float4 lam = d.man_cur.clam[slot * 8 + i];
    const v3 r1 = ld3(cp) - pa, r2 = ld3(cp) - pb;
    const v3 rn1 = cross(r1, n), rn2 = cross(r2, n);
    float kn = ima + imb;
    kn += dot(rn1, mulmv(Ia, rn1));
    kn += dot(rn2, mulmv(Ib, rn2));
    const float nm = kn == 0.0f ? 0.0f : 1.0f / kn;
    float tm[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const v3 tj = j == 0 ? t0 : t1;
      const v3 rt1 = cross(r1, tj), rt2 = cross(r2, tj);
      float kt = ima + imb;
      kt += dot(rt1, mulmv(Ia, rt1));
      kt += dot(rt2, mulmv(Ib, rt2));
      tm[j] = kt == 0.0f ? 0.0f : 1.0f / kt;
    }
    d.sv_r1[i * NP + k] = st3(r1, cp.w);
    d.sv_r2[i * NP + k] = st3(r2, nm);
    d.sv_l[i * NP + k] = make_float4(lam.x, lam.y, lam.z, tm[0]);
    d.sv_tm1[i * NP + k] = tm[1];
  }
  d.sv_hdr[k] = make_int4(h.x, h.y, h.w | (da ? 256 : 0) | (db ? 512 : 0) | (soft_static ? 1024 : 0),
                          __float_as_int(__builtin_sqrtf(fa * fb)));
  d.sv_n[k] = st3(n, 0.0f);
}

// Body state during the PGS sweeps: rows below `lcap` keep velocity, angular
// velocity and the three cached world inverse-inertia rows (inv_mass in
// row0.w) in LDS, kBodyLds float4 per row; the rest stay in HBM.
static constexpr uint32_t kBodyLds = 5;
struct VelCache {
  float4* lds;
  uint32_t off, lcap;
  BB_DI void load(const Dev& d, uint32_t i, v3& v, v3& w) const {
    if (i < lcap) {
      v = ld3(lds[kBodyLds * i]);
      w = ld3(lds[kBodyLds * i + 1]);
    } else {
      v = ld3(d.cur.vel[off + i]);
      w = ld3(d.cur.omg[off + i]);
    }
  }
  BB_DI void store(const Dev& d, uint32_t i, v3 v, v3 w) const {
    if (i < lcap) {
      lds[kBodyLds * i] = st3(v, 0.0f);
      lds[kBodyLds * i + 1] = st3(w, 0.0f);
    } else {
      d.cur.vel[off + i] = st3(v, 0.0f);
      d.cur.omg[off + i] = st3(w, 0.0f);
    }
  }
  BB_DI void load_iw(const Dev& d, uint32_t i, float4& r0, float4& r1, float4& r2) const {
    if (i < lcap) {
      r0 = lds[kBodyLds * i + 2];
      r1 = lds[kBodyLds * i + 3];
      r2 = lds[kBodyLds * i + 4];
    } else {
      const uint32_t g = off + i;
      r0 = d.iw[g * 3];
      r1 = d.iw[g * 3 + 1];
      r2 = d.iw[g * 3 + 2];
    }
  }
};

// Contacts per row whose solver data pgs_row loads up front.
static constexpr int kPgsPre = 4;
// A row's solver data that pgs_row loads before it touches the body state:
// the header, the normal and the first kPgsPre contacts.  None of it is
// written by any other row, so it can be loaded early: the colour sweeps
// fetch a thread's row of the next colour while it solves this colour's.
struct PgsRow {
  int4 h;
  float4 n;
  float4 r1[kPgsPre], r2[kPgsPre], l[kPgsPre];
  float tm[kPgsPre];
};
// The first kPgsPre contacts' data are loaded whatever the row's contact
// count: the contact arrays hold 8 entries per row slot, so the loads stay in
// bounds, and issuing them with the header instead of after it (the count is
// in the header) saves a dependent round trip per colour phase.  Entries past
// the count are never used.
BB_DI void pgs_fetch(const Dev& d, uint32_t k, int kind, PgsRow& R) {
  const uint32_t NP = d.np;
  R.h = d.sv_hdr[k];
  R.n = d.sv_n[k];
#pragma unroll
  for (int c = 0; c < kPgsPre; ++c) {
    R.r1[c] = d.sv_r1[c * NP + k];
    R.r2[c] = d.sv_r2[c * NP + k];
    R.l[c] = d.sv_l[c * NP + k];
    R.tm[c] = kind != 0 ? d.sv_tm1[c * NP + k] : 0.0f;
  }
}
// kind 0: warm-start impulses of pre_step; 1: solve with bias; 2: relax.
// kPlain: PGS (the per-island path) instead of PGS_SOFT.  R: the row's
// fetched data (pgs_fetch).
template <bool kPlain>
BB_DI void pgs_solve(const Dev& d, const VelCache& vc, uint32_t off, uint32_t k, const Pgs& s, int kind,
                     const PgsRow& R) {
  const uint32_t NP = d.np;
  const int4 h = R.h;
  const int count = h.z & 255;
  if (kind == 0 && !s.accum) return;
  const bool da = (h.z & 256) != 0, db = (h.z & 512) != 0;
  float4 pr1[kPgsPre], pr2[kPgsPre], pl[kPgsPre];
  float ptm[kPgsPre];
#pragma unroll
  for (int c = 0; c < kPgsPre; ++c) {
    pr1[c] = R.r1[c];
    pr2[c] = R.r2[c];
    pl[c] = R.l[c];
    ptm[c] = R.tm[c];
  }
  const v3 n = ld3(R.n);
  float4 ia0, ia1, ia2, ib0, ib1, ib2;
  vc.load_iw(d, (uint32_t)h.x, ia0, ia1, ia2);
  vc.load_iw(d, (uint32_t)h.y, ib0, ib1, ib2);
  const float ima = ia0.w, imb = ib0.w;
  const mat3 Ia = mrows(ld3(ia0), ld3(ia1), ld3(ia2));
  const mat3 Ib = mrows(ld3(ib0), ld3(ib1), ld3(ib2));
  v3 va, wa, vb, wb;
  vc.load(d, (uint32_t)h.x, va, wa);
  vc.load(d, (uint32_t)h.y, vb, wb);
  v3 t0, t1;
  contact_basis(n, t0, t1);
  if (kind == 0) {
    auto warm = [&](float4 l, float4 c1, float4 c2) {
      const v3 r1 = ld3(c1), r2 = ld3(c2);
      v3 P = n * l.x;
      if (s.friction) {
        P = P + t0 * l.y;
        P = P + t1 * l.z;
      }
      if (da) {
        va = va - P * ima;
        wa = wa - mulmv(Ia, cross(r1, P));
      }
      if (db) {
        vb = vb + P * imb;
        wb = wb + mulmv(Ib, cross(r2, P));
      }
    };
#pragma unroll
    for (int c = 0; c < kPgsPre; ++c)
      if (c < count) warm(pl[c], pr1[c], pr2[c]);
    for (int c = kPgsPre; c < count; ++c) warm(d.sv_l[c * NP + k], d.sv_r1[c * NP + k], d.sv_r2[c * NP + k]);
  } else {
    const bool use_bias = kind == 1;
    const float fr = __int_as_float(h.w);
    const Soft sc = (h.z & 1024) ? s.soft[1] : s.soft[0];
    const float bias_factor = sc.bias_factor, impulse_coef = sc.impulse_coef, mass_coef = sc.mass_coef;
    auto solve = [&](int c, float4 c1, float4 c2, float4 l, float tm1) {
      const v3 r1 = ld3(c1), r2 = ld3(c2);
      const float pen = c1.w, nm = c2.w;
      v3 rv = ((vb + cross(wb, r2)) - va) - cross(wa, r1);
      if (s.friction) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const v3 tj = j == 0 ? t0 : t1;
          const float tmass = j == 0 ? l.w : tm1;
          const float tv = dot(rv, tj);
          float ti = tmass * -tv;
          const float maxf = fr * l.x;
          if (s.accum) {
            const float old = j == 0 ? l.y : l.z;
            const float nw = fcl(old + ti, -maxf, maxf);
            if (j == 0) l.y = nw; else l.z = nw;
            ti = nw - old;
          } else {
            ti = fcl(ti, -maxf, maxf);
          }
          const v3 tiv = tj * ti;
          if (da) {
            va = va - tiv * ima;
            wa = wa - mulmv(Ia, cross(r1, tiv));
          }
          if (db) {
            vb = vb + tiv * imb;
            wb = wb + mulmv(Ib, cross(r2, tiv));
          }
        }
      }
      rv = ((vb + cross(wb, r2)) - va) - cross(wa, r1);
      const float nv = dot(rv, n);
      float imp;
      if (kPlain) {
        // PGS::solve (solvers.slang:106-117)
        imp = nm * (-nv + plain_bias(pen, s.inv_dt));
      } else {
        float bias = 0.0f, mass_scale = 1.0f, impulse_scale = 0.0f;
        if (pen > 0.0f) {
          bias = pen * s.inv_dt;
        } else if (use_bias) {
          bias = fmx(bias_factor * pen, -4.0f);
          mass_scale = mass_coef;
          impulse_scale = impulse_coef;
        }
        imp = -nm * mass_scale * (nv + bias) - impulse_scale * l.x;
      }
      if (s.accum) {
        const float old = l.x;
        l.x = fmx(old + imp, 0.0f);
        imp = l.x - old;
      } else {
        imp = fmx(imp, 0.0f);
      }
      const v3 niv = n * imp;
      if (da) {
        va = va - niv * ima;
        wa = wa - mulmv(Ia, cross(r1, niv));
      }
      if (db) {
        vb = vb + niv * imb;
        wb = wb + mulmv(Ib, cross(r2, niv));
      }
      d.sv_l[c * NP + k] = l;
    };
#pragma unroll
    for (int c = 0; c < kPgsPre; ++c)
      if (c < count) solve(c, pr1[c], pr2[c], pl[c], ptm[c]);
    for (int c = kPgsPre; c < count; ++c)
      solve(c, d.sv_r1[c * NP + k], d.sv_r2[c * NP + k], d.sv_l[c * NP + k], d.sv_tm1[c * NP + k]);
  }
  if (da) vc.store(d, (uint32_t)h.x, va, wa);
  if (db) vc.store(d, (uint32_t)h.y, vb, wb);
}
// One row, fetched and solved in turn.
template <bool kPlain>
BB_DI void pgs_row(const Dev& d, const VelCache& vc, uint32_t off, uint32_t k, const Pgs& s, int kind) {
  PgsRow R;
  pgs_fetch(d, k, kind, R);
  pgs_solve<kPlain>(d, vc, off, k, s, kind, R);
}

// RigidBody::apply_angular_velocity (src/shaders/extensions.slang:148-166)
BB_DI qt apply_angular_velocity(qt q, v3 om, float dt) {
  float angle = vlen(om) * dt;
  if (angle > 0.0001f) {
    v3 axis = vnorm(om);
    float half = angle / 2.0f;
    float sh, ch;
    sincos_c(half, &sh, &ch);
    return qnorm(qmul(qt{axis * sh, ch}, q));
  }
  v3 dr = (om * dt) * 0.5f;
  return qnorm(qmul(qt{dr, 1.0f}, q));
}

// kIsland: the instances for the per-island path (k_pgs_island), which
// sweep dependency levels; kPlain: PGS (per-island path only) rather than
// PGS_SOFT.  k_pgs, the colour sweeps, carries neither.
template <bool kIsland, bool kPlain, class T>
BB_DI void s_pgs(const Dev& d, uint32_t w, const T& t, float4* pgs_lds, uint32_t lcap) {
  const uint32_t tid = t.rank(), NT = t.size();
  const WorldInfo W = d.worlds[w];
  if (W.solver == BB_SOLVER_AVBD || W.n == 0) return;
  if (((W.flags & BB_SIM_USE_GRAPH_COLORING) == 0) != kIsland) return;
  if ((W.solver == BB_SOLVER_PGS) != kPlain) return;
  const uint32_t off = W.body_off, N = W.n, A = W.a, NP = d.np;
  Pgs s;
  s.accum = (W.flags & BB_SIM_ACCUM_IMPULSE) != 0;
  s.friction = (W.flags & BB_SIM_FRICTION) != 0;
  s.plain = kPlain;
  s.inv_dt = W.dt > 0.0f ? 1.0f / W.dt : 0.0f;
  {
    const float hertz = fmn(30.0f, s.inv_dt * 0.3333333333333333f);
    s.soft[0] = soft_coefs(W.dt, hertz);
    s.soft[1] = soft_coefs(W.dt, 2.0f * hertz);
  }
  const uint32_t* bo = d.bucket_off + w * (kBuckets + 1);
  // colours in use, or every capped level bucket of the per-island path
  // (empty buckets are skipped)
  const uint32_t used = kIsland ? kColors : d.stats[w * ST_COUNT + ST_GC_COUNT];
  const uint32_t base = W.pair_off;
  const uint32_t ovf_lo = bo[kColors], ovf_hi = bo[kBuckets];
  const VelCache vc{pgs_lds, off, N < lcap ? N : lcap};
  for (uint32_t i = tid; i < vc.lcap; i += NT) {
    const uint32_t g = off + i;
    pgs_lds[kBodyLds * i] = d.cur.vel[g];
    pgs_lds[kBodyLds * i + 1] = d.cur.omg[g];
    pgs_lds[kBodyLds * i + 2] = d.iw[g * 3];
    pgs_lds[kBodyLds * i + 3] = d.iw[g * 3 + 1];
    pgs_lds[kBodyLds * i + 4] = d.iw[g * 3 + 2];
  }
  t.sync();
  auto row = [&](uint32_t k, int kind) { pgs_row<kPlain>(d, vc, off, base + k, s, kind); };
  // exact dependency levels of the per-island path (0: buckets)
  const uint32_t nlev = kIsland ? d.nlev[w] : 0u;
  const uint32_t* lvoff = d.adj + 2 * W.pair_off;
  auto sweep = [&](int kind) {
    if (nlev) {
      for (uint32_t lv = 0; lv < nlev; ++lv) {
        for (uint32_t k = lvoff[lv] + tid; k < lvoff[lv + 1]; k += NT) row(k, kind);
        t.sync();
      }
      return;
    }
    if (d.pgs_prefetch) {
      // each thread fetches its first row of the next colour as soon as it
      // is done with this colour, before the barrier, so the loads are in
      // flight while the slower threads finish (rows beyond the first, when a
      // colour has more rows than threads, are fetched where they are solved)
      uint32_t c = 0;
      while (c < used && bo[c] == bo[c + 1]) ++c;
      PgsRow R;
      bool have = c < used && bo[c] + tid < bo[c + 1];
      if (have) pgs_fetch(d, base + bo[c] + tid, kind, R);
      while (c < used) {
        if (have) pgs_solve<kPlain>(d, vc, off, base + bo[c] + tid, s, kind, R);
        for (uint32_t k = bo[c] + tid + NT; k < bo[c + 1]; k += NT) row(k, kind);
        uint32_t cn = c + 1;
        while (cn < used && bo[cn] == bo[cn + 1]) ++cn;
        have = cn < used && bo[cn] + tid < bo[cn + 1];
        if (have) pgs_fetch(d, base + bo[cn] + tid, kind, R);
        t.sync();
        c = cn;
      }
    } else {
      // the next colour's end offset is loaded while this colour's rows are
      // solved, so no phase starts with a dependent load of its range
      uint32_t b0 = bo[0], b1 = used > 0 ? bo[1] : 0u;
      for (uint32_t c = 0; c < used; ++c) {
        const uint32_t b2 = c + 1 < used ? bo[c + 2] : 0u;
        if (b0 != b1) {
          for (uint32_t k = b0 + tid; k < b1; k += NT) row(k, kind);
          t.sync();
        }
        b0 = b1;
        b1 = b2;
      }
    }
    if (ovf_hi > ovf_lo) {
      if (tid == 0)
        for (uint32_t k = ovf_lo; k < ovf_hi; ++k) row(k, kind);
      t.sync();
    }
  };
  sweep(0);
  for (uint32_t it = 0; it < W.iterations; ++it) sweep(1);
  // integrate positions (RB_sim.slang:1287-1326), same row quirk as advect
  if (W.flags & BB_SIM_ADVECTION) {
    uint32_t lim = N;
    if (W.quirks & BB_QUIRK_ACTIVE_DISPATCH_ROWS) {
      uint32_t r = ((A + 31u) / 32u) * 32u;
      lim = r < N ? r : N;
    }
    for (uint32_t i = tid; i < lim; i += NT) {
      const uint32_t g = off + i;
      uint32_t fl = d.cur.flags[g];
      if (!((fl & BB_BODY_DYNAMIC) && !(fl & BB_BODY_SLEEPING))) continue;
      v3 v, om;
      vc.load(d, i, v, om);
      const v3 p = ld3(d.cur.pos[g]);
      d.cur.pos[g] = st3(p + v * W.dt, 0.0f);
      d.cur.rot[g] = stq(apply_angular_velocity(ldq(d.cur.rot[g]), om, W.dt));
    }
  }
  t.sync();
  iw_refresh(d, W, t);
  t.sync();
  for (uint32_t i = tid; i < vc.lcap; i += NT) {
    const uint32_t g = off + i;
    pgs_lds[kBodyLds * i + 2] = d.iw[g * 3];
    pgs_lds[kBodyLds * i + 3] = d.iw[g * 3 + 1];
    pgs_lds[kBodyLds * i + 4] = d.iw[g * 3 + 2];
  }
  t.sync();
  // relaxation sweeps: PGS_SOFT only (RB_sim.slang:1398-1402)
  if (!s.plain)
    for (uint32_t it = 0; it < W.iterations; ++it) sweep(2);
  for (uint32_t i = tid; i < vc.lcap; i += NT) {
    d.cur.vel[off + i] = pgs_lds[kBodyLds * i];
    d.cur.omg[off + i] = pgs_lds[kBodyLds * i + 1];
  }
  // impulses and pre-step quantities back into the manifold slots (warm start
  // of the next step, taps)
  const uint32_t* bucket = d.bucket + base;
  for (uint32_t k = tid; k < ovf_hi; k += NT) {
    const uint32_t row = base + k, slot = bucket[k];
    const int4 h = d.sv_hdr[row];
    const Soft sc = (h.z & 1024) ? s.soft[1] : s.soft[0];
    const float mass_coef = sc.mass_coef, bias_factor = sc.bias_factor, impulse_coef = sc.impulse_coef;
    for (int c = 0; c < (h.z & 255); ++c) {
      const float4 l = d.sv_l[c * NP + row];
      const float nm = d.sv_r2[c * NP + row].w, tm1 = d.sv_tm1[c * NP + row];
      if (s.plain) {
        // PGS::pre_step leaves the soft coefficients as the narrow phase set
        // them
        const float mc = d.man_cur.clam[slot * 8 + c].w;
        d.man_cur.clam[slot * 8 + c] = make_float4(l.x, l.y, l.z, mc);
        d.man_cur.cmas[slot * 8 + c] = make_float4(nm, l.w, tm1, plain_bias(d.sv_r1[c * NP + row].w, s.inv_dt));
        continue;
      }
      d.man_cur.clam[slot * 8 + c] = make_float4(l.x, l.y, l.z, mass_coef);
      d.man_cur.cmas[slot * 8 + c] = make_float4(nm, l.w, tm1, bias_factor);
      float4 a = d.man_cur.cana[slot * 8 + c];
      a.w = impulse_coef;
      d.man_cur.cana[slot * 8 + c] = a;
    }
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 4) void k_pgs(Dev d) {
  HIP_DYNAMIC_SHARED(float4, pgs_lds)
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{nullptr, nullptr};
  s_pgs<false, false>(d, w, t, pgs_lds, d.pgs_lcap);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 4) void k_pgs_island(Dev d) {
  HIP_DYNAMIC_SHARED(float4, pgs_lds)
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{nullptr, nullptr};
  if (d.worlds[w].solver == BB_SOLVER_PGS) s_pgs<true, true>(d, w, t, pgs_lds, d.pgs_lcap);
  else s_pgs<true, false>(d, w, t, pgs_lds, d.pgs_lcap);
}

// ================================================================== AVBD
// src/shaders/simulation/avbd.slang.  Adjacency = per-row list of incident
// manifold slots in ascending order (replaces the linked lists).
static constexpr float AV_BETA = 100000.0f, AV_GAMMA = 0.99f, AV_KMIN = 1.0f, AV_KMAX = 1000000000.0f;
static constexpr float AV_MARGIN = 0.0005f, AV_STICK_THRESH = 0.01f, AV_STICK_SLOP = 0.001f;
static constexpr uint32_t AV_ITERS = 10, AV_ROUNDS = 16, AV_POST_STAB = 4;
static constexpr uint32_t AV_NB = kAvbdColors * kAvbdLayers;
// a team's per-workgroup bucket counts and bases fit the sort's staging area
static_assert(2 * AV_NB <= (256 / 64) * 256, "AVBD bucket scratch exceeds SortLds::wcnt");

BB_DI uint32_t avbd_priority(uint32_t body, uint32_t round) {  // avbd.slang:58-63
  uint32_t x = body * 0x9e3779b9u + round * 0x85ebca6bu;
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return ((x & 0xFFFu) << 20) | (body & 0xFFFFFu);
}
BB_DI v3 small_angle_delta(qt now, qt start) {
  qt dq = qmul(now, qt{-start.v, start.w});
  float sg = dq.w < 0.0f ? -1.0f : 1.0f;
  return dq.v * (2.0f * sg);
}
BB_DI v3 avbd_q_rotate(qt q, v3 u) { return u + cross(q.v, cross(q.v, u) + u * q.w) * 2.0f; }
BB_DI mat3 avbd_inverse3(const mat3& a) {
  v3 c0 = mcol(a, 0), c1 = mcol(a, 1), c2 = mcol(a, 2);
  v3 r0 = cross(c1, c2), r1 = cross(c2, c0), r2 = cross(c0, c1);
  float det = dot(c0, r0);
  float inv = fab(det) > 1e-12f ? 1.0f / det : 0.0f;
  return mrows(r0 * inv, r1 * inv, r2 * inv);
}
// avbd_solve6 (avbd.slang:324-387): in-place LDL^T, diagonal >= 1e-8
BB_DI void avbd_solve6(float H[6][6], float b[6], float x[6]) {
#pragma unroll
  for (int jj = 0; jj < 6; ++jj) {
    float dd = H[jj][jj];
#pragma unroll
    for (int kk = 0; kk < jj; ++kk) dd -= H[jj][kk] * H[jj][kk] * H[kk][kk];
    dd = fmx(dd, 1e-8f);
    H[jj][jj] = dd;
#pragma unroll
    for (int ii = jj + 1; ii < 6; ++ii) {
      float s = H[ii][jj];
#pragma unroll
      for (int kk = 0; kk < jj; ++kk) s -= H[ii][kk] * H[jj][kk] * H[kk][kk];
      H[ii][jj] = s / dd;
    }
  }
#pragma unroll
  for (int ii = 0; ii < 6; ++ii) {
    float s = b[ii];
#pragma unroll
    for (int kk = 0; kk < ii; ++kk) s -= H[ii][kk] * b[kk];
    b[ii] = s;
  }
#pragma unroll
  for (int ii = 0; ii < 6; ++ii) b[ii] /= H[ii][ii];
#pragma unroll
  for (int ii = 5; ii >= 0; --ii) {
    float s = b[ii];
#pragma unroll
    for (int kk = ii + 1; kk < 6; ++kk) s -= H[kk][ii] * x[kk];
    x[ii] = s;
  }
}

struct AvBody {  // pose and step-start state of one side of a contact
  bool dyn;
  v3 pos;
  qt rot;
  v3 p0;
  qt q0;
};
BB_DI AvBody av_body(const Dev& d, uint32_t g) {
  AvBody b;
  uint32_t fl = d.cur.flags[g];
  b.dyn = (fl & BB_BODY_DYNAMIC) && !(fl & BB_BODY_SLEEPING);
  b.pos = ld3(d.cur.pos[g]);
  b.rot = ldq(d.cur.rot[g]);
  if (b.dyn) {
    b.p0 = ld3(d.s_pos0[g]);
    b.q0 = ldq(d.s_rot0[g]);
  } else {
    b.p0 = b.pos;
    b.q0 = b.rot;
  }
  return b;
}
// avbd_contact_C (avbd.slang:232-258)
BB_DI float av_contact_C(v3 n, v3 p, float pen, const AvBody& b1, const AvBody& b2, float alpha, v3& rr1,
                         v3& rr2) {
  rr1 = p - b1.p0;
  rr2 = p - b2.p0;
  float C0 = pen + AV_MARGIN;
  float C = C0 * (1.0f - alpha);
  if (b1.dyn) {
    v3 dp1 = b1.pos - b1.p0;
    v3 dth1 = small_angle_delta(b1.rot, b1.q0);
    C -= dot(n, dp1 + cross(dth1, rr1));
  }
  if (b2.dyn) {
    v3 dp2 = b2.pos - b2.p0;
    v3 dth2 = small_angle_delta(b2.rot, b2.q0);
    C += dot(n, dp2 + cross(dth2, rr2));
  }
  return C;
}
// avbd_dir_C (avbd.slang:274-305)
BB_DI float av_dir_C(v3 dir, v3 anc_a, v3 anc_b, const AvBody& b1, const AvBody& b2, float alpha, v3& ar1,
                     v3& ar2, float& C0t) {
  ar1 = avbd_q_rotate(b1.q0, anc_a);
  ar2 = avbd_q_rotate(b2.q0, anc_b);
  C0t = dot(dir, (b2.p0 + ar2) - (b1.p0 + ar1));
  float c0e = fsign(C0t) * fmx(0.0f, fab(C0t) - AV_STICK_SLOP);
  float C = c0e * (1.0f - alpha);
  if (b1.dyn) {
    v3 dp1 = b1.pos - b1.p0;
    v3 dth1 = small_angle_delta(b1.rot, b1.q0);
    C -= dot(dir, dp1 + cross(dth1, ar1));
  }
  if (b2.dyn) {
    v3 dp2 = b2.pos - b2.p0;
    v3 dth2 = small_angle_delta(b2.rot, b2.q0);
    C += dot(dir, dp2 + cross(dth2, ar2));
  }
  return C;
}

// entry_avbd_primal (avbd.slang:475-658) for row i, by a group of G lanes
// (lane = 0..G-1; G = 1, 2, 4 or 8 in the per-world kernels, a launch
// parameter, and 1 in teams).  The
// Hessian and gradient are the inertial terms plus one partial sum per
// incident manifold, added in ascending manifold order (the adjacency
// order); lane e % G assembles manifold e, so the dependent loads of G
// manifolds are in flight together.  Only the lower triangle of H is formed:
// avbd_solve6 reads no other entry.
static constexpr int kAvH = 21;  // lower-triangle entries of the 6x6 H
BB_DI void av_manifold_terms(const Dev& d, const WorldInfo& W, uint32_t i, uint32_t slot, float alpha,
                             float (&He)[kAvH], float (&ge)[6]) {
  const uint32_t off = W.body_off;
  const int4 h = d.man_cur.hdr[slot];
  const uint32_t b1 = (uint32_t)h.x, b2 = (uint32_t)h.y;
  const bool i1 = b1 == i;
  const AvBody B1 = av_body(d, off + b1), B2 = av_body(d, off + b2);
  const float f1 = d.props.mass[off + d.cur.id[off + b1]].z, f2 = d.props.mass[off + d.cur.id[off + b2]].z;
  const float mu = __builtin_sqrtf(fmx(f1, 0.0f) * fmx(f2, 0.0f));
  const v3 n = ld3(d.man_cur.nrm[slot]);
  const v3 tdir[2] = {ld3(d.man_cur.t0[slot]), ld3(d.man_cur.t1[slot])};
#pragma unroll
  for (int q = 0; q < kAvH; ++q) He[q] = 0.0f;
#pragma unroll
  for (int a = 0; a < 6; ++a) ge[a] = 0.0f;
  auto add_row = [&](float k, float f, const float (&J)[6]) {
#pragma unroll
    for (int a = 0; a < 6; ++a) ge[a] += f * J[a];
    int q = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = 0; b <= a; ++b) He[q++] += k * J[a] * J[b];
  };
  for (int j = 0; j < h.w; ++j) {
    const uint32_t cs = slot * 8 + j;
    const float4 cp = d.man_cur.cpos[cs];
    const float4 lm = d.man_cur.clam[cs];
    const float4 cmv = d.man_cur.cmas[cs];
    const v3 anc_a = ld3(d.man_cur.cana[cs]), anc_b = ld3(d.man_cur.canb[cs]);
    v3 rr1, rr2;
    const float C = av_contact_C(n, ld3(cp), cp.w, B1, B2, alpha, rr1, rr2);
    const float lam = lm.x, kk = lm.w;
    const float f = fmn(kk * C + lam, 0.0f);
    const v3 Jl = i1 ? -n : n;
    const v3 Ja = i1 ? -cross(rr1, n) : cross(rr2, n);
    const float J[6] = {Jl.x, Jl.y, Jl.z, Ja.x, Ja.y, Ja.z};
    add_row(kk, f, J);
    const float bound = mu * fab(lam);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const v3 td = tdir[t];
      v3 ar1, ar2;
      float C0t;
      const float Ct = av_dir_C(td, anc_a, anc_b, B1, B2, alpha, ar1, ar2, C0t);
      const float lt = t == 0 ? lm.y : lm.z;
      const float kt = t == 0 ? cmv.y : cmv.z;
      const float ft = fcl(kt * Ct + lt, -bound, bound);
      const v3 Jlt = i1 ? -td : td;
      const v3 Jat = i1 ? -cross(ar1, td) : cross(ar2, td);
      const float Jt[6] = {Jlt.x, Jlt.y, Jlt.z, Jat.x, Jat.y, Jat.z};
      add_row(kt, ft, Jt);
    }
  }
}

template <uint32_t G>
BB_DI void av_primal(const Dev& d, const WorldInfo& W, uint32_t i, float alpha, uint32_t lane) {
  const uint32_t off = W.body_off, g = off + i;
  const uint32_t id = d.cur.id[g];
  const bool quirk = (W.quirks & BB_QUIRK_TRANSPOSED_WORLD_INERTIA) != 0;
  const float dt = W.dt;
  const float inv_dt2 = dt > 0.0f ? 1.0f / (dt * dt) : 0.0f;
  v3 pos = ld3(d.cur.pos[g]);
  qt rot = ldq(d.cur.rot[g]);
  const float m_lin = d.props.mass[off + id].x * inv_dt2;
  const mat3 Iw = avbd_inverse3(world_inv_inertia(d, off + id, rot, quirk));
  // inertial terms (every lane of the group holds all of H and g)
  float Hi[kAvH], gi[6];
  {
    int q = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = 0; b <= a; ++b) {
        float v = 0.0f;
        if (a == b && a < 3) v = m_lin;
        if (a >= 3 && b >= 3) v = Iw.m[a - 3][b - 3] * inv_dt2;
        Hi[q++] = v;
      }
    const v3 dpos = pos - ld3(d.s_post[g]);
    const v3 dth = small_angle_delta(rot, ldq(d.s_rott[g]));
    const v3 ga = mulmv(Iw, dth) * inv_dt2;
    gi[0] = m_lin * dpos.x; gi[1] = m_lin * dpos.y; gi[2] = m_lin * dpos.z;
    gi[3] = ga.x; gi[4] = ga.y; gi[5] = ga.z;
  }
  const uint32_t a0 = 2 * W.pair_off + d.adj_off[g], deg = d.deg[g];
  for (uint32_t e0 = 0; e0 < deg; e0 += G) {
    float He[kAvH], ge[6];
    const uint32_t e = e0 + lane;
    if (e < deg) av_manifold_terms(d, W, i, d.adj[a0 + e], alpha, He, ge);
    if constexpr (G == 1) {
#pragma unroll
      for (int q = 0; q < kAvH; ++q) Hi[q] += He[q];
#pragma unroll
      for (int a = 0; a < 6; ++a) gi[a] += ge[a];
    } else {
      // add the group's partial sums in manifold order
#pragma unroll
      for (uint32_t src = 0; src < G; ++src) {
        const bool has = e0 + src < deg;
#pragma unroll
        for (int q = 0; q < kAvH; ++q) {
          const float v = __shfl(He[q], (int)src, (int)G);
          if (has) Hi[q] += v;
        }
#pragma unroll
        for (int a = 0; a < 6; ++a) {
          const float v = __shfl(ge[a], (int)src, (int)G);
          if (has) gi[a] += v;
        }
      }
    }
  }
  if (lane != 0) return;
  float H[6][6], rhs[6], dx[6];
  {
    int q = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = 0; b < 6; ++b) H[a][b] = 0.0f;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = 0; b <= a; ++b) H[a][b] = Hi[q++];
  }
#pragma unroll
  for (int a = 0; a < 6; ++a) rhs[a] = -gi[a];
  avbd_solve6(H, rhs, dx);
  pos = pos + V(dx[0], dx[1], dx[2]);
  v3 half = V(dx[3], dx[4], dx[5]) * 0.5f;
  rot = qnorm(qmul(qt{half, 1.0f}, rot));
  d.cur.pos[g] = st3(pos, 0.0f);
  d.cur.rot[g] = stq(rot);
}

// entry_avbd_dual (avbd.slang:726-808) for one manifold slot
BB_DI void av_dual(const Dev& d, const WorldInfo& W, uint32_t slot) {
  const uint32_t off = W.body_off;
  const int4 h = d.man_cur.hdr[slot];
  const AvBody B1 = av_body(d, off + h.x), B2 = av_body(d, off + h.y);
  if (!B1.dyn && !B2.dyn) return;
  const float f1 = d.props.mass[off + d.cur.id[off + h.x]].z, f2 = d.props.mass[off + d.cur.id[off + h.y]].z;
  const float mu = __builtin_sqrtf(fmx(f1, 0.0f) * fmx(f2, 0.0f));
  const v3 n = ld3(d.man_cur.nrm[slot]);
  const v3 tdir[2] = {ld3(d.man_cur.t0[slot]), ld3(d.man_cur.t1[slot])};
  for (int j = 0; j < h.w; ++j) {
    const uint32_t cs = slot * 8 + j;
    const float4 cp = d.man_cur.cpos[cs];
    float4 lm = d.man_cur.clam[cs];
    float4 cmv = d.man_cur.cmas[cs];
    float4 cb = d.man_cur.canb[cs];
    const v3 anc_a = ld3(d.man_cur.cana[cs]), anc_b = ld3(cb);
    v3 rr1, rr2;
    float C = av_contact_C(n, ld3(cp), cp.w, B1, B2, 1.0f, rr1, rr2);
    const float lam = lm.x, kk = lm.w;
    const float nl = fmn(kk * C + lam, 0.0f);
    lm.x = nl;
    if (nl < 0.0f) lm.w = fmn(kk + AV_BETA * fab(C), AV_KMAX);
    const float bound = mu * fab(lam);
    bool stick_ok = true;
    float drift2 = 0.0f;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      v3 ar1, ar2;
      float C0t;
      float Ct = av_dir_C(tdir[t], anc_a, anc_b, B1, B2, 1.0f, ar1, ar2, C0t);
      float lt = t == 0 ? lm.y : lm.z;
      float kt = t == 0 ? cmv.y : cmv.z;
      float nlt = fcl(kt * Ct + lt, -bound, bound);
      if (t == 0) lm.y = nlt; else lm.z = nlt;
      if (nlt > -bound && nlt < bound) {
        float nk = fmn(kt + AV_BETA * fab(Ct), AV_KMAX);
        if (t == 0) cmv.y = nk; else cmv.z = nk;
      } else {
        stick_ok = false;
      }
      drift2 += C0t * C0t;
    }
    cb.w = __uint_as_float((stick_ok && drift2 < AV_STICK_THRESH * AV_STICK_THRESH) ? 1u : 0u);
    d.man_cur.clam[cs] = lm;
    d.man_cur.cmas[cs] = cmv;
    d.man_cur.canb[cs] = cb;
  }
}

struct AvbdShared {
  uint32_t carry;
  uint32_t flag;
  uint32_t used;
  uint32_t bcount[AV_NB];
  uint32_t boff[AV_NB + 1];
};

template <class T, uint32_t kLanes = T::kAvGroup>
BB_DI void s_avbd(const Dev& d, uint32_t w, const T& t, AvbdShared& L) {
  const uint32_t tid = t.rank(), NT = t.size();
  constexpr uint32_t G = kLanes;  // lanes per body in the primal sweeps
  const WorldInfo W = d.worlds[w];
  if (W.solver != BB_SOLVER_AVBD || W.n == 0) return;
  const uint32_t N = W.n, off = W.body_off, P = d.npairs_cur[w];
  const float dt = W.dt;
  // ---- adjacency (CSR, ascending slot per row) for dynamic rows
  build_adj(d, W, P, t, L.carry);
  uint32_t* adj = d.adj + 2 * W.pair_off;
  for (uint32_t i = tid; i < N; i += NT) d.bcolor[off + i] = kMAX;
  t.sync();
  // ---- Jones-Plassmann body colouring, round-snapshot semantics
  // (entry_avbd_color_round, avbd.slang:846-911)
  for (uint32_t rd = 0; rd < AV_ROUNDS; ++rd) {
    if (tid == 0) L.flag = 0;
    t.sync();
    for (uint32_t i = tid; i < N; i += NT) {
      uint32_t c = d.bcolor[off + i];
      d.bcolor_snap[off + i] = c;
      if (c == kMAX && (d.cur.flags[off + i] & BB_BODY_DYNAMIC)) L.flag = 1;
    }
    t.sync();
    if (L.flag == 0) break;
    for (uint32_t i = tid; i < N; i += NT) {
      const uint32_t g = off + i;
      if (d.bcolor_snap[g] != kMAX || !(d.cur.flags[g] & BB_BODY_DYNAMIC)) continue;
      const uint32_t my = avbd_priority(d.cur.id[g], rd);
      uint32_t used = 0;
      bool lose = false;
      const uint32_t* l = adj + d.adj_off[g];
      for (uint32_t e = 0; e < d.deg[g]; ++e) {
        int4 h = d.man_cur.hdr[l[e]];
        uint32_t pr = (uint32_t)h.x == i ? (uint32_t)h.y : (uint32_t)h.x;
        if (pr >= N || pr == i) continue;
        if (!(d.cur.flags[off + pr] & BB_BODY_DYNAMIC)) continue;
        uint32_t pc = d.bcolor_snap[off + pr];
        if (pc == kMAX) {
          if (avbd_priority(d.cur.id[off + pr], rd) > my) {
            lose = true;
            break;
          }
        } else if (pc < kAvbdColors) {
          used |= 1u << pc;
        }
      }
      if (lose) continue;
      uint32_t avail = ~used;
      if (avail == 0) continue;
      uint32_t c = (uint32_t)(__ffs(avail) - 1);
      if (c < kAvbdColors) d.bcolor[g] = c;
    }
    t.sync();
  }
  // validator (entry_avbd_color_validate, avbd.slang:913-981)
  {
    uint32_t viol = 0, cnt = 0;
    for (uint32_t i = tid; i < N; i += NT) {
      const uint32_t g = off + i;
      if (!(d.cur.flags[g] & BB_BODY_DYNAMIC)) continue;
      const uint32_t mc = d.bcolor[g];
      bool has = false;
      const uint32_t* l = adj + d.adj_off[g];
      for (uint32_t e = 0; e < d.deg[g]; ++e) {
        int4 h = d.man_cur.hdr[l[e]];
        uint32_t pr = (uint32_t)h.x == i ? (uint32_t)h.y : (uint32_t)h.x;
        if (pr >= N || pr == i || !(d.cur.flags[off + pr] & BB_BODY_DYNAMIC)) continue;
        has = true;
        if (mc != kMAX && d.bcolor[off + pr] == mc && pr > i) ++viol;
      }
      if (mc == kMAX) {
        if (has) ++viol;
        continue;
      }
      cnt = cnt > mc + 1 ? cnt : mc + 1;
    }
    stat_add(d, w, ST_AV_VIOL, viol);
    if (cnt) stat_max(d, w, ST_AV_COUNT, cnt);
  }
  // ---- prepare (avbd.slang:70-128)
  for (uint32_t i = tid; i < N; i += NT) {
    const uint32_t g = off + i;
    const uint32_t fl = d.cur.flags[g];
    if (!(fl & BB_BODY_DYNAMIC) || (fl & BB_BODY_SLEEPING)) continue;
    v3 om = ld3(d.cur.omg[g]);
    float oml = vlen(om);
    if (oml > 50.0f) {
      om = om * (50.0f / oml);
      d.cur.omg[g] = st3(om, 0.0f);
    }
    const float grav = (fl & BB_BODY_GRAVITY) ? W.gravity : 0.0f;
    const v3 p = ld3(d.cur.pos[g]), v = ld3(d.cur.vel[g]), pv = ld3(d.cur.pvel[g]);
    const qt q = ldq(d.cur.rot[g]);
    d.s_pos0[g] = st3(p, 0.0f);
    d.s_rot0[g] = stq(q);
    d.s_post[g] = st3((p + v * dt) + V(0.0f, grav, 0.0f) * (dt * dt), 0.0f);
    qt qtl = apply_angular_velocity(q, om, dt);
    d.s_rott[g] = stq(qtl);
    float accel_w = 0.0f;
    if (fab(grav) > 1e-6f) {
      float accel_y = (v.y - pv.y) / dt;
      accel_w = fcl(accel_y * fsign(grav) / fab(grav), 0.0f, 1.0f);
    }
    d.cur.pos[g] = st3((p + v * dt) + V(0.0f, grav, 0.0f) * (accel_w * dt * dt), 0.0f);
    d.cur.rot[g] = stq(qtl);
  }
  t.sync();
  t.mark(kMarkAvColour);
  // ---- warm start (avbd.slang:389-473)
  {
    uint32_t sticks = 0;
    const bool ws_on = (W.flags & BB_SIM_WARM_STARTING) != 0;
    for (uint32_t p = tid; p < P; p += NT) {
      const uint32_t slot = W.pair_off + p;
      int4 h = d.man_cur.hdr[slot];
      if (h.w <= 0) continue;
      const AvBody B1 = av_body(d, off + h.x), B2 = av_body(d, off + h.y);
      const qt q1c = qconj(B1.q0), q2c = qconj(B2.q0);
      for (int j = 0; j < h.w; ++j) {
        const uint32_t cs = slot * 8 + j;
        float4 lm = d.man_cur.clam[cs];
        float4 cmv = d.man_cur.cmas[cs];
        float4 ca = d.man_cur.cana[cs];
        float4 cb = d.man_cur.canb[cs];
        const bool matched = d.man_cur.cws[cs] != 0 && ws_on;
        uint32_t stick = __float_as_uint(cb.w);
        if (!matched) {
          lm.x = 0.0f;
          lm.w = AV_KMIN;
          lm.y = lm.z = 0.0f;
          cmv.y = cmv.z = AV_KMIN;
          stick = 0;
        } else {
          lm.w = fcl(lm.w * AV_GAMMA, AV_KMIN, AV_KMAX);
          lm.x = fmn(lm.x, 0.0f);
          cmv.y = fcl(cmv.y * AV_GAMMA, AV_KMIN, AV_KMAX);
          cmv.z = fcl(cmv.z * AV_GAMMA, AV_KMIN, AV_KMAX);
        }
        const bool sticking = matched && stick != 0;
        if (!sticking) {
          const v3 pp = ld3(d.man_cur.cpos[cs]);
          ca = st3(avbd_q_rotate(q1c, pp - B1.p0), ca.w);
          cb = st3(avbd_q_rotate(q2c, pp - B2.p0), cb.w);
        } else {
          ++sticks;
        }
        cb.w = __uint_as_float(stick);
        d.man_cur.clam[cs] = lm;
        d.man_cur.cmas[cs] = cmv;
        d.man_cur.cana[cs] = ca;
        d.man_cur.canb[cs] = cb;
      }
    }
    stat_add(d, w, ST_AV_STICK, sticks);
  }
  // ---- support depth (avbd.slang:667-724), pass-snapshot semantics
  for (uint32_t i = tid; i < N; i += NT) {
    const uint32_t fl = d.cur.flags[off + i];
    const bool grounded = !(fl & BB_BODY_DYNAMIC) || (fl & BB_BODY_SLEEPING);
    d.depth[off + i] = grounded ? 0u : kMAX;
  }
  t.sync();
  for (uint32_t pass = 0; pass < kAvbdLayers; ++pass) {
    for (uint32_t i = tid; i < N; i += NT) d.depth_snap[off + i] = d.depth[off + i];
    t.sync();
    for (uint32_t p = tid; p < P; p += NT) {
      int4 h = d.man_cur.hdr[W.pair_off + p];
      if (h.w <= 0) continue;
      const uint32_t d1 = d.depth_snap[off + h.x], d2 = d.depth_snap[off + h.y];
      if (d2 != kMAX && d2 + 1u < d1) atomicMin(&d.depth[off + h.x], d2 + 1u);
      if (d1 != kMAX && d1 + 1u < d2) atomicMin(&d.depth[off + h.y], d1 + 1u);
    }
    t.sync();
  }
  // ---- bodies grouped by (colour, layer) for the primal sweeps.  A team's
  // counters live in global memory: each workgroup counts its rows in LDS
  // (the sort's staging area, lc[0, AV_NB)) and adds one total per bucket,
  // then takes its base in each bucket (lc[AV_NB, 2 AV_NB)) with one atomic.
  // Rows of a bucket are independent (one colour), so their order within
  // it does not change the sweeps.
  uint32_t* lc = nullptr;
  if constexpr (T::kGlobalScalars) {
    lc = t.sort_lds->s.wcnt;
    for (uint32_t k = threadIdx.x; k < 2 * AV_NB; k += blockDim.x) lc[k] = 0;
  }
  for (uint32_t k = tid; k < AV_NB; k += NT) L.bcount[k] = 0;
  if (tid == 0) L.used = 0;
  t.sync();
  uint32_t used_t = 0;  // colours used: per thread, per wave, then one atomic per wave
  for (uint32_t i = tid; i < N; i += NT) {
    const uint32_t g = off + i;
    const uint32_t fl = d.cur.flags[g], c = d.bcolor[g];
    if (!(fl & BB_BODY_DYNAMIC) || (fl & BB_BODY_SLEEPING) || c >= kAvbdColors) continue;
    uint32_t layer = d.depth[g] < kAvbdLayers - 1 ? d.depth[g] : kAvbdLayers - 1;
    atomicAdd(lc ? &lc[c * kAvbdLayers + layer] : &L.bcount[c * kAvbdLayers + layer], 1u);
    used_t = c + 1 > used_t ? c + 1 : used_t;
  }
  if constexpr (T::kGlobalScalars) {
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < AV_NB; k += blockDim.x)
      if (lc[k]) atomicAdd(&L.bcount[k], lc[k]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t u = __shfl_down(used_t, o, 64);
    used_t = u > used_t ? u : used_t;
  }
  if ((threadIdx.x & 63u) == 0u && used_t) atomicMax(&L.used, used_t);
  t.sync();
  if (tid == 0) {
    uint32_t s = 0;
    for (uint32_t k = 0; k < AV_NB; ++k) {
      L.boff[k] = s;
      s += L.bcount[k];
      L.bcount[k] = L.boff[k];  // fill cursor
    }
    L.boff[AV_NB] = s;
  }
  t.sync();
  if constexpr (T::kGlobalScalars) {
    for (uint32_t k = threadIdx.x; k < AV_NB; k += blockDim.x) {
      if (lc[k]) lc[AV_NB + k] = atomicAdd(&L.bcount[k], lc[k]);
      lc[k] = 0;  // now this workgroup's cursor in the bucket
    }
    __syncthreads();
  }
  uint32_t* bb = d.bbucket + off;
  for (uint32_t i = tid; i < N; i += NT) {
    const uint32_t g = off + i;
    const uint32_t fl = d.cur.flags[g], c = d.bcolor[g];
    if (!(fl & BB_BODY_DYNAMIC) || (fl & BB_BODY_SLEEPING) || c >= kAvbdColors) continue;
    uint32_t layer = d.depth[g] < kAvbdLayers - 1 ? d.depth[g] : kAvbdLayers - 1;
    const uint32_t k = c * kAvbdLayers + layer;
    const uint32_t o = lc ? lc[AV_NB + k] + atomicAdd(&lc[k], 1u) : atomicAdd(&L.bcount[k], 1u);
    if (o < N) bb[o] = i;  // the buckets hold at most the world's N rows
  }
  t.sync();
  t.mark(kMarkAvGroup);
  const uint32_t used = L.used;
  // ---- main sweeps: 10 x (colours, dual) (rigid_body_manager.cpp:1366-1373)
  for (uint32_t it = 0; it < AV_ITERS; ++it) {
    for (uint32_t c = 0; c < used; ++c) {
      for (uint32_t k = L.boff[c * kAvbdLayers] + tid / G; k < L.boff[(c + 1) * kAvbdLayers]; k += NT / G)
        av_primal<G>(d, W, bb[k], 1.0f, tid % G);
      t.sync();
    }
    for (uint32_t p = tid; p < P; p += NT)
      if (d.man_cur.hdr[W.pair_off + p].w > 0) av_dual(d, W, W.pair_off + p);
    t.sync();
  }
  t.mark(kMarkAvSweeps);
  // ---- finalize (avbd.slang:130-202)
  {
    const float inv_dt = dt > 0.0f ? 1.0f / dt : 0.0f;
    for (uint32_t i = tid; i < N; i += NT) {
      const uint32_t g = off + i;
      const uint32_t fl = d.cur.flags[g];
      if (!(fl & BB_BODY_DYNAMIC) || (fl & BB_BODY_SLEEPING)) continue;
      const v3 p = ld3(d.cur.pos[g]);
      const qt q = ldq(d.cur.rot[g]);
      const qt q0 = ldq(d.s_rot0[g]);
      d.cur.pvel[g] = d.cur.vel[g];
      v3 v = (p - ld3(d.s_pos0[g])) * inv_dt;
      qt dq = qmul(q, qt{-q0.v, q0.w});
      float sg = dq.w < 0.0f ? -1.0f : 1.0f;
      v3 om = dq.v * (2.0f * sg * inv_dt);
      float v2 = dot(v, v), w2 = dot(om, om);
      if (v2 < 0.09f && w2 < 0.16f && d.deg[g] >= 3u) {
        v = v * 0.9f;
        om = om * 0.9f;
      }
      d.cur.vel[g] = st3(v, 0.0f);
      d.cur.omg[g] = st3(om, 0.0f);
    }
  }
  t.sync();
  // ---- post-stabilisation: one cascade by support layer, then plain sweeps
  // (rigid_body_manager.cpp:1391-1423)
  for (uint32_t layer = 0; layer < kAvbdLayers; ++layer) {
    for (uint32_t c = 0; c < used; ++c) {
      const uint32_t k0 = L.boff[c * kAvbdLayers + layer], k1 = L.boff[c * kAvbdLayers + layer + 1];
      if (T::kSkipEmpty && k1 == k0) continue;  // an empty bucket needs no barrier
      for (uint32_t k = k0 + tid / G; k < k1; k += NT / G) av_primal<G>(d, W, bb[k], 0.0f, tid % G);
      t.sync();
    }
  }
  for (uint32_t ps = 1; ps < AV_POST_STAB; ++ps) {
    for (uint32_t c = 0; c < used; ++c) {
      for (uint32_t k = L.boff[c * kAvbdLayers] + tid / G; k < L.boff[(c + 1) * kAvbdLayers]; k += NT / G)
        av_primal<G>(d, W, bb[k], 0.0f, tid % G);
      t.sync();
    }
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_avbd(Dev d) {
  __shared__ AvbdShared L;
  __shared__ ScanLds<BLOCK> S;
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{&S, nullptr};
  switch (d.avbd_lanes) {  // the solver tile: lanes per body in the primal sweeps
    case 1: s_avbd<BlockTeam<BLOCK>, 1>(d, w, t, L); break;
    case 2: s_avbd<BlockTeam<BLOCK>, 2>(d, w, t, L); break;
    case 8: s_avbd<BlockTeam<BLOCK>, 8>(d, w, t, L); break;
    default: s_avbd<BlockTeam<BLOCK>, 4>(d, w, t, L); break;
  }
}

// ============================================================== k_finish
// End-of-step counters (RB_sim.slang:1655-1715): id checksum and max |v|.
template <class T>
BB_DI void s_finish(const Dev& d, uint32_t w, const T& t) {
  const uint32_t tid = t.rank(), NT = t.size();
  const WorldInfo W = d.worlds[w];
  uint32_t idsum = 0, maxv = 0;
  for (uint32_t i = tid; i < W.n; i += NT) {
    const uint32_t g = W.body_off + i;
    idsum += d.cur.id[g];
    if (d.cur.flags[g] & BB_BODY_DYNAMIC) {
      v3 v = ld3(d.cur.vel[g]);
      uint32_t mv = (uint32_t)fmn(__builtin_sqrtf(dot(v, v)) * 1000.0f, 1.0e9f);
      maxv = maxv > mv ? maxv : mv;
    }
  }
  stat_add(d, w, ST_IDSUM, idsum);
  if (maxv) stat_max(d, w, ST_MAXV, maxv);
  if (tid == 0) {
    d.stats[w * ST_COUNT + ST_N] = W.n;
    d.stats[w * ST_COUNT + ST_A] = W.a;
    d.stats[w * ST_COUNT + ST_NODES] = 2 * d.stats[w * ST_COUNT + ST_M];
    // overflow is kept until the host reads it: the world's sticky word, and
    // one word for the whole batch so that a check costs one 4-byte read
    const uint32_t ovf = d.stats[w * ST_COUNT + ST_OVERFLOW];
    if (ovf) {
      atomicOr(&d.stats[w * ST_COUNT + ST_OVF_STICKY], ovf);
      atomicOr(d.ovf_any, 1u);
    }
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_finish(Dev d) {
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{nullptr, nullptr};
  s_finish(d, w, t);
}

// ================================================================ k_team
// Large worlds: every stage of one step for each team, with team barriers in
// place of the kernel boundaries of the per-world path.  Teams always use
// kTeamBlock-thread workgroups, whatever the per-world workgroup size: the
// 1,024-thread build of this kernel is held to 128 VGPRs and spills heavily
// (about 2,400 scratch accesses), and it faulted the GPU on a 20,000-body
// AVBD world whose emulated run (every workgroup a host thread, every
// work-item a fiber, under AddressSanitizer) stays in bounds; that build is
// not used.
union TeamShared {
  BuildShared build;
  BroadShared broad;
  PostShared post;
  AvbdShared avbd;
};
static_assert(sizeof(TeamShared) <= kTeamSharedBytes, "team scalars exceed their slot");

// LDS of a team workgroup: the radix sort's in the build stage, the broad
// phase's shared walks after it
template <int BLOCK>
union TeamLds {
  TeamSortLds<BLOCK> sort;
  WalkLds<BLOCK> walk;
};

// kVox: the instance that also runs the voxel narrow phase, launched when a
// team world holds voxel bodies (its code would raise every team's register
// count otherwise).
template <int BLOCK, bool kVox>
__global__ __launch_bounds__(BLOCK) void k_team(Dev d) {
  __shared__ ScanLds<BLOCK> S;
  __shared__ TeamLds<BLOCK> Q;
  __shared__ uint32_t dead[3];
  const TeamTable& T = *d.teams;
  const uint32_t nt = T.count < kMaxTeams ? T.count : kMaxTeams;
  if (nt == 0 || blockIdx.x >= T.first[nt]) return;  // workgroup-uniform
  uint32_t j = 0;
  while (j + 1 < nt && blockIdx.x >= T.first[j + 1]) ++j;
  const uint32_t b0 = T.first[j], nb = T.first[j + 1] - b0;
  const uint32_t w = T.world[j];
  // stage boundaries: every workgroup has passed the barrier before rank 0
  // stamps the clock
  const GroupTeam<BLOCK> t{&S, &Q.sort, d.team_ctl + j, d.team_hist + (size_t)b0 * 256, blockIdx.x - b0, nb,
                           dead, d.stats + (size_t)w * ST_COUNT, d.ovf_any,
                           d.team_spin ? d.team_spin : kTeamSpinLimit,
                           d.team_clock ? d.team_clock + (size_t)j * kTeamMarks : nullptr,
                           d.sort_blocks ? d.sort_blocks : kTeamSortBlocks};
  TeamShared& G = *reinterpret_cast<TeamShared*>(d.team_shared + (size_t)j * kTeamSharedBytes);
  if (threadIdx.x == 0) dead[1] = dead[2] = 0u;  // thread 0's own words: no barrier needed
  auto mark = [&](uint32_t m) { t.mark(m); };
  mark(0);
  s_build(d, w, t, G.build);
  t.sync();
  mark(1);
  s_broad(d, w, t, G.broad, Q.walk);
  t.sync();
  mark(2);
  s_narrow<false>(d, w, t);
  if constexpr (kVox) s_narrow<true>(d, w, t);  // each instance returns at once on the other kind
  t.sync();
  s_contacts(d, w, t);
  t.sync();
  mark(3);
  s_post(d, w, t, G.post);
  t.sync();
  mark(4);
  s_levels(d, w, t, G.post, nullptr);
  t.sync();
  if (d.worlds[w].flags & BB_SIM_USE_GRAPH_COLORING) s_pgs<false, false>(d, w, t, nullptr, 0u);
  else if (d.worlds[w].solver == BB_SOLVER_PGS) s_pgs<true, true>(d, w, t, nullptr, 0u);
  else s_pgs<true, false>(d, w, t, nullptr, 0u);
  t.sync();
  mark(5);
  s_avbd(d, w, t, G.avbd);
  t.sync();
  mark(6);
  s_finish(d, w, t);
  t.sync();
  mark(7);
  // every barrier of the step was met (a workgroup that saw one fail has
  // left through GroupTeam::abandon)
  if (threadIdx.x == 0) t.leave();
}

// ============================================================== k_fused
// The whole step of a world in one launch, one workgroup per world: the
// stage functions in the order k_team runs them, a workgroup barrier between
// them.  A world's workgroup goes on to its next stage as soon as its own
// stage is done, where the stage kernels wait for the slowest world of the
// lane at every stage boundary.  Stage scalars share one LDS union (one stage
// runs at a time); the PGS body rows are the dynamic LDS, as in k_pgs.
template <int BLOCK>
struct FusedShared {
  union {
    struct {
      BuildShared b;
      SortLds<BLOCK> q;
    } build;
    struct {
      BroadShared b;
      WalkLds<BLOCK> w;
    } broad;
    struct {
      PostShared p;
      LevelLds lv;
    } post;
    AvbdShared avbd;
  };
};

template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 4) void k_fused(Dev d) {
  HIP_DYNAMIC_SHARED(float4, pgs_lds)
  __shared__ FusedShared<BLOCK> G;
  __shared__ ScanLds<BLOCK> S;
  const uint32_t w = d.order[blockIdx.x];
  if (d.worlds[w].mode & kModeTeam) return;
  const BlockTeam<BLOCK> t{&S, &G.build.q};
  s_build(d, w, t, G.build.b);
  __syncthreads();
  s_broad(d, w, t, G.broad.b, G.broad.w);
  __syncthreads();
  s_narrow<false>(d, w, t);
  if (d.any_voxel) s_narrow<true>(d, w, t);
  __syncthreads();
  s_contacts(d, w, t);
  __syncthreads();
  s_post(d, w, t, G.post.p);
  __syncthreads();
  if (d.any_island) {
    s_levels(d, w, t, G.post.p, &G.post.lv);
    __syncthreads();
  }
  s_pgs<false, false>(d, w, t, pgs_lds, d.pgs_lcap);
  if (d.any_island) {
    __syncthreads();
    if (d.worlds[w].solver == BB_SOLVER_PGS) s_pgs<true, true>(d, w, t, pgs_lds, d.pgs_lcap);
    else s_pgs<true, false>(d, w, t, pgs_lds, d.pgs_lcap);
  }
  __syncthreads();
  switch (d.avbd_lanes) {
    case 1: s_avbd<BlockTeam<BLOCK>, 1>(d, w, t, G.avbd); break;
    case 2: s_avbd<BlockTeam<BLOCK>, 2>(d, w, t, G.avbd); break;
    case 8: s_avbd<BlockTeam<BLOCK>, 8>(d, w, t, G.avbd); break;
    default: s_avbd<BlockTeam<BLOCK>, 4>(d, w, t, G.avbd); break;
  }
  __syncthreads();
  s_finish(d, w, t);
}

#ifndef BB_BLOCK_TU
// Render hand-off (entry_update_acceleration_structures, src/shaders/
// AS_mngr.slang:5-36, with RigidBody::get_instance_transform,
// include/shared.inl:377-419): for every body of world w, at its id, the
// 3x4 row-major affine transform [R | t] of its pose (the layout of a
// Vulkan TLAS instance), R from Quaternion::to_matrix (include/math.hpp:
// 292-309).
__global__ __launch_bounds__(256) void k_instances(Dev d, uint32_t w, float* __restrict__ dst) {
  const WorldInfo W = d.worlds[w];
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= W.n) return;
  const uint32_t g = W.body_off + i;
  const v3 p = ld3(d.cur.pos[g]);
  const mat3 R = qmat(ldq(d.cur.rot[g]));
  float4* o = reinterpret_cast<float4*>(dst + (size_t)d.cur.id[g] * 12u);
  o[0] = make_float4(R.m[0][0], R.m[0][1], R.m[0][2], p.x);
  o[1] = make_float4(R.m[1][0], R.m[1][1], R.m[1][2], p.y);
  o[2] = make_float4(R.m[2][0], R.m[2][1], R.m[2][2], p.z);
}

// LBVH hand-off (RigidBodyManager::get_lbvh_node_buffer, src/
// rigid_body_manager.cpp:1601-1604, consumed as procedural AABB geometry
// with a 40-byte stride by src/acceleration_structure_manager.cpp:490-535):
// the 2n-1 nodes of world w's last BVH in the reference's LBVHNode layout
// (include/shared.inl:867-897): min xyz, max xyz, left, right, primitive
// index (the leaf's row; 0xFFFFFFFF for an inner node) and a zero pad.  A
// leaf has left = right = -1.
__global__ __launch_bounds__(256) void k_lbvh_export(Dev d, uint32_t w, uint32_t* __restrict__ dst) {
  const WorldInfo W = d.worlds[w];
  const uint32_t nn = W.n ? 2u * W.n - 1u : 0u;
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= nn) return;
  const size_t g = 2 * (size_t)W.body_off + i;
  const float4 mn = d.nmin[g], mx = d.nmax[g];
  const int l = __float_as_int(mn.w), r = __float_as_int(mx.w);
  uint32_t* o = dst + (size_t)i * 10u;
  o[0] = __float_as_uint(mn.x);
  o[1] = __float_as_uint(mn.y);
  o[2] = __float_as_uint(mn.z);
  o[3] = __float_as_uint(mx.x);
  o[4] = __float_as_uint(mx.y);
  o[5] = __float_as_uint(mx.z);
  o[6] = l < 0 ? 0xFFFFFFFFu : (uint32_t)l;
  o[7] = l < 0 ? 0xFFFFFFFFu : (uint32_t)r;
  o[8] = l < 0 ? (uint32_t)r : kMAX;
  o[9] = 0u;
}

// Batch totals of one step's counters, one workgroup: out[0..7] = sum N,
// sum A, sum P, sum M, sum C, sum over worlds of (colours used x M), max
// penetration (mm), sum sleeping.
__global__ __launch_bounds__(1024) void k_sum_counts(const uint32_t* __restrict__ stats, uint32_t nw,
                                                     unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[8][1024 / 64];
  unsigned long long acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t w = threadIdx.x; w < nw; w += 1024) {
    const uint32_t* s = stats + (size_t)w * ST_COUNT;
    acc[0] += s[ST_N];
    acc[1] += s[ST_A];
    acc[2] += s[ST_P];
    acc[3] += s[ST_M];
    acc[4] += s[ST_C];
    acc[5] += (unsigned long long)s[ST_GC_COUNT] * s[ST_M];
    acc[6] = acc[6] > s[ST_PEN] ? acc[6] : s[ST_PEN];
    acc[7] += s[ST_SLEEPING];
  }
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int f = 0; f < 8; ++f) {
    unsigned long long v = acc[f];
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long u = __shfl_down(v, o, 64);
      v = f == 6 ? (v > u ? v : u) : v + u;
    }
    if (lane == 0) red[f][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int f = threadIdx.x;
    unsigned long long v = 0;
    for (int k = 0; k < 1024 / 64; ++k) v = f == 6 ? (v > red[f][k] ? v : red[f][k]) : v + red[f][k];
    out[f] = v;
  }
}

// Mechanical energy per world of the current rows, in double:
// sum over dynamic bodies of m|v|^2/2 + w_b.I_b.w_b/2 - m*g*y, with I_b the
// physical inertia of the shape (box: m/12 diag(ey^2+ez^2, ...) over full
// extents; sphere: 2/5 m r^2; voxel shape: the inverse of the stored inverse
// inertia) and w_b the angular velocity in the body frame.
#endif  // BB_BLOCK_TU

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_energy(Dev d, double* __restrict__ out) {
  __shared__ double red[BLOCK / 64];
  const uint32_t w = blockIdx.x, tid = threadIdx.x;
  const WorldInfo W = d.worlds[w];
  double e = 0.0;
  for (uint32_t i = tid; i < W.n; i += BLOCK) {
    const uint32_t g = W.body_off + i;
    if (!(d.cur.flags[g] & BB_BODY_DYNAMIC)) continue;
    const uint32_t pid = W.body_off + d.cur.id[g];
    const float4 mp = d.props.mass[pid], lo = d.props.pmin[pid], hi = d.props.pmax[pid];
    const double m = mp.x;
    const uint32_t kind = shape_word(hi.w) & 0xFFu;
    // body-frame inertia I (voxel bodies: the inverse of their inverse
    // inertia, a full symmetric 3x3)
    double I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (kind == BB_SHAPE_SPHERE) {
      I[0][0] = I[1][1] = I[2][2] = 0.4 * m * (double)lo.w * (double)lo.w;
    } else if (kind == BB_SHAPE_VOXEL) {
      const float4 r0 = d.props.ii0[pid], r1 = d.props.ii1[pid], r2 = d.props.ii2[pid];
      const double a[3][3] = {{r0.x, r0.y, r0.z}, {r1.x, r1.y, r1.z}, {r2.x, r2.y, r2.z}};
      const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                         a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                         a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
      if (det != 0.0) {
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            const int r1i = (c + 1) % 3, r2i = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
            I[r][c] = (a[r1i][c1] * a[r2i][c2] - a[r1i][c2] * a[r2i][c1]) / det;
          }
      }
    } else {
      const double ex = (double)hi.x - lo.x, ey = (double)hi.y - lo.y, ez = (double)hi.z - lo.z;
      I[0][0] = m / 12.0 * (ey * ey + ez * ez);
      I[1][1] = m / 12.0 * (ex * ex + ez * ez);
      I[2][2] = m / 12.0 * (ex * ex + ey * ey);
    }
    const float4 p = d.cur.pos[g], q = d.cur.rot[g], v = d.cur.vel[g], o = d.cur.omg[g];
    // body-frame angular velocity: conj(q) * w * q
    const double qx = q.x, qy = q.y, qz = q.z, qw = q.w, ox = o.x, oy = o.y, oz = o.z;
    const double tx = 2.0 * (qy * oz - qz * oy), ty = 2.0 * (qz * ox - qx * oz), tz = 2.0 * (qx * oy - qy * ox);
    const double bw[3] = {ox - qw * tx + (qy * tz - qz * ty), oy - qw * ty + (qz * tx - qx * tz),
                          oz - qw * tz + (qx * ty - qy * tx)};
    double rot = 0.0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) rot += bw[r] * I[r][c] * bw[c];
    e += 0.5 * m * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z) + 0.5 * rot -
         m * (double)W.gravity * (double)p.y;
  }
  for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o, 64);
  if ((tid & 63) == 0) red[tid >> 6] = e;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int k = 0; k < BLOCK / 64; ++k) s += red[k];
    out[w] = s;
  }
}

#ifndef BB_BLOCK_TU
// Grid-stride float4 copy of known size (byte-counter calibration).
__global__ __launch_bounds__(256) void k_copy_calibrate(const float4* __restrict__ src, float4* __restrict__ dst,
                                                        size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

#endif  // BB_BLOCK_TU

}  // namespace bb

// ============================================================ host launch
// The per-world kernels of each workgroup size and their launchers are
// compiled in a unit of their own (bb_step_block.hip includes this file with
// BB_BLOCK_TU and BB_BLOCK set), so the sizes build in parallel; this unit
// holds the team kernel, the launch sequence and the small utility kernels.
namespace bb {
static constexpr int kStages = 7;
template <int B>
void launch_stage_b(int stage, const Dev& d, hipStream_t st, uint32_t grid);
template <int B>
void launch_pgs_b(const Dev& d, hipStream_t st, uint32_t grid);
template <int B>
void launch_fused_b(const Lane& l, hipStream_t st);

#ifdef BB_BLOCK_TU
// Stages of the per-world path, in launch order (k_team, when there is one,
// goes before the first).  Stage k of the timed launch is bracketed by events
// ev[k] and ev[k + 1]; the narrow phase's two kernels share stage 2; the
// colour sweeps (k_pgs) of stage 4 are launched by the lane, with its own
// workgroup size.
template <int B>
void launch_stage_b(int stage, const Dev& d, hipStream_t st, uint32_t grid) {
  const dim3 g(grid), b(B);
  switch (stage) {
    case 0: hipLaunchKernelGGL(k_build<B>, g, b, 0, st, d); break;
    case 1: hipLaunchKernelGGL(k_broad<B>, g, b, 0, st, d); break;
    case 2:
      hipLaunchKernelGGL(k_narrow<B>, g, b, 0, st, d);
      if (d.any_voxel) hipLaunchKernelGGL(k_narrow_vox<B>, g, b, 0, st, d);
      hipLaunchKernelGGL(k_contacts<B>, g, b, 0, st, d);
      break;
    case 3: hipLaunchKernelGGL(k_post<B>, g, b, 0, st, d); break;
    case 4:
      if (d.any_island) hipLaunchKernelGGL(k_levels<B>, g, b, 0, st, d);
      if (d.any_island) hipLaunchKernelGGL(k_pgs_island<B>, g, b, (size_t)d.pgs_lcap * 16 * kBodyLds, st, d);
      break;
    case 5: hipLaunchKernelGGL(k_avbd<B>, g, b, 0, st, d); break;
    default: hipLaunchKernelGGL(k_finish<B>, g, b, 0, st, d); break;
  }
}
template <int B>
void launch_pgs_b(const Dev& d, hipStream_t st, uint32_t grid) {
  hipLaunchKernelGGL(k_pgs<B>, dim3(grid), dim3(B), (size_t)d.pgs_lcap * 16 * kBodyLds, st, d);
}
template <int B>
void launch_fused_b(const Lane& l, hipStream_t st) {
  hipLaunchKernelGGL(k_fused<B>, dim3(l.grid), dim3(B), (size_t)l.d.pgs_lcap * 16 * kBodyLds, st, l.d);
}
template void launch_stage_b<BB_BLOCK>(int, const Dev&, hipStream_t, uint32_t);
template void launch_pgs_b<BB_BLOCK>(const Dev&, hipStream_t, uint32_t);
template void launch_fused_b<BB_BLOCK>(const Lane&, hipStream_t);
#if BB_BLOCK == 128
// one-wave colour sweeps ride along with the 128-thread unit
template void launch_pgs_b<64>(const Dev&, hipStream_t, uint32_t);
#endif
#else
template __global__ void k_team<kTeamBlock, false>(Dev);
template __global__ void k_team<kTeamBlock, true>(Dev);

static void launch_stage(int stage, const Lane& l, hipStream_t st) {
  if (l.grid == 0) return;
  switch (l.block) {
    case 128: launch_stage_b<128>(stage, l.d, st, l.grid); break;
    case 512: launch_stage_b<512>(stage, l.d, st, l.grid); break;
    case 1024: launch_stage_b<1024>(stage, l.d, st, l.grid); break;
    default: launch_stage_b<256>(stage, l.d, st, l.grid); break;
  }
  if (stage == 4) {
    switch (l.pgs_block ? l.pgs_block : l.block) {
      case 64: launch_pgs_b<64>(l.d, st, l.grid); break;
      case 128: launch_pgs_b<128>(l.d, st, l.grid); break;
      case 512: launch_pgs_b<512>(l.d, st, l.grid); break;
      case 1024: launch_pgs_b<1024>(l.d, st, l.grid); break;
      default: launch_pgs_b<256>(l.d, st, l.grid); break;
    }
  }
}

static void launch_team(const Dev& d, hipStream_t st) {
  if (!d.team_grid) return;
  if (d.team_vox) hipLaunchKernelGGL((k_team<kTeamBlock, true>), dim3(d.team_grid), dim3(kTeamBlock), 0, st, d);
  else hipLaunchKernelGGL((k_team<kTeamBlock, false>), dim3(d.team_grid), dim3(kTeamBlock), 0, st, d);
}

void launch_step(const Lane& l, hipStream_t st, bool team) {
  if (team) launch_team(l.d, st);
  if (l.fused && l.grid) {
    switch (l.block) {
      case 128: launch_fused_b<128>(l, st); break;
      case 512: launch_fused_b<512>(l, st); break;
      case 1024: launch_fused_b<1024>(l, st); break;
      default: launch_fused_b<256>(l, st); break;
    }
    return;
  }
  for (int k = 0; k < kStages; ++k) launch_stage(k, l, st);
}

void launch_step_timed(const Lane* lanes, int nl, hipStream_t st, hipEvent_t* ev) {
  (void)hipEventRecord(ev[0], st);
  launch_team(lanes[0].d, st);
  for (int k = 0; k < kStages; ++k) {
    for (int i = 0; i < nl; ++i) launch_stage(k, lanes[i], st);
    (void)hipEventRecord(ev[k + 1], st);
  }
}

void launch_copy(const float4* src, float4* dst, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_calibrate, dim3(2048), dim3(256), 0, st, src, dst, n);
}

void launch_sum_counts(const Dev& d, unsigned long long* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_counts, dim3(1), dim3(1024), 0, st, d.stats, d.W, out);
}

void launch_instances(const Dev& d, uint32_t w, uint32_t n, float* dst, hipStream_t st) {
  if (n) hipLaunchKernelGGL(k_instances, dim3((n + 255u) / 256u), dim3(256), 0, st, d, w, dst);
}

void launch_lbvh_export(const Dev& d, uint32_t w, uint32_t nodes, void* dst, hipStream_t st) {
  if (nodes)
    hipLaunchKernelGGL(k_lbvh_export, dim3((nodes + 255u) / 256u), dim3(256), 0, st, d, w,
                       static_cast<uint32_t*>(dst));
}

void launch_energy(const Dev& d, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_energy<256>, dim3(d.W), dim3(256), 0, st, d, out);
}

// Workgroups of k_team that can be resident at once on the device.
uint32_t team_capacity() {
  int dev = 0, cus = 0, per = 0, per_vox = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  // the smaller of the two instances' residency: either may be launched
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_team<kTeamBlock, false>, kTeamBlock, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_vox, k_team<kTeamBlock, true>, kTeamBlock, 0);
  if (e != hipSuccess || per <= 0 || per_vox <= 0 || cus <= 0) return 0;
  return (uint32_t)std::min(per, per_vox) * (uint32_t)cus;
}

#endif  // BB_BLOCK_TU
}  // namespace bb
