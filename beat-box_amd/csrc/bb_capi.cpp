// bb_capi.cpp — host side of the C-ABI (include/bb_capi.h).
//
// A bb_world is a bb_batch holding one world.  A batch keeps every world's
// initial descriptors on the host and all simulation state on the device;
// the device layout (bb_dev.hpp) is rebuilt whenever worlds or bodies are
// added, carrying the current body state across the rebuild.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "bb_host.hpp"

using namespace bb;
using namespace bb::host;

struct bb_world {
  bb_batch* b;
};

namespace bb {
namespace host {

uint32_t pair_capacity(uint32_t n) { return 8u * n + 64u; }

bool has_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return false;
  return n > 0;
}

void inertia_of(const bb_world_desc& wd, const bb_body_desc& b, float inv[3]) {
  // SceneManager::load_scene (src/scene_manager.hpp:850-887):
  // inv_mass = mass == 0 ? 0 : 1/mass; cuboid_get_inverse_intertia
  // (include/math.hpp:106-130) called with inv_mass under the parity quirk.
  float mass = b.mass;
  float inv_mass = mass == 0.0f ? 0.0f : 1.0f / mass;
  float m_in = (wd.quirks & BB_QUIRK_INERTIA_FROM_INV_MASS) ? inv_mass : mass;
  inv[0] = inv[1] = inv[2] = 0.0f;
  if (m_in == 0.0f) return;
  if (b.shape == BB_SHAPE_SPHERE) {
    float I = (2.0f / 5.0f) * m_in * (b.radius * b.radius);
    inv[0] = inv[1] = inv[2] = 1.0f / I;
    return;
  }
  float w = b.maximum[0] - b.minimum[0], h = b.maximum[1] - b.minimum[1], d = b.maximum[2] - b.minimum[2];
  float Ixx = (1.0f / 12.0f) * m_in * (h * h + d * d);
  float Iyy = (1.0f / 12.0f) * m_in * (w * w + d * d);
  float Izz = (1.0f / 12.0f) * m_in * (w * w + h * h);
  inv[0] = 1.0f / Ixx;
  inv[1] = 1.0f / Iyy;
  inv[2] = 1.0f / Izz;
}

WorldInfo info_of(const HostWorld& hw) {
  WorldInfo wi{};
  wi.body_off = hw.body_off;
  wi.n = hw.n;
  wi.a = hw.a;
  wi.act_off = hw.act_off;
  wi.pair_off = hw.pair_off;
  wi.pair_cap = hw.pair_cap;
  wi.flags = hw.desc.flags;
  // the colour sweeps implement PGS_SOFT only (RB_sim.slang:1440-1441,
  // :1462-1463); the per-island path (USE_GRAPH_COLORING off) runs plain PGS
  const bool coloring = (hw.desc.flags & BB_SIM_USE_GRAPH_COLORING) != 0;
  wi.solver = (hw.desc.solver == BB_SOLVER_PGS && coloring) ? (uint32_t)BB_SOLVER_PGS_SOFT : hw.desc.solver;
  wi.iterations = hw.desc.iterations;
  wi.morton_bits = hw.desc.morton_bits == 63 ? 63u : 30u;
  wi.quirks = hw.desc.quirks;
  wi.prev_n = hw.n;
  wi.dt = hw.desc.dt;
  wi.gravity = hw.desc.gravity;
  wi.damp = std::exp(-hw.desc.dt * 0.1f);
  wi.mode = hw.has_voxel ? kModeVoxel : 0u;
  return wi;
}

// Reads the newest rows of the batch back into per-world host descriptors so a
// rebuild (new worlds/bodies) continues from the current state.
bb_status pull_state(bb_batch* b) {
  if (!b->built || !b->stepped) return BB_OK;
  const Rows& R = b->rows[b->latest];
  std::vector<uint32_t> id(b->NB), fl(b->NB);
  std::vector<float4> pos(b->NB), rot(b->NB), vel(b->NB), omg(b->NB);
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(id.data(), R.id, b->NB * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(fl.data(), R.flags, b->NB * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(pos.data(), R.pos, b->NB * 16, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(rot.data(), R.rot, b->NB * 16, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(vel.data(), R.vel, b->NB * 16, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(omg.data(), R.omg, b->NB * 16, hipMemcpyDeviceToHost));
  for (auto& hw : b->worlds) {
    for (uint32_t r = 0; r < hw.n; ++r) {
      uint32_t g = hw.body_off + r;
      bb_body_desc& bd = hw.bodies[id[g]];
      for (int k = 0; k < 3; ++k) {
        bd.position[k] = (&pos[g].x)[k];
        bd.velocity[k] = (&vel[g].x)[k];
        bd.omega[k] = (&omg[g].x)[k];
      }
      for (int k = 0; k < 4; ++k) bd.rotation[k] = (&rot[g].x)[k];
    }
    hw.drop_prev = true;
  }
  b->stepped = false;  // the state now lives in the host descriptors
  return BB_OK;
}

// Rows per world whose velocities the PGS sweeps keep in LDS: the tuning
// value, or the largest world when that is smaller.
uint32_t pgs_lds_rows(const bb_batch* b) {
  uint32_t mx = 0;
  for (auto& hw : b->worlds) mx = std::max(mx, (uint32_t)hw.bodies.size());
  return std::min(mx, b->pgs_lds);
}

bb_status alloc_arrays(bb_batch* b, Buffers& A, uint32_t W, uint32_t NB, uint32_t NA, uint32_t NP) {
  Dev& d = b->base;
  d = Dev{};
  d.W = W;
  d.worlds = A.alloc<WorldInfo>(W);
  d.order = A.alloc<uint32_t>(W);
  for (int s = 0; s < 2; ++s) {
    Rows& R = b->rows[s];
    R.id = A.alloc<uint32_t>(NB);
    R.flags = A.alloc<uint32_t>(NB);
    R.sleep = A.alloc<uint32_t>(NB);
    R.pos = A.alloc<float4>(NB);
    R.rot = A.alloc<float4>(NB);
    R.vel = A.alloc<float4>(NB);
    R.omg = A.alloc<float4>(NB);
    R.pvel = A.alloc<float4>(NB);
    b->entries[s] = A.alloc<uint32_t>(NB);
    b->pairs[s] = A.alloc<uint2>(NP);
    b->npairs[s] = A.alloc<uint32_t>(W);
    b->pstart[s] = A.alloc<uint32_t>(NB);
    Manifolds& M = b->man[s];
    M.hdr = A.alloc<int4>(NP);
    M.nrm = A.alloc<float4>(NP);
    M.t0 = A.alloc<float4>(NP);
    M.t1 = A.alloc<float4>(NP);
    M.anca = A.alloc<float4>(NP);
    M.ancb = A.alloc<float4>(NP);
    M.cpos = A.alloc<float4>((size_t)NP * 8);
    M.clam = A.alloc<float4>((size_t)NP * 8);
    M.cmas = A.alloc<float4>((size_t)NP * 8);
    M.cana = A.alloc<float4>((size_t)NP * 8);
    M.canb = A.alloc<float4>((size_t)NP * 8);
    M.cfp = A.alloc<uint4>((size_t)NP * 8);
    M.cws = A.alloc<uint32_t>((size_t)NP * 8);
  }
  Props& P = d.props;
  P.pmin = A.alloc<float4>(NB);
  P.pmax = A.alloc<float4>(NB);
  P.mass = A.alloc<float4>(NB);
  P.ii0 = A.alloc<float4>(NB);
  P.ii1 = A.alloc<float4>(NB);
  P.ii2 = A.alloc<float4>(NB);
  P.active_rank = A.alloc<uint32_t>(NB);
  P.active_ids = A.alloc<uint32_t>(NA);
  d.keys_a = A.alloc<uint64_t>(NB);
  d.keys_b = A.alloc<uint64_t>(NB);
  d.idx_a = A.alloc<uint32_t>(NB);
  d.idx_b = A.alloc<uint32_t>(NB);
  d.nmin = A.alloc<float4>(2 * (size_t)NB);
  d.nmax = A.alloc<float4>(2 * (size_t)NB);
  d.parent = A.alloc<uint32_t>(2 * (size_t)NB);
  d.visit = A.alloc<uint32_t>(2 * (size_t)NB);
  d.pcount = A.alloc<uint32_t>(NB);
  d.poff = A.alloc<uint32_t>(NB);
  d.link = A.alloc<uint32_t>(NA);
  d.deg = A.alloc<uint32_t>(NB);
  d.row_prev = A.alloc<uint32_t>(NB);
  d.row_rank = A.alloc<uint32_t>(NB);
  d.island = A.alloc<uint32_t>(NB);
  d.islot = A.alloc<uint32_t>(NB);
  d.ascratch = A.alloc<uint32_t>(NA);
  d.color = A.alloc<uint32_t>(NP);
  d.pick = A.alloc<uint32_t>(NP);
  d.mask = A.alloc<uint32_t>(NB);
  d.owner = A.alloc<uint32_t>((size_t)NB * kColors);
  d.bucket = A.alloc<uint32_t>(NP);
  d.bucket_off = A.alloc<uint32_t>((size_t)W * (kBuckets + 1));
  d.nlev = A.alloc<uint32_t>(W);
  d.iw = A.alloc<float4>((size_t)NB * 3);
  d.np = NP;
  d.pgs_lcap = 0;  // set after the layout (pgs_lds_rows)
  d.sv_hdr = A.alloc<int4>(NP);
  d.sv_n = A.alloc<float4>(NP);
  d.sv_r1 = A.alloc<float4>((size_t)NP * 8);
  d.sv_r2 = A.alloc<float4>((size_t)NP * 8);
  d.sv_l = A.alloc<float4>((size_t)NP * 8);
  d.sv_tm1 = A.alloc<float>((size_t)NP * 8);
  d.s_pos0 = A.alloc<float4>(NB);
  d.s_rot0 = A.alloc<float4>(NB);
  d.s_post = A.alloc<float4>(NB);
  d.s_rott = A.alloc<float4>(NB);
  d.depth = A.alloc<uint32_t>(NB);
  d.depth_snap = A.alloc<uint32_t>(NB);
  d.bcolor = A.alloc<uint32_t>(NB);
  d.bcolor_snap = A.alloc<uint32_t>(NB);
  d.adj_off = A.alloc<uint32_t>(NB);
  d.adj = A.alloc<uint32_t>(2 * (size_t)NP);
  d.bbucket = A.alloc<uint32_t>(NB);
  d.bbucket_off = A.alloc<uint32_t>((size_t)W * (kAvbdColors * kAvbdLayers + 1));
  d.stats = A.alloc<uint32_t>((size_t)W * ST_COUNT);
  d.ovf_any = A.alloc<uint32_t>(1);
  d.team_spin = b->team_spin;
  d.sort_blocks = b->sort_blocks;
  d.avbd_lanes = b->avbd_lanes;
  d.pgs_prefetch = b->pgs_prefetch;
  d.walk_budget = b->walk_budget;
  d.teams = A.alloc<TeamTable>(1);
  d.team_ctl = A.alloc<TeamCtl>(kMaxTeams);
  d.team_hist = A.alloc<uint32_t>((size_t)kMaxTeamBlocks * 256);
  d.team_shared = A.alloc<char>((size_t)kMaxTeams * kTeamSharedBytes);
  d.team_grid = 0;
  d.team_vox = 0;
  b->d_vshape = A.alloc<VoxShape>(BB_MAX_VOXEL_SHAPES);
  b->d_vocc = A.alloc<uint32_t>(BB_MAX_VOXEL_OCC_WORDS);
  b->d_vsurf = A.alloc<uint32_t>(BB_MAX_VOXEL_SURFACE);
  d.vshape = b->d_vshape;
  d.vocc = b->d_vocc;
  d.vsurf = b->d_vsurf;
  b->vox_dirty = true;
  if (A.count_only) return BB_OK;
  if (std::find(A.ptrs.begin(), A.ptrs.end(), nullptr) != A.ptrs.end()) return BB_ERR_CAPACITY;
  // a carved layout is cleared by its owner, on its own stream
  if (A.carve_base) return BB_OK;
  HIPCHK(hipMemset(d.teams, 0, sizeof(TeamTable)));
  HIPCHK(hipMemset(d.team_ctl, 0, sizeof(TeamCtl) * kMaxTeams));
  // tables that must start at zero: colour owner cells (zero between
  // rounds), counters, previous pair counts and pair-row offsets
  HIPCHK(hipMemset(d.owner, 0, (size_t)NB * kColors * 4));
  HIPCHK(hipMemset(d.stats, 0, (size_t)W * ST_COUNT * 4));
  HIPCHK(hipMemset(d.ovf_any, 0, 4));
  for (int s = 0; s < 2; ++s) {
    HIPCHK(hipMemset(b->npairs[s], 0, (size_t)W * 4));
    HIPCHK(hipMemset(b->pstart[s], 0, (size_t)NB * 4));
    HIPCHK(hipMemset(b->rows[s].sleep, 0, (size_t)NB * 4));
    HIPCHK(hipMemset(b->rows[s].pvel, 0, (size_t)NB * 16));
  }
  return BB_OK;
}

uint64_t device_bytes(uint32_t W, uint32_t NB, uint32_t NA, uint32_t NP) {
  bb_batch tmp;
  Buffers A;
  A.count_only = true;
  (void)alloc_arrays(&tmp, A, W, NB, NA, NP);
  return A.bytes;
}

Images carve_images(void* mem, uint32_t nb, uint32_t na) {
  char* p = (char*)mem;
  Images img;
  auto f4 = [&](float4*& dst) {
    dst = (float4*)p;
    p += (size_t)nb * 16;
  };
  f4(img.pos);
  f4(img.rot);
  f4(img.vel);
  f4(img.omg);
  f4(img.pmin);
  f4(img.pmax);
  f4(img.mass);
  f4(img.ii0);
  f4(img.ii1);
  f4(img.ii2);
  auto u4 = [&](uint32_t*& dst, size_t n) {
    dst = (uint32_t*)p;
    p += n * 4;
  };
  u4(img.id, nb);
  u4(img.flags, nb);
  u4(img.entries, nb);
  u4(img.rank, nb);
  u4(img.active_ids, na);
  return img;
}

void fill_images(const HostWorld& hw, uint32_t r0, uint32_t a0, const Images& img,
                 const std::vector<bb_voxel_shape_info>* vox) {
  uint32_t k = 0;
  for (uint32_t i = 0; i < hw.n; ++i) {
    const bb_body_desc& bd = hw.bodies[i];
    const uint32_t g = r0 + i;
    img.id[g] = i;
    img.flags[g] = bd.flags & (BB_BODY_DYNAMIC | BB_BODY_GRAVITY);
    img.entries[g] = i;
    img.pos[g] = make_float4(bd.position[0], bd.position[1], bd.position[2], 0.0f);
    img.rot[g] = make_float4(bd.rotation[0], bd.rotation[1], bd.rotation[2], bd.rotation[3]);
    img.vel[g] = make_float4(bd.velocity[0], bd.velocity[1], bd.velocity[2], 0.0f);
    img.omg[g] = make_float4(bd.omega[0], bd.omega[1], bd.omega[2], 0.0f);
    const float r = bd.radius;
    if (BB_SHAPE_KIND(bd.shape) == BB_SHAPE_VOXEL && vox) {
      // push_voxel_body (src/scene_manager.hpp:568-572): the shape's bounds,
      // mass and inverse inertia; immovable (mass 0) without DYNAMIC.  pmax.w
      // carries the shape word (exact in float below 2^24).
      const bb_voxel_shape_info& vi = (*vox)[BB_SHAPE_VOXEL_ID(bd.shape) - 1u];
      const bool dyn = (bd.flags & BB_BODY_DYNAMIC) != 0;
      img.pmin[g] = make_float4(vi.minimum[0], vi.minimum[1], vi.minimum[2], 0.0f);
      img.pmax[g] = make_float4(vi.maximum[0], vi.maximum[1], vi.maximum[2], (float)bd.shape);
      img.mass[g] = make_float4(dyn ? vi.mass : 0.0f, dyn ? 1.0f / vi.mass : 0.0f, bd.friction, bd.restitution);
      const float* I = vi.inv_inertia;
      img.ii0[g] = dyn ? make_float4(I[0], I[1], I[2], 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      img.ii1[g] = dyn ? make_float4(I[3], I[4], I[5], 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      img.ii2[g] = dyn ? make_float4(I[6], I[7], I[8], 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (dyn) {
        img.rank[g] = k;
        img.active_ids[a0 + k] = i;
        ++k;
      } else {
        img.rank[g] = kMAX;
      }
      continue;
    }
    if (bd.shape == BB_SHAPE_SPHERE) {
      img.pmin[g] = make_float4(-r, -r, -r, r);
      img.pmax[g] = make_float4(r, r, r, 1.0f);
    } else {
      img.pmin[g] = make_float4(bd.minimum[0], bd.minimum[1], bd.minimum[2], r);
      img.pmax[g] = make_float4(bd.maximum[0], bd.maximum[1], bd.maximum[2], 0.0f);
    }
    const float inv_mass = bd.mass == 0.0f ? 0.0f : 1.0f / bd.mass;
    img.mass[g] = make_float4(bd.mass, inv_mass, bd.friction, bd.restitution);
    float inv[3];
    inertia_of(hw.desc, bd, inv);
    img.ii0[g] = make_float4(inv[0], 0.0f, 0.0f, 0.0f);
    img.ii1[g] = make_float4(0.0f, inv[1], 0.0f, 0.0f);
    img.ii2[g] = make_float4(0.0f, 0.0f, inv[2], 0.0f);
    if (bd.flags & BB_BODY_DYNAMIC) {
      img.rank[g] = k;
      img.active_ids[a0 + k] = i;
      ++k;
    } else {
      img.rank[g] = kMAX;
    }
  }
}

bb_status upload_images(bb_batch* b, const Images& img, uint32_t body_lo, uint32_t nb, uint32_t act_lo,
                        uint32_t na, hipStream_t st) {
  if (nb == 0) return BB_OK;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
  };
  const size_t o = body_lo;
  for (int s = 0; s < 2; ++s) {
    Rows& R = b->rows[s];
    HIPCHK(up(R.id + o, img.id, nb * 4));
    HIPCHK(up(R.flags + o, img.flags, nb * 4));
    HIPCHK(up(R.pos + o, img.pos, nb * 16));
    HIPCHK(up(R.rot + o, img.rot, nb * 16));
    HIPCHK(up(R.vel + o, img.vel, nb * 16));
    HIPCHK(up(R.omg + o, img.omg, nb * 16));
    HIPCHK(up(b->entries[s] + o, img.entries, nb * 4));
  }
  Props& P = b->base.props;
  HIPCHK(up(P.pmin + o, img.pmin, nb * 16));
  HIPCHK(up(P.pmax + o, img.pmax, nb * 16));
  HIPCHK(up(P.mass + o, img.mass, nb * 16));
  HIPCHK(up(P.ii0 + o, img.ii0, nb * 16));
  HIPCHK(up(P.ii1 + o, img.ii1, nb * 16));
  HIPCHK(up(P.ii2 + o, img.ii2, nb * 16));
  HIPCHK(up(P.active_rank + o, img.rank, nb * 4));
  if (na) HIPCHK(up(P.active_ids + act_lo, img.active_ids, na * 4));
  return BB_OK;
}

bb_status reset_world_state(bb_batch* b, uint32_t w, hipStream_t st) {
  const HostWorld& hw = b->worlds[w];
  for (int s = 0; s < 2; ++s) {
    HIPCHK(hipMemsetAsync(b->npairs[s] + w, 0, 4, st));
    if (hw.n) {
      HIPCHK(hipMemsetAsync(b->pstart[s] + hw.body_off, 0, (size_t)hw.n * 4, st));
      HIPCHK(hipMemsetAsync(b->rows[s].sleep + hw.body_off, 0, (size_t)hw.n * 4, st));
      HIPCHK(hipMemsetAsync(b->rows[s].pvel + hw.body_off, 0, (size_t)hw.n * 16, st));
    }
  }
  HIPCHK(hipMemsetAsync(b->base.stats + (size_t)w * ST_COUNT, 0, ST_COUNT * 4, st));
  return BB_OK;
}

uint32_t assign_teams(bb_batch* b, uint32_t grid_cap) {
  TeamTable& T = b->team_table;
  T = TeamTable{};
  b->base.team_vox = 0;
  for (auto& wi : b->infos) wi.mode &= ~kModeTeam;
  if (b->team_bodies == 0) return 0;
  grid_cap = std::min(grid_cap, kMaxTeamBlocks);
  std::vector<uint32_t> cand;
  for (uint32_t w = 0; w < (uint32_t)b->worlds.size(); ++w)
    if (b->infos[w].n >= b->team_bodies && b->infos[w].n > 0 && !b->worlds[w].no_team) cand.push_back(w);
  // largest worlds first; each team gets at least two workgroups
  std::sort(cand.begin(), cand.end(), [&](uint32_t x, uint32_t y) {
    return b->infos[x].n != b->infos[y].n ? b->infos[x].n > b->infos[y].n : x < y;
  });
  while (cand.size() > kMaxTeams || 2 * cand.size() > grid_cap) cand.pop_back();
  if (cand.empty()) return 0;
  std::sort(cand.begin(), cand.end());
  uint64_t want_total = 0;
  std::vector<uint32_t> want(cand.size());
  for (size_t k = 0; k < cand.size(); ++k) {
    want[k] = std::max(2u, (b->infos[cand[k]].n + kTeamBlock - 1u) / kTeamBlock);
    want_total += want[k];
  }
  uint32_t first = 0;
  T.count = (uint32_t)cand.size();
  for (size_t k = 0; k < cand.size(); ++k) {
    uint32_t nb = want[k];
    if (want_total > grid_cap) nb = std::max(2u, (uint32_t)((uint64_t)want[k] * grid_cap / want_total));
    T.first[k] = first;
    T.world[k] = cand[k];
    first += nb;
    b->infos[cand[k]].mode |= kModeTeam;
    if (b->infos[cand[k]].mode & kModeVoxel) b->base.team_vox = 1;
  }
  T.first[cand.size()] = first;
  return first;
}

void world_order(const std::vector<WorldInfo>& infos, uint32_t* out) {
  const uint32_t W = (uint32_t)infos.size();
  for (uint32_t w = 0; w < W; ++w) out[w] = w;
  auto key = [&](uint32_t w) { return (infos[w].mode & kModeTeam) ? 0u : infos[w].n; };
  std::stable_sort(out, out + W, [&](uint32_t x, uint32_t y) { return key(x) > key(y); });
}

// Workgroups a k_team launch may use: BB_PARAM_TEAM_BLOCKS, never more than
// the device holds resident at once (every member of a team must run at the
// same time); by default half of that, leaving room for the lanes' kernels.
uint32_t team_grid_cap(const bb_batch* b) {
  const uint32_t resident = team_capacity();
  if (b->team_blocks) return std::min(b->team_blocks, resident);
  return resident / 2;
}

bb_status upload_teams(bb_batch* b) {
  const uint32_t cap = team_grid_cap(b);
  b->base.team_grid = assign_teams(b, cap);
  HIPCHK(hipMemcpy(b->base.teams, &b->team_table, sizeof(TeamTable), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(b->base.team_ctl, 0, sizeof(TeamCtl) * kMaxTeams));
  HIPCHK(hipMemcpy(b->base.worlds, b->infos.data(), b->infos.size() * sizeof(WorldInfo), hipMemcpyHostToDevice));
  std::vector<uint32_t> order(b->infos.size());
  world_order(b->infos, order.data());
  HIPCHK(hipMemcpy((void*)b->base.order, order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  b->n_wide = 0;
  while (b->lane_bodies && b->n_wide < order.size() && !(b->infos[order[b->n_wide]].mode & kModeTeam) &&
         b->infos[order[b->n_wide]].n >= b->lane_bodies)
    ++b->n_wide;
  return BB_OK;
}

bb_status build(bb_batch* b) {
  if (!b->dirty) return upload_voxels(b);
  if (!has_device()) return BB_ERR_NO_DEVICE;
  bb_status st = pull_state(b);
  if (st != BB_OK) return st;
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  if (!b->stream) HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  if (!b->stream2) HIPCHK(hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking));
  if (!b->ev_fork) HIPCHK(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  if (!b->ev_join) HIPCHK(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->buf.release();
  // layout: worlds back to back
  uint32_t NB = 0, NA = 0, NP = 0;
  for (auto& hw : b->worlds) {
    hw.n = (uint32_t)hw.bodies.size();
    hw.a = 0;
    for (auto& bd : hw.bodies)
      if (bd.flags & BB_BODY_DYNAMIC) ++hw.a;
    hw.body_off = NB;
    hw.act_off = NA;
    hw.pair_off = NP;
    hw.has_voxel = false;
    for (auto& bd : hw.bodies)
      if (BB_SHAPE_KIND(bd.shape) == BB_SHAPE_VOXEL) hw.has_voxel = true;
    // a pair with a voxel body takes six slots, one per contact-face code
    hw.pair_cap = pair_capacity(hw.n) * (hw.has_voxel ? 6u : 1u) * hw.pair_scale * b->pair_scale;
    NB += hw.n;
    NA += hw.a;
    NP += hw.pair_cap;
  }
  b->NB = NB;
  b->NA = NA;
  b->NP = NP;
  const uint32_t W = (uint32_t)b->worlds.size();
  st = alloc_arrays(b, b->buf, W, NB, NA, NP);
  if (st != BB_OK) return st;
  std::vector<char> mem(kImageBodyBytes * (size_t)std::max(NB, 1u) + kImageActiveBytes * std::max(NA, 1u) + 64);
  void* base = (void*)(((uintptr_t)mem.data() + 15) & ~(uintptr_t)15);
  const Images img = carve_images(base, std::max(NB, 1u), std::max(NA, 1u));
  b->infos.resize(W);
  for (uint32_t w = 0; w < W; ++w) {
    HostWorld& hw = b->worlds[w];
    WorldInfo wi = info_of(hw);
    wi.mode |= kModeDropPrev;  // fresh layout: no previous manifolds
    b->infos[w] = wi;
    fill_images(hw, hw.body_off, hw.act_off, img, &b->vox_info);
  }
  st = upload_images(b, img, 0, NB, 0, NA, b->stream);
  if (st != BB_OK) return st;
  HIPCHK(hipStreamSynchronize(b->stream));
  st = upload_voxels(b);
  if (st != BB_OK) return st;
  st = upload_teams(b);  // also uploads the world table
  if (st != BB_OK) return st;
  b->latest = 1;  // the first step reads set 1 and writes set 0
  b->base.pgs_lcap = pgs_lds_rows(b);
  b->base.any_voxel = 0;
  for (auto& hw : b->worlds)
    if (hw.has_voxel) b->base.any_voxel = 1;
  b->base.any_island = any_island(b);
  b->built = true;
  b->dirty = false;
  b->stepped = false;
  return BB_OK;
}

// The world's pair slots start at 0 in a single-world layout, so a larger
// buffer set laid out the same way keeps every slot index: each array of
// the old set is copied into the front of its counterpart (arrays sized by
// the pair capacity grow, the rest are the same size), then the old set is
// freed.  Arrays strided by the capacity (the PGS row scratch) land
// elsewhere, but a step writes them before it reads them.
bb_status grow_pairs(bb_batch* b) {
  if (!b->built || b->dirty || b->worlds.size() != 1) return BB_ERR_STATE;
  HostWorld& hw = b->worlds[0];
  if (hw.pair_cap > (1u << 26)) return BB_ERR_CAPACITY;
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipStreamSynchronize(b->stream2));
  // alloc_arrays repoints these; they are put back if it fails
  const Dev base0 = b->base;
  Rows rows0[2] = {b->rows[0], b->rows[1]};
  Manifolds man0[2] = {b->man[0], b->man[1]};
  uint32_t* entries0[2] = {b->entries[0], b->entries[1]};
  uint2* pairs0[2] = {b->pairs[0], b->pairs[1]};
  uint32_t* npairs0[2] = {b->npairs[0], b->npairs[1]};
  uint32_t* pstart0[2] = {b->pstart[0], b->pstart[1]};
  VoxShape* vshape0 = b->d_vshape;
  uint32_t *vocc0 = b->d_vocc, *vsurf0 = b->d_vsurf;
  const bool vdirty0 = b->vox_dirty;
  const uint32_t cap = hw.pair_cap * 2u;
  Buffers nb;
  bb_status st = alloc_arrays(b, nb, 1, b->NB, b->NA, cap);
  hipError_t rc = hipSuccess;
  if (st == BB_OK) {
    for (size_t i = 0; i < nb.ptrs.size() && rc == hipSuccess; ++i)
      rc = hipMemcpy(nb.ptrs[i], b->buf.ptrs[i], std::min(nb.sizes[i], b->buf.sizes[i]), hipMemcpyDeviceToDevice);
  }
  if (st != BB_OK || rc != hipSuccess) {
    nb.release();
    b->base = base0;
    for (int s = 0; s < 2; ++s) {
      b->rows[s] = rows0[s];
      b->man[s] = man0[s];
      b->entries[s] = entries0[s];
      b->pairs[s] = pairs0[s];
      b->npairs[s] = npairs0[s];
      b->pstart[s] = pstart0[s];
    }
    b->d_vshape = vshape0;
    b->d_vocc = vocc0;
    b->d_vsurf = vsurf0;
    b->vox_dirty = vdirty0;
    return st != BB_OK ? st : BB_ERR_DEVICE;
  }
  b->buf.release();
  b->buf = std::move(nb);
  b->base.pgs_lcap = base0.pgs_lcap;
  b->base.any_voxel = base0.any_voxel;
  b->base.any_island = base0.any_island;
  b->base.team_grid = base0.team_grid;
  b->base.team_vox = base0.team_vox;
  b->vox_dirty = false;  // the pools came over with the rest
  b->NP = cap;
  hw.pair_cap = cap;
  hw.pair_scale *= 2u;
  b->infos[0].pair_cap = cap;
  HIPCHK(hipMemcpy(b->base.worlds, b->infos.data(), sizeof(WorldInfo), hipMemcpyHostToDevice));
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  return BB_OK;
}

Dev dev_for(const bb_batch* b, int cur) {
  Dev d = b->base;
  int prev = cur ^ 1;
  d.cur = b->rows[cur];
  d.prev = b->rows[prev];
  d.entries_cur = b->entries[cur];
  d.entries_prev = b->entries[prev];
  d.pairs_cur = b->pairs[cur];
  d.pairs_prev = b->pairs[prev];
  d.npairs_cur = b->npairs[cur];
  d.npairs_prev = b->npairs[prev];
  d.pstart_cur = b->pstart[cur];
  d.pstart_prev = b->pstart[prev];
  d.man_cur = b->man[cur];
  d.man_prev = b->man[prev];
  return d;
}

uint32_t pick_block(const bb_batch* b) {
  if (b->block) return b->block;
  // worlds that teams or the wide lane step do not count: they have their own
  // workgroup sizes
  uint32_t mx = 0;
  for (auto& hw : b->worlds)
    if ((b->team_bodies == 0 || hw.n < b->team_bodies) && (b->lane_bodies == 0 || hw.n < b->lane_bodies))
      mx = std::max(mx, hw.n);
  if (mx >= 8192) return 1024;
  if (mx >= 2048 && b->worlds.size() < 64) return 512;
  return 256;
}

int lanes_of(const bb_batch* b, const Dev& d, Lane* out) {
  const uint32_t nw = std::min(b->n_wide, d.W);
  int nl = 0;
  if (nw) {
    Dev dw = d;
    uint32_t mx = 0;
    for (auto& hw : b->worlds) mx = std::max(mx, (uint32_t)hw.bodies.size());
    dw.pgs_lcap = std::min(mx, b->lane_lds);
    out[nl++] = Lane{dw, b->lane_block, nw, b->fused != 0};
  }
  Dev rest = d;
  rest.order = d.order + nw;
  out[nl++] = Lane{rest, pick_block(b), d.W - nw, b->fused != 0, b->pgs_block};
  return nl;
}

// One step on the batch stream.  With a wide lane, the two lanes run on two
// streams (fork and join by events), so the large worlds' longer workgroups
// overlap the rest; inside a graph capture the second stream joins the
// capture through the fork event.
bb_status launch_lanes(bb_batch* b, const Dev& d) {
  Lane L[2];
  const int nl = lanes_of(b, d, L);
  if (nl == 1) {
    launch_step(L[0], b->stream, true);
    return BB_OK;
  }
  HIPCHK(hipEventRecord(b->ev_fork, b->stream));
  HIPCHK(hipStreamWaitEvent(b->stream2, b->ev_fork, 0));
  launch_step(L[0], b->stream, true);
  launch_step(L[1], b->stream2, false);
  HIPCHK(hipEventRecord(b->ev_join, b->stream2));
  HIPCHK(hipStreamWaitEvent(b->stream, b->ev_join, 0));
  return BB_OK;
}

bb_status launch_one(bb_batch* b) {
  const int cur = b->latest ^ 1;
  Dev d = dev_for(b, cur);
  const bb_status st = launch_lanes(b, d);
  if (st != BB_OK) return st;
  HIPCHK(hipGetLastError());
  b->latest = cur;
  b->stepped = true;
  ++b->step_index;
  return BB_OK;
}

// One-shot per-world parameters (skip warm start / drop previous manifolds)
// apply to exactly one step: the world table is rewritten before and after.
bool any_one_shot(const bb_batch* b) {
  for (auto& hw : b->worlds)
    if (hw.skip_ws_once || hw.drop_prev) return true;
  for (auto& wi : b->infos)
    if (wi.mode & kModeDropPrev) return true;
  return false;
}

bb_status upload_infos(bb_batch* b, bool one_shot) {
  std::vector<WorldInfo> v = b->infos;
  for (size_t w = 0; w < v.size(); ++w) {
    HostWorld& hw = b->worlds[w];
    if (one_shot) {
      if (hw.skip_ws_once) v[w].flags &= ~(uint32_t)BB_SIM_WARM_STARTING;
      if (hw.drop_prev) v[w].mode |= kModeDropPrev;
    } else {
      v[w].mode &= ~kModeDropPrev;
    }
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(b->base.worlds, v.data(), v.size() * sizeof(WorldInfo), hipMemcpyHostToDevice));
  if (one_shot) {
    for (auto& hw : b->worlds) hw.skip_ws_once = hw.drop_prev = false;
  } else {
    for (auto& wi : b->infos) wi.mode &= ~kModeDropPrev;
  }
  return BB_OK;
}

// Whether any world runs the per-island solver path (k_levels and
// k_pgs_island are then part of the launch sequence).
uint32_t any_island(const bb_batch* b) {
  for (const WorldInfo& wi : b->infos)
    if (wi.solver != BB_SOLVER_AVBD && !(wi.flags & BB_SIM_USE_GRAPH_COLORING)) return 1u;
  return 0u;
}

bb_status refresh_infos(bb_batch* b) {
  for (size_t w = 0; w < b->worlds.size(); ++w) {
    WorldInfo wi = info_of(b->worlds[w]);
    wi.mode = b->infos[w].mode;
    b->infos[w] = wi;
  }
  // a solver or flag change can add or drop the per-island kernels: a
  // captured step graph holds the old launch sequence
  const uint32_t p = any_island(b);
  if (p != b->base.any_island) {
    b->base.any_island = p;
    if (b->graph) {
      (void)hipGraphExecDestroy(b->graph);
      b->graph = nullptr;
    }
  }
  return BB_OK;
}

bb_status step_batch(bb_batch* b, uint32_t nsteps, bool use_graph) {
  if (b->worlds.empty()) return BB_ERR_STATE;
  bb_status st = build(b);
  if (st != BB_OK) return st;
  uint32_t done = 0;
  while (done < nsteps && any_one_shot(b)) {
    st = upload_infos(b, true);
    if (st != BB_OK) return st;
    st = launch_one(b);
    if (st != BB_OK) return st;
    st = upload_infos(b, false);
    if (st != BB_OK) return st;
    ++done;
  }
  if (!use_graph) {
    for (; done < nsteps; ++done) {
      st = launch_one(b);
      if (st != BB_OK) return st;
    }
    return BB_OK;
  }
  // hipGraph: two consecutive steps (both parities) captured once, replayed
  const int first = b->latest ^ 1;
  if (b->graph && b->graph_parity != first) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  if (!b->graph && nsteps - done >= 2) {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    bb_status cst = launch_lanes(b, dev_for(b, first));
    if (cst == BB_OK) cst = launch_lanes(b, dev_for(b, first ^ 1));
    HIPCHK(hipStreamEndCapture(b->stream, &g));
    if (cst != BB_OK) {
      (void)hipGraphDestroy(g);
      return cst;
    }
    HIPCHK(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    b->graph_parity = first;
  }
  while (nsteps - done >= 2 && b->graph) {
    HIPCHK(hipGraphLaunch(b->graph, b->stream));
    b->stepped = true;
    b->step_index += 2;
    done += 2;
  }
  for (; done < nsteps; ++done) {
    st = launch_one(b);
    if (st != BB_OK) return st;
  }
  return BB_OK;
}

bb_status profile_batch(bb_batch* b, uint32_t nsteps, float* step_ms, float* kernel_ms, uint64_t* counts) {
  if (b->worlds.empty()) return BB_ERR_STATE;
  bb_status st = build(b);
  if (st != BB_OK) return st;
  while (any_one_shot(b)) {  // the one-shot step runs untimed
    st = step_batch(b, 1, false);
    if (st != BB_OK) return st;
  }
  if (kernel_ms)
    for (int k = 0; k < BB_KERNEL_COUNT; ++k) kernel_ms[k] = 0.0f;
  for (int k = 0; k < BB_BUILD_COUNT; ++k) b->build_ms[k] = 0.0f;
  for (int k = 0; k < BB_AVBD_COUNT; ++k) b->avbd_ms[k] = 0.0f;
  unsigned long long* dcounts = nullptr;
  if (counts && nsteps &&
      hipMalloc(&dcounts, (size_t)nsteps * BB_COUNT_FIELDS * sizeof(unsigned long long)) != hipSuccess)
    return BB_ERR_DEVICE;
  // team stage clocks: the team kernel runs before the stage kernels in a
  // timed step, so its whole time lands in the first stage's interval; the
  // slowest team's own marks split it by stage
  const size_t nclk = (size_t)kMaxTeams * kTeamMarks;
  unsigned long long* dclock = nullptr;
  if (kernel_ms && nsteps && b->base.team_grid) {
    if (hipMalloc(&dclock, nsteps * nclk * sizeof(unsigned long long)) != hipSuccess) {
      (void)hipFree(dcounts);
      return BB_ERR_DEVICE;
    }
    HIPCHK(hipMemset(dclock, 0, nsteps * nclk * sizeof(unsigned long long)));
  }
  std::vector<hipEvent_t> ev((size_t)nsteps * 8);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  for (uint32_t s = 0; s < nsteps; ++s) {
    const int cur = b->latest ^ 1;
    Dev d = dev_for(b, cur);
    if (dclock) d.team_clock = dclock + s * nclk;
    Lane L[2];
    const int nl = lanes_of(b, d, L);
    launch_step_timed(L, nl, b->stream, &ev[(size_t)s * 8]);
    if (dcounts) launch_sum_counts(d, dcounts + (size_t)s * BB_COUNT_FIELDS, b->stream);
    b->latest = cur;
    b->stepped = true;
    ++b->step_index;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  for (uint32_t s = 0; s < nsteps; ++s) {
    hipEvent_t* e = &ev[(size_t)s * 8];
    float total = 0.0f;
    HIPCHK(hipEventElapsedTime(&total, e[0], e[7]));
    if (step_ms) step_ms[s] = total;
    for (int k = 0; k < BB_KERNEL_COUNT && kernel_ms; ++k) {
      float t = 0.0f;
      HIPCHK(hipEventElapsedTime(&t, e[k], e[k + 1]));
      kernel_ms[k] += t;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (dclock) {
    std::vector<unsigned long long> clk(nsteps * nclk);
    int khz = 0;
    hipError_t rc = hipMemcpy(clk.data(), dclock, clk.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    (void)hipFree(dclock);
    int dev = 0;
    if (rc == hipSuccess) rc = hipGetDevice(&dev);
    if (rc == hipSuccess) rc = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
    if (rc != hipSuccess || khz <= 0) {
      (void)hipFree(dcounts);
      return BB_ERR_DEVICE;
    }
    const uint32_t nt = std::min(b->team_table.count, kMaxTeams);
    for (uint32_t s = 0; s < nsteps; ++s) {
      const unsigned long long* c = &clk[s * nclk];
      uint32_t slow = nt;
      unsigned long long span = 0;
      for (uint32_t j = 0; j < nt; ++j) {
        const unsigned long long* m = c + (size_t)j * kTeamMarks;
        const unsigned long long last = m[kTeamStageMarks - 1];
        if (last > m[0] && last - m[0] > span) {
          span = last - m[0];
          slow = j;
        }
      }
      if (slow == nt) continue;
      const unsigned long long* m = c + (size_t)slow * kTeamMarks;
      const float ms_per_tick = 1.0f / (float)khz;
      kernel_ms[0] = std::max(0.0f, kernel_ms[0] - (float)span * ms_per_tick);
      for (uint32_t k = 0; k + 1 < kTeamStageMarks; ++k) kernel_ms[k] += (float)(m[k + 1] - m[k]) * ms_per_tick;
      // inside the build stage (a world without bodies leaves these unset)
      const unsigned long long sub[BB_BUILD_COUNT + 1] = {m[0], m[kMarkKeys], m[kMarkSort], m[kMarkHierarchy],
                                                          m[kMarkBounds], m[kMarkClimb], m[1]};
      bool ordered = true;
      for (int k = 0; k < BB_BUILD_COUNT; ++k) ordered = ordered && sub[k + 1] >= sub[k];
      if (ordered)
        for (int k = 0; k < BB_BUILD_COUNT; ++k) b->build_ms[k] += (float)(sub[k + 1] - sub[k]) * ms_per_tick;
      // inside the AVBD stage (PGS worlds leave these unset)
      const unsigned long long av[BB_AVBD_COUNT + 1] = {m[5], m[kMarkAvColour], m[kMarkAvGroup], m[kMarkAvSweeps],
                                                         m[6]};
      ordered = true;
      for (int k = 0; k < BB_AVBD_COUNT; ++k) ordered = ordered && av[k + 1] >= av[k] && av[k] != 0;
      if (ordered)
        for (int k = 0; k < BB_AVBD_COUNT; ++k) b->avbd_ms[k] += (float)(av[k + 1] - av[k]) * ms_per_tick;
    }
  }
  if (dcounts) {
    const hipError_t rc =
        hipMemcpy(counts, dcounts, (size_t)nsteps * BB_COUNT_FIELDS * sizeof(uint64_t), hipMemcpyDeviceToHost);
    (void)hipFree(dcounts);
    if (rc != hipSuccess) return BB_ERR_DEVICE;
  }
  return BB_OK;
}

bb_status energy_batch(bb_batch* b, double* out, uint32_t n) {
  if (!out || n < b->worlds.size()) return BB_ERR_INVALID;
  if (b->worlds.empty()) return BB_OK;
  bb_status st = build(b);
  if (st != BB_OK) return st;
  double* dout = nullptr;
  if (hipMalloc(&dout, b->worlds.size() * sizeof(double)) != hipSuccess) return BB_ERR_DEVICE;
  launch_energy(dev_for(b, b->latest), dout, b->stream);
  hipError_t rc = hipGetLastError();
  if (rc == hipSuccess) rc = hipStreamSynchronize(b->stream);
  if (rc == hipSuccess) rc = hipMemcpy(out, dout, b->worlds.size() * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(dout);
  return rc == hipSuccess ? BB_OK : BB_ERR_DEVICE;
}

bb_status instances_batch(bb_batch* b, uint32_t w, float* dst, uint32_t n, int on_device) {
  if (!dst || w >= b->worlds.size()) return BB_ERR_INVALID;
  const uint32_t N = (uint32_t)b->worlds[w].bodies.size();
  if (n < N) return BB_ERR_CAPACITY;
  bb_status st = build(b);
  if (st != BB_OK) return st;
  const Dev d = dev_for(b, b->latest);
  if (on_device) {  // stream-ordered after the steps; bb_batch_sync waits for it
    launch_instances(d, w, N, dst, b->stream);
    HIPCHK(hipGetLastError());
    return BB_OK;
  }
  float* tmp = nullptr;
  if (hipMalloc(&tmp, (size_t)std::max(N, 1u) * 12 * sizeof(float)) != hipSuccess) return BB_ERR_DEVICE;
  launch_instances(d, w, N, tmp, b->stream);
  hipError_t rc = hipGetLastError();
  if (rc == hipSuccess) rc = hipStreamSynchronize(b->stream);
  if (rc == hipSuccess) rc = hipMemcpy(dst, tmp, (size_t)N * 12 * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(tmp);
  return rc == hipSuccess ? BB_OK : BB_ERR_DEVICE;
}

bb_status tap_world(bb_batch* b, uint32_t w, bb_tap kind, void* dst, size_t cap, size_t* written);

// The BVH of world w's last step as 2n-1 LBVHNode records (the debug tap's
// conversion, done by a kernel for a device destination).
bb_status lbvh_batch(bb_batch* b, uint32_t w, bb_lbvh_node* dst, uint32_t cap, uint32_t* count,
                     int on_device) {
  if (w >= b->worlds.size() || !count) return BB_ERR_INVALID;
  bb_status st = build(b);
  if (st != BB_OK) return st;
  const uint32_t N = b->worlds[w].n;
  const uint32_t nodes = (b->stepped && N) ? 2 * N - 1 : 0;
  *count = nodes;
  if (!dst) return BB_OK;
  if (cap < nodes) return BB_ERR_CAPACITY;
  if (!on_device) {
    size_t written = 0;
    return tap_world(b, w, BB_TAP_LBVH, dst, (size_t)cap * sizeof(bb_lbvh_node), &written);
  }
  launch_lbvh_export(dev_for(b, b->latest), w, nodes, dst, b->stream);
  HIPCHK(hipGetLastError());
  return BB_OK;
}

// ---- checkpoints.  A world's blob holds what its next step reads from the
// device (the newest parity's body rows in row order, the id -> row map, the
// pair list with its per-row offsets, the manifold slots that warm start the
// next step) plus its counters, island slots (for the read-back), step index
// and the world and body descriptors the device layout is rebuilt from.
struct SnapHeader {
  char magic[8];
  uint32_t version, n, npairs, warm;
  uint64_t step_index;
  bb_world_desc desc;
  uint32_t nshapes, shape_words;  // voxel shapes and their occupancy words
};
constexpr char kSnapMagic[8] = {'B', 'B', 'S', 'N', 'A', 'P', '0', '2'};
constexpr uint32_t kSnapVersion = 2;
// a voxel shape's record in the blob: dims, cell size, density, word count;
// its occupancy words follow
struct SnapShape {
  uint32_t dims[3];
  float voxel_size, density;
  uint32_t words;
};
// bytes of the arrays of n rows and p manifold slots, after the header and
// the voxel shape section
size_t snap_bytes(uint32_t n, uint32_t p, uint32_t nshapes = 0, uint32_t shape_words = 0) {
  return sizeof(SnapHeader) + (size_t)nshapes * sizeof(SnapShape) + (size_t)shape_words * 4 +
         (size_t)n * sizeof(bb_body_desc) + world_state_bytes(n, p);
}

size_t world_state_bytes(uint32_t n, uint32_t p) {
  return (size_t)n * (6 * 4 + 5 * 16) + (size_t)p * (8 + 6 * 16) + (size_t)p * 8 * (6 * 16 + 4) + ST_COUNT * 4;
}

// The state arrays of world slot w in parity `parity`, in blob order, each
// with its device address and size.
void world_state_arrays(bb_batch* b, uint32_t w, int parity, uint32_t n, uint32_t p,
                        const std::function<void(void*, size_t)>& fn) {
  const HostWorld& hw = b->worlds[w];
  const uint32_t g0 = hw.body_off, s0 = hw.pair_off;
  const Rows& R = b->rows[parity];
  const Manifolds& M = b->man[parity];
  fn((void*)(R.id + g0), (size_t)n * 4);
  fn((void*)(R.flags + g0), (size_t)n * 4);
  fn((void*)(R.sleep + g0), (size_t)n * 4);
  fn((void*)(b->entries[parity] + g0), (size_t)n * 4);
  fn((void*)(b->pstart[parity] + g0), (size_t)n * 4);
  fn((void*)(b->base.islot + g0), (size_t)n * 4);
  fn((void*)(R.pos + g0), (size_t)n * 16);
  fn((void*)(R.rot + g0), (size_t)n * 16);
  fn((void*)(R.vel + g0), (size_t)n * 16);
  fn((void*)(R.omg + g0), (size_t)n * 16);
  fn((void*)(R.pvel + g0), (size_t)n * 16);
  fn((void*)(b->pairs[parity] + s0), (size_t)p * 8);
  for (float4* a : {M.nrm, M.t0, M.t1, M.anca, M.ancb}) fn((void*)(a + s0), (size_t)p * 16);
  fn((void*)(M.hdr + s0), (size_t)p * 16);
  for (float4* a : {M.cpos, M.clam, M.cmas, M.cana, M.canb}) fn((void*)(a + (size_t)s0 * 8), (size_t)p * 8 * 16);
  fn((void*)(M.cfp + (size_t)s0 * 8), (size_t)p * 8 * 16);
  fn((void*)(M.cws + (size_t)s0 * 8), (size_t)p * 8 * 4);
  fn((void*)(b->base.stats + (size_t)w * ST_COUNT), (size_t)ST_COUNT * 4);
}

// The voxel shapes of a batch, rebuilt on restore by the same
// build_voxel_shape: per shape its SnapShape record, then its occupancy words.
char* put_shapes(const bb_batch* b, char* o) {
  for (size_t k = 0; k < b->vox_shapes.size(); ++k) {
    const VoxShape& vs = b->vox_shapes[k];
    const uint32_t cells = vs.dims[0] * vs.dims[1] * vs.dims[2];
    SnapShape r{{vs.dims[0], vs.dims[1], vs.dims[2]}, vs.vs, b->vox_density[k], (cells + 31u) / 32u};
    std::memcpy(o, &r, sizeof(r));
    o += sizeof(r);
    std::memcpy(o, b->vox_occ.data() + vs.occ_off, (size_t)r.words * 4);
    o += (size_t)r.words * 4;
  }
  return o;
}

// Adds a blob's shape section (nshapes records, shape_words occupancy words
// in all, within [in, end)) to b as shapes 1, 2, ... in blob order.  Returns
// the position after the section, or null for a section that does not add up
// or a shape that is refused.
const char* take_shapes(bb_batch* b, const char* in, const char* end, uint32_t nshapes, uint32_t shape_words) {
  if (nshapes > BB_MAX_VOXEL_SHAPES || shape_words > BB_MAX_VOXEL_OCC_WORDS) return nullptr;
  uint32_t words_seen = 0;
  for (uint32_t k = 0; k < nshapes; ++k) {
    SnapShape r;
    if ((size_t)(end - in) < sizeof(r)) return nullptr;
    std::memcpy(&r, in, sizeof(r));
    in += sizeof(r);
    const uint64_t cells = (uint64_t)r.dims[0] * r.dims[1] * r.dims[2];
    if (r.words > shape_words - words_seen || r.words != (cells + 31u) / 32u ||
        (size_t)(end - in) < (size_t)r.words * 4)
      return nullptr;
    words_seen += r.words;
    std::vector<uint32_t> occ(r.words);
    std::memcpy(occ.data(), in, (size_t)r.words * 4);
    in += (size_t)r.words * 4;
    bb_voxel_shape_desc vd{{r.dims[0], r.dims[1], r.dims[2]}, r.voxel_size, r.density, occ.data()};
    bb_voxel_shape_info info{};
    if (add_voxel_shape(b, &vd, &info) != BB_OK || info.shape_id != k + 1) return nullptr;
  }
  return words_seen == shape_words ? in : nullptr;
}

size_t shapes_bytes(const bb_batch* b) { return b->vox_shapes.size() * sizeof(SnapShape) + b->vox_occ.size() * 4; }

size_t shapes_section_bytes(const char* in, const char* end, uint32_t nshapes, uint32_t shape_words) {
  if (nshapes > BB_MAX_VOXEL_SHAPES || shape_words > BB_MAX_VOXEL_OCC_WORDS) return 0;
  const char* at = in;
  uint32_t words_seen = 0;
  for (uint32_t k = 0; k < nshapes; ++k) {
    SnapShape r;
    if ((size_t)(end - at) < sizeof(r)) return 0;
    std::memcpy(&r, at, sizeof(r));
    at += sizeof(r);
    const uint64_t cells = (uint64_t)r.dims[0] * r.dims[1] * r.dims[2];
    if (r.words > shape_words - words_seen || r.words != (cells + 31u) / 32u || (size_t)(end - at) < (size_t)r.words * 4)
      return 0;
    words_seen += r.words;
    at += (size_t)r.words * 4;
  }
  return words_seen == shape_words ? (size_t)(at - in) : 0;
}

bb_status snapshot_world(bb_batch* b, void* dst, size_t cap, size_t* size) {
  if (!size || b->worlds.size() != 1) return BB_ERR_INVALID;
  bb_status st = build(b);
  if (st != BB_OK) return st;
  HIPCHK(hipStreamSynchronize(b->stream));
  const HostWorld& hw = b->worlds[0];
  const uint32_t n = hw.n;
  uint32_t p = 0;
  // warm-start state exists once a step has run and no one-shot drop is pending
  const bool warm = b->stepped && !hw.drop_prev && !(b->infos[0].mode & kModeDropPrev);
  if (warm) HIPCHK(hipMemcpy(&p, b->npairs[b->latest], 4, hipMemcpyDeviceToHost));
  const uint32_t nshapes = (uint32_t)b->vox_shapes.size(), shape_words = (uint32_t)b->vox_occ.size();
  *size = snap_bytes(n, p, nshapes, shape_words);
  if (!dst) return BB_OK;
  if (cap < *size) return BB_ERR_CAPACITY;
  SnapHeader h{};
  std::memcpy(h.magic, kSnapMagic, 8);
  h.version = kSnapVersion;
  h.n = n;
  h.npairs = p;
  h.warm = (warm ? 1u : 0u) | (hw.skip_ws_once ? 2u : 0u);
  h.step_index = b->step_index;
  h.desc = hw.desc;
  h.nshapes = nshapes;
  h.shape_words = shape_words;
  char* o = (char*)dst;
  std::memcpy(o, &h, sizeof(h));
  o = put_shapes(b, o + sizeof(h));
  std::memcpy(o, hw.bodies.data(), (size_t)n * sizeof(bb_body_desc));
  o += (size_t)n * sizeof(bb_body_desc);
  hipError_t rc = hipSuccess;
  world_state_arrays(b, 0, b->latest, n, p, [&](void* src, size_t bytes) {
    if (rc == hipSuccess && bytes) rc = hipMemcpy(o, src, bytes, hipMemcpyDeviceToHost);
    o += bytes;
  });
  return rc == hipSuccess ? BB_OK : BB_ERR_DEVICE;
}

// Checks, on the host, every index of a blob's arrays (in = the first array,
// after the body descriptors) that the next step uses to address device
// memory: ids are a permutation of [0, n), the id -> row map agrees with
// them, pair rows are below n, and every manifold slot with contacts names
// rows below n and holds at most 8 contacts.  A blob that fails is refused
// before anything reaches the device.
bool valid_blob_indices(const char* in, uint32_t n, uint32_t p) {
  auto u32 = [&](size_t byte_off, size_t i) {
    uint32_t v;
    std::memcpy(&v, in + byte_off + i * 4, 4);
    return v;
  };
  const size_t o_id = 0, o_ent = (size_t)n * 12, o_pairs = (size_t)n * 104;
  const size_t o_hdr = o_pairs + (size_t)p * 88;
  std::vector<uint8_t> seen(n, 0);
  for (uint32_t r = 0; r < n; ++r) {
    const uint32_t id = u32(o_id, r);
    if (id >= n || seen[id]) return false;
    seen[id] = 1;
  }
  for (uint32_t r = 0; r < n; ++r) {
    const uint32_t id = u32(o_id, r);
    if (u32(o_ent, id) != r) return false;
  }
  for (uint32_t k = 0; k < p; ++k) {
    if (u32(o_pairs, 2 * (size_t)k) >= n || u32(o_pairs, 2 * (size_t)k + 1) >= n) return false;
    int32_t h[4];
    std::memcpy(h, in + o_hdr + (size_t)k * 16, 16);
    if (h[3] > (int32_t)kMaxContacts) return false;
    if (h[3] > 0 && ((uint32_t)h[0] >= n || (uint32_t)h[1] >= n)) return false;
  }
  return true;
}

bb_status restore_world(const void* src, size_t size, int warm_start, bb_world** out) {
  if (!src || !out || size < sizeof(SnapHeader)) return BB_ERR_INVALID;
  SnapHeader h;
  std::memcpy(&h, src, sizeof(h));
  if (std::memcmp(h.magic, kSnapMagic, 8) != 0 || h.version != kSnapVersion) return BB_ERR_INVALID;
  if (h.nshapes > BB_MAX_VOXEL_SHAPES || h.shape_words > BB_MAX_VOXEL_OCC_WORDS) return BB_ERR_INVALID;
  // at most six slots per body pair (voxel pairs)
  if (size != snap_bytes(h.n, h.npairs, h.nshapes, h.shape_words) ||
      (uint64_t)h.npairs > 3ull * h.n * (h.n ? h.n - 1ull : 0ull))
    return BB_ERR_INVALID;
  const char* in = (const char*)src + sizeof(h);
  bb_world* w = nullptr;
  bb_status st = bb_world_create(&h.desc, &w);
  if (st != BB_OK) return st;
  auto fail = [&](bb_status s) {
    bb_world_destroy(w);
    return s;
  };
  in = take_shapes(w->b, in, (const char*)src + size, h.nshapes, h.shape_words);
  if (!in) return fail(BB_ERR_INVALID);
  if (h.n) {
    st = bb_world_add_bodies(w, (const bb_body_desc*)in, h.n, nullptr);
    if (st != BB_OK) return fail(st);
  }
  in += (size_t)h.n * sizeof(bb_body_desc);
  if (!valid_blob_indices(in, h.n, h.npairs)) return fail(BB_ERR_INVALID);
  bb_batch* b = w->b;
  // a world whose pair capacity grew comes back with room for its pairs
  while ((uint64_t)pair_capacity(h.n) * b->worlds[0].pair_scale < h.npairs) b->worlds[0].pair_scale *= 2u;
  st = build(b);
  if (st != BB_OK) return fail(st);
  if (h.npairs > b->worlds[0].pair_cap) return fail(BB_ERR_INVALID);
  HIPCHK(hipStreamSynchronize(b->stream));
  hipError_t rc = hipSuccess;
  world_state_arrays(b, 0, b->latest, h.n, h.npairs, [&](void* dst, size_t bytes) {
    if (rc == hipSuccess && bytes) rc = hipMemcpy(dst, in, bytes, hipMemcpyHostToDevice);
    in += bytes;
  });
  if (rc != hipSuccess) return fail(BB_ERR_DEVICE);
  b->step_index = h.step_index;
  b->stepped = true;
  HostWorld& hw = b->worlds[0];
  const bool warm = warm_start && (h.warm & 1u);
  if (warm) {
    // the restored pairs and manifolds warm start the next step
    const uint32_t np = h.npairs;
    if (hipMemcpy(b->npairs[b->latest], &np, 4, hipMemcpyHostToDevice) != hipSuccess) return fail(BB_ERR_DEVICE);
    b->infos[0].mode &= ~kModeDropPrev;
    if (hipMemcpy(b->base.worlds, b->infos.data(), sizeof(WorldInfo), hipMemcpyHostToDevice) != hipSuccess)
      return fail(BB_ERR_DEVICE);
    hw.skip_ws_once = (h.warm & 2u) != 0;
  } else {
    // like a resume after a pause: the next step starts without warm start
    hw.skip_ws_once = true;
  }
  *out = w;
  return BB_OK;
}

// ---- batch checkpoints.  A batch's blob holds its voxel shapes and pair
// scale, then one record per world in world order: the world's counts, warm
// flags and pair scale, its world and body descriptors and its state arrays
// (world_state_arrays, as in a world's blob).
struct BatchSnapHeader {
  char magic[8];
  uint32_t version, nworlds, nshapes, shape_words, pair_scale, reserved;
  uint64_t step_index;
};
struct BatchSnapWorld {
  uint32_t n, npairs;
  uint32_t warm;  // bit 0: the state warm starts the next step; bit 1: skip_ws_once
  uint32_t pair_scale, no_team, reserved;
  bb_world_desc desc;
};
constexpr char kBatchSnapMagic[8] = {'B', 'B', 'B', 'A', 'T', 'C', 'H', '1'};
constexpr uint32_t kBatchSnapVersion = 1;
constexpr uint32_t kMaxWorldPairScale = 1u << 20;

bb_status snapshot_batch(bb_batch* b, void* dst, size_t cap, size_t* size) {
  if (!size) return BB_ERR_INVALID;
  bb_status st = build(b);
  if (st != BB_OK) return st;
  HIPCHK(hipStreamSynchronize(b->stream));
  const uint32_t nw = (uint32_t)b->worlds.size();
  std::vector<uint32_t> np(nw, 0);
  if (b->stepped && nw) HIPCHK(hipMemcpy(np.data(), b->npairs[b->latest], (size_t)nw * 4, hipMemcpyDeviceToHost));
  size_t total = sizeof(BatchSnapHeader) + b->vox_shapes.size() * sizeof(SnapShape) + b->vox_occ.size() * 4;
  for (uint32_t w = 0; w < nw; ++w) {
    const HostWorld& hw = b->worlds[w];
    // warm-start state exists once a step has run and no one-shot drop is pending
    if (!(b->stepped && !hw.drop_prev && !(b->infos[w].mode & kModeDropPrev))) np[w] = 0;
    total += sizeof(BatchSnapWorld) + (size_t)hw.n * sizeof(bb_body_desc) + world_state_bytes(hw.n, np[w]);
  }
  *size = total;
  if (!dst) return BB_OK;
  if (cap < total) return BB_ERR_CAPACITY;
  BatchSnapHeader h{};
  std::memcpy(h.magic, kBatchSnapMagic, 8);
  h.version = kBatchSnapVersion;
  h.nworlds = nw;
  h.nshapes = (uint32_t)b->vox_shapes.size();
  h.shape_words = (uint32_t)b->vox_occ.size();
  h.pair_scale = b->pair_scale;
  h.step_index = b->step_index;
  char* o = (char*)dst;
  std::memcpy(o, &h, sizeof(h));
  o = put_shapes(b, o + sizeof(h));
  hipError_t rc = hipSuccess;
  for (uint32_t w = 0; w < nw && rc == hipSuccess; ++w) {
    const HostWorld& hw = b->worlds[w];
    const bool warm = b->stepped && !hw.drop_prev && !(b->infos[w].mode & kModeDropPrev);
    BatchSnapWorld r{};
    r.n = hw.n;
    r.npairs = np[w];
    r.warm = (warm ? 1u : 0u) | (hw.skip_ws_once ? 2u : 0u);
    r.pair_scale = hw.pair_scale;
    r.no_team = hw.no_team ? 1u : 0u;
    r.desc = hw.desc;
    std::memcpy(o, &r, sizeof(r));
    o += sizeof(r);
    std::memcpy(o, hw.bodies.data(), (size_t)hw.n * sizeof(bb_body_desc));
    o += (size_t)hw.n * sizeof(bb_body_desc);
    world_state_arrays(b, w, b->latest, hw.n, np[w], [&](void* src, size_t bytes) {
      if (rc == hipSuccess && bytes) rc = hipMemcpy(o, src, bytes, hipMemcpyDeviceToHost);
      o += bytes;
    });
  }
  return rc == hipSuccess ? BB_OK : BB_ERR_DEVICE;
}

// Every record of the blob is read and checked on the host (descriptors by
// bb_batch_add_world, indices by valid_blob_indices, sizes against the blob's
// length) before anything reaches the device.
bb_status restore_batch(const void* src, size_t size, int warm_start, bb_batch** out) {
  if (!src || !out || size < sizeof(BatchSnapHeader)) return BB_ERR_INVALID;
  BatchSnapHeader h;
  std::memcpy(&h, src, sizeof(h));
  if (std::memcmp(h.magic, kBatchSnapMagic, 8) != 0 || h.version != kBatchSnapVersion) return BB_ERR_INVALID;
  const char* in = (const char*)src + sizeof(h);
  const char* end = (const char*)src + size;
  // every world record takes at least its fixed part
  if (h.nworlds > (size_t)(end - in) / sizeof(BatchSnapWorld)) return BB_ERR_INVALID;
  bb_batch* b = nullptr;
  bb_status st = bb_batch_create(h.nworlds, 0, &b);
  if (st != BB_OK) return st;
  auto fail = [&](bb_status s) {
    bb_batch_destroy(b);
    return s;
  };
  st = bb_batch_set_param(b, BB_PARAM_PAIR_SCALE, h.pair_scale);
  if (st != BB_OK) return fail(BB_ERR_INVALID);
  in = take_shapes(b, in, end, h.nshapes, h.shape_words);
  if (!in) return fail(BB_ERR_INVALID);
  struct Rec {
    BatchSnapWorld r;
    const char* state;
  };
  std::vector<Rec> recs(h.nworlds);
  for (uint32_t w = 0; w < h.nworlds; ++w) {
    BatchSnapWorld& r = recs[w].r;
    if ((size_t)(end - in) < sizeof(r)) return fail(BB_ERR_INVALID);
    std::memcpy(&r, in, sizeof(r));
    in += sizeof(r);
    // at most six slots per body pair (voxel pairs)
    if (r.pair_scale == 0 || r.pair_scale > kMaxWorldPairScale || (r.pair_scale & (r.pair_scale - 1)) ||
        r.no_team > 1 || r.warm > 3 || (uint64_t)r.npairs > 3ull * r.n * (r.n ? r.n - 1ull : 0ull) ||
        (size_t)(end - in) / sizeof(bb_body_desc) < r.n)
      return fail(BB_ERR_INVALID);
    const char* bodies = in;
    in += (size_t)r.n * sizeof(bb_body_desc);
    if ((size_t)(end - in) < world_state_bytes(r.n, r.npairs) || !valid_blob_indices(in, r.n, r.npairs))
      return fail(BB_ERR_INVALID);
    recs[w].state = in;
    in += world_state_bytes(r.n, r.npairs);
    std::vector<bb_body_desc> bd(r.n);
    if (r.n) std::memcpy(bd.data(), bodies, (size_t)r.n * sizeof(bb_body_desc));
    st = bb_batch_add_world(b, &r.desc, bd.data(), r.n, nullptr);
    if (st != BB_OK) return fail(st);
    b->worlds[w].pair_scale = r.pair_scale;
    b->worlds[w].no_team = r.no_team != 0;
  }
  if (in != end) return fail(BB_ERR_INVALID);
  st = build(b);
  if (st != BB_OK) return fail(st);
  for (uint32_t w = 0; w < h.nworlds; ++w)
    if (recs[w].r.npairs > b->worlds[w].pair_cap) return fail(BB_ERR_INVALID);
  HIPCHK(hipStreamSynchronize(b->stream));
  hipError_t rc = hipSuccess;
  std::vector<uint32_t> np(h.nworlds, 0);
  for (uint32_t w = 0; w < h.nworlds && rc == hipSuccess; ++w) {
    const Rec& R = recs[w];
    const char* at = R.state;
    world_state_arrays(b, w, b->latest, R.r.n, R.r.npairs, [&](void* dst, size_t bytes) {
      if (rc == hipSuccess && bytes) rc = hipMemcpy(dst, at, bytes, hipMemcpyHostToDevice);
      at += bytes;
    });
    HostWorld& hw = b->worlds[w];
    if (warm_start && (R.r.warm & 1u)) {
      // the restored pairs and manifolds warm start the next step
      np[w] = R.r.npairs;
      b->infos[w].mode &= ~kModeDropPrev;
      hw.skip_ws_once = (R.r.warm & 2u) != 0;
    } else {
      // like a resume after a pause: the next step starts without warm start
      hw.skip_ws_once = true;
    }
  }
  if (rc == hipSuccess && h.nworlds) {
    rc = hipMemcpy(b->npairs[b->latest], np.data(), (size_t)h.nworlds * 4, hipMemcpyHostToDevice);
    if (rc == hipSuccess)
      rc = hipMemcpy(b->base.worlds, b->infos.data(), sizeof(WorldInfo) * h.nworlds, hipMemcpyHostToDevice);
  }
  if (rc != hipSuccess) return fail(BB_ERR_DEVICE);
  b->step_index = h.step_index;
  b->stepped = true;
  *out = b;
  return BB_OK;
}

bb_status set_param(bb_batch* b, bb_param p, uint32_t value) {
  if (p == BB_PARAM_BLOCK) {
    if (value != 0 && value != 128 && value != 256 && value != 512 && value != 1024) return BB_ERR_INVALID;
    b->block = value;
  } else if (p == BB_PARAM_PGS_LDS_ROWS) {
    if (value > kMaxPgsLdsRows) return BB_ERR_INVALID;
    b->pgs_lds = value;
    b->base.pgs_lcap = pgs_lds_rows(b);
  } else if (p == BB_PARAM_TEAM_BODIES) {
    b->team_bodies = value;
  } else if (p == BB_PARAM_LANE_BODIES) {
    b->lane_bodies = value;
  } else if (p == BB_PARAM_LANE_BLOCK) {
    if (value != 256 && value != 512 && value != 1024) return BB_ERR_INVALID;
    b->lane_block = value;
  } else if (p == BB_PARAM_LANE_LDS_ROWS) {
    if (value > kMaxPgsLdsRows) return BB_ERR_INVALID;
    b->lane_lds = value;
  } else if (p == BB_PARAM_TEAM_SPIN) {
    b->team_spin = value;
    b->base.team_spin = value;
  } else if (p == BB_PARAM_AVBD_LANES) {
    if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8) return BB_ERR_INVALID;
    b->avbd_lanes = b->base.avbd_lanes = value;
  } else if (p == BB_PARAM_PGS_PREFETCH) {
    if (value > 1) return BB_ERR_INVALID;
    b->pgs_prefetch = b->base.pgs_prefetch = value;
  } else if (p == BB_PARAM_FUSED) {
    if (value > 1) return BB_ERR_INVALID;
    b->fused = value;
  } else if (p == BB_PARAM_SORT_BLOCKS) {
    if (value > kMaxTeamBlocks) return BB_ERR_INVALID;
    b->sort_blocks = b->base.sort_blocks = value;
  } else if (p == BB_PARAM_WALK_BUDGET) {
    b->walk_budget = b->base.walk_budget = value;
  } else if (p == BB_PARAM_TEAM_BLOCKS) {
    if (value > kMaxTeamBlocks) return BB_ERR_INVALID;
    b->team_blocks = value;
  } else if (p == BB_PARAM_PGS_BLOCK) {
    if (value != 0 && value != 64 && value != 128 && value != 256 && value != 512 && value != 1024)
      return BB_ERR_INVALID;
    b->pgs_block = value;
  } else if (p == BB_PARAM_PAIR_SCALE) {
    if (value < 1 || value > 16) return BB_ERR_INVALID;
    if (value != b->pair_scale) {
      bb_status st = pull_state(b);  // a new layout, carrying the current state
      if (st != BB_OK) return st;
      b->pair_scale = value;
      b->dirty = true;
    }
  } else {
    return BB_ERR_INVALID;
  }
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) {  // kernel arguments and launch shapes are baked into the graph
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  if (b->built && (p == BB_PARAM_BLOCK || p == BB_PARAM_TEAM_BODIES || p == BB_PARAM_LANE_BODIES ||
                   p == BB_PARAM_TEAM_BLOCKS))
    return upload_teams(b);
  return BB_OK;
}

// Every step reads only the newest parity (plus properties and the world
// table) and writes the other one, so the same step can be replayed from the
// same state: the tuner times each candidate on identical work and then
// restores the parity, step counter and counters it started from.
bb_status autotune_batch(bb_batch* b, uint32_t reps, float* best_ms) {
  if (b->worlds.empty() || reps == 0) return BB_ERR_INVALID;
  bb_status st = build(b);
  if (st != BB_OK) return st;
  while (any_one_shot(b)) {  // one-shot steps run for real first
    st = step_batch(b, 1, false);
    if (st != BB_OK) return st;
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  const int saved = b->latest;
  const uint64_t saved_index = b->step_index;
  const bool saved_stepped = b->stepped;
  const uint32_t saved_block = b->block, saved_lds = b->pgs_lds;
  const size_t stat_bytes = b->worlds.size() * ST_COUNT * sizeof(uint32_t);
  uint32_t* stat_copy = nullptr;
  HIPCHK(hipMalloc(&stat_copy, stat_bytes + 4));
  HIPCHK(hipMemcpyAsync(stat_copy, b->base.stats, stat_bytes, hipMemcpyDeviceToDevice, b->stream));
  HIPCHK(hipMemcpyAsync((char*)stat_copy + stat_bytes, b->base.ovf_any, 4, hipMemcpyDeviceToDevice, b->stream));
  uint32_t mx = 0;
  for (auto& hw : b->worlds) mx = std::max(mx, (uint32_t)hw.bodies.size());
  std::vector<uint32_t> lds_opts;
  for (uint32_t v : {0u, 256u, 512u, 768u}) {
    lds_opts.push_back(v);
    if (v >= mx) break;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // one candidate: an untimed first launch, then `reps` timed replays of the
  // same step from the same state
  auto measure = [&](float& ms) -> bb_status {
    b->latest = saved;
    bb_status r = launch_one(b);
    if (r != BB_OK) return r;
    (void)hipEventRecord(e0, b->stream);
    for (uint32_t k = 0; k < reps && r == BB_OK; ++k) {
      b->latest = saved;
      r = launch_one(b);
    }
    (void)hipEventRecord(e1, b->stream);
    if (r != BB_OK) return r;
    if (hipEventSynchronize(e1) != hipSuccess) return BB_ERR_DEVICE;
    ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= (float)reps;
    return BB_OK;
  };
  float best = 3.4e38f;
  uint32_t best_block = saved_block, best_lds = saved_lds;
  const uint32_t saved_fused = b->fused, saved_sort = b->sort_blocks, saved_lanes = b->avbd_lanes,
                 saved_prefetch = b->pgs_prefetch,
                 saved_team_blocks = b->team_blocks, saved_pgs_block = b->pgs_block;
  bb_status rc = BB_OK;
  // The wide lane is not tuned here: one replayed step early in a run does
  // not show how much the large worlds come to dominate later (over the
  // 240-step batch the lane gains 11% while the first steps favour one lane).
  for (uint32_t blk : {128u, 256u, 512u, 1024u}) {
    for (uint32_t lds : lds_opts) {
      b->block = blk;
      b->pgs_lds = lds;
      b->base.pgs_lcap = pgs_lds_rows(b);
      if ((rc = upload_teams(b)) != BB_OK) break;
      float ms = 0.0f;
      if ((rc = measure(ms)) != BB_OK) break;
      if (ms < best) {
        best = ms;
        best_block = blk;
        best_lds = lds;
      }
    }
    if (rc != BB_OK) break;
  }
  // then, one at a time from the best so far: one launch per step or per
  // stage; the solver tile of the per-world AVBD sweeps (lanes per body)
  // when a world runs AVBD in them; the sort's sub-team when a world runs
  // as a workgroup team
  bool any_avbd = false;
  for (const WorldInfo& wi : b->infos)
    if (wi.solver == BB_SOLVER_AVBD && !(wi.mode & kModeTeam) && wi.n) any_avbd = true;
  if (rc == BB_OK) {
    b->block = best_block;
    b->pgs_lds = best_lds;
    b->base.pgs_lcap = pgs_lds_rows(b);
    rc = upload_teams(b);
  }
  // `field` is the batch's setting, `mirror` the copy the kernels read
  auto tune = [&](uint32_t& field, uint32_t* mirror, std::initializer_list<uint32_t> values) {
    uint32_t keep = field;
    for (uint32_t v : values) {
      if (rc != BB_OK) return;
      if (v == keep) continue;  // measured already, as the best so far
      field = v;
      if (mirror) *mirror = v;
      float ms = 0.0f;
      if ((rc = measure(ms)) != BB_OK) break;
      if (ms < best) {
        best = ms;
        keep = v;
      }
    }
    field = keep;
    if (mirror) *mirror = keep;
  };
  // the colour sweeps of the small worlds in narrower workgroups, with their
  // own LDS rows (a 64-thread workgroup keeping 512 rows would hold a CU's
  // LDS for four waves)
  if (rc == BB_OK) {
    const uint32_t keep_pb = b->pgs_block, keep_lds = b->pgs_lds;
    uint32_t best_pb = keep_pb, best_pl = keep_lds;
    for (uint32_t pb : {64u, 128u}) {
      for (uint32_t lds : {128u, 256u, 512u}) {
        if (lds > kMaxPgsLdsRows) continue;
        b->pgs_block = pb;
        b->pgs_lds = lds;
        b->base.pgs_lcap = pgs_lds_rows(b);
        float ms = 0.0f;
        if ((rc = measure(ms)) != BB_OK) break;
        if (ms < best) {
          best = ms;
          best_pb = pb;
          best_pl = lds;
        }
      }
      if (rc != BB_OK) break;
    }
    b->pgs_block = best_pb;
    b->pgs_lds = best_pl;
    b->base.pgs_lcap = pgs_lds_rows(b);
  }
  tune(b->fused, nullptr, {0u, 1u});
  bool any_pgs = false;
  for (const WorldInfo& wi : b->infos)
    if (wi.solver != BB_SOLVER_AVBD && wi.n) any_pgs = true;
  if (any_pgs) tune(b->pgs_prefetch, &b->base.pgs_prefetch, {0u, 1u});
  if (any_avbd) {
    if (b->avbd_lanes == 0) b->avbd_lanes = b->base.avbd_lanes = 4;
    tune(b->avbd_lanes, &b->base.avbd_lanes, {1u, 2u, 4u, 8u});
  }
  if (b->base.team_grid) {
    if (b->sort_blocks == 0) b->sort_blocks = b->base.sort_blocks = kTeamSortBlocks;
    tune(b->sort_blocks, &b->base.sort_blocks, {8u, 16u, 32u, 64u, 128u});
  }
  // The walk budget is not tuned: the tuner replays an early step, before
  // the piles settle and a ground's walk grows long.  Chosen there, a budget
  // of 64 made the 240-step window 11% slower (many rows in slow shared
  // walks) and "never" 7% slower (the grounds' walks on one thread).
  // the team launch: half the resident workgroups (the default) or all of
  // them (the teams are laid out anew for each)
  if (b->base.team_grid && rc == BB_OK) {
    const uint32_t keep = b->team_blocks, all = team_capacity();
    b->team_blocks = all;
    if ((rc = upload_teams(b)) == BB_OK) {
      float ms = 0.0f;
      if ((rc = measure(ms)) == BB_OK && ms < best) best = ms;
      else b->team_blocks = keep;
      if (rc == BB_OK) rc = upload_teams(b);
    }
  }
  if (rc != BB_OK) {
    b->pgs_block = saved_pgs_block;
    b->team_blocks = saved_team_blocks;
    b->fused = saved_fused;
    b->sort_blocks = b->base.sort_blocks = saved_sort;
    b->avbd_lanes = b->base.avbd_lanes = saved_lanes;
    b->pgs_prefetch = b->base.pgs_prefetch = saved_prefetch;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->latest = saved;
  b->step_index = saved_index;
  b->stepped = saved_stepped;
  if (hipMemcpyAsync(b->base.stats, stat_copy, stat_bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
      hipMemcpyAsync(b->base.ovf_any, (char*)stat_copy + stat_bytes, 4, hipMemcpyDeviceToDevice, b->stream) !=
          hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    rc = rc != BB_OK ? rc : BB_ERR_DEVICE;
  (void)hipFree(stat_copy);
  b->block = rc == BB_OK ? best_block : saved_block;
  b->pgs_lds = rc == BB_OK ? best_lds : saved_lds;
  b->base.pgs_lcap = pgs_lds_rows(b);
  {
    const bb_status tr = upload_teams(b);
    if (rc == BB_OK) rc = tr;
  }
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  if (best_ms && rc == BB_OK) *best_ms = best;
  return rc;
}

bb_status read_overflow_any(bb_batch* b, uint32_t* out) {
  *out = 0;
  if (!b->built || !b->base.ovf_any) return BB_OK;
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(out, b->base.ovf_any, 4, hipMemcpyDeviceToHost));
  return BB_OK;
}

bb_status clear_overflow(bb_batch* b) {
  if (!b->built || !b->base.ovf_any) return BB_OK;
  HIPCHK(hipStreamSynchronize(b->stream));
  const uint32_t W = (uint32_t)b->worlds.size();
  for (uint32_t w = 0; w < W; ++w)
    HIPCHK(hipMemsetAsync(b->base.stats + (size_t)w * ST_COUNT + ST_OVF_STICKY, 0, 4, b->stream));
  HIPCHK(hipMemsetAsync(b->base.ovf_any, 0, 4, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return BB_OK;
}

// Waits for the batch's work; BB_ERR_OVERFLOW while a world's overflow
// (pairs past its capacity, or a team that did not meet at a barrier) has
// not been collected by bb_batch_overflow.
bb_status sync_batch(bb_batch* b) {
  if (!b->stream) return BB_OK;
  HIPCHK(hipStreamSynchronize(b->stream));
  uint32_t any = 0;
  bb_status st = read_overflow_any(b, &any);
  if (st != BB_OK) return st;
  return any ? BB_ERR_OVERFLOW : BB_OK;
}

bb_status overflow_worlds(bb_batch* b, uint32_t* worlds, uint32_t cap, uint32_t* count) {
  *count = 0;
  if (!b->built || !b->stepped) return BB_OK;
  HIPCHK(hipStreamSynchronize(b->stream));
  const uint32_t W = (uint32_t)b->worlds.size();
  std::vector<uint32_t> raw((size_t)W * ST_COUNT);
  HIPCHK(hipMemcpy(raw.data(), b->base.stats, raw.size() * 4, hipMemcpyDeviceToHost));
  uint32_t k = 0;
  for (uint32_t w = 0; w < W; ++w)
    if (raw[(size_t)w * ST_COUNT + ST_OVF_STICKY]) {
      if (worlds && k < cap) worlds[k] = w;
      ++k;
    }
  *count = k;
  return clear_overflow(b);
}

bb_status read_world(bb_batch* b, uint32_t w, bb_body_state* out, uint32_t n, bb_order order) {
  if (w >= b->worlds.size() || !out) return BB_ERR_INVALID;
  HostWorld& hw = b->worlds[w];
  if (n < hw.bodies.size()) return BB_ERR_INVALID;
  bb_status st = build(b);
  if (st != BB_OK) return st;
  HIPCHK(hipStreamSynchronize(b->stream));
  const uint32_t N = hw.n, g0 = hw.body_off;
  const Rows& R = b->rows[b->latest];
  std::vector<uint32_t> id(N), fl(N), isl(N, kMAX), ent(N);
  std::vector<float4> pos(N), rot(N), vel(N), omg(N);
  if (N) {
    HIPCHK(hipMemcpy(id.data(), R.id + g0, N * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fl.data(), R.flags + g0, N * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(pos.data(), R.pos + g0, N * 16, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rot.data(), R.rot + g0, N * 16, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(vel.data(), R.vel + g0, N * 16, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(omg.data(), R.omg + g0, N * 16, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ent.data(), b->entries[b->latest] + g0, N * 4, hipMemcpyDeviceToHost));
    if (b->stepped) HIPCHK(hipMemcpy(isl.data(), b->base.islot + g0, N * 4, hipMemcpyDeviceToHost));
  }
  for (uint32_t k = 0; k < N; ++k) {
    uint32_t row = order == BB_ORDER_BY_ID ? ent[k] : k;
    bb_body_state& o = out[k];
    o.id = id[row];
    o.flags = fl[row];
    o.island_index = isl[row];
    o.row = row;
    o.position[0] = pos[row].x; o.position[1] = pos[row].y; o.position[2] = pos[row].z;
    o.rotation[0] = rot[row].x; o.rotation[1] = rot[row].y; o.rotation[2] = rot[row].z; o.rotation[3] = rot[row].w;
    o.velocity[0] = vel[row].x; o.velocity[1] = vel[row].y; o.velocity[2] = vel[row].z;
    o.omega[0] = omg[row].x; o.omega[1] = omg[row].y; o.omega[2] = omg[row].z;
  }
  return BB_OK;
}

bb_status world_stats(bb_batch* b, uint32_t w, bb_step_stats* out) {
  if (w >= b->worlds.size() || !out) return BB_ERR_INVALID;
  std::memset(out, 0, sizeof(*out));
  if (!b->built || !b->stepped) {
    out->rigid_body_count = (uint32_t)b->worlds[w].bodies.size();
    return BB_OK;
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  uint32_t s[ST_COUNT];
  HIPCHK(hipMemcpy(s, b->base.stats + (size_t)w * ST_COUNT, sizeof(s), hipMemcpyDeviceToHost));
  stats_from_raw(s, b->step_index ? b->step_index - 1 : 0, out);
  return BB_OK;
}

void stats_from_raw(const uint32_t* s, uint64_t step_index, bb_step_stats* out) {
  out->rigid_body_count = s[ST_N];
  out->active_rigid_body_count = s[ST_A];
  out->broad_phase_collision_count = s[ST_P];
  out->collision_count = s[ST_M];
  out->manifold_node_count = s[ST_NODES];
  out->contact_count = s[ST_C];
  out->island_count = s[ST_ISLANDS];
  out->contact_island_count = s[ST_CISLANDS];
  out->sleeping_count = s[ST_SLEEPING];
  out->graph_color_count = s[ST_GC_COUNT];
  out->graph_color_violations = s[ST_GC_VIOL];
  out->graph_color_overflow = s[ST_GC_OVF];
  out->avbd_color_count = s[ST_AV_COUNT];
  out->avbd_violations = s[ST_AV_VIOL];
  out->avbd_stick_count = s[ST_AV_STICK];
  out->dbg_fresh = s[ST_FRESH];
  out->dbg_pen = s[ST_PEN];
  out->dbg_id_sum = s[ST_IDSUM];
  out->dbg_maxv = s[ST_MAXV];
  out->overflow = s[ST_OVERFLOW];
  out->step_index = step_index;
}

template <class T>
bb_status copy_down(std::vector<T>& v, const T* src, size_t n) {
  v.resize(n);
  if (n) HIPCHK(hipMemcpy(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
  return BB_OK;
}

bb_status tap_world(bb_batch* b, uint32_t w, bb_tap kind, void* dst, size_t cap, size_t* written) {
  if (w >= b->worlds.size()) return BB_ERR_INVALID;
  if (!b->built || !b->stepped) {
    if (written) *written = 0;
    return BB_OK;
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  const HostWorld& hw = b->worlds[w];
  const WorldInfo& wi = b->infos[w];
  const uint32_t N = hw.n, g0 = hw.body_off;
  const int cur = b->latest;
  std::vector<uint8_t> out;
  auto emit = [&](const void* p, size_t bytes) {
    out.assign((const uint8_t*)p, (const uint8_t*)p + bytes);
  };
  uint32_t P = 0;
  HIPCHK(hipMemcpy(&P, b->npairs[cur] + w, 4, hipMemcpyDeviceToHost));
  switch (kind) {
    case BB_TAP_MORTON:
    case BB_TAP_SORTED: {
      std::vector<uint64_t> keys;
      std::vector<uint32_t> idx;
      bb_status s = copy_down(keys, b->base.keys_a + g0, N);
      if (s != BB_OK) return s;
      s = copy_down(idx, b->base.idx_a + g0, N);
      if (s != BB_OK) return s;
      if (kind == BB_TAP_SORTED) {
        std::vector<uint32_t> v(4 * (size_t)N);
        for (uint32_t i = 0; i < N; ++i) {
          v[4 * i] = (uint32_t)keys[i];
          v[4 * i + 1] = (uint32_t)(keys[i] >> 32);
          v[4 * i + 2] = idx[i];
          v[4 * i + 3] = 0;
        }
        emit(v.data(), v.size() * 4);
      } else if (wi.morton_bits == 30) {
        std::vector<uint32_t> v(N);
        for (uint32_t i = 0; i < N; ++i) v[idx[i]] = (uint32_t)keys[i];
        emit(v.data(), v.size() * 4);
      } else {
        std::vector<uint64_t> v(N);
        for (uint32_t i = 0; i < N; ++i) v[idx[i]] = keys[i];
        emit(v.data(), v.size() * 8);
      }
      break;
    }
    case BB_TAP_LBVH: {
      std::vector<float4> mn, mx;
      size_t nn = N ? 2 * (size_t)N - 1 : 0;
      bb_status s = copy_down(mn, b->base.nmin + 2 * (size_t)g0, nn);
      if (s != BB_OK) return s;
      s = copy_down(mx, b->base.nmax + 2 * (size_t)g0, nn);
      if (s != BB_OK) return s;
      std::vector<bb_lbvh_node> v(nn);
      for (size_t i = 0; i < nn; ++i) {
        v[i].minimum[0] = mn[i].x; v[i].minimum[1] = mn[i].y; v[i].minimum[2] = mn[i].z;
        v[i].maximum[0] = mx[i].x; v[i].maximum[1] = mx[i].y; v[i].maximum[2] = mx[i].z;
        int l, r;
        std::memcpy(&l, &mn[i].w, 4);
        std::memcpy(&r, &mx[i].w, 4);
        if (l < 0) {
          v[i].left = -1;
          v[i].right = -1;
          v[i].primitive_idx = (uint32_t)r;
        } else {
          v[i].left = l;
          v[i].right = r;
          v[i].primitive_idx = kMAX;
        }
        v[i].pad = 0;
      }
      emit(v.data(), v.size() * sizeof(bb_lbvh_node));
      break;
    }
    case BB_TAP_PAIRS: {
      std::vector<uint2> v;
      bb_status s = copy_down(v, b->pairs[cur] + hw.pair_off, P);
      if (s != BB_OK) return s;
      emit(v.data(), v.size() * 8);
      break;
    }
    case BB_TAP_MANIFOLDS:
    case BB_TAP_EDGE_COLORS: {
      const Manifolds& M = b->man[cur];
      std::vector<int4> hdr;
      std::vector<float4> nrm, t0, t1, aa, ab, cpos, clam, cmas, cana, canb;
      std::vector<uint4> cfp;
      std::vector<uint32_t> cws, col;
      const size_t so = hw.pair_off;
      bb_status s = copy_down(hdr, M.hdr + so, P);
      if (s == BB_OK) s = copy_down(col, b->base.color + so, P);
      if (kind == BB_TAP_EDGE_COLORS) {
        if (s != BB_OK) return s;
        std::vector<uint32_t> v;
        for (uint32_t p = 0; p < P; ++p)
          if (hdr[p].w > 0) v.push_back(wi.solver == BB_SOLVER_AVBD ? kMAX : col[p]);
        emit(v.data(), v.size() * 4);
        break;
      }
      if (s == BB_OK) s = copy_down(nrm, M.nrm + so, P);
      if (s == BB_OK) s = copy_down(t0, M.t0 + so, P);
      if (s == BB_OK) s = copy_down(t1, M.t1 + so, P);
      if (s == BB_OK) s = copy_down(aa, M.anca + so, P);
      if (s == BB_OK) s = copy_down(ab, M.ancb + so, P);
      if (s == BB_OK) s = copy_down(cpos, M.cpos + so * 8, (size_t)P * 8);
      if (s == BB_OK) s = copy_down(clam, M.clam + so * 8, (size_t)P * 8);
      if (s == BB_OK) s = copy_down(cmas, M.cmas + so * 8, (size_t)P * 8);
      if (s == BB_OK) s = copy_down(cana, M.cana + so * 8, (size_t)P * 8);
      if (s == BB_OK) s = copy_down(canb, M.canb + so * 8, (size_t)P * 8);
      if (s == BB_OK) s = copy_down(cfp, M.cfp + so * 8, (size_t)P * 8);
      if (s == BB_OK) s = copy_down(cws, M.cws + so * 8, (size_t)P * 8);
      if (s != BB_OK) return s;
      std::vector<bb_manifold_tap> v;
      for (uint32_t p = 0; p < P; ++p) {
        if (hdr[p].w <= 0) continue;
        bb_manifold_tap t;
        std::memset(&t, 0, sizeof(t));
        t.obb1_index = (uint32_t)hdr[p].x;
        t.obb2_index = (uint32_t)hdr[p].y;
        t.key = hdr[p].z;
        t.contact_count = hdr[p].w;
        std::memcpy(&t.error, &nrm[p].w, 4);
        t.color = wi.solver == BB_SOLVER_AVBD ? kMAX : col[p];
        t.normal[0] = nrm[p].x; t.normal[1] = nrm[p].y; t.normal[2] = nrm[p].z;
        t.tangent[0][0] = t0[p].x; t.tangent[0][1] = t0[p].y; t.tangent[0][2] = t0[p].z;
        t.tangent[1][0] = t1[p].x; t.tangent[1][1] = t1[p].y; t.tangent[1][2] = t1[p].z;
        t.anchor_a[0] = aa[p].x; t.anchor_a[1] = aa[p].y; t.anchor_a[2] = aa[p].z;
        t.anchor_b[0] = ab[p].x; t.anchor_b[1] = ab[p].y; t.anchor_b[2] = ab[p].z;
        for (int c = 0; c < hdr[p].w && c < 8; ++c) {
          size_t k = (size_t)p * 8 + c;
          bb_contact_tap& o = t.contacts[c];
          o.position[0] = cpos[k].x; o.position[1] = cpos[k].y; o.position[2] = cpos[k].z;
          o.penetration = cpos[k].w;
          o.normal_impulse = clam[k].x;
          o.tangent_impulse[0] = clam[k].y;
          o.tangent_impulse[1] = clam[k].z;
          o.mass_coefficient = clam[k].w;
          o.tangent_mass[0] = cmas[k].y;
          o.tangent_mass[1] = cmas[k].z;
          o.anchor_a[0] = cana[k].x; o.anchor_a[1] = cana[k].y; o.anchor_a[2] = cana[k].z;
          o.anchor_b[0] = canb[k].x; o.anchor_b[1] = canb[k].y; o.anchor_b[2] = canb[k].z;
          std::memcpy(&o.stick, &canb[k].w, 4);
          o.fp[0] = cfp[k].x; o.fp[1] = cfp[k].y; o.fp[2] = cfp[k].z; o.fp[3] = cfp[k].w;
          o.warm_start = cws[k];
        }
        v.push_back(t);
      }
      emit(v.data(), v.size() * sizeof(bb_manifold_tap));
      break;
    }
    case BB_TAP_ISLANDS: {
      std::vector<uint32_t> v;
      bb_status s = copy_down(v, b->base.island + g0, N);
      if (s != BB_OK) return s;
      emit(v.data(), v.size() * 4);
      break;
    }
    case BB_TAP_BODY_COLORS: {
      std::vector<uint32_t> v(N, kMAX);
      if (wi.solver == BB_SOLVER_AVBD) {
        bb_status s = copy_down(v, b->base.bcolor + g0, N);
        if (s != BB_OK) return s;
      }
      emit(v.data(), v.size() * 4);
      break;
    }
    case BB_TAP_ROWS: {
      std::vector<bb_body_state> v(N);
      bb_status s = read_world(b, w, v.data(), N, BB_ORDER_BY_ROW);
      if (s != BB_OK) return s;
      emit(v.data(), v.size() * sizeof(bb_body_state));
      break;
    }
    default:
      return BB_ERR_INVALID;
  }
  if (written) *written = out.size();
  if (!dst) return BB_OK;
  if (out.size() > cap) return BB_ERR_INVALID;
  std::memcpy(dst, out.data(), out.size());
  return BB_OK;
}

bool valid_desc(const bb_world_desc* d) {
  if (!d) return false;
  if (d->solver > BB_SOLVER_AVBD) return false;
  if (!(d->dt > 0.0f) || !std::isfinite(d->dt) || !std::isfinite(d->gravity)) return false;
  if (d->morton_bits != 30 && d->morton_bits != 63) return false;
  return true;
}

bool over_reference_cap(const bb_world_desc& d, uint32_t n) {
  // parity mode keeps the reference's cap (MAX_RIGID_BODY_COUNT = 1024,
  // include/defines.hpp:357; build_accel_structs refuses larger scenes)
  return d.morton_bits == 30 && n > 1024 && (d.quirks & BB_QUIRKS_PARITY) == BB_QUIRKS_PARITY;
}

bool valid_body(const bb_body_desc& b) {
  if (b.shape > BB_SHAPE_SPHERE) return false;
  if (!(b.mass >= 0.0f) || !std::isfinite(b.mass)) return false;
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(b.position[k]) || !std::isfinite(b.velocity[k]) || !std::isfinite(b.omega[k])) return false;
  for (int k = 0; k < 4; ++k)
    if (!std::isfinite(b.rotation[k])) return false;
  if (b.shape == BB_SHAPE_SPHERE && !(b.radius > 0.0f)) return false;
  if (b.shape == BB_SHAPE_BOX)
    for (int k = 0; k < 3; ++k)
      if (!(b.maximum[k] > b.minimum[k])) return false;
  return true;
}

bool valid_body_with_voxels(const bb_body_desc& b, size_t shapes) {
  if (BB_SHAPE_KIND(b.shape) != BB_SHAPE_VOXEL) return valid_body(b);
  const uint32_t id = BB_SHAPE_VOXEL_ID(b.shape);
  if (id == 0 || id > shapes) return false;
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(b.position[k]) || !std::isfinite(b.velocity[k]) || !std::isfinite(b.omega[k])) return false;
  for (int k = 0; k < 4; ++k)
    if (!std::isfinite(b.rotation[k])) return false;
  return true;
}

// SceneManager::build_voxel_shape (src/scene_manager.hpp:478-566) with the
// solid(x,y,z) predicate given as occupancy bits.  Centre of mass in double;
// per-cell inertia (parallel-axis point term plus the cell's own cube term)
// accumulated in float in z, y, x order; glm::inverse of the 3x3 (adjugate
// over the determinant); the surface list keeps cells with an empty face
// neighbour, coded with the first empty direction (-x,+x,-y,+y,-z,+z).
bb_status add_voxel_shape(bb_batch* b, const bb_voxel_shape_desc* d, bb_voxel_shape_info* out) {
  if (!d || !out || !d->occupancy) return BB_ERR_INVALID;
  const uint32_t dx = d->dims[0], dy = d->dims[1], dz = d->dims[2];
  if (dx == 0 || dy == 0 || dz == 0 || dx > 255 || dy > 255 || dz > 255) return BB_ERR_INVALID;
  if (!(d->voxel_size > 0.0f) || !std::isfinite(d->voxel_size) || !(d->density > 0.0f) ||
      !std::isfinite(d->density))
    return BB_ERR_INVALID;
  const uint32_t* occ = d->occupancy;
  auto solid = [&](int x, int y, int z) -> bool {
    if (x < 0 || y < 0 || z < 0 || x >= (int)dx || y >= (int)dy || z >= (int)dz) return false;
    const uint32_t bit = (uint32_t)x + (uint32_t)y * dx + (uint32_t)z * dx * dy;
    return ((occ[bit >> 5] >> (bit & 31u)) & 1u) != 0u;
  };
  const float vs = d->voxel_size;
  double com[3] = {0.0, 0.0, 0.0};
  uint32_t count = 0;
  for (uint32_t z = 0; z < dz; ++z)
    for (uint32_t y = 0; y < dy; ++y)
      for (uint32_t x = 0; x < dx; ++x)
        if (solid((int)x, (int)y, (int)z)) {
          com[0] += x + 0.5;
          com[1] += y + 0.5;
          com[2] += z + 0.5;
          ++count;
        }
  if (count == 0) return BB_ERR_INVALID;
  for (int k = 0; k < 3; ++k) com[k] = com[k] / (double)count * (double)vs;
  const float voxel_mass = d->density * vs * vs * vs;
  const float mass = voxel_mass * (float)count;
  const float org[3] = {-(float)com[0], -(float)com[1], -(float)com[2]};
  const uint32_t cells = dx * dy * dz, words = (cells + 31u) / 32u;
  std::vector<uint32_t> surf;
  float I[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
  const float own = voxel_mass * vs * vs / 6.0f;
  static const int dirs[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  for (uint32_t z = 0; z < dz; ++z)
    for (uint32_t y = 0; y < dy; ++y)
      for (uint32_t x = 0; x < dx; ++x) {
        if (!solid((int)x, (int)y, (int)z)) continue;
        const float c[3] = {org[0] + ((float)x + 0.5f) * vs, org[1] + ((float)y + 0.5f) * vs,
                            org[2] + ((float)z + 0.5f) * vs};
        const float r2 = (c[0] * c[0] + c[1] * c[1]) + c[2] * c[2];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) I[i][j] = I[i][j] + voxel_mass * ((i == j ? r2 : 0.0f) - c[j] * c[i]);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) I[i][j] = I[i][j] + (i == j ? own : 0.0f);
        for (uint32_t k = 0; k < 6; ++k)
          if (!solid((int)x + dirs[k][0], (int)y + dirs[k][1], (int)z + dirs[k][2])) {
            surf.push_back(x | (y << 8) | (z << 16) | (k << 24));
            break;
          }
      }
  if (b->vox_shapes.size() >= BB_MAX_VOXEL_SHAPES || b->vox_occ.size() + words > BB_MAX_VOXEL_OCC_WORDS ||
      b->vox_surf.size() + surf.size() > BB_MAX_VOXEL_SURFACE)
    return BB_ERR_CAPACITY;
  // glm::inverse: m is column-major there; the inertia is symmetric, so its
  // columns are its rows
  const float(&m)[3][3] = I;
  const float det = (m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                     m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2])) +
                    m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
  const float k1 = 1.0f / det;
  float inv[3][3];  // [column][row]
  inv[0][0] = (m[1][1] * m[2][2] - m[2][1] * m[1][2]) * k1;
  inv[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * k1;
  inv[2][0] = (m[1][0] * m[2][1] - m[2][0] * m[1][1]) * k1;
  inv[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) * k1;
  inv[1][1] = (m[0][0] * m[2][2] - m[2][0] * m[0][2]) * k1;
  inv[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * k1;
  inv[0][2] = (m[0][1] * m[1][2] - m[1][1] * m[0][2]) * k1;
  inv[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * k1;
  inv[2][2] = (m[0][0] * m[1][1] - m[1][0] * m[0][1]) * k1;
  VoxShape vsh{};
  vsh.dims[0] = dx;
  vsh.dims[1] = dy;
  vsh.dims[2] = dz;
  vsh.occ_off = (uint32_t)b->vox_occ.size();
  vsh.vs = vs;
  vsh.ox = org[0];
  vsh.oy = org[1];
  vsh.oz = org[2];
  vsh.surf_off = (uint32_t)b->vox_surf.size();
  vsh.surf_count = (uint32_t)surf.size();
  const size_t o0 = b->vox_occ.size();
  b->vox_occ.insert(b->vox_occ.end(), occ, occ + words);
  if (cells % 32u) b->vox_occ[o0 + words - 1] &= (1u << (cells % 32u)) - 1u;
  b->vox_surf.insert(b->vox_surf.end(), surf.begin(), surf.end());
  b->vox_shapes.push_back(vsh);
  bb_voxel_shape_info info{};
  info.shape_id = (uint32_t)b->vox_shapes.size();
  info.solid_count = count;
  info.surface_count = vsh.surf_count;
  info.mass = mass;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) info.inv_inertia[r * 3 + c] = inv[r][c];
  for (int k = 0; k < 3; ++k) {
    info.minimum[k] = org[k];
    info.maximum[k] = org[k] + (float)d->dims[k] * vs;
  }
  b->vox_info.push_back(info);
  b->vox_density.push_back(d->density);
  b->vox_dirty = true;
  *out = info;
  return BB_OK;
}

bb_status upload_voxels(bb_batch* b) {
  if (!b->vox_dirty || !b->d_vshape) return BB_OK;
  if (!b->vox_shapes.empty())
    HIPCHK(hipMemcpyAsync(b->d_vshape, b->vox_shapes.data(), b->vox_shapes.size() * sizeof(VoxShape),
                          hipMemcpyHostToDevice, b->stream));
  if (!b->vox_occ.empty())
    HIPCHK(hipMemcpyAsync(b->d_vocc, b->vox_occ.data(), b->vox_occ.size() * 4, hipMemcpyHostToDevice, b->stream));
  if (!b->vox_surf.empty())
    HIPCHK(hipMemcpyAsync(b->d_vsurf, b->vox_surf.data(), b->vox_surf.size() * 4, hipMemcpyHostToDevice,
                          b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->vox_dirty = false;
  return BB_OK;
}

}  // namespace host
}  // namespace bb

// =================================================================== C ABI
extern "C" {

const char* bb_version(void) { return "beatbox-amd 0.1 (gfx950)"; }

bb_status bb_copy_bandwidth(uint64_t bytes, uint32_t reps, float* ms) {
  if (!ms || bytes < 16 || reps == 0) return BB_ERR_INVALID;
  if (!has_device()) return BB_ERR_NO_DEVICE;
  const size_t n = bytes / 16;
  float4 *src = nullptr, *dst = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bb_status rc = BB_OK;
  float total = 0.0f;
  if (hipMalloc(&src, n * 16) != hipSuccess || hipMalloc(&dst, n * 16) != hipSuccess ||
      hipMemset(src, 0, n * 16) != hipSuccess || hipStreamCreate(&st) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = BB_ERR_DEVICE;
  } else {
    launch_copy(src, dst, n, st);  // warm-up
    (void)hipEventRecord(e0, st);
    for (uint32_t r = 0; r < reps; ++r) launch_copy(src, dst, n, st);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&total, e0, e1) != hipSuccess)
      rc = BB_ERR_DEVICE;
  }
  *ms = total / (float)reps;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  return rc;
}

void bb_world_desc_default(bb_world_desc* d) {
  if (!d) return;
  d->capacity_bodies = 0;
  d->solver = BB_SOLVER_AVBD;
  d->iterations = 5;
  d->flags = BB_SIM_DEFAULT_FLAGS;
  d->dt = 0.01667f;
  d->gravity = -9.81f;
  d->morton_bits = 30;
  d->quirks = BB_QUIRKS_PARITY;
}

bb_status bb_batch_create(uint32_t capacity_worlds, uint32_t capacity_bodies, bb_batch** out) {
  if (!out) return BB_ERR_INVALID;
  bb_batch* b = new (std::nothrow) bb_batch();
  if (!b) return BB_ERR_CAPACITY;
  b->worlds.reserve(capacity_worlds);
  (void)capacity_bodies;
  *out = b;
  return BB_OK;
}

bb_status bb_batch_add_world(bb_batch* b, const bb_world_desc* desc, const bb_body_desc* bodies, uint32_t n,
                             uint32_t* out_world) {
  if (!b || !valid_desc(desc) || (n && !bodies)) return BB_ERR_INVALID;
  for (uint32_t i = 0; i < n; ++i)
    if (!valid_body_with_voxels(bodies[i], b->vox_shapes.size())) return BB_ERR_INVALID;
  if (desc->capacity_bodies && n > desc->capacity_bodies) return BB_ERR_CAPACITY;
  if (over_reference_cap(*desc, n)) return BB_ERR_CAPACITY;
  bb_status st = pull_state(b);
  if (st != BB_OK) return st;
  HostWorld hw;
  hw.desc = *desc;
  hw.bodies.assign(bodies, bodies + n);
  b->worlds.push_back(hw);
  b->dirty = true;
  if (out_world) *out_world = (uint32_t)b->worlds.size() - 1;
  return BB_OK;
}

bb_status bb_batch_add_voxel_shape(bb_batch* b, const bb_voxel_shape_desc* desc, bb_voxel_shape_info* out) {
  if (!b) return BB_ERR_INVALID;
  return add_voxel_shape(b, desc, out);
}

bb_status bb_batch_step(bb_batch* b, uint32_t nsteps, void* stream, int use_graph) {
  if (!b) return BB_ERR_INVALID;
  (void)stream;
  return step_batch(b, nsteps, use_graph != 0);
}

bb_status bb_batch_sync(bb_batch* b) { return b ? sync_batch(b) : BB_ERR_INVALID; }

bb_status bb_batch_overflow(bb_batch* b, uint32_t* worlds, uint32_t cap, uint32_t* count) {
  if (!b || !count || (cap && !worlds)) return BB_ERR_INVALID;
  return overflow_worlds(b, worlds, cap, count);
}

bb_status bb_batch_read_world(bb_batch* b, uint32_t w, bb_body_state* out, uint32_t n, bb_order order) {
  return b ? read_world(b, w, out, n, order) : BB_ERR_INVALID;
}

bb_status bb_batch_world_stats(bb_batch* b, uint32_t w, bb_step_stats* out) {
  return b ? world_stats(b, w, out) : BB_ERR_INVALID;
}

bb_status bb_batch_debug_tap(bb_batch* b, uint32_t w, bb_tap kind, void* dst, size_t cap, size_t* written) {
  return b ? tap_world(b, w, kind, dst, cap, written) : BB_ERR_INVALID;
}

bb_status bb_batch_profile(bb_batch* b, uint32_t nsteps, float* step_ms, float* kernel_ms, uint64_t* counts) {
  return b ? profile_batch(b, nsteps, step_ms, kernel_ms, counts) : BB_ERR_INVALID;
}

bb_status bb_batch_build_split(const bb_batch* b, float* build_ms) {
  if (!b || !build_ms) return BB_ERR_INVALID;
  for (int k = 0; k < BB_BUILD_COUNT; ++k) build_ms[k] = b->build_ms[k];
  return BB_OK;
}

bb_status bb_batch_avbd_split(const bb_batch* b, float* avbd_ms) {
  if (!b || !avbd_ms) return BB_ERR_INVALID;
  for (int k = 0; k < BB_AVBD_COUNT; ++k) avbd_ms[k] = b->avbd_ms[k];
  return BB_OK;
}

bb_status bb_batch_energy(bb_batch* b, double* per_world, uint32_t n) {
  return b ? energy_batch(b, per_world, n) : BB_ERR_INVALID;
}

bb_status bb_batch_instance_transforms(bb_batch* b, uint32_t world, float* dst, uint32_t n, int dst_on_device) {
  return b ? instances_batch(b, world, dst, n, dst_on_device) : BB_ERR_INVALID;
}

uint64_t bb_batch_device_bytes(const bb_batch* b) { return b ? b->buf.bytes : 0; }

uint32_t bb_batch_world_count(const bb_batch* b) { return b ? (uint32_t)b->worlds.size() : 0; }

uint32_t bb_batch_body_count(const bb_batch* b) {
  if (!b) return 0;
  uint32_t n = 0;
  for (auto& hw : b->worlds) n += (uint32_t)hw.bodies.size();
  return n;
}

bb_status bb_batch_set_block_size(bb_batch* b, uint32_t threads) {
  return bb_batch_set_param(b, BB_PARAM_BLOCK, threads);
}

bb_status bb_batch_set_param(bb_batch* b, bb_param p, uint32_t value) {
  if (!b) return BB_ERR_INVALID;
  return set_param(b, p, value);
}

uint32_t bb_batch_get_param(const bb_batch* b, bb_param p) {
  if (!b) return 0;
  if (p == BB_PARAM_BLOCK) return pick_block(b);
  if (p == BB_PARAM_PGS_LDS_ROWS) return pgs_lds_rows(b);
  if (p == BB_PARAM_TEAM_BODIES) return b->team_bodies;
  if (p == BB_PARAM_LANE_BODIES) return b->lane_bodies;
  if (p == BB_PARAM_LANE_BLOCK) return b->lane_block;
  if (p == BB_PARAM_LANE_LDS_ROWS) return b->lane_lds;
  if (p == BB_PARAM_TEAM_SPIN) return b->team_spin;
  if (p == BB_PARAM_PAIR_SCALE) return b->pair_scale;
  if (p == BB_PARAM_FUSED) return b->fused;
  if (p == BB_PARAM_PGS_PREFETCH) return b->pgs_prefetch;
  if (p == BB_PARAM_AVBD_LANES) return b->avbd_lanes ? b->avbd_lanes : 4u;
  if (p == BB_PARAM_SORT_BLOCKS) return b->sort_blocks ? b->sort_blocks : kTeamSortBlocks;
  if (p == BB_PARAM_WALK_BUDGET) return b->walk_budget ? b->walk_budget : kWalkBudgetDefault;
  if (p == BB_PARAM_TEAM_BLOCKS) return has_device() ? team_grid_cap(b) : b->team_blocks;
  if (p == BB_PARAM_PGS_BLOCK) return b->pgs_block;
  return 0;
}

bb_status bb_batch_autotune(bb_batch* b, uint32_t reps, float* best_ms) {
  return b ? autotune_batch(b, reps, best_ms) : BB_ERR_INVALID;
}

void bb_batch_destroy(bb_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->buf.release();
  if (b->stream2) (void)hipStreamSynchronize(b->stream2);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->stream2) (void)hipStreamDestroy(b->stream2);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  delete b;
}

// ------------------------------------------------------------ world API
bb_status bb_world_create(const bb_world_desc* desc, bb_world** out) {
  if (!out || !valid_desc(desc)) return BB_ERR_INVALID;
  bb_world* w = new (std::nothrow) bb_world();
  if (!w) return BB_ERR_CAPACITY;
  bb_status st = bb_batch_create(1, desc->capacity_bodies, &w->b);
  if (st != BB_OK) {
    delete w;
    return st;
  }
  HostWorld hw;
  hw.desc = *desc;
  w->b->worlds.push_back(hw);
  *out = w;
  return BB_OK;
}

bb_status bb_world_add_bodies(bb_world* w, const bb_body_desc* bodies, uint32_t n, uint32_t* out_ids) {
  if (!w || (n && !bodies)) return BB_ERR_INVALID;
  HostWorld& hw = w->b->worlds[0];
  for (uint32_t i = 0; i < n; ++i)
    if (!valid_body_with_voxels(bodies[i], w->b->vox_shapes.size())) return BB_ERR_INVALID;
  size_t total = hw.bodies.size() + n;
  if (hw.desc.capacity_bodies && total > hw.desc.capacity_bodies) return BB_ERR_CAPACITY;
  if (over_reference_cap(hw.desc, (uint32_t)total)) return BB_ERR_CAPACITY;
  bb_status st = pull_state(w->b);
  if (st != BB_OK) return st;
  for (uint32_t i = 0; i < n; ++i) {
    if (out_ids) out_ids[i] = (uint32_t)hw.bodies.size();
    hw.bodies.push_back(bodies[i]);
  }
  w->b->dirty = true;
  return BB_OK;
}

bb_status bb_world_add_voxel_shape(bb_world* w, const bb_voxel_shape_desc* desc, bb_voxel_shape_info* out) {
  if (!w) return BB_ERR_INVALID;
  return add_voxel_shape(w->b, desc, out);
}

bb_status bb_world_set_flags(bb_world* w, uint32_t flags) {
  if (!w) return BB_ERR_INVALID;
  w->b->worlds[0].desc.flags |= flags;
  return w->b->built ? (refresh_infos(w->b), upload_infos(w->b, false)) : BB_OK;
}

bb_status bb_world_clear_flags(bb_world* w, uint32_t flags) {
  if (!w) return BB_ERR_INVALID;
  w->b->worlds[0].desc.flags &= ~flags;
  return w->b->built ? (refresh_infos(w->b), upload_infos(w->b, false)) : BB_OK;
}

uint32_t bb_world_get_flags(const bb_world* w) { return w ? w->b->worlds[0].desc.flags : 0u; }

bb_status bb_world_set_solver(bb_world* w, uint32_t solver) {
  if (!w || solver > BB_SOLVER_AVBD) return BB_ERR_INVALID;
  w->b->worlds[0].desc.solver = solver;
  w->b->worlds[0].drop_prev = true;
  return w->b->built ? (refresh_infos(w->b), upload_infos(w->b, false)) : BB_OK;
}

bb_status bb_world_set_iterations(bb_world* w, uint32_t iterations) {
  if (!w) return BB_ERR_INVALID;
  w->b->worlds[0].desc.iterations = iterations;
  return w->b->built ? (refresh_infos(w->b), upload_infos(w->b, false)) : BB_OK;
}

bb_status bb_world_skip_warm_start_once(bb_world* w) {
  if (!w) return BB_ERR_INVALID;
  w->b->worlds[0].skip_ws_once = true;
  return BB_OK;
}

// One step at a time, each checked for a pair overflow (counter bit 0: the
// broad phase found more pairs than the world's pair capacity and dropped
// the rest).  Such a step is run again with twice the capacity
// (grow_pairs keeps every device array), from the state it read: a step
// writes only the other parity, so that state, the manifolds it warm-starts
// from included, is still there.  The replay sees the same world table and
// one-shot settings as the first run, so the result equals a step with
// enough capacity from the start, bit for bit.
//
// A world stepped by a workgroup team whose workgroups did not all meet at a
// barrier (counter bit 1: the team left the step unfinished, see k_team) is
// stepped again the same way from the same state, by one workgroup: teams
// are switched off for this world from then on.  The two paths give
// bit-identical results.
bb_status bb_world_step(bb_world* w, uint32_t nsteps, void* stream) {
  if (!w) return BB_ERR_INVALID;
  constexpr int kPairGrowths = 16;
  bb_batch* b = w->b;
  if (b->worlds.empty()) return BB_ERR_STATE;
  for (uint32_t s = 0; s < nsteps; ++s) {
    bb_status st = build(b);
    if (st != BB_OK) return st;
    HostWorld& hw = b->worlds[0];
    const std::vector<WorldInfo> infos0 = b->infos;
    const bool skip0 = hw.skip_ws_once, drop0 = hw.drop_prev;
    st = bb_batch_step(b, 1, stream, 0);
    int g = 0, team_retry = 0;
    while (st == BB_OK) {
      uint32_t ovf = 0;
      HIPCHK(hipStreamSynchronize(b->stream));
      HIPCHK(hipMemcpy(&ovf, b->base.stats + ST_OVERFLOW, 4, hipMemcpyDeviceToHost));
      if (!(ovf & 3u)) break;
      // undo: the step wrote only the other parity
      b->latest ^= 1;
      b->step_index -= 1;
      b->infos = infos0;
      hw.skip_ws_once = skip0;
      hw.drop_prev = drop0;
      if (ovf & 2u) {
        if (team_retry++) return BB_ERR_DEVICE;
        b->team_bodies = 0;
        for (auto& wi : b->infos) wi.mode &= ~kModeTeam;
        st = upload_teams(b);  // also uploads the world table
        if (b->graph) {
          (void)hipGraphExecDestroy(b->graph);
          b->graph = nullptr;
        }
      } else {
        if (g++ == kPairGrowths) return BB_ERR_CAPACITY;
        st = grow_pairs(b);  // also uploads the world table
      }
      if (st == BB_OK) st = bb_batch_step(b, 1, stream, 0);
    }
    if (st != BB_OK) return st;
    // the step that stands has no overflow; the marks of the replayed runs go
    if (g || team_retry) {
      st = clear_overflow(b);
      if (st != BB_OK) return st;
    }
  }
  return BB_OK;
}

bb_status bb_world_sync(bb_world* w) { return w ? sync_batch(w->b) : BB_ERR_INVALID; }

bb_status bb_world_set_param(bb_world* w, bb_param p, uint32_t value) {
  return w ? set_param(w->b, p, value) : BB_ERR_INVALID;
}

bb_status bb_world_read_bodies(bb_world* w, bb_body_state* out, uint32_t n, bb_order order) {
  return w ? read_world(w->b, 0, out, n, order) : BB_ERR_INVALID;
}

bb_status bb_world_read_stats(bb_world* w, bb_step_stats* out) {
  return w ? world_stats(w->b, 0, out) : BB_ERR_INVALID;
}

bb_status bb_world_energy(bb_world* w, double* out) {
  return w ? energy_batch(w->b, out, 1) : BB_ERR_INVALID;
}

bb_status bb_world_instance_transforms(bb_world* w, float* dst, uint32_t n, int dst_on_device) {
  return w ? instances_batch(w->b, 0, dst, n, dst_on_device) : BB_ERR_INVALID;
}

bb_status bb_world_lbvh_nodes(bb_world* w, bb_lbvh_node* dst, uint32_t cap, uint32_t* count, int dst_on_device) {
  return w ? lbvh_batch(w->b, 0, dst, cap, count, dst_on_device) : BB_ERR_INVALID;
}

bb_status bb_batch_lbvh_nodes(bb_batch* b, uint32_t world, bb_lbvh_node* dst, uint32_t cap, uint32_t* count,
                              int dst_on_device) {
  return b ? lbvh_batch(b, world, dst, cap, count, dst_on_device) : BB_ERR_INVALID;
}


bb_status bb_world_snapshot(bb_world* w, void* dst, size_t cap, size_t* size) {
  if (!w) return BB_ERR_INVALID;
  return snapshot_world(w->b, dst, cap, size);
}


bb_status bb_world_restore(const void* src, size_t size, int warm_start, bb_world** out) {
  return restore_world(src, size, warm_start, out);
}

bb_status bb_batch_snapshot(bb_batch* b, void* dst, size_t cap, size_t* size) {
  return b ? snapshot_batch(b, dst, cap, size) : BB_ERR_INVALID;
}

bb_status bb_batch_restore(const void* src, size_t size, int warm_start, bb_batch** out) {
  return restore_batch(src, size, warm_start, out);
}

bb_status bb_world_debug_tap(bb_world* w, bb_tap kind, void* dst, size_t cap, size_t* written) {
  return w ? tap_world(w->b, 0, kind, dst, cap, written) : BB_ERR_INVALID;
}

uint32_t bb_world_body_count(const bb_world* w) { return w ? (uint32_t)w->b->worlds[0].bodies.size() : 0; }

void bb_world_destroy(bb_world* w) {
  if (!w) return;
  bb_batch_destroy(w->b);
  delete w;
}

}  // extern "C"
