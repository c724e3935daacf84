// bb_dev.hpp — device data layout of a batch of worlds (host + device).
//
// All worlds of a batch live in one set of arrays.  World w owns body rows
// [body_off, body_off + n), LBVH node slots [2*body_off, 2*body_off + 2n),
// pair/manifold slots [pair_off, pair_off + pair_cap) and active slots
// [act_off, act_off + a).  Body state is structure-of-arrays, double
// buffered: step K reads parity K^1 ("prev", the last step's Morton rows) and
// gathers it into parity K ("cur") in the new Morton order, so the row copy at
// the end of the reference step (entry_update_rigid_bodies) is the reorder
// itself.  Immutable per-body properties are indexed by id and never move.
#pragma once
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace bb {

static constexpr uint32_t kMaxContacts = 8;   // MAX_CONTACT_POINT_COUNT
static constexpr uint32_t kColors = 32;       // BB_MAX_COLORS
static constexpr uint32_t kBuckets = kColors + 1;  // + overflow bucket
static constexpr uint32_t kAvbdColors = 32;
static constexpr uint32_t kAvbdLayers = 12;
static constexpr uint32_t kMAX = 0xFFFFFFFFu;
// Broad phase: nodes one thread walks for its row before the row is handed to
// a whole workgroup (BB_PARAM_WALK_BUDGET's default)
static constexpr uint32_t kWalkBudgetDefault = 128;

// per-world statistics slots (bb_step_stats order)
enum StatSlot : uint32_t {
  ST_N = 0, ST_A, ST_P, ST_M, ST_NODES, ST_C, ST_ISLANDS, ST_CISLANDS, ST_SLEEPING,
  ST_GC_COUNT, ST_GC_VIOL, ST_GC_OVF, ST_AV_COUNT, ST_AV_VIOL, ST_AV_STICK, ST_FRESH,
  ST_PEN, ST_IDSUM, ST_MAXV, ST_OVERFLOW, ST_USED_COLORS,
  // OR of ST_OVERFLOW over every step since the host last read and cleared it
  // (the step's own reset leaves it alone)
  ST_OVF_STICKY, ST_COUNT
};

struct WorldInfo {  // 64 bytes
  uint32_t body_off, n, a, act_off;
  uint32_t pair_off, pair_cap, flags, solver;
  uint32_t iterations, morton_bits, quirks, prev_n;
  float dt, gravity, damp;
  uint32_t mode;  // kModeDropPrev | kModeTeam | kModeVoxel
};
// WorldInfo::mode bits.  kModeDropPrev: the previous step left nothing to
// warm start from.  kModeTeam: the world is stepped by a group of workgroups
// in k_team instead of one workgroup in the per-world stage kernels.
// kModeVoxel: the world has voxel bodies; each pair with a voxel body takes
// six consecutive pair slots, one per contact-face code (keys 90..95).
static constexpr uint32_t kModeDropPrev = 1u, kModeTeam = 2u, kModeVoxel = 4u;

// A voxel shape of the batch's pools (VoxelShape, include/shared.inl:
// 938-952): occupancy bits at vocc[occ_off], bit x + y*dx + z*dx*dy, and
// surface cells at vsurf[surf_off], x | y<<8 | z<<16 | normal code<<24.
struct VoxShape {  // 48 bytes
  uint32_t dims[3];
  uint32_t occ_off;
  float vs;
  float ox, oy, oz;  // corner of cell (0,0,0) in the body frame
  uint32_t surf_off, surf_count, pad0, pad1;
};

// Control block of a workgroup team (k_team): barrier arrival counter and
// generation, an error flag (a barrier wait that ran out of spins), a
// team-wide flag for the radix sort, the count of workgroups that have left
// the step (the last one resets count, error and exited for the next launch),
// and per-workgroup scan partials.
static constexpr uint32_t kMaxTeamBlocks = 2048;
// Threads per workgroup of k_team (independent of the per-world kernels').
static constexpr uint32_t kTeamBlock = 512;
// The radix sort runs on the first workgroups of the team only (a sub-team,
// sort_blocks of them), with a barrier of its own (scount, sgen); `sorted`
// hands the side the sorted keys ended on to the rest of the team.
struct TeamCtl {
  uint32_t count;   // team barrier arrivals in this launch
  uint32_t error;
  uint32_t diff;
  uint32_t exited;
  uint32_t scount;  // the sort sub-team's barrier arrivals in this launch
  uint32_t sorted;
  uint32_t part[kMaxTeamBlocks];
};
// Workgroups of a team that run its radix sort when the launch parameter is
// 0: fewer than the team's own, so the sort's barriers stay cheap.
static constexpr uint32_t kTeamSortBlocks = 32;
// Team j steps world world[j] with workgroups [first[j], first[j+1]) of the
// k_team launch; workgroups from first[count] on have nothing to do.
static constexpr uint32_t kMaxTeams = 256, kTeamSharedBytes = 4096;
// Stage clock marks of a team step (profiling): start, then the end of the
// build, broad, narrow (with contacts), post, PGS (with levels), AVBD and
// finish stages (marks 0-7, kTeamStageMarks); then marks inside the build
// stage: Morton keys done, sort done, LBVH hierarchy done, the teamwide part
// of the bounds pass done, the lead workgroup's climb of the top nodes done
// (stamped by rank 0, a member of the lead workgroup; the reorder follows,
// up to mark 1); and inside the AVBD stage (from mark 5): adjacency,
// colouring and prepare done, warm start, support depth and the (colour,
// layer) grouping done, the main sweeps with their dual updates done (the
// finalize and the post-stabilisation follow, up to mark 6).
static constexpr uint32_t kTeamStageMarks = 8;
static constexpr uint32_t kMarkKeys = 8, kMarkSort = 9, kMarkHierarchy = 10, kMarkBounds = 11, kMarkClimb = 12;
static constexpr uint32_t kMarkAvColour = 13, kMarkAvGroup = 14, kMarkAvSweeps = 15;
static constexpr uint32_t kTeamMarks = 16;
struct TeamTable {
  uint32_t count;
  uint32_t first[kMaxTeams + 1];
  uint32_t world[kMaxTeams];
};

struct Rows {  // one parity of body rows
  uint32_t* id;
  uint32_t* flags;
  uint32_t* sleep;
  float4* pos;
  float4* rot;
  float4* vel;
  float4* omg;
  float4* pvel;
};

struct Props {  // by global id slot (body_off + id)
  float4* pmin;  // minimum xyz, w = radius
  float4* pmax;  // maximum xyz, w = shape (0 box, 1 sphere) as float
  float4* mass;  // mass, inv_mass, friction, restitution
  float4* ii0;   // inv_inertia rows
  float4* ii1;
  float4* ii2;
  uint32_t* active_rank;  // active index (ascending id among DYNAMIC), MAX if static
  uint32_t* active_ids;   // [act_off + k] = id
};

struct Manifolds {  // pair-indexed slots; contact arrays are [slot*8 + c]
  int4* hdr;     // a, b, key, count (count 0: no manifold)
  float4* nrm;   // normal xyz, error (int bits)
  float4* t0;
  float4* t1;
  float4* anca;  // body a position at detection
  float4* ancb;
  float4* cpos;  // contact position xyz, penetration
  float4* clam;  // normal impulse, tangent impulse 0/1, mass coefficient
  float4* cmas;  // normal mass, tangent mass 0/1, bias factor
  float4* cana;  // anchor a xyz, impulse coefficient
  float4* canb;  // anchor b xyz, stick (uint bits)
  uint4* cfp;    // feature pair in_ref, out_ref, in_inc, out_inc
  uint32_t* cws; // warm start counter
};

struct Dev {
  WorldInfo* worlds;
  uint32_t W;
  // voxel shape pools, shared by the batch's worlds; any_voxel: a world has
  // voxel bodies (k_narrow_vox is launched); any_island: a world runs the
  // per-island solver path (k_levels and k_pgs_island are launched)
  uint32_t any_voxel, any_island;
  const VoxShape* vshape;
  const uint32_t* vocc;
  const uint32_t* vsurf;
  Rows cur, prev;
  Props props;
  uint32_t* entries_cur;   // [body_off + id] = row
  uint32_t* entries_prev;
  // sort scratch [body]
  uint64_t* keys_a;
  uint64_t* keys_b;
  uint32_t* idx_a;
  uint32_t* idx_b;
  // LBVH [2*body]
  float4* nmin;  // xyz, w = left child (int bits, -1 for a leaf)
  float4* nmax;  // xyz, w = right child (int bits) or the leaf row
  uint32_t* parent;
  uint32_t* visit;
  // broad phase [body]
  uint32_t* pcount;
  uint32_t* poff;
  uint32_t* pstart_cur;   // [body] first pair of each row (uncapped offset), kept
  uint32_t* pstart_prev;  // for the next step's warm-start lookup
  // pairs / manifolds [pair slot]
  uint2* pairs_cur;
  uint2* pairs_prev;
  uint32_t* npairs_cur;  // [W]
  uint32_t* npairs_prev;
  Manifolds man_cur, man_prev;
  // islands / sleep / colouring
  uint32_t* link;     // [active slot]
  uint32_t* deg;      // [body] manifolds per row
  uint32_t* row_prev; // [body] previous step's row of this row's body (entries_prev by id)
  uint32_t* row_rank; // [body] active rank of this row's body (kMAX: not dynamic)
  uint32_t* island;   // [body] island root (tap)
  uint32_t* islot;    // [body] island slot (ascending root order)
  uint32_t* ascratch; // [active slot] per-island scratch
  uint32_t* color;    // [pair slot]
  uint32_t* pick;     // [pair slot]
  uint32_t* mask;     // [body]
  uint32_t* owner;    // [body*32], zero between rounds
  uint32_t* bucket;   // [pair slot] manifold slots grouped by colour
  uint32_t* bucket_off;  // [W * (kBuckets+1)]
  uint32_t* nlev;        // [W] exact dependency levels of the per-island path (0: none)
  float4* iw;         // [body*3] cached world inverse inertia rows, row0.w = inv_mass
  // PGS solver rows (colour-sorted, [pair slot]; contact c of row k at c*np+k)
  uint32_t np;        // pair slots in the batch
  uint32_t pgs_lcap;  // rows per world whose body state k_pgs keeps in LDS (80 B each)
  int4* sv_hdr;       // a, b, count | 256 a dynamic | 512 b dynamic | 1024 one immovable, friction bits
  float4* sv_n;       // normal
  float4* sv_r1;      // [8*np] arm a, penetration
  float4* sv_r2;      // [8*np] arm b, normal mass
  float4* sv_l;       // [8*np] normal/tangent impulses, tangent mass 0
  float* sv_tm1;      // [8*np] tangent mass 1
  // AVBD
  float4* s_pos0;     // [body] pos_start, w = support depth bits
  float4* s_rot0;
  float4* s_post;
  float4* s_rott;
  uint32_t* depth;    // [body]
  uint32_t* depth_snap;
  uint32_t* bcolor;   // [body]
  uint32_t* bcolor_snap;
  uint32_t* adj_off;  // [body] CSR offsets (world-local, exclusive)
  uint32_t* adj;      // [2*pair slot]
  uint32_t* bbucket;  // [body] rows grouped by (colour, layer)
  uint32_t* bbucket_off;  // [W * (32*12+1)]
  uint32_t* stats;    // [W * ST_COUNT]
  // teams (k_team): table, per-team control blocks, radix histograms
  // [launch workgroups * 256] and stage scalars [kMaxTeams * kTeamSharedBytes]
  TeamTable* teams;
  TeamCtl* team_ctl;
  uint32_t* team_hist;
  char* team_shared;
  uint32_t team_grid;     // workgroups of the k_team launch
  uint32_t team_vox;      // a team world holds voxel bodies (the k_team instance with their narrow phase)
  uint32_t team_spin;     // spins before a team barrier wait gives up (0: default)
  uint32_t sort_blocks;   // workgroups of a team that sort (0: kTeamSortBlocks)
  uint32_t avbd_lanes;    // lanes per body in the per-world AVBD primal sweeps (1, 2, 4, 8; 0: 4)
  uint32_t pgs_prefetch;  // 1: the colour sweeps fetch each thread's next-colour row early
  uint32_t walk_budget;   // nodes a broad-phase thread walks before its workgroup takes the row
                          // (0: kWalkBudgetDefault)
  // set (bit 0) by any world whose step flags an overflow; the host reads
  // and clears it
  uint32_t* ovf_any;
  // [kMaxTeams * kTeamMarks] wall-clock marks of each team's stages, written
  // when non-null (profiling launches only)
  unsigned long long* team_clock;
  // [W] world stepped by workgroup b of the per-world kernels: worlds in
  // descending body count, so the longest workgroups start first
  const uint32_t* order;
};

// One launch lane of the per-world kernels: `grid` workgroups of `block`
// threads stepping worlds d.order[0..grid).
struct Lane {
  Dev d;
  uint32_t block, grid;
  // the whole step of each world in one launch (k_fused) instead of one
  // launch per stage
  bool fused = false;
  // threads per workgroup of this lane's colour sweeps (k_pgs); 0: block.
  // Small worlds have a few dozen rows per colour: one wave per world turns
  // each colour's barrier into a wave's and lets more worlds share a CU.
  uint32_t pgs_block = 0;
};

}  // namespace bb
