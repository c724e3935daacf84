// Emulation driver: steps the worlds of a scene file through the C-ABI.
// usage: emu_main scene.bin steps team_bodies block graph out.bin [lane_bodies [single [team_spin]]]
// With single = 1 the scene's one world runs through the single-world API
// (bb_world_*: pair-capacity growth on overflow) instead of a batch.
// team_spin sets BB_PARAM_TEAM_SPIN (spins before a team barrier wait gives
// up).  The per-step line reports the step's overflow word; a batch's
// overflow marks are collected after each step and listed.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "../../include/bb_capi.h"

struct Shape {
  bb_voxel_shape_desc vd{};
  std::vector<uint32_t> words;
};
struct World {
  bb_world_desc d{};
  std::vector<bb_body_desc> bodies;
};

int main(int argc, char** argv) {
  if (argc < 7) { std::fprintf(stderr, "usage\n"); return 2; }
  FILE* f = std::fopen(argv[1], "rb");
  uint32_t nw = 0;
  if (std::fread(&nw, 4, 1, f) != 1) return 3;
  const bool has_shapes = (nw & 0x80000000u) != 0;
  nw &= 0x7FFFFFFFu;
  std::vector<Shape> shapes;
  if (has_shapes) {
    // voxel shapes: count, then dims[3], voxel size, density, word count, words
    uint32_t ns = 0;
    if (std::fread(&ns, 4, 1, f) != 1) return 3;
    shapes.resize(ns);
    for (Shape& s : shapes) {
      uint32_t nwords = 0;
      if (std::fread(s.vd.dims, 4, 3, f) != 3 || std::fread(&s.vd.voxel_size, 4, 1, f) != 1 ||
          std::fread(&s.vd.density, 4, 1, f) != 1 || std::fread(&nwords, 4, 1, f) != 1)
        return 3;
      s.words.resize(nwords);
      if (nwords && std::fread(s.words.data(), 4, nwords, f) != nwords) return 3;
    }
  }
  std::vector<World> worlds(nw);
  for (World& w : worlds) {
    uint32_t dsz = 0, n = 0;
    if (std::fread(&dsz, 4, 1, f) != 1 || dsz != sizeof(w.d)) { std::fprintf(stderr, "desc size %u vs %zu\n", dsz, sizeof(w.d)); return 3; }
    if (std::fread(&w.d, sizeof(w.d), 1, f) != 1) return 3;
    if (std::fread(&n, 4, 1, f) != 1) return 3;
    w.bodies.resize(n);
    if (n && std::fread(w.bodies.data(), sizeof(bb_body_desc), n, f) != n) return 3;
  }
  std::fclose(f);
  const int steps = std::atoi(argv[2]);
  const bool single = argc > 8 && std::atoi(argv[8]) != 0;
  if (single && nw != 1) { std::fprintf(stderr, "single-world mode takes one world\n"); return 2; }

  bb_batch* b = nullptr;
  bb_world* sw = nullptr;
  if (single) {
    if (bb_world_create(&worlds[0].d, &sw)) return 4;
  } else {
    bb_batch_create(nw, 0, &b);
  }
  for (Shape& s : shapes) {
    s.vd.occupancy = s.words.data();
    bb_voxel_shape_info info{};
    int st = single ? bb_world_add_voxel_shape(sw, &s.vd, &info) : bb_batch_add_voxel_shape(b, &s.vd, &info);
    if (st) { std::fprintf(stderr, "add_voxel_shape %d\n", st); return 4; }
  }
  for (World& w : worlds) {
    const uint32_t n = (uint32_t)w.bodies.size();
    uint32_t idx = 0;
    int st = single ? bb_world_add_bodies(sw, w.bodies.data(), n, nullptr)
                    : bb_batch_add_world(b, &w.d, w.bodies.data(), n, &idx);
    if (st) { std::fprintf(stderr, "add_world %d\n", st); return 4; }
  }
  const uint32_t spin = argc > 9 ? (uint32_t)std::strtoul(argv[9], nullptr, 0) : 0u;
  if (!single) {
    bb_batch_set_param(b, BB_PARAM_TEAM_BODIES, (uint32_t)std::strtoul(argv[3], nullptr, 0));
    bb_batch_set_param(b, BB_PARAM_BLOCK, (uint32_t)std::atoi(argv[4]));
    if (argc > 7) bb_batch_set_param(b, BB_PARAM_LANE_BODIES, (uint32_t)std::strtoul(argv[7], nullptr, 0));
    if (spin) bb_batch_set_param(b, BB_PARAM_TEAM_SPIN, spin);
    if (argc > 10) bb_batch_set_param(b, BB_PARAM_SORT_BLOCKS, (uint32_t)std::strtoul(argv[10], nullptr, 0));
    if (argc > 11) bb_batch_set_param(b, BB_PARAM_WALK_BUDGET, (uint32_t)std::strtoul(argv[11], nullptr, 0));
    if (argc > 12) bb_batch_set_param(b, BB_PARAM_PGS_BLOCK, (uint32_t)std::strtoul(argv[12], nullptr, 0));
    if (argc > 13) bb_batch_set_param(b, BB_PARAM_PGS_LDS_ROWS, (uint32_t)std::strtoul(argv[13], nullptr, 0));
  } else {
    bb_world_set_param(sw, BB_PARAM_TEAM_BODIES, (uint32_t)std::strtoul(argv[3], nullptr, 0));
    if (spin) bb_world_set_param(sw, BB_PARAM_TEAM_SPIN, spin);
    if (argc > 10) bb_world_set_param(sw, BB_PARAM_SORT_BLOCKS, (uint32_t)std::strtoul(argv[10], nullptr, 0));
    if (argc > 11) bb_world_set_param(sw, BB_PARAM_WALK_BUDGET, (uint32_t)std::strtoul(argv[11], nullptr, 0));
  }
  const int graph = std::atoi(argv[5]);
  for (int s = 0; s < steps; ++s) {
    int st = single ? bb_world_step(sw, 1, nullptr) : bb_batch_step(b, 1, nullptr, graph);
    if (st) { std::fprintf(stderr, "step %d\n", st); return 5; }
    if (!single) {
      st = bb_batch_sync(b);
      if (st == BB_ERR_OVERFLOW) {
        std::vector<uint32_t> ow(nw);
        uint32_t cnt = 0;
        if (bb_batch_overflow(b, ow.data(), nw, &cnt)) return 5;
        for (uint32_t k = 0; k < cnt && k < nw; ++k) std::printf("step %d overflow world %u\n", s, ow[k]);
      } else if (st) {
        std::fprintf(stderr, "sync %d\n", st);
        return 5;
      }
    }
    for (uint32_t w = 0; w < nw; ++w) {
      bb_step_stats ss{};
      if (single) bb_world_read_stats(sw, &ss);
      else bb_batch_world_stats(b, w, &ss);
      std::printf("step %d world %u pairs %u manifolds %u contacts %u overflow %u\n", s, w,
                  ss.broad_phase_collision_count, ss.collision_count,
                  ss.contact_count, ss.overflow);
    }
    std::fflush(stdout);
  }
  FILE* o = std::fopen(argv[6], "wb");
  for (uint32_t w = 0; w < nw; ++w) {
    const uint32_t n = (uint32_t)worlds[w].bodies.size();
    std::vector<bb_body_state> out(n);
    if (single) bb_world_read_bodies(sw, out.data(), n, BB_ORDER_BY_ID);
    else bb_batch_read_world(b, w, out.data(), n, BB_ORDER_BY_ID);
    std::fwrite(out.data(), sizeof(bb_body_state), n, o);
  }
  std::fclose(o);
  if (single) bb_world_destroy(sw);
  else bb_batch_destroy(b);
  return 0;
}
