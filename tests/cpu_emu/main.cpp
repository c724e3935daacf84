// Emulation driver: steps the worlds of a scene file through the C-ABI.
// usage: emu_main scene.bin steps team_bodies block graph out.bin [lane_bodies]
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "../../include/bb_capi.h"

int main(int argc, char** argv) {
  if (argc < 7) { std::fprintf(stderr, "usage\n"); return 2; }
  FILE* f = std::fopen(argv[1], "rb");
  uint32_t nw = 0;
  if (std::fread(&nw, 4, 1, f) != 1) return 3;
  bb_batch* b = nullptr;
  const bool shapes = (nw & 0x80000000u) != 0;
  nw &= 0x7FFFFFFFu;
  bb_batch_create(nw, 0, &b);
  if (shapes) {
    // voxel shapes: count, then dims[3], voxel size, density, word count, words
    uint32_t ns = 0;
    if (std::fread(&ns, 4, 1, f) != 1) return 3;
    for (uint32_t k = 0; k < ns; ++k) {
      bb_voxel_shape_desc vd{};
      uint32_t nwords = 0;
      if (std::fread(vd.dims, 4, 3, f) != 3 || std::fread(&vd.voxel_size, 4, 1, f) != 1 ||
          std::fread(&vd.density, 4, 1, f) != 1 || std::fread(&nwords, 4, 1, f) != 1)
        return 3;
      std::vector<uint32_t> words(nwords);
      if (nwords && std::fread(words.data(), 4, nwords, f) != nwords) return 3;
      vd.occupancy = words.data();
      bb_voxel_shape_info info{};
      int st = bb_batch_add_voxel_shape(b, &vd, &info);
      if (st) { std::fprintf(stderr, "add_voxel_shape %d\n", st); return 4; }
    }
  }
  std::vector<uint32_t> ns;
  for (uint32_t w = 0; w < nw; ++w) {
    uint32_t dsz = 0, n = 0;
    bb_world_desc d{};
    if (std::fread(&dsz, 4, 1, f) != 1 || dsz != sizeof(d)) { std::fprintf(stderr, "desc size %u vs %zu\n", dsz, sizeof(d)); return 3; }
    if (std::fread(&d, sizeof(d), 1, f) != 1) return 3;
    if (std::fread(&n, 4, 1, f) != 1) return 3;
    std::vector<bb_body_desc> bodies(n);
    if (n && std::fread(bodies.data(), sizeof(bb_body_desc), n, f) != n) return 3;
    uint32_t idx = 0;
    int st = bb_batch_add_world(b, &d, bodies.data(), n, &idx);
    if (st) { std::fprintf(stderr, "add_world %d\n", st); return 4; }
    ns.push_back(n);
  }
  std::fclose(f);
  const int steps = std::atoi(argv[2]);
  bb_batch_set_param(b, BB_PARAM_TEAM_BODIES, (uint32_t)std::strtoul(argv[3], nullptr, 0));
  bb_batch_set_param(b, BB_PARAM_BLOCK, (uint32_t)std::atoi(argv[4]));
  const int graph = std::atoi(argv[5]);
  if (argc > 7) bb_batch_set_param(b, BB_PARAM_LANE_BODIES, (uint32_t)std::strtoul(argv[7], nullptr, 0));
  for (int s = 0; s < steps; ++s) {
    int st = bb_batch_step(b, 1, nullptr, graph);
    if (st) { std::fprintf(stderr, "step %d\n", st); return 5; }
    bb_batch_sync(b);
    for (uint32_t w = 0; w < nw; ++w) {
      bb_step_stats ss{};
      bb_batch_world_stats(b, w, &ss);
      std::printf("step %d world %u pairs %u manifolds %u contacts %u overflow %u\n", s, w,
                  ss.broad_phase_collision_count, ss.collision_count,
                  ss.contact_count, ss.overflow);
    }
    std::fflush(stdout);
  }
  FILE* o = std::fopen(argv[6], "wb");
  for (uint32_t w = 0; w < nw; ++w) {
    std::vector<bb_body_state> out(ns[w]);
    bb_batch_read_world(b, w, out.data(), ns[w], BB_ORDER_BY_ID);
    std::fwrite(out.data(), sizeof(bb_body_state), ns[w], o);
  }
  std::fclose(o);
  bb_batch_destroy(b);
  return 0;
}
