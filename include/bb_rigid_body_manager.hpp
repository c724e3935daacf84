// bb_rigid_body_manager.hpp — C++ host mirror of the reference simulation
// manager over the C-ABI (bb_capi.h).
//
// The method names, argument meaning and error behaviour follow the
// reference's RigidBodyManager (src/rigid_body_manager.hpp:14-116) and the
// add-body path of SceneManager/AccelerationStructureManager
// (src/scene_manager.hpp:825-929, src/acceleration_structure_manager.hpp:49),
// so a caller written against the reference keeps its call sequence:
//
//   bb::RigidBodyManager rbm;
//   rbm.create("sim", /*iterations=*/8);
//   rbm.build_accel_structs(bodies);        // load_scene's upload
//   while (running) { rbm.simulate(); rbm.read_back_sim_config(); }
//
// Header only; the device work happens in libbeatbox_amd.so.
#pragma once

#include <cstdint>
#include <vector>

#include "bb_capi.h"

namespace bb {

class RigidBodyManager {
 public:
  RigidBodyManager() { bb_world_desc_default(&desc_); }
  ~RigidBodyManager() { destroy(); }
  RigidBodyManager(const RigidBodyManager&) = delete;
  RigidBodyManager& operator=(const RigidBodyManager&) = delete;

  // create(name, renderer, gui, iterations): false when already created or on
  // a device error (reference returns false when initialized).
  bool create(const char* /*name*/, uint32_t iterations = 5) {
    if (world_) return false;
    desc_.iterations = iterations;
    last_ = bb_world_create(&desc_, &world_);
    return last_ == BB_OK;
  }
  // Parity mode (reference keys, caps and quirks) or extended mode.
  void configure(uint32_t morton_bits, uint32_t quirks) {
    desc_.morton_bits = morton_bits;
    desc_.quirks = quirks;
  }
  void destroy() {
    if (world_) bb_world_destroy(world_);
    world_ = nullptr;
  }
  // simulate(): one step.  Like the reference, returns true when called
  // before create() (src/rigid_body_manager.cpp:1608-1611).
  bool simulate() {
    if (!world_) return true;
    last_ = bb_world_step(world_, 1, nullptr);
    return last_ == BB_OK;
  }
  bool read_back_sim_config() {
    if (!world_) return true;
    last_ = bb_world_read_stats(world_, &stats_);
    return last_ == BB_OK;
  }
  const bb_step_stats& get_sim_config_reference() const { return stats_; }

  // build_accel_structs(bodies): uploads the scene with the load_scene body
  // rules; false (nothing added) when the capacity is exceeded.
  bool build_accel_structs(const std::vector<bb_body_desc>& bodies, std::vector<uint32_t>* ids = nullptr) {
    if (!world_) return false;
    std::vector<uint32_t> tmp(bodies.size());
    last_ = bb_world_add_bodies(world_, bodies.data(), (uint32_t)bodies.size(), tmp.data());
    if (ids) *ids = tmp;
    return last_ == BB_OK;
  }
  // SceneManager::build_voxel_shape (src/scene_manager.hpp:478-566): a
  // concave shape for voxel bodies (shape = BB_SHAPE_VOXEL_OF(info.shape_id),
  // push_voxel_body's mass and inertia come from the shape).
  bool build_voxel_shape(const bb_voxel_shape_desc& desc, bb_voxel_shape_info* info) {
    if (!world_ || !info) return false;
    last_ = bb_world_add_voxel_shape(world_, &desc, info);
    return last_ == BB_OK;
  }
  // Body readback through the id -> row map (the TLAS-update consumer).
  bool read_bodies(std::vector<bb_body_state>& out, bb_order order = BB_ORDER_BY_ID) {
    if (!world_) return false;
    out.resize(bb_world_body_count(world_));
    last_ = bb_world_read_bodies(world_, out.data(), (uint32_t)out.size(), order);
    return last_ == BB_OK;
  }
  // entry_update_acceleration_structures (src/shaders/AS_mngr.slang:5-36):
  // per body id, the 3x4 [R | t] instance transform, 12 floats each, into
  // device memory (queued on the world's stream) or host memory.
  bool update_instances(float* dst, uint32_t capacity_bodies, bool dst_on_device) {
    if (!world_) return false;
    last_ = bb_world_instance_transforms(world_, dst, capacity_bodies, dst_on_device ? 1 : 0);
    return last_ == BB_OK;
  }
  // Checkpoint of the simulation (bb_world_snapshot) and a manager resumed
  // from one (bb_world_restore): warm_start keeps the exact continuation,
  // otherwise the next step starts cold as after the reference's pause.
  bool snapshot(std::vector<char>& blob) {
    if (!world_) return false;
    size_t size = 0;
    last_ = bb_world_snapshot(world_, nullptr, 0, &size);
    if (last_ != BB_OK) return false;
    blob.resize(size);
    last_ = bb_world_snapshot(world_, blob.data(), blob.size(), &size);
    return last_ == BB_OK;
  }
  bool restore(const std::vector<char>& blob, bool warm_start = true) {
    destroy();
    last_ = bb_world_restore(blob.data(), blob.size(), warm_start ? 1 : 0, &world_);
    return last_ == BB_OK;
  }

  uint32_t get_sim_flags() const { return world_ ? bb_world_get_flags(world_) : desc_.flags; }
  uint32_t set_sim_flags(uint32_t flags) {
    if (world_) bb_world_set_flags(world_, flags);
    else desc_.flags |= flags;
    return get_sim_flags();
  }
  uint32_t clear_sim_flags(uint32_t flags) {
    if (world_) bb_world_clear_flags(world_, flags);
    else desc_.flags &= ~flags;
    return get_sim_flags();
  }
  void set_sim_type(uint32_t solver) {
    desc_.solver = solver;
    if (world_) bb_world_set_solver(world_, solver);
  }
  uint32_t get_sim_type() const { return desc_.solver; }
  void skip_warm_starting_once() {
    if (world_) bb_world_skip_warm_start_once(world_);
  }
  bb_status last_status() const { return last_; }
  bb_world* handle() const { return world_; }

 private:
  bb_world_desc desc_{};
  bb_world* world_ = nullptr;
  bb_step_stats stats_{};
  bb_status last_ = BB_OK;
};

}  // namespace bb
