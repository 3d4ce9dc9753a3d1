// C++ host facade over the C ABI, shaped like the reference's device-accelerator slot:
//   nova::aggregate::DeviceAcceleratorInterface / DeviceIntersectorInterface
//   (corelib/nova/aggregate/device_acceleration_interface.h:11-38).
// build() uploads the scene and builds the BVH, copyParamsToDevice() stores the frame
// parameters (device_traversal_param_s, corelib/nova/gpu/nova_gpu.h:20-29) and
// traverse() renders one frame (1 spp) into the partial buffer and folds it into the
// accumulator, as gpu_draw -> device_start_integrator does once per frame.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/nova_hip.h"

namespace nova {

// Per-frame launch parameters (the fields of device_traversal_param_s the path uses).
struct device_traversal_param_s {
  float *partial_buffer = nullptr;      // HdrBufferStruct::partial_buffer, RGBA32F, device
  float *accumulator_buffer = nullptr;  // running sum of frames, device
  float *final_buffer = nullptr;        // accumulator / (frame + 1), device
  float inv_P[16] = {}, inv_V[16] = {};
  uint32_t width = 0, height = 0, depth = 0;
  uint32_t sample_index = 0;  // frame index
  uint32_t sample_max = 0;    // renderer_max_samples offset added to the sampler index
  bool vertical_invert = false;
  int32_t rr_start = -1;
  void *stream = nullptr;
};

namespace aggregate {

// primitive_aggregate_data_s stand-in: the flattened C-ABI scene descriptor.
using primitive_aggregate_data_s = nova_hip_scene_desc;

class DeviceIntersectorInterface {
 public:
  virtual ~DeviceIntersectorInterface() = default;
  virtual void traverse(const device_traversal_param_s &params) const = 0;
};

class DeviceAcceleratorInterface {
 public:
  virtual ~DeviceAcceleratorInterface() = default;
  virtual void build(primitive_aggregate_data_s primitive_data_list) = 0;
  virtual void cleanup() = 0;
  virtual unsigned getMaxRecursiveDepth() const = 0;
  virtual void copyParamsToDevice(const device_traversal_param_s &params) const = 0;
  virtual std::unique_ptr<DeviceIntersectorInterface> getIntersectorObject() const = 0;
  /* Replaces the factory's OptiX branch (aggregate/api/device_accelerator_factory.cpp:9-17). */
  static std::unique_ptr<DeviceAcceleratorInterface> make(int device = 0);
};

// Device errors surface as exceptions on the host side, as GenericException does in the
// reference's device layer (engine.cpp:478-486 maps them to RENDERER_EXCEPTION).
class HipDeviceError : public std::runtime_error {
 public:
  HipDeviceError(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

class HipAccelerator final : public DeviceAcceleratorInterface {
 public:
  explicit HipAccelerator(int device = 0);
  ~HipAccelerator() override;
  void build(primitive_aggregate_data_s scene) override;
  void cleanup() override;
  unsigned getMaxRecursiveDepth() const override { return 8; } /* MAX_REC_DEPTH (optix/host.cpp:25) */
  void copyParamsToDevice(const device_traversal_param_s &params) const override;
  std::unique_ptr<DeviceIntersectorInterface> getIntersectorObject() const override;
  nova_hip_ctx *context() const { return ctx_; }
  const nova_hip_bvh_info &bvhInfo() const { return info_; }

 private:
  nova_hip_ctx *ctx_ = nullptr;
  nova_hip_bvh_info info_{};
  mutable device_traversal_param_s params_{};
};

}  // namespace aggregate
}  // namespace nova
