// HipAccelerator: the C++ device-accelerator facade over the C ABI.
#include "hip_accelerator.h"

namespace nova::aggregate {

static void throw_if(nova_hip_ctx *ctx, int code) {
  if (code != NOVA_HIP_SUCCESS) throw HipDeviceError(code, ctx ? nova_hip_last_error(ctx) : "no context");
}

HipAccelerator::HipAccelerator(int device) {
  int code = nova_hip_create(device, &ctx_);
  if (code != NOVA_HIP_SUCCESS) throw HipDeviceError(code, "nova_hip_create failed");
}

HipAccelerator::~HipAccelerator() {
  if (ctx_) nova_hip_destroy(ctx_);
}

void HipAccelerator::build(primitive_aggregate_data_s scene) {
  throw_if(ctx_, nova_hip_scene_upload(ctx_, &scene, nullptr));
  throw_if(ctx_, nova_hip_bvh_build(ctx_, 4, &info_, nullptr));
}

void HipAccelerator::cleanup() {
  if (!ctx_) return;
  int device = 0;
  nova_hip_destroy(ctx_);
  ctx_ = nullptr;
  throw_if(nullptr, nova_hip_create(device, &ctx_));
}

void HipAccelerator::copyParamsToDevice(const device_traversal_param_s &params) const { params_ = params; }

namespace {
class HipIntersector final : public DeviceIntersectorInterface {
 public:
  HipIntersector(nova_hip_ctx *ctx, const device_traversal_param_s *params) : ctx_(ctx), params_(params) {}
  void traverse(const device_traversal_param_s &p) const override {
    nova_hip_render_params rp{};
    rp.width = (int32_t)p.width;
    rp.height = (int32_t)p.height;
    rp.max_depth = (int32_t)p.depth;
    rp.frame_start = (int32_t)p.sample_index;
    rp.frame_count = 1;
    rp.max_samples_offset = p.sample_max;
    rp.vertical_invert = p.vertical_invert ? 1 : 0;
    rp.rr_start = p.rr_start;
    rp.sampling_roughness_texture = -1;
    for (int i = 0; i < 16; i++) {
      rp.inv_P[i] = p.inv_P[i];
      rp.inv_V[i] = p.inv_V[i];
    }
    throw_if(ctx_, nova_hip_render_frames(ctx_, &rp, p.accumulator_buffer, p.final_buffer, p.partial_buffer, p.stream));
  }

 private:
  nova_hip_ctx *ctx_;
  const device_traversal_param_s *params_;
};
}  // namespace

std::unique_ptr<DeviceIntersectorInterface> HipAccelerator::getIntersectorObject() const {
  return std::make_unique<HipIntersector>(ctx_, &params_);
}

std::unique_ptr<DeviceAcceleratorInterface> DeviceAcceleratorInterface::make(int device) { return std::make_unique<HipAccelerator>(device); }

}  // namespace nova::aggregate
