// C++ host-side mirror of ToyTrace's rendering-algorithm interface over the
// C ABI of libttgpu.so (header only).
//
// RenderingAlgo follows src/renderingalgorithms_interface.hxx:13-30 (Run,
// RequestInterrupt, RequestFullStop, GenerateImage, SetInterruptCallback,
// InitializeScene, a noexcept(false) destructor); RenderAlgorithmFactory follows
// src/renderingalgorithms.cxx:27-42 for the keys "pt", "pt2", "ptg" and "bdpt".  Errors the
// library reports as status codes are thrown as the exceptions the reference
// throws: std::invalid_argument for bad scenes and parameters,
// std::runtime_error otherwise.
#pragma once
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ttgpu.h"

namespace ttgpu {

struct Image8 {
  int width = 0, height = 0;
  std::vector<uint8_t> rgb;  // row-major, 3 bytes per pixel, top row first
};

class RenderingAlgo {
 public:
  virtual ~RenderingAlgo() noexcept(false) {}
  virtual void InitializeScene() {}
  virtual void Run() = 0;
  virtual void RequestInterrupt() = 0;
  virtual void RequestFullStop() = 0;
  virtual std::unique_ptr<Image8> GenerateImage() = 0;
  void SetInterruptCallback(std::function<void(bool)> cb) { irq_cb_ = std::move(cb); }

 protected:
  void CallInterruptCb(bool is_complete_pass) {
    if (irq_cb_) irq_cb_(is_complete_pass);
  }

 private:
  std::function<void(bool)> irq_cb_;
};

inline void ThrowOnError(int rc, const char *msg) {
  if (rc == TTGPU_OK) return;
  const std::string m = msg ? msg : "";
  if (rc == TTGPU_ERR_INVALID_ARGUMENT) throw std::invalid_argument(m);
  throw std::runtime_error(m);
}

// pt / pt2 / ptg / bdpt on the GPU.  The scene description and its arrays are borrowed: they must outlive
// InitializeScene (the library copies what it needs to the device there).
class GpuRenderingAlgo : public RenderingAlgo {
 public:
  GpuRenderingAlgo(const ttgpu_scene &scene, const ttgpu_params &params) : scene_(&scene), params_(params) {}
  ~GpuRenderingAlgo() noexcept(false) override {
    if (h_) ttgpu_destroy(h_);
  }
  GpuRenderingAlgo(const GpuRenderingAlgo &) = delete;
  GpuRenderingAlgo &operator=(const GpuRenderingAlgo &) = delete;

  void InitializeScene() override {
    if (h_) return;
    // The message is read only after the call: argument evaluation order is unspecified.
    const int rc = ttgpu_create(scene_, &params_, &h_);
    if (rc != TTGPU_OK) ThrowOnError(rc, ttgpu_create_error());
    Check(ttgpu_set_pass_callback(h_, &GpuRenderingAlgo::PassCb, this));
    if (stop_requested_) ttgpu_request_stop(h_);
  }
  void Run() override {
    InitializeScene();
    const int rc = ttgpu_run(h_, 0);
    if (rc != TTGPU_ERR_INTERRUPTED) Check(rc);
  }
  void RequestInterrupt() override {
    if (h_) ttgpu_request_interrupt(h_);
  }
  void RequestFullStop() override {
    stop_requested_ = true;
    if (h_) ttgpu_request_stop(h_);
  }
  std::unique_ptr<Image8> GenerateImage() override {
    InitializeScene();
    auto img = std::make_unique<Image8>();
    img->width = params_.width;
    img->height = params_.height;
    img->rgb.resize(3 * (size_t)params_.width * params_.height);
    Check(ttgpu_generate_image(h_, img->rgb.data()));
    return img;
  }
  // Linear RGB sums and samples per pixel (for HDR output and tests).
  void ReadFramebuffer(std::vector<double> &rgb, std::vector<uint64_t> &spp) {
    InitializeScene();
    const size_t n = (size_t)params_.width * params_.height;
    rgb.resize(3 * n);
    spp.resize(n);
    Check(ttgpu_read_framebuffer(h_, rgb.data(), nullptr, spp.data()));
  }
  ttgpu_handle *handle() { return h_; }

 private:
  void Check(int rc) {
    if (rc != TTGPU_OK) ThrowOnError(rc, ttgpu_last_error(h_));
  }
  static void PassCb(void *self, int32_t complete) { static_cast<GpuRenderingAlgo *>(self)->CallInterruptCb(complete != 0); }
  const ttgpu_scene *scene_;
  ttgpu_params params_;
  ttgpu_handle *h_ = nullptr;
  bool stop_requested_ = false;
};

// RenderAlgorithmFactory: "pt" (the random-walk forward path tracer, surfaces), "pt2" (forward path
// tracing with NEE and null-path tracking), "ptg" (guided; params.guide_kind selects the quadtree,
// the vMF tree or the learned guide) and "bdpt" (bidirectional, surfaces).
inline std::unique_ptr<RenderingAlgo> RenderAlgorithmFactory(const ttgpu_scene &scene, ttgpu_params params,
                                                             const std::string &name) {
  if (name == "pt2") params.algo = TTGPU_ALGO_PT2;
  else if (name == "ptg") params.algo = TTGPU_ALGO_PTG;
  else if (name == "pt") params.algo = TTGPU_ALGO_PT;
  else if (name == "bdpt") params.algo = TTGPU_ALGO_BDPT;
  else throw std::invalid_argument("unknown or unsupported algorithm " + name + "; available: pt, pt2, ptg, bdpt");
  return std::make_unique<GpuRenderingAlgo>(scene, params);
}

}  // namespace ttgpu
