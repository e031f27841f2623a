// PMHuberStereo.h -- C++ drop-in for cvt::PMHuberStereo's plain
// PatchMatch path (cvt/vision/PMHuberStereo.h:36-110), implemented over the
// C ABI in pm_stereo.h.
//
// Same names, argument meaning and defaults as the reference class:
// disparityMap(dmap, left, right, patchsize, depthmax, iterations, pp,
// normalmap, initdisparity, initnormalmap).  The reference keeps init and
// propagate private; here setup / iterate / planes are public, which is the
// setup-then-iterate interface a caller drives directly.  Errors throw
// cvt_amd::Exception (the reference throws CVTException).
//
// Pipeline: by default the reference's disparityMap one, PatchMatch coupled
// to a diffusion-tensor ROF smoothing of the (normal, depth) fields after
// every iteration (PM_SMOOTH_HUBER, PMHuberStereo.cpp:406-581); constructed
// with PM_SMOOTH_NONE it runs plain PatchMatch (theta = 0, no smoothing).
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "pm_stereo.h"

namespace cvt_amd {

enum DisparityPostProcessing {
  DISPARITY_PP_NONE = PM_PP_NONE,
  DISPARITY_PP_LRCHECK = PM_PP_LRCHECK,
  DISPARITY_PP_LRCHECK_FILL = PM_PP_LRCHECK_FILL
};

class Exception : public std::runtime_error {
 public:
  Exception(int code, const std::string& msg)
      : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

// The data contract of cvt::Image on this path: W x H float32, GRAY_FLOAT
// (1 channel) or RGBA_FLOAT (4 channels), row stride in floats.
class Image {
 public:
  Image() = default;
  Image(size_t width, size_t height, size_t channels) {
    reallocate(width, height, channels);
  }
  void reallocate(size_t width, size_t height, size_t channels) {
    if (channels != 1 && channels != 4)
      throw Exception(PM_ERR_ARG, "Image: channels must be 1 or 4");
    w_ = width;
    h_ = height;
    c_ = channels;
    stride_ = width * channels;
    data_.assign(stride_ * height, 0.0f);
  }
  size_t width() const { return w_; }
  size_t height() const { return h_; }
  size_t channels() const { return c_; }
  size_t stride() const { return stride_; }  // floats per row
  float* ptr() { return data_.data(); }
  const float* ptr() const { return data_.data(); }
  float& at(size_t x, size_t y, size_t ch = 0) {
    return data_[y * stride_ + x * c_ + ch];
  }
  float at(size_t x, size_t y, size_t ch = 0) const {
    return data_[y * stride_ + x * c_ + ch];
  }

 private:
  size_t w_ = 0, h_ = 0, c_ = 0, stride_ = 0;
  std::vector<float> data_;
};

class PMHuberStereo {
 public:
  explicit PMHuberStereo(int device = 0,
                         pm_smoothing smoothing = PM_SMOOTH_HUBER);
  ~PMHuberStereo();
  PMHuberStereo(const PMHuberStereo&) = delete;
  PMHuberStereo& operator=(const PMHuberStereo&) = delete;

  // PMHuberStereo.h:47-48.  Like disparityMapAll (PMHuberStereo.cpp:490),
  // the iteration count is made odd.
  void disparityMap(Image& dmap, const Image& left, const Image& right,
                    size_t patchsize, float depthmax, size_t iterations,
                    DisparityPostProcessing pp = DISPARITY_PP_LRCHECK,
                    Image* normalmap = nullptr,
                    const Image* initdisparity = nullptr,
                    const Image* initnormalmap = nullptr);

  // PMHuberStereo.h:50-51, PMHuberStereo.cpp:147-347: depthMap's own
  // coupling (PM_SMOOTH_DEPTHMAP: Huber smoothing of filled fields, theta
  // from 1e-3), exactly `iterations` iterations, the disparity scaled by
  // dscale (default 1 / depthmax).
  void depthMap(Image& dmap, const Image& left, const Image& right,
                size_t patchsize, float depthmax, size_t iterations,
                float dscale = -1.0f,
                DisparityPostProcessing pp = DISPARITY_PP_LRCHECK,
                Image* normalmap = nullptr,
                const Image* initdisparity = nullptr,
                const Image* initnormalmap = nullptr);

  // setup / iterate interface
  void setup(const Image& left, const Image& right, size_t patchsize,
             float depthmax, const Image* initdisparity = nullptr,
             const Image* initnormalmap = nullptr);
  void iterate(size_t n);
  size_t iteration() const;
  // plane field of a view (0 = left, 1 = right) as RGBA (a, b, c, cost)
  void planes(Image& out, int view);
  void setPlanes(const Image& in, int view, size_t iteration);
  // PM_SMOOTH_HUBER: the smoothed (normal, disparity / depthmax) field of a
  // view
  void smoothField(Image& out, int view);
  // LR check at (maxdispdiff px, maxanglediff degrees) + conversion
  void disparity(Image& dmap, DisparityPostProcessing pp,
                 Image* normalmap = nullptr, float maxdispdiff = 1.0f,
                 float maxanglediff = 5.0f);

  pm_ctx* handle() { return ctx_; }

 private:
  void check(int rc) const;
  void configure(size_t patchsize, float depthmax);

  int device_;
  pm_smoothing smoothing_;
  pm_smoothing mode_;      // coupling of the next context (depthMap switches)
  pm_smoothing ctx_mode_ = PM_SMOOTH_NONE;
  pm_ctx* ctx_ = nullptr;
  int radius_ = -1;
  float dmax_ = -1.0f;
  size_t w_ = 0, h_ = 0;
};

}  // namespace cvt_amd
