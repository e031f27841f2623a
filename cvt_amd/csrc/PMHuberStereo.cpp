// PMHuberStereo.cpp -- C++ drop-in class over the engine's C ABI
// (mirrors cvt/vision/PMHuberStereo.cpp:75-143).
#include "cvt_amd/PMHuberStereo.h"

namespace cvt_amd {

PMHuberStereo::PMHuberStereo(int device, pm_smoothing smoothing)
    : device_(device), smoothing_(smoothing), mode_(smoothing) {}

PMHuberStereo::~PMHuberStereo() {
  if (ctx_) pm_destroy(ctx_);
}

void PMHuberStereo::check(int rc) const {
  if (rc != PM_OK)
    throw Exception(rc, std::string("PMHuberStereo: ") + pm_last_error(ctx_));
}

// (re)create the context when the patch size or disparity range change
void PMHuberStereo::configure(size_t patchsize, float depthmax) {
  if (ctx_ && (int)patchsize == radius_ && depthmax == dmax_ &&
      mode_ == ctx_mode_)
    return;
  if (ctx_) {
    pm_destroy(ctx_);
    ctx_ = nullptr;
  }
  pm_params p;
  pm_params_default(&p);
  p.radius = (int)patchsize;
  p.disparity_max = depthmax;
  p.smoothing = mode_;
  p.device = device_;
  pm_ctx* c = nullptr;
  int rc = pm_create(&p, &c);
  if (rc != PM_OK) {
    std::string msg = c ? pm_last_error(c) : "pm_create failed";
    if (c) pm_destroy(c);
    throw Exception(rc, "PMHuberStereo: " + msg);
  }
  ctx_ = c;
  ctx_mode_ = mode_;
  radius_ = (int)patchsize;
  dmax_ = depthmax;
}

void PMHuberStereo::setup(const Image& left, const Image& right,
                          size_t patchsize, float depthmax,
                          const Image* initdisparity,
                          const Image* initnormalmap) {
  // PMHuberStereo.cpp:412-414
  if (left.width() != right.width() || left.height() != right.height() ||
      left.channels() != right.channels())
    throw Exception(PM_ERR_ARG,
                    "Left/Right stereo images inconsistent or incompatible");
  configure(patchsize, depthmax);
  check(pm_set_images(ctx_, left.ptr(), right.ptr(), (int)left.width(),
                      (int)left.height(), (int)left.channels(),
                      (int)left.stride()));
  w_ = left.width();
  h_ = left.height();
  if (initdisparity) {
    if (initdisparity->width() != w_ || initdisparity->height() != h_ ||
        initdisparity->channels() != 1)
      throw Exception(PM_ERR_ARG, "initdisparity must be W x H GRAY_FLOAT");
    if (initnormalmap && (initnormalmap->width() != w_ ||
                          initnormalmap->height() != h_ ||
                          initnormalmap->channels() != 4))
      throw Exception(PM_ERR_ARG, "initnormalmap must be W x H RGBA_FLOAT");
    check(pm_setup_prior(ctx_, initdisparity->ptr(),
                         initnormalmap ? initnormalmap->ptr() : nullptr));
  } else {
    check(pm_setup(ctx_));
  }
}

void PMHuberStereo::iterate(size_t n) {
  if (!ctx_) throw Exception(PM_ERR_STATE, "PMHuberStereo: call setup first");
  check(pm_iterate(ctx_, (int)n));
}

size_t PMHuberStereo::iteration() const {
  return ctx_ ? (size_t)pm_iteration(ctx_) : 0;
}

void PMHuberStereo::planes(Image& out, int view) {
  if (!ctx_) throw Exception(PM_ERR_STATE, "PMHuberStereo: call setup first");
  out.reallocate(w_, h_, 4);
  check(pm_get_planes(ctx_, view, out.ptr()));
}

void PMHuberStereo::setPlanes(const Image& in, int view, size_t iteration) {
  if (!ctx_) throw Exception(PM_ERR_STATE, "PMHuberStereo: call setup first");
  if (in.width() != w_ || in.height() != h_ || in.channels() != 4)
    throw Exception(PM_ERR_ARG, "setPlanes: W x H RGBA_FLOAT expected");
  check(pm_set_planes(ctx_, view, in.ptr(), (int)iteration));
}

void PMHuberStereo::smoothField(Image& out, int view) {
  if (!ctx_) throw Exception(PM_ERR_STATE, "PMHuberStereo: call setup first");
  out.reallocate(w_, h_, 4);
  check(pm_get_smooth(ctx_, view, out.ptr()));
}

void PMHuberStereo::disparity(Image& dmap, DisparityPostProcessing pp,
                              Image* normalmap, float maxdispdiff,
                              float maxanglediff) {
  if (!ctx_) throw Exception(PM_ERR_STATE, "PMHuberStereo: call setup first");
  dmap.reallocate(w_, h_, 1);
  check(pm_disparity(ctx_, (int)pp, maxdispdiff, maxanglediff, dmap.ptr(),
                     nullptr));
  if (normalmap) {
    normalmap->reallocate(w_, h_, 4);
    check(pm_normals(ctx_, (int)pp, normalmap->ptr()));
  }
}

void PMHuberStereo::disparityMap(Image& dmap, const Image& left,
                                 const Image& right, size_t patchsize,
                                 float depthmax, size_t iterations,
                                 DisparityPostProcessing pp, Image* normalmap,
                                 const Image* initdisparity,
                                 const Image* initnormalmap) {
  const float maxdispdiff = 1.0f;   // PMHuberStereo.cpp:79
  const float maxanglediff = 5.0f;  // PMHuberStereo.cpp:80
  setup(left, right, patchsize, depthmax, initdisparity, initnormalmap);
  iterate(iterations | 1);  // PMHuberStereo.cpp:490
  disparity(dmap, pp, normalmap, maxdispdiff, maxanglediff);
}

void PMHuberStereo::depthMap(Image& dmap, const Image& left,
                             const Image& right, size_t patchsize,
                             float depthmax, size_t iterations, float dscale,
                             DisparityPostProcessing pp, Image* normalmap,
                             const Image* initdisparity,
                             const Image* initnormalmap) {
  if (dscale <= 0.0f) dscale = 1.0f / depthmax;  // PMHuberStereo.cpp:168-169
  const float maxdispdiff = 1.0f;   // PMHuberStereo.cpp:156
  const float maxanglediff = 5.0f;  // PMHuberStereo.cpp:157
  mode_ = PM_SMOOTH_DEPTHMAP;
  try {
    setup(left, right, patchsize, depthmax, initdisparity, initnormalmap);
    iterate(iterations);  // no `| 1` here (PMHuberStereo.cpp:214)
    disparity(dmap, pp, normalmap, maxdispdiff, maxanglediff);
  } catch (...) {
    mode_ = smoothing_;
    throw;
  }
  mode_ = smoothing_;
  for (size_t y = 0; y < dmap.height(); ++y)
    for (size_t x = 0; x < dmap.width(); ++x) dmap.at(x, y) *= dscale;
}

}  // namespace cvt_amd
