// Disparity.h -- cvt::Disparity's error metrics (cvt/vision/Disparity.h:
// 30-56) for the engine's disparity maps, over the C ABI in pm_io.h.
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "cvt_amd/PMHuberStereo.h"
#include "pm_io.h"

namespace cvt_amd {

class Disparity {
 public:
  // Disparity.cpp:201-233
  static void evaluateError(size_t& error, size_t& valid, size_t& all,
                            const Image& disparity, const Image& dispgt,
                            float abs_thres, float rel_thres) {
    check_gray(disparity, dispgt);
    uint64_t e = 0, v = 0, a = 0;
    check(pm_eval_error(disparity.ptr(), (int)disparity.stride(),
                        dispgt.ptr(), (int)dispgt.stride(),
                        (int)dispgt.width(), (int)dispgt.height(), abs_thres,
                        rel_thres, &e, &v, &a));
    error = (size_t)e;
    valid = (size_t)v;
    all = (size_t)a;
  }

  // Disparity.h:49-56
  static void evaluateError(float& error, float& density,
                            const Image& disparity, const Image& dispgt,
                            float abs_thres, float rel_thres) {
    size_t wrong, valid, all;
    evaluateError(wrong, valid, all, disparity, dispgt, abs_thres, rel_thres);
    error = (float)wrong / (float)all;
    density = (float)valid / (float)all;
  }

  // Disparity.cpp:236-318: histogram holds (bin edge, value) pairs; the
  // values are replaced by the fraction of pixels in each bin
  static void histogramErrorAbsolute(
      const Image& gt, const Image& test,
      std::vector<std::pair<float, float>>& histogram) {
    histogram_error(gt, test, histogram, 0);
  }
  static void histogramErrorRelative(
      const Image& gt, const Image& test,
      std::vector<std::pair<float, float>>& histogram) {
    histogram_error(gt, test, histogram, 1);
  }

  // Disparity.cpp:320-355
  static void defaultHistogramAbsolute(
      std::vector<std::pair<float, float>>& histogram) {
    default_histogram(histogram, 0);
  }
  static void defaultHistogramRelative(
      std::vector<std::pair<float, float>>& histogram) {
    default_histogram(histogram, 1);
  }

 private:
  static void check(int rc) {
    if (rc != PM_OK) throw Exception(rc, pm_io_last_error());
  }
  static void check_gray(const Image& a, const Image& b) {
    if (a.channels() != 1 || b.channels() != 1 || a.width() != b.width() ||
        a.height() != b.height())
      throw Exception(PM_ERR_ARG, "Invalid formats provided!");
  }
  static void histogram_error(const Image& gt, const Image& test,
                              std::vector<std::pair<float, float>>& h,
                              int relative) {
    check_gray(gt, test);
    std::vector<float> bins(h.size()), out(h.size());
    for (size_t i = 0; i < h.size(); ++i) bins[i] = h[i].first;
    check(pm_error_histogram(gt.ptr(), (int)gt.stride(), test.ptr(),
                             (int)test.stride(), (int)gt.width(),
                             (int)gt.height(), relative, bins.data(),
                             (int)bins.size(), out.data()));
    for (size_t i = 0; i < h.size(); ++i) h[i].second = out[i];
  }
  static void default_histogram(std::vector<std::pair<float, float>>& h,
                                int relative) {
    int n = 0;
    check(pm_default_histogram(relative, nullptr, &n));
    std::vector<float> bins((size_t)n);
    check(pm_default_histogram(relative, bins.data(), &n));
    h.clear();
    for (float b : bins) h.emplace_back(b, 0.0f);
  }
};

// PFM disparity maps (plugins/iloader/PFM/PFMLoader.cpp,
// plugins/isaver/PFM/PFMSaver.cpp): GRAY_FLOAT for "Pf", RGBA_FLOAT for
// "PF"
inline void loadPFM(Image& img, const char* path) {
  int W = 0, H = 0, C = 0;
  int rc = pm_pfm_info(path, &W, &H, &C);
  if (rc != PM_OK) throw Exception(rc, pm_io_last_error());
  img.reallocate((size_t)W, (size_t)H, (size_t)C);
  rc = pm_pfm_read(path, img.ptr(), W, H, C, (int)img.stride());
  if (rc != PM_OK) throw Exception(rc, pm_io_last_error());
}

inline void savePFM(const char* path, const Image& img) {
  int rc = pm_pfm_write(path, img.ptr(), (int)img.width(), (int)img.height(),
                        (int)img.channels(), (int)img.stride());
  if (rc != PM_OK) throw Exception(rc, pm_io_last_error());
}

}  // namespace cvt_amd
