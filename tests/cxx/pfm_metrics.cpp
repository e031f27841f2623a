// pfm_metrics.cpp -- the PFM and Disparity mirrors used from C++ the way a
// cvt caller uses cvt::Disparity (cvt/vision/Disparity.h:30-56): write a
// ground truth and an estimate as PFM, read them back, evaluate.
//
// usage: pfm_metrics <dir>; prints "error valid all" and the absolute
// histogram.
#include <cstdio>
#include <string>

#include "cvt_amd/Disparity.h"

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  const std::string dir = argv[1];
  try {
    cvt_amd::Image gt(6, 4, 1), est(6, 4, 1);
    for (size_t y = 0; y < 4; ++y)
      for (size_t x = 0; x < 6; ++x) {
        gt.at(x, y) = (x == 0) ? 0.0f : 10.0f + (float)x;
        est.at(x, y) = gt.at(x, y) + (float)(y) * 0.75f;
      }
    cvt_amd::savePFM((dir + "/gt.pfm").c_str(), gt);
    cvt_amd::savePFM((dir + "/est.pfm").c_str(), est);
    cvt_amd::Image g2, e2;
    cvt_amd::loadPFM(g2, (dir + "/gt.pfm").c_str());
    cvt_amd::loadPFM(e2, (dir + "/est.pfm").c_str());
    size_t error, valid, all;
    cvt_amd::Disparity::evaluateError(error, valid, all, e2, g2, 1.0f, 0.0f);
    std::printf("%zu %zu %zu\n", error, valid, all);
    std::vector<std::pair<float, float>> h;
    cvt_amd::Disparity::defaultHistogramAbsolute(h);
    cvt_amd::Disparity::histogramErrorAbsolute(g2, e2, h);
    for (const auto& b : h) std::printf("%g:%g ", b.first, b.second);
    std::printf("\n");
    try {
      cvt_amd::loadPFM(g2, (dir + "/absent.pfm").c_str());
      return 1;
    } catch (const cvt_amd::Exception& e) {
      std::printf("missing: %d\n", e.code());
    }
  } catch (const cvt_amd::Exception& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
