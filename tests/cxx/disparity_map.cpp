// disparity_map.cpp -- drives the C++ drop-in class the way a cvt caller
// drives cvt::PMHuberStereo::disparityMap (cvt/vision/PMHuberStereo.h:47-48).
//
// usage: disparity_map W H C radius dmax iterations pp left.f32 right.f32
//                      disp_out.f32 planes_out.f32 normals_out.f32
//                      [plain | depth]
// ("plain": PatchMatch without the PM-Huber smoothing coupling; "depth":
// depthMap with its default dscale instead of disparityMap)
// Images are raw row-major float32 files; outputs are W x H (disparity)
// and W x H x 4 (left plane field after the iterations, normal map).
#include <cstdio>
#include <cstdlib>
#include <string>

#include "cvt_amd/PMHuberStereo.h"

static bool read_raw(const char* path, cvt_amd::Image& im) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  const size_t n = im.stride() * im.height();
  const size_t got = std::fread(im.ptr(), sizeof(float), n, f);
  std::fclose(f);
  return got == n;
}

static bool write_raw(const char* path, const cvt_amd::Image& im) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  const size_t n = im.stride() * im.height();
  const size_t put = std::fwrite(im.ptr(), sizeof(float), n, f);
  std::fclose(f);
  return put == n;
}

int main(int argc, char** argv) {
  if (argc != 13 && argc != 14) {
    std::fprintf(stderr, "usage: %s W H C radius dmax iterations pp left "
                         "right disp planes normals\n", argv[0]);
    return 2;
  }
  const size_t W = std::strtoul(argv[1], nullptr, 10);
  const size_t H = std::strtoul(argv[2], nullptr, 10);
  const size_t C = std::strtoul(argv[3], nullptr, 10);
  const size_t radius = std::strtoul(argv[4], nullptr, 10);
  const float dmax = std::strtof(argv[5], nullptr);
  const size_t iters = std::strtoul(argv[6], nullptr, 10);
  const int pp = std::atoi(argv[7]);
  try {
    cvt_amd::Image left(W, H, C), right(W, H, C), dmap, normals, planes;
    if (!read_raw(argv[8], left) || !read_raw(argv[9], right)) {
      std::fprintf(stderr, "cannot read the input images\n");
      return 2;
    }
    const std::string mode = argc == 14 ? argv[13] : "";
    cvt_amd::PMHuberStereo pm(0, mode == "plain" ? PM_SMOOTH_NONE
                                                 : PM_SMOOTH_HUBER);
    const auto post = static_cast<cvt_amd::DisparityPostProcessing>(pp);
    if (mode == "depth")
      pm.depthMap(dmap, left, right, radius, dmax, iters, -1.0f, post,
                  &normals);
    else
      pm.disparityMap(dmap, left, right, radius, dmax, iters, post,
                      &normals);
    pm.planes(planes, 0);
    if (!write_raw(argv[10], dmap) || !write_raw(argv[11], planes) ||
        !write_raw(argv[12], normals)) {
      std::fprintf(stderr, "cannot write the outputs\n");
      return 2;
    }
    std::printf("iterations run: %zu\n", pm.iteration());
    // a mismatched pair must throw, as the reference does
    cvt_amd::Image small(W / 2, H, C);
    try {
      pm.disparityMap(dmap, left, small, radius, dmax, iters);
      std::fprintf(stderr, "mismatched pair did not throw\n");
      return 1;
    } catch (const cvt_amd::Exception& e) {
      std::printf("mismatch: %s\n", e.what());
    }
  } catch (const cvt_amd::Exception& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
