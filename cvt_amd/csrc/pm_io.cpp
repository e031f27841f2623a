// pm_io.cpp -- PFM disparity-map I/O and cvt's disparity error metrics
// (include/pm_io.h).  Host code: no GPU involved.
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>

#include "pm_engine.h"
#include "pm_io.h"

namespace {

thread_local std::string g_err;

int io_fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace

namespace pm {
// the message pm_io_last_error returns on this thread, for the library's
// other context-free entry points
int io_set_error(int code, const char* msg) {
  g_err = msg;
  return code;
}
}  // namespace pm

namespace {

bool host_little_endian() {
  const uint32_t one = 1;
  unsigned char b;
  std::memcpy(&b, &one, 1);
  return b == 1;
}

uint32_t bswap32(uint32_t v) {
  return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

// A PFM file read into memory, with the header parsed.
struct PfmFile {
  std::vector<unsigned char> bytes;
  size_t data = 0;       // offset of the first sample
  int W = 0, H = 0;
  int file_channels = 0;  // 1 ("Pf") or 3 ("PF")
  float scale = 0.0f;
};

// Next header line with at least one token whose first token does not
// start with '#' (PFMLoader::nextDataLine, PFMLoader.cpp:58-66); tokens
// are separated by spaces and tabs.
bool next_data_line(const PfmFile& f, size_t* pos,
                    std::vector<std::string>* tokens) {
  const std::vector<unsigned char>& b = f.bytes;
  while (*pos < b.size()) {
    size_t end = *pos;
    while (end < b.size() && b[end] != '\n') ++end;
    tokens->clear();
    std::string tok;
    for (size_t i = *pos; i < end; ++i) {
      const char ch = (char)b[i];
      if (ch == ' ' || ch == '\t') {
        if (!tok.empty()) tokens->push_back(tok);
        tok.clear();
      } else {
        tok.push_back(ch);
      }
    }
    if (!tok.empty()) tokens->push_back(tok);
    *pos = end < b.size() ? end + 1 : end;
    if (!tokens->empty() && (*tokens)[0][0] != '#') return true;
  }
  tokens->clear();
  return false;
}

// A header size as String::toInteger reads it (cvt/util/String.h:510-513:
// strtol with base 0, so 0x.. is hexadecimal and 0.. octal) and the
// header's int field keeps it (PFMLoader.cpp:38-39,94-95): the low 32
// bits as two's complement.
int header_int(const std::string& tok) {
  const long v = std::strtol(tok.c_str(), nullptr, 0);
  const uint32_t low = (uint32_t)(unsigned long)v;
  int out;
  std::memcpy(&out, &low, sizeof out);
  return out;
}

// PFMLoader::parseHeader (PFMLoader.cpp:68-105)
int parse_pfm(const char* path, PfmFile* f) {
  if (!path) return io_fail(PM_ERR_ARG, "path is NULL");
  std::ifstream in(path, std::ios::binary);
  if (!in) return io_fail(PM_ERR_IO, "cannot open %s", path);
  f->bytes.assign(std::istreambuf_iterator<char>(in),
                  std::istreambuf_iterator<char>());
  size_t pos = 0;
  std::vector<std::string> t;
  if (!next_data_line(*f, &pos, &t))
    return io_fail(PM_ERR_IO, "%s: PFM header corrupt", path);
  if (t[0] == "Pf")
    f->file_channels = 1;
  else if (t[0] == "PF")
    f->file_channels = 3;
  else
    return io_fail(PM_ERR_IO, "%s: unknown PFM type %s", path, t[0].c_str());
  if (!next_data_line(*f, &pos, &t) || t.size() < 2)
    return io_fail(PM_ERR_IO, "%s: PFM header corrupt, expected <width> "
                   "<height>", path);
  f->W = header_int(t[0]);
  f->H = header_int(t[1]);
  if (!next_data_line(*f, &pos, &t))
    return io_fail(PM_ERR_IO, "%s: PFM header corrupt, expected <scale>",
                   path);
  f->scale = std::strtof(t[0].c_str(), nullptr);
  if (f->W < 1 || f->H < 1)
    return io_fail(PM_ERR_IO, "%s: invalid PFM size %dx%d", path, f->W, f->H);
  f->data = pos;
  // whole pixels present, compared row by row: W * H * channels * 4 can
  // exceed size_t for header sizes up to INT_MAX and wrap to a small count
  const size_t have = f->bytes.size() - f->data;
  const size_t pixels = have / (sizeof(float) * f->file_channels);
  if (pixels / (size_t)f->W < (size_t)f->H)
    return io_fail(PM_ERR_IO, "%s: not enough PFM data (%zu bytes for %dx%d "
                   "with %d channels)", path, have, f->W, f->H,
                   f->file_channels);
  return PM_OK;
}

}  // namespace

extern "C" {

const char* pm_io_last_error(void) { return g_err.c_str(); }

int pm_pfm_info(const char* path, int* W, int* H, int* channels) try {
  if (!W || !H || !channels) return io_fail(PM_ERR_ARG, "output is NULL");
  PfmFile f;
  int rc = parse_pfm(path, &f);
  if (rc) return rc;
  *W = f.W;
  *H = f.H;
  *channels = f.file_channels == 1 ? 1 : 4;
  return PM_OK;
} catch (const std::bad_alloc&) {
  return io_fail(PM_ERR_NOMEM, "out of host memory");
}

// PFMLoader::load, parseBinary, parseBinarySwap (PFMLoader.cpp:107-205)
int pm_pfm_read(const char* path, float* out, int W, int H, int channels,
                int stride) try {
  if (!out) return io_fail(PM_ERR_ARG, "out is NULL");
  PfmFile f;
  int rc = parse_pfm(path, &f);
  if (rc) return rc;
  const int C = f.file_channels == 1 ? 1 : 4;
  if (W != f.W || H != f.H || channels != C)
    return io_fail(PM_ERR_ARG, "%s holds %dx%d with %d channels, not %dx%d "
                   "with %d", path, f.W, f.H, C, W, H, channels);
  if ((long long)stride < (long long)W * C)
    return io_fail(PM_ERR_ARG, "stride %d < width*channels", stride);
  // little-endian samples when the scale is negative on a little-endian
  // host, byte-swapped samples otherwise (PFMLoader.cpp:122-125)
  const bool swap = !(host_little_endian() && f.scale < 0.0f);
  const unsigned char* src = f.bytes.data() + f.data;
  const size_t n = (size_t)W * f.file_channels;
  for (int row = 0; row < H; ++row) {  // file rows run bottom to top
    float* dst = out + (size_t)(H - 1 - row) * stride;
    for (int x = 0; x < W; ++x) {
      for (int c = 0; c < f.file_channels; ++c) {
        uint32_t v;
        std::memcpy(&v, src + ((size_t)row * n + (size_t)x * f.file_channels +
                               c) * 4, 4);
        if (swap) v = bswap32(v);
        std::memcpy(dst + (size_t)x * C + c, &v, 4);
      }
      if (C == 4) dst[(size_t)x * 4 + 3] = 1.0f;
    }
  }
  return PM_OK;
} catch (const std::bad_alloc&) {
  return io_fail(PM_ERR_NOMEM, "out of host memory");
}

// PFMSaver::save (plugins/isaver/PFM/PFMSaver.cpp:117-181)
int pm_pfm_write(const char* path, const float* data, int W, int H,
                 int channels, int stride) try {
  if (!path || !data) return io_fail(PM_ERR_ARG, "path or data is NULL");
  if (W < 1 || H < 1) return io_fail(PM_ERR_ARG, "invalid size %dx%d", W, H);
  if (channels != 1 && channels != 4)
    return io_fail(PM_ERR_ARG, "channels must be 1 (GRAY) or 4 (RGBA)");
  if ((long long)stride < (long long)W * channels)
    return io_fail(PM_ERR_ARG, "stride %d < width*channels", stride);
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  if (!out) return io_fail(PM_ERR_IO, "cannot open %s for writing", path);
  out << (channels == 4 ? "PF" : "Pf") << "\n" << W << " " << H << "\n"
      << (host_little_endian() ? "-" : "") << "1.0" << "\n";
  const int fc = channels == 4 ? 3 : 1;
  std::vector<float> line((size_t)W * fc);
  for (int row = H - 1; row >= 0; --row) {
    const float* src = data + (size_t)row * stride;
    for (int x = 0; x < W; ++x)
      for (int c = 0; c < fc; ++c) line[(size_t)x * fc + c] =
          src[(size_t)x * channels + c];
    out.write(reinterpret_cast<const char*>(line.data()),
              (std::streamsize)(line.size() * sizeof(float)));
  }
  out.close();
  if (!out) return io_fail(PM_ERR_IO, "write to %s failed", path);
  return PM_OK;
} catch (const std::bad_alloc&) {
  return io_fail(PM_ERR_NOMEM, "out of host memory");
}

// Disparity::evaluateError (cvt/vision/Disparity.cpp:201-233)
int pm_eval_error(const float* disparity, int stride_d, const float* gt,
                  int stride_gt, int W, int H, float abs_thres,
                  float rel_thres, uint64_t* error, uint64_t* valid,
                  uint64_t* all) {
  if (!disparity || !gt || !error || !valid || !all)
    return io_fail(PM_ERR_ARG, "NULL argument");
  if (W < 0 || H < 0 || stride_d < W || stride_gt < W)
    return io_fail(PM_ERR_ARG, "invalid size or stride");
  uint64_t e = 0, v = 0, a = 0;
  for (int y = 0; y < H; ++y) {
    const float* pg = gt + (size_t)y * stride_gt;
    const float* pd = disparity + (size_t)y * stride_d;
    for (int x = 0; x < W; ++x) {
      if (pg[x] > 0.0f) {
        const float d_gt = pg[x], d_est = pd[x];
        const bool err = std::fabs(d_gt - d_est) > abs_thres &&
                         (std::fabs(d_gt - d_est) / std::fabs(d_gt)) >
                             rel_thres;
        if (d_est > 0.0f) ++v;
        if (err) ++e;
        ++a;
      }
    }
  }
  *error = e;
  *valid = v;
  *all = a;
  return PM_OK;
}

int pm_eval_error_device(const float* d_disparity, int stride_d,
                         const float* d_gt, int stride_gt, int W, int H,
                         float abs_thres, float rel_thres, uint64_t* d_counts,
                         void* stream) {
  if (!d_disparity || !d_gt || !d_counts)
    return io_fail(PM_ERR_ARG, "NULL argument");
  if (W < 1 || H < 1 || stride_d < W || stride_gt < W)
    return io_fail(PM_ERR_ARG, "invalid size or stride");
  const hipError_t e = pm::launch_eval_error(
      static_cast<hipStream_t>(stream), d_disparity, stride_d, d_gt,
      stride_gt, W, H, abs_thres, rel_thres,
      reinterpret_cast<unsigned long long*>(d_counts));
  if (e != hipSuccess)
    return io_fail(PM_ERR_HIP, "evaluate error: %s", hipGetErrorString(e));
  return PM_OK;
}

// Disparity::histogramErrorAbsolute / histogramErrorRelative
// (Disparity.cpp:236-318)
int pm_error_histogram(const float* gt, int stride_gt, const float* test,
                       int stride_test, int W, int H, int relative,
                       const float* bins, int nbins, float* out) {
  if (!gt || !test || !out || (nbins > 0 && !bins))
    return io_fail(PM_ERR_ARG, "NULL argument");
  if (W < 0 || H < 0 || nbins < 0 || stride_gt < W || stride_test < W)
    return io_fail(PM_ERR_ARG, "invalid size, stride or bin count");
  for (int i = 0; i < nbins; ++i) out[i] = 0.0f;
  size_t n = 0;
  for (int y = 0; y < H; ++y) {
    const float* pg = gt + (size_t)y * stride_gt;
    const float* pt = test + (size_t)y * stride_test;
    for (int x = 0; x < W; ++x) {
      if (std::isfinite(pg[x]) && pg[x] > 0.0f && std::isfinite(pt[x]) &&
          pt[x] > 0.0f) {
        ++n;
        float diff = std::fabs(pt[x] - pg[x]);
        if (relative) diff = diff / std::fabs(pg[x]);
        float last = 0.0f;
        for (int i = 0; i < nbins; ++i) {
          if (diff <= bins[i] && diff > last) out[i] += 1.0f;
          last = bins[i];
        }
      }
    }
  }
  for (int i = 0; i < nbins; ++i) out[i] /= (float)n;
  return PM_OK;
}

// Disparity::defaultHistogramAbsolute / defaultHistogramRelative
// (Disparity.cpp:320-355)
int pm_default_histogram(int relative, float* bins, int* nbins) {
  static const float kAbs[] = {0.50f, 0.75f, 1.00f, 2.00f,  3.00f,
                               4.00f, 5.00f, 6.00f, 7.00f,  8.00f,
                               9.00f, 10.0f, 20.0f, 40.0f,  1e10f};
  static const float kRel[] = {0.005f, 0.01f, 0.02f, 0.03f, 0.04f,
                               0.05f,  0.10f, 0.20f, 0.40f, 1e10f};
  if (!nbins) return io_fail(PM_ERR_ARG, "nbins is NULL");
  const float* src = relative ? kRel : kAbs;
  const int n = relative ? (int)(sizeof kRel / sizeof kRel[0])
                         : (int)(sizeof kAbs / sizeof kAbs[0]);
  *nbins = n;
  if (bins) std::memcpy(bins, src, n * sizeof(float));
  return PM_OK;
}

}  // extern "C"
