// pfm_fuzz.cpp -- the PFM reader of include/pm_io.h on malformed files.
// Built from the library's host source (cvt_amd/csrc/pm_io.cpp) together
// with this file under the host address and undefined-behaviour
// sanitizers; nothing here runs on a GPU (the one device launch the source
// references is stubbed and never reached).
//
// Seeds are valid gray and RGB files written by pm_pfm_write; each case
// mutates one (truncation, byte changes, a replaced header, inserted
// comment and blank lines) and reads it back through pm_pfm_info and
// pm_pfm_read.  A file the reader accepts must be read in full into a
// buffer of exactly the size pm_pfm_info reports; one it rejects must
// fail with PM_ERR_IO and a message.  The metrics entry points run on
// random maps with NaN, infinities and non-positive values.
//
// usage: pfm_fuzz <dir> <cases>; prints "cases N accepted A rejected R".
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pm_engine.h"
#include "pm_io.h"
#include "pm_stereo.h"

namespace pm {
hipError_t launch_eval_error(hipStream_t, const float*, int, const float*,
                             int, int, int, float, float,
                             unsigned long long*) {
  return hipErrorNotSupported;
}
}  // namespace pm

const char* hipGetErrorString(hipError_t) { return "not built"; }

namespace {

uint64_t g_rng = 0x9e3779b97f4a7c15ull;

uint64_t next() {  // xorshift64*
  g_rng ^= g_rng >> 12;
  g_rng ^= g_rng << 25;
  g_rng ^= g_rng >> 27;
  return g_rng * 0x2545f4914f6cdd1dull;
}

size_t below(size_t n) { return n ? (size_t)(next() % n) : 0; }

float any_float() {
  static const float special[] = {0.0f, -0.0f, 1.0f, -1.0f, 1e30f, -1e30f,
                                  1e-40f, __builtin_inff(), -__builtin_inff(),
                                  __builtin_nanf("")};
  if (below(4) == 0) return special[below(sizeof special / sizeof *special)];
  return ((float)(next() >> 40) / (float)(1u << 24)) * 200.0f - 50.0f;
}

std::vector<unsigned char> slurp(const std::string& path) {
  std::vector<unsigned char> b;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return b;
  unsigned char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0)
    b.insert(b.end(), buf, buf + n);
  std::fclose(f);
  return b;
}

bool spit(const std::string& path, const std::vector<unsigned char>& b) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok =
      b.empty() || std::fwrite(b.data(), 1, b.size(), f) == b.size();
  return std::fclose(f) == 0 && ok;
}

// offset of the first sample: after the third line that holds a token
// not starting with '#'
size_t header_end(const std::vector<unsigned char>& b) {
  size_t pos = 0;
  int lines = 0;
  while (pos < b.size() && lines < 3) {
    size_t end = pos;
    while (end < b.size() && b[end] != '\n') ++end;
    size_t i = pos;
    while (i < end && (b[i] == ' ' || b[i] == '\t')) ++i;
    if (i < end && b[i] != '#') ++lines;
    pos = end < b.size() ? end + 1 : end;
  }
  return pos;
}

const char* const kHeaders[] = {
    "PF\n2147380029 715862424\n-1.0\n",  // W*H*12 wraps to 11936 bytes
    "Pf\n2147483647 2147483647\n-1\n",
    "Pf\n-3 4\n-1.0\n",
    "Pf\n0 0\n-1.0\n",
    "Pf\n4294967299 1\n-1\n",  // low 32 bits: 3
    "Pf\n0x3 02\n-1.0\n",
    "PF\n99999999999999999999 2\n-1\n",
    "Pf\n 2\t2 \n# comment\n\t-1.0\n",
    "Pf\n2\n-1\n",
    "Pf\n2 2\n",
    "Pf\n2 2",
    "PF",
    "Pf\n# only comments\n#\n#\n",
    "#Pf\nPf\n1 1\n-1\n",
    "P\n1 1\n-1\n",
    "Pf\r\n1 1\r\n-1\r\n",
    "Pf\n1 1\nnan\n",
    "Pf\n1 1\n1e40\n",
    "\n\n\n",
    "",
};

std::vector<unsigned char> mutate(const std::vector<unsigned char>& seed) {
  std::vector<unsigned char> b = seed;
  const size_t hdr = header_end(seed);
  switch (below(5)) {
    case 0:  // truncate
      b.resize(below(b.size() + 1));
      break;
    case 1: {  // change bytes, mostly in the header
      const size_t k = 1 + below(4);
      for (size_t i = 0; i < k && !b.empty(); ++i) {
        const size_t at = below(2) ? below(hdr + 1) : below(b.size());
        if (at < b.size()) b[at] = (unsigned char)next();
      }
      break;
    }
    case 2: {  // another header before the seed's samples
      const char* h = kHeaders[below(sizeof kHeaders / sizeof *kHeaders)];
      std::vector<unsigned char> out(h, h + std::strlen(h));
      const size_t keep = below(b.size() - hdr + 1);
      out.insert(out.end(), b.begin() + hdr, b.begin() + hdr + keep);
      b.swap(out);
      break;
    }
    case 3: {  // comment, blank or whitespace-only lines in the header
      static const char* const ins[] = {"# c\n", "\n", " \t \n", "#\n",
                                        "#x y z\n"};
      const char* s = ins[below(sizeof ins / sizeof *ins)];
      const size_t at = below(hdr + 1);
      b.insert(b.begin() + at, s, s + std::strlen(s));
      break;
    }
    default: {  // drop a byte range
      if (b.empty()) break;
      const size_t at = below(b.size());
      const size_t n = 1 + below(std::min<size_t>(16, b.size() - at));
      b.erase(b.begin() + at, b.begin() + at + n);
      break;
    }
  }
  return b;
}

int fail(const char* what, size_t i) {
  std::printf("case %zu: %s (%s)\n", i, what, pm_io_last_error());
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  const std::string dir = argv[1];
  const size_t cases = std::strtoul(argv[2], nullptr, 10);

  std::vector<std::vector<unsigned char>> seeds;
  const int shapes[][3] = {{7, 5, 1}, {6, 3, 4}, {1, 1, 1}, {2, 9, 4}};
  for (const auto& s : shapes) {
    const int W = s[0], H = s[1], C = s[2];
    std::vector<float> img((size_t)W * H * C);
    for (float& v : img) v = any_float();
    const std::string p = dir + "/seed.pfm";
    if (pm_pfm_write(p.c_str(), img.data(), W, H, C, W * C) != PM_OK)
      return fail("seed write", 0);
    seeds.push_back(slurp(p));
  }
  for (size_t h = 0; h < sizeof kHeaders / sizeof *kHeaders; ++h) {
    std::vector<unsigned char> b(kHeaders[h], kHeaders[h] +
                                 std::strlen(kHeaders[h]));
    b.resize(b.size() + 64, 0x3f);
    seeds.push_back(b);
  }
  {  // the wrapping size with the 11936 bytes its wrapped count asks for
    std::vector<unsigned char> b(kHeaders[0],
                                 kHeaders[0] + std::strlen(kHeaders[0]));
    b.resize(b.size() + 11936, 0);
    seeds.push_back(b);
  }

  size_t accepted = 0, rejected = 0;
  const std::string path = dir + "/case.pfm";
  for (size_t i = 0; i < cases; ++i) {
    const std::vector<unsigned char>& seed = seeds[below(seeds.size())];
    const std::vector<unsigned char> b = below(8) ? mutate(seed) : seed;
    if (!spit(path, b)) return fail("case write", i);
    int W = -1, H = -1, C = -1;
    const int rc = pm_pfm_info(path.c_str(), &W, &H, &C);
    if (rc == PM_OK) {
      if (W < 1 || H < 1 || (C != 1 && C != 4)) return fail("bad info", i);
      const size_t n = (size_t)W * H * C;
      if (n > b.size()) return fail("size beyond the file", i);
      std::vector<float> out(n);
      if (pm_pfm_read(path.c_str(), out.data(), W, H, C, W * C) != PM_OK)
        return fail("accepted by info, refused by read", i);
      if (W > 1 && pm_pfm_read(path.c_str(), out.data(), W - 1, H, C,
                               W * C) != PM_ERR_ARG)
        return fail("wrong size read", i);
      ++accepted;
    } else {
      if (rc != 5 || pm_io_last_error()[0] == '\0')
        return fail("rejected without PM_ERR_IO and a message", i);
      float one[4];
      if (pm_pfm_read(path.c_str(), one, 1, 1, 1, 1) == PM_OK)
        return fail("refused by info, accepted by read", i);
      ++rejected;
    }
  }

  // the metrics on maps with special values and uneven strides
  for (size_t i = 0; i < cases / 8; ++i) {
    const int W = (int)below(9), H = (int)below(7);
    const int sd = W + (int)below(3), sg = W + (int)below(3);
    std::vector<float> d((size_t)sd * H + 1), g((size_t)sg * H + 1);
    for (float& v : d) v = any_float();
    for (float& v : g) v = any_float();
    uint64_t e = 0, v = 0, a = 0;
    if (pm_eval_error(d.data(), sd, g.data(), sg, W, H, 1.0f, 0.0f, &e, &v,
                      &a) != PM_OK || e > a || v > a ||
        a > (uint64_t)W * H)
      return fail("evaluate error", i);
    float bins[16];
    int nb = 0;
    if (pm_default_histogram((int)(i & 1), bins, &nb) != PM_OK || nb > 16)
      return fail("default histogram", i);
    float hist[16];
    if (pm_error_histogram(g.data(), sg, d.data(), sd, W, H, (int)(i & 1),
                           bins, nb, hist) != PM_OK)
      return fail("histogram", i);
  }
  std::printf("cases %zu accepted %zu rejected %zu\n", cases, accepted,
              rejected);
  return 0;
}
