// Drives the C++ FasTC-style API (include/fasthip/fastc.hpp) the way a FasTC
// caller would.  Usage:
//   facade_test validate                 -> checks error behaviour (no GPU needed)
//   facade_test encode in.raw w h outdir -> BPTC(q=0), DXT1, DXT5, ETC1 via the
//                                           codec entry points and CompressImageData
//   facade_test select in.raw w h outdir -> BPTCC::Compress (q=0) with a shape
//                                           selection function
//   facade_test stats in.raw w h outdir  -> BPTCC::CompressWithStats (q=0) on a
//                                           sub-job and CompressImageData with a
//                                           log stream: blocks and log text
//   facade_test image in.raw w h outdir  -> CompressImage of the top-left
//                                           (w-2) x (h-2) pixels (zero padding),
//                                           its blocks and DecompressImage
//   facade_test atomic in.raw w h outdir -> BPTCC::CompressAtomic from two
//                                           threads over three jobs
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "fasthip/fastc.hpp"

using namespace fasthip;

// The reference Image's interface as CompressImage reads it: a w x h image
// of RGBA8 words.
struct TestPixel {
  uint32_t v;
  uint32_t Pack() const { return v; }
};
struct TestImage {
  uint32_t w, h;
  std::vector<uint32_t> px;  // stride `stride`
  uint32_t stride;
  bool computed = false;
  uint32_t GetWidth() const { return w; }
  uint32_t GetHeight() const { return h; }
  void ComputePixels() { computed = true; }
  TestPixel operator()(uint32_t x, uint32_t y) const { return TestPixel{px[(size_t)y * stride + x]}; }
};

static int validate() {
  std::vector<uint8_t> img(8 * 8 * 4, 0), out(8 * 8, 0);
  SCompressionSettings s;
  int fails = 0;
  // width not a multiple of the block size (TexComp.cpp:474-476)
  if (CompressImageData(img.data(), 6, 8, out.data(), (unsigned)out.size(), s)) fails++;
  // output buffer too small (TexComp.cpp:493-495)
  if (CompressImageData(img.data(), 8, 8, out.data(), 10, s)) fails++;
  // no data (TexComp.cpp:454-457)
  if (CompressImageData(nullptr, 8, 8, out.data(), (unsigned)out.size(), s)) fails++;
  SCompressionSettings cpu = s;
  cpu.bUseGPU = false;
  if (CompressImageData(img.data(), 8, 8, out.data(), (unsigned)out.size(), cpu)) fails++;
  if (FasTC::GetCompressedSize(8, 8, FasTC::eCompressionFormat_DXT1) != 32) fails++;
  // CompressImage of no image; a job list that is full
  if (CompressImage(static_cast<TestImage *>(nullptr), s) != nullptr) fails++;
  FasTC::CompressionJobList list(1);
  FasTC::CompressionJob j0(FasTC::eCompressionFormat_BPTC, img.data(), out.data(), 8, 8);
  if (!list.AddJob(j0) || list.AddJob(j0) || list.GetNumJobs() != 1) fails++;
  if (list.GetJob(1) != nullptr || list.GetFinishedFlag(1) != nullptr) fails++;
  if (*list.GetFinishedFlag(0) != 0 || list.GetTotalNumJobs() != 1) fails++;
  if (!fh_device_available()) {
    // a valid job must report the missing device, never fall back to CPU
    FasTC::CompressionJob cj(FasTC::eCompressionFormat_BPTC, img.data(), out.data(), 8, 8);
    bool threw = false;
    try {
      BPTCC::Compress(cj);
    } catch (const Error &e) {
      threw = e.status == FH_ERR_NO_DEVICE;
    }
    if (!threw) fails++;
  }
  printf("validate: %d failures\n", fails);
  return fails ? 1 : 0;
}

static void save(const std::string &path, const std::vector<uint8_t> &v) {
  FILE *f = fopen(path.c_str(), "wb");
  fwrite(v.data(), 1, v.size(), f);
  fclose(f);
}

static int encode(const char *in, unsigned w, unsigned h, const std::string &dir) {
  std::vector<uint8_t> img((size_t)w * h * 4);
  FILE *f = fopen(in, "rb");
  if (!f || fread(img.data(), 1, img.size(), f) != img.size()) return 2;
  fclose(f);
  const size_t n16 = (size_t)w * h, n8 = n16 / 2;
  std::vector<uint8_t> bptc(n16), d1(n8), d5(n16), e1(n8), via(n16);
  BPTCC::CompressionSettings bs;
  bs.m_NumSimulatedAnnealingSteps = 0;
  BPTCC::Compress(FasTC::CompressionJob(FasTC::eCompressionFormat_BPTC, img.data(), bptc.data(), w, h), bs);
  DXTC::CompressImageDXT1(FasTC::CompressionJob(FasTC::eCompressionFormat_DXT1, img.data(), d1.data(), w, h));
  DXTC::CompressImageDXT5(FasTC::CompressionJob(FasTC::eCompressionFormat_DXT5, img.data(), d5.data(), w, h));
  ETCC::Compress_RG(FasTC::CompressionJob(FasTC::eCompressionFormat_ETC1, img.data(), e1.data(), w, h));
  SCompressionSettings s;
  s.format = FasTC::eCompressionFormat_BPTC;
  s.iQuality = 0;
  if (!CompressImageData(img.data(), w, h, via.data(), (unsigned)via.size(), s)) return 3;
  save(dir + "/bptc.bin", bptc);
  save(dir + "/dxt1.bin", d1);
  save(dir + "/dxt5.bin", d5);
  save(dir + "/etc1.bin", e1);
  save(dir + "/bptc_via_cid.bin", via);
  printf("encode ok\n");
  return 0;
}

// A selection function: the two-subset shape (x / 4 + y / 4) % 64, all modes.
static BPTCC::ShapeSelection diagonal(uint32_t x, uint32_t y, const uint32_t *, const void *ud) {
  BPTCC::ShapeSelection s;
  s.m_NumShapesToSearch = 1;
  s.m_Shapes[0].m_NumPartitions = 2;
  s.m_Shapes[0].m_Index = (x / 4 + y / 4 + *static_cast<const uint32_t *>(ud)) % 64;
  s.m_SelectedModes = 0xFF;
  return s;
}

static int select(const char *in, unsigned w, unsigned h, const std::string &dir) {
  std::vector<uint8_t> img((size_t)w * h * 4);
  FILE *f = fopen(in, "rb");
  if (!f || fread(img.data(), 1, img.size(), f) != img.size()) return 2;
  fclose(f);
  std::vector<uint8_t> out((size_t)w * h, 0);
  const uint32_t offset = 5;
  BPTCC::CompressionSettings bs;
  bs.m_NumSimulatedAnnealingSteps = 0;
  bs.m_ShapeSelectionFn = diagonal;
  bs.m_ShapeSelectionUserData = &offset;
  BPTCC::Compress(FasTC::CompressionJob(FasTC::eCompressionFormat_BPTC, img.data(), out.data(),
                                        w, h), bs);
  save(dir + "/select.bin", out);
  printf("select ok\n");
  return 0;
}

static int stats(const char *in, unsigned w, unsigned h, const std::string &dir) {
  std::vector<uint8_t> img((size_t)w * h * 4);
  FILE *f = fopen(in, "rb");
  if (!f || fread(img.data(), 1, img.size(), f) != img.size()) return 2;
  fclose(f);
  std::vector<uint8_t> sub((size_t)w * h, 0), whole((size_t)w * h, 0);
  BPTCC::CompressionSettings bs;
  bs.m_NumSimulatedAnnealingSteps = 0;
  std::ostringstream log_sub, log_whole;
  BPTCC::CompressWithStats(FasTC::CompressionJob(FasTC::eCompressionFormat_BPTC, img.data(),
                                                 sub.data(), w, h, 8, 4, 40, 48),
                           &log_sub, bs);
  SCompressionSettings s;
  s.format = FasTC::eCompressionFormat_BPTC;
  s.iQuality = 0;
  s.logStream = &log_whole;
  if (!CompressImageData(img.data(), w, h, whole.data(), (unsigned)whole.size(), s)) return 3;
  save(dir + "/stats_sub.bin", sub);
  save(dir + "/stats_whole.bin", whole);
  const std::string a = log_sub.str(), b = log_whole.str();
  save(dir + "/stats_sub.log", std::vector<uint8_t>(a.begin(), a.end()));
  save(dir + "/stats_whole.log", std::vector<uint8_t>(b.begin(), b.end()));
  printf("stats ok\n");
  return 0;
}

static bool load(const char *in, unsigned w, unsigned h, std::vector<uint8_t> &img) {
  img.resize((size_t)w * h * 4);
  FILE *f = fopen(in, "rb");
  if (!f) return false;
  const bool ok = fread(img.data(), 1, img.size(), f) == img.size();
  fclose(f);
  return ok;
}

static int image(const char *in, unsigned w, unsigned h, const std::string &dir) {
  std::vector<uint8_t> raw;
  if (!load(in, w, h, raw)) return 2;
  TestImage ti;
  ti.w = w - 2;
  ti.h = h - 2;
  ti.stride = w;
  ti.px.resize((size_t)w * h);
  memcpy(ti.px.data(), raw.data(), raw.size());
  SCompressionSettings s;
  s.format = FasTC::eCompressionFormat_BPTC;
  s.iQuality = 0;
  CompressedImage *ci = CompressImage(&ti, s);
  if (!ci || !ti.computed || ci->GetWidth() != w || ci->GetHeight() != h) return 3;
  std::vector<uint8_t> dec(ci->GetUncompressedSize());
  if (ci->DecompressImage(dec.data(), 16)) return 4;  // too small
  if (!ci->DecompressImage(dec.data(), (uint32_t)dec.size())) return 5;
  save(dir + "/image.bin",
       std::vector<uint8_t>(ci->GetCompressedData(), ci->GetCompressedData() + ci->GetCompressedSize()));
  save(dir + "/image_dec.bin", dec);
  delete ci;
  printf("image ok\n");
  return 0;
}

static int atomic(const char *in, unsigned w, unsigned h, const std::string &dir) {
  std::vector<uint8_t> img;
  if (!load(in, w, h, img)) return 2;
  // three jobs: the image, its top half, its left-half columns as a narrow image
  const unsigned h2 = h / 2, w2 = w / 2;
  std::vector<uint8_t> top(img.begin(), img.begin() + (size_t)w * h2 * 4), left((size_t)w2 * h * 4);
  for (unsigned y = 0; y < h; y++) memcpy(&left[(size_t)y * w2 * 4], &img[(size_t)y * w * 4], w2 * 4);
  std::vector<uint8_t> o0((size_t)w * h), o1((size_t)w * h2), o2((size_t)w2 * h);
  FasTC::CompressionJobList list(3);
  list.AddJob(FasTC::CompressionJob(FasTC::eCompressionFormat_BPTC, img.data(), o0.data(), w, h));
  list.AddJob(FasTC::CompressionJob(FasTC::eCompressionFormat_BPTC, top.data(), o1.data(), w, h2));
  list.AddJob(FasTC::CompressionJob(FasTC::eCompressionFormat_BPTC, left.data(), o2.data(), w2, h));
  std::thread t1([&] { BPTCC::CompressAtomic(list); });
  std::thread t2([&] { BPTCC::CompressAtomic(list); });
  t1.join();
  t2.join();
  for (uint32_t i = 0; i < 3; i++)
    if (*list.GetFinishedFlag(i) != 1) return 3;
  BPTCC::CompressAtomic(list);  // nothing left: returns at once
  save(dir + "/atomic0.bin", o0);
  save(dir + "/atomic1.bin", o1);
  save(dir + "/atomic2.bin", o2);
  printf("atomic ok\n");
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 2 && !strcmp(argv[1], "validate")) return validate();
  if (argc >= 6 && !strcmp(argv[1], "encode"))
    return encode(argv[2], (unsigned)atoi(argv[3]), (unsigned)atoi(argv[4]), argv[5]);
  if (argc >= 6 && !strcmp(argv[1], "select"))
    return select(argv[2], (unsigned)atoi(argv[3]), (unsigned)atoi(argv[4]), argv[5]);
  if (argc >= 6 && !strcmp(argv[1], "stats"))
    return stats(argv[2], (unsigned)atoi(argv[3]), (unsigned)atoi(argv[4]), argv[5]);
  if (argc >= 6 && !strcmp(argv[1], "image"))
    return image(argv[2], (unsigned)atoi(argv[3]), (unsigned)atoi(argv[4]), argv[5]);
  if (argc >= 6 && !strcmp(argv[1], "atomic"))
    return atomic(argv[2], (unsigned)atoi(argv[3]), (unsigned)atoi(argv[4]), argv[5]);
  fprintf(stderr, "usage: facade_test validate | encode|select|stats in.raw w h outdir\n");
  return 2;
}
