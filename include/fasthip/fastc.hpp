// C++ mirror of FasTC's encode API on top of the fasthip C ABI.
//
// Same names, argument meaning and error behaviour as the reference:
//   fasthip::FasTC::ECompressionFormat   Base/include/FasTC/CompressionFormat.h:26-55
//   fasthip::FasTC::CompressionJob       Base/include/FasTC/CompressionJob.h:40-140
//   fasthip::FasTC::DecompressionJob     Base/include/FasTC/CompressionJob.h:144-169
//   fasthip::SCompressionSettings        Core/include/FasTC/TexComp.h:30-76
//   fasthip::CompressImageData           Core/include/FasTC/TexComp.h:81-88
//   fasthip::CompressImage               Core/include/FasTC/TexComp.h:78-79
//   fasthip::CompressedImage             Core/include/FasTC/CompressedImage.h:25-75
//   fasthip::FasTC::CompressionJobList   Base/include/FasTC/CompressionJob.h:172-209
//   fasthip::BPTCC::CompressAtomic       BPTCEncoder/include/FasTC/BPTCCompressor.h:184-189
//   fasthip::BPTCC::CompressionSettings  BPTCEncoder/include/FasTC/BPTCCompressor.h:123-158
//   fasthip::BPTCC::Compress/Decompress  BPTCCompressor.h:166-167, 216
//   fasthip::BPTCC::CompressWithStats    BPTCCompressor.h:173-174
//   fasthip::DXTC::CompressImageDXT1/5   DXTEncoder/include/FasTC/DXTCompressor.h:24-25
//   fasthip::ETCC::Compress_RG           ETCEncoder/include/FasTC/ETCCompressor.h:35
// Everything lives in namespace fasthip so it can be linked next to FasTC;
// the only new setting is SCompressionSettings::bUseGPU, mirroring bUseNVTT.
//
// CompressImageData returns false and prints "TexComp -- <message>" on
// stderr like the reference (Core/src/TexComp.cpp:157-159).  The codec entry
// points return void in the reference and have no error channel; here they
// throw fasthip::Error instead of silently producing nothing.  There is no
// CPU fallback: without a HIP device every entry point reports an error.
#ifndef FASTHIP_FASTC_HPP_
#define FASTHIP_FASTC_HPP_

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <ostream>
#include <thread>
#include <stdexcept>
#include <string>
#include <vector>

#include "../fasthip.h"

namespace fasthip {

class Error : public std::runtime_error {
 public:
  Error(fh_status s, const std::string &what) : std::runtime_error(what), status(s) {}
  fh_status status;
};

namespace FasTC {

enum ECompressionFormat {
  eCompressionFormat_DXT1 = FH_FORMAT_DXT1,
  eCompressionFormat_DXT5 = FH_FORMAT_DXT5,
  eCompressionFormat_ETC1 = FH_FORMAT_ETC1,
  eCompressionFormat_BPTC = FH_FORMAT_BPTC
};

class CompressionJob {
 public:
  CompressionJob(ECompressionFormat fmt, const uint8_t *in, uint8_t *out, uint32_t w, uint32_t h)
      : m_Format(fmt), m_InBuf(in), m_OutBuf(out), m_Width(w), m_Height(h), m_XStart(0),
        m_XEnd(w), m_YStart(0), m_YEnd(h) {}
  CompressionJob(ECompressionFormat fmt, const uint8_t *in, uint8_t *out, uint32_t w, uint32_t h,
                 uint32_t x0, uint32_t y0)
      : m_Format(fmt), m_InBuf(in), m_OutBuf(out), m_Width(w), m_Height(h), m_XStart(x0),
        m_XEnd(w), m_YStart(y0), m_YEnd(h) {}
  CompressionJob(ECompressionFormat fmt, const uint8_t *in, uint8_t *out, uint32_t w, uint32_t h,
                 uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1)
      : m_Format(fmt), m_InBuf(in), m_OutBuf(out), m_Width(w), m_Height(h), m_XStart(x0),
        m_XEnd(x1), m_YStart(y0), m_YEnd(y1) {}
  ECompressionFormat Format() const { return m_Format; }
  const uint8_t *InBuf() const { return m_InBuf; }
  uint8_t *OutBuf() const { return m_OutBuf; }
  uint32_t Width() const { return m_Width; }
  uint32_t Height() const { return m_Height; }
  uint32_t XStart() const { return m_XStart; }
  uint32_t XEnd() const { return m_XEnd; }
  uint32_t YStart() const { return m_YStart; }
  uint32_t YEnd() const { return m_YEnd; }
  uint32_t CoordsToBlockIdx(uint32_t x, uint32_t y) const { return (y / 4) * (m_Width / 4) + x / 4; }

 private:
  ECompressionFormat m_Format;
  const uint8_t *m_InBuf;
  uint8_t *m_OutBuf;
  uint32_t m_Width, m_Height;
  uint32_t m_XStart, m_XEnd, m_YStart, m_YEnd;
};

class DecompressionJob {
 public:
  DecompressionJob(ECompressionFormat fmt, const uint8_t *in, uint8_t *out, uint32_t w, uint32_t h)
      : m_Format(fmt), m_InBuf(in), m_OutBuf(out), m_Width(w), m_Height(h) {}
  const uint8_t *InBuf() const { return m_InBuf; }
  uint8_t *OutBuf() const { return m_OutBuf; }
  uint32_t Width() const { return m_Width; }
  uint32_t Height() const { return m_Height; }
  ECompressionFormat Format() const { return m_Format; }

 private:
  ECompressionFormat m_Format;
  const uint8_t *m_InBuf;
  uint8_t *m_OutBuf;
  uint32_t m_Width, m_Height;
};

// CompressedImage::GetCompressedSize (Core/src/CompressedImage.cpp:136-149).
inline size_t GetCompressedSize(uint32_t w, uint32_t h, ECompressionFormat fmt) {
  size_t n = 0;
  return fh_compressed_size((uint32_t)fmt, w, h, &n) == FH_OK ? n : 0;
}

// A list of textures for BPTCC::CompressAtomic (CompressionJob.h:172-209,
// Base/src/CompressionJob.cpp:29-106): room for nJobs jobs, AddJob false
// once full, a finished flag per job, and the shared job / block indices the
// callers advance.  GetJob(idx) returns job idx (the reference's returns the
// job at m_CurrentJobIndex whatever idx is, CompressionJob.cpp:92-98).
class CompressionJobList {
 public:
  explicit CompressionJobList(uint32_t nJobs)
      : m_TotalNumJobs(nJobs), m_Finished(nJobs, 0), m_CurrentJobIndex(0),
        m_CurrentBlockIndex(0) {
    m_Jobs.reserve(nJobs);
  }
  CompressionJobList(const CompressionJobList &) = default;
  CompressionJobList &operator=(const CompressionJobList &) = default;

  bool AddJob(const CompressionJob &cj) {
    if (m_Jobs.size() == m_TotalNumJobs) return false;
    m_Jobs.push_back(cj);
    return true;
  }
  uint32_t GetTotalNumJobs() const { return m_TotalNumJobs; }
  uint32_t GetNumJobs() const { return (uint32_t)m_Jobs.size(); }
  const CompressionJob *GetJob(uint32_t idx) const {
    return idx < m_Jobs.size() ? &m_Jobs[idx] : nullptr;
  }
  uint32_t *GetFinishedFlag(uint32_t idx) const {
    return idx < m_Jobs.size() ? const_cast<uint32_t *>(&m_Finished[idx]) : nullptr;
  }

 private:
  std::vector<CompressionJob> m_Jobs;
  uint32_t m_TotalNumJobs;
  std::vector<uint32_t> m_Finished;

 public:
  uint32_t m_CurrentJobIndex;
  uint32_t m_CurrentBlockIndex;
};

}  // namespace FasTC

namespace detail {

inline void check(fh_status s) {
  if (s != FH_OK) throw Error(s, fh_last_error());
}

inline fh_job make_job(const FasTC::CompressionJob &cj) {
  fh_job j = fh_job();
  j.format = (uint32_t)cj.Format();
  j.in = cj.InBuf();
  j.out = cj.OutBuf();
  j.out_size = FasTC::GetCompressedSize(cj.Width(), cj.Height(), cj.Format());
  j.width = cj.Width();
  j.height = cj.Height();
  j.x_start = cj.XStart();
  j.y_start = cj.YStart();
  j.x_end = cj.XEnd();
  j.y_end = cj.YEnd();
  j.quality = 50;
  j.block_modes = 0xFF;
  return j;
}

// Blocks a job processes, by Compress's row rule (Compressor.cpp:1478-1482).
inline uint32_t job_blocks(const FasTC::CompressionJob &cj) {
  if (cj.Height() < 4) return 0;
  const uint32_t end_y = cj.YEnd() < cj.Height() - 4 ? cj.YEnd() : cj.Height() - 4;
  uint32_t n = 0, sx = cj.XStart();
  for (uint32_t y = cj.YStart(); y <= end_y; y += 4, sx = 0) {
    const uint32_t end_x = y == cj.YEnd() ? cj.XEnd() : cj.Width();
    if (end_x > sx) n += (end_x - sx + 3) / 4;
  }
  return n;
}

}  // namespace detail

namespace BPTCC {

enum ErrorMetric { eErrorMetric_Uniform, eErrorMetric_Nonuniform, kNumErrorMetrics };

// Shape, ShapeSelection and ShapeSelectionFn (BPTCCompressor.h:80-120).
struct Shape {
  uint32_t m_NumPartitions;
  uint32_t m_Index;
};

struct ShapeSelection {
  uint32_t m_NumShapesToSearch = 0;
  Shape m_Shapes[10] = {};
  uint32_t m_SelectedModes = 0xFF;
};

typedef ShapeSelection (*ShapeSelectionFn)(uint32_t x, uint32_t y, const uint32_t pixels[16],
                                           const void *userData);

struct CompressionSettings {
  // Called on the host for every block that is neither one colour nor fully
  // transparent, as the reference calls it (Compressor.cpp:1840-1858); the
  // encode itself runs on the GPU.  Null: BoxSelection.
  ShapeSelectionFn m_ShapeSelectionFn = nullptr;
  const void *m_ShapeSelectionUserData = nullptr;
  uint32_t m_BlockModes = 0xFF;
  ErrorMetric m_ErrorMetric = eErrorMetric_Uniform;
  uint32_t m_NumSimulatedAnnealingSteps = 50;
  // Key of the annealing random stream (the reference seeds a process-wide
  // LCG with time(NULL), Compressor.cpp:358).
  uint64_t m_Seed = 0;
};

inline void Compress(const FasTC::CompressionJob &cj,
                     CompressionSettings settings = CompressionSettings()) {
  fh_job j = detail::make_job(cj);
  j.quality = settings.m_NumSimulatedAnnealingSteps;
  j.block_modes = settings.m_BlockModes;
  j.error_metric = (uint32_t)settings.m_ErrorMetric;
  j.seed = settings.m_Seed;
  if (!settings.m_ShapeSelectionFn) {
    detail::check(fh_compress(&j));
    return;
  }
  // the caller's selection of every block, in the raster loop's order
  // (Compressor.cpp:1476-1490); solid and transparent blocks keep an empty
  // entry, as the reference never asks for theirs
  static_assert(sizeof(ShapeSelection) == sizeof(fh_shape_selection), "ShapeSelection layout");
  std::vector<fh_shape_selection> sels(detail::job_blocks(cj) + 1);
  const uint32_t *px = reinterpret_cast<const uint32_t *>(cj.InBuf());
  const uint32_t end_y = cj.YEnd() < cj.Height() - 4 ? cj.YEnd() : cj.Height() - 4;
  size_t k = 0;
  uint32_t sx = cj.XStart();
  for (uint32_t y = cj.YStart(); y <= end_y && cj.Height() >= 4; y += 4, sx = 0) {
    const uint32_t end_x = y == cj.YEnd() ? cj.XEnd() : cj.Width();
    for (uint32_t x = sx; x < end_x; x += 4, k++) {
      uint32_t blk[16];
      bool solid = true, transparent = true;
      for (int i = 0; i < 16; i++) {
        blk[i] = px[(size_t)(y + i / 4) * cj.Width() + x + i % 4];
        solid = solid && blk[i] == blk[0];
        transparent = transparent && (blk[i] >> 24) == 0;
      }
      sels[k] = fh_shape_selection();
      if (solid || transparent) continue;
      const ShapeSelection g =
          settings.m_ShapeSelectionFn(x, y, blk, settings.m_ShapeSelectionUserData);
      sels[k].num_shapes = g.m_NumShapesToSearch;
      for (int e = 0; e < 10; e++) {
        sels[k].shapes[e].num_partitions = g.m_Shapes[e].m_NumPartitions;
        sels[k].shapes[e].index = g.m_Shapes[e].m_Index;
      }
      sels[k].selected_modes = g.m_SelectedModes;
    }
  }
  detail::check(fh_compress_with_selection(&j, sels.data(), sels.size()));
}

// BPTCC::CompressWithStats (BPTCCompressor.h:173-174, Compressor.cpp:1577-1623):
// the statistics search, and with a log stream the BlockLogger lines of
// every block (fh_bc7_stats_format).  Rows follow Compress's rule: the
// reference's statistics loop has no min(YEnd, Height - 4) clamp
// (Compressor.cpp:1583) and would run a row past a whole-image job.
inline void CompressWithStats(const FasTC::CompressionJob &cj, std::ostream *logStream,
                              CompressionSettings settings = CompressionSettings()) {
  fh_job j = detail::make_job(cj);
  j.quality = settings.m_NumSimulatedAnnealingSteps;
  j.block_modes = settings.m_BlockModes;
  j.error_metric = (uint32_t)settings.m_ErrorMetric;
  j.seed = settings.m_Seed;
  const uint32_t n = detail::job_blocks(cj);
  std::vector<fh_bc7_block_stats> st(n > 0 ? n : 1);
  detail::check(fh_compress_with_stats(&j, st.data(), st.size()));
  if (!logStream) return;
  std::vector<char> text(16384);
  uint64_t idx = cj.CoordsToBlockIdx(cj.XStart(), cj.YStart());
  for (uint32_t k = 0; k < n; k++, idx++) {
    size_t len = 0;
    detail::check(fh_bc7_stats_format(&st[k], idx, text.data(), text.size(), &len));
    logStream->write(text.data(), (std::streamsize)len);
  }
}

inline void Decompress(const FasTC::DecompressionJob &dj) {
  const size_t n = FasTC::GetCompressedSize(dj.Width(), dj.Height(), dj.Format());
  detail::check(fh_decompress((uint32_t)dj.Format(), dj.InBuf(), n, dj.OutBuf(), dj.Width(),
                              dj.Height()));
}

// BPTCC::CompressAtomic (Compressor.cpp:1541-1574): threads that call it with
// the same list work through it together and each returns once every job is
// done.  A caller claims a whole job at a time through m_CurrentJobIndex
// (the reference shares the blocks of the current job through
// m_CurrentBlockIndex; here one job is one GPU encode) and encodes it with
// the default settings, as the reference's CompressBC7Block does, then sets
// its finished flag.
inline void CompressAtomic(FasTC::CompressionJobList &cjl) {
  const uint32_t n = cjl.GetNumJobs();
  for (;;) {
    uint32_t idx = __atomic_load_n(&cjl.m_CurrentJobIndex, __ATOMIC_ACQUIRE);
    if (idx >= n) break;
    if (!__atomic_compare_exchange_n(&cjl.m_CurrentJobIndex, &idx, idx + 1, false,
                                     __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE))
      continue;
    Compress(*cjl.GetJob(idx));
    __atomic_store_n(cjl.GetFinishedFlag(idx), 1u, __ATOMIC_RELEASE);
  }
  // wait for the jobs other callers claimed
  for (uint32_t i = 0; i < n; i++)
    while (!__atomic_load_n(cjl.GetFinishedFlag(i), __ATOMIC_ACQUIRE)) std::this_thread::yield();
}

}  // namespace BPTCC

namespace DXTC {
inline void CompressImageDXT1(const FasTC::CompressionJob &cj) {
  fh_job j = detail::make_job(cj);
  j.format = FH_FORMAT_DXT1;
  detail::check(fh_compress(&j));
}
inline void CompressImageDXT5(const FasTC::CompressionJob &cj) {
  fh_job j = detail::make_job(cj);
  j.format = FH_FORMAT_DXT5;
  detail::check(fh_compress(&j));
}
}  // namespace DXTC

namespace ETCC {
inline void Compress_RG(const FasTC::CompressionJob &cj) {
  fh_job j = detail::make_job(cj);
  j.format = FH_FORMAT_ETC1;
  detail::check(fh_compress(&j));
}
}  // namespace ETCC

// SCompressionSettings (Core/include/FasTC/TexComp.h:30-76) with the same
// defaults as its constructor (Core/src/TexComp.cpp:77-85) plus bUseGPU.
struct SCompressionSettings {
  FasTC::ECompressionFormat format = FasTC::eCompressionFormat_BPTC;
  bool bUseSIMD = false;
  int iNumThreads = 1;   // ignored: the GPU path is one launch
  int iQuality = 50;     // BPTC annealing steps, clamped to [0, 256]
  int iNumCompressions = 1;
  int iJobSize = 0;      // ignored
  bool bUseAtomics = false;
  bool bUsePVRTexLib = false;
  bool bUseNVTT = false;
  bool bUseGPU = true;   // route the encode to the MI355X kernels
  std::ostream *logStream = nullptr;  // BPTC: CompressWithStats and its log
  uint64_t seed = 0;     // annealing stream key
};

// CompressImageData (Core/src/TexComp.cpp:427-525): validates, encodes the
// whole image and returns false with a "TexComp -- ..." message on failure.
inline bool CompressImageData(const unsigned char *data, unsigned int width, unsigned int height,
                              unsigned char *cmpData, unsigned int cmpDataSz,
                              const SCompressionSettings &settings) {
  if (!settings.bUseGPU) {
    fprintf(stderr, "TexComp -- %s\n", "this library only implements the GPU path (bUseGPU)");
    return false;
  }
  if (settings.bUseSIMD) {
    fprintf(stderr, "TexComp -- %s\n", "Platform does not support SIMD!\n");
    return false;
  }
  fh_job j = fh_job();
  j.format = (uint32_t)settings.format;
  j.in = data;
  j.out = cmpData;
  j.out_size = cmpDataSz;
  j.width = width;
  j.height = height;
  j.x_end = width;
  j.y_end = height;
  int q = settings.iQuality;
  j.quality = (uint32_t)(q < 0 ? 0 : (q > 256 ? 256 : q));
  j.block_modes = 0xFF;
  j.seed = settings.seed;
  // with a log stream BPTC runs CompressWithStats (TexComp.cpp:165-181);
  // the reference has no statistics for the other formats and encodes them
  const bool stats = settings.logStream && settings.format == FasTC::eCompressionFormat_BPTC;
  const uint32_t nblocks = (width / 4) * (height / 4);
  std::vector<fh_bc7_block_stats> st(stats && nblocks ? nblocks : 0);
  for (int i = 0; i < (settings.iNumCompressions > 0 ? settings.iNumCompressions : 1); i++) {
    const fh_status s = stats ? fh_compress_with_stats(&j, st.data(), st.size()) : fh_compress(&j);
    if (s != FH_OK) {
      fprintf(stderr, "TexComp -- %s\n", fh_last_error());
      return false;
    }
    std::vector<char> text(16384);
    for (size_t k = 0; k < st.size(); k++) {
      size_t len = 0;
      if (fh_bc7_stats_format(&st[k], k, text.data(), text.size(), &len) == FH_OK)
        settings.logStream->write(text.data(), (std::streamsize)len);
    }
  }
  return true;
}

// CompressedImage (Core/include/FasTC/CompressedImage.h:25-75,
// Core/src/CompressedImage.cpp): the compressed blocks of a width x height
// image in one format, and their decode.  The reference derives it from its
// Image<Pixel> class; this one keeps the members a caller of CompressImage
// reads.
class CompressedImage {
 public:
  CompressedImage(uint32_t width, uint32_t height, FasTC::ECompressionFormat format,
                  const uint8_t *data)
      : m_Width(width), m_Height(height), m_Format(format),
        m_Data(data, data + GetCompressedSize(width, height, format)) {}
  static uint32_t GetCompressedSize(uint32_t width, uint32_t height,
                                    FasTC::ECompressionFormat format) {
    return (uint32_t)FasTC::GetCompressedSize(width, height, format);
  }
  uint32_t GetCompressedSize() const { return GetCompressedSize(m_Width, m_Height, m_Format); }
  uint32_t GetUncompressedSize() const { return m_Width * m_Height * 4; }
  uint32_t GetWidth() const { return m_Width; }
  uint32_t GetHeight() const { return m_Height; }
  const uint8_t *GetCompressedData() const { return m_Data.data(); }
  FasTC::ECompressionFormat GetFormat() const { return m_Format; }
  // CompressedImage::DecompressImage: false when outBuf is smaller than the
  // RGBA8 image; decoded on the GPU.
  bool DecompressImage(uint8_t *outBuf, uint32_t outBufSz) const {
    if (outBufSz < GetUncompressedSize()) return false;
    return fh_decompress((uint32_t)m_Format, m_Data.data(), m_Data.size(), outBuf, m_Width,
                         m_Height) == FH_OK;
  }

 private:
  uint32_t m_Width, m_Height;
  FasTC::ECompressionFormat m_Format;
  std::vector<uint8_t> m_Data;
};

// CompressImage (Core/src/TexComp.cpp:367-424) for any image type with the
// reference Image's interface: GetWidth(), GetHeight(), ComputePixels() and
// operator()(x, y) returning a pixel whose Pack() is the RGBA8 word.  A size
// that is not a multiple of the block size is padded with zeros (with the
// reference's warning); returns null when the encode fails.  The caller
// owns the result.
template <typename ImageT>
CompressedImage *CompressImage(ImageT *img, const SCompressionSettings &settings) {
  if (!img) return nullptr;
  uint32_t width = img->GetWidth();
  uint32_t height = img->GetHeight();
  if (width % 4 || height % 4) {
    fprintf(stderr, "TexComp -- %s\n",
            "WARNING - Image size is not a multiple of block size. Padding with zeros...");
    width = (width + 3) / 4 * 4;
    height = (height + 3) / 4 * 4;
  }
  std::vector<uint32_t> data((size_t)width * height, 0);
  img->ComputePixels();
  for (uint32_t j = 0; j < img->GetHeight(); j++)
    for (uint32_t i = 0; i < img->GetWidth(); i++)
      data[(size_t)j * width + i] = (uint32_t)(*img)(i, j).Pack();
  const uint32_t sz = CompressedImage::GetCompressedSize(width, height, settings.format);
  std::vector<uint8_t> cmp(sz);
  if (!CompressImageData(reinterpret_cast<const unsigned char *>(data.data()), width, height,
                         cmp.data(), sz, settings))
    return nullptr;
  return new CompressedImage(width, height, settings.format, cmp.data());
}

}  // namespace fasthip

#endif  // FASTHIP_FASTC_HPP_
