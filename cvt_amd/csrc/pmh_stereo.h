// pmh_stereo.h -- C++ facade with the reference's class API over the C-ABI.
//
// cvt::PMHuberStereo keeps the signature of cvt/vision/PMHuberStereo.h:35-41;
// cvt::Image is a minimal image container with the two formats the path uses
// (GRAY_FLOAT, RGBA_FLOAT; cvt/gfx/IFormat.h:302,312) and host or device (HBM)
// storage.  IALLOCATOR_CL is kept as a spelling of device storage so that
// callers written against the reference compile unchanged.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>

#include "pmh_capi.h"

namespace cvt {

// cvt/util/Exception.h:35 (CVTException)
class Exception : public std::runtime_error {
  public:
    explicit Exception(const std::string& what) : std::runtime_error(what) {}
};

enum IAllocatorType { IALLOCATOR_MEM = 0, IALLOCATOR_HIP = 1, IALLOCATOR_CL = IALLOCATOR_HIP };

class IFormat {
  public:
    enum Id { GRAY_FLOAT_ID, RGBA_FLOAT_ID };
    static const IFormat GRAY_FLOAT;
    static const IFormat RGBA_FLOAT;
    size_t channels() const { return _id == GRAY_FLOAT_ID ? 1 : 4; }
    Id id() const { return _id; }
    bool operator==(const IFormat& o) const { return _id == o._id; }
    bool operator!=(const IFormat& o) const { return _id != o._id; }

  private:
    explicit constexpr IFormat(Id id) : _id(id) {}
    Id _id;
};

class Image {
  public:
    Image();
    Image(size_t w, size_t h, const IFormat& fmt, IAllocatorType mem = IALLOCATOR_MEM);
    Image(const Image& src, IAllocatorType mem);  // copy into the given storage
    Image(const Image& src);
    Image& operator=(const Image& src);
    ~Image();

    size_t width() const { return _w; }
    size_t height() const { return _h; }
    size_t channels() const { return _fmt->channels(); }
    const IFormat& format() const { return *_fmt; }
    IAllocatorType memType() const { return _mem; }
    void reallocate(size_t w, size_t h, const IFormat& fmt, IAllocatorType mem);

    float* data() { return _data; }
    const float* data() const { return _data; }
    size_t bytes() const { return _w * _h * channels() * sizeof(float); }
    // copies between this image and host memory of bytes() bytes
    void upload(const float* host);
    void download(float* host) const;
    // cvt/gfx/Image.h load/save for the formats of this path: ".cvtraw"
    // (load and save, GRAY_FLOAT / RGBA_FLOAT) and ".pfm" (save); the image
    // keeps its storage type (a device image is staged through the host)
    void load(const std::string& path);
    void save(const std::string& path) const;

  private:
    void release();
    size_t _w = 0, _h = 0;
    const IFormat* _fmt = &IFormat::GRAY_FLOAT;
    IAllocatorType _mem = IALLOCATOR_MEM;
    float* _data = nullptr;
};

// cvt/vision/PMHuberStereo.h:35-59
class PMHuberStereo {
  public:
    PMHuberStereo();
    ~PMHuberStereo();
    PMHuberStereo(const PMHuberStereo&) = delete;
    PMHuberStereo& operator=(const PMHuberStereo&) = delete;

    void depthMap(Image& dmap, const Image& left, const Image& right, size_t patchsize,
                  float depthmax, size_t iterations, size_t viewsamples, float dscale = -1.0f,
                  Image* normalmap = NULL);
    // PMHuberStereo.h:41; dmap in px
    void depthMapInpaint(Image& dmap, const Image& left, const Image& right, size_t patchsize,
                         float depthmax, size_t iterations, size_t viewsamples);

    // not in the reference: the knobs it hard-codes (PMHuberStereo.cpp:70-72, and
    // :454,466 for depthMapInpaint)
    pmh_options& options() { return _opts; }
    pmh_options& inpaintOptions() { return _iopts; }

  private:
    pmh_ctx* _ctx = nullptr;
    pmh_options _opts;
    pmh_options _iopts;
};

}  // namespace cvt
