// pmh_stereo.cpp -- cvt::Image / cvt::PMHuberStereo facade over the C-ABI.
#include "pmh_stereo.h"

#include <hip/hip_runtime.h>

#include <cstring>

namespace cvt {

const IFormat IFormat::GRAY_FLOAT(IFormat::GRAY_FLOAT_ID);
const IFormat IFormat::RGBA_FLOAT(IFormat::RGBA_FLOAT_ID);

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Exception(std::string(what) + ": " + hipGetErrorString(e));
}

Image::Image() = default;

Image::Image(size_t w, size_t h, const IFormat& fmt, IAllocatorType mem) {
    reallocate(w, h, fmt, mem);
}

Image::Image(const Image& src, IAllocatorType mem) {
    reallocate(src._w, src._h, *src._fmt, mem);
    if (!bytes()) return;
    if (src._mem == IALLOCATOR_MEM && mem == IALLOCATOR_MEM) {
        std::memcpy(_data, src._data, bytes());
        return;
    }
    hipMemcpyKind k = src._mem == IALLOCATOR_MEM
                          ? (mem == IALLOCATOR_MEM ? hipMemcpyHostToHost : hipMemcpyHostToDevice)
                          : (mem == IALLOCATOR_MEM ? hipMemcpyDeviceToHost
                                                   : hipMemcpyDeviceToDevice);
    hip_check(hipMemcpy(_data, src._data, bytes(), k), "Image copy");
}

Image::Image(const Image& src) : Image(src, src._mem) {}

Image& Image::operator=(const Image& src) {
    if (this != &src) {
        Image tmp(src, src._mem);
        std::swap(_w, tmp._w);
        std::swap(_h, tmp._h);
        std::swap(_fmt, tmp._fmt);
        std::swap(_mem, tmp._mem);
        std::swap(_data, tmp._data);
    }
    return *this;
}

Image::~Image() { release(); }

void Image::release() {
    if (_data) {
        if (_mem == IALLOCATOR_MEM)
            delete[] _data;
        else
            (void)hipFree(_data);
    }
    _data = nullptr;
}

void Image::reallocate(size_t w, size_t h, const IFormat& fmt, IAllocatorType mem) {
    if (_data && w == _w && h == _h && fmt == *_fmt && mem == _mem) return;
    release();
    _w = w;
    _h = h;
    _fmt = fmt == IFormat::GRAY_FLOAT ? &IFormat::GRAY_FLOAT : &IFormat::RGBA_FLOAT;
    _mem = mem;
    if (!bytes()) return;
    if (mem == IALLOCATOR_MEM) {
        _data = new float[w * h * channels()]();
    } else {
        void* p = nullptr;
        hip_check(hipMalloc(&p, bytes()), "Image allocation");
        _data = static_cast<float*>(p);
    }
}

void Image::upload(const float* host) {
    if (_mem == IALLOCATOR_MEM)
        std::memcpy(_data, host, bytes());
    else
        hip_check(hipMemcpy(_data, host, bytes(), hipMemcpyHostToDevice), "Image upload");
}

void Image::download(float* host) const {
    if (_mem == IALLOCATOR_MEM)
        std::memcpy(host, _data, bytes());
    else
        hip_check(hipMemcpy(host, _data, bytes(), hipMemcpyDeviceToHost), "Image download");
}

static bool ends_with(const std::string& s, const char* suffix) {
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

void Image::save(const std::string& path) const {
    Image host(*this, IALLOCATOR_MEM);
    const int fmt = *_fmt == IFormat::GRAY_FLOAT ? PMH_FORMAT_GRAY : PMH_FORMAT_RGBA;
    int st;
    if (ends_with(path, ".pfm"))
        st = pmh_pfm_save(path.c_str(), host.data(), (int)_w, (int)_h, fmt);
    else if (ends_with(path, ".cvtraw"))
        st = pmh_cvtraw_save(path.c_str(), host.data(), (int)_w, (int)_h,
                             _w * channels() * sizeof(float), fmt == PMH_FORMAT_GRAY ? 4 : 12);
    else
        throw Exception("Image::save: no saver for " + path + " (.pfm, .cvtraw)");
    if (st != PMH_OK) throw Exception(pmh_io_last_error());
}

void Image::load(const std::string& path) {
    int w = 0, h = 0, fid = 0;
    size_t stride = 0;
    if (!ends_with(path, ".cvtraw"))
        throw Exception("Image::load: no loader for " + path + " (.cvtraw)");
    if (pmh_cvtraw_info(path.c_str(), &w, &h, &stride, &fid) != PMH_OK)
        throw Exception(pmh_io_last_error());
    if (fid != 4 && fid != 12)
        throw Exception("Image::load: only GRAY_FLOAT and RGBA_FLOAT images are supported");
    Image host((size_t)w, (size_t)h, fid == 4 ? IFormat::GRAY_FLOAT : IFormat::RGBA_FLOAT,
               IALLOCATOR_MEM);
    if (pmh_cvtraw_load(path.c_str(), host.data(), host._w * host.channels() * sizeof(float)) !=
        PMH_OK)
        throw Exception(pmh_io_last_error());
    *this = Image(host, _mem);
}

PMHuberStereo::PMHuberStereo() {
    pmh_default_options(&_opts);
    pmh_default_inpaint_options(&_iopts);
    if (pmh_create(&_ctx, nullptr) != PMH_OK) throw Exception(pmh_last_error());
}

PMHuberStereo::~PMHuberStereo() { pmh_destroy(_ctx); }

namespace {

// Device copies of host images (the reference requires CL images, :65-68).
struct DeviceInputs {
    const Image* L;
    const Image* R;
    Image lcopy, rcopy;
    DeviceInputs(const Image& left, const Image& right) : L(&left), R(&right) {
        if (left.width() != right.width() || left.height() != right.height())
            throw Exception("Left/Right stereo images inconsistent or incompatible memory type");
        if (left.format() != right.format())
            throw Exception("PMHuberStereo: left and right images must have the same format");
        if (left.memType() != IALLOCATOR_HIP) {
            lcopy = Image(left, IALLOCATOR_HIP);
            L = &lcopy;
        }
        if (right.memType() != IALLOCATOR_HIP) {
            rcopy = Image(right, IALLOCATOR_HIP);
            R = &rcopy;
        }
    }
};

}  // namespace

// cvt/vision/PMHuberStereo.cpp:293-507
void PMHuberStereo::depthMapInpaint(Image& dmap, const Image& left, const Image& right,
                                    size_t patchsize, float depthmax, size_t iterations,
                                    size_t viewsamples) {
    (void)viewsamples;
    DeviceInputs in(left, right);
    dmap.reallocate(left.width(), left.height(), IFormat::GRAY_FLOAT, IALLOCATOR_HIP);  // :499
    pmh_pair pr{};
    pr.left = in.L->data();
    pr.right = in.R->data();
    pr.width = (int)left.width();
    pr.height = (int)left.height();
    pr.format = left.format() == IFormat::GRAY_FLOAT ? PMH_FORMAT_GRAY : PMH_FORMAT_RGBA;
    pr.depthmax = depthmax;
    pr.dmap = dmap.data();
    if (pmh_batch_depthmap_inpaint(_ctx, &pr, 1, (int)patchsize, (int)iterations, &_iopts) !=
        PMH_OK)
        throw Exception(pmh_last_error());
}

// cvt/vision/PMHuberStereo.cpp:64-291.  Host images are uploaded instead of
// rejected; sizes must match as in :65-68.
void PMHuberStereo::depthMap(Image& dmap, const Image& left, const Image& right,
                             size_t patchsize, float depthmax, size_t iterations,
                             size_t viewsamples, float dscale, Image* normalmap) {
    if (left.width() != right.width() || left.height() != right.height())
        throw Exception("Left/Right stereo images inconsistent or incompatible memory type");
    if (left.format() != right.format())
        throw Exception("PMHuberStereo: left and right images must have the same format");
    const Image* L = &left;
    const Image* R = &right;
    Image lcopy, rcopy;
    if (left.memType() != IALLOCATOR_HIP) {
        lcopy = Image(left, IALLOCATOR_HIP);
        L = &lcopy;
    }
    if (right.memType() != IALLOCATOR_HIP) {
        rcopy = Image(right, IALLOCATOR_HIP);
        R = &rcopy;
    }
    // :278 -- dmap becomes a device GRAY_FLOAT image of the input size
    dmap.reallocate(left.width(), left.height(), IFormat::GRAY_FLOAT, IALLOCATOR_HIP);
    float* nm = nullptr;
    if (normalmap) {
        normalmap->reallocate(left.width(), left.height(), IFormat::RGBA_FLOAT, IALLOCATOR_HIP);
        nm = normalmap->data();
    }
    int st;
    if (left.format() == IFormat::GRAY_FLOAT) {
        st = pmh_depthmap(_ctx, L->data(), R->data(), (int)left.width(), (int)left.height(),
                          (int)patchsize, depthmax, (int)iterations, (int)viewsamples, dscale,
                          dmap.data(), nm, &_opts, nullptr, nullptr);
    } else {  // RGBA_FLOAT: colour sums over R, G, B (pmhstereo.cl:199,217)
        pmh_pair pr{};
        pr.left = L->data();
        pr.right = R->data();
        pr.width = (int)left.width();
        pr.height = (int)left.height();
        pr.format = PMH_FORMAT_RGBA;
        pr.depthmax = depthmax;
        pr.dscale = dscale;
        pr.dmap = dmap.data();
        pr.normalmap = nm;
        st = pmh_batch_depthmap(_ctx, &pr, 1, (int)patchsize, (int)iterations, &_opts);
    }
    if (st != PMH_OK) throw Exception(pmh_last_error());
}

}  // namespace cvt
