// pmh_io.cpp -- on-disk formats for disparity maps and plane fields: the PFM
// writer of plugins/isaver/PFM/PFMSaver.cpp:117-160 and the CVTRaw reader and
// writer of plugins/iloader/CVTRaw/CVTRawLoader.cpp and
// plugins/isaver/CVTRaw/CVTRawSaver.cpp.  Host memory only.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "pmh_capi.h"
#include "pmh_internal.hpp"

using namespace pmh;

namespace {

// bytes per pixel of the image format ids (cvt/gfx/IFormat.h:36-63,
// cvt/gfx/IFormat.cpp:32-52: channels x sizeof(type))
int bytes_per_pixel(uint32_t id) {
    switch (id) {
    case 1: return 1;                    // GRAY_UINT8
    case 2: case 3: return 2;            // GRAY_UINT16, GRAY_INT16
    case 4: return 4;                    // GRAY_FLOAT
    case 5: return 2;                    // GRAYALPHA_UINT8
    case 6: case 7: return 4;            // GRAYALPHA_UINT16, _INT16
    case 8: return 8;                    // GRAYALPHA_FLOAT
    case 9: case 13: return 4;           // RGBA_UINT8, BGRA_UINT8
    case 10: case 11: case 14: case 15: return 8;
    case 12: case 16: return 16;         // RGBA_FLOAT, BGRA_FLOAT
    case 17: case 18: case 19: return 1; // BAYER_RGGB / GRBG / GBRG _UINT8
    case 20: case 21: return 2;          // YUYV_UINT8, UYVY_UINT8
    default: return 0;
    }
}

struct File {
    FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {
        if (!f) throw ArgError(std::string("cannot open ") + path);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void write(const void* p, size_t n) {
        if (n && std::fwrite(p, 1, n, f) != n) throw ArgError("short write");
    }
    void read(void* p, size_t n) {
        if (n && std::fread(p, 1, n, f) != n) throw ArgError("file shorter than its header says");
    }
};

struct RawHeader {
    uint32_t width, height, stride, format;
};

RawHeader read_header(File& in) {
    RawHeader h;
    in.read(&h, sizeof(h));  // CVTRawLoader.cpp: width, height, stride, IFormat id
    const int bpp = bytes_per_pixel(h.format);
    if (!bpp) throw ArgError("unknown image format id " + std::to_string(h.format));
    if ((uint64_t)h.width * bpp > h.stride) throw ArgError("stride shorter than a row");
    return h;
}

thread_local std::string g_io_error;

template <class F>
int io_guarded(F&& f) {
    try {
        f();
        return PMH_OK;
    } catch (const std::exception& e) {
        g_io_error = e.what();
        return PMH_ERR_INVALID;
    }
}

}  // namespace

extern "C" {

const char* pmh_io_last_error(void) { return g_io_error.c_str(); }

int pmh_pfm_save(const char* path, const float* img, int width, int height, int format) {
    return io_guarded([&] {
        if (!path || !img) throw ArgError("null path or image");
        if (width <= 0 || height <= 0) throw ArgError("image size must be positive");
        if (format != PMH_FORMAT_GRAY && format != PMH_FORMAT_RGBA)
            throw ArgError("format must be PMH_FORMAT_GRAY or PMH_FORMAT_RGBA");
        const int nc = format == PMH_FORMAT_RGBA ? 4 : 1;
        File out(path, "wb");
        // PFMSaver.cpp:143-158: "PF" (RGB) or "Pf" (gray), "W H", scale with
        // a leading '-' on little-endian hosts
        const std::string head = std::string(nc == 4 ? "PF" : "Pf") + "\n" +
                                 std::to_string(width) + " " + std::to_string(height) +
                                 "\n-1.0\n";
        out.write(head.data(), head.size());
        std::vector<float> row((size_t)width * (nc == 4 ? 3 : 1));
        for (int y = height - 1; y >= 0; --y) {  // bottom to top (:160-175)
            const float* src = img + (size_t)y * width * nc;
            if (nc == 1) {
                out.write(src, (size_t)width * sizeof(float));
                continue;
            }
            for (int x = 0; x < width; ++x)  // alpha removed (AlphaRemover, :41-80)
                for (int c = 0; c < 3; ++c) row[(size_t)x * 3 + c] = src[(size_t)x * 4 + c];
            out.write(row.data(), row.size() * sizeof(float));
        }
    });
}

int pmh_cvtraw_save(const char* path, const void* img, int width, int height, size_t stride,
                    int format_id) {
    return io_guarded([&] {
        if (!path || !img) throw ArgError("null path or image");
        const int bpp = bytes_per_pixel((uint32_t)format_id);
        if (!bpp) throw ArgError("unknown image format id " + std::to_string(format_id));
        if (width <= 0 || height <= 0) throw ArgError("image size must be positive");
        if (stride < (size_t)width * bpp || stride > 0xffffffffull)
            throw ArgError("bad stride");
        File out(path, "wb");
        const RawHeader h{(uint32_t)width, (uint32_t)height, (uint32_t)stride,
                          (uint32_t)format_id};
        out.write(&h, sizeof(h));  // CVTRawSaver.cpp: header then height x stride bytes
        out.write(img, (size_t)height * stride);
    });
}

int pmh_cvtraw_info(const char* path, int* width, int* height, size_t* stride, int* format_id) {
    return io_guarded([&] {
        if (!path || !width || !height || !stride || !format_id) throw ArgError("null argument");
        File in(path, "rb");
        const RawHeader h = read_header(in);
        *width = (int)h.width;
        *height = (int)h.height;
        *stride = h.stride;
        *format_id = (int)h.format;
    });
}

int pmh_cvtraw_load(const char* path, void* img, size_t stride) {
    return io_guarded([&] {
        if (!path || !img) throw ArgError("null path or image");
        File in(path, "rb");
        const RawHeader h = read_header(in);
        const size_t row = (size_t)h.width * bytes_per_pixel(h.format);
        if (stride < row) throw ArgError("destination stride shorter than a row");
        std::vector<unsigned char> buf(h.stride);
        auto* dst = static_cast<unsigned char*>(img);
        for (uint32_t y = 0; y < h.height; ++y) {  // rows copied at the caller's stride
            in.read(buf.data(), h.stride);
            std::memcpy(dst + (size_t)y * stride, buf.data(), row);
        }
    });
}

}  // extern "C"
