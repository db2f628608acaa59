// Film output: the OpenEXR file SORT writes at the end of a render and the display tiles it
// streams while rendering.
//
// * EXR: Texture2DBase::Output (texture/texturebase.cpp:27-57) calls tinyexr's SaveEXR(rgb, w, h,
//   3, save_as_fp16 = 1): a single-part scanline file, channels B, G, R as half floats, no
//   compression (InitEXRHeader zero-fills the header), one scanline per chunk. The bytes here are
//   that file's: the same attributes in the same order, the same offset table, and tinyexr's
//   float-to-half rounding (the ISPC "float_to_half_full" conversion: round half away from zero
//   on the dropped bits, float denormals to zero, overflow to infinity, NaN to a quiet NaN).
//   A device film is converted to half scanlines by a kernel, so only 6 bytes per pixel cross to
//   the host.
// * Display tiles (core/display_mgr.cpp:86-140): Blender mode sends per tile [length without
//   itself, w, h, x, y] then RGBA float32 rows from the tile's bottom row up (alpha 1); TEV mode
//   sends per channel an UpdateImage packet [length including itself, type 3, grab 0, title,
//   channel, x, y, w, h, floats row by row], and CreateImage [length, type 4, grab 1, title,
//   width, height, 3, "R", "G", "B"] once per image.
#include <hip/hip_runtime.h>

#include <arpa/inet.h>
#include <netdb.h>
#include <sys/socket.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "api_internal.h"
#include "sortgpu.h"

namespace sortgpu {
namespace {

// tinyexr float_to_half_full (tinyexr.h:7133-7171), on the bit patterns.
__host__ __device__ inline uint16_t exr_half(float f) {
    const uint32_t b = __builtin_bit_cast(uint32_t, f);
    const uint32_t sign = b >> 31, exp = (b >> 23) & 0xffu, mant = b & 0x7fffffu;
    uint32_t o = 0;  // exponent and mantissa; the sign is set last
    if (exp == 0) {
        o = 0;
    } else if (exp == 255) {
        o = (31u << 10) | (mant ? 0x200u : 0u);
    } else {
        const int newexp = (int)exp - 127 + 15;
        if (newexp >= 31) {
            o = 31u << 10;
        } else if (newexp <= 0) {
            if (14 - newexp <= 24) {
                const uint32_t m = mant | 0x800000u;
                o = (m >> (14 - newexp)) & 0x3ffu;
                if ((m >> (13 - newexp)) & 1u) o++;
            }
        } else {
            o = ((uint32_t)newexp << 10) | (mant >> 13);
            if (mant & 0x1000u) o++;
        }
    }
    return (uint16_t)((o & 0x7fffu) | (sign << 15));
}

// The file image of a device film from its offset table on (the header is written on the host):
// body[0, 8h) the offset table, then one chunk per scanline, (y, byte count, B[w], G[w], R[w]).
// One thread per pixel writes its three halves; the first thread of a line also writes the line's
// chunk header and offset-table entry (file offset = head + body offset). A chunk is 8 + 6w bytes,
// an even count, so every half sits at an even body offset; the chunk header's 32-bit words need
// not be 4-byte aligned and are written byte by byte.
__global__ void k_exr_body(const float* __restrict__ rgb, uint32_t w, uint32_t h, uint64_t head,
                           uint8_t* __restrict__ body) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)w * h) return;
    const uint64_t y = i / w, x = i - y * w;
    const uint64_t chunk = 8ull + 6ull * w;
    uint8_t* c = body + 8ull * h + y * chunk;
    if (x == 0) {
        reinterpret_cast<uint64_t*>(body)[y] = head + 8ull * h + y * chunk;
        const uint32_t yy = (uint32_t)y, len = 6u * w;
        for (int b = 0; b < 4; b++) {
            c[b] = (uint8_t)(yy >> (8 * b));
            c[4 + b] = (uint8_t)(len >> (8 * b));
        }
    }
    const float* p = rgb + 3 * i;
    uint16_t* line = reinterpret_cast<uint16_t*>(c + 8);
    line[x] = exr_half(p[2]);
    line[w + x] = exr_half(p[1]);
    line[2 * (uint64_t)w + x] = exr_half(p[0]);
}

void put(std::vector<uint8_t>& v, const void* p, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(p);
    v.insert(v.end(), b, b + n);
}
void put_i32(std::vector<uint8_t>& v, int32_t x) { put(v, &x, 4); }
void put_f32(std::vector<uint8_t>& v, float x) { put(v, &x, 4); }
void put_str(std::vector<uint8_t>& v, const char* s) { put(v, s, std::strlen(s) + 1); }
void put_attr(std::vector<uint8_t>& v, const char* name, const char* type, const std::vector<uint8_t>& data) {
    put_str(v, name);
    put_str(v, type);
    put_i32(v, (int32_t)data.size());
    put(v, data.data(), data.size());
}

// Header of SaveEXRImageToMemory (tinyexr.h:11474-11604) for 3 half channels, no compression.
std::vector<uint8_t> exr_header(uint32_t w, uint32_t h) {
    std::vector<uint8_t> v = {0x76, 0x2f, 0x31, 0x01, 2, 0, 0, 0};
    std::vector<uint8_t> ch;
    for (const char* name : {"B", "G", "R"}) {
        put_str(ch, name);
        put_i32(ch, 1);  // HALF
        const uint8_t plinear[4] = {0, 0, 0, 0};
        put(ch, plinear, 4);
        put_i32(ch, 1);  // x sampling
        put_i32(ch, 1);  // y sampling
    }
    ch.push_back(0);
    put_attr(v, "channels", "chlist", ch);
    put_attr(v, "compression", "compression", {0});
    std::vector<uint8_t> box;
    put_i32(box, 0);
    put_i32(box, 0);
    put_i32(box, (int32_t)w - 1);
    put_i32(box, (int32_t)h - 1);
    put_attr(v, "dataWindow", "box2i", box);
    put_attr(v, "displayWindow", "box2i", box);
    put_attr(v, "lineOrder", "lineOrder", {0});
    std::vector<uint8_t> f;
    put_f32(f, 1.0f);
    put_attr(v, "pixelAspectRatio", "float", f);
    std::vector<uint8_t> c;
    put_f32(c, 0.0f);
    put_f32(c, 0.0f);
    put_attr(v, "screenWindowCenter", "v2f", c);
    std::vector<uint8_t> sw;
    put_f32(sw, (float)w);
    put_attr(v, "screenWindowWidth", "float", sw);
    v.push_back(0);
    return v;
}

uint64_t exr_bytes(uint32_t w, uint32_t h) {
    return exr_header(w, h).size() + 8ull * h + (8ull + 6ull * w) * h;
}

// Header, offsets, chunks around half scanlines (3 * w halves per line).
void exr_assemble(uint32_t w, uint32_t h, const uint16_t* lines, uint8_t* out) {
    const std::vector<uint8_t> head = exr_header(w, h);
    uint8_t* p = out;
    std::memcpy(p, head.data(), head.size());
    p += head.size();
    const uint64_t chunk = 8ull + 6ull * w;
    uint64_t off = head.size() + 8ull * h;
    for (uint32_t y = 0; y < h; y++, off += chunk) {
        std::memcpy(p, &off, 8);
        p += 8;
    }
    const uint32_t len = 6u * w;
    for (uint32_t y = 0; y < h; y++) {
        const int32_t yy = (int32_t)y;
        std::memcpy(p, &yy, 4);
        std::memcpy(p + 4, &len, 4);
        std::memcpy(p + 8, lines + 3ull * w * y, len);
        p += chunk;
    }
}

int check_size(uint32_t w, uint32_t h, const char* what) {
    // SaveEXR refuses images below 16 x 16 (tinyexr.h:12862-12864)
    if (w < 16 || h < 16 || (uint64_t)w * h > (1ull << 32)) {
        set_error(nullptr, "%s: image %ux%u outside 16x16 .. 2^32 pixels", what, w, h);
        return SORTGPU_ERR_ARG;
    }
    return SORTGPU_OK;
}

// The half scanlines of a film in host memory.
int film_lines_host(const float* film, uint32_t w, uint32_t h, std::vector<uint16_t>& lines) {
    const uint64_t n = (uint64_t)w * h;
    try {
        lines.resize(3 * n);
    } catch (const std::bad_alloc&) {
        set_error(nullptr, "film output: host allocation failed");
        return SORTGPU_ERR_OOM;
    }
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t y = i / w, x = i - y * w;
        uint16_t* line = lines.data() + 3 * (uint64_t)w * y;
        line[x] = exr_half(film[3 * i + 2]);
        line[w + x] = exr_half(film[3 * i + 1]);
        line[2 * (uint64_t)w + x] = exr_half(film[3 * i + 0]);
    }
    return SORTGPU_OK;
}

// The EXR file of a device film into `out`: the file image after the header is built on the
// device and copied into place in one transfer; the host writes only the header.
int film_exr_device(const float* film, uint32_t w, uint32_t h, uint8_t* out, void* stream) {
    const uint64_t n = (uint64_t)w * h;
    const std::vector<uint8_t> head = exr_header(w, h);
    const uint64_t body_bytes = 8ull * h + (8ull + 6ull * w) * h;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint8_t* d = nullptr;
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&d), body_bytes, s);
    if (e != hipSuccess) return hip_fail(nullptr, e, "film output: hipMallocAsync");
    const uint32_t block = 256;
    hipLaunchKernelGGL(k_exr_body, dim3((uint32_t)((n + block - 1) / block)), dim3(block), 0, s, film, w, h,
                       (uint64_t)head.size(), d);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out + head.size(), d, body_bytes, hipMemcpyDeviceToHost, s);
    const hipError_t ef = hipFreeAsync(d, s);
    if (e == hipSuccess) e = ef;
    std::memcpy(out, head.data(), head.size());
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(nullptr, e, "film output");
    return SORTGPU_OK;
}

bool send_all(int fd, const uint8_t* p, size_t n) {
    while (n > 0) {
        const ssize_t k = send(fd, p, n, MSG_NOSIGNAL);
        if (k < 0 && errno == EINTR) continue;  // interrupted before any byte went out
        if (k <= 0) return false;
        p += k;
        n -= (size_t)k;
    }
    return true;
}

}  // namespace
}  // namespace sortgpu

using namespace sortgpu;

struct sortgpu_display {
    int fd = -1;
    int mode = 0;
};

extern "C" {

uint64_t sortgpu_exr_size(uint32_t width, uint32_t height) { return exr_bytes(width, height); }

int sortgpu_film_encode_exr(const float* film_rgb, int on_device, uint32_t width, uint32_t height, uint8_t* out,
                            uint64_t cap, void* stream) {
    if (!film_rgb || !out) {
        set_error(nullptr, "film_encode_exr: null film or output");
        return SORTGPU_ERR_ARG;
    }
    int rc = check_size(width, height, "film_encode_exr");
    if (rc != SORTGPU_OK) return rc;
    if (cap < exr_bytes(width, height)) {
        set_error(nullptr, "film_encode_exr: %llu bytes needed, %llu given",
                  (unsigned long long)exr_bytes(width, height), (unsigned long long)cap);
        return SORTGPU_ERR_ARG;
    }
    if (on_device) return film_exr_device(film_rgb, width, height, out, stream);
    std::vector<uint16_t> lines;
    rc = film_lines_host(film_rgb, width, height, lines);
    if (rc != SORTGPU_OK) return rc;
    exr_assemble(width, height, lines.data(), out);
    return SORTGPU_OK;
}

int sortgpu_film_write_exr(const float* film_rgb, int on_device, uint32_t width, uint32_t height, const char* path,
                           void* stream) {
    if (!path) {
        set_error(nullptr, "film_write_exr: null path");
        return SORTGPU_ERR_ARG;
    }
    int rc = check_size(width, height, "film_write_exr");
    if (rc != SORTGPU_OK) return rc;
    std::vector<uint8_t> buf;
    try {
        buf.resize(exr_bytes(width, height));
    } catch (const std::bad_alloc&) {
        set_error(nullptr, "film_write_exr: host allocation failed");
        return SORTGPU_ERR_OOM;
    }
    rc = sortgpu_film_encode_exr(film_rgb, on_device, width, height, buf.data(), buf.size(), stream);
    if (rc != SORTGPU_OK) return rc;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) {
        set_error(nullptr, "film_write_exr: cannot open %s", path);
        return SORTGPU_ERR_ARG;
    }
    const size_t wrote = std::fwrite(buf.data(), 1, buf.size(), fp);
    const int closed = std::fclose(fp);
    if (wrote != buf.size() || closed != 0) {
        set_error(nullptr, "film_write_exr: write to %s failed", path);
        return SORTGPU_ERR_ARG;
    }
    return SORTGPU_OK;
}

int64_t sortgpu_display_packet(int mode, const char* title, const float* film_rgb, uint32_t width, uint32_t height,
                               uint32_t x, uint32_t y, uint32_t tw, uint32_t th, uint8_t* out, uint64_t cap) {
    if ((mode != SORTGPU_DISPLAY_BLENDER && mode != SORTGPU_DISPLAY_TEV && mode != SORTGPU_DISPLAY_TEV_CREATE) ||
        (mode != SORTGPU_DISPLAY_BLENDER && !title) || (mode != SORTGPU_DISPLAY_TEV_CREATE && !film_rgb)) {
        set_error(nullptr, "display_packet: bad mode or null title / film");
        return SORTGPU_ERR_ARG;
    }
    if (mode != SORTGPU_DISPLAY_TEV_CREATE &&
        (tw == 0 || th == 0 || x >= width || y >= height || tw > width - x || th > height - y)) {
        set_error(nullptr, "display_packet: tile %u,%u %ux%u outside the %ux%u film", x, y, tw, th, width, height);
        return SORTGPU_ERR_ARG;
    }
    // the packets' length fields are 32-bit signed integers
    if (mode != SORTGPU_DISPLAY_TEV_CREATE && (uint64_t)tw * th * 16u + 4096u > 0x7fffffffull) {
        set_error(nullptr, "display_packet: tile %ux%u too large for one packet", tw, th);
        return SORTGPU_ERR_ARG;
    }
    std::vector<uint8_t> v;
    try {
        if (mode == SORTGPU_DISPLAY_BLENDER) {
            put_i32(v, (int32_t)(tw * th * 16u + 16u));
            put_i32(v, (int32_t)tw);
            put_i32(v, (int32_t)th);
            put_i32(v, (int32_t)x);
            put_i32(v, (int32_t)y);
            for (uint32_t ly = th; ly-- > 0;)  // the tile's bottom row first
                for (uint32_t lx = 0; lx < tw; lx++) {
                    const float* p = film_rgb + 3 * ((uint64_t)(y + ly) * width + x + lx);
                    put_f32(v, p[0]);
                    put_f32(v, p[1]);
                    put_f32(v, p[2]);
                    put_f32(v, 1.0f);
                }
        } else if (mode == SORTGPU_DISPLAY_TEV) {
            for (int c = 0; c < 3; c++) {
                const size_t start = v.size();
                put_i32(v, 0);
                v.push_back(3);  // UpdateImage
                v.push_back(0);  // the current image
                put_str(v, title);
                put_str(v, c == 0 ? "R" : c == 1 ? "G" : "B");
                put_i32(v, (int32_t)x);
                put_i32(v, (int32_t)y);
                put_i32(v, (int32_t)tw);
                put_i32(v, (int32_t)th);
                for (uint32_t ly = 0; ly < th; ly++)
                    for (uint32_t lx = 0; lx < tw; lx++)
                        put_f32(v, film_rgb[3 * ((uint64_t)(y + ly) * width + x + lx) + c]);
                const int32_t len = (int32_t)(v.size() - start);
                std::memcpy(v.data() + start, &len, 4);
            }
        } else {
            put_i32(v, 0);
            v.push_back(4);  // CreateImage
            v.push_back(1);
            put_str(v, title);
            put_i32(v, (int32_t)width);
            put_i32(v, (int32_t)height);
            put_i32(v, 3);
            for (const char* name : {"R", "G", "B"}) put_str(v, name);
            const int32_t len = (int32_t)v.size();
            std::memcpy(v.data(), &len, 4);
        }
    } catch (const std::bad_alloc&) {
        set_error(nullptr, "display_packet: host allocation failed");
        return SORTGPU_ERR_OOM;
    }
    if (out) {
        if (cap < v.size()) {
            set_error(nullptr, "display_packet: %zu bytes needed, %llu given", v.size(), (unsigned long long)cap);
            return SORTGPU_ERR_ARG;
        }
        std::memcpy(out, v.data(), v.size());
    }
    return (int64_t)v.size();
}

int sortgpu_display_connect(const char* host, uint16_t port, sortgpu_display** out) {
    if (!host || !out) {
        set_error(nullptr, "display_connect: null argument");
        return SORTGPU_ERR_ARG;
    }
    *out = nullptr;
    char port_s[16];
    std::snprintf(port_s, sizeof(port_s), "%u", (unsigned)port);
    addrinfo hints{};
    hints.ai_family = AF_UNSPEC;
    hints.ai_socktype = SOCK_STREAM;
    addrinfo* res = nullptr;
    if (getaddrinfo(host, port_s, &hints, &res) != 0 || !res) {
        set_error(nullptr, "display_connect: cannot resolve %s", host);
        return SORTGPU_ERR_ARG;
    }
    int fd = -1;
    for (addrinfo* a = res; a; a = a->ai_next) {
        fd = socket(a->ai_family, a->ai_socktype, a->ai_protocol);
        if (fd < 0) continue;
        if (connect(fd, a->ai_addr, a->ai_addrlen) == 0) break;
        close(fd);
        fd = -1;
    }
    freeaddrinfo(res);
    if (fd < 0) {
        set_error(nullptr, "display_connect: cannot connect to %s:%u", host, (unsigned)port);
        return SORTGPU_ERR_ARG;
    }
    sortgpu_display* d = new (std::nothrow) sortgpu_display();
    if (!d) {
        close(fd);
        set_error(nullptr, "display_connect: host allocation failed");
        return SORTGPU_ERR_OOM;
    }
    d->fd = fd;
    *out = d;
    return SORTGPU_OK;
}

int sortgpu_display_send_film(sortgpu_display* disp, int mode, const char* title, const float* film_rgb, int on_device,
                              uint32_t width, uint32_t height, uint32_t tile, void* stream) {
    if (!disp || !film_rgb || (mode != SORTGPU_DISPLAY_BLENDER && mode != SORTGPU_DISPLAY_TEV) ||
        (mode == SORTGPU_DISPLAY_TEV && !title) || width == 0 || height == 0) {
        set_error(nullptr, "display_send_film: bad argument");
        return SORTGPU_ERR_ARG;
    }
    if (tile == 0) tile = 64;  // IMAGE_TILE_SIZE (image_evaluation.cpp:44)
    const uint64_t n = 3ull * width * height;
    std::vector<float> host;
    const float* film = film_rgb;
    if (on_device) {
        try {
            host.resize(n);
        } catch (const std::bad_alloc&) {
            set_error(nullptr, "display_send_film: host allocation failed");
            return SORTGPU_ERR_OOM;
        }
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipError_t e = hipMemcpyAsync(host.data(), film_rgb, n * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(nullptr, e, "display_send_film: copy");
        film = host.data();
    }
    std::vector<uint8_t> buf;
    auto send_packet = [&](int m, uint32_t x, uint32_t y, uint32_t tw, uint32_t th) -> int {
        const int64_t k = sortgpu_display_packet(m, title, film, width, height, x, y, tw, th, nullptr, 0);
        if (k < 0) return (int)k;
        buf.resize((size_t)k);
        sortgpu_display_packet(m, title, film, width, height, x, y, tw, th, buf.data(), buf.size());
        if (!send_all(disp->fd, buf.data(), buf.size())) {
            set_error(nullptr, "display_send_film: the display server closed the connection");
            return SORTGPU_ERR_STATE;
        }
        return SORTGPU_OK;
    };
    if (mode == SORTGPU_DISPLAY_TEV) {
        const int rc = send_packet(SORTGPU_DISPLAY_TEV_CREATE, 0, 0, 0, 0);
        if (rc != SORTGPU_OK) return rc;
    }
    for (uint32_t y = 0; y < height; y += tile)
        for (uint32_t x = 0; x < width; x += tile) {
            const uint32_t tw = width - x < tile ? width - x : tile;
            const uint32_t th = height - y < tile ? height - y : tile;
            const int rc = send_packet(mode, x, y, tw, th);
            if (rc != SORTGPU_OK) return rc;
        }
    return SORTGPU_OK;
}

void sortgpu_display_close(sortgpu_display* disp) {
    if (!disp) return;
    if (disp->fd >= 0) close(disp->fd);
    delete disp;
}

}  // extern "C"
