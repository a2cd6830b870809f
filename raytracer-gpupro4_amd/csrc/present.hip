// Presentation: the frame as the reference's swap chain shows it and its
// screenshot key saves it (Renderer/DX11/RendererDX11.inl:60-100, 134-150).
// The colour stage writes float4 results into an R8G8B8A8_UNORM back buffer,
// so each channel is stored as D3D converts a float to UNORM8: NaN -> 0,
// clamp to [0, 1], times 255, rounded to nearest.  The screenshot is a PNG
// of that buffer (D3DX11SaveTextureToFile).  Here one kernel does the
// conversion on the device and the host writes PNG (RGBA, zlib stream of
// stored deflate blocks, no compression library) or binary PPM (RGB).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rt_internal.h"

namespace {

__device__ __forceinline__ uint32_t unorm8(float f)
{
    if (!(f == f)) return 0u;  // NaN
    const float c = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
    return (uint32_t)floorf(c * 255.0f + 0.5f);
}

// rgba8 = UNORM8(rgba); with average, rgb are divided by the sample count in
// w first and alpha is 1 (an accumulation buffer of rt_render_jobs).
__global__ void k_present(const float4 *__restrict__ src, uint32_t n, int average, uint32_t *__restrict__ dst)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c = src[i];
    if (average) {
        const float s = c.w;
        c = make_float4(c.x / s, c.y / s, c.z / s, 1.0f);
    }
    dst[i] = unorm8(c.x) | (unorm8(c.y) << 8) | (unorm8(c.z) << 16) | (unorm8(c.w) << 24);
}

uint32_t crc32_update(uint32_t crc, const uint8_t *p, size_t n)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t c = k;
            for (int j = 0; j < 8; ++j) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[k] = c;
        }
        init = true;
    }
    crc = ~crc;
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

void put_be32(std::vector<uint8_t> &v, uint32_t x)
{
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void png_chunk(FILE *f, const char *type, const std::vector<uint8_t> &data, bool &ok)
{
    std::vector<uint8_t> head;
    put_be32(head, (uint32_t)data.size());
    head.insert(head.end(), type, type + 4);
    uint32_t crc = crc32_update(0, head.data() + 4, 4);
    crc = crc32_update(crc, data.data(), data.size());
    std::vector<uint8_t> tail;
    put_be32(tail, crc);
    ok = ok && fwrite(head.data(), 1, head.size(), f) == head.size();
    ok = ok && (data.empty() || fwrite(data.data(), 1, data.size(), f) == data.size());
    ok = ok && fwrite(tail.data(), 1, 4, f) == 4;
}

// PNG (ISO/IEC 15948): 8-bit RGBA, filter 0 on every row, the zlib stream
// (RFC 1950) of stored deflate blocks (RFC 1951 section 3.2.4) with Adler-32.
bool write_png(FILE *f, const uint8_t *rgba, uint32_t w, uint32_t h)
{
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    bool ok = fwrite(sig, 1, 8, f) == 8;
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr.push_back(8);  // bit depth
    ihdr.push_back(6);  // colour type RGBA
    ihdr.push_back(0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    png_chunk(f, "IHDR", ihdr, ok);
    const size_t row = (size_t)w * 4 + 1;
    std::vector<uint8_t> raw(row * h);
    for (uint32_t y = 0; y < h; ++y) {
        raw[y * row] = 0;
        memcpy(&raw[y * row + 1], rgba + (size_t)y * w * 4, (size_t)w * 4);
    }
    std::vector<uint8_t> z;
    z.push_back(0x78);  // deflate, 32 K window
    z.push_back(0x01);  // no preset dictionary, check bits
    size_t pos = 0;
    do {
        const size_t len = raw.size() - pos < 65535 ? raw.size() - pos : 65535;
        z.push_back(pos + len == raw.size() ? 1 : 0);  // BFINAL, BTYPE = 00
        z.push_back((uint8_t)(len & 0xFF));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xFF));
        z.push_back((uint8_t)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    png_chunk(f, "IDAT", z, ok);
    png_chunk(f, "IEND", std::vector<uint8_t>(), ok);
    return ok;
}

// Binary PPM (P6), 8-bit RGB; alpha is dropped.
bool write_ppm(FILE *f, const uint8_t *rgba, uint32_t w, uint32_t h)
{
    bool ok = fprintf(f, "P6\n%u %u\n255\n", w, h) > 0;
    std::vector<uint8_t> rgb((size_t)w * 3);
    for (uint32_t y = 0; y < h && ok; ++y) {
        for (uint32_t x = 0; x < w; ++x) memcpy(&rgb[3 * (size_t)x], rgba + ((size_t)y * w + x) * 4, 3);
        ok = fwrite(rgb.data(), 1, rgb.size(), f) == rgb.size();
    }
    return ok;
}

bool ends_with_ci(const std::string &s, const char *suffix)
{
    const size_t n = strlen(suffix);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i) {
        char c = s[s.size() - n + i];
        if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
        if (c != suffix[i]) return false;
    }
    return true;
}

thread_local std::string g_image_err;

}  // namespace

extern "C" {

int rt_present(rt_ctx *ctx, const float *d_rgba, uint32_t width, uint32_t height, uint32_t flags, uint8_t *rgba8)
{
    if (!ctx || !rgba8) return RT_ERR_ARG;
    const float4 *src = reinterpret_cast<const float4 *>(d_rgba);
    if (!src) {
        if (!ctx->frame.rgba || ctx->frame.width != width || ctx->frame.height != height) {
            ctx->err = "rt_present: no frame of that size in the context";
            return RT_ERR_STATE;
        }
        src = ctx->frame.rgba;
    }
    const uint64_t n = (uint64_t)width * height;
    if (n == 0 || n > (1ull << 28)) {
        ctx->err = "rt_present: bad image size";
        return RT_ERR_ARG;
    }
    hipSetDevice(ctx->device);
    void *dst = nullptr;
    hipError_t e = hipMalloc(&dst, 4 * n);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_present, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, ctx->stream, src, (uint32_t)n,
                           (flags & RT_PRESENT_AVERAGE) ? 1 : 0, static_cast<uint32_t *>(dst));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(rgba8, dst, 4 * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(dst);
    if (e != hipSuccess) {
        ctx->err = std::string("rt_present: ") + hipGetErrorString(e);
        return RT_ERR_HIP;
    }
    return RT_OK;
}

int rt_image_write(const char *path, const uint8_t *rgba8, uint32_t width, uint32_t height)
{
    g_image_err.clear();
    if (!path || !rgba8 || width == 0 || height == 0) {
        g_image_err = "rt_image_write: bad argument";
        return RT_ERR_ARG;
    }
    const std::string p(path);
    const bool png = ends_with_ci(p, ".png"), ppm = ends_with_ci(p, ".ppm");
    if (!png && !ppm) {
        g_image_err = "rt_image_write: the file name must end in .png or .ppm";
        return RT_ERR_ARG;
    }
    FILE *f = fopen(path, "wb");
    if (!f) {
        g_image_err = "rt_image_write: cannot open " + p;
        return RT_ERR_ARG;
    }
    const bool ok = png ? write_png(f, rgba8, width, height) : write_ppm(f, rgba8, width, height);
    const bool closed = fclose(f) == 0;
    if (!ok || !closed) {
        g_image_err = "rt_image_write: write failed for " + p;
        return RT_ERR_ARG;
    }
    return RT_OK;
}

const char *rt_image_last_error(void) { return g_image_err.c_str(); }

}  // extern "C"
