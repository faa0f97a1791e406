// Binary STL reading and writing (src/io/stl.rs).
#include "stl.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>

namespace bs {

namespace {

struct File {
    std::FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

// little-endian loads and stores, whatever the host order
uint32_t load_u32(const unsigned char* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
float load_f32(const unsigned char* p) {
    const uint32_t u = load_u32(p);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
void store_u32(unsigned char* p, uint32_t u) {
    p[0] = (unsigned char)u;
    p[1] = (unsigned char)(u >> 8);
    p[2] = (unsigned char)(u >> 16);
    p[3] = (unsigned char)(u >> 24);
}
void store_f32(unsigned char* p, float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    store_u32(p, u);
}

constexpr size_t HEADER = 80;
constexpr size_t FACE = 50;  // normal, 3 vertices (12 floats), u16 attribute

}  // namespace

std::string stl_read(const char* path, std::vector<float>& xyz) {
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return std::string("cannot open ") + path;
    unsigned char head[HEADER + 4];
    if (std::fread(head, 1, sizeof(head), fh.f) != sizeof(head)) return "file ends inside the STL header";
    const uint32_t n = load_u32(head + HEADER);
    xyz.reserve(xyz.size() + (size_t)n * 9);
    std::vector<unsigned char> buf(FACE * 4096);
    for (uint32_t done = 0; done < n;) {
        const uint32_t chunk = std::min<uint32_t>(n - done, 4096);
        if (std::fread(buf.data(), FACE, chunk, fh.f) != chunk)
            return "file ends after " + std::to_string(done) + " of " + std::to_string(n) + " faces";
        for (uint32_t i = 0; i < chunk; ++i) {
            const unsigned char* f = buf.data() + FACE * i + 12;  // skip the normal
            for (int k = 0; k < 9; ++k) xyz.push_back(load_f32(f + 4 * k));
        }
        done += chunk;
    }
    return {};
}

std::string stl_write(const char* path, const float* xyz, size_t n_tris) {
    if (n_tris > 0xFFFFFFFFull) return "mesh is too big for STL";  // io/stl.rs:163-166
    File fh;
    fh.f = std::fopen(path, "wb");
    if (!fh.f) return std::string("cannot create ") + path;
    unsigned char head[HEADER + 4] = {};
    store_u32(head + HEADER, (uint32_t)n_tris);
    if (std::fwrite(head, 1, sizeof(head), fh.f) != sizeof(head)) return "write failed";
    std::vector<unsigned char> buf(FACE * 4096);
    for (size_t done = 0; done < n_tris;) {
        const size_t chunk = std::min<size_t>(n_tris - done, 4096);
        for (size_t i = 0; i < chunk; ++i) {
            const float* t = xyz + 9 * (done + i);
            const float e1[3] = {t[3] - t[0], t[4] - t[1], t[5] - t[2]};
            const float e2[3] = {t[6] - t[0], t[7] - t[1], t[8] - t[2]};
            float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            const float n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
            if (n2 == 0.0f) {
                n[0] = n[1] = n[2] = 0.0f;
            } else {
                const float len = std::sqrt(n2);
                for (float& c : n) c /= len;
            }
            unsigned char* f = buf.data() + FACE * i;
            for (int k = 0; k < 3; ++k) store_f32(f + 4 * k, n[k]);
            for (int k = 0; k < 9; ++k) store_f32(f + 12 + 4 * k, t[k]);
            f[48] = f[49] = 0;
        }
        if (std::fwrite(buf.data(), FACE, chunk, fh.f) != chunk) return "write failed";
        done += chunk;
    }
    if (std::fflush(fh.f) != 0) return "write failed";
    return {};
}

}  // namespace bs
