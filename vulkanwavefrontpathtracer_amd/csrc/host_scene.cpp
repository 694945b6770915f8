// Camera matrices, primary-ray pixel order and the seeded procedural scenes.
#include <cmath>
#include <cstring>
#include <algorithm>
#include <vector>

#include "common.h"
#include "host.h"

namespace wfpt {

// ------------------------------------------------------------------ camera
// Column-major 4x4 helpers (glm layout: m[col*4 + row]).
namespace {
struct M4 {
    float m[16];
};
M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int c = 0; c < 4; ++c)
        for (int rr = 0; rr < 4; ++rr) {
            float s = 0.f;
            for (int k = 0; k < 4; ++k) s += a.m[k * 4 + rr] * b.m[c * 4 + k];
            r.m[c * 4 + rr] = s;
        }
    return r;
}
// General inverse by cofactors, evaluated in double and rounded once.
M4 inverse(const M4& a) {
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = a.m[i];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    M4 r;
    for (int i = 0; i < 16; ++i) r.m[i] = (float)(inv[i] / det);
    return r;
}

struct F3 {
    float x, y, z;
};
F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
F3 norm(F3 a) {
    float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / l, a.y / l, a.z / l};
}
F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 scale(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }

// camera.hpp:47-64: rotate v about u; n becomes u x v.
void rotate_axis(float rad, F3* u, F3* v, F3* n) {
    float s = std::sin(rad), c = std::cos(rad);
    *v = norm(add(scale(*v, c), scale(cross(*u, *v), s)));
    *n = norm(cross(*u, *v));
}
}  // namespace

void camera_matrices(const float pos[3], const float rot_deg[3], float fov_deg, float aspect,
                     float znear, float zfar, CameraMatrices& out) {
    const float deg2rad = 0.01745329251994329576923690768489f;  // glm::radians
    // camera.hpp:23-45 (updateViewMatrix)
    F3 u{1, 0, 0}, v{0, 1, 0}, n{0, 0, 1};
    if (rot_deg[1] * deg2rad != 0.f) rotate_axis(rot_deg[1] * deg2rad, &v, &n, &u);
    if (rot_deg[0] * deg2rad != 0.f) rotate_axis(rot_deg[0] * deg2rad, &u, &v, &n);
    if (rot_deg[2] * deg2rad != 0.f) rotate_axis(rot_deg[2] * deg2rad, &n, &u, &v);
    // view = translate(transpose(mat4(u, v, n, e3)), -position)
    M4 rt{};
    F3 cols[3] = {u, v, n};
    for (int c = 0; c < 3; ++c) {  // transpose: row c of rt = column c of rot
        rt.m[0 * 4 + c] = cols[c].x;
        rt.m[1 * 4 + c] = cols[c].y;
        rt.m[2 * 4 + c] = cols[c].z;
    }
    rt.m[15] = 1.f;
    M4 view = rt;
    for (int r = 0; r < 4; ++r)
        view.m[12 + r] = rt.m[0 + r] * -pos[0] + rt.m[4 + r] * -pos[1] + rt.m[8 + r] * -pos[2] + rt.m[12 + r];
    // camera.hpp:108-120: glm::perspective (right-handed, depth zero-to-one), Y flipped
    M4 P{};
    float th = std::tan(fov_deg * deg2rad / 2.0f);
    P.m[0] = 1.0f / (aspect * th);
    P.m[5] = 1.0f / th;
    P.m[10] = zfar / (znear - zfar);
    P.m[11] = -1.0f;
    P.m[14] = -(zfar * znear) / (zfar - znear);
    P.m[5] *= -1.0f;
    // RayGen.cpp:93: screenToWorld = inverse(P * V)
    M4 s2w = inverse(mul(P, view));
    std::memcpy(out.view, view.m, sizeof(view.m));
    std::memcpy(out.perspective, P.m, sizeof(P.m));
    std::memcpy(out.screen_to_world, s2w.m, sizeof(s2w.m));
}

// PixelTable::recalculate (Ray/PixelTable.cpp:21-126).  The slot order:
//   1. the 8x8 blocks of the 8-aligned part of the image (bw x bh), in
//      ascending 2-D Morton code of the block coordinates (x in the even
//      bits), and inside each block its 64 pixels in ascending Morton code
//      of the in-block coordinates;
//   2. the bottom stripe (rows bh..height-1 under the aligned part), column
//      by column;
//   3. the right stripe (columns bw..width-1, with the corner), row by row.
// i2p[slot] = pixel (y * width + x) and p2i is its inverse.
namespace {
uint32_t spread_even_bits(uint32_t v) {   // bit k of v -> bit 2k
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
uint32_t gather_even_bits(uint32_t v) {   // bit 2k of v -> bit k
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0f0f0f0fu;
    v = (v | (v >> 4)) & 0x00ff00ffu;
    v = (v | (v >> 8)) & 0x0000ffffu;
    return v;
}
}  // namespace

void pixel_table(int width, int height, std::vector<int32_t>& i2p, std::vector<int32_t>& p2i) {
    const size_t npx = (size_t)width * height;
    i2p.assign(npx, 0);
    p2i.assign(npx, 0);
    int32_t slot = 0;
    auto emit = [&](int x, int y) {
        const int32_t pos = y * width + x;
        i2p[slot] = pos;
        p2i[pos] = slot++;
    };
    const int bw = width & ~7, bh = height & ~7;
    std::vector<uint64_t> blocks;              // Morton code << 32 | block x << 16 | block y
    blocks.reserve((size_t)(bw / 8) * (bh / 8));
    for (uint32_t by = 0; by < (uint32_t)bh / 8; ++by)
        for (uint32_t bx = 0; bx < (uint32_t)bw / 8; ++bx)
            blocks.push_back((uint64_t)(spread_even_bits(bx) | spread_even_bits(by) << 1) << 32 |
                             bx << 16 | by);
    std::sort(blocks.begin(), blocks.end());
    for (uint64_t b : blocks) {
        const int x0 = 8 * (int)((b >> 16) & 0xffffu), y0 = 8 * (int)(b & 0xffffu);
        for (uint32_t k = 0; k < 64; ++k)
            emit(x0 + (int)gather_even_bits(k), y0 + (int)gather_even_bits(k >> 1));
    }
    for (int x = 0; x < bw; ++x)
        for (int y = bh; y < height; ++y) emit(x, y);
    for (int y = 0; y < height; ++y)
        for (int x = bw; x < width; ++x) emit(x, y);
}

// ------------------------------------------------------- procedural scene
namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float u01() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
    float uni(float a, float b) { return a + (b - a) * u01(); }
    double normal() {
        double u1 = ((next() >> 11) + 1.0) * (1.0 / 9007199254740993.0);
        double u2 = (next() >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Builder {
    HostScene& s;
    explicit Builder(HostScene& sc) : s(sc) {}
    uint32_t vertex(F3 p, F3 n) {
        uint32_t i = (uint32_t)(s.vertices.size() / 6);
        s.vertices.insert(s.vertices.end(), {p.x, p.y, p.z, n.x, n.y, n.z});
        return i;
    }
    void tri(uint32_t a, uint32_t b, uint32_t c, int mat) {
        s.indices.insert(s.indices.end(), {a, b, c});
        s.tri_material.push_back(mat);
    }
    void begin_geom() { s.geom_first.push_back((int32_t)(s.indices.size() / 3)); }
    // nu x nv grid of quads spanning p0 + [0,1]eu + [0,1]ev with normal nrm.
    void grid(F3 p0, F3 eu, F3 ev, int nu, int nv, F3 nrm, int mat) {
        uint32_t base = (uint32_t)(s.vertices.size() / 6);
        for (int j = 0; j <= nv; ++j)
            for (int i = 0; i <= nu; ++i)
                vertex(add(p0, add(scale(eu, (float)i / nu), scale(ev, (float)j / nv))), nrm);
        auto id = [&](int i, int j) { return base + (uint32_t)(j * (nu + 1) + i); };
        for (int j = 0; j < nv; ++j)
            for (int i = 0; i < nu; ++i) {
                tri(id(i, j), id(i + 1, j), id(i + 1, j + 1), mat);
                tri(id(i, j), id(i + 1, j + 1), id(i, j + 1), mat);
            }
    }
    // Axis-aligned box with outward faces, each an n x n grid.
    void box(F3 c, F3 h, int n, int mat) {
        F3 X{2 * h.x, 0, 0}, Y{0, 2 * h.y, 0}, Z{0, 0, 2 * h.z};
        grid({c.x + h.x, c.y - h.y, c.z - h.z}, Y, Z, n, n, {1, 0, 0}, mat);
        grid({c.x - h.x, c.y - h.y, c.z - h.z}, Z, Y, n, n, {-1, 0, 0}, mat);
        grid({c.x - h.x, c.y + h.y, c.z - h.z}, Z, X, n, n, {0, 1, 0}, mat);
        grid({c.x - h.x, c.y - h.y, c.z - h.z}, X, Z, n, n, {0, -1, 0}, mat);
        grid({c.x - h.x, c.y - h.y, c.z + h.z}, X, Y, n, n, {0, 0, 1}, mat);
        grid({c.x - h.x, c.y - h.y, c.z - h.z}, Y, X, n, n, {0, 0, -1}, mat);
    }
    // Vertical cylinder (no caps): seg around, rings along the height.
    void cylinder(F3 base_c, float r, float height, int seg, int rings, int mat) {
        uint32_t b = (uint32_t)(s.vertices.size() / 6);
        for (int j = 0; j <= rings; ++j)
            for (int i = 0; i < seg; ++i) {
                float a = 6.2831853f * (float)i / seg;
                F3 nrm{std::cos(a), 0, std::sin(a)};
                vertex({base_c.x + r * nrm.x, base_c.y + height * (float)j / rings, base_c.z + r * nrm.z}, nrm);
            }
        for (int j = 0; j < rings; ++j)
            for (int i = 0; i < seg; ++i) {
                uint32_t i1 = (uint32_t)((i + 1) % seg);
                uint32_t a = b + j * seg + i, bb = b + j * seg + i1;
                uint32_t c = b + (j + 1) * seg + i1, d = b + (j + 1) * seg + i;
                tri(a, bb, c, mat);
                tri(a, c, d, mat);
            }
    }
    // Half torus in the x-y plane over the arc angle [0, pi].
    void half_torus(F3 c, float R, float r, int m, int nt, int mat) {
        uint32_t b = (uint32_t)(s.vertices.size() / 6);
        for (int i = 0; i <= m; ++i) {
            float u = 3.14159265f * (float)i / m;
            F3 ring{std::cos(u), std::sin(u), 0};
            for (int j = 0; j < nt; ++j) {
                float v = 6.2831853f * (float)j / nt;
                F3 nrm{ring.x * std::cos(v), ring.y * std::cos(v), std::sin(v)};
                vertex({c.x + R * ring.x + r * nrm.x, c.y + R * ring.y + r * nrm.y, c.z + r * nrm.z}, nrm);
            }
        }
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < nt; ++j) {
                uint32_t j1 = (uint32_t)((j + 1) % nt);
                uint32_t a = b + i * nt + j, bb = b + (i + 1) * nt + j;
                uint32_t cc = b + (i + 1) * nt + j1, d = b + i * nt + j1;
                tri(a, bb, cc, mat);
                tri(a, cc, d, mat);
            }
    }
    void random_tri(SplitMix64& g, F3 c, float size, int mat) {
        F3 p[3];
        for (int k = 0; k < 3; ++k) {
            F3 d{(float)g.normal(), (float)g.normal(), (float)g.normal()};
            p[k] = add(c, scale(norm(d), size * 0.57735f));
        }
        F3 nrm = norm(cross(add(p[1], scale(p[0], -1)), add(p[2], scale(p[0], -1))));
        uint32_t a = vertex(p[0], nrm), b = vertex(p[1], nrm), cc = vertex(p[2], nrm);
        tri(a, b, cc, mat);
    }
    void material(F3 albedo, F3 emission) {
        s.albedo.insert(s.albedo.end(), {albedo.x, albedo.y, albedo.z});
        s.emission.insert(s.emission.end(), {emission.x, emission.y, emission.z});
    }
};

void finish(HostScene& s) {
    s.geom_first.push_back((int32_t)(s.indices.size() / 3));
    for (int a = 0; a < 3; ++a) {
        s.bmin[a] = INFINITY;
        s.bmax[a] = -INFINITY;
    }
    for (size_t t = 0; t < s.indices.size(); ++t) {
        const float* v = &s.vertices[6 * (size_t)s.indices[t]];
        for (int a = 0; a < 3; ++a) {
            s.bmin[a] = std::fmin(s.bmin[a], v[a]);
            s.bmax[a] = std::fmax(s.bmax[a], v[a]);
        }
    }
}

// 64 boxes (centres U[-10,10]^3, edges U(0.2,2), 6 faces x 4x4 x 2 = 192
// triangles, albedo U(0.1,0.9)), a 20x20 ground quad and an emissive quad
// light of radiance 10: 12,292 triangles.
void scene_boxes(uint64_t seed, HostScene& s, int num_boxes = 64) {
    SplitMix64 g{seed};
    Builder b(s);
    for (int k = 0; k < num_boxes; ++k) {
        F3 c{g.uni(-10, 10), g.uni(-10, 10), g.uni(-10, 10)};
        F3 e{g.uni(0.2f, 2.f), g.uni(0.2f, 2.f), g.uni(0.2f, 2.f)};
        F3 alb{g.uni(0.1f, 0.9f), g.uni(0.1f, 0.9f), g.uni(0.1f, 0.9f)};
        b.material(alb, {0, 0, 0});
        b.begin_geom();
        b.box(c, scale(e, 0.5f), 4, k);
    }
    b.material({0.5f, 0.5f, 0.5f}, {0, 0, 0});
    b.begin_geom();
    b.grid({-10, -11, -10}, {0, 0, 20}, {20, 0, 0}, 1, 1, {0, 1, 0}, num_boxes);
    b.material({0, 0, 0}, {10, 10, 10});
    b.begin_geom();
    b.grid({-2, 12, -2}, {4, 0, 0}, {0, 0, 4}, 1, 1, {0, -1, 0}, num_boxes + 1);
    finish(s);
}

// Interior: 60x25x35 room, 2 rows of 16 columns (32 segments x 64 rings),
// 24 half-torus arches, 40,000 clutter triangles with edge U(0.05,0.5),
// 8 albedo materials, 3 emissive ceiling quads: 263,250 triangles.
void scene_interior(uint64_t seed, HostScene& s) {
    SplitMix64 g{seed};
    Builder b(s);
    for (int m = 0; m < 8; ++m) b.material({g.uni(0.1f, 0.9f), g.uni(0.1f, 0.9f), g.uni(0.1f, 0.9f)}, {0, 0, 0});
    b.material({0, 0, 0}, {10, 10, 10});  // 8: light
    const float X0 = -30, X1 = 30, Y0 = 0, Y1 = 25, Z0 = -17.5f, Z1 = 17.5f;
    b.begin_geom();  // walls, inward normals
    int wm = (int)(g.next() % 8);
    b.grid({X0, Y0, Z0}, {X1 - X0, 0, 0}, {0, 0, Z1 - Z0}, 1, 1, {0, 1, 0}, wm);
    b.grid({X0, Y1, Z0}, {0, 0, Z1 - Z0}, {X1 - X0, 0, 0}, 1, 1, {0, -1, 0}, wm);
    b.grid({X0, Y0, Z0}, {0, Y1 - Y0, 0}, {X1 - X0, 0, 0}, 1, 1, {0, 0, 1}, wm);
    b.grid({X0, Y0, Z1}, {X1 - X0, 0, 0}, {0, Y1 - Y0, 0}, 1, 1, {0, 0, -1}, wm);
    b.grid({X0, Y0, Z0}, {0, 0, Z1 - Z0}, {0, Y1 - Y0, 0}, 1, 1, {1, 0, 0}, wm);
    b.grid({X1, Y0, Z0}, {0, Y1 - Y0, 0}, {0, 0, Z1 - Z0}, 1, 1, {-1, 0, 0}, wm);
    const float zrow[2] = {-8.f, 8.f};
    for (int r = 0; r < 2; ++r)
        for (int k = 0; k < 16; ++k) {
            b.begin_geom();
            b.cylinder({-27.f + 3.6f * k, Y0, zrow[r]}, 0.6f, Y1 - Y0, 32, 64, (int)(g.next() % 8));
        }
    for (int r = 0; r < 2; ++r)
        for (int k = 0; k < 12; ++k) {
            b.begin_geom();
            b.half_torus({-27.f + 3.6f * k + 1.8f, 18.f, zrow[r]}, 1.8f, 0.3f, 48, 40, (int)(g.next() % 8));
        }
    b.begin_geom();
    for (int k = 0; k < 40000; ++k) {
        F3 c{g.uni(X0 + 0.5f, X1 - 0.5f), g.uni(Y0 + 0.05f, Y1 - 0.5f), g.uni(Z0 + 0.5f, Z1 - 0.5f)};
        b.random_tri(g, c, g.uni(0.05f, 0.5f), (int)(g.next() % 8));
    }
    b.begin_geom();
    for (int k = 0; k < 3; ++k)
        b.grid({-22.f + 20.f * k, Y1 - 0.1f, -2.f}, {4, 0, 0}, {0, 0, 4}, 1, 1, {0, -1, 0}, 8);
    finish(s);
}

// City: 4,000 buildings on a jittered grid (height lognormal(3, 0.5),
// facades of 8 window quads per floor of 3 units), 500,000 thin foliage
// quads, a ground plane; about 2.8M triangles.
void scene_city(uint64_t seed, HostScene& s) {
    SplitMix64 g{seed};
    Builder b(s);
    for (int m = 0; m < 8; ++m) b.material({g.uni(0.1f, 0.9f), g.uni(0.1f, 0.9f), g.uni(0.1f, 0.9f)}, {0, 0, 0});
    b.material({0.2f, 0.5f, 0.15f}, {0, 0, 0});  // 8: foliage
    b.material({0.35f, 0.35f, 0.35f}, {0, 0, 0});  // 9: ground
    const float cell = 12.f;
    const int gx = 64, gz = 63;
    b.begin_geom();
    b.grid({-gx * cell / 2 - 20, 0, -gz * cell / 2 - 20}, {0, 0, gz * cell + 40}, {gx * cell + 40, 0, 0}, 1, 1, {0, 1, 0}, 9);
    int built = 0;
    for (int j = 0; j < gz && built < 4000; ++j)
        for (int i = 0; i < gx && built < 4000; ++i, ++built) {
            float cx = (i - gx / 2 + 0.5f) * cell + g.uni(-2, 2), cz = (j - gz / 2 + 0.5f) * cell + g.uni(-2, 2);
            float w = g.uni(5, 9), d = g.uni(5, 9);
            float h = (float)std::exp(3.0 + 0.5 * g.normal());
            int floors = std::max(1, (int)(h / 3.0f));
            int mat = (int)(g.next() % 8);
            b.begin_geom();
            F3 p{cx - w / 2, 0, cz - d / 2};
            b.grid({p.x + w, 0, p.z}, {0, h, 0}, {0, 0, d}, floors, 8, {1, 0, 0}, mat);
            b.grid(p, {0, 0, d}, {0, h, 0}, 8, floors, {-1, 0, 0}, mat);
            b.grid({p.x, 0, p.z + d}, {w, 0, 0}, {0, h, 0}, 8, floors, {0, 0, 1}, mat);
            b.grid(p, {0, h, 0}, {w, 0, 0}, floors, 8, {0, 0, -1}, mat);
            b.grid({p.x, h, p.z}, {0, 0, d}, {w, 0, 0}, 1, 1, {0, 1, 0}, mat);
        }
    b.begin_geom();
    for (int k = 0; k < 500000; ++k) {
        F3 c{g.uni(-gx * cell / 2, gx * cell / 2), g.uni(0.f, 3.f), g.uni(-gz * cell / 2, gz * cell / 2)};
        float len = g.uni(0.2f, 1.0f), wid = g.uni(0.02f, 0.05f);
        F3 a = norm({(float)g.normal(), (float)g.normal() * 0.3f, (float)g.normal()});
        F3 up = norm(cross(a, {0.3f, 1.f, 0.2f}));
        F3 nrm = norm(cross(a, up));
        b.grid(add(c, add(scale(a, -len / 2), scale(up, -wid / 2))), scale(a, len), scale(up, wid), 1, 1, nrm, 8);
    }
    finish(s);
}

}  // namespace

int procedural_scene(int kind, uint64_t seed, HostScene& out, int num_boxes) {
    out = HostScene();
    switch (kind) {
        case WF_SCENE_BOXES:
            if (num_boxes < 1 || num_boxes > (1 << 20))
                return fail(WF_ERR_INVALID_ARG, "box count must be in [1, 2^20]");
            scene_boxes(seed, out, num_boxes);
            break;
        case WF_SCENE_INTERIOR: scene_interior(seed, out); break;
        case WF_SCENE_CITY: scene_city(seed, out); break;
        default: return fail(WF_ERR_INVALID_ARG, "procedural scene kind must be 0, 1 or 2");
    }
    return WF_OK;
}

// Emissive materials keep their own slots (first); the non-emissive
// triangles, in order, are cut into groups - n_emissive contiguous runs, each
// with the albedo of its first triangle's original material.
int regroup_materials(HostScene& s, int groups) {
    const int nm = (int)(s.albedo.size() / 3);
    std::vector<int> emissive_slot(nm, -1);
    std::vector<float> alb, em;
    int ne = 0;
    for (int m = 0; m < nm; ++m)
        if (s.emission[3 * m] > 0 || s.emission[3 * m + 1] > 0 || s.emission[3 * m + 2] > 0) {
            emissive_slot[m] = ne++;
            alb.insert(alb.end(), {s.albedo[3 * m], s.albedo[3 * m + 1], s.albedo[3 * m + 2]});
            em.insert(em.end(), {s.emission[3 * m], s.emission[3 * m + 1], s.emission[3 * m + 2]});
        }
    const int g = groups - ne;
    if (g < 1) return fail(WF_ERR_INVALID_ARG, "material_groups must exceed the emissive materials");
    int64_t plain = 0;
    for (int32_t m : s.tri_material)
        if (m < 0 || m >= nm || emissive_slot[m] < 0) ++plain;
    alb.resize(3 * (size_t)groups, 0.8f);
    em.resize(3 * (size_t)groups, 0.0f);
    std::vector<char> seen(g, 0);
    int64_t rank = 0;
    for (auto& m : s.tri_material) {
        if (m >= 0 && m < nm && emissive_slot[m] >= 0) {
            m = emissive_slot[m];
            continue;
        }
        const int grp = (int)(rank * g / std::max<int64_t>(plain, 1));
        ++rank;
        const int slot = ne + grp;
        if (!seen[grp]) {
            seen[grp] = 1;
            const int src = (m >= 0 && m < nm) ? m : 0;
            for (int c = 0; c < 3; ++c) alb[3 * slot + c] = s.albedo[3 * src + c];
        }
        m = slot;
    }
    s.albedo = alb;
    s.emission = em;
    return WF_OK;
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int wf_pixel_table(int width, int height, int32_t* index_to_pixel,
                              int32_t* pixel_to_index) {
    WF_CHECK_ARG(width > 0 && height > 0, "wf_pixel_table: empty image");
    std::vector<int32_t> a, b;
    pixel_table(width, height, a, b);
    if (index_to_pixel) std::memcpy(index_to_pixel, a.data(), a.size() * 4);
    if (pixel_to_index) std::memcpy(pixel_to_index, b.data(), b.size() * 4);
    return WF_OK;
}

extern "C" int wf_camera_matrices(const float position[3], const float rotation_deg[3],
                                  float fov_deg, float aspect, float znear, float zfar,
                                  float view[16], float perspective[16],
                                  float screen_to_world[16]) {
    WF_CHECK_ARG(position && rotation_deg, "wf_camera_matrices: null position/rotation");
    WF_CHECK_ARG(aspect > 0 && znear > 0 && zfar > znear && fov_deg > 0 && fov_deg < 180,
                 "wf_camera_matrices: bad projection parameters");
    CameraMatrices m;
    camera_matrices(position, rotation_deg, fov_deg, aspect, znear, zfar, m);
    if (view) std::memcpy(view, m.view, 64);
    if (perspective) std::memcpy(perspective, m.perspective, 64);
    if (screen_to_world) std::memcpy(screen_to_world, m.screen_to_world, 64);
    return WF_OK;
}

static int export_scene(const HostScene& s, int64_t* num_vertices, int64_t* num_tris,
                        int* num_geoms, int* num_materials, float* vertices, uint32_t* indices,
                        int32_t* geom_first_tri, int32_t* tri_material, float* albedo,
                        float* emission, float scene_min[3], float scene_max[3]);

extern "C" int wf_procedural_boxes(uint64_t seed, int num_boxes, int64_t* num_vertices,
                                   int64_t* num_tris, int* num_geoms, int* num_materials,
                                   float* vertices, uint32_t* indices, int32_t* geom_first_tri,
                                   int32_t* tri_material, float* albedo, float* emission,
                                   float scene_min[3], float scene_max[3]) {
    HostScene s;
    WF_TRY(procedural_scene(WF_SCENE_BOXES, seed, s, num_boxes));
    return export_scene(s, num_vertices, num_tris, num_geoms, num_materials, vertices, indices,
                        geom_first_tri, tri_material, albedo, emission, scene_min, scene_max);
}

extern "C" int wf_procedural_scene(int kind, uint64_t seed, int material_groups,
                                   int64_t* num_vertices, int64_t* num_tris, int* num_geoms,
                                   int* num_materials,
                                   float* vertices, uint32_t* indices, int32_t* geom_first_tri,
                                   int32_t* tri_material, float* albedo, float* emission,
                                   float scene_min[3], float scene_max[3]) {
    HostScene s;
    WF_TRY(procedural_scene(kind, seed, s));
    WF_CHECK_ARG(material_groups >= 0, "wf_procedural_scene: material_groups < 0");
    if (material_groups > 0) WF_TRY(regroup_materials(s, material_groups));
    return export_scene(s, num_vertices, num_tris, num_geoms, num_materials, vertices, indices,
                        geom_first_tri, tri_material, albedo, emission, scene_min, scene_max);
}

static int export_scene(const HostScene& s, int64_t* num_vertices, int64_t* num_tris,
                        int* num_geoms, int* num_materials, float* vertices, uint32_t* indices,
                        int32_t* geom_first_tri, int32_t* tri_material, float* albedo,
                        float* emission, float scene_min[3], float scene_max[3]) {
    if (num_vertices) *num_vertices = (int64_t)(s.vertices.size() / 6);
    if (num_tris) *num_tris = (int64_t)(s.indices.size() / 3);
    if (num_geoms) *num_geoms = (int)s.geom_first.size() - 1;
    if (num_materials) *num_materials = (int)(s.albedo.size() / 3);
    if (vertices) std::memcpy(vertices, s.vertices.data(), s.vertices.size() * 4);
    if (indices) std::memcpy(indices, s.indices.data(), s.indices.size() * 4);
    if (geom_first_tri) std::memcpy(geom_first_tri, s.geom_first.data(), s.geom_first.size() * 4);
    if (tri_material) std::memcpy(tri_material, s.tri_material.data(), s.tri_material.size() * 4);
    if (albedo) std::memcpy(albedo, s.albedo.data(), s.albedo.size() * 4);
    if (emission) std::memcpy(emission, s.emission.data(), s.emission.size() * 4);
    if (scene_min) std::memcpy(scene_min, s.bmin, 12);
    if (scene_max) std::memcpy(scene_max, s.bmax, 12);
    return WF_OK;
}
