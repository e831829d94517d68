// scene.cpp -- seeded synthetic meshes and camera matrices (host code).
//
// Stand-ins for Mesh::LoadSphere / Mesh::LoadPlane (EDXRaster/Utils/Mesh.cpp:36-70),
// which delegate to EDXUtil's ObjMesh (not available).  Geometry is computed in
// double and rounded once to fp32; random numbers come from SplitMix64 so every
// consumer (HIP pipeline, CPU reference, tests, benchmark) sees identical bytes.
// Winding: front faces are clockwise on screen (det > 0 in
// RasterTriangle::Setup, EDXRaster/Core/RasterTriangle.h:49-51).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/edxraster.h"

namespace {

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double u01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
inline float u01f(uint64_t& s) { return (float)((splitmix64(s) >> 40) * (1.0 / 16777216.0)); }

bool alloc_mesh(edxr_host_mesh* m, uint64_t V, uint64_t T) {
    memset(m, 0, sizeof(*m));
    if (V > 0xFFFFFFFFull || T > 0x55555555ull) return false;
    float** planes[9] = {&m->px, &m->py, &m->pz, &m->nx, &m->ny, &m->nz, &m->cr, &m->cg, &m->cb};
    for (auto p : planes) {
        *p = (float*)malloc(sizeof(float) * (V ? V : 1));
        if (!*p) return false;
    }
    m->indices = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (T ? T : 1));
    if (!m->indices) return false;
    m->num_vertices = (uint32_t)V;
    m->num_triangles = (uint32_t)T;
    return true;
}

void fill_colors(edxr_host_mesh* m, uint64_t seed) {
    uint64_t s = seed;
    for (uint32_t i = 0; i < m->num_vertices; i++) {
        m->cr[i] = u01f(s);
        m->cg[i] = u01f(s);
        m->cb[i] = u01f(s);
    }
}

// Lattice value in [-1, 1] for octave o at integer point (ix, iz).
inline double lattice(uint64_t seed, int o, int64_t ix, int64_t iz) {
    uint64_t s = seed ^ (0x632BE59BD9B4E019ull * (uint64_t)(o + 1)) ^ (0x8CB92BA72F3D8DD7ull * (uint64_t)ix) ^
                 (0xD6E8FEB86659FD93ull * (uint64_t)iz);
    splitmix64(s);
    return u01(s) * 2.0 - 1.0;
}
inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

double value_noise(uint64_t seed, int o, double x, double z) {
    double fx = std::floor(x), fz = std::floor(z);
    int64_t ix = (int64_t)fx, iz = (int64_t)fz;
    double tx = smooth(x - fx), tz = smooth(z - fz);
    double a = lattice(seed, o, ix, iz), b = lattice(seed, o, ix + 1, iz);
    double c = lattice(seed, o, ix, iz + 1), d = lattice(seed, o, ix + 1, iz + 1);
    double ab = a + (b - a) * tx, cd = c + (d - c) * tx;
    return ab + (cd - ab) * tz;
}

void build_sphere(edxr_host_mesh* m, uint32_t voff, uint32_t toff, uint32_t slices, uint32_t stacks, double radius,
                  const double c[3]) {
    const uint32_t row = slices + 1;
    for (uint32_t i = 0; i <= stacks; i++) {
        const double theta = M_PI * (double)i / (double)stacks;
        for (uint32_t j = 0; j <= slices; j++) {
            const double phi = 2.0 * M_PI * (double)j / (double)slices;
            const double n[3] = {std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi)};
            const uint32_t v = voff + i * row + j;
            m->px[v] = (float)(c[0] + radius * n[0]);
            m->py[v] = (float)(c[1] + radius * n[1]);
            m->pz[v] = (float)(c[2] + radius * n[2]);
            m->nx[v] = (float)n[0];
            m->ny[v] = (float)n[1];
            m->nz[v] = (float)n[2];
        }
    }
    uint32_t t = toff;
    for (uint32_t i = 0; i < stacks; i++)
        for (uint32_t j = 0; j < slices; j++) {
            const uint32_t a = voff + i * row + j, b = a + 1, cc = a + row, d = cc + 1;
            uint32_t* I = m->indices + 3ull * t;
            I[0] = a; I[1] = b; I[2] = cc;
            I[3] = b; I[4] = d; I[5] = cc;
            t += 2;
        }
}

}  // namespace

extern "C" {

int edxr_scene_sphere(edxr_host_mesh* out, uint32_t slices, uint32_t stacks, float radius, const float center[3],
                      uint64_t color_seed) {
    if (!out || slices < 3 || stacks < 2) return EDXR_E_INVALID;
    const uint64_t V = (uint64_t)(stacks + 1) * (slices + 1), T = 2ull * slices * stacks;
    if (!alloc_mesh(out, V, T)) return EDXR_E_NOMEM;
    const double c[3] = {center ? center[0] : 0.0, center ? center[1] : 0.0, center ? center[2] : 0.0};
    build_sphere(out, 0, 0, slices, stacks, radius, c);
    fill_colors(out, color_seed);
    return EDXR_OK;
}

int edxr_scene_heightfield(edxr_host_mesh* out, uint32_t n, float amplitude, uint64_t height_seed,
                           uint64_t color_seed) {
    if (!out || n < 2) return EDXR_E_INVALID;
    const uint64_t V = (uint64_t)n * n, T = 2ull * (n - 1) * (n - 1);
    if (!alloc_mesh(out, V, T)) return EDXR_E_NOMEM;
    std::vector<double> h(V);
    double wsum = 0.0;
    for (int o = 0; o < 8; o++) wsum += std::ldexp(1.0, -o);
    const double step = 2.0 / (double)(n - 1);
    auto rows = [&](uint32_t r0, uint32_t r1) {
        for (uint32_t i = r0; i < r1; i++)
            for (uint32_t j = 0; j < n; j++) {
                const double x = -1.0 + step * (double)j, z = -1.0 + step * (double)i;
                double s = 0.0;
                for (int o = 0; o < 8; o++) {
                    const double f = 4.0 * std::ldexp(1.0, o);
                    s += std::ldexp(1.0, -o) * value_noise(height_seed, o, f * (x + 1.0), f * (z + 1.0));
                }
                h[(size_t)i * n + j] = (double)amplitude * s / wsum;
            }
    };
    {
        const uint32_t nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (uint32_t k = 0; k < nt; k++) pool.emplace_back(rows, (uint32_t)((uint64_t)n * k / nt), (uint32_t)((uint64_t)n * (k + 1) / nt));
        for (auto& th : pool) th.join();
    }
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < n; j++) {
            const size_t v = (size_t)i * n + j;
            out->px[v] = (float)(-1.0 + step * (double)j);
            out->pz[v] = (float)(-1.0 + step * (double)i);
            out->py[v] = (float)h[v];
            const uint32_t j0 = j ? j - 1 : j, j1 = j + 1 < n ? j + 1 : j;
            const uint32_t i0 = i ? i - 1 : i, i1 = i + 1 < n ? i + 1 : i;
            const double dhdx = (h[(size_t)i * n + j1] - h[(size_t)i * n + j0]) / (step * (double)(j1 - j0));
            const double dhdz = (h[(size_t)i1 * n + j] - h[(size_t)i0 * n + j]) / (step * (double)(i1 - i0));
            const double len = std::sqrt(dhdx * dhdx + 1.0 + dhdz * dhdz);
            out->nx[v] = (float)(-dhdx / len);
            out->ny[v] = (float)(1.0 / len);
            out->nz[v] = (float)(-dhdz / len);
        }
    uint64_t t = 0;
    for (uint32_t i = 0; i + 1 < n; i++)
        for (uint32_t j = 0; j + 1 < n; j++) {
            const uint32_t a = i * n + j, b = a + 1, c = a + n, d = c + 1;
            uint32_t* I = out->indices + 3 * t;
            I[0] = a; I[1] = c; I[2] = b;
            I[3] = b; I[4] = c; I[5] = d;
            t += 2;
        }
    fill_colors(out, color_seed);
    return EDXR_OK;
}

int edxr_scene_sphere_grid(edxr_host_mesh* out, uint32_t grid, uint32_t segments, uint64_t seed) {
    if (!out || grid == 0 || segments < 3) return EDXR_E_INVALID;
    const uint64_t vs = (uint64_t)(segments + 1) * (segments + 1), ts = 2ull * segments * segments;
    const uint64_t count = (uint64_t)grid * grid;
    if (!alloc_mesh(out, vs * count, ts * count)) return EDXR_E_NOMEM;
    uint64_t s = seed;
    for (uint32_t i = 0; i < grid; i++)
        for (uint32_t j = 0; j < grid; j++) {
            const uint64_t k = (uint64_t)i * grid + j;
            const double radius = 0.5 + 1.5 * u01(s);
            const double c[3] = {5.0 * i - 2.5 * (grid - 1), 0.0, 5.0 * j - 2.5 * (grid - 1)};
            build_sphere(out, (uint32_t)(vs * k), (uint32_t)(ts * k), segments, segments, radius, c);
            for (uint64_t v = vs * k; v < vs * (k + 1); v++) {
                out->cr[v] = u01f(s);
                out->cg[v] = u01f(s);
                out->cb[v] = u01f(s);
            }
        }
    return EDXR_OK;
}

int edxr_scene_merge(edxr_host_mesh* a, const edxr_host_mesh* b) {
    if (!a || !b) return EDXR_E_INVALID;
    edxr_host_mesh m;
    const uint64_t V = (uint64_t)a->num_vertices + b->num_vertices;
    const uint64_t T = (uint64_t)a->num_triangles + b->num_triangles;
    if (!alloc_mesh(&m, V, T)) return EDXR_E_NOMEM;
    float* const* pa[9] = {&a->px, &a->py, &a->pz, &a->nx, &a->ny, &a->nz, &a->cr, &a->cg, &a->cb};
    float* const* pb[9] = {&b->px, &b->py, &b->pz, &b->nx, &b->ny, &b->nz, &b->cr, &b->cg, &b->cb};
    float** pm[9] = {&m.px, &m.py, &m.pz, &m.nx, &m.ny, &m.nz, &m.cr, &m.cg, &m.cb};
    for (int k = 0; k < 9; k++) {
        memcpy(*pm[k], *pa[k], sizeof(float) * a->num_vertices);
        memcpy(*pm[k] + a->num_vertices, *pb[k], sizeof(float) * b->num_vertices);
    }
    memcpy(m.indices, a->indices, 12ull * a->num_triangles);
    for (uint64_t i = 0; i < 3ull * b->num_triangles; i++)
        m.indices[3ull * a->num_triangles + i] = b->indices[i] + a->num_vertices;
    edxr_scene_free(a);
    *a = m;
    return EDXR_OK;
}

void edxr_scene_free(edxr_host_mesh* m) {
    if (!m) return;
    float** planes[9] = {&m->px, &m->py, &m->pz, &m->nx, &m->ny, &m->nz, &m->cr, &m->cg, &m->cb};
    for (auto p : planes) {
        free(*p);
        *p = nullptr;
    }
    free(m->indices);
    m->indices = nullptr;
    free(m->tu);
    free(m->tv);
    free(m->tex_ids);
    m->tu = m->tv = nullptr;
    m->tex_ids = nullptr;
    m->num_vertices = m->num_triangles = 0;
}

void edxr_look_at(const float eye[3], const float at[3], const float up[3], float out[16]) {
    double z[3] = {(double)at[0] - eye[0], (double)at[1] - eye[1], (double)at[2] - eye[2]};
    double l = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    for (double& v : z) v /= l;
    // x = up x z (left-handed basis: x right, y up, z forward)
    double x[3] = {up[1] * z[2] - up[2] * z[1], up[2] * z[0] - up[0] * z[2], up[0] * z[1] - up[1] * z[0]};
    l = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (double& v : x) v /= l;
    const double y[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
    const double e[3] = {eye[0], eye[1], eye[2]};
    const double m[16] = {x[0], x[1], x[2], -(x[0] * e[0] + x[1] * e[1] + x[2] * e[2]),
                          y[0], y[1], y[2], -(y[0] * e[0] + y[1] * e[1] + y[2] * e[2]),
                          z[0], z[1], z[2], -(z[0] * e[0] + z[1] * e[1] + z[2] * e[2]),
                          0.0,  0.0,  0.0,  1.0};
    for (int i = 0; i < 16; i++) out[i] = (float)m[i];
}

void edxr_perspective(float fov_y_deg, float aspect, float znear, float zfar, float out[16]) {
    const double ys = 1.0 / std::tan((double)fov_y_deg * M_PI / 360.0), xs = ys / (double)aspect;
    const double q = (double)zfar / ((double)zfar - (double)znear);
    const double m[16] = {xs, 0, 0, 0, 0, ys, 0, 0, 0, 0, q, -q * (double)znear, 0, 0, 1, 0};
    for (int i = 0; i < 16; i++) out[i] = (float)m[i];
}

void edxr_raster_matrix(uint32_t width, uint32_t height, float out[16]) {
    const double hw = 0.5 * width, hh = 0.5 * height;
    const double m[16] = {hw, 0, 0, hw, 0, -hh, 0, hh, 0, 0, 1, 0, 0, 0, 0, 1};
    for (int i = 0; i < 16; i++) out[i] = (float)m[i];
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Wavefront OBJ input: Mesh::LoadMesh (EDXRaster/Utils/Mesh.cpp:11-34) through
// EDXUtil's ObjMesh::LoadFromObj, which is not available.  Declared semantics:
//   - 'v', 'vn', 'vt', 'f' (v, v/t, v//n, v/t/n; 1-based or negative
//     relative indices), 'mtllib' and 'usemtl'; the .mtl file supplies
//     'newmtl', 'Kd' (the material colour, ConstantTexture2D in the
//     reference) and 'map_Kd' (its image, ImageTexture); other statements
//     are skipped;
//   - a polygon face is split into a fan (0, i, i + 1), keeping the file's
//     vertex order (front faces of a right-handed export are clockwise on
//     screen in the left-handed, y-down raster space);
//   - one output vertex per distinct (position, texture coordinate, normal,
//     material); faces without normals get area-weighted vertex normals of
//     their position (computed in double); missing texture coordinates are
//     (0, 0);
//   - world transform p' = T(pos) Rz Ry Rx S(scl) p (rotation in degrees),
//     normals n' = normalize(Rz Ry Rx (n / scl));
//   - vertex colour = the face material's Kd, (0.9, 0.9, 0.9) without one
//     (the constant albedo Mesh::LoadPlane gives, Utils/Mesh.cpp:46).
// ---------------------------------------------------------------------------
namespace {

// The placement of a loaded mesh (ObjMesh::LoadFromObj / LoadSphere /
// LoadPlane take pos, scl, rot): p' = T(pos) Rz Ry Rx S(scl) p with the
// rotation in degrees, n' = normalize(Rz Ry Rx (n / scl)), in double.
struct Placement {
    double R[3][3], pos[3], scl[3];
    Placement(const float p[3], const float s[3], const float r[3]) {
        const double d2r = 3.14159265358979323846 / 180.0;
        const double cx = cos(r[0] * d2r), sx = sin(r[0] * d2r), cy = cos(r[1] * d2r), sy = sin(r[1] * d2r),
                     cz = cos(r[2] * d2r), sz = sin(r[2] * d2r);
        const double M[3][3] = {{cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx},
                                {sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx},
                                {-sy, cy * sx, cy * cx}};
        memcpy(R, M, sizeof(R));
        for (int i = 0; i < 3; i++) {
            pos[i] = p[i];
            scl[i] = s[i];
        }
    }
    // Places model-space (p, n) and stores it as vertex v of out.
    void apply(const double p[3], const double n[3], edxr_host_mesh* out, uint32_t v) const {
        double sp[3], sn[3], wp[3], wn[3];
        for (int j = 0; j < 3; j++) {
            sp[j] = p[j] * scl[j];
            sn[j] = n[j] / scl[j];
        }
        for (int i = 0; i < 3; i++) {
            wp[i] = R[i][0] * sp[0] + R[i][1] * sp[1] + R[i][2] * sp[2] + pos[i];
            wn[i] = R[i][0] * sn[0] + R[i][1] * sn[1] + R[i][2] * sn[2];
        }
        const double len = sqrt(wn[0] * wn[0] + wn[1] * wn[1] + wn[2] * wn[2]);
        const double inv = len > 0.0 ? 1.0 / len : 0.0;
        out->px[v] = (float)wp[0];
        out->py[v] = (float)wp[1];
        out->pz[v] = (float)wp[2];
        out->nx[v] = (float)(wn[0] * inv);
        out->ny[v] = (float)(wn[1] * inv);
        out->nz[v] = (float)(wn[2] * inv);
    }
};

// Texture coordinates, white vertex colours and texture slot 0 for every
// triangle: the vertex and texture layout of a generated Mesh
// (Vertex_PositionNormalTex, one ConstantTexture2D, Utils/Mesh.cpp:36-70).
bool add_generated_extras(edxr_host_mesh* m) {
    m->tu = (float*)malloc(sizeof(float) * (m->num_vertices ? m->num_vertices : 1));
    m->tv = (float*)malloc(sizeof(float) * (m->num_vertices ? m->num_vertices : 1));
    m->tex_ids = (uint32_t*)calloc(m->num_triangles ? m->num_triangles : 1, sizeof(uint32_t));
    if (!m->tu || !m->tv || !m->tex_ids) return false;
    for (uint32_t v = 0; v < m->num_vertices; v++) m->cr[v] = m->cg[v] = m->cb[v] = 1.0f;
    return true;
}

}  // namespace

extern "C" {

int edxr_load_sphere(const float pos[3], const float scl[3], const float rot[3], float radius, uint32_t slices,
                     uint32_t stacks, edxr_host_mesh* out) {
    if (!out || !pos || !scl || !rot || slices < 3 || stacks < 2) return EDXR_E_INVALID;
    const uint64_t V = (uint64_t)(stacks + 1) * (slices + 1), T = 2ull * slices * stacks;
    if (!alloc_mesh(out, V, T) || !add_generated_extras(out)) {
        edxr_scene_free(out);
        return EDXR_E_NOMEM;
    }
    const double origin[3] = {0.0, 0.0, 0.0};
    build_sphere(out, 0, 0, slices, stacks, radius, origin);
    const Placement X(pos, scl, rot);
    const uint32_t row = slices + 1;
    for (uint32_t v = 0; v < out->num_vertices; v++) {
        const double p[3] = {out->px[v], out->py[v], out->pz[v]}, n[3] = {out->nx[v], out->ny[v], out->nz[v]};
        X.apply(p, n, out, v);
        out->tu[v] = (float)(v % row) / (float)slices;
        out->tv[v] = (float)(v / row) / (float)stacks;
    }
    return EDXR_OK;
}

int edxr_load_plane(const float pos[3], const float scl[3], const float rot[3], float length, edxr_host_mesh* out) {
    if (!out || !pos || !scl || !rot || !(length > 0.0f)) return EDXR_E_INVALID;
    if (!alloc_mesh(out, 4, 2) || !add_generated_extras(out)) {
        edxr_scene_free(out);
        return EDXR_E_NOMEM;
    }
    const double h = 0.5 * (double)length;
    const double P[4][3] = {{-h, 0.0, -h}, {h, 0.0, -h}, {-h, 0.0, h}, {h, 0.0, h}};
    const double n[3] = {0.0, 1.0, 0.0};
    const Placement X(pos, scl, rot);
    for (uint32_t v = 0; v < 4; v++) {
        X.apply(P[v], n, out, v);
        out->tu[v] = (float)(v & 1u);
        out->tv[v] = (float)(v >> 1);
    }
    const uint32_t I[6] = {0, 2, 1, 1, 2, 3};  // clockwise on screen seen from +y
    memcpy(out->indices, I, sizeof(I));
    return EDXR_OK;
}

}  // extern "C"

namespace {

struct ObjMaterial {
    std::string name;
    float kd[3] = {0.9f, 0.9f, 0.9f};
    std::string map;  // diffuse map path as written in the library
};

bool read_file(const std::string& path, std::string& out) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, n);
    fclose(f);
    return true;
}

std::string dir_of(const std::string& path) {
    const size_t s = path.find_last_of("/\\");
    return s == std::string::npos ? std::string() : path.substr(0, s + 1);
}

// Splits text into lines with leading blanks removed and comments cut.
template <class F>
void for_each_line(const std::string& text, F f) {
    size_t i = 0;
    while (i < text.size()) {
        size_t e = text.find('\n', i);
        if (e == std::string::npos) e = text.size();
        std::string line = text.substr(i, e - i);
        i = e + 1;
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t')) line.pop_back();
        size_t b = 0;
        while (b < line.size() && (line[b] == ' ' || line[b] == '\t')) b++;
        if (b < line.size()) f(line.substr(b));
    }
}

void load_mtl(const std::string& path, std::vector<ObjMaterial>& mats) {
    const std::string dir = dir_of(path);
    std::string text;
    if (!read_file(path, text)) return;  // a missing library leaves the default colour
    for_each_line(text, [&](const std::string& l) {
        char key[64] = {0}, name[512] = {0};
        if (sscanf(l.c_str(), "%63s", key) != 1) return;
        if (!strcmp(key, "newmtl") && sscanf(l.c_str(), "%*s %511s", name) == 1) {
            mats.push_back(ObjMaterial());
            mats.back().name = name;
        } else if (!strcmp(key, "map_Kd") && !mats.empty()) {
            char file[512] = {0};
            if (sscanf(l.c_str(), "%*s %511s", file) == 1) mats.back().map = file[0] == '/' ? file : dir + file;
        } else if (!strcmp(key, "Kd") && !mats.empty()) {
            float r, g, b;
            if (sscanf(l.c_str(), "%*s %f %f %f", &r, &g, &b) == 3) {
                mats.back().kd[0] = r;
                mats.back().kd[1] = g;
                mats.back().kd[2] = b;
            }
        }
    });
}

// Resolves an OBJ index (1-based, or negative = relative to the end).
bool obj_index(long v, size_t count, uint32_t& out) {
    if (v > 0 && (size_t)v <= count) {
        out = (uint32_t)(v - 1);
        return true;
    }
    if (v < 0 && (size_t)(-v) <= count) {
        out = (uint32_t)(count + v);
        return true;
    }
    return false;
}

}  // namespace

int edxr_load_obj(const char* path, const float pos[3], const float scl[3], const float rot[3],
                  edxr_host_mesh* out) {
    if (!path || !out) return EDXR_E_INVALID;
    memset(out, 0, sizeof(*out));
    std::string text;
    if (!read_file(path, text)) return EDXR_E_IO;
    std::vector<double> P, N;  // file positions and normals, xyz each
    std::vector<float> UV;     // file texture coordinates, uv each
    std::vector<ObjMaterial> mats;
    struct Corner {
        uint32_t p, t, n, m;  // t, n = UINT32_MAX: none
    };
    std::vector<Corner> corners;  // triangle corners, 3 per triangle
    uint32_t cur = UINT32_MAX;    // current material (UINT32_MAX: default)
    bool bad = false;
    const std::string dir = dir_of(path);
    for_each_line(text, [&](const std::string& l) {
        if (bad) return;
        char key[64] = {0};
        if (sscanf(l.c_str(), "%63s", key) != 1) return;
        if (!strcmp(key, "v") || !strcmp(key, "vn")) {
            double x, y, z;
            if (sscanf(l.c_str(), "%*s %lf %lf %lf", &x, &y, &z) != 3) {
                bad = true;
                return;
            }
            std::vector<double>& d = key[1] ? N : P;
            d.push_back(x);
            d.push_back(y);
            d.push_back(z);
        } else if (!strcmp(key, "vt")) {
            float u, v = 0.0f;
            if (sscanf(l.c_str(), "%*s %f %f", &u, &v) < 1) {
                bad = true;
                return;
            }
            UV.push_back(u);
            UV.push_back(v);
        } else if (!strcmp(key, "mtllib")) {
            char name[512] = {0};
            if (sscanf(l.c_str(), "%*s %511s", name) == 1) load_mtl(dir + name, mats);
        } else if (!strcmp(key, "usemtl")) {
            char name[512] = {0};
            cur = UINT32_MAX;
            if (sscanf(l.c_str(), "%*s %511s", name) == 1)
                for (size_t i = 0; i < mats.size(); i++)
                    if (mats[i].name == name) cur = (uint32_t)i;
        } else if (!strcmp(key, "f")) {
            std::vector<Corner> poly;
            const char* s = l.c_str() + 1;
            while (*s) {
                while (*s == ' ' || *s == '\t') s++;
                if (!*s) break;
                char* e;
                const long vi = strtol(s, &e, 10);
                if (e == s) {
                    bad = true;
                    return;
                }
                s = e;
                long ni = 0, ti = 0;
                if (*s == '/') {
                    s++;
                    if (*s != '/') {
                        ti = strtol(s, &e, 10);
                        if (e == s) {
                            bad = true;
                            return;
                        }
                        s = e;
                    }
                    if (*s == '/') {
                        s++;
                        ni = strtol(s, &e, 10);
                        if (e == s) {
                            bad = true;
                            return;
                        }
                        s = e;
                    }
                }
                Corner c{0, UINT32_MAX, UINT32_MAX, cur};
                if (!obj_index(vi, P.size() / 3, c.p) || (ni && !obj_index(ni, N.size() / 3, c.n)) ||
                    (ti && !obj_index(ti, UV.size() / 2, c.t))) {
                    bad = true;
                    return;
                }
                poly.push_back(c);
            }
            if (poly.size() < 3) {
                bad = true;
                return;
            }
            for (size_t i = 1; i + 1 < poly.size(); i++) {
                corners.push_back(poly[0]);
                corners.push_back(poly[i]);
                corners.push_back(poly[i + 1]);
            }
        }
    });
    if (bad) return EDXR_E_FORMAT;

    // Area-weighted vertex normals for corners without one.
    std::vector<double> AN(P.size(), 0.0);
    for (size_t t = 0; t + 2 < corners.size(); t += 3) {
        const double* a = &P[3 * corners[t].p];
        const double* b = &P[3 * corners[t + 1].p];
        const double* c = &P[3 * corners[t + 2].p];
        const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++) AN[3 * corners[t + k].p + j] += n[j];
    }

    // Distinct (position, normal, material) triples become output vertices,
    // numbered by first appearance in the file.
    std::vector<uint32_t> vid(corners.size());
    std::vector<uint32_t> first;  // corner index of each output vertex
    {
        std::vector<uint32_t> order(corners.size());
        for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            const Corner &x = corners[a], &y = corners[b];
            if (x.p != y.p) return x.p < y.p;
            if (x.t != y.t) return x.t < y.t;
            if (x.n != y.n) return x.n < y.n;
            return x.m < y.m;
        });
        std::vector<uint32_t> rep(corners.size());
        for (size_t k = 0; k < order.size(); k++) {
            const Corner& c = corners[order[k]];
            const Corner& d = corners[order[k ? k - 1 : 0]];
            const bool same = k && d.p == c.p && d.t == c.t && d.n == c.n && d.m == c.m;
            rep[order[k]] = same ? rep[order[k - 1]] : order[k];
        }
        std::vector<uint32_t> idOf(corners.size(), UINT32_MAX);
        for (size_t i = 0; i < corners.size(); i++) {
            if (idOf[rep[i]] == UINT32_MAX) {
                idOf[rep[i]] = (uint32_t)first.size();
                first.push_back((uint32_t)i);
            }
            vid[i] = idOf[rep[i]];
        }
    }
    const size_t T = corners.size() / 3;
    if (!alloc_mesh(out, first.size(), T)) {
        edxr_scene_free(out);
        return EDXR_E_NOMEM;
    }
    out->tu = (float*)malloc(sizeof(float) * (first.size() ? first.size() : 1));
    out->tv = (float*)malloc(sizeof(float) * (first.size() ? first.size() : 1));
    out->tex_ids = (uint32_t*)malloc(sizeof(uint32_t) * (T ? T : 1));
    if (!out->tu || !out->tv || !out->tex_ids) {
        edxr_scene_free(out);
        return EDXR_E_NOMEM;
    }
    for (size_t t = 0; t < T; t++)
        out->tex_ids[t] = corners[3 * t].m != UINT32_MAX ? corners[3 * t].m : (uint32_t)mats.size();
    const Placement X(pos, scl, rot);
    for (size_t v = 0; v < first.size(); v++) {
        const Corner& c = corners[first[v]];
        double p[3], n[3];
        for (int j = 0; j < 3; j++) {
            p[j] = P[3 * c.p + j];
            n[j] = c.n != UINT32_MAX ? N[3 * c.n + j] : AN[3 * c.p + j];
        }
        X.apply(p, n, out, (uint32_t)v);
        static const float kDefaultKd[3] = {0.9f, 0.9f, 0.9f};
        const float* kd = c.m != UINT32_MAX ? mats[c.m].kd : kDefaultKd;
        out->cr[v] = kd[0];
        out->cg[v] = kd[1];
        out->cb[v] = kd[2];
        out->tu[v] = c.t != UINT32_MAX ? UV[2 * c.t] : 0.0f;
        out->tv[v] = c.t != UINT32_MAX ? UV[2 * c.t + 1] : 0.0f;
    }
    for (size_t i = 0; i < corners.size(); i++) out->indices[i] = vid[i];
    return EDXR_OK;
}

// Bitmap::SaveBitmapFile (used by Renderer::WriteFrameToFile,
// EDXRaster/Core/Renderer.cpp:352-358), restated: an uncompressed 24-bit BMP
// (BITMAPINFOHEADER, BI_RGB), rows bottom-up as the back buffer holds them,
// BGR byte order, each row padded to 4 bytes; alpha is dropped.
int edxr_write_bmp(const char* path, const uint8_t* rgba, uint32_t width, uint32_t height) {
    if (!path || !rgba || !width || !height) return EDXR_E_INVALID;
    const uint32_t row = (3 * width + 3) & ~3u;
    const uint64_t image = (uint64_t)row * height;
    if (image + 54 > 0xFFFFFFFFull) return EDXR_E_LIMIT;
    uint8_t h[54] = {0};
    auto put32 = [&](int at, uint32_t v) {
        for (int k = 0; k < 4; k++) h[at + k] = (uint8_t)(v >> (8 * k));
    };
    h[0] = 'B';
    h[1] = 'M';
    put32(2, (uint32_t)(54 + image));
    put32(10, 54);
    put32(14, 40);
    put32(18, width);
    put32(22, height);
    h[26] = 1;   // planes
    h[28] = 24;  // bits per pixel
    put32(34, (uint32_t)image);
    put32(38, 2835);  // 72 dpi
    put32(42, 2835);
    FILE* f = fopen(path, "wb");
    if (!f) return EDXR_E_IO;
    bool ok = fwrite(h, 1, 54, f) == 54;
    std::vector<uint8_t> line(row, 0);
    for (uint32_t y = 0; ok && y < height; y++) {
        const uint8_t* s = rgba + 4ull * width * y;
        for (uint32_t x = 0; x < width; x++) {
            line[3 * x] = s[4 * x + 2];
            line[3 * x + 1] = s[4 * x + 1];
            line[3 * x + 2] = s[4 * x];
        }
        ok = fwrite(line.data(), 1, row, f) == row;
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? EDXR_OK : EDXR_E_IO;
}

int64_t edxr_obj_materials(const char* path, edxr_obj_material* out, uint32_t cap) {
    if (!path || (cap && !out)) return EDXR_E_INVALID;
    std::string text;
    if (!read_file(path, text)) return EDXR_E_IO;
    std::vector<ObjMaterial> mats;
    const std::string dir = dir_of(path);
    for_each_line(text, [&](const std::string& l) {
        char key[64] = {0}, name[512] = {0};
        if (sscanf(l.c_str(), "%63s %511s", key, name) == 2 && !strcmp(key, "mtllib")) load_mtl(dir + name, mats);
    });
    for (size_t i = 0; i < mats.size() && i < cap; i++) {
        memset(&out[i], 0, sizeof(out[i]));
        strncpy(out[i].name, mats[i].name.c_str(), sizeof(out[i].name) - 1);
        for (int k = 0; k < 3; k++) out[i].kd[k] = mats[i].kd[k];
        strncpy(out[i].map_kd, mats[i].map.c_str(), sizeof(out[i].map_kd) - 1);
    }
    return (int64_t)mats.size();
}

// Uncompressed BMP input (BITMAPINFOHEADER or later, BI_RGB, 24 or 32 bits,
// bottom-up rows as stored; a negative height means top-down rows, which are
// reversed so that the result is bottom-up in every case).
int edxr_load_bmp(const char* path, uint32_t* width, uint32_t* height, uint8_t** rgba) {
    if (!path || !width || !height || !rgba) return EDXR_E_INVALID;
    *rgba = nullptr;
    std::string d;
    if (!read_file(path, d)) return EDXR_E_IO;
    auto u16 = [&](size_t at) { return (uint32_t)(uint8_t)d[at] | ((uint32_t)(uint8_t)d[at + 1] << 8); };
    auto u32 = [&](size_t at) { return u16(at) | (u16(at + 2) << 16); };
    if (d.size() < 54 || d[0] != 'B' || d[1] != 'M') return EDXR_E_FORMAT;
    const uint32_t off = u32(10), hdr = u32(14), bpp = u16(28), comp = u32(30);
    const int32_t w = (int32_t)u32(18), h = (int32_t)u32(22);
    if (hdr < 40 || u16(26) != 1 || (bpp != 24 && bpp != 32) || comp != 0 || w <= 0 || h == 0 || w > 16384 ||
        h > 16384 || h < -16384)
        return EDXR_E_FORMAT;
    const uint32_t H = (uint32_t)(h < 0 ? -h : h), W = (uint32_t)w;
    const size_t row = ((size_t)W * (bpp / 8) + 3) & ~(size_t)3;
    if ((size_t)off + row * H > d.size()) return EDXR_E_FORMAT;
    uint8_t* img = (uint8_t*)malloc(4ull * W * H);
    if (!img) return EDXR_E_NOMEM;
    for (uint32_t y = 0; y < H; y++) {
        const uint8_t* src = (const uint8_t*)d.data() + off + row * (h < 0 ? H - 1 - y : y);
        for (uint32_t x = 0; x < W; x++) {
            const uint8_t* p = src + (size_t)x * (bpp / 8);
            uint8_t* q = img + 4ull * ((size_t)y * W + x);
            q[0] = p[2];
            q[1] = p[1];
            q[2] = p[0];
            q[3] = bpp == 32 ? p[3] : 255;
        }
    }
    *width = W;
    *height = H;
    *rgba = img;
    return EDXR_OK;
}

// Truevision TGA: image types 2 (uncompressed true colour) and 10 (run-length
// encoded true colour), 24 or 32 bits per pixel, BGR(A) bytes, no colour map;
// rows bottom-up unless bit 5 of the descriptor (top-left origin) is set, and
// returned bottom-up as the BMP reader returns them.
int edxr_load_tga(const char* path, uint32_t* width, uint32_t* height, uint8_t** rgba) {
    if (!path || !width || !height || !rgba) return EDXR_E_INVALID;
    *rgba = nullptr;
    std::string d;
    if (!read_file(path, d)) return EDXR_E_IO;
    if (d.size() < 18) return EDXR_E_FORMAT;
    const uint8_t* b = (const uint8_t*)d.data();
    const uint32_t idLen = b[0], cmapType = b[1], type = b[2];
    const uint32_t W = b[12] | (b[13] << 8), H = b[14] | (b[15] << 8), bpp = b[16], desc = b[17];
    if (cmapType != 0 || (type != 2 && type != 10) || (bpp != 24 && bpp != 32) || !W || !H || W > 16384 ||
        H > 16384 || (desc & 0x10))
        return EDXR_E_FORMAT;
    const size_t px = bpp / 8, n = (size_t)W * H;
    size_t at = 18 + idLen;
    std::vector<uint8_t> raw(n * px);
    if (type == 2) {
        if (at + raw.size() > d.size()) return EDXR_E_FORMAT;
        memcpy(raw.data(), b + at, raw.size());
    } else {
        for (size_t i = 0; i < n;) {
            if (at >= d.size()) return EDXR_E_FORMAT;
            const uint32_t h = b[at++], count = (h & 0x7F) + 1;
            if (i + count > n) return EDXR_E_FORMAT;
            if (h & 0x80) {  // one pixel repeated
                if (at + px > d.size()) return EDXR_E_FORMAT;
                for (uint32_t k = 0; k < count; k++) memcpy(&raw[(i + k) * px], b + at, px);
                at += px;
            } else {         // count literal pixels
                if (at + count * px > d.size()) return EDXR_E_FORMAT;
                memcpy(&raw[i * px], b + at, count * px);
                at += count * px;
            }
            i += count;
        }
    }
    uint8_t* img = (uint8_t*)malloc(4 * n);
    if (!img) return EDXR_E_NOMEM;
    const bool topDown = (desc & 0x20) != 0;
    for (uint32_t y = 0; y < H; y++) {
        const uint8_t* src = raw.data() + (size_t)(topDown ? H - 1 - y : y) * W * px;
        for (uint32_t x = 0; x < W; x++) {
            const uint8_t* p = src + x * px;
            uint8_t* q = img + 4ull * ((size_t)y * W + x);
            q[0] = p[2];
            q[1] = p[1];
            q[2] = p[0];
            q[3] = px == 4 ? p[3] : 255;
        }
    }
    *width = W;
    *height = H;
    *rgba = img;
    return EDXR_OK;
}

// Binary PPM (P6, maxval 255): rows top-down in the file, returned bottom-up.
int edxr_load_ppm(const char* path, uint32_t* width, uint32_t* height, uint8_t** rgba) {
    if (!path || !width || !height || !rgba) return EDXR_E_INVALID;
    *rgba = nullptr;
    std::string d;
    if (!read_file(path, d)) return EDXR_E_IO;
    size_t at = 0;
    // the header's four tokens: magic, width, height, maxval ('#' comments skipped)
    auto token = [&](std::string& t) {
        t.clear();
        while (at < d.size()) {
            if (d[at] == '#') {
                while (at < d.size() && d[at] != '\n') at++;
            } else if (isspace((unsigned char)d[at])) {
                at++;
            } else {
                break;
            }
        }
        while (at < d.size() && !isspace((unsigned char)d[at])) t += d[at++];
        return !t.empty();
    };
    std::string magic, ws, hs, ms;
    if (!token(magic) || magic != "P6" || !token(ws) || !token(hs) || !token(ms)) return EDXR_E_FORMAT;
    const long W = strtol(ws.c_str(), nullptr, 10), H = strtol(hs.c_str(), nullptr, 10);
    if (ms != "255" || W <= 0 || H <= 0 || W > 16384 || H > 16384 || at >= d.size()) return EDXR_E_FORMAT;
    at++;  // the single whitespace byte after maxval
    const size_t n = (size_t)W * H;
    if (at + 3 * n > d.size()) return EDXR_E_FORMAT;
    uint8_t* img = (uint8_t*)malloc(4 * n);
    if (!img) return EDXR_E_NOMEM;
    const uint8_t* b = (const uint8_t*)d.data() + at;
    for (long y = 0; y < H; y++)
        for (long x = 0; x < W; x++) {
            const uint8_t* p = b + 3 * ((size_t)(H - 1 - y) * W + x);
            uint8_t* q = img + 4 * ((size_t)y * W + x);
            q[0] = p[0];
            q[1] = p[1];
            q[2] = p[2];
            q[3] = 255;
        }
    *width = (uint32_t)W;
    *height = (uint32_t)H;
    *rgba = img;
    return EDXR_OK;
}

// The diffuse-map reader of ImageTexture (Utils/Mesh.cpp:26-27): BMP, TGA or
// PPM by the file's first bytes ("BM", "P6", else TGA).
int edxr_load_image(const char* path, uint32_t* width, uint32_t* height, uint8_t** rgba) {
    if (!path || !width || !height || !rgba) return EDXR_E_INVALID;
    FILE* f = fopen(path, "rb");
    if (!f) return EDXR_E_IO;
    char m[2] = {0, 0};
    const size_t got = fread(m, 1, 2, f);
    fclose(f);
    if (got == 2 && m[0] == 'B' && m[1] == 'M') return edxr_load_bmp(path, width, height, rgba);
    if (got == 2 && m[0] == 'P' && m[1] == '6') return edxr_load_ppm(path, width, height, rgba);
    return edxr_load_tga(path, width, height, rgba);
}

void edxr_free(void* p) { free(p); }
