/*
 * edxr_oracle.cpp -- CPU oracle: a restatement of the EDXRaster CPU pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).  The product
 * library never links, loads or calls this file.
 *
 * The stages keep the reference's structure: per-core vertex and triangle
 * buffers, per-(tile, core) reference lists, a 4-wide SSE 2x2-quad rasteriser
 * with an early depth test emitting fragments, fragment-parallel shading and an
 * ordered per-tile framebuffer update.  PPL parallel_for is replaced by a small
 * std::thread pool.  Every function cites the reference lines it restates
 * (paths relative to EDXRaster/).
 *
 * EDXUtil is not available, so its semantics are fixed by declaration:
 *   TransformPoint(v4, M)  out_r = ((m_r0*x + m_r1*y) + m_r2*z) + m_r3*w, no FMA
 *   TransformPoint(v3, M)  the same with w = 1, then x/w, y/w, z/w
 *   HomogeneousProject     x/w, y/w, z/w (true division)
 *   Matrix product         float, k = 0..3 left to right, no FMA
 *   Matrix::Inverse        cofactors in double, rounded once
 *   Rsqrt(x)               1.0f / sqrtf(x)
 *   Normalize(v)           v * (1.0f / sqrtf(dot(v, v)))
 *   Dot(a, b)              (a.x*b.x + a.y*b.y) + a.z*b.z
 *   Pow(x, s)              exponentiation by squaring for integer s, else powf
 *   Color4b::FromFloats    (uint8)(int)(clamp(c, 0, 1) * 255 + 0.5), alpha 255;
 *                          clamp is fminf(fmaxf(c, 0), 1), so NaN gives 0
 *   IntSSE(EDX_INFINITY)   INT_MAX
 *   BoolSSE true           all ones (-1 when added to an IntSSE)
 *   Color(Color4b b)       (b.r / 255.0f, b.g / 255.0f, b.b / 255.0f, b.a / 255.0f)
 *   Color4b(Color c)       as FromFloats, per channel including alpha
 *   Color4b(r, g, b)       alpha 255
 * Parity unpinned: these declared EDXUtil semantics (and the LoadSphere /
 * LoadPlane stand-in topology) have no check against EDXUtil itself, which is
 * absent; the tests pin them only against closed forms and brute force.
 * Additions beyond the reference: a triangle-ID buffer (input triangle index,
 * clipped pieces inherit it, background 0xFFFFFFFF), a vertex-colour
 * Blinn-Phong shader and a settable light direction.  Pixels outside the
 * screen are never written.
 *
 * Build: g++ -O2 -msse4.1 -ffp-contract=off -fwrapv (int32 wrap is defined).
 */
#include "edxr_oracle.h"

#include <smmintrin.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace orc {

/* ------------------------------------------------------------------ */
/* Scalar math with the declared EDXUtil semantics                      */
/* ------------------------------------------------------------------ */
struct Vec2i { int x, y; };
struct Vec3 { float x, y, z; };
struct Vec4 { float x, y, z, w; };
struct Mat4 { float m[4][4]; };

static Mat4 load_mat(const float* a) {
    Mat4 r;
    for (int i = 0; i < 16; ++i) r.m[i / 4][i % 4] = a[i];
    return r;
}

static inline Vec4 TransformPoint(const Vec4& v, const Mat4& M) {
    Vec4 r;
    r.x = ((M.m[0][0] * v.x + M.m[0][1] * v.y) + M.m[0][2] * v.z) + M.m[0][3] * v.w;
    r.y = ((M.m[1][0] * v.x + M.m[1][1] * v.y) + M.m[1][2] * v.z) + M.m[1][3] * v.w;
    r.z = ((M.m[2][0] * v.x + M.m[2][1] * v.y) + M.m[2][2] * v.z) + M.m[2][3] * v.w;
    r.w = ((M.m[3][0] * v.x + M.m[3][1] * v.y) + M.m[3][2] * v.z) + M.m[3][3] * v.w;
    return r;
}

static inline Vec3 TransformPoint(const Vec3& v, const Mat4& M) {
    Vec4 r = TransformPoint(Vec4{v.x, v.y, v.z, 1.0f}, M);
    return Vec3{r.x / r.w, r.y / r.w, r.z / r.w};
}

static inline Vec3 HomogeneousProject(const Vec4& v) {
    return Vec3{v.x / v.w, v.y / v.w, v.z / v.w};
}

static Mat4 MatMul(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = ((a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j]) +
                        a.m[i][3] * b.m[3][j];
    return r;
}

static Mat4 Inverse(const Mat4& a) {
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = a.m[i / 4][i % 4];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    Mat4 r;
    for (int i = 0; i < 16; ++i) r.m[i / 4][i % 4] = (float)(det != 0.0 ? inv[i] / det : 0.0);
    return r;
}

static inline float Dot(const Vec3& a, const Vec3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline Vec3 Normalize(const Vec3& v) {
    float s = 1.0f / sqrtf(Dot(v, v));
    return Vec3{v.x * s, v.y * s, v.z * s};
}

/* Exponentiation by squaring (declared Math::Pow for integer exponents). */
static inline float PowInt(float x, unsigned n) {
    float result = 1.0f, base = x;
    while (n) {
        if (n & 1u) result *= base;
        n >>= 1;
        if (n) base *= base;
    }
    return result;
}
static inline float PowDecl(float x, float s) {
    if (s >= 0.0f && s <= 4096.0f && s == floorf(s)) return PowInt(x, (unsigned)s);
    return powf(x, s);
}

static inline uint8_t ToByte(float c) {
    c = fminf(fmaxf(c, 0.0f), 1.0f);
    return (uint8_t)(int)(c * 255.0f + 0.5f);
}

/* ------------------------------------------------------------------ */
/* Texture slots: Texture2D<Color>::Sample of EDXUtil (absent).  Declared   */
/* semantics (parity unpinned against EDXUtil): texel (i, j) centred at     */
/* ((i + 0.5) / w, (j + 0.5) / h); coordinates wrap; byte b reads b / 255;  */
/* mip level k + 1 halves the sides (floor, at least 1), each texel the     */
/* 0.25 * (((a + b) + c) + d) mean of its 2x2 parents (a side of 1 repeats); */
/* float-to-index conversion maps NaN to 0 and clamps to +-2^30.            */
/* ------------------------------------------------------------------ */
struct TextureSlot {
    bool constant = true;
    float color[3] = {0.0f, 0.0f, 0.0f};
    std::vector<std::vector<float>> levels;  /* RGBA floats per level */
    std::vector<int> lw, lh;
};

static TextureSlot MakeSlot(const oracle_texture& t) {
    TextureSlot s;
    for (int k = 0; k < 3; k++) s.color[k] = t.color[k];
    if (!t.rgba) return s;
    s.constant = false;
    int w = (int)t.width, h = (int)t.height;
    std::vector<float> l0((size_t)w * h * 4);
    for (size_t i = 0; i < l0.size(); i++) l0[i] = (float)t.rgba[i] / 255.0f;
    s.levels.push_back(std::move(l0));
    s.lw.push_back(w);
    s.lh.push_back(h);
    while (w > 1 || h > 1) {
        const int w2 = std::max(1, w / 2), h2 = std::max(1, h / 2);
        const std::vector<float>& a = s.levels.back();
        std::vector<float> n((size_t)w2 * h2 * 4);
        for (int y = 0; y < h2; y++)
            for (int x = 0; x < w2; x++) {
                const int x0 = 2 * x, x1 = std::min(2 * x + 1, w - 1), y0 = 2 * y, y1 = std::min(2 * y + 1, h - 1);
                for (int ch = 0; ch < 4; ch++) {
                    const float p00 = a[((size_t)y0 * w + x0) * 4 + ch], p10 = a[((size_t)y0 * w + x1) * 4 + ch];
                    const float p01 = a[((size_t)y1 * w + x0) * 4 + ch], p11 = a[((size_t)y1 * w + x1) * 4 + ch];
                    n[((size_t)y * w2 + x) * 4 + ch] = 0.25f * (((p00 + p10) + p01) + p11);
                }
            }
        s.levels.push_back(std::move(n));
        s.lw.push_back(w2);
        s.lh.push_back(h2);
        w = w2;
        h = h2;
    }
    return s;
}

static inline int TexIndex(float x) {
    if (x != x) return 0;
    return (int)std::min(std::max(x, -1073741824.0f), 1073741823.0f);
}

static inline const float* Texel(const TextureSlot& s, int level, int i, int j) {
    const int w = s.lw[level], h = s.lh[level];
    i %= w;
    if (i < 0) i += w;
    j %= h;
    if (j < 0) j += h;
    return &s.levels[level][((size_t)j * w + i) * 4];
}

static Vec3 Bilinear(const TextureSlot& s, int level, float u, float v) {
    const float x = u * (float)s.lw[level] - 0.5f, y = v * (float)s.lh[level] - 0.5f;
    const float x0 = std::floor(x), y0 = std::floor(y);
    const float fx = x - x0, fy = y - y0, gx = 1.0f - fx, gy = 1.0f - fy;
    const int i = TexIndex(x0), j = TexIndex(y0);
    const float *c00 = Texel(s, level, i, j), *c10 = Texel(s, level, i + 1, j);
    const float *c01 = Texel(s, level, i, j + 1), *c11 = Texel(s, level, i + 1, j + 1);
    float r[3];
    for (int ch = 0; ch < 3; ch++)
        r[ch] = gy * (gx * c00[ch] + fx * c10[ch]) + fy * (gx * c01[ch] + fx * c11[ch]);
    return Vec3{r[0], r[1], r[2]};
}

/* Sample(texCoord, differentials[2]) with the state's filter. */
/* The trilinear blend of the two mip levels around lod (clamped to the chain). */
static Vec3 TrilinearAt(const TextureSlot& s, float lod, float u, float v) {
    const int nl = (int)s.levels.size();
    lod = std::min(std::max(lod, 0.0f), (float)(nl - 1));
    const float fl = std::floor(lod);
    const int l0 = (int)fl, l1 = std::min(l0 + 1, nl - 1);
    const float t = lod - fl, g = 1.0f - t;
    const Vec3 a = Bilinear(s, l0, u, v), b = Bilinear(s, l1, u, v);
    return Vec3{g * a.x + t * b.x, g * a.y + t * b.y, g * a.z + t * b.z};
}

static Vec3 SampleSlot(const TextureSlot& s, int filter, float u, float v, float dux, float dvx, float duy, float dvy) {
    if (s.constant) return Vec3{s.color[0], s.color[1], s.color[2]};
    const int w = s.lw[0], h = s.lh[0];
    if (filter == ORACLE_FILTER_NEAREST) {
        const float* c = Texel(s, 0, TexIndex(std::floor(u * (float)w)), TexIndex(std::floor(v * (float)h)));
        return Vec3{c[0], c[1], c[2]};
    }
    if (filter == ORACLE_FILTER_LINEAR) return Bilinear(s, 0, u, v);
    const float ax = dux * (float)w, ay = dvx * (float)h, bx = duy * (float)w, by = dvy * (float)h;
    const float la2 = ax * ax + ay * ay, lb2 = bx * bx + by * by;
    if (filter == ORACLE_FILTER_TRILINEAR) {
        /* lod = log2 of the longer texel-space differential */
        const float rho2 = std::max(la2, lb2);
        return TrilinearAt(s, rho2 > 0.0f ? 0.5f * std::log2(rho2) : 0.0f, u, v);
    }
    /* N:1 anisotropic: N trilinear taps spread evenly along the longer
     * differential (tap k at offset (2k + 1 - N) / 2N of it), at the lod of
     * the shorter one raised to at least 1/N of the longer, averaged in tap
     * order. */
    const int N = 4 << (filter - ORACLE_FILTER_ANISO_4X);
    const bool xMajor = la2 >= lb2;
    const float mu = xMajor ? dux : duy, mv = xMajor ? dvx : dvy;
    const float maj2 = xMajor ? la2 : lb2, n2 = (float)(N * N);
    float min2 = xMajor ? lb2 : la2;
    if (min2 * n2 < maj2) min2 = maj2 / n2;
    const float lod = min2 > 0.0f ? 0.5f * std::log2(min2) : 0.0f;
    Vec3 acc{0.0f, 0.0f, 0.0f};
    for (int k = 0; k < N; k++) {
        const float sk = (float)(2 * k + 1 - N) / (float)(2 * N);
        const Vec3 c = TrilinearAt(s, lod, u + sk * mu, v + sk * mv);
        acc.x += c.x;
        acc.y += c.y;
        acc.z += c.z;
    }
    const float inv = 1.0f / (float)N;
    return Vec3{acc.x * inv, acc.y * inv, acc.z * inv};
}

/* 28.4 snapping: RasterTriangle.h:35-40 converts x * 16.0 to int by
 * truncation.  The product is exact; out-of-range values saturate and NaN
 * gives 0 (declared: the reference has undefined behaviour there). */
static inline int Snap(float a) {
    float s = a * 16.0f;
    if (!(s == s)) return 0;
    if (s >= 2147483648.0f) return INT_MAX;
    if (s <= -2147483648.0f) return INT_MIN;
    return (int)s;
}

/* ------------------------------------------------------------------ */
/* Minimal SSE wrappers (4 lanes = one 2x2 quad)                         */
/* ------------------------------------------------------------------ */
struct IntSSE {
    __m128i v;
    IntSSE() : v(_mm_setzero_si128()) {}
    explicit IntSSE(__m128i a) : v(a) {}
    IntSSE(int a) : v(_mm_set1_epi32(a)) {}
    IntSSE(int a, int b, int c, int d) : v(_mm_setr_epi32(a, b, c, d)) {}
    int operator[](int i) const { alignas(16) int t[4]; _mm_store_si128((__m128i*)t, v); return t[i]; }
};
static inline IntSSE operator+(IntSSE a, IntSSE b) { return IntSSE(_mm_add_epi32(a.v, b.v)); }
static inline IntSSE operator-(IntSSE a, IntSSE b) { return IntSSE(_mm_sub_epi32(a.v, b.v)); }
static inline IntSSE operator*(IntSSE a, IntSSE b) { return IntSSE(_mm_mullo_epi32(a.v, b.v)); }
static inline IntSSE operator|(IntSSE a, IntSSE b) { return IntSSE(_mm_or_si128(a.v, b.v)); }
static inline IntSSE& operator+=(IntSSE& a, IntSSE b) { a = a + b; return a; }
/* BoolSSE: all-ones lanes for true. */
typedef IntSSE BoolSSE;
static inline BoolSSE operator&(BoolSSE a, BoolSSE b) { return IntSSE(_mm_and_si128(a.v, b.v)); }
static inline BoolSSE LessZero(IntSSE a) { return IntSSE(_mm_cmplt_epi32(a.v, _mm_setzero_si128())); }
static inline BoolSSE GreaterEqZero(IntSSE a) {
    return IntSSE(_mm_xor_si128(_mm_cmplt_epi32(a.v, _mm_setzero_si128()), _mm_set1_epi32(-1)));
}
static inline bool Any(BoolSSE a) { return _mm_movemask_ps(_mm_castsi128_ps(a.v)) != 0; }

struct FloatSSE {
    __m128 v;
    FloatSSE() : v(_mm_setzero_ps()) {}
    explicit FloatSSE(__m128 a) : v(a) {}
    FloatSSE(float a) : v(_mm_set1_ps(a)) {}
    explicit FloatSSE(IntSSE a) : v(_mm_cvtepi32_ps(a.v)) {}
    float operator[](int i) const { alignas(16) float t[4]; _mm_store_ps(t, v); return t[i]; }
};
static inline FloatSSE operator+(FloatSSE a, FloatSSE b) { return FloatSSE(_mm_add_ps(a.v, b.v)); }
static inline FloatSSE operator-(FloatSSE a, FloatSSE b) { return FloatSSE(_mm_sub_ps(a.v, b.v)); }
static inline FloatSSE operator*(FloatSSE a, FloatSSE b) { return FloatSSE(_mm_mul_ps(a.v, b.v)); }
static inline FloatSSE operator/(FloatSSE a, FloatSSE b) { return FloatSSE(_mm_div_ps(a.v, b.v)); }
static inline FloatSSE& operator*=(FloatSSE& a, FloatSSE b) { a = a * b; return a; }
static inline BoolSSE LessEq(FloatSSE a, FloatSSE b) { return IntSSE(_mm_castps_si128(_mm_cmple_ps(a.v, b.v))); }
static inline FloatSSE Select(BoolSSE m, FloatSSE a, FloatSSE b) {
    return FloatSSE(_mm_blendv_ps(b.v, a.v, _mm_castsi128_ps(m.v)));
}
/* SSE::Rsqrt, declared exact. */
static inline FloatSSE Rsqrt(FloatSSE a) { return FloatSSE(_mm_div_ps(_mm_set1_ps(1.0f), _mm_sqrt_ps(a.v))); }

struct Vec3SSE {
    FloatSSE x, y, z;
};
static inline Vec3SSE operator+(const Vec3SSE& a, const Vec3SSE& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline Vec3SSE operator-(const Vec3SSE& a, const Vec3SSE& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline Vec3SSE operator*(const FloatSSE& s, const Vec3SSE& a) { return {s * a.x, s * a.y, s * a.z}; }
static inline Vec3SSE operator*(const Vec3SSE& a, const FloatSSE& s) { return {a.x * s, a.y * s, a.z * s}; }
static inline FloatSSE Dot(const Vec3SSE& a, const Vec3SSE& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline Vec3SSE Splat(const Vec3& v) { return {FloatSSE(v.x), FloatSSE(v.y), FloatSSE(v.z)}; }

/* ------------------------------------------------------------------ */
/* Thread pool standing in for concurrency::parallel_for                 */
/* ------------------------------------------------------------------ */
class Pool {
  public:
    explicit Pool(int n) : nthreads_(std::max(1, n)) {
        for (int i = 1; i < nthreads_; ++i) workers_.emplace_back([this] { loop(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    int size() const { return nthreads_; }
    /* Dynamic scheduling like PPL; every index runs exactly once. */
    void parallel_for(int begin, int end, const std::function<void(int)>& fn) {
        if (end <= begin) return;
        if (nthreads_ == 1) {
            for (int i = begin; i < end; ++i) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            fn_ = &fn;
            next_.store(begin);
            end_ = end;
            grain_ = std::max(1, (end - begin) / (nthreads_ * 16));
            pending_ = nthreads_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return pending_ == 0; });
        fn_ = nullptr;
    }

  private:
    void work() {
        for (;;) {
            int i = next_.fetch_add(grain_);
            if (i >= end_) break;
            int e = std::min(end_, i + grain_);
            for (int k = i; k < e; ++k) (*fn_)(k);
        }
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
            }
            work();
            {
                std::lock_guard<std::mutex> g(mu_);
                if (--pending_ == 0) done_cv_.notify_one();
            }
        }
    }
    int nthreads_;
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)>* fn_ = nullptr;
    std::atomic<int> next_{0};
    int end_ = 0, grain_ = 1, pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

/* ------------------------------------------------------------------ */
/* Pipeline types                                                        */
/* ------------------------------------------------------------------ */
/* Core/Shader.h:13-25, with an RGB vertex colour beside the UV pair. */
struct ProjectedVertex {
    Vec4 projectedPos;
    float invW = 0.0f;
    Vec3 position, normal, color;
    float u = 0.0f, v = 0.0f;
};

/* Core/RasterTriangle.h:11-24 plus the input triangle index and fan index. */
struct RasterTriangle {
    Vec2i v0, v1, v2;
    int B0, C0, B1, C1, B2, C2;
    int stepB0, stepC0, stepB1, stepC1, stepB2, stepC2;
    float invDet;
    uint32_t vId0, vId1, vId2, coreId;
    uint32_t triId, sub;
    uint8_t rejectCorner0, rejectCorner1, rejectCorner2;
    uint8_t acceptCorner0, acceptCorner1, acceptCorner2;

    static void Corners(int B, int C, uint8_t& rej, uint8_t& acc) {
        /* Core/RasterTriangle.h:66-150: slope B/float(C) and the sign of C. */
        float slope = (float)B / (float)C;
        if (slope >= 0.0f) {
            if ((float)C >= 0.0f) { rej = 3; acc = 0; } else { rej = 0; acc = 3; }
        } else {
            if ((float)C >= 0.0f) { rej = 2; acc = 1; } else { rej = 1; acc = 2; }
        }
    }

    /* Core/RasterTriangle.h:27-153. */
    bool Setup(Vec3 a, Vec3 b, Vec3 c, const uint32_t* pIdx, uint32_t cId, const Mat4& raster, int mutation) {
        a = TransformPoint(a, raster);
        b = TransformPoint(b, raster);
        c = TransformPoint(c, raster);
        v0 = {Snap(a.x), Snap(a.y)};
        v1 = {Snap(b.x), Snap(b.y)};
        v2 = {Snap(c.x), Snap(c.y)};
        B0 = v0.y - v1.y; C0 = v1.x - v0.x;
        B1 = v1.y - v2.y; C1 = v2.x - v1.x;
        B2 = v2.y - v0.y; C2 = v0.x - v2.x;
        int det = C2 * B1 - C1 * B2;
        if (det <= 0) return false;
        stepB0 = 16 * B0; stepC0 = 16 * C0;
        stepB1 = 16 * B1; stepC1 = 16 * C1;
        stepB2 = 16 * B2; stepC2 = 16 * C2;
        invDet = 1.0f / float(det < 0 ? -det : det);
        vId0 = pIdx[0]; vId1 = pIdx[1]; vId2 = pIdx[2];
        coreId = cId;
        Corners(B0, C0, rejectCorner0, acceptCorner0);
        Corners(B1, C1, rejectCorner1, acceptCorner1);
        Corners(B2, C2, rejectCorner2, acceptCorner2);
        (void)mutation;
        return true;
    }

    /* Scalar top-left bias, Core/RasterTriangle.h:155-158 (inverted vs SSE). */
    static inline int TopLeftEdge(const Vec2i& a, const Vec2i& b) {
        return ((b.y > a.y) || (a.y == b.y && a.x > b.x)) ? 0 : -1;
    }
    /* Core/RasterTriangle.h:160-171. */
    inline int EdgeFunc0(const Vec2i& p) const { return B0 * (p.x - v0.x) + C0 * (p.y - v0.y) + TopLeftEdge(v0, v1); }
    inline int EdgeFunc1(const Vec2i& p) const { return B1 * (p.x - v1.x) + C1 * (p.y - v1.y) + TopLeftEdge(v1, v2); }
    inline int EdgeFunc2(const Vec2i& p) const { return B2 * (p.x - v2.x) + C2 * (p.y - v2.y) + TopLeftEdge(v2, v0); }

    /* Core/RasterTriangle.h:192-255. */
    static inline IntSSE Step(int corner, int B, int C, int stepSize) {
        switch (corner) {
            case 0: return IntSSE(stepSize) * IntSSE(0, B, C, B + C);
            case 1: return IntSSE(stepSize) * IntSSE(-B, 0, C - B, C);
            case 2: return IntSSE(stepSize) * IntSSE(-C, B - C, 0, B);
            default: return IntSSE(stepSize) * IntSSE(-B - C, -C, -B, 0);
        }
    }
};

/* Core/RasterTriangle.h:258-338. */
struct TriangleSSE {
    IntSSE v0x, v0y, v1x, v1y, v2x, v2y;
    IntSSE B0, C0, B1, C1, B2, C2;
    IntSSE stepB0, stepC0, stepB1, stepC1, stepB2, stepC2;
    FloatSSE invDet;
    IntSSE bias0, bias1, bias2;
    FloatSSE lambda0, lambda1;

    static IntSSE TopLeftEdge(const Vec2i& a, const Vec2i& b, int mutation) {
        /* Core/RasterTriangle.h:296-299: mask (b.y > a.y) | (a.y == b.y & a.x > b.x). */
        bool t = (b.y > a.y) || (a.y == b.y && a.x > b.x);
        if (mutation == 1) t = !t;  /* planted mistake: inverted top-left rule */
        return IntSSE(t ? -1 : 0);
    }

    TriangleSSE(const RasterTriangle& t, int mutation)
        : v0x(t.v0.x), v0y(t.v0.y), v1x(t.v1.x), v1y(t.v1.y), v2x(t.v2.x), v2y(t.v2.y),
          B0(t.B0), C0(t.C0), B1(t.B1), C1(t.C1), B2(t.B2), C2(t.C2),
          stepB0(2 * t.stepB0), stepC0(2 * t.stepC0), stepB1(2 * t.stepB1), stepC1(2 * t.stepC1),
          stepB2(2 * t.stepB2), stepC2(2 * t.stepC2), invDet(t.invDet) {
        mutation_ = mutation;
        bias0 = TopLeftEdge(t.v0, t.v1, mutation);
        bias1 = TopLeftEdge(t.v1, t.v2, mutation);
        bias2 = TopLeftEdge(t.v2, t.v0, mutation);
    }
    inline IntSSE EdgeFunc0(const IntSSE& px, const IntSSE& py) const { return B0 * (px - v0x) + C0 * (py - v0y) + bias0; }
    inline IntSSE EdgeFunc1(const IntSSE& px, const IntSSE& py) const { return B1 * (px - v1x) + C1 * (py - v1y) + bias1; }
    inline IntSSE EdgeFunc2(const IntSSE& px, const IntSSE& py) const { return B2 * (px - v2x) + C2 * (py - v2y) + bias2; }

    /* Core/RasterTriangle.h:333-337. */
    int mutation_ = 0;
    inline void CalcBarycentricCoord(const IntSSE& x, const IntSSE& y) {
        if (mutation_ == 6) {  /* planted: wrong edge in lambda0 */
            lambda0 = FloatSSE(B0 * (x - v2x) + C0 * (y - v2y)) * invDet;
            lambda1 = FloatSSE(B2 * (x - v2x) + C2 * (y - v2y)) * invDet;
            return;
        }
        lambda0 = FloatSSE(B1 * (x - v2x) + C1 * (y - v2y)) * invDet;
        lambda1 = FloatSSE(B2 * (x - v2x) + C2 * (y - v2y)) * invDet;
    }
    /* Core/RasterTriangle.h:324-331. */
    inline FloatSSE GetDepth(float z0, float z1, float z2) const {
        const FloatSSE one(1.0f);
        FloatSSE lambda2 = one - lambda0 - lambda1;
        return lambda0 * FloatSSE(z0) + lambda1 * FloatSSE(z1) + lambda2 * FloatSSE(z2);
    }
};

/* Core/Tile.h:10-39 (per-core lists sized by the thread count). */
struct TriangleRef {
    uint32_t triId;
    bool acceptEdge0, acceptEdge1, acceptEdge2, trivialAccept, big;
    TriangleRef(uint32_t id, bool a0 = false, bool a1 = false, bool a2 = false, bool b = false)
        : triId(id), acceptEdge0(a0), acceptEdge1(a1), acceptEdge2(a2), trivialAccept(a0 && a1 && a2), big(b) {}
};

/* CoverageMask, Core/Shader.h:52-103: bit (sample << 2) + lane, up to 32 samples. */
struct CoverageMask {
    uint32_t bits[4] = {0, 0, 0, 0};
    void SetBit(int i) { bits[i >> 5] |= 1u << (i & 31); }
    void SetBit(BoolSSE m, int sampleId) {
        for (int k = 0; k < 4; k++)
            if (m[k] != 0) SetBit((sampleId << 2) + k);
    }
    bool GetBit(int i) const { return (bits[i >> 5] >> (i & 31)) & 1u; }
};

/* Core/Shader.h:105-140 plus the triangle ID. */
struct Fragment {
    FloatSSE lambda0, lambda1;
    CoverageMask coverage;
    int x, y;
    uint32_t vId0, vId1, vId2, coreId, triId;
    uint32_t tileId, intraTileIdx;
};

struct Tile {
    static const int SIZE_LOG_2 = 5;
    static const int SIZE = 1 << SIZE_LOG_2;
    Vec2i minCoord, maxCoord;
    uint32_t tileId;
    std::vector<std::vector<TriangleRef>> triangleRefs;
    std::vector<Fragment> fragmentBuf;
};

/* Core/Clipper.h:12-35. */
struct Polygon {
    struct Vertex { Vec4 pos; Vec3 clipWeights; };
    std::vector<Vertex> vertices;
    void FromTriangle(const Vec4& a, const Vec4& b, const Vec4& c) {
        vertices.assign(3, Vertex{});
        vertices[0] = {a, {1, 0, 0}};
        vertices[1] = {b, {0, 1, 0}};
        vertices[2] = {c, {0, 0, 1}};
    }
};

/* Planted-mistake switch for the clipper (pinning tests only). */
static thread_local int g_mutation = 0;

/* Core/Clipper.h:37-68. */
static const uint32_t LEFT_BIT = 1 << 0, RIGHT_BIT = 1 << 1, BOTTOM_BIT = 1 << 2, TOP_BIT = 1 << 3,
                      NEAR_BIT = 1 << 4, FAR_BIT = 1 << 5;
static inline uint32_t ComputeClipCode(const Vec4& v) {
    uint32_t code = 0;
    if (v.x < -v.w) code |= LEFT_BIT;
    if (v.x > v.w) code |= RIGHT_BIT;
    if (v.y < -v.w) code |= BOTTOM_BIT;
    if (v.y > v.w) code |= TOP_BIT;
    if (v.z > v.w) code |= FAR_BIT;
    if (v.z < 0.0f) code |= NEAR_BIT;
    return code;
}

/* Core/Clipper.h:192-232; pos = v0*(1-t) + v1*t per component. */
template <typename Pred, typename CompT, typename ClipF>
static void ClipByPlane(Polygon*& in, Polygon*& buf, Pred pred, CompT computeT, ClipF clip) {
    buf->vertices.clear();
    const int n = (int)in->vertices.size();
    for (int i = 0; i < n; i++) {
        int i1 = i + 1;
        if (i1 == n) i1 = 0;
        const Vec4 v0 = in->vertices[i].pos;
        const Vec4 v1 = in->vertices[i1].pos;
        const Vec3 w0 = in->vertices[i].clipWeights, w1 = in->vertices[i1].clipWeights;
        auto lerp = [&](float t) {
            const float s = 1 - t;
            Vec4 p{v0.x * s + v1.x * t, v0.y * s + v1.y * t, v0.z * s + v1.z * t, v0.w * s + v1.w * t};
            clip(p);
            Vec3 w{w0.x * s + w1.x * t, w0.y * s + w1.y * t, w0.z * s + w1.z * t};
            return Polygon::Vertex{p, w};
        };
        if (pred(v0)) {
            if (pred(v1)) buf->vertices.push_back({v1, w1});
            else buf->vertices.push_back(lerp(computeT(v0, v1)));
        } else if (pred(v1)) {
            buf->vertices.push_back(lerp(computeT(v0, v1)));
            buf->vertices.push_back({v1, w1});
        }
    }
    std::swap(in, buf);
}

/* Core/Clipper.h:235-288 (CLIP_ALL_PLANES = 1, plane order L R B T F N). */
static void ClipPolygon(Polygon*& in, Polygon*& buf, uint32_t planeCode) {
    if (planeCode & LEFT_BIT)
        ClipByPlane(in, buf, [](const Vec4& v) { return v.x >= -v.w; },
                    [](const Vec4& a, const Vec4& b) { return (a.w + a.x) / ((a.x + a.w) - (b.x + b.w)); },
                    [](Vec4& v) { v.x = -v.w; });
    if (planeCode & RIGHT_BIT)
        ClipByPlane(in, buf, [](const Vec4& v) { return v.x <= v.w; },
                    [](const Vec4& a, const Vec4& b) { return (-a.w + a.x) / ((a.x - a.w) - (b.x - b.w)); },
                    [](Vec4& v) { v.x = v.w; });
    if (planeCode & BOTTOM_BIT)
        ClipByPlane(in, buf, [](const Vec4& v) { return v.y >= -v.w; },
                    [](const Vec4& a, const Vec4& b) { return (a.w + a.y) / ((a.y + a.w) - (b.y + b.w)); },
                    [](Vec4& v) { v.y = -v.w; });
    if (planeCode & TOP_BIT)
        ClipByPlane(in, buf, [](const Vec4& v) { return v.y <= v.w; },
                    [](const Vec4& a, const Vec4& b) { return (-a.w + a.y) / ((a.y - a.w) - (b.y - b.w)); },
                    [](Vec4& v) { v.y = v.w; });
    if (planeCode & FAR_BIT)
        ClipByPlane(in, buf, [](const Vec4& v) { return v.z <= v.w; },
                    [](const Vec4& a, const Vec4& b) {
                        /* planted when 8: wrong far-plane t (numerator sign) */
                        return g_mutation == 8 ? (a.w - a.z) / ((a.z - a.w) - (b.z - b.w))
                                               : (-a.w + a.z) / ((a.z - a.w) - (b.z - b.w));
                    },
                    [](Vec4& v) { v.z = v.w; });
    if (planeCode & NEAR_BIT)
        ClipByPlane(in, buf, [](const Vec4& v) { return v.z >= 0.0f; },
                    [](const Vec4& a, const Vec4& b) { return g_mutation == 5 ? a.z / (b.z - a.z) : a.z / (a.z - b.z); },
                    [](Vec4& v) { v.z = 0.0f; });
    for (const auto& v : in->vertices)
        if (v.pos.w <= 0.0f) {
            in->vertices.clear();
            return;
        }
}

/* ------------------------------------------------------------------ */
/* Renderer                                                              */
/* ------------------------------------------------------------------ */
struct PerCoreStats {
    uint64_t rejected = 0, clipped = 0, near_clipped = 0, dropped = 0, culled = 0, clip_vertices = 0;
};

class Renderer {
  public:
    ~Renderer() { delete pool_; }

    int Render(const oracle_params& p, const oracle_mesh& mesh, oracle_stats* st);

    /* outputs */
    std::vector<uint8_t> color_;      /* bottom-up RGBA8 */
    std::vector<uint32_t> ids_;       /* bottom-up */
    std::vector<float> depthLinear_;  /* bottom-up */
    std::vector<float> depthMS_;      /* [sample][row][x] */
    std::vector<uint32_t> idsMS_;     /* [sample][row][x] */
    std::vector<uint8_t> colorMS_;    /* [sample][row][x] RGBA8 */
    int sampleCount_ = 1, msLevel_ = 0;
    std::vector<oracle_record> records_;
    std::vector<uint64_t> binOffsets_;
    std::vector<oracle_ref> binRefs_;

  private:
    void Initialize(uint32_t w, uint32_t h, int threads);
    void VertexProcessing(const oracle_mesh& mesh);
    void Clipping(const oracle_mesh& mesh);
    void TiledRasterization();
    void RasterizeTile(Tile& tile);
    void FragmentProcessing();
    void UpdateFrameBuffer();

    /* Rasterizer (Core/Rasterizer.h) */
    void CoarseRasterize(Tile& tile, const TriangleRef& ref, int blockSize, Vec2i bmin, Vec2i bmax,
                         const RasterTriangle& tri);
    void FineRasterize(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri);
    void FineRasterizeSingle(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri);
    void FineRasterizeMulti(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri);
    void TrivialAcceptTriangle(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri);
    void TrivialAcceptSingle(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri);
    void TrivialAcceptMulti(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri);
    void EmitFragment(Tile& tile, const TriangleSSE& t, const RasterTriangle& tri, int x, int y,
                      const CoverageMask& mask);
    /* FrameBuffer (Core/FrameBuffer.cpp:54-68) */
    BoolSSE ZTestQuad(const FloatSSE& d, int x, int y, int sampleId, BoolSSE mask);
    void Resolve();

    oracle_params p_{};
    const uint32_t* texIds_ = nullptr;   /* Mesh::GetTextureIds (textured shader) */
    std::vector<TextureSlot> slots_;     /* RenderStates::TextureSlots */
    Mat4 mvp_{}, mvInv_{}, raster_{};
    uint32_t W_ = 0, H_ = 0;
    Vec2i tileDim_{0, 0};
    int numCores_ = 0;
    Pool* pool_ = nullptr;
    std::vector<Tile> tiles_;
    std::vector<ProjectedVertex> projVB_;
    std::vector<std::vector<ProjectedVertex>> coreVB_;
    std::vector<std::vector<RasterTriangle>> coreTris_;
    std::vector<PerCoreStats> coreStats_;
    std::vector<FloatSSE> depth_;  /* tiled: [tile][16 x 16 quads], y-flipped quad rows */
    std::vector<Fragment*> fragPtrs_;
    std::vector<std::vector<uint32_t>> shaded_;  /* per tile, per fragment: 4 packed RGBA8 */
};

/* Core/Renderer.cpp:22-62 and 64-83. */
void Renderer::Initialize(uint32_t w, uint32_t h, int threads) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    if (!pool_ || pool_->size() != threads) {
        delete pool_;
        pool_ = new Pool(threads);
    }
    numCores_ = threads;
    if (w != W_ || h != H_ || (int)coreVB_.size() != threads) {
        W_ = w;
        H_ = h;
        tileDim_.x = (int)((w + Tile::SIZE - 1) >> Tile::SIZE_LOG_2);
        tileDim_.y = (int)((h + Tile::SIZE - 1) >> Tile::SIZE_LOG_2);
        tiles_.clear();
        uint32_t tId = 0;
        for (int i = 0; i < (int)h; i += Tile::SIZE)
            for (int j = 0; j < (int)w; j += Tile::SIZE) {
                Tile t;
                t.minCoord = {j, i};
                t.maxCoord = {std::min(j + Tile::SIZE, (int)w), std::min(i + Tile::SIZE, (int)h)};
                t.tileId = tId++;
                tiles_.push_back(std::move(t));
            }
        depth_.assign(tiles_.size() * 256, FloatSSE(1.0f));
        coreVB_.assign(threads, {});
        coreTris_.assign(threads, {});
    }
    for (auto& t : tiles_) t.triangleRefs.assign(threads, {});
    coreStats_.assign(threads, {});
}

/* Core/Renderer.cpp:120-127 with DefaultVertexShader::Execute (Core/Shader.h:40-49). */
void Renderer::VertexProcessing(const oracle_mesh& m) {
    projVB_.resize(m.num_vertices);
    pool_->parallel_for(0, (int)m.num_vertices, [&](int i) {
        ProjectedVertex& o = projVB_[i];
        o.projectedPos = TransformPoint(Vec4{m.px[i], m.py[i], m.pz[i], 1.0f}, mvp_);
        o.position = {m.px[i], m.py[i], m.pz[i]};
        o.normal = {m.nx[i], m.ny[i], m.nz[i]};
        o.color = {m.cr[i], m.cg[i], m.cb[i]};
        if (m.tu) o.u = m.tu[i];
        if (m.tv) o.v = m.tv[i];
        o.invW = 0.0f;
    });
}

/* Core/Renderer.cpp:129-148 with Clipper::Clip (Core/Clipper.h:71-189). */
void Renderer::Clipping(const oracle_mesh& m) {
    const int numCores = numCores_;
    pool_->parallel_for(0, numCores, [&](int c) {
        coreVB_[c].clear();
        coreTris_[c].clear();
    });
    const uint32_t T = m.num_triangles;
    pool_->parallel_for(0, numCores, [&](int coreId) {
        const uint32_t interval = (T + numCores - 1) / numCores;
        const uint32_t startIdx = coreId * interval;
        const uint32_t endIdx = std::min<uint64_t>((uint64_t)(coreId + 1) * interval, T);
        auto& vb = coreVB_[coreId];
        auto& tris = coreTris_[coreId];
        PerCoreStats& cs = coreStats_[coreId];
        Polygon polygon0, polygon1;
        g_mutation = p_.mutation;
        for (uint32_t i = startIdx; i < endIdx; i++) {
            const uint32_t* pIndex = m.indices + 3 * (size_t)i;
            const Vec4 v0 = projVB_[pIndex[0]].projectedPos;
            const Vec4 v1 = projVB_[pIndex[1]].projectedPos;
            const Vec4 v2 = projVB_[pIndex[2]].projectedPos;
            const uint32_t idx0 = (uint32_t)vb.size();
            vb.push_back(projVB_[pIndex[0]]);
            const uint32_t idx1 = (uint32_t)vb.size();
            vb.push_back(projVB_[pIndex[1]]);
            const uint32_t idx2 = (uint32_t)vb.size();
            vb.push_back(projVB_[pIndex[2]]);

            const uint32_t c0 = ComputeClipCode(v0), c1 = ComputeClipCode(v1), c2 = ComputeClipCode(v2);
            if (c0 | c1 | c2) {
                if (!(c0 & c1 & c2)) {
                    cs.clipped++;
                    if ((c0 | c1 | c2) & NEAR_BIT) cs.near_clipped++;
                    uint32_t clipVertIds[12];
                    polygon0.FromTriangle(v0, v1, v2);
                    Polygon* cur = &polygon0;
                    Polygon* buf = &polygon1;
                    ClipPolygon(cur, buf, (c0 ^ c1) | (c1 ^ c2) | (c2 ^ c0));
                    if (cur->vertices.empty()) cs.dropped++;
                    for (int j = 0; j < (int)cur->vertices.size(); j++) {
                        const Vec3 wt = cur->vertices[j].clipWeights;
                        if (wt.x == 1.0f) clipVertIds[j] = idx0;
                        else if (wt.y == 1.0f) clipVertIds[j] = idx1;
                        else if (wt.z == 1.0f) clipVertIds[j] = idx2;
                        else {
                            clipVertIds[j] = (uint32_t)vb.size();
                            const ProjectedVertex& a = vb[idx0];
                            const ProjectedVertex& b = vb[idx1];
                            const ProjectedVertex& c = vb[idx2];
                            ProjectedVertex t;
                            t.projectedPos = cur->vertices[j].pos;
                            /* planted when 7: the first two clip weights swapped */
                            const Vec3 w = g_mutation == 7 ? Vec3{wt.y, wt.x, wt.z} : wt;
                            auto comb = [&](const Vec3& pa, const Vec3& pb, const Vec3& pc) {
                                return Vec3{(w.x * pa.x + w.y * pb.x) + w.z * pc.x,
                                            (w.x * pa.y + w.y * pb.y) + w.z * pc.y,
                                            (w.x * pa.z + w.y * pb.z) + w.z * pc.z};
                            };
                            t.position = comb(a.position, b.position, c.position);
                            t.normal = comb(a.normal, b.normal, c.normal);
                            t.color = comb(a.color, b.color, c.color);
                            t.u = (w.x * a.u + w.y * b.u) + w.z * c.u;
                            t.v = (w.x * a.v + w.y * b.v) + w.z * c.v;
                            vb.push_back(t);
                            cs.clip_vertices++;
                        }
                    }
                    /* Fan triangulation, Core/Clipper.h:155-170. */
                    for (int k = 2; k < (int)cur->vertices.size(); k++) {
                        uint32_t idx[3] = {clipVertIds[0], clipVertIds[k - 1], clipVertIds[k]};
                        RasterTriangle tri;
                        if (tri.Setup(HomogeneousProject(vb[idx[0]].projectedPos),
                                      HomogeneousProject(vb[idx[1]].projectedPos),
                                      HomogeneousProject(vb[idx[2]].projectedPos), idx, coreId, raster_,
                                      p_.mutation)) {
                            tri.triId = i;
                            tri.sub = (uint32_t)(k - 2);
                            tris.push_back(tri);
                        } else {
                            cs.culled++;
                        }
                    }
                } else {
                    cs.rejected++;
                }
                continue;
            }
            const uint32_t index[3] = {idx0, idx1, idx2};
            RasterTriangle tri;
            if (tri.Setup(HomogeneousProject(vb[idx0].projectedPos), HomogeneousProject(vb[idx1].projectedPos),
                          HomogeneousProject(vb[idx2].projectedPos), index, coreId, raster_, p_.mutation)) {
                tri.triId = i;
                tri.sub = 0;
                tris.push_back(tri);
            } else {
                cs.culled++;
            }
        }
    });
    /* Core/Renderer.cpp:139-147: invW and the z divide on every post-clip vertex. */
    pool_->parallel_for(0, numCores, [&](int coreId) {
        for (auto& v : coreVB_[coreId]) {
            v.invW = 1.0f / v.projectedPos.w;
            v.projectedPos.z *= v.invW;
        }
    });
}

/* Core/FrameBuffer.cpp:54-68: tiled depth with a y-flipped quad row, one
 * 16x16-quad plane per sample. */
BoolSSE Renderer::ZTestQuad(const FloatSSE& d, int x, int y, int sampleId, BoolSSE mask) {
    const int tileX = x >> Tile::SIZE_LOG_2, tileY = y >> Tile::SIZE_LOG_2;
    const int ix = x & (Tile::SIZE - 1), iy = y & (Tile::SIZE - 1);
    FloatSSE& cur = depth_[((size_t)(tileY * tileDim_.x + tileX) * sampleCount_ + sampleId) * 256 +
                           ((Tile::SIZE - 1 - iy) >> 1) * 16 + (ix >> 1)];
    BoolSSE ret = LessEq(d, cur);
    if (p_.mutation == 2) ret = IntSSE(_mm_castps_si128(_mm_cmplt_ps(d.v, cur.v)));  /* planted: LESS */
    cur = Select(ret & mask, d, cur);
    return ret;
}

void Renderer::EmitFragment(Tile& tile, const TriangleSSE& t, const RasterTriangle& tri, int x, int y,
                            const CoverageMask& mask) {
    Fragment f;
    f.lambda0 = t.lambda0;
    f.lambda1 = t.lambda1;
    f.coverage = mask;
    f.x = x;
    f.y = y;
    f.vId0 = tri.vId0; f.vId1 = tri.vId1; f.vId2 = tri.vId2;
    f.coreId = tri.coreId;
    f.triId = tri.triId;
    f.tileId = tile.tileId;
    f.intraTileIdx = (uint32_t)tile.fragmentBuf.size();
    tile.fragmentBuf.push_back(f);
}

/* Pixel-centre offsets, Core/Rasterizer.h:23. */
static const IntSSE kCenterX(8, 24, 8, 24), kCenterY(8, 8, 24, 24);

/* Sample offsets in 28.4 units, (x, y) pairs per MSAA level: the table of
 * FrameBuffer::MultiSampleOffsets (Core/FrameBuffer.cpp:107-191). */
static const int kSampleOffsets[6][64] = {
    {0, 0},
    {4, 4, -4, -4},
    {-2, -6, 6, -2, -6, 2, 2, 6},
    {1, -3, -1, 3, 5, 1, -3, -5, -5, 5, -7, -1, 3, 7, 7, -7},
    {1, 1, -1, -3, -3, 2, 4, -1, -5, -2, 2, 5, 5, 3, 3, -5, -2, 6, 0, -7, -4, -6, -6, 4, -8, 0, 7, -4, 6, 7, -7, -8},
    {1, 1, -1, -3, -3, 2, 4, -1, -5, -2, 2, 5, 5, 3, 3, -5, -2, 6, 0, -7, -4, -6, -6, 4, -8, 0, 7, -4, 6, 7, -7, -8,
     1, 3, -3, -3, -3, 0, 6, -2, -7, -1, 3, 4, 7, 3, 3, -6, -2, 7, 0, -4, -2, -5, -7, 6, -8, 3, 4, -1, 2, 7, 4, -8}};

/* Core/Rasterizer.h:113-124 (FineRasterize dispatch on the sample count). */
void Renderer::FineRasterize(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri) {
    if (sampleCount_ == 1) FineRasterizeSingle(tile, bmin, bmax, tri);
    else FineRasterizeMulti(tile, bmin, bmax, tri);
}

/* Core/Rasterizer.h:202-292 (FineRasterize_MultiSample): coverage and depth per
 * sample at centre + offset; the fragment's barycentrics at the pixel centre. */
void Renderer::FineRasterizeMulti(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri) {
    int minX = std::max(bmin.x, std::min(tri.v0.x, std::min(tri.v1.x, tri.v2.x)) >> 4);
    int maxX = std::min(bmax.x - 1, std::max(tri.v0.x, std::max(tri.v1.x, tri.v2.x)) >> 4);
    int minY = std::max(bmin.y, std::min(tri.v0.y, std::min(tri.v1.y, tri.v2.y)) >> 4);
    int maxY = std::min(bmax.y - 1, std::max(tri.v0.y, std::max(tri.v1.y, tri.v2.y)) >> 4);
    minX -= minX % 2;
    minY -= minY % 2;
    if (maxX < minX || maxY < minY) return;

    TriangleSSE t(tri, p_.mutation);
    const ProjectedVertex& pv0 = coreVB_[tri.coreId][tri.vId0];
    const ProjectedVertex& pv1 = coreVB_[tri.coreId][tri.vId1];
    const ProjectedVertex& pv2 = coreVB_[tri.coreId][tri.vId2];
    const IntSSE B0(tri.B0), C0(tri.C0), B1(tri.B1), C1(tri.C1), B2(tri.B2), C2(tri.C2);

    IntSSE baseX(minX << 4), baseY(minY << 4);
    IntSSE e0 = t.EdgeFunc0(baseX + kCenterX, baseY + kCenterY);
    IntSSE e1 = t.EdgeFunc1(baseX + kCenterX, baseY + kCenterY);
    IntSSE e2 = t.EdgeFunc2(baseX + kCenterX, baseY + kCenterY);
    const IntSSE step(32);
    for (int y = minY; y <= maxY; y += 2, baseY += step) {
        IntSSE ey0 = e0, ey1 = e1, ey2 = e2;
        baseX = IntSSE(minX << 4);
        for (int x = minX; x <= maxX; x += 2, baseX += step) {
            const IntSSE cx = baseX + kCenterX, cy = baseY + kCenterY;
            bool genFragment = false;
            CoverageMask mask;
            for (int sId = 0; sId < sampleCount_; sId++) {
                const int ox = kSampleOffsets[msLevel_][2 * sId], oy = kSampleOffsets[msLevel_][2 * sId + 1];
                const IntSSE se0 = e0 + IntSSE(ox) * B0 + IntSSE(oy) * C0;
                const IntSSE se1 = e1 + IntSSE(ox) * B1 + IntSSE(oy) * C1;
                const IntSSE se2 = e2 + IntSSE(ox) * B2 + IntSSE(oy) * C2;
                const BoolSSE covered = GreaterEqZero(se0 | se1 | se2);
                if (Any(covered)) {
                    t.CalcBarycentricCoord(cx + IntSSE(ox), cy + IntSSE(oy));
                    BoolSSE z = ZTestQuad(t.GetDepth(pv0.projectedPos.z, pv1.projectedPos.z, pv2.projectedPos.z), x,
                                          y, sId, covered);
                    BoolSSE visible = z & covered;
                    if (Any(visible)) {
                        mask.SetBit(visible, sId);
                        genFragment = true;
                    }
                }
            }
            if (genFragment) {
                t.CalcBarycentricCoord(cx, cy);
                EmitFragment(tile, t, tri, x, y, mask);
            }
            e0 += t.stepB0;
            e1 += t.stepB1;
            e2 += t.stepB2;
        }
        e0 = ey0 + t.stepC0;
        e1 = ey1 + t.stepC1;
        e2 = ey2 + t.stepC2;
    }
}

/* Core/Rasterizer.h:126-200 (FineRasterize_SingleSample). */
void Renderer::FineRasterizeSingle(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri) {
    int minX = std::max(bmin.x, std::min(tri.v0.x, std::min(tri.v1.x, tri.v2.x)) >> 4);
    int maxX = std::min(bmax.x - 1, std::max(tri.v0.x, std::max(tri.v1.x, tri.v2.x)) >> 4);
    int minY = std::max(bmin.y, std::min(tri.v0.y, std::min(tri.v1.y, tri.v2.y)) >> 4);
    int maxY = std::min(bmax.y - 1, std::max(tri.v0.y, std::max(tri.v1.y, tri.v2.y)) >> 4);
    minX -= minX % 2;
    minY -= minY % 2;
    if (maxX < minX || maxY < minY) return;

    TriangleSSE t(tri, p_.mutation);
    const ProjectedVertex& pv0 = coreVB_[tri.coreId][tri.vId0];
    const ProjectedVertex& pv1 = coreVB_[tri.coreId][tri.vId1];
    const ProjectedVertex& pv2 = coreVB_[tri.coreId][tri.vId2];

    IntSSE baseX(minX << 4), baseY(minY << 4);
    IntSSE e0 = t.EdgeFunc0(baseX + kCenterX, baseY + kCenterY);
    IntSSE e1 = t.EdgeFunc1(baseX + kCenterX, baseY + kCenterY);
    IntSSE e2 = t.EdgeFunc2(baseX + kCenterX, baseY + kCenterY);
    const IntSSE step(32);
    for (int y = minY; y <= maxY; y += 2, baseY += step) {
        IntSSE ey0 = e0, ey1 = e1, ey2 = e2;
        baseX = IntSSE(minX << 4);
        for (int x = minX; x <= maxX; x += 2, baseX += step) {
            BoolSSE covered = GreaterEqZero(e0 | e1 | e2);
            if (Any(covered)) {
                t.CalcBarycentricCoord(baseX + kCenterX, baseY + kCenterY);
                BoolSSE z = ZTestQuad(t.GetDepth(pv0.projectedPos.z, pv1.projectedPos.z, pv2.projectedPos.z), x, y, 0,
                                      covered);
                BoolSSE visible = z & covered;
                if (Any(visible)) {
                    CoverageMask m;
                    m.SetBit(visible, 0);
                    EmitFragment(tile, t, tri, x, y, m);
                }
            }
            e0 += t.stepB0;
            e1 += t.stepB1;
            e2 += t.stepB2;
        }
        e0 = ey0 + t.stepC0;
        e1 = ey1 + t.stepC1;
        e2 = ey2 + t.stepC2;
    }
}

/* Core/Rasterizer.h:302-308 (TrivialAcceptTriangle dispatch). */
void Renderer::TrivialAcceptTriangle(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri) {
    if (sampleCount_ == 1) TrivialAcceptSingle(tile, bmin, bmax, tri);
    else TrivialAcceptMulti(tile, bmin, bmax, tri);
}

/* Core/Rasterizer.h:355-414 (TrivialAcceptTriangle_MultiSample). */
void Renderer::TrivialAcceptMulti(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri) {
    int minX = bmin.x, maxX = bmax.x - 1, minY = bmin.y, maxY = bmax.y - 1;
    minX -= minX % 2;
    minY -= minY % 2;
    TriangleSSE t(tri, p_.mutation);
    const ProjectedVertex& pv0 = coreVB_[tri.coreId][tri.vId0];
    const ProjectedVertex& pv1 = coreVB_[tri.coreId][tri.vId1];
    const ProjectedVertex& pv2 = coreVB_[tri.coreId][tri.vId2];
    const IntSSE allTrue(-1);
    for (int y = minY; y <= maxY; y += 2) {
        for (int x = minX; x <= maxX; x += 2) {
            const IntSSE cx = IntSSE(x << 4) + kCenterX, cy = IntSSE(y << 4) + kCenterY;
            CoverageMask mask;
            bool genFragment = false;
            for (int sId = 0; sId < sampleCount_; sId++) {
                const int ox = kSampleOffsets[msLevel_][2 * sId], oy = kSampleOffsets[msLevel_][2 * sId + 1];
                t.CalcBarycentricCoord(cx + IntSSE(ox), cy + IntSSE(oy));
                BoolSSE z = ZTestQuad(t.GetDepth(pv0.projectedPos.z, pv1.projectedPos.z, pv2.projectedPos.z), x, y,
                                      sId, allTrue);
                if (Any(z)) {
                    mask.SetBit(z, sId);
                    genFragment = true;
                }
            }
            if (genFragment) {
                t.CalcBarycentricCoord(cx, cy);
                EmitFragment(tile, t, tri, x, y, mask);
            }
        }
    }
}

/* Core/Rasterizer.h:310-353 (TrivialAcceptTriangle_SingleSample). */
void Renderer::TrivialAcceptSingle(Tile& tile, Vec2i bmin, Vec2i bmax, const RasterTriangle& tri) {
    int minX = bmin.x, maxX = bmax.x - 1, minY = bmin.y, maxY = bmax.y - 1;
    minX -= minX % 2;
    minY -= minY % 2;
    TriangleSSE t(tri, p_.mutation);
    const ProjectedVertex& pv0 = coreVB_[tri.coreId][tri.vId0];
    const ProjectedVertex& pv1 = coreVB_[tri.coreId][tri.vId1];
    const ProjectedVertex& pv2 = coreVB_[tri.coreId][tri.vId2];
    const IntSSE allTrue(-1);
    for (int y = minY; y <= maxY; y += 2) {
        for (int x = minX; x <= maxX; x += 2) {
            t.CalcBarycentricCoord(IntSSE(x << 4) + kCenterX, IntSSE(y << 4) + kCenterY);
            BoolSSE z = ZTestQuad(t.GetDepth(pv0.projectedPos.z, pv1.projectedPos.z, pv2.projectedPos.z), x, y, 0,
                                  allTrue);
            if (Any(z)) {
                CoverageMask m;
                m.SetBit(z, 0);
                EmitFragment(tile, t, tri, x, y, m);
            }
        }
    }
}

/* Core/Rasterizer.h:32-111 (CoarseRasterize, single sample). */
void Renderer::CoarseRasterize(Tile& tile, const TriangleRef& ref, int blockSize, Vec2i bmin, Vec2i bmax,
                               const RasterTriangle& tri) {
    const Vec2i base{bmin.x << 4, bmin.y << 4};
    const int stepSize = blockSize << 3;
    const int off = (blockSize << 4) - 1;
    auto corner = [&](int c) { return Vec2i{base.x + (c % 2) * off, base.y + (c >> 1) * off}; };
    IntSSE rej0 = IntSSE(tri.EdgeFunc0(corner(tri.rejectCorner0))) + RasterTriangle::Step(tri.rejectCorner0, tri.B0, tri.C0, stepSize);
    IntSSE rej1 = IntSSE(tri.EdgeFunc1(corner(tri.rejectCorner1))) + RasterTriangle::Step(tri.rejectCorner1, tri.B1, tri.C1, stepSize);
    IntSSE rej2 = IntSSE(tri.EdgeFunc2(corner(tri.rejectCorner2))) + RasterTriangle::Step(tri.rejectCorner2, tri.B2, tri.C2, stepSize);
    IntSSE acc0(INT_MAX), acc1(INT_MAX), acc2(INT_MAX);
    if (!ref.acceptEdge0)
        acc0 = IntSSE(tri.EdgeFunc0(corner(tri.acceptCorner0))) + RasterTriangle::Step(tri.acceptCorner0, tri.B0, tri.C0, stepSize);
    if (!ref.acceptEdge1)
        acc1 = IntSSE(tri.EdgeFunc1(corner(tri.acceptCorner1))) + RasterTriangle::Step(tri.acceptCorner1, tri.B1, tri.C1, stepSize);
    if (!ref.acceptEdge2)
        acc2 = IntSSE(tri.EdgeFunc2(corner(tri.acceptCorner2))) + RasterTriangle::Step(tri.acceptCorner2, tri.B2, tri.C2, stepSize);
    BoolSSE rejMask = LessZero(rej0) | LessZero(rej1) | LessZero(rej2);
    BoolSSE accMask = GreaterEqZero(acc0) & GreaterEqZero(acc1) & GreaterEqZero(acc2);
    for (int i = 0; i < 4; i++) {
        if (rejMask[i] != 0) continue;
        int minX = !(i % 2) ? bmin.x : bmin.x + (blockSize >> 1);
        int maxX = !(i % 2) ? bmin.x + (blockSize >> 1) : bmax.x;
        int minY = !(i >> 1) ? bmin.y : bmin.y + (blockSize >> 1);
        int maxY = !(i >> 1) ? bmin.y + (blockSize >> 1) : bmax.y;
        if (accMask[i] != 0) {
            TrivialAcceptTriangle(tile, {minX, minY}, {maxX, maxY}, tri);
            continue;
        }
        FineRasterize(tile, {minX, minY}, {maxX, maxY}, tri);
    }
}

/* Core/Renderer.cpp:248-270. */
void Renderer::RasterizeTile(Tile& tile) {
    for (int coreId = 0; coreId < numCores_; coreId++) {
        for (const TriangleRef& ref : tile.triangleRefs[coreId]) {
            const RasterTriangle& tri = coreTris_[coreId][ref.triId];
            if (ref.trivialAccept) {
                TrivialAcceptTriangle(tile, tile.minCoord, tile.maxCoord, tri);
                continue;
            }
            if (p_.hierarchical && ref.big)
                CoarseRasterize(tile, ref, Tile::SIZE, tile.minCoord, tile.maxCoord, tri);
            else
                FineRasterize(tile, tile.minCoord, tile.maxCoord, tri);
        }
    }
}

/* Core/Renderer.cpp:150-246. */
void Renderer::TiledRasterization() {
    pool_->parallel_for(0, (int)tiles_.size(), [&](int i) {
        for (auto& l : tiles_[i].triangleRefs) l.clear();
        tiles_[i].fragmentBuf.clear();
    });
    const int Shift = Tile::SIZE_LOG_2 + 4;
    pool_->parallel_for(0, numCores_, [&](int coreId) {
        const auto& tris = coreTris_[coreId];
        for (uint32_t i = 0; i < (uint32_t)tris.size(); i++) {
            const RasterTriangle& tri = tris[i];
            int minX = std::max(0, std::min(tri.v0.x, std::min(tri.v1.x, tri.v2.x)) >> Shift);
            int maxX = std::min(tileDim_.x - 1, std::max(tri.v0.x, std::max(tri.v1.x, tri.v2.x)) >> Shift);
            int minY = std::max(0, std::min(tri.v0.y, std::min(tri.v1.y, tri.v2.y)) >> Shift);
            int maxY = std::min(tileDim_.y - 1, std::max(tri.v0.y, std::max(tri.v1.y, tri.v2.y)) >> Shift);
            if (maxX - minX < 2 && maxY - minY < 2) {
                for (int y = minY; y <= maxY; y++)
                    for (int x = minX; x <= maxX; x++) tiles_[y * tileDim_.x + x].triangleRefs[coreId].emplace_back(i);
            } else {
                for (int y = minY; y <= maxY; y++) {
                    for (int x = minX; x <= maxX; x++) {
                        auto cornerOf = [&](int c) { return Vec2i{(x + c % 2) << Shift, (y + c / 2) << Shift}; };
                        if (tri.EdgeFunc0(cornerOf(tri.rejectCorner0)) < 0 || tri.EdgeFunc1(cornerOf(tri.rejectCorner1)) < 0 ||
                            tri.EdgeFunc2(cornerOf(tri.rejectCorner2)) < 0)
                            continue;
                        tiles_[y * tileDim_.x + x].triangleRefs[coreId].emplace_back(
                            i, tri.EdgeFunc0(cornerOf(tri.acceptCorner0)) >= 0,
                            tri.EdgeFunc1(cornerOf(tri.acceptCorner1)) >= 0,
                            tri.EdgeFunc2(cornerOf(tri.acceptCorner2)) >= 0, true);
                    }
                }
            }
        }
    });
    pool_->parallel_for(0, (int)tiles_.size(), [&](int i) { RasterizeTile(tiles_[i]); });

    /* Core/Renderer.cpp:238-245: serial concatenation of the tile fragment lists. */
    fragPtrs_.clear();
    shaded_.resize(tiles_.size());
    for (size_t i = 0; i < tiles_.size(); i++) {
        shaded_[i].resize(tiles_[i].fragmentBuf.size() * 4);
        for (auto& f : tiles_[i].fragmentBuf) fragPtrs_.push_back(&f);
    }
}

/* Core/Renderer.cpp:272-303 with Fragment::Interpolate (Core/Shader.h:142-170) and the
 * pixel shaders (Core/Shader.h:209-282). */
void Renderer::FragmentProcessing() {
    const Vec3 eyePos = TransformPoint(Vec3{0, 0, 0}, mvInv_);
    const Vec3 lightDir{p_.light[0], p_.light[1], p_.light[2]};
    pool_->parallel_for(0, (int)fragPtrs_.size(), [&](int i) {
        const Fragment& f = *fragPtrs_[i];
        const ProjectedVertex& v0 = coreVB_[f.coreId][f.vId0];
        const ProjectedVertex& v1 = coreVB_[f.coreId][f.vId1];
        const ProjectedVertex& v2 = coreVB_[f.coreId][f.vId2];
        /* Interpolate */
        const FloatSSE one(1.0f);
        FloatSSE b0 = f.lambda0, b1 = f.lambda1;
        FloatSSE b2 = one - b0 - b1;
        if (p_.mutation != 3) {  /* planted when 3: affine (no 1/w) interpolation */
            b0 *= FloatSSE(v0.invW);
            b1 *= FloatSSE(v1.invW);
            b2 *= FloatSSE(v2.invW);
        }
        FloatSSE invB = one / (b0 + b1 + b2);
        b0 *= invB;
        b1 *= invB;
        b2 = one - b0 - b1;
        const Vec3SSE position = b0 * Splat(v0.position) + b1 * Splat(v1.position) + b2 * Splat(v2.position);
        const Vec3SSE normal = b0 * Splat(v0.normal) + b1 * Splat(v1.normal) + b2 * Splat(v2.normal);
        const Vec3SSE color = b0 * Splat(v0.color) + b1 * Splat(v1.color) + b2 * Splat(v2.color);
        const FloatSSE tu = b0 * FloatSSE(v0.u) + b1 * FloatSSE(v1.u) + b2 * FloatSSE(v2.u);
        const FloatSSE tv = b0 * FloatSSE(v0.v) + b1 * FloatSSE(v1.v) + b2 * FloatSSE(v2.v);

        /* Shade */
        FloatSSE w = Rsqrt(Dot(normal, normal));
        const Vec3SSE n = normal * w;
        const Vec3SSE L = Splat(Normalize(lightDir));
        FloatSSE diffuseAmount = Dot(L, n);
        BoolSSE neg = IntSSE(_mm_castps_si128(_mm_cmplt_ps(diffuseAmount.v, _mm_setzero_ps())));
        diffuseAmount = Select(neg, FloatSSE(0.0f), diffuseAmount);
        const FloatSSE diffuse = (diffuseAmount + FloatSSE(p_.ambient)) * FloatSSE(p_.kd) * FloatSSE((float)(1.0 / M_PI));
        Vec3SSE out;
        if (p_.shader == ORACLE_SHADER_LAMBERT_ALBEDO) {
            out = {diffuse * FloatSSE(p_.albedo[0]), diffuse * FloatSSE(p_.albedo[1]), diffuse * FloatSSE(p_.albedo[2])};
        } else if (p_.shader == ORACLE_SHADER_TEXTURED_LAMBERT) {
            /* Core/Shader.h:226-241: differentials from lanes 1 - 0 and 2 - 0,
             * one Sample per lane of the fragment's texture slot. */
            alignas(16) float u[4], v[4], a[3][4];
            _mm_store_ps(u, tu.v);
            _mm_store_ps(v, tv.v);
            const TextureSlot& slot = slots_[texIds_[f.triId]];
            for (int k = 0; k < 4; k++) {
                const Vec3 c = SampleSlot(slot, p_.tex_filter, u[k], v[k], u[1] - u[0], v[1] - v[0], u[2] - u[0],
                                          v[2] - v[0]);
                a[0][k] = c.x;
                a[1][k] = c.y;
                a[2][k] = c.z;
            }
            out = {diffuse * FloatSSE(_mm_load_ps(a[0])), diffuse * FloatSSE(_mm_load_ps(a[1])),
                   diffuse * FloatSSE(_mm_load_ps(a[2]))};
        } else {
            Vec3SSE eyeDir = Splat(eyePos) - position;
            w = Rsqrt(Dot(eyeDir, eyeDir));
            eyeDir = eyeDir * w;
            Vec3SSE halfVec = L + eyeDir;
            w = Rsqrt(Dot(halfVec, halfVec));
            halfVec = halfVec * w;
            FloatSSE specAmount = Dot(n, halfVec);
            alignas(16) float sa[4];
            _mm_store_ps(sa, specAmount.v);
            for (int k = 0; k < 4; k++) {
                float s = sa[k];
                if (p_.clamp_specular) s = fmaxf(s, 0.0f);
                sa[k] = PowDecl(s, p_.shininess);
            }
            const FloatSSE spec = FloatSSE(_mm_load_ps(sa)) * FloatSSE(p_.ks);
            if (p_.shader == ORACLE_SHADER_BLINN_PHONG_REF)
                out = {diffuse + spec, diffuse + spec, diffuse + spec};
            else
                out = {diffuse * color.x + spec, diffuse * color.y + spec, diffuse * color.z + spec};
        }
        uint32_t* dst = &shaded_[f.tileId][(size_t)f.intraTileIdx * 4];
        for (int k = 0; k < 4; k++)
            dst[k] = (uint32_t)ToByte(out.x[k]) | ((uint32_t)ToByte(out.y[k]) << 8) |
                     ((uint32_t)ToByte(out.z[k]) << 16) | (255u << 24);
    });
}

/* Core/Renderer.cpp:305-350 and FrameBuffer::SetPixel (Core/FrameBuffer.cpp:39-42). */
void Renderer::UpdateFrameBuffer() {
    const size_t P = (size_t)W_ * H_;
    pool_->parallel_for(0, (int)tiles_.size(), [&](int i) {
        const Tile& tile = tiles_[i];
        for (size_t j = 0; j < tile.fragmentBuf.size(); j++) {
            const Fragment& f = tile.fragmentBuf[j];
            for (int sId = 0; sId < sampleCount_; sId++) {
                for (int k = 0; k < 4; k++) {
                    if (!f.coverage.GetBit((sId << 2) + k)) continue;
                    const int x = f.x + (k & 1), y = f.y + (k >> 1);
                    if (x >= (int)W_ || y >= (int)H_) continue;
                    const size_t o = sId * P + (size_t)(H_ - 1 - y) * W_ + x;
                    memcpy(&colorMS_[o * 4], &shaded_[i][j * 4 + k], 4);
                    idsMS_[o] = f.triId;
                }
            }
        }
    });
    Resolve();
}

/* FrameBuffer::Resolve (Core/FrameBuffer.cpp:70-87): mean of the samples'
 * colours, with the declared Color <-> Color4b conversions.  One sample: the
 * sample buffer is the back buffer (Core/FrameBuffer.h:55-58). */
void Renderer::Resolve() {
    const size_t P = (size_t)W_ * H_;
    if (sampleCount_ == 1) {
        memcpy(color_.data(), colorMS_.data(), P * 4);
        return;
    }
    const float invSampleCount = 1.0f / float(sampleCount_);
    pool_->parallel_for(0, (int)H_, [&](int row) {
        for (uint32_t x = 0; x < W_; x++) {
            const size_t o = (size_t)row * W_ + x;
            float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int s = 0; s < sampleCount_; s++)
                for (int ch = 0; ch < 4; ch++) c[ch] += (float)colorMS_[(s * P + o) * 4 + ch] / 255.0f;
            for (int ch = 0; ch < 4; ch++) color_[o * 4 + ch] = ToByte(c[ch] * invSampleCount);
        }
    });
}

int Renderer::Render(const oracle_params& p, const oracle_mesh& m, oracle_stats* st) {
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    if (p.width == 0 || p.height == 0) return -1;
    for (uint64_t i = 0; i < 3ull * m.num_triangles; i++)
        if (m.indices[i] >= m.num_vertices) return -2;
    if (p.msaa_log2 < 0 || p.msaa_log2 > 5) return -1;
    if (p.shader == ORACLE_SHADER_TEXTURED_LAMBERT) {
        if (!m.tu || !m.tv || (m.num_triangles && !m.tex_ids) || !p.num_textures) return -1;
        for (uint32_t t = 0; t < m.num_triangles; t++)
            if (m.tex_ids[t] >= p.num_textures) return -2;
        slots_.clear();
        for (uint32_t k = 0; k < p.num_textures; k++) slots_.push_back(MakeSlot(p.textures[k]));
        texIds_ = m.tex_ids;
    }
    p_ = p;
    Initialize(p.width, p.height, p.num_threads);
    msLevel_ = p.msaa_log2;
    sampleCount_ = 1 << msLevel_;
    /* Core/Renderer.cpp:85-92: SetTransform. */
    const Mat4 mv = load_mat(p.mv), proj = load_mat(p.proj);
    raster_ = load_mat(p.raster);
    mvInv_ = Inverse(mv);
    mvp_ = p.mutation == 4 ? MatMul(mv, proj) : MatMul(proj, mv);  /* planted when 4: operand order */

    auto t0 = clk::now();
    /* FrameBuffer::Clear, Core/FrameBuffer.cpp:89-105. */
    color_.assign((size_t)W_ * H_ * 4, 0);
    colorMS_.assign((size_t)sampleCount_ * W_ * H_ * 4, 0);
    idsMS_.assign((size_t)sampleCount_ * W_ * H_, 0xFFFFFFFFu);
    depth_.resize(tiles_.size() * 256 * sampleCount_);
    pool_->parallel_for(0, (int)tiles_.size(), [&](int i) {
        for (int j = 0; j < 256 * sampleCount_; j++) depth_[(size_t)i * 256 * sampleCount_ + j] = FloatSSE(1.0f);
    });
    VertexProcessing(m);
    auto t1 = clk::now();
    Clipping(m);
    auto t2 = clk::now();
    TiledRasterization();
    auto t3 = clk::now();
    FragmentProcessing();
    auto t4 = clk::now();
    UpdateFrameBuffer();
    auto t5 = clk::now();

    /* Linear depth output and introspection (not timed). */
    const size_t P = (size_t)W_ * H_;
    depthMS_.assign(P * sampleCount_, 1.0f);
    for (int sId = 0; sId < sampleCount_; sId++)
        for (uint32_t y = 0; y < H_; y++)
            for (uint32_t x = 0; x < W_; x++) {
                const int tX = x >> 5, tY = y >> 5, ix = x & 31, iy = y & 31;
                const FloatSSE& q = depth_[((size_t)(tY * tileDim_.x + tX) * sampleCount_ + sId) * 256 +
                                           ((31 - (iy & ~1)) >> 1) * 16 + (ix >> 1)];
                depthMS_[sId * P + (size_t)(H_ - 1 - y) * W_ + x] = q[(ix & 1) + 2 * (iy & 1)];
            }
    depthLinear_.assign(depthMS_.begin(), depthMS_.begin() + P);
    ids_.assign(idsMS_.begin(), idsMS_.begin() + P);
    records_.clear();
    std::vector<uint64_t> coreBase(numCores_, 0);
    for (int c = 0; c < numCores_; c++) {
        coreBase[c] = records_.size();
        for (const auto& t : coreTris_[c]) {
            oracle_record r;
            r.x[0] = t.v0.x; r.y[0] = t.v0.y; r.x[1] = t.v1.x; r.y[1] = t.v1.y; r.x[2] = t.v2.x; r.y[2] = t.v2.y;
            r.inv_det = t.invDet;
            const uint32_t vid[3] = {t.vId0, t.vId1, t.vId2};
            for (int k = 0; k < 3; k++) {
                r.z[k] = coreVB_[c][vid[k]].projectedPos.z;
                r.inv_w[k] = coreVB_[c][vid[k]].invW;
            }
            r.tri_id = t.triId;
            r.sub = t.sub;
            r.corners = t.rejectCorner0 | (t.rejectCorner1 << 2) | (t.rejectCorner2 << 4) | (t.acceptCorner0 << 6) |
                        (t.acceptCorner1 << 8) | (t.acceptCorner2 << 10);
            records_.push_back(r);
        }
    }
    binOffsets_.assign(tiles_.size() + 1, 0);
    binRefs_.clear();
    for (size_t i = 0; i < tiles_.size(); i++) {
        binOffsets_[i] = binRefs_.size();
        for (int c = 0; c < numCores_; c++)
            for (const auto& r : tiles_[i].triangleRefs[c]) {
                oracle_ref o;
                o.record = (uint32_t)(coreBase[c] + r.triId);
                o.flags = (r.acceptEdge0 ? 1u : 0u) | (r.acceptEdge1 ? 2u : 0u) | (r.acceptEdge2 ? 4u : 0u) |
                          (r.trivialAccept ? 8u : 0u) | (r.big ? 16u : 0u);
                binRefs_.push_back(o);
            }
    }
    binOffsets_[tiles_.size()] = binRefs_.size();

    if (st) {
        memset(st, 0, sizeof(*st));
        st->triangles_in = m.num_triangles;
        for (const auto& cs : coreStats_) {
            st->trivially_rejected += cs.rejected;
            st->clipped += cs.clipped;
            st->near_clipped += cs.near_clipped;
            st->clip_dropped += cs.dropped;
            st->culled += cs.culled;
            st->clip_vertices += cs.clip_vertices;
        }
        st->setup_records = records_.size();
        st->refs = binRefs_.size();
        st->fragments = fragPtrs_.size();
        st->ms_vs = ms(t0, t1);
        st->ms_clip = ms(t1, t2);
        st->ms_bin = ms(t2, t3);  /* binning + raster (the reference runs them in one stage) */
        st->ms_raster = 0.0;
        st->ms_shade = ms(t3, t4);
        st->ms_write = ms(t4, t5);
        st->ms_total = ms(t0, t5);
        st->threads = numCores_;
    }
    return 0;
}

}  // namespace orc

/* ------------------------------------------------------------------ */
/* C interface                                                           */
/* ------------------------------------------------------------------ */
struct oracle_ctx {
    orc::Renderer r;
};

extern "C" {

oracle_ctx* oracle_create(void) { return new oracle_ctx(); }
void oracle_destroy(oracle_ctx* c) { delete c; }

int oracle_render(oracle_ctx* c, const oracle_params* p, const oracle_mesh* m, oracle_stats* st) {
    if (!c || !p || !m) return -1;
    return c->r.Render(*p, *m, st);
}

int oracle_read_color(const oracle_ctx* c, uint8_t* rgba, uint64_t bytes) {
    if (bytes < c->r.color_.size()) return -1;
    memcpy(rgba, c->r.color_.data(), c->r.color_.size());
    return 0;
}
int oracle_read_depth(const oracle_ctx* c, float* d, uint64_t count) {
    if (count < c->r.depthLinear_.size()) return -1;
    memcpy(d, c->r.depthLinear_.data(), c->r.depthLinear_.size() * 4);
    return 0;
}
int oracle_read_ids(const oracle_ctx* c, uint32_t* ids, uint64_t count) {
    if (count < c->r.ids_.size()) return -1;
    memcpy(ids, c->r.ids_.data(), c->r.ids_.size() * 4);
    return 0;
}
int oracle_read_sample_depth(const oracle_ctx* c, float* depth, uint64_t count) {
    if (count < c->r.depthMS_.size()) return -1;
    memcpy(depth, c->r.depthMS_.data(), c->r.depthMS_.size() * 4);
    return 0;
}
int oracle_read_sample_ids(const oracle_ctx* c, uint32_t* ids, uint64_t count) {
    if (count < c->r.idsMS_.size()) return -1;
    memcpy(ids, c->r.idsMS_.data(), c->r.idsMS_.size() * 4);
    return 0;
}
uint64_t oracle_num_records(const oracle_ctx* c) { return c->r.records_.size(); }
int oracle_read_records(const oracle_ctx* c, oracle_record* out, uint64_t count) {
    if (count < c->r.records_.size()) return -1;
    memcpy(out, c->r.records_.data(), c->r.records_.size() * sizeof(oracle_record));
    return 0;
}
uint64_t oracle_num_tiles(const oracle_ctx* c) {
    return c->r.binOffsets_.empty() ? 0 : c->r.binOffsets_.size() - 1;
}
int oracle_read_bins(const oracle_ctx* c, uint64_t* offsets, oracle_ref* refs, uint64_t cap) {
    if (cap < c->r.binRefs_.size()) return -1;
    memcpy(offsets, c->r.binOffsets_.data(), c->r.binOffsets_.size() * 8);
    memcpy(refs, c->r.binRefs_.data(), c->r.binRefs_.size() * sizeof(oracle_ref));
    return 0;
}

void oracle_mat_mul(const float* a, const float* b, float* out) {
    orc::Mat4 r = orc::MatMul(orc::load_mat(a), orc::load_mat(b));
    for (int i = 0; i < 16; i++) out[i] = r.m[i / 4][i % 4];
}
void oracle_mat_inverse(const float* m, float* out) {
    orc::Mat4 r = orc::Inverse(orc::load_mat(m));
    for (int i = 0; i < 16; i++) out[i] = r.m[i / 4][i % 4];
}
void oracle_transform_point4(const float* m, const float* v, float* out) {
    orc::Vec4 r = orc::TransformPoint(orc::Vec4{v[0], v[1], v[2], v[3]}, orc::load_mat(m));
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

}  // extern "C"
