/*
 * edxraster.hpp -- header-only C++ facade over the C ABI (edxraster.h) that
 * re-creates the reference's host API, so that EDXRaster host code (the
 * RealtimeViewer's OnInit / OnRender, RealtimeViewer/Main.cpp:32-131)
 * compiles unchanged apart from its includes:
 *
 *   EDX::RasterRenderer::Renderer       Core/Renderer.h:15-59
 *   EDX::RasterRenderer::Mesh           Utils/Mesh.h:17-69
 *   EDX::RasterRenderer::IVertexBuffer, VertexBuffer<>, IndexBuffer,
 *       Vertex_PositionNormalTex, CreateVertexBuffer, CreateIndexBuffer
 *                                       Utils/InputBuffer.h:10-194
 *   EDX::RasterRenderer::PixelShader    Core/Shader.h:173-282 (shader objects)
 *   EDX::Vector2/3, Matrix, BoundingBox, Camera, Array, Color, TextureFilter
 *                                       the EDXUtil types the viewer uses
 *
 * EDXUtil is not available, so its types are declared here with the
 * semantics stated on each (the C ABI's conventions: row-major matrices with
 * column vectors, a left-handed look-at, a D3D projection, the raster matrix
 * x_r = (x + 1) W / 2, y_r = (1 - y) H / 2).  AoS vertex buffers are
 * converted to the ABI's SoA planes when a mesh is first rendered.  The
 * reference reports no errors; this facade throws std::runtime_error with
 * edxr_last_error() text when the C ABI returns a negative status.
 */
#ifndef EDXRASTER_HPP
#define EDXRASTER_HPP

#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "edxraster.h"

namespace EDX {

typedef unsigned int uint;
typedef unsigned char _byte;

struct Vector2 {
    float u = 0.0f, v = 0.0f;
    Vector2() = default;
    Vector2(float a, float b) : u(a), v(b) {}
    static const Vector2 ZERO;
};
inline const Vector2 Vector2::ZERO{0.0f, 0.0f};

struct Vector3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    Vector3() = default;
    Vector3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit Vector3(float s) : x(s), y(s), z(s) {}
    Vector3 operator+(const Vector3& o) const { return Vector3(x + o.x, y + o.y, z + o.z); }
    Vector3 operator-(const Vector3& o) const { return Vector3(x - o.x, y - o.y, z - o.z); }
    Vector3 operator-() const { return Vector3(-x, -y, -z); }
    Vector3 operator*(float s) const { return Vector3(x * s, y * s, z * s); }
    float Length() const { return std::sqrt(x * x + y * y + z * z); }
    const float* Data() const { return &x; }
    static const Vector3 ZERO, UNIT_X, UNIT_Y, UNIT_Z, UNIT_SCALE;
};
inline Vector3 operator*(float s, const Vector3& v) { return v * s; }
inline const Vector3 Vector3::ZERO{0.0f, 0.0f, 0.0f};
inline const Vector3 Vector3::UNIT_X{1.0f, 0.0f, 0.0f};
inline const Vector3 Vector3::UNIT_Y{0.0f, 1.0f, 0.0f};
inline const Vector3 Vector3::UNIT_Z{0.0f, 0.0f, 1.0f};
inline const Vector3 Vector3::UNIT_SCALE{1.0f, 1.0f, 1.0f};

/* Linear RGBA colour (EDXUtil Color). */
struct Color {
    float r = 0.0f, g = 0.0f, b = 0.0f, a = 1.0f;
    Color() = default;
    Color(float R, float G, float B, float A = 1.0f) : r(R), g(G), b(B), a(A) {}
    static const Color WHITE;
};
inline const Color Color::WHITE{1.0f, 1.0f, 1.0f, 1.0f};

/* EDXUtil's Array: a growable array with its member names. */
template <class T>
class Array {
  public:
    uint Size() const { return (uint)v_.size(); }
    T* Data() { return v_.data(); }
    const T* Data() const { return v_.data(); }
    void Add(const T& x) { v_.push_back(x); }
    void Resize(uint n) { v_.resize(n); }
    void Clear() { v_.clear(); }
    T& operator[](uint i) { return v_[i]; }
    const T& operator[](uint i) const { return v_[i]; }
    std::vector<T>& Vector() { return v_; }
    const std::vector<T>& Vector() const { return v_; }

  private:
    std::vector<T> v_;
};

/* Row-major 4x4, column-vector convention: out_r = sum_c m[r][c] * v_c. */
struct Matrix {
    float m[4][4];
    static Matrix Identity() {
        Matrix r{};
        for (int i = 0; i < 4; i++) r.m[i][i] = 1.0f;
        return r;
    }
    static Matrix LookAt(const float eye[3], const float at[3], const float up[3]) {
        Matrix r;
        edxr_look_at(eye, at, up, &r.m[0][0]);
        return r;
    }
    static Matrix LookAt(const Vector3& eye, const Vector3& at, const Vector3& up) {
        return LookAt(eye.Data(), at.Data(), up.Data());
    }
    static Matrix Perspective(float fovYDeg, float aspect, float znear, float zfar) {
        Matrix r;
        edxr_perspective(fovYDeg, aspect, znear, zfar, &r.m[0][0]);
        return r;
    }
    static Matrix ToRaster(uint32_t width, uint32_t height) {
        Matrix r;
        edxr_raster_matrix(width, height, &r.m[0][0]);
        return r;
    }
};

/* Axis-aligned box (EDXUtil BoundingBox). */
struct BoundingBox {
    Vector3 mMin{1e30f, 1e30f, 1e30f}, mMax{-1e30f, -1e30f, -1e30f};
    void Expand(const Vector3& p) {
        mMin = Vector3(std::fmin(mMin.x, p.x), std::fmin(mMin.y, p.y), std::fmin(mMin.z, p.z));
        mMax = Vector3(std::fmax(mMax.x, p.x), std::fmax(mMax.y, p.y), std::fmax(mMax.z, p.z));
    }
    /* Centre = the box's midpoint, radius = the distance from it to mMax
     * (0 for an empty box). */
    void BoundingSphere(Vector3* center, float* radius) const {
        *center = 0.5f * (mMin + mMax);
        *radius = mMin.x <= mMax.x ? (mMax - *center).Length() : 0.0f;
    }
};

/* The viewer's camera (EDXUtil Camera) without input devices: Init sets a
 * left-handed look-at from pos to target, a D3D perspective of `fov` degrees
 * vertically with aspect resX / resY, and the raster matrix of resX x resY;
 * Transform() (mouse and keyboard motion in the viewer) changes nothing. */
class Camera {
  public:
    float mMoveScaler = 1.0f, mRotateScaler = 1.0f;

    void Init(const Vector3& pos = Vector3::ZERO, const Vector3& target = Vector3::UNIT_Z,
              const Vector3& up = Vector3::UNIT_Y, int resX = 1280, int resY = 720, float fov = 65.0f,
              float nearClip = 0.1f, float farClip = 1000.0f) {
        pos_ = pos;
        target_ = target;
        up_ = up;
        fov_ = fov;
        near_ = nearClip;
        far_ = farClip;
        Resize(resX, resY);
    }
    void Resize(int resX, int resY) {
        resX_ = resX;
        resY_ = resY;
        view_ = Matrix::LookAt(pos_, target_, up_);
        proj_ = Matrix::Perspective(fov_, (float)resX / (float)resY, near_, far_);
        raster_ = Matrix::ToRaster((uint32_t)resX, (uint32_t)resY);
    }
    void Transform() {}
    const Matrix& GetViewMatrix() const { return view_; }
    const Matrix& GetProjMatrix() const { return proj_; }
    const Matrix& GetRasterMatrix() const { return raster_; }

  private:
    Vector3 pos_, target_{0.0f, 0.0f, 1.0f}, up_{0.0f, 1.0f, 0.0f};
    float fov_ = 65.0f, near_ = 0.1f, far_ = 1000.0f;
    int resX_ = 1280, resY_ = 720;
    Matrix view_ = Matrix::Identity(), proj_ = Matrix::Identity(), raster_ = Matrix::Identity();
};

/* Texture filters (EDXUtil TextureFilter; RenderStates::TexFilter). */
enum class TextureFilter { Nearest = 0, Linear = 1, TriLinear = 2, Anisotropic4x = 3, Anisotropic8x = 4, Anisotropic16x = 5 };

namespace RasterRenderer {

/* Utils/InputBuffer.h:10-41.  The byte layouts are the structures' own
 * (Size = sizeof): 32 bytes for Vertex_PositionNormalTex, 40 for
 * Vertex_PositionNormalColor with its four-float Color. */
enum class VertexFormat { PositionNormalTex, PositionNormalColor };

struct Vertex_PositionNormalTex {
    Vector3 Position;
    Vector3 Normal;
    Vector2 TexCoord;
    static const VertexFormat Format = VertexFormat::PositionNormalTex;
    static const int Size = 32;
    static const int PosOffset = 0;
    static const int NormalOffset = 12;
    static const int TexOffset = 24;
    static const int ColorOffset = -1;
};
static_assert(sizeof(Vertex_PositionNormalTex) == 32, "Vertex_PositionNormalTex layout");

struct Vertex_PositionNormalColor {
    Vector3 Position;
    Vector3 Normal;
    EDX::Color Color;
    static const VertexFormat Format = VertexFormat::PositionNormalColor;
    static const int Size = 40;
    static const int PosOffset = 0;
    static const int NormalOffset = 12;
    static const int TexOffset = -1;
    static const int ColorOffset = 24;
};
static_assert(sizeof(Vertex_PositionNormalColor) == 40, "Vertex_PositionNormalColor layout");

/* IVertexBuffer (Utils/InputBuffer.h:43-66): AoS vertices behind virtual
 * per-attribute fetches. */
class IVertexBuffer {
  protected:
    size_t mBufSize = 0;
    uint mVertexCount = 0;

  public:
    virtual ~IVertexBuffer() {}
    virtual void NewBuffer(const uint vertexCount) = 0;
    virtual void* GetBuffer() const = 0;
    virtual VertexFormat GetVertexFormat() const = 0;
    virtual int GetVertexSize() const = 0;
    virtual size_t GetBufferSize() const = 0;
    virtual Vector3 GetPosition(const uint idx) const = 0;
    virtual Vector3 GetNormal(const uint idx) const = 0;
    virtual Vector2 GetTexCoord(const uint idx) const = 0;
    virtual Color GetColor(const uint idx) const = 0;
    virtual void Release() = 0;
    uint GetVertexCount() const { return mVertexCount; }
};

/* VertexBuffer<VertexType> (Utils/InputBuffer.h:68-134). */
template <typename VertexType = Vertex_PositionNormalTex>
class VertexBuffer : public IVertexBuffer {
  public:
    ~VertexBuffer() { Release(); }
    void NewBuffer(const uint vertexCount) {
        Release();
        mVertexCount = vertexCount;
        mBufSize = (size_t)vertexCount * VertexType::Size;
        buf_.reset(new _byte[mBufSize ? mBufSize : 1]);
    }
    void* GetBuffer() const { return buf_.get(); }
    VertexFormat GetVertexFormat() const { return VertexType::Format; }
    int GetVertexSize() const { return VertexType::Size; }
    size_t GetBufferSize() const { return mBufSize; }
    Vector3 GetPosition(const uint idx) const { return at<Vector3>(idx, VertexType::PosOffset, Vector3::ZERO); }
    Vector3 GetNormal(const uint idx) const { return at<Vector3>(idx, VertexType::NormalOffset, Vector3::ZERO); }
    Vector2 GetTexCoord(const uint idx) const { return at<Vector2>(idx, VertexType::TexOffset, Vector2::ZERO); }
    Color GetColor(const uint idx) const { return at<Color>(idx, VertexType::ColorOffset, Color::WHITE); }
    void Release() {
        buf_.reset();
        mVertexCount = 0;
        mBufSize = 0;
    }

  private:
    template <class T>
    T at(uint idx, int offset, const T& absent) const {
        if (offset < 0) return absent;
        T v;
        std::memcpy(&v, buf_.get() + (size_t)idx * VertexType::Size + offset, sizeof(T));
        return v;
    }
    std::unique_ptr<_byte[]> buf_;
};

template <typename VertexType = Vertex_PositionNormalTex>
inline IVertexBuffer* CreateVertexBuffer(const void* pData, const size_t vertexCount) {
    IVertexBuffer* ret = new VertexBuffer<VertexType>;
    ret->NewBuffer((uint)vertexCount);
    std::memcpy(ret->GetBuffer(), pData, ret->GetBufferSize());
    return ret;
}

/* IndexBuffer (Utils/InputBuffer.h:148-181): 3 uint per triangle. */
class IndexBuffer {
  public:
    void ResizeBuffer(const uint triCount) { buf_.resize(3ull * triCount); }
    uint* GetBuffer() { return buf_.data(); }
    const uint* GetBuffer() const { return buf_.data(); }
    uint GetTriangleCount() const { return (uint)(buf_.size() / 3); }
    size_t GetBufferSize() const { return buf_.size(); }
    const uint* GetIndex(const uint idx) const {
        if (3ull * idx >= buf_.size()) throw std::out_of_range("IndexBuffer::GetIndex");
        return &buf_[3ull * idx];
    }
    void AppendTriangle(const int a, const int b, const int c) {
        buf_.push_back((uint)a);
        buf_.push_back((uint)b);
        buf_.push_back((uint)c);
    }
    void CopyFrom(const IndexBuffer& other) { buf_ = other.buf_; }
    void Release() {}

  private:
    std::vector<uint> buf_;
};

inline IndexBuffer* CreateIndexBuffer(const void* pData, const size_t triCount) {
    IndexBuffer* ret = new IndexBuffer;
    ret->ResizeBuffer((uint)triCount);
    std::memcpy(ret->GetBuffer(), pData, ret->GetBufferSize() * sizeof(uint));
    return ret;
}

/* Pixel shader objects (Core/Shader.h:173-282) as shading descriptors. */
class PixelShader {
  public:
    virtual ~PixelShader() {}
    int32_t Kind() const { return kind_; }
    edxr_material Material() const { return material_; }

  protected:
    explicit PixelShader(int32_t kind) : kind_(kind) { edxr_default_material(kind, &material_); }
    int32_t kind_;
    edxr_material material_;
};

/* LambertianAlbedoPixelShader with a constant albedo (Core/Shader.h:209-244). */
class LambertianAlbedoPixelShader : public PixelShader {
  public:
    explicit LambertianAlbedoPixelShader(float r = 0.9f, float g = 0.9f, float b = 0.9f)
        : PixelShader(EDXR_SHADER_LAMBERT_ALBEDO) {
        material_.albedo[0] = r;
        material_.albedo[1] = g;
        material_.albedo[2] = b;
    }
};

/* BlinnPhongPixelShader (Core/Shader.h:246-282). */
class BlinnPhongPixelShader : public PixelShader {
  public:
    BlinnPhongPixelShader() : PixelShader(EDXR_SHADER_BLINN_PHONG_REF) {}
};

/* LambertianAlbedoPixelShader sampling the mesh's texture slots
 * (Core/Shader.h:209-244): the reference Renderer's own shader. */
class TexturedLambertianPixelShader : public PixelShader {
  public:
    TexturedLambertianPixelShader() : PixelShader(EDXR_SHADER_TEXTURED_LAMBERT) {}
};

/* Vertex-colour Blinn-Phong: diffuse * interpolated RGB + specular. */
class VertexColorBlinnPhongPixelShader : public PixelShader {
  public:
    explicit VertexColorBlinnPhongPixelShader(float shininess = 32.0f) : PixelShader(EDXR_SHADER_VC_BLINN_PHONG) {
        material_.shininess = shininess;
    }
};

/* One texture slot of a mesh (Texture2D<Color>): an RGBA8 image or a
 * constant colour, in the C ABI's edxr_texture form. */
typedef edxr_texture Texture2D;

/* Mesh (Utils/Mesh.h:17-69): an AoS vertex buffer, an index buffer, texture
 * slots with a slot per triangle, and bounds.  The SoA planes the C ABI reads
 * are built from the vertex buffer on first render and kept. */
class Mesh {
  public:
    Mesh() = default;
    ~Mesh() { Release(); }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;

    /* Mesh::LoadMesh (Utils/Mesh.cpp:11-34): a Wavefront OBJ with its
     * texture slots: an image for a material whose diffuse map reads as BMP,
     * TGA or PPM (edxr_load_image), a constant colour (Kd) otherwise, and a
     * last 0.9-grey slot for faces without a material. */
    void LoadMesh(const Vector3& pos, const Vector3& scl, const Vector3& rot, const char* path) {
        Release();
        edxr_host_mesh hm;
        check(edxr_load_obj(path, pos.Data(), scl.Data(), rot.Data(), &hm));
        adopt(hm, true);
        const int64_t n = edxr_obj_materials(path, nullptr, 0);
        check(n < 0 ? (int)n : EDXR_OK);
        std::vector<edxr_obj_material> mats((size_t)n);
        if (n) edxr_obj_materials(path, mats.data(), (uint32_t)n);
        for (const edxr_obj_material& mt : mats) {
            Texture2D t = {};
            uint8_t* img = nullptr;
            if (mt.map_kd[0] && edxr_load_image(mt.map_kd, &t.width, &t.height, &img) == EDXR_OK) {
                images_.push_back(img);
                t.rgba = img;
            } else {
                t.color[0] = mt.kd[0];
                t.color[1] = mt.kd[1];
                t.color[2] = mt.kd[2];
            }
            textures_.push_back(t);
        }
        textures_.push_back(grey());
    }
    /* Mesh::LoadPlane (Utils/Mesh.cpp:36-51): one 0.9-white constant slot. */
    void LoadPlane(const Vector3& pos, const Vector3& scl, const Vector3& rot, const float length) {
        Release();
        edxr_host_mesh hm;
        check(edxr_load_plane(pos.Data(), scl.Data(), rot.Data(), length, &hm));
        adopt(hm, true);
        textures_.push_back(grey());
    }
    /* Mesh::LoadSphere (Utils/Mesh.cpp:53-70): one 0.9-white constant slot. */
    void LoadSphere(const Vector3& pos, const Vector3& scl, const Vector3& rot, const float radius,
                    const int slices = 64, const int stacks = 64) {
        Release();
        edxr_host_mesh hm;
        check(edxr_load_sphere(pos.Data(), scl.Data(), rot.Data(), radius, (uint32_t)slices, (uint32_t)stacks, &hm));
        adopt(hm, true);
        textures_.push_back(grey());
    }
    /* Extension: a UV sphere at `center` with seeded vertex colours and no
     * texture slots (the synthetic benchmark scenes). */
    void LoadColoredSphere(const float center[3], float radius, int slices = 64, int stacks = 64,
                           uint64_t colorSeed = 1) {
        Release();
        edxr_host_mesh hm;
        check(edxr_scene_sphere(&hm, (uint32_t)slices, (uint32_t)stacks, radius, center, colorSeed));
        adopt(hm, false);
    }
    /* Extension: a seeded noise heightfield on [-1,1]^2, seeded colours. */
    void LoadHeightfield(uint32_t n, float amplitude, uint64_t heightSeed, uint64_t colorSeed) {
        Release();
        edxr_host_mesh hm;
        check(edxr_scene_heightfield(&hm, n, amplitude, heightSeed, colorSeed));
        adopt(hm, false);
    }
    /* Extension: the caller's own buffers (ownership passes to the mesh),
     * texture slots and a slot per triangle (empty: none). */
    void SetBuffers(IVertexBuffer* vertices, IndexBuffer* indices, const std::vector<Texture2D>& textures = {},
                    const std::vector<uint>& textureIds = {}) {
        Release();
        vb_.reset(vertices);
        ib_.reset(indices);
        textures_ = textures;
        for (uint t : textureIds) texIdx_.Add(t);
        for (uint v = 0; v < vb_->GetVertexCount(); v++) bounds_.Expand(vb_->GetPosition(v));
        ++version_;
    }
    void Release() {
        vb_.reset();
        ib_.reset();
        for (uint8_t* p : images_) edxr_free(p);
        images_.clear();
        textures_.clear();
        texIdx_.Clear();
        bounds_ = BoundingBox();
        for (auto& p : soa_) p.clear();
        soaVersion_ = 0;
        ++version_;
    }

    const IVertexBuffer* GetVertexBuffer() const { return vb_.get(); }
    IndexBuffer* GetIndexBuffer() const { return ib_.get(); }
    const std::vector<Texture2D>& GetTextures() const { return textures_; }
    const Array<uint>& GetTextureIds() const { return texIdx_; }
    const BoundingBox GetBounds() const { return bounds_; }
    /* Changes whenever the mesh is loaded, set or released. */
    uint64_t Version() const { return version_; }

    /* The mesh as the C ABI reads it: SoA planes converted from the vertex
     * buffer (positions, normals, colours, texture coordinates) on first use
     * after a load, indices and slots in place (host memory). */
    edxr_mesh View() const {
        edxr_mesh v;
        std::memset(&v, 0, sizeof(v));
        if (!vb_ || !ib_) {
            v.memory = EDXR_MEM_HOST;
            return v;
        }
        if (soaVersion_ != version_) toSoA();
        float* f[11];
        for (int k = 0; k < 11; k++) f[k] = soa_[k].data();
        v.px = f[0]; v.py = f[1]; v.pz = f[2];
        v.nx = f[3]; v.ny = f[4]; v.nz = f[5];
        v.cr = f[6]; v.cg = f[7]; v.cb = f[8];
        v.tu = f[9]; v.tv = f[10];
        v.indices = ib_->GetBuffer();
        v.num_vertices = vb_->GetVertexCount();
        v.num_triangles = ib_->GetTriangleCount();
        v.memory = EDXR_MEM_HOST;
        v.tex_ids = texIdx_.Size() == v.num_triangles && v.num_triangles ? texIdx_.Data() : nullptr;
        return v;
    }

  private:
    static void check(int rc) {
        if (rc != EDXR_OK) throw std::runtime_error("edxraster: mesh loading failed (" + std::to_string(rc) + ")");
    }
    static Texture2D grey() {
        Texture2D t = {};
        t.color[0] = t.color[1] = t.color[2] = 0.9f;
        return t;
    }
    // Takes a generated or loaded host mesh: AoS vertex buffer (texture
    // coordinates when the mesh has them, else vertex colours), index
    // buffer, slots, bounds; the SoA planes are kept as they came.
    void adopt(edxr_host_mesh& hm, bool textured) {
        const uint V = hm.num_vertices, T = hm.num_triangles;
        if (textured && hm.tu) {
            std::vector<Vertex_PositionNormalTex> a(V);
            for (uint i = 0; i < V; i++)
                a[i] = {Vector3(hm.px[i], hm.py[i], hm.pz[i]), Vector3(hm.nx[i], hm.ny[i], hm.nz[i]),
                        Vector2(hm.tu[i], hm.tv[i])};
            vb_.reset(CreateVertexBuffer<Vertex_PositionNormalTex>(a.data(), V));
        } else {
            std::vector<Vertex_PositionNormalColor> a(V);
            for (uint i = 0; i < V; i++)
                a[i] = {Vector3(hm.px[i], hm.py[i], hm.pz[i]), Vector3(hm.nx[i], hm.ny[i], hm.nz[i]),
                        Color(hm.cr[i], hm.cg[i], hm.cb[i])};
            vb_.reset(CreateVertexBuffer<Vertex_PositionNormalColor>(a.data(), V));
        }
        ib_.reset(CreateIndexBuffer(hm.indices, T));
        if (hm.tex_ids)
            for (uint t = 0; t < T; t++) texIdx_.Add(hm.tex_ids[t]);
        for (uint i = 0; i < V; i++) bounds_.Expand(Vector3(hm.px[i], hm.py[i], hm.pz[i]));
        // the loaders' SoA planes are what toSoA would rebuild: keep them
        float* planes[11] = {hm.px, hm.py, hm.pz, hm.nx, hm.ny, hm.nz, hm.cr, hm.cg, hm.cb, hm.tu, hm.tv};
        for (int k = 0; k < 11; k++) {
            if (planes[k]) soa_[k].assign(planes[k], planes[k] + V);
            else soa_[k].assign(V, 0.0f);
        }
        edxr_scene_free(&hm);
        ++version_;
        soaVersion_ = version_;
    }
    // AoS -> SoA through the buffer's own fetches (IVertexBuffer::Get*).
    void toSoA() const {
        const uint V = vb_->GetVertexCount();
        for (auto& p : soa_) p.resize(V);
        for (uint i = 0; i < V; i++) {
            const Vector3 p = vb_->GetPosition(i), n = vb_->GetNormal(i);
            const Vector2 uv = vb_->GetTexCoord(i);
            const Color c = vb_->GetColor(i);
            const float vals[11] = {p.x, p.y, p.z, n.x, n.y, n.z, c.r, c.g, c.b, uv.u, uv.v};
            for (int k = 0; k < 11; k++) soa_[k][i] = vals[k];
        }
        soaVersion_ = version_;
    }

    std::unique_ptr<IVertexBuffer> vb_;
    std::unique_ptr<IndexBuffer> ib_;
    std::vector<Texture2D> textures_;
    std::vector<uint8_t*> images_;
    Array<uint> texIdx_;
    BoundingBox bounds_;
    mutable std::vector<float> soa_[11];
    mutable uint64_t soaVersion_ = 0;
    uint64_t version_ = 0;
};

/* Renderer (Core/Renderer.h:15-59) on one MI355X.  As in the reference, the
 * pixel shader is LambertianAlbedoPixelShader over the rendered mesh's
 * texture slots (Core/Renderer.cpp:41,105-106) unless SetPixelShader chose
 * another (an extension). */
class Renderer {
  public:
    Renderer() = default;
    ~Renderer() {
        if (ctx_) edxr_destroy(ctx_);
    }
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    void Initialize(uint32_t width, uint32_t height) {
        if (ctx_) edxr_destroy(ctx_);
        ctx_ = nullptr;
        const int rc = edxr_create(&ctx_, width, height);
        if (rc != EDXR_OK) throw std::runtime_error("edxraster: edxr_create failed (" + std::to_string(rc) + ")");
        width_ = width;
        height_ = height;
        shaderKind_ = -1;
        boundMesh_ = nullptr;
        boundVersion_ = 0;
    }
    void Resize(uint32_t width, uint32_t height) {
        check(edxr_resize(ctx_, width, height));
        width_ = width;
        height_ = height;
        backBufferValid_ = false;
    }
    void SetTransform(const Matrix& modelView, const Matrix& proj, const Matrix& toRaster) {
        check(edxr_set_transform(ctx_, &modelView.m[0][0], &proj.m[0][0], &toRaster.m[0][0]));
    }
    void SetLight(float x, float y, float z) {
        const float d[3] = {x, y, z};
        check(edxr_set_light(ctx_, d));
    }
    void SetPixelShader(const PixelShader& ps) {
        const edxr_material m = ps.Material();
        check(edxr_set_pixel_shader(ctx_, ps.Kind(), &m));
        shaderKind_ = ps.Kind();
    }
    void SetHierarchicalRasterize(const bool on) { check(edxr_set_state(ctx_, EDXR_STATE_HIERARCHICAL, on ? 1 : 0)); }
    /* Renderer::SetWriteFrames (Core/Renderer.h:50): every RenderMesh writes
     * Frames/Frame%05i.bmp under the frame directory ("." by default). */
    void SetWriteFrames(const bool on) { check(edxr_set_state(ctx_, EDXR_STATE_WRITE_FRAMES, on ? 1 : 0)); }
    void SetFrameDirectory(const char* dir) { check(edxr_set_frame_directory(ctx_, dir)); }
    /* Renderer::WriteFrameToFile (Core/Renderer.cpp:352-358). */
    void WriteFrameToFile() { check(edxr_write_frame_to_file(ctx_)); }
    /* Renderer::SetMSAAMode (Core/Renderer.cpp:94-98): 0 (1 sample) .. 5 (32).
     * Setting the current level again keeps the frame buffers. */
    void SetMSAAMode(const int sampleCountLog2) {
        if (sampleCountLog2 == msaa_) return;
        check(edxr_set_state(ctx_, EDXR_STATE_MSAA, sampleCountLog2));
        msaa_ = sampleCountLog2;
        backBufferValid_ = false;
    }
    /* Renderer::SetTextureFilter (Core/Renderer.h:48). */
    void SetTextureFilter(const TextureFilter filter) {
        check(edxr_set_state(ctx_, EDXR_STATE_TEXTURE_FILTER, (int)filter));
    }
    /* Renderer::RenderMesh (Core/Renderer.cpp:100-118): binds the mesh's
     * texture slots and renders. */
    void RenderMesh(const Mesh& mesh) {
        const std::vector<Texture2D>& t = mesh.GetTextures();
        if (&mesh != boundMesh_ || mesh.Version() != boundVersion_) {
            check(edxr_set_textures(ctx_, t.data(), (uint32_t)t.size()));
            boundMesh_ = &mesh;
            boundVersion_ = mesh.Version();
        }
        const edxr_mesh v = mesh.View();
        if (shaderKind_ < 0) {
            // the reference's shader: Lambertian over the mesh's slots
            const bool slots = !t.empty() && v.tex_ids && v.tu;
            const int32_t kind = slots ? EDXR_SHADER_TEXTURED_LAMBERT : EDXR_SHADER_LAMBERT_ALBEDO;
            check(edxr_set_pixel_shader(ctx_, kind, nullptr));
        }
        check(edxr_render_mesh(ctx_, &v));
        backBufferValid_ = false;
    }
    void RenderMesh(const edxr_mesh& mesh) {
        check(edxr_render_mesh(ctx_, &mesh));
        backBufferValid_ = false;
    }
    /* Renderer::GetBackBuffer: RGBA8, W x H, bottom-up rows; valid until the
     * next RenderMesh, Resize or SetMSAAMode. */
    const _byte* GetBackBuffer() {
        if (!backBufferValid_) {
            backBuffer_.resize((size_t)width_ * height_ * 4);
            check(edxr_read_color(ctx_, backBuffer_.data(), backBuffer_.size()));
            backBufferValid_ = true;
        }
        return backBuffer_.data();
    }
    void ReadDepth(std::vector<float>& out) {
        out.resize((size_t)width_ * height_);
        check(edxr_read_depth(ctx_, out.data(), out.size()));
    }
    void ReadTriangleIds(std::vector<uint32_t>& out) {
        out.resize((size_t)width_ * height_);
        check(edxr_read_triangle_ids(ctx_, out.data(), out.size()));
    }
    edxr_stats GetStats() {
        edxr_stats s;
        check(edxr_get_stats(ctx_, &s));
        return s;
    }
    uint32_t GetWidth() const { return width_; }
    uint32_t GetHeight() const { return height_; }

  private:
    void check(int rc) {
        if (rc != EDXR_OK) throw std::runtime_error(std::string("edxraster: ") + edxr_last_error(ctx_));
    }
    edxr_ctx* ctx_ = nullptr;
    uint32_t width_ = 0, height_ = 0;
    int msaa_ = 0;
    int32_t shaderKind_ = -1;  // -1: the reference's Lambertian over the mesh's slots
    std::vector<uint8_t> backBuffer_;
    bool backBufferValid_ = false;
    const Mesh* boundMesh_ = nullptr;
    uint64_t boundVersion_ = 0;
};

}  // namespace RasterRenderer
}  // namespace EDX

#endif
