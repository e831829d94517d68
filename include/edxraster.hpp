/*
 * edxraster.hpp -- header-only C++ facade over the C ABI (edxraster.h) that
 * re-creates the reference's host API so EDXRaster-style code keeps its shape:
 *
 *   EDX::RasterRenderer::Renderer     Core/Renderer.h:15-59
 *   EDX::RasterRenderer::PixelShader  Core/Shader.h:173-282 (shader objects)
 *   EDX::RasterRenderer::Mesh         Utils/Mesh.h:17-69 (SoA instead of AoS)
 *   EDX::Matrix                       EDXUtil's Matrix (row-major, column vectors)
 *
 * The reference reports no errors; this facade throws std::runtime_error with
 * edxr_last_error() text when the C ABI returns a negative status.
 */
#ifndef EDXRASTER_HPP
#define EDXRASTER_HPP

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "edxraster.h"

namespace EDX {

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

namespace RasterRenderer {

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

/* LambertianAlbedoPixelShader with a constant texture (Core/Shader.h:209-244). */
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
 * (Core/Shader.h:209-244), as the reference binds it to loaded meshes. */
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

/* Mesh (Utils/Mesh.h:17-69): SoA fp32 positions, normals, colours, uint32 indices. */
class Mesh {
  public:
    Mesh() { std::memset(&m_, 0, sizeof(m_)); }
    ~Mesh() { Release(); }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;

    /* Mesh::LoadSphere (Utils/Mesh.cpp:53-70); colours seeded. */
    void LoadSphere(const float center[3], float radius, int slices = 64, int stacks = 64, uint64_t colorSeed = 1) {
        Release();
        check(edxr_scene_sphere(&m_, (uint32_t)slices, (uint32_t)stacks, radius, center, colorSeed));
    }
    /* A seeded noise heightfield on [-1,1]^2 (stands in for Mesh::LoadPlane). */
    void LoadHeightfield(uint32_t n, float amplitude, uint64_t heightSeed, uint64_t colorSeed) {
        Release();
        check(edxr_scene_heightfield(&m_, n, amplitude, heightSeed, colorSeed));
    }
    /* Mesh::LoadMesh (Utils/Mesh.cpp:11-34): a Wavefront OBJ placed by
     * translation pos, rotation rot (degrees) and scale scl, with its texture
     * slots: an image texture for a material with a .bmp diffuse map, a
     * constant colour (Kd) otherwise, and a last 0.9-grey slot for faces
     * without a material. */
    void LoadMesh(const float pos[3], const float scl[3], const float rot[3], const char* path) {
        Release();
        check(edxr_load_obj(path, pos, scl, rot, &m_));
        const int64_t n = edxr_obj_materials(path, nullptr, 0);
        check(n < 0 ? (int)n : EDXR_OK);
        std::vector<edxr_obj_material> mats((size_t)n);
        if (n) edxr_obj_materials(path, mats.data(), (uint32_t)n);
        for (const edxr_obj_material& mt : mats) {
            edxr_texture t = {};
            const std::string map = mt.map_kd;
            const bool bmp = map.size() > 4 && (map.substr(map.size() - 4) == ".bmp" || map.substr(map.size() - 4) == ".BMP");
            uint8_t* img = nullptr;
            if (bmp && edxr_load_bmp(map.c_str(), &t.width, &t.height, &img) == EDXR_OK) {
                images_.push_back(img);
                t.rgba = img;
            } else {
                t.color[0] = mt.kd[0];
                t.color[1] = mt.kd[1];
                t.color[2] = mt.kd[2];
            }
            textures_.push_back(t);
        }
        edxr_texture grey = {};
        grey.color[0] = grey.color[1] = grey.color[2] = 0.9f;
        textures_.push_back(grey);
        static uint64_t loads = 0;
        textureSetId_ = ++loads;
    }
    void Release() {
        textureSetId_ = 0;
        edxr_scene_free(&m_);
        for (uint8_t* p : images_) edxr_free(p);
        images_.clear();
        textures_.clear();
    }
    /* Mesh::GetTextures: the texture slots tex_ids refer to (empty for generated meshes). */
    const std::vector<edxr_texture>& GetTextures() const { return textures_; }
    /* Identifies the texture set of the last LoadMesh (0: none). */
    uint64_t TextureSetId() const { return textureSetId_; }

    uint32_t GetVertexCount() const { return m_.num_vertices; }
    uint32_t GetTriangleCount() const { return m_.num_triangles; }
    edxr_mesh View() const {
        edxr_mesh v = {};  // texture coordinates and slots stay NULL
        v.px = m_.px; v.py = m_.py; v.pz = m_.pz;
        v.nx = m_.nx; v.ny = m_.ny; v.nz = m_.nz;
        v.cr = m_.cr; v.cg = m_.cg; v.cb = m_.cb;
        v.indices = m_.indices;
        v.num_vertices = m_.num_vertices;
        v.num_triangles = m_.num_triangles;
        v.memory = EDXR_MEM_HOST;
        v.tu = m_.tu;
        v.tv = m_.tv;
        v.tex_ids = m_.tex_ids;
        return v;
    }

  private:
    static void check(int rc) {
        if (rc != EDXR_OK) throw std::runtime_error("edxraster: mesh generation failed (" + std::to_string(rc) + ")");
    }
    edxr_host_mesh m_;
    std::vector<edxr_texture> textures_;
    std::vector<uint8_t*> images_;
    uint64_t textureSetId_ = 0;
};

/* Renderer (Core/Renderer.h:15-59) on one MI355X. */
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
    }
    void Resize(uint32_t width, uint32_t height) {
        check(edxr_resize(ctx_, width, height));
        width_ = width;
        height_ = height;
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
    }
    void SetHierarchicalRasterize(bool on) { check(edxr_set_state(ctx_, EDXR_STATE_HIERARCHICAL, on ? 1 : 0)); }
    /* Renderer::SetWriteFrames (Core/Renderer.h:50): every RenderMesh writes
     * Frames/Frame%05i.bmp under the frame directory ("." by default). */
    void SetWriteFrames(bool on) { check(edxr_set_state(ctx_, EDXR_STATE_WRITE_FRAMES, on ? 1 : 0)); }
    void SetFrameDirectory(const char* dir) { check(edxr_set_frame_directory(ctx_, dir)); }
    /* Renderer::WriteFrameToFile (Core/Renderer.cpp:352-358). */
    void WriteFrameToFile() { check(edxr_write_frame_to_file(ctx_)); }
    /* Renderer::SetMSAAMode (Core/Renderer.cpp:94-98): 0 (1 sample) .. 5 (32). */
    void SetMSAAMode(int sampleCountLog2) { check(edxr_set_state(ctx_, EDXR_STATE_MSAA, sampleCountLog2)); }
    /* Renderer::RenderMesh binds the mesh's texture slots
     * (RenderStates::TextureSlots, Core/Renderer.cpp:105-106) and renders. */
    void RenderMesh(const Mesh& mesh) {
        const std::vector<edxr_texture>& t = mesh.GetTextures();
        if (!t.empty() && mesh.TextureSetId() != boundTextureSet_) {
            check(edxr_set_textures(ctx_, t.data(), (uint32_t)t.size()));
            boundTextureSet_ = mesh.TextureSetId();
        }
        const edxr_mesh v = mesh.View();
        check(edxr_render_mesh(ctx_, &v));
        backBufferValid_ = false;
    }
    /* Renderer::SetTextureFilter (RenderStates::TexFilter): EDXR_FILTER_*. */
    void SetTextureFilter(int filter) { check(edxr_set_state(ctx_, EDXR_STATE_TEXTURE_FILTER, filter)); }
    void RenderMesh(const edxr_mesh& mesh) {
        check(edxr_render_mesh(ctx_, &mesh));
        backBufferValid_ = false;
    }
    /* Renderer::GetBackBuffer: RGBA8, W x H, bottom-up rows; valid until the
     * next RenderMesh or Resize. */
    const uint8_t* GetBackBuffer() {
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
    std::vector<uint8_t> backBuffer_;
    bool backBufferValid_ = false;
    uint64_t boundTextureSet_ = 0;
};

}  // namespace RasterRenderer
}  // namespace EDX

#endif
