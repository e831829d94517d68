// Reference-style host code over the C++ facade: the flow of the reference
// viewer's OnInit / OnRender (RealtimeViewer/Main.cpp:32-75) -- Initialize,
// LoadSphere or LoadMesh, SetTransform, RenderMesh, GetBackBuffer -- without
// the window.  Writes the back buffer (RGBA8, bottom-up rows) to argv[1].
// Without an OBJ path it renders the seeded 64x32 colour sphere with the
// vertex-colour Blinn-Phong shader (an extension); with an OBJ path as
// argv[2] it renders that mesh (320 x 240) with its texture slots and the
// reference's textured Lambertian shader.
#include <cstdio>
#include <vector>

#include "edxraster.hpp"

using namespace EDX;
using namespace EDX::RasterRenderer;

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: facade_demo out.rgba\n");
        return 2;
    }
    const bool obj = argc >= 3;
    const uint32_t W = obj ? 320 : 800, H = obj ? 240 : 600;
    try {
        Renderer renderer;
        renderer.Initialize(W, H);
        Mesh mesh;
        const float center[3] = {0.0f, 0.0f, 0.0f};
        const float up[3] = {0.0f, 1.0f, 0.0f};
        if (obj) {
            mesh.LoadMesh(Vector3::ZERO, Vector3::UNIT_SCALE, Vector3::ZERO, argv[2]);
            const float eye[3] = {0.0f, 1.2f, -2.2f};
            renderer.SetTransform(Matrix::LookAt(eye, center, up),
                                  Matrix::Perspective(60.0f, (float)W / H, 0.1f, 100.0f), Matrix::ToRaster(W, H));
            renderer.SetPixelShader(TexturedLambertianPixelShader());
        } else {
            mesh.LoadColoredSphere(center, 1.0f, 64, 32, 1);
            const float eye[3] = {0.0f, 0.0f, -3.0f};
            renderer.SetTransform(Matrix::LookAt(eye, center, up),
                                  Matrix::Perspective(60.0f, (float)W / H, 0.1f, 100.0f), Matrix::ToRaster(W, H));
            renderer.SetPixelShader(VertexColorBlinnPhongPixelShader(32.0f));
        }
        renderer.RenderMesh(mesh);
        const uint8_t* bb = renderer.GetBackBuffer();
        FILE* f = std::fopen(argv[1], "wb");
        if (!f) return 3;
        std::fwrite(bb, 1, (size_t)W * H * 4, f);
        std::fclose(f);
        const edxr_stats st = renderer.GetStats();
        std::printf("triangles %llu setup %llu refs %llu\n", (unsigned long long)st.triangles_in,
                    (unsigned long long)st.setup_records, (unsigned long long)st.refs);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
