// Reference-style host code over the C++ facade: the flow of the reference
// viewer's OnInit / OnRender (RealtimeViewer/Main.cpp:32-75) -- Initialize,
// LoadSphere, SetTransform, RenderMesh, GetBackBuffer -- without the window.
// Writes the back buffer (RGBA8, bottom-up rows) to the file given as argv[1].
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
    const uint32_t W = 800, H = 600;
    try {
        Renderer renderer;
        renderer.Initialize(W, H);
        Mesh mesh;
        const float center[3] = {0.0f, 0.0f, 0.0f};
        mesh.LoadSphere(center, 1.0f, 64, 32, 1);
        const float eye[3] = {0.0f, 0.0f, -3.0f}, at[3] = {0.0f, 0.0f, 0.0f}, up[3] = {0.0f, 1.0f, 0.0f};
        renderer.SetTransform(Matrix::LookAt(eye, at, up), Matrix::Perspective(60.0f, (float)W / H, 0.1f, 100.0f),
                              Matrix::ToRaster(W, H));
        renderer.SetPixelShader(VertexColorBlinnPhongPixelShader(32.0f));
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
