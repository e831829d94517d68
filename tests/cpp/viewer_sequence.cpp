// The reference viewer's host code (RealtimeViewer/Main.cpp:18-131) over the
// C++ facade, with its window, OpenGL and GUI calls left out and nothing else
// changed: the same globals, OnInit (Initialize, Camera::Init, LoadSphere,
// GetBounds().BoundingSphere), OnRender (Camera::Transform, SetTransform,
// RenderMesh, GetBackBuffer, GetIndexBuffer()->GetTriangleCount(),
// SetMSAAMode, SetTextureFilter(TextureFilter(id))) and the "Load Scene"
// branch (Release, LoadMesh, a camera fitted to the bounds).
//
// usage: viewer_sequence out_sphere.rgba [scene.obj out_scene.rgba [filter]]
// Writes each frame's back buffer (RGBA8, bottom-up rows, 1280 x 720) and
// prints the triangle count the viewer's GUI would show.
#include <cstdio>
#include <cstdlib>

#include "edxraster.hpp"

using namespace EDX;
using namespace EDX::RasterRenderer;

int giWindowWidth = 1280;
int giWindowHeight = 720;

int gTexFilterId = 2;
int gMSAAId = 0;
bool gHRas = true;
bool gRecord = false;

// Global variables
Renderer*		gpRenderer;
Camera			gCamera;
Mesh			gMesh;

void OnInit()
{
	gpRenderer = new Renderer;
	gpRenderer->Initialize(giWindowWidth, giWindowHeight);
	gCamera.Init(-5.0f * Vector3::UNIT_Z, Vector3::ZERO, Vector3::UNIT_Y, giWindowWidth, giWindowHeight, 65, 0.01f);

	gMesh.LoadSphere(Vector3::ZERO, Vector3::UNIT_SCALE, Vector3::ZERO, 1.2f);

	Vector3 center;
	float radius;
	gMesh.GetBounds().BoundingSphere(&center, &radius);
	gCamera.mMoveScaler = radius / 50.0f;
}

void OnRender()
{
	gCamera.Transform();

	gpRenderer->SetTransform(gCamera.GetViewMatrix(), gCamera.GetProjMatrix(), gCamera.GetRasterMatrix());
	gpRenderer->RenderMesh(gMesh);

	const _byte* pixels = gpRenderer->GetBackBuffer();
	(void)pixels;

	std::printf("Triangle Count: %i\n", gMesh.GetIndexBuffer()->GetTriangleCount());
	gpRenderer->SetHierarchicalRasterize(gHRas);
	gpRenderer->SetMSAAMode(gMSAAId);
	gpRenderer->SetTextureFilter(TextureFilter(gTexFilterId));
}

void LoadScene(const char* filePath)
{
	gMesh.Release();
	gMesh.LoadMesh(Vector3(0, 0, 0), 0.01f * Vector3::UNIT_SCALE, Vector3(0, 0, 0), filePath);

	Vector3 center;
	float radius;
	gMesh.GetBounds().BoundingSphere(&center, &radius);

	gCamera.Init(Vector3::ZERO, Vector3::UNIT_Z, Vector3::UNIT_Y, giWindowWidth, giWindowHeight, 65, radius * 0.02f, radius * 5.0f);
	gCamera.mMoveScaler = radius / 50.0f;
}

static bool Save(const char* path)
{
	FILE* f = std::fopen(path, "wb");
	if (!f) return false;
	const size_t n = (size_t)giWindowWidth * giWindowHeight * 4;
	const bool ok = std::fwrite(gpRenderer->GetBackBuffer(), 1, n, f) == n;
	std::fclose(f);
	return ok;
}

int main(int argc, char** argv)
{
	if (argc < 2) {
		std::fprintf(stderr, "usage: viewer_sequence out_sphere.rgba [scene.obj out_scene.rgba [filter]]\n");
		return 2;
	}
	try {
		OnInit();
		OnRender();
		if (!Save(argv[1])) return 3;
		if (argc >= 4) {
			if (argc >= 5) gTexFilterId = std::atoi(argv[4]);
			LoadScene(argv[2]);
			OnRender();  // applies the filter for the next frame, as the viewer's combo box does
			OnRender();
			if (!Save(argv[3])) return 3;
			Vector3 center;
			float radius;
			gMesh.GetBounds().BoundingSphere(&center, &radius);
			std::printf("scene radius %a\n", radius);
		}
		delete gpRenderer;
	} catch (const std::exception& e) {
		std::fprintf(stderr, "%s\n", e.what());
		return 1;
	}
	return 0;
}
