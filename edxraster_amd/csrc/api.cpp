// api.cpp -- C ABI of the MI355X EDXRaster pipeline (host orchestration).
//
// Replaces EDXRaster's Renderer (Core/Renderer.h:37-50, Core/Renderer.cpp):
// the per-context state replaces the RenderStates singleton
// (Core/RenderStates.h:38-52), device buffers replace the per-core arrays, and
// RenderMesh becomes one ordered sequence of kernel launches on the context's
// stream (pipeline.hip).  Fixed-capacity device buffers report overflow through
// counters; the frame is then re-run with larger buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/edxraster.h"
#include "pipeline.h"

using namespace edxr;

namespace {

// Matrix helpers with the reference's declared semantics (see FrameParams).
void mat_mul(const float* a, const float* b, float* out) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            out[4 * i + j] = ((a[4 * i + 0] * b[0 + j] + a[4 * i + 1] * b[4 + j]) + a[4 * i + 2] * b[8 + j]) +
                             a[4 * i + 3] * b[12 + j];
}

// Matrix::Inverse: cofactors in double, rounded once.
void mat_inverse(const float* a, float* out) {
    double m[16], inv[16];
    for (int i = 0; i < 16; i++) m[i] = a[i];
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
    const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    for (int i = 0; i < 16; i++) out[i] = (float)(det != 0.0 ? inv[i] / det : 0.0);
}

// TransformPoint of a Vector3 (w = 1, then divided by w).
void transform_point3(const float* m, const float* v, float* out) {
    float r[4];
    for (int i = 0; i < 4; i++)
        r[i] = ((m[4 * i] * v[0] + m[4 * i + 1] * v[1]) + m[4 * i + 2] * v[2]) + m[4 * i + 3] * 1.0f;
    out[0] = r[0] / r[3];
    out[1] = r[1] / r[3];
    out[2] = r[2] / r[3];
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n) {
        if (n <= cap && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t bytes = sizeof(T) * std::max<size_t>(n, 1);
        hipError_t e = hipMalloc((void**)&p, bytes);
        if (e == hipSuccess) cap = std::max<size_t>(n, 1);
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

bool is_integer_exponent(float s) { return s >= 0.0f && s <= 4096.0f && s == std::floor(s); }

}  // namespace

struct edxr_ctx {
    uint32_t W = 0, H = 0, tilesX = 0, tilesY = 0, nTiles = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    float mv[16], proj[16], raster[16], mvp[16], mvInv[16], eye[3];
    bool haveTransform = false;
    float lightDir[3] = {1.0f, 1.0f, -1.0f};
    int32_t shader = EDXR_SHADER_LAMBERT_ALBEDO;
    edxr_material mat;
    int32_t hierarchical = 1;
    uint32_t binChunksSetting = 0;
    int32_t profile = 0;
    hipEvent_t events[kNumStages + 1] = {};
    double stageMs[kNumStages] = {};
    uint64_t profiledFrames = 0;

    // owned copies of host meshes
    DevBuf<float> mesh[9];
    DevBuf<uint32_t> meshIdx;
    // pipeline buffers
    DevBuf<RasterRec> rrec;
    DevBuf<ShadeRec> srec;
    DevBuf<uint32_t> desc, clipQueue, binCounts, tileTotals, tileStart, refs, counters, color, ids, work, tileItem;
    DevBuf<float> depth;
    DevBuf<uint64_t> partials;
    DevBuf<uint16_t> slotRefs;
    DevBuf<uint32_t> groupSums;
    uint32_t numCUs = 256;
    DevBuf<ShadeVert> verts;  // V input + extraVertCap clip-created shading vertices
    uint32_t clipQueueCap = 0, extraRecCap = 0, extraVertCap = 0, refCap = 0;
    uint32_t* hostCounters = nullptr;

    // last frame
    bool rendered = false;
    uint32_t lastT = 0, lastV = 0, lastChunks = 0;
    MeshPtrs lastMesh{};
    edxr_stats stats{};
    std::string err;

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return EDXR_E_HIP;
    }
};

namespace {

int set_size(edxr_ctx* c, uint32_t w, uint32_t h) {
    if (w == 0 || h == 0 || w > kMaxDim || h > kMaxDim) return c->fail(EDXR_E_LIMIT, "resolution out of range");
    const uint32_t tx = (w + kTileSize - 1) >> kTileLog2, ty = (h + kTileSize - 1) >> kTileLog2;
    if ((uint64_t)tx * ty > kMaxTiles) return c->fail(EDXR_E_LIMIT, "more than 16384 tiles");
    c->W = w;
    c->H = h;
    c->tilesX = tx;
    c->tilesY = ty;
    c->nTiles = tx * ty;
    const size_t P = (size_t)w * h;
    hipError_t e;
    if ((e = c->color.ensure(P)) != hipSuccess || (e = c->depth.ensure(P)) != hipSuccess ||
        (e = c->ids.ensure(P)) != hipSuccess || (e = c->tileTotals.ensure(c->nTiles)) != hipSuccess ||
        (e = c->tileStart.ensure(c->nTiles + 1)) != hipSuccess ||
        (e = c->tileItem.ensure(c->nTiles)) != hipSuccess)
        return c->hip_fail(e, "hipMalloc(frame buffers)");
    c->rendered = false;
    return EDXR_OK;
}

// Binning chunks: about 1024 triangles each so the count/scatter grids fill
// the GPU, capped so the chunk x tile count matrix stays within 64 MB.
uint32_t choose_chunks(const edxr_ctx* c, uint32_t T) {
    if (c->binChunksSetting) return std::max(1u, std::min(std::min(c->binChunksSetting, kMaxChunks), std::max(1u, T)));
    const uint32_t byMatrix = (uint32_t)std::max<uint64_t>(1, (64ull << 20) / (4ull * c->nTiles));
    const uint32_t C = (T + 1023) / 1024;
    return std::max(1u, std::min(std::min(C, byMatrix), kMaxChunks));
}

}  // namespace

extern "C" {

void edxr_default_material(int32_t kind, edxr_material* m) {
    if (!m) return;
    m->ambient = 0.2f;
    m->kd = 3.0f;
    m->ks = 3.0f;
    m->clamp_specular = 0;
    if (kind == EDXR_SHADER_LAMBERT_ALBEDO) {
        m->shininess = 0.0f;
        m->albedo[0] = m->albedo[1] = m->albedo[2] = 0.9f;
    } else {
        m->shininess = kind == EDXR_SHADER_BLINN_PHONG_REF ? 200.0f : 32.0f;
        m->albedo[0] = m->albedo[1] = m->albedo[2] = 1.0f;
    }
}

int edxr_create(edxr_ctx** out, uint32_t width, uint32_t height) {
    if (!out) return EDXR_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return EDXR_E_HIP;
    edxr_ctx* c = new edxr_ctx();
    edxr_default_material(c->shader, &c->mat);
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->hostCounters, sizeof(uint32_t) * kCntSlots, hipHostMallocDefault);
    if (e == hipSuccess) e = c->counters.ensure(kCntSlots);
    if (e == hipSuccess) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && cus > 0)
            c->numCUs = (uint32_t)cus;
    }
    if (e != hipSuccess) {
        edxr_destroy(c);
        return EDXR_E_HIP;
    }
    const int r = set_size(c, width, height);
    if (r != EDXR_OK) {
        edxr_destroy(c);
        return r;
    }
    *out = c;
    return EDXR_OK;
}

void edxr_destroy(edxr_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& b : c->mesh) b.release();
    c->meshIdx.release();
    c->rrec.release();
    c->srec.release();
    for (DevBuf<uint32_t>* b : {&c->desc, &c->clipQueue, &c->binCounts, &c->tileTotals, &c->tileStart, &c->refs,
                                &c->counters, &c->color, &c->ids, &c->work, &c->tileItem})
        b->release();
    c->depth.release();
    c->partials.release();
    c->slotRefs.release();
    c->groupSums.release();
    c->verts.release();
    if (c->hostCounters) (void)hipHostFree(c->hostCounters);
    for (auto& ev : c->events)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int edxr_resize(edxr_ctx* c, uint32_t w, uint32_t h) {
    if (!c) return EDXR_E_INVALID;
    return set_size(c, w, h);
}

int edxr_set_transform(edxr_ctx* c, const float mv[16], const float proj[16], const float raster[16]) {
    if (!c || !mv || !proj || !raster) return EDXR_E_INVALID;
    memcpy(c->mv, mv, 64);
    memcpy(c->proj, proj, 64);
    memcpy(c->raster, raster, 64);
    mat_mul(proj, mv, c->mvp);
    mat_inverse(mv, c->mvInv);
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    transform_point3(c->mvInv, zero, c->eye);
    c->haveTransform = true;
    return EDXR_OK;
}

int edxr_set_light(edxr_ctx* c, const float dir[3]) {
    if (!c || !dir) return EDXR_E_INVALID;
    memcpy(c->lightDir, dir, 12);
    return EDXR_OK;
}

int edxr_set_pixel_shader(edxr_ctx* c, int32_t kind, const edxr_material* m) {
    if (!c || kind < EDXR_SHADER_LAMBERT_ALBEDO || kind > EDXR_SHADER_VC_BLINN_PHONG) return EDXR_E_INVALID;
    c->shader = kind;
    if (m) c->mat = *m;
    else edxr_default_material(kind, &c->mat);
    return EDXR_OK;
}

int edxr_set_state(edxr_ctx* c, int32_t key, int32_t value) {
    if (!c) return EDXR_E_INVALID;
    switch (key) {
        case EDXR_STATE_HIERARCHICAL: c->hierarchical = value ? 1 : 0; return EDXR_OK;
        case EDXR_STATE_PROFILE:
            if (value && !c->events[0])
                for (auto& ev : c->events)
                    if (hipEventCreate(&ev) != hipSuccess) return c->fail(EDXR_E_HIP, "hipEventCreate");
            c->profile = value ? 1 : 0;
            return EDXR_OK;
        case EDXR_STATE_BIN_CHUNKS:
            if (value < 0 || value > 65536) return c->fail(EDXR_E_INVALID, "bin chunks out of range");
            c->binChunksSetting = (uint32_t)value;
            return EDXR_OK;
        default: return c->fail(EDXR_E_INVALID, "unknown state key");
    }
}

int edxr_render_mesh(edxr_ctx* c, const edxr_mesh* m) {
    if (!c || !m) return EDXR_E_INVALID;
    if (!c->haveTransform) return c->fail(EDXR_E_STATE, "edxr_set_transform was not called");
    if (!m->px || !m->py || !m->pz || (m->num_triangles && !m->indices))
        return c->fail(EDXR_E_INVALID, "mesh positions/indices missing");
    if (!m->nx || !m->ny || !m->nz) return c->fail(EDXR_E_INVALID, "mesh normals missing");
    if (c->shader == EDXR_SHADER_VC_BLINN_PHONG && (!m->cr || !m->cg || !m->cb))
        return c->fail(EDXR_E_INVALID, "vertex colours required by EDXR_SHADER_VC_BLINN_PHONG");
    const uint32_t V = m->num_vertices, T = m->num_triangles;
    // Record indices (T + clip records) must fit the 28-bit reference payload.
    if ((uint64_t)T + 65536 > kRecordMask) return c->fail(EDXR_E_LIMIT, "too many triangles");
    hipError_t e;

    MeshPtrs M;
    if (m->memory == EDXR_MEM_DEVICE) {
        M = MeshPtrs{m->px, m->py, m->pz, m->nx, m->ny, m->nz, m->cr, m->cg, m->cb, m->indices};
    } else {
        const float* src[9] = {m->px, m->py, m->pz, m->nx, m->ny, m->nz, m->cr, m->cg, m->cb};
        const float* dst[9];
        for (int k = 0; k < 9; k++) {
            dst[k] = nullptr;
            if (!src[k]) continue;
            if ((e = c->mesh[k].ensure(V)) != hipSuccess) return c->hip_fail(e, "hipMalloc(mesh)");
            if (V && (e = hipMemcpyAsync(c->mesh[k].p, src[k], sizeof(float) * V, hipMemcpyHostToDevice, c->stream)) !=
                         hipSuccess)
                return c->hip_fail(e, "hipMemcpyAsync(mesh)");
            dst[k] = c->mesh[k].p;
        }
        if ((e = c->meshIdx.ensure(3ull * T)) != hipSuccess) return c->hip_fail(e, "hipMalloc(indices)");
        if (T && (e = hipMemcpyAsync(c->meshIdx.p, m->indices, 12ull * T, hipMemcpyHostToDevice, c->stream)) !=
                     hipSuccess)
            return c->hip_fail(e, "hipMemcpyAsync(indices)");
        M = MeshPtrs{dst[0], dst[1], dst[2], dst[3], dst[4], dst[5], dst[6], dst[7], dst[8], c->meshIdx.p};
    }

    FrameParams P;
    memset(&P, 0, sizeof(P));
    memcpy(P.mvp, c->mvp, 64);
    memcpy(P.raster, c->raster, 64);
    memcpy(P.eye, c->eye, 12);
    {
        // Math::Normalize(lightDir) as the pixel shader does (Core/Shader.h:221).
        const float* L = c->lightDir;
        const float s = 1.0f / sqrtf((L[0] * L[0] + L[1] * L[1]) + L[2] * L[2]);
        P.light[0] = L[0] * s;
        P.light[1] = L[1] * s;
        P.light[2] = L[2] * s;
    }
    P.ambient = c->mat.ambient;
    P.kd = c->mat.kd;
    P.ks = c->mat.ks;
    P.shininess = c->mat.shininess;
    memcpy(P.albedo, c->mat.albedo, 12);
    P.shader = c->shader;
    P.shin_int = is_integer_exponent(c->mat.shininess) ? (int32_t)c->mat.shininess : -1;
    P.clamp_spec = c->mat.clamp_specular;
    P.hierarchical = c->hierarchical;
    P.width = c->W;
    P.height = c->H;
    P.tiles_x = c->tilesX;
    P.tiles_y = c->tilesY;
    P.num_tiles = c->nTiles;
    P.num_vertices = V;
    P.num_triangles = T;
    P.num_chunks = choose_chunks(c, T);
    P.chunk_slots = (T + P.num_chunks - 1) / P.num_chunks;

    if (!c->clipQueueCap) c->clipQueueCap = std::max<uint32_t>(1024, std::min<uint32_t>(T, 1u << 16));
    if (!c->extraRecCap) c->extraRecCap = 4096;
    if (!c->extraVertCap) c->extraVertCap = 8192;
    if (!c->refCap) c->refCap = (uint32_t)std::min<uint64_t>(4ull * T + 4ull * c->nTiles + 1024, 0xFFFFFFF0ull);

    c->stats = edxr_stats{};
    for (uint32_t attempt = 0; attempt < 8; attempt++) {
        if ((uint64_t)T + c->extraRecCap > kRecordMask) return c->fail(EDXR_E_LIMIT, "record capacity exceeded");
        if ((e = c->rrec.ensure((size_t)T + c->extraRecCap)) != hipSuccess ||
            (e = c->srec.ensure((size_t)T + c->extraRecCap)) != hipSuccess || (e = c->desc.ensure(T)) != hipSuccess ||
            (e = c->slotRefs.ensure(T)) != hipSuccess ||
            (e = c->groupSums.ensure((size_t)((P.num_chunks + kChunkGroup - 1) / kChunkGroup) * c->nTiles)) !=
                hipSuccess ||
            (e = c->clipQueue.ensure(c->clipQueueCap)) != hipSuccess ||
            (e = c->binCounts.ensure((size_t)P.num_chunks * c->nTiles)) != hipSuccess ||
            (e = c->refs.ensure(c->refCap)) != hipSuccess ||
            (e = c->work.ensure((size_t)c->nTiles + c->refCap / kSegRefs + 2)) != hipSuccess ||
            (e = c->partials.ensure(c->work.cap * kTileSize * kTileSize)) != hipSuccess)
            return c->hip_fail(e, "hipMalloc(pipeline)");
        if ((e = c->verts.ensure((size_t)V + c->extraVertCap)) != hipSuccess)
            return c->hip_fail(e, "hipMalloc(vertices)");

        DeviceBuffers B;
        B.rrec = c->rrec.p;
        B.srec = c->srec.p;
        B.desc = c->desc.p;
        B.slot_refs = c->slotRefs.p;
        B.group_sums = c->groupSums.p;
        B.clip_queue = c->clipQueue.p;
        B.clip_queue_cap = c->clipQueueCap;
        B.extra_rec_cap = c->extraRecCap;
        B.verts = c->verts.p;
        B.extra_vert_cap = c->extraVertCap;
        B.bin_counts = c->binCounts.p;
        B.tile_totals = c->tileTotals.p;
        B.tile_start = c->tileStart.p;
        B.refs = c->refs.p;
        B.ref_cap = c->refCap;
        B.work = c->work.p;
        B.work_cap = (uint32_t)c->work.cap;
        B.tile_item = c->tileItem.p;
        B.partials = c->partials.p;
        // Persistent raster grid: enough resident workgroups per CU to hide latency.
        B.raster_blocks = c->numCUs * 8;
        B.counters = c->counters.p;
        B.color = c->color.p;
        B.depth = c->depth.p;
        B.ids = c->ids.p;

        if ((e = launch_frame(P, M, B, c->stream, c->profile ? c->events : nullptr)) != hipSuccess)
            return c->hip_fail(e, "kernel launch");
        if ((e = hipMemcpyAsync(c->hostCounters, c->counters.p, sizeof(uint32_t) * kCntSlots, hipMemcpyDeviceToHost,
                                c->stream)) != hipSuccess)
            return c->hip_fail(e, "hipMemcpyAsync(counters)");
        if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return c->hip_fail(e, "hipStreamSynchronize");
        const uint32_t* h = c->hostCounters;
        const uint32_t ov = h[kCntOverflow];
        if (ov & kOvPolygon) return c->fail(EDXR_E_LIMIT, "clip polygon exceeded 9 vertices");
        if (!ov) {
            if (h[kCntBadIndex]) return c->fail(EDXR_E_INDEX, "triangle index >= num_vertices");
            uint64_t rejected = 0, mainRecords = 0;
            for (uint32_t s = 0; s < kStripes; s++) {
                rejected += h[kCntStriped + kStRejected * kStripes + s];
                mainRecords += h[kCntStriped + kStMainRecords * kStripes + s];
            }
            c->stats.triangles_in = T;
            c->stats.trivially_rejected = rejected;
            c->stats.clipped = h[kCntClipped];
            c->stats.near_clipped = h[kCntNearClipped];
            c->stats.setup_records = mainRecords + h[kCntExtraRecords];
            c->stats.refs = h[kCntTotalRefs];
            c->stats.clip_vertices = h[kCntExtraVertices];
            c->stats.retries = attempt;
            c->stats.bin_chunks = P.num_chunks;
            c->rendered = true;
            c->lastT = T;
            c->lastV = V;
            c->lastChunks = P.num_chunks;
            c->lastMesh = M;
            if (c->profile) {
                for (int k = 0; k < kNumStages; k++) {
                    float ms = 0.0f;
                    if (hipEventElapsedTime(&ms, c->events[k], c->events[k + 1]) == hipSuccess) c->stageMs[k] += ms;
                }
                c->profiledFrames++;
            }
            return EDXR_OK;
        }
        auto grow = [](uint32_t need) {
            return (uint32_t)std::min<uint64_t>((uint64_t)need + need / 4 + 1024, 0xFFFFFFF0ull);
        };
        if (ov & kOvClipQueue) c->clipQueueCap = grow(h[kCntClipQueue]);
        if (ov & kOvExtraRecords) c->extraRecCap = grow(std::max(h[kCntExtraRecords], 2 * c->extraRecCap));
        if (ov & kOvExtraVertices) c->extraVertCap = grow(std::max(h[kCntExtraVertices], 2 * c->extraVertCap));
        if (ov & kOvRefs) c->refCap = grow(h[kCntTotalRefs]);
        // A clip-queue overflow hides the clip-stage demand: grow those too.
        if (ov & kOvClipQueue) {
            c->extraRecCap = std::max(c->extraRecCap, grow(2 * c->extraRecCap));
            c->extraVertCap = std::max(c->extraVertCap, grow(2 * c->extraVertCap));
        }
    }
    return c->fail(EDXR_E_LIMIT, "capacity retries exhausted");
}

int edxr_read_color(edxr_ctx* c, uint8_t* rgba, size_t bytes) {
    if (!c || !rgba) return EDXR_E_INVALID;
    if (!c->rendered) return c->fail(EDXR_E_STATE, "nothing rendered");
    const size_t need = (size_t)c->W * c->H * 4;
    if (bytes < need) return c->fail(EDXR_E_INVALID, "buffer too small");
    hipError_t e = hipMemcpyAsync(rgba, c->color.p, need, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? EDXR_OK : c->hip_fail(e, "read_color");
}

int edxr_read_depth(edxr_ctx* c, float* depth, size_t count) {
    if (!c || !depth) return EDXR_E_INVALID;
    if (!c->rendered) return c->fail(EDXR_E_STATE, "nothing rendered");
    const size_t need = (size_t)c->W * c->H;
    if (count < need) return c->fail(EDXR_E_INVALID, "buffer too small");
    hipError_t e = hipMemcpyAsync(depth, c->depth.p, need * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? EDXR_OK : c->hip_fail(e, "read_depth");
}

int edxr_read_triangle_ids(edxr_ctx* c, uint32_t* ids, size_t count) {
    if (!c || !ids) return EDXR_E_INVALID;
    if (!c->rendered) return c->fail(EDXR_E_STATE, "nothing rendered");
    const size_t need = (size_t)c->W * c->H;
    if (count < need) return c->fail(EDXR_E_INVALID, "buffer too small");
    hipError_t e = hipMemcpyAsync(ids, c->ids.p, need * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? EDXR_OK : c->hip_fail(e, "read_triangle_ids");
}

int edxr_device_buffers(edxr_ctx* c, void** color, void** depth, void** ids) {
    if (!c) return EDXR_E_INVALID;
    if (color) *color = c->color.p;
    if (depth) *depth = c->depth.p;
    if (ids) *ids = c->ids.p;
    return EDXR_OK;
}

int edxr_get_stats(edxr_ctx* c, edxr_stats* out) {
    if (!c || !out) return EDXR_E_INVALID;
    *out = c->stats;
    return EDXR_OK;
}

int edxr_profile_read(edxr_ctx* c, double* stage_ms, uint64_t* frames) {
    if (!c) return EDXR_E_INVALID;
    if (stage_ms)
        for (int k = 0; k < kNumStages; k++) stage_ms[k] = c->stageMs[k];
    if (frames) *frames = c->profiledFrames;
    return EDXR_OK;
}

int edxr_profile_reset(edxr_ctx* c) {
    if (!c) return EDXR_E_INVALID;
    for (double& v : c->stageMs) v = 0.0;
    c->profiledFrames = 0;
    return EDXR_OK;
}

void* edxr_stream(edxr_ctx* c) { return c ? (void*)c->stream : nullptr; }

const char* edxr_last_error(const edxr_ctx* c) { return c ? c->err.c_str() : "null context"; }

int64_t edxr_debug_read_records(edxr_ctx* c, uint32_t* out, uint64_t cap) {
    if (!c) return EDXR_E_INVALID;
    if (!c->rendered) return c->fail(EDXR_E_STATE, "nothing rendered");
    const uint32_t T = c->lastT;
    std::vector<uint32_t> desc(T);
    const uint64_t nrec = (uint64_t)T + c->extraRecCap;
    std::vector<RasterRec> rr(nrec);
    std::vector<ShadeRec> sr(nrec);
    std::vector<uint32_t> idx(3ull * T);
    hipError_t e = hipSuccess;
    if (T) e = hipMemcpy(desc.data(), c->desc.p, 4ull * T, hipMemcpyDeviceToHost);
    if (T && e == hipSuccess) e = hipMemcpy(idx.data(), c->lastMesh.idx, 12ull * T, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(rr.data(), c->rrec.p, sizeof(RasterRec) * nrec, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(sr.data(), c->srec.p, sizeof(ShadeRec) * nrec, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "debug_read_records");
    uint64_t n = 0;
    for (uint32_t t = 0; t < T; t++) {
        const bool compact = (desc[t] & kDescCompact) != 0;
        const uint32_t first = compact ? t : desc[t] >> 3, cnt = compact ? 1u : desc[t] & 7u;
        for (uint32_t k = 0; k < cnt; k++, n++) {
            if (n >= cap || !out) continue;
            const RasterRec& r = rr[first + k];
            // Unclipped triangles keep their input vertex ids (no ShadeRec).
            const ShadeRec s = (r.order & 1u) ? sr[first + k] : ShadeRec{idx[3 * t], idx[3 * t + 1], idx[3 * t + 2], t};
            uint32_t* o = out + 16 * n;
            o[0] = (uint32_t)r.x0; o[1] = (uint32_t)r.y0; o[2] = (uint32_t)r.x1; o[3] = (uint32_t)r.y1;
            o[4] = (uint32_t)r.x2; o[5] = (uint32_t)r.y2;
            memcpy(&o[6], &r.inv_det, 4);
            memcpy(&o[7], &r.z0, 4);
            memcpy(&o[8], &r.z1, 4);
            memcpy(&o[9], &r.z2, 4);
            o[10] = s.tri_id;
            o[11] = (r.info >> 12) & 7u;
            o[12] = r.info & 0xFFFu;
            o[13] = s.v0; o[14] = s.v1; o[15] = s.v2;
        }
    }
    return (int64_t)n;
}

int64_t edxr_debug_read_bins(edxr_ctx* c, uint64_t* offsets, uint32_t* refs3, uint64_t cap) {
    if (!c) return EDXR_E_INVALID;
    if (!c->rendered) return c->fail(EDXR_E_STATE, "nothing rendered");
    std::vector<uint32_t> ts(c->nTiles + 1);
    hipError_t e = hipMemcpy(ts.data(), c->tileStart.p, 4ull * (c->nTiles + 1), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "debug_read_bins");
    const uint64_t total = ts[c->nTiles];
    std::vector<uint32_t> refs(total);
    const uint64_t nrec = (uint64_t)c->lastT + c->extraRecCap;
    std::vector<RasterRec> rr(nrec);
    std::vector<ShadeRec> sr(nrec);
    if (total) e = hipMemcpy(refs.data(), c->refs.p, 4ull * total, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(rr.data(), c->rrec.p, sizeof(RasterRec) * nrec, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(sr.data(), c->srec.p, sizeof(ShadeRec) * nrec, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "debug_read_bins");
    if (offsets)
        for (uint32_t i = 0; i <= c->nTiles; i++) offsets[i] = ts[i];
    if (refs3 && total <= cap)
        for (uint64_t i = 0; i < total; i++) {
            const uint32_t p = refs[i], ri = p & kRecordMask;
            const uint32_t acc = (p >> kRecordBits) & 7u;
            const uint32_t big = p >> 31;
            refs3[3 * i] = rr[ri].order >> 4;
            refs3[3 * i + 1] = (rr[ri].info >> 12) & 7u;
            refs3[3 * i + 2] = acc | ((acc == 7u) ? 8u : 0u) | (big ? 16u : 0u);
        }
    return (int64_t)total;
}

}  // extern "C"
