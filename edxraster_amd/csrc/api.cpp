// api.cpp -- C ABI of the MI355X EDXRaster pipeline (host orchestration).
//
// Replaces EDXRaster's Renderer (Core/Renderer.h:37-50, Core/Renderer.cpp):
// the per-context state replaces the RenderStates singleton
// (Core/RenderStates.h:38-52), device buffers replace the per-core arrays, and
// RenderMesh becomes one ordered sequence of kernel launches on the context's
// stream (pipeline.hip).  Fixed-capacity device buffers report overflow through
// counters; the frame is then re-run with larger buffers.
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
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

// Pinned host staging buffer (hipHostMalloc), grown on demand.
template <class T>
struct HostBuf {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n) {
        if (n <= cap && p) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc((void**)&p, sizeof(T) * std::max<size_t>(n, 1), hipHostMallocDefault);
        if (e == hipSuccess) cap = std::max<size_t>(n, 1);
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Host description of one job before placement in the batch arrays.
struct JobInput {
    MeshPtrs mesh;
    const TexDesc* tex = nullptr;
    const float4* texels = nullptr;
    uint32_t numTex = 0;
    int32_t texFilter = 0;
    uint32_t V, T, W, H;
    uint32_t sampleLevel;
    float mvp[16], raster[16], eye[3], light[3];
    edxr_material mat;
    int32_t shader, hierarchical;
};

// A batch laid out for launch_frame: job table, batch sizes, block -> job
// tables (vertex blocks | setup blocks | chunks | tiles) and buffer sizes.
struct Plan {
    std::vector<FrameParams> jobs;
    BatchParams Q{};
    std::vector<uint32_t> tables;
    uint64_t countEntries = 0, groupEntries = 0, pixels = 0;
};

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
    uint32_t trisPerChunk = 0, rasterPerCU = 0;
    int32_t profile = 0;
    int32_t useGraph = 0;
    int32_t floatColor = 0;
    int32_t msaaLog2 = 0;
    int32_t writeFrames = 0;      // Renderer::SetWriteFrames
    int64_t frameCount = 0;       // RenderStates::FrameCount
    std::string frameDir = ".";   // Application::GetBaseDirectory
    hipEvent_t events[kNumStages + 1] = {};
    double stageMs[kNumStages] = {};
    uint64_t profiledFrames = 0;

    // owned copies of host meshes
    DevBuf<float> mesh[11];   // the 9 SoA planes, then tu, tv
    DevBuf<uint32_t> meshIdx, meshTex;
    DevBuf<TexDesc> texDesc;  // texture slots (edxr_set_textures)
    DevBuf<float4> texels;
    uint32_t numTex = 0;
    int32_t texFilter = kFilterTrilinear;
    DevBuf<uint8_t> batchMesh;
    // pipeline buffers
    DevBuf<RasterRec> rrec;
    DevBuf<ShadeRec> srec;
    DevBuf<uint32_t> desc, clipQueue, binCounts, tileTotals, tileStart, refs, counters, color, ids, work, tileItem,
        tileBig;
    DevBuf<float> depth;
    DevBuf<uint64_t> partials;
    DevBuf<uint32_t> chunkBox;
    DevBuf<uint32_t> groupSums;
    DevBuf<ShadeVert> verts;  // shading vertices + extraVertCap clip-created ones
    DevBuf<ScreenVert> sverts;
    DevBuf<FrameParams> jobsDev;
    DevBuf<uint32_t> jobTables;
    HostBuf<FrameParams> jobsHost;
    HostBuf<uint32_t> tablesHost;
    // batch outputs
    DevBuf<uint32_t> bcolor, bids;
    DevBuf<float> bdepth;
    DevBuf<float> colorF, bcolorF;  // float RGB images (EDXR_STATE_FLOAT_COLOR)
    // adjoint
    Plan lastPlan;                  // layout of the last render (tables stay resident)
    bool lastPlanValid = false, lastFloat = false;
    DevBuf<GradPtrs> gradsDev;
    HostBuf<GradPtrs> gradsHost;
    DevBuf<double> adjAcc, adjVertAcc;
    std::vector<uint64_t> batchPix;
    std::vector<uint32_t> batchW, batchH;
    bool batchRendered = false;

    uint32_t numCUs = 256;
    uint32_t clipQueueCap = 0, extraRecCap = 0, extraVertCap = 0, refCap = 0;
    uint32_t debugCapacity = 0;     // EDXR_STATE_DEBUG_CAPACITY
    // The one-pass column scan gave up waiting once (or debug capacities make
    // it give up at once): three-kernel column scans from then on.
    bool legacyScans = false;
    uint8_t* scratch = nullptr;     // device address of this device's no-allocation scratch
    uint64_t renderGeneration = 0;  // advanced by every render_mesh / render_batch
    uint32_t* hostCounters = nullptr;

    // captured frame
    hipGraphExec_t graphExec = nullptr;
    std::vector<uint8_t> graphKey;
    // Single-job graphs: the captured graph (kept for its first node, which
    // takes the job by value) and the job its first node holds now.
    hipGraph_t graph = nullptr;
    hipGraphNode_t graphFirst = nullptr;
    FrameParams graphJob{};

    // last single frame
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
        return e == hipErrorOutOfMemory ? EDXR_E_NOMEM : EDXR_E_HIP;
    }
};

namespace {

bool tile_grid(uint32_t w, uint32_t h, uint32_t& tx, uint32_t& ty) {
    if (w == 0 || h == 0 || w > kMaxDim || h > kMaxDim) return false;
    tx = (w + kTileSize - 1) >> kTileLog2;
    ty = (h + kTileSize - 1) >> kTileLog2;
    return (uint64_t)tx * ty <= kMaxTiles;
}

int set_size(edxr_ctx* c, uint32_t w, uint32_t h) {
    uint32_t tx, ty;
    if (!tile_grid(w, h, tx, ty)) return c->fail(EDXR_E_LIMIT, "resolution out of range (<= 8192, <= 16384 tiles)");
    c->W = w;
    c->H = h;
    c->tilesX = tx;
    c->tilesY = ty;
    c->nTiles = tx * ty;
    const size_t P = (size_t)w * h, S = (size_t)1 << c->msaaLog2;
    hipError_t e;
    if ((e = c->color.ensure(P)) != hipSuccess || (e = c->depth.ensure(S * P)) != hipSuccess ||
        (e = c->ids.ensure(S * P)) != hipSuccess)
        return c->hip_fail(e, "hipMalloc(frame buffers)");
    c->rendered = false;
    return EDXR_OK;
}

// Binning chunks of a job: about trisPerChunk (1024) triangles each so the
// count/scatter grids fill the GPU, capped so the job's chunk x tile count
// matrix (allocated in full, touched only inside each chunk's tile box) stays
// within matrixBytes.
// Sums of the striped counters of a frame (h: the counters read back).  The
// clip stage allocates from kAllocStripes regions; a capacity that holds its
// fullest stripe's demand in every region is that demand times the stripes.
struct CounterSums {
    uint64_t rejected = 0, mainRecords = 0, clipped = 0, nearClipped = 0, extraVerts = 0, extraRecs = 0;
    uint32_t vertDemand = 0, recDemand = 0;
};
CounterSums counter_sums(const uint32_t* h) {
    CounterSums cs;
    for (uint32_t s = 0; s < kStripes; s++) {
        cs.rejected += h[kCntStriped + kStRejected * kStripes + s];
        cs.mainRecords += h[kCntStriped + kStMainRecords * kStripes + s];
        cs.clipped += h[kCntStriped + kStClipped * kStripes + s];
        cs.nearClipped += h[kCntStriped + kStNearClipped * kStripes + s];
    }
    uint32_t vmax = 0, rmax = 0;
    for (uint32_t s = 0; s < kAllocStripes; s++) {
        const uint32_t v = h[kCntAllocStriped + s], r = h[kCntAllocStriped + kAllocStripes + s];
        cs.extraVerts += v;
        cs.extraRecs += r;
        vmax = std::max(vmax, v);
        rmax = std::max(rmax, r);
    }
    cs.vertDemand = (uint32_t)std::min<uint64_t>((uint64_t)vmax * kAllocStripes, 0xFFFFFFF0ull);
    cs.recDemand = (uint32_t)std::min<uint64_t>((uint64_t)rmax * kAllocStripes, 0xFFFFFFF0ull);
    return cs;
}

uint32_t choose_chunks(const edxr_ctx* c, uint32_t T, uint32_t nTiles, bool single,
                       uint64_t matrixBytes = 256ull << 20) {
    if (T == 0) return 0;
    const uint32_t byMatrix = (uint32_t)std::max<uint64_t>(1, matrixBytes / (4ull * nTiles));
    if (single && c->binChunksSetting)
        return std::max(1u, std::min(std::min(std::min(c->binChunksSetting, kMaxChunks), T), byMatrix));
    const uint32_t tpc = c->trisPerChunk ? c->trisPerChunk : 1024;
    const uint32_t C = (T + tpc - 1) / tpc;
    return std::max(1u, std::min(std::min(C, byMatrix), kMaxChunks));
}

int check_mesh(edxr_ctx* c, const edxr_mesh* m, int32_t shader) {
    if (!m->px || !m->py || !m->pz || (m->num_triangles && !m->indices))
        return c->fail(EDXR_E_INVALID, "mesh positions/indices missing");
    if (!m->nx || !m->ny || !m->nz) return c->fail(EDXR_E_INVALID, "mesh normals missing");
    if (shader == EDXR_SHADER_VC_BLINN_PHONG && (!m->cr || !m->cg || !m->cb))
        return c->fail(EDXR_E_INVALID, "vertex colours required by EDXR_SHADER_VC_BLINN_PHONG");
    if (m->memory != EDXR_MEM_HOST && m->memory != EDXR_MEM_DEVICE) return c->fail(EDXR_E_INVALID, "bad mesh memory");
    if (shader == EDXR_SHADER_TEXTURED_LAMBERT) {
        if (!m->tu || !m->tv || (m->num_triangles && !m->tex_ids))
            return c->fail(EDXR_E_INVALID, "texture coordinates and slots required by EDXR_SHADER_TEXTURED_LAMBERT");
        if (!c->numTex) return c->fail(EDXR_E_STATE, "no textures set (edxr_set_textures)");
    }
    return EDXR_OK;
}

MeshPtrs mesh_ptrs(const float* const f[11], const uint32_t* idx, const uint32_t* tex) {
    return MeshPtrs{f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], idx, f[9], f[10], tex};
}

// Per-job constants; the light is normalised on the host as Math::Normalize
// in the pixel shader does (Core/Shader.h:221).
FrameParams job_params(const JobInput& in, uint32_t tilesX, uint32_t tilesY) {
    FrameParams P;
    memset(&P, 0, sizeof(P));
    memcpy(P.mvp, in.mvp, 64);
    memcpy(P.raster, in.raster, 64);
    memcpy(P.eye, in.eye, 12);
    const float* L = in.light;
    const float s = 1.0f / sqrtf((L[0] * L[0] + L[1] * L[1]) + L[2] * L[2]);
    P.light[0] = L[0] * s;
    P.light[1] = L[1] * s;
    P.light[2] = L[2] * s;
    P.light_len = 1.0f / s;
    P.ambient = in.mat.ambient;
    P.kd = in.mat.kd;
    P.ks = in.mat.ks;
    P.shininess = in.mat.shininess;
    memcpy(P.albedo, in.mat.albedo, 12);
    P.shader = in.shader;
    P.shin_int = is_integer_exponent(in.mat.shininess) ? (int32_t)in.mat.shininess : -1;
    P.clamp_spec = in.mat.clamp_specular;
    P.hierarchical = in.hierarchical;
    P.width = in.W;
    P.height = in.H;
    P.tiles_x = tilesX;
    P.tiles_y = tilesY;
    P.num_tiles = tilesX * tilesY;
    P.num_vertices = in.V;
    P.num_triangles = in.T;
    P.sample_level = in.sampleLevel;
    P.samples = 1u << in.sampleLevel;
    P.mesh = in.mesh;
    P.tex = in.tex;
    P.texels = in.texels;
    P.num_tex = in.numTex;
    P.tex_filter = (uint32_t)in.texFilter;
    return P;
}

// Places the jobs in the batch arrays.  Vertex and triangle slots start on
// kJobAlign boundaries so each block of k_vertex / k_setup belongs to one job;
// the block -> job tables have one entry per kJobTableBlock slots.
int build_plan(edxr_ctx* c, const std::vector<JobInput>& in, Plan& pl) {
    const bool single = in.size() == 1;
    pl.jobs.clear();
    pl.tables.clear();
    uint64_t vert = 0, slot = 0, tile = 0, chunk = 0;
    pl.countEntries = pl.groupEntries = pl.pixels = 0;
    BatchParams& Q = pl.Q;
    memset(&Q, 0, sizeof(Q));  // padding included: part of the graph key
    for (const JobInput& j : in) {
        uint32_t tx, ty;
        if (!tile_grid(j.W, j.H, tx, ty)) return c->fail(EDXR_E_LIMIT, "job resolution out of range");
        FrameParams P = job_params(j, tx, ty);
        vert = (vert + kJobAlign - 1) / kJobAlign * kJobAlign;
        slot = (slot + kJobAlign - 1) / kJobAlign * kJobAlign;
        P.vert_base = (uint32_t)vert;
        P.slot_base = (uint32_t)slot;
        P.tile_base = (uint32_t)tile;
        P.chunk_base = (uint32_t)chunk;
        P.num_chunks = choose_chunks(c, j.T, P.num_tiles, single);
        P.chunk_slots = P.num_chunks ? (j.T + P.num_chunks - 1) / P.num_chunks : 0;
        P.count_base = pl.countEntries;
        P.group_base = pl.groupEntries;
        P.pix_base = pl.pixels;
        vert += j.V;
        slot += j.T;
        tile += P.num_tiles;
        chunk += P.num_chunks;
        pl.countEntries += (uint64_t)P.num_chunks * P.num_tiles;
        pl.groupEntries += (uint64_t)((P.num_chunks + kChunkGroup - 1) / kChunkGroup) * P.num_tiles;
        pl.pixels += (uint64_t)j.W * j.H;
        Q.max_job_tiles = std::max(Q.max_job_tiles, P.num_tiles);
        Q.max_job_chunks = std::max(Q.max_job_chunks, P.num_chunks);
        Q.max_samples = std::max(Q.max_samples, P.samples);
        if (P.shader == kTexturedLambert) Q.any_textured = 1;
        pl.jobs.push_back(P);
        if (vert > 0xF0000000ull || slot + 65536 > kRecordMask) return c->fail(EDXR_E_LIMIT, "too many triangles");
    }
    if (tile > kMaxBatchTiles) return c->fail(EDXR_E_LIMIT, "batch exceeds 262144 tiles");
    Q.num_jobs = (uint32_t)in.size();
    Q.total_verts = (uint32_t)vert;
    Q.total_slots = (uint32_t)slot;
    Q.total_tiles = (uint32_t)tile;
    Q.total_chunks = (uint32_t)chunk;
    Q.vert_blocks = (uint32_t)((vert + kJobTableBlock - 1) / kJobTableBlock);
    Q.slot_blocks = (uint32_t)((slot + kJobTableBlock - 1) / kJobTableBlock);
    if (!single) {
        pl.tables.resize((size_t)Q.vert_blocks + Q.slot_blocks + Q.total_chunks + Q.total_tiles);
        uint32_t* vb = pl.tables.data();
        uint32_t* sb = vb + Q.vert_blocks;
        uint32_t* cj = sb + Q.slot_blocks;
        uint32_t* tj = cj + Q.total_chunks;
        // Padding blocks up to the next job's base belong to the job before it.
        for (uint32_t k = 0; k < Q.num_jobs; k++) {
            const FrameParams& P = pl.jobs[k];
            const uint32_t vEnd = k + 1 < Q.num_jobs ? pl.jobs[k + 1].vert_base : Q.total_verts;
            const uint32_t sEnd = k + 1 < Q.num_jobs ? pl.jobs[k + 1].slot_base : Q.total_slots;
            for (uint32_t b = P.vert_base / kJobTableBlock; b < (vEnd + kJobTableBlock - 1) / kJobTableBlock; b++)
                vb[b] = k;
            for (uint32_t b = P.slot_base / kJobTableBlock; b < (sEnd + kJobTableBlock - 1) / kJobTableBlock; b++)
                sb[b] = k;
            for (uint32_t i = 0; i < P.num_chunks; i++) cj[P.chunk_base + i] = k;
            for (uint32_t i = 0; i < P.num_tiles; i++) tj[P.tile_base + i] = k;
        }
    }
    return EDXR_OK;
}

// Bytes of the launch layout: a captured graph is replayed while they repeat.
std::vector<uint8_t> launch_key(const BatchParams& Q, const DeviceBuffers& B, const FrameParams* jobs) {
    std::vector<uint8_t> k(sizeof(Q) + sizeof(B) + sizeof(jobs));
    memcpy(k.data(), &Q, sizeof(Q));
    memcpy(k.data() + sizeof(Q), &B, sizeof(B));
    memcpy(k.data() + sizeof(Q) + sizeof(B), &jobs, sizeof(jobs));
    return k;
}

void drop_graph(edxr_ctx* c) {
    if (c->graphExec) (void)hipGraphExecDestroy(c->graphExec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graphExec = nullptr;
    c->graph = nullptr;
    c->graphFirst = nullptr;
    c->graphKey.clear();
}

// Runs a planned batch (sized, launched, checked; re-run with larger buffers
// after a capacity overflow).  For a single job the job table is written by
// the frame's first kernel, which takes the job by value (in a graph replay,
// by an update of that node's arguments); otherwise it is uploaded from
// pinned memory.
int run_plan(edxr_ctx* c, Plan& pl, uint32_t* color, float* depth, uint32_t* ids, float* colorF) {
    hipError_t e;
    BatchParams& Q = pl.Q;
    const bool single = Q.num_jobs == 1;
    const bool graph = c->useGraph != 0;
    if ((e = c->jobsDev.ensure(Q.num_jobs)) != hipSuccess ||
        (e = c->tileTotals.ensure(Q.total_tiles)) != hipSuccess ||
        (e = c->tileStart.ensure((size_t)Q.total_tiles + 1)) != hipSuccess ||
        (e = c->tileItem.ensure(Q.total_tiles)) != hipSuccess || (e = c->tileBig.ensure(Q.total_tiles)) != hipSuccess)
        return c->hip_fail(e, "hipMalloc(job tables)");
    if (!single) {
        if ((e = c->jobsHost.ensure(Q.num_jobs)) != hipSuccess) return c->hip_fail(e, "hipHostMalloc(jobs)");
        memcpy(c->jobsHost.p, pl.jobs.data(), sizeof(FrameParams) * Q.num_jobs);
        if ((e = hipMemcpyAsync(c->jobsDev.p, c->jobsHost.p, sizeof(FrameParams) * Q.num_jobs,
                                hipMemcpyHostToDevice, c->stream)) != hipSuccess)
            return c->hip_fail(e, "hipMemcpyAsync(jobs)");
    }
    if (!pl.tables.empty()) {
        if ((e = c->jobTables.ensure(pl.tables.size())) != hipSuccess ||
            (e = c->tablesHost.ensure(pl.tables.size())) != hipSuccess)
            return c->hip_fail(e, "hipMalloc(job tables)");
        memcpy(c->tablesHost.p, pl.tables.data(), 4 * pl.tables.size());
        if ((e = hipMemcpyAsync(c->jobTables.p, c->tablesHost.p, 4 * pl.tables.size(), hipMemcpyHostToDevice,
                                c->stream)) != hipSuccess)
            return c->hip_fail(e, "hipMemcpyAsync(job tables)");
        Q.vblock_job = c->jobTables.p;
        Q.sblock_job = Q.vblock_job + Q.vert_blocks;
        Q.chunk_job = Q.sblock_job + Q.slot_blocks;
        Q.tile_job = Q.chunk_job + Q.total_chunks;
    }
    const uint32_t T = Q.total_slots, V = Q.total_verts;
    if (c->debugCapacity) {
        // Every frame starts from tiny capacities, so the overflow flags and
        // the grow-and-rerun loop below run (EDXR_STATE_DEBUG_CAPACITY).
        c->clipQueueCap = c->extraRecCap = c->extraVertCap = c->refCap = c->debugCapacity;
    } else {
        if (!c->clipQueueCap) c->clipQueueCap = std::max<uint32_t>(1024, std::min<uint32_t>(T, 1u << 16));
        // (the clip stage takes them in kAllocStripes regions)
        if (!c->extraRecCap) c->extraRecCap = 16384;
        if (!c->extraVertCap) c->extraVertCap = 16384;
        const uint32_t refWant = (uint32_t)std::min<uint64_t>(4ull * T + 4ull * Q.total_tiles + 1024, 0xFFFFFFF0ull);
        c->refCap = std::max(c->refCap, refWant);
    }

    // Raster work items per frame the segment size aims for (tuning).
    const uint32_t rasterBlocks = c->numCUs * (c->rasterPerCU ? c->rasterPerCU : 8);
    c->stats = edxr_stats{};
    for (uint32_t attempt = 0; attempt < 8; attempt++) {
        if ((uint64_t)T + c->extraRecCap > kRecordMask) return c->fail(EDXR_E_LIMIT, "record capacity exceeded");
        if ((e = c->rrec.ensure((size_t)T + c->extraRecCap)) != hipSuccess ||
            (e = c->srec.ensure((size_t)T + c->extraRecCap)) != hipSuccess ||
            (e = c->desc.ensure(std::max<uint32_t>(T, 1))) != hipSuccess ||
            (e = c->chunkBox.ensure(2 * std::max<uint32_t>(Q.total_chunks, 1))) != hipSuccess ||
            (e = c->groupSums.ensure(std::max<uint64_t>(pl.groupEntries, 1))) != hipSuccess ||
            (e = c->clipQueue.ensure(c->clipQueueCap)) != hipSuccess ||
            (e = c->binCounts.ensure(std::max<uint64_t>(pl.countEntries, 1))) != hipSuccess ||
            (e = c->refs.ensure(c->refCap)) != hipSuccess ||
            (e = c->work.ensure((size_t)Q.total_tiles + std::max(rasterBlocks, c->refCap / kSegRefs) + 2)) !=
                hipSuccess ||
            (e = c->partials.ensure(c->work.cap * kTileSize * kTileSize * Q.max_samples)) != hipSuccess ||
            (e = c->verts.ensure((size_t)V + c->extraVertCap)) != hipSuccess ||
            (e = c->sverts.ensure(std::max<uint32_t>(V, 1))) != hipSuccess)
            return c->hip_fail(e, "hipMalloc(pipeline)");

        // Zeroed as a whole: the structure's bytes, padding included, are
        // the graph key (launch_key).
        DeviceBuffers B;
        memset(&B, 0, sizeof(B));
        B.rrec = c->rrec.p;
        B.srec = c->srec.p;
        B.desc = c->desc.p;
        B.chunk_box = c->chunkBox.p;
        B.group_sums = c->groupSums.p;
        B.group_entries = pl.groupEntries;
        B.scan_spin_limit = c->debugCapacity ? 0u : kScanSpinLimit;
        B.legacy_scans = c->legacyScans ? 1u : 0u;
        B.clip_queue = c->clipQueue.p;
        B.clip_queue_cap = c->clipQueueCap;
        B.extra_rec_cap = c->extraRecCap;
        B.verts = c->verts.p;
        B.sverts = c->sverts.p;
        B.extra_vert_cap = c->extraVertCap;
        B.bin_counts = c->binCounts.p;
        B.tile_totals = c->tileTotals.p;
        B.tile_start = c->tileStart.p;
        B.refs = c->refs.p;
        B.ref_cap = c->refCap;
        B.work = c->work.p;
        B.work_cap = (uint32_t)c->work.cap;
        B.tile_item = c->tileItem.p;
        B.tile_big = c->tileBig.p;
        B.partials = c->partials.p;
        B.raster_blocks = rasterBlocks;
        // The persistent raster grid is the resident capacity: workgroups
        // beyond it would start only when a resident one exits, i.e. after the
        // work queue has drained.  The tuning's workgroups per CU set the
        // segment size only (raster_blocks).
        B.raster_grid = c->numCUs * raster_resident_per_cu(Q.max_samples > 1);
        B.num_cus = c->numCUs;
        B.counters = c->counters.p;
        B.host_counters = c->hostCounters;
        B.color = color;
        B.depth = depth;
        B.ids = ids;
        B.color_f32 = colorF;

        const bool prof = c->profile && !graph;
        if (graph) {
            const std::vector<uint8_t> key = launch_key(Q, B, c->jobsDev.p);
            const FrameParams* job0 = single ? &pl.jobs[0] : nullptr;
            if (!c->graphExec || key != c->graphKey) {
                drop_graph(c);
                if ((e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                    return c->hip_fail(e, "hipStreamBeginCapture");
                const hipError_t le = launch_frame(Q, c->jobsDev.p, job0, B, c->stream, nullptr);
                e = hipStreamEndCapture(c->stream, &c->graph);
                if (le != hipSuccess) e = le;
                if (e == hipSuccess && job0) {
                    // the first kernel is the capture's one root
                    size_t n = 1;
                    hipGraphNodeType type = hipGraphNodeTypeEmpty;
                    e = hipGraphGetRootNodes(c->graph, &c->graphFirst, &n);
                    if (e == hipSuccess) e = hipGraphNodeGetType(c->graphFirst, &type);
                    if (e == hipSuccess && (n != 1 || type != hipGraphNodeTypeKernel)) e = hipErrorUnknown;
                    c->graphJob = *job0;
                }
                if (e == hipSuccess) e = hipGraphInstantiate(&c->graphExec, c->graph, nullptr, nullptr, 0);
                if (e != hipSuccess) {
                    drop_graph(c);
                    return c->hip_fail(e, "graph capture");
                }
                c->graphKey = key;
            } else if (job0 && memcmp(job0, &c->graphJob, sizeof(FrameParams)) != 0) {
                if ((e = set_frame_job(c->graphExec, c->graphFirst, Q, c->jobsDev.p, B, *job0)) != hipSuccess)
                    return c->hip_fail(e, "hipGraphExecKernelNodeSetParams");
                c->graphJob = *job0;
            }
            if ((e = hipGraphLaunch(c->graphExec, c->stream)) != hipSuccess) return c->hip_fail(e, "hipGraphLaunch");
        } else {
            const FrameParams* job0 = single ? &pl.jobs[0] : nullptr;
            if ((e = launch_frame(Q, c->jobsDev.p, job0, B, c->stream, prof ? c->events : nullptr)) != hipSuccess)
                return c->hip_fail(e, "kernel launch");
        }
        if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return c->hip_fail(e, "hipStreamSynchronize");
        const uint32_t* h = c->hostCounters;
        const uint32_t ov = h[kCntOverflow];
        if (ov & kOvPolygon) return c->fail(EDXR_E_LIMIT, "clip polygon exceeded 9 vertices");
        if (!ov) {
            if (h[kCntBadIndex])
                return c->fail(EDXR_E_INDEX, "triangle index >= num_vertices or texture slot >= texture count");
            uint64_t tin = 0;
            for (const FrameParams& P : pl.jobs) tin += P.num_triangles;
            const CounterSums cs = counter_sums(h);
            c->stats.triangles_in = tin;
            c->stats.trivially_rejected = cs.rejected;
            c->stats.clipped = cs.clipped;
            c->stats.near_clipped = cs.nearClipped;
            c->stats.setup_records = cs.mainRecords + cs.extraRecs;
            c->stats.refs = h[kCntTotalRefs];
            c->stats.clip_vertices = cs.extraVerts;
            c->stats.retries = attempt;
            c->stats.bin_chunks = Q.total_chunks;
            if (prof) {
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
        const CounterSums cs = counter_sums(h);
        if (ov & kOvExtraRecords) c->extraRecCap = grow(std::max(cs.recDemand, 2 * c->extraRecCap));
        if (ov & kOvExtraVertices) c->extraVertCap = grow(std::max(cs.vertDemand, 2 * c->extraVertCap));
        if (ov & kOvRefs) c->refCap = grow(std::max(h[kCntTotalRefs], c->refCap));
        if (ov & kOvScanWait) c->legacyScans = true;
        // A clip-queue overflow hides the clip-stage demand: grow those too.
        if (ov & kOvClipQueue) {
            c->extraRecCap = std::max(c->extraRecCap, grow(2 * c->extraRecCap));
            c->extraVertCap = std::max(c->extraVertCap, grow(2 * c->extraVertCap));
        }
        c->stats.retries = attempt + 1;
    }
    return c->fail(EDXR_E_LIMIT, "capacity retries exhausted");
}

// The job of a single-frame render: the mesh with the context's render target,
// transform, light, shader, states and texture slots.
JobInput context_job(const edxr_ctx* c, const MeshPtrs& M, uint32_t V, uint32_t T) {
    JobInput in;
    in.mesh = M;
    in.V = V;
    in.T = T;
    in.W = c->W;
    in.H = c->H;
    memcpy(in.mvp, c->mvp, 64);
    memcpy(in.raster, c->raster, 64);
    memcpy(in.eye, c->eye, 12);
    memcpy(in.light, c->lightDir, 12);
    in.mat = c->mat;
    in.shader = c->shader;
    in.hierarchical = c->hierarchical;
    in.sampleLevel = (uint32_t)c->msaaLog2;
    in.tex = c->texDesc.p;
    in.texels = c->texels.p;
    in.numTex = c->numTex;
    in.texFilter = c->texFilter;
    return in;
}

// Sub-ranges of the no-allocation scratch, 256-byte aligned, taken in order;
// `used` keeps counting past the end so the caller can report the need.
struct ScratchCarver {
    uint8_t* base;
    size_t used = 0;
    template <class T>
    T* take(uint64_t n) {
        const size_t at = (used + 255) & ~size_t(255);
        used = at + sizeof(T) * std::max<uint64_t>(n, 1);
        return used <= kScratchBytes ? reinterpret_cast<T*>(base + at) : nullptr;
    }
    bool fits() const { return used <= kScratchBytes; }
};

// Uploads the host-memory meshes of a batch into one device buffer and
// rewrites their pointers (device meshes are used in place).
int stage_batch_meshes(edxr_ctx* c, const edxr_job* jobs, uint32_t n, std::vector<JobInput>& in) {
    uint64_t bytes = 0;
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    for (uint32_t k = 0; k < n; k++) {
        const edxr_mesh& m = jobs[k].mesh;
        if (m.memory != EDXR_MEM_HOST) continue;
        const float* planes[11] = {m.px, m.py, m.pz, m.nx, m.ny, m.nz, m.cr, m.cg, m.cb, m.tu, m.tv};
        for (const float* p : planes)
            if (p) bytes += al(4ull * m.num_vertices);
        bytes += al(12ull * m.num_triangles);
        if (m.tex_ids) bytes += al(4ull * m.num_triangles);
    }
    if (!bytes) return EDXR_OK;
    hipError_t e;
    if ((e = c->batchMesh.ensure(bytes)) != hipSuccess) return c->hip_fail(e, "hipMalloc(batch meshes)");
    uint64_t off = 0;
    for (uint32_t k = 0; k < n; k++) {
        const edxr_mesh& m = jobs[k].mesh;
        if (m.memory != EDXR_MEM_HOST) continue;
        const float* planes[11] = {m.px, m.py, m.pz, m.nx, m.ny, m.nz, m.cr, m.cg, m.cb, m.tu, m.tv};
        const float* dst[11];
        for (int q = 0; q < 11; q++) {
            dst[q] = nullptr;
            if (!planes[q]) continue;
            uint8_t* d = c->batchMesh.p + off;
            if (m.num_vertices &&
                (e = hipMemcpyAsync(d, planes[q], 4ull * m.num_vertices, hipMemcpyHostToDevice, c->stream)) !=
                    hipSuccess)
                return c->hip_fail(e, "hipMemcpyAsync(batch mesh)");
            dst[q] = (const float*)d;
            off += al(4ull * m.num_vertices);
        }
        uint8_t* di = c->batchMesh.p + off;
        if (m.num_triangles &&
            (e = hipMemcpyAsync(di, m.indices, 12ull * m.num_triangles, hipMemcpyHostToDevice, c->stream)) !=
                hipSuccess)
            return c->hip_fail(e, "hipMemcpyAsync(batch indices)");
        off += al(12ull * m.num_triangles);
        const uint32_t* dt = nullptr;
        if (m.tex_ids) {
            uint8_t* d = c->batchMesh.p + off;
            if (m.num_triangles &&
                (e = hipMemcpyAsync(d, m.tex_ids, 4ull * m.num_triangles, hipMemcpyHostToDevice, c->stream)) !=
                    hipSuccess)
                return c->hip_fail(e, "hipMemcpyAsync(batch texture slots)");
            dt = (const uint32_t*)d;
            off += al(4ull * m.num_triangles);
        }
        in[k].mesh = mesh_ptrs(dst, (const uint32_t*)di, dt);
    }
    return EDXR_OK;
}

}  // namespace

extern "C" {

void edxr_default_material(int32_t kind, edxr_material* m) {
    if (!m) return;
    m->ambient = 0.2f;
    m->kd = 3.0f;
    m->ks = 3.0f;
    m->clamp_specular = 0;
    if (kind == EDXR_SHADER_LAMBERT_ALBEDO || kind == EDXR_SHADER_TEXTURED_LAMBERT) {
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
    if (e == hipSuccess) e = init_kernels();
    // The no-allocation scratch of this device (the current one): its
    // address is looked up here, once per device, never in a render call.
    if (e == hipSuccess) e = scratch_address(&c->scratch);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    // coherent: the frame's last kernel writes it, the host reads it after a stream synchronisation
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->hostCounters, sizeof(uint32_t) * kCntSlots, hipHostMallocCoherent);
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
    drop_graph(c);
    for (auto& b : c->mesh) b.release();
    c->meshIdx.release();
    c->meshTex.release();
    c->texDesc.release();
    c->texels.release();
    c->batchMesh.release();
    c->rrec.release();
    c->srec.release();
    for (DevBuf<uint32_t>* b : {&c->desc, &c->clipQueue, &c->binCounts, &c->tileTotals, &c->tileStart, &c->refs,
                                &c->counters, &c->color, &c->ids, &c->work, &c->tileItem, &c->tileBig, &c->groupSums,
                                &c->jobTables, &c->bcolor, &c->bids})
        b->release();
    c->depth.release();
    c->bdepth.release();
    c->colorF.release();
    c->bcolorF.release();
    c->gradsDev.release();
    c->gradsHost.release();
    c->adjAcc.release();
    c->adjVertAcc.release();
    c->partials.release();
    c->chunkBox.release();
    c->verts.release();
    c->sverts.release();
    c->jobsDev.release();
    c->jobsHost.release();
    c->tablesHost.release();
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
    if (!c || kind < EDXR_SHADER_LAMBERT_ALBEDO || kind > EDXR_SHADER_TEXTURED_LAMBERT) return EDXR_E_INVALID;
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
        case EDXR_STATE_GRAPH:
            c->useGraph = value ? 1 : 0;
            if (!value) drop_graph(c);
            return EDXR_OK;
        case EDXR_STATE_TRIS_PER_CHUNK:
            if (value < 0 || (value > 0 && value < 64) || value > (1 << 20))
                return c->fail(EDXR_E_INVALID, "triangles per chunk out of range (0 or 64..2^20)");
            c->trisPerChunk = (uint32_t)value;
            return EDXR_OK;
        case EDXR_STATE_FLOAT_COLOR: c->floatColor = value ? 1 : 0; return EDXR_OK;
        case EDXR_STATE_WRITE_FRAMES: c->writeFrames = value ? 1 : 0; return EDXR_OK;
        case EDXR_STATE_TEXTURE_FILTER:
            if (value < EDXR_FILTER_NEAREST || value > EDXR_FILTER_ANISO_16X)
                return c->fail(EDXR_E_INVALID, "unknown texture filter");
            c->texFilter = value;
            return EDXR_OK;
        case EDXR_STATE_MSAA: {
            if (value < 0 || value > (int32_t)kMaxSampleLevel)
                return c->fail(EDXR_E_INVALID, "MSAA level out of range (0..5)");
            // Renderer::SetMSAAMode resizes the frame buffer (Core/Renderer.cpp:94-98).
            c->msaaLog2 = value;
            return set_size(c, c->W, c->H);
        }
        case EDXR_STATE_DEBUG_CAPACITY:
            if (value < 0 || value > 65536) return c->fail(EDXR_E_INVALID, "debug capacity out of range (0..65536)");
            c->debugCapacity = (uint32_t)value;
            if (!value) {  // back to the defaults
                c->clipQueueCap = c->extraRecCap = c->extraVertCap = 0;
                c->legacyScans = false;
            }
            return EDXR_OK;
        case EDXR_STATE_RASTER_BLOCKS_PER_CU:
            if (value < 0 || value > 32) return c->fail(EDXR_E_INVALID, "raster blocks per CU out of range (0..32)");
            c->rasterPerCU = (uint32_t)value;
            return EDXR_OK;
        default: return c->fail(EDXR_E_INVALID, "unknown state key");
    }
}

int edxr_render_mesh(edxr_ctx* c, const edxr_mesh* m) {
    if (!c || !m) return EDXR_E_INVALID;
    if (!c->haveTransform) return c->fail(EDXR_E_STATE, "edxr_set_transform was not called");
    int r = check_mesh(c, m, c->shader);
    if (r != EDXR_OK) return r;
    const uint32_t V = m->num_vertices, T = m->num_triangles;
    // Record indices (T + clip records) must fit the 28-bit reference payload.
    if ((uint64_t)T + 65536 > kRecordMask) return c->fail(EDXR_E_LIMIT, "too many triangles");
    hipError_t e;

    MeshPtrs M;
    if (m->memory == EDXR_MEM_DEVICE) {
        const float* const f[11] = {m->px, m->py, m->pz, m->nx, m->ny, m->nz, m->cr, m->cg, m->cb, m->tu, m->tv};
        M = mesh_ptrs(f, m->indices, m->tex_ids);
    } else {
        const float* src[11] = {m->px, m->py, m->pz, m->nx, m->ny, m->nz, m->cr, m->cg, m->cb, m->tu, m->tv};
        const float* dst[11];
        for (int k = 0; k < 11; k++) {
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
        const uint32_t* dt = nullptr;
        if (m->tex_ids) {
            if ((e = c->meshTex.ensure(T)) != hipSuccess) return c->hip_fail(e, "hipMalloc(texture slots)");
            if (T && (e = hipMemcpyAsync(c->meshTex.p, m->tex_ids, 4ull * T, hipMemcpyHostToDevice, c->stream)) !=
                         hipSuccess)
                return c->hip_fail(e, "hipMemcpyAsync(texture slots)");
            dt = c->meshTex.p;
        }
        M = mesh_ptrs(dst, c->meshIdx.p, dt);
    }

    const JobInput in = context_job(c, M, V, T);
    Plan pl;
    if ((r = build_plan(c, std::vector<JobInput>{in}, pl)) != EDXR_OK) return r;
    c->renderGeneration++;
    c->rendered = false;
    c->lastPlanValid = false;
    if (c->floatColor && (e = c->colorF.ensure(3ull * c->W * c->H)) != hipSuccess)
        return c->hip_fail(e, "hipMalloc(float colour)");
    if ((r = run_plan(c, pl, c->color.p, c->depth.p, c->ids.p, c->floatColor ? c->colorF.p : nullptr)) != EDXR_OK)
        return r;
    c->lastPlan = pl;
    c->lastPlanValid = true;
    c->lastFloat = c->floatColor != 0;
    c->rendered = true;
    c->batchRendered = false;
    c->lastT = T;
    c->lastV = V;
    c->lastChunks = pl.Q.total_chunks;
    c->lastMesh = M;
    // RenderMesh's epilogue (Core/Renderer.cpp:114-117).
    if (c->writeFrames && (r = edxr_write_frame_to_file(c)) != EDXR_OK) return r;
    c->frameCount++;
    return EDXR_OK;
}

// The frame of edxr_render_mesh without allocating: job table, setup records,
// bins, work list and key images are carved from the device's static scratch
// (kScratchBytes), the outputs are the caller's.  Eager launches only (no
// graph, no profiling events); a capacity overflow re-runs with larger
// capacities while they fit the scratch.
int edxr_render_mesh_noalloc(edxr_ctx* c, const edxr_mesh* m, void* color, void* depth, void* ids) {
    if (!c || !m) return EDXR_E_INVALID;
    if (!color || !depth || !ids) return c->fail(EDXR_E_INVALID, "colour, depth and ID outputs are required");
    if (!c->haveTransform) return c->fail(EDXR_E_STATE, "edxr_set_transform was not called");
    if (m->memory != EDXR_MEM_DEVICE) return c->fail(EDXR_E_INVALID, "the allocation-free render needs a device mesh");
    if (c->msaaLog2) return c->fail(EDXR_E_STATE, "the allocation-free render is single-sample (EDXR_STATE_MSAA 0)");
    int r = check_mesh(c, m, c->shader);
    if (r != EDXR_OK) return r;
    if (!c->scratch) return c->fail(EDXR_E_HIP, "no scratch address for this device");
    const uint32_t V = m->num_vertices, T = m->num_triangles;
    if ((uint64_t)T + 65536 > kRecordMask) return c->fail(EDXR_E_LIMIT, "too many triangles");

    const float* const f[11] = {m->px, m->py, m->pz, m->nx, m->ny, m->nz, m->cr, m->cg, m->cb, m->tu, m->tv};
    const JobInput in = context_job(c, mesh_ptrs(f, m->indices, m->tex_ids), V, T);
    // The single job of the frame, as build_plan places it (every base 0),
    // with the chunk x tile matrix held to 1/16 of the scratch.
    FrameParams P = job_params(in, c->tilesX, c->tilesY);
    P.num_chunks = choose_chunks(c, T, P.num_tiles, true, kScratchBytes / 16);
    P.chunk_slots = P.num_chunks ? (T + P.num_chunks - 1) / P.num_chunks : 0;
    BatchParams Q;
    memset(&Q, 0, sizeof(Q));
    Q.num_jobs = 1;
    Q.total_verts = V;
    Q.total_slots = T;
    Q.total_tiles = P.num_tiles;
    Q.total_chunks = P.num_chunks;
    Q.vert_blocks = (V + kJobTableBlock - 1) / kJobTableBlock;
    Q.slot_blocks = (T + kJobTableBlock - 1) / kJobTableBlock;
    Q.max_job_tiles = P.num_tiles;
    Q.max_job_chunks = P.num_chunks;
    Q.max_samples = 1;
    Q.any_textured = P.shader == kTexturedLambert ? 1u : 0u;
    const uint64_t countEntries = (uint64_t)P.num_chunks * P.num_tiles;
    const uint64_t groupEntries = (uint64_t)((P.num_chunks + kChunkGroup - 1) / kChunkGroup) * P.num_tiles;

    // Capacities of this call (never the context's, which may be larger).
    uint32_t clipCap = std::max<uint32_t>(1024, std::min<uint32_t>(T, 1u << 16)), recCap = 16384, vertCap = 16384;
    uint32_t refCap = (uint32_t)std::min<uint64_t>(4ull * T + 4ull * P.num_tiles + 1024, 0xFFFFFFF0ull);
    // Few raster items: each holds an 8 KB key image in the scratch.
    const uint32_t rasterBlocks = c->numCUs * 2;
    bool legacyScans = c->legacyScans;
    hipError_t e;
    c->stats = edxr_stats{};
    for (uint32_t attempt = 0; attempt < 8; attempt++) {
        if ((uint64_t)T + recCap > kRecordMask) return c->fail(EDXR_E_LIMIT, "record capacity exceeded");
        const uint64_t workCap = (uint64_t)P.num_tiles + std::max(rasterBlocks, refCap / kSegRefs) + 2;
        ScratchCarver s{c->scratch};
        FrameParams* jobs = s.take<FrameParams>(1);
        DeviceBuffers B;
        memset(&B, 0, sizeof(B));
        B.tile_totals = s.take<uint32_t>(P.num_tiles);
        B.tile_start = s.take<uint32_t>((uint64_t)P.num_tiles + 1);
        B.tile_item = s.take<uint32_t>(P.num_tiles);
        B.tile_big = s.take<uint32_t>(P.num_tiles);
        B.rrec = s.take<RasterRec>((uint64_t)T + recCap);
        B.srec = s.take<ShadeRec>((uint64_t)T + recCap);
        B.desc = s.take<uint32_t>(T);
        B.chunk_box = s.take<uint32_t>(2ull * std::max<uint32_t>(P.num_chunks, 1));
        B.group_sums = s.take<uint32_t>(groupEntries);
        B.group_entries = groupEntries;
        B.scan_spin_limit = kScanSpinLimit;
        B.legacy_scans = legacyScans ? 1u : 0u;
        B.clip_queue = s.take<uint32_t>(clipCap);
        B.bin_counts = s.take<uint32_t>(countEntries);
        B.refs = s.take<uint32_t>(refCap);
        B.work = s.take<uint32_t>(workCap);
        B.partials = s.take<uint64_t>(workCap * kTileSize * kTileSize);
        B.verts = s.take<ShadeVert>((uint64_t)V + vertCap);
        B.sverts = s.take<ScreenVert>(V);
        if (!s.fits()) {
            char msg[160];
            snprintf(msg, sizeof(msg),
                     "the frame needs %.1f MiB of device scratch, more than the %llu MiB the allocation-free "
                     "render has (use edxr_render_mesh)",
                     s.used / 1048576.0, (unsigned long long)(kScratchBytes >> 20));
            return c->fail(EDXR_E_LIMIT, msg);
        }
        B.clip_queue_cap = clipCap;
        B.extra_rec_cap = recCap;
        B.extra_vert_cap = vertCap;
        B.ref_cap = refCap;
        B.work_cap = (uint32_t)workCap;
        B.raster_blocks = rasterBlocks;
        B.raster_grid = c->numCUs * raster_resident_per_cu(false);
        B.num_cus = c->numCUs;
        B.counters = c->counters.p;
        B.host_counters = c->hostCounters;
        B.color = static_cast<uint32_t*>(color);
        B.depth = static_cast<float*>(depth);
        B.ids = static_cast<uint32_t*>(ids);
        B.color_f32 = nullptr;
        // The first kernel writes the job table from its argument: no upload.
        if ((e = launch_frame(Q, jobs, &P, B, c->stream, nullptr)) != hipSuccess) return c->hip_fail(e, "kernel launch");
        if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return c->hip_fail(e, "hipStreamSynchronize");
        const uint32_t* h = c->hostCounters;
        const uint32_t ov = h[kCntOverflow];
        if (ov & kOvPolygon) return c->fail(EDXR_E_LIMIT, "clip polygon exceeded 9 vertices");
        if (!ov) {
            if (h[kCntBadIndex])
                return c->fail(EDXR_E_INDEX, "triangle index >= num_vertices or texture slot >= texture count");
            const CounterSums cs = counter_sums(h);
            c->stats.triangles_in = T;
            c->stats.trivially_rejected = cs.rejected;
            c->stats.clipped = cs.clipped;
            c->stats.near_clipped = cs.nearClipped;
            c->stats.setup_records = cs.mainRecords + cs.extraRecs;
            c->stats.refs = h[kCntTotalRefs];
            c->stats.clip_vertices = cs.extraVerts;
            c->stats.retries = attempt;
            c->stats.bin_chunks = P.num_chunks;
            return EDXR_OK;
        }
        auto grow = [](uint32_t need) {
            return (uint32_t)std::min<uint64_t>((uint64_t)need + need / 4 + 1024, 0xFFFFFFF0ull);
        };
        if (ov & kOvClipQueue) {
            clipCap = grow(h[kCntClipQueue]);
            recCap = std::max(recCap, grow(2 * recCap));
            vertCap = std::max(vertCap, grow(2 * vertCap));
        }
        const CounterSums cs = counter_sums(h);
        if (ov & kOvExtraRecords) recCap = grow(std::max(cs.recDemand, 2 * recCap));
        if (ov & kOvExtraVertices) vertCap = grow(std::max(cs.vertDemand, 2 * vertCap));
        if (ov & kOvRefs) refCap = grow(std::max(h[kCntTotalRefs], refCap));
        if (ov & kOvScanWait) legacyScans = c->legacyScans = true;
    }
    return c->fail(EDXR_E_LIMIT, "capacity retries exhausted");
}

// Texture slots with their mip chains (RGBA floats, b / 255 at level 0; each
// further level halves the sides, floor, at least 1, texel = 0.25 * (((a + b)
// + c) + d) of its 2x2 parent texels, a side of 1 repeating its texel).
int edxr_set_textures(edxr_ctx* c, const edxr_texture* tex, uint32_t n) {
    if (!c || (n && !tex)) return EDXR_E_INVALID;
    std::vector<TexDesc> desc(n);
    std::vector<float4> texels;
    for (uint32_t k = 0; k < n; k++) {
        const edxr_texture& t = tex[k];
        TexDesc& d = desc[k];
        memset(&d, 0, sizeof(d));
        d.color[0] = t.color[0];
        d.color[1] = t.color[1];
        d.color[2] = t.color[2];
        if (!t.rgba) continue;  // ConstantTexture2D
        if (!t.width || !t.height || t.width > 16384 || t.height > 16384)
            return c->fail(EDXR_E_LIMIT, "texture size out of range");
        d.width = t.width;
        d.height = t.height;
        d.offset = (uint32_t)texels.size();
        uint32_t w = t.width, h = t.height;
        const size_t base = texels.size();
        texels.resize(base + (size_t)w * h);
        for (size_t i = 0; i < (size_t)w * h; i++)
            texels[base + i] = make_float4(t.rgba[4 * i] / 255.0f, t.rgba[4 * i + 1] / 255.0f,
                                           t.rgba[4 * i + 2] / 255.0f, t.rgba[4 * i + 3] / 255.0f);
        d.levels = 1;
        size_t prev = base;
        while (w > 1 || h > 1) {
            const uint32_t w2 = std::max(1u, w >> 1), h2 = std::max(1u, h >> 1);
            const size_t at = texels.size();
            texels.resize(at + (size_t)w2 * h2);
            for (uint32_t j = 0; j < h2; j++)
                for (uint32_t i = 0; i < w2; i++) {
                    const uint32_t i0 = 2 * i, i1 = std::min(2 * i + 1, w - 1);
                    const uint32_t j0 = 2 * j, j1 = std::min(2 * j + 1, h - 1);
                    const float4 a = texels[prev + (size_t)j0 * w + i0], b = texels[prev + (size_t)j0 * w + i1];
                    const float4 cc = texels[prev + (size_t)j1 * w + i0], e = texels[prev + (size_t)j1 * w + i1];
                    texels[at + (size_t)j * w2 + i] =
                        make_float4(0.25f * (((a.x + b.x) + cc.x) + e.x), 0.25f * (((a.y + b.y) + cc.y) + e.y),
                                    0.25f * (((a.z + b.z) + cc.z) + e.z), 0.25f * (((a.w + b.w) + cc.w) + e.w));
                }
            prev = at;
            w = w2;
            h = h2;
            d.levels++;
        }
        if (texels.size() > 0xFFFFFFFFull) return c->fail(EDXR_E_LIMIT, "texture memory out of range");
    }
    hipError_t e;
    if ((e = c->texDesc.ensure(std::max<uint32_t>(n, 1))) != hipSuccess ||
        (e = c->texels.ensure(std::max<size_t>(texels.size(), 1))) != hipSuccess)
        return c->hip_fail(e, "hipMalloc(textures)");
    if (n && (e = hipMemcpyAsync(c->texDesc.p, desc.data(), sizeof(TexDesc) * n, hipMemcpyHostToDevice, c->stream)) !=
                 hipSuccess)
        return c->hip_fail(e, "hipMemcpyAsync(texture slots)");
    if (!texels.empty() && (e = hipMemcpyAsync(c->texels.p, texels.data(), sizeof(float4) * texels.size(),
                                               hipMemcpyHostToDevice, c->stream)) != hipSuccess)
        return c->hip_fail(e, "hipMemcpyAsync(texels)");
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return c->hip_fail(e, "set_textures");
    c->numTex = n;
    drop_graph(c);  // texture pointers are part of the captured job table
    return EDXR_OK;
}

int edxr_set_frame_directory(edxr_ctx* c, const char* dir) {
    if (!c || !dir || !dir[0]) return EDXR_E_INVALID;
    c->frameDir = dir;
    return EDXR_OK;
}

int64_t edxr_frame_count(edxr_ctx* c) { return c ? c->frameCount : EDXR_E_INVALID; }

int edxr_write_frame_to_file(edxr_ctx* c) {
    if (!c) return EDXR_E_INVALID;
    if (!c->rendered) return c->fail(EDXR_E_STATE, "nothing rendered");
    const std::string dir = c->frameDir + "/Frames";
    struct stat st;
    if (stat(dir.c_str(), &st) != 0 && mkdir(dir.c_str(), 0755) != 0)
        return c->fail(EDXR_E_IO, "cannot create the Frames directory");
    char name[32];
    snprintf(name, sizeof(name), "/Frame%05lld.bmp", (long long)c->frameCount);
    std::vector<uint8_t> rgba(4ull * c->W * c->H);
    int r = edxr_read_color(c, rgba.data(), rgba.size());
    if (r != EDXR_OK) return r;
    r = edxr_write_bmp((dir + name).c_str(), rgba.data(), c->W, c->H);
    return r == EDXR_OK ? r : c->fail(r, "cannot write the frame bitmap");
}

int edxr_render_batch(edxr_ctx* c, const edxr_job* jobs, uint32_t n) {
    if (!c || !jobs || n == 0) return EDXR_E_INVALID;
    if (n > 65536) return c->fail(EDXR_E_LIMIT, "more than 65536 jobs");
    std::vector<JobInput> in(n);
    int r;
    for (uint32_t k = 0; k < n; k++) {
        const edxr_job& j = jobs[k];
        if (j.shader < EDXR_SHADER_LAMBERT_ALBEDO || j.shader > EDXR_SHADER_TEXTURED_LAMBERT)
            return c->fail(EDXR_E_INVALID, "bad shader kind in job " + std::to_string(k));
        if ((r = check_mesh(c, &j.mesh, j.shader)) != EDXR_OK) {
            c->err += " (job " + std::to_string(k) + ")";
            return r;
        }
        JobInput& d = in[k];
        const edxr_mesh& m = j.mesh;
        const float* const f[11] = {m.px, m.py, m.pz, m.nx, m.ny, m.nz, m.cr, m.cg, m.cb, m.tu, m.tv};
        d.mesh = mesh_ptrs(f, m.indices, m.tex_ids);
        d.tex = c->texDesc.p;
        d.texels = c->texels.p;
        d.numTex = c->numTex;
        d.texFilter = c->texFilter;
        d.V = m.num_vertices;
        d.T = m.num_triangles;
        d.W = j.width;
        d.H = j.height;
        mat_mul(j.proj, j.mv, d.mvp);
        float inv[16];
        mat_inverse(j.mv, inv);
        const float zero[3] = {0.0f, 0.0f, 0.0f};
        transform_point3(inv, zero, d.eye);
        memcpy(d.raster, j.raster, 64);
        memcpy(d.light, j.light, 12);
        d.mat = j.material;
        d.shader = j.shader;
        d.hierarchical = j.hierarchical ? 1 : 0;
        d.sampleLevel = 0;
    }
    if ((r = stage_batch_meshes(c, jobs, n, in)) != EDXR_OK) return r;
    Plan pl;
    if ((r = build_plan(c, in, pl)) != EDXR_OK) return r;
    hipError_t e;
    if ((e = c->bcolor.ensure(pl.pixels)) != hipSuccess || (e = c->bdepth.ensure(pl.pixels)) != hipSuccess ||
        (e = c->bids.ensure(pl.pixels)) != hipSuccess)
        return c->hip_fail(e, "hipMalloc(batch frame buffers)");
    if (c->floatColor && (e = c->bcolorF.ensure(3 * pl.pixels)) != hipSuccess)
        return c->hip_fail(e, "hipMalloc(batch float colour)");
    c->renderGeneration++;
    c->batchRendered = false;
    c->rendered = false;
    c->lastPlanValid = false;
    if ((r = run_plan(c, pl, c->bcolor.p, c->bdepth.p, c->bids.p, c->floatColor ? c->bcolorF.p : nullptr)) !=
        EDXR_OK)
        return r;
    c->lastPlan = pl;
    c->lastPlanValid = true;
    c->lastFloat = c->floatColor != 0;
    c->batchPix.resize(n + 1);
    c->batchW.resize(n);
    c->batchH.resize(n);
    for (uint32_t k = 0; k < n; k++) {
        c->batchPix[k] = pl.jobs[k].pix_base;
        c->batchW[k] = pl.jobs[k].width;
        c->batchH[k] = pl.jobs[k].height;
    }
    c->batchPix[n] = pl.pixels;
    c->batchRendered = true;
    return EDXR_OK;
}

int edxr_batch_read(edxr_ctx* c, uint32_t job, uint8_t* rgba, float* depth, uint32_t* ids) {
    if (!c) return EDXR_E_INVALID;
    if (!c->batchRendered) return c->fail(EDXR_E_STATE, "no batch rendered");
    if (job >= c->batchW.size()) return c->fail(EDXR_E_INVALID, "job out of range");
    const uint64_t o = c->batchPix[job], n = (uint64_t)c->batchW[job] * c->batchH[job];
    hipError_t e = hipSuccess;
    if (rgba) e = hipMemcpyAsync(rgba, c->bcolor.p + o, 4 * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && depth) e = hipMemcpyAsync(depth, c->bdepth.p + o, 4 * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && ids) e = hipMemcpyAsync(ids, c->bids.p + o, 4 * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? EDXR_OK : c->hip_fail(e, "batch_read");
}

int edxr_batch_device_buffers(edxr_ctx* c, void** color, void** depth, void** ids, uint64_t* offsets) {
    if (!c) return EDXR_E_INVALID;
    if (!c->batchRendered) return c->fail(EDXR_E_STATE, "no batch rendered");
    if (color) *color = c->bcolor.p;
    if (depth) *depth = c->bdepth.p;
    if (ids) *ids = c->bids.p;
    if (offsets)
        for (size_t k = 0; k < c->batchPix.size(); k++) offsets[k] = c->batchPix[k];
    return EDXR_OK;
}

int edxr_float_color(edxr_ctx* c, void** rgb) {
    if (!c || !rgb) return EDXR_E_INVALID;
    if (!c->lastPlanValid || !c->lastFloat)
        return c->fail(EDXR_E_STATE, "no render with EDXR_STATE_FLOAT_COLOR on");
    *rgb = c->rendered ? (void*)c->colorF.p : (void*)c->bcolorF.p;
    return EDXR_OK;
}

int edxr_copy_float_color(edxr_ctx* c, void* dst, size_t bytes) {
    if (!c || !dst) return EDXR_E_INVALID;
    void* src = nullptr;
    const int r = edxr_float_color(c, &src);
    if (r != EDXR_OK) return r;
    const uint64_t need = 12ull * c->lastPlan.pixels;
    if (bytes < need) return c->fail(EDXR_E_INVALID, "buffer too small");
    hipError_t e = hipMemcpyAsync(dst, src, need, hipMemcpyDefault, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? EDXR_OK : c->hip_fail(e, "copy_float_color");
}

int edxr_backward(edxr_ctx* c, const float* d_rgb, const edxr_grads* g, uint32_t n) {
    if (!c || !d_rgb || !g) return EDXR_E_INVALID;
    if (!c->lastPlanValid) return c->fail(EDXR_E_STATE, "backward before a render");
    const Plan& pl = c->lastPlan;
    if (n != pl.Q.num_jobs) return c->fail(EDXR_E_INVALID, "one edxr_grads per job of the last render");
    if (pl.Q.max_samples > 1) return c->fail(EDXR_E_STATE, "backward of a multisampled frame is not supported");
    for (const FrameParams& P : pl.jobs)
        if (P.shader == kTexturedLambert)
            return c->fail(EDXR_E_STATE, "backward of the textured shader is not supported");
    for (const FrameParams& P : pl.jobs) {
        const float* R = P.raster;
        if (R[2] != 0.0f || R[6] != 0.0f || R[12] != 0.0f || R[13] != 0.0f || R[14] != 0.0f || R[15] != 1.0f ||
            R[0] * R[5] - R[1] * R[4] == 0.0f)
            return c->fail(EDXR_E_INVALID, "backward needs a raster matrix affine in x and y");
    }
    hipError_t e;
    if ((e = c->gradsDev.ensure(n)) != hipSuccess || (e = c->gradsHost.ensure(n)) != hipSuccess ||
        (e = c->adjAcc.ensure(8ull * n)) != hipSuccess)
        return c->hip_fail(e, "hipMalloc(adjoint)");
    // Jobs whose nine vertex outputs are the same arrays share one double
    // accumulator (as many vertices as the largest of their meshes).
    std::vector<uint32_t> groupOf(n), groupV, groupFirst;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t gi = 0;
        for (; gi < groupFirst.size(); gi++) {
            const edxr_grads& f = g[groupFirst[gi]];
            if (f.d_px == g[k].d_px && f.d_py == g[k].d_py && f.d_pz == g[k].d_pz && f.d_nx == g[k].d_nx &&
                f.d_ny == g[k].d_ny && f.d_nz == g[k].d_nz && f.d_cr == g[k].d_cr && f.d_cg == g[k].d_cg &&
                f.d_cb == g[k].d_cb)
                break;
        }
        if (gi == groupFirst.size()) {
            groupFirst.push_back(k);
            groupV.push_back(0);
        }
        groupOf[k] = gi;
        groupV[gi] = std::max(groupV[gi], pl.jobs[k].num_vertices);
    }
    std::vector<uint64_t> groupBase(groupFirst.size());
    uint64_t vaccSize = 0;
    for (size_t gi = 0; gi < groupFirst.size(); gi++) {
        groupBase[gi] = vaccSize;
        vaccSize += 9ull * groupV[gi];
    }
    if ((e = c->adjVertAcc.ensure(std::max<uint64_t>(vaccSize, 1))) != hipSuccess)
        return c->hip_fail(e, "hipMalloc(adjoint accumulators)");
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t gi = groupOf[k];
        c->gradsHost.p[k] = GradPtrs{g[k].d_px, g[k].d_py, g[k].d_pz, g[k].d_nx, g[k].d_ny, g[k].d_nz,
                                     g[k].d_cr, g[k].d_cg, g[k].d_cb, g[k].d_light, g[k].d_material,
                                     groupBase[gi], groupV[gi], groupFirst[gi] == k ? 1u : 0u};
    }
    if ((e = hipMemcpyAsync(c->gradsDev.p, c->gradsHost.p, sizeof(GradPtrs) * n, hipMemcpyHostToDevice,
                            c->stream)) != hipSuccess)
        return c->hip_fail(e, "hipMemcpyAsync(grads)");
    const uint32_t* ids = c->rendered ? c->ids.p : c->bids.p;
    if ((e = launch_adjoint(pl.Q, c->jobsDev.p, c->gradsDev.p, ids, d_rgb, c->adjAcc.p,
                            c->adjVertAcc.p, vaccSize, c->stream)) != hipSuccess)
        return c->hip_fail(e, "adjoint launch");
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return c->hip_fail(e, "hipStreamSynchronize");
    return EDXR_OK;
}

int edxr_autotune(edxr_ctx* c, const edxr_job* jobs, uint32_t n, edxr_tuning* out) {
    if (!c || !jobs || n == 0) return EDXR_E_INVALID;
    static const uint32_t kTpc[] = {512, 1024, 2048, 4096};
    static const uint32_t kRpc[] = {4, 8, 16};
    const uint32_t saveT = c->trisPerChunk, saveR = c->rasterPerCU;
    double best = 1e30;
    uint32_t bestT = 1024, bestR = 8;
    for (uint32_t tpc : kTpc)
        for (uint32_t rpc : kRpc) {
            c->trisPerChunk = tpc;
            c->rasterPerCU = rpc;
            int r = edxr_render_batch(c, jobs, n);  // warm-up (allocation, capture)
            if (r != EDXR_OK) {
                c->trisPerChunk = saveT;
                c->rasterPerCU = saveR;
                return r;
            }
            // the median of seven frames: candidates often differ by 1-2 %
            double times[7];
            for (double& t : times) {
                const auto t0 = std::chrono::steady_clock::now();
                if ((r = edxr_render_batch(c, jobs, n)) != EDXR_OK) {
                    c->trisPerChunk = saveT;
                    c->rasterPerCU = saveR;
                    return r;
                }
                t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
            std::sort(times, times + 7);
            if (times[3] < best) {
                best = times[3];
                bestT = tpc;
                bestR = rpc;
            }
        }
    c->trisPerChunk = bestT;
    c->rasterPerCU = bestR;
    if (out) {
        out->tris_per_chunk = bestT;
        out->raster_blocks_per_cu = bestR;
        out->frame_ms = best;
    }
    return EDXR_OK;
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

int edxr_read_sample_buffers(edxr_ctx* c, float* depth, uint32_t* ids, size_t count) {
    if (!c) return EDXR_E_INVALID;
    if (!c->rendered) return c->fail(EDXR_E_STATE, "nothing rendered");
    const size_t need = ((size_t)1 << c->msaaLog2) * c->W * c->H;
    if (count < need) return c->fail(EDXR_E_INVALID, "buffer too small");
    hipError_t e = hipSuccess;
    if (depth) e = hipMemcpyAsync(depth, c->depth.p, need * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && ids) e = hipMemcpyAsync(ids, c->ids.p, need * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? EDXR_OK : c->hip_fail(e, "read_sample_buffers");
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

uint64_t edxr_render_generation(const edxr_ctx* c) { return c ? c->renderGeneration : 0; }

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
    std::vector<ScreenVert> sv(c->lastV);
    hipError_t e = hipSuccess;
    if (T) e = hipMemcpy(desc.data(), c->desc.p, 4ull * T, hipMemcpyDeviceToHost);
    if (T && e == hipSuccess) e = hipMemcpy(idx.data(), c->lastMesh.idx, 12ull * T, hipMemcpyDeviceToHost);
    if (c->lastV && e == hipSuccess)
        e = hipMemcpy(sv.data(), c->sverts.p, sizeof(ScreenVert) * c->lastV, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(rr.data(), c->rrec.p, sizeof(RasterRec) * nrec, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(sr.data(), c->srec.p, sizeof(ShadeRec) * nrec, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "debug_read_records");
    uint64_t n = 0;
    for (uint32_t t = 0; t < T; t++) {
        const bool compact = (desc[t] & kDescCompact) != 0;
        const uint32_t first = compact ? t : desc[t] >> 3, cnt = compact ? 1u : desc[t] & 7u;
        for (uint32_t k = 0; k < cnt; k++, n++) {
            if (n >= cap || !out) continue;
            RasterRec r = rr[first + k];
            if (compact) {
                // Not stored on the device: rebuilt from the vertices' screen
                // values as the raster kernel rebuilds it.
                const ScreenVert& a = sv[idx[3 * t]];
                const ScreenVert& b = sv[idx[3 * t + 1]];
                const ScreenVert& d = sv[idx[3 * t + 2]];
                r.x0 = a.x; r.y0 = a.y; r.x1 = b.x; r.y1 = b.y; r.x2 = d.x; r.y2 = d.y;
                const int B0 = r.y0 - r.y1, C0 = r.x1 - r.x0, B1 = r.y1 - r.y2, C1 = r.x2 - r.x1;
                const int B2 = r.y2 - r.y0, C2 = r.x0 - r.x2;
                r.inv_det = 1.0f / (float)(C2 * B1 - C1 * B2);
                r.z0 = a.z; r.z1 = b.z; r.z2 = d.z;
                r.info = corner_info(B0, C0, B1, C1, B2, C2);
                r.order = make_order(t, 0, false);
            }
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
    if (refs3 && total <= cap) {
        for (uint64_t i = 0; i < total; i++) {
            const uint32_t p = refs[i], ri = p & kRecordMask;
            const uint32_t acc = (p >> kRecordBits) & 7u;
            const uint32_t big = p >> 31;
            const bool compact = !big && ri < c->lastT;  // record slot = triangle, not stored
            refs3[3 * i] = compact ? ri : rr[ri].order >> 4;
            refs3[3 * i + 1] = compact ? 0u : (rr[ri].info >> 12) & 7u;
            refs3[3 * i + 2] = acc | ((acc == 7u) ? 8u : 0u) | (big ? 16u : 0u);
        }
        // Submission order within each tile: (triangle, fan piece) ascending;
        // a record is at most once in a tile's list.
        struct Ref3 {
            uint32_t w[3];
        };
        Ref3* r3 = reinterpret_cast<Ref3*>(refs3);
        for (uint32_t t = 0; t < c->nTiles; t++)
            std::sort(r3 + ts[t], r3 + ts[t + 1], [](const Ref3& a, const Ref3& b) {
                return a.w[0] != b.w[0] ? a.w[0] < b.w[0] : a.w[1] < b.w[1];
            });
    }
    return (int64_t)total;
}

}  // extern "C"
