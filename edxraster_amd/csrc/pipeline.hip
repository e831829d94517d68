// pipeline.hip -- gfx950 kernels of the EDXRaster hot path.
//
// Every kernel reads its job's constants from a device table (FrameParams):
// a plain frame is a batch of one job, a batch renders many independent jobs
// with the same launches.  Stages (EDXRaster/Core/Renderer.cpp:100-118 order):
//   k_vertex     one lane per vertex: SoA in, 48-byte shading vertex out
//                (VertexProcessing, Core/Renderer.cpp:120-127); it also clears
//                the counters and writes a single job's table entry (k_reset
//                does that alone for a frame without vertices).
//   k_setup      vertex transform + outcodes + triangle setup, one lane per
//                triangle (DefaultVertexShader::Execute Core/Shader.h:40-49,
//                Clipper::Clip Core/Clipper.h:71-189, RasterTriangle::Setup
//                Core/RasterTriangle.h:27-153).  Straddling triangles are
//                queued for k_clip.
//   k_clip       Sutherland-Hodgman clipping + fan setup of queued triangles
//                (Core/Clipper.h:109-170,192-288).
//   k_bin_count  per-chunk tile histograms (binning, Core/Renderer.cpp:161-230)
//   k_bin_colpass (or k_bin_colsum / k_bin_groupscan / k_bin_colscan) and
//   k_bin_tilescan
//                exclusive scans -> per-(tile, chunk) offsets and the raster
//                work list of (tile, segment) items.
//   k_bin_scatter scatter through LDS cursors: a tile's list holds the chunks'
//                references in chunk order (the reference concatenates its
//                per-core lists in core order, Core/Renderer.cpp:250-253); the
//                depth keys carry the submission order, so the list order
//                within a chunk does not matter.
//   k_raster     persistent: hierarchical edge-function coverage of one (tile,
//                segment) item at a time (CoarseRasterize / FineRasterize /
//                TrivialAcceptTriangle, Core/Rasterizer.h:32-200,310-353) and
//                the depth test as a (depth, submission order) key minimum.
//   k_shade      per tile: minimum key over the tile's items, perspective-correct
//                interpolation + pixel shading of the winner and one write-back
//                of colour, depth and triangle ID (Core/Renderer.cpp:272-350).
//
// Arithmetic follows the reference operation by operation: int32 wrap
// (-fwrapv), truncating 28.4 snap, round-to-nearest int->float, correctly
// rounded division and sqrt, no FMA contraction (-ffp-contract=off).
#include <hip/hip_runtime.h>

#include <mutex>

#include "pipeline.h"

// Scratch of the allocation-free frame (edxr_render_mesh_noalloc), one copy
// per device.  Host code never uses this name as an address: the device
// address comes from hipGetSymbolAddress (scratch_address below).
static __device__ __align__(256) uint8_t scratch[edxr::kScratchBytes];

namespace edxr {

namespace {

constexpr uint32_t kWave = 64;
constexpr uint32_t kBlock = 256;
constexpr uint32_t kSetupSlots = 4;  // triangle slots per k_setup thread
static_assert(kJobAlign % (kBlock * kSetupSlots) == 0, "a vertex/setup block must not span two jobs");
static_assert(kBlock == kJobTableBlock, "block -> job tables are indexed by 256-slot block");

__device__ __forceinline__ float4 xform4(const float* m, float x, float y, float z, float w) {
    float4 r;
    r.x = ((m[0] * x + m[1] * y) + m[2] * z) + m[3] * w;
    r.y = ((m[4] * x + m[5] * y) + m[6] * z) + m[7] * w;
    r.z = ((m[8] * x + m[9] * y) + m[10] * z) + m[11] * w;
    r.w = ((m[12] * x + m[13] * y) + m[14] * z) + m[15] * w;
    return r;
}

// Clipper::ComputeClipCode (Core/Clipper.h:48-68).
__device__ __forceinline__ uint32_t clip_code(const float4& v) {
    uint32_t c = 0;
    if (v.x < -v.w) c |= 1u;
    if (v.x > v.w) c |= 2u;
    if (v.y < -v.w) c |= 4u;
    if (v.y > v.w) c |= 8u;
    if (v.z > v.w) c |= 32u;
    if (v.z < 0.0f) c |= 16u;
    return c;
}

// 28.4 snap by truncation (Core/RasterTriangle.h:35-40); saturating, NaN -> 0.
__device__ __forceinline__ int snap(float a) {
    const float s = a * 16.0f;
    if (!(s == s)) return 0;
    if (s >= 2147483648.0f) return 0x7FFFFFFF;
    if (s <= -2147483648.0f) return (int)0x80000000;
    return (int)s;
}

// The raster matrix applied to a homogeneous-projected point and the 28.4 snap
// (Core/RasterTriangle.h:29-40).  An affine raster matrix gives w == 1, where
// v / 1 == v exactly and the divide is skipped.
__device__ __forceinline__ int2 snap_point(const FrameParams& P, float ax, float ay, float az) {
    const float4 A = xform4(P.raster, ax, ay, az, 1.0f);
    if (A.w == 1.0f) return make_int2(snap(A.x), snap(A.y));
    return make_int2(snap(A.x / A.w), snap(A.y / A.w));
}

// RasterTriangle::Setup from snapped vertices (Core/RasterTriangle.h:42-153);
// false when culled (det <= 0).
__device__ __forceinline__ bool setup_snapped(int x0, int y0, int x1, int y1, int x2, int y2, RasterRec& r) {
    const int B0 = y0 - y1, C0 = x1 - x0, B1 = y1 - y2, C1 = x2 - x1, B2 = y2 - y0, C2 = x0 - x2;
    const int det = C2 * B1 - C1 * B2;
    if (det <= 0) return false;
    r.x0 = x0; r.y0 = y0; r.x1 = x1; r.y1 = y1; r.x2 = x2; r.y2 = y2;
    r.inv_det = 1.0f / (float)det;
    r.info = corner_info(B0, C0, B1, C1, B2, C2);
    return true;
}

// RasterTriangle::Setup on homogeneous-projected points (Core/RasterTriangle.h:27-153).
__device__ __forceinline__ bool setup_record(const FrameParams& P, float ax, float ay, float az, float bx, float by,
                                             float bz, float cx, float cy, float cz, RasterRec& r) {
    const int2 a = snap_point(P, ax, ay, az), b = snap_point(P, bx, by, bz), c = snap_point(P, cx, cy, cz);
    return setup_snapped(a.x, a.y, b.x, b.y, c.x, c.y, r);
}

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }
// Inclusive prefix sum over the wave, in registers: DPP row shifts by 1, 2,
// 4 and 8 scan each row of 16 lanes (lanes shifted in from outside the row
// read 0), then row broadcasts of lane 15 (into rows 1 and 3) and lane 31
// (into rows 2 and 3) carry the rows' totals; six v_add_u32_dpp instead of
// six LDS permutes.  Every lane of the wave must be active (DPP reads the
// neighbours' registers whatever the exec mask).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}
// op over the wave's 64 lanes, the same DPP steps with lanes that have no
// source keeping the identity; the result (lane 63's) is returned to every
// lane.  Every lane must be active.
template <class Op>
__device__ __forceinline__ uint32_t wave_reduce(uint32_t v, uint32_t identity, Op op) {
    const int id = (int)identity;
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x111, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x112, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x114, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x118, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x142, 0xa, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    return wave_reduce(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
    return wave_reduce(v, 0xFFFFFFFFu, [](uint32_t a, uint32_t b) { return a < b ? a : b; });
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    return wave_reduce(v, 0u, [](uint32_t a, uint32_t b) { return a > b ? a : b; });
}


__device__ __forceinline__ uint32_t wave_count(bool p) { return (uint32_t)__popcll(__ballot(p)); }

}  // namespace

// ---------------------------------------------------------------------------
// Counters
// ---------------------------------------------------------------------------
__global__ void k_reset(uint32_t* counters, FrameParams* jobs, FrameParams job0, int writeJob) {
    for (uint32_t i = threadIdx.x; i < kCntSlots; i += blockDim.x) counters[i] = 0;
    if (writeJob && threadIdx.x == 0) jobs[0] = job0;
}

// The frame's last launch: the counters to the host's pinned buffer, which
// the host reads once the stream has drained (no copy queued behind the
// frame, whose start would wait for the frame's end and then for its own
// launch).
__global__ void k_publish(const uint32_t* counters, uint32_t* host) {
    for (uint32_t i = threadIdx.x; i < kCntSlots; i += blockDim.x) host[i] = counters[i];
    __threadfence_system();
}


// Sample offsets in 28.4 units per MSAA level, (x, y) per sample: the table of
// FrameBuffer::MultiSampleOffsets (Core/FrameBuffer.cpp:107-191).
__constant__ int2 c_sample_offsets[kMaxSampleLevel + 1][32] = {
    {{0, 0}},
    {{4, 4}, {-4, -4}},
    {{-2, -6}, {6, -2}, {-6, 2}, {2, 6}},
    {{1, -3}, {-1, 3}, {5, 1}, {-3, -5}, {-5, 5}, {-7, -1}, {3, 7}, {7, -7}},
    {{1, 1}, {-1, -3}, {-3, 2}, {4, -1}, {-5, -2}, {2, 5}, {5, 3}, {3, -5}, {-2, 6}, {0, -7}, {-4, -6}, {-6, 4},
     {-8, 0}, {7, -4}, {6, 7}, {-7, -8}},
    {{1, 1}, {-1, -3}, {-3, 2}, {4, -1}, {-5, -2}, {2, 5}, {5, 3}, {3, -5}, {-2, 6}, {0, -7}, {-4, -6}, {-6, 4},
     {-8, 0}, {7, -4}, {6, 7}, {-7, -8}, {1, 3}, {-3, -3}, {-3, 0}, {6, -2}, {-7, -1}, {3, 4}, {7, 3}, {3, -6},
     {-2, 7}, {0, -4}, {-2, -5}, {-7, 6}, {-8, 3}, {4, -1}, {2, 7}, {4, -8}}};

namespace {
// Job of a block-uniform index i (the block -> job tables are null for a
// single job).  The result is wave-uniform, so job fields load as scalars.
__device__ __forceinline__ uint32_t job_of(const BatchParams& Q, const uint32_t* tbl, uint32_t i) {
    return Q.num_jobs == 1 ? 0u : __builtin_amdgcn_readfirstlane(tbl[i]);
}
}  // namespace

// ---------------------------------------------------------------------------
// Vertex processing (Core/Renderer.cpp:120-127): one lane per vertex, SoA
// input read coalesced, one 48-byte shading record out.  invW = 1/w of
// DefaultVertexShader's MVP transform (Core/Shader.h:45) as the z-divide loop
// computes it (Core/Renderer.cpp:144).
// It also does k_reset's work, one launch fewer per frame: block 0 clears
// the counters, and with kArgJob (a single job passed by value) every block
// reads the job from the argument and block 0 writes it to the job table for
// the later kernels.
// ---------------------------------------------------------------------------
template <bool kArgJob>
__global__ __launch_bounds__(kBlock) void k_vertex(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                   ShadeVert* __restrict__ verts, ScreenVert* __restrict__ sverts,
                                                   uint32_t* counters, FrameParams* jobTable, FrameParams job0) {
    if (blockIdx.x == 0) {
        for (uint32_t i = threadIdx.x; i < kCntSlots; i += kBlock) counters[i] = 0;
        if (kArgJob && threadIdx.x == 0) jobTable[0] = job0;
    }
    const FrameParams& P = kArgJob ? job0 : jobs[job_of(Q, Q.vblock_job, blockIdx.x)];
    const MeshPtrs& M = P.mesh;
    const uint32_t slot = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t v = slot - P.vert_base;
    if (v >= P.num_vertices) return;
    const float px = M.px[v], py = M.py[v], pz = M.pz[v];
    // DefaultVertexShader::Execute: projectedPos = MVP * (p, 1); the outcode,
    // the homogeneous projection and the snapped raster position are per
    // vertex, so every unclipped triangle's setup reuses them.
    const float4 c = xform4(P.mvp, px, py, pz, 1.0f);
    const uint32_t code = clip_code(c);
    // Shading vertices are written for vertices inside the frustum only: an
    // unclipped triangle has no other, and a clipped triangle's pieces use
    // vertices made by k_clip (which reads its inputs from the mesh) or
    // original vertices that lie inside every plane.
    if (code == 0) {
        ShadeVert o;
        o.px = px;
        o.py = py;
        o.pz = pz;
        o.nx = M.nx[v];
        o.ny = M.ny[v];
        o.nz = M.nz[v];
        o.cr = M.cr ? M.cr[v] : 0.0f;
        o.cg = M.cg ? M.cg[v] : 0.0f;
        o.cb = M.cb ? M.cb[v] : 0.0f;
        o.inv_w = 1.0f / c.w;
        o.u = M.tu ? M.tu[v] : 0.0f;
        o.v = M.tv ? M.tv[v] : 0.0f;
        verts[slot] = o;
    }
    ScreenVert sv;
    const int2 xy = snap_point(P, c.x / c.w, c.y / c.w, c.z / c.w);
    sv.x = xy.x;
    sv.y = xy.y;
    sv.z = c.z * (1.0f / c.w);
    sv.code = code;
    sverts[slot] = sv;
}

namespace {
// Descriptor of an unclipped triangle's record in slot t.  When its clamped
// tile bbox spans at most 2x2 tiles, binning lists it in every bbox tile with
// no corner test (Core/Renderer.cpp:168-180), so the bbox alone (packed here,
// in the job's tile coordinates) enumerates its references; otherwise
// (t << 3) | 1.
__device__ __forceinline__ uint32_t compact_desc(const FrameParams& P, const RasterRec& r, uint32_t t) {
    const int Shift = kTileLog2 + 4;
    const int minTX = max(0, min(r.x0, min(r.x1, r.x2)) >> Shift);
    const int maxTX = min((int)P.tiles_x - 1, max(r.x0, max(r.x1, r.x2)) >> Shift);
    const int minTY = max(0, min(r.y0, min(r.y1, r.y2)) >> Shift);
    const int maxTY = min((int)P.tiles_y - 1, max(r.y0, max(r.y1, r.y2)) >> Shift);
    const int dx = maxTX - minTX, dy = maxTY - minTY;
    if (dx < 0 || dy < 0 || dx > 1 || dy > 1) return (t << 3) | 1u;
    // a pixel bounding box of kHeavyPixels or more: a whole wave's work in the raster
    const int bw = (max(r.x0, max(r.x1, r.x2)) - min(r.x0, min(r.x1, r.x2))) >> 4;
    const int bh = (max(r.y0, max(r.y1, r.y2)) - min(r.y0, min(r.y1, r.y2))) >> 4;
    const uint32_t heavy = bw * bh >= (int)kHeavyPixels ? kDescHeavy : 0u;
    return kDescCompact | heavy | (uint32_t)minTX | ((uint32_t)minTY << 14) | ((uint32_t)dx << 28) |
           ((uint32_t)dy << 29);
}
}  // namespace

// ---------------------------------------------------------------------------
// Setup: one lane per input triangle.
// ---------------------------------------------------------------------------
// Each thread takes kSetupSlots slots kBlock apart; all their index and
// screen-vertex loads are issued before any is used, so one wave keeps
// several dependent-load chains in flight.
// Held to 64 VGPRs: eight waves per SIMD, so that a 2M-triangle frame's
// 8,192 waves are resident at once (one round of the load chain).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_setup(
    BatchParams Q, const FrameParams* __restrict__ jobs, DeviceBuffers B) {
    const FrameParams& P = jobs[job_of(Q, Q.sblock_job, blockIdx.x * kSetupSlots)];
    const MeshPtrs& M = P.mesh;
    // k_bin_count adds the big references of every tile to zeroed counts;
    // k_bin_colpass's look-back starts from cleared status words.
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < Q.total_tiles; t += gridDim.x * kBlock)
        B.tile_big[t] = 0u;
    if (!B.legacy_scans)
        for (uint64_t i = blockIdx.x * kBlock + threadIdx.x; i < B.group_entries; i += gridDim.x * kBlock)
            B.group_sums[i] = 0u;
    uint32_t s[kSetupSlots], i0[kSetupSlots], i1[kSetupSlots], i2[kSetupSlots];
    bool valid[kSetupSlots];
#pragma unroll
    for (uint32_t u = 0; u < kSetupSlots; u++) {
        s[u] = (blockIdx.x * kSetupSlots + u) * kBlock + threadIdx.x;  // triangle slot
        const uint32_t t = s[u] - P.slot_base;                        // input triangle of the job
        valid[u] = t < P.num_triangles;
        i0[u] = i1[u] = i2[u] = 0;
        if (valid[u]) {
            i0[u] = M.idx[3 * (size_t)t];
            i1[u] = M.idx[3 * (size_t)t + 1];
            i2[u] = M.idx[3 * (size_t)t + 2];
        }
    }
    // The vertices' snapped positions and outcodes; z is loaded only for the
    // few triangles whose record is stored (a compact record is rebuilt from
    // the screen vertices by its readers).
    int2 xy[kSetupSlots][3];
    uint32_t code[kSetupSlots][3];
    bool bad[kSetupSlots];
#pragma unroll
    for (uint32_t u = 0; u < kSetupSlots; u++) {
        bad[u] = valid[u] && (i0[u] >= P.num_vertices || i1[u] >= P.num_vertices || i2[u] >= P.num_vertices);
        if (P.shader == kTexturedLambert && valid[u] && M.tex[s[u] - P.slot_base] >= P.num_tex) bad[u] = true;
        if (valid[u] && !bad[u]) {
            const ScreenVert* v0 = &B.sverts[P.vert_base + i0[u]];
            const ScreenVert* v1 = &B.sverts[P.vert_base + i1[u]];
            const ScreenVert* v2 = &B.sverts[P.vert_base + i2[u]];
            xy[u][0] = *reinterpret_cast<const int2*>(v0);
            xy[u][1] = *reinterpret_cast<const int2*>(v1);
            xy[u][2] = *reinterpret_cast<const int2*>(v2);
            code[u][0] = v0->code;
            code[u][1] = v1->code;
            code[u][2] = v2->code;
        }
    }
    uint32_t nrej = 0, nbad = 0, nmain = 0;
    bool queuedU[kSetupSlots];
    uint64_t qmask[kSetupSlots];
#pragma unroll
    for (uint32_t u = 0; u < kSetupSlots; u++) {
        bool rejected = false, queued = false, main = false;
        RasterRec rec;
        if (valid[u] && !bad[u]) {
            if (code[u][0] | code[u][1] | code[u][2]) {
                if (code[u][0] & code[u][1] & code[u][2]) rejected = true;
                else queued = true;
            } else {
                // the cull test of RasterTriangle::Setup (setup_snapped)
                const int2 a = xy[u][0], b = xy[u][1], c = xy[u][2];
                main = (a.x - c.x) * (b.y - c.y) - (c.x - b.x) * (c.y - a.y) > 0;
                rec.x0 = a.x; rec.y0 = a.y; rec.x1 = b.x; rec.y1 = b.y; rec.x2 = c.x; rec.y2 = c.y;
            }
        }
        queuedU[u] = queued;
        qmask[u] = __ballot(queued);
        if (valid[u]) {
            uint32_t d = 0;
            if (main) {
                // A compact triangle's record is not stored: it is the three
                // vertices' screen values, which the raster re-reads.
                d = compact_desc(P, rec, s[u]);
                if (!(d & kDescCompact)) {
                    setup_snapped(rec.x0, rec.y0, rec.x1, rec.y1, rec.x2, rec.y2, rec);
                    // invW and the z divide (Core/Renderer.cpp:139-147).
                    rec.z0 = B.sverts[P.vert_base + i0[u]].z;
                    rec.z1 = B.sverts[P.vert_base + i1[u]].z;
                    rec.z2 = B.sverts[P.vert_base + i2[u]].z;
                    rec.order = make_order(s[u], 0, false);
                    B.rrec[s[u]] = rec;
                }
            }
            B.desc[s[u]] = d;
        }
        nrej += wave_count(rejected);
        nbad += wave_count(bad[u]);
        nmain += wave_count(main);
    }
    // Per-workgroup totals, one striped atomic per counter per workgroup, and
    // one clip-queue allocation per workgroup (same-address atomics from every
    // wave would serialise in L2).
    __shared__ uint32_t wgCount[3];
    __shared__ uint32_t sQueued[kBlock / kWave], sQueueBase[kBlock / kWave];
    const uint32_t wave = threadIdx.x / kWave;
    uint32_t nq = 0;
#pragma unroll
    for (uint32_t u = 0; u < kSetupSlots; u++) nq += (uint32_t)__popcll(qmask[u]);
    if (__lane_id() == 0) sQueued[wave] = nq;
    if (threadIdx.x < 3) wgCount[threadIdx.x] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tq = 0;
        for (uint32_t w = 0; w < kBlock / kWave; w++) {
            sQueueBase[w] = tq;
            tq += sQueued[w];
        }
        const uint32_t qb = tq ? atomicAdd(&B.counters[kCntClipQueue], tq) : 0u;
        for (uint32_t w = 0; w < kBlock / kWave; w++) sQueueBase[w] += qb;
    }
    if (__lane_id() == 0) {
        if (nrej) atomicAdd(&wgCount[0], nrej);
        if (nbad) atomicAdd(&wgCount[1], nbad);
        if (nmain) atomicAdd(&wgCount[2], nmain);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t stripe = blockIdx.x & (kStripes - 1);
        if (wgCount[0]) atomicAdd(&B.counters[kCntStriped + kStRejected * kStripes + stripe], wgCount[0]);
        if (wgCount[1]) atomicAdd(&B.counters[kCntBadIndex], wgCount[1]);
        if (wgCount[2]) atomicAdd(&B.counters[kCntStriped + kStMainRecords * kStripes + stripe], wgCount[2]);
    }
    if (nq) {
        uint32_t slot = sQueueBase[wave];
#pragma unroll
        for (uint32_t u = 0; u < kSetupSlots; u++) {
            if (queuedU[u]) {
                const uint32_t at = slot + (uint32_t)__popcll(qmask[u] & lanemask_lt());
                if (at < B.clip_queue_cap) B.clip_queue[at] = s[u];
                else atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvClipQueue);
            }
            slot += (uint32_t)__popcll(qmask[u]);
        }
    }
}

// ---------------------------------------------------------------------------
// Clipping of queued straddlers (rare path).
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ bool plane_inside(const float4& v, int p) {
    switch (p) {
        case 0: return v.x >= -v.w;
        case 1: return v.x <= v.w;
        case 2: return v.y >= -v.w;
        case 3: return v.y <= v.w;
        case 4: return v.z <= v.w;
        default: return v.z >= 0.0f;
    }
}
__device__ __forceinline__ float plane_t(const float4& a, const float4& b, int p) {
    switch (p) {
        case 0: return (a.w + a.x) / ((a.x + a.w) - (b.x + b.w));
        case 1: return (-a.w + a.x) / ((a.x - a.w) - (b.x - b.w));
        case 2: return (a.w + a.y) / ((a.y + a.w) - (b.y + b.w));
        case 3: return (-a.w + a.y) / ((a.y - a.w) - (b.y - b.w));
        case 4: return (-a.w + a.z) / ((a.z - a.w) - (b.z - b.w));
        default: return a.z / (a.z - b.z);
    }
}
__device__ __forceinline__ void plane_snap(float4& v, int p) {
    switch (p) {
        case 0: v.x = -v.w; break;
        case 1: v.x = v.w; break;
        case 2: v.y = -v.w; break;
        case 3: v.y = v.w; break;
        case 4: v.z = v.w; break;
        default: v.z = 0.0f; break;
    }
}

// Plane order of ClipPolygon (Core/Clipper.h:238-278): L R B T F N.
__device__ const uint32_t kPlaneBits[6] = {1u, 2u, 4u, 8u, 32u, 16u};

}  // namespace

// One wave per queued triangle, one polygon vertex per lane (kMaxPoly <= 64),
// 16 triangles per workgroup pass.  Clipper::ClipByPlane (Core/Clipper.h:
// 192-232) per plane: lane i takes the edge (i, i + 1) and emits 0, 1 or 2
// vertices, and a wave prefix sum places them in the reference's output
// order; the arithmetic per vertex is the serial algorithm's.  The passes'
// vertex and record allocations are summed in LDS and taken with one atomic
// each per workgroup pass (one per triangle would serialise in L2).
namespace {
__device__ __forceinline__ float4 shfl4(const float4& v, int src) {
    return make_float4(__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src), __shfl(v.w, src));
}
constexpr uint32_t kClipWaves = 16;  // waves (triangles) per workgroup pass
// Input vertex v of a job's mesh with its shading attributes (the caller sets inv_w).
__device__ __forceinline__ ShadeVert input_vertex(const MeshPtrs& M, uint32_t v) {
    ShadeVert o;
    o.px = M.px[v];
    o.py = M.py[v];
    o.pz = M.pz[v];
    o.nx = M.nx[v];
    o.ny = M.ny[v];
    o.nz = M.nz[v];
    o.cr = M.cr ? M.cr[v] : 0.0f;
    o.cg = M.cg ? M.cg[v] : 0.0f;
    o.cb = M.cb ? M.cb[v] : 0.0f;
    o.inv_w = 0.0f;
    o.u = M.tu ? M.tu[v] : 0.0f;
    o.v = M.tv ? M.tv[v] : 0.0f;
    return o;
}
}  // namespace

__global__ __launch_bounds__(kClipWaves * 64) void k_clip(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                           DeviceBuffers B) {
    __shared__ uint32_t sNew[kClipWaves], sRec[kClipWaves], sVBase[kClipWaves], sRBase[kClipWaves];
    __shared__ uint32_t sStat[2];
    const int lane = (int)__lane_id();
    const uint32_t wave = threadIdx.x / 64;
    const uint32_t n = min(B.counters[kCntClipQueue], B.clip_queue_cap);
    uint32_t clipped = 0, nearClipped = 0;  // this wave's statistics (lane 0)
    // The workgroup's allocation stripe and the stripes' regions.
    const uint32_t stripe = blockIdx.x % kAllocStripes;
    const uint32_t vRegion = B.extra_vert_cap / kAllocStripes, rRegion = B.extra_rec_cap / kAllocStripes;
    for (uint32_t base = blockIdx.x * kClipWaves; base < n; base += gridDim.x * kClipWaves) {
        const uint32_t q = base + wave;
        const bool live = q < n;
        // ---- per wave: clip its triangle (no global writes yet) ----
        uint32_t t = 0, vi[3] = {0, 0, 0};
        uint32_t outside = 0;  // bit k: input vertex k has a nonzero outcode
        ShadeVert pa, pb, pc;
        float4 pos = make_float4(0.0f, 0.0f, 0.0f, 1.0f), wgt = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        int cnt = 0;
        bool ok = true;
        const FrameParams* Pp = jobs;
        if (live) {
            t = __builtin_amdgcn_readfirstlane(B.clip_queue[q]);  // triangle slot
            Pp = &jobs[Q.num_jobs == 1 ? 0u : __builtin_amdgcn_readfirstlane(Q.sblock_job[t / kBlock])];
            const FrameParams& P = *Pp;
            const uint32_t* idx = P.mesh.idx + 3 * (size_t)(t - P.slot_base);
            vi[0] = P.vert_base + __builtin_amdgcn_readfirstlane(idx[0]);
            vi[1] = P.vert_base + __builtin_amdgcn_readfirstlane(idx[1]);
            vi[2] = P.vert_base + __builtin_amdgcn_readfirstlane(idx[2]);
            // from the mesh: k_vertex stores shading vertices inside the frustum only
            pa = input_vertex(P.mesh, vi[0] - P.vert_base);
            pb = input_vertex(P.mesh, vi[1] - P.vert_base);
            pc = input_vertex(P.mesh, vi[2] - P.vert_base);
            const float4 c0 = xform4(P.mvp, pa.px, pa.py, pa.pz, 1.0f);
            const float4 c1 = xform4(P.mvp, pb.px, pb.py, pb.pz, 1.0f);
            const float4 c2 = xform4(P.mvp, pc.px, pc.py, pc.pz, 1.0f);
            const uint32_t k0 = clip_code(c0), k1 = clip_code(c1), k2 = clip_code(c2);
            pa.inv_w = 1.0f / c0.w;
            pb.inv_w = 1.0f / c1.w;
            pc.inv_w = 1.0f / c2.w;
            outside = (k0 ? 1u : 0u) | (k1 ? 2u : 0u) | (k2 ? 4u : 0u);
            const uint32_t planeCode = (k0 ^ k1) | (k1 ^ k2) | (k2 ^ k0);
            clipped++;
            if ((k0 | k1 | k2) & 16u) nearClipped++;
            // The polygon: vertex `lane` of cnt (clip-space position, clip weights).
            pos = lane == 0 ? c0 : lane == 1 ? c1 : c2;
            wgt = make_float4(lane == 0 ? 1.0f : 0.0f, lane == 1 ? 1.0f : 0.0f, lane == 2 ? 1.0f : 0.0f, 0.0f);
            cnt = 3;
            for (int p = 0; p < 6; p++) {
                if (!(planeCode & kPlaneBits[p])) continue;
                const bool act = lane < cnt;
                const int i1 = lane + 1 >= cnt ? 0 : lane + 1;
                const float4 v0 = pos, w0 = wgt, v1 = shfl4(pos, i1), w1 = shfl4(wgt, i1);
                const bool in0 = plane_inside(v0, p), in1 = plane_inside(v1, p);
                const int emit = !act ? 0 : (in0 && in1) ? 1 : (in0 != in1) ? (in0 ? 1 : 2) : 0;
                // Intersection of the edge with the plane.
                float4 ip = v1, iw = w1;
                if (act && in0 != in1) {
                    const float tt = plane_t(v0, v1, p);
                    const float s = 1 - tt;
                    ip = make_float4(v0.x * s + v1.x * tt, v0.y * s + v1.y * tt, v0.z * s + v1.z * tt,
                                     v0.w * s + v1.w * tt);
                    plane_snap(ip, p);
                    iw = make_float4(w0.x * s + w1.x * tt, w0.y * s + w1.y * tt, w0.z * s + w1.z * tt, 0.0f);
                }
                // Inclusive prefix of the emitted counts over the lanes.
                const int incl = (int)wave_incl_scan((uint32_t)emit);
                const int total = __builtin_amdgcn_readlane(incl, 63);
                if (total > kMaxPoly) {
                    ok = false;
                    break;
                }
                // Output vertex `lane` comes from the edge lane whose range holds
                // it: its first item (the intersection, or v1 when both ends are
                // inside) or its second (v1 after an entering intersection).
                const int excl = incl - emit;
                int src = 0;
                for (int l = 1; l < cnt; l++) {
                    // both shuffles run with every lane active (a lane switched
                    // off by a condition reads as 0 to the others)
                    const int el = __shfl(excl, l), ml = __shfl(emit, l);
                    if (el <= lane && ml > 0) src = l;
                }
                const int k = lane - __shfl(excl, src);
                const float4 a = shfl4(ip, src), aw = shfl4(iw, src), b = shfl4(v1, src), bw = shfl4(w1, src);
                if (lane < total) {
                    pos = k == 0 ? a : b;
                    wgt = k == 0 ? aw : bw;
                }
                cnt = total;
            }
            if (!ok || cnt > 9) {
                if (lane == 0) atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvPolygon);
                cnt = 0;
            }
            if (__ballot(lane < cnt && pos.w <= 0.0f) != 0) cnt = 0;
            if (cnt < 3) cnt = 0;
        }
        // Vertex ids: weight == 1 reuses a parent vertex (Core/Clipper.h:121-153);
        // new vertices are numbered in polygon order.
        const int which = (wgt.x == 1.0f) ? 0 : (wgt.y == 1.0f) ? 1 : (wgt.z == 1.0f) ? 2 : -1;
        const uint64_t newMask = __ballot(lane < cnt && which < 0);
        // Projected vertices (x/w, y/w, z/w, z * (1/w)) and the fans
        // (Core/Clipper.h:155-170): fan k = lane + 2 uses vertices (0, k - 1, k).
        const float4 h = make_float4(pos.x / pos.w, pos.y / pos.w, pos.z / pos.w, pos.z * (1.0f / pos.w));
        const int fk = lane + 2;
        const float4 h0 = shfl4(h, 0), h1 = shfl4(h, fk - 1 < 64 ? fk - 1 : 0), h2 = shfl4(h, fk < 64 ? fk : 0);
        RasterRec r;
        const bool fan = fk < cnt && setup_record(*Pp, h0.x, h0.y, h0.z, h1.x, h1.y, h1.z, h2.x, h2.y, h2.z, r);
        const uint64_t fanMask = __ballot(fan);
        const uint32_t nNew = (uint32_t)__popcll(newMask);
        const uint32_t nrec = (uint32_t)__popcll(fanMask);
        // ---- per workgroup: one allocation of vertices and records ----
        if (lane == 0) {
            sNew[wave] = nNew;
            sRec[wave] = nrec;
        }
        if (threadIdx.x == 0) sStat[0] = sStat[1] = 0;
        __syncthreads();
        if (lane == 0 && (clipped | nearClipped)) {
            atomicAdd(&sStat[0], clipped);
            atomicAdd(&sStat[1], nearClipped);
            clipped = nearClipped = 0;
        }
        if (threadIdx.x == 0) {
            uint32_t tv = 0, tr = 0;
            for (uint32_t w = 0; w < kClipWaves; w++) {
                sVBase[w] = tv;
                sRBase[w] = tr;
                tv += sNew[w];
                tr += sRec[w];
            }
            // One allocation per workgroup pass, from its stripe's region of
            // the extra vertices and records: every live workgroup reaches
            // this point at about the same time, and atomics on one address
            // would serialise.
            const uint32_t vb = tv ? atomicAdd(&B.counters[kCntAllocStriped + stripe], tv) : 0u;
            const uint32_t rb = tr ? atomicAdd(&B.counters[kCntAllocStriped + kAllocStripes + stripe], tr) : 0u;
            for (uint32_t w = 0; w < kClipWaves; w++) {
                sVBase[w] += stripe * vRegion + vb;
                sRBase[w] += stripe * rRegion + rb;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0 && (sStat[0] | sStat[1])) {
            const uint32_t st = blockIdx.x & (kStripes - 1);
            atomicAdd(&B.counters[kCntStriped + kStClipped * kStripes + st], sStat[0]);
            if (sStat[1]) atomicAdd(&B.counters[kCntStriped + kStNearClipped * kStripes + st], sStat[1]);
        }
        // ---- per wave: write the new vertices, the records and the descriptor ----
        if (live) {
            const FrameParams& P = *Pp;
            const uint32_t vbase = sVBase[wave], rbase = sRBase[wave];
            // As the serial clipper: the polygon's new vertices are created
            // whenever it survives, its records when it has fans.
            const bool vfit = vbase + nNew <= (stripe + 1) * vRegion;
            if (!vfit && lane == 0) atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvExtraVertices);
            bool keep = vfit && nrec > 0;
            if (keep && rbase + nrec > (stripe + 1) * rRegion) {
                if (lane == 0) atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvExtraRecords);
                keep = false;
            }
            uint32_t vid = which == 0 ? vi[0] : which == 1 ? vi[1] : vi[2];
            if (vfit && lane < cnt && which < 0) {
                vid = Q.total_verts + vbase + (uint32_t)__popcll(newMask & lanemask_lt());
                auto comb = [&](float fa, float fb, float fc) { return (wgt.x * fa + wgt.y * fb) + wgt.z * fc; };
                ShadeVert o;
                o.px = comb(pa.px, pb.px, pc.px);
                o.py = comb(pa.py, pb.py, pc.py);
                o.pz = comb(pa.pz, pb.pz, pc.pz);
                o.nx = comb(pa.nx, pb.nx, pc.nx);
                o.ny = comb(pa.ny, pb.ny, pc.ny);
                o.nz = comb(pa.nz, pb.nz, pc.nz);
                o.cr = comb(pa.cr, pb.cr, pc.cr);
                o.cg = comb(pa.cg, pb.cg, pc.cg);
                o.cb = comb(pa.cb, pb.cb, pc.cb);
                o.inv_w = 1.0f / pos.w;
                o.u = comb(pa.u, pb.u, pc.u);
                o.v = comb(pa.v, pb.v, pc.v);
                B.verts[vid] = o;
            } else if (lane < cnt && which >= 0 && ((outside >> which) & 1u)) {
                // A parent vertex with a nonzero outcode that the polygon kept
                // (an intersection weight that rounds to 1): k_vertex did not
                // store it.  Every writer stores the same value.
                // (three stores, not a selected struct, so that the three
                // vertices are not spilled for a dynamic index)
                if (which == 0) B.verts[vid] = pa;
                else if (which == 1) B.verts[vid] = pb;
                else B.verts[vid] = pc;
            }
            const uint32_t vid0 = __shfl(vid, 0), vid1 = __shfl(vid, fk - 1 < 64 ? fk - 1 : 0),
                           vid2 = __shfl(vid, fk < 64 ? fk : 0);
            const uint32_t first = Q.total_slots + rbase;
            if (keep && fan) {
                const uint32_t sidx = (uint32_t)__popcll(fanMask & lanemask_lt());
                r.z0 = h0.w;
                r.z1 = h1.w;
                r.z2 = h2.w;
                r.info |= (uint32_t)(fk - 2) << 12;
                r.order = make_order(t, sidx, true);
                B.rrec[first + sidx] = r;
                B.srec[first + sidx] = ShadeRec{vid0, vid1, vid2, t};
            }
            if (lane == 0) B.desc[t] = keep ? (first << 3) | nrec : 0u;
        }
        __syncthreads();  // the LDS tables are rewritten by the next pass
    }
}

// ---------------------------------------------------------------------------
// Binning (Core/Renderer.cpp:161-230): tile sets of a record.
// ---------------------------------------------------------------------------
namespace {

struct BinTri {
    int x0, y0, x1, y1, x2, y2;
    int B0, C0, B1, C1, B2, C2;
    int sb0, sb1, sb2;   // scalar top-left bias (Core/RasterTriangle.h:155-158)
    uint32_t info;
    int minTX, minTY, w, h;
    bool small;
};

__device__ __forceinline__ int scalar_bias(int ax, int ay, int bx, int by) {
    return ((by > ay) || (ay == by && ax > bx)) ? 0 : -1;
}

__device__ __forceinline__ void load_bin_tri(const FrameParams& P, const RasterRec& r, BinTri& b) {
    b.x0 = r.x0; b.y0 = r.y0; b.x1 = r.x1; b.y1 = r.y1; b.x2 = r.x2; b.y2 = r.y2;
    b.B0 = b.y0 - b.y1; b.C0 = b.x1 - b.x0;
    b.B1 = b.y1 - b.y2; b.C1 = b.x2 - b.x1;
    b.B2 = b.y2 - b.y0; b.C2 = b.x0 - b.x2;
    b.sb0 = scalar_bias(b.x0, b.y0, b.x1, b.y1);
    b.sb1 = scalar_bias(b.x1, b.y1, b.x2, b.y2);
    b.sb2 = scalar_bias(b.x2, b.y2, b.x0, b.y0);
    b.info = r.info;
    const int Shift = kTileLog2 + 4;
    const int minTX = max(0, min(b.x0, min(b.x1, b.x2)) >> Shift);
    const int maxTX = min((int)P.tiles_x - 1, max(b.x0, max(b.x1, b.x2)) >> Shift);
    const int minTY = max(0, min(b.y0, min(b.y1, b.y2)) >> Shift);
    const int maxTY = min((int)P.tiles_y - 1, max(b.y0, max(b.y1, b.y2)) >> Shift);
    b.minTX = minTX;
    b.minTY = minTY;
    b.w = max(0, maxTX - minTX + 1);
    b.h = max(0, maxTY - minTY + 1);
    b.small = (maxTX - minTX < 2 && maxTY - minTY < 2);
    if (b.w == 0 || b.h == 0) b.w = b.h = 0;
}

// Tile (tx, ty) test of a big triangle.  Returns -1 for reject, otherwise the
// accept-edge flags (bits 0..2).
__device__ __forceinline__ int bin_test(const BinTri& b, int tx, int ty) {
    const int Shift = kTileLog2 + 4;
    auto cx = [&](uint32_t c) { return (tx + (int)(c & 1u)) << Shift; };
    auto cy = [&](uint32_t c) { return (ty + (int)(c >> 1)) << Shift; };
    const uint32_t r0 = b.info & 3u, r1 = (b.info >> 2) & 3u, r2 = (b.info >> 4) & 3u;
    const uint32_t a0 = (b.info >> 6) & 3u, a1 = (b.info >> 8) & 3u, a2 = (b.info >> 10) & 3u;
    if (b.B0 * (cx(r0) - b.x0) + b.C0 * (cy(r0) - b.y0) + b.sb0 < 0) return -1;
    if (b.B1 * (cx(r1) - b.x1) + b.C1 * (cy(r1) - b.y1) + b.sb1 < 0) return -1;
    if (b.B2 * (cx(r2) - b.x2) + b.C2 * (cy(r2) - b.y2) + b.sb2 < 0) return -1;
    int f = 0;
    if (b.B0 * (cx(a0) - b.x0) + b.C0 * (cy(a0) - b.y0) + b.sb0 >= 0) f |= 1;
    if (b.B1 * (cx(a1) - b.x1) + b.C1 * (cy(a1) - b.y1) + b.sb1 >= 0) f |= 2;
    if (b.B2 * (cx(a2) - b.x2) + b.C2 * (cy(a2) - b.y2) + b.sb2 >= 0) f |= 4;
    return f;
}

// Reference k (row-major over the bbox) of a compact descriptor: its tile.
__device__ __forceinline__ uint32_t compact_tile(uint32_t desc, uint32_t k, uint32_t tilesX) {
    const uint32_t w = 1u + ((desc >> 28) & 1u);
    return ((desc >> 14) & 0x3FFFu) * tilesX + (desc & 0x3FFFu) + (k % w) + (k / w) * tilesX;
}
__device__ __forceinline__ uint32_t compact_refs(uint32_t desc) {
    return (1u + ((desc >> 28) & 1u)) * (1u + ((desc >> 29) & 1u));
}

// A slot's references: its records, the tiles of a record in its tile bbox.
// A compact descriptor is its record's tile bbox (at most 2x2 tiles, no corner
// tests; the record is the slot).  Otherwise the descriptor is
// (first record << 3) | count: the binning kernels list such descriptors per
// step and hand them to block_enum_listed.

// Whether any of a thread's descriptors of a step is non-compact.
template <uint32_t kN>
__device__ __forceinline__ int any_non_compact(const uint32_t (&d)[kN]) {
    bool any = false;
#pragma unroll
    for (uint32_t u = 0; u < kN; u++) any |= d[u] != 0 && !(d[u] & kDescCompact);
    return any ? 1 : 0;
}

// The references of record rec (its tile bbox loaded in b) whose bbox tiles
// are k0, k0 + dk, ... (row-major over the bbox), passed to emit(tile, tx,
// ty, payload).
template <class Emit>
__device__ __forceinline__ void enum_record_tiles(const FrameParams& P, const BinTri& b, uint32_t rec, uint32_t k0,
                                                  uint32_t dk, Emit&& emit) {
    const uint32_t n = (uint32_t)(b.w * b.h);
    if (n == 0) return;
    const uint32_t w = (uint32_t)b.w;
    uint32_t cx = k0 % w, cy = k0 / w;
    const uint32_t dx = dk % w, dy = dk / w;
    for (uint32_t k = k0; k < n; k += dk) {
        const int tx = b.minTX + (int)cx, ty = b.minTY + (int)cy;
        const int f = b.small ? 0 : bin_test(b, tx, ty);
        if (f >= 0)
            emit((uint32_t)ty * P.tiles_x + (uint32_t)tx, tx, ty,
                 b.small ? rec : rec | ((uint32_t)f << kRecordBits) | (1u << 31));
        cx += dx;
        cy += dy;
        if (cx >= w) {
            cx -= w;
            cy++;
        }
    }
}

// Records of a record's bbox this size or smaller are enumerated by one
// thread; larger ones by a whole wave, 64 tiles at a time.
constexpr uint32_t kThreadTiles = 16;
// Wide records listed per step: beyond this many, a thread enumerates its
// record itself.
constexpr uint32_t kWideCap = 256;

// All references of the nBig listed non-compact descriptors, for the whole
// workgroup (every thread must call it; sScan holds kBlock + kBlock / kWave
// words): the descriptors' records are numbered, kBlock descriptors at a
// time, and each thread takes records, so that the records' loads are in
// flight together.  A thread enumerates a record of up to kThreadTiles bbox
// tiles itself and lists a wider one, which a wave then enumerates.  There is
// no order among the emitted references.
template <class Emit>
__device__ __forceinline__ void block_enum_listed(const FrameParams& P, const RasterRec* rrec, const uint32_t* sBig,
                                                  uint32_t nBig, uint32_t* sScan, uint32_t* sWide, uint32_t* sWideN,
                                                  Emit&& emit) {
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    uint32_t* sPre = sScan;            // record number of the batch's first record of each descriptor
    uint32_t* sWaveSum = sScan + kBlock;
    for (uint32_t b0 = 0; b0 < nBig; b0 += kBlock) {
        const uint32_t nb = min(nBig - b0, kBlock);
        const uint32_t c = threadIdx.x < nb ? sBig[b0 + threadIdx.x] & 7u : 0u;
        const uint32_t incl = wave_incl_scan(c);
        if (lane == kWave - 1) sWaveSum[wave] = incl;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < kBlock / kWave; w++) {
            before += w < wave ? sWaveSum[w] : 0u;
            total += sWaveSum[w];
        }
        sPre[threadIdx.x] = before + incl - c;
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < total; j += kBlock) {
            // the descriptor whose records hold j: every listed count is >= 1,
            // so the numbers increase strictly over [0, nb)
            uint32_t lo = 0, hi = nb - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (sPre[mid] <= j) lo = mid;
                else hi = mid - 1;
            }
            const uint32_t rec = (sBig[b0 + lo] >> 3) + (j - sPre[lo]);
            BinTri b;
            load_bin_tri(P, rrec[rec], b);
            const uint32_t n = (uint32_t)(b.w * b.h);
            uint32_t slot = kWideCap;
            if (n > kThreadTiles) {
                slot = atomicAdd(sWideN, 1u);
                if (slot < kWideCap) sWide[slot] = rec;
            }
            if (slot >= kWideCap) enum_record_tiles(P, b, rec, 0, 1, emit);
        }
        __syncthreads();  // the numbering is rewritten by the next batch
    }
    const uint32_t nWide = min(*sWideN, kWideCap);
    for (uint32_t i = wave; i < nWide; i += kBlock / kWave) {
        const uint32_t rec = sWide[i];
        BinTri b;
        load_bin_tri(P, rrec[rec], b);
        enum_record_tiles(P, b, rec, lane, kWave, emit);
    }
}

}  // namespace

__global__ __launch_bounds__(kBlock) void k_bin_count(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                      DeviceBuffers B) {
    constexpr uint32_t kUnroll = 4;
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t sBox[4];  // min x, min y, max x, max y of the chunk's tiles
    __shared__ uint32_t sBig[kUnroll * kBlock], sBigN;  // a step's non-compact descriptors
    __shared__ uint32_t sWide[kWideCap], sWideN;         // and their wide records
    __shared__ uint32_t sScan[kBlock + kBlock / kWave];
    const FrameParams& P = jobs[job_of(Q, Q.chunk_job, blockIdx.x)];
    const uint32_t chunk = blockIdx.x - P.chunk_base;  // chunk of the job
    for (uint32_t i = threadIdx.x; i < P.num_tiles; i += kBlock) hist[i] = 0;
    if (threadIdx.x == 0) {
        sBox[0] = sBox[1] = 0xFFFFu;
        sBox[2] = sBox[3] = 0u;
        sBigN = 0;
    }
    __syncthreads();
    uint32_t bx0 = 0xFFFFu, by0 = 0xFFFFu, bx1 = 0u, by1 = 0u;  // this thread's references
    const uint32_t s0 = P.slot_base + chunk * P.chunk_slots;
    const uint32_t s1 = min(P.slot_base + P.num_triangles, s0 + P.chunk_slots);
    // Four slots per thread and step, their descriptors loaded together so
    // the loads' latencies overlap.  Compact descriptors (at most 2x2 tiles)
    // are counted by their lanes; the others are listed and enumerated by the
    // whole workgroup at the end of the step.
    const uint32_t tilesX = P.tiles_x;
    // A tile's count of heavy references rides in the high half of its
    // histogram entry while a chunk cannot reach 2^16 references per tile (at
    // most 7 records per slot); one global atomic per tile then adds it to
    // tile_big, instead of one per reference.
    const bool packBig = (uint64_t)P.chunk_slots * 7u <= 0xFFFFu;
    auto bump = [&](uint32_t tile, bool heavy) {
        if (packBig) {
            atomicAdd(&hist[tile], heavy ? 0x10001u : 1u);
        } else {
            atomicAdd(&hist[tile], 1u);
            if (heavy) atomicAdd(&B.tile_big[P.tile_base + tile], 1u);
        }
    };
    // The next step's descriptors are loaded before this step's work.
    uint32_t dn[kUnroll];
#pragma unroll
    for (uint32_t u = 0; u < kUnroll; u++) {
        const uint32_t s = s0 + threadIdx.x + u * kBlock;
        dn[u] = s < s1 ? B.desc[s] : 0u;
    }
    for (uint32_t sb = s0 + threadIdx.x; sb - threadIdx.x < s1; sb += kUnroll * kBlock) {
        uint32_t d[kUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; u++) {
            d[u] = dn[u];
            const uint32_t sn = sb + (kUnroll + u) * kBlock;
            dn[u] = sn < s1 ? B.desc[sn] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; u++) {
            if (d[u] & kDescCompact) {
                // at most 2x2 tiles, no corner tests
                const uint32_t R = compact_refs(d[u]);
                for (uint32_t k = 0; k < R; k++) bump(compact_tile(d[u], k, tilesX), (d[u] & kDescHeavy) != 0);
                const uint32_t tx = d[u] & 0x3FFFu, ty = (d[u] >> 14) & 0x3FFFu;
                bx0 = min(bx0, tx);
                by0 = min(by0, ty);
                bx1 = max(bx1, tx + ((d[u] >> 28) & 1u));
                by1 = max(by1, ty + ((d[u] >> 29) & 1u));
            } else if (d[u]) {
                sBig[atomicAdd(&sBigN, 1u)] = d[u];
            }
        }
        if (__syncthreads_or(any_non_compact(d))) {
            const uint32_t nBig = sBigN;
            __syncthreads();  // every thread has read the count
            if (threadIdx.x == 0) {
                sBigN = 0;
                sWideN = 0;
            }
            __syncthreads();
            block_enum_listed(P, B.rrec, sBig, nBig, sScan, sWide, &sWideN,
                              [&](uint32_t tile, int tx, int ty, uint32_t payload) {
                                  bump(tile, (payload >> 31) != 0);
                                  bx0 = min(bx0, (uint32_t)tx);
                                  by0 = min(by0, (uint32_t)ty);
                                  bx1 = max(bx1, (uint32_t)tx);
                                  by1 = max(by1, (uint32_t)ty);
                              });
            __syncthreads();  // the lists are rewritten by the next step
        }
    }
    // The chunk's tile box (k_bin_scatter sets up cursors only inside it).
    bx0 = wave_min(bx0);
    by0 = wave_min(by0);
    bx1 = wave_max(bx1);
    by1 = wave_max(by1);
    if (__lane_id() == 0) {
        atomicMin(&sBox[0], bx0);
        atomicMin(&sBox[1], by0);
        atomicMax(&sBox[2], bx1);
        atomicMax(&sBox[3], by1);
    }
    __syncthreads();
    const uint32_t x0 = sBox[0], y0 = sBox[1], x1 = sBox[2], y1 = sBox[3];
    if (threadIdx.x == 0) {
        B.chunk_box[2 * blockIdx.x] = x0 | (y0 << 16);
        B.chunk_box[2 * blockIdx.x + 1] = x1 | (y1 << 16);
    }
    // Only the box part of the chunk's row is written: the column scans and
    // the scatter read an entry only when its tile lies in the chunk's box.
    if (x0 <= x1 && y0 <= y1) {
        uint32_t* row = B.bin_counts + P.count_base + (size_t)chunk * P.num_tiles;
        const uint32_t bw = x1 - x0 + 1, nb = bw * (y1 - y0 + 1);
        for (uint32_t i = threadIdx.x; i < nb; i += kBlock) {
            const uint32_t t = (y0 + i / bw) * tilesX + x0 + i % bw;
            const uint32_t h = hist[t];
            row[t] = packBig ? h & 0xFFFFu : h;
            if (packBig && (h >> 16)) atomicAdd(&B.tile_big[P.tile_base + t], h >> 16);
        }
    }
}

// Each job's chunk x tile count matrix is scanned down each tile's column in
// three steps over blocks of 64 tiles x kChunkGroup chunks (lanes = tiles, each
// wave a quarter of the group's chunks), so every step has C/64 x tiles/64
// blocks per job (grid z = job; blocks beyond a smaller job's matrix exit).
namespace {
constexpr uint32_t kWaveChunks = kChunkGroup / (kBlock / kWave);  // chunks of a group per wave

// The wave's counts of one tile column (all loads issued before any use).
// Entries whose tile lies outside the chunk's tile box are zero and are not
// stored: they are neither read nor written.  Bit k of the returned mask is
// set when chunk c0 + k's entry is stored.
__device__ __forceinline__ uint32_t colgroup_load(const FrameParams& P, const DeviceBuffers& B, uint32_t tile,
                                                  uint32_t c0, uint32_t c1, uint32_t (&v)[kWaveChunks]) {
    const bool in = tile < P.num_tiles;
    const uint32_t ty = in ? tile / P.tiles_x : 0u, tx = in ? tile - ty * P.tiles_x : 0u;
    const uint32_t* box = B.chunk_box + 2 * (size_t)(P.chunk_base + __builtin_amdgcn_readfirstlane(c0));
    uint32_t mask = 0;
#pragma unroll
    for (uint32_t k = 0; k < kWaveChunks; k++) {
        bool stored = false;
        if (c0 + k < c1) {
            const uint32_t b0 = box[2 * k], b1 = box[2 * k + 1];
            stored = in && tx >= (b0 & 0xFFFFu) && tx <= (b1 & 0xFFFFu) && ty >= (b0 >> 16) && ty <= (b1 >> 16);
        }
        mask |= stored ? 1u << k : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kWaveChunks; k++)
        v[k] = (mask >> k) & 1u ? B.bin_counts[P.count_base + (size_t)(c0 + k) * P.num_tiles + tile] : 0u;
    return mask;
}

// Per-wave partial sums of one (64-tile, chunk-group) block, combined in LDS;
// returns this wave's exclusive base within the group and the group total.
__device__ __forceinline__ void colgroup_sums(const uint32_t (&v)[kWaveChunks], uint32_t (*part)[kWave],
                                              uint32_t& waveBase, uint32_t& groupTotal) {
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kWaveChunks; k++) sum += v[k];
    part[wave][lane] = sum;
    __syncthreads();
    waveBase = 0;
    groupTotal = 0;
    for (uint32_t w = 0; w < kBlock / kWave; w++) {
        if (w < wave) waveBase += part[w][lane];
        groupTotal += part[w][lane];
    }
}
__device__ __forceinline__ void colgroup_range(const FrameParams& P, uint32_t& c0, uint32_t& c1) {
    const uint32_t wave = threadIdx.x / kWave;
    c0 = min(P.num_chunks, blockIdx.y * kChunkGroup + wave * kWaveChunks);
    c1 = min(P.num_chunks, c0 + kWaveChunks);
}
}  // namespace

// Whether block (x, y) of the column-scan grid lies inside job P's matrix.
__device__ __forceinline__ bool colgroup_in_job(const FrameParams& P) {
    return blockIdx.x * kWave < P.num_tiles && blockIdx.y * kChunkGroup < P.num_chunks;
}

// Step 1: per (tile, chunk group) column sums.
__global__ __launch_bounds__(kBlock) void k_bin_colsum(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                       DeviceBuffers B) {
    __shared__ uint32_t part[kBlock / kWave][kWave];
    const FrameParams& P = jobs[blockIdx.z];
    if (!colgroup_in_job(P)) return;
    const uint32_t tile = blockIdx.x * kWave + __lane_id();
    uint32_t c0, c1, waveBase, groupTotal, v[kWaveChunks];
    colgroup_range(P, c0, c1);
    colgroup_load(P, B, tile, c0, c1, v);
    colgroup_sums(v, part, waveBase, groupTotal);
    if (threadIdx.x < kWave && tile < P.num_tiles)
        B.group_sums[P.group_base + (size_t)blockIdx.y * P.num_tiles + tile] = groupTotal;
}

// Step 2: per tile, exclusive scan over the chunk groups and the tile total.
__global__ __launch_bounds__(kBlock) void k_bin_groupscan(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                          DeviceBuffers B) {
    const FrameParams& P = jobs[blockIdx.z];
    const uint32_t tile = blockIdx.x * kBlock + threadIdx.x;
    if (tile >= P.num_tiles) return;
    const uint32_t G = (P.num_chunks + kChunkGroup - 1) / kChunkGroup;
    uint32_t* col = &B.group_sums[P.group_base + tile];
    uint32_t run = 0;
    // Batches of 16 groups: the loads of a batch are issued before its stores.
    constexpr uint32_t kBatch = 16;
    for (uint32_t g0 = 0; g0 < G; g0 += kBatch) {
        uint32_t v[kBatch];
#pragma unroll
        for (uint32_t k = 0; k < kBatch; k++) v[k] = g0 + k < G ? col[(size_t)(g0 + k) * P.num_tiles] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kBatch; k++) {
            if (g0 + k < G) col[(size_t)(g0 + k) * P.num_tiles] = run;
            run += v[k];
        }
    }
    B.tile_totals[P.tile_base + tile] = run;
}

// Step 3: rewrite the counts in place as offsets within the tile's list.
__global__ __launch_bounds__(kBlock) void k_bin_colscan(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                        DeviceBuffers B) {
    __shared__ uint32_t part[kBlock / kWave][kWave];
    const FrameParams& P = jobs[blockIdx.z];
    if (!colgroup_in_job(P)) return;
    const uint32_t tile = blockIdx.x * kWave + __lane_id();
    uint32_t c0, c1, waveBase, groupTotal, v[kWaveChunks];
    colgroup_range(P, c0, c1);
    const uint32_t stored = colgroup_load(P, B, tile, c0, c1, v);
    colgroup_sums(v, part, waveBase, groupTotal);
    if (tile >= P.num_tiles) return;
    uint32_t run = B.group_sums[P.group_base + (size_t)blockIdx.y * P.num_tiles + tile] + waveBase;
#pragma unroll
    for (uint32_t k = 0; k < kWaveChunks; k++) {
        if ((stored >> k) & 1u) B.bin_counts[P.count_base + (size_t)(c0 + k) * P.num_tiles + tile] = run;
        run += v[k];
    }
}

// The three column-scan steps in one launch, by a decoupled look-back over
// the chunk groups of each tile column.  A workgroup sums its entries per
// tile, publishes the group sums in group_sums (value | kAggregate), adds the
// earlier groups' published values walking back to the nearest inclusive
// prefix, publishes its own inclusive prefix (value | kInclusive) and
// rewrites its entries as offsets within the tile's list; the last group
// writes the tile totals.  Each status word is written whole, so its flag and
// value are seen together.  A wait is bounded by scan_spin_limit: past it the
// frame is flagged (kOvScanWait) and run again with the three kernels; the
// workgroup still publishes, so no other workgroup waits longer on it.
namespace {
constexpr uint32_t kAggregate = 1u << 30, kInclusive = 2u << 30, kStatusValue = (1u << 30) - 1;
__device__ __forceinline__ uint32_t status_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
}  // namespace

__global__ __launch_bounds__(kBlock) void k_bin_colpass(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                        DeviceBuffers B) {
    __shared__ uint32_t part[kBlock / kWave][kWave];
    __shared__ uint32_t sExcl[kWave];
    // Workgroups are dispatched in linear index order, x fastest, then the
    // chunk group y: a workgroup waits only on lower indices (same job,
    // earlier groups), which have started.
    const uint32_t x = blockIdx.x, g = blockIdx.y;
    const FrameParams& P = jobs[blockIdx.z];
    if (x * kWave >= P.num_tiles) return;
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    const uint32_t tile = x * kWave + lane;
    const bool in = tile < P.num_tiles;
    const uint32_t G = (P.num_chunks + kChunkGroup - 1) / kChunkGroup;
    if (G == 0) {  // a job without triangles: its tiles are empty
        if (g == 0 && wave == 0 && in) B.tile_totals[P.tile_base + tile] = 0u;
        return;
    }
    if (g >= G) return;
    const uint32_t c0 = min(P.num_chunks, g * kChunkGroup + wave * kWaveChunks);
    const uint32_t c1 = min(P.num_chunks, c0 + kWaveChunks);
    uint32_t v[kWaveChunks], waveBase, groupTotal;
    const uint32_t stored = colgroup_load(P, B, tile, c0, c1, v);
    colgroup_sums(v, part, waveBase, groupTotal);
    if (wave == 0) {
        uint32_t excl = 0;
        if (in) {
            uint32_t* st = B.group_sums + P.group_base + tile;
            if (g > 0) {
                status_store(st + (size_t)g * P.num_tiles, kAggregate | groupTotal);
                bool failed = B.scan_spin_limit == 0;
                uint32_t spins = 0;
                // kLook predecessors are read at once, nearest first; the
                // walk takes them up to the first inclusive prefix or the
                // first one not yet published, where the next round resumes.
                constexpr uint32_t kLook = 16;
                for (uint32_t gp = g; gp > 0 && !failed;) {
                    uint32_t sv[kLook];
#pragma unroll
                    for (uint32_t k = 0; k < kLook; k++)
                        sv[k] = k < gp ? status_load(st + (size_t)(gp - 1 - k) * P.num_tiles) : kInclusive;
                    uint32_t taken = 0;
                    bool done = false;
#pragma unroll
                    for (uint32_t k = 0; k < kLook; k++) {
                        if (done || taken < k || !(sv[k] >> 30)) continue;
                        if (k < gp) excl += sv[k] & kStatusValue;
                        taken = k + 1;
                        done = (sv[k] & kInclusive) != 0;
                    }
                    if (done) break;
                    gp -= taken;
                    if (taken < kLook && gp > 0) {  // stopped at one not yet published
                        if (++spins > B.scan_spin_limit) failed = true;
                        else __builtin_amdgcn_s_sleep(1);
                    }
                }
                if (failed) atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvScanWait);
            }
            // (a prefix beyond the 30-bit status value: the three kernels count in 32 bits)
            if ((uint64_t)excl + groupTotal > kStatusValue) atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvScanWait);
            status_store(st + (size_t)g * P.num_tiles, kInclusive | ((excl + groupTotal) & kStatusValue));
            if (g == G - 1) B.tile_totals[P.tile_base + tile] = excl + groupTotal;
        }
        sExcl[lane] = excl;
    }
    __syncthreads();
    if (!in) return;
    uint32_t run = sExcl[lane] + waveBase;
#pragma unroll
    for (uint32_t k = 0; k < kWaveChunks; k++) {
        if ((stored >> k) & 1u) B.bin_counts[P.count_base + (size_t)(c0 + k) * P.num_tiles + tile] = run;
        run += v[k];
    }
}

namespace {
// Exclusive scans of kN values at once over the 1024 threads of a block:
// one pair of barriers for all of them.
template <int kN>
__device__ __forceinline__ void scan1024_n(uint32_t (&v)[kN], uint32_t (*wsum)[16], uint32_t (&total)[kN]) {
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    uint32_t x[kN];
#pragma unroll
    for (int i = 0; i < kN; i++) x[i] = wave_incl_scan(v[i]);
    if (lane == kWave - 1) {
#pragma unroll
        for (int i = 0; i < kN; i++) wsum[i][wave] = x[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kN; i++) {
        uint32_t base = 0, tot = 0;
        for (uint32_t w = 0; w < 16; w++) {
            if (w < wave) base += wsum[i][w];
            tot += wsum[i][w];
        }
        total[i] = tot;
        v[i] = base + x[i] - v[i];
    }
    __syncthreads();
}

// Raster work items of a tile with n references (an empty tile has none).
__device__ __forceinline__ uint32_t tile_segments(uint32_t n, uint32_t segLog2) {
    return (n + (1u << segLog2) - 1) >> segLog2;
}
}  // namespace

// Exclusive scan of the tile totals over every job's tiles (global tiles <=
// 2^18, one block of 1024), and the raster work list: one item per (tile,
// segment of <= 2^segLog2 refs).  The segment size is the smallest of 256 ..
// 4096 that gives no more items than the raster grid has workgroups (beyond
// one per tile), so a frame with few references is still spread over the
// whole grid instead of being bound by one workgroup walking a deep tile.
// Items are listed longest first, so that the last ones claimed are short and
// the pass does not end on one workgroup finishing a full segment: every full
// segment, then the tiles' last (partial) segments in four size classes,
// largest first.  A segment's key image is at the tile's image base
// (tile_item) plus its segment index, whatever its place in the list.
__global__ __launch_bounds__(1024) void k_bin_tilescan(BatchParams Q, DeviceBuffers B) {
    // staged: [nt] totals, then [nt] big counts (overwritten by each tile's
    // start and item entry), then up to kTilescanWork work items
    extern __shared__ uint32_t sTiles[];
    __shared__ uint32_t sScan6[6][16];
    __shared__ uint32_t sMax, sTotal, sSegs, sExtra;
    const uint32_t nt = Q.total_tiles;
    const uint32_t per = (nt + 1023) / 1024;
    const uint32_t t0 = min(nt, threadIdx.x * per), t1 = min(nt, t0 + per);
    if (threadIdx.x == 0) sMax = sTotal = sSegs = sExtra = 0;
    // Up to kTilescanStaged tiles, the totals and big counts are staged in LDS
    // with coalesced loads that are all in flight together; the passes below
    // walk each thread's contiguous tiles several times.
    const bool staged = nt <= kTilescanStaged;
    if (staged) {
        for (uint32_t t = threadIdx.x; t < nt; t += 1024) {
            sTiles[t] = B.tile_totals[t];
            sTiles[nt + t] = B.tile_big[t];
        }
    }
    __syncthreads();
    auto total_of = [&](uint32_t t) { return staged ? sTiles[t] : B.tile_totals[t]; };
    auto big_of = [&](uint32_t t) { return staged ? sTiles[nt + t] : B.tile_big[t]; };
    uint32_t localRefs = 0, localMax = 0;
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t n = total_of(t);
        localRefs += n;
        localMax = max(localMax, n);
    }
    // Total and deepest tile: wave reductions, then one LDS atomic per wave.
    const uint32_t waveRefs = wave_sum(localRefs), waveMax = wave_max(localMax);
    if (__lane_id() == 0) {
        atomicAdd(&sTotal, waveRefs);
        atomicMax(&sMax, waveMax);
    }
    __syncthreads();
    const uint32_t total = sTotal, maxN = sMax;
    uint32_t segLog2 = kSegLog2Min;
    while (segLog2 < kSegLog2Max && (total >> segLog2) > B.raster_blocks) segLog2++;
    // a tile's segment index must fit above the tile bits of a work item
    while (segLog2 < kSegLog2Max && tile_segments(maxN, segLog2) > (1u << kSegIndexBits)) segLog2++;
    const bool segsFit = tile_segments(maxN, segLog2) <= (1u << kSegIndexBits);
    // A tile's segment log2: the frame's, or kSegLog2Heavy for a tile whose
    // references are mostly tile-level triangles (a whole wave each), so that
    // such a tile is spread over several workgroups.  Finer segments never
    // make a tile's segment index overflow its bits.
    auto tile_log2 = [&](uint32_t n, uint32_t big, bool split) {
        if (!split || big < kHeavyBigMin || 4 * big < n) return segLog2;
        uint32_t s = kSegLog2Heavy;
        while (s < segLog2 && tile_segments(n, s) > (1u << kSegIndexBits)) s++;
        return s;
    };
    // The finer segments of heavy tiles are taken only while the work list
    // has room for them: segments with and without them, summed over the block.
    {
        uint32_t segs = 0, extra = 0;
        auto add = [&](uint32_t n, uint32_t big) {
            const uint32_t plain = tile_segments(n, segLog2);
            segs += plain;
            extra += tile_segments(n, tile_log2(n, big, true)) - plain;
        };
        for (uint32_t t = t0; t < t1; t++) add(total_of(t), big_of(t));
        segs = wave_sum(segs);
        extra = wave_sum(extra);
        if (__lane_id() == 0) {
            atomicAdd(&sSegs, segs);
            atomicAdd(&sExtra, extra);
        }
        __syncthreads();
    }
    const bool split = (uint64_t)sSegs + sExtra <= B.work_cap;
    // Per thread: full segments (cnt[0]) and partial last segments by size
    // class (cnt[1 + c], class c = remainder >> (the tile's segment log2 - 2)).
    uint32_t cnt[5] = {0, 0, 0, 0, 0};
    auto count = [&](uint32_t n, uint32_t big) {
        const uint32_t s = tile_log2(n, big, split);
        cnt[0] += n >> s;
        const uint32_t r = n & ((1u << s) - 1);
        if (r) cnt[1 + (r >> (s - 2))]++;
    };
    for (uint32_t t = t0; t < t1; t++) count(total_of(t), big_of(t));
    // the tile-reference prefix rides along as a sixth value
    uint32_t pos[6] = {cnt[0], cnt[1], cnt[2], cnt[3], cnt[4], localRefs}, tot[6];
    scan1024_n<6>(pos, sScan6, tot);
    uint32_t run = pos[5];
    const uint32_t totalSegs = tot[0] + tot[1] + tot[2] + tot[3] + tot[4];
    // the thread's image base: segments of all earlier threads' tiles
    uint32_t seg = pos[0] + pos[1] + pos[2] + pos[3] + pos[4];
    // list order: full segments, then classes 3, 2, 1, 0
    pos[4] += tot[0];
    pos[3] += tot[0] + tot[4];
    pos[2] += tot[0] + tot[4] + tot[3];
    pos[1] += tot[0] + tot[4] + tot[3] + tot[2];
    const bool fits = total <= B.ref_cap && totalSegs <= B.work_cap && totalSegs <= kImageMask && segsFit;
    // A thread's outputs are its contiguous tiles' entries: written straight to
    // memory, each store instruction touches 64 lines.  Beyond two tiles per
    // thread they go to LDS instead (a tile's start and item entry over its
    // total and big count, which only this thread reads) and are copied out
    // with coalesced stores (10M frame: scans 47 -> 44 us).
    const bool stageOut = staged && nt > 2048;
    const bool stageWork = stageOut && totalSegs <= kTilescanWork;
    uint32_t* sWork = sTiles + 2 * nt;
    auto place = [&](uint32_t t, uint32_t n, uint32_t big) {
        if (stageOut) sTiles[t] = run;
        else B.tile_start[t] = run;
        run += n;
        if (fits) {
            const uint32_t s = tile_log2(n, big, split);
            const uint32_t nf = n >> s, r = n & ((1u << s) - 1);
            if (stageOut) sTiles[nt + t] = seg | (s << kImageBits);
            else B.tile_item[t] = seg | (s << kImageBits);
            // a tile of one segment is marked (kSingleSegment): its image is
            // the 32-bit winner orders
            const uint32_t single = nf + (r ? 1u : 0u) == 1 ? kSingleSegment : 0u;
            uint32_t* work = stageWork ? sWork : B.work;
            for (uint32_t i = 0; i < nf; i++) work[pos[0]++] = t | (i << kTileBits) | single;
            if (r) {
                const uint32_t c = 1 + (r >> (s - 2));
                work[pos[c]++] = t | (nf << kTileBits) | single;
            }
            seg += nf + (r ? 1u : 0u);
        }
    };
    for (uint32_t t = t0; t < t1; t++) place(t, total_of(t), big_of(t));
    if (stageOut) {
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < nt; t += 1024) {
            B.tile_start[t] = sTiles[t];
            if (fits) B.tile_item[t] = sTiles[nt + t];
        }
        if (fits && stageWork)
            for (uint32_t i = threadIdx.x; i < totalSegs; i += 1024) B.work[i] = sWork[i];
    }
    if (threadIdx.x == 0) {
        B.tile_start[nt] = total;
        B.counters[kCntTotalRefs] = total;
        B.counters[kCntWorkItems] = totalSegs;
        B.counters[kCntSegLog2] = segLog2;
        if (!fits) atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvRefs);
    }
}

// k_bin_tilescan for batches beyond kTilescanStaged tiles, with the same
// outputs (tile starts, image bases, the same work list in the same order).
// The three reductions (references and deepest tile; segments with and
// without the heavy tiles' finer ones; the size classes' totals) stream the
// counts with coalesced loads.  The placement then takes the tiles in rounds
// of kTilescanStaged, staged in LDS: each thread takes kRoundPer contiguous
// tiles of a round, and the six prefixes (five classes, references) carry
// from round to round.  (Walking a thread's contiguous share of all tiles
// instead, 113 tiles per thread for 256 jobs, took 0.7 ms: every load
// instruction touches 64 lines.)
__global__ __launch_bounds__(1024) void k_bin_tilescan_rounds(BatchParams Q, DeviceBuffers B) {
    constexpr uint32_t kRound = kTilescanStaged, kRoundPer = kRound / 1024;
    extern __shared__ uint32_t sTiles[];  // [kRound] totals, then [kRound] big counts
    __shared__ uint32_t sScan6[6][16];
    __shared__ uint32_t sRed[4], sCls[5];  // references, deepest, segments, extra; class totals
    const uint32_t nt = Q.total_tiles;
    if (threadIdx.x < 4) sRed[threadIdx.x] = 0;
    if (threadIdx.x < 5) sCls[threadIdx.x] = 0;
    __syncthreads();
    uint32_t refs = 0, deepest = 0;
    for (uint32_t t = threadIdx.x; t < nt; t += 1024) {
        const uint32_t n = B.tile_totals[t];
        refs += n;
        deepest = max(deepest, n);
    }
    refs = wave_sum(refs);
    deepest = wave_max(deepest);
    if (__lane_id() == 0) {
        atomicAdd(&sRed[0], refs);
        atomicMax(&sRed[1], deepest);
    }
    __syncthreads();
    const uint32_t total = sRed[0], maxN = sRed[1];
    uint32_t segLog2 = kSegLog2Min;
    while (segLog2 < kSegLog2Max && (total >> segLog2) > B.raster_blocks) segLog2++;
    while (segLog2 < kSegLog2Max && tile_segments(maxN, segLog2) > (1u << kSegIndexBits)) segLog2++;
    const bool segsFit = tile_segments(maxN, segLog2) <= (1u << kSegIndexBits);
    auto tile_log2 = [&](uint32_t n, uint32_t big, bool split) -> uint32_t {
        if (!split || big < kHeavyBigMin || 4 * big < n) return segLog2;
        uint32_t s = kSegLog2Heavy;
        while (s < segLog2 && tile_segments(n, s) > (1u << kSegIndexBits)) s++;
        return s;
    };
    {
        uint32_t segs = 0, extra = 0;
        for (uint32_t t = threadIdx.x; t < nt; t += 1024) {
            const uint32_t n = B.tile_totals[t], big = B.tile_big[t];
            const uint32_t plain = tile_segments(n, segLog2);
            segs += plain;
            extra += tile_segments(n, tile_log2(n, big, true)) - plain;
        }
        segs = wave_sum(segs);
        extra = wave_sum(extra);
        if (__lane_id() == 0) {
            atomicAdd(&sRed[2], segs);
            atomicAdd(&sRed[3], extra);
        }
        __syncthreads();
    }
    const bool split = (uint64_t)sRed[2] + sRed[3] <= B.work_cap;
    auto count = [&](uint32_t n, uint32_t big, uint32_t (&cnt)[5]) {
        const uint32_t s = tile_log2(n, big, split);
        cnt[0] += n >> s;
        const uint32_t r = n & ((1u << s) - 1);
        if (r) cnt[1 + (r >> (s - 2))]++;
    };
    {
        uint32_t cnt[5] = {0, 0, 0, 0, 0};
        for (uint32_t t = threadIdx.x; t < nt; t += 1024) count(B.tile_totals[t], B.tile_big[t], cnt);
#pragma unroll
        for (int c = 0; c < 5; c++) {
            const uint32_t v = wave_sum(cnt[c]);
            if (__lane_id() == 0 && v) atomicAdd(&sCls[c], v);
        }
        __syncthreads();
    }
    const uint32_t totalSegs = sCls[0] + sCls[1] + sCls[2] + sCls[3] + sCls[4];
    const bool fits = total <= B.ref_cap && totalSegs <= B.work_cap && totalSegs <= kImageMask && segsFit;
    // list order: full segments, then classes 3, 2, 1, 0
    uint32_t pos[5] = {0, sCls[0] + sCls[4] + sCls[3] + sCls[2], sCls[0] + sCls[4] + sCls[3], sCls[0] + sCls[4],
                       sCls[0]};
    uint32_t run = 0, seg = 0;  // references and segments of the earlier rounds
    for (uint32_t r0 = 0; r0 < nt; r0 += kRound) {
        const uint32_t rn = min(kRound, nt - r0);
        __syncthreads();  // the previous round's LDS entries are copied out
        for (uint32_t i = threadIdx.x; i < rn; i += 1024) {
            sTiles[i] = B.tile_totals[r0 + i];
            sTiles[kRound + i] = B.tile_big[r0 + i];
        }
        __syncthreads();
        const uint32_t a0 = min(rn, threadIdx.x * kRoundPer), a1 = min(rn, a0 + kRoundPer);
        uint32_t cnt[5] = {0, 0, 0, 0, 0}, mine = 0;
        for (uint32_t i = a0; i < a1; i++) {
            mine += sTiles[i];
            count(sTiles[i], sTiles[kRound + i], cnt);
        }
        uint32_t x[6] = {cnt[0], cnt[1], cnt[2], cnt[3], cnt[4], mine}, tot[6];
        scan1024_n<6>(x, sScan6, tot);
        uint32_t p[5], tRun = run + x[5], tSeg = seg + x[0] + x[1] + x[2] + x[3] + x[4];
#pragma unroll
        for (int c = 0; c < 5; c++) p[c] = pos[c] + x[c];
        // a tile's start and item entry over its total and big count, which
        // only this thread reads; copied out with coalesced stores below
        for (uint32_t i = a0; i < a1; i++) {
            const uint32_t t = r0 + i, n = sTiles[i], big = sTiles[kRound + i];
            sTiles[i] = tRun;
            tRun += n;
            if (fits) {
                const uint32_t s = tile_log2(n, big, split);
                const uint32_t nf = n >> s, r = n & ((1u << s) - 1);
                sTiles[kRound + i] = tSeg | (s << kImageBits);
                const uint32_t single = nf + (r ? 1u : 0u) == 1 ? kSingleSegment : 0u;
                for (uint32_t k = 0; k < nf; k++) B.work[p[0]++] = t | (k << kTileBits) | single;
                if (r) {
                    const uint32_t c = 1 + (r >> (s - 2));
                    B.work[p[c]++] = t | (nf << kTileBits) | single;
                }
                tSeg += nf + (r ? 1u : 0u);
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < rn; i += 1024) {
            B.tile_start[r0 + i] = sTiles[i];
            if (fits) B.tile_item[r0 + i] = sTiles[kRound + i];
        }
#pragma unroll
        for (int c = 0; c < 5; c++) pos[c] += tot[c];
        run += tot[5];
        seg += tot[0] + tot[1] + tot[2] + tot[3] + tot[4];
    }
    if (threadIdx.x == 0) {
        B.tile_start[nt] = total;
        B.counters[kCntTotalRefs] = total;
        B.counters[kCntWorkItems] = totalSegs;
        B.counters[kCntSegLog2] = segLog2;
        if (!fits) atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvRefs);
    }
}

// Scatter: each reference goes to its tile's cursor, which starts at the
// tile's list offset plus the counts of the earlier chunks (k_bin_colscan), so
// a tile's list holds the chunks in submission order.  Within a chunk the
// order is that of the LDS cursor increments: the raster's depth keys carry
// the submission order, so a tile's winner does not depend on its list order.
__global__ __launch_bounds__(kBlock) void k_bin_scatter(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                        DeviceBuffers B) {
    constexpr uint32_t kUnroll = 4;
    extern __shared__ uint32_t cursor[];
    __shared__ uint32_t sBig[kUnroll * kBlock], sBigN;  // a step's non-compact descriptors
    __shared__ uint32_t sWide[kWideCap], sWideN;         // and their wide records
    __shared__ uint32_t sScan[kBlock + kBlock / kWave];
    if (B.counters[kCntOverflow] & kOvNoLists) return;
    const FrameParams& P = jobs[job_of(Q, Q.chunk_job, blockIdx.x)];
    const uint32_t chunk = blockIdx.x - P.chunk_base;
    const uint32_t b0 = B.chunk_box[2 * blockIdx.x], b1 = B.chunk_box[2 * blockIdx.x + 1];
    const uint32_t x0 = b0 & 0xFFFFu, y0 = b0 >> 16, x1 = b1 & 0xFFFFu, y1 = b1 >> 16;
    if (x0 > x1 || y0 > y1) return;  // the chunk has no references
    {
        // Cursors of the tiles inside the chunk's tile box: every reference of
        // the chunk falls in it (k_bin_count).
        const uint32_t* offs = B.bin_counts + P.count_base + (size_t)chunk * P.num_tiles;
        const uint32_t bw = x1 - x0 + 1, nb = bw * (y1 - y0 + 1);
        for (uint32_t i = threadIdx.x; i < nb; i += kBlock) {
            const uint32_t t = (y0 + i / bw) * P.tiles_x + x0 + i % bw;
            cursor[t] = B.tile_start[P.tile_base + t] + offs[t];
        }
    }
    if (threadIdx.x == 0) sBigN = 0;
    __syncthreads();
    const uint32_t s0 = P.slot_base + chunk * P.chunk_slots;
    const uint32_t s1 = min(P.slot_base + P.num_triangles, s0 + P.chunk_slots);
    const uint32_t tilesX = P.tiles_x;
    // The next step's descriptors are loaded before this step's work.
    uint32_t dn[kUnroll];
#pragma unroll
    for (uint32_t u = 0; u < kUnroll; u++) {
        const uint32_t s = s0 + threadIdx.x + u * kBlock;
        dn[u] = s < s1 ? B.desc[s] : 0u;
    }
    for (uint32_t sb = s0 + threadIdx.x; sb - threadIdx.x < s1; sb += kUnroll * kBlock) {
        uint32_t d[kUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; u++) {
            d[u] = dn[u];
            const uint32_t sn = sb + (kUnroll + u) * kBlock;
            dn[u] = sn < s1 ? B.desc[sn] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; u++) {
            const uint32_t s = sb + u * kBlock;
            if (d[u] & kDescCompact) {
                const uint32_t R = compact_refs(d[u]);
                for (uint32_t k = 0; k < R; k++) B.refs[atomicAdd(&cursor[compact_tile(d[u], k, tilesX)], 1u)] = s;
            } else if (d[u]) {
                sBig[atomicAdd(&sBigN, 1u)] = d[u];
            }
        }
        // Other descriptors: listed and enumerated by the whole workgroup.
        if (__syncthreads_or(any_non_compact(d))) {
            const uint32_t nBig = sBigN;
            __syncthreads();  // every thread has read the count
            if (threadIdx.x == 0) {
                sBigN = 0;
                sWideN = 0;
            }
            __syncthreads();
            block_enum_listed(P, B.rrec, sBig, nBig, sScan, sWide, &sWideN,
                              [&](uint32_t tile, int, int, uint32_t payload) {
                                  B.refs[atomicAdd(&cursor[tile], 1u)] = payload;
                              });
            __syncthreads();  // the lists are rewritten by the next step
        }
    }
}

// ---------------------------------------------------------------------------
// Tile rasterisation + deferred shading.
// ---------------------------------------------------------------------------
namespace {

// Medium references with at least this many quads are walked by the whole
// wave (the triangle in scalar registers, 64 quads per pass); fewer are packed
// with other references' quads into shared passes.
constexpr uint32_t kWaveQuads = 40;
// Packed quads of one wave's slice: at most 64 references of < kWaveQuads.
constexpr uint32_t kOwnerCap = kWave * (kWaveQuads - 1);

struct RTri {
    int x0, y0, x1, y1, x2, y2;
    int B0, C0, B1, C1, B2, C2;
    int pb0, pb1, pb2;   // pixel-level top-left bias (Core/RasterTriangle.h:296-299)
    float inv_det, z0, z1, z2;
    uint32_t info, order;
};

__device__ __forceinline__ int pixel_bias(int ax, int ay, int bx, int by) {
    return ((by > ay) || (ay == by && ax > bx)) ? -1 : 0;
}

__device__ __forceinline__ void load_rtri(const RasterRec& r, RTri& t) {
    t.x0 = r.x0; t.y0 = r.y0; t.x1 = r.x1; t.y1 = r.y1; t.x2 = r.x2; t.y2 = r.y2;
    t.B0 = t.y0 - t.y1; t.C0 = t.x1 - t.x0;
    t.B1 = t.y1 - t.y2; t.C1 = t.x2 - t.x1;
    t.B2 = t.y2 - t.y0; t.C2 = t.x0 - t.x2;
    t.pb0 = pixel_bias(t.x0, t.y0, t.x1, t.y1);
    t.pb1 = pixel_bias(t.x1, t.y1, t.x2, t.y2);
    t.pb2 = pixel_bias(t.x2, t.y2, t.x0, t.y0);
    t.inv_det = r.inv_det;
    t.z0 = r.z0; t.z1 = r.z1; t.z2 = r.z2;
    t.info = r.info;
    t.order = r.order;
}

// Record of an unclipped triangle in slot r from its vertices' screen values
// (the values k_setup used); corner codes are not needed below tile level.
__device__ __forceinline__ void screen_rtri_from(const ScreenVert& s0, const ScreenVert& s1, const ScreenVert& s2,
                                                 uint32_t r, RTri& t) {
    t.x0 = s0.x; t.y0 = s0.y; t.x1 = s1.x; t.y1 = s1.y; t.x2 = s2.x; t.y2 = s2.y;
    t.B0 = t.y0 - t.y1; t.C0 = t.x1 - t.x0;
    t.B1 = t.y1 - t.y2; t.C1 = t.x2 - t.x1;
    t.B2 = t.y2 - t.y0; t.C2 = t.x0 - t.x2;
    t.pb0 = pixel_bias(t.x0, t.y0, t.x1, t.y1);
    t.pb1 = pixel_bias(t.x1, t.y1, t.x2, t.y2);
    t.pb2 = pixel_bias(t.x2, t.y2, t.x0, t.y0);
    t.inv_det = 1.0f / (float)(t.C2 * t.B1 - t.C1 * t.B2);
    t.z0 = s0.z; t.z1 = s1.z; t.z2 = s2.z;
    t.info = 0;
    t.order = make_order(r, 0, false);
}
__device__ __forceinline__ void screen_rtri(const FrameParams& P, const DeviceBuffers& B, uint32_t r, RTri& t) {
    const uint32_t* idx = P.mesh.idx + 3 * (size_t)(r - P.slot_base);
    screen_rtri_from(B.sverts[P.vert_base + idx[0]], B.sverts[P.vert_base + idx[1]], B.sverts[P.vert_base + idx[2]],
                     r, t);
}

// A reference's index-buffer entries when it is a compact unclipped triangle
// (otherwise the job's first entries, a valid address whose value is unused).
__device__ __forceinline__ uint3 ref_indices(const FrameParams& P, uint32_t total_slots, uint32_t pay) {
    const uint32_t rid = pay & kRecordMask;
    const bool compact = !(pay >> 31) && rid < total_slots && rid >= P.slot_base;
    const uint32_t* idx = P.mesh.idx + (compact ? 3 * (size_t)(rid - P.slot_base) : 0);
    return make_uint3(idx[0], idx[1], idx[2]);
}

// Ordered depth key: (depth, submission order) min = the winner of the
// sequential d <= D test (Core/FrameBuffer.cpp:54-68).  -0 and +0 compare equal.
__device__ __forceinline__ uint32_t depth_bits(float d) {
    const uint32_t b = __float_as_uint(d + 0.0f);  // -0 + 0 = +0; any other d is unchanged
    return b ^ ((uint32_t)((int)b >> 31) | 0x80000000u);  // negative: ~b, otherwise b | sign
}
__device__ __forceinline__ uint64_t bg_key() { return ((uint64_t)0xBF800000u << 32) | 0xFFFFFFFFull; }

// Barycentrics and depth at a pixel centre (Core/RasterTriangle.h:324-337).
__device__ __forceinline__ void bary_depth(const RTri& t, int px, int py, float& l0, float& l1, float& d) {
    l0 = (float)(t.B1 * (px - t.x2) + t.C1 * (py - t.y2)) * t.inv_det;
    l1 = (float)(t.B2 * (px - t.x2) + t.C2 * (py - t.y2)) * t.inv_det;
    const float l2 = 1.0f - l0 - l1;
    d = l0 * t.z0 + l1 * t.z1 + l2 * t.z2;
}

// LDS key image rows are one key longer than a tile row, so that vertically
// adjacent pixels fall in different LDS banks.
// (Single-sample only: with four padded planes the multisample kernel ran with
// one resident workgroup per CU fewer, and slower.)
template <bool kMS>
constexpr int kKeyStride = kMS ? kTileSize : kTileSize + 1;
template <bool kMS>
constexpr uint32_t kKeyPlane = kTileSize * kKeyStride<kMS>;

// Samples of one raster pass: MSAA level, first sample and count; sample
// k's keys live in plane k (kKeyPlane<true> keys apart).
struct SamplePass {
    uint32_t level, base, n;
};

// Depth test of one pixel from its edge values e0..e2 (coverage when
// useEdges, Core/Rasterizer.h:162), then the (depth, order) key update; klo is
// the key's low word, 0xFFFFFFFE - order.  The barycentric numerators
// n0 = B1 (x - x2) + C1 (y - y2) and n1 = B2 (x - x2) + C2 (y - y2)
// (Core/RasterTriangle.h:324-337) are e1 and e2 without their top-left bias:
// e1 - pb1 - n0 = B1 (x2 - x1) + C1 (y2 - y1) = 0 as a polynomial identity, so
// it holds in wrapping int32 as well.  With multisampling each sample of the
// pass is tested at centre + offset (Core/Rasterizer.h:243-266):
// e + ox B + oy C, wrapping as the reference.
template <bool kMS>
__device__ __forceinline__ void pixel_key(const RTri& t, int e0, int e1, int e2, uint32_t klo, bool useEdges,
                                          uint64_t* slot, const SamplePass& sp) {
    const int n0 = e1 - t.pb1, n1 = e2 - t.pb2;
    if (!kMS) {
        if (useEdges && (e0 | e1 | e2) < 0) return;
        const float l0 = (float)n0 * t.inv_det;
        const float l1 = (float)n1 * t.inv_det;
        const float l2 = 1.0f - l0 - l1;
        const float d = l0 * t.z0 + l1 * t.z1 + l2 * t.z2;
        if (!(d <= 1.0f)) return;   // never passes d <= D with D <= 1 (NaN included)
        const uint64_t key = ((uint64_t)depth_bits(d) << 32) | (uint64_t)klo;
        // (the atomic alone is cheaper than a read, a compare and the atomic)
        atomicMin((unsigned long long*)slot, (unsigned long long)key);
        return;
    }
    for (uint32_t k = 0; k < sp.n; k++) {
        const int2 o = c_sample_offsets[sp.level][sp.base + k];
        if (useEdges) {
            const int s0 = e0 + o.x * t.B0 + o.y * t.C0;
            const int s1 = e1 + o.x * t.B1 + o.y * t.C1;
            const int s2 = e2 + o.x * t.B2 + o.y * t.C2;
            if ((s0 | s1 | s2) < 0) continue;
        }
        const float l0 = (float)(n0 + o.x * t.B1 + o.y * t.C1) * t.inv_det;
        const float l1 = (float)(n1 + o.x * t.B2 + o.y * t.C2) * t.inv_det;
        const float l2 = 1.0f - l0 - l1;
        const float d = l0 * t.z0 + l1 * t.z1 + l2 * t.z2;
        if (!(d <= 1.0f)) continue;
        const uint64_t key = ((uint64_t)depth_bits(d) << 32) | (uint64_t)klo;
        atomicMin((unsigned long long*)(slot + k * kKeyPlane<kMS>), (unsigned long long)key);
    }
}

// Pixels (qx + dx, qy + dy), dx, dy in {0, 1}, that lie within (maxX, maxY):
// the 2x2 quad at (qx, qy) (one SSE lane per pixel in the reference).  Edge
// functions are evaluated at (qx, qy)'s centre (16x+8, 16y+8) and stepped by
// 16 B and 16 C to the other pixels, which is
// exact in wrapping int32 arithmetic (the reference steps the same way,
// Core/Rasterizer.h:190-198).
template <bool kMS>
__device__ __forceinline__ void quad_depth(const RTri& t, int qx, int qy, bool useEdges, int tminX, int tminY,
                                           uint64_t* keys, int maxX, int maxY, const SamplePass& sp) {
    const int px = qx * 16 + 8, py = qy * 16 + 8;
    const int e0 = t.B0 * (px - t.x0) + t.C0 * (py - t.y0) + t.pb0;
    const int e1 = t.B1 * (px - t.x1) + t.C1 * (py - t.y1) + t.pb1;
    const int e2 = t.B2 * (px - t.x2) + t.C2 * (py - t.y2) + t.pb2;
    const int sB0 = t.B0 * 16, sB1 = t.B1 * 16, sB2 = t.B2 * 16;
    const int sC0 = t.C0 * 16, sC1 = t.C1 * 16, sC2 = t.C2 * 16;
    const uint32_t klo = 0xFFFFFFFEu - t.order;
    uint64_t* row = &keys[(qy - tminY) * kKeyStride<kMS> + (qx - tminX)];
    const bool right = qx + 1 <= maxX, down = qy + 1 <= maxY;
    pixel_key<kMS>(t, e0, e1, e2, klo, useEdges, row, sp);
    if (right) pixel_key<kMS>(t, e0 + sB0, e1 + sB1, e2 + sB2, klo, useEdges, row + 1, sp);
    if (down) pixel_key<kMS>(t, e0 + sC0, e1 + sC1, e2 + sC2, klo, useEdges, row + kKeyStride<kMS>, sp);
    if (right && down)
        pixel_key<kMS>(t, e0 + sB0 + sC0, e1 + sB1 + sC1, e2 + sB2 + sC2, klo, useEdges, row + kKeyStride<kMS> + 1, sp);
}
template <bool kMS>
__device__ __forceinline__ void quad_depth(const RTri& t, int qx, int qy, bool useEdges, int tminX, int tminY,
                                           uint64_t* keys, const SamplePass& sp) {
    quad_depth<kMS>(t, qx, qy, useEdges, tminX, tminY, keys, qx + 1, qy + 1, sp);
}

__device__ __forceinline__ float pow_decl(float x, int n, float s) {
    if (n < 0) return powf(x, s);
    float result = 1.0f, base = x;
    unsigned u = (unsigned)n;
    while (u) {
        if (u & 1u) result *= base;
        u >>= 1;
        if (u) base *= base;
    }
    return result;
}

// Color4b::FromFloats clamp (NaN -> 0) and its byte conversion.
__device__ __forceinline__ float clamp01(float c) { return fminf(fmaxf(c, 0.0f), 1.0f); }
__device__ __forceinline__ uint32_t to_byte(float c) { return (uint32_t)(int)(clamp01(c) * 255.0f + 0.5f); }

// Fragment::Interpolate (Core/Shader.h:142-170) + pixel shader
// (Core/Shader.h:209-282); unclamped RGB.
__device__ __forceinline__ float3 shade_pixel(const FrameParams& P, float b0, float b1, const ShadeVert& v0,
                                              const ShadeVert& v1, const ShadeVert& v2) {
    float b2 = 1.0f - b0 - b1;
    b0 *= v0.inv_w;
    b1 *= v1.inv_w;
    b2 *= v2.inv_w;
    const float invB = 1.0f / (b0 + b1 + b2);
    b0 *= invB;
    b1 *= invB;
    b2 = 1.0f - b0 - b1;
    const float Px = b0 * v0.px + b1 * v1.px + b2 * v2.px;
    const float Py = b0 * v0.py + b1 * v1.py + b2 * v2.py;
    const float Pz = b0 * v0.pz + b1 * v1.pz + b2 * v2.pz;
    const float Nx = b0 * v0.nx + b1 * v1.nx + b2 * v2.nx;
    const float Ny = b0 * v0.ny + b1 * v1.ny + b2 * v2.ny;
    const float Nz = b0 * v0.nz + b1 * v1.nz + b2 * v2.nz;

    float w = 1.0f / sqrtf(Nx * Nx + Ny * Ny + Nz * Nz);
    const float nx = Nx * w, ny = Ny * w, nz = Nz * w;
    const float Lx = P.light[0], Ly = P.light[1], Lz = P.light[2];
    float diffuseAmount = Lx * nx + Ly * ny + Lz * nz;
    diffuseAmount = diffuseAmount < 0.0f ? 0.0f : diffuseAmount;
    const float diffuse = (diffuseAmount + P.ambient) * P.kd * (float)(1.0 / 3.14159265358979323846);
    float r, g, b;
    if (P.shader == kLambertAlbedo) {
        r = diffuse * P.albedo[0];
        g = diffuse * P.albedo[1];
        b = diffuse * P.albedo[2];
    } else {
        float ex = P.eye[0] - Px, ey = P.eye[1] - Py, ez = P.eye[2] - Pz;
        w = 1.0f / sqrtf(ex * ex + ey * ey + ez * ez);
        ex *= w; ey *= w; ez *= w;
        float hx = Lx + ex, hy = Ly + ey, hz = Lz + ez;
        w = 1.0f / sqrtf(hx * hx + hy * hy + hz * hz);
        hx *= w; hy *= w; hz *= w;
        float sa = nx * hx + ny * hy + nz * hz;
        if (P.clamp_spec) sa = fmaxf(sa, 0.0f);
        const float spec = pow_decl(sa, P.shin_int, P.shininess) * P.ks;
        if (P.shader == kBlinnPhongRef) {
            r = g = b = diffuse + spec;
        } else {
            const float Cr = b0 * v0.cr + b1 * v1.cr + b2 * v2.cr;
            const float Cg = b0 * v0.cg + b1 * v1.cg + b2 * v2.cg;
            const float Cb = b0 * v0.cb + b1 * v1.cb + b2 * v2.cb;
            r = diffuse * Cr + spec;
            g = diffuse * Cg + spec;
            b = diffuse * Cb + spec;
        }
    }
    return make_float3(r, g, b);
}

// ---- Texture slots (Texture2D<Color>::Sample, EDXUtil; declared semantics:
// texel (i, j) centred at ((i + 0.5) / w, (j + 0.5) / h), coordinates wrap,
// float-to-index conversion with NaN -> 0 and a clamp to +-2^30). ----
__device__ __forceinline__ int tex_index(float x) {
    return x != x ? 0 : (int)fminf(fmaxf(x, -1073741824.0f), 1073741823.0f);
}
__device__ __forceinline__ float4 tex_texel(const FrameParams& P, uint32_t off, int w, int h, int i, int j) {
    i %= w;
    if (i < 0) i += w;
    j %= h;
    if (j < 0) j += h;
    return P.texels[off + (uint32_t)j * (uint32_t)w + (uint32_t)i];
}
// Bilinear filter over one level of size w x h at offset off.
__device__ __forceinline__ float3 tex_bilinear(const FrameParams& P, uint32_t off, int w, int h, float u, float v) {
    const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const int i = tex_index(x0), j = tex_index(y0);
    const float4 c00 = tex_texel(P, off, w, h, i, j), c10 = tex_texel(P, off, w, h, i + 1, j);
    const float4 c01 = tex_texel(P, off, w, h, i, j + 1), c11 = tex_texel(P, off, w, h, i + 1, j + 1);
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    return make_float3(gy * (gx * c00.x + fx * c10.x) + fy * (gx * c01.x + fx * c11.x),
                       gy * (gx * c00.y + fx * c10.y) + fy * (gx * c01.y + fx * c11.y),
                       gy * (gx * c00.z + fx * c10.z) + fy * (gx * c01.z + fx * c11.z));
}
// Offset of mip level l of a slot.
__device__ __forceinline__ uint32_t tex_level(const TexDesc& d, uint32_t l, int& w, int& h) {
    uint32_t off = d.offset;
    w = (int)d.width;
    h = (int)d.height;
    for (uint32_t k = 0; k < l; k++) {
        off += (uint32_t)w * (uint32_t)h;
        w = max(1, w >> 1);
        h = max(1, h >> 1);
    }
    return off;
}
// The trilinear blend of the two mip levels around lod (clamped to the chain).
__device__ __forceinline__ float3 tex_trilinear_at(const FrameParams& P, const TexDesc& d, float lod, float u,
                                                   float v) {
    lod = fminf(fmaxf(lod, 0.0f), (float)(d.levels - 1));
    const float fl = floorf(lod);
    const uint32_t l0 = (uint32_t)fl, l1 = min(l0 + 1, d.levels - 1);
    const float t = lod - fl;
    int w0, h0, w1, h1;
    const uint32_t o0 = tex_level(d, l0, w0, h0), o1 = tex_level(d, l1, w1, h1);
    const float3 a = tex_bilinear(P, o0, w0, h0, u, v), b = tex_bilinear(P, o1, w1, h1, u, v);
    const float s = 1.0f - t;
    return make_float3(s * a.x + t * b.x, s * a.y + t * b.y, s * a.z + t * b.z);
}
// Sample(uv, differentials): nearest / bilinear on level 0; trilinear with
// lod = log2 of the longer of the two texel-space differentials; or N:1
// anisotropic (N = 4, 8, 16): N trilinear taps spread evenly along the longer
// differential (tap k at offset (2k + 1 - N) / 2N of it) at the lod of the
// shorter one, raised to at least 1/N of the longer, averaged in tap order.
__device__ __forceinline__ float3 tex_sample(const FrameParams& P, const TexDesc& d, float u, float v, float dux,
                                             float dvx, float duy, float dvy) {
    if (d.levels == 0) return make_float3(d.color[0], d.color[1], d.color[2]);
    const int w = (int)d.width, h = (int)d.height;
    if (P.tex_filter == kFilterNearest) {
        const float4 c = tex_texel(P, d.offset, w, h, tex_index(floorf(u * (float)w)), tex_index(floorf(v * (float)h)));
        return make_float3(c.x, c.y, c.z);
    }
    if (P.tex_filter == kFilterLinear) return tex_bilinear(P, d.offset, w, h, u, v);
    const float ax = dux * (float)w, ay = dvx * (float)h, bx = duy * (float)w, by = dvy * (float)h;
    const float la2 = ax * ax + ay * ay, lb2 = bx * bx + by * by;
    if (P.tex_filter == kFilterTrilinear) {
        const float rho2 = fmaxf(la2, lb2);
        return tex_trilinear_at(P, d, rho2 > 0.0f ? 0.5f * log2f(rho2) : 0.0f, u, v);
    }
    const int N = 4 << (P.tex_filter - kFilterAniso4);
    const bool xMajor = la2 >= lb2;
    const float mu = xMajor ? dux : duy, mv = xMajor ? dvx : dvy;
    const float maj2 = xMajor ? la2 : lb2, n2 = (float)(N * N);
    float min2 = xMajor ? lb2 : la2;
    if (min2 * n2 < maj2) min2 = maj2 / n2;
    const float lod = min2 > 0.0f ? 0.5f * log2f(min2) : 0.0f;
    float3 acc = make_float3(0.0f, 0.0f, 0.0f);
    for (int k = 0; k < N; k++) {
        const float sk = (float)(2 * k + 1 - N) / (float)(2 * N);
        const float3 c = tex_trilinear_at(P, d, lod, u + sk * mu, v + sk * mv);
        acc.x += c.x;
        acc.y += c.y;
        acc.z += c.z;
    }
    const float inv = 1.0f / (float)N;
    return make_float3(acc.x * inv, acc.y * inv, acc.z * inv);
}

// Perspective-correct texture coordinate of the triangle at pixel (x, y)
// (Fragment::Interpolate, Core/Shader.h:142-170, at that pixel's centre).
__device__ __forceinline__ float2 interp_uv(const RTri& t, int x, int y, const ShadeVert& v0, const ShadeVert& v1,
                                           const ShadeVert& v2) {
    const int px = x * 16 + 8, py = y * 16 + 8;
    float b0 = (float)(t.B1 * (px - t.x2) + t.C1 * (py - t.y2)) * t.inv_det;
    float b1 = (float)(t.B2 * (px - t.x2) + t.C2 * (py - t.y2)) * t.inv_det;
    float b2 = 1.0f - b0 - b1;
    b0 *= v0.inv_w;
    b1 *= v1.inv_w;
    b2 *= v2.inv_w;
    const float invB = 1.0f / (b0 + b1 + b2);
    b0 *= invB;
    b1 *= invB;
    b2 = 1.0f - b0 - b1;
    return make_float2(b0 * v0.u + b1 * v1.u + b2 * v2.u, b0 * v0.v + b1 * v1.v + b2 * v2.v);
}

// LambertianAlbedoPixelShader (Core/Shader.h:209-244) for pixel (x, y) of the
// 2x2 quad at (x & ~1, y & ~1): the albedo is the triangle's texture sampled
// at the pixel's coordinate with the quad's differentials (lane 1 - lane 0,
// lane 2 - lane 0; uncovered lanes extrapolate the triangle's plane).
__device__ __forceinline__ float3 shade_textured(const FrameParams& P, const RTri& t, int x, int y, float l0, float l1,
                                                 const ShadeVert& v0, const ShadeVert& v1, const ShadeVert& v2,
                                                 uint32_t slot) {
    const int qx = x & ~1, qy = y & ~1;
    const float2 q0 = interp_uv(t, qx, qy, v0, v1, v2), q1 = interp_uv(t, qx + 1, qy, v0, v1, v2),
                 q2 = interp_uv(t, qx, qy + 1, v0, v1, v2);
    const uint32_t lane = (uint32_t)(x & 1) | ((uint32_t)(y & 1) << 1);
    const float2 uv = lane == 0 ? q0 : lane == 1 ? q1 : lane == 2 ? q2 : interp_uv(t, x, y, v0, v1, v2);
    const float3 albedo = tex_sample(P, P.tex[slot], uv.x, uv.y, q1.x - q0.x, q1.y - q0.y, q2.x - q0.x, q2.y - q0.y);
    // the normal and the diffuse term as shade_pixel's Lambertian path
    float b2 = 1.0f - l0 - l1;
    float b0 = l0 * v0.inv_w, b1 = l1 * v1.inv_w;
    b2 *= v2.inv_w;
    const float invB = 1.0f / (b0 + b1 + b2);
    b0 *= invB;
    b1 *= invB;
    b2 = 1.0f - b0 - b1;
    const float Nx = b0 * v0.nx + b1 * v1.nx + b2 * v2.nx;
    const float Ny = b0 * v0.ny + b1 * v1.ny + b2 * v2.ny;
    const float Nz = b0 * v0.nz + b1 * v1.nz + b2 * v2.nz;
    const float w = 1.0f / sqrtf(Nx * Nx + Ny * Ny + Nz * Nz);
    const float nx = Nx * w, ny = Ny * w, nz = Nz * w;
    float diffuseAmount = P.light[0] * nx + P.light[1] * ny + P.light[2] * nz;
    diffuseAmount = diffuseAmount < 0.0f ? 0.0f : diffuseAmount;
    const float diffuse = (diffuseAmount + P.ambient) * P.kd * (float)(1.0 / 3.14159265358979323846);
    return make_float3(diffuse * albedo.x, diffuse * albedo.y, diffuse * albedo.z);
}

// The pixel shader of the job for the winner at pixel (x, y); kTex: a job of
// the launch may use the textured shader (its code is only compiled in then,
// keeping the other shaders' register budget).
template <bool kTex>
__device__ __forceinline__ float3 shade_any(const FrameParams& P, const RTri& t, uint32_t order, int x, int y,
                                            float l0, float l1, const ShadeVert& v0, const ShadeVert& v1,
                                            const ShadeVert& v2) {
    if (kTex && P.shader == kTexturedLambert)
        return shade_textured(P, t, x, y, l0, l1, v0, v1, v2, P.mesh.tex[(order >> 4) - P.slot_base]);
    return shade_pixel(P, l0, l1, v0, v1, v2);
}

}  // namespace

namespace {

// Tile-level references for one wave: the lanes are the 8x8 quads of a
// 16x16-pixel quadrant and the wave walks the tile's four quadrants.  Covers
// TrivialAcceptTriangle over the tile and CoarseRasterize's stepped quadrant
// tests with FineRasterize inside partly covered quadrants
// (Core/Rasterizer.h:32-200,310-353).
template <bool kMS>
__device__ __forceinline__ void raster_large_wave(const FrameParams& P, const RTri& t, uint32_t pay,
                                                  int tminX, int tminY, int tmaxX, int tmaxY, uint64_t* keys,
                                                  const SamplePass& sp) {
    const uint32_t lane = __lane_id();
    const int lx = (int)(lane & 7u) * 2, ly = (int)(lane >> 3) * 2;
    const bool trivial = ((pay >> kRecordBits) & 7u) == 7u;
    const bool big = (pay >> 31) != 0;
    if (trivial) {
        // TrivialAcceptTriangle over the tile (Core/Rasterizer.h:310-353).
        for (int q = 0; q < 4; q++) {
            const int qx = tminX + 16 * (q & 1) + lx, qy = tminY + 16 * (q >> 1) + ly;
            if (qx <= tmaxX - 1 && qy <= tmaxY - 1) quad_depth<kMS>(t, qx, qy, false, tminX, tminY, keys, sp);
        }
    } else {
        // CoarseRasterize (Core/Rasterizer.h:32-111): stepped corner tests per 16x16 quadrant.
        const int baseX = tminX << 4, baseY = tminY << 4;
        const int off = (kTileSize << 4) - 1, stepSize = kTileSize << 3;
        const int sb0 = scalar_bias(t.x0, t.y0, t.x1, t.y1);
        const int sb1 = scalar_bias(t.x1, t.y1, t.x2, t.y2);
        const int sb2 = scalar_bias(t.x2, t.y2, t.x0, t.y0);
        auto ef = [&](int e, uint32_t c) {
            const int cxp = baseX + (int)(c & 1u) * off, cyp = baseY + (int)(c >> 1) * off;
            if (e == 0) return t.B0 * (cxp - t.x0) + t.C0 * (cyp - t.y0) + sb0;
            if (e == 1) return t.B1 * (cxp - t.x1) + t.C1 * (cyp - t.y1) + sb1;
            return t.B2 * (cxp - t.x2) + t.C2 * (cyp - t.y2) + sb2;
        };
        // GenStepVectors (Core/RasterTriangle.h:192-255), lane q of the vector.
        auto step = [&](uint32_t c, int Bk, int Ck, int q) {
            int v;
            if (c == 0) v = (q == 0) ? 0 : (q == 1) ? Bk : (q == 2) ? Ck : Bk + Ck;
            else if (c == 1) v = (q == 0) ? -Bk : (q == 1) ? 0 : (q == 2) ? Ck - Bk : Ck;
            else if (c == 2) v = (q == 0) ? -Ck : (q == 1) ? Bk - Ck : (q == 2) ? 0 : Bk;
            else v = (q == 0) ? -Bk - Ck : (q == 1) ? -Ck : (q == 2) ? -Bk : 0;
            return stepSize * v;
        };
        const uint32_t r0 = t.info & 3u, r1 = (t.info >> 2) & 3u, r2 = (t.info >> 4) & 3u;
        const uint32_t a0 = (t.info >> 6) & 3u, a1 = (t.info >> 8) & 3u, a2 = (t.info >> 10) & 3u;
        const int rj0 = ef(0, r0), rj1 = ef(1, r1), rj2 = ef(2, r2);
        const bool ae0 = (pay >> kRecordBits) & 1u, ae1 = (pay >> (kRecordBits + 1)) & 1u,
                   ae2 = (pay >> (kRecordBits + 2)) & 1u;
        const int ac0 = ae0 ? 0 : ef(0, a0), ac1 = ae1 ? 0 : ef(1, a1), ac2 = ae2 ? 0 : ef(2, a2);
        const int fminX = min(t.x0, min(t.x1, t.x2)) >> 4, fmaxX = max(t.x0, max(t.x1, t.x2)) >> 4;
        const int fminY = min(t.y0, min(t.y1, t.y2)) >> 4, fmaxY = max(t.y0, max(t.y1, t.y2)) >> 4;
        for (int q = 0; q < 4; q++) {
            const bool rej = (rj0 + step(r0, t.B0, t.C0, q) < 0) || (rj1 + step(r1, t.B1, t.C1, q) < 0) ||
                             (rj2 + step(r2, t.B2, t.C2, q) < 0);
            if (rej) continue;
            const bool acc = (ae0 || ac0 + step(a0, t.B0, t.C0, q) >= 0) &&
                             (ae1 || ac1 + step(a1, t.B1, t.C1, q) >= 0) &&
                             (ae2 || ac2 + step(a2, t.B2, t.C2, q) >= 0);
            const int bminX = (q & 1) ? tminX + 16 : tminX, bmaxX = (q & 1) ? tmaxX : tminX + 16;
            const int bminY = (q >> 1) ? tminY + 16 : tminY, bmaxY = (q >> 1) ? tmaxY : tminY + 16;
            const int qx = tminX + 16 * (q & 1) + lx, qy = tminY + 16 * (q >> 1) + ly;
            if (acc) {
                if (qx >= bminX && qx <= bmaxX - 1 && qy >= bminY && qy <= bmaxY - 1)
                    quad_depth<kMS>(t, qx, qy, false, tminX, tminY, keys, sp);
            } else {
                int minX = max(bminX, fminX), minY = max(bminY, fminY);
                const int maxX = min(bmaxX - 1, fmaxX), maxY = min(bmaxY - 1, fmaxY);
                minX -= minX % 2;
                minY -= minY % 2;
                if (qx >= minX && qx <= maxX && qy >= minY && qy <= maxY)
                    quad_depth<kMS>(t, qx, qy, true, tminX, tminY, keys, sp);
            }
        }
    }
}

// The record of lane l, made wave-uniform.
__device__ __forceinline__ RTri broadcast_rtri(const RTri& t, int l) {
    RTri u;
    u.x0 = __builtin_amdgcn_readlane(t.x0, l);
    u.y0 = __builtin_amdgcn_readlane(t.y0, l);
    u.x1 = __builtin_amdgcn_readlane(t.x1, l);
    u.y1 = __builtin_amdgcn_readlane(t.y1, l);
    u.x2 = __builtin_amdgcn_readlane(t.x2, l);
    u.y2 = __builtin_amdgcn_readlane(t.y2, l);
    u.B0 = u.y0 - u.y1; u.C0 = u.x1 - u.x0;
    u.B1 = u.y1 - u.y2; u.C1 = u.x2 - u.x1;
    u.B2 = u.y2 - u.y0; u.C2 = u.x0 - u.x2;
    u.pb0 = __builtin_amdgcn_readlane(t.pb0, l);
    u.pb1 = __builtin_amdgcn_readlane(t.pb1, l);
    u.pb2 = __builtin_amdgcn_readlane(t.pb2, l);
    u.inv_det = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t.inv_det), l));
    u.z0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t.z0), l));
    u.z1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t.z1), l));
    u.z2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t.z2), l));
    u.info = __builtin_amdgcn_readlane(t.info, l);
    u.order = __builtin_amdgcn_readlane(t.order, l);
    return u;
}

// One raster work item: references [seg << segLog2, (seg + 1) << segLog2) of
// a tile, segLog2 being the tile's (tile_item).  Each wave streams its own 64-reference slices through the tile's LDS
// key buffer (no workgroup barrier per slice): small triangles take one lane
// each, medium ones one quad per lane, tile-level ones the whole wave.  The
// item's keys go to partials[workIndex].
template <bool kMS>
__device__ __forceinline__ void raster_item(const BatchParams& Q, const FrameParams* __restrict__ jobs,
                                            const DeviceBuffers& B, uint32_t workIndex,
                                            uint64_t* keys, uint16_t (*sOwner)[kOwnerCap]) {
    const uint32_t workItem = __builtin_amdgcn_readfirstlane(B.work[workIndex]);
    // segFlag: the segment index, with kSingleSegment >> kTileBits above it
    const uint32_t gtile = workItem & (kMaxBatchTiles - 1), segFlag = workItem >> kTileBits;
    const uint32_t seg = segFlag & ((1u << kSegIndexBits) - 1);
    const FrameParams& P = jobs[job_of(Q, Q.tile_job, gtile)];
    const uint32_t tile = gtile - P.tile_base;
    const int tx = (int)(tile % P.tiles_x), ty = (int)(tile / P.tiles_x);
    const int tminX = tx * kTileSize, tminY = ty * kTileSize;
    const int tmaxX = min(tminX + kTileSize, (int)P.width), tmaxY = min(tminY + kTileSize, (int)P.height);
    const uint32_t start = B.tile_start[gtile];
    const uint32_t n = B.tile_start[gtile + 1] - start;
    const uint32_t tileItem = B.tile_item[gtile], segLog2 = tileItem >> kImageBits;
    const uint32_t segBegin = seg << segLog2, segEnd = min(n, (seg + 1) << segLog2);

    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    // Multisampling keeps kPassSamples samples' keys in LDS and takes the
    // item's references once per group of samples.
    const uint32_t S = kMS ? P.samples : 1u;
    for (uint32_t pass = 0; pass * kPassSamples < S; pass++) {
        const SamplePass sp{P.sample_level, pass * kPassSamples, min(S - pass * kPassSamples, kPassSamples)};
        for (uint32_t i = threadIdx.x; i < sp.n * kKeyPlane<kMS>; i += kBlock) keys[i] = bg_key();
        __syncthreads();

        // Each wave takes slices of L references (64, or a quarter of a short
        // segment so that every wave has work).  Loads run ahead of the work:
        // references three slices ahead, index-buffer entries two and screen
        // vertices one, so no slice waits on a chain of dependent loads.  All
        // are unconditional, at clamped addresses, so that the in-order
        // vector-memory wait count stays exact.
        const uint32_t segLen = segEnd - segBegin;
        const uint32_t L = segLen >= kBlock ? kWave : (segLen + kBlock / kWave - 1) / (kBlock / kWave);
        const uint32_t stride = L * (kBlock / kWave);
        const uint32_t first = segBegin + wave * L + lane;
        const uint32_t* segRefs = B.refs + start;
        auto sv_of = [&](uint32_t pay, const uint3& idx, ScreenVert& a, ScreenVert& b2, ScreenVert& c) {
            const uint32_t rid = pay & kRecordMask;
            const bool compact = !(pay >> 31) && rid < Q.total_slots;
            a = B.sverts[P.vert_base + (compact ? idx.x : 0u)];
            b2 = B.sverts[P.vert_base + (compact ? idx.y : 0u)];
            c = B.sverts[P.vert_base + (compact ? idx.z : 0u)];
        };
        uint32_t pay1 = segRefs[min(first, segEnd - 1)];
        uint32_t pay2 = segRefs[min(first + stride, segEnd - 1)];
        uint32_t pay3 = segRefs[min(first + 2 * stride, segEnd - 1)];
        ScreenVert n0, n1, n2;
        sv_of(pay1, ref_indices(P, Q.total_slots, pay1), n0, n1, n2);
        uint3 idx2 = ref_indices(P, Q.total_slots, pay2);
        for (uint32_t b = segBegin + wave * L; b < segEnd; b += stride) {
            const uint32_t p = b + lane;
            const bool valid = lane < L && p < segEnd;
            const uint32_t pay = pay1;
            const uint32_t rid = pay & kRecordMask;
            const bool trivial = ((pay >> kRecordBits) & 7u) == 7u;
            const bool big = (pay >> 31) != 0;
            const bool compact = !big && rid < Q.total_slots;
            const ScreenVert s0 = n0, s1 = n1, s2 = n2;
            sv_of(pay2, idx2, n0, n1, n2);
            idx2 = ref_indices(P, Q.total_slots, pay3);
            pay1 = pay2;
            pay2 = pay3;
            pay3 = segRefs[min(p + 3 * stride, segEnd - 1)];
            // Tile-level paths (RasterizeTile, Core/Renderer.cpp:255-268):
            // trivially accepted, or big with hierarchical rasterisation on.
            const bool large = valid && (trivial || (big && P.hierarchical));
            // Every lane holds its own record; large ones are broadcast below.
            RTri t;
            screen_rtri_from(s0, s1, s2, rid, t);
            if (valid && !compact) load_rtri(B.rrec[rid], t);
            uint32_t myQ = 0;  // quads of a FineRasterize reference beyond the tiny path
            int fx = 0, fy = 0, fnx = 1;
            if (valid && !large) {
                const int bx0 = min(t.x0, min(t.x1, t.x2)) >> 4, bx1 = max(t.x0, max(t.x1, t.x2)) >> 4;
                const int by0 = min(t.y0, min(t.y1, t.y2)) >> 4, by1 = max(t.y0, max(t.y1, t.y2)) >> 4;
                // FineRasterize_SingleSample bbox (Core/Rasterizer.h:132-140).
                int minX = max(tminX, bx0);
                const int maxX = min(tmaxX - 1, bx1);
                int minY = max(tminY, by0);
                const int maxY = min(tmaxY - 1, by1);
                if (bx1 - bx0 <= 1 && by1 - by0 <= 1) {
                    // Pixel bbox of at most 2x2: test only its pixels.  The
                    // reference's even-aligned quads add pixels whose centres lie
                    // outside the bbox, hence outside the triangle, and with edges
                    // this short no edge function can wrap.
                    if (maxX >= minX && maxY >= minY)
                        quad_depth<kMS>(t, minX, minY, true, tminX, tminY, keys, maxX, maxY, sp);
                } else {
                    minX -= minX % 2;
                    minY -= minY % 2;
                    if (maxX >= minX && maxY >= minY) {
                        const int nqx = (maxX - minX) / 2 + 1, nqy = (maxY - minY) / 2 + 1;
                        myQ = (uint32_t)(nqx * nqy);
                        fx = minX;
                        fy = minY;
                        fnx = nqx;
                    }
                }
            }
            // Medium references of kWaveQuads quads or more: the whole wave walks
            // each one's FineRasterize quads (bbox clamped to the tile,
            // Core/Rasterizer.h:126-200), 64 per pass.
            uint64_t wm = __ballot(myQ >= kWaveQuads);
            while (wm) {
                const int l = __ffsll((unsigned long long)wm) - 1;
                wm &= wm - 1;
                const RTri u = broadcast_rtri(t, l);
                const int ux = __builtin_amdgcn_readlane(fx, l), uy = __builtin_amdgcn_readlane(fy, l);
                const uint32_t unx = __builtin_amdgcn_readlane((uint32_t)fnx, l);
                const uint32_t uq = __builtin_amdgcn_readlane(myQ, l);
                // quad k = lane + 64 i at (k % unx, k / unx), stepped per pass
                uint32_t cx = lane % unx, cy = lane / unx;
                const uint32_t dx = kWave % unx, dy = kWave / unx;
                for (uint32_t k = lane; k < uq; k += kWave) {
                    quad_depth<kMS>(u, ux + 2 * (int)cx, uy + 2 * (int)cy, true, tminX, tminY, keys, sp);
                    cx += dx;
                    cy += dy;
                    if (cx >= unx) {
                        cx -= unx;
                        cy++;
                    }
                }
            }
            if (myQ >= kWaveQuads) myQ = 0;
            // Other references with quads: the wave's quads are expanded into
            // (reference, quad) pairs and tested 64 per pass, one quad per lane.
            // Pair k's entry in the wave's owner table holds the owner lane and
            // the quad's column and row in the owner's bbox.
            const uint32_t incl = wave_incl_scan(myQ);
            const uint32_t totalQ = __builtin_amdgcn_readlane(incl, kWave - 1);
            if (totalQ) {
                {
                    uint16_t* own = &sOwner[wave][incl - myQ];
                    uint32_t cx = 0, cy = 0;
                    for (uint32_t i = 0; __ballot(i < myQ); i++) {
                        if (i < myQ) {
                            own[i] = (uint16_t)(lane | (cx << 6) | (cy << 10));
                            if (++cx == (uint32_t)fnx) {
                                cx = 0;
                                cy++;
                            }
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (uint32_t base = 0; base < totalQ; base += kWave) {
                    const uint32_t k = base + lane;
                    const uint32_t e = k < totalQ ? sOwner[wave][k] : 0u;
                    const int j = (int)(e & 63u);
                    RTri u;
                    u.x0 = __shfl(t.x0, j); u.y0 = __shfl(t.y0, j);
                    u.x1 = __shfl(t.x1, j); u.y1 = __shfl(t.y1, j);
                    u.x2 = __shfl(t.x2, j); u.y2 = __shfl(t.y2, j);
                    u.B0 = u.y0 - u.y1; u.C0 = u.x1 - u.x0;
                    u.B1 = u.y1 - u.y2; u.C1 = u.x2 - u.x1;
                    u.B2 = u.y2 - u.y0; u.C2 = u.x0 - u.x2;
                    u.pb0 = __shfl(t.pb0, j); u.pb1 = __shfl(t.pb1, j); u.pb2 = __shfl(t.pb2, j);
                    u.inv_det = __shfl(t.inv_det, j);
                    u.z0 = __shfl(t.z0, j); u.z1 = __shfl(t.z1, j); u.z2 = __shfl(t.z2, j);
                    u.order = __shfl(t.order, j);
                    const int ux = __shfl(fx, j), uy = __shfl(fy, j);
                    if (k < totalQ)
                        quad_depth<kMS>(u, ux + 2 * (int)((e >> 6) & 15u), uy + 2 * (int)(e >> 10), true, tminX,
                                        tminY, keys, sp);
                }
            }
            // Large references: the whole wave takes them one at a time, the
            // owner lane's record broadcast with readlane (no memory round trip).
            uint64_t lm = __ballot(large);
            while (lm) {
                const int l = __ffsll((unsigned long long)lm) - 1;
                lm &= lm - 1;
                const uint32_t ppay = __builtin_amdgcn_readlane(pay, l);
                raster_large_wave<kMS>(P, broadcast_rtri(t, l), ppay, tminX, tminY, tmaxX, tmaxY, keys, sp);
            }
        }
        __syncthreads();
        uint64_t* g = B.partials + ((size_t)((tileItem & kImageMask) + seg) * S + sp.base) * (kTileSize * kTileSize);
        // A tile of one segment has no other image to merge with, and k_shade
        // recomputes the depth: the low words of the keys (the winners'
        // orders; 0xFFFFFFFF where nothing was drawn) are its image.
        const bool orders = !kMS && (segFlag >> kSegIndexBits);
        for (uint32_t i = threadIdx.x; i < sp.n * kTileSize * kTileSize; i += kBlock) {
            const uint32_t plane = i / (kTileSize * kTileSize), p = i % (kTileSize * kTileSize);
            const uint64_t k = keys[plane * kKeyPlane<kMS> + (p / kTileSize) * kKeyStride<kMS> + p % kTileSize];
            if (orders) reinterpret_cast<uint32_t*>(g)[i] = (uint32_t)k;
            else g[i] = k;
        }
        __syncthreads();
    }
}

// Cleared pixels of an empty tile (as k_shade writes uncovered ones).
__device__ __forceinline__ void clear_tile(const BatchParams& Q, const FrameParams* __restrict__ jobs,
                                           const DeviceBuffers& B, uint32_t gtile) {
    const FrameParams& P = jobs[job_of(Q, Q.tile_job, gtile)];
    const uint32_t tile = gtile - P.tile_base;
    const int tminX = (int)(tile % P.tiles_x) * kTileSize, tminY = (int)(tile / P.tiles_x) * kTileSize;
    for (uint32_t i = threadIdx.x; i < kTileSize * kTileSize; i += kBlock) {
        const int x = tminX + (int)(i % kTileSize), y = tminY + (int)(i / kTileSize);
        if (x < (int)P.width && y < (int)P.height) {
            const size_t o = P.pix_base + (size_t)(P.height - 1 - y) * P.width + x;
            B.color[o] = 0u;
            B.depth[o] = 1.0f;
            B.ids[o] = 0xFFFFFFFFu;
            if (B.color_f32) B.color_f32[3 * o] = B.color_f32[3 * o + 1] = B.color_f32[3 * o + 2] = 0.0f;
        }
    }
}

}  // namespace

// Persistent raster kernel: workgroups pull (tile, segment) items from the
// list built by k_bin_tilescan until it is exhausted.  The single-sample
// kernel is held to 96 VGPRs so that five workgroups per CU, as many as its LDS
// allows, are resident.
template <bool kMS>
__global__ __launch_bounds__(kBlock, kMS ? 3 : 5) void k_raster(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                   DeviceBuffers B) {
    __shared__ uint64_t keys[(kMS ? kPassSamples : 1) * kKeyPlane<kMS>];
    __shared__ uint16_t sOwner[kBlock / kWave][kOwnerCap];
    __shared__ uint32_t sItem;
    if (B.counters[kCntOverflow] & kOvNoLists) return;
    const uint32_t nItems = B.counters[kCntWorkItems];
    // Single-sample batches: the empty tiles are cleared here, a share per
    // workgroup, before any item is taken (k_shade skips them).  Lane l looks
    // at tile blockIdx.x + l * gridDim.x of each pass; every wave computes the
    // same mask.
    if (!kMS) {
        for (uint32_t base = blockIdx.x; base < Q.total_tiles; base += gridDim.x * kWave) {
            const uint32_t t = base + __lane_id() * gridDim.x;
            const bool empty = t < Q.total_tiles && B.tile_start[t + 1] == B.tile_start[t];
            uint64_t m = __ballot(empty);
            while (m) {
                const uint32_t l = (uint32_t)__ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                clear_tile(Q, jobs, B, base + l * gridDim.x);
            }
        }
    }
    // The first item is the workgroup's index (no burst of same-address
    // atomics at launch); later ones come from the shared counter.
    for (uint32_t round = 0;; round++) {
        if (threadIdx.x == 0) sItem = round ? gridDim.x + atomicAdd(&B.counters[kCntWorkNext], 1u) : blockIdx.x;
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(sItem);
        if (item >= nItems) break;
        raster_item<kMS>(Q, jobs, B, item, keys, sOwner);
        __syncthreads();
    }
}

// Deferred shading, one pixel per thread and two 512-thread workgroups per
// tile (16 rows each; smaller workgroups than one per tile schedule better and
// cost less where the tile is empty): the
// winner is the minimum (depth, order) key over the tile's segments (the
// sequential d <= D test is an associative minimum), then perspective-correct
// interpolation, the pixel shader and one write of colour, depth and
// triangle ID (Core/Renderer.cpp:272-350).  Empty tiles of single-sample
// batches are cleared by k_raster.
namespace {
// Record and shading vertices of the winner with submission order `order`:
// clipped pieces from their stored record, unclipped triangles rebuilt from
// their vertices' screen values (1/det as setup computed it).  With
// recompute, the screen values come from the shading vertices' positions
// (transform, divide and snap exactly as k_vertex computes them) instead of
// a second gather: cheaper where neighbouring pixels have different winners
// (deep tiles of small triangles), dearer where they share them.
__device__ __forceinline__ void winner_tri(const FrameParams& P, const DeviceBuffers& B, uint32_t order, bool recompute,
                                           RTri& t, ShadeVert& v0, ShadeVert& v1, ShadeVert& v2) {
    const uint32_t tri = order >> 4;  // triangle slot
    if (order & 1u) {
        const uint32_t ri = (B.desc[tri] >> 3) + ((order >> 1) & 7u);
        const ShadeRec s = B.srec[ri];
        v0 = B.verts[s.v0];
        v1 = B.verts[s.v1];
        v2 = B.verts[s.v2];
        load_rtri(B.rrec[ri], t);
    } else {
        const uint32_t* idx = P.mesh.idx + 3 * (size_t)(tri - P.slot_base);
        const uint32_t v0i = P.vert_base + idx[0], v1i = P.vert_base + idx[1], v2i = P.vert_base + idx[2];
        if (recompute) {
            v0 = B.verts[v0i];
            v1 = B.verts[v1i];
            v2 = B.verts[v2i];
            auto screen = [&](const ShadeVert& v, int& x, int& y, float& z) {
                const float4 c = xform4(P.mvp, v.px, v.py, v.pz, 1.0f);
                const int2 xy = snap_point(P, c.x / c.w, c.y / c.w, c.z / c.w);
                x = xy.x;
                y = xy.y;
                z = c.z * (1.0f / c.w);
            };
            screen(v0, t.x0, t.y0, t.z0);
            screen(v1, t.x1, t.y1, t.z1);
            screen(v2, t.x2, t.y2, t.z2);
        } else {
            const ScreenVert s0 = B.sverts[v0i], s1 = B.sverts[v1i], s2 = B.sverts[v2i];
            t.x0 = s0.x; t.y0 = s0.y; t.x1 = s1.x; t.y1 = s1.y; t.x2 = s2.x; t.y2 = s2.y;
            t.z0 = s0.z; t.z1 = s1.z; t.z2 = s2.z;
            v0 = B.verts[v0i];
            v1 = B.verts[v1i];
            v2 = B.verts[v2i];
        }
        t.B1 = t.y1 - t.y2; t.C1 = t.x2 - t.x1;
        t.B2 = t.y2 - t.y0; t.C2 = t.x0 - t.x2;
        t.inv_det = 1.0f / (float)(t.C2 * t.B1 - t.C1 * t.B2);
    }
}

// Minimum key of sample s over the tile's segment images (S samples per image).
__device__ __forceinline__ uint64_t tile_key(const uint64_t* g, uint32_t nseg, uint32_t S, uint32_t s) {
    constexpr uint32_t kPix = kTileSize * kTileSize;
    uint64_t key = g[(size_t)s * kPix];
    uint32_t k = 1;
    for (; k + 4 <= nseg; k += 4) {  // four independent loads in flight
        const uint64_t a = g[((size_t)k * S + s) * kPix], b = g[((size_t)(k + 1) * S + s) * kPix];
        const uint64_t c = g[((size_t)(k + 2) * S + s) * kPix], e = g[((size_t)(k + 3) * S + s) * kPix];
        const uint64_t ab = a < b ? a : b, ce = c < e ? c : e;
        const uint64_t m = ab < ce ? ab : ce;
        key = m < key ? m : key;
    }
    for (; k < nseg; k++) {
        const uint64_t v = g[((size_t)k * S + s) * kPix];
        key = v < key ? v : key;
    }
    return key;
}
}  // namespace

template <bool kTex>
__device__ __forceinline__ void shade_body(const BatchParams& Q, const FrameParams* __restrict__ jobs,
                                           const DeviceBuffers& B) {
    constexpr uint32_t kPix = kTileSize * kTileSize;
    if (B.counters[kCntOverflow] & kOvNoLists) return;
    const uint32_t gtile = blockIdx.x >> 1;
    const uint32_t pix = threadIdx.x + (blockIdx.x & 1u) * 512u;  // pixel of the tile
    const uint32_t start = B.tile_start[gtile];
    const uint32_t n = B.tile_start[gtile + 1] - start;
    if (n == 0 && Q.max_samples == 1) return;  // an empty tile, cleared by k_raster
    const FrameParams& P = jobs[job_of(Q, Q.tile_job, gtile)];
    const uint32_t tile = gtile - P.tile_base;
    const int x = (int)(tile % P.tiles_x) * kTileSize + (int)(pix & 31u);
    const int y = (int)(tile / P.tiles_x) * kTileSize + (int)(pix >> 5);
    if (x >= (int)P.width || y >= (int)P.height) return;
    const uint32_t tileItem = B.tile_item[gtile];
    const uint32_t nseg = tile_segments(n, tileItem >> kImageBits), S = P.samples;
    const uint64_t* g = B.partials + (size_t)(tileItem & kImageMask) * S * kPix + pix;
    const size_t o = P.pix_base + (size_t)(P.height - 1 - y) * P.width + x;
    if (S == 1) {
        // A single-segment tile's image holds 32-bit low words (k_raster).
        const uint64_t key = !n ? bg_key()
                             : nseg == 1 ? (uint64_t)reinterpret_cast<const uint32_t*>(g - pix)[pix] | 0xBF80000000000000ull
                                         : tile_key(g, nseg, 1, 0);
        if (key == bg_key()) {
            B.color[o] = 0u;
            B.depth[o] = 1.0f;
            B.ids[o] = 0xFFFFFFFFu;
            if (B.color_f32) B.color_f32[3 * o] = B.color_f32[3 * o + 1] = B.color_f32[3 * o + 2] = 0.0f;
            return;
        }
        // The key's low word is the winner's submission order (make_order).
        const uint32_t order = 0xFFFFFFFEu - (uint32_t)key;
        RTri t;
        ShadeVert v0, v1, v2;
        winner_tri(P, B, order, n >= kShadeRecomputeRefs, t, v0, v1, v2);
        float l0, l1, d;
        bary_depth(t, x * 16 + 8, y * 16 + 8, l0, l1, d);
        const float3 c = shade_any<kTex>(P, t, order, x, y, l0, l1, v0, v1, v2);
        B.color[o] = to_byte(c.x) | (to_byte(c.y) << 8) | (to_byte(c.z) << 16) | (255u << 24);
        if (B.color_f32) {
            B.color_f32[3 * o] = clamp01(c.x);
            B.color_f32[3 * o + 1] = clamp01(c.y);
            B.color_f32[3 * o + 2] = clamp01(c.z);
        }
        B.depth[o] = d;
        B.ids[o] = (order >> 4) - P.slot_base;
        return;
    }
    // Multisampling: per-sample winners and depths (the sequential LEQUAL test
    // per sample, Core/FrameBuffer.cpp:54-68); a winner is shaded once at the
    // pixel centre (Core/Rasterizer.h:270-285, Core/Renderer.cpp:305-350); the
    // resolve averages the samples' colours (Core/FrameBuffer.cpp:70-87) as
    // Color(b) = b / 255 per channel, summed in sample order, times 1 / S, then
    // FromFloats.
    const size_t plane = (size_t)P.width * P.height;
    const int cx = x * 16 + 8, cy = y * 16 + 8;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // The last winner's record and colour are kept: neighbouring samples
    // mostly share their winner.
    uint32_t cachedOrder = 0xFFFFFFFFu, cachedColor = 0u;
    RTri t;
    for (uint32_t s = 0; s < S; s++) {
        const uint64_t key = n ? tile_key(g, nseg, S, s) : bg_key();
        uint32_t rgba = 0u;
        if (key == bg_key()) {
            B.depth[o + s * plane] = 1.0f;
            B.ids[o + s * plane] = 0xFFFFFFFFu;
        } else {
            const uint32_t order = 0xFFFFFFFEu - (uint32_t)key;
            float l0, l1, d;
            if (order != cachedOrder) {
                ShadeVert v0, v1, v2;
                winner_tri(P, B, order, false, t, v0, v1, v2);
                bary_depth(t, cx, cy, l0, l1, d);
                const float3 c = shade_any<kTex>(P, t, order, x, y, l0, l1, v0, v1, v2);
                cachedColor = to_byte(c.x) | (to_byte(c.y) << 8) | (to_byte(c.z) << 16) | (255u << 24);
                cachedOrder = order;
            }
            const int2 off = c_sample_offsets[P.sample_level][s];
            bary_depth(t, cx + off.x, cy + off.y, l0, l1, d);
            B.depth[o + s * plane] = d;
            B.ids[o + s * plane] = (order >> 4) - P.slot_base;
            rgba = cachedColor;
        }
        for (int ch = 0; ch < 4; ch++) acc[ch] += (float)((rgba >> (8 * ch)) & 255u) / 255.0f;
    }
    const float invS = 1.0f / (float)S;
    B.color[o] = to_byte(acc[0] * invS) | (to_byte(acc[1] * invS) << 8) | (to_byte(acc[2] * invS) << 16) |
                 (to_byte(acc[3] * invS) << 24);
    if (B.color_f32) {
        B.color_f32[3 * o] = clamp01(acc[0] * invS);
        B.color_f32[3 * o + 1] = clamp01(acc[1] * invS);
        B.color_f32[3 * o + 2] = clamp01(acc[2] * invS);
    }
}

// Untextured launches are held to 64 VGPRs, so that eight waves per SIMD are
// resident (two of them spill to scratch outside the pixel's gathers); the
// texture samplers need more.
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_shade(
    BatchParams Q, const FrameParams* __restrict__ jobs, DeviceBuffers B) {
    shade_body<false>(Q, jobs, B);
}
__global__ __launch_bounds__(512) void k_shade_tex(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                   DeviceBuffers B) {
    shade_body<true>(Q, jobs, B);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace {
int g_rasterResident[2] = {0, 0};  // resident k_raster workgroups per CU: single-sample, multisample
}

hipError_t init_kernels() {
    // Bin cursors for up to kMaxTiles tiles live in LDS (160 KB per CU on gfx950).
    const int maxDyn = (int)(sizeof(uint32_t) * kMaxTiles);
    hipError_t e = hipFuncSetAttribute((const void*)k_bin_count, hipFuncAttributeMaxDynamicSharedMemorySize, maxDyn);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_bin_scatter, hipFuncAttributeMaxDynamicSharedMemorySize, maxDyn);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_bin_tilescan, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(uint32_t) * (2 * kTilescanStaged + kTilescanWork)));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_bin_tilescan_rounds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(uint32_t) * 2 * kTilescanStaged));
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&g_rasterResident[0],
                                                         reinterpret_cast<const void*>(&k_raster<false>), kBlock, 0);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&g_rasterResident[1],
                                                         reinterpret_cast<const void*>(&k_raster<true>), kBlock, 0);
    return e;
}

uint32_t raster_resident_per_cu(bool multisample) {
    return (uint32_t)std::max(1, g_rasterResident[multisample ? 1 : 0]);
}

namespace {
std::mutex g_scratchLock;
uint8_t* g_scratch[kMaxScratchDevices] = {};  // device address of each device's scratch copy
}  // namespace

hipError_t scratch_address(uint8_t** out) {
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchLock);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<uint8_t*>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

hipError_t set_frame_job(hipGraphExec_t exec, hipGraphNode_t first, const BatchParams& Q, FrameParams* jobs,
                         const DeviceBuffers& B, const FrameParams& job0) {
    // the arguments of launch_frame's first launch, in its order
    BatchParams q = Q;
    FrameParams* table = jobs;
    ShadeVert* verts = B.verts;
    ScreenVert* sverts = B.sverts;
    uint32_t* counters = B.counters;
    FrameParams job = job0;
    int writeJob = 1;
    hipKernelNodeParams p{};
    p.sharedMemBytes = 0;
    p.extra = nullptr;
    if (Q.vert_blocks) {
        void* args[] = {&q, &table, &verts, &sverts, &counters, &table, &job};
        p.func = reinterpret_cast<void*>(&k_vertex<true>);
        p.gridDim = dim3(Q.vert_blocks);
        p.blockDim = dim3(kBlock);
        p.kernelParams = args;
        return hipGraphExecKernelNodeSetParams(exec, first, &p);
    }
    void* args[] = {&counters, &table, &job, &writeJob};
    p.func = reinterpret_cast<void*>(&k_reset);
    p.gridDim = dim3(1);
    p.blockDim = dim3(64);
    p.kernelParams = args;
    return hipGraphExecKernelNodeSetParams(exec, first, &p);
}

hipError_t launch_frame(const BatchParams& Q, FrameParams* jobs, const FrameParams* job0, const DeviceBuffers& B,
                        hipStream_t s, hipEvent_t* events) {
    const size_t lds = sizeof(uint32_t) * Q.max_job_tiles;
    auto mark = [&](int i) {
        if (events) (void)hipEventRecord(events[i], s);
    };
    const FrameParams none{};
    mark(0);
    if (!Q.vert_blocks)
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, B.counters, jobs, job0 ? *job0 : none, job0 ? 1 : 0);
    else if (job0)
        hipLaunchKernelGGL(k_vertex<true>, dim3(Q.vert_blocks), dim3(kBlock), 0, s, Q, jobs, B.verts, B.sverts,
                           B.counters, jobs, *job0);
    else
        hipLaunchKernelGGL(k_vertex<false>, dim3(Q.vert_blocks), dim3(kBlock), 0, s, Q, jobs, B.verts, B.sverts,
                           B.counters, jobs, none);
    if (Q.slot_blocks) {
        hipLaunchKernelGGL(k_setup, dim3((Q.slot_blocks + kSetupSlots - 1) / kSetupSlots), dim3(kBlock), 0, s, Q,
                           jobs, B);
        // One workgroup per CU at most (its occupancy): the queue's length is
        // known only on the device, and each workgroup loops over it.
        const uint32_t clipBlocks = max(1u, min(B.num_cus, (B.clip_queue_cap + kClipWaves - 1) / kClipWaves));
        hipLaunchKernelGGL(k_clip, dim3(clipBlocks), dim3(kClipWaves * 64), 0, s, Q, jobs, B);
    }
    mark(1);
    if (Q.total_chunks) hipLaunchKernelGGL(k_bin_count, dim3(Q.total_chunks), dim3(kBlock), lds, s, Q, jobs, B);
    mark(2);
    const dim3 colGrid((Q.max_job_tiles + kWave - 1) / kWave, (Q.max_job_chunks + kChunkGroup - 1) / kChunkGroup,
                       Q.num_jobs);
    if (B.legacy_scans) {
        // The group scan also writes the tile totals of jobs without triangles.
        if (Q.total_chunks) hipLaunchKernelGGL(k_bin_colsum, colGrid, dim3(kBlock), 0, s, Q, jobs, B);
        hipLaunchKernelGGL(k_bin_groupscan, dim3((Q.max_job_tiles + kBlock - 1) / kBlock, 1, Q.num_jobs),
                           dim3(kBlock), 0, s, Q, jobs, B);
        if (Q.total_chunks) hipLaunchKernelGGL(k_bin_colscan, colGrid, dim3(kBlock), 0, s, Q, jobs, B);
    } else {
        // (also writes the tile totals of jobs without triangles: a grid of
        // one chunk group at least)
        const dim3 passGrid(colGrid.x, max(1u, colGrid.y), colGrid.z);
        hipLaunchKernelGGL(k_bin_colpass, passGrid, dim3(kBlock), 0, s, Q, jobs, B);
    }
    if (Q.total_tiles <= kTilescanStaged)
        hipLaunchKernelGGL(k_bin_tilescan, dim3(1), dim3(1024),
                           sizeof(uint32_t) * (2 * (size_t)Q.total_tiles + kTilescanWork), s, Q, B);
    else
        hipLaunchKernelGGL(k_bin_tilescan_rounds, dim3(1), dim3(1024), sizeof(uint32_t) * 2 * kTilescanStaged, s, Q,
                           B);
    mark(3);
    if (Q.total_chunks) hipLaunchKernelGGL(k_bin_scatter, dim3(Q.total_chunks), dim3(kBlock), lds, s, Q, jobs, B);
    mark(4);
    if (Q.max_samples > 1) hipLaunchKernelGGL((k_raster<true>), dim3(B.raster_grid), dim3(kBlock), 0, s, Q, jobs, B);
    else hipLaunchKernelGGL((k_raster<false>), dim3(B.raster_grid), dim3(kBlock), 0, s, Q, jobs, B);
    if (Q.any_textured) hipLaunchKernelGGL(k_shade_tex, dim3(2 * Q.total_tiles), dim3(512), 0, s, Q, jobs, B);
    else hipLaunchKernelGGL(k_shade, dim3(2 * Q.total_tiles), dim3(512), 0, s, Q, jobs, B);
    mark(5);
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(256), 0, s, B.counters, B.host_counters);
    return hipGetLastError();
}

}  // namespace edxr
