// pipeline.h -- structures shared by the host orchestration (api.cpp) and the
// gfx950 kernels (pipeline.hip).  Internal: not part of the C ABI.
#pragma once
#include <stdint.h>

#include <hip/hip_runtime_api.h>

namespace edxr {

constexpr int kTileLog2 = 5;                 // Tile::SIZE_LOG_2 (EDXRaster/Core/Tile.h:12-13)
constexpr int kTileSize = 1 << kTileLog2;    // 32 x 32 pixels
constexpr uint32_t kMaxTiles = 16384;        // LDS-resident bin cursors (per job)
constexpr uint32_t kTileBits = 18;           // global tile index bits in a work item
constexpr uint32_t kSegIndexBits = 31 - kTileBits;  // segment index bits in a work item, above the tile
constexpr uint32_t kSingleSegment = 1u << 31;       // work item flag: the tile has one segment
constexpr uint32_t kMaxBatchTiles = 1u << kTileBits;
constexpr uint32_t kMaxDim = 8192;
constexpr uint32_t kRecordBits = 28;         // record index bits in a tile reference
constexpr uint32_t kRecordMask = (1u << kRecordBits) - 1;
constexpr int kMaxPoly = 16;                 // clip polygon capacity (convex input needs <= 9)
constexpr uint32_t kMaxSampleLevel = 5;      // up to 32 samples per pixel
constexpr uint32_t kPassSamples = 4;         // samples whose keys one raster pass keeps in LDS
constexpr uint32_t kSegLog2Min = 8;          // raster work items hold 256 .. 4096 tile references,
constexpr uint32_t kSegLog2Max = 12;         // chosen per frame from the total (k_bin_tilescan)
constexpr uint32_t kSegRefs = 1u << kSegLog2Max;
// Tiles whose references are mostly tile-level triangles (each takes a whole
// wave in the raster) are cut into segments of 2^kSegLog2Heavy references when
// at least kHeavyBigMin of their references, and a quarter of them, are such.
constexpr uint32_t kSegLog2Heavy = 5;
constexpr uint32_t kHeavyBigMin = 32;
// tile_item: first key image of the tile (low kImageBits) | the tile's
// segment log2 << kImageBits.
constexpr uint32_t kImageBits = 26;
constexpr uint32_t kImageMask = (1u << kImageBits) - 1;
constexpr uint32_t kChunkGroup = 64;         // chunks per block of the column scan
constexpr uint32_t kTilescanStaged = 8192;   // batch tiles whose counts k_bin_tilescan stages in LDS
constexpr uint32_t kTilescanWork = 8192;     // work items k_bin_tilescan stages in LDS (with the tiles)
constexpr uint32_t kScanSpinLimit = 1u << 20;  // k_bin_colpass look-back waits (about 30 ms) before giving up
// Single-sample shading rebuilds a winner's screen values from its shading
// vertices, instead of gathering them, in tiles of this many references or
// more (about one triangle per two pixels of a 32x32 tile).
constexpr uint32_t kShadeRecomputeRefs = 512;
constexpr uint32_t kJobAlign = 1024;         // job slot alignment: vertex/setup blocks never span jobs
constexpr uint32_t kJobTableBlock = 256;     // slots (vertices) per entry of the block -> job tables
constexpr uint32_t kMaxChunks = 4096;
// Per-triangle descriptor: bit 31 set = compact small triangle (record in slot
// t; tile bbox minTX | minTY << 14 | dx << 28 | dy << 29, dx, dy in {0, 1});
// otherwise (first record << 3) | record count (0..7).
constexpr uint32_t kDescCompact = 1u << 31;
// Compact descriptor flag: the triangle's pixel bounding box has kHeavyPixels
// or more pixels (it takes a whole wave in the raster; counted in tile_big).
constexpr uint32_t kDescHeavy = 1u << 30;
constexpr uint32_t kHeavyPixels = 256;

enum ShaderKind { kLambertAlbedo = 0, kBlinnPhongRef = 1, kVcBlinnPhong = 2, kTexturedLambert = 3 };
enum TextureFilter {
    kFilterNearest = 0,
    kFilterLinear = 1,
    kFilterTrilinear = 2,
    kFilterAniso4 = 3,  // N:1 anisotropic, N = 4 << (filter - kFilterAniso4)
    kFilterAniso8 = 4,
    kFilterAniso16 = 5
};

struct MeshPtrs {
    const float *px, *py, *pz, *nx, *ny, *nz, *cr, *cg, *cb;
    const uint32_t* idx;
    const float *tu, *tv;      // texture coordinates (textured shader only)
    const uint32_t* tex;       // texture slot per triangle (textured shader only)
};

// One texture slot: its mip chain in the context's texel array (RGBA floats,
// level l of size max(1, w >> l) x max(1, h >> l) at offset + level_offset(l)),
// or a constant colour (levels == 0).
struct TexDesc {
    uint32_t width, height, levels, offset;
    float color[4];
};

// Constants of one render job (one mesh, camera, shader and render target),
// read by the kernels from a device table: a plain frame is a batch of one.
// Batch-wide arrays hold every job's data at the job's bases: shading vertices
// from vert_base, triangle slots (setup records of input triangles) from
// slot_base, tiles from tile_base, binning chunks from chunk_base, its
// chunk x tile count matrix at count_base, column-scan group sums at
// group_base and its pixels at pix_base.  vert_base and slot_base are
// multiples of the 256-lane block, so every block of the vertex and setup
// passes belongs to one job.
struct FrameParams {
    float mvp[16];        // proj * mv, row-major, column-vector convention
    float raster[16];
    float eye[3];         // TransformPoint(0, MV^-1)
    float light[3];       // Normalize(light direction)
    float light_len;      // length of the light direction as given
    float ambient, kd, ks, shininess;
    float albedo[3];
    int32_t shader;
    int32_t shin_int;     // integer exponent, or -1 for powf
    int32_t clamp_spec;
    int32_t hierarchical;
    uint32_t width, height, tiles_x, tiles_y, num_tiles;
    uint32_t num_vertices, num_triangles;
    uint32_t chunk_slots, num_chunks;
    uint32_t samples, sample_level;  // MSAA: 1 << sample_level samples per pixel
    uint32_t vert_base, slot_base, tile_base, chunk_base;
    uint64_t count_base, group_base, pix_base;
    MeshPtrs mesh;
    const TexDesc* tex;   // texture slots (textured shader), num_tex of them
    const float4* texels;
    uint32_t num_tex, tex_filter;
};

// Batch-wide sizes and the block -> job tables (null for a single job).
struct BatchParams {
    uint32_t num_jobs;
    uint32_t total_slots;    // triangle slots; clipped pieces' records follow
    uint32_t total_verts;    // shading-vertex slots; clip-created vertices follow
    uint32_t total_tiles;
    uint32_t total_chunks;
    uint32_t vert_blocks, slot_blocks;
    uint32_t max_job_tiles;     // max over jobs of num_tiles
    uint32_t max_job_chunks;    // max over jobs of num_chunks
    uint32_t max_samples;       // max over jobs of samples
    uint32_t any_textured;      // 1 if a job uses the textured shader
    const uint32_t* vblock_job;
    const uint32_t* sblock_job;
    const uint32_t* chunk_job;
    const uint32_t* tile_job;
};

// Post-transform vertex for shading (ProjectedVertex, Core/Shader.h:13-25,
// with an RGB colour beside the texture coordinate): one 48-byte AoS record so
// a vertex fetch touches one cache line instead of eleven SoA planes.  Indices < V are the input
// vertices (written by k_vertex); V + e are vertices created by clipping
// (Core/Clipper.h:136-151).
struct __attribute__((aligned(16))) ShadeVert {
    float px, py, pz, nx;
    float ny, nz, cr, cg;
    float cb, inv_w, u, v;
};
static_assert(sizeof(ShadeVert) == 48, "ShadeVert layout");

// Per input vertex, the values every unclipped triangle's setup uses: the 28.4
// snapped raster position of the homogeneous-projected vertex, z/w as
// Renderer::RenderMesh's z-divide computes it (Core/Renderer.cpp:139-147) and
// the outcode (Core/Clipper.h:48-68).
struct __attribute__((aligned(16))) ScreenVert {
    int32_t x, y;
    float z;
    uint32_t code;
};

// Rasterisation part of a setup record (RasterTriangle, Core/RasterTriangle.h:13-23).
struct __attribute__((aligned(16))) RasterRec {
    int32_t x0, y0, x1, y1;
    int32_t x2, y2;
    float inv_det, z0;
    float z1, z2;
    uint32_t info;        // corners (12 bits) | fan index << 12
    uint32_t order;       // submission order: tri << 4 | survivor << 1 | clipped
};
static_assert(sizeof(RasterRec) == 48, "RasterRec layout");

// Trivial reject/accept corner choice (Core/RasterTriangle.h:66-150).  The
// reference branches on slope = float(B) / float(C) >= 0; int->float keeps the
// sign and the quotient of nonzero ints cannot underflow, so the test is a sign
// test: C == 0 gives +-inf (or NaN for B == 0), B == 0 with C != 0 gives +-0.
__host__ __device__ inline uint32_t corner_pair(int B, int C) {
    const bool slopeNonNeg = (C == 0) ? (B > 0) : (B == 0 || ((B > 0) == (C > 0)));
    uint32_t rej, acc;
    if (slopeNonNeg) {
        if (C >= 0) { rej = 3; acc = 0; } else { rej = 0; acc = 3; }
    } else {
        if (C >= 0) { rej = 2; acc = 1; } else { rej = 1; acc = 2; }
    }
    return rej | (acc << 8);
}

// RasterRec::info corner codes: reject corners in bits 0-5, accept corners in
// bits 6-11, two bits per edge.
__host__ __device__ inline uint32_t corner_info(int B0, int C0, int B1, int C1, int B2, int C2) {
    const uint32_t c0 = corner_pair(B0, C0), c1 = corner_pair(B1, C1), c2 = corner_pair(B2, C2);
    return (c0 & 3u) | ((c1 & 3u) << 2) | ((c2 & 3u) << 4) | ((c0 >> 8) << 6) | ((c1 >> 8) << 8) | ((c2 >> 8) << 10);
}

// Submission order of a record: triangle slot (the job's slot_base + input
// index), then the index among the triangle's surviving fan pieces
// (Core/Clipper.h:155-170).  Unclipped triangles live in record slot `tri`
// (stored only when their descriptor is not compact); clipped ones at
// desc[tri] >> 3 + survivor.
__host__ __device__ inline uint32_t make_order(uint32_t tri, uint32_t survivor, bool clipped) {
    return (tri << 4) | (survivor << 1) | (clipped ? 1u : 0u);
}

// Shading part of a clipped piece: shading-vertex slots (input vertices at the
// job's vert_base, clip-created ones from total_verts) and the triangle slot.
// Unclipped triangles take their vertex ids from the job's index buffer.
struct __attribute__((aligned(16))) ShadeRec {
    uint32_t v0, v1, v2, tri_id;
};

enum CounterSlot {
    kCntClipQueue = 0,
    kCntOverflow,        // bit0 clip queue, bit1 extra records, bit2 extra vertices, bit3 refs, bit4 polygon
    kCntBadIndex,
    kCntTotalRefs,
    kCntWorkItems,       // raster work items (tile segments)
    kCntWorkNext,        // next work item to claim
    kCntSegLog2,         // log2 of the tile references per work item of this frame
    kCntStriped = 16,    // high-rate counters: kStripes slots each, summed on the host
    kCntAllocStriped = kCntStriped + 4 * 32,  // k_clip's allocations: kAllocStripes vertices, then records
    kCntSlots = kCntAllocStriped + 2 * 16
};

// Counters bumped by many workgroups are striped over kStripes addresses so
// the atomics do not serialise on one cache line.
constexpr uint32_t kStripes = 32;
enum StripedCounter { kStRejected = 0, kStMainRecords = 1, kStClipped = 2, kStNearClipped = 3 };
// k_clip's workgroups allocate extra vertices and records from kAllocStripes
// counters, each over its own region (1 / kAllocStripes of the capacity).
constexpr uint32_t kAllocStripes = 16;

enum OverflowBit {
    kOvClipQueue = 1,
    kOvExtraRecords = 2,
    kOvExtraVertices = 4,
    kOvRefs = 8,
    kOvPolygon = 16,
    kOvScanWait = 32     // a k_bin_colpass look-back gave up waiting: re-run with the three-kernel scan
};
// Overflows after which the tile lists are not valid: the scatter, raster
// and shading kernels return at once (the frame is run again).
constexpr uint32_t kOvNoLists = kOvRefs | kOvScanWait;

struct DeviceBuffers {
    RasterRec* rrec;
    ShadeRec* srec;
    uint32_t* desc;        // per input triangle, see kDescCompact
    uint32_t* clip_queue;
    uint32_t clip_queue_cap;
    uint32_t extra_rec_cap;
    ShadeVert* verts;      // V + extra_vert_cap shading vertices
    ScreenVert* sverts;    // V screen vertices (input vertices only)
    uint32_t extra_vert_cap;
    uint32_t* chunk_box;   // per chunk of the batch: tile box of its references,
                           // (min x | min y << 16, max x | max y << 16); empty: min > max
    uint32_t* bin_counts;  // num_chunks * num_tiles, scanned in place
    uint32_t* group_sums;  // (num_chunks / kChunkGroup) * num_tiles; k_bin_colpass's look-back status
    uint64_t group_entries;  // entries of group_sums over all jobs
    uint32_t scan_spin_limit;  // k_bin_colpass: waits before giving up (0: give up at once)
    uint32_t legacy_scans;     // 1: column sums, group scan and column scan as three kernels
    uint32_t* tile_totals;
    uint32_t* tile_start;  // num_tiles + 1
    uint32_t* refs;
    uint32_t ref_cap;
    uint32_t* work;        // raster work items: global tile | segment << kTileBits | kSingleSegment
    uint32_t work_cap;
    uint32_t* tile_item;   // per global tile: first key image (one per segment) | segment log2 << kImageBits
    uint32_t* tile_big;    // per global tile: heavy references (big, or compact with kDescHeavy)
    uint64_t* partials;    // work_cap * samples * 1024 (depth, order) keys, one tile image per item and sample
    uint32_t raster_blocks;  // work items the segment size is chosen for
    uint32_t raster_grid;    // launched raster workgroups: the resident capacity
    uint32_t num_cus;        // compute units of the device
    uint32_t* counters;
    uint32_t* host_counters;  // pinned, coherent host copy of counters, written by the frame's last launch
    uint32_t* color;       // every job's pixels at its pix_base (resolved)
    float* depth;          // samples planes of the job's pixels (single frame), sample-major
    uint32_t* ids;
    float* color_f32;      // optional clamped float RGB (3 per pixel) for the adjoint
};

// Gradient outputs of one job (device pointers, accumulated with +=; any may
// be null).  Positions, normals and colours are per input vertex of the job's
// mesh; light is w.r.t. the light direction as given, material is (ambient,
// kd, ks, shininess).
// The per-vertex sums are accumulated in double first: jobs that share their
// vertex outputs share one accumulator of 9 x acc_v doubles (value-major) at
// acc_base, which the first of them (group_first) adds to the outputs at the
// end.
struct GradPtrs {
    float *d_px, *d_py, *d_pz, *d_nx, *d_ny, *d_nz, *d_cr, *d_cg, *d_cb;
    float* d_light;
    float* d_material;
    uint64_t acc_base;
    uint32_t acc_v;
    uint32_t group_first;
};

// Stage boundaries recorded when profiling: events[0] before the counter reset,
// then after setup+clip, bin count, bin scans, bin scatter and raster.
constexpr int kNumStages = 5;

// Host-side launchers (pipeline.hip).  events may be null (no profiling).
// With job0 non-null the reset kernel also stores *job0 as job 0 of the table
// (a single job, no separate upload); otherwise the table is already resident.
hipError_t launch_frame(const BatchParams& Q, FrameParams* jobs, const FrameParams* job0, const DeviceBuffers& B,
                        hipStream_t s, hipEvent_t* events);
// A graph captured from launch_frame with job0 non-null starts with the kernel
// that takes the job by value: set_frame_job gives that node (first, the root
// of the captured graph) another job's constants, launch layout unchanged.
hipError_t set_frame_job(hipGraphExec_t exec, hipGraphNode_t first, const BatchParams& Q, FrameParams* jobs,
                         const DeviceBuffers& B, const FrameParams& job0);
// Adjoint of the shading of a rendered batch with visibility held fixed (the
// triangle-ID buffer): d_rgb holds dLoss/d(clamped float RGB) per pixel in the
// layout of color_f32; acc is 8 doubles of scratch per job, vacc the vertex
// accumulators (vacc_size doubles, the GradPtrs' acc_base ranges).
hipError_t launch_adjoint(const BatchParams& Q, const FrameParams* jobs, const GradPtrs* grads,
                          const uint32_t* ids, const float* d_rgb, double* acc,
                          double* vacc, uint64_t vacc_size, hipStream_t s);
// Kernel attributes (dynamic LDS limits) and occupancy; once per process before launches.
hipError_t init_kernels();

// Device scratch of the allocation-free frame (edxr_render_mesh_noalloc): a
// file-scope __device__ array in pipeline.hip, one copy per device, held until
// the process ends.  Every buffer of a frame other than its inputs, outputs
// and the context's counters is carved from it per call.
constexpr size_t kScratchBytes = 64ull << 20;
constexpr int kMaxScratchDevices = 64;
// The device address of the current device's copy, looked up once per device
// (with that device current) and cached; *out is null on error.
hipError_t scratch_address(uint8_t** out);
// Workgroups of the persistent raster kernel that one CU holds at once.
uint32_t raster_resident_per_cu(bool multisample);

}  // namespace edxr
