// pipeline.h -- structures shared by the host orchestration (api.cpp) and the
// gfx950 kernels (pipeline.hip).  Internal: not part of the C ABI.
#pragma once
#include <stdint.h>

#include <hip/hip_runtime_api.h>

namespace edxr {

constexpr int kTileLog2 = 5;                 // Tile::SIZE_LOG_2 (EDXRaster/Core/Tile.h:12-13)
constexpr int kTileSize = 1 << kTileLog2;    // 32 x 32 pixels
constexpr uint32_t kMaxTiles = 16384;        // LDS-resident bin cursors
constexpr uint32_t kMaxDim = 8192;
constexpr uint32_t kRecordBits = 28;         // record index bits in a tile reference
constexpr uint32_t kRecordMask = (1u << kRecordBits) - 1;
constexpr int kMaxPoly = 16;                 // clip polygon capacity (convex input needs <= 9)

enum ShaderKind { kLambertAlbedo = 0, kBlinnPhongRef = 1, kVcBlinnPhong = 2 };

// Per-frame constants, passed by value to every kernel.
struct FrameParams {
    float mvp[16];        // proj * mv, row-major, column-vector convention
    float raster[16];
    float eye[3];         // TransformPoint(0, MV^-1)
    float light[3];       // Normalize(light direction)
    float ambient, kd, ks, shininess;
    float albedo[3];
    int32_t shader;
    int32_t shin_int;     // integer exponent, or -1 for powf
    int32_t clamp_spec;
    int32_t hierarchical;
    uint32_t width, height, tiles_x, tiles_y, num_tiles;
    uint32_t num_vertices, num_triangles;
    uint32_t chunk_slots, num_chunks;
};

struct MeshPtrs {
    const float *px, *py, *pz, *nx, *ny, *nz, *cr, *cg, *cb;
    const uint32_t* idx;
};

// Vertices created by clipping (Clipper.h:136-151), SoA.
struct ExtraVerts {
    float *px, *py, *pz, *nx, *ny, *nz, *cr, *cg, *cb, *inv_w;
    uint32_t capacity;
};

// Rasterisation part of a setup record (RasterTriangle, Core/RasterTriangle.h:13-23).
struct __attribute__((aligned(16))) RasterRec {
    int32_t x0, y0, x1, y1;
    int32_t x2, y2;
    float inv_det, z0;
    float z1, z2;
    uint32_t info;        // corners (12 bits) | sub << 12
    uint32_t pad;
};
static_assert(sizeof(RasterRec) == 48, "RasterRec layout");

// Shading part: unified vertex ids (< V input, >= V clip-created) and the
// input triangle index.
struct __attribute__((aligned(16))) ShadeRec {
    uint32_t v0, v1, v2, tri_id;
};

enum CounterSlot {
    kCntClipQueue = 0,
    kCntExtraRecords,
    kCntExtraVertices,
    kCntOverflow,        // bit0 clip queue, bit1 extra records, bit2 extra vertices, bit3 refs, bit4 polygon
    kCntBadIndex,
    kCntRejected,
    kCntClipped,
    kCntNearClipped,
    kCntMainRecords,
    kCntTotalRefs,
    kCntSlots = 16
};

enum OverflowBit {
    kOvClipQueue = 1,
    kOvExtraRecords = 2,
    kOvExtraVertices = 4,
    kOvRefs = 8,
    kOvPolygon = 16
};

struct DeviceBuffers {
    RasterRec* rrec;
    ShadeRec* srec;
    uint32_t* desc;        // per input triangle: (first record << 3) | count
    uint32_t* clip_queue;
    uint32_t clip_queue_cap;
    uint32_t extra_rec_cap;
    ExtraVerts extra;
    uint32_t* bin_counts;  // num_chunks * num_tiles, scanned in place
    uint32_t* tile_totals;
    uint32_t* tile_start;  // num_tiles + 1
    uint32_t* refs;
    uint32_t ref_cap;
    uint32_t* counters;
    uint32_t* color;
    float* depth;
    uint32_t* ids;
};

// Stage boundaries recorded when profiling: events[0] before the counter reset,
// then after setup+clip, bin count, bin scans, bin scatter and raster.
constexpr int kNumStages = 5;

// Host-side launchers (pipeline.hip).  events may be null (no profiling).
hipError_t launch_frame(const FrameParams& P, const MeshPtrs& M, const DeviceBuffers& B, hipStream_t s,
                        hipEvent_t* events);
hipError_t launch_reset(uint32_t* counters, hipStream_t s);

}  // namespace edxr
