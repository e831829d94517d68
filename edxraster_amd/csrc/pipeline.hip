// pipeline.hip -- gfx950 kernels of the EDXRaster hot path.
//
// Stages (EDXRaster/Core/Renderer.cpp:100-118 order):
//   k_setup      vertex transform + outcodes + triangle setup, one lane per
//                triangle (DefaultVertexShader::Execute Core/Shader.h:40-49,
//                Clipper::Clip Core/Clipper.h:71-189, RasterTriangle::Setup
//                Core/RasterTriangle.h:27-153).  Straddling triangles are
//                queued for k_clip.
//   k_clip       Sutherland-Hodgman clipping + fan setup of queued triangles
//                (Core/Clipper.h:109-170,192-288).
//   k_bin_count  per-chunk tile histograms (binning, Core/Renderer.cpp:161-230)
//   k_bin_colscan / k_bin_tilescan   exclusive scans -> per-(tile, chunk) offsets
//   k_bin_scatter order-preserving scatter: refs of a tile are in ascending
//                (triangle, fan index) order, like the reference's per-core lists
//                concatenated in core order (Core/Renderer.cpp:250-253).
//   k_raster     one workgroup per 32x32 tile: hierarchical edge-function
//                coverage (CoarseRasterize / FineRasterize / TrivialAcceptTriangle,
//                Core/Rasterizer.h:32-200,310-353), depth test with the (depth,
//                submission order) winner rule, then deferred perspective-correct
//                interpolation + pixel shading of the winner and one write-back of
//                colour, depth and triangle ID.
//
// Arithmetic follows the reference operation by operation: int32 wrap
// (-fwrapv), truncating 28.4 snap, round-to-nearest int->float, correctly
// rounded division and sqrt, no FMA contraction (-ffp-contract=off).
#include <hip/hip_runtime.h>

#include "pipeline.h"

namespace edxr {

namespace {

constexpr uint32_t kWave = 64;
constexpr uint32_t kBlock = 256;

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

// Trivial reject/accept corner choice (Core/RasterTriangle.h:66-150).
__device__ __forceinline__ uint32_t corner_pair(int B, int C) {
    const float slope = (float)B / (float)C;
    uint32_t rej, acc;
    if (slope >= 0.0f) {
        if ((float)C >= 0.0f) { rej = 3; acc = 0; } else { rej = 0; acc = 3; }
    } else {
        if ((float)C >= 0.0f) { rej = 2; acc = 1; } else { rej = 1; acc = 2; }
    }
    return rej | (acc << 8);
}

// RasterTriangle::Setup on homogeneous-projected points (Core/RasterTriangle.h:27-153).
__device__ __forceinline__ bool setup_record(const FrameParams& P, float ax, float ay, float az, float bx, float by,
                                             float bz, float cx, float cy, float cz, RasterRec& r) {
    const float4 A = xform4(P.raster, ax, ay, az, 1.0f);
    const float4 Bv = xform4(P.raster, bx, by, bz, 1.0f);
    const float4 Cv = xform4(P.raster, cx, cy, cz, 1.0f);
    const int x0 = snap(A.x / A.w), y0 = snap(A.y / A.w);
    const int x1 = snap(Bv.x / Bv.w), y1 = snap(Bv.y / Bv.w);
    const int x2 = snap(Cv.x / Cv.w), y2 = snap(Cv.y / Cv.w);
    const int B0 = y0 - y1, C0 = x1 - x0, B1 = y1 - y2, C1 = x2 - x1, B2 = y2 - y0, C2 = x0 - x2;
    const int det = C2 * B1 - C1 * B2;
    if (det <= 0) return false;
    r.x0 = x0; r.y0 = y0; r.x1 = x1; r.y1 = y1; r.x2 = x2; r.y2 = y2;
    r.inv_det = 1.0f / (float)det;
    const uint32_t c0 = corner_pair(B0, C0), c1 = corner_pair(B1, C1), c2 = corner_pair(B2, C2);
    r.info = (c0 & 3u) | ((c1 & 3u) << 2) | ((c2 & 3u) << 4) | ((c0 >> 8) << 6) | ((c1 >> 8) << 8) | ((c2 >> 8) << 10);
    r.pad = 0;
    return true;
}

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

__device__ __forceinline__ uint32_t wave_count(bool p) { return (uint32_t)__popcll(__ballot(p)); }

}  // namespace

// ---------------------------------------------------------------------------
// Counters
// ---------------------------------------------------------------------------
__global__ void k_reset(uint32_t* counters) {
    if (threadIdx.x < kCntSlots) counters[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// Setup: one lane per input triangle.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_setup(FrameParams P, MeshPtrs M, DeviceBuffers B) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = t < P.num_triangles;
    bool rejected = false, queued = false, bad = false, main = false;
    uint32_t i0 = 0, i1 = 0, i2 = 0;
    RasterRec rec;
    if (valid) {
        i0 = M.idx[3 * (size_t)t];
        i1 = M.idx[3 * (size_t)t + 1];
        i2 = M.idx[3 * (size_t)t + 2];
        if (i0 >= P.num_vertices || i1 >= P.num_vertices || i2 >= P.num_vertices) {
            bad = true;
        } else {
            // DefaultVertexShader::Execute: projectedPos = MVP * (p, 1).
            const float4 c0 = xform4(P.mvp, M.px[i0], M.py[i0], M.pz[i0], 1.0f);
            const float4 c1 = xform4(P.mvp, M.px[i1], M.py[i1], M.pz[i1], 1.0f);
            const float4 c2 = xform4(P.mvp, M.px[i2], M.py[i2], M.pz[i2], 1.0f);
            const uint32_t k0 = clip_code(c0), k1 = clip_code(c1), k2 = clip_code(c2);
            if (k0 | k1 | k2) {
                if (k0 & k1 & k2) rejected = true;
                else queued = true;
            } else if (setup_record(P, c0.x / c0.w, c0.y / c0.w, c0.z / c0.w, c1.x / c1.w, c1.y / c1.w,
                                    c1.z / c1.w, c2.x / c2.w, c2.y / c2.w, c2.z / c2.w, rec)) {
                // invW and the z divide (Core/Renderer.cpp:139-147).
                rec.z0 = c0.z * (1.0f / c0.w);
                rec.z1 = c1.z * (1.0f / c1.w);
                rec.z2 = c2.z * (1.0f / c2.w);
                main = true;
            }
        }
    }
    // Wave-aggregated clip-queue append.
    const uint64_t qmask = __ballot(queued);
    uint32_t qbase = 0;
    if (qmask) {
        const int leader = __ffsll((unsigned long long)qmask) - 1;
        if ((int)__lane_id() == leader) qbase = atomicAdd(&B.counters[kCntClipQueue], (uint32_t)__popcll(qmask));
        qbase = __shfl(qbase, leader);
    }
    if (valid) {
        uint32_t d = 0;
        if (main) {
            B.rrec[t] = rec;
            B.srec[t] = ShadeRec{i0, i1, i2, t};
            d = (t << 3) | 1u;
        } else if (queued) {
            const uint32_t slot = qbase + (uint32_t)__popcll(qmask & lanemask_lt());
            if (slot < B.clip_queue_cap) B.clip_queue[slot] = t;
            else atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvClipQueue);
        }
        B.desc[t] = d;
    }
    const uint32_t nrej = wave_count(rejected), nbad = wave_count(bad), nmain = wave_count(main);
    if (__lane_id() == 0) {
        if (nrej) atomicAdd(&B.counters[kCntRejected], nrej);
        if (nbad) atomicAdd(&B.counters[kCntBadIndex], nbad);
        if (nmain) atomicAdd(&B.counters[kCntMainRecords], nmain);
    }
}

// ---------------------------------------------------------------------------
// Clipping of queued straddlers (rare path, one lane per triangle).
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

__global__ __launch_bounds__(64) void k_clip(FrameParams P, MeshPtrs M, DeviceBuffers B) {
    const uint32_t n = min(B.counters[kCntClipQueue], B.clip_queue_cap);
    for (uint32_t q = blockIdx.x * 64 + threadIdx.x; q < n; q += gridDim.x * 64) {
        const uint32_t t = B.clip_queue[q];
        const uint32_t vi[3] = {M.idx[3 * (size_t)t], M.idx[3 * (size_t)t + 1], M.idx[3 * (size_t)t + 2]};
        float4 c[3];
        for (int k = 0; k < 3; k++) c[k] = xform4(P.mvp, M.px[vi[k]], M.py[vi[k]], M.pz[vi[k]], 1.0f);
        const uint32_t k0 = clip_code(c[0]), k1 = clip_code(c[1]), k2 = clip_code(c[2]);
        const uint32_t planeCode = (k0 ^ k1) | (k1 ^ k2) | (k2 ^ k0);
        atomicAdd(&B.counters[kCntClipped], 1u);
        if ((k0 | k1 | k2) & 16u) atomicAdd(&B.counters[kCntNearClipped], 1u);

        float4 pp[2][kMaxPoly];
        float3 ww[2][kMaxPoly];
        int cnt = 3, cur = 0;
        bool ok = true;
        for (int k = 0; k < 3; k++) {
            pp[0][k] = c[k];
            ww[0][k] = make_float3(k == 0 ? 1.0f : 0.0f, k == 1 ? 1.0f : 0.0f, k == 2 ? 1.0f : 0.0f);
        }
        // Clipper::ClipByPlane (Core/Clipper.h:192-232).
        for (int p = 0; p < 6 && ok; p++) {
            if (!(planeCode & kPlaneBits[p])) continue;
            int m = 0;
            const int nx = cur ^ 1;
            for (int i = 0; i < cnt; i++) {
                int i1 = i + 1;
                if (i1 == cnt) i1 = 0;
                const float4 v0 = pp[cur][i], v1 = pp[cur][i1];
                const float3 w0 = ww[cur][i], w1 = ww[cur][i1];
                const bool in0 = plane_inside(v0, p), in1 = plane_inside(v1, p);
                if (in0 && in1) {
                    if (m >= kMaxPoly) { ok = false; break; }
                    pp[nx][m] = v1; ww[nx][m] = w1; m++;
                } else if (in0 != in1) {
                    const float tt = plane_t(v0, v1, p);
                    const float s = 1 - tt;
                    float4 pos = make_float4(v0.x * s + v1.x * tt, v0.y * s + v1.y * tt, v0.z * s + v1.z * tt,
                                             v0.w * s + v1.w * tt);
                    plane_snap(pos, p);
                    const float3 w = make_float3(w0.x * s + w1.x * tt, w0.y * s + w1.y * tt, w0.z * s + w1.z * tt);
                    if (m >= kMaxPoly) { ok = false; break; }
                    pp[nx][m] = pos; ww[nx][m] = w; m++;
                    if (!in0) {
                        if (m >= kMaxPoly) { ok = false; break; }
                        pp[nx][m] = v1; ww[nx][m] = w1; m++;
                    }
                }
            }
            cnt = m;
            cur = nx;
        }
        if (!ok || cnt > 9) {
            atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvPolygon);
            B.desc[t] = 0;
            continue;
        }
        for (int i = 0; i < cnt; i++)
            if (pp[cur][i].w <= 0.0f) { cnt = 0; break; }
        if (cnt < 3) {
            B.desc[t] = 0;
            continue;
        }
        // Vertex ids: weight == 1 reuses a parent vertex (Core/Clipper.h:121-153).
        int which[kMaxPoly];
        int nNew = 0;
        for (int j = 0; j < cnt; j++) {
            const float3 w = ww[cur][j];
            which[j] = (w.x == 1.0f) ? 0 : (w.y == 1.0f) ? 1 : (w.z == 1.0f) ? 2 : -1;
            if (which[j] < 0) nNew++;
        }
        uint32_t vbase = 0;
        if (nNew) {
            vbase = atomicAdd(&B.counters[kCntExtraVertices], (uint32_t)nNew);
            if (vbase + (uint32_t)nNew > B.extra.capacity) {
                atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvExtraVertices);
                B.desc[t] = 0;
                continue;
            }
        }
        uint32_t vid[kMaxPoly];
        float hx[kMaxPoly], hy[kMaxPoly], hz[kMaxPoly], zz[kMaxPoly];
        int e = 0;
        for (int j = 0; j < cnt; j++) {
            float4 pos;
            if (which[j] >= 0) {
                vid[j] = vi[which[j]];
                pos = c[which[j]];
            } else {
                const uint32_t xi = vbase + (uint32_t)e++;
                vid[j] = P.num_vertices + xi;
                pos = pp[cur][j];
                const float3 w = ww[cur][j];
                auto comb = [&](const float* a) {
                    return (w.x * a[vi[0]] + w.y * a[vi[1]]) + w.z * a[vi[2]];
                };
                B.extra.px[xi] = comb(M.px);
                B.extra.py[xi] = comb(M.py);
                B.extra.pz[xi] = comb(M.pz);
                B.extra.nx[xi] = M.nx ? comb(M.nx) : 0.0f;
                B.extra.ny[xi] = M.ny ? comb(M.ny) : 0.0f;
                B.extra.nz[xi] = M.nz ? comb(M.nz) : 0.0f;
                B.extra.cr[xi] = M.cr ? comb(M.cr) : 0.0f;
                B.extra.cg[xi] = M.cg ? comb(M.cg) : 0.0f;
                B.extra.cb[xi] = M.cb ? comb(M.cb) : 0.0f;
                B.extra.inv_w[xi] = 1.0f / pos.w;
            }
            hx[j] = pos.x / pos.w;
            hy[j] = pos.y / pos.w;
            hz[j] = pos.z / pos.w;
            zz[j] = pos.z * (1.0f / pos.w);
        }
        // Fan triangulation (Core/Clipper.h:155-170): count survivors, reserve, write.
        uint32_t nrec = 0;
        RasterRec r;
        for (int k = 2; k < cnt; k++)
            if (setup_record(P, hx[0], hy[0], hz[0], hx[k - 1], hy[k - 1], hz[k - 1], hx[k], hy[k], hz[k], r)) nrec++;
        if (nrec == 0) {
            B.desc[t] = 0;
            continue;
        }
        const uint32_t rbase = atomicAdd(&B.counters[kCntExtraRecords], nrec);
        if (rbase + nrec > B.extra_rec_cap) {
            atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvExtraRecords);
            B.desc[t] = 0;
            continue;
        }
        const uint32_t first = P.num_triangles + rbase;
        uint32_t s = 0;
        for (int k = 2; k < cnt; k++) {
            if (!setup_record(P, hx[0], hy[0], hz[0], hx[k - 1], hy[k - 1], hz[k - 1], hx[k], hy[k], hz[k], r)) continue;
            r.z0 = zz[0];
            r.z1 = zz[k - 1];
            r.z2 = zz[k];
            r.info |= (uint32_t)(k - 2) << 12;
            B.rrec[first + s] = r;
            B.srec[first + s] = ShadeRec{vid[0], vid[k - 1], vid[k], t};
            s++;
        }
        B.desc[t] = (first << 3) | nrec;
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

// Enumerates the (tile, ref) pairs of one slot in submission order: records in
// fan order, tiles of a record in row-major order.
struct RefEnum {
    uint32_t first, count, r;
    int k;
    bool loaded;
    BinTri b;

    __device__ __forceinline__ void init(uint32_t desc) {
        first = desc >> 3;
        count = desc & 7u;
        r = 0;
        k = 0;
        loaded = false;
    }
    // Returns false when exhausted.
    __device__ __forceinline__ bool next(const FrameParams& P, const RasterRec* rrec, uint32_t& tile,
                                         uint32_t& payload) {
        while (r < count) {
            if (!loaded) {
                load_bin_tri(P, rrec[first + r], b);
                loaded = true;
                k = 0;
            }
            const int n = b.w * b.h;
            while (k < n) {
                const int tx = b.minTX + k % b.w, ty = b.minTY + k / b.w;
                k++;
                if (b.small) {
                    tile = (uint32_t)ty * P.tiles_x + (uint32_t)tx;
                    payload = first + r;
                    return true;
                }
                const int f = bin_test(b, tx, ty);
                if (f < 0) continue;
                tile = (uint32_t)ty * P.tiles_x + (uint32_t)tx;
                payload = (first + r) | ((uint32_t)f << kRecordBits) | (1u << 31);
                return true;
            }
            r++;
            loaded = false;
        }
        return false;
    }
};

// Block-wide exclusive scan of one value per thread (256 threads).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum, uint32_t& total) {
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == kWave - 1) wsum[wave] = x;
    __syncthreads();
    uint32_t base = 0;
    total = 0;
    for (uint32_t w = 0; w < kBlock / kWave; w++) {
        if (w < wave) base += wsum[w];
        total += wsum[w];
    }
    __syncthreads();
    return base + x - v;
}

}  // namespace

__global__ __launch_bounds__(kBlock) void k_bin_count(FrameParams P, DeviceBuffers B) {
    extern __shared__ uint32_t hist[];
    const uint32_t chunk = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < P.num_tiles; i += kBlock) hist[i] = 0;
    __syncthreads();
    const uint32_t s0 = chunk * P.chunk_slots;
    const uint32_t s1 = min(P.num_triangles, s0 + P.chunk_slots);
    for (uint32_t s = s0 + threadIdx.x; s < s1; s += kBlock) {
        RefEnum e;
        e.init(B.desc[s]);
        uint32_t tile, payload;
        while (e.next(P, B.rrec, tile, payload)) atomicAdd(&hist[tile], 1u);
    }
    __syncthreads();
    uint32_t* row = B.bin_counts + (size_t)chunk * P.num_tiles;
    for (uint32_t i = threadIdx.x; i < P.num_tiles; i += kBlock) row[i] = hist[i];
}

// Column scan: for 64 tiles per block, exclusive prefix over chunks (in place)
// and the per-tile totals.
__global__ __launch_bounds__(kBlock) void k_bin_colscan(FrameParams P, DeviceBuffers B) {
    __shared__ uint32_t part[kBlock / kWave][kWave];
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    const uint32_t tile = blockIdx.x * kWave + lane;
    const bool ok = tile < P.num_tiles;
    const uint32_t C = P.num_chunks, per = (C + 3) / 4;
    const uint32_t c0 = min(C, wave * per), c1 = min(C, c0 + per);
    uint32_t sum = 0;
    if (ok)
        for (uint32_t c = c0; c < c1; c++) sum += B.bin_counts[(size_t)c * P.num_tiles + tile];
    part[wave][lane] = sum;
    __syncthreads();
    uint32_t run = 0, total = 0;
    for (uint32_t w = 0; w < kBlock / kWave; w++) {
        if (w < wave) run += part[w][lane];
        total += part[w][lane];
    }
    if (ok) {
        for (uint32_t c = c0; c < c1; c++) {
            const size_t o = (size_t)c * P.num_tiles + tile;
            const uint32_t v = B.bin_counts[o];
            B.bin_counts[o] = run;
            run += v;
        }
        if (wave == 0) B.tile_totals[tile] = total;
    }
}

// Exclusive scan of the tile totals (num_tiles <= 16384, one block of 1024).
__global__ __launch_bounds__(1024) void k_bin_tilescan(FrameParams P, DeviceBuffers B) {
    __shared__ uint32_t wsum[16];
    const uint32_t per = (P.num_tiles + 1023) / 1024;
    const uint32_t t0 = min(P.num_tiles, threadIdx.x * per), t1 = min(P.num_tiles, t0 + per);
    uint32_t local = 0;
    for (uint32_t t = t0; t < t1; t++) local += B.tile_totals[t];
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    uint32_t x = local;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == kWave - 1) wsum[wave] = x;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (uint32_t w = 0; w < 16; w++) {
        if (w < wave) base += wsum[w];
        total += wsum[w];
    }
    uint32_t run = base + x - local;
    for (uint32_t t = t0; t < t1; t++) {
        B.tile_start[t] = run;
        run += B.tile_totals[t];
    }
    if (threadIdx.x == 0) {
        B.tile_start[P.num_tiles] = total;
        B.counters[kCntTotalRefs] = total;
        if (total > B.ref_cap) atomicOr(&B.counters[kCntOverflow], (uint32_t)kOvRefs);
    }
}

// Order-preserving scatter.  Each block walks its chunk in rounds of 256
// slots; the refs of a round are enumerated in submission order in windows of
// 256, ranked per tile with wave ballots and written at the tile's cursor.
__global__ __launch_bounds__(kBlock) void k_bin_scatter(FrameParams P, DeviceBuffers B) {
    extern __shared__ uint32_t cursor[];
    __shared__ uint32_t stageTile[kBlock], stagePay[kBlock];
    __shared__ uint32_t wsum[kBlock / kWave];
    if (B.counters[kCntOverflow] & kOvRefs) return;
    const uint32_t chunk = blockIdx.x;
    const uint32_t* offs = B.bin_counts + (size_t)chunk * P.num_tiles;
    for (uint32_t i = threadIdx.x; i < P.num_tiles; i += kBlock) cursor[i] = B.tile_start[i] + offs[i];
    __syncthreads();
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    const uint32_t s0 = chunk * P.chunk_slots;
    const uint32_t s1 = min(P.num_triangles, s0 + P.chunk_slots);
    for (uint32_t rb = s0; rb < s1; rb += kBlock) {
        const uint32_t s = rb + threadIdx.x;
        const uint32_t desc = s < s1 ? B.desc[s] : 0u;
        RefEnum e;
        e.init(desc);
        uint32_t R = 0, tile, payload;
        while (e.next(P, B.rrec, tile, payload)) R++;
        uint32_t total;
        const uint32_t base = block_exclusive_scan(R, wsum, total);
        e.init(desc);
        for (uint32_t win = 0; win < total; win += kBlock) {
            // Emit this thread's refs that fall in [win, win + 256).
            const uint32_t lo = max(base, win), hi = min(base + R, win + kBlock);
            for (uint32_t pos = lo; pos < hi; pos++) {
                e.next(P, B.rrec, tile, payload);
                stageTile[pos - win] = tile;
                stagePay[pos - win] = payload;
            }
            __syncthreads();
            const bool valid = win + threadIdx.x < total;
            const uint32_t key = valid ? stageTile[threadIdx.x] : 0xFFFFFFFFu;
            const uint32_t pay = valid ? stagePay[threadIdx.x] : 0u;
            // Wave-level match: leader lane and rank among equal keys.
            uint64_t remaining = __ballot(valid);
            int myLeader = -1;
            uint32_t myRank = 0, myCnt = 0;
            while (remaining) {
                const int leader = __ffsll((unsigned long long)remaining) - 1;
                const uint32_t k = __shfl(key, leader);
                const uint64_t m = __ballot(valid && key == k);
                if (valid && key == k) {
                    myLeader = leader;
                    myRank = (uint32_t)__popcll(m & lanemask_lt());
                    myCnt = (uint32_t)__popcll(m);
                }
                remaining &= ~m;
            }
            // Waves take their turn in order so ranks follow submission order.
            uint32_t myBase = 0;
            for (uint32_t w = 0; w < kBlock / kWave; w++) {
                if (w == wave && valid && myLeader == (int)lane) {
                    myBase = cursor[key];
                    cursor[key] = myBase + myCnt;
                }
                __syncthreads();
            }
            const uint32_t bse = __shfl(myBase, myLeader < 0 ? 0 : myLeader);
            if (valid) B.refs[bse + myRank] = pay;
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Tile rasterisation + deferred shading.
// ---------------------------------------------------------------------------
namespace {

constexpr int kSmallQuads = 4;

struct RTri {
    int x0, y0, x1, y1, x2, y2;
    int B0, C0, B1, C1, B2, C2;
    int pb0, pb1, pb2;   // pixel-level top-left bias (Core/RasterTriangle.h:296-299)
    float inv_det, z0, z1, z2;
    uint32_t info;
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
}

// Ordered depth key: (depth, submission order) min = the winner of the
// sequential d <= D test (Core/FrameBuffer.cpp:54-68).  -0 and +0 compare equal.
__device__ __forceinline__ uint32_t depth_bits(float d) {
    uint32_t b = __float_as_uint(d);
    if ((b & 0x7FFFFFFFu) == 0) b = 0;
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ uint64_t bg_key() { return ((uint64_t)0xBF800000u << 32) | 0xFFFFFFFFull; }

// Barycentrics and depth at a pixel centre (Core/RasterTriangle.h:324-337).
__device__ __forceinline__ void bary_depth(const RTri& t, int px, int py, float& l0, float& l1, float& d) {
    l0 = (float)(t.B1 * (px - t.x2) + t.C1 * (py - t.y2)) * t.inv_det;
    l1 = (float)(t.B2 * (px - t.x2) + t.C2 * (py - t.y2)) * t.inv_det;
    const float l2 = 1.0f - l0 - l1;
    d = l0 * t.z0 + l1 * t.z1 + l2 * t.z2;
}

// Depth-tests the 4 pixels of the quad at (qx, qy); edges tested when useEdges.
__device__ __forceinline__ void quad_depth(const RTri& t, int qx, int qy, bool useEdges, int tminX, int tminY,
                                           uint32_t pos, uint64_t* keys) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int x = qx + (k & 1), y = qy + (k >> 1);
        const int px = x * 16 + 8, py = y * 16 + 8;
        if (useEdges) {
            const int e0 = t.B0 * (px - t.x0) + t.C0 * (py - t.y0) + t.pb0;
            const int e1 = t.B1 * (px - t.x1) + t.C1 * (py - t.y1) + t.pb1;
            const int e2 = t.B2 * (px - t.x2) + t.C2 * (py - t.y2) + t.pb2;
            if ((e0 | e1 | e2) < 0) continue;
        }
        float l0, l1, d;
        bary_depth(t, px, py, l0, l1, d);
        if (!(d <= 1.0f)) continue;   // never passes d <= D with D <= 1 (NaN included)
        const uint64_t key = ((uint64_t)depth_bits(d) << 32) | (uint64_t)(0xFFFFFFFEu - pos);
        uint64_t* slot = &keys[(y - tminY) * kTileSize + (x - tminX)];
        if (key < *slot) atomicMin((unsigned long long*)slot, (unsigned long long)key);
    }
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

__device__ __forceinline__ uint32_t to_byte(float c) {
    c = fminf(fmaxf(c, 0.0f), 1.0f);
    return (uint32_t)(int)(c * 255.0f + 0.5f);
}

struct VAttr {
    float px, py, pz, nx, ny, nz, cr, cg, cb, inv_w;
};

__device__ __forceinline__ VAttr fetch_vertex(const FrameParams& P, const MeshPtrs& M, const ExtraVerts& X,
                                              uint32_t v) {
    VAttr a;
    if (v < P.num_vertices) {
        a.px = M.px[v]; a.py = M.py[v]; a.pz = M.pz[v];
        a.nx = M.nx[v]; a.ny = M.ny[v]; a.nz = M.nz[v];
        if (M.cr) { a.cr = M.cr[v]; a.cg = M.cg[v]; a.cb = M.cb[v]; }
        else { a.cr = a.cg = a.cb = 0.0f; }
        const float w = ((P.mvp[12] * a.px + P.mvp[13] * a.py) + P.mvp[14] * a.pz) + P.mvp[15] * 1.0f;
        a.inv_w = 1.0f / w;
    } else {
        const uint32_t e = v - P.num_vertices;
        a.px = X.px[e]; a.py = X.py[e]; a.pz = X.pz[e];
        a.nx = X.nx[e]; a.ny = X.ny[e]; a.nz = X.nz[e];
        a.cr = X.cr[e]; a.cg = X.cg[e]; a.cb = X.cb[e];
        a.inv_w = X.inv_w[e];
    }
    return a;
}

// Fragment::Interpolate (Core/Shader.h:142-170) + pixel shader
// (Core/Shader.h:209-282) + Color4b::FromFloats.
__device__ __forceinline__ uint32_t shade_pixel(const FrameParams& P, float b0, float b1, const VAttr& v0,
                                                const VAttr& v1, const VAttr& v2) {
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
    return to_byte(r) | (to_byte(g) << 8) | (to_byte(b) << 16) | (255u << 24);
}

}  // namespace

__global__ __launch_bounds__(kBlock) void k_raster(FrameParams P, MeshPtrs M, DeviceBuffers B) {
    __shared__ uint64_t keys[kTileSize * kTileSize];
    __shared__ uint32_t bigq[kBlock];
    __shared__ uint32_t nbig;
    if (B.counters[kCntOverflow] & kOvRefs) return;
    const uint32_t tile = blockIdx.x;
    const int tx = (int)(tile % P.tiles_x), ty = (int)(tile / P.tiles_x);
    const int tminX = tx * kTileSize, tminY = ty * kTileSize;
    const int tmaxX = min(tminX + kTileSize, (int)P.width), tmaxY = min(tminY + kTileSize, (int)P.height);
    for (uint32_t i = threadIdx.x; i < kTileSize * kTileSize; i += kBlock) keys[i] = bg_key();
    const uint32_t start = B.tile_start[tile];
    const uint32_t n = B.tile_start[tile + 1] - start;
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    __syncthreads();

    for (uint32_t b = 0; b < n; b += kBlock) {
        if (threadIdx.x == 0) nbig = 0;
        __syncthreads();
        const uint32_t p = b + threadIdx.x;
        if (p < n) {
            const uint32_t pay = B.refs[start + p];
            const bool trivial = ((pay >> kRecordBits) & 7u) == 7u;
            const bool big = (pay >> 31) != 0;
            bool large = trivial || (big && P.hierarchical);
            if (!large) {
                RTri t;
                load_rtri(B.rrec[pay & kRecordMask], t);
                // FineRasterize_SingleSample bbox (Core/Rasterizer.h:132-140).
                int minX = max(tminX, min(t.x0, min(t.x1, t.x2)) >> 4);
                const int maxX = min(tmaxX - 1, max(t.x0, max(t.x1, t.x2)) >> 4);
                int minY = max(tminY, min(t.y0, min(t.y1, t.y2)) >> 4);
                const int maxY = min(tmaxY - 1, max(t.y0, max(t.y1, t.y2)) >> 4);
                minX -= minX % 2;
                minY -= minY % 2;
                if (maxX >= minX && maxY >= minY) {
                    const int nqx = (maxX - minX) / 2 + 1, nqy = (maxY - minY) / 2 + 1;
                    if (nqx * nqy <= kSmallQuads) {
                        for (int qy = 0; qy < nqy; qy++)
                            for (int qx = 0; qx < nqx; qx++)
                                quad_depth(t, minX + 2 * qx, minY + 2 * qy, true, tminX, tminY, p, keys);
                    } else {
                        large = true;
                    }
                }
            }
            if (large) bigq[atomicAdd(&nbig, 1u)] = p;
        }
        __syncthreads();
        const uint32_t nb = nbig;
        for (uint32_t e = wave; e < nb; e += kBlock / kWave) {
            const uint32_t pp = __builtin_amdgcn_readfirstlane(bigq[e]);
            const uint32_t pay = __builtin_amdgcn_readfirstlane(B.refs[start + pp]);
            RTri t;
            load_rtri(B.rrec[pay & kRecordMask], t);
            const bool trivial = ((pay >> kRecordBits) & 7u) == 7u;
            const bool big = (pay >> 31) != 0;
            const int lx = (int)(lane & 7u) * 2, ly = (int)(lane >> 3) * 2;
            if (trivial) {
                // TrivialAcceptTriangle over the tile (Core/Rasterizer.h:310-353).
                for (int q = 0; q < 4; q++) {
                    const int qx = tminX + 16 * (q & 1) + lx, qy = tminY + 16 * (q >> 1) + ly;
                    if (qx <= tmaxX - 1 && qy <= tmaxY - 1) quad_depth(t, qx, qy, false, tminX, tminY, pp, keys);
                }
            } else if (big && P.hierarchical) {
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
                            quad_depth(t, qx, qy, false, tminX, tminY, pp, keys);
                    } else {
                        int minX = max(bminX, fminX), minY = max(bminY, fminY);
                        const int maxX = min(bmaxX - 1, fmaxX), maxY = min(bmaxY - 1, fmaxY);
                        minX -= minX % 2;
                        minY -= minY % 2;
                        if (qx >= minX && qx <= maxX && qy >= minY && qy <= maxY)
                            quad_depth(t, qx, qy, true, tminX, tminY, pp, keys);
                    }
                }
            } else {
                // FineRasterize over the tile with the bbox clamp.
                int minX = max(tminX, min(t.x0, min(t.x1, t.x2)) >> 4);
                const int maxX = min(tmaxX - 1, max(t.x0, max(t.x1, t.x2)) >> 4);
                int minY = max(tminY, min(t.y0, min(t.y1, t.y2)) >> 4);
                const int maxY = min(tmaxY - 1, max(t.y0, max(t.y1, t.y2)) >> 4);
                minX -= minX % 2;
                minY -= minY % 2;
                for (int q = 0; q < 4; q++) {
                    const int qx = tminX + 16 * (q & 1) + lx, qy = tminY + 16 * (q >> 1) + ly;
                    if (qx >= minX && qx <= maxX && qy >= minY && qy <= maxY)
                        quad_depth(t, qx, qy, true, tminX, tminY, pp, keys);
                }
            }
        }
        __syncthreads();
    }

    // Deferred shading of the winner and one write-back per pixel.
    const int row = (int)(threadIdx.x >> 3), xg = (int)(threadIdx.x & 7u) * 4;
    const int y = tminY + row;
    if (y >= (int)P.height) return;
    const ExtraVerts X = B.extra;
    for (int i = 0; i < 4; i++) {
        const int x = tminX + xg + i;
        if (x >= (int)P.width) continue;
        const uint64_t key = keys[row * kTileSize + xg + i];
        const size_t o = (size_t)(P.height - 1 - y) * P.width + x;
        if (key == bg_key()) {
            B.color[o] = 0u;
            B.depth[o] = 1.0f;
            B.ids[o] = 0xFFFFFFFFu;
            continue;
        }
        const uint32_t pos = 0xFFFFFFFEu - (uint32_t)key;
        const uint32_t ri = B.refs[start + pos] & kRecordMask;
        RTri t;
        load_rtri(B.rrec[ri], t);
        const ShadeRec s = B.srec[ri];
        float l0, l1, d;
        bary_depth(t, x * 16 + 8, y * 16 + 8, l0, l1, d);
        const VAttr v0 = fetch_vertex(P, M, X, s.v0), v1 = fetch_vertex(P, M, X, s.v1), v2 = fetch_vertex(P, M, X, s.v2);
        B.color[o] = shade_pixel(P, l0, l1, v0, v1, v2);
        B.depth[o] = d;
        B.ids[o] = s.tri_id;
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
hipError_t launch_reset(uint32_t* counters, hipStream_t s) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, counters);
    return hipGetLastError();
}

hipError_t launch_frame(const FrameParams& P, const MeshPtrs& M, const DeviceBuffers& B, hipStream_t s,
                        hipEvent_t* events) {
    const size_t lds = sizeof(uint32_t) * P.num_tiles;
    static bool ldsAttr = false;
    if (!ldsAttr) {
        // Bin cursors for up to kMaxTiles tiles live in LDS (160 KB per CU on gfx950).
        const int maxDyn = (int)(sizeof(uint32_t) * kMaxTiles);
        (void)hipFuncSetAttribute((const void*)k_bin_count, hipFuncAttributeMaxDynamicSharedMemorySize, maxDyn);
        (void)hipFuncSetAttribute((const void*)k_bin_scatter, hipFuncAttributeMaxDynamicSharedMemorySize, maxDyn);
        ldsAttr = true;
    }
    auto mark = [&](int i) {
        if (events) (void)hipEventRecord(events[i], s);
    };
    mark(0);
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, B.counters);
    if (P.num_triangles) {
        hipLaunchKernelGGL(k_setup, dim3((P.num_triangles + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, M, B);
        const uint32_t clipBlocks = max(1u, min(1024u, (B.clip_queue_cap + 63) / 64));
        hipLaunchKernelGGL(k_clip, dim3(clipBlocks), dim3(64), 0, s, P, M, B);
    }
    mark(1);
    hipLaunchKernelGGL(k_bin_count, dim3(P.num_chunks), dim3(kBlock), lds, s, P, B);
    mark(2);
    hipLaunchKernelGGL(k_bin_colscan, dim3((P.num_tiles + kWave - 1) / kWave), dim3(kBlock), 0, s, P, B);
    hipLaunchKernelGGL(k_bin_tilescan, dim3(1), dim3(1024), 0, s, P, B);
    mark(3);
    hipLaunchKernelGGL(k_bin_scatter, dim3(P.num_chunks), dim3(kBlock), lds, s, P, B);
    mark(4);
    hipLaunchKernelGGL(k_raster, dim3(P.num_tiles), dim3(kBlock), 0, s, P, M, B);
    mark(5);
    return hipGetLastError();
}

}  // namespace edxr
