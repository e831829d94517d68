// adjoint.hip -- hand-written adjoint of shading and perspective-correct
// interpolation with visibility held fixed (no counterpart in the reference).
//
// For every covered pixel the winning input triangle comes from the forward
// pass's triangle-ID buffer; its vertices are re-read from the job's mesh and
// re-projected with the job's MVP.  The pixel's perspective-correct
// barycentrics are recomputed in continuous (unsnapped) clip space:
//   A_i = X_i - x W_i,  B_i = Y_i - y W_i,  q = A x B,  b = q / (q0 + q1 + q2)
// for the pixel centre (x, y) in NDC (the weights that put the clip-space
// point sum b_i c_i on the pixel's ray; the same weights Fragment::Interpolate,
// Core/Shader.h:142-170, forms from screen barycentrics and 1/w).  Using the
// input triangle, not a clipped piece, is exact: clipping is linear in clip
// space, so a clipped piece interpolates the same attributes.  The colour is
// the pixel shader of the forward pass (Core/Shader.h:209-282: diffuse =
// (max(N.L, 0) + ambient) kd / pi, spec = pow(N.H, s) ks) clamped to [0, 1];
// the adjoint back-propagates dL/dcolour to per-vertex positions (through
// shading and the barycentrics), normals, colours, the light direction and
// (ambient, kd, ks, shininess).
//
// Everything after the fp32 inputs is computed in double: a vertex's
// position gradient is the sum of many per-pixel terms that largely cancel
// (moving a vertex moves the barycentrics of every pixel of its triangles),
// so fp32 per-pixel arithmetic leaves relative errors of about 1e-3 at
// 512 x 512.  Per-vertex sums of a tile go to an LDS table keyed by vertex
// (double LDS atomics), then once per tile to the job's double accumulators,
// which a last pass adds to the float outputs.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pipeline.h"

namespace edxr {

namespace {

constexpr uint32_t kWave = 64;
constexpr uint32_t kAdjBlock = 256;  // threads per tile workgroup (each wave takes 4 of the 16 8x8 blocks)
constexpr double kInvPi = 1.0 / 3.14159265358979323846;

struct V3 {
    double x, y, z;
};
__device__ __forceinline__ V3 v3(double x, double y, double z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// d/dv of normalize(v) applied to the upstream gradient g (u = v / |v|).
__device__ __forceinline__ V3 normalize_back(V3 u, double invLen, V3 g) { return (g - u * dot(u, g)) * invLen; }

__device__ __forceinline__ double pow_int(double x, int n) {
    double r = 1.0, b = x;
    unsigned u = (unsigned)n;
    while (u) {
        if (u & 1u) r *= b;
        u >>= 1;
        if (u) b *= b;
    }
    return r;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

constexpr uint32_t kVertSlots = 1024;  // per-tile vertex table (open addressing)
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

// Slot of vertex v in the tile's table (inserted if absent), or kVertSlots
// when the table is full.
__device__ __forceinline__ uint32_t vertex_slot(uint32_t* keys, uint32_t v) {
    uint32_t h = (v * 2654435761u) >> 22;  // 10-bit multiplicative hash
    for (uint32_t probe = 0; probe < kVertSlots; probe++, h = (h + 1) & (kVertSlots - 1)) {
        const uint32_t k = keys[h];
        if (k == v) return h;
        if (k == kEmpty) {
            const uint32_t old = atomicCAS(&keys[h], kEmpty, v);
            if (old == kEmpty || old == v) return h;
        }
    }
    return kVertSlots;
}

__device__ __forceinline__ uint32_t job_of(const BatchParams& Q, const uint32_t* tbl, uint32_t i) {
    return Q.num_jobs == 1 ? 0u : __builtin_amdgcn_readfirstlane(tbl[i]);
}

// The adjoint of one pixel: per-vertex gradients gv[k][0..8] (position,
// normal, colour of vertex k) and the scalars gs[0..6] (light direction after
// normalisation, ambient, kd, ks, shininess).  Returns false for a pixel that
// contributes nothing.
__device__ __forceinline__ bool pixel_adjoint(const FrameParams& P,
                                              const uint32_t vi[3], int x, int y, const float* gpix,
                                              double gv[3][9], double gs[7]) {
    double X[3], Y[3], W[3];
    V3 p[3], n[3], c[3];
    double M[16];
    for (int i = 0; i < 16; i++) M[i] = (double)P.mvp[i];
    // the input vertices, read from the mesh (the forward stores shading
    // vertices only for vertices inside the frustum)
    const MeshPtrs& Mh = P.mesh;
    for (int k = 0; k < 3; k++) {
        const uint32_t v = vi[k];
        p[k] = v3(Mh.px[v], Mh.py[v], Mh.pz[v]);
        n[k] = v3(Mh.nx[v], Mh.ny[v], Mh.nz[v]);
        c[k] = Mh.cr ? v3(Mh.cr[v], Mh.cg[v], Mh.cb[v]) : v3(0.0, 0.0, 0.0);
        X[k] = ((M[0] * p[k].x + M[1] * p[k].y) + M[2] * p[k].z) + M[3];
        Y[k] = ((M[4] * p[k].x + M[5] * p[k].y) + M[6] * p[k].z) + M[7];
        W[k] = ((M[12] * p[k].x + M[13] * p[k].y) + M[14] * p[k].z) + M[15];
    }
    // Pixel centre in NDC through the inverse of the raster matrix's 2D
    // affine part (x_r = R0 . (x, y) + R3, y_r likewise).
    const float* R = P.raster;
    const double r0 = R[0], r1 = R[1], r4 = R[4], r5 = R[5];
    const double rdet = r0 * r5 - r1 * r4;
    const double xr = (double)x + 0.5 - (double)R[3], yr = (double)y + 0.5 - (double)R[7];
    const double xn = (r5 * xr - r1 * yr) / rdet, yn = (r0 * yr - r4 * xr) / rdet;
    double A[3], Bc[3];
    for (int k = 0; k < 3; k++) {
        A[k] = X[k] - xn * W[k];
        Bc[k] = Y[k] - yn * W[k];
    }
    const double q0 = A[1] * Bc[2] - A[2] * Bc[1];
    const double q1 = A[2] * Bc[0] - A[0] * Bc[2];
    const double q2 = A[0] * Bc[1] - A[1] * Bc[0];
    const double S = q0 + q1 + q2;
    if (!(S != 0.0 && isfinite(S))) return false;
    const double invS = 1.0 / S;
    const double b[3] = {q0 * invS, q1 * invS, q2 * invS};
    const V3 Pw = (p[0] * b[0] + p[1] * b[1]) + p[2] * b[2];
    const V3 Nw = (n[0] * b[0] + n[1] * b[1]) + n[2] * b[2];
    const V3 Cw = (c[0] * b[0] + c[1] * b[1]) + c[2] * b[2];
    const double invN = 1.0 / sqrt(dot(Nw, Nw));
    const V3 nh = Nw * invN;
    const V3 L = v3(P.light[0], P.light[1], P.light[2]);
    const double nl = dot(nh, L);
    const double dAmt = nl < 0.0 ? 0.0 : nl;
    const double ambient = P.ambient, kd = P.kd, ks = P.ks;
    const double diffuse = (dAmt + ambient) * kd * kInvPi;
    const bool lambert = P.shader == kLambertAlbedo;
    V3 base = P.shader == kVcBlinnPhong ? Cw : v3(1.0, 1.0, 1.0);
    if (lambert) base = v3(P.albedo[0], P.albedo[1], P.albedo[2]);
    double spec = 0.0, sa = 0.0, sk = 0.0, dsk = 0.0, invV = 0.0, invH = 0.0;
    V3 vh = v3(0, 0, 0), hh = v3(0, 0, 0);
    if (!lambert) {
        const V3 ev = v3(P.eye[0], P.eye[1], P.eye[2]) - Pw;
        invV = 1.0 / sqrt(dot(ev, ev));
        vh = ev * invV;
        const V3 h = L + vh;
        invH = 1.0 / sqrt(dot(h, h));
        hh = h * invH;
        sa = dot(nh, hh);
        if (P.clamp_spec) sa = fmax(sa, 0.0);
        if (P.shin_int >= 0) {
            sk = pow_int(sa, P.shin_int);
            dsk = P.shin_int > 0 ? (double)P.shin_int * pow_int(sa, P.shin_int - 1) : 0.0;
        } else {
            const double s = P.shininess;
            sk = pow(sa, s);
            dsk = s * pow(sa, s - 1.0);
        }
        spec = sk * ks;
    }
    const double col[3] = {diffuse * base.x + spec, diffuse * base.y + spec, diffuse * base.z + spec};
    double g[3];
    for (int ch = 0; ch < 3; ch++) {
        g[ch] = gpix[ch];
        if (!(col[ch] > 0.0 && col[ch] < 1.0)) g[ch] = 0.0;  // clamp
    }
    // colour = diffuse * base + spec
    const double dDiffuse = (g[0] * base.x + g[1] * base.y) + g[2] * base.z;
    const double dSpec = lambert ? 0.0 : (g[0] + g[1]) + g[2];
    const V3 dC = P.shader == kVcBlinnPhong ? v3(g[0] * diffuse, g[1] * diffuse, g[2] * diffuse) : v3(0.0, 0.0, 0.0);
    // diffuse = (max(n.L, 0) + ambient) kd / pi
    gs[3] = dDiffuse * kd * kInvPi;
    gs[4] = dDiffuse * (dAmt + ambient) * kInvPi;
    const double dnl = nl > 0.0 ? dDiffuse * kd * kInvPi : 0.0;
    V3 dnh = L * dnl;
    V3 dL = nh * dnl;
    V3 dP = v3(0.0, 0.0, 0.0);
    gs[5] = gs[6] = 0.0;
    if (!lambert) {
        // spec = pow(sa, s) ks
        gs[5] = dSpec * sk;
        gs[6] = sa > 0.0 ? dSpec * ks * sk * log(sa) : 0.0;
        double dsa = dSpec * ks * dsk;
        if (P.clamp_spec && !(dot(nh, hh) > 0.0)) dsa = 0.0;
        dnh = dnh + hh * dsa;
        const V3 dh = normalize_back(hh, invH, nh * dsa);  // h = L + v
        dL = dL + dh;
        dP = v3(0.0, 0.0, 0.0) - normalize_back(vh, invV, dh);  // v = normalize(eye - P)
    }
    gs[0] = dL.x;
    gs[1] = dL.y;
    gs[2] = dL.z;
    const V3 dN = normalize_back(nh, invN, dnh);
    // Interpolated attributes: d(attr_k) = b_k d(attr); d b_k = d(attr) . attr_k
    double db[3];
    for (int k = 0; k < 3; k++) {
        const V3 a = dP * b[k], bn = dN * b[k], bc = dC * b[k];
        gv[k][0] = a.x; gv[k][1] = a.y; gv[k][2] = a.z;
        gv[k][3] = bn.x; gv[k][4] = bn.y; gv[k][5] = bn.z;
        gv[k][6] = bc.x; gv[k][7] = bc.y; gv[k][8] = bc.z;
        db[k] = (dot(dP, p[k]) + dot(dN, n[k])) + dot(dC, c[k]);
    }
    // b = q / S
    const double bdb = (b[0] * db[0] + b[1] * db[1]) + b[2] * db[2];
    const double dq[3] = {(db[0] - bdb) * invS, (db[1] - bdb) * invS, (db[2] - bdb) * invS};
    // q = A x B
    const double dA[3] = {Bc[1] * dq[2] - Bc[2] * dq[1], Bc[2] * dq[0] - Bc[0] * dq[2], Bc[0] * dq[1] - Bc[1] * dq[0]};
    const double dB[3] = {A[2] * dq[1] - A[1] * dq[2], A[0] * dq[2] - A[2] * dq[0], A[1] * dq[0] - A[0] * dq[1]};
    // A = X - x W, B = Y - y W, (X, Y, W) = MVP rows 0, 1, 3 . (p, 1)
    for (int k = 0; k < 3; k++) {
        const double dX = dA[k], dY = dB[k], dW = -xn * dA[k] - yn * dB[k];
        gv[k][0] += (M[0] * dX + M[4] * dY) + M[12] * dW;
        gv[k][1] += (M[1] * dX + M[5] * dY) + M[13] * dW;
        gv[k][2] += (M[2] * dX + M[6] * dY) + M[14] * dW;
    }
    return true;
}

}  // namespace

__global__ void k_adjoint_zero(double* acc, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        acc[i] = 0.0;
}

// One workgroup per 32x32 tile, each wave over four of its 8x8 blocks (lane = pixel).
__global__ __launch_bounds__(kAdjBlock) void k_adjoint(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                       const GradPtrs* __restrict__ grads,
                                                       const uint32_t* __restrict__ ids,
                                                       const float* __restrict__ d_rgb, double* acc, double* vacc) {
    __shared__ double wsum[kAdjBlock / kWave][8];
    __shared__ uint32_t vkeys[kVertSlots];
    __shared__ double vsum[9][kVertSlots];
    for (uint32_t i = threadIdx.x; i < kVertSlots; i += blockDim.x) {
        vkeys[i] = kEmpty;
        for (int j = 0; j < 9; j++) vsum[j][i] = 0.0;
    }
    __syncthreads();
    const uint32_t gtile = blockIdx.x;
    const uint32_t job = job_of(Q, Q.tile_job, gtile);
    const FrameParams& P = jobs[job];
    const GradPtrs& G = grads[job];
    const uint32_t tile = gtile - P.tile_base;
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    double gsum[7] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t blk = wave; blk < 16; blk += kAdjBlock / kWave) {
        const int x = (int)(tile % P.tiles_x) * kTileSize + (int)(blk & 3u) * 8 + (int)(lane & 7u);
        const int y = (int)(tile / P.tiles_x) * kTileSize + (int)(blk >> 2) * 8 + (int)(lane >> 3);
        const bool inside = x < (int)P.width && y < (int)P.height;
        const size_t o = P.pix_base + (size_t)(P.height - 1 - (inside ? y : 0)) * P.width + (inside ? x : 0);
        const uint32_t id = inside ? ids[o] : 0xFFFFFFFFu;
        if (id == 0xFFFFFFFFu) continue;
        const uint32_t* idx = P.mesh.idx + 3 * (size_t)id;
        const uint32_t vi[3] = {idx[0], idx[1], idx[2]};
        double gv[3][9], gs[7];
        if (!pixel_adjoint(P, vi, x, y, d_rgb + 3 * o, gv, gs)) continue;
        for (int j = 0; j < 7; j++) gsum[j] += gs[j];
        // Per-vertex sums: LDS atomics into the tile's vertex table; a vertex
        // the full table cannot take goes straight to the accumulators.
        for (int k = 0; k < 3; k++) {
            const uint32_t slot = vertex_slot(vkeys, vi[k]);
            for (int j = 0; j < 9; j++) {
                if (slot < kVertSlots) atomicAdd(&vsum[j][slot], gv[k][j]);
                else atomicAdd(&vacc[G.acc_base + (uint64_t)j * G.acc_v + vi[k]], gv[k][j]);
            }
        }
    }
    // Scalars: wave sums, then one double atomic per workgroup and value.
    for (int j = 0; j < 7; j++) {
        const double s = wave_sum(gsum[j]);
        if (lane == 0) wsum[wave][j] = s;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        double t = 0.0;
        for (uint32_t w = 0; w < kAdjBlock / kWave; w++) t += wsum[w][threadIdx.x];
        if (t != 0.0) atomicAdd(&acc[8 * job + threadIdx.x], t);
    }
    for (uint32_t i = threadIdx.x; i < kVertSlots; i += blockDim.x) {
        const uint32_t v = vkeys[i];
        if (v == kEmpty) continue;
        for (int j = 0; j < 9; j++)
            if (vsum[j][i] != 0.0) atomicAdd(&vacc[G.acc_base + (uint64_t)j * G.acc_v + v], vsum[j][i]);
    }
}

// Per job: the light gradient through Normalize (L = l / |l|) and the
// material gradients, added to the caller's arrays.
__global__ void k_adjoint_finish(BatchParams Q, const FrameParams* __restrict__ jobs,
                                 const GradPtrs* __restrict__ grads, const double* acc) {
    const uint32_t job = blockIdx.x * blockDim.x + threadIdx.x;
    if (job >= Q.num_jobs) return;
    const FrameParams& P = jobs[job];
    const GradPtrs& G = grads[job];
    const double* a = acc + 8 * job;
    if (G.d_light) {
        const double L[3] = {P.light[0], P.light[1], P.light[2]};
        const double LdL = L[0] * a[0] + L[1] * a[1] + L[2] * a[2];
        for (int k = 0; k < 3; k++) atomicAdd(&G.d_light[k], (float)((a[k] - L[k] * LdL) / (double)P.light_len));
    }
    if (G.d_material)
        for (int k = 0; k < 4; k++) atomicAdd(&G.d_material[k], (float)a[3 + k]);
}

// Per group of jobs sharing vertex outputs (its first job, blockIdx.x): the
// double accumulators rounded once and added to the float outputs.
__global__ void k_adjoint_vertices(const GradPtrs* __restrict__ grads, const double* vacc) {
    const GradPtrs& G = grads[blockIdx.x];
    if (!G.group_first) return;
    float* const outs[9] = {G.d_px, G.d_py, G.d_pz, G.d_nx, G.d_ny, G.d_nz, G.d_cr, G.d_cg, G.d_cb};
    const uint64_t n = 9ull * G.acc_v;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t j = (uint32_t)(i / G.acc_v), v = (uint32_t)(i % G.acc_v);
        const double a = vacc[G.acc_base + i];
        if (outs[j] && a != 0.0) atomicAdd(outs[j] + v, (float)a);
    }
}

hipError_t launch_adjoint(const BatchParams& Q, const FrameParams* jobs, const GradPtrs* grads,
                          const uint32_t* ids, const float* d_rgb, double* acc,
                          double* vacc, uint64_t vacc_size, hipStream_t s) {
    const uint32_t nacc = 8 * Q.num_jobs;
    hipLaunchKernelGGL(k_adjoint_zero, dim3((nacc + 255) / 256), dim3(256), 0, s, acc, (uint64_t)nacc);
    if (vacc_size)
        hipLaunchKernelGGL(k_adjoint_zero, dim3((uint32_t)std::min<uint64_t>((vacc_size + 255) / 256, 4096)), dim3(256),
                           0, s, vacc, vacc_size);
    if (Q.total_tiles)
        hipLaunchKernelGGL(k_adjoint, dim3(Q.total_tiles), dim3(kAdjBlock), 0, s, Q, jobs, grads, ids, d_rgb,
                           acc, vacc);
    hipLaunchKernelGGL(k_adjoint_finish, dim3((Q.num_jobs + 63) / 64), dim3(64), 0, s, Q, jobs, grads, acc);
    hipLaunchKernelGGL(k_adjoint_vertices, dim3(Q.num_jobs), dim3(256), 0, s, grads, vacc);
    return hipGetLastError();
}

}  // namespace edxr
