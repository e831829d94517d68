// adjoint.hip -- hand-written adjoint of shading and perspective-correct
// interpolation with visibility held fixed (no counterpart in the reference).
//
// For every covered pixel the winning input triangle comes from the forward
// pass's triangle-ID buffer; its vertices are re-read from the shading-vertex
// buffer and re-projected with the job's MVP.  The pixel's perspective-correct
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
// (ambient, kd, ks, shininess).  Per-vertex sums of a tile go to an LDS table
// keyed by vertex (LDS float atomics), which is added to the outputs once per
// tile; scalar sums per workgroup into doubles.
#include <hip/hip_runtime.h>

#include "pipeline.h"

namespace edxr {

namespace {

constexpr uint32_t kWave = 64;
constexpr float kInvPi = (float)(1.0 / 3.14159265358979323846);

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// d/dv of normalize(v) applied to the upstream gradient g (u = v / |v|).
__device__ __forceinline__ V3 normalize_back(V3 u, float invLen, V3 g) { return (g - u * dot(u, g)) * invLen; }

__device__ __forceinline__ float pow_int(float x, int n) {
    float r = 1.0f, b = x;
    unsigned u = (unsigned)n;
    while (u) {
        if (u & 1u) r *= b;
        u >>= 1;
        if (u) b *= b;
    }
    return r;
}

__device__ __forceinline__ float wave_sum(float v) {
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

}  // namespace

__global__ void k_adjoint_zero(double* acc, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) acc[i] = 0.0;
}

// One workgroup per 32x32 tile, one wave per 8x8 block (lane = pixel).
__global__ __launch_bounds__(1024) void k_adjoint(BatchParams Q, const FrameParams* __restrict__ jobs,
                                                  const GradPtrs* __restrict__ grads,
                                                  const ShadeVert* __restrict__ verts,
                                                  const uint32_t* __restrict__ ids, const float* __restrict__ d_rgb,
                                                  double* acc) {
    __shared__ float wsum[16][8];
    __shared__ uint32_t vkeys[kVertSlots];
    __shared__ float vsum[9][kVertSlots];
    for (uint32_t i = threadIdx.x; i < kVertSlots; i += blockDim.x) {
        vkeys[i] = kEmpty;
        for (int j = 0; j < 9; j++) vsum[j][i] = 0.0f;
    }
    __syncthreads();
    const uint32_t gtile = blockIdx.x;
    const uint32_t job = job_of(Q, Q.tile_job, gtile);
    const FrameParams& P = jobs[job];
    const uint32_t tile = gtile - P.tile_base;
    const uint32_t lane = __lane_id(), wave = threadIdx.x / kWave;
    const int x = (int)(tile % P.tiles_x) * kTileSize + (int)(wave & 3u) * 8 + (int)(lane & 7u);
    const int y = (int)(tile / P.tiles_x) * kTileSize + (int)(wave >> 2) * 8 + (int)(lane >> 3);
    const bool inside = x < (int)P.width && y < (int)P.height;
    const size_t o = P.pix_base + (size_t)(P.height - 1 - (inside ? y : 0)) * P.width + (inside ? x : 0);
    const uint32_t id = inside ? ids[o] : 0xFFFFFFFFu;
    bool has = id != 0xFFFFFFFFu;

    // Per-lane results: 9 values per vertex (position, normal, colour) and
    // the scalars (light direction after normalisation, ambient, kd, ks, s).
    float gv[3][9];
    float gs[7];
    for (int k = 0; k < 3; k++)
        for (int j = 0; j < 9; j++) gv[k][j] = 0.0f;
    for (int j = 0; j < 7; j++) gs[j] = 0.0f;
    uint32_t vi[3] = {0, 0, 0};

    if (has) {
        const uint32_t* idx = P.mesh.idx + 3 * (size_t)id;
        vi[0] = idx[0];
        vi[1] = idx[1];
        vi[2] = idx[2];
        ShadeVert sv[3];
        float X[3], Y[3], W[3];
        for (int k = 0; k < 3; k++) {
            sv[k] = verts[P.vert_base + vi[k]];
            X[k] = ((P.mvp[0] * sv[k].px + P.mvp[1] * sv[k].py) + P.mvp[2] * sv[k].pz) + P.mvp[3];
            Y[k] = ((P.mvp[4] * sv[k].px + P.mvp[5] * sv[k].py) + P.mvp[6] * sv[k].pz) + P.mvp[7];
            W[k] = ((P.mvp[12] * sv[k].px + P.mvp[13] * sv[k].py) + P.mvp[14] * sv[k].pz) + P.mvp[15];
        }
        // Pixel centre in NDC through the inverse of the raster matrix's 2D
        // affine part (x_r = R0 . (x, y) + R3, y_r likewise).
        const float* R = P.raster;
        const float rdet = R[0] * R[5] - R[1] * R[4];
        const float xr = (float)x + 0.5f - R[3], yr = (float)y + 0.5f - R[7];
        const float xn = (R[5] * xr - R[1] * yr) / rdet, yn = (R[0] * yr - R[4] * xr) / rdet;
        float A[3], Bc[3];
        for (int k = 0; k < 3; k++) {
            A[k] = X[k] - xn * W[k];
            Bc[k] = Y[k] - yn * W[k];
        }
        const float q0 = A[1] * Bc[2] - A[2] * Bc[1];
        const float q1 = A[2] * Bc[0] - A[0] * Bc[2];
        const float q2 = A[0] * Bc[1] - A[1] * Bc[0];
        const float S = q0 + q1 + q2;
        has = S != 0.0f && isfinite(S);
        if (has) {
            const float invS = 1.0f / S;
            const float b[3] = {q0 * invS, q1 * invS, q2 * invS};
            V3 p[3], n[3], c[3];
            for (int k = 0; k < 3; k++) {
                p[k] = v3(sv[k].px, sv[k].py, sv[k].pz);
                n[k] = v3(sv[k].nx, sv[k].ny, sv[k].nz);
                c[k] = v3(sv[k].cr, sv[k].cg, sv[k].cb);
            }
            const V3 Pw = (p[0] * b[0] + p[1] * b[1]) + p[2] * b[2];
            const V3 Nw = (n[0] * b[0] + n[1] * b[1]) + n[2] * b[2];
            const V3 Cw = (c[0] * b[0] + c[1] * b[1]) + c[2] * b[2];
            const float invN = 1.0f / sqrtf(dot(Nw, Nw));
            const V3 nh = Nw * invN;
            const V3 L = v3(P.light[0], P.light[1], P.light[2]);
            const float nl = dot(nh, L);
            const float dAmt = nl < 0.0f ? 0.0f : nl;
            const float diffuse = (dAmt + P.ambient) * P.kd * kInvPi;
            const bool lambert = P.shader == kLambertAlbedo;
            V3 base = P.shader == kVcBlinnPhong ? Cw : v3(1.0f, 1.0f, 1.0f);
            if (lambert) base = v3(P.albedo[0], P.albedo[1], P.albedo[2]);
            float spec = 0.0f, sa = 0.0f, sk = 0.0f, dsk = 0.0f, invV = 0.0f, invH = 0.0f;
            V3 vh = v3(0, 0, 0), hh = v3(0, 0, 0);
            if (!lambert) {
                const V3 ev = v3(P.eye[0], P.eye[1], P.eye[2]) - Pw;
                invV = 1.0f / sqrtf(dot(ev, ev));
                vh = ev * invV;
                const V3 h = L + vh;
                invH = 1.0f / sqrtf(dot(h, h));
                hh = h * invH;
                sa = dot(nh, hh);
                if (P.clamp_spec) sa = fmaxf(sa, 0.0f);
                if (P.shin_int >= 0) {
                    sk = pow_int(sa, P.shin_int);
                    dsk = P.shin_int > 0 ? (float)P.shin_int * pow_int(sa, P.shin_int - 1) : 0.0f;
                } else {
                    sk = powf(sa, P.shininess);
                    dsk = P.shininess * powf(sa, P.shininess - 1.0f);
                }
                spec = sk * P.ks;
            }
            const float col[3] = {diffuse * base.x + spec, diffuse * base.y + spec, diffuse * base.z + spec};
            float g[3];
            for (int ch = 0; ch < 3; ch++) {
                g[ch] = d_rgb[3 * o + ch];
                if (!(col[ch] > 0.0f && col[ch] < 1.0f)) g[ch] = 0.0f;  // clamp
            }
            // colour = diffuse * base + spec
            const float dDiffuse = (g[0] * base.x + g[1] * base.y) + g[2] * base.z;
            const float dSpec = lambert ? 0.0f : (g[0] + g[1]) + g[2];
            const V3 dC = P.shader == kVcBlinnPhong ? v3(g[0] * diffuse, g[1] * diffuse, g[2] * diffuse)
                                                   : v3(0.0f, 0.0f, 0.0f);
            // diffuse = (max(n.L, 0) + ambient) kd / pi
            gs[3] = dDiffuse * P.kd * kInvPi;
            gs[4] = dDiffuse * (dAmt + P.ambient) * kInvPi;
            const float dnl = nl > 0.0f ? dDiffuse * P.kd * kInvPi : 0.0f;
            V3 dnh = L * dnl;
            V3 dL = nh * dnl;
            V3 dP = v3(0.0f, 0.0f, 0.0f);
            if (!lambert) {
                // spec = pow(sa, s) ks
                gs[5] = dSpec * sk;
                gs[6] = sa > 0.0f ? dSpec * P.ks * sk * logf(sa) : 0.0f;
                float dsa = dSpec * P.ks * dsk;
                if (P.clamp_spec && !(dot(nh, hh) > 0.0f)) dsa = 0.0f;
                dnh = dnh + hh * dsa;
                const V3 dh = normalize_back(hh, invH, nh * dsa);  // h = L + v
                dL = dL + dh;
                dP = v3(0.0f, 0.0f, 0.0f) - normalize_back(vh, invV, dh);  // v = normalize(eye - P)
            }
            gs[0] = dL.x;
            gs[1] = dL.y;
            gs[2] = dL.z;
            const V3 dN = normalize_back(nh, invN, dnh);
            // Interpolated attributes: d(attr_k) = b_k d(attr); d b_k = d(attr) . attr_k
            float db[3];
            for (int k = 0; k < 3; k++) {
                const V3 a = dP * b[k], bn = dN * b[k], bc = dC * b[k];
                gv[k][0] = a.x; gv[k][1] = a.y; gv[k][2] = a.z;
                gv[k][3] = bn.x; gv[k][4] = bn.y; gv[k][5] = bn.z;
                gv[k][6] = bc.x; gv[k][7] = bc.y; gv[k][8] = bc.z;
                db[k] = (dot(dP, p[k]) + dot(dN, n[k])) + dot(dC, c[k]);
            }
            // b = q / S
            const float bdb = (b[0] * db[0] + b[1] * db[1]) + b[2] * db[2];
            const float dq[3] = {(db[0] - bdb) * invS, (db[1] - bdb) * invS, (db[2] - bdb) * invS};
            // q = A x B
            const float dA[3] = {Bc[1] * dq[2] - Bc[2] * dq[1], Bc[2] * dq[0] - Bc[0] * dq[2],
                                 Bc[0] * dq[1] - Bc[1] * dq[0]};
            const float dB[3] = {A[2] * dq[1] - A[1] * dq[2], A[0] * dq[2] - A[2] * dq[0],
                                 A[1] * dq[0] - A[0] * dq[1]};
            // A = X - x W, B = Y - y W, (X, Y, W) = MVP rows 0, 1, 3 . (p, 1)
            for (int k = 0; k < 3; k++) {
                const float dX = dA[k], dY = dB[k], dW = -xn * dA[k] - yn * dB[k];
                gv[k][0] += (P.mvp[0] * dX + P.mvp[4] * dY) + P.mvp[12] * dW;
                gv[k][1] += (P.mvp[1] * dX + P.mvp[5] * dY) + P.mvp[13] * dW;
                gv[k][2] += (P.mvp[2] * dX + P.mvp[6] * dY) + P.mvp[14] * dW;
            }
        }
    }

    // Per-vertex sums: LDS atomics into the tile's vertex table; a vertex the
    // full table cannot take goes straight to the outputs.
    const GradPtrs& G = grads[job];
    float* const outs[9] = {G.d_px, G.d_py, G.d_pz, G.d_nx, G.d_ny, G.d_nz, G.d_cr, G.d_cg, G.d_cb};
    if (has) {
        for (int k = 0; k < 3; k++) {
            const uint32_t slot = vertex_slot(vkeys, vi[k]);
            for (int j = 0; j < 9; j++) {
                if (slot < kVertSlots) atomicAdd(&vsum[j][slot], gv[k][j]);
                else if (outs[j]) atomicAdd(outs[j] + vi[k], gv[k][j]);
            }
        }
    }
    // Scalars: wave sums, then one double atomic per workgroup and value.
    for (int j = 0; j < 7; j++) {
        const float s = wave_sum(has ? gs[j] : 0.0f);
        if (lane == 0) wsum[wave][j] = s;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        double t = 0.0;
        for (int w = 0; w < 16; w++) t += (double)wsum[w][threadIdx.x];
        if (t != 0.0) atomicAdd(&acc[8 * job + threadIdx.x], t);
    }
    for (uint32_t i = threadIdx.x; i < kVertSlots; i += blockDim.x) {
        const uint32_t v = vkeys[i];
        if (v == kEmpty) continue;
        for (int j = 0; j < 9; j++)
            if (outs[j] && vsum[j][i] != 0.0f) atomicAdd(outs[j] + v, vsum[j][i]);
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

hipError_t launch_adjoint(const BatchParams& Q, const FrameParams* jobs, const GradPtrs* grads,
                          const ShadeVert* verts, const uint32_t* ids, const float* d_rgb, double* acc,
                          hipStream_t s) {
    const uint32_t nacc = 8 * Q.num_jobs;
    hipLaunchKernelGGL(k_adjoint_zero, dim3((nacc + 255) / 256), dim3(256), 0, s, acc, nacc);
    if (Q.total_tiles)
        hipLaunchKernelGGL(k_adjoint, dim3(Q.total_tiles), dim3(1024), 0, s, Q, jobs, grads, verts, ids, d_rgb, acc);
    hipLaunchKernelGGL(k_adjoint_finish, dim3((Q.num_jobs + 63) / 64), dim3(64), 0, s, Q, jobs, grads, acc);
    return hipGetLastError();
}

}  // namespace edxr
