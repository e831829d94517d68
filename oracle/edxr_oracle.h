/*
 * edxr_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a C++/SSE restatement of the
 * EDXRaster CPU pipeline (EDXRaster/Core/Renderer.cpp, Clipper.h,
 * RasterTriangle.h, Rasterizer.h, Tile.h, FrameBuffer.cpp, Shader.h).  It is
 * the checker for the HIP library and the timed CPU baseline; nothing in the
 * product library links, loads or calls it.
 *
 * Matrices are row-major float[16] with the column-vector convention
 * (out_r = sum_c m[4r+c] * v_c), as EDXUtil's Matrix::TransformPoint.
 */
#ifndef EDXR_ORACLE_H
#define EDXR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORACLE_SHADER_LAMBERT_ALBEDO = 0, /* LambertianAlbedoPixelShader, constant texture */
    ORACLE_SHADER_BLINN_PHONG_REF = 1, /* BlinnPhongPixelShader (albedo 1) */
    ORACLE_SHADER_VC_BLINN_PHONG = 2, /* vertex-colour Blinn-Phong */
    ORACLE_SHADER_TEXTURED_LAMBERT = 3 /* LambertianAlbedoPixelShader, texture slots */
};

enum {
    ORACLE_FILTER_NEAREST = 0,
    ORACLE_FILTER_LINEAR = 1,
    ORACLE_FILTER_TRILINEAR = 2,
    ORACLE_FILTER_ANISO_4X = 3, /* N:1 anisotropic, N = 4, 8, 16 */
    ORACLE_FILTER_ANISO_8X = 4,
    ORACLE_FILTER_ANISO_16X = 5
};

/* A texture slot (Texture2D<Color>, EDXUtil, absent; the declared semantics
 * of the library header, restated independently here): RGBA8 rows from v = 0,
 * or with rgba NULL a constant colour.  Parity unpinned against EDXUtil. */
typedef struct oracle_texture {
    uint32_t width, height;
    const uint8_t* rgba;
    float color[3];
} oracle_texture;

typedef struct oracle_params {
    uint32_t width, height;
    float mv[16], proj[16], raster[16];
    float light[3];        /* unnormalised, reference default (1,1,-1) */
    int32_t shader;        /* ORACLE_SHADER_* */
    float shininess;       /* integer values use exponentiation by squaring */
    float ambient, kd, ks; /* reference: 0.2, 3, 3 */
    float albedo[3];       /* constant albedo (LoadSphere texture: 0.9 white) */
    int32_t clamp_specular;/* 0 = reference (pow of the unclamped N.H) */
    int32_t hierarchical;  /* Renderer::SetHierarchicalRasterize */
    int32_t num_threads;   /* 0 = hardware concurrency */
    int32_t mutation;      /* 0; non-zero plants a known mistake (pinning tests) */
    int32_t msaa_log2;     /* Renderer::SetMSAAMode: 0 (1 sample) .. 5 (32 samples) */
    const oracle_texture* textures; /* texture slots (textured shader) */
    uint32_t num_textures;
    int32_t tex_filter;    /* ORACLE_FILTER_* */
} oracle_params;

typedef struct oracle_mesh {
    const float *px, *py, *pz;
    const float *nx, *ny, *nz;
    const float *cr, *cg, *cb;
    const uint32_t *indices; /* 3 per triangle */
    uint32_t num_vertices, num_triangles;
    const float *tu, *tv;    /* texture coordinates (textured shader) */
    const uint32_t *tex_ids; /* texture slot per triangle (textured shader) */
} oracle_mesh;

typedef struct oracle_stats {
    uint64_t triangles_in;
    uint64_t trivially_rejected;  /* c0 & c1 & c2 != 0 */
    uint64_t clipped;             /* went through ClipPolygon */
    uint64_t near_clipped;        /* straddled the near plane */
    uint64_t clip_dropped;        /* polygon removed by the w <= 0 rule */
    uint64_t setup_records;       /* RasterTriangles that survived setup */
    uint64_t culled;              /* det <= 0 */
    uint64_t refs;                /* tile references (N_ref) */
    uint64_t fragments;           /* emitted 2x2 fragments */
    uint64_t clip_vertices;       /* vertices created by clipping */
    double ms_vs, ms_clip, ms_bin, ms_raster, ms_shade, ms_write, ms_total;
    int32_t threads;
} oracle_stats;

/* One setup record, in submission order (triangle, then fan index). */
typedef struct oracle_record {
    int32_t x[3], y[3];  /* 28.4 fixed point */
    float inv_det;
    float z[3];          /* z_clip * (1/w) of the referenced vertices */
    float inv_w[3];
    uint32_t tri_id;     /* input triangle index */
    uint32_t sub;        /* fan index (0 for unclipped triangles) */
    uint32_t corners;    /* rej0|rej1<<2|rej2<<4|acc0<<6|acc1<<8|acc2<<10 */
} oracle_record;

/* One tile reference as the rasteriser consumes it, per tile in order. */
typedef struct oracle_ref {
    uint32_t record;     /* index into the oracle_record array */
    uint32_t flags;      /* bit0-2 accept edge 0..2, bit3 trivial accept, bit4 big */
} oracle_ref;

typedef struct oracle_ctx oracle_ctx;

oracle_ctx* oracle_create(void);
void oracle_destroy(oracle_ctx*);
/* Renders one mesh (Renderer::RenderMesh).  Returns 0 or a negative error. */
int oracle_render(oracle_ctx*, const oracle_params*, const oracle_mesh*, oracle_stats*);
/* Outputs in GetBackBuffer layout: row (H-1-y) holds raster row y. */
int oracle_read_color(const oracle_ctx*, uint8_t* rgba, uint64_t bytes);
int oracle_read_depth(const oracle_ctx*, float* depth, uint64_t count);
int oracle_read_ids(const oracle_ctx*, uint32_t* ids, uint64_t count);
/* Per-sample depth and winning triangle IDs, [sample][row][x] (rows as
 * above); count = samples * W * H.  With one sample these equal the above. */
int oracle_read_sample_depth(const oracle_ctx*, float* depth, uint64_t count);
int oracle_read_sample_ids(const oracle_ctx*, uint32_t* ids, uint64_t count);
uint64_t oracle_num_records(const oracle_ctx*);
int oracle_read_records(const oracle_ctx*, oracle_record*, uint64_t count);
/* Per-tile reference lists: offsets has num_tiles+1 entries. */
uint64_t oracle_num_tiles(const oracle_ctx*);
int oracle_read_bins(const oracle_ctx*, uint64_t* offsets, oracle_ref* refs, uint64_t cap);

/* Host helpers with the declared EDXUtil semantics, exposed for tests. */
void oracle_mat_mul(const float* a, const float* b, float* out);     /* a*b */
void oracle_mat_inverse(const float* m, float* out);                 /* double, rounded once */
void oracle_transform_point4(const float* m, const float* v, float* out);

#ifdef __cplusplus
}
#endif
#endif
