/*
 * edxraster.h -- C ABI of the MI355X (gfx950) EDXRaster pipeline.
 *
 * Each entry point replaces one member of the reference's C++ renderer API
 * (EDXRaster/Core/Renderer.h:37-50 and the shader objects of Core/Shader.h);
 * the replaced member is named in the comment above each declaration.
 *
 * Conventions
 *   - Status: 0 = OK, negative = error (EDXR_E_*); edxr_last_error() gives text.
 *     No C++ exception crosses this boundary.
 *   - Matrices are row-major float[16], column-vector convention
 *     (out_r = sum_c m[4r+c] * v_c), the convention of EDXUtil's
 *     Matrix::TransformPoint used by the reference.
 *   - Output images use GetBackBuffer's layout (Core/FrameBuffer.h:55-58):
 *     W x H, row (H-1-y) holds raster row y (y grows downwards on screen).
 *   - One context = one renderer = one HIP stream on the current device.  The
 *     context owns its device memory; callers own every host array they pass.
 *     Calls on one context are serialised; separate contexts are independent.
 *   - There is no CPU fallback: every render call runs HIP kernels and fails
 *     with EDXR_E_HIP when no usable gfx950 device is present.
 */
#ifndef EDXRASTER_H
#define EDXRASTER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EDXR_VERSION_MAJOR 0
#define EDXR_VERSION_MINOR 1

enum edxr_status {
    EDXR_OK = 0,
    EDXR_E_INVALID = -1,   /* bad argument */
    EDXR_E_HIP = -2,       /* HIP runtime error or no device */
    EDXR_E_NOMEM = -3,     /* device allocation failed */
    EDXR_E_INDEX = -4,     /* an index >= num_vertices */
    EDXR_E_LIMIT = -5,     /* size beyond a documented limit */
    EDXR_E_STATE = -6,     /* call out of order (e.g. backward before forward) */
    EDXR_E_IO = -7,        /* a file could not be read or written */
    EDXR_E_FORMAT = -8     /* malformed input file */
};

enum edxr_shader_kind {
    /* LambertianAlbedoPixelShader (Core/Shader.h:209-244) with a constant
     * texture, as Mesh::LoadSphere/LoadPlane bind (Utils/Mesh.cpp:44,66). */
    EDXR_SHADER_LAMBERT_ALBEDO = 0,
    /* BlinnPhongPixelShader (Core/Shader.h:246-282): albedo 1, shininess 200. */
    EDXR_SHADER_BLINN_PHONG_REF = 1,
    /* Vertex-colour Blinn-Phong: diffuse * interpolated RGB + specular. */
    EDXR_SHADER_VC_BLINN_PHONG = 2,
    /* LambertianAlbedoPixelShader (Core/Shader.h:209-244) with the albedo
     * sampled from the triangle's texture slot (edxr_set_textures) at the
     * interpolated texture coordinate, with the 2x2 quad's differentials and
     * the EDXR_STATE_TEXTURE_FILTER filter.  The mesh needs tu, tv, tex_ids. */
    EDXR_SHADER_TEXTURED_LAMBERT = 3
};

/* Texture filters (RenderStates::TexFilter, Core/RenderStates.h:22). */
enum edxr_texture_filter {
    EDXR_FILTER_NEAREST = 0,
    EDXR_FILTER_LINEAR = 1,     /* bilinear, level 0 */
    EDXR_FILTER_TRILINEAR = 2,  /* default (RenderStates::DefaultSettings):
                                 * lod = log2 of the longer texel-space
                                 * differential, two bilinear levels blended */
    /* N:1 anisotropic, N = 4, 8, 16 (the viewer's filter choices,
     * RealtimeViewer/Main.cpp:99-108): N trilinear taps spread evenly along
     * the longer texel-space differential (tap k at offset (2k + 1 - N) / 2N
     * of its uv vector), each at lod = log2 of the shorter differential's
     * length raised to at least 1/N of the longer, averaged in tap order.
     * EDXUtil's filter is not available: this is the library's declaration. */
    EDXR_FILTER_ANISO_4X = 3,
    EDXR_FILTER_ANISO_8X = 4,
    EDXR_FILTER_ANISO_16X = 5
};

/* Pixel-shader parameters.  The reference values are ambient 0.2, kd 3,
 * ks 3, shininess 200 (Core/Shader.h:264,275-278). */
typedef struct edxr_material {
    float ambient;
    float kd;
    float ks;
    float shininess;     /* integer values use exponentiation by squaring */
    float albedo[3];     /* constant albedo of EDXR_SHADER_LAMBERT_ALBEDO */
    int32_t clamp_specular; /* 0 = reference (pow of the unclamped N.H) */
} edxr_material;

enum edxr_state_key {
    /* Renderer::SetHierarchicalRasterize (Core/Renderer.h:49); default 1. */
    EDXR_STATE_HIERARCHICAL = 1,
    /* Bin chunks used by the order-preserving binning (0 = automatic). */
    EDXR_STATE_BIN_CHUNKS = 2,
    /* 1 = record HIP events around each pipeline stage of every frame. */
    EDXR_STATE_PROFILE = 3,
    /* 1 = capture each frame's launches in a HIP graph and replay it while the
     * launch layout (sizes, capacities, tuning) repeats; default 0. */
    EDXR_STATE_GRAPH = 4,
    /* Triangles per binning chunk (0 = automatic, 1024). */
    EDXR_STATE_TRIS_PER_CHUNK = 5,
    /* Raster work items per compute unit that the tile-segment size aims for
     * (0 = automatic, 8); the raster grid itself is always as many
     * workgroups as the GPU holds at once. */
    EDXR_STATE_RASTER_BLOCKS_PER_CU = 6,
    /* 1 = also write the clamped float RGB image (3 floats per pixel, the
     * function edxr_backward differentiates); default 0. */
    EDXR_STATE_FLOAT_COLOR = 7,
    /* Renderer::SetMSAAMode (Core/Renderer.cpp:94-98): log2 of the samples per
     * pixel, 0 (1 sample) .. 5 (32 samples); default 0.  Sample positions are
     * the reference's table (Core/FrameBuffer.cpp:107-191); the colour buffer
     * is resolved, depth and IDs are kept per sample. */
    EDXR_STATE_MSAA = 8,
    /* Renderer::SetWriteFrames (Core/Renderer.h:50): after every render_mesh
     * write <frame directory>/Frames/Frame%05i.bmp, numbered by the frame
     * counter (Core/Renderer.cpp:114-117, 352-358).  Default off. */
    EDXR_STATE_WRITE_FRAMES = 9,
    /* Renderer::SetTextureFilter: EDXR_FILTER_*; default trilinear. */
    EDXR_STATE_TEXTURE_FILTER = 10,
    /* Testing aid (no counterpart in the reference): n > 0 starts every frame
     * with capacity for only n clip-queue entries, clip-created records and
     * vertices, and tile references, so that the capacity-overflow path runs
     * (flags raised on the device, buffers grown, frame re-run; edxr_stats
     * .retries counts the re-runs).  It also makes the one-pass binning
     * column scan give up at once wherever it would wait, so that the frame
     * is re-run with the three-kernel scan, which the context then keeps.
     * 0 (default) restores automatic sizes and the one-pass scan. */
    EDXR_STATE_DEBUG_CAPACITY = 11
};

/* Pipeline stages timed when EDXR_STATE_PROFILE is on. */
enum edxr_stage {
    EDXR_STAGE_SETUP = 0,       /* vertex transform, outcodes, setup, clipping */
    EDXR_STAGE_BIN_COUNT = 1,
    EDXR_STAGE_BIN_SCAN = 2,
    EDXR_STAGE_BIN_SCATTER = 3,
    EDXR_STAGE_RASTER = 4,      /* rasterisation, depth test, shading, write-back */
    EDXR_NUM_STAGES = 5
};

enum edxr_memory { EDXR_MEM_HOST = 0, EDXR_MEM_DEVICE = 1 };

/* One mesh (Utils/Mesh.h:17-69, Utils/InputBuffer.h:43-194), SoA fp32.
 * Positions, normals and indices are required; colours may be NULL except for
 * EDXR_SHADER_VC_BLINN_PHONG (they are then read as (0,0,0)).  Texture
 * coordinates and per-triangle texture slots (Mesh::GetTextureIds) are
 * needed by EDXR_SHADER_TEXTURED_LAMBERT only; they follow `memory` so that
 * initialisers of the earlier fields leave them NULL. */
typedef struct edxr_mesh {
    const float *px, *py, *pz;
    const float *nx, *ny, *nz;
    const float *cr, *cg, *cb;
    const uint32_t *indices;     /* 3 per triangle (IndexBuffer) */
    uint32_t num_vertices;
    uint32_t num_triangles;
    int32_t memory;              /* EDXR_MEM_HOST or EDXR_MEM_DEVICE */
    const float *tu, *tv;        /* texture coordinate per vertex */
    const uint32_t *tex_ids;     /* texture slot per triangle */
} edxr_mesh;

/* One texture slot (Texture2D<Color>): an RGBA8 image, rows from v = 0 up,
 * (ImageTexture<Color, Color4b>), or, with rgba NULL, a constant colour
 * (ConstantTexture2D<Color>).  Texel (i, j) is centred at
 * ((i + 0.5) / width, (j + 0.5) / height); coordinates wrap; a byte b reads
 * as b / 255.  Mip levels halve each side (floor, at least 1) with 2x2 box
 * averages (edge texels repeated for odd sizes). */
typedef struct edxr_texture {
    uint32_t width, height;
    const uint8_t *rgba;         /* host memory, width * height * 4 bytes */
    float color[3];
} edxr_texture;

/* Per-frame counters (filled after every render). */
typedef struct edxr_stats {
    uint64_t triangles_in;
    uint64_t trivially_rejected;
    uint64_t clipped;            /* straddled at least one plane */
    uint64_t near_clipped;       /* straddled the near plane */
    uint64_t setup_records;      /* triangles (and clipped pieces) after setup */
    uint64_t refs;               /* tile references (N_ref) */
    uint64_t clip_vertices;      /* vertices created by clipping */
    uint32_t retries;            /* re-runs after a capacity overflow */
    uint32_t bin_chunks;
} edxr_stats;

/* One independent render job of a batch: a mesh, a camera, a light, a pixel
 * shader and a W x H render target (the state one Renderer::RenderMesh call
 * uses).  Jobs may share one mesh. */
typedef struct edxr_job {
    edxr_mesh mesh;
    uint32_t width, height;          /* <= 8192, at most 16384 tiles */
    float mv[16], proj[16], raster[16];
    float light[3];
    int32_t shader;                  /* EDXR_SHADER_* */
    edxr_material material;
    int32_t hierarchical;            /* as EDXR_STATE_HIERARCHICAL */
} edxr_job;

/* Gradient outputs of one job for edxr_backward: device arrays that are
 * accumulated into (+=), any of which may be NULL.  Positions, normals and
 * colours hold one value per input vertex of the job's mesh; d_light (3) is
 * w.r.t. the light direction as given, d_material (4) w.r.t. (ambient, kd,
 * ks, shininess).  Jobs that share a mesh may share the arrays. */
typedef struct edxr_grads {
    float *d_px, *d_py, *d_pz;
    float *d_nx, *d_ny, *d_nz;
    float *d_cr, *d_cg, *d_cb;
    float *d_light;
    float *d_material;
} edxr_grads;

/* Launch parameters picked by edxr_autotune. */
typedef struct edxr_tuning {
    uint32_t tris_per_chunk;
    uint32_t raster_blocks_per_cu;
    double frame_ms;                 /* best measured time */
} edxr_tuning;

typedef struct edxr_ctx edxr_ctx;

/* Renderer::Initialize (Core/Renderer.cpp:22-62).  w, h <= 8192 and the
 * tile grid (32x32-pixel tiles) holds at most 16384 tiles. */
int edxr_create(edxr_ctx** out, uint32_t width, uint32_t height);
/* Renderer::~Renderer. */
void edxr_destroy(edxr_ctx* ctx);
/* Renderer::Resize (Core/Renderer.cpp:64-83). */
int edxr_resize(edxr_ctx* ctx, uint32_t width, uint32_t height);
/* Renderer::SetTransform (Core/Renderer.cpp:85-92): MVP = proj * mv and
 * MV^-1 are derived on the host. */
int edxr_set_transform(edxr_ctx* ctx, const float mv[16], const float proj[16], const float raster[16]);
/* Light direction passed to PixelShader::Shade (reference: the constant
 * (1,1,-1) at Core/Renderer.cpp:290).  Normalised by the shader. */
int edxr_set_light(edxr_ctx* ctx, const float dir[3]);
/* Replaces Renderer's pixel-shader object (Core/Renderer.cpp:41). */
int edxr_set_pixel_shader(edxr_ctx* ctx, int32_t kind, const edxr_material* material);
/* Fills the reference's shader constants for a kind. */
void edxr_default_material(int32_t kind, edxr_material* out);
/* RenderStates flags (Core/RenderStates.h:55-61). */
int edxr_set_state(edxr_ctx* ctx, int32_t key, int32_t value);
/* Renderer::RenderMesh (Core/Renderer.cpp:100-118).  Synchronous: returns
 * after the frame is complete on the device. */
int edxr_render_mesh(edxr_ctx* ctx, const edxr_mesh* mesh);
/* No counterpart in the reference: edxr_render_mesh for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Renders a device-resident mesh
 * (memory = EDXR_MEM_DEVICE) with the context's state into the caller's device
 * buffers: color (W*H RGBA8 as uint32), depth (W*H float), ids (W*H uint32),
 * in edxr_read_color's layout, identical to what edxr_render_mesh produces.
 * Every intermediate buffer comes from one static device scratch of 64 MiB
 * per device, held until the process ends; a frame that needs more (about
 * 200k triangles at 1920x1080, fewer at higher resolutions) fails with
 * EDXR_E_LIMIT before anything is launched.  Single-sample only
 * (EDXR_STATE_MSAA 0: else EDXR_E_STATE), eager launches (EDXR_STATE_GRAPH
 * and EDXR_STATE_PROFILE do not apply), no float colour output.  It calls no
 * allocator; edxr_create has already set up the scratch's address.  All calls
 * on one device share the scratch: calls that overlap on one device (two
 * contexts, two threads) race.  Synchronous.  Fills edxr_get_stats; what
 * edxr_read_*, edxr_backward and edxr_debug_* return is unchanged (they refer
 * to the last edxr_render_mesh / edxr_render_batch). */
int edxr_render_mesh_noalloc(edxr_ctx* ctx, const edxr_mesh* mesh, void* color, void* depth, void* ids);
/* Renderer::GetBackBuffer (Core/Renderer.cpp:360-363), copied to the caller:
 * RGBA8, W*H*4 bytes, bottom-up rows; background 0, alpha 255 where drawn. */
int edxr_read_color(edxr_ctx* ctx, uint8_t* rgba, size_t bytes);
/* Final depth (fp32, cleared to 1.0), same layout. */
int edxr_read_depth(edxr_ctx* ctx, float* depth, size_t count);
/* Winning input-triangle index per pixel (0xFFFFFFFF = none), same layout. */
int edxr_read_triangle_ids(edxr_ctx* ctx, uint32_t* ids, size_t count);
/* Per-sample depth and winning triangle IDs of the last single frame,
 * [sample][row][x] with rows as edxr_read_depth; count >= samples * W * H.
 * Either pointer may be NULL.  Sample 0 is what edxr_read_depth and
 * edxr_read_triangle_ids return. */
int edxr_read_sample_buffers(edxr_ctx* ctx, float* depth, uint32_t* ids, size_t count);
/* Device addresses of the three frame buffers (valid until resize/destroy). */
int edxr_device_buffers(edxr_ctx* ctx, void** color, void** depth, void** ids);
int edxr_get_stats(edxr_ctx* ctx, edxr_stats* out);
/* Profiling: summed milliseconds per stage (EDXR_NUM_STAGES values) and the
 * number of frames measured since the last reset. */
int edxr_profile_read(edxr_ctx* ctx, double* stage_ms, uint64_t* frames);
int edxr_profile_reset(edxr_ctx* ctx);
/* The HIP stream (hipStream_t) the context launches on. */
void* edxr_stream(edxr_ctx* ctx);
/* Number of renders (edxr_render_mesh and edxr_render_batch calls that got
 * as far as launching) made on the context; no counterpart in the reference.
 * edxr_backward differentiates the last render, so a caller that keeps a
 * frame for a later backward records this value and checks it is unchanged. */
uint64_t edxr_render_generation(const edxr_ctx* ctx);
const char* edxr_last_error(const edxr_ctx* ctx);

/* ---- batched rendering (no counterpart in the reference) ---- */
/* Renders num_jobs independent jobs with one sequence of launches: every
 * kernel covers all jobs at once through a device job table, so a batch costs
 * about as many launches as one frame.  Synchronous.  Each job's output has
 * the layout of edxr_read_color / _depth / _triangle_ids at its own size. */
int edxr_render_batch(edxr_ctx* ctx, const edxr_job* jobs, uint32_t num_jobs);
/* Copies job `job` of the last batch to caller memory; any pointer may be
 * NULL.  Sizes are width * height (* 4 bytes for rgba). */
int edxr_batch_read(edxr_ctx* ctx, uint32_t job, uint8_t* rgba, float* depth, uint32_t* ids);
/* Device addresses of the last batch's colour, depth and ID arrays and, when
 * pixel_offsets is non-NULL, the first pixel of every job (num_jobs + 1). */
int edxr_batch_device_buffers(edxr_ctx* ctx, void** color, void** depth, void** ids, uint64_t* pixel_offsets);
/* Times candidate (triangles per chunk, raster work items per CU) pairs on the
 * given batch (or single job, num_jobs = 1), keeps the fastest in the
 * context's state and returns it in *out (may be NULL). */
int edxr_autotune(edxr_ctx* ctx, const edxr_job* jobs, uint32_t num_jobs, edxr_tuning* out);

/* ---- differentiation (no counterpart in the reference) ---- */
/* Device address of the last render's float RGB image (EDXR_STATE_FLOAT_COLOR
 * must have been on): 3 floats per pixel, the layout of edxr_read_color for a
 * single frame, every job at 3 * its pixel offset for a batch. */
int edxr_float_color(edxr_ctx* ctx, void** rgb);
/* Copies the last render's float RGB image (all jobs) to dst, host or device
 * memory, bytes = 12 * total pixels. */
int edxr_copy_float_color(edxr_ctx* ctx, void* dst, size_t bytes);
/* Adjoint of the last render (edxr_render_mesh: num_jobs = 1; or
 * edxr_render_batch) with visibility held fixed: d_rgb (device, the layout
 * of edxr_float_color) is dLoss/dRGB; gradients are added to grads[job].
 * Barycentrics are those of the continuous (unsnapped) projection of the
 * winning input triangle; clamped channels pass no gradient.  Requires a
 * raster matrix that is affine in x and y (as edxr_raster_matrix's). */
int edxr_backward(edxr_ctx* ctx, const float* d_rgb, const edxr_grads* grads, uint32_t num_jobs);

/* ---- introspection for parity tests (no counterpart in the reference) ---- */
/* Setup records in submission order: per record 16 uint32 words
 * {x0,y0,x1,y1,x2,y2, inv_det, z0,z1,z2, tri_id, sub, corners, v0,v1,v2}
 * (floats as bit patterns).  Returns the record count or a negative error;
 * writes at most cap records. */
int64_t edxr_debug_read_records(edxr_ctx* ctx, uint32_t* out, uint64_t cap);
/* Per-tile reference lists: offsets[num_tiles+1] and per reference
 * {tri_id, sub, flags} (flags: bit0-2 accept edge, bit3 trivial, bit4 big),
 * each tile's references in submission order ((tri_id, sub) ascending, as the
 * reference's per-core lists concatenated in core order).  The device lists
 * keep one binning chunk's references of a tile in no fixed order; the depth
 * keys carry the submission order, so the frame does not depend on it.
 * Returns the reference count or a negative error. */
int64_t edxr_debug_read_bins(edxr_ctx* ctx, uint64_t* offsets, uint32_t* refs3, uint64_t cap);

/* Texture slots of the context (Mesh::GetTextures, bound by RenderMesh as
 * RenderStates::TextureSlots): copied to the device with their mip chains.
 * Replaces the previous set; count 0 clears it. */
int edxr_set_textures(edxr_ctx* ctx, const edxr_texture* textures, uint32_t count);

/* Directory that WriteFrameToFile writes under (Application::GetBaseDirectory
 * in the reference): frames go to <dir>/Frames/, created when missing.
 * Default ".". */
int edxr_set_frame_directory(edxr_ctx* ctx, const char* dir);
/* Renderer::WriteFrameToFile (Core/Renderer.cpp:352-358): the last frame's
 * back buffer as <dir>/Frames/Frame%05i.bmp with the current frame counter,
 * which render_mesh advances after every frame (Core/Renderer.cpp:117). */
int edxr_write_frame_to_file(edxr_ctx* ctx);
/* Number of frames rendered by render_mesh on this context (RenderStates::FrameCount). */
int64_t edxr_frame_count(edxr_ctx* ctx);

/* ---- scenes (Mesh::LoadSphere / LoadPlane stand-ins; host memory) ---- */
typedef struct edxr_host_mesh {
    float *px, *py, *pz, *nx, *ny, *nz, *cr, *cg, *cb;
    uint32_t* indices;
    uint32_t num_vertices, num_triangles;
    float *tu, *tv;        /* texture coordinates (OBJ input; NULL for generated scenes) */
    uint32_t* tex_ids;     /* material slot per triangle (OBJ input; NULL otherwise) */
} edxr_host_mesh;

/* UV sphere: (stacks+1)*(slices+1) vertices, 2*slices*stacks triangles,
 * outward faces clockwise on screen; colours U[0,1] from SplitMix64(seed). */
int edxr_scene_sphere(edxr_host_mesh* out, uint32_t slices, uint32_t stacks, float radius,
                      const float center[3], uint64_t color_seed);
/* n x n heightfield on [-1,1]^2 (y up), 8-octave value noise * amplitude. */
int edxr_scene_heightfield(edxr_host_mesh* out, uint32_t n, float amplitude, uint64_t height_seed,
                           uint64_t color_seed);
/* 16x16 grid of spheres (segments x segments), radii U[0.5,2], one mesh. */
int edxr_scene_sphere_grid(edxr_host_mesh* out, uint32_t grid, uint32_t segments, uint64_t seed);
/* Mesh::LoadSphere (Utils/Mesh.cpp:53-70) with ObjMesh::LoadSphere's
 * arguments: a UV sphere of the given radius at the origin (the topology of
 * edxr_scene_sphere: (stacks+1)*(slices+1) vertices, 2*slices*stacks
 * triangles, outward faces clockwise on screen), texture coordinates
 * (j / slices, i / stacks) for slice j of stack i, white vertex colours,
 * texture slot 0 on every triangle, placed as edxr_load_obj places a file:
 * p' = T(pos) Rz Ry Rx S(scl) p, rotation in degrees.  Free with
 * edxr_scene_free.  EDXUtil's ObjMesh is not available, so the topology and
 * the texture coordinates are this library's declaration. */
int edxr_load_sphere(const float pos[3], const float scl[3], const float rot[3], float radius, uint32_t slices,
                     uint32_t stacks, edxr_host_mesh* out);
/* Mesh::LoadPlane (Utils/Mesh.cpp:36-51): a square of side `length` in the
 * y = 0 plane centred at the origin, normal +y, corners (-l/2, 0, -l/2),
 * (l/2, 0, -l/2), (-l/2, 0, l/2), (l/2, 0, l/2) with texture coordinates
 * (0,0), (1,0), (0,1), (1,1), triangles (0, 2, 1) and (1, 2, 3) (front faces
 * seen from +y), placed as edxr_load_sphere.  Declared, as above. */
int edxr_load_plane(const float pos[3], const float scl[3], const float rot[3], float length, edxr_host_mesh* out);
/* Appends b to a (a's indices unchanged, b's shifted). */
int edxr_scene_merge(edxr_host_mesh* a, const edxr_host_mesh* b);
void edxr_scene_free(edxr_host_mesh* m);
/* Mesh::LoadMesh (Utils/Mesh.cpp:11-34): a Wavefront OBJ file (with its .mtl
 * colours) placed by p' = T(pos) Rz Ry Rx S(scl) p, rotation in degrees.
 * Polygons are fanned in file order; one vertex per distinct (position,
 * texture coordinate, normal, material); missing normals are area-weighted;
 * missing texture coordinates are (0, 0); colour = the material's Kd, 0.9
 * grey without one.  tex_ids holds each triangle's material slot in
 * edxr_obj_materials order, or the material count for faces without one.
 * Free with edxr_scene_free. */
int edxr_load_obj(const char* path, const float pos[3], const float scl[3], const float rot[3],
                  edxr_host_mesh* out);
/* Bitmap::SaveBitmapFile: 24-bit uncompressed BMP of a bottom-up RGBA8 image
 * (rows kept in order, BGR bytes, alpha dropped). */
int edxr_write_bmp(const char* path, const uint8_t* rgba, uint32_t width, uint32_t height);

/* One material of an OBJ's libraries (Mesh::LoadMesh's material list): the
 * diffuse colour (ConstantTexture2D) and the diffuse map path (ImageTexture),
 * resolved against the .obj file's directory, or "" for none. */
typedef struct edxr_obj_material {
    char name[128];
    float kd[3];
    char map_kd[512];
} edxr_obj_material;
/* The materials of an OBJ file's mtllib statements in slot order.  Returns
 * the material count (at most cap are written) or a negative error. */
int64_t edxr_obj_materials(const char* path, edxr_obj_material* out, uint32_t cap);
/* Reads an uncompressed 24- or 32-bit BMP (the image input of ImageTexture
 * for .bmp diffuse maps) as RGBA8 with its rows in file order (bottom-up,
 * which is v = 0 first).  *rgba is allocated here: release with edxr_free. */
int edxr_load_bmp(const char* path, uint32_t* width, uint32_t* height, uint8_t** rgba);
/* Truevision TGA, image types 2 (uncompressed) and 10 (run-length encoded),
 * 24- or 32-bit BGR(A), either row order; output as edxr_load_bmp's. */
int edxr_load_tga(const char* path, uint32_t* width, uint32_t* height, uint8_t** rgba);
/* Binary PPM (P6, maxval 255); output as edxr_load_bmp's (alpha 255). */
int edxr_load_ppm(const char* path, uint32_t* width, uint32_t* height, uint8_t** rgba);
/* The diffuse-map reader of ImageTexture (Utils/Mesh.cpp:26-27): BMP, PPM or
 * TGA, told apart by the file's first bytes ("BM", "P6", otherwise TGA). */
int edxr_load_image(const char* path, uint32_t* width, uint32_t* height, uint8_t** rgba);
void edxr_free(void* p);
/* Left-handed look-at view matrix and D3D-style projection (z in [0,1]),
 * raster matrix x_r = (x+1)W/2, y_r = (1-y)H/2.  Computed in double. */
void edxr_look_at(const float eye[3], const float at[3], const float up[3], float out[16]);
void edxr_perspective(float fov_y_deg, float aspect, float znear, float zfar, float out[16]);
void edxr_raster_matrix(uint32_t width, uint32_t height, float out[16]);

#ifdef __cplusplus
}
#endif
#endif
