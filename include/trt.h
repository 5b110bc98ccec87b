/* trt — MI355X (gfx950) wavefront path tracer, C ABI.
 *
 * One opaque context owns a HIP stream and every device buffer. Scene data is
 * copied in from caller host arrays; results are copied out to caller host arrays.
 * Calls are synchronous, never abort, and report failures through trt_status plus
 * trt_last_error(). A context is not thread-safe (callers serialise access).
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * reference repository root).
 */
#ifndef TRT_H
#define TRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum trt_status {
  TRT_OK = 0,
  TRT_INVALID_ARG = 1,
  TRT_OUT_OF_MEMORY = 2,
  TRT_DEVICE = 3,
  TRT_NOT_BUILT = 4,
  TRT_UNSUPPORTED = 5
} trt_status;

typedef struct trt_context trt_context;

/* Material (replaces CudaMatarial's 13 properties, Tracer/Common/CommonStructs.h:76-91,
 * 288-307; Disney parameters as in Tracer/Resources/Material.cpp:98-115). `emission`
 * is radiance; model 0 = Disney, 1 = Lambert. */
typedef struct trt_material {
  float base_color[3];
  float roughness;
  float emission[3];
  float metallic;
  float specular;
  float specular_tint;
  float sheen;
  float sheen_tint;
  float clearcoat;
  float clearcoat_gloss;
  float subsurface;
  float anisotropic;
  uint32_t model;
} trt_material;

/* Instance of a mesh (replaces Instance::InstanceData, Tracer/Resources/Instance.cpp:39-49):
 * row-major 3x4 object-to-world transform. */
typedef struct trt_instance {
  float transform[12];
  uint32_t mesh;
  uint32_t material_offset;
} trt_instance;

/* Camera (replaces DeviceRenderer::SetCamera, Tracer/Renderer/DeviceRenderer.cpp:25-40,
 * and the CameraNode fields it reads). fov is horizontal, in radians. */
typedef struct trt_camera {
  float position[3];
  float target[3];
  float up[3];
  float fov;
  float aperture;
  float focal_dist;
  float distortion;
  int32_t bokeh_sides;
  float bokeh_rotation;
} trt_camera;

/* BVH build options (replaces the OptiX build inputs, Tracer/Resources/Model.cpp:240-306,
 * Tracer/Optix/OptixRenderer.cpp:90-123). */
typedef struct trt_build_opts {
  int32_t bins;      /* binned-SAH bins per axis (32) */
  int32_t max_leaf;  /* primitives per leaf, 1..7 (4) */
  int32_t threads;   /* host build threads (0 = OMP_NUM_THREADS if set, else the CPUs of the process affinity mask) */
  float presplit;    /* spatial pre-splitting budget: extra BVH references per triangle
                        (0 = off; 1.0 roughly quarters the triangle tests on scenes of
                        large overlapping triangles). Results do not depend on it. */
} trt_build_opts;

/* Render options (replaces KernelSettings, Tracer/Common/CommonStructs.h:145-153, plus the
 * film size). One trt_render call traces `spp` samples per pixel (the reference's
 * multiSample) and adds them to the accumulator. */
typedef struct trt_render_opts {
  uint32_t width;
  uint32_t height;
  uint32_t spp;
  uint32_t max_depth;
  uint32_t rr_depth;   /* Russian roulette applies at vertices of depth >= rr_depth */
  uint32_t seed;
  float ray_epsilon;
  float clamp;         /* firefly clamp on a contribution's max component; 0 = off */
  uint32_t mode;       /* render mode (replaces Renderer::SetRenderMode, Tracer/Renderer/Renderer.h:48-49,
                          with the kernels of the Tracer/CUDA/GPU/Shade headers): TRT_MODE_ values below */
  uint32_t mode_flags; /* material property for TRT_MODE_MATERIAL_PROPERTY (the reference's
                          MaterialPropertyIds order: anisotropic 0, clearcoat, clearcoat gloss, diffuse,
                          emissive, metallic, normal, roughness, sheen, sheen tint, specular,
                          specular tint, subsurface 12) */
  float ao_dist;       /* ambient-occlusion distance (KernelSettings::aoDist) */
  float zdepth_max;    /* depth shown as white by TRT_MODE_ZDEPTH (KernelSettings::zDepthMax) */
} trt_render_opts;

/* Render modes. Path tracing is 0; the others write one colour per camera sample
 * (misses black), averaged like radiance. AO: fraction min(t, ao_dist) / ao_dist of a
 * cosine-distributed ray from the first hit (AO_SHADING: times the base colour). Normals,
 * tangent frames (built from the shading normal) and texture coordinates (barycentric
 * (b1, b2, 0): the scene data has no UVs) are shown as (v + 1) / 2 or raw; ids as
 * distinct colours; WIREFRAME marks hits near triangle edges; ZDEPTH = view depth /
 * zdepth_max clamped to [0, 1]. */
enum {
  TRT_MODE_PATH_TRACING = 0,
  TRT_MODE_AMBIENT_OCCLUSION = 1,
  TRT_MODE_AMBIENT_OCCLUSION_SHADING = 2,
  TRT_MODE_BITANGENT = 3,
  TRT_MODE_GEOMETRIC_NORMAL = 4,
  TRT_MODE_MATERIAL_ID = 5,
  TRT_MODE_MATERIAL_PROPERTY = 6,
  TRT_MODE_OBJECT_ID = 7,
  TRT_MODE_SHADING_NORMAL = 8,
  TRT_MODE_TANGENT = 9,
  TRT_MODE_TEXTURE_COORDINATE = 10,
  TRT_MODE_WIREFRAME = 11,
  TRT_MODE_ZDEPTH = 12
};

/* Frame statistics (replaces RenderStatistics::DeviceStatistics,
 * Tracer/Renderer/RenderStatistics.h:11-26). Times are HIP-event milliseconds. */
typedef struct trt_stats {
  uint64_t paths;
  uint64_t extend_rays;   /* all closest-hit rays, primary included */
  uint64_t shadow_rays;
  uint64_t primary_rays;
  float raygen_ms;
  float extend_ms;        /* sum over bounces of Extend kernel time (from launch to end) */
  float shade_ms;         /* sum over bounces of Shade, which runs beside the previous bounce's
                             Shadow launch */
  float shadow_ms;        /* Shadow of bounce d (its emission list included) overlaps Extend and
                             Shade of d + 1: from its launch to end */
  float finalize_ms;
  float frame_ms;         /* first to last event of the frame */
  uint32_t extend_launches;
  uint32_t shadow_launches;
  uint32_t sample_count;  /* accumulated samples per pixel after this frame */
  uint32_t overflow;      /* nonzero if a traversal stack overflowed (must stay 0) */
  float trace_ms;         /* frame minus RayGen, Shade and Finalize: the wall time of the Extend
                             and Shadow phases, less the part of Shadow that runs beside Shade */
} trt_stats;

typedef struct trt_bvh_info {
  uint64_t nodes;
  uint64_t leaves;
  uint32_t depth;
  uint32_t stack_entries;
  double sah_cost;
  double build_ms;
  uint64_t device_bytes;
  uint32_t two_level;
  uint32_t tlas_depth;
  uint64_t references;  /* triangle references in the leaves (> triangles when pre-split) */
} trt_bvh_info;

typedef struct trt_pick_result {
  float origin[3];
  float direction[3];
  float t;
  uint32_t instance;
  uint32_t primitive;
} trt_pick_result;

/* Context life cycle (replaces Renderer::Renderer / ~Renderer, Tracer/Renderer/Renderer.cpp:41-66,
 * and CudaDevice(0), Tracer/CUDA/CudaDevice.cpp). */
trt_status trt_create(int device, trt_context** out);
void trt_destroy(trt_context* ctx);
const char* trt_last_error(const trt_context* ctx);
const char* trt_version(void);

/* Scene (replaces Renderer::UpdateScene's geometry/material upload, Tracer/Renderer/Renderer.cpp:79-143,
 * 442-504, with Model::Build packing, Tracer/Resources/Model.cpp:93-167). Triangle soup:
 * positions 9 floats per triangle, normals 9 floats per triangle or NULL, one material
 * id per triangle. */
trt_status trt_clear_scene(trt_context* ctx);
trt_status trt_set_materials(trt_context* ctx, const trt_material* mats, uint32_t n);
trt_status trt_add_mesh(trt_context* ctx, const float* positions, const float* normals,
                        const uint32_t* material_ids, uint32_t n_tris, uint32_t* mesh_id);
/* Two-level scenes: when instances are set, every mesh is traced through them. */
trt_status trt_set_instances(trt_context* ctx, const trt_instance* inst, uint32_t n);
/* Builds the BVH(s), the light table, and uploads everything (replaces Model::Upload +
 * OptixRenderer::BuildAccel + Scene::GatherLights, Tracer/Renderer/Scene.cpp:275-302).
 * TRT_INVALID_ARG when the scene has no triangles or no materials. */
trt_status trt_build(trt_context* ctx, const trt_build_opts* opts);
trt_status trt_bvh_info_get(trt_context* ctx, trt_bvh_info* info);

/* Scene files (replaces SceneFile::Load, Tracer/FileIO/SceneFile.cpp:683-731, and
 * ModelFile::Import, Tracer/FileIO/ModelFile.cpp:331-448, for Wavefront OBJ/MTL models):
 * the reference's JSON scene format with its "models", "instances" (transform
 * operations matrix / translate / rotate-x|y|z / scale), "materials", "camera",
 * "renderer" and "sky" sections. Instance matrices are composed like the reference and
 * read row-major, as its traversal reads them (Tracer/Resources/Instance.cpp:39-49).
 * Materials become Disney materials (model 0) with emission = emissive x diffuse, the
 * radiance the reference shades with (Tracer/CUDA/GPU/CudaMaterial.h:158). Texture maps
 * are not supported: each one is reported as a warning. */
typedef struct trt_scene_file_info {
  uint32_t n_models;      /* models used by instances */
  uint32_t n_instances;
  uint32_t n_materials;   /* all models' materials, concatenated in model order */
  uint32_t n_triangles;   /* all models' triangles (each model counted once) */
  uint32_t has_camera;
  trt_camera camera;      /* fov converted to radians */
  uint32_t spp;           /* renderer.multisample */
  uint32_t max_depth;     /* renderer.maxdepth */
  float ray_epsilon;      /* renderer.ray-epsilon */
  float ao_dist;          /* renderer.aodist */
  float zdepth_max;       /* renderer.zdepthmax */
  uint32_t sky_enabled;
  uint32_t sky_draw_sun;
  float sky_dir[3];
  float sky_angular_diameter;  /* degrees (the reference's unit) */
  float sky_intensity;
  float sky_turbidity;
  float sky_bias;
  uint32_t n_warnings;
} trt_scene_file_info;

/* Parses a scene file and its OBJ models on the host (no device use) and copies the
 * result out. Pass NULL arrays to query the sizes in *info first. Arrays: positions and
 * normals (9 floats per triangle, models concatenated, object space), mat_ids (model-
 * local), model_tri_offset / model_tri_count / model_mat_offset (one per model),
 * materials (n_materials), instances (n_instances; mesh = model index, material_offset =
 * its model's first material). `msg` (msg_len bytes, may be NULL) receives the error,
 * or on success the warnings, one per line. */
trt_status trt_scene_file_read(const char* path, trt_scene_file_info* info, float* positions, float* normals,
                               uint32_t* mat_ids, uint32_t* model_tri_offset, uint32_t* model_tri_count,
                               uint32_t* model_mat_offset, trt_material* materials, trt_instance* instances,
                               char* msg, size_t msg_len);

/* Loads a scene file into the context: replaces the scene (meshes, materials,
 * instances), builds it with `opts` (NULL = defaults) and sets the file's camera. The
 * file's render settings come back in *info for the caller's trt_render_opts. */
trt_status trt_load_scene_file(trt_context* ctx, const char* path, const trt_build_opts* opts,
                               trt_scene_file_info* info);

/* Display transform (replaces the tone-mapping post pass, Tracer/glsl/Tonemap.frag, with
 * its uniforms exposure / gamma / tonemapMethod, Tracer/glsl/Tonemap.json): colour x
 * exposure, tone curve `method` (0 ACES, 1 filmic, 2 Lottes, 3 Reinhard, 4 Reinhard2,
 * 5 Uchimura, 6 Uncharted 2, other: none), gamma, 8-bit RGBA (alpha 255), rows bottom-up
 * like the colour buffer. source 0: the mean colour, 1: the last denoised frame. */
typedef struct trt_tonemap_opts {
  uint32_t method;
  float exposure;  /* 1 */
  float gamma;     /* 2.2 */
  uint32_t source;
} trt_tonemap_opts;
trt_status trt_tonemap(trt_context* ctx, const trt_tonemap_opts* opts, uint8_t* rgba8);

/* Image files (replaces the screenshot/export path of the reference's GUI, which saves
 * through FreeImage): portable float map (RGB, bottom-up rows as stored) and PNG (8-bit
 * RGBA, written top-down). Buffers are width x height x 4, rows bottom-up. No context. */
trt_status trt_write_pfm(const char* path, uint32_t width, uint32_t height, const float* rgba);
trt_status trt_write_png(const char* path, uint32_t width, uint32_t height, const uint8_t* rgba8);

/* Sky seen by rays that leave the scene (replaces Sky and SampleSky,
 * Tracer/Renderer/Sky.h:10-66, Tracer/CUDA/GPU/CudaSky.h:44-156): the reference's
 * Preetham-style daylight model. With sun_nee = 0 (the reference's setting: it compiles
 * sky NEE out, Tracer/CUDA/GPU/CudaLights.h:11) the sky is reached only by rays that
 * miss and the sun disc is drawn for camera rays and specular bounces only. With
 * sun_nee = 1 the sun disc is also a light of next-event estimation, chosen against
 * the emissive triangles by energy (the reference's disabled branch,
 * Tracer/CUDA/GPU/CudaLights.h:24, 66-88), and rays that reach the disc add it with the
 * balance-heuristic MIS weight. Resets accumulation. Disabled by default
 * (trt_clear_scene disables it). */
typedef struct trt_sky {
  uint32_t enabled;
  uint32_t draw_sun;
  float sun_dir[3];          /* towards the sun (normalised here) */
  float angular_diameter;    /* sun disc, arc minutes (32) */
  float intensity;           /* sun intensity (1000) */
  float turbidity;           /* atmosphere turbidity (4) */
  uint32_t sun_nee;          /* 1: the sun disc is a light of next-event estimation */
} trt_sky;
trt_status trt_set_sky(trt_context* ctx, const trt_sky* sky);

/* Camera and frame (replaces Renderer::SetCamera / RenderFrame / Reset,
 * Tracer/Renderer/Renderer.h:45-61; DeviceRenderer::RenderFrame, Tracer/Renderer/DeviceRenderer.cpp:97-133). */
trt_status trt_set_camera(trt_context* ctx, const trt_camera* cam);
trt_status trt_render(trt_context* ctx, const trt_render_opts* opts, trt_stats* stats);
/* Per-phase timing of trt_render (no counterpart in the reference, which brackets its
 * frame with one pair of CUDA events, Tracer/Renderer/DeviceRenderer.cpp:119-125): on by
 * default. Off, a frame records only the events its two streams need to stay ordered and
 * the two that bracket it, since each event between two kernels delays the second by a
 * few microseconds; trt_stats then holds the counts and frame_ms, and the per-phase
 * times (raygen_ms ... finalize_ms, trace_ms) are 0. */
trt_status trt_set_phase_timing(trt_context* ctx, int enabled);
trt_status trt_reset(trt_context* ctx);
uint32_t trt_sample_count(const trt_context* ctx);
/* Number of frames this context has rendered (trt_render and trt_measure_traversal each
 * count one); lets a caller tell whether the frame it rendered is still the last one
 * (the reference keeps the equivalent per-frame counters in RenderStatistics). */
uint64_t trt_frame_id(const trt_context* ctx);

/* Outputs (replace the accumulator / colour / albedo / normal buffers,
 * Tracer/Renderer/DeviceRenderer.h:92-95, FinalizeFrame, Tracer/CUDA/GPU/CudaUtilityKernels.h:27-44).
 * Each writes width*height*4 floats; accum = (rgb sum, sample count), color = mean,
 * albedo = (first-hit albedo rgb, 1), normal = (first-hit shading normal xyz, view depth).
 * Albedo and normal are means over the pixel's accumulated samples (the reference keeps
 * whichever sample wrote last, PathTracing.h:48-70); the normal mean is not renormalised.
 * A camera ray that misses contributes the sky's radiance as albedo and a zero normal.
 * trt_write_accum restarts these means with the next frame. */
trt_status trt_read_accum(trt_context* ctx, float* rgba);
trt_status trt_read_color(trt_context* ctx, float* rgba);
trt_status trt_read_albedo(trt_context* ctx, float* rgba);
trt_status trt_read_normal(trt_context* ctx, float* rgba);
/* Checkpoint / resume of progressive rendering (no counterpart in the reference, which
 * restarts accumulation on every reset, Tracer/Renderer/Renderer.cpp:186-190): restores the
 * accumulator of a width x height film from rgba as trt_read_accum wrote it (rgb sums, the
 * sample count in alpha, equal in every pixel); the next trt_render of that size
 * continues with the following samples, so the resumed render equals the uninterrupted
 * one. The scene and camera must be set as they were (setting them restarts
 * accumulation). */
trt_status trt_write_accum(trt_context* ctx, uint32_t width, uint32_t height, const float* rgba);
/* Per-path radiance of the last frame, 3 floats per path (path = sample*W*H + pixel). */
trt_status trt_read_path_radiance(trt_context* ctx, float* rgb, uint64_t n_paths);

/* Ray queries on the built scene (replace optixTrace closest / terminate-on-first-hit,
 * Tracer/Optix/GPU/optix.cu:264-265, 333-335). rays: 8 floats per ray
 * (ox, oy, oz, tmin, dx, dy, dz, tmax). hits: 4 floats per ray (t, b1, b2, prim as uint
 * bits; prim = 0xffffffff on a miss); inst: instance per ray (0xffffffff single-level).
 * Both run the renderer's own Extend / Shadow traversal kernels over the rays.
 * TRT_INVALID_ARG, before anything is launched, for a ray with a non-finite origin or
 * direction component, a NaN tmin/tmax (tmax = +inf is allowed) or a negative tmin (the
 * renderer's rays start at tmin = ray_epsilon >= 0); n = 0 does nothing. */
trt_status trt_trace_closest(trt_context* ctx, const float* rays, uint32_t n, float* hits, uint32_t* inst);
trt_status trt_trace_any(trt_context* ctx, const float* rays, uint32_t n, uint8_t* occluded);

/* trt_trace_closest without device allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory during the call (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it). Same arguments, checks and
 * results as trt_trace_closest, at most TRT_NOALLOC_MAX_RAYS rays per call
 * (TRT_INVALID_ARG beyond). Its staging (the ray records, their queue form and the hit
 * records, 96 bytes per ray) is one fixed device array per GPU, held until the process
 * ends, instead of context buffers grown on demand; the context's own buffers (queue
 * counters, traversal stack spill area) come from trt_create and trt_build. Every call
 * on a GPU shares that GPU's array: calls that overlap on one GPU, from any context or
 * thread, race and must be serialised by the caller. */
#define TRT_NOALLOC_MAX_RAYS 65536u
trt_status trt_trace_closest_noalloc(trt_context* ctx, const float* rays, uint32_t n, float* hits, uint32_t* inst);

/* Ray pick (replaces DeviceRenderer::PickRay, Tracer/Renderer/DeviceRenderer.cpp:71-93, with the
 * RayPick ray-gen mode, Tracer/Optix/GPU/optix.cu:345-377): camera ray through the
 * centre of pixel (x, y) (y = 0 bottom row), no jitter. */
trt_status trt_pick(trt_context* ctx, uint32_t width, uint32_t height, uint32_t x, uint32_t y,
                    trt_pick_result* out);

/* Measurement hook: renders one frame with `opts` through counting variants of the
 * Extend and Shadow kernels (nothing is accumulated) and reports average node visits
 * and triangle tests per ray (the bytes every ray loads), how many distinct nodes and
 * triangle records each launch read (summed over the frame's launches: the bytes a
 * launch must bring from memory at least; bottom-level trees only in instanced scenes),
 * and how many traversal-stack entries were moved from the on-chip ring to the global
 * spill area. */
typedef struct trt_trace_counts {
  double closest_nodes_per_ray;
  double closest_tris_per_ray;
  double any_nodes_per_ray;
  double any_tris_per_ray;
  uint64_t closest_rays;
  uint64_t any_rays;
  uint32_t node_bytes;
  uint32_t tri_bytes;
  uint64_t closest_spills;  /* stack entries spilled by the closest-hit traversals */
  uint64_t any_spills;      /* and by the any-hit traversals */
  uint64_t closest_distinct_nodes;  /* distinct nodes read per Extend launch, summed over launches */
  uint64_t closest_distinct_tris;   /* distinct triangle records, likewise */
  uint64_t any_distinct_nodes;      /* the same for the Shadow launches */
  uint64_t any_distinct_tris;
  uint32_t launches;                /* Extend launches of the frame (= Shadow launches) */
} trt_trace_counts;
trt_status trt_measure_traversal(trt_context* ctx, const trt_render_opts* opts, trt_trace_counts* out);

/* Measurement hook (no counterpart in the reference): the vector-memory ceiling of the
 * traversal's loads. Lanes read 48-byte records (a node or a triangle record, three
 * 16-byte loads) at pseudo-random indices of a device buffer of `buffer_bytes`, `iters`
 * records per lane, with independent loads; *gbs = bytes read per second (GB/s, best of
 * three timed launches). Over a buffer that fits the L2 this is the highest rate such
 * gathers reach on the device. No context: it runs on `device` with its own buffers. */
trt_status trt_measure_gather(int device, uint64_t buffer_bytes, uint32_t iters, double* gbs);
/* Measurement hook (no counterpart in the reference): the same loads timed per
 * wave-level instruction, with only the first `active_lanes` (1..64) lanes of each wave
 * reading, each lane its own pseudo-random record (same_record = 0) or every active
 * lane of a wave the same one (1); *ns_per_load = nanoseconds per wave-level 16-byte load
 * instruction per CU (best of three timed launches). */
trt_status trt_measure_gather_lanes(int device, uint64_t buffer_bytes, uint32_t iters, uint32_t active_lanes,
                                    uint32_t same_record, double* ns_per_load);

/* Differentiable rendering. No counterpart in the reference (its PathTracingKernel,
 * Tracer/CUDA/GPU/Shade/PathTracing.h:8-175, has no adjoint): the adjoint of one frame
 * by path replay. Parameters per material: base_color rgb, roughness, metallic
 * (5 floats). which: 1 = values used for BSDF evaluation, 2 = values used for sampling
 * decisions and MIS weights (held fixed for finite differences), 3 = both. */
trt_status trt_set_material_params(trt_context* ctx, const float* params, uint32_t n_materials, uint32_t which);
/* Gradient of a loss with respect to the evaluation parameters of every material for
 * the last frame rendered by trt_render (same options): dL_dimage = dLoss/d(frame mean
 * image), 3 floats per pixel; out dL_dparams = 5 floats per material. Records are reduced
 * in a fixed order (no float atomics): repeated calls return identical bits.
 * TRT_INVALID_ARG when the options differ from that frame's, or when the scene, camera,
 * sky or materials changed after it (the replay would not match the frame). */
trt_status trt_backward(trt_context* ctx, const trt_render_opts* opts, const float* dL_dimage, float* dL_dparams);

/* Learned denoiser (replaces Denoiser, Tracer/Optix/Denoiser.h:14-38 and its
 * DenoiseFrame, Tracer/Optix/Denoiser.cpp:47-115, which run the closed OptiX AI
 * denoiser on colour + albedo + normal). A 4-level U-Net (32/64/96/128 channels, 3x3
 * convolutions, BF16 weights/activations, FP32 accumulation on MFMA) over 10 input
 * channels: log(1 + radiance), first-hit albedo, normal, log(1 + view depth). */
typedef struct trt_denoise_stats {
  float features_ms;
  float network_ms;
  float total_ms;
} trt_denoise_stats;
trt_status trt_denoiser_init(trt_context* ctx, uint32_t seed);
uint64_t trt_denoiser_num_params(void);
trt_status trt_denoiser_get_params(trt_context* ctx, float* out, uint64_t n);
trt_status trt_denoiser_set_params(trt_context* ctx, const float* in, uint64_t n);
trt_status trt_denoiser_save(trt_context* ctx, const char* path);
trt_status trt_denoiser_load(trt_context* ctx, const char* path);
/* Training set of n (input, target) pairs of width x height, captured from frames the
 * context renders: kind 0 stores the current frame's input features, kind 1 its mean
 * colour as the target (render the target at high spp). */
trt_status trt_denoiser_dataset(trt_context* ctx, uint32_t n, uint32_t width, uint32_t height);
trt_status trt_denoiser_capture(trt_context* ctx, uint32_t slot, uint32_t kind);
/* `steps` Adam steps on random crop x crop windows (crop % 128 == 0), batch crops per
 * step, on the mean absolute log-space error (L1); losses (nullable) receives that loss
 * per step. */
trt_status trt_denoiser_train(trt_context* ctx, uint32_t steps, uint32_t batch, uint32_t crop, float lr, uint32_t seed,
                              float* losses);
/* Denoises the accumulated frame (needs width % 128 == 0, height % 8 == 0). */
trt_status trt_denoise(trt_context* ctx, trt_denoise_stats* stats);
/* The reference's denoising policy (Denoiser::SetEnabled / SetSampleTreshold,
 * Tracer/Optix/Denoiser.h:24-31; Renderer::ShouldDenoise and DenoiseFrame,
 * Tracer/Renderer/Renderer.cpp:291-304, 546-569): trt_denoise_frame, called after a frame,
 * denoises when the policy is enabled, the last frame was path traced (or ambient
 * occlusion), at least `sample_threshold` samples are accumulated and, after a first
 * denoise, the sample count reached 1.5 x the one last denoised; the result is blended
 * towards the noisy frame by 1 / (spp + 1) (Tracer/Optix/Denoiser.cpp:108). *ran = 1 when
 * it denoised (stats then filled). Restarting accumulation restarts the policy. Off by
 * default. */
trt_status trt_denoiser_set_policy(trt_context* ctx, uint32_t enabled, uint32_t sample_threshold);
trt_status trt_denoise_frame(trt_context* ctx, trt_denoise_stats* stats, uint32_t* ran);
trt_status trt_read_denoised(trt_context* ctx, float* rgba);
/* Test hooks: network forward on host features [N][H][W][16] -> log-space output
 * [N][H][W][32]; backward of sum(dout * output) -> parameter gradients. */
trt_status trt_denoiser_forward(trt_context* ctx, const float* features, uint32_t n, uint32_t height, uint32_t width,
                                float* out);
trt_status trt_denoiser_backward(trt_context* ctx, const float* dout, uint32_t n, uint32_t height, uint32_t width,
                                 float* grads);
/* Post-activation output of convolution `index` (0..13) of the last forward. */
trt_status trt_denoiser_activation(trt_context* ctx, uint32_t index, float* out, uint64_t n);
/* Test hook: the inference path of trt_denoise on host features [N][H][W][16] ->
 * radiance [N][H][W][4] = (expm1(max(feature c + output c, 0)) for c < 3, 1). */
trt_status trt_denoiser_apply(trt_context* ctx, const float* features, uint32_t n, uint32_t height, uint32_t width,
                              float* radiance);

#ifdef __cplusplus
}
#endif

#endif /* TRT_H */
