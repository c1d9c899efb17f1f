/*
 * ever_amd — C ABI of the MI355X (gfx950) exact volumetric ellipsoid tracer.
 *
 * Plain C, plain pointers and sizes; every device pointer is HBM memory owned by the caller (PyTorch
 * tensors in the Python binding). Every entry point enqueues on the given hipStream_t (passed as void*)
 * and never synchronizes the host, allocates memory or copies device data to the host, so a call sequence
 * can be captured into a hipGraph. Errors are returned as status codes; ever_last_error() gives the
 * message of the last failing call on the calling thread.
 *
 * The entry points replace these reference interfaces (paths relative to the reference repository root):
 *   ever_build_scene      splinetracers/fast_ellipsoid_splinetracer/py_binding.cpp:99-132 (Primitives.add_primitives)
 *                         + create_aabbs.cu:117-132 + GAS.cpp:56-153 (GAS::build, OptiX BVH -> LBVH)
 *                         + splinetracers/slang/sh_kernel.slang:19-42 (per-primitive SH colour, color_mode 1)
 *   ever_update_colors    py_binding.cpp:207-209 (Forward.update_model) -> Forward.cpp:252-255 (reset_features):
 *                         new colour features for an already built scene, geometry and tree kept
 *   ever_trace_forward    py_binding.cpp:210-258 (Forward.trace_rays) -> Forward.cpp:23-82,
 *                         initialize_density.cu:152-208, slang/shaders.slang:171-235
 *   ever_trace_backward   splinetracers/fast_ellipsoid_splinetracer.py:105-219 (SplineTracer.backward) ->
 *                         slang/backwards_kernel.slang:259-367 (backwards_kernel), :374-428
 *                         (backwards_initial_drgb_kernel), gradient clamps fast_ellipsoid_splinetracer.py:201-211
 *   ever_trace_backward_recompute  the same adjoint without the stored event list of py_binding.cpp:221-223
 *   ever_eval_sh          eval_sh.py:31-55 (EvalSH.forward) -> splinetracers/slang/sh_kernel.slang:19-42
 *   ever_eval_sh_backward eval_sh.py:57-68 (EvalSH.backward) -> splinetracers/slang/sh_kernel.slang:45-98
 *   ever_scene_from_records  splinetracers/fast_ellipsoid_splinetracer/ply_file_loader.cc:150-293 (ReadSceneFromFile,
 *                         NormalizeSceneData) and the layout of utils/math_util.py:152-202 (save_ply)
 *   ever_morton_order     ply_file_loader.cc:72-148 (SortComponentsToMortonOrder)
 *   ever_camera_rays      integration_utils/gaussian_renderer/ever.py:29-121 (get_ray_directions, get_rays,
 *                         camera2rays, camera2rays_full) -> integration_utils/utils/camera_utils_zipnerf.py:816-1040
 *                         (pixels_to_rays: radial/tangential undistortion, fisheye, panoramic)
 */
#ifndef EVER_AMD_H
#define EVER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EVER_ABI_VERSION 9

/* status codes */
#define EVER_OK 0
#define EVER_ERR_INVALID 1   /* bad argument (shape, null pointer, size) */
#define EVER_ERR_HIP 2       /* a HIP runtime call or kernel launch failed */
#define EVER_ERR_WORKSPACE 3 /* workspace smaller than the *_workspace_bytes() query */

/* colour modes */
#define EVER_COLOR_DC 0 /* reference: colour = SH_C0 * features[i,0,:] + 0.5 (shaders.slang:65-82) */
#define EVER_COLOR_SH 1 /* colour = SH(features[i], normalize(mean_i - sh_origin)), fused prepass */

/* Primitive set. Layout of each array as in the reference's Primitives (structs.h:56-68). */
typedef struct ever_prims {
  const float* means;     /* (N,3) */
  const float* scales;    /* (N,3) */
  const float* quats;     /* (N,4), (w,x,y,z) */
  const float* densities; /* (N) */
  const float* features;  /* (N,d,3) SH coefficients */
  int64_t num_prims;
  int32_t feature_size; /* d = (deg+1)^2 */
  int32_t color_mode;   /* EVER_COLOR_* */
  int32_t sh_degree;    /* color_mode EVER_COLOR_SH only */
  int32_t activation;   /* color_mode EVER_COLOR_SH only: 0 none, 1 softplus(beta=10) */
  float sh_origin[3];   /* color_mode EVER_COLOR_SH only */
  /* Several views in one batch (color_mode EVER_COLOR_SH only; not in the reference, where each camera is its own
   * call): num_views > 1 colours every primitive once per view origin (view_origins, (V,3)) into view_colors
   * ((V,N,4) caller buffer written by ever_build_scene as (density, R, G, B)), and each ray uses its view's
   * colours (ever_rays.view_index). num_views 0 or 1: one view at sh_origin. V * N < 2^32. */
  int32_t num_views;
  const float* view_origins;
  float* view_colors;
} ever_prims;

/* Ray batch and trace options (Forward::trace_rays arguments, Forward.cpp:23-82). */
/* Tree walk of the forward pass. AUTO takes WAVE when the call sorts the rays (at least 65536 rays, so that a
   wave holds neighbouring rays) and RAY otherwise; both give the same events and results. */
#define EVER_TRAVERSAL_AUTO 0
#define EVER_TRAVERSAL_RAY 1
#define EVER_TRAVERSAL_WAVE 2

typedef struct ever_rays {
  const float* origins;    /* (R,3) */
  const float* directions; /* (R,3), need not be unit */
  int64_t num_rays;
  float tmin;
  float tmax;
  int32_t max_iters;
  /* events gathered per traversal: 8, 16 (reference), 19 or 32. Results do not depend on it: they follow the
   * reference's 16-event rounds, which matter only for exits beyond a tmax that cuts through primitives (an
   * exit beyond tmax is integrated when its entry lies in the same 16-event round, shaders.slang:96-159). */
  int32_t kbuf;
  int32_t traversal; /* EVER_TRAVERSAL_*: who walks the tree, one lane per ray or the 64 rays of a wave */
  /* (R) proposed termination depth per ray, nullable. Depths are measured along the unit direction from
   * origin + tmin * direction. A ray first gathers only events up to its proposal (nodes beyond it are not
   * entered); if it has not terminated by then it continues with the plain walk, so results equal the plain
   * trace's. Not in the reference: the learned-proposal form. */
  const float* proposal;
  const int32_t* view_index; /* (R) view of each ray when prims.num_views > 1, else ignored */
  /* Persistent launch of the per-wave walk (not in the reference): 0 = one wave per block over the whole grid;
   * t > 0 = one block per wave slot of the GPU, each fetching t consecutive 64-ray tiles at a time from a counter
   * until the batch is done. Results do not depend on it. */
  int32_t tile;
  /* Event-list layout (not in the reference, which has only the first): EVER_EVENTS_BY_RAY puts event k of ray r
   * at tri_collection[r + k*R]; EVER_EVENTS_BY_SLOT puts it at [s + k*R], s = the ray's position in the forward's
   * ray_order, so that the 64 rays a wave traces together write (and the backward reads) one contiguous row per
   * event. ever_trace_backward must be given the value the forward used. */
  int32_t event_layout;
} ever_rays;

#define EVER_EVENTS_BY_RAY 0
#define EVER_EVENTS_BY_SLOT 1

typedef struct ever_forward_out {
  float* color;            /* (R,5): R,G,B, opacity, distortion */
  float* states;           /* (R,16) final SplineState (structs.h:41-54) */
  int32_t* iters;          /* (R) events processed */
  int32_t* tri_collection; /* (R*max_iters), event id 2*prim+entry (layout: ever_rays.event_layout); nullable */
  int32_t* touch_count;    /* (N) events per primitive (exact); nullable; zeroed by the call */
  float* initial_drgb;     /* (R,4) density/colour already open at the ray start */
  double* ray_state64;     /* (R,8) FP64 final accumulators: density and colour (4), log-transmittance, depth of the
                              last event, and the log-transmittance and depth that ever_trace_backward unwinds (the
                              same accumulation with its safe roots) */
  int32_t* ray_order;      /* (R) processing order (rays sorted for coherence), read by ever_trace_backward */
  void* workspace;         /* stack spill + ray sort, ever_forward_workspace_bytes_for(R, traversal) bytes */
  size_t workspace_bytes;
  float* last_dirac; /* (R,4) control point of the last event integrated, (+-density, +-density * colour)
                        (shaders.slang:229-232), zeros for a ray without events; nullable */
} ever_forward_out;

typedef struct ever_backward_io {
  const float* grad_out;      /* (R,5) */
  const int32_t* iters;       /* forward outputs */
  const int32_t* tri_collection;
  const float* states;
  const double* ray_state64;
  const int32_t* ray_order; /* forward's processing order; nullable (identity) unless the events are stored by slot */
  const float* wcts; /* (W,4,4) world-to-camera matrices for the means2D statistic; nullable */
  int64_t num_wcts;
  float* dmeans;     /* (N,3) */
  float* dscales;    /* (N,3) */
  float* dquats;     /* (N,4) */
  float* ddensities; /* (N) */
  float* dfeatures;  /* (N,d,3) */
  float* drayo;      /* (R,3) */
  float* drayd;      /* (R,3) */
  float* dmeans2D;   /* (N,2); written only when wcts != NULL */
  float* dinitial_drgb; /* (R,4) adjoint of the start-inside density; nullable */
  void* workspace;      /* ever_backward_workspace_bytes(N, R, W) bytes */
  size_t workspace_bytes;
  float* view_color_grads; /* (V,N,4) scratch when prims.num_views > 1, else ignored */
} ever_backward_io;

int ever_abi_version(void);
/* sizeof() of each ABI struct, in the order ever_prims, ever_rays, ever_forward_out, ever_backward_io. */
int ever_struct_sizes(int64_t* out4);
const char* ever_last_error(void);

/* Byte offsets of the scene workspace regions, for inspection by tests (order: header, aabb, fprim, keys0,
 * vals0, keys1, vals1, hist, nodes, nodes4, range, sbox, pre, suf, tab, rec, colr, cand, bpart, total). */
int ever_scene_layout(int64_t num_prims, int64_t* offsets20);

/* Scene workspace: LBVH nodes, sort buffers, per-primitive records. Lives from build to backward. */
size_t ever_scene_workspace_bytes(int64_t num_prims);
int ever_build_scene(const ever_prims* prims, void* workspace, size_t workspace_bytes, void* stream);
/* Rewrites only the per-primitive colours of a scene built from the same geometry (means, scales, quats,
 * densities and num_prims unchanged; features, color_mode, SH origin/degree and view colours may differ):
 * the tree, the primitive records' geometry rows and the bounds stay as built. */
int ever_update_colors(const ever_prims* prims, void* workspace, size_t workspace_bytes, void* stream);

/* Forward workspace: the per-ray walk's traversal-stack spill (160 entries per ray beyond 32 in LDS; the per-wave
 * walk keeps its 192 entries on chip; either holds every entry a tree can need: at most 180, from the key and
 * index bits of the hierarchy) and the ray sort. ever_forward_workspace_bytes_for sizes it for a traversal mode (EVER_TRAVERSAL_*);
 * ever_forward_workspace_bytes(n) = ever_forward_workspace_bytes_for(n, EVER_TRAVERSAL_AUTO).
 * The scene workspace is also written: the start-inside candidate list of this ray batch. */
size_t ever_forward_workspace_bytes(int64_t num_rays);
size_t ever_forward_workspace_bytes_for(int64_t num_rays, int32_t traversal);
int ever_trace_forward(const ever_prims* prims, void* workspace, size_t workspace_bytes,
                       const ever_rays* rays, const ever_forward_out* out, void* stream);

size_t ever_backward_workspace_bytes(int64_t num_prims, int64_t num_rays, int64_t num_wcts);
int ever_trace_backward(const ever_prims* prims, void* workspace, size_t workspace_bytes,
                        const ever_rays* rays, const ever_backward_io* io, void* stream);

/* The adjoint without a stored event list. Replaces the R x max_iters tri_collection that Forward.trace_rays
 * allocates and keeps for the backward (py_binding.cpp:221-223, read back at backwards_kernel.slang:327-361): the
 * forward is called with tri_collection = NULL, and this call traces the rays again, chunk_rays at a time, into one
 * chunk-sized event list inside its workspace, running the adjoint on each chunk before the next. Same inputs and
 * outputs as ever_trace_backward except that io->tri_collection and io->ray_order are ignored; rays must be those of
 * the forward call, and the scene workspace must still hold that call's start-inside candidates (no forward on it in
 * between, as for ever_trace_backward). Gradients equal the stored-list adjoint's up to the order of float atomic
 * additions. The workspace, ever_backward_recompute_workspace_bytes(...) bytes, grows with chunk_rays x max_iters
 * instead of R x max_iters. */
size_t ever_backward_recompute_workspace_bytes(int64_t num_prims, int64_t num_rays, int64_t num_wcts,
                                               int32_t max_iters, int32_t traversal, int64_t chunk_rays);
int ever_trace_backward_recompute(const ever_prims* prims, void* workspace, size_t workspace_bytes,
                                  const ever_rays* rays, const ever_backward_io* io, int64_t chunk_rays,
                                  void* stream);

/* Per-primitive SH colour at normalize(mean - origin), unclamped (sh_kernel.slang:19-42). */
int ever_eval_sh(const float* means, const float* features, const float* origin, int64_t num_prims,
                 int32_t feature_size, int32_t sh_degree, float* out_rgb, void* stream);
/* Kernel timing for measurement tools (no counterpart in the reference). While enabled, every launch of the
 * forward trace kernel (EVER_KERNEL_FORWARD) and of the backward trace kernel (EVER_KERNEL_BACKWARD) is bracketed
 * by a pair of HIP events recorded on its stream. ever_profile_read waits for the recorded pairs of one kernel,
 * returns their summed milliseconds and count, and forgets them. Not thread-safe; host-side only. */
#define EVER_KERNEL_FORWARD 0
#define EVER_KERNEL_BACKWARD 1
int ever_profile_enable(int on);
int ever_profile_read(int kernel, double* total_ms, int64_t* count);

/* Termination-depth proposal network (no counterpart in the reference: the learned-proposal form). A 63-dim
 * encoding of the ray (origin + tmin * direction, unit direction) -> 3 x 128 ReLU -> 1 output y = log(depth),
 * BF16 weights and activations on MFMA with FP32 accumulation. params: FP32 master block of
 * ever_mlp_param_count() floats (W1 [64][128], b1, W2 [128][128], b2, W3 [128][128], b3, w4 [128], b4);
 * params_bf16: its bf16 mirror (same offsets), refreshed by ever_mlp_sync_bf16 and by every training step.
 * Workspaces: ever_mlp_workspace_bytes(rows). */
int64_t ever_mlp_param_count(void);
size_t ever_mlp_workspace_bytes(int64_t num_rays);
int ever_mlp_sync_bf16(const float* params, uint16_t* params_bf16, void* stream);
/* log_depth (R) and/or proposal (R) = exp(log_depth) * margin; either may be null. */
int ever_mlp_infer(const float* params, const uint16_t* params_bf16, const float* origins, const float* directions,
                   int64_t num_rays, float tmin, float margin, float* log_depth, float* proposal, void* workspace,
                   size_t workspace_bytes, void* stream);
/* One Adam step (lr > 0; lr == 0 computes the gradient only) on the pinball loss (quantile tau) of y against
 * log(target_depth); rays with target_depth <= 0 are ignored. loss (1 float, device) and grad (param_count
 * floats, device) are optional outputs. step counts from 1 (bias correction). */
int ever_mlp_train_step(float* params, uint16_t* params_bf16, float* adam_m, float* adam_v, const float* origins,
                        const float* directions, const float* target_depth, int64_t num_rays, float tmin, float tau,
                        float lr, int32_t step, void* workspace, size_t workspace_bytes, float* loss, float* grad,
                        void* stream);
/* The MFMA product the network is built on, exposed for its tests: C[M][N] (FP32, row-major) =
 * sum_k A[m*sam + k*sak] * B[k*sbk + n*sbn] (bf16) + bias[n], optionally ReLU. M, N multiples of 64, K of 16;
 * splits > 1 splits K over blocks through part (splits * M * N floats). */
int ever_gemm_bf16(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn,
                   int64_t M, int64_t N, int64_t K, const float* bias, int32_t relu, float* C, int32_t splits,
                   float* part, void* stream);

/* Camera rays of one image, one ray per pixel in row-major pixel order (r = y * width + x). */
#define EVER_CAMERA_PERSPECTIVE 0
#define EVER_CAMERA_FISHEYE 1   /* equidistant: angle to the optical axis = |camera-plane point|, clipped at pi */
#define EVER_CAMERA_PANORAMIC 2 /* camera-plane point (u, v) = (azimuth, polar angle) */

typedef struct ever_camera {
  int32_t width;
  int32_t height;
  int32_t model;     /* EVER_CAMERA_* */
  int32_t distorted; /* 1: undo the radial (k1..k4) and tangential (p1, p2) distortion (10 Newton steps) */
  float fx, fy;      /* focal lengths in pixels */
  float cx, cy;      /* principal point in pixels; camera-plane point ((x + jx - cx) / fx, (y + jy - cy) / fy) */
  float dist[6];     /* k1, k2, k3, k4, p1, p2 */
  const float* c2w;  /* device, 12 floats: row-major 3x4 camera-to-world [R | t] (OpenCV axes: +z forward) */
} ever_camera;

/* origins (W*H,3) = t; directions (W*H,3) = unit R * camera direction. jitter: device (W*H,2) pixel offsets, or null
 * for pixel centres (0.5, 0.5). */
int ever_camera_rays(const ever_camera* camera, const float* jitter, float* origins, float* directions, void* stream);

/* Primitive arrays from PLY records: records is a device (N, num_columns) float32 array, one row per primitive
 * (the file's vertex payload); the layout names the column of each field. f_dc: 3 columns; f_rest: 3 * num_rest
 * consecutive columns from f_rest0, channel-major (f_rest[c * num_rest + k], as 3D Gaussian Splatting files and
 * save_ply store them). Outputs (N,3) means, (N,3) scales, (N,4) quats, (N) densities, (N, 1 + num_rest, 3)
 * features. EVER_PLY_ACTIVATE_NONE copies stored values (density column); EVER_PLY_ACTIVATE_3DGS applies the
 * viewer's activations to a 3DGS file (opacity column): scales min(softplus(s) + 3e-4, 25), density
 * max(0, -log(max(1 - 0.99 sigmoid(opacity), 1e-10))) / min(2 s), quaternions divided by their squared length.
 * order: (N) source row of each output row (e.g. ever_morton_order), nullable for identity. */
#define EVER_PLY_ACTIVATE_NONE 0
#define EVER_PLY_ACTIVATE_3DGS 1

typedef struct ever_ply_layout {
  int32_t num_columns;
  int32_t mean[3];
  int32_t scale[3];
  int32_t rot[4];
  int32_t f_dc[3];
  int32_t f_rest0;
  int32_t num_rest; /* SH coefficients per channel beyond the DC one: 0, 3, 8 or 15 */
  int32_t opacity;  /* EVER_PLY_ACTIVATE_3DGS */
  int32_t density;  /* EVER_PLY_ACTIVATE_NONE */
  int32_t activation;
} ever_ply_layout;

int ever_scene_from_records(const float* records, int64_t num_prims, const ever_ply_layout* layout,
                            const int32_t* order, float* means, float* scales, float* quats, float* densities,
                            float* features, void* stream);

/* Morton order of the primitives: order[j] = source index of the j-th primitive sorted by the 63-bit Morton code
 * of its mean (21 bits per axis over the means' bounding box); ties keep index order. */
size_t ever_morton_workspace_bytes(int64_t num_prims);
int ever_morton_order(const float* means, int64_t num_prims, int32_t* order, void* workspace, size_t workspace_bytes,
                      void* stream);

/* dL/dfeatures of ever_eval_sh (sh_kernel.slang:45-98); dfeatures is overwritten. */
int ever_eval_sh_backward(const float* means, const float* features, const float* origin, const float* drgb,
                          int64_t num_prims, int32_t feature_size, int32_t sh_degree, float* dfeatures,
                          void* stream);

/* ever_morton_order without a workspace argument, for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). No counterpart in the reference. It uses a scratch array the
 * library holds on each GPU it runs on until the process ends (about 21 MB), so it allocates nothing; num_prims
 * above EVER_MORTON_NOALLOC_MAX_PRIMS is rejected (EVER_ERR_INVALID). Calls that overlap on one GPU share that
 * array and race: serialize them (one stream, or events). Results equal ever_morton_order's. */
#define EVER_MORTON_NOALLOC_MAX_PRIMS 1048576
int ever_morton_order_noalloc(const float* means, int64_t num_prims, int32_t* order, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EVER_AMD_H */
