// C ABI entry points (include/ever_amd.h): argument validation, workspace checks, status codes.
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "ever_internal.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_status(hipError_t e, const char* where) {
  if (e == hipSuccess) return EVER_OK;
  return fail(EVER_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

int check_prims(const ever_prims* P) {
  if (!P) return fail(EVER_ERR_INVALID, "prims is null");
  if (P->num_prims < 0) return fail(EVER_ERR_INVALID, "num_prims must be >= 0");
  if (P->num_prims >= (int64_t(1) << 28))
    return fail(EVER_ERR_INVALID, "num_prims must be < 2^28 (4 leaf fragments per primitive, int32 leaf indices)");
  if (P->num_prims > 0 && (!P->means || !P->scales || !P->quats || !P->densities || !P->features))
    return fail(EVER_ERR_INVALID, "a primitive array is null");
  if (P->feature_size < 1) return fail(EVER_ERR_INVALID, "feature_size must be >= 1");
  if (P->color_mode != EVER_COLOR_DC && P->color_mode != EVER_COLOR_SH)
    return fail(EVER_ERR_INVALID, "unknown color_mode");
  if (P->color_mode == EVER_COLOR_SH) {
    if (P->sh_degree < 0 || P->sh_degree > 3) return fail(EVER_ERR_INVALID, "sh_degree must be in [0, 3]");
    if ((P->sh_degree + 1) * (P->sh_degree + 1) > P->feature_size)
      return fail(EVER_ERR_INVALID, "feature_size is smaller than (sh_degree+1)^2");
    if (P->activation != 0 && P->activation != 1) return fail(EVER_ERR_INVALID, "unknown activation");
  }
  if (P->num_views < 0) return fail(EVER_ERR_INVALID, "num_views must be >= 0");
  if (P->num_views > 1) {
    if (P->color_mode != EVER_COLOR_SH) return fail(EVER_ERR_INVALID, "several views need color_mode EVER_COLOR_SH");
    if (!P->view_origins || (P->num_prims > 0 && !P->view_colors))
      return fail(EVER_ERR_INVALID, "view_origins and view_colors are needed for several views");
    if (P->num_views > 64) return fail(EVER_ERR_INVALID, "num_views must be <= 64");
    if ((int64_t)P->num_views * P->num_prims >= (int64_t(1) << 32))
      return fail(EVER_ERR_INVALID, "num_views * num_prims must be < 2^32");
  }
  return EVER_OK;
}

int check_rays(const ever_rays* R) {
  if (!R) return fail(EVER_ERR_INVALID, "rays is null");
  if (R->num_rays < 0) return fail(EVER_ERR_INVALID, "num_rays must be >= 0");
  if (R->num_rays > 0 && (!R->origins || !R->directions)) return fail(EVER_ERR_INVALID, "ray arrays are null");
  if (R->max_iters < 0) return fail(EVER_ERR_INVALID, "max_iters must be >= 0");
  if (R->kbuf != 8 && R->kbuf != 16 && R->kbuf != 19 && R->kbuf != 32)
    return fail(EVER_ERR_INVALID, "kbuf must be 8, 16, 19 or 32");
  if (R->traversal != EVER_TRAVERSAL_AUTO && R->traversal != EVER_TRAVERSAL_RAY && R->traversal != EVER_TRAVERSAL_WAVE)
    return fail(EVER_ERR_INVALID, "traversal must be EVER_TRAVERSAL_AUTO, _RAY or _WAVE");
  if (R->tile < 0 || R->tile > 64) return fail(EVER_ERR_INVALID, "tile must be in [0, 64]");
  if (R->event_layout != EVER_EVENTS_BY_RAY && R->event_layout != EVER_EVENTS_BY_SLOT)
    return fail(EVER_ERR_INVALID, "event_layout must be EVER_EVENTS_BY_RAY or EVER_EVENTS_BY_SLOT");
  if (R->num_rays > INT32_MAX) return fail(EVER_ERR_INVALID, "num_rays must be < 2^31");
  return EVER_OK;
}

// kernel timing state (ever_profile_enable / ever_profile_read)
bool g_profile_on = false;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_profile_events[2];
hipEvent_t g_profile_open[2] = {nullptr, nullptr};

}  // namespace

namespace ever {

void profile_begin(int kernel, hipStream_t st) {
  if (!g_profile_on) return;
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, st);
  g_profile_open[kernel] = e;
}

void profile_end(int kernel, hipStream_t st) {
  if (!g_profile_on || !g_profile_open[kernel]) return;
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, st);
  g_profile_events[kernel].emplace_back(g_profile_open[kernel], e);
  g_profile_open[kernel] = nullptr;
}

}  // namespace ever

extern "C" {

int ever_profile_enable(int on) {
  g_profile_on = on != 0;
  return EVER_OK;
}

int ever_profile_read(int kernel, double* total_ms, int64_t* count) {
  if (kernel != EVER_KERNEL_FORWARD && kernel != EVER_KERNEL_BACKWARD)
    return fail(EVER_ERR_INVALID, "kernel must be EVER_KERNEL_FORWARD or EVER_KERNEL_BACKWARD");
  if (!total_ms || !count) return fail(EVER_ERR_INVALID, "output pointers are null");
  double sum = 0.0;
  int64_t n = 0;
  int rc = EVER_OK;
  for (auto& pr : g_profile_events[kernel]) {
    float ms = 0.f;
    if (rc == EVER_OK) {
      hipError_t e = hipEventSynchronize(pr.second);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, pr.first, pr.second);
      rc = hip_status(e, "ever_profile_read");
      sum += ms;
      ++n;
    }
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  g_profile_events[kernel].clear();
  *total_ms = sum;
  *count = n;
  return rc;
}


int ever_abi_version(void) { return EVER_ABI_VERSION; }

int ever_struct_sizes(int64_t* out4) {
  if (!out4) return fail(EVER_ERR_INVALID, "out4 is null");
  out4[0] = sizeof(ever_prims);
  out4[1] = sizeof(ever_rays);
  out4[2] = sizeof(ever_forward_out);
  out4[3] = sizeof(ever_backward_io);
  return EVER_OK;
}

const char* ever_last_error(void) { return g_last_error.c_str(); }

int ever_scene_layout(int64_t num_prims, int64_t* o) {
  if (!o) return fail(EVER_ERR_INVALID, "offsets is null");
  ever::SceneLayout L = ever::scene_layout(num_prims > 0 ? num_prims : 0);
  const size_t v[20] = {L.header, L.aabb, L.fprim, L.keys0, L.vals0, L.keys1, L.vals1, L.hist, L.nodes, L.nodes4,
                        L.range, L.sbox, L.pre, L.suf, L.tab, L.rec, L.colr, L.cand, L.bpart, L.total};
  for (int i = 0; i < 20; ++i) o[i] = (int64_t)v[i];
  return EVER_OK;
}

size_t ever_scene_workspace_bytes(int64_t num_prims) {
  return ever::scene_layout(num_prims > 0 ? num_prims : 0).total;
}

size_t ever_forward_workspace_bytes(int64_t num_rays) {
  return ever::forward_layout(num_rays, ever::per_ray_walk(num_rays, EVER_TRAVERSAL_AUTO)).total;
}

size_t ever_forward_workspace_bytes_for(int64_t num_rays, int32_t traversal) {
  return ever::forward_layout(num_rays, ever::per_ray_walk(num_rays, traversal)).total;
}

size_t ever_backward_workspace_bytes(int64_t num_prims, int64_t num_rays, int64_t num_wcts) {
  return ever::backward_layout(num_prims, num_rays, num_wcts).total;
}

int ever_build_scene(const ever_prims* prims, void* workspace, size_t workspace_bytes, void* stream) {
  if (int rc = check_prims(prims)) return rc;
  if (!workspace) return fail(EVER_ERR_INVALID, "workspace is null");
  if (workspace_bytes < ever_scene_workspace_bytes(prims->num_prims))
    return fail(EVER_ERR_WORKSPACE, "scene workspace too small");
  return hip_status(ever::launch_build(*prims, workspace, (hipStream_t)stream), "ever_build_scene");
}

int ever_update_colors(const ever_prims* prims, void* workspace, size_t workspace_bytes, void* stream) {
  if (int rc = check_prims(prims)) return rc;
  if (!workspace) return fail(EVER_ERR_INVALID, "workspace is null");
  if (workspace_bytes < ever_scene_workspace_bytes(prims->num_prims))
    return fail(EVER_ERR_WORKSPACE, "scene workspace too small");
  return hip_status(ever::launch_update_colors(*prims, workspace, (hipStream_t)stream), "ever_update_colors");
}

int ever_trace_forward(const ever_prims* prims, void* workspace, size_t workspace_bytes, const ever_rays* rays,
                       const ever_forward_out* out, void* stream) {
  if (int rc = check_prims(prims)) return rc;
  if (int rc = check_rays(rays)) return rc;
  if (!out) return fail(EVER_ERR_INVALID, "out is null");
  if (!workspace) return fail(EVER_ERR_INVALID, "workspace is null");
  if (workspace_bytes < ever_scene_workspace_bytes(prims->num_prims))
    return fail(EVER_ERR_WORKSPACE, "scene workspace too small");
  if (prims->num_views > 1 && rays->num_rays > 0 && !rays->view_index)
    return fail(EVER_ERR_INVALID, "view_index is needed when num_views > 1");
  if (rays->num_rays > 0) {
    if (!out->color || !out->states || !out->iters || !out->initial_drgb || !out->ray_state64 || !out->ray_order)
      return fail(EVER_ERR_INVALID, "a required forward output is null");
    if (!out->workspace || out->workspace_bytes < ever_forward_workspace_bytes_for(rays->num_rays, rays->traversal))
      return fail(EVER_ERR_WORKSPACE, "forward workspace too small");
  }
  return hip_status(ever::launch_forward(*prims, workspace, *rays, *out, (hipStream_t)stream),
                    "ever_trace_forward");
}

// Checks shared by both adjoints; `stored`: the event list (and, by slot, the ray order) comes from the caller.
static int check_backward(const ever_prims* prims, void* workspace, size_t workspace_bytes, const ever_rays* rays,
                          const ever_backward_io* io, bool stored) {
  if (int rc = check_prims(prims)) return rc;
  if (int rc = check_rays(rays)) return rc;
  if (!io) return fail(EVER_ERR_INVALID, "io is null");
  if (stored && rays->event_layout == EVER_EVENTS_BY_SLOT && rays->num_rays > 0 && !io->ray_order)
    return fail(EVER_ERR_INVALID, "event lists stored by slot need the forward's ray_order");
  if (prims->num_views > 1 && prims->num_prims > 0 && rays->num_rays > 0 && (!rays->view_index || !io->view_color_grads))
    return fail(EVER_ERR_INVALID, "view_index and view_color_grads are needed when num_views > 1");
  if (!workspace) return fail(EVER_ERR_INVALID, "workspace is null");
  if (workspace_bytes < ever_scene_workspace_bytes(prims->num_prims))
    return fail(EVER_ERR_WORKSPACE, "scene workspace too small");
  if (io->wcts && io->num_wcts < 1) return fail(EVER_ERR_INVALID, "num_wcts must be >= 1 when wcts is given");
  if (io->wcts && !io->dmeans2D) return fail(EVER_ERR_INVALID, "dmeans2D is null while wcts is given");
  if (prims->num_prims > 0 && (!io->dmeans || !io->dscales || !io->dquats || !io->ddensities || !io->dfeatures))
    return fail(EVER_ERR_INVALID, "a primitive gradient output is null");
  if (rays->num_rays > 0 && (!io->grad_out || !io->iters || !io->states || !io->ray_state64 || !io->drayo ||
                             !io->drayd || (stored && !io->tri_collection)))
    return fail(EVER_ERR_INVALID, stored ? "a ray input/output of the backward is null (stored events are required)"
                                         : "a ray input/output of the backward is null");
  return EVER_OK;
}

int ever_trace_backward(const ever_prims* prims, void* workspace, size_t workspace_bytes, const ever_rays* rays,
                        const ever_backward_io* io, void* stream) {
  if (int rc = check_backward(prims, workspace, workspace_bytes, rays, io, true)) return rc;
  const int64_t nw = io->wcts ? io->num_wcts : 0;
  if (!io->workspace ||
      io->workspace_bytes < ever_backward_workspace_bytes(prims->num_prims, rays->num_rays, nw))
    return fail(EVER_ERR_WORKSPACE, "backward workspace too small");
  return hip_status(ever::launch_backward(*prims, workspace, *rays, *io, (hipStream_t)stream), "ever_trace_backward");
}

size_t ever_backward_recompute_workspace_bytes(int64_t num_prims, int64_t num_rays, int64_t num_wcts,
                                               int32_t max_iters, int32_t traversal, int64_t chunk_rays) {
  if (num_prims < 0 || num_rays < 0 || num_wcts < 0 || max_iters < 0 || chunk_rays < 1) return 0;
  return ever::recompute_layout(num_prims, num_rays, num_wcts, max_iters, traversal, chunk_rays).total;
}

int ever_trace_backward_recompute(const ever_prims* prims, void* workspace, size_t workspace_bytes,
                                  const ever_rays* rays, const ever_backward_io* io, int64_t chunk_rays,
                                  void* stream) {
  if (int rc = check_backward(prims, workspace, workspace_bytes, rays, io, false)) return rc;
  if (chunk_rays < 1) return fail(EVER_ERR_INVALID, "chunk_rays must be >= 1");
  const int64_t nw = io->wcts ? io->num_wcts : 0;
  if (!io->workspace || io->workspace_bytes < ever_backward_recompute_workspace_bytes(
                                                  prims->num_prims, rays->num_rays, nw, rays->max_iters,
                                                  rays->traversal, chunk_rays))
    return fail(EVER_ERR_WORKSPACE, "backward workspace too small");
  return hip_status(ever::launch_backward_recompute(*prims, workspace, *rays, *io, chunk_rays, (hipStream_t)stream),
                    "ever_trace_backward_recompute");
}

int ever_scene_from_records(const float* records, int64_t num_prims, const ever_ply_layout* layout,
                            const int32_t* order, float* means, float* scales, float* quats, float* densities,
                            float* features, void* stream) {
  if (!layout) return fail(EVER_ERR_INVALID, "layout is null");
  const ever_ply_layout& L = *layout;
  if (num_prims < 0) return fail(EVER_ERR_INVALID, "num_prims must be >= 0");
  if (L.num_columns < 1) return fail(EVER_ERR_INVALID, "num_columns must be >= 1");
  if (L.num_rest != 0 && L.num_rest != 3 && L.num_rest != 8 && L.num_rest != 15)
    return fail(EVER_ERR_INVALID, "num_rest must be 0, 3, 8 or 15");
  if (L.activation != EVER_PLY_ACTIVATE_NONE && L.activation != EVER_PLY_ACTIVATE_3DGS)
    return fail(EVER_ERR_INVALID, "activation must be EVER_PLY_ACTIVATE_NONE or _3DGS");
  auto col_ok = [&](int32_t c) { return c >= 0 && c < L.num_columns; };
  for (int a = 0; a < 3; ++a)
    if (!col_ok(L.mean[a]) || !col_ok(L.scale[a]) || !col_ok(L.f_dc[a]))
      return fail(EVER_ERR_INVALID, "a mean, scale or f_dc column is out of range");
  for (int a = 0; a < 4; ++a)
    if (!col_ok(L.rot[a])) return fail(EVER_ERR_INVALID, "a rot column is out of range");
  if (L.num_rest > 0 && (!col_ok(L.f_rest0) || !col_ok(L.f_rest0 + 3 * L.num_rest - 1)))
    return fail(EVER_ERR_INVALID, "the f_rest columns are out of range");
  if (!col_ok(L.activation == EVER_PLY_ACTIVATE_3DGS ? L.opacity : L.density))
    return fail(EVER_ERR_INVALID, "the opacity/density column is out of range");
  if (num_prims > 0 && (!records || !means || !scales || !quats || !densities || !features))
    return fail(EVER_ERR_INVALID, "null pointer");
  if (num_prims > INT32_MAX) return fail(EVER_ERR_INVALID, "num_prims must be < 2^31");
  return hip_status(ever::launch_scene_from_records(records, num_prims, L, order, means, scales, quats, densities,
                                                    features, (hipStream_t)stream),
                    "ever_scene_from_records");
}

size_t ever_morton_workspace_bytes(int64_t num_prims) {
  return num_prims < 0 ? 0 : ever::morton_layout(num_prims).total;
}

int ever_morton_order(const float* means, int64_t num_prims, int32_t* order, void* workspace, size_t workspace_bytes,
                      void* stream) {
  if (num_prims < 0) return fail(EVER_ERR_INVALID, "num_prims must be >= 0");
  if (num_prims > INT32_MAX) return fail(EVER_ERR_INVALID, "num_prims must be < 2^31");
  if (num_prims > 0 && (!means || !order || !workspace)) return fail(EVER_ERR_INVALID, "null pointer");
  if (workspace_bytes < ever_morton_workspace_bytes(num_prims))
    return fail(EVER_ERR_WORKSPACE, "morton workspace too small");
  return hip_status(ever::launch_morton_order(means, num_prims, order, workspace, (hipStream_t)stream),
                    "ever_morton_order");
}

int ever_morton_order_noalloc(const float* means, int64_t num_prims, int32_t* order, void* stream) {
  if (num_prims < 0) return fail(EVER_ERR_INVALID, "num_prims must be >= 0");
  if (num_prims > EVER_MORTON_NOALLOC_MAX_PRIMS)
    return fail(EVER_ERR_INVALID, "num_prims exceeds EVER_MORTON_NOALLOC_MAX_PRIMS (use ever_morton_order with a workspace)");
  if (num_prims > 0 && (!means || !order)) return fail(EVER_ERR_INVALID, "null pointer");
  if (num_prims == 0) return EVER_OK;
  return hip_status(ever::launch_morton_order_noalloc(means, num_prims, order, (hipStream_t)stream),
                    "ever_morton_order_noalloc");
}

int ever_camera_rays(const ever_camera* camera, const float* jitter, float* origins, float* directions,
                     void* stream) {
  if (!camera) return fail(EVER_ERR_INVALID, "camera is null");
  const ever_camera& C = *camera;
  if (C.width < 0 || C.height < 0) return fail(EVER_ERR_INVALID, "width and height must be >= 0");
  if ((int64_t)C.width * C.height > INT32_MAX) return fail(EVER_ERR_INVALID, "width * height must be < 2^31");
  if (C.model != EVER_CAMERA_PERSPECTIVE && C.model != EVER_CAMERA_FISHEYE && C.model != EVER_CAMERA_PANORAMIC)
    return fail(EVER_ERR_INVALID, "model must be EVER_CAMERA_PERSPECTIVE, _FISHEYE or _PANORAMIC");
  if (C.distorted != 0 && C.distorted != 1) return fail(EVER_ERR_INVALID, "distorted must be 0 or 1");
  if (!(C.fx != 0.f && C.fy != 0.f) || !std::isfinite(C.fx) || !std::isfinite(C.fy))
    return fail(EVER_ERR_INVALID, "focal lengths must be finite and non-zero");
  if ((int64_t)C.width * C.height > 0 && (!C.c2w || !origins || !directions))
    return fail(EVER_ERR_INVALID, "null pointer");
  return hip_status(ever::launch_camera_rays(C, jitter, origins, directions, (hipStream_t)stream),
                    "ever_camera_rays");
}

int ever_eval_sh(const float* means, const float* features, const float* origin, int64_t num_prims,
                 int32_t feature_size, int32_t sh_degree, float* out_rgb, void* stream) {
  if (num_prims < 0) return fail(EVER_ERR_INVALID, "num_prims must be >= 0");
  if (sh_degree < 0 || sh_degree > 3) return fail(EVER_ERR_INVALID, "sh_degree must be in [0, 3]");
  if ((sh_degree + 1) * (sh_degree + 1) > feature_size)
    return fail(EVER_ERR_INVALID, "feature_size is smaller than (sh_degree+1)^2");
  if (num_prims > 0 && (!means || !features || !origin || !out_rgb)) return fail(EVER_ERR_INVALID, "null pointer");
  return hip_status(ever::launch_eval_sh(means, features, origin, num_prims, feature_size, sh_degree, out_rgb,
                                         (hipStream_t)stream),
                    "ever_eval_sh");
}

int ever_eval_sh_backward(const float* means, const float* features, const float* origin, const float* drgb,
                          int64_t num_prims, int32_t feature_size, int32_t sh_degree, float* dfeatures,
                          void* stream) {
  if (num_prims < 0) return fail(EVER_ERR_INVALID, "num_prims must be >= 0");
  if (sh_degree < 0 || sh_degree > 3) return fail(EVER_ERR_INVALID, "sh_degree must be in [0, 3]");
  if ((sh_degree + 1) * (sh_degree + 1) > feature_size)
    return fail(EVER_ERR_INVALID, "feature_size is smaller than (sh_degree+1)^2");
  if (num_prims > 0 && (!means || !origin || !drgb || !dfeatures)) return fail(EVER_ERR_INVALID, "null pointer");
  return hip_status(ever::launch_eval_sh_bwd(means, features, origin, drgb, num_prims, feature_size, sh_degree,
                                             dfeatures, (hipStream_t)stream),
                    "ever_eval_sh_backward");
}

int64_t ever_mlp_param_count(void) { return ever::mlp_layout().count; }

size_t ever_mlp_workspace_bytes(int64_t num_rays) {
  if (num_rays < 0) return 0;
  return (size_t)ever::mlp_work((num_rays + 63) / 64 * 64).total;
}

int ever_mlp_sync_bf16(const float* params, uint16_t* params_bf16, void* stream) {
  if (!params || !params_bf16) return fail(EVER_ERR_INVALID, "null pointer");
  return hip_status(ever::launch_mlp_sync_bf16(params, params_bf16, (hipStream_t)stream), "ever_mlp_sync_bf16");
}

int ever_mlp_infer(const float* params, const uint16_t* params_bf16, const float* origins, const float* directions,
                   int64_t num_rays, float tmin, float margin, float* log_depth, float* proposal, void* workspace,
                   size_t workspace_bytes, void* stream) {
  if (num_rays < 0) return fail(EVER_ERR_INVALID, "num_rays must be >= 0");
  if (num_rays == 0) return EVER_OK;
  if (!params || !params_bf16 || !origins || !directions || !workspace) return fail(EVER_ERR_INVALID, "null pointer");
  if (workspace_bytes < ever_mlp_workspace_bytes(num_rays))
    return fail(EVER_ERR_WORKSPACE, "mlp workspace too small");
  return hip_status(ever::launch_mlp_infer(params, params_bf16, origins, directions, num_rays, tmin, margin,
                                           log_depth, proposal, workspace, (hipStream_t)stream),
                    "ever_mlp_infer");
}

int ever_mlp_train_step(float* params, uint16_t* params_bf16, float* adam_m, float* adam_v, const float* origins,
                        const float* directions, const float* target_depth, int64_t num_rays, float tmin, float tau,
                        float lr, int32_t step, void* workspace, size_t workspace_bytes, float* loss, float* grad,
                        void* stream) {
  if (num_rays <= 0) return fail(EVER_ERR_INVALID, "num_rays must be > 0");
  if (!params || !params_bf16 || !origins || !directions || !target_depth || !workspace)
    return fail(EVER_ERR_INVALID, "null pointer");
  if (lr > 0.f && (!adam_m || !adam_v || step < 1)) return fail(EVER_ERR_INVALID, "Adam state and step >= 1 needed");
  if (!(tau > 0.f && tau < 1.f)) return fail(EVER_ERR_INVALID, "tau must be in (0, 1)");
  if (workspace_bytes < ever_mlp_workspace_bytes(num_rays))
    return fail(EVER_ERR_WORKSPACE, "mlp workspace too small");
  return hip_status(ever::launch_mlp_train_step(params, params_bf16, adam_m, adam_v, origins, directions,
                                                target_depth, num_rays, tmin, tau, lr, step, workspace, loss, grad,
                                                (hipStream_t)stream),
                    "ever_mlp_train_step");
}

int ever_gemm_bf16(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn,
                   int64_t M, int64_t N, int64_t K, const float* bias, int32_t relu, float* C, int32_t splits,
                   float* part, void* stream) {
  if (M <= 0 || N <= 0 || K <= 0 || M % 64 || N % 64 || K % 16)
    return fail(EVER_ERR_INVALID, "M, N must be positive multiples of 64 and K of 16");
  if (!A || !B || !C || splits < 1 || (splits > 1 && !part)) return fail(EVER_ERR_INVALID, "null pointer or splits");
  return hip_status(ever::launch_gemm_public(A, sam, sak, B, sbk, sbn, M, N, K, bias, relu, C, splits, part,
                                             (hipStream_t)stream),
                    "ever_gemm_bf16");
}

}  // extern "C"
