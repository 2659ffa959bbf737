/*
 * qsdvr_hip.h -- C ABI of the MI355X (gfx950) QSDVR render hot path.
 *
 * One entry point per function the reference binds in kernel/qsdvr.cpp:5-50,
 * plus the fused renderer and the proposal network, which have no reference
 * counterpart.  Conventions shared by every entry point:
 *
 *  - Plain pointers and sizes only.  Pointers are device pointers unless the
 *    argument says "host".  Sizes are 64-bit element counts.
 *  - The caller owns all memory.  Entry points never allocate; scratch comes
 *    in through a (workspace, workspace_bytes) pair sized by the matching
 *    *_workspace_size() call.
 *  - Forward ops overwrite their outputs.  The three backward ops the
 *    reference documents as accumulating (GridInterpolationBackward,
 *    SampleSDFBackward, SampleNormalBackward) add into a caller-zeroed grid;
 *    the others overwrite.
 *  - Work is enqueued on `stream` (a hipStream_t; NULL = the null stream).
 *    Nothing synchronises the host except qsdvr_ray_cut (it must return the
 *    two sample totals) and qsdvr_layer_to_grid_backward with host outputs.
 *  - Every call returns QSDVR_OK (0) or a negative status; the message of
 *    the calling thread's last failure is available from qsdvr_last_error().
 *  - Sample / cell indices read from caller buffers are range-checked in the
 *    kernels: an out-of-range index yields NaN outputs (forward) or no
 *    scatter (backward) instead of an out-of-bounds access.
 *  - No global mutable state besides the thread-local error message;
 *    calls on different streams may run concurrently.
 */
#ifndef QSDVR_HIP_H
#define QSDVR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSDVR_ABI_VERSION 1

enum {
    QSDVR_OK = 0,
    QSDVR_EINVAL = -1,    /* bad argument (null pointer, size, shape)    */
    QSDVR_EHIP = -2,      /* a HIP runtime call or launch failed          */
    QSDVR_ERANGE = -3,    /* a total or size exceeds what the format holds */
    QSDVR_EWORKSPACE = -4 /* workspace missing or too small               */
};

/* Pinhole camera, the fields of CameraArgs (include/camera.h:4-17):
 * row-major 3x3 camera-to-world rotation, translation, focal lengths,
 * principal point (cx = -W/2, cy = -H/2 in the reference's datasets) and
 * near/far camera-space depths. */
typedef struct qsdvr_camera {
    uint32_t width;
    uint32_t height;
    float rotation[9];
    float translation[3];
    float fx, fy, cx, cy;
    float near_depth, far_depth;
} qsdvr_camera;

typedef void *qsdvr_stream; /* hipStream_t */

int qsdvr_abi_version(void);
const char *qsdvr_last_error(void);

/* ------------------------------------------------------------ camera --- */

/* Replaces RayCut (kernel/camera.cu:98-130, bound at kernel/qsdvr.cpp:32,
 * prototype include/camera.h:57).  Per pixel: ray direction, first in-cube
 * step (origin) and the count of steps strictly inside +-0.9999; then, over
 * the masked pixels in order, sdfCount (<2 -> 0), renderCount =
 * max(sdfCount-1, 0) and both exclusive offsets.  ray_list is int32 [P,4]
 * {sdfCount, sdfOffset, renderCount, renderOffset}; unmasked rows receive
 * only column 0 (the raw count).  mask_host: host bool/uint8 [P] or NULL
 * (all pixels).  Writes the two totals to *sdf_total / *render_total (host)
 * after synchronising the stream. */
size_t qsdvr_ray_cut_workspace_size(int64_t pixel_count);
int qsdvr_ray_cut(const qsdvr_camera *cam, const uint8_t *mask_host,
                  int32_t *ray_list, float *origin, float *dir,
                  float interval, int64_t pixel_count,
                  void *workspace, size_t workspace_bytes,
                  int64_t *sdf_total, int64_t *render_total,
                  qsdvr_stream stream);

/* No counterpart in the reference: qsdvr_ray_cut for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, semantics and results as
 * qsdvr_ray_cut, but the workspace is a static device array of the library
 * (8.07 MiB per GPU, held until the process ends) instead of a caller
 * buffer; the call allocates nothing.  At most
 * qsdvr_ray_cut_noalloc_max_pixels() = 2^23 pixels (larger: QSDVR_ERANGE,
 * nothing written).  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other host threads) race. */
int64_t qsdvr_ray_cut_noalloc_max_pixels(void);
int qsdvr_ray_cut_noalloc(const qsdvr_camera *cam, const uint8_t *mask_host,
                          int32_t *ray_list, float *origin, float *dir,
                          float interval, int64_t pixel_count,
                          int64_t *sdf_total, int64_t *render_total,
                          qsdvr_stream stream);

/* Replaces GenerateRayPoints (kernel/camera.cu:132-146, bound at
 * kernel/qsdvr.cpp:34, prototype include/camera.h:59).  For each of the
 * ray_count rows of ray_list: sdf samples ori + (dir*j)*interval and render
 * samples ori + dir*((j+0.5)*interval), mapped to the cell index
 * x + yR + zR^2 and local point in [-1,1)^3 with R = base_reso.  Bit-exact
 * float32 (no contraction).  sdf_total / render_total are the output
 * lengths (the RayCut totals) and bound every write. */
int qsdvr_generate_ray_points(float *sdf_point, int32_t *sdf_index,
                              float *render_point, int32_t *render_index,
                              float *view_dir, const int32_t *ray_list,
                              const float *origin, const float *dir,
                              int64_t ray_count, int64_t sdf_total,
                              int64_t render_total, float interval,
                              int32_t base_reso, qsdvr_stream stream);

/* ----------------------------------------------------------- surface --- */

/* Replaces LayerToGridForward (kernel/surface.cu:148-229, bound at
 * kernel/qsdvr.cpp:12, prototype include/surface.h:4).  Layers x,y,z [R]
 * and offset [4] are host pointers when layers_on_host != 0 (as in the
 * reference), device pointers otherwise.  grid_out: [R,R,R,7]. */
size_t qsdvr_layer_to_grid_workspace_size(int32_t reso);
int qsdvr_layer_to_grid_forward(float *grid_out, const float *x_layer,
                                const float *y_layer, const float *z_layer,
                                const float *offset, int32_t reso,
                                int layers_on_host, void *workspace,
                                size_t workspace_bytes, qsdvr_stream stream);

/* Replaces LayerToGridBackward (kernel/surface.cu:231-341, bound at
 * kernel/qsdvr.cpp:14, prototype include/surface.h:6).  grad_grid
 * [R,R,R,7] (device) -> gradients of x,y,z [R] and offset [4] (overwrite;
 * host pointers when outputs_on_host != 0, then the stream is synchronised).
 * offset[3] receives the full sum of the G gradients. Deterministic. */
int qsdvr_layer_to_grid_backward(const float *grad_grid, float *gx, float *gy,
                                 float *gz, float *goffset, int32_t reso,
                                 int outputs_on_host, void *workspace,
                                 size_t workspace_bytes, qsdvr_stream stream);

/* Replaces SampleSDFForward (kernel/surface.cu:343-350, bound at
 * kernel/qsdvr.cpp:16, prototype include/surface.h:8):
 * out[i] = f(p)/||grad f(p)||/reso for the quadric of cell index[i]. */
int qsdvr_sample_sdf_forward(float *out, const float *sdf_grid,
                             const float *point, const int32_t *index,
                             int64_t n, int64_t cell_count, float reso,
                             qsdvr_stream stream);

/* Replaces SampleSDFBackward (kernel/surface.cu:352-359, bound at
 * kernel/qsdvr.cpp:18, prototype include/surface.h:10): accumulates
 * d sdf / d(A..G) * grad_out into the caller-zeroed grad_grid. */
int qsdvr_sample_sdf_backward(const float *grad_out, const float *sdf_grid,
                              const float *point, const int32_t *index,
                              float *grad_grid, int64_t n, int64_t cell_count,
                              float reso, qsdvr_stream stream);

/* Replaces SampleNormalForward (kernel/surface.cu:361-367, bound at
 * kernel/qsdvr.cpp:20, prototype include/surface.h:12):
 * out = grad f / max(||grad f||, 1e-12), [n,3]. */
int qsdvr_sample_normal_forward(float *out, const float *sdf_grid,
                                const float *point, const int32_t *index,
                                int64_t n, int64_t cell_count,
                                qsdvr_stream stream);

/* Replaces SampleNormalBackward (kernel/surface.cu:369-375, bound at
 * kernel/qsdvr.cpp:22, prototype include/surface.h:14): accumulates into
 * the caller-zeroed grad_grid (A..F; G receives nothing). */
int qsdvr_sample_normal_backward(const float *grad_out, const float *sdf_grid,
                                 const float *point, const int32_t *index,
                                 float *grad_grid, int64_t n,
                                 int64_t cell_count, qsdvr_stream stream);

/* ------------------------------------------------------------ render --- */

/* Replaces GridInterpolationForward (kernel/render.cu:317-323, bound at
 * kernel/qsdvr.cpp:8, prototype include/render.h:4).  vertex_reso is what
 * the reference passes (dataGrid.size(0) = R+1); rows are addressed as in
 * models/render_torch.py:128-143: base row = cell index, neighbours +1, +R,
 * +R^2, weights = raw local point.  data_grid [(R+1)^3, channels],
 * out [n, channels]. */
int qsdvr_grid_interpolation_forward(const float *data_grid, const float *point,
                                     const int32_t *index, float *out,
                                     int64_t n, int32_t vertex_reso,
                                     int32_t channels, qsdvr_stream stream);

/* Replaces GridInterpolationBackward (kernel/render.cu:325-331, bound at
 * kernel/qsdvr.cpp:10, prototype include/render.h:6).  Same argument roles
 * as the reference: grad_grid is the caller-zeroed accumulator,
 * grad_out [n, channels] the upstream gradient. */
int qsdvr_grid_interpolation_backward(float *grad_grid, const float *point,
                                      const int32_t *index,
                                      const float *grad_out, int64_t n,
                                      int32_t vertex_reso, int32_t channels,
                                      qsdvr_stream stream);

/* Replaces ShaderForward (kernel/render.cu:333-339, bound at
 * kernel/qsdvr.cpp:24, prototype include/render.h:8).  rgb = ao*(diffuse*
 * (1-metallic) + SH(reflect(view, n))*metallic); specular = SH(...).
 * channels 32 (SH degree <= 2, coefficient k colour c at 3k+c) or 53
 * (degree 3). */
int qsdvr_shader_forward(float *out, const float *normal, const float *view_dir,
                         const float *data, float *specular, int64_t n,
                         int32_t channels, qsdvr_stream stream);

/* Replaces ShaderBackward (kernel/render.cu:341-347, bound at
 * kernel/qsdvr.cpp:26, prototype include/render.h:10).  Overwrites
 * normal_grad [n,3] and data_grad [n,channels]. */
int qsdvr_shader_backward(const float *grad_out, const float *normal,
                          const float *view_dir, const float *data,
                          const float *specular, float *normal_grad,
                          float *data_grad, int64_t n, int32_t channels,
                          qsdvr_stream stream);

/* Replaces RayAggregateForward (kernel/render.cu:349-355, bound at
 * kernel/qsdvr.cpp:28, prototype include/render.h:12).  NeuS logistic
 * compositing per ray of ray_list; writes colour [rays,3] and the per
 * render-sample alpha and transmittance T. */
int qsdvr_ray_aggregate_forward(float *out, const float *rgb, const float *sdf,
                                const int32_t *ray_list, float *alpha, float *T,
                                int64_t ray_count, int64_t sdf_total,
                                int64_t render_total, float logistic_coef,
                                qsdvr_stream stream);

/* Replaces RayAggregateBackward (kernel/render.cu:357-363, bound at
 * kernel/qsdvr.cpp:30, prototype include/render.h:14).  Overwrites
 * rgb_grad [S_r,3] and sdf_grad [S_sdf] on the rows of non-empty rays; the
 * alpha adjoint uses the division-free suffix recurrence. */
int qsdvr_ray_aggregate_backward(const float *grad_out, const float *rgb,
                                 const float *sdf, const int32_t *ray_list,
                                 const float *alpha, const float *T,
                                 float *rgb_grad, float *sdf_grad,
                                 int64_t ray_count, int64_t sdf_total,
                                 int64_t render_total, float logistic_coef,
                                 qsdvr_stream stream);

/* ---------------------------------------------- sorted-segment scatters --- */

/* Sorted-segment forms of the three accumulating backward ops, with the same
 * arguments and the same result (added into the caller-zeroed grid) as
 * qsdvr_grid_interpolation_backward (kernel/render.cu:325-331),
 * qsdvr_sample_sdf_backward (kernel/surface.cu:352-359) and
 * qsdvr_sample_normal_backward (kernel/surface.cu:369-375).  The samples
 * are radix-sorted by cell (SampleSDF sorts its point and upstream gradient
 * with the keys, GridInterpolation its point and sample index, SampleNormal
 * sample indices; GridInterpolation's keys put the cells in 8 parity classes)
 * and every cell's samples are summed in a fixed order before one
 * read-modify-write of the cell's rows (the appearance in 8 cell-parity
 * passes), instead of one float atomic per sample and coefficient: bitwise
 * reproducible.  Workspace:
 * qsdvr_sorted_scatter_workspace_size(n, cells) bytes (about 41 B per
 * sample plus the sort's temporary storage) with cells = R^3
 * (cell_count for the quadric ops, (vertex_reso - 1)^3 for the
 * interpolation).  GridInterpolation: channels 32 or 53 and an even cell
 * resolution (with the stride-R vertex rows of the reference, cells of one
 * parity class would share a row at an odd resolution); the atomic op
 * covers the rest. */
size_t qsdvr_sorted_scatter_workspace_size(int64_t n, int64_t cell_count);
int qsdvr_grid_interpolation_backward_sorted(float *grad_grid, const float *point,
                                             const int32_t *index,
                                             const float *grad_out, int64_t n,
                                             int32_t vertex_reso, int32_t channels,
                                             void *workspace, size_t workspace_bytes,
                                             qsdvr_stream stream);
int qsdvr_sample_sdf_backward_sorted(const float *grad_out, const float *sdf_grid,
                                     const float *point, const int32_t *index,
                                     float *grad_grid, int64_t n, int64_t cell_count,
                                     float reso, void *workspace,
                                     size_t workspace_bytes, qsdvr_stream stream);
int qsdvr_sample_normal_backward_sorted(const float *grad_out, const float *sdf_grid,
                                        const float *point, const int32_t *index,
                                        float *grad_grid, int64_t n,
                                        int64_t cell_count, void *workspace,
                                        size_t workspace_bytes, qsdvr_stream stream);
/* SampleNormalBackward (kernel/surface.cu:369-375, bound at kernel/qsdvr.cpp:22)
 * on the sort a preceding qsdvr_grid_interpolation_backward_sorted call made
 * of the same point and index lists: `workspace` is that call's workspace
 * (n samples, cell resolution reso = vertex_reso - 1, even), which holds
 * the samples' points and indices in (parity class, cell) order -- equal
 * cells grouped as the cell order groups them.  The points come from there,
 * so the reference's pointList and indexList arguments are not taken.  Same
 * sums as qsdvr_sample_normal_backward_sorted, without its own sort; the
 * two calls must run on one stream, and the workspace must not change in
 * between. */
int qsdvr_sample_normal_backward_presorted(const float *grad_out, const float *sdf_grid,
                                           float *grad_grid, int64_t n, int32_t reso,
                                           void *workspace, size_t workspace_bytes,
                                           qsdvr_stream stream);
/* GridInterpolationForward (kernel/render.cu:317-323, bound at
 * kernel/qsdvr.cpp:8, prototype include/render.h:4) that first makes GridInterpolationBackward's sort of
 * its samples (the sort qsdvr_grid_interpolation_backward_sorted makes) in
 * `workspace` (qsdvr_sorted_scatter_workspace_size(n, (vertex_reso - 1)^3)
 * bytes), then interpolates one run of a cell's samples at a time in that
 * order.  Outputs equal qsdvr_grid_interpolation_forward's bit for bit (NaN
 * rows for invalid indices).  32 channels, an even cell resolution, 16-B
 * aligned data_grid and out. */
int qsdvr_grid_interpolation_forward_sorted(const float *data_grid, const float *point,
                                            const int32_t *index, float *out, int64_t n,
                                            int32_t vertex_reso, int32_t channels,
                                            void *workspace, size_t workspace_bytes,
                                            qsdvr_stream stream);
/* GridInterpolationBackward (kernel/render.cu:325-331, bound at
 * kernel/qsdvr.cpp:10, prototype include/render.h:6) on the sort that a preceding
 * qsdvr_grid_interpolation_forward_sorted call made of the same point and
 * index lists: `workspace` is that call's, unchanged since (the points and
 * indices come from there, so the reference's pointList and indexList are
 * not taken).  Same sums as qsdvr_grid_interpolation_backward_sorted; the
 * workspace then also serves qsdvr_sample_normal_backward_presorted. */
int qsdvr_grid_interpolation_backward_presorted(float *grad_grid, const float *grad_out,
                                                int64_t n, int32_t vertex_reso,
                                                int32_t channels, void *workspace,
                                                size_t workspace_bytes, qsdvr_stream stream);

/* ------------------------------------------------------------- fused --- */

/* No reference counterpart.  The whole plain render of kernel/camera.cu
 * ray generation + SampleSDF + SampleNormal + GridInterpolation + Shader +
 * RayAggregate in one pass.  Rays: when ray_pixel == NULL, every pixel of
 * the n_cams cameras (all width x height), in row-major order per camera;
 * otherwise ray_pixel[i] = cam * W*H + pixel for n_rays rays.  Outputs per
 * ray: color [n_rays,3], depth [n_rays] (sum of T_i alpha_i times the
 * distance from the camera centre to render sample i; nullable).
 * cams is a DEVICE array of n_cams qsdvr_camera; a camera whose size differs
 * from width x height renders black.  Sample positions and cell indices are
 * those of RayCut + GenerateRayPoints, bit for bit. */
int qsdvr_fused_forward(const qsdvr_camera *cams, int32_t n_cams,
                        int32_t width, int32_t height,
                        const int32_t *ray_pixel, int64_t n_rays,
                        const float *sdf_grid, const float *data_grid,
                        int32_t reso, int32_t channels, float interval,
                        float logistic_coef, float *color, float *depth,
                        qsdvr_stream stream);

/* Adjoint of qsdvr_fused_forward.  Recomputes alpha and T front to back
 * from the saved per-ray colour/depth (no per-sample storage).  Accumulates
 * into the caller-zeroed sdf_grid_grad [R^3,7] and data_grid_grad
 * [(R+1)^3, channels].  grad_depth may be NULL.  scatter_mode:
 * 0 = tile-reduced atomics, 1 = plain per-sample atomics. */
int qsdvr_fused_backward(const qsdvr_camera *cams, int32_t n_cams,
                         int32_t width, int32_t height,
                         const int32_t *ray_pixel, int64_t n_rays,
                         const float *sdf_grid, const float *data_grid,
                         int32_t reso, int32_t channels, float interval,
                         float logistic_coef, const float *color,
                         const float *depth, const float *grad_color,
                         const float *grad_depth, float *sdf_grid_grad,
                         float *data_grid_grad, int32_t scatter_mode,
                         qsdvr_stream stream);


/* Like qsdvr_fused_forward, and also accumulates every ray's compositing
 * weights T_i alpha_i into 64 equal bins over its sampled segment:
 * bins [n_rays, 64] (caller-zeroed).  These are the proposal's training
 * targets.  No reference counterpart. */
int qsdvr_fused_forward_bins(const qsdvr_camera *cams, int32_t n_cams,
                             int32_t width, int32_t height,
                             const int32_t *ray_pixel, int64_t n_rays,
                             const float *sdf_grid, const float *data_grid,
                             int32_t reso, int32_t channels, float interval,
                             float logistic_coef, float *color, float *depth,
                             float *bins, qsdvr_stream stream);

/* The bins of qsdvr_fused_forward_bins from the SDF alone (no appearance,
 * no colour): bins [n_rays, 64] (overwritten) of the rays whose RayCut
 * state is ray_state [n_rays, 8] (qsdvr_ray_state).  Transmittance is
 * multiplied in a different order than the sequential walk, so the
 * weights agree with qsdvr_fused_forward_bins to float32 rounding.  No
 * reference counterpart (the proposal's training targets). */
int qsdvr_plain_bins(const float *ray_state, int64_t n_rays,
                     const float *sdf_grid, int32_t reso, float interval,
                     float logistic_coef, float *bins, qsdvr_stream stream);

/* Per-ray state of the RayCut of every ray (same ray selection as the
 * fused entry points): ray_state [n_rays, 8] float = (origin xyz, in-cube
 * step count as uint32 bits, dir xyz, 0).  No reference counterpart. */
int qsdvr_ray_state(const qsdvr_camera *cams, int32_t n_cams, int32_t width,
                    int32_t height, const int32_t *ray_pixel, int64_t n_rays,
                    float interval, float *ray_state, qsdvr_stream stream);

/* Fused render with explicit sample positions: for every ray with at least
 * two in-cube steps, K sdf samples at origin + dir * tlist[ray, k] (tlist
 * [n_rays, K], increasing) and K-1 render samples at the interval midpoints;
 * compositing as in RayAggregate.  tlist == NULL renders the uniform plain
 * samples from the ray state.  No reference counterpart. */
int qsdvr_fused_forward_listed(const qsdvr_camera *cams, int32_t n_cams,
                               int32_t width, int32_t height,
                               const int32_t *ray_pixel, int64_t n_rays,
                               const float *ray_state, const float *tlist,
                               int32_t K, const float *sdf_grid,
                               const float *data_grid, int32_t reso,
                               int32_t channels, float interval,
                               float logistic_coef, float *color,
                               float *depth, qsdvr_stream stream);

int qsdvr_fused_backward_listed(const qsdvr_camera *cams, int32_t n_cams,
                                int32_t width, int32_t height,
                                const int32_t *ray_pixel, int64_t n_rays,
                                const float *ray_state, const float *tlist,
                                int32_t K, const float *sdf_grid,
                                const float *data_grid, int32_t reso,
                                int32_t channels, float interval,
                                float logistic_coef, const float *color,
                                const float *depth, const float *grad_color,
                                const float *grad_depth, float *sdf_grid_grad,
                                float *data_grid_grad, int32_t scatter_mode,
                                qsdvr_stream stream);

/* Cost-ordered tile adjoint of the fused renderer (no reference counterpart;
 * the reference has no fused renderer).  The tile adjoint's time per 8x8
 * tile varies about 4x across a frame, and a launch in a fixed order leaves
 * its heaviest tiles to finish last.  qsdvr_fused_forward_costed is
 * qsdvr_fused_forward on whole frames (or qsdvr_fused_forward_listed, when
 * ray_state and tlist are given) that also writes each tile's expected
 * backward cost to tile_cost[n_cams * ceil(W/8) * ceil(H/8)] (index =
 * camera * tiles + row-major tile).  qsdvr_fused_backward_ordered is the
 * tile-mode qsdvr_fused_backward (or _listed) that sorts the tiles by that
 * cost in its workspace (qsdvr_fused_order_workspace_size bytes) and runs
 * them heaviest first: the same gradient sums, with its float atomics in
 * another order. */
size_t qsdvr_fused_order_workspace_size(int32_t n_cams, int32_t width, int32_t height);
int qsdvr_fused_forward_costed(const qsdvr_camera *cams, int32_t n_cams,
                               int32_t width, int32_t height,
                               const float *ray_state, const float *tlist,
                               int32_t K, const float *sdf_grid,
                               const float *data_grid, int32_t reso,
                               int32_t channels, float interval,
                               float logistic_coef, float *color, float *depth,
                               uint32_t *tile_cost, qsdvr_stream stream);
int qsdvr_fused_backward_ordered(const qsdvr_camera *cams, int32_t n_cams,
                                 int32_t width, int32_t height,
                                 const float *ray_state, const float *tlist,
                                 int32_t K, const float *sdf_grid,
                                 const float *data_grid, int32_t reso,
                                 int32_t channels, float interval,
                                 float logistic_coef, const float *color,
                                 const float *depth, const float *grad_color,
                                 const float *grad_depth, float *sdf_grid_grad,
                                 float *data_grid_grad, const uint32_t *tile_cost,
                                 void *workspace, size_t workspace_bytes,
                                 qsdvr_stream stream);

/* Deterministic adjoint (sorted-segment gradient scatter).  No reference
 * counterpart: the reference's backward kernels add every sample's gradient
 * with float atomics (kernel/render.cu:121-152, kernel/surface.cu:74-146),
 * whose order, and so whose rounding, varies run to run.  Same rays,
 * samples and result as qsdvr_fused_backward_listed (tlist == NULL: the
 * uniform samples), but the sums into sdf_grid_grad and data_grid_grad are
 * bitwise reproducible: per-sample gradient records are radix-sorted by
 * cell and every cell's run is summed in a fixed order (appearance runs in
 * 8 cell-parity passes without atomics).  ray_state comes from
 * qsdvr_ray_state.  sample_capacity bounds the total sdf samples of all
 * rays (sum over rays of count, or of K when listed, for rays with count
 * >= 2); the workspace holds qsdvr_fused_det_workspace_size() bytes for it.
 * reso even and <= 512.  Synchronises the host once (reads back the sample total);
 * QSDVR_EWORKSPACE if the total exceeds sample_capacity. */
size_t qsdvr_fused_det_workspace_size(int64_t n_rays, int64_t sample_capacity,
                                      int32_t channels, int32_t reso);
int qsdvr_fused_backward_det(const qsdvr_camera *cams, int32_t n_cams,
                             int32_t width, int32_t height,
                             const int32_t *ray_pixel, int64_t n_rays,
                             const float *ray_state, const float *tlist,
                             int32_t K, const float *sdf_grid,
                             const float *data_grid, int32_t reso,
                             int32_t channels, float interval,
                             float logistic_coef, const float *color,
                             const float *depth, const float *grad_color,
                             const float *grad_depth, float *sdf_grid_grad,
                             float *data_grid_grad, int64_t sample_capacity,
                             void *workspace, size_t workspace_bytes,
                             qsdvr_stream stream);

/* ---------------------------------------------------------- proposal --- */
/* No reference counterpart.  Per-point MLP PE(39) -> 64 -> 64 -> 64 -> 1
 * (ReLU), bfloat16 weights on MFMA with float32 accumulation, evaluated at
 * the 64 bin centres of every ray.  theta: float32 master parameters
 * [qsdvr_proposal_param_count()] = W1[64,39] b1 W2[64,64] b2 W3[64,64] b3
 * w4[64] b4. */
int32_t qsdvr_proposal_param_count(void);
size_t qsdvr_proposal_packed_bytes(void);
/* theta -> packed bf16 MFMA fragments + biases (device buffer of
 * qsdvr_proposal_packed_bytes()). */
int qsdvr_proposal_pack(const float *theta, void *packed, qsdvr_stream stream);
/* logits [n_rays, 64]; with x0, x1, x2, h3 [n_rays*64, 64] non-NULL the
 * activations are stored for qsdvr_proposal_backward. */
int qsdvr_proposal_forward(const float *ray_state, int64_t n_rays,
                           float interval, const void *packed, float *logits,
                           float *x0, float *x1, float *x2, float *h3,
                           qsdvr_stream stream);
size_t qsdvr_proposal_workspace_size(int64_t n_rays);
/* Gradient [param_count] of the per-ray log-likelihood of the bin targets
 * q under the composited proposal weights: with d = softplus(logits),
 * w_b = exp(-sum_{c<b} d_c) (1 - exp(-d_b)), T_end = exp(-sum d),
 * loss = -sum_b q_b log w_b - (1 - sum_b q_b) log T_end.  loss [n_rays]
 * per-ray values (nullable).  Deterministic. */
int qsdvr_proposal_backward(const float *logits, const float *target,
                            int64_t n_rays, const float *theta,
                            const float *x0, const float *x1, const float *x2,
                            const float *h3, void *workspace,
                            size_t workspace_bytes, float *grad, float *loss,
                            qsdvr_stream stream);
/* Adam update of n float32 parameters (bias-corrected, step >= 1). */
int qsdvr_adam_step(float *theta, const float *grad, float *m, float *v,
                    int64_t n, float lr, float beta1, float beta2, float eps,
                    int32_t step, qsdvr_stream stream);
/* Fine sample positions tlist [n_rays, K]: t = 0, t = L and K-2 inverse-CDF
 * samples of (1-eps) w^gamma/sum(w^gamma) + eps/64 (w the composited
 * proposal weights above, gamma in (0, 16] a sharpening exponent) over the
 * 64 bins of [0, L], L = (count-1) * interval. */
int qsdvr_proposal_sample(const float *logits, const float *ray_state,
                          int64_t n_rays, float interval, int32_t K, float eps,
                          float gamma, float *tlist, qsdvr_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* QSDVR_HIP_H */
