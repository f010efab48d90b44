/*
 * gs_hip.h — the C ABI of the MI355X (gfx950) 3D Gaussian Splatting trainer.
 *
 * Plain C: opaque handles, POD structs, device pointers as plain pointers,
 * sizes as integers, int status codes (0 = ok, otherwise GS_HIP_E_*) and a
 * thread-local message from gs_hip_last_error().  Nothing here throws.
 *
 * Layers:
 *   1. device and memory (gs_hip_init ... gs_hip_synchronize)
 *   2. the workspace: every device buffer one view-step needs, sized for a
 *      slot capacity and an image size (the role of the reference's
 *      VkContext-owned Trainer buffers, engine/src/trainer.cpp:303-386)
 *   3. the stages of one training step (engine/src/trainer.cpp:665-776)
 *   4. launch shapes (the reference's reduction / alternative choices)
 *   5. batches of scenes trained together (no counterpart in the reference)
 *   6. scene files on the device (SPZ load, PLY rows)
 *   7. the trainer handle: the gs::Trainer C++ class behind a C surface, the
 *      shape a JNI/ctypes binding uses (app/.../NativeTrainer.kt:8-47)
 *
 * Threading: a workspace or trainer is used by one thread at a time; calls
 * on one handle are not thread-safe (engine/src/vk_context.cpp:583-586).
 */
#ifndef GS_HIP_H
#define GS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define GS_HIP_OK 0
#define GS_HIP_E_HIP 1         /* a HIP runtime call failed */
#define GS_HIP_E_ARGUMENT 2    /* invalid argument (std::invalid_argument in C++) */
#define GS_HIP_E_CAPACITY 3    /* more than a buffer or format limit holds */
#define GS_HIP_E_IO 4          /* file or format error (std::runtime_error in C++) */
#define GS_HIP_E_STATE 5       /* call out of order (std::logic_error in C++) */

/* ---- constants (engine/shaders/train_common.glsl:59-70) ------------------ */
#define GS_HIP_TILE 16
#define GS_HIP_MAX_SH_REST 15     /* degree 3: 15 coefficients beyond DC */
#define GS_HIP_PARAM_FLOATS 16    /* per slot: [pos, logit | logscale, alive | quat wxyz | dc, -] */
#define GS_HIP_RECORD_FLOATS 12   /* projected raster record per slot */
#define GS_HIP_GRAD2D_FLOATS 9    /* dL/d(mean xy, conic abc, opacity, colour rgb), per slot */
#define GS_HIP_MAX_VIEWS 8        /* views of one batched optimizer step */

/* frame flags */
#define GS_HIP_FLAG_MASKED_TARGET 1u  /* target = mix(bg, photo, alpha) (train_forward.comp:124-127) */
#define GS_HIP_FLAG_SPARSE_ADAM 2u    /* gaussians that drew nothing keep value and moments (train_step.glsl:223-230) */
#define GS_HIP_FLAG_DEPTH16 4u        /* order pairs by the reference's 16-bit log depth (train_project.comp:110)
                                         instead of the float bits of z */
#define GS_HIP_FLAG_NO_UPDATE 8u      /* run the step but leave parameters, moments and statistics alone */
#define GS_HIP_FLAG_POSE_GRAD 16u     /* this view's camera pose gradient (train_step.glsl:196-201, 253-269), read
                                         with gs_hip_pose_gradient; with NO_UPDATE the reference's kFlagPoseOnly */
#define GS_HIP_FLAG_EXPOSURE 32u      /* photo rgb times exposure[] (train_forward.comp:123) and the view's gain
                                         fit sums (142-151), read with gs_hip_exposure_sums */
#define GS_HIP_FLAG_DETERMINISTIC 64u /* the backward sums each slot's GRAD2D in an order independent of scheduling:
                                         a tile's four quadrant waves in a fixed order, then the tiles as 64-bit fixed
                                         point (2^-44 resolution, |value| < 2^18), so a view's GRAD2D is bitwise the
                                         same in every run and in any batch (the reference's pair-sum pass,
                                         train_pair_sum.comp:17-32, is atomic-free); set on any view of a step, it
                                         applies to all; 72 B per slot per view, allocated on first use */

/* Pinhole camera, scene -> camera, OpenCV axes (x right, y down, z forward).
 * Same fields as gs::Camera (engine/include/gs/camera.h:10-16). */
typedef struct gs_hip_camera {
  float rotation[9];    /* row-major */
  float translation[3];
  float position[3];    /* camera centre in scene coordinates */
  float fx, fy, cx, cy;
  uint32_t width, height;
} gs_hip_camera;

/* Everything one view-step's kernels read besides buffers: the role of the
 * reference's FrameData uniform (engine/src/trainer.cpp:47-64, filled by
 * Trainer::write_frame, 440-547).  Learning rates are final (the position
 * rate already scaled by the camera extent and decayed); bias1/bias2 are
 * 1/(1 - beta^t) computed on the host in double. */
typedef struct gs_hip_frame {
  gs_hip_camera cam;
  uint32_t tiles_x, tiles_y;
  uint32_t slots;        /* slots in use (alive + free) */
  uint32_t sh_degree;    /* active SH degree, <= the stored degree */
  uint32_t flags;        /* GS_HIP_FLAG_* */
  uint32_t seed;         /* split sampling seed (train_densify_spawn.comp:113) */
  float background[3];
  float lambda_dssim;    /* loss = (1 - lambda) L1 + lambda (1 - SSIM) */
  float ssim_c1, ssim_c2;
  float lr_position, lr_scale, lr_rotation, lr_opacity, lr_sh_dc, lr_sh_rest;
  float beta1, beta2, epsilon, bias1, bias2;
  float grad_threshold;   /* densify: mean view-space gradient to clone or split */
  float clone_max_scale;  /* clone at or below this max scale, split above */
  float prune_opacity;    /* prune below this opacity */
  float prune_max_scale;  /* > 0: also prune larger world scales */
  float prune_max_radius; /* > 0: also prune larger screen radii */
  float split_divisor;    /* split halves: scale / this (1.6) */
  float exposure[3];      /* with GS_HIP_FLAG_EXPOSURE: the photo's rgb is multiplied by these (1 / the view's
                             gain: the photo at the model's brightness, trainer.cpp:442-444) */
} gs_hip_frame;

/* ---- 1. device and memory ------------------------------------------------ */
int gs_hip_init(int device);
int gs_hip_device_name(char* buffer, size_t size);
int gs_hip_malloc(void** ptr, size_t bytes);
int gs_hip_free(void* ptr);
int gs_hip_upload(void* device_dst, const void* host_src, size_t bytes);
int gs_hip_download(void* host_dst, const void* device_src, size_t bytes);
int gs_hip_memset(void* device_dst, int value, size_t bytes);
/* Device-to-device copy by the library's own streaming kernel (16-byte
 * vectors when both pointers and the size allow); waits for completion. */
int gs_hip_copy(void* device_dst, const void* device_src, size_t bytes);
int gs_hip_synchronize(void);
/* Copies the calling thread's last error message (NUL-terminated); returns its status code. */
int gs_hip_last_error(char* buffer, size_t size);

/* ---- 2. workspace --------------------------------------------------------- */
typedef struct gs_hip_workspace gs_hip_workspace;

/* Buffers of a workspace, for gs_hip_workspace_buffer.  Layouts:
 *  PARAMS, M_PARAMS, V_PARAMS   float[capacity][16]
 *  SH, M_SH, V_SH               float[sh_coeffs][capacity][3]   (coefficient-major planes)
 *  STATS                        float[capacity][4]  (grad norm sum, views, max radius, -)
 *  RECORDS                      float[capacity][12] (mean xy, conic abc, opacity, rgb, bbox lo/hi, radius|clamp)
 *  FOOTPRINT                    float[capacity][8]  (dilated cov2D abc, threshold | mean xy, tiles as u32 bits, 0):
 *                               everything the pair emitter reads of a slot, in one 32 B line
 *  TILES, DEPTH_KEYS, DEPTH_ORDER, OFFSETS  uint32[capacity]  (TILES: bins touched; OFFSETS in depth order; with
 *                               GS_HIP_FLAG_DEPTH16 a key is the 16-bit depth | tiles << 16, sorted by
 *                               its lower half, 0xffff for a slot that touches no tile)
 *  GRAD2D                       float[capacity][9]  (the nine sums of a slot together)
 *  PAIR_TILES, PAIR_GAUSSIANS   uint32[pair_capacity] (sorted)
 *  RANGES                       uint32[bins + 1][2] (bins = the tiles with bin_shift 0)
 *  IMAGE                        float[H*W][4]  (rendered rgb, final transmittance)
 *  LAST                         uint32[H*W]    (1-based position of the last contributor)
 *  TARGET                       uint32[H*W]    (RGBA8)
 *  DLDC                         float[H*W][4]  (dL/dC rgb, -)
 *  TILE_LOSS                    float[tiles][4] (sum |e|, sum e^2, sum SSIM, -)
 *  TOTALS                       uint32[16]     ([0] pairs of the last bin, [1] emit mismatches)
 *  TILE_ORDER                   uint32[views][4 quads] (the rasterizers' launch order: view << 16 | tile
 *                               in 2x2 quads of tiles, 0xffffffff for a member outside the grid, the
 *                               quads with the longest lists first)
 *  TILE_FIT                     float[tiles][8] (with GS_HIP_FLAG_EXPOSURE: sum C^2 rgb, fitted pixels,
 *                               sum C photo rgb, subject pixels; C the model's colour before the background)
 *  POSE_GRAD                    float[ceil(capacity / 256)][8] (with GS_HIP_FLAG_POSE_GRAD, per 256 slots:
 *                               dL/dd xyz, 0, dL/dw xyz, 0)
 * The model buffers (PARAMS .. STATS) are shared; the others belong to a view. */
enum gs_hip_buffer {
  GS_HIP_BUF_PARAMS = 0,
  GS_HIP_BUF_SH,
  GS_HIP_BUF_M_PARAMS,
  GS_HIP_BUF_V_PARAMS,
  GS_HIP_BUF_M_SH,
  GS_HIP_BUF_V_SH,
  GS_HIP_BUF_STATS,
  GS_HIP_BUF_RECORDS,
  GS_HIP_BUF_FOOTPRINT,
  GS_HIP_BUF_TILES,
  GS_HIP_BUF_DEPTH_KEYS,
  GS_HIP_BUF_DEPTH_ORDER,
  GS_HIP_BUF_OFFSETS,
  GS_HIP_BUF_GRAD2D,
  GS_HIP_BUF_PAIR_TILES,
  GS_HIP_BUF_PAIR_GAUSSIANS,
  GS_HIP_BUF_RANGES,
  GS_HIP_BUF_IMAGE,
  GS_HIP_BUF_LAST,
  GS_HIP_BUF_TARGET,
  GS_HIP_BUF_DLDC,
  GS_HIP_BUF_TILE_LOSS,
  GS_HIP_BUF_TOTALS,
  GS_HIP_BUF_TILE_ORDER,
  GS_HIP_BUF_TILE_FIT,
  GS_HIP_BUF_POSE_GRAD,
  GS_HIP_BUF_COUNT
};

/* sh_degree: the stored SH degree (0..3), fixing the per-slot SH planes. */
int gs_hip_workspace_create(gs_hip_workspace** ws, uint32_t capacity, int sh_degree, uint32_t width,
                            uint32_t height);
int gs_hip_workspace_destroy(gs_hip_workspace* ws);
int gs_hip_workspace_buffer(gs_hip_workspace* ws, int which, void** device_ptr, size_t* bytes);
/* Views: independent sets of per-view buffers over the workspace's one model
 * (view 0 exists from creation); a batched step runs every stage of all its
 * views in one launch.  Grows to `count` (<= GS_HIP_MAX_VIEWS) views. */
int gs_hip_workspace_views(gs_hip_workspace* ws, uint32_t count);
/* A buffer of view `view` (model buffers are the same for every view). */
int gs_hip_workspace_view_buffer(gs_hip_workspace* ws, uint32_t view, int which, void** device_ptr, size_t* bytes);
/* Pair buffers for at least `pairs` (tile, gaussian) pairs; grows, never shrinks. */
int gs_hip_workspace_reserve_pairs(gs_hip_workspace* ws, uint64_t pairs);
/* Copies `count` slots of parameters (16 floats each) and SH rest coefficients
 * ([slot][coeff][rgb], `sh_coeffs_in` per slot) in at slot 0; zeroes moments and stats. */
int gs_hip_workspace_load(gs_hip_workspace* ws, const float* params, const float* sh_rest, uint32_t count,
                          uint32_t sh_coeffs_in);
/* Copies slots [0, count) out: params (16 per slot), SH [slot][coeff][rgb]; either may be NULL. */
int gs_hip_workspace_store(gs_hip_workspace* ws, float* params, float* sh_rest, uint32_t count);

/* ---- 3. stages of one view-step ------------------------------------------ */
/* EWA projection, SH colour, footprint tile count and depth key per slot;
 * zeroes GRAD2D (train_project.comp:26-112). */
int gs_hip_project(gs_hip_workspace* ws, const gs_hip_frame* frame);
/* Depth sort of the slots, prefix of their tile counts in depth order, pair
 * emission, stable tile sort of the pairs, tile ranges (trainer.cpp:665-711).
 * Grows the pair buffers when needed; *pairs = P. */
int gs_hip_bin(gs_hip_workspace* ws, const gs_hip_frame* frame, uint64_t* pairs);
/* Front-to-back compositing per 16x16 tile: IMAGE, LAST (train_forward.comp:59-141). */
int gs_hip_forward(gs_hip_workspace* ws, const gs_hip_frame* frame);
/* Uploads a RGBA8 target (width*height*4 bytes) into TARGET. */
int gs_hip_set_target(gs_hip_workspace* ws, const uint8_t* rgba8);
/* Copies a RGBA8 target already in device memory into TARGET, in stream order. */
int gs_hip_set_target_device(gs_hip_workspace* ws, const void* device_rgba8);
/* Loss: SSIM maps and tile sums, then dL/dC of (1-lambda) L1 + lambda (1-SSIM)
 * into DLDC (train_forward.comp:119-140, train_ssim.comp, train_ssim_grad.comp).
 * losses (may be NULL): sum |e|, sum e^2, sum SSIM over pixels and channels. */
int gs_hip_loss(gs_hip_workspace* ws, const gs_hip_frame* frame, double losses[3]);
/* Backward rasterizer: re-walks each tile back to front from the final
 * transmittance and accumulates dL/d(2D splat) into GRAD2D (train_backward_serial.glsl). */
int gs_hip_backward(gs_hip_workspace* ws, const gs_hip_frame* frame);
/* Projection adjoint, densify statistics, fused Adam on the 16 parameters and
 * the SH coefficients (train_step.glsl:72-240, train_step_sh.comp:21-44). */
int gs_hip_adam(gs_hip_workspace* ws, const gs_hip_frame* frame);
/* Densification: prune / clone / split by stream compaction.  *slots is the
 * slots in use (updated); counts = {freed, spawned placed, pruned}
 * (train_densify_mark.comp, train_densify_spawn.comp, trainer.cpp:1083-1119). */
int gs_hip_densify(gs_hip_workspace* ws, const gs_hip_frame* frame, uint32_t* slots, uint32_t counts[3]);
/* Caps every alive opacity logit at max_logit and clears its Adam moments
 * (train_reset_opacity.comp:17-23). */
int gs_hip_reset_opacity(gs_hip_workspace* ws, const gs_hip_frame* frame, float max_logit);
/* The loss stage of gs_hip_loss on caller-owned device buffers, allocating
 * nothing.  No counterpart in the reference: for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  image: float4 [H*W] (C rgb, T), target: RGBA8 [H*W],
 * dldc: float4 [H*W] written (dL/dC rgb, 0); the frame supplies the size,
 * tile grid, background, lambda, SSIM constants and flags (MASKED_TARGET,
 * EXPOSURE).  The SSIM maps and per-tile sums live in a file-scope device
 * array of the library (about 73 MiB, held by each GPU it runs on until the
 * process ends), so at most 2^21 pixels and 2^14 tiles (GS_HIP_E_CAPACITY
 * beyond).  stream: a hipStream_t, or NULL for the null stream; every launch
 * goes there.  losses (may be NULL: then nothing waits, and the call can be
 * captured): sum |e|, sum e^2, sum SSIM, copied back after a wait on the
 * stream.  Calls that overlap on one GPU share the array and race. */
int gs_hip_loss_noalloc(const gs_hip_frame* frame, const void* image, const void* target, void* dldc, void* stream,
                        double losses[3]);
/* One full training view-step: project, bin, forward, loss, backward, adam. */
int gs_hip_train_step(gs_hip_workspace* ws, const gs_hip_frame* frame, uint64_t* pairs, double losses[3]);
/* One batched optimizer step over `count` views of the scene (count <= the
 * workspace's views): project, bin, forward, loss and backward of all views,
 * one launch per stage; then one projection adjoint + Adam on the mean of the
 * views' gradients.  Pair counts stay on the device; a step whose pairs
 * overflow the buffers leaves the model alone and runs again with larger
 * buffers (trainer.cpp:819-836).  frames[v] differ only in camera (the
 * optimizer reads frames[0]'s rates and step).  targets[v]: device RGBA8 of
 * view v, read in place while the call runs (TARGET is left as it was), or
 * NULL (targets or an entry) to use what view v's TARGET holds.
 * pairs[count], losses[3 * count] may be NULL. */
int gs_hip_train_views(gs_hip_workspace* ws, const gs_hip_frame* frames, uint32_t count, const void* const* targets,
                       uint64_t* pairs, double* losses);
/* Camera pose gradient of view `view`'s latest step that carried GS_HIP_FLAG_POSE_GRAD: the loss's
 * derivative by a shift d of the camera ([R | t] -> [R | t + d]) and by a turn w about its centre
 * ([R | t] -> [E R | E t], E = exp([w]x)), in camera axes: grad = {dL/dd xyz, dL/dw xyz}, the
 * per-256-slot partials of POSE_GRAD summed in double (Trainer::pose_gradient, trainer.cpp:1319-1330).
 * Colour's dependence on the view direction is left out, as in the reference.  Waits for the stream. */
int gs_hip_pose_gradient(gs_hip_workspace* ws, uint32_t view, double grad[6]);
/* Exposure fit sums of view `view`'s latest step that carried GS_HIP_FLAG_EXPOSURE, over pixels whose
 * photo is opaque (alpha > 0.95) and render nearly so (T < 0.05): sums[0..2] = sum C^2, sums[3] = those
 * pixels, sums[4..6] = sum C photo (the photo as stored, before exposure[]), sums[7] = pixels with
 * alpha > 0; the least-squares gain per channel is sums[4 + c] / sums[c] (trainer.cpp:786-792,
 * 1697-1711).  Summed in double over TILE_FIT.  Waits for the stream. */
int gs_hip_exposure_sums(gs_hip_workspace* ws, uint32_t view, double sums[8]);
/* Forward only (render): project, bin, forward; image as RGBA8 (clamped rgb, alpha = 1 - T). */
int gs_hip_render(gs_hip_workspace* ws, const gs_hip_frame* frame, uint8_t* rgba8_out);

/* The viewer's image contract (gs_render, the app viewer:
 * engine/shaders/splat_preprocess.comp:159-291, splat.vert, splat.frag,
 * engine/src/splat_renderer.cpp:243-348): quads at +-2 sqrt(2 lambda) along
 * the 2D covariance's axes, alpha = min(o exp(-|local|^2), 0.99) dropped
 * below 1/255, colour and opacity at fp16 precision, a 16-bit log-depth key,
 * drawn back to front with premultiplied "over" into an RGBA8 target cleared
 * to `clear` (rgba in 0..1).  The SH rest coefficients are used as stored:
 * load them rounded to fp16 for the reference's upload (gs::upload_scene
 * does).  Writes width*height*4 bytes; *visible = splats drawn (on screen).
 * Works in view 0's per-view buffers (a training step rewrites them) and
 * grows the pair buffers when a view needs more. */
#define GS_HIP_VIEW_ANTIALIASED 1u /* mip-splat opacity compensation sqrt(det_raw / det) */
int gs_hip_view_render(gs_hip_workspace* ws, const gs_hip_frame* frame, const float clear[4], uint32_t flags,
                       uint8_t* rgba8_out, uint32_t* visible);

/* ---- 4. launch shapes ------------------------------------------------------ */
/* Knobs of the batched step (no counterpart in the reference, whose shaders
 * fix them: trainer.h:351 kTileSize, radix_sort.cpp 4-bit digits).  The tile
 * stays 16x16 (GS_HIP_TILE), the reference's size, so pair order and tile
 * ranges stay comparable with it. */
typedef struct gs_hip_tuning {
  uint32_t raster_groups_per_cu; /* persistent rasterizers: workgroups per CU pulling (view, tile) work;
                                    0 = one workgroup per work item */
  uint32_t depth_digit_bits;     /* radix digit width of the depth sort, 4..8 */
  uint32_t tile_digit_bits;      /* tile sort: 4..8 = LSD radix passes of that digit width (default 8);
                                    0 = one stable counting pass over tile grids up to 16383 tiles (larger
                                    ones fall back to 8-bit digits) */
  uint32_t graph;                /* 1: capture the batched step once and replay it as a HIP graph */
  uint32_t bin_shift;            /* pairs are binned by squares of (16 << bin_shift) pixels, 0..2: 0 = the
                                    16x16 raster tiles themselves (the reference's pairs and ranges); 1, 2 =
                                    32x32 / 64x64 bins, each raster tile walking its bin's list (fewer pairs to
                                    emit and sort; the same splats drawn at every pixel in the same order, so
                                    the same image and gradients) */
  uint32_t stored_state;         /* backward: 0 = each pixel re-walks its tile's list back to front from its
                                    final T (nothing stored per pixel but T, LAST and dL/dC); 1 = the forward
                                    stores every pixel's (C, T) at each 64th list position and a wave per
                                    64-splat bucket sweeps the pixels front to back, one splat per lane
                                    (train_backward_bucket.glsl's scheme; needs per-view state memory of
                                    4 KB per 64 pairs).  Steps with GS_HIP_FLAG_DETERMINISTIC or coarse bins
                                    (bin_shift > 0) re-walk. */
} gs_hip_tuning;
int gs_hip_workspace_get_tuning(gs_hip_workspace* ws, gs_hip_tuning* out);
int gs_hip_workspace_set_tuning(gs_hip_workspace* ws, const gs_hip_tuning* tuning);
/* Times `repeats` batched steps of the given views (GS_HIP_FLAG_NO_UPDATE
 * added, so the model is left alone) for every candidate shape, keeps the
 * fastest in the workspace and returns it (and its ms per step). */
int gs_hip_autotune(gs_hip_workspace* ws, const gs_hip_frame* frames, uint32_t count, const void* const* targets,
                    uint32_t repeats, gs_hip_tuning* best, double* best_ms);
/* Pair-buffer overflows so far: steps that found more (tile, gaussian) pairs
 * than the buffers held, left the model alone and ran again with larger
 * buffers (trainer.cpp:819-836). */
int gs_hip_workspace_overflows(gs_hip_workspace* ws, uint64_t* steps);

/* Kernel timing of the last stage calls on the workspace's stream, in ms
 * (hipEvents): names[i] / ms[i] for i < *count (at most `capacity`). */
int gs_hip_workspace_timings(gs_hip_workspace* ws, int enable, const char** names, double* ms, int capacity,
                             int* count);

/* ---- 5. batches of scenes ---------------------------------------------------- */
/* Several independent scenes (workspaces of any capacity, SH degree and image
 * size) trained together, one launch per stage over every view of every
 * scene: grids are sized for the largest scene, pair buffer and tile grid,
 * and each block finds its scene and view in device tables.  No counterpart
 * in the reference, which trains one scene at a time.  The batch borrows the
 * scenes' models and views (it owns only its tables, stream and work lists),
 * so a scene stays usable on its own (densify, render, store) between batch
 * steps.  At most 32 scenes and 64 views per step; tiles per view below 2^16. */
typedef struct gs_hip_batch gs_hip_batch;
int gs_hip_batch_create(gs_hip_batch** out, gs_hip_workspace* const* scenes, uint32_t count);
int gs_hip_batch_destroy(gs_hip_batch* batch);
/* One optimizer step of every scene: views[s] frames of scene s (1 .. that
 * workspace's views), all frames concatenated in scene order; targets[i] is
 * the device RGBA8 of frame i, read in place (NULL, or a NULL entry: that
 * view's TARGET buffer).  Each scene's Adam applies the mean gradient of its
 * own views (its frames[first].rates and step); a step any view of which
 * overflows its pair buffers leaves every model alone, grows the buffers
 * and runs again (each overflowing scene's overflow count goes up).
 * pairs[n] and losses[3 n] (n = sum of views[]) may be NULL. */
int gs_hip_batch_train(gs_hip_batch* batch, const gs_hip_frame* frames, const uint32_t* views,
                       const void* const* targets, uint64_t* pairs, double* losses);
int gs_hip_batch_get_tuning(gs_hip_batch* batch, gs_hip_tuning* out);
int gs_hip_batch_set_tuning(gs_hip_batch* batch, const gs_hip_tuning* tuning);
/* Kernel timing of the last batch steps (as gs_hip_workspace_timings). */
int gs_hip_batch_timings(gs_hip_batch* batch, int enable, const char** names, double* ms, int capacity, int* count);

/* ---- 6. scene files on the device ------------------------------------------- */
/* SPZ v2/v3 (the reference's load_spz, engine/src/splat_scene.cpp:367-448): the
 * gzip stream is inflated on the host, the fixed-point attributes are unpacked
 * on the device into slots [0, n) of the workspace (params, SH planes; moments
 * and statistics zeroed), every value equal to gs::load_spz's.  The file's
 * convention (RUB) is kept.  Requires capacity >= n; SH rest beyond the
 * workspace's degree is dropped, missing coefficients are zero.  gs_hip_spz_info
 * reads only the header (count, SH degree <= 3, antialiased flag), to size the
 * workspace first.  SPZ v4 (zstd) is rejected. */
int gs_hip_spz_info(const uint8_t* spz, size_t bytes, uint32_t* count, int* sh_degree, int* antialiased);
int gs_hip_workspace_load_spz(gs_hip_workspace* ws, const uint8_t* spz, size_t bytes, uint32_t* count,
                              int* sh_degree, int* antialiased);
/* The alive slots among [0, slots) as 3DGS PLY vertex rows, in slot order (the
 * reference's export_scene -> convert_convention -> save_ply,
 * trainer.cpp:1773-1800, splat_scene.cpp:336-365): per row 17 + 3 K floats, x y z,
 * nx ny nz = 0, f_dc_0..2, f_rest channel-major (K = the workspace's SH rest
 * coefficients), opacity, scale_0..2, rot_0..3 (normalised); rdf != 0 applies the
 * half turn about x (RUB -> RDF).  *count = alive slots; rows == NULL asks for the
 * count only (rows needs count * (17 + 3 K) * 4 bytes). */
int gs_hip_export_ply(gs_hip_workspace* ws, uint32_t slots, int rdf, uint8_t* rows, size_t bytes, uint32_t* count);

/* ---- 7. the trainer handle ------------------------------------------------- */
/* A whole training session behind one handle: the reference's
 * NativeTrainer (app/src/main/java/com/ondevice/gsplat/NativeTrainer.kt:8-47,
 * app/src/main/cpp/jni_bridge.cpp:815-1108) with a status code per call in
 * place of its JSON strings.  One handle per session; a handle is used by
 * one thread at a time. */
typedef struct gs_hip_trainer gs_hip_trainer;

/* NativeTrainer.prepare's arguments (jni_bridge.cpp:816-819). */
typedef struct gs_hip_trainer_options {
  int device;               /* HIP device index */
  uint32_t width;           /* images are resampled so their short side is this (never enlarged) */
  uint32_t iterations;      /* views to train on; the schedule scales to it (gs::scaled_options) */
  uint32_t max_gaussians;   /* slot capacity; the initial gaussians take at most half, capped at 150000 */
  int sh_degree;            /* 0..3 */
  int masked;               /* train on the masked subject when the capture has masks */
  int refine_poses;         /* camera pose refinement */
  int refine_exposure;      /* per-view exposure and white balance gains */
  uint32_t views_per_step;  /* views per optimizer step, 1..GS_HIP_MAX_VIEWS (0 means 1); the reference trains one */
  uint32_t seed;
} gs_hip_trainer_options;

/* What prepare reports about the folder and the initial gaussians (jni_bridge.cpp:886-903). */
typedef struct gs_hip_trainer_info {
  uint32_t frames, holdout, skipped; /* frames loaded, held out (every 8th), left out as unreadable */
  uint32_t width, height;            /* training resolution */
  int masks, depth;                  /* the capture has masks / depth maps */
  uint32_t points;                   /* sparse points in the capture */
  uint64_t init_points;              /* points the initial gaussians were built from */
  uint32_t init_gaussians;
  char init_source[16];              /* "depth", "points" or "random" */
} gs_hip_trainer_info;

/* The training statistics step() reports (jni_bridge.cpp:926-960, write_stats). */
typedef struct gs_hip_trainer_stats {
  uint32_t step, iterations;  /* views trained on so far, of the run's */
  uint32_t gaussians;         /* alive */
  uint32_t optimizer_steps;
  int sh_degree;              /* active */
  uint32_t densify_events, overflow_steps;
  uint64_t pairs;             /* (tile, gaussian) pairs of the last step */
  double l1, psnr, ssim, loss; /* of the last step's view(s) */
  double step_ms;             /* moving average of the wall-clock step time */
  int finished;
} gs_hip_trainer_stats;

/* GPU milliseconds of each pass of a profiled step (hipEvents). */
typedef struct gs_hip_pass_times {
  double project, sort, rasterize, ssim, backward, optimize;
  int valid;
} gs_hip_pass_times;

/* NativeTrainer.evaluate: the held-out views over black (jni_bridge.cpp:1005-1058). */
typedef struct gs_hip_trainer_eval {
  double psnr, ssim, l1;
  double psnr_unaligned;  /* at the given poses, when the views were aligned first (refine_poses) */
  uint32_t views;
  int aligned;
} gs_hip_trainer_eval;

/* NativeTrainer.prepare: loads `dataset_dir`/transforms.json, builds the
 * initial gaussians and the trainer.  GS_HIP_E_IO for an unreadable folder. */
int gs_hip_trainer_create(const char* dataset_dir, const gs_hip_trainer_options* options, gs_hip_trainer** out,
                          gs_hip_trainer_info* info);
/* NativeTrainer.release. */
int gs_hip_trainer_destroy(gs_hip_trainer* trainer);
/* NativeTrainer.step: up to `count` steps (fewer at the end of the run);
 * profile != 0 times the first one pass by pass into *times.  done, stats
 * and times may be NULL. */
int gs_hip_trainer_steps(gs_hip_trainer* trainer, uint32_t count, int profile, uint32_t* done,
                         gs_hip_trainer_stats* stats, gs_hip_pass_times* times);
int gs_hip_trainer_get_stats(gs_hip_trainer* trainer, gs_hip_trainer_stats* stats);
/* NativeTrainer.evaluate. */
int gs_hip_trainer_evaluate(gs_hip_trainer* trainer, gs_hip_trainer_eval* out);
/* NativeTrainer.preview: held-out view `index` (modulo their count; frame 0
 * when none is held out) as RGBA8, the render followed by the photo, both over
 * the background (0.08, 0.09, 0.11): 2 W H 4 bytes.  rgba == NULL asks only
 * for width and height. */
int gs_hip_trainer_preview(gs_hip_trainer* trainer, uint32_t index, uint8_t* rgba, size_t bytes, uint32_t* width,
                           uint32_t* height);
/* NativeTrainer.exportPly: the alive gaussians as a 3DGS PLY file (RDF),
 * rows written on the device.  count and bytes may be NULL. */
int gs_hip_trainer_export_ply(gs_hip_trainer* trainer, const char* path, uint32_t* count, uint64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* GS_HIP_H */
