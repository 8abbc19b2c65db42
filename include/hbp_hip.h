/* hbp_hip.h - C ABI of the MI355X (gfx950) HBP-ISP stereo library, libhbp_hip.so.
 *
 * Plain C: opaque handles, pointers and sizes, int status (0 = success, < 0 = error, the
 * message from hbp_last_error()).  Host buffers are caller-owned; device memory is owned by
 * the library (a caching pool per context).  All work of a context runs in order on the
 * context's own HIP stream.  Each entry point names the reference interface it replaces
 * (file:line in the HBP-ISP sources); the C++ drop-in classes in hbp-isp_amd/api call these.
 *
 * Layouts: images are BGR u8 interleaved [H][W][3] (cv::Mat CV_8UC3); disparities f32 [H][W]
 * (CV_32FC1); planes f32 [H][W][3] (CV_32FC3, a b c of d = a*x + b*y + c); the cost volume
 * is u8 [D][H][pitch] on the device (pitch = W rounded up to 64) and f32 [D][H][W] through
 * hbp_volume_dev_f32 / the float* getter.
 */
#ifndef HBP_HIP_H
#define HBP_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hbp_ctx hbp_ctx;
typedef struct hbp_volume hbp_volume;
typedef struct hbp_pyramid hbp_pyramid;
typedef struct hbp_bp hbp_bp;

/* ---- context: replaces the implicit device/legacy stream of the reference
 *      (cudaSetDevice(0), main_stereo_hbp_isp.cu:31) */
int hbp_ctx_create(int device, hbp_ctx **out);
/* fails (and destroys nothing) while volumes, pyramids, BP objects or servers made on the
 * context are still alive: free those first */
int hbp_ctx_destroy(hbp_ctx *ctx);
int hbp_ctx_sync(hbp_ctx *ctx);
void *hbp_ctx_stream(hbp_ctx *ctx); /* hipStream_t of the context */
const char *hbp_last_error(void);
long long hbp_ctx_pool_bytes(hbp_ctx *ctx); /* device bytes reserved by the context's pool */
/* kernel timing with HIP events on the context's stream (no counterpart in the reference, which
 * times stages with host clocks, main_stereo_hbp_isp.cu:36-72): enable / reset, then read the
 * summed event time, launch count and algorithmic HBM bytes of one kernel name; names lists the
 * recorded names separated by ',' (returns the length needed) */
int hbp_profile_enable(hbp_ctx *ctx, int on);
int hbp_profile_reset(hbp_ctx *ctx);
int hbp_profile_read(hbp_ctx *ctx, const char *kernel, double *total_ms, long long *launches,
                     double *bytes);
int hbp_profile_names(hbp_ctx *ctx, char *buf, int buflen);
/* time only launches whose name starts with prefix (NULL or "" times all), and of those one
 * launch in `every` (1 = all) */
int hbp_profile_only(hbp_ctx *ctx, const char *prefix, int every);
/* a float4 copy of `bytes` (a multiple of 16), `launches` times (no counterpart in the
 * reference): the kernel of known HBM traffic that a profiler's FETCH_SIZE / WRITE_SIZE
 * counters are calibrated on (bench.py --traffic-calibration) */
int hbp_profile_copy(hbp_ctx *ctx, long long bytes, int launches);

/* ---- the device primitives on host arrays, for their tests (no counterpart in the reference,
 *      whose edge sorts and scans are Thrust calls: segment.cu:1058-1085, CRTrees.cu:395).
 *      hbp_primitive_sort: the project's stable LSD radix sort of n keys (key_bytes 4 or 8) on
 *      bits [0, end_bit), in place, vals (n u32, may be NULL for 8-byte keys) permuted with them;
 *      keys with bits at or above end_bit are refused.  hbp_primitive_scan: out[i] = in[0] + ...
 *      + in[i-1] (int32; the sums must fit). */
int hbp_primitive_sort(hbp_ctx *ctx, void *keys, int key_bytes, uint32_t *vals, long long n, int end_bit);
int hbp_primitive_scan(hbp_ctx *ctx, const int32_t *in, int32_t *out, long long n);

/* ---- kernel-form tuning of a context (no counterpart in the reference: its kernels have one
 *      form each).  Every value gives the same results bit for bit; only the speed changes.
 *      "coarse_min_pix" (>= 1, default 96): average pixels per superpixel from which a BP level
 *      costs planes in the per-superpixel segment form rather than in chunks; "single_hb",
 *      "batch_hb" (2 or 4; defaults 2, 2): steps per batch of the segment wavefronts in one-pair
 *      and in batched launches; "hamming_cols" (4 or 8, default 4): columns per lane of the
 *      Hamming cost kernel; "seg_len" (a multiple of 64 in [64, 512], default 512): pixels per
 *      segment wavefront of the segment form; "batch_graph" (0 or 1, default 0): hbp_bp_run_batch queues its
 *      rounds one by one from the host (0) or captures the whole schedule of rounds in one HIP
 *      graph and replays it (1); "fuse" (0, 1, 2): one-pair BP on segment-form levels
 *      runs each RANSAC sweep that follows another non-spatial sweep of its level as one launch,
 *      the previous sweep's node update recomputed where it is read (1), or every sweep as two
 *      launches (0); 2 (the default) as 1, with every RANSAC sweep's plane costs computed at the
 *      level's start on a second stream of the context; "pipe_batch" (1..16, default 4),
 *      "pipe_blocks" (8..65536, default 512), "pipe_hb" (2 or 4, default 2): RANSAC sweeps per
 *      such launch, its persistent blocks and the steps per batch of its segment wavefronts;
 *      "pipe_window" (0..1024, default 4): such launches issued ahead of the one the first
 *      stream waits for (0: all of a level's at its start);
 *      "crt_rounds" (1..5), "crt_jumps" (1..8), "crt_max_jumps" (1..8): the schedule of the
 *      next pyramid level's CRTrees phase kernels (levels above 8000 nodes); the context
 *      keeps adapting it to the steps its levels use (a level that needs more reruns with the
 *      full schedule);
 *      "level_form.<l>" (l < 24; -1 default, 0, 1): BP level l costs
 *      planes by coarse_min_pix's rule, in chunks of sorted pixels, or in per-superpixel
 *      segments.  Applies to BP objects created afterwards (and to the batched
 *      rounds run afterwards).  hbp_amd.tune picks them by timing the caller's workload. */
int hbp_ctx_set_tuning(hbp_ctx *ctx, const char *name, int value);
int hbp_ctx_get_tuning(hbp_ctx *ctx, const char *name, int *value);

/* ---- (a) cost volume: StereoCost(imgL, imgR, labels, vol_dir) + compute_stereo_cost(algo)
 *      (stereo_cost.hpp:90-116, stereo_cost.cu:421-445).  algo "Census" (u8 volume, tau 25)
 *      or "AD" (ad_cost, stereo_cost.cu:9-43: f32 volume of |grey difference| truncated at
 *      tau 30, zero in the columns x < D-1; read it with hbp_volume_download_f32). */
int hbp_cost_volume(hbp_ctx *ctx, const uint8_t *bgr_left, const uint8_t *bgr_right, int width,
                    int height, int labels, const char *algo, hbp_volume **out);
/* same, images already in device memory */
int hbp_cost_volume_dev(hbp_ctx *ctx, const uint8_t *d_bgr_left, const uint8_t *d_bgr_right,
                        int width, int height, int labels, const char *algo, hbp_volume **out);
/* StereoCost::cost_to_right (stereo_cost.cu:488-516): the volume and WTA of the right view */
int hbp_cost_to_right(hbp_ctx *ctx, hbp_volume *vol);
/* compute_stereo_cost("MC-CNN") (stereo_cost.cu:399-445): a given f32 volume [D][H][W] (the
 * .acrt file of read_volume_from_file, headerless), its left out-of-view columns filled, every
 * entry truncated at tau = 0.5, then WTA; host buffer (algo must be "MC-CNN") */
int hbp_cost_volume_f32(hbp_ctx *ctx, const float *vol, int width, int height, int labels, const char *algo,
                        hbp_volume **out);
/* same from a device buffer (not kept: the library holds its own truncated copy) */
int hbp_cost_volume_f32_dev(hbp_ctx *ctx, const float *d_vol, int width, int height, int labels,
                            const char *algo, hbp_volume **out);
/* the volume as f32 [D][H][W] (the reference's float getter layout) and the WTA disparity;
 * either pointer may be NULL */
int hbp_volume_download_f32(hbp_ctx *ctx, hbp_volume *vol, float *vol_host, float *wta_host);
/* getters g_labels / g_im_size / g_tau (stereo_cost.hpp:27-41) and the layout */
int hbp_volume_info(const hbp_volume *vol, int *width, int *height, int *labels, float *tau,
                    int *pitch);
/* g_img_D_d (stereo_cost.hpp:47-51): device f32 WTA disparity [H][W] */
float *hbp_volume_dev_wta(hbp_volume *vol);
/* g_img_cost_d (stereo_cost.hpp:42-46): device f32 [D][H][W] volume, built on first call */
float *hbp_volume_dev_f32(hbp_ctx *ctx, hbp_volume *vol);
/* device u8 volume [D][H][pitch] */
const uint8_t *hbp_volume_dev_u8(const hbp_volume *vol);
/* copy to host: vol u8 [D][H][W] (may be NULL), wta f32 [H][W] (may be NULL) */
int hbp_volume_download(hbp_ctx *ctx, const hbp_volume *vol, uint8_t *vol_host, float *wta_host);
int hbp_volume_free(hbp_ctx *ctx, hbp_volume *vol);

/* ---- (b) superpixel pyramid: SegHAC(src, link, num_nb, sigma, LabelISPFormat, 1, false) +
 *      run_ms() (segment.hpp:248-309, segment.cu:1279-1321).  link 0 MinLink, 1 MaxLink,
 *      2 CentroidLink (the reference's default), 3 WardLink; num_nb 4 or 8.  sigma > FLT_MIN
 *      blurs the image first (img_CPU_to_GPU, segment.cu:795-809; imutil.cu:399-498): a
 *      (2S+1)^2 Gaussian, S = round(3 sigma) <= 8 (larger is an error, as the reference's launch
 *      of (16+2S)^2 threads fails), clamped borders; 0 (every script) skips it. */
int hbp_pyramid_build(hbp_ctx *ctx, const uint8_t *bgr, int width, int height, int link,
                      int num_nb, double sigma, hbp_pyramid **out);
int hbp_pyramid_build_dev(hbp_ctx *ctx, const uint8_t *d_bgr, int width, int height, int link,
                          int num_nb, double sigma, hbp_pyramid **out);
/* the same with SegHAC's target_clus (segment.cu:954-1041, 1299-1306): 1 <= target_clus <
 * width*height; above 1 the pyramid ends at the first level of at most target_clus clusters,
 * re-labelled to exactly target_clus (the clusters whose children have the largest summed
 * 1-NN distances are broken into their children). */
int hbp_pyramid_build_target(hbp_ctx *ctx, const uint8_t *bgr, int width, int height, int link, int num_nb,
                             double sigma, int target_clus, hbp_pyramid **out);
int hbp_pyramid_build_target_dev(hbp_ctx *ctx, const uint8_t *d_bgr, int width, int height, int link,
                                 int num_nb, double sigma, int target_clus, hbp_pyramid **out);
/* g_num_clus_isp().size() */
int hbp_pyramid_levels(const hbp_pyramid *pyr);
/* g_num_clus_isp()[level], g_num_bd_isp()[level] (segment.hpp:127-136) */
int hbp_pyramid_level_size(const hbp_pyramid *pyr, int level, int *num_clus, int *num_bd);
/* g_isp_clus_d / g_isp_mean_d / g_isp_bd_d / g_isp_bdl_d (segment.hpp:137-156): packed device
 * arrays (int, float [6][V] per level, int2 as int pairs, int) */
int *hbp_pyramid_dev_clus(hbp_pyramid *pyr);
float *hbp_pyramid_dev_mean(hbp_pyramid *pyr);
int *hbp_pyramid_dev_bd(hbp_pyramid *pyr);
int *hbp_pyramid_dev_bdl(hbp_pyramid *pyr);
/* copy one level to host: clus [V], mean [6][V] (sums), bd [E][2], bdl [E]; any may be NULL */
int hbp_pyramid_download_level(hbp_ctx *ctx, const hbp_pyramid *pyr, int level, int *clus,
                               float *mean, int *bd, int *bdl);
int hbp_pyramid_free(hbp_ctx *ctx, hbp_pyramid *pyr);

/* ---- CRTrees on device arrays (CRTrees.hpp:16-78) */
/* CRTrees::get_clus(nn_d, clus_d, N): cluster representative per node, nn_d unchanged */
int hbp_crtrees_get_clus(hbp_ctx *ctx, int *d_nn, int *d_clus, int n);
/* CRTrees::compact_clus_label(clus_d, N): ranks in place, returns the cluster count */
int hbp_crtrees_compact(hbp_ctx *ctx, int *d_clus, int n, int *num_clus);

/* ---- (c) hierarchical BP: HBP_ISP(isp_data, stereo_data, alpha, tau_smooth)
 *      (hbp_isp.hpp:288-386) with StereoData(stereo_cost, use_cost_vol) (hbp_isp.hpp:91-105).
 *      The pyramid and volume are borrowed and must outlive the BP object. */
int hbp_bp_create(hbp_ctx *ctx, const hbp_pyramid *pyr, const hbp_volume *vol, float alpha,
                  float tau_smooth, int min_level, int use_cost_vol, hbp_bp **out);
/* run up to max_sweeps sweeps of HBP_ISP::run_ms (hbp_isp.cu:1313-1346); *done = sweeps run */
int hbp_bp_step(hbp_ctx *ctx, hbp_bp *bp, int max_sweeps, int *done);
/* total sweeps the schedule holds, the level being worked on (min_level-1 when finished) */
int hbp_bp_total_sweeps(const hbp_bp *bp);
int hbp_bp_level(const hbp_bp *bp);
int hbp_bp_level_size(const hbp_bp *bp, int level, int *num_clus, int *num_bd);
/* map_msg (hbp_isp.cu:1081-1105): disp f32 [H][W] and plane f32 [H][W][3] on the host
 * (either may be NULL); runs the remaining sweeps first */
int hbp_bp_run(hbp_ctx *ctx, hbp_bp *bp, float *disp, float *plane3);
/* ---- many pairs' cost volumes, pyramids and BP objects at once (no counterpart in the
 *      reference, which builds one pair per run_ms; for hbp_bp_run_batch).  Each pair: device
 *      BGR images (u8 [height][width][3]), the Census cost with `labels` disparities,
 *      hbp_pyramid_build_dev(linkage, num_nb, sigma 0), hbp_bp_create(alpha, tau_smooth,
 *      min_level, use_cost_vol) with min_level < 0 meaning BP on -min_level levels below the
 *      top and min_level clamped into [1, levels - 1].  `workers` (1-16) host threads build
 *      the pairs, each on a stream and memory pool of its own that the context keeps, so the
 *      pairs' latency-bound pyramid levels and set-up waits overlap; the objects are the same
 *      bits as the one-pair calls, belong to ctx (free them with ctx) and are complete when the
 *      call returns.  A pair whose CRTrees needs the cooperative fallback is redone on ctx's
 *      own stream after the workers finish.  On error nothing is returned (all NULL). */
typedef struct {
    const uint8_t *left, *right;
    int width, height, labels, linkage, num_nb, min_level;
} hbp_pair_dev;
int hbp_pairs_prepare_dev(hbp_ctx *ctx, int n, const hbp_pair_dev *pairs, float alpha, float tau_smooth,
                          int use_cost_vol, int workers, hbp_volume **vols, hbp_pyramid **pyrs, hbp_bp **bps);

/* batched BP (no counterpart in the reference, which runs one pair per HBP_ISP::run_ms): runs
 * the remaining sweeps of n BP objects of one context together, sweep i of every pair in the
 * same launches, then the map of each.  Every pair's disparity and planes are bit-identical to
 * its own hbp_bp_run; read them with hbp_bp_run (which then only maps and copies) or the
 * device pointers below.  All n objects must use the same view and use_cost_vol. */
int hbp_bp_run_batch(hbp_ctx *ctx, int n, hbp_bp *const *bps);
/* device pointers of the last map: disp [H][W], plane [H][W][3] */
float *hbp_bp_dev_disp(hbp_bp *bp);
float *hbp_bp_dev_plane(hbp_bp *bp);
/* state of the current level after the last sweep, for checks: planes [12][V][4], cost and
 * belief [12][V], messages [E][8]; slots 8..11 of planes may already hold the next sweep's
 * candidates */
int hbp_bp_get_state(hbp_ctx *ctx, hbp_bp *bp, float *planes, float *cost, float *belief,
                     float *msg);
int hbp_bp_free(hbp_ctx *ctx, hbp_bp *bp);
/* sweeps of the schedule for `levels` pyramid levels, min_level and D labels */
int hbp_schedule_sweeps(int levels, int min_level, int labels);

/* ---- request server: continuous batching (no counterpart in the reference, which runs one
 *      pair per call).  One host thread drives a server: hbp_submit copies a pair into the
 *      queue and returns a ticket; every hbp_server_step hands queued pairs to the server's
 *      admission thread while a batch slot is free and their estimated device memory fits
 *      hbm_budget_bytes (cost volume, pyramid and BP set-up run there, on a stream of their
 *      own, without blocking the step), lets admitted pairs join, runs one BP sweep of every
 *      pair on the device in shared launches, and queues the map and copies out of the pairs
 *      that finished, behind an event (no step waits for the device).  hbp_poll returns 0 and
 *      copies disp [H][W] / plane [H][W][3] (either may be NULL) once the pair is done, 1
 *      while it is queued or running, < 0 on error (a pair that failed or does not fit).
 *      Each pair's result equals its own hbp_bp_run with the same parameters (num_nb 4, Census,
 *      left view, min_level clamped to [1, levels-1]; a negative min_level -n asks for BP on n
 *      levels, min_level = levels - 1 - n).  Destroy the server before its context. */
typedef struct hbp_server hbp_server;
int hbp_server_create(hbp_ctx *ctx, int max_active, long long hbm_budget_bytes, float alpha,
                      float tau_smooth, hbp_server **out);
int hbp_server_destroy(hbp_server *srv);
int hbp_submit(hbp_server *srv, const uint8_t *bgr_left, const uint8_t *bgr_right, int width,
               int height, int labels, int min_level, long long *ticket);
/* n_active: pairs in the BP rounds; n_pending: pairs queued or being admitted; n_done: pairs
 * whose results became ready in this step (any may be NULL) */
int hbp_server_step(hbp_server *srv, int *n_active, int *n_pending, int *n_done);
int hbp_poll(hbp_server *srv, long long ticket, float *disp, float *plane3);
/* backpressure: with max_pending > 0, hbp_submit refuses (error "server queue full") while that
 * many requests wait for admission; 0 (the default) queues without limit */
int hbp_server_set_max_pending(hbp_server *srv, int max_pending);
/* the device bytes the server accounts to its pairs now and at most (admission included), and
 * the submissions refused by max_pending (any may be NULL) */
int hbp_server_stats(hbp_server *srv, long long *bytes_in_use, long long *peak_bytes, long long *rejected);
/* withdraw a request: a queued one leaves at once, one being admitted or on the device leaves
 * the rounds and its memory returns once the queued work has run; it then polls as an error
 * ("cancelled").  A finished result is discarded.  Unknown tickets are an error. */
int hbp_cancel(hbp_server *srv, long long ticket);

/* ---- host-forced failures for tests of the error paths (no counterpart in the reference,
 *      which checks no errors).  what = "alloc": the after-th next device allocation of any
 *      context fails; "round": the after-th next batched round of a server fails before it
 *      launches anything.  after = 0 disarms.  No kernel runs in a failing state. */
int hbp_inject_fault(const char *what, long long after);

/* ---- (d) PostProcess(imL, imR) (post_process.hpp:47-55)
 * left_right_check (post_process.cu:47-110): host buffers; plane_left is filled in place;
 * fail_left / fail_right / fail2_left (u8 [H][W], 255 = failed) may be NULL. */
int hbp_lr_check_fill(hbp_ctx *ctx, const float *disp_left, const float *disp_right,
                      float *plane_left, int width, int height, uint8_t *fail_left,
                      uint8_t *fail_right, uint8_t *fail2_left);
/* same on device buffers */
int hbp_lr_check_fill_dev(hbp_ctx *ctx, const float *d_disp_left, const float *d_disp_right,
                          float *d_plane_left, int width, int height, uint8_t *d_fail_left,
                          uint8_t *d_fail_right, uint8_t *d_fail2_left);
/* the same as hbp_lr_check_fill_dev, allocating nothing (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture, tight serving
 * loops).  Its scratch is a static device array of the library, 88 MiB per GPU, held until the
 * process ends; images of more than 2^23 pixels are an error.  Calls that overlap on one GPU
 * share that array and race: order them (one stream, or events). */
int hbp_lr_check_fill_dev_noalloc(hbp_ctx *ctx, const float *d_disp_left, const float *d_disp_right,
                                  float *d_plane_left, int width, int height, uint8_t *d_fail_left,
                                  uint8_t *d_fail_right, uint8_t *d_fail2_left);
/* plane_label_to_disp (post_process.cu:373-390), host buffers */
int hbp_plane_to_disp(hbp_ctx *ctx, const float *plane3, int width, int height, float *disp);
/* median_filter / median_filter_gpu (post_process.cu:127-185, 187-371): every pixel with
 * fail != 0 takes the weighted median of the planes in its 41x41 window, in the order of their
 * disparities at the pixel, weights exp(-|I(p) - I(q)|_1 / 10) on the BGR image; planes are
 * read as they were before the call.  Host buffers: bgr u8 [H][W][3], fail u8 [H][W],
 * plane3 f32 [H][W][3] in place. */
int hbp_weighted_median(hbp_ctx *ctx, const uint8_t *bgr, const uint8_t *fail, float *plane3, int width,
                        int height);
/* same on device buffers */
int hbp_weighted_median_dev(hbp_ctx *ctx, const uint8_t *d_bgr, const uint8_t *d_fail, float *d_plane3,
                            int width, int height);
/* medianFilter of the CLI (main_stereo_hbp_isp.cu:120-193): horizontal then vertical 7-tap
 * median of a disparity image f32 [H][W], in place (host buffer) */
int hbp_median7(hbp_ctx *ctx, float *disp, int width, int height);
/* same on a device buffer */
int hbp_median7_dev(hbp_ctx *ctx, float *d_disp, int width, int height);

/* ---- Middlebury evaluation: Evaluator::evaluate (Evaluator.h:46-52, 66-78).  Over
 * valid = (GT > 0 and GT != +inf), nonocc = (mask != 0) and occ = (valid and not nonocc):
 * counts[6] = {valid, valid and good, nonocc, nonocc and good, occ, occ and good}, a pixel
 * being good when |disp - GT| <= (float)threshold, in float.  Host buffers f32 [H][W] and
 * u8 [H][W]; the bad-pixel rates are 100 (1 - good / pixels) per mask. */
int hbp_eval_bad_pixels(hbp_ctx *ctx, const float *disp, const float *gt, const uint8_t *nonocc, int width,
                        int height, double threshold, long long *counts);

#ifdef __cplusplus
}
#endif
#endif
