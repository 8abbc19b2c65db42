/* Oracle C API (test infrastructure only).
 *
 * A host C++17 restatement of the HBP-ISP stereo path, used by tests/, by the smoke check
 * and by bench.py's cpu_baseline leg as the checker.  It is never linked into, loaded by or
 * called from the product library.  Each function names the reference file:line it follows.
 *
 * Layouts: images are BGR interleaved u8 [H][W][3]; cost volumes are u8 [D][H][W] holding
 * min(Hamming, tau) (the reference's f32 [D][H][W] holds the same integers); planes are
 * float4 (a, b, c, 0) in slot-major [K][V] order like the reference's plane_d.
 */
#ifndef HBP_ORACLE_H
#define HBP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* OpenMP threads of the oracle's parallel loops (results do not depend on it) */
void orc_set_threads(int n);
int orc_get_threads(void);

/* pinned scalar rules (exported so tests can pin them against libm) */
double orc_pin_cbrt(double x);
double orc_pin_exp(double x);
double orc_pin_pow24(double t);
float orc_i64_to_f32(int64_t q);
int orc_f2i_rd(float x);
int orc_fixed_shift(float lambda, float tau, int use_cost_vol, int npix);
float orc_amin_fold(float stored, float v);

/* (a) census + Hamming cost volume */
void orc_grey(const uint8_t *bgr, int W, int H, uint8_t *grey);
void orc_census(const uint8_t *grey, int W, int H, uint32_t *census);
void orc_hamming(const uint32_t *cl, const uint32_t *cr, int W, int H, int D, int tau,
                 uint8_t *vol);
void orc_wta(const uint8_t *vol, int W, int H, int D, float *dwta);
void orc_vol_to_right(uint8_t *vol, int W, int H, int D);
/* the "AD" cost (stereo_cost.cu:9-43): f32 volume, tau 30; right_view applies cost_to_right */
int orc_cost_volume_ad(const uint8_t *bgrL, const uint8_t *bgrR, int W, int H, int D, int right_view, float *vol,
                       float *dwta);
int orc_cost_volume_f32(const float *raw, int W, int H, int D, float tau, int right_view, float *vol,
                        float *dwta);
int orc_cost_volume(const uint8_t *bgrL, const uint8_t *bgrR, int W, int H, int D,
                    int right_view, uint8_t *vol, float *dwta);

/* (b) CRTrees superpixel pyramid */
void orc_lab(const uint8_t *bgr, int W, int H, float *lab4);
void orc_nn_grid(const float *lab4, int W, int H, int num_nb, int *nn, float *dist_min,
                 int *bd /* [num_nb*N][2] */);
void orc_cc_max(const int *nn, int V, int *label);
void orc_crtrees_literal(const int *nn, int V, int order, int *label);
int orc_compact(int *label, int V);
void *orc_pyramid_build(const uint8_t *bgr, int W, int H, int num_nb);
/* the same with a linkage: 0 MinLink, 1 MaxLink, 2 CentroidLink, 3 WardLink (segment.hpp:13) */
void *orc_pyramid_build_link(const uint8_t *bgr, int W, int H, int num_nb, int link);
/* ... and target_clus (1 <= target_clus < W*H; get_target_clus, segment.cu:954-1041) */
void *orc_pyramid_build_target(const uint8_t *bgr, int W, int H, int num_nb, int link, int target_clus);
/* ... and the Gaussian pre-blur (sigma > FLT_MIN; imutil.cu:399-498, segment.cu:795-809);
 * NULL for round(3 sigma) > 8 */
void *orc_pyramid_build_sigma(const uint8_t *bgr, int W, int H, int num_nb, int link, int target_clus,
                              double sigma);
/* the blur's weights ((2S+1)^2 floats, row-major) and the blurred float4 (B, G, R, w) image;
 * both return S = round(3 sigma), or -1 above 8 */
int orc_gauss_weights(double sigma, float *w);
int orc_gauss_blur(const uint8_t *bgr, int W, int H, double sigma, float *out4);
/* BGR_to_LAB on float4 pixels (w passes through) */
void orc_lab_f(const float *in4, long n, float *lab4);
int orc_pyr_levels(void *p);
void orc_pyr_level_size(void *p, int level, int *V, int *E);
void orc_pyr_get_level(void *p, int level, int *clus, float *mean, int *bd, int *bdl);
void orc_pyr_free(void *p);

/* (c) hierarchical BP on the pyramid */
void *orc_bp_new(void *pyr, const uint8_t *vol, const float *dwta, int W, int H, int D,
                 float tau, float alpha, float tau_s, int min_level, int use_cost_vol);
void *orc_bp_new_f32(void *pyr, const float *vol, const float *dwta, int W, int H, int D, float tau,
                     float alpha, float tau_s, int min_level, int use_cost_vol);
int orc_bp_step(void *bp, int max_sweeps);
int orc_bp_level(void *bp);
int orc_bp_sweeps_done(void *bp);
void orc_bp_level_size(void *bp, int level, int *V, int *E);
void orc_bp_get_state(void *bp, float *planes /* [12][V][4] */, float *cost /* [12][V] */,
                      float *belief /* [12][V] */, float *msg /* [E][8] */);
void orc_bp_map(void *bp, float *disp, float *plane3);
void orc_bp_free(void *bp);
int orc_bp_total_sweeps(int levels, int min_level, int D);

/* (d) left-right check and fill */
void orc_lr_consistency(const float *dispL, const float *dispR, int W, int H, uint8_t *failL,
                        uint8_t *failR);
void orc_dilate3(const uint8_t *in, int W, int H, uint8_t *out);
void orc_lr_fill(const uint8_t *fail, const uint8_t *fail2, float *plane3, int W, int H);
void orc_plane_to_disp(const float *plane3, int W, int H, float *disp);
void orc_left_right_check(const float *dispL, const float *dispR, float *planeL, float *planeR,
                          int W, int H, uint8_t *failL, uint8_t *failR, uint8_t *fail2L);

/* refinements after the LR check: weighted median of the failed pixels' planes (in place) and
 * the CLI's 7-tap separable median of a disparity image (in place) */
float orc_wm_weight(float absdiff);
void orc_weighted_median(const uint8_t *bgr, const uint8_t *fail, float *plane3, int W, int H);
void orc_median7(float *disp, int W, int H);
/* Evaluator.h:49-76: counts {valid, good & valid, nonocc, good & nonocc, occ, good & occ} */
void orc_eval_bad_pixels(const float *disp, const float *gt, const uint8_t *nonocc, int W, int H, double thr,
                         long long *counts);

/* whole plain path (left view) for the CPU baseline */
int orc_plain_pair(const uint8_t *bgrL, const uint8_t *bgrR, int W, int H, int D, float alpha,
                   float tau_s, int min_level, int bp_levels, float *disp, float *plane3,
                   int *levels_out);

#ifdef __cplusplus
}
#endif
#endif
