// Implementation of the C++ drop-in classes over the C ABI of libhbp_hip.so.
#include <cmath>
#include <cstring>
#include <fstream>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "CRTrees.hpp"
#include "hbp_isp.hpp"
#include "post_process.hpp"
#include "segment.hpp"
#include "stereo_cost.hpp"

namespace hbp_api {

hbp_ctx *default_ctx() {
    static std::once_flag once;
    static hbp_ctx *ctx = nullptr;
    std::call_once(once, [] { check(hbp_ctx_create(0, &ctx), "hbp_ctx_create"); });
    return ctx;
}

static void require_bgr(const cv::Mat &m) {
    if (m.empty() || m.type() != CV_8UC3 || !m.isContinuous())
        throw std::invalid_argument("images must be continuous CV_8UC3 (BGR)");
}

}  // namespace hbp_api

using hbp_api::check;
using hbp_api::default_ctx;

// ------------------------------------------------------------------ CRTrees
int get_num_from_scan(const int *predicate_d, const int *pos_scan_d, const int size) {
    int a = 0, b = 0;
    if (hipMemcpy(&a, pos_scan_d + size - 1, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&b, predicate_d + size - 1, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("get_num_from_scan: copy failed");
    return a + b;
}

void CRTrees::get_clus(int *nn_d, int *clus_d, const int N, int *mask_cycle_d) {
    if (N > size) throw std::invalid_argument("CRTrees::get_clus: N exceeds the size given at construction");
    if (mask_cycle_d) throw std::invalid_argument("CRTrees::get_clus: the cycle mask is not provided");
    check(hbp_crtrees_get_clus(default_ctx(), nn_d, clus_d, N), "CRTrees::get_clus");
}

int CRTrees::compact_clus_label(int *clus_d, const int N) {
    if (N > size) throw std::invalid_argument("CRTrees::compact_clus_label: N exceeds the size");
    int n = 0;
    check(hbp_crtrees_compact(default_ctx(), clus_d, N, &n), "CRTrees::compact_clus_label");
    return n;
}

// ------------------------------------------------------------------ SegHAC
SegHAC::SegHAC(const cv::Mat &src, const Linkage link, const int num_nb_, const double sigma,
               const SegFormat seg_format, const int target_clus, const bool show_cycle)
    : num_nb(num_nb_), m_link(link), m_target_clus(target_clus), m_sigma(sigma), width(src.cols), height(src.rows) {
    hbp_api::require_bgr(src);
    if (link < MinLink || link > WardLink) throw std::invalid_argument("SegHAC: unknown linkage");
    if (seg_format != LabelISPFormat) throw std::invalid_argument("SegHAC: only LabelISPFormat is implemented");
    if (target_clus < 1 || (long)target_clus >= (long)src.cols * src.rows)
        throw std::invalid_argument("SegHAC: target_clus must be in [1, width*height)");
    if (sigma > 1.17549435e-38 && std::round(3 * sigma) > 8)
        throw std::invalid_argument("SegHAC: round(3*sigma) must be at most 8 (the Gaussian pre-blur's window)");
    if (show_cycle) throw std::invalid_argument("SegHAC: show_cycle is not implemented");
    if (num_nb != 4 && num_nb != 8) throw std::invalid_argument("SegHAC: num_nb must be 4 or 8");
    img = src.clone();
}

void SegHAC::set_frame(const cv::Mat &frame) {
    hbp_api::require_bgr(frame);
    if (frame.cols != width || frame.rows != height) throw std::invalid_argument("set_frame: size differs");
    img = frame.clone();
}

void SegHAC::set_seq_img(const cv::Mat &m) {
    hbp_api::require_bgr(m);
    if (m.cols != width && !(m.cols == height && m.rows == width))
        throw std::invalid_argument("set_seq_img: size differs");
    width = m.cols;
    height = m.rows;
    img = m.clone();
}

void SegHAC::run_ms(std::vector<cv::Mat> *) {
    auto h = std::make_shared<hbp_api::PyramidHandle>();
    check(hbp_pyramid_build_target(default_ctx(), img.data, width, height, (int)m_link, num_nb, m_sigma, m_target_clus,
                                   &h->p),
          "SegHAC::run_ms");
    pyr = h;
    int L = hbp_pyramid_levels(pyr->p);
    num_clus_isp.assign(L, 0);
    num_bd_isp.assign(L, 0);
    for (int l = 0; l < L; ++l) check(hbp_pyramid_level_size(pyr->p, l, &num_clus_isp[l], &num_bd_isp[l]), "level");
}

int *SegHAC::g_isp_clus_d() const { return pyr ? hbp_pyramid_dev_clus(pyr->p) : nullptr; }
float *SegHAC::g_isp_mean_d() const { return pyr ? hbp_pyramid_dev_mean(pyr->p) : nullptr; }
int2 *SegHAC::g_isp_bd_d() const { return pyr ? (int2 *)hbp_pyramid_dev_bd(pyr->p) : nullptr; }
int *SegHAC::g_isp_bdl_d() const { return pyr ? hbp_pyramid_dev_bdl(pyr->p) : nullptr; }

// ------------------------------------------------------------------ StereoCost
StereoCost::StereoCost(const cv::Mat &img_left, const cv::Mat &img_right, const int labels, const std::string vd)
    : width(img_left.cols), height(img_left.rows), m_labels(labels), vol_dir(vd), im_size(width * height) {
    hbp_api::require_bgr(img_left);
    hbp_api::require_bgr(img_right);
    if (img_right.cols != width || img_right.rows != height) throw std::invalid_argument("image sizes differ");
    imgL = img_left.clone();
    imgR = img_right.clone();
}

void StereoCost::compute_stereo_cost(const std::string algo) {
    auto h = std::make_shared<hbp_api::VolumeHandle>();
    if (algo == "Census" || algo == "AD") {
        check(hbp_cost_volume(default_ctx(), imgL.data, imgR.data, width, height, m_labels, algo.c_str(), &h->v),
              "StereoCost::compute_stereo_cost");
        m_tau = algo == "AD" ? 30.f : 25.f;
    } else if (algo == "MC-CNN") {
        // read_volume_from_file (stereo_cost.cu:399-419): vol_dir names a headerless f32
        // [labels][height][width] file (loadMatBinary(..., readHeader = false))
        const size_t n = (size_t)m_labels * height * width;
        std::vector<float> raw(n);
        std::ifstream f(vol_dir, std::ios::binary);
        if (!f) throw std::runtime_error("StereoCost: cost volume file " + vol_dir + " not found");
        f.read((char *)raw.data(), (std::streamsize)(n * sizeof(float)));
        if ((size_t)f.gcount() != n * sizeof(float))
            throw std::runtime_error("StereoCost: " + vol_dir + " holds fewer than labels*height*width floats");
        check(hbp_cost_volume_f32(default_ctx(), raw.data(), width, height, m_labels, "MC-CNN", &h->v),
              "StereoCost::compute_stereo_cost");
        m_tau = 0.5f;
    } else {
        throw std::invalid_argument("StereoCost: the costs are \"Census\", \"AD\" and \"MC-CNN\" (not \"" + algo +
                                    "\")");
    }
    vol = h;
}

void StereoCost::cost_to_right() {
    if (!vol) throw std::logic_error("cost_to_right before compute_stereo_cost");
    check(hbp_cost_to_right(default_ctx(), vol->v), "StereoCost::cost_to_right");
}

float *StereoCost::g_img_cost_d() const { return vol ? hbp_volume_dev_f32(default_ctx(), vol->v) : nullptr; }
float *StereoCost::g_img_D_d() const { return vol ? hbp_volume_dev_wta(vol->v) : nullptr; }

// ------------------------------------------------------------------ ISPData / StereoData
ISPData::ISPData(const SegHAC *s)
    : width(s->g_width()), height(s->g_height()), mean_channels(s->g_mean_channels()),
      num_clus_isp(s->g_num_clus_isp()), num_bd_isp(s->g_num_bd_isp()), pyr(s->pyramid()) {
    if (!pyr) throw std::logic_error("ISPData: SegHAC::run_ms has not run");
}

StereoData::StereoData(const StereoCost *c, const bool use)
    : m_labels(c->g_labels()), im_size(c->g_im_size()), use_cost_vol(use), m_tau(c->g_tau()), vol(c->volume()) {
    if (!vol) throw std::logic_error("StereoData: compute_stereo_cost has not run");
}

// ------------------------------------------------------------------ HBP_ISP
HBP_ISP::HBP_ISP(const ISPData *isp_data, const StereoData *stereo_data, const float alpha, const float tau_smooth)
    : width(isp_data->g_width()), height(isp_data->g_height()), im_size(width * height),
      m_labels(stereo_data->g_labels()), isp_levels((int)isp_data->g_num_clus_isp().size()), isp(isp_data),
      stereo(stereo_data), m_alpha(alpha), m_tau_smooth(tau_smooth) {}

std::pair<cv::Mat, cv::Mat> HBP_ISP::run_ms(const int min_level) {
    hbp_bp *bp = nullptr;
    check(hbp_bp_create(default_ctx(), isp->handle(), stereo->handle(), m_alpha, m_tau_smooth, min_level,
                        stereo->g_use_cost_vol() ? 1 : 0, &bp),
          "HBP_ISP::run_ms");
    cv::Mat disp(height, width, CV_32FC1), label(height, width, CV_32FC3);
    int rc = hbp_bp_run(default_ctx(), bp, (float *)disp.data, (float *)label.data);
    hbp_bp_free(default_ctx(), bp);
    check(rc, "HBP_ISP::run_ms");
    return {disp, label};
}

// Lab -> BGR u8 (the cv::cvtColor(COLOR_Lab2BGR) + convertTo(CV_8UC3, 255) of draw_spx_mean,
// hbp_isp.cu:1126-1127): inverse CIE Lab with the D65 white, then sRGB companding.
static void lab_to_bgr8(float L, float a, float b, uint8_t *out) {
    double fy = (L + 16.0) / 116.0, fx = fy + a / 500.0, fz = fy - b / 200.0;
    auto finv = [](double t) { return t > 6.0 / 29.0 ? t * t * t : 3.0 * (6.0 / 29.0) * (6.0 / 29.0) * (t - 4.0 / 29.0); };
    double X = 0.950456 * finv(fx), Y = finv(fy), Z = 1.088754 * finv(fz);
    double r = 3.240479 * X - 1.53715 * Y - 0.498535 * Z;
    double g = -0.969256 * X + 1.875991 * Y + 0.041556 * Z;
    double bl = 0.055648 * X - 0.204043 * Y + 1.057311 * Z;
    auto comp = [](double c) {
        c = c <= 0.0031308 ? 12.92 * c : 1.055 * std::pow(c, 1.0 / 2.4) - 0.055;
        c = c < 0 ? 0 : (c > 1 ? 1 : c);
        return (uint8_t)std::lround(c * 255.0);
    };
    out[0] = comp(bl);
    out[1] = comp(g);
    out[2] = comp(r);
}

cv::Mat HBP_ISP::spx_mean_image(int level) const {
    int N = im_size;
    std::vector<int> im(N);
    for (int i = 0; i < N; ++i) im[i] = i;
    auto V = isp->g_num_clus_isp();
    for (int l = 0; l < level; ++l) {
        std::vector<int> clus(V[l]);
        check(hbp_pyramid_download_level(default_ctx(), isp->handle(), l, clus.data(), nullptr, nullptr, nullptr),
              "download");
        for (int i = 0; i < N; ++i) im[i] = clus[im[i]];
    }
    std::vector<float> mean(6 * (size_t)V[level]);
    check(hbp_pyramid_download_level(default_ctx(), isp->handle(), level, nullptr, mean.data(), nullptr, nullptr),
          "download");
    int Vl = V[level];
    cv::Mat out(height, width, CV_8UC3);
    for (int i = 0; i < N; ++i) {
        int c = im[i];
        float w = mean[5 * (size_t)Vl + c];
        lab_to_bgr8(mean[c] / w, mean[(size_t)Vl + c] / w, mean[2 * (size_t)Vl + c] / w, out.data + 3 * (size_t)i);
    }
    return out;
}

std::vector<std::tuple<cv::Mat, cv::Mat, cv::Mat>> HBP_ISP::run_ms_vis(const int min_level) {
    std::vector<std::tuple<cv::Mat, cv::Mat, cv::Mat>> out;
    for (int l = isp_levels - 1; l >= min_level; --l) {
        auto r = run_ms(l);
        out.emplace_back(r.first, r.second, spx_mean_image(l));
    }
    return out;
}

// ------------------------------------------------------------------ PostProcess
void PostProcess::left_right_check(std::pair<cv::Mat, cv::Mat> &L, std::pair<cv::Mat, cv::Mat> &R) {
    for (const cv::Mat *m : {&L.first, &R.first})
        if (m->type() != CV_32FC1 || m->cols != width || m->rows != height)
            throw std::invalid_argument("left_right_check: disparities must be CV_32FC1 of the image size");
    if (L.second.type() != CV_32FC3) throw std::invalid_argument("left_right_check: planes must be CV_32FC3");
    fail[0] = cv::Mat(height, width, CV_8UC1);
    fail[1] = cv::Mat(height, width, CV_8UC1);
    check(hbp_lr_check_fill(default_ctx(), (const float *)L.first.data, (const float *)R.first.data,
                            (float *)L.second.data, width, height, fail[0].data, fail[1].data, nullptr),
          "PostProcess::left_right_check");
}

void PostProcess::plane_label_to_disp(std::pair<cv::Mat, cv::Mat> &L, std::pair<cv::Mat, cv::Mat> &) {
    check(hbp_plane_to_disp(default_ctx(), (const float *)L.second.data, width, height, (float *)L.first.data),
          "PostProcess::plane_label_to_disp");
}

// post_process.cu:127-185: the left view's failed pixels (i < 1 in the reference's loop)
void PostProcess::median_filter(std::pair<cv::Mat, cv::Mat> &L, std::pair<cv::Mat, cv::Mat> &) {
    if (fail[0].empty()) throw std::logic_error("PostProcess::median_filter: run left_right_check first");
    if (L.second.type() != CV_32FC3 || L.second.cols != width || L.second.rows != height)
        throw std::invalid_argument("median_filter: planes must be CV_32FC3 of the image size");
    if (I[0].type() != CV_8UC3) throw std::invalid_argument("median_filter: the left image must be CV_8UC3");
    check(hbp_weighted_median(default_ctx(), I[0].data, fail[0].data, (float *)L.second.data, width, height),
          "PostProcess::median_filter");
}

void medianFilter(cv::Mat &D) {
    if (D.type() != CV_32FC1) throw std::invalid_argument("medianFilter: CV_32FC1 expected");
    check(hbp_median7(default_ctx(), (float *)D.data, D.cols, D.rows), "medianFilter");
}

// post_process.cu:187-371, the device twin of median_filter: the same computation
void PostProcess::median_filter_gpu(std::pair<cv::Mat, cv::Mat> &L, std::pair<cv::Mat, cv::Mat> &R) {
    median_filter(L, R);
}
