// Drop-in for hbp_isp.hpp of HBP-ISP (hbp_isp.hpp:9-442): ISPData, StereoData, HBP_ISP,
// backed by hbp_bp_create / hbp_bp_run.
//   * ISPData and StereoData share the immutable device objects of SegHAC / StereoCost
//     instead of deep-copying them; the source objects may still be deleted first.
//   * HBP_ISP does not normalise ISPData's means in place, so run_ms may be called again.
#ifndef _HBP_ISP_HPP_
#define _HBP_ISP_HPP_

#include <tuple>
#include <utility>
#include <vector>

#include "hbp_api_common.hpp"
#include "segment.hpp"
#include "stereo_cost.hpp"

class ISPData {
public:
    explicit ISPData(const SegHAC *seg_hac);
    int g_width() const { return width; }
    int g_height() const { return height; }
    int g_mean_channels() const { return mean_channels; }
    std::vector<int> g_num_clus_isp() const { return num_clus_isp; }
    std::vector<int> g_num_bd_isp() const { return num_bd_isp; }
    int *g_isp_clus_d() const { return hbp_pyramid_dev_clus(pyr->p); }
    float *g_isp_mean_d() const { return hbp_pyramid_dev_mean(pyr->p); }
    int2 *g_isp_bd_d() const { return (int2 *)hbp_pyramid_dev_bd(pyr->p); }
    int *g_isp_bdl_d() const { return hbp_pyramid_dev_bdl(pyr->p); }
    hbp_pyramid *handle() const { return pyr->p; }

private:
    int width, height, mean_channels;
    std::vector<int> num_clus_isp, num_bd_isp;
    std::shared_ptr<hbp_api::PyramidHandle> pyr;
};

class StereoData {
public:
    StereoData(const StereoCost *stereo_cost, const bool use_cost_vol);
    int g_labels() const { return m_labels; }
    int g_im_size() const { return im_size; }
    float *g_im_D_d() const { return hbp_volume_dev_wta(vol->v); }
    bool g_use_cost_vol() const { return use_cost_vol; }
    float *g_img_cost_d() const { return use_cost_vol ? hbp_volume_dev_f32(hbp_api::default_ctx(), vol->v) : nullptr; }
    float g_tau() const { return m_tau; }
    hbp_volume *handle() const { return vol->v; }

private:
    int m_labels, im_size;
    bool use_cost_vol;
    float m_tau;
    std::shared_ptr<hbp_api::VolumeHandle> vol;
};

class HBP_ISP {
public:
    HBP_ISP(const ISPData *isp_data, const StereoData *stereo_data, const float alpha, const float tau_smooth);
    ~HBP_ISP() = default;
    // hbp_isp.cu:1313-1346: (CV_32FC1 disparity, CV_32FC3 planes a b c)
    std::pair<cv::Mat, cv::Mat> run_ms(const int min_level);
    // hbp_isp.cu:1348-1383: per level from the top down to min_level, (disparity, planes,
    // superpixel mean colour as CV_8UC3 BGR); the graph overlay of draw_spx_mean is not drawn
    std::vector<std::tuple<cv::Mat, cv::Mat, cv::Mat>> run_ms_vis(const int min_level);

protected:
    int width, height;
    const int im_size;
    const int m_labels;
    const int isp_levels;

private:
    const ISPData *isp;
    const StereoData *stereo;
    const float m_alpha, m_tau_smooth;
    cv::Mat spx_mean_image(int level) const;
};

#endif
