// Drop-in for stereo_cost.hpp of HBP-ISP (stereo_cost.hpp:8-132), backed by hbp_cost_volume.
// compute_stereo_cost implements "Census" (tau 25), "AD" (tau 30) and "MC-CNN" (tau 0.5: the
// headerless f32 volume file named by vol_dir); another name throws std::invalid_argument (the
// reference leaves the volume unset).  g_img_cost_d() returns
// the f32 [D][H][W] volume of the reference (for Census built on first use from the u8 volume
// the library computes with).
#ifndef _STEREO_COST_HPP_
#define _STEREO_COST_HPP_

#include <string>

#include "hbp_api_common.hpp"

class StereoCost {
public:
    StereoCost(const cv::Mat &img_left, const cv::Mat &img_right, const int labels, const std::string vol_dir);
    ~StereoCost() = default;

    void compute_stereo_cost(const std::string algo);
    void cost_to_right();

    int g_labels() const { return m_labels; }
    int g_im_size() const { return im_size; }
    float g_tau() const { return m_tau; }
    float *g_img_cost_d() const;
    float *g_img_D_d() const;

    // the volume object behind the getters (shared with StereoData)
    std::shared_ptr<hbp_api::VolumeHandle> volume() const { return vol; }

protected:
    const int width;
    const int height;
    const int m_labels;
    cv::Mat imgL, imgR;

private:
    const std::string vol_dir;
    int im_size;
    float m_tau = 25.f;
    std::shared_ptr<hbp_api::VolumeHandle> vol;
};

#endif
