// Drop-in for segment.hpp of HBP-ISP (segment.hpp:8-340): the CRTrees superpixel pyramid,
// backed by hbp_pyramid_build_target.  Implemented configuration: the four linkages
// (CentroidLink, the one every script of the reference uses, MinLink, MaxLink, WardLink),
// LabelISPFormat, any target_clus in [1, width*height) (1 in every script), sigma with
// round(3*sigma) <= 8 (0 in every script; above FLT_MIN the Gaussian pre-blur runs), num_nb 4
// or 8; other values throw std::invalid_argument.
#ifndef _SEGMENT_HPP_
#define _SEGMENT_HPP_

#include <vector>

#include "CRTrees.hpp"
#include "hbp_api_common.hpp"

enum Linkage { MinLink, MaxLink, CentoridLink, WardLink };

enum SegFormat { LabelUchar3Format, LabelIntFormat, BoundaryFormat, LabelDeviceIntFormat, LabelISPFormat, EmptyFormat };

class SegHAC {
public:
    SegHAC(const cv::Mat &src, const Linkage link, const int num_nb, const double sigma,
           const SegFormat seg_format, const int target_clus, const bool show_cycle);
    ~SegHAC() = default;

    // segment.cu:1279-1321; segs is unused for LabelISPFormat
    void run_ms(std::vector<cv::Mat> *segs = nullptr);
    // segment.hpp:87-108: replace the image (same size, or transposed for set_seq_img)
    void set_frame(const cv::Mat &frame);
    void set_seq_img(const cv::Mat &img);

    int g_width() const { return width; }
    int g_height() const { return height; }
    int g_mean_channels() const { return mean_channels; }
    std::vector<int> g_num_clus_isp() const { return num_clus_isp; }
    std::vector<int> g_num_bd_isp() const { return num_bd_isp; }
    int *g_isp_clus_d() const;
    float *g_isp_mean_d() const;
    int2 *g_isp_bd_d() const;
    int *g_isp_bdl_d() const;

    // the pyramid object behind the getters (shared with ISPData)
    std::shared_ptr<hbp_api::PyramidHandle> pyramid() const { return pyr; }

protected:
    std::vector<int> num_clus_isp;
    std::vector<int> num_bd_isp;

private:
    cv::Mat img;
    const int num_nb;
    const Linkage m_link;
    const int m_target_clus;
    const double m_sigma;
    int width, height;
    const int mean_channels = 6;
    std::shared_ptr<hbp_api::PyramidHandle> pyr;
};

#endif
