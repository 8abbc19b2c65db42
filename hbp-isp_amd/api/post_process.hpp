// Drop-in for post_process.hpp of HBP-ISP (post_process.hpp:8-61): left-right check and fill,
// plane_label_to_disp and the weighted-median refinement (median_filter and its twin
// median_filter_gpu, the same device computation) on the device; the reference runs the
// check and the refinement on the host with OpenCV.  median_filter uses the failure masks of
// the last left_right_check.
#ifndef _POST_PROCESS_HPP_
#define _POST_PROCESS_HPP_

#include <utility>

#include "hbp_api_common.hpp"

class PostProcess {
public:
    PostProcess(const cv::Mat &imL, const cv::Mat &imR)
        : I{imL.clone(), imR.clone()}, width(imL.cols), height(imL.rows), im_size(width * height) {}
    ~PostProcess() = default;

    void left_right_check(std::pair<cv::Mat, cv::Mat> &results_left, std::pair<cv::Mat, cv::Mat> &results_right);
    void median_filter(std::pair<cv::Mat, cv::Mat> &, std::pair<cv::Mat, cv::Mat> &);
    void median_filter_gpu(std::pair<cv::Mat, cv::Mat> &, std::pair<cv::Mat, cv::Mat> &);
    void plane_label_to_disp(std::pair<cv::Mat, cv::Mat> &results_left, std::pair<cv::Mat, cv::Mat> &results_right);

    // failure masks of the last check (u8, 255 = failed)
    cv::Mat fail[2];

private:
    cv::Mat I[2];   // the BGR images (the weights of the refinement)
    const double lrc_thresh = 1.5;
    const int width;
    const int height;
    const int im_size;
};

// medianFilter of the reference's CLI (main_stereo_hbp_isp.cu:120-193): 7-tap horizontal then
// vertical median of a CV_32FC1 disparity image, in place, on the device
void medianFilter(cv::Mat &D);

#endif
