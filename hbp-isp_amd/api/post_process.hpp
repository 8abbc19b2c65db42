// Drop-in for post_process.hpp of HBP-ISP (post_process.hpp:8-61): left-right check and fill
// and plane_label_to_disp on the device (the reference runs them on the host with OpenCV).
// median_filter / median_filter_gpu (the weighted-median refinement) are not provided and
// throw std::logic_error.
#ifndef _POST_PROCESS_HPP_
#define _POST_PROCESS_HPP_

#include <utility>

#include "hbp_api_common.hpp"

class PostProcess {
public:
    PostProcess(const cv::Mat &imL, const cv::Mat &imR) : width(imL.cols), height(imL.rows), im_size(width * height) {}
    ~PostProcess() = default;

    void left_right_check(std::pair<cv::Mat, cv::Mat> &results_left, std::pair<cv::Mat, cv::Mat> &results_right);
    void median_filter(std::pair<cv::Mat, cv::Mat> &, std::pair<cv::Mat, cv::Mat> &);
    void median_filter_gpu(std::pair<cv::Mat, cv::Mat> &, std::pair<cv::Mat, cv::Mat> &);
    void plane_label_to_disp(std::pair<cv::Mat, cv::Mat> &results_left, std::pair<cv::Mat, cv::Mat> &results_right);

    // failure masks of the last check (u8, 255 = failed)
    cv::Mat fail[2];

private:
    const double lrc_thresh = 1.5;
    const int width;
    const int height;
    const int im_size;
};

#endif
