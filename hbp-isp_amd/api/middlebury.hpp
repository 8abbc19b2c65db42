// The Middlebury harness pieces of HBP-ISP's benchmark driver (main_bench_middv3.cu), with the
// reference's names: Calib (ArgsParser.h:27-71: calib.txt), cvutils::io::read_pfm_file /
// save_pfm_file (cvutils.hpp:23-144), cvutils::contains, TimeStamper (TimeStamper.h) and
// Evaluator (Evaluator.h: bad-pixel rates over all valid, non-occluded and occluded pixels,
// counted on the device).  hbp_api::read_png replaces cv::imread for the PNG images and masks
// when OpenCV is absent.
//
// Deviations: TimeStamper measures wall time (the reference's clock() counts host CPU time,
// which misses time the host spends blocked on the device); Evaluator writes its log but not
// the error-map images (no image encoder here).
#ifndef _HBP_MIDDLEBURY_HPP_
#define _HBP_MIDDLEBURY_HPP_

#include <chrono>
#include <cstdio>
#include <string>

#include "hbp_api_common.hpp"

// calib.txt of a Middlebury scene (ArgsParser.h:27-71)
class Calib {
public:
    float cam0[3][3], cam1[3][3];
    float doffs, baseline;
    int width, height, ndisp, isint, vmin, vmax;
    float dyavg, dymax;
    Calib();
    explicit Calib(std::string filename);
};

namespace cvutils {
namespace io {
// PFM ("Pf" one channel, "PF" three), rows stored bottom to top; a negative scale means
// little endian (cvutils.hpp:23-144).  An absent or malformed file gives an empty Mat.
cv::Mat read_pfm_file(const std::string &filename);
void save_pfm_file(const std::string &filename, const cv::Mat &image);
}  // namespace io
inline bool contains(const std::string &str1, const std::string &str2) { return str1.find(str2) != std::string::npos; }
}  // namespace cvutils

namespace hbp_api {
// 8-bit, non-interlaced PNG: grey, grey + alpha, RGB, RGBA or palette.  color = true gives
// CV_8UC3 in BGR order (alpha dropped, grey replicated), false gives CV_8UC1 from a grey
// image.  An absent file gives an empty Mat; an unsupported one throws.
cv::Mat read_png(const std::string &path, bool color);
}  // namespace hbp_api

class TimeStamper {
    using clk = std::chrono::steady_clock;
    clk::time_point previous;
    double totalTime = 0.0;
    bool _isTicking = false;

public:
    void start() {
        if (!_isTicking) {
            _isTicking = true;
            previous = clk::now();
        }
    }
    void stop() {
        if (_isTicking) {
            totalTime += std::chrono::duration<double>(clk::now() - previous).count();
            _isTicking = false;
        }
    }
    bool isTicking() const { return _isTicking; }
    double getCurrentTime() const {
        double t = totalTime;
        if (_isTicking) t += std::chrono::duration<double>(clk::now() - previous).count();
        return t;
    }
};

// Evaluator.h: bad-pixel rates (percent) of a disparity image against the ground truth, over
// valid = (GT > 0 and GT != +inf), nonocc (the mask) and occ = (valid and not nonocc); a pixel
// is bad when |disp - GT| > the error threshold (0.5 by default).
class Evaluator {
protected:
    TimeStamper timer;
    const cv::Mat dispGT;
    const cv::Mat nonoccMask;
    std::string saveDir;
    std::string header;
    FILE *fp_output = nullptr;
    double errorThreshold = 0.5;

public:
    bool showProgress, saveProgress, printProgress;
    double all = 0.0, nonocc = 0.0, occ = 0.0;   // the rates of the last evaluate

    Evaluator(cv::Mat dispGT, cv::Mat nonoccMask, std::string header = "result", std::string saveDir = "./",
              bool show = true, bool print = true, bool save = true);
    ~Evaluator();
    Evaluator(const Evaluator &) = delete;
    Evaluator &operator=(const Evaluator &) = delete;

    std::string getSaveDirectory() { return saveDir; }
    void setErrorThreshold(double t) { errorThreshold = t; }
    void evaluate(cv::Mat disp, bool save, bool print, int level = -1);
    void start() { timer.start(); }
    void stop() { timer.stop(); }
    double getCurrentTime() { return timer.getCurrentTime(); }
};

#endif
