// Middlebury v3 benchmark driver with the argument order, files and call sequence of
// main_bench_middv3.cu: reads <input_dir>calib.txt, im0.png, im1.png, disp0GT.pfm and
// mask0nocc.png (the MC-CNN volume from <input_dir>im0.acrt), runs SegHAC -> ISPData ->
// StereoCost -> StereoData -> HBP_ISP::run_ms (MidV3, lines 50-130) and writes
// <output_dir>disp0HBP_ISP.pfm, timeHBP_ISP.txt and debug/log_output.txt; a "training" scene
// is evaluated with the bad-pixel threshold 1.0 (0.5 for Q, 2.0 for F).  With a trailing "lr"
// argument it runs MidV3_LR (lines 132-238) instead: the right view, the left-right check and
// fill, the weighted-median refinement.  The reference's main calls MidV3 only and keeps
// MidV3_LR unused; the "lr" switch is this driver's addition.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "hbp_isp.hpp"
#include "middlebury.hpp"
#include "post_process.hpp"
#include "segment.hpp"
#include "stereo_cost.hpp"

// loadData, main_bench_middv3.cu:12-48
static bool loadData(const std::string &input_dir, cv::Mat &im0, cv::Mat &im1, cv::Mat &dispGT, cv::Mat &nonocc,
                     Calib &calib) {
    calib = Calib(input_dir + "calib.txt");
    if (calib.ndisp <= 0) {
        std::printf("ndisp is not speficied.\n");
        return false;
    }
    im0 = hbp_api::read_png(input_dir + "im0.png", true);
    if (im0.empty()) {
        std::printf("Image im0.png not found in\n%s\n", input_dir.c_str());
        return false;
    }
    im1 = hbp_api::read_png(input_dir + "im1.png", true);
    if (im1.empty()) {
        std::printf("Image im1.png not found in\n%s\n", input_dir.c_str());
        return false;
    }
    dispGT = cvutils::io::read_pfm_file(input_dir + "disp0GT.pfm");
    if (dispGT.empty()) dispGT = cv::Mat::zeros(im0.rows, im0.cols, CV_32FC1);
    nonocc = hbp_api::read_png(input_dir + "mask0nocc.png", false);
    if (!nonocc.empty()) {
        for (size_t i = 0; i < nonocc.total(); ++i) nonocc.data[i] = nonocc.data[i] == 255 ? 255 : 0;
    } else {
        nonocc = cv::Mat(im0.rows, im0.cols, CV_8UC1);
        for (size_t i = 0; i < nonocc.total(); ++i) nonocc.data[i] = 255;
    }
    return true;
}

struct Params {
    Linkage link;
    int num_nb;
    double sigma;
    float alpha, tau_smooth;
    std::string cost_name;
    bool use_cost_vol;
    int min_level;
};

static double error_threshold(const std::string &input_dir) {
    double t = 1.0;
    if (cvutils::contains(input_dir, "trainingQ") || cvutils::contains(input_dir, "testQ"))
        t /= 2.0;
    else if (cvutils::contains(input_dir, "trainingF") || cvutils::contains(input_dir, "testF"))
        t *= 2.0;
    return t;
}

// one view: SegHAC(image) -> ISPData, StereoCost(im0, im1) [-> cost_to_right] -> StereoData, run_ms
static std::pair<cv::Mat, cv::Mat> run_view(const cv::Mat &seg_image, const cv::Mat &im0, const cv::Mat &im1,
                                            int maxdisp, const std::string &vol_dir, const Params &p, bool right,
                                            bool report) {
    auto *seg = new SegHAC(seg_image, p.link, p.num_nb, p.sigma, LabelISPFormat, 1, false);
    seg->run_ms();
    auto *isp = new ISPData(seg);
    delete seg;
    auto *sc = new StereoCost(im0, im1, maxdisp, vol_dir);
    sc->compute_stereo_cost(p.cost_name);
    if (right) sc->cost_to_right();
    auto *sd = new StereoData(sc, p.use_cost_vol);
    delete sc;
    auto c0 = std::chrono::steady_clock::now();
    auto *hbp = new HBP_ISP(isp, sd, p.alpha, p.tau_smooth);
    auto result = hbp->run_ms(p.min_level);
    auto c1 = std::chrono::steady_clock::now();
    if (report)
        std::cout << "Time for HBP-ISP inference: "
                  << std::chrono::duration_cast<std::chrono::microseconds>(c1 - c0).count() / 1e3 << " ms"
                  << std::endl;
    delete hbp;
    delete sd;
    delete isp;
    return result;
}

static void write_time(const std::string &path, const char *fmt, double t) {
    FILE *fp = std::fopen(path.c_str(), "w");
    if (fp) {
        std::fprintf(fp, fmt, t);
        std::fclose(fp);
    }
}

// MidV3 (main_bench_middv3.cu:50-130) and MidV3_LR (132-238)
static int midv3(const std::string &input_dir, const std::string &output_dir, const Params &p, bool lr) {
    cv::Mat im0, im1, dispGT, nonocc;
    Calib calib;
    if (!loadData(input_dir, im0, im1, dispGT, nonocc, calib)) return 1;
    std::printf("ndisp = %d\n", calib.ndisp);
    const int maxdisp = calib.ndisp;
    const double errorThresh = error_threshold(input_dir);
    if (!lr) std::printf("errorThresh = %f\n", errorThresh);
    const std::string vol_dir = input_dir + "im0.acrt";
    const bool training = cvutils::contains(input_dir, "training");

    mkdir((output_dir + "debug").c_str(), 0755);
    Evaluator eval(dispGT, nonocc, "result", output_dir + "debug/");
    eval.setErrorThreshold(errorThresh);
    eval.start();

    if (!lr) {
        cv::Mat disp_img = run_view(im0, im0, im1, maxdisp, vol_dir, p, false, true).first;
        cvutils::io::save_pfm_file(output_dir + "disp0HBP_ISP.pfm", disp_img);
        write_time(output_dir + "timeHBP_ISP.txt", "%lf\n", eval.getCurrentTime());
        if (training) eval.evaluate(disp_img, true, true);
        return 0;
    }
    auto results_left = run_view(im0, im0, im1, maxdisp, vol_dir, p, false, false);
    auto results_right = run_view(im1, im0, im1, maxdisp, vol_dir, p, true, false);
    if (training) eval.evaluate(results_left.first, true, true);
    PostProcess post_process(im0, im1);
    post_process.left_right_check(results_left, results_right);
    post_process.plane_label_to_disp(results_left, results_right);
    if (training) eval.evaluate(results_left.first, true, true);
    post_process.median_filter(results_left, results_right);
    post_process.plane_label_to_disp(results_left, results_right);
    cv::Mat disp_img = results_left.first.clone();
    cvutils::io::save_pfm_file(output_dir + "disp0HBP_ISP.pfm", disp_img);
    write_time(output_dir + "timeHBP_ISP.txt", "runtime %lf\n", eval.getCurrentTime());
    if (training) eval.evaluate(disp_img, true, true);
    return 0;
}

int main(int argc, char **argv) {
    const bool lr = argc == 12 && std::string(argv[11]) == "lr";
    if (argc != 11 && !lr) {
        std::cerr << "usage: " << argv[0]
                  << " input_dir output_dir linkage num_nb sigma alpha tau_smooth cost_name use_cost_vol min_level "
                     "[lr]\n"
                     "linkage: 0 - MinLink, 1 - MaxLink, 2- CentoridLink, 3 - WardLink\n";
        return 1;
    }
    const std::string input_dir(argv[1]), output_dir(argv[2]);
    std::cout << "processing: " << input_dir << std::endl;
    if (output_dir.length()) mkdir(output_dir.c_str(), 0755);
    Params p;
    p.link = static_cast<Linkage>(std::atoi(argv[3]));
    p.num_nb = std::atoi(argv[4]);
    p.sigma = std::atof(argv[5]);
    p.alpha = (float)std::atof(argv[6]);
    p.tau_smooth = (float)std::atof(argv[7]);
    p.cost_name = argv[8];
    p.use_cost_vol = std::atoi(argv[9]) != 0;
    p.min_level = std::atoi(argv[10]);
    try {
        return midv3(input_dir, output_dir, p, lr);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 2;
    }
}
