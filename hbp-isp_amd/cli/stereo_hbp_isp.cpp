// Single-pair driver with the argument order and call sequence of main_stereo_hbp_isp.cu
// (lines 8-118): SegHAC -> run_ms -> ISPData -> StereoCost -> StereoData -> HBP_ISP::run_ms ->
// medianFilter -> PFM.  Images are binary PPM (P6) files; the output is
// <out_name>_hbp_isp_disp.pfm.  With a trailing "lr" argument it also runs the right view, the
// left-right check + fill and the weighted-median refinement (the sequence the reference keeps
// commented out at main_stereo_hbp_isp.cu:84-113, and MidV3_LR, main_bench_middv3.cu:132-238)
// and writes <out_name>_hbp_isp_disp_lr.pfm.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "hbp_isp.hpp"
#include "post_process.hpp"
#include "segment.hpp"
#include "stereo_cost.hpp"

static cv::Mat read_ppm(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string magic;
    f >> magic;
    auto skip = [&]() {
        while (f >> std::ws && f.peek() == '#') {
            std::string line;
            std::getline(f, line);
        }
    };
    int w, h, maxv;
    skip();
    f >> w;
    skip();
    f >> h;
    skip();
    f >> maxv;
    f.get();
    if (magic != "P6" || maxv != 255) throw std::runtime_error(path + ": expected an 8-bit binary PPM (P6)");
    std::vector<unsigned char> rgb((size_t)w * h * 3);
    f.read((char *)rgb.data(), (std::streamsize)rgb.size());
    cv::Mat m(h, w, CV_8UC3);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        m.data[3 * i] = rgb[3 * i + 2];
        m.data[3 * i + 1] = rgb[3 * i + 1];
        m.data[3 * i + 2] = rgb[3 * i];
    }
    return m;
}

// cvutils::io::save_pfm_file (cvutils.hpp:92-144), one channel
static void save_pfm(const std::string &path, const cv::Mat &img) {
    FILE *s = std::fopen(path.c_str(), "wb");
    if (!s) throw std::runtime_error("cannot write " + path);
    std::fprintf(s, "Pf\n%d %d\n%lf\n", img.cols, img.rows, -1.0 / 255.0);
    for (int y = img.rows - 1; y >= 0; --y) std::fwrite(img.ptr<float>(y), sizeof(float), (size_t)img.cols, s);
    std::fclose(s);
}

int main(int argc, char **argv) {
    if (argc != 13 && argc != 14 && argc != 15) {
        std::cerr << "usage: " << argv[0]
                  << " img_left.ppm img_right.ppm disp_labels linkage num_nb sigma alpha tau_smooth "
                     "out_name use_cost_vol min_level cost_name [cost_vol] [lr]\n";
        return 1;
    }
    try {
        cv::Mat L = read_ppm(argv[1]), R = read_ppm(argv[2]);
        int labels = std::atoi(argv[3]);
        Linkage link = static_cast<Linkage>(std::atoi(argv[4]));
        int num_nb = std::atoi(argv[5]);
        double sigma = std::atof(argv[6]);
        float alpha = (float)std::atof(argv[7]);
        float tau_s = (float)std::atof(argv[8]);
        std::string out = argv[9];
        bool use_cost_vol = std::atoi(argv[10]) != 0;
        int min_level = std::atoi(argv[11]);
        std::string cost_name = argv[12];
        bool lr = std::string(argv[argc - 1]) == "lr";
        std::string vol_dir = (argc >= 14 && std::string(argv[13]) != "lr") ? argv[13] : "";
        using clk = std::chrono::steady_clock;
        auto ms = [](clk::time_point a, clk::time_point b) {
            return std::chrono::duration_cast<std::chrono::microseconds>(b - a).count() / 1000.0;
        };

        auto t0 = clk::now();
        auto *seg = new SegHAC(L, link, num_nb, sigma, LabelISPFormat, 1, false);
        seg->run_ms();
        auto t1 = clk::now();
        std::cout << "time for hierarchical segmentation: " << ms(t0, t1) << " ms, "
                  << seg->g_num_clus_isp().size() << " levels" << std::endl;
        auto *isp = new ISPData(seg);
        delete seg;
        auto *sc = new StereoCost(L, R, labels, vol_dir);
        t0 = clk::now();
        sc->compute_stereo_cost(cost_name);
        t1 = clk::now();
        std::cout << "time for matching cost computation: " << ms(t0, t1) << " ms" << std::endl;
        auto *sd = new StereoData(sc, use_cost_vol);
        delete sc;
        auto *hbp = new HBP_ISP(isp, sd, alpha, tau_s);
        t0 = clk::now();
        auto left = hbp->run_ms(min_level);
        t1 = clk::now();
        std::cout << "time for belief propagation on ISP: " << ms(t0, t1) << " ms" << std::endl;
        cv::Mat disp = left.first.clone();
        medianFilter(disp);
        save_pfm(out + "_hbp_isp_disp.pfm", disp);
        delete hbp;
        delete sd;
        delete isp;

        if (lr) {
            auto *seg_r = new SegHAC(R, link, num_nb, sigma, LabelISPFormat, 1, false);
            seg_r->run_ms();
            auto *isp_r = new ISPData(seg_r);
            delete seg_r;
            auto *sc_r = new StereoCost(L, R, labels, vol_dir);
            sc_r->compute_stereo_cost(cost_name);
            sc_r->cost_to_right();
            auto *sd_r = new StereoData(sc_r, use_cost_vol);
            delete sc_r;
            auto *hbp_r = new HBP_ISP(isp_r, sd_r, alpha, tau_s);
            auto right = hbp_r->run_ms(min_level);
            delete hbp_r;
            delete sd_r;
            delete isp_r;
            // main_stereo_hbp_isp.cu:107-113: check + fill, weighted-median refinement
            PostProcess pp(L, R);
            pp.left_right_check(left, right);
            pp.plane_label_to_disp(left, right);
            pp.median_filter(left, right);
            pp.plane_label_to_disp(left, right);
            save_pfm(out + "_hbp_isp_disp_lr.pfm", left.first);
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
