// Test driver of the drop-in classes (built into hbp-isp_amd/bin/api_check by its Makefile; run
// by tests/test_gpu_api.py): SegHAC's video hooks and StereoCost's float getter.
//   api_check OUT W H D A.bgr B.bgr T.bgr
// A, B: BGR u8 [H][W][3]; T: BGR u8 [W][H][3] (a transposed frame).  Writes into OUT:
//   seg_<tag>_<level>.bin  (clus int32 [V], then bd int32 [E][2]) for tag a (A), b (after
//   set_frame(B)), t (after set_seq_img(T)); vol.bin: g_img_cost_d() of StereoCost(A, B, D)
//   with "Census", f32 [D][H][W]; wta.bin: g_img_D_d(), f32 [H][W].
#include <hip/hip_runtime.h>

#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "segment.hpp"
#include "stereo_cost.hpp"

static cv::Mat read_bgr(const std::string &path, int rows, int cols) {
    cv::Mat m(rows, cols, CV_8UC3);
    std::ifstream f(path, std::ios::binary);
    f.read((char *)m.data, (std::streamsize)rows * cols * 3);
    if (!f) throw std::runtime_error("cannot read " + path);
    return m;
}

template <typename T>
static std::vector<T> from_device(const T *d, size_t n) {
    std::vector<T> h(n);
    if (hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("hipMemcpy failed");
    return h;
}

static void write_levels(const std::string &dir, const std::string &tag, const SegHAC &s) {
    const std::vector<int> V = s.g_num_clus_isp(), E = s.g_num_bd_isp();
    size_t vo = 0, eo = 0;
    for (size_t l = 0; l < V.size(); ++l) {
        std::vector<int> clus = from_device(s.g_isp_clus_d() + vo, V[l]);
        std::vector<int2> bd = from_device(s.g_isp_bd_d() + eo, E[l]);
        std::ofstream f(dir + "/seg_" + tag + "_" + std::to_string(l) + ".bin", std::ios::binary);
        f.write((const char *)clus.data(), clus.size() * 4);
        f.write((const char *)bd.data(), bd.size() * 8);
        vo += V[l];
        eo += E[l];
    }
}

int main(int argc, char **argv) {
    if (argc != 8) {
        std::fprintf(stderr, "usage: api_check OUT W H D A.bgr B.bgr T.bgr\n");
        return 2;
    }
    try {
        const std::string out = argv[1];
        const int W = std::stoi(argv[2]), H = std::stoi(argv[3]), D = std::stoi(argv[4]);
        cv::Mat A = read_bgr(argv[5], H, W), B = read_bgr(argv[6], H, W), T = read_bgr(argv[7], W, H);
        SegHAC seg(A, CentoridLink, 4, 0.0, LabelISPFormat, 1, false);
        seg.run_ms();
        write_levels(out, "a", seg);
        seg.set_frame(B);
        seg.run_ms();
        write_levels(out, "b", seg);
        seg.set_seq_img(T);
        seg.run_ms();
        write_levels(out, "t", seg);
        StereoCost sc(A, B, D, "");
        sc.compute_stereo_cost("Census");
        std::vector<float> vol = from_device(sc.g_img_cost_d(), (size_t)D * H * W);
        std::vector<float> wta = from_device(sc.g_img_D_d(), (size_t)H * W);
        std::ofstream(out + "/vol.bin", std::ios::binary).write((const char *)vol.data(), vol.size() * 4);
        std::ofstream(out + "/wta.bin", std::ios::binary).write((const char *)wta.data(), wta.size() * 4);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "api_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
