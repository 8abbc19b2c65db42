// Oracle (test infrastructure only): pyramid data shared by the oracle's stages.
#pragma once
#include <cstdint>
#include <vector>

namespace orc {

// One pushed level in the reference's packed ISP format (segment.cu:884-905):
// clus[V] -> id at the next level, mean[6][V] = per-node sums of (L, a, b, x, y, w)
// (level 0: the pixel values), bd[E][2] edges (s, t), bdl[E] boundary lengths.
struct Level {
    int V = 0;
    long E = 0;
    std::vector<int> clus;
    std::vector<float> mean;
    std::vector<int> bd;
    std::vector<int> bdl;
};

struct Pyramid {
    int W = 0, H = 0;
    int top = 1;   // clusters the last level maps to: 1, or target_clus
    std::vector<Level> levels;
};

void lab_image(const uint8_t *bgr, int W, int H, float *lab4);
void lab_image_f(const float *in4, long n, float *lab4);
int gauss_weights(double sigma, std::vector<float> &w);
int gauss_blur(const uint8_t *bgr, int W, int H, double sigma, float *out4);
void nn_grid(const float *lab4, int W, int H, int num_nb, int *nn, float *dist_min, int *bd,
             float *edist = nullptr);
void cc_max(const int *nn, int V, int *label);
int compact(int *label, int V);
void crtrees_literal(const int *nn, int V, int order, int *out);
// linkages as in segment.hpp:13: 0 MinLink, 1 MaxLink, 2 CentroidLink, 3 WardLink
enum { kMinLink = 0, kMaxLink = 1, kCentroidLink = 2, kWardLink = 3 };
bool pyramid_build(const uint8_t *bgr, int W, int H, int num_nb, Pyramid &P, int link = kCentroidLink,
                   int target_clus = 1, double sigma = 0.0);

}  // namespace orc
