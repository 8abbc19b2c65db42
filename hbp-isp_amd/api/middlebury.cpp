// Middlebury harness pieces (see middlebury.hpp): calib.txt, PFM, PNG and the evaluator.
#include "middlebury.hpp"

#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <vector>

// ------------------------------------------------------------------ Calib (ArgsParser.h:27-71)
Calib::Calib()
    : cam0{}, cam1{}, doffs(0), baseline(0), width(0), height(0), ndisp(0), isint(0), vmin(0), vmax(0), dyavg(0),
      dymax(0) {}

Calib::Calib(std::string filename) : Calib() {
    FILE *fp = std::fopen(filename.c_str(), "r");
    if (!fp) return;
    char buff[512];
    // one line per field, in the file's order
    if (std::fgets(buff, sizeof(buff), fp))
        std::sscanf(buff, "cam0 = [%f %f %f; %f %f %f; %f %f %f]", &cam0[0][0], &cam0[0][1], &cam0[0][2], &cam0[1][0],
                    &cam0[1][1], &cam0[1][2], &cam0[2][0], &cam0[2][1], &cam0[2][2]);
    if (std::fgets(buff, sizeof(buff), fp))
        std::sscanf(buff, "cam1 = [%f %f %f; %f %f %f; %f %f %f]", &cam1[0][0], &cam1[0][1], &cam1[0][2], &cam1[1][0],
                    &cam1[1][1], &cam1[1][2], &cam1[2][0], &cam1[2][1], &cam1[2][2]);
    if (std::fgets(buff, sizeof(buff), fp)) std::sscanf(buff, "doffs = %f", &doffs);
    if (std::fgets(buff, sizeof(buff), fp)) std::sscanf(buff, "baseline = %f", &baseline);
    if (std::fgets(buff, sizeof(buff), fp)) std::sscanf(buff, "width = %d", &width);
    if (std::fgets(buff, sizeof(buff), fp)) std::sscanf(buff, "height = %d", &height);
    if (std::fgets(buff, sizeof(buff), fp)) std::sscanf(buff, "ndisp = %d", &ndisp);
    if (std::fgets(buff, sizeof(buff), fp)) std::sscanf(buff, "isint = %d", &isint);
    if (std::fgets(buff, sizeof(buff), fp)) std::sscanf(buff, "vmin = %d", &vmin);
    if (std::fgets(buff, sizeof(buff), fp)) std::sscanf(buff, "vmax = %d", &vmax);
    if (std::fgets(buff, sizeof(buff), fp)) std::sscanf(buff, "dyavg = %f", &dyavg);
    if (std::fgets(buff, sizeof(buff), fp)) std::sscanf(buff, "dymax = %f", &dymax);
    std::fclose(fp);
}

// ------------------------------------------------------------------ PFM (cvutils.hpp:23-144)
namespace cvutils {
namespace io {

static bool host_little_endian() {
    const uint16_t v = 1;
    uint8_t b;
    std::memcpy(&b, &v, 1);
    return b == 1;
}

cv::Mat read_pfm_file(const std::string &filename) {
    FILE *f = std::fopen(filename.c_str(), "rb");
    if (!f) return cv::Mat();
    char magic[8] = {0};
    int w = 0, h = 0;
    double scale = 1.0;
    if (std::fscanf(f, "%7s", magic) != 1 || std::fscanf(f, "%d %d", &w, &h) != 2 ||
        std::fscanf(f, "%lf", &scale) != 1 || w <= 0 || h <= 0) {
        std::fclose(f);
        return cv::Mat();
    }
    int channels = std::strcmp(magic, "Pf") == 0 ? 1 : (std::strcmp(magic, "PF") == 0 ? 3 : 0);
    if (!channels) {
        std::fclose(f);
        return cv::Mat();
    }
    // the data are the file's last w*h*channels floats
    const size_t count = (size_t)w * h * channels;
    std::vector<uint8_t> data(count * 4);
    if (std::fseek(f, -(long)(count * 4), SEEK_END) != 0 || std::fread(data.data(), 4, count, f) != count) {
        std::fclose(f);
        return cv::Mat();
    }
    std::fclose(f);
    const bool file_le = scale < 0.0;
    cv::Mat image(h, w, channels == 1 ? CV_32FC1 : CV_32FC3);
    for (size_t i = 0; i < count; ++i) {
        uint8_t *p = &data[4 * i];
        if (file_le != host_little_endian()) {
            std::swap(p[0], p[3]);
            std::swap(p[1], p[2]);
        }
        const size_t px = i / channels, ch = i % channels;
        const size_t x = px % w, row = px / w;
        float v;
        std::memcpy(&v, p, 4);
        image.ptr<float>((int)(h - 1 - row))[x * channels + ch] = v;
    }
    return image;
}

void save_pfm_file(const std::string &filename, const cv::Mat &image) {
    const int channels = image.channels();
    if (channels != 1 && channels != 3) throw std::invalid_argument("save_pfm_file: 1 or 3 channels expected");
    FILE *s = std::fopen(filename.c_str(), "wb");
    if (!s) throw std::runtime_error("cannot write " + filename);
    std::fprintf(s, "%s\n%d %d\n%lf\n", channels == 1 ? "Pf" : "PF", image.cols, image.rows,
                 host_little_endian() ? -1.0 / 255.0 : 1.0 / 255.0);
    for (int y = image.rows - 1; y >= 0; --y)
        std::fwrite(image.ptr<float>(y), sizeof(float), (size_t)image.cols * channels, s);
    std::fclose(s);
}

}  // namespace io
}  // namespace cvutils

// ------------------------------------------------------------------ PNG
namespace hbp_api {

static uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

cv::Mat read_png(const std::string &path, bool color) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return cv::Mat();
    std::vector<uint8_t> file((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (file.size() < 8 || std::memcmp(file.data(), sig, 8) != 0) throw std::runtime_error(path + ": not a PNG file");
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, palette;
    for (size_t pos = 8; pos + 12 <= file.size();) {
        const uint32_t len = be32(&file[pos]);
        if (pos + 12 + (size_t)len > file.size()) throw std::runtime_error(path + ": truncated PNG");
        const char *type = (const char *)&file[pos + 4];
        const uint8_t *d = &file[pos + 8];
        if (std::memcmp(type, "IHDR", 4) == 0 && len >= 13) {
            W = be32(d);
            H = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (std::memcmp(type, "PLTE", 4) == 0) {
            palette.assign(d, d + len);
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), d, d + len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            break;
        }
        pos += 12 + (size_t)len;
    }
    const int nch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (depth != 8 || nch == 0 || interlace != 0 || W == 0 || H == 0 || (uint64_t)W * H > (1ull << 28))
        throw std::runtime_error(path + ": only 8-bit, non-interlaced PNG images are read");
    const size_t stride = (size_t)W * nch;
    std::vector<uint8_t> raw((stride + 1) * H);
    uLongf out_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &out_len, idat.data(), (uLong)idat.size()) != Z_OK || out_len != raw.size())
        throw std::runtime_error(path + ": corrupt PNG data");
    // undo the row filters (None, Sub, Up, Average, Paeth)
    std::vector<uint8_t> img(stride * H);
    for (uint32_t y = 0; y < H; ++y) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t *src = &raw[y * (stride + 1) + 1];
        uint8_t *cur = &img[y * stride];
        const uint8_t *up = y ? &img[(y - 1) * stride] : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= (size_t)nch ? cur[i - nch] : 0, b = up ? up[i] : 0,
                      c = (up && i >= (size_t)nch) ? up[i - nch] : 0;
            int v = src[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: {
                    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: throw std::runtime_error(path + ": bad PNG filter");
            }
            cur[i] = (uint8_t)v;
        }
    }
    if (!color) {
        if (ctype != 0 && ctype != 4) throw std::runtime_error(path + ": a grey PNG is expected");
        cv::Mat m((int)H, (int)W, CV_8UC1);
        for (size_t i = 0; i < (size_t)W * H; ++i) m.data[i] = img[i * nch];
        return m;
    }
    cv::Mat m((int)H, (int)W, CV_8UC3);
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        uint8_t r, g, b;
        if (ctype == 3) {
            const size_t k = 3 * (size_t)img[i];
            if (k + 2 >= palette.size()) throw std::runtime_error(path + ": bad palette");
            r = palette[k];
            g = palette[k + 1];
            b = palette[k + 2];
        } else if (nch <= 2) {
            r = g = b = img[i * nch];
        } else {
            r = img[i * nch];
            g = img[i * nch + 1];
            b = img[i * nch + 2];
        }
        m.data[3 * i] = b;
        m.data[3 * i + 1] = g;
        m.data[3 * i + 2] = r;
    }
    return m;
}

}  // namespace hbp_api

// ------------------------------------------------------------------ Evaluator (Evaluator.h)
Evaluator::Evaluator(cv::Mat gt, cv::Mat mask, std::string hdr, std::string dir, bool show, bool print, bool save)
    : dispGT(gt), nonoccMask(mask), saveDir(dir), header(hdr), showProgress(show), saveProgress(save),
      printProgress(print) {
    if (dispGT.type() != CV_32FC1 || nonoccMask.type() != CV_8UC1 || dispGT.rows != nonoccMask.rows ||
        dispGT.cols != nonoccMask.cols)
        throw std::invalid_argument("Evaluator: a CV_32FC1 ground truth and a CV_8UC1 mask of one size are expected");
    if (save) {
        fp_output = std::fopen((saveDir + "log_output.txt").c_str(), "w");
        if (fp_output) {
            std::fprintf(fp_output, "%s\t%s\t%s\n", "Time", "all", "nonocc");
            std::fflush(fp_output);
        }
    }
}

Evaluator::~Evaluator() {
    if (fp_output) std::fclose(fp_output);
}

void Evaluator::evaluate(cv::Mat disp, bool save, bool print, int) {
    if (disp.type() != CV_32FC1 || disp.rows != dispGT.rows || disp.cols != dispGT.cols)
        throw std::invalid_argument("Evaluator::evaluate: a CV_32FC1 disparity of the ground truth's size is expected");
    long long c[6];
    hbp_api::check(hbp_eval_bad_pixels(hbp_api::default_ctx(), (const float *)disp.data, (const float *)dispGT.data,
                                       nonoccMask.data, disp.cols, disp.rows, errorThreshold, c),
                   "Evaluator::evaluate");
    all = 100.0 * (1.0 - (double)c[1] / (double)c[0]);
    nonocc = 100.0 * (1.0 - (double)c[3] / (double)c[2]);
    occ = 100.0 * (1.0 - (double)c[5] / (double)c[4]);
    if (saveProgress && save && fp_output) {
        std::fprintf(fp_output, "%lf\t%lf\t%lf\t%lf\n", getCurrentTime(), all, nonocc, occ);
        std::fflush(fp_output);
    }
    if (printProgress && print) std::printf("%5.1lf\t%4.2lf\t%4.2lf\t%4.2lf\n", getCurrentTime(), all, nonocc, occ);
}
