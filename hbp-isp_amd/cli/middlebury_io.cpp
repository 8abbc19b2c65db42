// Format utility for Middlebury scene files, built on the harness's readers and writers (no
// device work):
//   middlebury_io png <in.png> <color 0|1> <out.raw>   decoded pixels, row-major (BGR when color)
//   middlebury_io pfm <in.pfm> <out.pfm>               read then write (rows top to bottom on read)
//   middlebury_io calib <calib.txt>                    the parsed fields, one "name value" a line
// It prints "<rows> <cols> <channels>" for png and pfm.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "middlebury.hpp"

int main(int argc, char **argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    try {
        if (mode == "png" && argc == 5) {
            cv::Mat m = hbp_api::read_png(argv[2], std::atoi(argv[3]) != 0);
            if (m.empty()) throw std::runtime_error(std::string(argv[2]) + ": not found");
            FILE *f = std::fopen(argv[4], "wb");
            if (!f) throw std::runtime_error(std::string("cannot write ") + argv[4]);
            std::fwrite(m.data, 1, m.total() * m.elemSize(), f);
            std::fclose(f);
            std::printf("%d %d %d\n", m.rows, m.cols, m.channels());
            return 0;
        }
        if (mode == "pfm" && argc == 4) {
            cv::Mat m = cvutils::io::read_pfm_file(argv[2]);
            if (m.empty()) throw std::runtime_error(std::string(argv[2]) + ": not a readable PFM file");
            cvutils::io::save_pfm_file(argv[3], m);
            std::printf("%d %d %d\n", m.rows, m.cols, m.channels());
            return 0;
        }
        if (mode == "calib" && argc == 3) {
            Calib c(argv[2]);
            for (int k = 0; k < 2; ++k) {
                std::printf("cam%d", k);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) std::printf(" %.9g", k ? c.cam1[i][j] : c.cam0[i][j]);
                std::printf("\n");
            }
            std::printf("doffs %.9g\nbaseline %.9g\nwidth %d\nheight %d\nndisp %d\nisint %d\nvmin %d\nvmax %d\n"
                        "dyavg %.9g\ndymax %.9g\n",
                        c.doffs, c.baseline, c.width, c.height, c.ndisp, c.isint, c.vmin, c.vmax, c.dyavg, c.dymax);
            return 0;
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 2;
    }
    std::cerr << "usage: " << argv[0] << " png <in.png> <color 0|1> <out.raw> | pfm <in.pfm> <out.pfm> | calib <file>\n";
    return 1;
}
