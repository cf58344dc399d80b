// pct_cli - command-line front end with the reference GPU program's argument
// surface (src/GPU/PatchBasedPseudoCT.cu:1876-1883, 1931-1983):
//
//   pct_cli <data> <result> <EXEC_CHOICE> <NH> [sigma beta] [device] [--seed S]
//
//   data         configuration file (cpp:908-960): random flag, then either
//                num/nx/ny/nz (seeded uniform data) or atlas/label/image
//                MatrixIO paths; or a .mat file (the MATLAB build's input:
//                structs atlas, atlas_label and the image)
//   result       output name; the post-processed pseudo-CT (HU) is written to
//                <result>.mat: a MAT-file when the data was one (the MATLAB
//                build), else the MatrixIO text format, as the reference
//                CPU program's non-MATLAB build does (cpp:1741-1748)
//   EXEC_CHOICE  0 GM, 1 GM2 (direct kernel), 2 SM (tiled kernel),
//                4 GpuTestTime (timing sweep, cu:2028-2148)
//   NH           search window edge (7, 9, 11, ...)
//   sigma beta   read only when both are given (cu:1979-1982); default 1 1
//   device       HIP device ordinal (cu:1973-1975), default 0
//   --seed S     seed of the random-data mode (the reference seeds rand()
//                from the clock, cpp:918); default 0
//
// GpuTestTime options (the reference fixes them in code, cu:2033-2043):
//   --sweep-nh A,B,...  window edges to sweep (default: NH)
//   --many K            K samples of K calls per point (default 3; the
//                       reference uses its `many`)
//   --results FILE      append the averaged table in the reference's final
//                       results format (cu:2105-2116)
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pct/pct.h"

extern "C" int pct_set_device(int device);

static void usage() {
    std::printf(
        "Usage: pct_cli Data EXEC_CHOICE-result-name EXEC_CHOICE NH [sigma] [beta] [device] "
        "[--seed S]\n"
        "[1] Data: configuration file (random flag; num nx ny nz | atlas label image paths)\n"
        "[2] Name for the result file (<name>.mat, MatrixIO text)\n"
        "[3] EXEC_CHOICE {GM = 0, GM2 = 1, SM = 2, GpuTestTime = 4}\n"
        "[4] NH {7,9,11}\n"
        "[5] sigma 1 default (2.5)\n"
        "[6] beta 1 default\n"
        "[7] HIP device, default 0\n");
}

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// The reference's effective-bandwidth byte count for one call (cu:2193-2198).
static double naive_gbs(int nx, int ny, int nz, int num, int nh, double ms) {
    const double hw = (nh - 1) / 2;
    const double vox = (nx - hw) * (ny - hw) * (nz - hw);
    const double rB = 4.0 * vox * ((double)num * nh * nh * nh + 2.0 * num * 27 * nh * nh * nh);
    const double wB = vox * 4.0;
    return (rB + wB) / (ms * 1e-3) / 1e9;
}

int main(int argc, char** argv) {
    std::vector<char*> args;
    unsigned long long seed = 0;
    std::vector<int> sweep_nh;
    int many = 3;
    std::string results_path;
    for (int i = 0; i < argc; ++i) {
        if (std::strcmp(argv[i], "--seed") == 0 && i + 1 < argc) {
            seed = std::strtoull(argv[++i], nullptr, 10);
            continue;
        }
        if (std::strcmp(argv[i], "--sweep-nh") == 0 && i + 1 < argc) {
            for (char* t = std::strtok(argv[++i], ","); t; t = std::strtok(nullptr, ","))
                sweep_nh.push_back(std::atoi(t));
            continue;
        }
        if (std::strcmp(argv[i], "--many") == 0 && i + 1 < argc) {
            many = std::max(1, std::atoi(argv[++i]));
            continue;
        }
        if (std::strcmp(argv[i], "--results") == 0 && i + 1 < argc) {
            results_path = argv[++i];
            continue;
        }
        args.push_back(argv[i]);
    }
    const int ac = (int)args.size();
    if (ac < 5) {
        usage();
        return EXIT_SUCCESS;  // the reference prints usage and succeeds (cu:2213-2224)
    }
    double sigma = 1, beta = 1;
    int device = 0;
    if (ac >= 8) device = std::atoi(args[7]);
    if (ac >= 7) {
        sigma = std::atof(args[5]);
        beta = std::atof(args[6]);
    }
    const std::string result_name = args[2];
    const std::string data_path = args[1];
    const bool mat_input =
        data_path.size() > 4 && data_path.compare(data_path.size() - 4, 4, ".mat") == 0;
    // a MATLAB variable name: the result name's last path component
    std::string result_var = result_name.substr(result_name.find_last_of('/') + 1);
    if (result_var.empty()) result_var = "result";
    const int choice = std::atoi(args[3]);
    const int nh = std::atoi(args[4]);

    pct_volume_set vs;
    int rc = pct_read_config(args[1], seed, &vs);
    if (rc) {
        std::fprintf(stderr, "cannot read %s: %s\n", args[1], pct_strerror(rc));
        return EXIT_FAILURE;
    }
    std::printf("nx=%d ny=%d nz=%d num=%d\n", vs.nx, vs.ny, vs.nz, vs.n_atlas);
    rc = pct_set_device(device);
    if (rc) {
        std::fprintf(stderr, "device %d: %s\n", device, pct_strerror(rc));
        pct_volume_set_free(&vs);
        return EXIT_FAILURE;
    }
    std::printf("[3] Choice %d\n[4] NH %d\n[5] sigma %g\n[6] beta %g\n[7] Device %d\n", choice,
                nh, sigma, beta, device);
    const size_t N = (size_t)vs.nx * vs.ny * vs.nz;
    pct_normalize_labels(vs.label, N * vs.n_atlas);   // cu:2009-2020

    pct_problem pb;
    pb.nx = vs.nx;
    pb.ny = vs.ny;
    pb.nz = vs.nz;
    pb.n_atlas = vs.n_atlas;
    pb.nh = nh;
    pb.patch = 3;
    pb.sigma = sigma;
    pb.beta = beta;
    std::vector<float> out(N, 0.0f);
    int status = EXIT_SUCCESS;

    if (choice == PCT_CHOICE_GM || choice == PCT_CHOICE_GM2 || choice == PCT_CHOICE_SM ||
        choice == PCT_CHOICE_DIRECT) {
        // timed region as the reference's cu:2152-2188: fusion incl. copies,
        // plus the post-processing (on the device, bit-identical to the host's)
        const double t0 = now_ms();
        rc = pct_pseudo_ct_host(&pb, vs.img, vs.atlas, vs.label, out.data(), choice);
        const double t1 = now_ms();
        if (rc) {
            std::fprintf(stderr, "fusion failed: %s\n", pct_strerror(rc));
            status = EXIT_FAILURE;
        } else {
            const double ms = t1 - t0;
            std::printf("Effective BandWidth = %.4f GB/s, Elapsed Time %f ms %f seconds\n",
                        naive_gbs(vs.nx, vs.ny, vs.nz, vs.n_atlas, nh, ms), ms, ms / 1000.0);
            const std::string path = result_name + ".mat";
            // the MATLAB build (data from a .mat file) writes a MAT-file with
            // the result name as variable (cpp:1743-1745); the text build
            // writes MatrixIO text under the same name (cpp:1747)
            if (mat_input)
                rc = pct_mat_write(path.c_str(), result_var.c_str(), out.data(), vs.nx, vs.ny,
                                   vs.nz);
            else
                rc = pct_matrixio_write(path.c_str(), out.data(), 1, vs.nx, vs.ny, vs.nz);
            if (rc) {
                std::fprintf(stderr, "cannot write %s: %s\n", path.c_str(), pct_strerror(rc));
                status = EXIT_FAILURE;
            } else {
                std::printf("Saving %s\n", path.c_str());
            }
        }
    } else if (choice == PCT_CHOICE_GPU_TEST_TIME) {
        // cu:2028-2148: for every window edge and every atlas count from num
        // down to 1, `many` samples of `many` calls of the host wrapper (SM);
        // each sample's mean call time is printed, and the per-point means
        // form the final table
        if (sweep_nh.empty()) sweep_nh.push_back(nh);
        std::vector<std::vector<double>> table;   // rows: nh, atlas, mean ms
        for (int snh : sweep_nh) {
            pb.nh = snh;
            for (int at = vs.n_atlas; at >= 1 && !rc; --at) {
                pb.n_atlas = at;
                double sum = 0.0;
                for (int i = 0; i < many && !rc; ++i) {
                    const double t0 = now_ms();
                    for (int j = 0; j < many && !rc; ++j)
                        rc = pct_fuse_host(&pb, vs.img, vs.atlas, vs.label, out.data(),
                                           PCT_CHOICE_SM);
                    const double ms = (now_ms() - t0) / many;
                    if (rc) break;
                    sum += ms;
                    std::printf("Iter %d Atlas %d Choice 2 Nh %d Elapsed Time %f ms %f minutes\n",
                                i, at, snh, ms, ms / 60000.0);
                }
                if (rc) break;
                const double mean = sum / many;
                std::printf("Effective BandWidth = %.4f GB/s, Elapsed Time %f ms %f seconds\n",
                            naive_gbs(vs.nx, vs.ny, vs.nz, at, snh, mean), mean, mean / 1000.0);
                table.push_back({(double)snh, (double)at, mean});
            }
        }
        if (rc) {
            std::fprintf(stderr, "fusion failed: %s\n", pct_strerror(rc));
            status = EXIT_FAILURE;
        } else if (!results_path.empty()) {
            FILE* f = std::fopen(results_path.c_str(), "a");
            if (!f) {
                std::fprintf(stderr, "cannot write %s\n", results_path.c_str());
                status = EXIT_FAILURE;
            } else {
                std::fprintf(f, "Num of executions to calculate a result %d sampled %d times\n",
                             many, many);
                for (const auto& r : table)
                    std::fprintf(f, "Nh\t%d Atlas\t%d Elapsed Time\t%g\t%g\t%g\n", (int)r[0],
                                 (int)r[1], r[2], r[2] / 1000.0, r[2] / 60000.0);
                std::fclose(f);
            }
        }
    } else {
        std::fprintf(stderr, "EXEC_CHOICE %d is not handled\n", choice);
        status = EXIT_FAILURE;
    }
    pct_volume_set_free(&vs);
    pct_release();
    return status;
}
