// pct_cli - command-line front end with the reference GPU program's argument
// surface (src/GPU/PatchBasedPseudoCT.cu:1876-1883, 1931-1983):
//
//   pct_cli <data> <result> <EXEC_CHOICE> <NH> [sigma beta] [device] [--seed S]
//
//   data         configuration file (cpp:908-960): random flag, then either
//                num/nx/ny/nz (seeded uniform data) or atlas/label/image
//                MatrixIO paths
//   result       output name; the post-processed pseudo-CT (HU) is written to
//                <result>.mat in the MatrixIO text format, as the reference
//                CPU program's non-MATLAB build does (cpp:1741-1748)
//   EXEC_CHOICE  0 GM, 1 GM2 (direct kernel), 2 SM (tiled kernel),
//                4 GpuTestTime (timing sweep, cu:2028-2148)
//   NH           search window edge (7, 9, 11, ...)
//   sigma beta   read only when both are given (cu:1979-1982); default 1 1
//   device       HIP device ordinal (cu:1973-1975), default 0
//   --seed S     seed of the random-data mode (the reference seeds rand()
//                from the clock, cpp:918); default 0
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

int main(int argc, char** argv) {
    std::vector<char*> args;
    unsigned long long seed = 0;
    for (int i = 0; i < argc; ++i) {
        if (std::strcmp(argv[i], "--seed") == 0 && i + 1 < argc) {
            seed = std::strtoull(argv[++i], nullptr, 10);
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

    if (choice == PCT_CHOICE_GM || choice == PCT_CHOICE_GM2 || choice == PCT_CHOICE_SM) {
        // timed region as the reference's cu:2152-2188: fusion incl. copies,
        // plus the host post-processing
        const double t0 = now_ms();
        rc = pct_fuse_host(&pb, vs.img, vs.atlas, vs.label, out.data(), choice);
        if (!rc) rc = pct_postprocess(nh, vs.nx, vs.ny, vs.nz, out.data());
        const double t1 = now_ms();
        if (rc) {
            std::fprintf(stderr, "fusion failed: %s\n", pct_strerror(rc));
            status = EXIT_FAILURE;
        } else {
            const double ms = t1 - t0;
            // the reference's naive byte count (cu:2193-2198)
            const double hw = (nh - 1) / 2;
            const double vox = (vs.nx - hw) * (vs.ny - hw) * (vs.nz - hw);
            const double rB = 4.0 * vox * ((double)vs.n_atlas * nh * nh * nh +
                                           2.0 * vs.n_atlas * 27 * nh * nh * nh);
            const double wB = vox * 4.0;
            std::printf("Effective BandWidth = %.4f GB/s, Elapsed Time %f ms %f seconds\n",
                        (rB + wB) / (ms * 1e-3) / 1e9, ms, ms / 1000.0);
            const std::string path = result_name + ".mat";
            rc = pct_matrixio_write(path.c_str(), out.data(), 1, vs.nx, vs.ny, vs.nz);
            if (rc) {
                std::fprintf(stderr, "cannot write %s: %s\n", path.c_str(), pct_strerror(rc));
                status = EXIT_FAILURE;
            } else {
                std::printf("Saving %s\n", path.c_str());
            }
        }
    } else if (choice == PCT_CHOICE_GPU_TEST_TIME) {
        // cu:2028-2148 sweeps atlas counts for one window with 10 x 10 calls;
        // here: every atlas count from num down to 1 for the given NH, 3
        // samples of 3 calls of the full host wrapper (SM kernel).
        const int many = 3;
        for (int at = vs.n_atlas; at >= 1; --at) {
            pb.n_atlas = at;
            for (int i = 0; i < many && !rc; ++i) {
                const double t0 = now_ms();
                for (int j = 0; j < many && !rc; ++j)
                    rc = pct_fuse_host(&pb, vs.img, vs.atlas, vs.label, out.data(), PCT_CHOICE_SM);
                const double ms = (now_ms() - t0) / many;
                if (!rc)
                    std::printf("Iter %d Atlas %d Choice 2 Nh %d Elapsed Time %f ms %f minutes\n",
                                i, at, nh, ms, ms / 60000.0);
            }
            if (rc) {
                std::fprintf(stderr, "fusion failed: %s\n", pct_strerror(rc));
                status = EXIT_FAILURE;
                break;
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
