// pct_cli - command-line front end with the reference programs' argument
// surfaces.
//
// GPU program surface (src/GPU/PatchBasedPseudoCT.cu:1876-1883, 1931-1983):
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
//   EXEC_CHOICE  0 GM, 1 GM2, 2 SM (all the tiled kernel), 4 GpuTestTime
//                (timing sweep, cu:2028-2148), 5 the direct kernel
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
//
// CPU program surface (src/CPU/NewPatchedBrainPureC.cpp:1484-1491,
// 1517-1565), selected with --cpu-surface:
//
//   pct_cli --cpu-surface <data> <result> <EXEC_CHOICE> <NH> [TH] [sigma beta] [--seed S]
//
//   EXEC_CHOICE  0 SingleCore, 1 MultiCore: one fusion, timed (cpp:1606-1655);
//                2 SingleCoreTestTime, 3 MultiCoreTestTime: 10 samples of 10
//                fusions, each sample's mean time printed (cpp:1660-1711).
//                The fusion runs on the GPU (device 0) in every case.
//   TH           threads of the host parts (OpenMP), default 1 (cpp:1552-1554).
//                The reference's MultiCoreTestTime takes its thread count from
//                the EXEC_CHOICE argument (cpp:1690); here TH counts.
//   sigma beta   read only when TH is given too (cpp:1555-1558); default 1 1.
//   In random mode the generated inputs are written to atlas.txt, label.txt
//   and im.txt in the working directory, as cpp:974-978 does.
//   The result is written as on the GPU surface, after the NaN fill, the
//   rounding and the HU rescale (cpp:1721-1748).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

#include "pct/pct.h"

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

static void usage_cpu() {
    std::printf("Usage: pct_cli --cpu-surface Data EXEC_CHOICE-result-name EXEC_CHOICE NH [TH] "
                "[sigma] [beta] [--seed S]\n"
                "[3] EXEC_CHOICE {SingleCore = 0, MultiCore = 1, SingleCoreTestTime = 2, "
                "MultiCoreTestTime = 3}\n"
                "[4] NH {7,9,11}\n"
                "[5] TH host threads, default 1\n"
                "[6] sigma 1 default (2.5), read with TH and beta\n"
                "[7] beta 1 default\n");
}

static const char* cpu_choice_name(int c) {
    switch (c) {
        case 0: return "SingleCore";
        case 1: return "MultiCore";
        case 2: return "SingleCoreTestTime";
        case 3: return "MultiCoreTestTime";
        default: return "Unknown";
    }
}

// First line of a text configuration file: the random-data flag (cpp:924-927).
static bool config_is_random(const std::string& path) {
    if (path.size() > 4 && path.compare(path.size() - 4, 4, ".mat") == 0) return false;
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    char line[64] = {0};
    const bool ok = std::fgets(line, sizeof(line), f) != nullptr;
    std::fclose(f);
    return ok && std::atoi(line) != 0;
}

// <result>.mat: a MAT-file when the data was one, else MatrixIO text.
static int write_result(const std::string& result_name, bool mat_input, const float* data,
                        int nx, int ny, int nz) {
    std::string var = result_name.substr(result_name.find_last_of('/') + 1);
    if (var.empty()) var = "result";
    const std::string path = result_name + ".mat";
    const int rc = mat_input ? pct_mat_write(path.c_str(), var.c_str(), data, nx, ny, nz)
                             : pct_matrixio_write(path.c_str(), data, 1, nx, ny, nz);
    if (rc)
        std::fprintf(stderr, "cannot write %s: %s\n", path.c_str(), pct_strerror(rc));
    else
        std::printf("Saving %s\n", path.c_str());
    return rc;
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

// The CPU program's main (cpp:1517-1748) with the fusion on the GPU.
static int cpu_surface_main(const std::vector<char*>& args, unsigned long long seed) {
    const int ac = (int)args.size();
    if (ac < 5) {
        usage_cpu();
        return EXIT_SUCCESS;   // the reference prints its usage and succeeds (cpp:1751-1754)
    }
    int threads = 1;
    double sigma = 1, beta = 1;
    if (ac >= 6) threads = std::atoi(args[5]);
    if (ac >= 8) {
        sigma = std::atof(args[6]);
        beta = std::atof(args[7]);
    }
    const std::string data_path = args[1], result_name = args[2];
    const int choice = std::atoi(args[3]);
    const int nh = std::atoi(args[4]);
    const bool mat_input =
        data_path.size() > 4 && data_path.compare(data_path.size() - 4, 4, ".mat") == 0;
    std::printf("[2] Anatomy Atlas %s\n[3] Result %s\n[3] Choice %s\n[4] NH %d\n[5] TH %d\n"
                "[6] sigma %g\n[7] beta %g\n",
                data_path.c_str(), result_name.c_str(), cpu_choice_name(choice), nh, threads,
                sigma, beta);
    std::fflush(stdout);
    if (choice < 0 || choice > 3) {
        std::fprintf(stderr, "EXEC_CHOICE %d is not handled\n", choice);
        return EXIT_FAILURE;
    }
    if (threads >= 1) omp_set_num_threads(threads);   // host parts (cpp:1626)
    pct_volume_set vs;
    int rc = pct_read_config(data_path.c_str(), seed, &vs);
    if (rc) {
        std::fprintf(stderr, "cannot read %s: %s\n", data_path.c_str(), pct_strerror(rc));
        return EXIT_FAILURE;
    }
    std::printf("nx=%d ny=%d nz=%d num=%d\n", vs.nx, vs.ny, vs.nz, vs.n_atlas);
    const size_t N = (size_t)vs.nx * vs.ny * vs.nz;
    int status = EXIT_SUCCESS;
    if (config_is_random(data_path)) {
        // cpp:974-978: the generated inputs, before the CT normalisation
        if (pct_matrixio_write("atlas.txt", vs.atlas, vs.n_atlas, vs.nx, vs.ny, vs.nz) ||
            pct_matrixio_write("label.txt", vs.label, vs.n_atlas, vs.nx, vs.ny, vs.nz) ||
            pct_matrixio_write("im.txt", vs.img, 1, vs.nx, vs.ny, vs.nz)) {
            std::fprintf(stderr, "cannot write the random inputs\n");
            pct_volume_set_free(&vs);
            return EXIT_FAILURE;
        }
    }
    std::printf("normalizeFlag=1 cleanNaNFlag=1\n");
    pct_normalize_labels(vs.label, N * vs.n_atlas);   // cpp:1583-1593
    float lo = vs.label[0], hi = vs.label[0];
    for (size_t i = 1; i < N * vs.n_atlas; ++i) {
        lo = std::min(lo, vs.label[i]);
        hi = std::max(hi, vs.label[i]);
    }
    std::printf("maxLabels %g minLabels %g\n", hi, lo);
    rc = pct_set_device(0);
    if (rc) {
        std::fprintf(stderr, "device 0: %s\n", pct_strerror(rc));
        pct_volume_set_free(&vs);
        return EXIT_FAILURE;
    }
    pct_problem pb;
    pb.nx = vs.nx;
    pb.ny = vs.ny;
    pb.nz = vs.nz;
    pb.n_atlas = vs.n_atlas;
    pb.nh = nh;
    pb.patch = 3;
    pb.sigma = sigma;
    pb.beta = beta;
    pb.div = 0.0;   // the library forms 2*27*beta*sigma^2 as cpp:1601 does
    std::vector<float> out(N, 0.0f);
    const double div = pct_weight_divisor(sigma, beta);
    if (choice <= 1) {
        std::printf("%s div=%g\n",
                    choice == 0 ? "MultiPatchBasedSegmentation" : "MultiPatchBasedSegmentationOmp",
                    div);
        const double t0 = now_ms();
        rc = pct_fuse_host(&pb, vs.img, vs.atlas, vs.label, out.data(), PCT_CHOICE_SM);
        const double ms = now_ms() - t0;
        if (!rc)
            std::printf("Atlas %d Nh %d Elapsed Time %g ms  minutes %g\n", vs.n_atlas, nh, ms,
                        ms / 60000.0);
    } else {
        // cpp:1660-1711: 10 samples of 10 calls, each sample's mean printed
        std::printf("%s div=%g\n", cpu_choice_name(choice), div);
        const int many = 10;
        for (int i = 0; i < many && !rc; ++i) {
            const double t0 = now_ms();
            for (int j = 0; j < many && !rc; ++j)
                rc = pct_fuse_host(&pb, vs.img, vs.atlas, vs.label, out.data(), PCT_CHOICE_SM);
            const double ms = (now_ms() - t0) / many;
            if (!rc) std::printf("Iter %d Elapsed Time %g ms  minutes %g\n", i, ms, ms / 60000.0);
        }
    }
    if (rc) {
        std::fprintf(stderr, "fusion failed: %s\n", pct_strerror(rc));
        status = EXIT_FAILURE;
    } else {
        // cpp:1721-1739: NaN fill, rounding, HU rescale, then the result file
        std::printf("Regularization step\n");
        rc = pct_postprocess(nh, vs.nx, vs.ny, vs.nz, out.data());
        if (rc < 0) {
            std::fprintf(stderr, "post-processing failed: %s\n", pct_strerror(rc));
            status = EXIT_FAILURE;
        } else if (write_result(result_name, mat_input, out.data(), vs.nx, vs.ny, vs.nz)) {
            status = EXIT_FAILURE;
        }
    }
    pct_volume_set_free(&vs);
    pct_release();
    return status;
}

int main(int argc, char** argv) {
    std::vector<char*> args;
    unsigned long long seed = 0;
    std::vector<int> sweep_nh;
    int many = 3;
    std::string results_path;
    bool cpu_surface = false;
    for (int i = 0; i < argc; ++i) {
        if (std::strcmp(argv[i], "--cpu-surface") == 0) {
            cpu_surface = true;
            continue;
        }
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
    if (cpu_surface) return cpu_surface_main(args, seed);
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
    pb.div = 0.0;
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
            // the MATLAB build (data from a .mat file) writes a MAT-file with
            // the result name as variable (cpp:1743-1745); the text build
            // writes MatrixIO text under the same name (cpp:1747)
            if (write_result(result_name, mat_input, out.data(), vs.nx, vs.ny, vs.nz))
                status = EXIT_FAILURE;
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
