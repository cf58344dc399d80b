// Test-infrastructure harness around the reference CPU program.
//
// This translation unit compiles the reference's own source file in place
// (path given by -DPCT_REF_CPP, see oracle/Makefile) and exposes its hot-path
// functions through extern "C" wrappers so that the test-suite and the CPU
// baseline leg of bench.py can call them.  Nothing here is part of the
// product: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load the resulting oracle/_ref/libpctref.so.
//
// The reference's main() is renamed by -Dmain and is never called; the
// reference file's non-MATLAB build has three misspelled MatrixIO calls
// (src/CPU/NewPatchedBrainPureC.cpp:976-978, 987-989, 1747) that the
// Makefile maps onto the existing methods with -Dstore=Store -Dread_=Read,
// and an undeclared FindDimension call inside main() (cpp:1519) that is
// neutralised by a macro.  main() is then discarded by --gc-sections.
#define main pctref_unused_main
#include PCT_REF_CPP
#undef main

#define PCTREF_API extern "C" __attribute__((visibility("default")))

// MultiAtlasPatchBasedSegmentation[OpenMP] read the globals nx, ny, nz, num,
// nh and the target through the global `im` (cpp:1316, cpp:1428).
static void set_globals(int nx_, int ny_, int nz_, int num_, int nh_,
                        const float* img, const float* atl, const float* lab) {
    nx = nx_; ny = ny_; nz = nz_; num = num_; nh = nh_;
    im = const_cast<float*>(img);
    atlas = const_cast<float*>(atl);
    label = const_cast<float*>(lab);
}

// cpp:1228-1337 (multicore == 0) or cpp:1350-1449 (multicore != 0).
// `out` must be pre-zeroed by the caller, as main() does (cpp:1606, 1635).
PCTREF_API void pctref_fuse(int nx_, int ny_, int nz_, int num_, int nh_,
                            const float* img, const float* atl, const float* lab,
                            float* out, double div, int threads, int multicore) {
    set_globals(nx_, ny_, nz_, num_, nh_, img, atl, lab);
    if (multicore) {
        omp_set_num_threads(threads);
        MultiAtlasPatchBasedSegmentationOpenMP(img, atl, lab, out, div);
    } else {
        MultiAtlasPatchBasedSegmentation(img, atl, lab, out, div);
    }
}

// cpp:1184-1212 for one candidate: target patch centred at (z,x,y) of `img`,
// atlas patch centred at (w,u,v) of atlas number `at`.
PCTREF_API double pctref_patch_similarity(int nx_, int ny_, int nz_,
                                          const float* img, const float* atl, int at,
                                          int z, int x, int y, int w, int u, int v,
                                          double div) {
    matrixMask mi, ma;
    mi.c = y - 1; mi.cC = y + 1; mi.r = x - 1; mi.rR = x + 1; mi.z = z - 1; mi.zZ = z + 1;
    mi.cols = ny_; mi.rows = nx_; mi.mats = nz_;
    ma.c = v - 1; ma.cC = v + 1; ma.r = u - 1; ma.rR = u + 1; ma.z = w - 1; ma.zZ = w + 1;
    ma.cols = ny_; ma.rows = nx_; ma.mats = nz_;
    return PatchSimilarity(mi, ma, img, atl, nx_ * ny_ * nz_ * at, div);
}

// cpp:1020-1043
PCTREF_API double pctref_median(const float* patch, int length) {
    return Median(patch, length);
}

// cpp:1059-1106
PCTREF_API void pctref_regularize(int half_window, int half_patch, int nx_, int ny_, int nz_,
                                  float* arr) {
    RegularizationStep(half_window, half_patch, nx_, ny_, nz_, arr);
}

// cpp:1221-1227
PCTREF_API void pctref_round(float* arr, int n) {
    RoundMatrix(arr, n);
}

// cpp:252-320 (MatrixIO<float>::Read with dimensions, allocates with malloc)
PCTREF_API int pctref_matrixio_read(const char* path, float** data, int* dims4) {
    MatrixIO<float> io;
    std::vector<int> d;
    float* v = nullptr;
    io.Read(std::string(path), v, d);
    if (d.size() != 4) return -1;
    for (int i = 0; i < 4; ++i) dims4[i] = d[i];
    *data = v;
    return 0;
}

// cpp:289-320 (MatrixIO<float>::Store)
PCTREF_API void pctref_matrixio_store(const char* path, const float* v,
                                      int num_, int rows, int cols, int depth) {
    MatrixIO<float> io;
    io.Store(std::string(path), v, num_, rows, cols, depth);
}

// cpp:908-1018: text configuration (non-random mode only: the random mode
// seeds rand() from the wall clock and writes files into the working
// directory).  Fills dims4 = {nx, ny, nz, num} and hands out the buffers the
// reference allocated.
PCTREF_API int pctref_read_config(const char* path, int* dims4,
                                  float** img, float** atl, float** lab) {
    nx = ny = nz = num = -1;
    im = atlas = label = nullptr;
    readMatFile(path);
    dims4[0] = nx; dims4[1] = ny; dims4[2] = nz; dims4[3] = num;
    *img = im; *atl = atlas; *lab = label;
    return (im && atlas && label && nx > 0) ? 0 : -1;
}

PCTREF_API void pctref_free(void* p) { free(p); }
