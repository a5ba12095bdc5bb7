// Drives the C++ LMultiSlice mirror (include/msx/lmultislice.hpp) the way the
// reference's Dispatcher::Run does (Dispatcher.cu:211-223): build a model,
// ModelPotential(model, nx, ny, nz, dpa, radius).calculatePotentialGrid(),
// savePotential(out), ModelSimulated(...).imageCalculation(result, microscope).
// Writes <out>_pt.mrc, <out>_potential.bin (double [nz][ny][nx]) and
// <out>_image.bin (interleaved complex double [ny][nx]) for tests/test_gpu_cpp.py.
#define MSX_LMS_GLOBAL_NAMES
#include <cstdio>
#include <string>

#include "msx/lmultislice.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: lms_check <output prefix>\n");
        return 2;
    }
    const std::string out = argv[1];
    // Si [001], 2x2x1 conventional cells, fractional coordinates, planes at slice centres
    const float a = 5.431f;
    msx::lms::amodel::Model model(2 * a, 2 * a, a);
    const float base[8][3] = {{0, 0, 0}, {0, .5f, .5f}, {.5f, 0, .5f}, {.5f, .5f, 0},
                              {.25f, .25f, .25f}, {.25f, .75f, .75f}, {.75f, .25f, .75f}, {.75f, .75f, .25f}};
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 8; k++)
                model.add("Si", (base[k][0] + i) / 2, (base[k][1] + j) / 2, base[k][2] + 0.125f);
    const size_t nx = 128, ny = 128, nz = 4;
    ModelPotential mp(&model, nx, ny, nz, nx / (2 * a), 4.0);
    mp.set_formula(MSX_FORMULA_KIRKLAND);
    if (mp.calculatePotentialGrid() != 0) {
        std::fprintf(stderr, "calculatePotentialGrid failed: %s\n", msx_last_error());
        return 1;
    }
    if (mp.savePotential(out.c_str()) != 0) return 1;
    FILE* f = std::fopen((out + "_potential.bin").c_str(), "wb");
    std::fwrite(mp.potential, sizeof(double), nx * ny * nz, f);
    std::fclose(f);
    Microscope mic(200.0, 1.2, 20.0, -60.0);  // keV, Cs mm, aperture mrad, defocus nm
    ModelSimulated ms(&mp, nx, ny, nz, nx / (2 * a));
    msx::lms::Image result(nx, ny, 1, 8, 2);
    if (ms.imageCalculation(&result, &mic) != 0) {
        std::fprintf(stderr, "imageCalculation failed: %s\n", msx_last_error());
        return 1;
    }
    f = std::fopen((out + "_image.bin").c_str(), "wb");
    std::fwrite(result.data(), sizeof(double), 2 * nx * ny, f);
    std::fclose(f);
    CPotentialBuilder pb;
    std::printf("lms_check ok atoms=%zu potential_sum=%.6e builder=%d\n", model.getNumberAtoms(),
                [&] { double s = 0; for (size_t i = 0; i < nx * ny * nz; i++) s += mp.potential[i]; return s; }(),
                pb.fnPotentialBuilder());
    return 0;
}
