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

// The reference's call sequence on the Microscope and PotentialBuilder
// surfaces (Microscope.h:3-34, PotentialBuilder.h:14-22), written to those
// headers' signatures; prints Es(k) and alpha(k) for tests/test_lms_surface.py.
// Needs no GPU.
static int surface() {
    Microscope set;  // Microscope(void), then the setters
    set.setKeV(300.0);
    set.setCs(1.2);
    set.setAperture(20.0);
    set.setDefocus(-60.0);
    // keV, Cs, aperture, defocus as the reference takes them (no unit conversion)
    Microscope mics[] = {Microscope(200.0, 0.5, 10.0, 35.0), Microscope(200.0, 0.0, 10.0, 35.0),
                         Microscope(200.0, 0.5, 10.0, 0.0), set};
    const double ks[] = {0.0, 0.1, 0.5, 1.0, 2.5};
    for (int m = 0; m < 4; m++) {
        Microscope& mic = mics[m];
        std::printf("mic %d keV=%.17g cs=%.17g aperture=%.17g defocus=%.17g lambda=%.17g sigma=%.17g\n", m,
                    mic.getKeV(), mic.getCs(), mic.getAperture(), mic.getDefocus(), mic.getLambda(), mic.getSigma());
        for (double k : ks) std::printf("val %d k=%.17g Es=%.17g alpha=%.17g\n", m, k, mic.Es(k), mic.alpha(k));
    }
    CPotentialBuilder builder;  // CPotentialBuilder(void)
    (void)builder;
    const int r = fnPotentialBuilder();
    nPotentialBuilder += 1;
    ModelPotential mp0;  // ModelPotential(void): no model
    ModelSimulated ms0;  // ModelSimulated(void): no potential
    msx::lms::Image img(4, 4, 1, 8, 2);
    std::printf("fnPotentialBuilder=%d nPotentialBuilder=%d empty_potential=%d empty_image=%d\n", r,
                nPotentialBuilder, mp0.calculatePotentialGrid(), ms0.imageCalculation(&img, &mics[0]));
    return 0;
}

// Writes <prefix>_m<mode>.mrc for every MRC mode (MRC.cpp:206-345) from
// images of the mode's channel count (1, 2 or 3) whose values cover
// fractions, negatives, values beyond each integer type and large floats;
// tests/test_lms_surface.py reads them back.  Needs no GPU.
static int mrc_modes(const std::string& out) {
    msx::lms::amodel::Model model(7.5f, 6.25f, 3.5f);
    model.setAlpha(90.f);
    model.setBeta(95.5f);
    model.setGamma(120.f);
    const int modes[] = {0, 1, 2, 3, 4, 6, 16};
    for (int mode : modes) {
        const int ch = msx::lms::mrc_channels(mode);
        msx::lms::Image img(5, 3, 2, 8, ch);
        double* d = img.data();
        for (int i = 0; i < 5 * 3 * 2 * ch; i++) d[i] = (i % 7 - 3) * 37.25 + (i % 5) * 1000.5 * ((i % 3) - 1) + 0.75;
        const std::string p = out + "_m" + std::to_string(mode);
        if (img.saveMRC(p.c_str(), &model, 5, 3, 2, mode) != 0) return 1;
    }
    // an unknown mode and an image too small for its mode are refused
    msx::lms::Image one(5, 3, 2, 8, 1);
    if (one.saveMRC((out + "_bad").c_str(), &model, 5, 3, 2, 5) != -1) return 1;
    if (one.saveMRC((out + "_bad").c_str(), &model, 5, 3, 2, 4) != -1) return 1;
    std::printf("mrc modes written\n");
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: lms_check <output prefix> | --surface\n");
        return 2;
    }
    if (std::string(argv[1]) == "--surface") return surface();
    if (std::string(argv[1]) == "--mrc" && argc > 2) return mrc_modes(argv[2]);
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
    (void)pb;
    std::printf("lms_check ok atoms=%zu potential_sum=%.6e builder=%d\n", model.getNumberAtoms(),
                [&] { double s = 0; for (size_t i = 0; i < nx * ny * nz; i++) s += mp.potential[i]; return s; }(),
                fnPotentialBuilder());
    return 0;
}
