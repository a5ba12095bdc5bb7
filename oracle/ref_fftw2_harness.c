/*
 * ref_fftw2_harness.c -- C entry point over the FFTW 2.1.5 library vendored by
 * the reference (ImageSimulation/fftw-2.1.5), built in float mode by
 * oracle/build_ref.sh into oracle/_ref/.  TEST INFRASTRUCTURE ONLY: tests/
 * use it to check the oracle's FFT (and the engine's) against the FFT
 * library the reference ships.
 *
 * ref_fftw2_2d(data, nx, ny, sign): in-place unscaled 2D transform of an
 * nx*ny complex float image stored y fastest; sign = +1 is FFTW_BACKWARD
 * (exponent +i), sign = -1 is FFTW_FORWARD (exponent -i).
 */
#include "fftw.h"

__attribute__((visibility("default"))) int ref_fftw2_2d(float *data, int nx, int ny, int sign) {
    fftwnd_plan p = fftw2d_create_plan(nx, ny, sign > 0 ? FFTW_BACKWARD : FFTW_FORWARD,
                                       FFTW_ESTIMATE | FFTW_IN_PLACE);
    if (!p) return -1;
    fftwnd_one(p, (fftw_complex *)data, 0);
    fftwnd_destroy_plan(p);
    return 0;
}
