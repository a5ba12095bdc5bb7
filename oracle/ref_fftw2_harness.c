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

/*
 * ref_fftw2_loop: the coherent slice loop of Kirkland's mulslice
 * (ImageSimulation/multislice/mulslice.cpp:703-739) with its transforms done
 * by the reference's FFTW 2.1.5 (float): per slice wave *= trans[layer]
 * (cfpix::operator*=, cfpix.cpp:296-318), fft (FFTW_BACKWARD, exponent +i,
 * cfpix.cpp:177-195), the separable propagator inside the band limit and
 * zero outside (slicelib.cpp:744-782), ifft (FFTW_FORWARD) and the 1/(nx ny)
 * scale, then the integrated intensity the reference prints (:728-734).
 * mulslice.cpp and cfpix.cpp themselves need the FFTW 3 float library, which
 * the image lacks; this harness is the same loop over the FFTW the reference
 * vendors, for timing the CPU path (oracle/time_ref_loop.py).  Independent
 * waves run on `nthreads` OpenMP threads sharing one plan made with
 * FFTW_THREADSAFE.  Planning: FFTW_MEASURE as cfpix::init
 * (cfpix.cpp:185-197) with measure = 1, else FFTW_ESTIMATE; seconds[0] is the
 * planning time, seconds[1] the loop's.
 */
#include <math.h>
#include <stdlib.h>
#include <omp.h>

__attribute__((visibility("default"))) int ref_fftw2_loop(const float *trans, int nlayer, const int *layer,
                                                          int nslice, int nx, int ny, const float *pxr,
                                                          const float *pxi, const float *pyr, const float *pyi,
                                                          const float *kx2, const float *ky2, float k2max,
                                                          float *waves, int nwave, int nthreads, double *inten,
                                                          int measure, double *seconds) {
    /* FFTW_THREADSAFE: an in-place multi-dimensional plan keeps a work array;
     * this flag makes the plan read-only so the threads can share it */
    const int flags = (measure ? FFTW_MEASURE : FFTW_ESTIMATE) | FFTW_IN_PLACE | FFTW_THREADSAFE;
    double t0 = omp_get_wtime();
    fftwnd_plan pf = fftw2d_create_plan(nx, ny, FFTW_BACKWARD, flags);
    fftwnd_plan pi = fftw2d_create_plan(nx, ny, FFTW_FORWARD, flags);
    if (!pf || !pi) return -1;
    const double t_plan = omp_get_wtime() - t0;
    t0 = omp_get_wtime();
    const float scale = 1.0F / (((float)nx) * ((float)ny));
    int bad = 0;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int b = 0; b < nwave; b++) {
        fftw_complex *w = (fftw_complex *)(waves + 2 * (size_t)b * nx * ny);
        for (int is = 0; is < nslice; is++) {
            const int l = layer[is];
            if (l < 0 || l >= nlayer) { bad = 1; break; }
            const fftw_complex *t = (const fftw_complex *)(trans + 2 * (size_t)l * nx * ny);
            for (size_t j = 0; j < (size_t)nx * ny; j++) {
                const float wr = w[j].re, wi = w[j].im, tr = t[j].re, ti = t[j].im;
                w[j].re = wr * tr - wi * ti;
                w[j].im = wr * ti + wi * tr;
            }
            fftwnd_one(pf, w, 0);
            for (int ix = 0; ix < nx; ix++) {
                fftw_complex *row = w + (size_t)ix * ny;
                const float xr = pxr[(size_t)l * nx + ix], xi = pxi[(size_t)l * nx + ix];
                for (int iy = 0; iy < ny; iy++) {
                    if (kx2[ix] + ky2[iy] < k2max) {
                        const float yr = pyr[(size_t)l * ny + iy], yi = pyi[(size_t)l * ny + iy];
                        const float wr = row[iy].re, wi = row[iy].im;
                        const float tr = wr * yr - wi * yi, ti = wr * yi + wi * yr;
                        row[iy].re = tr * xr - ti * xi;
                        row[iy].im = tr * xi + ti * xr;
                    } else {
                        row[iy].re = 0.0F;
                        row[iy].im = 0.0F;
                    }
                }
            }
            fftwnd_one(pi, w, 0);
            double sum = 0.0;
            for (size_t j = 0; j < (size_t)nx * ny; j++) {
                w[j].re *= scale;
                w[j].im *= scale;
                sum += w[j].re * w[j].re + w[j].im * w[j].im;
            }
            if (inten) inten[(size_t)b * nslice + is] = sum;
        }
    }
    if (seconds) {
        seconds[0] = t_plan;
        seconds[1] = omp_get_wtime() - t0;
    }
    fftwnd_destroy_plan(pf);
    fftwnd_destroy_plan(pi);
    return bad ? -1 : 0;
}
