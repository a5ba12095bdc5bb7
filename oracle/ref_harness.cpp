/*
 * ref_harness.cpp -- C entry points over the reference's own CPU code, built
 * by oracle/build_ref.sh into oracle/_ref/ (never committed, never shipped).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle
 * (oracle/msx_oracle.cpp) against the compiled reference.
 *
 * Reference units compiled from their sources under the reference tree,
 * unmodified:
 *   ImageSimulation/multislice/slicelib.cpp  (Kirkland slicelib: wavelength,
 *       sigma, bessi0/bessk0, vzatom, featom, freqn, chi, parlay, ranflat,
 *       rangauss)
 *   ImageSimulation/multislice/floatTIFF.cpp (Kirkland's float TIFF files)
 *   ImageSimulation/fftw-2.1.5/fftw/*.c     (the FFTW 2.1.5 library the
 *       reference vendors, built in float mode)
 * This file only forwards C-callable symbols to them.
 */
#include <cstdio>

#include "slicelib.hpp"
#include "floatTIFF.hpp"

#define REF_API extern "C" __attribute__((visibility("default")))

REF_API double ref_wavelength(double kev) { return wavelength(kev); }
REF_API double ref_sigma(double kev) { return sigma(kev); }
REF_API double ref_bessi0(double x) { return bessi0(x); }
REF_API double ref_bessk0(double x) { return bessk0(x); }
REF_API double ref_vzatom(int Z, double r) { return vzatom(Z, r); }
REF_API double ref_featom(int Z, double k2) { return featom(Z, k2); }
REF_API void ref_freqn(float *ko, float *ko2, float *xo, int nk, double ak) { freqn(ko, ko2, xo, nk, ak); }
/* chi with only Cs and defocus set (multiMode 0): slicelib.cpp:258-318 */
REF_API double ref_chi_csdf(double cs, double df, double alx, double aly) {
    float p[64];  /* covers every slicelib.hpp parameter index (max 56) */
    for (int i = 0; i < 64; i++) p[i] = 0.0f;
    p[pCS] = (float)cs;
    p[pDEFOCUS] = (float)df;
    return chi(p, alx, aly, 0);
}
/* chi with Cs, defocus and the multipole two- and three-fold astigmatism
 * coefficients (multiMode 1): slicelib.cpp:258-318 */
REF_API double ref_chi_multi(double cs, double df, double c12a, double c12b, double c23a, double c23b, double alx,
                             double aly) {
    float p[64];
    for (int i = 0; i < 64; i++) p[i] = 0.0f;
    p[pCS] = (float)cs;
    p[pDEFOCUS] = (float)df;
    p[pC12a] = (float)c12a;
    p[pC12b] = (float)c12b;
    p[pC23a] = (float)c23a;
    p[pC23b] = (float)c23b;
    return chi(p, alx, aly, 1);
}
/* the stacking-sequence parser (slicelib.cpp:558-713); exitflag 0 returns the
 * error code instead of exiting (its messages still print) */
REF_API int ref_parlay(const char *c, int *islice, int nsmax, int lmax, int *nslice) {
    const int rc = parlay(c, islice, nsmax, lmax, nslice, 0);
    std::fflush(stdout);  // its error messages go out while the caller still captures them
    return rc;
}
/* floatTIFF write / read (floatTIFF.cpp:1263-1838, 324-346) on a plain
 * float array [ny][nx]; returns the class's own codes (1 = success). */
REF_API int ref_tiff_write(const char *path, const float *pix, int nx, int ny, int npix, const float *params,
                           int nparam, float rmin, float rmax, float imin, float imax, float dx, float dy) {
    floatTIFF t;
    t.resize(nx, ny);
    t.setnpix(npix);
    for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) t(ix, iy) = pix[ix + iy * nx];
    t.zeroParam();
    for (int i = 0; i < nparam && i < t.maxParam(); i++) t.setParam(i, params[i]);
    return t.write(path, rmin, rmax, imin, imax, dx, dy);
}
REF_API int ref_tiff_read(const char *path, float *pix, long cap, int *nx, int *ny, int *npix, float *params) {
    floatTIFF t;
    const int rc = t.read(path);
    if (rc != 1) return rc;
    *nx = t.nx();
    *ny = t.ny();
    *npix = t.getnpix();
    if ((long)t.nx() * t.ny() > cap) return -100;
    for (int iy = 0; iy < t.ny(); iy++)
        for (int ix = 0; ix < t.nx(); ix++) pix[ix + iy * t.nx()] = t(ix, iy);
    for (int i = 0; i < t.maxParam(); i++) params[i] = t.getParam(i);
    return rc;
}
REF_API double ref_rangauss(unsigned long *seed) { return rangauss(seed); }
REF_API double ref_ranflat(unsigned long *seed) { return ranflat(seed); }
