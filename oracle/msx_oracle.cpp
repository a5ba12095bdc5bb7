/*
 * msx_oracle.cpp -- CPU restatement of the reference's multislice path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * engine in multislice_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * Every function restates one reference routine (file:line given with each).
 * Two arithmetic modes are provided where the reference has a precision
 * choice:
 *   *_f64 : everything in double / complex128 (the "FP64 run" of the path);
 *   *_f32 : Kirkland mulslice's float arithmetic, statement for statement
 *           (float wave, float transmission, float propagator phase,
 *           float 1/(nx*ny) inverse scale).
 *
 * Data layout (same as Kirkland's cfpix, cfpix.hpp:99-126): a 2D complex
 * image of nx*ny pixels is stored x-major, y fastest: element (ix,iy) at
 * offset iy + ix*ny, interleaved (re,im).
 *
 * FFT sign convention (cfpix.cpp:115-145,177-195): fft() is FFTW_BACKWARD
 * (exponent +i, unscaled) and ifft() is FFTW_FORWARD (exponent -i) followed
 * by a multiply by 1/(nx*ny).  Here sign=+1 is the exponent +i transform.
 *
 * Pinning (each function and what fixes it, all in tests/):
 *   wavelength, sigma, bessi0, bessk0, vzatom(kirkland), featom, freqn, chi:
 *       compiled reference slicelib.cpp (oracle/_ref) on many arguments, and
 *       closed forms (lambda, sigma at 100/200/300 kV; scipy k0/i0).
 *   vzatom(kirkland): the Hankel identity 2*pi*int vz(r) J0(2 pi k r) r dr =
 *       2 pi a0 e f_e(k^2)  (pins al, ag and the exp(-pi^2 r^2/d) term).
 *   vzatom(lmultislice): direct evaluation of ModelPotential.cu:232-254 with
 *       table rows checked against FPARAMS.DAT.
 *   fft2d: brute-force DFT (tiny N), numpy.fft (FP64), the FFTW 2.1.5 float
 *       library vendored by the reference (oracle/_ref), Parseval, round trip.
 *   potential: brute force over all atoms and pixels with periodic images.
 *   transmission / propagator / multislice: |t|=1, vacuum identity, Fresnel
 *       semigroup, strict band-limit edge, translation covariance, weak-phase
 *       limit, intensity conservation.
 *   probe / detector: Parseval normalisation, integer shift = circular
 *       shift, annulus partition sums to the total.
 *   ctf: closed-form chi, aperture mask.
 *   illum_angles: the Gauss circle count of lattice points (exact integer
 *       arithmetic) for cone radii away from lattice points; rings as
 *       differences of disks.
 *   defocus_steps: closed-form Gaussian weights, step count floor(2.5 s/d).
 *   chi_astig: the compiled reference's multipole chi() (slicelib.cpp:
 *       258-318, multiMode 1) with C12a = -dfa2 sin(2 phi2), C12b = dfa2
 *       cos(2 phi2), C23a = -dfa3 sin(3 phi3), C23b = dfa3 cos(3 phi3), times
 *       2 pi/lambda -- another parametrisation of the same function.
 *   partial_image: one angle and one defocus reduce it to ctf_image of the
 *       coherent exit wave; vacuum gives 1; the defocus sum equals the
 *       weighted mean of single-defocus images; a four-angle ring equals
 *       the mean of coherent images of incident waves built from the
 *       compiled reference's freqn positions.
 *   potential_rec (Doyle-Turner, reciprocal space): one atom at the origin
 *       gives the real radial f_e(|g|/2) exp(-B|g|^2/4)/area at every
 *       reflection; translation theorem; Hermitian symmetry without
 *       absorption; the Fourier sum equals the closed-form real-space
 *       Gaussian potential of Curves.cpp:23-34 summed over periodic images;
 *       table rows against the reference's Curves.cpp bank.
 *   potential_rs (Doyle-Turner, real-space boxes): brute force over all
 *       pixels with the box rule written as explicit interval tests; a box
 *       inside the cell gives the same sum in cut and periodic forms; the
 *       periodic form of a box crossing an edge equals the cut form of the
 *       shifted atom plus its wrapped remainder; integral of one atom's
 *       potential = 4 pi^2 sum_j a_j (DT forward scattering f_e(0) / pi).
 *   potential_rec_curve / potential_rs_curve (tabulated curves): with a
 *       fine table sampled from the Doyle-Turner formula they reproduce the
 *       Doyle-Turner potentials to the linear-interpolation error (h^2 f''/8).
 */
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../multislice_amd/csrc/fparams_tables.inc"
#include "../multislice_amd/csrc/doyle_turner.inc"

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

/* ------------------------------------------------------------------ */
/* scalars: slicelib.cpp:2795-2806 (wavelength), 2387-2401 (sigma)     */
/* ------------------------------------------------------------------ */
double wavelength(double kev) {
    const double emass = 510.99906, hc = 12.3984244;
    return hc / std::sqrt(kev * (2 * emass + kev));
}
double sigma(double kev) {
    const double emass = 510.99906;
    double x = (emass + kev) / (2.0 * emass + kev);
    double pi = 4.0 * std::atan(1.0);
    return 2.0 * pi * x / (wavelength(kev) * kev);
}

/* slicelib.cpp:168-194 (bessi0), 210-237 (bessk0); the same coefficients
 * are ModelPotential.h:57-75 (Abramowitz & Stegun 9.8.1-9.8.6). */
double bessi0(double x) {
    static const double i0a[] = {1.0, 3.5156229, 3.0899424, 1.2067492,
                                 0.2659732, 0.0360768, 0.0045813};
    static const double i0b[] = {0.39894228, 0.01328592, 0.00225319,
                                 -0.00157565, 0.00916281, -0.02057706,
                                 0.02635537, -0.01647633, 0.00392377};
    double ax = std::fabs(x), sum, t;
    if (ax <= 3.75) {
        t = x / 3.75;
        t = t * t;
        sum = i0a[6];
        for (int i = 5; i >= 0; i--) sum = sum * t + i0a[i];
    } else {
        t = 3.75 / ax;
        sum = i0b[8];
        for (int i = 7; i >= 0; i--) sum = sum * t + i0b[i];
        sum = std::exp(ax) * sum / std::sqrt(ax);
    }
    return sum;
}
double bessk0(double x) {
    static const double k0a[] = {-0.57721566, 0.42278420, 0.23069756,
                                 0.03488590, 0.00262698, 0.00010750,
                                 0.00000740};
    static const double k0b[] = {1.25331414, -0.07832358, 0.02189568,
                                 -0.01062446, 0.00587872, -0.00251540,
                                 0.00053208};
    double ax = std::fabs(x), x2, sum;
    if (ax > 0.0 && ax <= 2.0) {
        x2 = ax / 2.0;
        x2 = x2 * x2;
        sum = k0a[6];
        for (int i = 5; i >= 0; i--) sum = sum * x2 + k0a[i];
        sum = -std::log(ax / 2.0) * bessi0(x) + sum;
    } else if (ax > 2.0) {
        x2 = 2.0 / ax;
        sum = k0b[6];
        for (int i = 5; i >= 0; i--) sum = sum * x2 + k0b[i];
        sum = std::exp(-ax) * sum / std::sqrt(ax);
    } else {
        sum = 1.0e20;
    }
    return sum;
}

/* Projected atomic potential in V*Angstrom.
 * formula 0 (kirkland): slicelib.cpp:2666-2699 vzatom:
 *     al*sum a_i K0(2 pi r sqrt(b_i)) + ag*sum (c_i/d_i) exp(-pi^2 r^2/d_i),
 *     al=300.8242834, ag=150.4121417, pi=3.141592654 (the literal used there)
 * formula 1 (lmultislice): ModelPotential.cu:232-254
 *     300.73079394295*sum a_i K0(6.2831853071796 r sqrt(b_i))
 *   + 150.36539697148*sum (c_i/d_i) exp(-6.2831853071796 r^2/d_i)
 * r is clamped to >= 1e-10 (slicelib.cpp:2686, ModelPotential.cu:188). */
double vzatom(int Z, double radius, int table, int formula) {
    if (Z < 1 || Z > 103 || table < 0 || table > 1) return 0.0;
    const double *p = MSX_FPARAMS[table][Z - 1];
    double r = std::fabs(radius);
    if (r < 1.0e-10) r = 1.0e-10;
    double suml = 0.0, sumg = 0.0;
    if (formula == 0) {
        const double al = 300.8242834, ag = 150.4121417, pi = 3.141592654;
        double x = 2.0 * pi * r;
        for (int i = 0; i < 6; i += 2) suml += p[i] * bessk0(x * std::sqrt(p[i + 1]));
        x = pi * r;
        x = x * x;
        for (int i = 6; i < 12; i += 2) sumg += p[i] * std::exp(-x / p[i + 1]) / p[i + 1];
        return al * suml + ag * sumg;
    }
    double dR1 = 6.2831853071796 * r;
    for (int k = 0; k < 3; k++) suml += p[2 * k] * bessk0(dR1 * std::sqrt(p[2 * k + 1]));
    suml *= 300.73079394295;
    for (int k = 0; k < 3; k++)
        sumg += (p[6 + 2 * k] / p[7 + 2 * k]) * std::exp(-(6.2831853071796 * r * r) / p[7 + 2 * k]);
    sumg *= 150.36539697148;
    return suml + sumg;
}

/* slicelib.cpp:350-374 featom: f_e(k^2) = sum a/(k^2+b) + sum c exp(-d k^2) */
double featom(int Z, double k2, int table) {
    if (Z < 1 || Z > 103 || table < 0 || table > 1) return 0.0;
    const double *p = MSX_FPARAMS[table][Z - 1];
    double sum = 0.0;
    for (int i = 0; i < 6; i += 2) sum += p[i] / (k2 + p[i + 1]);
    for (int i = 6; i < 12; i += 2) sum += p[i] * std::exp(-k2 * p[i + 1]);
    return sum;
}

/* slicelib.cpp:433-450 freqn, float arithmetic as there. */
void freqn(float *ko, float *ko2, float *xo, int nk, double ak) {
    int imid = (int)(nk / 2.0 + 0.5);
    for (int i = 0; i < nk; i++) {
        if (xo) xo[i] = ((float)(i * ak)) / ((float)(nk - 1));
        if (i > imid) ko[i] = ((float)(i - nk)) / ((float)ak);
        else ko[i] = ((float)i) / ((float)ak);
        ko2[i] = ko[i] * ko[i];
    }
}
/* signed frequency index with freqn's rule */
inline int freq_index(int i, int nk) {
    int imid = (int)(nk / 2.0 + 0.5);
    return (i > imid) ? i - nk : i;
}

/* ------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time (any N; prime factors  */
/* handled by direct sums).  Unscaled; exponent sign*2*pi*i*j*k/N.      */
/* The reference's own transform is FFTW (cfpix.cpp:115-145); this is a  */
/* textbook Cooley-Tukey, checked against a brute-force DFT and numpy.  */
/* ------------------------------------------------------------------ */
template <typename R>
struct Fft1d {
    int n;
    std::vector<std::complex<R>> w;  // w[j] = exp(sign*2*pi*i*j/n), from double
    std::vector<std::complex<R>> tmp;
    Fft1d(int n_, int sign) : n(n_), w(n_), tmp(n_) {
        for (int j = 0; j < n; j++) {
            double a = 2.0 * M_PI * (double)j / (double)n * (double)sign;
            w[j] = std::complex<R>((R)std::cos(a), (R)std::sin(a));
        }
    }
    void rec(const std::complex<R> *in, std::complex<R> *out, int m, int stride) {
        if (m == 1) { out[0] = in[0]; return; }
        int p = 2;
        while (m % p) p++;
        if (m % 4 == 0) p = 4;
        int q = m / p;
        for (int r = 0; r < p; r++) rec(in + r * stride, out + r * q, q, stride * p);
        // combine: X[k + s*q] = sum_r out[r*q + k] * W_m^{r*(k+s*q)}
        int wstep = n / m;
        std::complex<R> v[64];
        for (int k = 0; k < q; k++) {
            for (int r = 0; r < p; r++) v[r] = out[r * q + k];
            for (int s = 0; s < p; s++) {
                std::complex<R> acc(0, 0);
                int e = k + s * q;
                for (int r = 0; r < p; r++) {
                    long idx = ((long)r * e) % m;
                    acc += v[r] * w[(size_t)(idx * wstep)];
                }
                tmp[k + s * q] = acc;
            }
        }
        for (int j = 0; j < m; j++) out[j] = tmp[j];
    }
    void run(std::complex<R> *data, int stride) {
        std::vector<std::complex<R>> in(n), out(n);
        for (int j = 0; j < n; j++) in[j] = data[(size_t)j * stride];
        rec(in.data(), out.data(), n, 1);
        for (int j = 0; j < n; j++) data[(size_t)j * stride] = out[j];
    }
};

template <typename R>
void fft2d(std::complex<R> *d, int nx, int ny, int sign) {
    Fft1d<R> fy(ny, sign), fx(nx, sign);
    for (int ix = 0; ix < nx; ix++) fy.run(d + (size_t)ix * ny, 1);
    for (int iy = 0; iy < ny; iy++) fx.run(d + iy, ny);
}

/* Kirkland cfpix::ifft scale (cfpix.cpp:135-144) */
template <typename R>
void scale_inv(std::complex<R> *d, int nx, int ny) {
    R s = (R)1 / ((R)nx * (R)ny);
    for (size_t j = 0; j < (size_t)nx * ny; j++) d[j] *= s;
}

/* ------------------------------------------------------------------ */
/* band limit k2max: mulslice.cpp:481-489 (float arithmetic)           */
/* ------------------------------------------------------------------ */
float k2max_of(int nx, int ny, float ax, float by, float bw) {
    float k2max = nx / (2.0F * ax);
    float tctx = ny / (2.0F * by);
    if (tctx < k2max) k2max = tctx;
    k2max = bw * k2max;
    return k2max * k2max;
}

/* mask pass of mulslice.cpp:515-522 and slicelib.cpp:755-778: zero all
 * pixels with kx2[ix]+ky2[iy] >= k2max (float sum, strict <). */
template <typename R>
void bandlimit(std::complex<R> *d, int nx, int ny, const float *kx2, const float *ky2, float k2max) {
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            float k2 = ky2[iy] + kx2[ix];
            if (!(k2 < k2max)) d[(size_t)ix * ny + iy] = 0;
        }
}

/* ------------------------------------------------------------------ */
/* propagator tables: mulslice.cpp:494-513                              */
/*   t = pi*cz*(k^2*lambda - k*2tan(tilt)); P = (cos t, -sin t)          */
/* f32 mode: the float expression of the reference.  f64 mode: the same  */
/* in double with k = index/a exact.                                     */
/* ------------------------------------------------------------------ */
void prop_f32(float *pr, float *pi_, const float *k, const float *k2, int n, float cz, float wavlen, float tct) {
    float pi = (float)(4.0 * std::atan(1.0));  /* mulslice.cpp:174 sets pi = (float)(4*atan(1)) */
    float scale = pi * cz;
    for (int i = 0; i < n; i++) {
        double t = scale * (k2[i] * wavlen - k[i] * tct);
        pr[i] = (float)std::cos(t);
        pi_[i] = (float)-std::sin(t);
    }
}
void prop_f64(double *pr, double *pi_, int n, double a, double cz, double wavlen, double tct) {
    for (int i = 0; i < n; i++) {
        double k = (double)freq_index(i, n) / a;
        double t = M_PI * cz * (k * k * wavlen - k * tct);
        pr[i] = std::cos(t);
        pi_[i] = -std::sin(t);
    }
}

/* slicelib.cpp:744-782 propagate (separable propagator, zero outside k2max) */
template <typename R>
void propagate(std::complex<R> *w, const R *pxr, const R *pxi, const R *pyr, const R *pyi,
               const float *kx2, const float *ky2, float k2max, int nx, int ny) {
    for (int ix = 0; ix < nx; ix++) {
        std::complex<R> *row = w + (size_t)ix * ny;
        if (kx2[ix] < k2max) {
            R xr = pxr[ix], xi = pxi[ix];
            for (int iy = 0; iy < ny; iy++) {
                if ((kx2[ix] + ky2[iy]) < k2max) {
                    R wr = row[iy].real(), wi = row[iy].imag();
                    R yr = pyr[iy], yi = pyi[iy];
                    R tr = wr * yr - wi * yi;
                    R ti = wr * yi + wi * yr;
                    row[iy] = std::complex<R>(tr * xr - ti * xi, tr * xi + ti * xr);
                } else {
                    row[iy] = 0;
                }
            }
        } else {
            for (int iy = 0; iy < ny; iy++) row[iy] = 0;
        }
    }
}

struct Freq {
    std::vector<float> kx, kx2, ky, ky2;
    Freq(int nx, int ny, double ax, double by) : kx(nx), kx2(nx), ky(ny), ky2(ny) {
        freqn(kx.data(), kx2.data(), nullptr, nx, ax);
        freqn(ky.data(), ky2.data(), nullptr, ny, by);
    }
};

/* chi(k) of slicelib.cpp:258-318 restricted to Cs and defocus (multiMode 0),
 * times 2*pi/lambda:  chi = pi*lambda*k^2*(0.5*Cs*lambda^2*k^2 - df)   */
inline double chi_k2(double k2, double lambda, double cs, double df) {
    return M_PI * lambda * k2 * (0.5 * cs * lambda * lambda * k2 - df);
}

}  // namespace

/* ====================================================================== */
/* exported API                                                           */
/* ====================================================================== */

ORC_API double orc_wavelength(double kev) { return wavelength(kev); }
ORC_API double orc_sigma(double kev) { return sigma(kev); }
ORC_API double orc_bessi0(double x) { return bessi0(x); }
ORC_API double orc_bessk0(double x) { return bessk0(x); }
ORC_API double orc_vzatom(int Z, double r, int table, int formula) { return vzatom(Z, r, table, formula); }
ORC_API double orc_featom(int Z, double k2, int table) { return featom(Z, k2, table); }
ORC_API void orc_freqn(float *ko, float *ko2, float *xo, int nk, double ak) { freqn(ko, ko2, xo, nk, ak); }
ORC_API float orc_k2max(int nx, int ny, double ax, double by, double bw) {
    return k2max_of(nx, ny, (float)ax, (float)by, (float)bw);
}
ORC_API double orc_fparam(int table, int Z, int i) { return MSX_FPARAMS[table][Z - 1][i]; }

/* brute-force 2D DFT (FP64), for tiny grids only */
ORC_API void orc_dft2d_f64(double *data, int nx, int ny, int sign) {
    std::vector<std::complex<double>> in((size_t)nx * ny), out((size_t)nx * ny);
    std::memcpy(in.data(), data, sizeof(double) * 2 * nx * ny);
    for (int kx = 0; kx < nx; kx++)
        for (int ky = 0; ky < ny; ky++) {
            std::complex<double> acc(0, 0);
            for (int x = 0; x < nx; x++)
                for (int y = 0; y < ny; y++) {
                    double a = 2.0 * M_PI * sign *
                               ((double)((long)kx * x % nx) / nx + (double)((long)ky * y % ny) / ny);
                    acc += in[(size_t)x * ny + y] * std::complex<double>(std::cos(a), std::sin(a));
                }
            out[(size_t)kx * ny + ky] = acc;
        }
    std::memcpy(data, out.data(), sizeof(double) * 2 * nx * ny);
}
ORC_API void orc_fft2d_f64(double *data, int nx, int ny, int sign) {
    fft2d(reinterpret_cast<std::complex<double> *>(data), nx, ny, sign);
}
ORC_API void orc_fft2d_f32(float *data, int nx, int ny, int sign) {
    fft2d(reinterpret_cast<std::complex<float> *>(data), nx, ny, sign);
}

/* Projected potential of a slab of atoms on an nx*ny grid per slice.
 * ModelPotential.cu:27-99 (slicing) and :125-230 (per-pixel sum):
 *   slice of an atom: kz*dz <= z < (kz+1)*dz, dz = c/nslices;
 *   pixel (ix,iy) sits at (ix*dx, iy*dy), dx=a/nx, dy=b/ny;
 *   minimum image per axis: d=|x_atom - x_pix|; if d >= a/2 then d -= a;
 *   only r < radius contributes; r clamped >= 1e-10; sum in FP64.
 * Atom occupancy multiplies the contribution (Kirkland atompot convention;
 * LMultiSlice ignores occupancy, i.e. occ=1).
 * Output V[kz][ix][iy] (V*Angstrom), y fastest. */
ORC_API int orc_potential(const int *Z, const double *xyz, const double *occ, long natom,
                          double a, double b, double c, int nx, int ny, int nslices,
                          double radius, int table, int formula, double *V) {
    if (nx <= 0 || ny <= 0 || nslices <= 0 || radius <= 0) return -1;
    const double dx = a / nx, dy = b / ny, dz = c / nslices;
    std::memset(V, 0, sizeof(double) * (size_t)nslices * nx * ny);
    for (long i = 0; i < natom; i++) {
        double z = xyz[3 * i + 2];
        long kz = (long)std::floor(z / dz);
        if (kz < 0 || kz >= nslices) continue;
        if (!(kz * dz <= z && z < (kz + 1) * dz)) {  // exact comparison of the reference
            if (z < kz * dz) kz--; else kz++;
            if (kz < 0 || kz >= nslices) continue;
        }
        // periodic supercell: positions wrapped into [0,a) x [0,b) (the
        // reference's fractional coordinates already lie there)
        double xa = xyz[3 * i + 0] - a * std::floor(xyz[3 * i + 0] / a);
        double ya = xyz[3 * i + 1] - b * std::floor(xyz[3 * i + 1] / b);
        if (xa >= a) xa -= a;
        if (ya >= b) ya -= b;
        double w = occ ? occ[i] : 1.0;
        double *Vs = V + (size_t)kz * nx * ny;
        for (int ix = 0; ix < nx; ix++) {
            double dX = std::fabs(xa - ix * dx);
            if (dX >= a / 2.0) dX -= a;
            if (std::fabs(dX) >= radius) continue;
            for (int iy = 0; iy < ny; iy++) {
                double dY = std::fabs(ya - iy * dy);
                if (dY >= b / 2.0) dY -= b;
                double r = std::sqrt(dX * dX + dY * dY);
                if (r < radius) Vs[(size_t)ix * ny + iy] += w * vzatom(Z[i], r, table, formula);
            }
        }
    }
    return 0;
}

/* Transmission function of one layer, mulslice.cpp:414-421 + 515-522:
 *   t = exp(i*phase), phase = sigma(kev)*V/1000 (V in V*Angstrom; the same
 *   phase LMultiSlice's phaseObject applies, kernel.cu:4-18), then
 *   fft -> zero k^2 >= k2max -> ifft (x 1/(nx ny)).
 * f32: float V, float phase product as in mulslice (scale*V in float),
 * cos/sin in double rounded to float, float FFTs.  */
ORC_API void orc_transmission_f32(const float *V, int nlayer, int nx, int ny, double ax, double by,
                                  double kev, double bw, float *t) {
    Freq f(nx, ny, ax, by);
    float k2max = k2max_of(nx, ny, (float)ax, (float)by, (float)bw);
    float scale = (float)(sigma(kev) / 1000.0);
    for (int l = 0; l < nlayer; l++) {
        std::complex<float> *tl = reinterpret_cast<std::complex<float> *>(t) + (size_t)l * nx * ny;
        const float *Vl = V + (size_t)l * nx * ny;
        for (size_t j = 0; j < (size_t)nx * ny; j++) {
            float vz = scale * Vl[j];
            tl[j] = std::complex<float>((float)std::cos(vz), (float)std::sin(vz));
        }
        if (bw > 0) {
            fft2d(tl, nx, ny, +1);
            bandlimit(tl, nx, ny, f.kx2.data(), f.ky2.data(), k2max);
            fft2d(tl, nx, ny, -1);
            scale_inv(tl, nx, ny);
        }
    }
}
ORC_API void orc_transmission_f64(const double *V, int nlayer, int nx, int ny, double ax, double by,
                                  double kev, double bw, double *t) {
    Freq f(nx, ny, ax, by);
    float k2max = k2max_of(nx, ny, (float)ax, (float)by, (float)bw);
    double scale = sigma(kev) / 1000.0;
    for (int l = 0; l < nlayer; l++) {
        std::complex<double> *tl = reinterpret_cast<std::complex<double> *>(t) + (size_t)l * nx * ny;
        const double *Vl = V + (size_t)l * nx * ny;
        for (size_t j = 0; j < (size_t)nx * ny; j++) {
            double vz = scale * Vl[j];
            tl[j] = std::complex<double>(std::cos(vz), std::sin(vz));
        }
        if (bw > 0) {
            fft2d(tl, nx, ny, +1);
            bandlimit(tl, nx, ny, f.kx2.data(), f.ky2.data(), k2max);
            fft2d(tl, nx, ny, -1);
            scale_inv(tl, nx, ny);
        }
    }
}

/* Propagator tables for one layer of thickness cz (mulslice.cpp:494-513).
 * out: pxr,pxi[nx], pyr,pyi[ny]. */
ORC_API void orc_propagator_f64(int nx, int ny, double ax, double by, double kev, double cz,
                                double ctiltx, double ctilty, double *pxr, double *pxi, double *pyr,
                                double *pyi) {
    double lam = wavelength(kev);
    prop_f64(pxr, pxi, nx, ax, cz, lam, 2.0 * std::tan(ctiltx));
    prop_f64(pyr, pyi, ny, by, cz, lam, 2.0 * std::tan(ctilty));
}

/* Coherent multislice loop, mulslice.cpp:703-739:
 *   for islice: wave *= trans[layer[islice]]; wave.fft();
 *               propagate(wave, prop[layer], kx2, ky2, k2max); wave.ifft();
 *               intensity[islice] = sum |wave|^2 / (nx*ny)   (printed there;
 *               here sum|wave|^2 without the extra 1/(nx ny) of :735, i.e.
 *               the integrated intensity of a wave normalised to sum 1 stays 1)
 * trans: nlayer complex images; cz: thickness per layer; layer: nslice
 * indices.  wave: in/out complex image. */
ORC_API int orc_multislice_f32(const float *trans, int nlayer, const float *cz, const int *layer,
                               int nslice, int nx, int ny, double ax, double by, double kev,
                               double ctiltx, double ctilty, double bw, float *wave, double *intensity) {
    Freq f(nx, ny, ax, by);
    float k2max = k2max_of(nx, ny, (float)ax, (float)by, (float)bw);
    float wavlen = (float)wavelength(kev);
    float tctx = (float)(2.0 * std::tan(ctiltx)), tcty = (float)(2.0 * std::tan(ctilty));
    std::vector<float> pxr((size_t)nlayer * nx), pxi((size_t)nlayer * nx), pyr((size_t)nlayer * ny),
        pyi((size_t)nlayer * ny);
    for (int l = 0; l < nlayer; l++) {
        prop_f32(&pxr[(size_t)l * nx], &pxi[(size_t)l * nx], f.kx.data(), f.kx2.data(), nx, cz[l], wavlen, tctx);
        prop_f32(&pyr[(size_t)l * ny], &pyi[(size_t)l * ny], f.ky.data(), f.ky2.data(), ny, cz[l], wavlen, tcty);
    }
    std::complex<float> *w = reinterpret_cast<std::complex<float> *>(wave);
    const std::complex<float> *T = reinterpret_cast<const std::complex<float> *>(trans);
    Fft1d<float> fyp(ny, +1), fxp(nx, +1), fym(ny, -1), fxm(nx, -1);
    for (int is = 0; is < nslice; is++) {
        int l = layer[is];
        if (l < 0 || l >= nlayer) return -1;
        const std::complex<float> *tl = T + (size_t)l * nx * ny;
        for (size_t j = 0; j < (size_t)nx * ny; j++) {  // cfpix::operator*= (cfpix.cpp:296-318)
            float wr = w[j].real(), wi = w[j].imag(), tr = tl[j].real(), ti = tl[j].imag();
            w[j] = std::complex<float>(wr * tr - wi * ti, wr * ti + wi * tr);
        }
        for (int ix = 0; ix < nx; ix++) fyp.run(w + (size_t)ix * ny, 1);
        for (int iy = 0; iy < ny; iy++) fxp.run(w + iy, ny);
        propagate(w, &pxr[(size_t)l * nx], &pxi[(size_t)l * nx], &pyr[(size_t)l * ny], &pyi[(size_t)l * ny],
                  f.kx2.data(), f.ky2.data(), k2max, nx, ny);
        for (int ix = 0; ix < nx; ix++) fym.run(w + (size_t)ix * ny, 1);
        for (int iy = 0; iy < ny; iy++) fxm.run(w + iy, ny);
        scale_inv(w, nx, ny);
        if (intensity) {
            double s = 0;
            for (size_t j = 0; j < (size_t)nx * ny; j++)
                s += (double)w[j].real() * w[j].real() + (double)w[j].imag() * w[j].imag();
            intensity[is] = s;
        }
    }
    return 0;
}

/* Oblique-cell slice loop: Musli's ApplyPropagator (Musli console 1.3
 * (STEM) 2/Muslib/Propagator.cpp:55-107) in Kirkland's sign convention (the
 * reference applies exp(+i con) with its FFT signs reversed; here exp(-i con),
 * the transform pair of cfpix):
 *   x = i/(a sin gam), y = j/(b sin gam)  (freqn indices; Musli's i < N/2
 *      rule differs only at the Nyquist index, which the mask removes)
 *   s = x (x - 2 y cos gam) + y^2                     (Propagator.cpp:86)
 *   con = pi lambda dz s - 2 pi dz (x tan tx + y tan ty) (:67-72, :94)
 *   P = exp(-i con) where s < smax, else 0, smax = (bw min(N/(2a), M/(2b)))^2
 *      (:69-70 with bw = 1; bw = 2/3 gives Kirkland's band limit radius)
 * in FP64 throughout, with Kirkland's lambda (the engine's, not Musli's
 * 0.387818/sqrt(...) fit, :67).  Slice: w *= t; fft; w *= P; ifft. */
ORC_API int orc_multislice_oblique_f64(const double *trans, int nlayer, const double *cz, const int *layer,
                                       int nslice, int nx, int ny, double ax, double by, double gam, double kev,
                                       double ctiltx, double ctilty, double bw, double *wave) {
    const double sg = std::sin(gam), cg = std::cos(gam);
    const double lam = wavelength(kev);
    double kmax = std::min(nx / (2.0 * ax), ny / (2.0 * by)) * bw;
    const double smax = kmax * kmax;
    std::complex<double> *w = reinterpret_cast<std::complex<double> *>(wave);
    const std::complex<double> *T = reinterpret_cast<const std::complex<double> *>(trans);
    Fft1d<double> fyp(ny, +1), fxp(nx, +1), fym(ny, -1), fxm(nx, -1);
    for (int is = 0; is < nslice; is++) {
        const int l = layer[is];
        if (l < 0 || l >= nlayer) return -1;
        const std::complex<double> *tl = T + (size_t)l * nx * ny;
        for (size_t j = 0; j < (size_t)nx * ny; j++) w[j] *= tl[j];
        for (int ix = 0; ix < nx; ix++) fyp.run(w + (size_t)ix * ny, 1);
        for (int iy = 0; iy < ny; iy++) fxp.run(w + iy, ny);
        const double gk = M_PI * lam * cz[l], xtg = 2.0 * M_PI * cz[l] * std::tan(ctiltx),
                     ytg = 2.0 * M_PI * cz[l] * std::tan(ctilty);
        for (int ix = 0; ix < nx; ix++) {
            const double x = (double)freq_index(ix, nx) / (ax * sg);
            for (int iy = 0; iy < ny; iy++) {
                const double y = (double)freq_index(iy, ny) / (by * sg);
                const double s2 = x * (x - 2.0 * y * cg) + y * y;
                std::complex<double> &v = w[(size_t)ix * ny + iy];
                if (!(s2 < smax)) {
                    v = 0;
                } else {
                    const double con = gk * s2 - xtg * x - ytg * y;
                    v *= std::complex<double>(std::cos(con), -std::sin(con));
                }
            }
        }
        for (int ix = 0; ix < nx; ix++) fym.run(w + (size_t)ix * ny, 1);
        for (int iy = 0; iy < ny; iy++) fxm.run(w + iy, ny);
        scale_inv(w, nx, ny);
    }
    return 0;
}

/* FP64 slice loop with optional beam records (mulslice.cpp:682-724): after
 * propagate, before the inverse transform, beams[is][ib] = (1/(nx ny)) *
 * wave(h, k) for the beam indices hk[ib] = (h, k), wrapped once when negative
 * and clamped to the grid as mulslice.cpp:693-700 does. */
static int multislice_f64(const double *trans, int nlayer, const double *cz, const int *layer, int nslice, int nx,
                          int ny, double ax, double by, double kev, double ctiltx, double ctilty, double bw,
                          double *wave, double *intensity, const int *hk, int nbeams, double *beams) {
    Freq f(nx, ny, ax, by);
    float k2max = k2max_of(nx, ny, (float)ax, (float)by, (float)bw);
    double wavlen = wavelength(kev);
    std::vector<double> pxr((size_t)nlayer * nx), pxi((size_t)nlayer * nx), pyr((size_t)nlayer * ny),
        pyi((size_t)nlayer * ny);
    for (int l = 0; l < nlayer; l++) {
        prop_f64(&pxr[(size_t)l * nx], &pxi[(size_t)l * nx], nx, ax, cz[l], wavlen, 2.0 * std::tan(ctiltx));
        prop_f64(&pyr[(size_t)l * ny], &pyi[(size_t)l * ny], ny, by, cz[l], wavlen, 2.0 * std::tan(ctilty));
    }
    std::complex<double> *w = reinterpret_cast<std::complex<double> *>(wave);
    const std::complex<double> *T = reinterpret_cast<const std::complex<double> *>(trans);
    Fft1d<double> fyp(ny, +1), fxp(nx, +1), fym(ny, -1), fxm(nx, -1);
    for (int is = 0; is < nslice; is++) {
        int l = layer[is];
        if (l < 0 || l >= nlayer) return -1;
        const std::complex<double> *tl = T + (size_t)l * nx * ny;
        for (size_t j = 0; j < (size_t)nx * ny; j++) w[j] *= tl[j];
        for (int ix = 0; ix < nx; ix++) fyp.run(w + (size_t)ix * ny, 1);
        for (int iy = 0; iy < ny; iy++) fxp.run(w + iy, ny);
        propagate(w, &pxr[(size_t)l * nx], &pxi[(size_t)l * nx], &pyr[(size_t)l * ny], &pyi[(size_t)l * ny],
                  f.kx2.data(), f.ky2.data(), k2max, nx, ny);
        for (int ib = 0; ib < nbeams; ib++) {
            int h = hk[2 * ib], k = hk[2 * ib + 1];
            if (h < 0) h = nx + h;
            if (k < 0) k = ny + k;
            h = std::min(std::max(h, 0), nx - 1);
            k = std::min(std::max(k, 0), ny - 1);
            const std::complex<double> v = w[(size_t)h * ny + k] / ((double)nx * ny);
            beams[2 * ((size_t)is * nbeams + ib)] = v.real();
            beams[2 * ((size_t)is * nbeams + ib) + 1] = v.imag();
        }
        for (int ix = 0; ix < nx; ix++) fym.run(w + (size_t)ix * ny, 1);
        for (int iy = 0; iy < ny; iy++) fxm.run(w + iy, ny);
        scale_inv(w, nx, ny);
        if (intensity) {
            double s = 0;
            for (size_t j = 0; j < (size_t)nx * ny; j++) s += std::norm(w[j]);
            intensity[is] = s;
        }
    }
    return 0;
}

ORC_API int orc_multislice_f64(const double *trans, int nlayer, const double *cz, const int *layer,
                               int nslice, int nx, int ny, double ax, double by, double kev,
                               double ctiltx, double ctilty, double bw, double *wave, double *intensity) {
    return multislice_f64(trans, nlayer, cz, layer, nslice, nx, ny, ax, by, kev, ctiltx, ctilty, bw, wave, intensity,
                          nullptr, 0, nullptr);
}

/* The thickness series of the coherent loop (mulslice.cpp:682-737): beams
 * [nslice][nbeams] complex as above, intensity[nslice] = sum |wave|^2 (the
 * reference prints it times 1/(nx ny), :735). */
ORC_API int orc_multislice_series_f64(const double *trans, int nlayer, const double *cz, const int *layer,
                                      int nslice, int nx, int ny, double ax, double by, double kev, double ctiltx,
                                      double ctilty, double bw, double *wave, double *intensity, const int *hk,
                                      int nbeams, double *beams) {
    return multislice_f64(trans, nlayer, cz, layer, nslice, nx, ny, ax, by, kev, ctiltx, ctilty, bw, wave, intensity,
                          hk, nbeams, beams);
}

/* STEM probe, after Musli ConvergedBeam + Xi (MSProject.cs:458-514) in
 * Kirkland's conventions (orthogonal cell, exp(-i chi), chi of
 * slicelib.cpp:258-318):
 *   Psi0(k) = A(k) exp(-i chi(k)) exp(+2 pi i (kx xp + ky yp)),
 *   A(k) = 1 for kmin2 <= kx2+ky2 <= kmax2 (float sum of freqn tables),
 *   kmax = alpha_max/lambda, kmin = alpha_min/lambda,
 *   psi0 = ifft(Psi0) (cfpix::ifft, exponent -i, 1/(nx ny)), then scaled so
 *   that sum |psi0|^2 = 1 (Musli leaves the probe unnormalised; the scale is
 *   a global constant per probe).
 * The phase -chi + 2 pi k.x_p is evaluated in FP64 with k = index/a. */
ORC_API int orc_probe_f64(int nx, int ny, double ax, double by, double kev, double alpha_max,
                          double alpha_min, double cs, double df, double xp, double yp, double *wave) {
    Freq f(nx, ny, ax, by);
    double lam = wavelength(kev);
    float kmax2 = (float)((alpha_max / lam) * (alpha_max / lam));
    float kmin2 = (float)((alpha_min / lam) * (alpha_min / lam));
    std::complex<double> *w = reinterpret_cast<std::complex<double> *>(wave);
    long nap = 0;
    for (int ix = 0; ix < nx; ix++) {
        double kx = (double)freq_index(ix, nx) / ax;
        for (int iy = 0; iy < ny; iy++) {
            double ky = (double)freq_index(iy, ny) / by;
            float k2f = f.kx2[ix] + f.ky2[iy];
            std::complex<double> v(0, 0);
            if (k2f <= kmax2 && k2f >= kmin2) {
                double ph = -chi_k2(kx * kx + ky * ky, lam, cs, df) + 2.0 * M_PI * (kx * xp + ky * yp);
                v = std::complex<double>(std::cos(ph), std::sin(ph));
                nap++;
            }
            w[(size_t)ix * ny + iy] = v;
        }
    }
    if (nap == 0) return -2;
    fft2d(w, nx, ny, -1);
    scale_inv(w, nx, ny);
    double s = 0;
    for (size_t j = 0; j < (size_t)nx * ny; j++) s += std::norm(w[j]);
    double g = 1.0 / std::sqrt(s);
    for (size_t j = 0; j < (size_t)nx * ny; j++) w[j] *= g;
    return 0;
}

/* STEM scan with annular detectors (Musli StartCalculations /
 * RunCalculation / SaveSTEM: MSProjectUseOnce.cs:100-291,
 * MSUseOnceCalculator.cs:125-277, MSProject.cs:792-842) using the Kirkland
 * slice step above.  After the propagator of slice s (k-space, before the
 * inverse FFT), if s+1 is a checkpoint thickness (in slices):
 *   out[ck][d][p] = sum_{kin2_d <= k2 <= kout2_d} |Psi(k)|^2 / (nx*ny)
 * with k2 the float sum kx2[ix]+ky2[iy] and bounds rounded to float.
 * The probe is normalised (sum |psi0|^2 = 1), so out is the fraction of the
 * incident intensity (Musli accumulates unnormalised sums).
 * det: ndet pairs (k_inner, k_outer) in 1/Angstrom.  pos: nprobe (x,y) in
 * Angstrom.  ck: nck slice counts (1..nslice).  OpenMP over probes. */
ORC_API int orc_stem_f64(const double *trans, int nlayer, const double *cz, const int *layer, int nslice,
                         int nx, int ny, double ax, double by, double kev, double bw, double alpha_max,
                         double alpha_min, double cs, double df, const double *pos, int nprobe,
                         const double *det, int ndet, const int *ck, int nck, double *out, int nthreads) {
    Freq f(nx, ny, ax, by);
    float k2max = k2max_of(nx, ny, (float)ax, (float)by, (float)bw);
    double wavlen = wavelength(kev);
    std::vector<double> pxr((size_t)nlayer * nx), pxi((size_t)nlayer * nx), pyr((size_t)nlayer * ny),
        pyi((size_t)nlayer * ny);
    for (int l = 0; l < nlayer; l++) {
        prop_f64(&pxr[(size_t)l * nx], &pxi[(size_t)l * nx], nx, ax, cz[l], wavlen, 0.0);
        prop_f64(&pyr[(size_t)l * ny], &pyi[(size_t)l * ny], ny, by, cz[l], wavlen, 0.0);
    }
    std::vector<float> dlo(ndet), dhi(ndet);
    for (int d = 0; d < ndet; d++) {
        dlo[d] = (float)(det[2 * d] * det[2 * d]);
        dhi[d] = (float)(det[2 * d + 1] * det[2 * d + 1]);
    }
    std::memset(out, 0, sizeof(double) * (size_t)nck * ndet * nprobe);
    const std::complex<double> *T = reinterpret_cast<const std::complex<double> *>(trans);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int p = 0; p < nprobe; p++) {
        std::vector<std::complex<double>> w((size_t)nx * ny);
        if (orc_probe_f64(nx, ny, ax, by, kev, alpha_max, alpha_min, cs, df, pos[2 * p], pos[2 * p + 1],
                          reinterpret_cast<double *>(w.data())) != 0) {
            bad = 1;
            continue;
        }
        Fft1d<double> fyp(ny, +1), fxp(nx, +1), fym(ny, -1), fxm(nx, -1);
        int ick = 0;
        for (int is = 0; is < nslice && ick < nck; is++) {
            int l = layer[is];
            const std::complex<double> *tl = T + (size_t)l * nx * ny;
            for (size_t j = 0; j < (size_t)nx * ny; j++) w[j] *= tl[j];
            for (int ix = 0; ix < nx; ix++) fyp.run(w.data() + (size_t)ix * ny, 1);
            for (int iy = 0; iy < ny; iy++) fxp.run(w.data() + iy, ny);
            propagate(w.data(), &pxr[(size_t)l * nx], &pxi[(size_t)l * nx], &pyr[(size_t)l * ny],
                      &pyi[(size_t)l * ny], f.kx2.data(), f.ky2.data(), k2max, nx, ny);
            if (ck[ick] == is + 1) {
                for (int d = 0; d < ndet; d++) {
                    double s = 0;
                    for (int ix = 0; ix < nx; ix++)
                        for (int iy = 0; iy < ny; iy++) {
                            float k2 = f.kx2[ix] + f.ky2[iy];
                            if (k2 >= dlo[d] && k2 <= dhi[d]) s += std::norm(w[(size_t)ix * ny + iy]);
                        }
                    out[((size_t)ick * ndet + d) * nprobe + p] = s / ((double)nx * ny);
                }
                ick++;
            }
            if (ick >= nck) break;
            for (int ix = 0; ix < nx; ix++) fym.run(w.data() + (size_t)ix * ny, 1);
            for (int iy = 0; iy < ny; iy++) fxm.run(w.data() + iy, ny);
            scale_inv(w.data(), nx, ny);
        }
    }
    return bad ? -2 : 0;
}

/* FP32 STEM variant for the CPU baseline timing (same steps, float wave and
 * Kirkland float propagator tables). */
ORC_API int orc_stem_f32(const float *trans, int nlayer, const float *cz, const int *layer, int nslice,
                         int nx, int ny, double ax, double by, double kev, double bw, double alpha_max,
                         double alpha_min, double cs, double df, const double *pos, int nprobe,
                         const double *det, int ndet, const int *ck, int nck, double *out, int nthreads) {
    Freq f(nx, ny, ax, by);
    float k2max = k2max_of(nx, ny, (float)ax, (float)by, (float)bw);
    float wavlen = (float)wavelength(kev);
    std::vector<float> pxr((size_t)nlayer * nx), pxi((size_t)nlayer * nx), pyr((size_t)nlayer * ny),
        pyi((size_t)nlayer * ny);
    for (int l = 0; l < nlayer; l++) {
        prop_f32(&pxr[(size_t)l * nx], &pxi[(size_t)l * nx], f.kx.data(), f.kx2.data(), nx, cz[l], wavlen, 0.f);
        prop_f32(&pyr[(size_t)l * ny], &pyi[(size_t)l * ny], f.ky.data(), f.ky2.data(), ny, cz[l], wavlen, 0.f);
    }
    std::vector<float> dlo(ndet), dhi(ndet);
    for (int d = 0; d < ndet; d++) {
        dlo[d] = (float)(det[2 * d] * det[2 * d]);
        dhi[d] = (float)(det[2 * d + 1] * det[2 * d + 1]);
    }
    std::memset(out, 0, sizeof(double) * (size_t)nck * ndet * nprobe);
    const std::complex<float> *T = reinterpret_cast<const std::complex<float> *>(trans);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int p = 0; p < nprobe; p++) {
        std::vector<std::complex<double>> w64((size_t)nx * ny);
        if (orc_probe_f64(nx, ny, ax, by, kev, alpha_max, alpha_min, cs, df, pos[2 * p], pos[2 * p + 1],
                          reinterpret_cast<double *>(w64.data())) != 0) {
            bad = 1;
            continue;
        }
        std::vector<std::complex<float>> w((size_t)nx * ny);
        for (size_t j = 0; j < w.size(); j++) w[j] = std::complex<float>((float)w64[j].real(), (float)w64[j].imag());
        Fft1d<float> fyp(ny, +1), fxp(nx, +1), fym(ny, -1), fxm(nx, -1);
        int ick = 0;
        for (int is = 0; is < nslice && ick < nck; is++) {
            int l = layer[is];
            const std::complex<float> *tl = T + (size_t)l * nx * ny;
            for (size_t j = 0; j < (size_t)nx * ny; j++) w[j] *= tl[j];
            for (int ix = 0; ix < nx; ix++) fyp.run(w.data() + (size_t)ix * ny, 1);
            for (int iy = 0; iy < ny; iy++) fxp.run(w.data() + iy, ny);
            propagate(w.data(), &pxr[(size_t)l * nx], &pxi[(size_t)l * nx], &pyr[(size_t)l * ny],
                      &pyi[(size_t)l * ny], f.kx2.data(), f.ky2.data(), k2max, nx, ny);
            if (ck[ick] == is + 1) {
                for (int d = 0; d < ndet; d++) {
                    double s = 0;
                    for (int ix = 0; ix < nx; ix++)
                        for (int iy = 0; iy < ny; iy++) {
                            float k2 = f.kx2[ix] + f.ky2[iy];
                            if (k2 >= dlo[d] && k2 <= dhi[d]) s += std::norm(w[(size_t)ix * ny + iy]);
                        }
                    out[((size_t)ick * ndet + d) * nprobe + p] = s / ((double)nx * ny);
                }
                ick++;
            }
            if (ick >= nck) break;
            for (int ix = 0; ix < nx; ix++) fym.run(w.data() + (size_t)ix * ny, 1);
            for (int iy = 0; iy < ny; iy++) fxm.run(w.data() + iy, ny);
            scale_inv(w.data(), nx, ny);
        }
    }
    return bad ? -2 : 0;
}

/* Objective-lens image of an exit wave (mulslice.cpp:612-656, one defocus,
 * no illumination spread):  Psi = fft(psi); Psi *= exp(-i chi) for
 * kx2+ky2 <= k2maxo (aperture, inclusive), else 0; psi_img = ifft(Psi);
 * image = |psi_img|^2.  chi = pi*lambda*k2*(0.5*Cs*lambda^2*k2 - df),
 * Cs and df in Angstrom, aobj in rad.  FP64. */
ORC_API void orc_ctf_image_f64(const double *wave, int nx, int ny, double ax, double by, double kev,
                               double cs, double df, double aobj, double *img_wave, double *image) {
    Freq f(nx, ny, ax, by);
    double lam = wavelength(kev);
    float k2maxo = (float)((aobj / lam) * (aobj / lam));
    std::vector<std::complex<double>> w((size_t)nx * ny);
    std::memcpy(w.data(), wave, sizeof(double) * 2 * nx * ny);
    fft2d(w.data(), nx, ny, +1);
    for (int ix = 0; ix < nx; ix++) {
        double kx = (double)freq_index(ix, nx) / ax;
        for (int iy = 0; iy < ny; iy++) {
            double ky = (double)freq_index(iy, ny) / by;
            float k2f = f.kx2[ix] + f.ky2[iy];
            std::complex<double> &v = w[(size_t)ix * ny + iy];
            if (k2f <= k2maxo) {
                double c = chi_k2(kx * kx + ky * ky, lam, cs, df);
                v *= std::complex<double>(std::cos(c), -std::sin(c));
            } else {
                v = 0;
            }
        }
    }
    fft2d(w.data(), nx, ny, -1);
    scale_inv(w.data(), nx, ny);
    if (img_wave) std::memcpy(img_wave, w.data(), sizeof(double) * 2 * nx * ny);
    if (image)
        for (size_t j = 0; j < (size_t)nx * ny; j++) image[j] = std::norm(w[j]);
}

/* ------------------------------------------------------------------ */
/* Partial-coherence TEM image (mulslice.cpp:539-676)                  */
/* ------------------------------------------------------------------ */
/* Illumination angles, mulslice.cpp:556-583, in the reference's float
 * arithmetic (the discrete choices -- how many angles, which defocus steps --
 * must come out identical):  rx = 1/ax, nacx = (int)(acmax/(wavlen rx) + 1.5),
 * q2max = (acmax/wavlen)^2, q2min = (acmin/wavlen)^2; loop iqy outer,
 * iqx inner over [-nac, nac]; keep q2 = (iqx rx)^2 + (iqy ry)^2 with
 * q2min <= q2 <= q2max.  Writes (iqx, iqy) pairs in loop order; returns
 * the count (entries beyond maxn are counted, not written). */
ORC_API int orc_illum_angles(int nx, int ny, double ax, double by, double kev, double acmin, double acmax,
                             int *iq, int maxn) {
    (void)nx;
    (void)ny;
    const float axf = (float)ax, byf = (float)by;
    const float wavlen = (float)wavelength((double)(float)kev);
    const float rx = 1.0F / axf, ry = 1.0F / byf;
    const float acmaxf = (float)acmax, acminf = (float)acmin;
    const int nacx = (int)((acmaxf / (wavlen * rx)) + 1.5F);
    const int nacy = (int)((acmaxf / (wavlen * ry)) + 1.5F);
    float q2max = acmaxf / wavlen;
    q2max = q2max * q2max;
    float q2min = acminf / wavlen;
    q2min = q2min * q2min;
    int n = 0;
    for (int iqy = -nacy; iqy <= nacy; iqy++) {
        const float qy = iqy * ry;
        const float qy2 = qy * qy;
        for (int iqx = -nacx; iqx <= nacx; iqx++) {
            const float qx = iqx * rx;
            const float q2 = qx * qx + qy2;
            if ((q2 <= q2max) && (q2 >= q2min)) {
                if (n < maxn) {
                    iq[2 * n] = iqx;
                    iq[2 * n + 1] = iqy;
                }
                n++;
            }
        }
    }
    return n;
}

/* Defocus steps and weights, mulslice.cpp:556,617-645: ndf =
 * (int)(2.5 sigmaf / dfdelt) in float; df_i = df0 + i dfdelt for
 * i = -ndf..ndf with weight exp(-0.5 ((df_i - df0)/sigmaf)^2) (evaluated in
 * double here).  sigmaf <= 0 gives the single step df0 with weight 1 (the
 * reference divides 0 by 0 there).  Writes offsets i*dfdelt and weights;
 * returns 2 ndf + 1, or -1 for dfdelt <= 0 with sigmaf > 0. */
ORC_API int orc_defocus_steps(double sigmaf, double dfdelt, double *offset, double *weight, int maxn) {
    if (!(sigmaf > 0)) {
        if (maxn > 0) {
            offset[0] = 0.0;
            weight[0] = 1.0;
        }
        return 1;
    }
    if (!(dfdelt > 0)) return -1;
    const int ndf = (int)((2.5F * (float)sigmaf) / (float)dfdelt);
    int n = 0;
    for (int idf = -ndf; idf <= ndf; idf++, n++) {
        if (n >= maxn) continue;
        const double x = (idf * dfdelt) / sigmaf;
        offset[n] = idf * dfdelt;
        weight[n] = std::exp(-0.5 * x * x);
    }
    return n;
}

/* Aberration phase of the partial-coherence lens, mulslice.cpp:620-628:
 *   chi = pi lambda k^2 (0.5 Cs lambda^2 k^2 - df
 *         + dfa2 sin(2 (phi - dfa2phi)) + 2 dfa3 lambda |k| sin(3 (phi - dfa3phi)) / 3),
 * phi = atan2(ky, kx).  FP64. */
inline double chi_astig(double kx, double ky, double lam, double cs, double df, double dfa2, double dfa2phi,
                        double dfa3, double dfa3phi) {
    const double k2 = kx * kx + ky * ky;
    const double phi = std::atan2(ky, kx);
    return M_PI * lam * k2 *
           (0.5 * cs * lam * lam * k2 - df + dfa2 * std::sin(2.0 * (phi - dfa2phi)) +
            2.0 * dfa3 * lam * std::sqrt(k2) * std::sin(3.0 * (phi - dfa3phi)) / 3.0);
}
ORC_API double orc_chi_astig(double kx, double ky, double kev, double cs, double df, double dfa2, double dfa2phi,
                             double dfa3, double dfa3phi) {
    return chi_astig(kx, ky, wavelength(kev), cs, df, dfa2, dfa2phi, dfa3, dfa3phi);
}

/* The partial-coherence image, FP64 (mulslice.cpp:539-676): for every
 * illumination angle (iqx, iqy) the incident wave
 *   exp(2 pi i (qx xpos[ix] + qy ypos[iy])),  qx = iqx/ax,
 *   xpos[ix] = ix ax/(nx-1)  (freqn, slicelib.cpp:433-450: note nx-1),
 * runs through the coherent slice loop; its spectrum Psi = fft(psi) is then,
 * for every defocus step, multiplied by exp(-i chi) inside the objective
 * aperture (kx2+ky2 <= (aobj/lambda)^2, float sum as the reference) and
 * transformed back, and w_i |ifft(...)|^2 is summed.  The sum is divided by
 * (number of angles) * (sum of weights).  image: nx*ny doubles.  Returns the
 * number of angles (0: none inside the illumination cone; -1: bad input). */
ORC_API int orc_partial_image_f64(const double *trans, int nlayer, const double *cz, const int *layer, int nslice,
                                  int nx, int ny, double ax, double by, double kev, double ctiltx, double ctilty,
                                  double bw, double cs, double df0, double aobj, double dfa2, double dfa2phi,
                                  double dfa3, double dfa3phi, double acmin, double acmax, double sigmaf,
                                  double dfdelt, double *image) {
    if (nx < 2 || ny < 2) return -1;
    const int nill = orc_illum_angles(nx, ny, ax, by, kev, acmin, acmax, nullptr, 0);
    std::vector<int> iq(2 * (size_t)std::max(nill, 1));
    orc_illum_angles(nx, ny, ax, by, kev, acmin, acmax, iq.data(), nill);
    const int ndfs = orc_defocus_steps(sigmaf, dfdelt, nullptr, nullptr, 0);
    if (ndfs < 1) return -1;
    std::vector<double> off(ndfs), wt(ndfs);
    orc_defocus_steps(sigmaf, dfdelt, off.data(), wt.data(), ndfs);
    double wsum = 0;
    for (double w : wt) wsum += w;
    Freq f(nx, ny, ax, by);
    const double lam = wavelength(kev);
    const float k2maxo = (float)((aobj / lam) * (aobj / lam));
    const size_t img = (size_t)nx * ny;
    std::vector<double> pix(img, 0.0);
    std::vector<std::complex<double>> w(img), Psi(img);
    for (int a = 0; a < nill; a++) {
        const long iqx = iq[2 * a], iqy = iq[2 * a + 1];
        for (int ix = 0; ix < nx; ix++)
            for (int iy = 0; iy < ny; iy++) {
                // 2 pi (iqx ix/(nx-1) + iqy iy/(ny-1)), reduced exactly in integers
                const long mx = ((iqx * ix) % (nx - 1) + (nx - 1)) % (nx - 1);
                const long my = ((iqy * iy) % (ny - 1) + (ny - 1)) % (ny - 1);
                const double t = 2.0 * M_PI * ((double)mx / (nx - 1) + (double)my / (ny - 1));
                w[(size_t)ix * ny + iy] = std::complex<double>(std::cos(t), std::sin(t));
            }
        if (orc_multislice_f64(trans, nlayer, cz, layer, nslice, nx, ny, ax, by, kev, ctiltx, ctilty, bw,
                               reinterpret_cast<double *>(w.data()), nullptr) != 0)
            return -1;
        fft2d(w.data(), nx, ny, +1);
        for (int d = 0; d < ndfs; d++) {
            const double df = df0 + off[d];
            for (int ix = 0; ix < nx; ix++) {
                const double kx = (double)freq_index(ix, nx) / ax;
                for (int iy = 0; iy < ny; iy++) {
                    const double ky = (double)freq_index(iy, ny) / by;
                    const size_t j = (size_t)ix * ny + iy;
                    if (f.kx2[ix] + f.ky2[iy] <= k2maxo) {
                        const double c = chi_astig(kx, ky, lam, cs, df, dfa2, dfa2phi, dfa3, dfa3phi);
                        Psi[j] = w[j] * std::complex<double>(std::cos(c), -std::sin(c));
                    } else {
                        Psi[j] = 0;
                    }
                }
            }
            fft2d(Psi.data(), nx, ny, -1);
            scale_inv(Psi.data(), nx, ny);
            for (size_t j = 0; j < img; j++) pix[j] += wt[d] * std::norm(Psi[j]);
        }
    }
    const double sc = nill > 0 ? 1.0 / ((double)nill * wsum) : 0.0;
    for (size_t j = 0; j < img; j++) image[j] = pix[j] * sc;
    return nill;
}

/* ------------------------------------------------------------------ */
/* Reciprocal-space potential (Muslib)                                 */
/* ------------------------------------------------------------------ */
/* scat_ampl_rec_DT_DW (Curves.cpp:3-11): f_e(s) = sum_j a_j exp(-b_j s^2) */
ORC_API double orc_dt_fe(int Z, double s) {
    const double *p = MSX_DOYLE_TURNER[Z];
    double f = 0.0;
    for (int j = 0; j < 4; j++) f += p[2 * j] * std::exp(-p[2 * j + 1] * s * s);
    return f;
}
/* real-space counterpart, scat_ampl_rs_DT_DW (Curves.cpp:12-21) at distance r:
 * 4 pi sum_j a_j/(b_j+dw) exp(-4 pi^2 r^2/(b_j+dw)) (terms with b_j = 0 skipped) */
ORC_API double orc_dt_rs(int Z, double r, double dw) {
    const double *p = MSX_DOYLE_TURNER[Z];
    double f = 0.0;
    for (int j = 0; j < 4; j++)
        if (p[2 * j + 1] > 0) f += p[2 * j] / (p[2 * j + 1] + dw) * std::exp(-4.0 * M_PI * M_PI * r * r / (p[2 * j + 1] + dw));
    return 4.0 * M_PI * f;
}

/* Potential_rec_DT_DW (Pot_rec_DT_DW.cpp:3-132), whole slice (SL = 0, EL = M):
 * adds to gli, in the reference's layout (row jj of N interleaved complex
 * values, x fastest), the Fourier coefficients
 *   V(g) += f_e(|g|/2) exp(-dw |g|^2/4) occ / area * exp(-2 pi i (h xx + k yy)) * (1 + i ads)
 * for every reflection (h, k) with |h|*nmult < N/2 and |k|*mmult < M/2 (integer
 * halves), stored at column h*nmult mod N and row k*mmult mod M; reciprocal
 * metric of the oblique cell |g|^2 = (hA)^2 + (kB)^2 - 2hkAB cos(gam),
 * A = 1/(a sin gam), B = 1/(b sin gam), area = a b sin gam (:6-12, :30-35,
 * :69-84).  xx, yy are fractional coordinates.  The reference enumerates the
 * four sign quadrants explicitly; this loop visits the same set of signed
 * reflections. */
ORC_API void orc_potential_rec_dt(double a, double b, double gam, int nmult, int mmult, int N, int M, long na,
                                  const double *xx, const double *yy, const double *dw, const double *ads,
                                  const double *occ, const int *Z, double *gli) {
    const double A = 1.0 / (a * std::sin(gam)), B = 1.0 / (b * std::sin(gam));
    const double inv_area = 1.0 / (a * b * std::sin(gam));
    const double cg = std::cos(gam);
    const int hmax = (N / 2 - 1) / nmult, kmax = (M / 2 - 1) / mmult;  /* largest |h| with |h|*nmult < N/2 */
    for (long t = 0; t < na; t++) {
        for (int k = -kmax; k <= kmax; k++) {
            const int row = ((k * mmult) % M + M) % M;
            for (int h = -hmax; h <= hmax; h++) {
                const int col = ((h * nmult) % N + N) % N;
                const double x = h * A, y = k * B;
                const double g2 = x * x + y * y - 2.0 * x * y * cg;
                const double ss = 0.25 * g2;
                const double res = orc_dt_fe(Z[t], std::sqrt(ss)) * std::exp(-dw[t] * ss) * inv_area * occ[t];
                const double ph = -2.0 * M_PI * (h * xx[t] + k * yy[t]);
                const double rr = std::cos(ph) * res, ri = std::sin(ph) * res;
                double *o = gli + 2 * ((size_t)row * N + col);
                o[0] += rr - ads[t] * ri;
                o[1] += ri + ads[t] * rr;
            }
        }
    }
}

/* scat_ampl_rs2_DT_DW (Curves.cpp:23-34), argument r^2 */
ORC_API double orc_dt_rs2(int Z, double r2, double dw) {
    const double *p = MSX_DOYLE_TURNER[Z];
    double f = 0.0;
    for (int j = 0; j < 4; j++)
        if (p[2 * j + 1] > 0) f += p[2 * j] / (p[2 * j + 1] + dw) * std::exp(-4.0 * M_PI * M_PI * r2 / (p[2 * j + 1] + dw));
    return 4.0 * M_PI * f;
}

/* Potential_rs_DT_DW_cutpot (Pot_RS.cpp:111-155) and Potential_rs_DT_DW_periodic
 * (Pot_RS.cpp:350-486): adds occ * scat_ampl_rs2_DT_DW(Z, r^2, dw) (and times
 * ads to the imaginary part) over a box of pixels around every atom.  Box of
 * atom (xx, yy) (fractional): ibb = (long)((xx - drax) N), iee = (long)((xx +
 * drax) N) with drax = atrad sin(gam)/a (C truncation toward zero), the same
 * along y; pixels i in [max(ibb,0), min(iee,N)), j likewise, distance to the
 * atom at (xx a, yy b) with r^2 = x (x + 2 y cos gam) + y^2.  The periodic form
 * adds the eight wrapped pieces of the box with the atom image shifted by a
 * and/or b; a piece wrapped from the far edge keeps one more pixel (je =
 * jee - M + 1, ie = iee - N + 1).  Layout: row j of N interleaved complex. */
ORC_API void orc_potential_rs_dt(double a, double b, double gam, double atrad, int N, int M, long na,
                                 const double *xx, const double *yy, const double *dw, const double *ads,
                                 const double *occ, const int *Z, int periodic, double *gli) {
    const double sg = std::sin(gam), cg = std::cos(gam);
    const double drax = atrad * sg / a, dray = atrad * sg / b;
    const double dxa = a / (double)N, dyb = b / (double)M;
    for (long k = 0; k < na; k++) {
        const long ibb = (long)((xx[k] - drax) * N), iee = (long)((xx[k] + drax) * N);
        const long jbb = (long)((yy[k] - dray) * M), jee = (long)((yy[k] + dray) * M);
        for (int l = 0; l < (periodic ? 9 : 1); l++) {
            long ib = std::max(ibb, 0L), ie = std::min(iee, (long)N);
            long jb = std::max(jbb, 0L), je = std::min(jee, (long)M);
            double x0 = xx[k] * a, y0 = yy[k] * b;
            const bool lo_i = ibb < 0, hi_i = iee >= N, lo_j = jbb < 0, hi_j = jee >= M;
            /* x part: 0 = main, +1 = wrapped from the left (atom image at +a),
             * -1 = wrapped from the right (image at -a); same for y */
            static const int sx[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
            static const int sy[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
            if ((sx[l] == 1 && !lo_i) || (sx[l] == -1 && !hi_i) || (sy[l] == 1 && !lo_j) || (sy[l] == -1 && !hi_j))
                continue;
            if (sx[l] == 1) { ib = ibb + N; ie = N; x0 += a; }
            if (sx[l] == -1) { ib = 0; ie = iee - N + 1; x0 -= a; }
            if (sy[l] == 1) { jb = jbb + M; je = M; y0 += b; }
            if (sy[l] == -1) { jb = 0; je = jee - M + 1; y0 -= b; }
            for (long j = jb; j < je; j++) {
                const double y = j * dyb - y0;
                for (long i = ib; i < ie; i++) {
                    const double x = i * dxa - x0;
                    const double r2 = x * (x + 2.0 * y * cg) + y * y;
                    const double res = orc_dt_rs2(Z[k], r2, dw[k]) * occ[k];
                    double *o = gli + 2 * ((size_t)j * N + i);
                    o[0] += res;
                    o[1] += res * ads[k];
                }
            }
        }
    }
}

/* linear interpolation in a curve table with spacing dds; 0 past the end
 * (Pot_RS.cpp:40-46, Pot_rec_curve_DW.cpp:38-44) */
static bool curve_at(const double *c, int npoints, double u, double *out) {
    const double f = std::floor(u);
    if (f >= (double)(npoints - 1)) return false;
    const long j = (long)f;
    const double rem = u - (double)j;
    *out = c[j] + (c[j + 1] - c[j]) * rem;
    return true;
}

/* Potential_rec_curve_DW (Pot_rec_curve_DW.cpp:7-164): as orc_potential_rec_dt
 * with the factor f_e(|g|/2) exp(-dw |g|^2/4) / area replaced by the curve of
 * the atom's type at s = |g|/2 (table spacing dds; the table already holds the
 * Debye-Waller factor and 1/area, MSProject.cs:850-863), times occ. */
ORC_API void orc_potential_rec_curve(double a, double b, double gam, int nmult, int mmult, int N, int M, long na,
                                     const double *xx, const double *yy, const double *ads, const double *occ,
                                     const int *type, const double *curves, int npoints, double dds, double *gli) {
    const double A = 1.0 / (a * std::sin(gam)), B = 1.0 / (b * std::sin(gam));
    const double cg = std::cos(gam);
    const int hmax = (N / 2 - 1) / nmult, kmax = (M / 2 - 1) / mmult;
    for (long t = 0; t < na; t++) {
        const double *c = curves + (size_t)type[t] * npoints;
        for (int k = -kmax; k <= kmax; k++) {
            const int row = ((k * mmult) % M + M) % M;
            for (int h = -hmax; h <= hmax; h++) {
                const int col = ((h * nmult) % N + N) % N;
                const double x = h * A, y = k * B;
                const double g2 = x * x + y * y - 2.0 * x * y * cg;
                double v;
                if (!curve_at(c, npoints, 0.5 * std::sqrt(g2) / dds, &v)) continue;
                const double res = v * occ[t];
                const double ph = -2.0 * M_PI * (h * xx[t] + k * yy[t]);
                const double rr = std::cos(ph) * res, ri = std::sin(ph) * res;
                double *o = gli + 2 * ((size_t)row * N + col);
                o[0] += rr - ads[t] * ri;
                o[1] += ri + ads[t] * rr;
            }
        }
    }
}

/* Potential_rs_curve_DW_cutpot / _periodic (Pot_RS.cpp:5-55, 209-349): the
 * boxes of orc_potential_rs_dt with the value curve(type, r / dds) * occ. */
ORC_API void orc_potential_rs_curve(double a, double b, double gam, double atrad, int N, int M, long na,
                                    const double *xx, const double *yy, const double *ads, const double *occ,
                                    const int *type, const double *curves, int npoints, double dds, int periodic,
                                    double *gli) {
    const double sg = std::sin(gam), cg = std::cos(gam);
    const double drax = atrad * sg / a, dray = atrad * sg / b;
    const double dxa = a / (double)N, dyb = b / (double)M;
    static const int sx[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
    static const int sy[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
    for (long k = 0; k < na; k++) {
        const double *c = curves + (size_t)type[k] * npoints;
        const long ibb = (long)((xx[k] - drax) * N), iee = (long)((xx[k] + drax) * N);
        const long jbb = (long)((yy[k] - dray) * M), jee = (long)((yy[k] + dray) * M);
        for (int l = 0; l < (periodic ? 9 : 1); l++) {
            long ib = std::max(ibb, 0L), ie = std::min(iee, (long)N);
            long jb = std::max(jbb, 0L), je = std::min(jee, (long)M);
            double x0 = xx[k] * a, y0 = yy[k] * b;
            if ((sx[l] == 1 && !(ibb < 0)) || (sx[l] == -1 && !(iee >= N)) || (sy[l] == 1 && !(jbb < 0)) ||
                (sy[l] == -1 && !(jee >= M)))
                continue;
            if (sx[l] == 1) { ib = ibb + N; ie = N; x0 += a; }
            if (sx[l] == -1) { ib = 0; ie = iee - N + 1; x0 -= a; }
            if (sy[l] == 1) { jb = jbb + M; je = M; y0 += b; }
            if (sy[l] == -1) { jb = 0; je = jee - M + 1; y0 -= b; }
            for (long j = jb; j < je; j++) {
                const double y = j * dyb - y0;
                for (long i = ib; i < ie; i++) {
                    const double x = i * dxa - x0;
                    double v;
                    if (!curve_at(c, npoints, std::sqrt(x * (x + 2.0 * y * cg) + y * y) / dds, &v)) continue;
                    const double res = v * occ[k];
                    double *o = gli + 2 * ((size_t)j * N + i);
                    o[0] += res;
                    o[1] += res * ads[k];
                }
            }
        }
    }
}
