/*
 * msx.h -- C ABI of the msx MI355X multislice engine (libmsx.so).
 *
 * This is the drop-in boundary for the reference's multislice hot path:
 * the projected-potential build (LMultiSlice ModelPotential), the transmission
 * function, the slice loop (LMultiSlice ModelSimulated / Kirkland mulslice),
 * STEM probes and annular detectors (Musli), and the objective-lens image.
 * Every entry point names the reference interface it replaces.
 *
 * Conventions (all entry points):
 *   - Plain C types and pointers only; no C++ or framework types.
 *   - Return 0 on success, a negative MSX_E* code on failure; msx_last_error()
 *     returns a per-thread description of the last failure.  Nothing calls
 *     exit() or waits for input (the reference's CUERR macro did,
 *     LMultiSlice/stdafx.h:72-76).
 *   - Images are complex, stored x-major with y fastest, element (ix,iy) at
 *     offset iy + ix*ny, (re,im) interleaved -- the layout of the reference's
 *     cfpix (multislice/cfpix.hpp:99-126).  Batches of images are contiguous.
 *   - prec = MSX_F32 (complex64 / float) or MSX_F64 (complex128 / double) for
 *     every complex and real buffer of the call.
 *   - where = MSX_HOST: buffers are host memory, the call copies in and out and
 *     returns when the result is in host memory (as the reference's synchronous
 *     calls do).  where = MSX_DEVICE: buffers are device memory on the
 *     context's GPU; work is queued on the context stream and the call returns
 *     after that stream has completed it.
 *   - FFT sign: the "forward" transform (real space -> k-space) has exponent
 *     +2*pi*i*k.x and is unscaled; the inverse has exponent -2*pi*i*k.x and is
 *     scaled by 1/(nx*ny) -- cfpix::fft()/ifft() (cfpix.cpp:115-145).
 *   - Units: Angstrom, keV, radians.  Potentials in volt*Angstrom.
 *   - Grid sizes per axis: 16, 32, 64, 128, 256, 384, 512, 768, 1024, 2048.
 */
#ifndef MSX_H
#define MSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSX_API __attribute__((visibility("default")))

enum { MSX_F32 = 0, MSX_F64 = 1 };
enum { MSX_HOST = 0, MSX_DEVICE = 1 };
enum {
    MSX_OK = 0,
    MSX_EINVAL = -1,   /* invalid argument (sizes, Z, pointers) */
    MSX_ENOMEM = -2,   /* device allocation failed */
    MSX_EHIP = -3,     /* HIP runtime error */
    MSX_ESIZE = -4,    /* unsupported grid size */
    MSX_ESTATE = -5,   /* context misuse */
};
/* fparams tables: 0 = LMultiSlice / FPARAMS.DAT fit (ModelPotential.h:80-595),
 * 1 = Kirkland slicelib fit (multislice/slicelib.cpp:903-2138). */
enum { MSX_TABLE_LMULTISLICE = 0, MSX_TABLE_KIRKLAND = 1 };
/* projected-potential formula: 0 = Kirkland vzatom (slicelib.cpp:2666-2699),
 * 1 = LMultiSlice calculateProjectedPotential (ModelPotential.cu:232-254). */
enum { MSX_FORMULA_KIRKLAND = 0, MSX_FORMULA_LMULTISLICE = 1 };

typedef struct msx_ctx_s* msx_ctx;

/* ---- context ---------------------------------------------------------- */
/* Replaces the device selection of Dispatcher::Run (Dispatcher.cu:191-207):
 * binds one GPU and creates the context's stream.  One context per host
 * thread; calls on one context are serialised on its stream. */
MSX_API int msx_create(int device, msx_ctx* out);
MSX_API int msx_destroy(msx_ctx ctx);
/* Use an external HIP stream (hipStream_t passed as void*; NULL restores the
 * context's own stream). */
MSX_API int msx_set_stream(msx_ctx ctx, void* hip_stream);
/* Order the context stream after all work queued so far on hip_stream (a
 * hipStream_t passed as void*; NULL = the null stream): an event recorded on
 * hip_stream that the context stream waits for, with no host wait.  Call it
 * before an MSX_DEVICE call whose inputs another stream produced (for example
 * a framework's current stream); the MSX_DEVICE calls themselves return only
 * after the context stream has finished, so their outputs need no wait.  No
 * reference counterpart: the reference ran everything on one synchronous
 * stream (Dispatcher.cu:204-207). */
MSX_API int msx_wait_stream(msx_ctx ctx, void* hip_stream);
MSX_API const char* msx_last_error(void);
MSX_API int msx_version(void);
/* Replay slice loops (msx_multislice, msx_stem batches) from cached HIP graphs
 * (the reference relaunched every kernel and rebuilt FFT plans each slice,
 * ModelSimulated.cu:60-120).  The first call with a given set of launch
 * arguments captures the per-slice sequence; identical later calls replay it
 * with one graph launch.  Off by default; ignored while profiling. */
MSX_API int msx_set_graphs(msx_ctx ctx, int enable);
/* Autotuner: times the slice loop of an nx x ny grid on synthetic waves for
 * candidate launch shapes and keeps the fastest in the context (used by every
 * later slice loop and msx_stem of that grid and precision): waves per
 * work-group of the row and column sweeps (row: 0 = built-in heuristic;
 * column: -1 to -8 select the balanced schedule, in which -col_wpb
 * work-groups per resident slot share the (column group, wave) units, and 0
 * is that schedule with one work-group per slot), the
 * column-sweep tile width (columns per work-group; 0 = built-in width, else
 * the built-in one or its alternative for this line length) and the STEM
 * batch (at most max_batch, 0: 512; of the batches within 10% of the best
 * per-sweep rate, the one whose slice loop runs fastest back to back, the
 * smaller within 1%);
 * then, at that batch, the FFT plan of each sweep (0 = the built-in radix
 * sequence, 1 = the alternative where the line length has one: both sweeps
 * at 512 (rows 16 x 16 x 2, columns 4 x 4 x 4 x 8, against the built-in
 * 8 x 8 x 8), the row sweep at 256,
 * 384, 768, 1024 and 2048 (other thread counts per line and radix
 * sequences)).  ms_per_slice is the time of one wave
 * through one slice with the chosen shapes.  msx_set_tuning installs a saved
 * result (NULL clears it). */
typedef struct msx_tuning {
    int32_t row_wpb, col_wpb, batch;
    double ms_per_slice;
    int32_t col_tile;
    int32_t row_plan, col_plan;
} msx_tuning;
MSX_API int msx_autotune(msx_ctx ctx, int nx, int ny, int prec, int max_batch, msx_tuning* best);
MSX_API int msx_set_tuning(msx_ctx ctx, int nx, int ny, int prec, const msx_tuning* t);
MSX_API int msx_graph_stats(msx_ctx ctx, int64_t* captures, int64_t* replays);

/* Kernel timing (the reference timed its loops with cudaEvents and printed
 * them, ModelPotential.cu:46-50,92-96 and ModelSimulated.cu:43-47,178-182).
 * With profiling on, every kernel launch of the context is bracketed by HIP
 * events on the context stream and waited for, and its time and algorithmic
 * HBM bytes are added to per-kind totals (enabling resets them).  This
 * serialises launches: use it for measurement passes, not production runs. */
enum { MSX_PROF_ROW = 0, MSX_PROF_COL = 1, MSX_PROF_SPLAT = 2, MSX_PROF_DET = 3, MSX_PROF_KINDS = 4 };
/* Bandwidth probe (no counterpart in the reference; measurement only): one
 * launch copying device buffers src -> dst of nimg nx x ny complex64 images,
 * timed with HIP events on the context stream (*ms).  mode 0: a contiguous
 * copy (nimg*nx*ny a multiple of 4096, 16-byte aligned buffers); mode 8 or 16
 * (nx = 512, ny a multiple of it): the column sweep's access pattern (that
 * many adjacent columns per work-group, 16 strided rows per thread, XCD
 * order) without the transforms -- the ceiling of that pattern for the
 * roofline of the column sweep; mode 1008 (nx = 512, ny a multiple of 8): the
 * same 8-column pattern on images stored in 8-row bands (element (ix, iy) at
 * ((ix >> 3) * ny + iy) * 8 + (ix & 7)), each instruction one contiguous
 * 512-byte chunk. */
MSX_API int msx_copy_probe(msx_ctx ctx, int mode, const void* src, void* dst, int nimg, int nx, int ny,
                           double* ms);
MSX_API int msx_profile(msx_ctx ctx, int enable);
MSX_API int msx_profile_read(msx_ctx ctx, int kind, int64_t* launches, double* total_ms, double* total_bytes);

/* ---- scalars ---------------------------------------------------------- */
/* Microscope::wavelength / Sigma (Microscope.cpp:69-86) = Kirkland
 * wavelength / sigma (slicelib.cpp:2795-2806, 2387-2401). */
MSX_API double msx_wavelength(double kev);
MSX_API double msx_sigma(double kev);
/* calculateProjectedPotential (ModelPotential.cu:232-254) or vzatom
 * (slicelib.cpp:2666-2699): projected potential of atom Z at radius r. */
MSX_API double msx_vzatom(int Z, double r, int table, int formula);
/* featom (slicelib.cpp:350-374): electron scattering factor f_e(k^2). */
MSX_API double msx_featom(int Z, double k2, int table);

/* ---- atoms ------------------------------------------------------------ */
/* One atom: the fields of AModel::Record (amodel/include/BaseStructures.h:
 * 120-135) the hot path uses, with Cartesian coordinates in Angstrom
 * (Record holds fractional xsCoordinate; x = xs.x * a etc.). */
typedef struct msx_atom {
    int32_t Z;   /* atomic number 1..103 */
    float occ;   /* occupancy weight (LMultiSlice ignores it: use 1) */
    double x, y, z;
} msx_atom;

/* ---- frozen phonons ------------------------------------------------------ */
/* Isotropic Gaussian displacement of every atom for one frozen-phonon
 * configuration: x += u*gx, y += u*gy, z += u*gz with u = u_rms[Z] (Angstrom,
 * per axis) and g ~ N(0,1) from Philox-4x32-10 (counter = atom index and
 * configuration, key = seed) through Box-Muller in FP64 -- the same numbers on
 * the host (where = MSX_HOST; ctx may be NULL) and on the GPU (MSX_DEVICE;
 * in/out device arrays).  Replaces the reference's sequential generators
 * (Kirkland ReadXYZcoord wobble, slicelib.cpp:2197-2320 with rangauss,
 * :792-830; Musli Vobble, MusliRandom.cs:43-55), whose draws depend on the
 * visiting order.  u_rms: host array of 104 doubles indexed by Z. */
MSX_API int msx_displace(msx_ctx ctx, const msx_atom* in, int64_t natoms, const double* u_rms, uint64_t seed,
                         int32_t config, msx_atom* out, int where);
/* The Philox-4x32-10 block function itself (counter, key -> 4 words), exposed
 * for known-answer tests of the generator. */
MSX_API void msx_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- specimen generator (no reference counterpart) ---------------------- */
/* Amorphous specimen for synthetic workloads (the uneven-job set: amorphous
 * carbon at a given density with a minimum spacing): random sequential
 * addition of up to natoms atoms of element Z, uniform in the periodic box
 * a x b x c (Angstrom), each at least min_dist from every earlier one
 * (minimum image); stops after max_candidates draws.  Deterministic in seed
 * (Philox-4x32-10 per candidate).  out: host array of natoms; *nplaced
 * receives the count placed.  Host only (no context). */
MSX_API int msx_amorphous(double a, double b, double c, int32_t Z, int64_t natoms, double min_dist, uint64_t seed,
                          int64_t max_candidates, msx_atom* out, int64_t* nplaced);

/* ---- layer stacking sequence -------------------------------------------- */
/* Replaces Kirkland's parlay (slicelib.cpp:558-713, called by mulslice.cpp:
 * 197): "2(abc)d" -> layer indices 0 1 2 0 1 2 3 (a = 0 ... z = 25, A = 26 ...
 * Z = 51; n(...) repeats a group, nested up to 100 deep).  islice: host
 * array of nsmax entries; lmax: largest allowed layer index; *nslice gets the
 * count.  Returns parlay's codes: 0 ok, -1 layer out of range, -2 missing
 * left parenthesis, -3 nested too deep, -5 unmatched right parenthesis, -6
 * invalid character (the end of the string after white space counts as one,
 * as in the reference), -7 too many layers, -8 incomplete sequence; the
 * message is in msx_last_error().  NULL pointers give MSX_EINVAL (-1). */
MSX_API int msx_parlay(const char* c, int32_t* islice, int nsmax, int lmax, int32_t* nslice);

/* ---- Kirkland floating-point TIFF files ----------------------------------- */
/* Replaces floatTIFF::write / floatTIFF::read (floatTIFF.cpp:1263-1838,
 * 324-346), the layer and wave files of Kirkland's programs: an 8-bit preview
 * (square pixels, scaled by rmin..rmax and, for the imaginary half,
 * imin..imax), the FP32 image and 512 FP32 parameters (param[0] = npix).
 * pix: ny rows of nx floats, x fastest (floatTIFF's operator()(ix, iy)); a
 * complex image has npix = 2 and its real and imaginary halves side by side
 * (nx = 2 * width).  dx, dy: pixel size (<= 0: square).  The file matches the
 * reference's byte for byte except the 20-byte date.  Returns 0, or
 * floatTIFF::write's codes: -1 rmax == rmin, -2 npix not 1 or 2, -6 cannot
 * open, -8 cannot write; MSX_EINVAL for arrays below 2 x 2. */
MSX_API int msx_tiff_write(const char* path, const float* pix, int nx, int ny, int npix, const float* params,
                           int nparam, float rmin, float rmax, float imin, float imax, float dx, float dy);
/* Reads the first FP32 image of a TIFF file (either byte order) and the
 * parameter image after it: *nx, *ny, *npix (rounded param[0]; 1 without
 * parameters) and params (512 floats, may be NULL).  pix == NULL only
 * queries the size; else capacity (floats) must hold nx * ny.  Returns 0,
 * or -1 cannot open, -2/-3 not a TIFF file, -4 more than one sample per
 * pixel, -5 no FP32 image, -7 short read. */
MSX_API int msx_tiff_read(const char* path, float* pix, int64_t capacity, int32_t* nx, int32_t* ny, int32_t* npix,
                          float* params);

/* ---- Musli output files (no reference C ABI: Musli's C# writes them) ------ */
/* A STEM image as Musli's FlushSTEMfile (MSProject.cs:697-790) writes it:
 * stem_n * stem_m doubles in scan order, then the 160-byte trailer "~~~
 * scan_x scan_y thickness pi/2 pi/2 gamma kV stem_n stem_m thickness 0"
 * (scan extents and thickness in Angstrom, gamma in rad, kV = E/1000). */
MSX_API int msx_musli_stem_write(const char* path, const double* image, int stem_n, int stem_m, double scan_x,
                                 double scan_y, double thickness, double gamma, double kev);
/* A wave as Musli's SaveWav (MSProject.cs:522-596): n * m complex doubles
 * (interleaved, x fastest: Musli's layout), then the trailer "___ a b
 * thickness pi/2 pi/2 gamma kV n m thickness 0". */
MSX_API int msx_musli_wave_write(const char* path, const double* wave, int n, int m, double a, double b,
                                 double thickness, double gamma, double kev);

/* ---- 2D FFT ----------------------------------------------------------- */
/* cfpix::fft() / cfpix::ifft() (cfpix.cpp:115-145), batched: sign = +1 is
 * fft() (exponent +i, unscaled), sign = -1 is ifft() (exponent -i, scaled by
 * 1/(nx*ny)).  In place on nbatch images. */
MSX_API int msx_fft2d(msx_ctx ctx, void* data, int nbatch, int nx, int ny, int sign, int prec, int where);

/* ---- projected potential ---------------------------------------------- */
/* ModelPotential::calculatePotentialGrid (ModelPotential.cu:27-99) with the
 * per-pixel sum of calculatePotentialGridGPU (:125-230):
 *   atom slice kz: kz*dz <= z < (kz+1)*dz, dz = c/nslices;
 *   pixel (ix,iy) at (ix*a/nx, iy*b/ny); minimum image per axis;
 *   atoms with r < radius contribute occ*vz(r) (r clamped >= 1e-10), summed
 *   in FP64.
 * V: nslices images of nx*ny reals [kz][ix][iy] (float for MSX_F32, double for
 * MSX_F64), host or device per `where`.  An atom with Z outside 1..103 or a
 * coordinate that is not finite is refused (MSX_EINVAL, naming the first such
 * atom). */
MSX_API int msx_potential(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, double a, double b, double c,
                          int nx, int ny, int nslices, double radius, int table, int formula, void* V,
                          int prec, int where);
/* msx_potential for atoms already in device memory (atoms: a device array,
 * e.g. the output of msx_displace with MSX_DEVICE): the atoms are binned on
 * the GPU (cell of every atom, stable radix sort by cell) instead of on the
 * host, and the potential is identical to msx_potential's on the same atoms.
 * No counterpart in the reference, whose atoms always start on the host
 * (ModelPotential.cu:27-99); for frozen-phonon loops that displace on the
 * GPU.  At most INT_MAX - 1 atoms and slice x bin cells; atoms must lie in
 * one device allocation of the context's GPU (checked: MSX_EINVAL). */
MSX_API int msx_potential_device_atoms(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, double a, double b,
                                       double c, int nx, int ny, int nslices, double radius, int table, int formula,
                                       void* V, int prec, int where);
/* Reciprocal-space projected potential from structure factors: Muslib
 * Potential_rec_DT_DW (Muslib/Pot_rec_DT_DW.cpp:3-132) followed by the inverse
 * transform of the phase-grating build (MSProjectReuse.cs:425-431).  Per slice
 * (atoms assigned as in msx_potential):
 *   V(g) = sum_atoms f_e(|g|/2) exp(-dw |g|^2/4) occ / area exp(-2 pi i g.r) (1 + i ads)
 * with Doyle-Turner f_e (Curves.cpp:3-11, table Curves.cpp:83-179, Z 1..92),
 * dw the Debye-Waller B (A^2) and ads the absorptive fraction of each atom
 * (NULL: 0), over the reflections g = (h/a, k/b) of the unit cell a x b (angle
 * gam, a along x) whose grid indices |h|*nmult < nx/2 and |k|*mmult < ny/2;
 * the grid spans the supercell nmult*a x mmult*b (pass a, b of the UNIT cell
 * and atoms inside it).  Then V(r) = sum_g V(g) exp(+2 pi i g.r) in FP64, and
 * V = 2 pi a0 e Re V(r) (V*A, the units of msx_potential), Vabs = 2 pi a0 e
 * Im V(r) (may be NULL).  Layout as msx_potential: [slice][ix][iy]. */
/* Real-space projected potential from Doyle-Turner Gaussians over pixel boxes:
 * Muslib Potential_rs_DT_DW_cutpot (periodic = 0, Pot_RS.cpp:111-155) and
 * Potential_rs_DT_DW_periodic (periodic = 1, :350-486).  Every pixel of the box
 * of an atom (fractional position xf, yf in the cell a x b at angle gam):
 *   i in [max(ibb,0), min(iee,nx)), ibb = (long)((xf - atrad sin(gam)/a) nx),
 *   iee = (long)((xf + atrad sin(gam)/a) nx), likewise j along b,
 * gets occ * 4 pi sum_j a_j/(b_j+dw) exp(-4 pi^2 r^2/(b_j+dw)) (Curves.cpp:23-34)
 * with r^2 = x (x + 2 y cos gam) + y^2 to the atom, and times ads in Vabs.  The
 * periodic form adds the eight pieces of the box wrapped across the cell edges
 * (atom image shifted by a and/or b; a piece wrapped from the high edge keeps
 * one pixel more, as the reference does).  Scaled by 2 pi a0 e to V*A, layout
 * [slice][ix][iy] as msx_potential; any nx, ny >= 1. */
MSX_API int msx_potential_rs(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, const double* dw,
                             const double* ads, double a, double b, double c, double gam, double atrad, int periodic,
                             int nx, int ny, int nslices, void* V, void* Vabs, int prec, int where);
MSX_API int msx_potential_rec(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, const double* dw,
                              const double* ads, double a, double b, double c, double gam, int nmult, int mmult,
                              int nx, int ny, int nslices, void* V, void* Vabs, int prec, int where);

/* The curve tables of the tabulated-curve potentials, built from the
 * Doyle-Turner factors (host arithmetic; no context).  Rows nt = 0..ntypes-1
 * of sc[ntypes][npoints] belong to element types[nt] (0..92; row 0 of the
 * table is empty), sampled at s = 0, dds, 2 dds, ... (accumulated by
 * repeated addition, as the reference):
 *   msx_create_pots_rec_dt: CreatePots_rec_DT_DW (Muslib/Curves.cpp:45-63),
 *     sc = f_e(s) exp(-dw[nt] s^2) vol with f_e(s) = sum_j a_j exp(-b_j s^2)
 *     (scat_ampl_rec_DT_DW, :3-11), s = sin(theta)/lambda;
 *   msx_create_pots_rs_dt: CreatePots_rs_DT_DW (:64-81), sc = 4 pi sum_j
 *     a_j/(b_j + dw[nt]) exp(-4 pi^2 s^2/(b_j + dw[nt])) over the terms with
 *     b_j > 0 (scat_ampl_rs_DT_DW, :12-21), s = distance in Angstrom. */
MSX_API int msx_create_pots_rec_dt(int ntypes, int npoints, double dds, const int32_t* types, const double* dw,
                                   double vol, double* sc);
MSX_API int msx_create_pots_rs_dt(int ntypes, int npoints, double dds, const int32_t* types, const double* dw,
                                  double* sc);

/* The tabulated-curve forms of the two potentials above: Muslib
 * Potential_rec_curve_DW (Pot_rec_curve_DW.cpp:7-164) and
 * Potential_rs_curve_DW_cutpot / _periodic (Pot_RS.cpp:5-55, 209-349).  Each
 * atom has a type (0..ntypes-1) selecting a row of curves[ntypes][npoints],
 * sampled at spacing dds and interpolated linearly; nothing is added past the
 * last point.  Reciprocal space: the row is sampled at s = |g|/2 and already
 * holds the Debye-Waller factor and 1/area (MSProject.cs:850-863); real space:
 * sampled at r.  Output and scaling as msx_potential_rec / msx_potential_rs. */
MSX_API int msx_potential_rec_curve(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, const int32_t* type,
                                    const double* curves, int ntypes, int npoints, double dds, const double* ads,
                                    double a, double b, double c, double gam, int nmult, int mmult, int nx, int ny,
                                    int nslices, void* V, void* Vabs, int prec, int where);
MSX_API int msx_potential_rs_curve(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, const int32_t* type,
                                   const double* curves, int ntypes, int npoints, double dds, const double* ads,
                                   double a, double b, double c, double gam, double atrad, int periodic, int nx,
                                   int ny, int nslices, void* V, void* Vabs, int prec, int where);

/* ---- transmission ----------------------------------------------------- */
/* Transmission functions of nlayer potential layers (mulslice.cpp:414-421 and
 * the anti-aliasing pass :515-522; LMultiSlice phaseObject kernel.cu:4-18):
 *   t = exp(i*sigma(kev)*V/1000), then if bw > 0: fft, zero every pixel with
 *   kx2+ky2 >= k2max, ifft; k2max = (bw*min(nx/2ax, ny/2by))^2 in float and
 *   kx2, ky2 the float tables of freqn (slicelib.cpp:433-450).
 * V: nlayer real images; t: nlayer complex images. */
MSX_API int msx_transmission(msx_ctx ctx, const void* V, int nlayer, int nx, int ny, double ax, double by,
                             double kev, double bw, void* t, int prec, int where);
/* As msx_transmission with an absorptive potential (Muslib phase grating,
 * MSProjectReuse.cs:438-457): t = exp(i sigma V/1000) exp(-sigma Vabs/1000)
 * before the band limit.  Vabs may be NULL. */
MSX_API int msx_transmission_abs(msx_ctx ctx, const void* V, const void* Vabs, int nlayer, int nx, int ny,
                                 double ax, double by, double kev, double bw, void* t, int prec, int where);

/* ---- slice stack -------------------------------------------------------- */
/* The specimen/optics description of one multislice run: the layers, their
 * thicknesses and the stacking sequence of mulslice.cpp:190-206 (parlay). */
typedef struct msx_stack {
    int32_t nx, ny;          /* grid */
    double ax, by;           /* supercell size (Angstrom) */
    double kev;              /* beam energy */
    double ctilt_x, ctilt_y; /* crystal tilt (rad), propagator term of mulslice.cpp:494-513 */
    double bw;               /* band-limit fraction of Nyquist (Kirkland BW = 2/3) */
    int32_t nlayer;          /* number of distinct transmission layers */
    const double* cz;        /* host array [nlayer]: layer thickness (Angstrom) */
    int32_t nslice;          /* number of slices */
    const int32_t* layer;    /* host array [nslice]: layer index of each slice */
    /* angle between the cell axes a (x) and b (rad); 0 or pi/2: an orthogonal
     * cell.  Otherwise (an oblique cell: the slice loop of msx_multislice;
     * the other entry points refuse it) the propagator is Musli's
     * (Propagator.cpp:55-107, in this engine's sign convention):
     *   P = exp(-i pi dz (lambda s - 2 (x tan(tilt_x) + y tan(tilt_y))))
     * with x = fx/(ax sin gam), y = fy/(by sin gam) for the frequency indices
     * fx, fy of freqn, s = x^2 + y^2 - 2 x y cos(gam), all in FP64, and P = 0
     * where s >= (bw min(nx/(2 ax), ny/(2 by)))^2 (bw = 1: Musli's full
     * Nyquist disk). */
    double gam;
} msx_stack;

/* ---- multislice (HRTEM / any incident waves) ------------------------------ */
/* The coherent slice loop of mulslice.cpp:703-739 (ModelSimulated::
 * imageCalculation, ModelSimulated.cu:52-119, without its k^4/fftshift/1/nx
 * quirks): for each slice: wave *= t[layer]; fft; multiply by the Fresnel
 * propagator exp(-i pi dz (lambda k^2 - k.2tan(tilt))) inside the band limit
 * (strict kx2+ky2 < k2max), zero outside; ifft.
 * trans: nlayer complex images (host or device per `where`).
 * psi: nwave complex images, incident waves in, exit waves out.  If psi0 is
 * not NULL it is the incident wave for every wave (one image, e.g. a plane
 * wave) and psi is output only. */
MSX_API int msx_multislice(msx_ctx ctx, const msx_stack* stack, const void* trans, const void* psi0,
                           void* psi, int nwave, int prec, int where);

/* What msx_multislice_series records while the slice loop runs: the thickness
 * series of mulslice.cpp's coherent loop (mulslice.cpp:682-737: the "(real,
 * imag)" values of chosen beams after every slice, and the integrated
 * intensity it prints after every slice) and exit waves at intermediate
 * thicknesses (the checkpoints a run is continued from: mulslice.cpp:255-340
 * "start from previous result", Musli's last.wav/IsContinue,
 * MSProject.cs:423-440; pass such a wave back as psi0 to continue).  Every
 * output lives where `where` says (host or device), C being the complex type
 * of prec. */
typedef struct msx_series {
    const int32_t* at;  /* host [nat]: slice counts (1..nslice, strictly ascending) after which */
    int32_t nat;        /*   the exit wave is kept */
    void* waves;        /* [nat][nwave][nx][ny] C: exit waves after slice at[i] (NULL if nat = 0) */
    const int32_t* hk;  /* host [nbeams][2]: beam indices (h, k) of kx and ky; as mulslice.cpp:693-700 */
    int32_t nbeams;     /*   a negative index counts from the end (n + h), then clamps to [0, n-1] */
    void* beams;        /* [nslice][nwave][nbeams] C: after slice s (propagated, before the inverse
                         * transform) the wave's k-space value at (h, k) times 1/(nx ny), i.e. the
                         * 2D transform (exponent +i) of the exit wave of slice s over nx ny;
                         * exactly 0 outside the band limit of an orthogonal cell */
    double* intensity;  /* [nslice][nwave] or NULL: sum |psi|^2 / (nx ny) of the exit wave of each
                         * slice (mulslice.cpp:727-737) */
} msx_series;
/* msx_multislice that also fills `series` (maximum 4096 beams).  The exit
 * waves psi are those msx_multislice returns, bit for bit; a kept wave at
 * slice count nslice equals psi. */
MSX_API int msx_multislice_series(msx_ctx ctx, const msx_stack* stack, const void* trans, const void* psi0,
                                  void* psi, int nwave, int prec, int where, const msx_series* series);

/* Size of the fixed device array behind msx_multislice_noalloc. */
#define MSX_NOALLOC_SCRATCH_BYTES (4u << 20)
/* msx_multislice for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * No counterpart in the reference.  Device buffers only (trans, psi0, psi as
 * msx_multislice with MSX_DEVICE); the call allocates nothing (no hipMalloc,
 * no hipFree, no graph), launches the same kernels with the same tables as
 * msx_multislice, and so returns the same exit waves.  The slice tables
 * (freqn k^2 tables, the propagator factors of every layer) and the FFT
 * twiddles go into one array of MSX_NOALLOC_SCRATCH_BYTES bytes per GPU, a
 * static device array that each GPU the library runs on holds until the
 * process ends.  It needs, with C the complex type of prec and A(b) = b
 * rounded up to 256 bytes,
 *   A(4 nx) + A(4 ny) + A(sizeof(C) nlayer nx) + A(sizeof(C) nlayer ny + 1)
 *     + A(sizeof(C) nx) + A(sizeof(C) ny)  <=  MSX_NOALLOC_SCRATCH_BYTES
 * (510 layers at 512^2 FP32, 126 at 2048^2 FP32, 62 at 2048^2 FP64);
 * larger stacks are refused with MSX_ESIZE before anything runs.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * threads, contexts or streams) race.  The call copies its tables, waits for
 * the copies, launches, and returns when the context stream has finished. */
MSX_API int msx_multislice_noalloc(msx_ctx ctx, const msx_stack* stack, const void* trans, const void* psi0,
                                   void* psi, int nwave, int prec);

/* Many independent HRTEM jobs of uneven sizes and thicknesses in as few
 * launches as possible (the reference ran one job per process,
 * statistics/getStatistics.py:34 -> LMultiSlice main).  Each job is a full
 * msx_multislice of one wave, with the same exit wave bit for bit.  Jobs
 * with the same grid, cell and band limit advance together whatever their
 * kV, tilt and slice thicknesses: one row and one column sweep per slice for
 * all of their waves still running, each wave with its own transmission and
 * propagator, the column sweep's persistent work-groups taking (wave, column
 * group) units from device-side work counters (one per XCD, so that
 * neighbouring column groups stay on one L2).  A tuned grid's FFT plans
 * apply to the batch as to msx_multislice.  Device buffers only (where
 * = MSX_DEVICE); psi0 may be NULL (psi holds the incident wave). */
typedef struct msx_job {
    const msx_stack* stack;
    const void* trans;  /* [nlayer][nx][ny] complex */
    const void* psi0;   /* incident wave [nx][ny], or NULL */
    void* psi;          /* exit wave out (and incident wave in when psi0 is NULL) */
} msx_job;
MSX_API int msx_multislice_batch(msx_ctx ctx, const msx_job* jobs, int njobs, int prec, int where);

/* ---- STEM --------------------------------------------------------------- */
/* Probe of Musli ConvergedBeam/Xi (MSProject.cs:458-514) in Kirkland's
 * conventions: Psi0(k) = A exp(-i chi(k)) exp(+2 pi i k.r_p) for
 * kmin2 <= kx2+ky2 <= kmax2 (kmax = alpha_max/lambda), chi = pi lambda k^2
 * (0.5 Cs lambda^2 k^2 - df) (slicelib.cpp:258-318), normalised to
 * sum |psi0(x)|^2 = 1.  Cs and df in Angstrom. */
typedef struct msx_probe {
    double alpha_max, alpha_min; /* aperture (rad) */
    double cs, df;               /* spherical aberration, defocus (Angstrom) */
} msx_probe;
/* Annular detector (SaveSTEM, MSProject.cs:792-842): k_inner <= |k| <=
 * k_outer in 1/Angstrom (theta = lambda*k). */
typedef struct msx_detector {
    double k_inner, k_outer;
} msx_detector;

/* STEM scan (MSProjectUseOnce.StartCalculations / MSUseOnceCalculator.
 * RunCalculation, MSProjectUseOnce.cs:100-291, MSUseOnceCalculator.cs:
 * 125-277): every probe position runs the slice loop above; after the
 * propagator of each checkpoint slice the detector sums
 *   out[ck][d][p] += weight * sum_{annulus d} |Psi(k)|^2 / (nx*ny)
 * are taken (the fraction of the incident intensity; Musli sums unnormalised
 * values, MSProject.cs:834-838).  Probes run `batch` at a time (0 = choose).
 * pos: host [nprobe][2] (x, y) in Angstrom; checkpoints: host [nck] slice
 * counts 1..nslice, increasing; out: host [nck][ndet][nprobe] doubles,
 * overwritten if accumulate == 0, else added to.  trans per `where`. */
MSX_API int msx_stem(msx_ctx ctx, const msx_stack* stack, const void* trans, const msx_probe* probe,
                     const double* pos, int nprobe, const msx_detector* det, int ndet,
                     const int32_t* checkpoints, int nck, double weight, int accumulate, double* out,
                     int batch, int prec, int where);

/* ---- objective-lens image ------------------------------------------------ */
/* Image of exit waves through the objective lens (mulslice.cpp:612-656 at one
 * defocus; LMultiSlice objectLens kernel.cu:49-64 with the Kirkland chi):
 * Psi = fft(psi); Psi *= exp(-i chi(k)) for kx2+ky2 <= (aobj/lambda)^2, else
 * 0; psi_img = ifft(Psi).  psi_img (complex) and/or intensity (real |psi_img|^2)
 * may be NULL.  Cs and df in Angstrom, aobj in rad. */
MSX_API int msx_ctf_image(msx_ctx ctx, const void* psi, int nwave, int nx, int ny, double ax, double by,
                          double kev, double cs, double df, double aobj, void* psi_img, void* intensity,
                          int prec, int where);

/* ---- partial-coherence TEM image ------------------------------------------ */
/* Objective lens with the astigmatism terms of mulslice.cpp:620-628:
 * chi = pi lambda k^2 (0.5 Cs lambda^2 k^2 - df + dfa2 sin(2(phi - dfa2phi))
 *       + 2 dfa3 lambda |k| sin(3(phi - dfa3phi))/3), phi = atan2(ky, kx);
 * Cs, df, dfa2, dfa3 in Angstrom, angles and aobj in rad. */
typedef struct msx_lens {
    double cs, df, aobj;
    double dfa2, dfa2phi;
    double dfa3, dfa3phi;
} msx_lens;

/* Replaces Kirkland mulslice's partial-coherence branch (mulslice.cpp:
 * 539-676, lpartl == 1): the image integrated over illumination angles
 * (acmin..acmax, rad) and a Gaussian defocus spread (sigmaf, sampled every
 * dfdelt over +-2.5 sigmaf; Angstrom).  Every angle (iqx/ax, iqy/by) of the
 * grid inside the cone starts a wave exp(2 pi i (qx xpos + qy ypos)) (freqn
 * positions, xpos = ix ax/(nx-1)); the waves run `batch` at a time (0 =
 * choose) through the slice loop of msx_multislice; each exit wave is imaged
 * through the lens at every defocus step and the weighted intensities are
 * averaged:  image = sum_angles sum_steps w |psi_img|^2 / (nangles sum w).
 * The angle set and step count follow the reference's float arithmetic.
 * sigmaf == 0 images at df alone (the reference divides 0 by 0 there).
 * image: nx*ny doubles (host or device per `where`, like trans).  nangles
 * and nsteps (may be NULL) receive the number of angles and defocus steps. */
MSX_API int msx_partial_coherence(msx_ctx ctx, const msx_stack* stack, const void* trans, const msx_lens* lens,
                                  double acmin, double acmax, double sigmaf, double dfdelt, int batch,
                                  double* image, int32_t* nangles, int32_t* nsteps, int prec, int where);

/* ---- frozen-phonon surrogate (no reference counterpart) --------------------
 * A U-Net that maps the mean ADF image of a few frozen-phonon configurations
 * to the converged many-configuration mean; trained and run on the GPU by the
 * library's own BF16 matrix-core convolution kernels (FP32 accumulation).
 * Architecture: `levels` resolutions with base_channels << l channels, two
 * 3x3 conv+ReLU per level, 2x2 average pooling, nearest upsampling with skip
 * concatenation, a 1x1 head and a residual connection (output = input + head;
 * with several input channels, their mean + head).  Input images are FP32,
 * n x in_channels x h x w (one channel per frozen-phonon configuration, or one
 * mean image: msx_unet_create), outputs and targets n x h x w; h and w
 * multiples of 2^(levels-1).  Parameters are one packed FP32 vector: conv by conv (encoder
 * levels 0..L-1 two each, decoder levels L-2..0 two each, head), each as
 * W[Cout][Cin][k*k] followed by bias[Cout].  Loss: mean squared error. */
typedef struct msx_unet_s* msx_unet;
MSX_API int msx_unet_create(msx_ctx ctx, int levels, int base_channels, uint64_t seed, msx_unet* out);
/* a network whose first convolution takes in_channels (1..64) images */
MSX_API int msx_unet_create_in(msx_ctx ctx, int levels, int base_channels, int in_channels, uint64_t seed,
                               msx_unet* out);
MSX_API int msx_unet_in_channels(msx_unet net);
MSX_API int msx_unet_destroy(msx_unet net);
MSX_API int64_t msx_unet_param_count(msx_unet net);
MSX_API int msx_unet_params(msx_unet net, float* dst, int where);
/* replaces the parameters and resets the optimiser state */
MSX_API int msx_unet_set_params(msx_unet net, const float* src, int where);
MSX_API int msx_unet_forward(msx_unet net, const float* x, float* y, int n, int h, int w, int where);
/* loss (may be NULL) and its gradient w.r.t. the packed parameters (grad may be NULL) */
MSX_API int msx_unet_gradient(msx_unet net, const float* x, const float* target, int n, int h, int w,
                              double* loss, float* grad, int where);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch; loss may be NULL */
MSX_API int msx_unet_train_step(msx_unet net, const float* x, const float* target, int n, int h, int w, float lr,
                                double* loss, int where);
/* One convolution of the network's kernels on device buffers (stride 1, zero
 * padding ks/2, ks = 1 or 3, NCHW FP32, BF16 operands, FP32 accumulation):
 *   mode 0: out[n,co] = relu?(sum W[co,ci,tap] x[n,ci,p+off(tap)] + bias[co])
 *   mode 1: out[n,ci] = sum W[co,ci,tap] G[n,co,p-off(tap)]           (input gradient)
 *   mode 2: out[co,ci,tap] += sum_{n,p} G[n,co,p] x[n,ci,p+off(tap)]  (weight gradient)
 * with G = g masked by (y > 0) when relu is set (y = the forward output). */
MSX_API int msx_unet_conv(msx_ctx ctx, int mode, int ks, const float* x, const float* w, const float* bias,
                          const float* g, const float* y, float* out, int n, int cin, int cout, int h, int wd,
                          int relu);

#ifdef __cplusplus
}
#endif

#endif /* MSX_H */
