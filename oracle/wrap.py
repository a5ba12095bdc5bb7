"""ctypes wrappers over liboracle.so and the oracle/_ref reference builds.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  All images use the
layout of the reference's cfpix (x-major, y fastest): a complex image is a
numpy array of shape (nx, ny).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None
_fftw2 = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    """Load liboracle.so (built by `make -C oracle`)."""
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so not built (run make -C oracle)")
        L = C.CDLL(path)
        d, i, f, l = C.c_double, C.c_int, C.c_float, C.c_long
        for name, res, args in [
            ("orc_wavelength", d, [d]), ("orc_sigma", d, [d]),
            ("orc_bessi0", d, [d]), ("orc_bessk0", d, [d]),
            ("orc_vzatom", d, [i, d, i, i]), ("orc_featom", d, [i, d, i]),
            ("orc_fparam", d, [i, i, i]),
            ("orc_freqn", None, [_fp, _fp, _fp, i, d]),
            ("orc_k2max", f, [i, i, d, d, d]),
            ("orc_dft2d_f64", None, [_dp, i, i, i]),
            ("orc_fft2d_f64", None, [_dp, i, i, i]),
            ("orc_fft2d_f32", None, [_fp, i, i, i]),
            ("orc_potential", i, [_ip, _dp, C.c_void_p, l, d, d, d, i, i, i, d, i, i, _dp]),
            ("orc_transmission_f32", None, [_fp, i, i, i, d, d, d, d, _fp]),
            ("orc_transmission_f64", None, [_dp, i, i, i, d, d, d, d, _dp]),
            ("orc_propagator_f64", None, [i, i, d, d, d, d, d, d, _dp, _dp, _dp, _dp]),
            ("orc_multislice_f32", i, [_fp, i, _fp, _ip, i, i, i, d, d, d, d, d, d, _fp, C.c_void_p]),
            ("orc_multislice_f64", i, [_dp, i, _dp, _ip, i, i, i, d, d, d, d, d, d, _dp, C.c_void_p]),
            ("orc_multislice_oblique_f64", i, [_dp, i, _dp, _ip, i, i, i, d, d, d, d, d, d, d, _dp]),
            ("orc_multislice_series_f64", i, [_dp, i, _dp, _ip, i, i, i, d, d, d, d, d, d, _dp, C.c_void_p, _ip, i,
                                              _dp]),
            ("orc_probe_f64", i, [i, i, d, d, d, d, d, d, d, d, d, _dp]),
            ("orc_stem_f64", i, [_dp, i, _dp, _ip, i, i, i, d, d, d, d, d, d, d, d, _dp, i, _dp, i, _ip, i, _dp, i]),
            ("orc_stem_f32", i, [_fp, i, _fp, _ip, i, i, i, d, d, d, d, d, d, d, d, _dp, i, _dp, i, _ip, i, _dp, i]),
            ("orc_ctf_image_f64", None, [_dp, i, i, d, d, d, d, d, d, C.c_void_p, C.c_void_p]),
            ("orc_dt_fe", d, [i, d]), ("orc_dt_rs", d, [i, d, d]),
            ("orc_potential_rec_dt", None, [d, d, d, i, i, i, i, l, _dp, _dp, _dp, _dp, _dp, _ip, _dp]),
            ("orc_dt_rs2", d, [i, d, d]),
            ("orc_potential_rs_dt", None, [d, d, d, d, i, i, l, _dp, _dp, _dp, _dp, _dp, _ip, i, _dp]),
            ("orc_potential_rec_curve", None, [d, d, d, i, i, i, i, l, _dp, _dp, _dp, _dp, _ip, _dp, i, d, _dp]),
            ("orc_potential_rs_curve", None, [d, d, d, d, i, i, l, _dp, _dp, _dp, _dp, _ip, _dp, i, d, i, _dp]),
            ("orc_illum_angles", i, [i, i, d, d, d, d, d, C.c_void_p, i]),
            ("orc_defocus_steps", i, [d, d, C.c_void_p, C.c_void_p, i]),
            ("orc_chi_astig", d, [d, d, d, d, d, d, d, d, d]),
            ("orc_partial_image_f64", i, [_dp, i, _dp, _ip, i, i, i, d, d, d, d, d, d, d, d, d, d, d, d, d,
                                          d, d, d, d, _dp]),
        ]:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ref():
    """Load oracle/_ref/libkirkland_ref.so (reference slicelib.cpp), or None."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libkirkland_ref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        d, i = C.c_double, C.c_int
        for name, res, args in [
            ("ref_wavelength", d, [d]), ("ref_sigma", d, [d]),
            ("ref_bessi0", d, [d]), ("ref_bessk0", d, [d]),
            ("ref_vzatom", d, [i, d]), ("ref_featom", d, [i, d]),
            ("ref_freqn", None, [_fp, _fp, _fp, i, d]),
            ("ref_chi_csdf", d, [d, d, d, d]),
            ("ref_chi_multi", d, [d, d, d, d, d, d, d, d]),
            ("ref_parlay", i, [C.c_char_p, _ip, i, i, C.POINTER(C.c_int)]),
            ("ref_tiff_write", i, [C.c_char_p, _fp, i, i, i, _fp, i, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_float]),
            ("ref_tiff_read", i, [C.c_char_p, _fp, C.c_long, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), _fp]),
            ("ref_rangauss", d, [C.POINTER(C.c_ulong)]),
            ("ref_ranflat", d, [C.POINTER(C.c_ulong)]),
        ]:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _ref = L
    return _ref


def ref_fftw2():
    """Load oracle/_ref/libfftw2f_ref.so (reference-vendored FFTW 2.1.5), or None."""
    global _fftw2
    if _fftw2 is None:
        path = os.path.join(_HERE, "_ref", "libfftw2f_ref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_fftw2_2d.restype = C.c_int
        L.ref_fftw2_2d.argtypes = [_fp, C.c_int, C.c_int, C.c_int]
        _fftw2 = L
    return _fftw2


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.complex64)


class Oracle:
    """numpy-level interface to the oracle (all static)."""

    @staticmethod
    def wavelength(kev):
        return lib().orc_wavelength(kev)

    @staticmethod
    def sigma(kev):
        return lib().orc_sigma(kev)

    @staticmethod
    def vzatom(Z, r, table=1, formula=0):
        return lib().orc_vzatom(int(Z), float(r), int(table), int(formula))

    @staticmethod
    def featom(Z, k2, table=1):
        return lib().orc_featom(int(Z), float(k2), int(table))

    @staticmethod
    def freqn(n, a):
        ko = np.zeros(n, np.float32)
        ko2 = np.zeros(n, np.float32)
        xo = np.zeros(n, np.float32)
        lib().orc_freqn(ko, ko2, xo, n, a)
        return ko, ko2, xo

    @staticmethod
    def k2max(nx, ny, ax, by, bw=2.0 / 3.0):
        return lib().orc_k2max(nx, ny, ax, by, bw)

    @staticmethod
    def dft2(x, sign):
        a = _c128(x).copy()
        lib().orc_dft2d_f64(a.view(np.float64).reshape(-1), a.shape[0], a.shape[1], sign)
        return a

    @staticmethod
    def fft2(x, sign, f32=False):
        if f32:
            a = _c64(x).copy()
            lib().orc_fft2d_f32(a.view(np.float32).reshape(-1), a.shape[0], a.shape[1], sign)
        else:
            a = _c128(x).copy()
            lib().orc_fft2d_f64(a.view(np.float64).reshape(-1), a.shape[0], a.shape[1], sign)
        return a

    @staticmethod
    def potential(Z, xyz, cell, nx, ny, nslices, radius, table=0, formula=0, occ=None):
        Z = np.ascontiguousarray(Z, np.int32)
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        V = np.zeros((nslices, nx, ny), np.float64)
        occ_p = None
        if occ is not None:
            occ = np.ascontiguousarray(occ, np.float64)
            occ_p = occ.ctypes.data
        rc = lib().orc_potential(Z, xyz, occ_p, len(Z), cell[0], cell[1], cell[2], nx, ny,
                                 nslices, radius, table, formula, V.reshape(-1))
        if rc != 0:
            raise ValueError("orc_potential failed %d" % rc)
        return V

    @staticmethod
    def transmission(V, ax, by, kev, bw=2.0 / 3.0, f32=False):
        V = np.asarray(V)
        nl, nx, ny = V.shape
        if f32:
            t = np.zeros((nl, nx, ny), np.complex64)
            lib().orc_transmission_f32(np.ascontiguousarray(V, np.float32).reshape(-1), nl, nx, ny,
                                       ax, by, kev, bw, t.view(np.float32).reshape(-1))
        else:
            t = np.zeros((nl, nx, ny), np.complex128)
            lib().orc_transmission_f64(np.ascontiguousarray(V, np.float64).reshape(-1), nl, nx, ny,
                                       ax, by, kev, bw, t.view(np.float64).reshape(-1))
        return t

    @staticmethod
    def propagator(nx, ny, ax, by, kev, cz, tilt=(0.0, 0.0)):
        pxr, pxi = np.zeros(nx), np.zeros(nx)
        pyr, pyi = np.zeros(ny), np.zeros(ny)
        lib().orc_propagator_f64(nx, ny, ax, by, kev, cz, tilt[0], tilt[1], pxr, pxi, pyr, pyi)
        return pxr + 1j * pxi, pyr + 1j * pyi

    @staticmethod
    def multislice(trans, cz, layer, wave, ax, by, kev, tilt=(0.0, 0.0), bw=2.0 / 3.0,
                   f32=False, want_intensity=False):
        trans = np.asarray(trans)
        nl, nx, ny = trans.shape
        layer = np.ascontiguousarray(layer, np.int32)
        inten = np.zeros(len(layer), np.float64)
        ip = inten.ctypes.data if want_intensity else None
        if f32:
            w = _c64(wave).copy()
            rc = lib().orc_multislice_f32(_c64(trans).view(np.float32).reshape(-1), nl,
                                          np.ascontiguousarray(cz, np.float32), layer, len(layer),
                                          nx, ny, ax, by, kev, tilt[0], tilt[1], bw,
                                          w.view(np.float32).reshape(-1), ip)
        else:
            w = _c128(wave).copy()
            rc = lib().orc_multislice_f64(_c128(trans).view(np.float64).reshape(-1), nl,
                                          np.ascontiguousarray(cz, np.float64), layer, len(layer),
                                          nx, ny, ax, by, kev, tilt[0], tilt[1], bw,
                                          w.view(np.float64).reshape(-1), ip)
        if rc != 0:
            raise ValueError("orc_multislice failed %d" % rc)
        return (w, inten) if want_intensity else w

    @staticmethod
    def multislice_series(trans, cz, layer, wave, ax, by, kev, hk, tilt=(0.0, 0.0), bw=2.0 / 3.0):
        """FP64 slice loop recording, after every slice, the beams hk [(h, k)]
        (mulslice.cpp:682-724) and sum |wave|^2; returns (exit wave, beams
        [nslice][nbeams] complex, intensity [nslice])."""
        trans = np.asarray(trans)
        nl, nx, ny = trans.shape
        layer = np.ascontiguousarray(layer, np.int32)
        hk = np.ascontiguousarray(np.asarray(hk, np.int32).reshape(-1, 2))
        beams = np.zeros((len(layer), len(hk)), np.complex128)
        inten = np.zeros(len(layer), np.float64)
        w = _c128(wave).copy()
        rc = lib().orc_multislice_series_f64(_c128(trans).view(np.float64).reshape(-1), nl,
                                             np.ascontiguousarray(cz, np.float64), layer, len(layer),
                                             nx, ny, ax, by, kev, tilt[0], tilt[1], bw,
                                             w.view(np.float64).reshape(-1), inten.ctypes.data, hk.reshape(-1),
                                             len(hk), beams.view(np.float64).reshape(-1))
        if rc != 0:
            raise ValueError("orc_multislice_series failed %d" % rc)
        return w, beams, inten

    @staticmethod
    def multislice_oblique(trans, cz, layer, wave, ax, by, gam, kev, tilt=(0.0, 0.0), bw=2.0 / 3.0):
        """Slice loop of an oblique cell (angle gam between a and b): Musli's
        propagator (Propagator.cpp:55-107) in Kirkland's convention, FP64."""
        trans = np.asarray(trans)
        nl, nx, ny = trans.shape
        layer = np.ascontiguousarray(layer, np.int32)
        w = _c128(wave).copy()
        rc = lib().orc_multislice_oblique_f64(_c128(trans).view(np.float64).reshape(-1), nl,
                                              np.ascontiguousarray(cz, np.float64), layer, len(layer), nx, ny,
                                              ax, by, gam, kev, tilt[0], tilt[1], bw,
                                              w.view(np.float64).reshape(-1))
        if rc != 0:
            raise ValueError("orc_multislice_oblique failed %d" % rc)
        return w

    @staticmethod
    def probe(nx, ny, ax, by, kev, alpha_max, alpha_min=0.0, cs=0.0, df=0.0, pos=(0.0, 0.0)):
        w = np.zeros((nx, ny), np.complex128)
        rc = lib().orc_probe_f64(nx, ny, ax, by, kev, alpha_max, alpha_min, cs, df, pos[0], pos[1],
                                 w.view(np.float64).reshape(-1))
        if rc != 0:
            raise ValueError("orc_probe failed %d" % rc)
        return w

    @staticmethod
    def stem(trans, cz, layer, ax, by, kev, alpha_max, positions, detectors, checkpoints,
             alpha_min=0.0, cs=0.0, df=0.0, bw=2.0 / 3.0, nthreads=1, f32=False):
        trans = np.asarray(trans)
        nl, nx, ny = trans.shape
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
        det = np.ascontiguousarray(detectors, np.float64).reshape(-1)
        ck = np.ascontiguousarray(checkpoints, np.int32)
        npr, nd, nck = len(pos) // 2, len(det) // 2, len(ck)
        out = np.zeros((nck, nd, npr), np.float64)
        layer = np.ascontiguousarray(layer, np.int32)
        if f32:
            rc = lib().orc_stem_f32(_c64(trans).view(np.float32).reshape(-1), nl,
                                    np.ascontiguousarray(cz, np.float32), layer, len(layer), nx, ny,
                                    ax, by, kev, bw, alpha_max, alpha_min, cs, df, pos, npr, det, nd,
                                    ck, nck, out.reshape(-1), nthreads)
        else:
            rc = lib().orc_stem_f64(_c128(trans).view(np.float64).reshape(-1), nl,
                                    np.ascontiguousarray(cz, np.float64), layer, len(layer), nx, ny,
                                    ax, by, kev, bw, alpha_max, alpha_min, cs, df, pos, npr, det, nd,
                                    ck, nck, out.reshape(-1), nthreads)
        if rc != 0:
            raise ValueError("orc_stem failed %d" % rc)
        return out

    @staticmethod
    def ctf_image(wave, ax, by, kev, cs, df, aobj):
        w = _c128(wave)
        nx, ny = w.shape
        img_w = np.zeros((nx, ny), np.complex128)
        img = np.zeros((nx, ny), np.float64)
        lib().orc_ctf_image_f64(w.view(np.float64).reshape(-1), nx, ny, ax, by, kev, cs, df, aobj,
                                img_w.ctypes.data, img.ctypes.data)
        return img_w, img

    @staticmethod
    def illum_angles(nx, ny, ax, by, kev, acmin, acmax):
        """(iqx, iqy) pairs of the partial-coherence illumination integral, in the reference's loop order."""
        n = lib().orc_illum_angles(nx, ny, ax, by, kev, acmin, acmax, None, 0)
        iq = np.zeros((max(n, 1), 2), np.int32)
        lib().orc_illum_angles(nx, ny, ax, by, kev, acmin, acmax, iq.ctypes.data, n)
        return iq[:n]

    @staticmethod
    def defocus_steps(sigmaf, dfdelt):
        """(offsets from df0, weights) of the defocus integral."""
        n = lib().orc_defocus_steps(sigmaf, dfdelt, None, None, 0)
        if n < 1:
            raise ValueError("dfdelt must be > 0")
        off = np.zeros(n)
        w = np.zeros(n)
        lib().orc_defocus_steps(sigmaf, dfdelt, off.ctypes.data, w.ctypes.data, n)
        return off, w

    @staticmethod
    def chi_astig(kx, ky, kev, cs, df, dfa2=0.0, dfa2phi=0.0, dfa3=0.0, dfa3phi=0.0):
        return lib().orc_chi_astig(kx, ky, kev, cs, df, dfa2, dfa2phi, dfa3, dfa3phi)

    @staticmethod
    def partial_image(trans, cz, layer, ax, by, kev, lens, acmin, acmax, sigmaf, dfdelt, ctilt=(0.0, 0.0), bw=2.0 / 3.0):
        """Partial-coherence image (FP64); lens: dict(cs, df, aobj, dfa2, dfa2phi, dfa3, dfa3phi).
        Returns (image, number of illumination angles)."""
        t = _c128(trans)
        nl, nx, ny = t.shape
        img = np.zeros((nx, ny), np.float64)
        g = lambda k: float(lens.get(k, 0.0))  # noqa: E731
        n = lib().orc_partial_image_f64(t.view(np.float64).reshape(-1), nl, np.ascontiguousarray(cz, np.float64),
                                        np.ascontiguousarray(layer, np.int32), len(layer), nx, ny, ax, by, kev,
                                        ctilt[0], ctilt[1], bw, g("cs"), g("df"), g("aobj"), g("dfa2"), g("dfa2phi"),
                                        g("dfa3"), g("dfa3phi"), acmin, acmax, sigmaf, dfdelt, img.reshape(-1))
        if n < 0:
            raise ValueError("orc_partial_image_f64 rejected its input")
        return img, n

    @staticmethod
    def dt_fe(Z, s):
        return lib().orc_dt_fe(int(Z), float(s))

    @staticmethod
    def dt_rs(Z, r, dw=0.0):
        return lib().orc_dt_rs(int(Z), float(r), float(dw))

    @staticmethod
    def potential_rec_g(Z, xf, yf, a, b, N, M, gam=np.pi / 2, nmult=1, mmult=1, dw=None, ads=None, occ=None):
        """Potential_rec_DT_DW Fourier coefficients of one slice in the
        reference's layout: complex array [M][N] (x fastest).  xf, yf are
        fractional coordinates."""
        n = len(Z)
        z = lambda v, d=0.0: np.ascontiguousarray(np.full(n, d) if v is None else v, np.float64)  # noqa: E731
        g = np.zeros((M, N), np.complex128)
        lib().orc_potential_rec_dt(a, b, gam, nmult, mmult, N, M, n, z(xf), z(yf), z(dw), z(ads), z(occ, 1.0),
                                   np.ascontiguousarray(Z, np.int32), g.view(np.float64).reshape(-1))
        return g

    @staticmethod
    def potential_rs_g(Z, xf, yf, a, b, N, M, atrad, periodic=True, gam=np.pi / 2, dw=None, ads=None, occ=None):
        """Potential_rs_DT_DW_{periodic,cutpot} of one slice in the reference's
        layout: complex array [M][N] (x fastest), Muslib units."""
        n = len(Z)
        z = lambda v, d=0.0: np.ascontiguousarray(np.full(n, d) if v is None else v, np.float64)  # noqa: E731
        g = np.zeros((M, N), np.complex128)
        lib().orc_potential_rs_dt(a, b, gam, atrad, N, M, n, z(xf), z(yf), z(dw), z(ads), z(occ, 1.0),
                                  np.ascontiguousarray(Z, np.int32), int(periodic), g.view(np.float64).reshape(-1))
        return g

    @staticmethod
    def dt_curves_rec(Zs, dws, npoints, dds, area):
        """CreatePots_rec_DT_DW (Curves.cpp:47-62): per type f_e(s) exp(-dw s^2) / area at s = i dds."""
        s = np.arange(npoints) * dds
        return np.stack([np.array([Oracle.dt_fe(z, v) for v in s]) * np.exp(-dw * s * s) / area
                         for z, dw in zip(Zs, dws)])

    @staticmethod
    def dt_curves_rs(Zs, dws, npoints, dds):
        """CreatePots_rs_DT_DW (Curves.cpp:63-81): per type the real-space Gaussian sum at r = i dds."""
        r = np.arange(npoints) * dds
        return np.stack([np.array([Oracle.dt_rs(z, v, dw) for v in r]) for z, dw in zip(Zs, dws)])

    @staticmethod
    def potential_rec_curve_g(types, xf, yf, a, b, N, M, curves, dds, gam=np.pi / 2, nmult=1, mmult=1, ads=None,
                              occ=None):
        n = len(types)
        z = lambda v, d=0.0: np.ascontiguousarray(np.full(n, d) if v is None else v, np.float64)  # noqa: E731
        curves = np.ascontiguousarray(curves, np.float64)
        g = np.zeros((M, N), np.complex128)
        lib().orc_potential_rec_curve(a, b, gam, nmult, mmult, N, M, n, z(xf), z(yf), z(ads), z(occ, 1.0),
                                      np.ascontiguousarray(types, np.int32), curves.reshape(-1), curves.shape[1],
                                      dds, g.view(np.float64).reshape(-1))
        return g

    @staticmethod
    def potential_rs_curve_g(types, xf, yf, a, b, N, M, atrad, curves, dds, periodic=True, gam=np.pi / 2,
                             ads=None, occ=None):
        n = len(types)
        z = lambda v, d=0.0: np.ascontiguousarray(np.full(n, d) if v is None else v, np.float64)  # noqa: E731
        curves = np.ascontiguousarray(curves, np.float64)
        g = np.zeros((M, N), np.complex128)
        lib().orc_potential_rs_curve(a, b, gam, atrad, N, M, n, z(xf), z(yf), z(ads), z(occ, 1.0),
                                     np.ascontiguousarray(types, np.int32), curves.reshape(-1), curves.shape[1], dds,
                                     int(periodic), g.view(np.float64).reshape(-1))
        return g
