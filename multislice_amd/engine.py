"""ctypes binding of libmsx.so (include/msx.h) and the host-side mirror of the
reference's LMultiSlice call surface.

Functional API (numpy arrays = host buffers, torch CUDA tensors = device
buffers; complex images have shape (..., nx, ny), y contiguous):

    fft2d, potential, transmission, multislice, stem, ctf_image

Class API mirroring the reference (ImageSimulation/LMultiSlice):

    Microscope(keV, Cs, aperture, defocus)          Microscope.h:3-34
    ModelPotential(model, nx, ny, nz, dpa, radius)  ModelPotential.h:27-51
    ModelSimulated(potential, nx, ny, nz, dpa)      ModelSimulated.h:7-28
    PotentialBuilder                                PotentialBuilder.h:14-22

There is no CPU fallback: every call goes to the HIP library, and importing
this module fails if libmsx.so has not been built.
"""
import ctypes as C
import importlib.util
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MSX_LIB selects an alternative build of the same library (tuning variants)
LIB_PATH = os.environ.get("MSX_LIB") or os.path.join(_HERE, "libmsx.so")

F32, F64 = 0, 1
HOST, DEVICE = 0, 1
TABLE_LMULTISLICE, TABLE_KIRKLAND = 0, 1
FORMULA_KIRKLAND, FORMULA_LMULTISLICE = 0, 1
SIZES = (16, 32, 64, 128, 256, 384, 512, 768, 1024, 2048)
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ESIZE", -5: "ESTATE"}


class MsxError(RuntimeError):
    """A libmsx call returned a negative status."""

    def __init__(self, code, msg):
        super().__init__(f"msx error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Atom(C.Structure):
    _fields_ = [("Z", C.c_int32), ("occ", C.c_float), ("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


ATOM_DTYPE = np.dtype([("Z", np.int32), ("occ", np.float32), ("x", np.float64), ("y", np.float64),
                       ("z", np.float64)], align=True)


class Stack(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("ax", C.c_double), ("by", C.c_double),
                ("kev", C.c_double), ("ctilt_x", C.c_double), ("ctilt_y", C.c_double), ("bw", C.c_double),
                ("nlayer", C.c_int32), ("cz", C.POINTER(C.c_double)), ("nslice", C.c_int32),
                ("layer", C.POINTER(C.c_int32)), ("gam", C.c_double)]


class Series(C.Structure):
    _fields_ = [("at", C.c_void_p), ("nat", C.c_int32), ("waves", C.c_void_p), ("hk", C.c_void_p),
                ("nbeams", C.c_int32), ("beams", C.c_void_p), ("intensity", C.c_void_p)]


class Probe(C.Structure):
    _fields_ = [("alpha_max", C.c_double), ("alpha_min", C.c_double), ("cs", C.c_double), ("df", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("row_wpb", C.c_int32), ("col_wpb", C.c_int32), ("batch", C.c_int32), ("ms_per_slice", C.c_double),
                ("col_tile", C.c_int32), ("row_plan", C.c_int32), ("col_plan", C.c_int32)]


class Job(C.Structure):
    _fields_ = [("stack", C.c_void_p), ("trans", C.c_void_p), ("psi0", C.c_void_p), ("psi", C.c_void_p)]


class Detector(C.Structure):
    _fields_ = [("k_inner", C.c_double), ("k_outer", C.c_double)]


class Lens(C.Structure):
    _fields_ = [("cs", C.c_double), ("df", C.c_double), ("aobj", C.c_double), ("dfa2", C.c_double),
                ("dfa2phi", C.c_double), ("dfa3", C.c_double), ("dfa3phi", C.c_double)]


PROF_ROW, PROF_COL, PROF_SPLAT, PROF_DET = 0, 1, 2, 3

EXPORTS = ["msx_create", "msx_destroy", "msx_set_stream", "msx_wait_stream", "msx_set_graphs", "msx_graph_stats", "msx_autotune",
           "msx_set_tuning", "msx_last_error", "msx_version",
           "msx_profile", "msx_profile_read", "msx_copy_probe", "msx_displace", "msx_philox4x32_10",
           "msx_wavelength", "msx_sigma", "msx_vzatom", "msx_featom", "msx_fft2d", "msx_potential",
           "msx_transmission", "msx_multislice", "msx_stem", "msx_ctf_image",
           "msx_unet_create", "msx_unet_destroy", "msx_unet_param_count", "msx_unet_params",
           "msx_unet_set_params", "msx_unet_forward", "msx_unet_gradient", "msx_unet_train_step",
           "msx_unet_conv", "msx_potential_rec", "msx_transmission_abs", "msx_potential_rs",
           "msx_multislice_batch", "msx_potential_rec_curve", "msx_potential_rs_curve",
           "msx_partial_coherence", "msx_amorphous", "msx_parlay", "msx_tiff_write", "msx_tiff_read",
           "msx_musli_stem_write", "msx_musli_wave_write", "msx_potential_device_atoms", "msx_multislice_noalloc",
           "msx_create_pots_rec_dt", "msx_create_pots_rs_dt", "msx_multislice_series", "msx_unet_create_in",
           "msx_unet_in_channels"]

_lib = None
_lock = threading.Lock()


def _preload_hip_runtime():
    """Make the process use ONE HIP runtime.  PyTorch-ROCm bundles its own
    libamdhip64.so (same SONAME as /opt/rocm's); if libmsx.so pulled in the
    system copy first, a later torch import would bring a second runtime that
    cannot see the GPU.  Loading torch's copy globally first lets libmsx.so
    (NEEDED libamdhip64.so.7) and torch share it.  MSX_HIP_RUNTIME=system
    keeps the system runtime."""
    if os.environ.get("MSX_HIP_RUNTIME") == "system":
        return
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        C.CDLL(path, mode=C.RTLD_GLOBAL)


def load():
    """Load libmsx.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    _preload_hip_runtime()
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run python -m multislice_amd.build")
    L = C.CDLL(LIB_PATH)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    sig = {
        "msx_create": (i, [i, C.POINTER(vp)]), "msx_destroy": (i, [vp]), "msx_set_stream": (i, [vp, vp]),
        "msx_wait_stream": (i, [vp, vp]),
        "msx_set_graphs": (i, [vp, i]),
        "msx_autotune": (i, [vp, i, i, i, i, C.POINTER(Tuning)]), "msx_set_tuning": (i, [vp, i, i, i, C.POINTER(Tuning)]), "msx_graph_stats": (i, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "msx_last_error": (C.c_char_p, []), "msx_version": (i, []),
        "msx_profile": (i, [vp, i]),
        "msx_profile_read": (i, [vp, i, C.POINTER(C.c_int64), C.POINTER(d), C.POINTER(d)]),
        "msx_wavelength": (d, [d]), "msx_sigma": (d, [d]), "msx_vzatom": (d, [i, d, i, i]),
        "msx_featom": (d, [i, d, i]),
        "msx_displace": (i, [vp, vp, C.c_int64, vp, C.c_uint64, C.c_int32, vp, i]),
        "msx_philox4x32_10": (None, [vp, vp, vp]),
        "msx_fft2d": (i, [vp, vp, i, i, i, i, i, i]),
        "msx_potential": (i, [vp, vp, C.c_int64, d, d, d, i, i, i, d, i, i, vp, i, i]),
        "msx_potential_device_atoms": (i, [vp, vp, C.c_int64, d, d, d, i, i, i, d, i, i, vp, i, i]),
        "msx_transmission": (i, [vp, vp, i, i, i, d, d, d, d, vp, i, i]),
        "msx_multislice": (i, [vp, C.POINTER(Stack), vp, vp, vp, i, i, i]),
        "msx_multislice_noalloc": (i, [vp, C.POINTER(Stack), vp, vp, vp, i, i]),
        "msx_multislice_series": (i, [vp, C.POINTER(Stack), vp, vp, vp, i, i, i, C.POINTER(Series)]),
        "msx_create_pots_rec_dt": (i, [i, i, d, vp, vp, d, vp]),
        "msx_create_pots_rs_dt": (i, [i, i, d, vp, vp, vp]),
        "msx_stem": (i, [vp, C.POINTER(Stack), vp, C.POINTER(Probe), vp, i, C.POINTER(Detector), i, vp, i, d, i,
                         vp, i, i, i]),
        "msx_ctf_image": (i, [vp, vp, i, i, i, d, d, d, d, d, d, vp, vp, i, i]),
        "msx_partial_coherence": (i, [vp, vp, vp, vp, d, d, d, d, i, vp, vp, vp, i, i]),
        "msx_amorphous": (i, [d, d, d, C.c_int32, C.c_int64, d, C.c_uint64, C.c_int64, vp, vp]),
        "msx_parlay": (i, [C.c_char_p, vp, i, i, vp]),
        "msx_tiff_write": (i, [C.c_char_p, vp, i, i, i, vp, i, C.c_float, C.c_float, C.c_float, C.c_float,
                               C.c_float, C.c_float]),
        "msx_tiff_read": (i, [C.c_char_p, vp, C.c_int64, vp, vp, vp, vp]),
        "msx_musli_stem_write": (i, [C.c_char_p, vp, i, i, d, d, d, d, d]),
        "msx_musli_wave_write": (i, [C.c_char_p, vp, i, i, d, d, d, d, d]),
        "msx_unet_create": (i, [vp, i, i, C.c_uint64, C.POINTER(vp)]), "msx_unet_destroy": (i, [vp]),
        "msx_unet_create_in": (i, [vp, i, i, i, C.c_uint64, C.POINTER(vp)]), "msx_unet_in_channels": (i, [vp]),
        "msx_unet_param_count": (C.c_int64, [vp]), "msx_unet_params": (i, [vp, vp, i]),
        "msx_unet_set_params": (i, [vp, vp, i]), "msx_unet_forward": (i, [vp, vp, vp, i, i, i, i]),
        "msx_unet_gradient": (i, [vp, vp, vp, i, i, i, C.POINTER(d), vp, i]),
        "msx_unet_train_step": (i, [vp, vp, vp, i, i, i, C.c_float, C.POINTER(d), i]),
        "msx_unet_conv": (i, [vp, i, i, vp, vp, vp, vp, vp, vp, i, i, i, i, i, i]),
        "msx_potential_rec": (i, [vp, vp, C.c_int64, vp, vp, d, d, d, d, i, i, i, i, i, vp, vp, i, i]),
        "msx_transmission_abs": (i, [vp, vp, vp, i, i, i, d, d, d, d, vp, i, i]),
        "msx_potential_rs": (i, [vp, vp, C.c_int64, vp, vp, d, d, d, d, d, i, i, i, i, vp, vp, i, i]),
        "msx_multislice_batch": (i, [vp, vp, i, i, i]),
        "msx_copy_probe": (i, [vp, i, vp, vp, i, i, i, C.POINTER(d)]),
        "msx_potential_rec_curve": (i, [vp, vp, C.c_int64, vp, vp, i, i, d, vp, d, d, d, d, i, i, i, i, i, vp, vp, i,
                                        i]),
        "msx_potential_rs_curve": (i, [vp, vp, C.c_int64, vp, vp, i, i, d, vp, d, d, d, d, d, i, i, i, i, vp, vp, i,
                                       i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise MsxError(rc, load().msx_last_error().decode(errors="replace"))


# ------------------------------------------------------------------ buffers
def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _buf(x, prec, complex_=True, writable=False):
    """(pointer, where, keepalive, host_array_or_None) for a numpy array or a
    torch CUDA tensor."""
    if _is_torch(x):
        if not x.is_cuda:
            raise ValueError("torch tensors must be on the GPU (use numpy for host buffers)")
        if not x.is_contiguous():
            raise ValueError("torch tensor must be contiguous")
        want = ({F32: "complex64", F64: "complex128"} if complex_ else {F32: "float32", F64: "float64"})[prec]
        if str(x.dtype).split(".")[-1] != want:
            raise ValueError(f"tensor dtype {x.dtype} != {want}")
        return C.c_void_p(x.data_ptr()), DEVICE, x, None
    dt = ({F32: np.complex64, F64: np.complex128} if complex_ else {F32: np.float32, F64: np.float64})[prec]
    a = np.asarray(x)
    if a.dtype != dt or not a.flags.c_contiguous:
        if writable:
            raise ValueError(f"output array must be C-contiguous {np.dtype(dt)}")
        a = np.ascontiguousarray(a, dtype=dt)
    return C.c_void_p(a.ctypes.data), HOST, a, a


def _prec_of(x):
    s = str(getattr(x, "dtype", ""))
    return F64 if ("complex128" in s or "float64" in s) else F32


def _ptr_where(args):
    wh = {w for (_, w, _, _) in args}
    if len(wh) > 1:
        raise ValueError("mix of host and device buffers in one call")
    return wh.pop()


def _need(cond, msg):
    """Argument check of a wrapper: raise ValueError before anything reaches
    the C ABI, whose kernels size their reads and writes from these shapes."""
    if not cond:
        raise ValueError(msg)


def _numel(x):
    return int(np.prod(tuple(x.shape))) if len(tuple(x.shape)) else 1


def _check_layers(nl, cz, layer):
    """cz: thickness per layer; layer: layer index per slice, each below both
    len(cz) and the number of transmission images nl."""
    cz = np.asarray(cz)
    layer = np.asarray(layer)
    _need(cz.ndim == 1 and len(cz) >= 1, "cz must be a non-empty 1-D sequence of slice thicknesses")
    _need(layer.ndim == 1 and len(layer) >= 1, "layer must be a non-empty 1-D sequence of layer indices")
    _need(np.issubdtype(layer.dtype, np.integer), "layer indices must be integers")
    top = min(int(nl), len(cz))
    _need(int(layer.min()) >= 0 and int(layer.max()) < top,
          f"layer indices must lie in [0, {top}) (transmission images {nl}, thicknesses {len(cz)})")


def _check_trans(trans):
    _need(len(tuple(trans.shape)) == 3, "trans must be (nlayer, nx, ny)")
    _need(trans.shape[0] >= 1, "trans holds no layer")
    return tuple(trans.shape)


def _check_waves(name, x, nx, ny):
    shp = tuple(x.shape)
    _need(len(shp) >= 2 and shp[-2:] == (nx, ny), f"{name} must end in ({nx}, {ny}), got {shp}")


class Context:
    """An msx context bound to one GPU (msx_create / msx_destroy)."""

    def __init__(self, device=0):
        L = load()
        h = C.c_void_p()
        _check(L.msx_create(int(device), C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            load().msx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _order(self, *xs):
        """Make the context stream wait for the work torch has queued on its
        current stream (msx_wait_stream), before a call that reads device
        tensors which that work may still be producing."""
        for x in xs:
            if x is not None and _is_torch(x) and x.is_cuda:
                import torch
                raw = torch.cuda.current_stream(x.device).cuda_stream
                _check(load().msx_wait_stream(self.h, C.c_void_p(raw) if raw else None))
                return

    def set_stream(self, stream):
        """Queue work on an external stream (torch.cuda.Stream or raw handle)."""
        raw = getattr(stream, "cuda_stream", stream)
        _check(load().msx_set_stream(self.h, C.c_void_p(raw) if raw else None))

    def autotune(self, nx, ny=None, prec=F32, max_batch=0):
        """Pick launch shapes for an nx x ny grid (msx_autotune); returns them as a dict."""
        t = Tuning()
        _check(load().msx_autotune(self.h, nx, ny or nx, prec, max_batch, C.byref(t)))
        return dict(row_wpb=t.row_wpb, col_wpb=t.col_wpb, batch=t.batch, ms_per_slice=t.ms_per_slice,
                    col_tile=t.col_tile, row_plan=t.row_plan, col_plan=t.col_plan)

    def set_tuning(self, nx, ny=None, prec=F32, tuning=None):
        """Install a saved autotuner result (dict as returned by autotune), or clear it (None).
        Missing keys: col_tile and the row / column FFT plans 0 (built-in)."""
        ny = ny or nx
        t = None if tuning is None else C.byref(Tuning(tuning["row_wpb"], tuning["col_wpb"], tuning["batch"],
                                                       tuning.get("ms_per_slice", 0.0), tuning.get("col_tile", 0),
                                                       tuning.get("row_plan", 0), tuning.get("col_plan", 0)))
        _check(load().msx_set_tuning(self.h, nx, ny or nx, prec, t))

    def set_graphs(self, enable=True):
        """Replay slice loops from cached HIP graphs (msx_set_graphs)."""
        _check(load().msx_set_graphs(self.h, int(enable)))

    def graph_stats(self):
        """(captures, replays) of slice-loop graphs so far."""
        c, r = C.c_int64(), C.c_int64()
        _check(load().msx_graph_stats(self.h, C.byref(c), C.byref(r)))
        return c.value, r.value

    def profile(self, enable=True):
        """Turn per-kernel event timing on/off (resets the totals)."""
        _check(load().msx_profile(self.h, int(bool(enable))))

    def profile_read(self, kind):
        """(launches, total_ms, total_algorithmic_bytes) for a kernel kind."""
        n, ms, by = C.c_int64(), C.c_double(), C.c_double()
        _check(load().msx_profile_read(self.h, int(kind), C.byref(n), C.byref(ms), C.byref(by)))
        return n.value, ms.value, by.value

    def copy_probe(self, mode, src, dst, nimg, nx, ny):
        """msx_copy_probe: ms of one copy of nimg nx x ny complex64 images
        (torch CUDA tensors src -> dst); mode 0 contiguous, 8 / 16 the column
        sweep's access pattern with that tile width (nx = 512)."""
        if src.numel() < nimg * nx * ny or dst.numel() < nimg * nx * ny:
            raise ValueError("copy_probe: buffers smaller than nimg * nx * ny")
        ms = C.c_double()
        self._order(src, dst)
        _check(load().msx_copy_probe(self.h, int(mode), C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()),
                                     int(nimg), int(nx), int(ny), C.byref(ms)))
        return ms.value

    # -------------------------------------------------------------- calls
    def fft2d(self, data, sign):
        """In-place cfpix::fft (sign=+1) / cfpix::ifft (sign=-1) on (..., nx, ny)."""
        prec = _prec_of(data)
        p, where, keep, arr = _buf(data, prec, writable=True)
        nx, ny = data.shape[-2], data.shape[-1]
        _need(len(tuple(data.shape)) >= 2, "fft2d needs (..., nx, ny)")
        nb = int(np.prod(data.shape[:-2])) if len(data.shape) > 2 else 1
        self._order(data)
        _check(load().msx_fft2d(self.h, p, nb, nx, ny, int(sign), prec, where))
        return data

    def potential(self, atoms, cell, nx, ny, nslices, radius, table=TABLE_LMULTISLICE,
                  formula=FORMULA_KIRKLAND, prec=F32, out=None):
        """ModelPotential::calculatePotentialGrid -> V[nslices][nx][ny] (V*Angstrom).
        atoms: structured ATOM_DTYPE numpy array (binned on the host,
        msx_potential) or a contiguous torch CUDA uint8 view of one, e.g. the
        output of displace() on the GPU (binned on the GPU,
        msx_potential_device_atoms; the same potential)."""
        if out is None:
            out = np.zeros((nslices, nx, ny), np.float64 if prec == F64 else np.float32)
        _need(tuple(out.shape) == (nslices, nx, ny), f"out must be ({nslices}, {nx}, {ny}), got {tuple(out.shape)}")
        p, where, keep, arr = _buf(out, prec, complex_=False, writable=True)
        geo = (float(cell[0]), float(cell[1]), float(cell[2]), nx, ny, nslices, float(radius), table, formula, p,
               prec, where)
        if _is_torch(atoms):
            if not atoms.is_cuda or not atoms.is_contiguous() or atoms.numel() % ATOM_DTYPE.itemsize:
                raise ValueError("device atoms: a contiguous CUDA tensor of whole ATOM_DTYPE records")
            self._order(atoms, out)
            _check(load().msx_potential_device_atoms(self.h, C.c_void_p(atoms.data_ptr()),
                                                     atoms.numel() // ATOM_DTYPE.itemsize, *geo))
            return out
        at = np.ascontiguousarray(atoms, dtype=ATOM_DTYPE)
        self._order(out)
        _check(load().msx_potential(self.h, C.c_void_p(at.ctypes.data), len(at), *geo))
        return out

    def potential_rs(self, atoms, cell, nx, ny, nslices, atrad, periodic=True, dw=None, ads=None,
                     gamma=np.pi / 2, prec=F32, out=None, out_abs=None, absorptive=False):
        """Real-space Doyle-Turner potential over pixel boxes, msx_potential_rs.
        Returns V (and Vabs when absorptive=True), [nslices][nx][ny] in V*A."""
        return self._dt_potential("rs", atoms, cell, nx, ny, nslices, dw, ads, gamma, (float(atrad), int(periodic)),
                                  prec, out, out_abs, absorptive)

    def potential_curve(self, kind, atoms, types, curves, dds, cell, nx, ny, nslices, ads=None, gamma=np.pi / 2,
                        nmult=1, mmult=1, atrad=None, periodic=True, prec=F32, out=None, out_abs=None,
                        absorptive=False):
        """Tabulated-curve potentials: kind "rec" (msx_potential_rec_curve) or "rs"
        (msx_potential_rs_curve); types: per-atom row of curves[ntypes][npoints]."""
        tp = np.ascontiguousarray(types, np.int32)
        cv = np.ascontiguousarray(curves, np.float64)
        if cv.ndim != 2:
            raise ValueError("curves must be (ntypes, npoints)")
        _need(tp.ndim == 1 and len(tp) == len(np.atleast_1d(atoms)), "types: one curve row per atom")
        self._keep_curves = (tp, cv)
        head = (C.c_void_p(tp.ctypes.data), C.c_void_p(cv.ctypes.data), cv.shape[0], cv.shape[1], float(dds))
        extra = (int(nmult), int(mmult)) if kind == "rec" else (float(atrad), int(periodic))
        return self._dt_potential(kind + "_curve", atoms, cell, nx, ny, nslices, None, ads, gamma, extra, prec, out,
                                  out_abs, absorptive, head=head)

    def _dt_potential(self, kind, atoms, cell, nx, ny, nslices, dw, ads, gamma, extra, prec, out, out_abs,
                      absorptive, head=None):
        at = np.ascontiguousarray(atoms, dtype=ATOM_DTYPE)
        n = len(at)
        arr = lambda v: None if v is None else np.ascontiguousarray(np.broadcast_to(v, (n,)), np.float64)  # noqa
        dwa, ada = arr(dw), arr(ads)
        dt = np.float64 if prec == F64 else np.float32
        if out is None:
            out = np.zeros((nslices, nx, ny), dt)
        if absorptive and out_abs is None:
            out_abs = np.zeros((nslices, nx, ny), dt) if not _is_torch(out) else out.new_empty(out.shape)
        for nm, o in (("out", out), ("out_abs", out_abs)):
            _need(o is None or tuple(o.shape) == (nslices, nx, ny),
                  f"{nm} must be ({nslices}, {nx}, {ny}), got {None if o is None else tuple(o.shape)}")
        pv = _buf(out, prec, complex_=False, writable=True)
        pa = _buf(out_abs, prec, complex_=False, writable=True) if out_abs is not None else (None, pv[1], None, None)
        where = _ptr_where([pv, pa])
        ap = None if ada is None else C.c_void_p(ada.ctypes.data)
        if head is None:
            args = [self.h, C.c_void_p(at.ctypes.data), n, None if dwa is None else C.c_void_p(dwa.ctypes.data), ap]
        else:
            args = [self.h, C.c_void_p(at.ctypes.data), n, *head, ap]
        args += [float(cell[0]), float(cell[1]), float(cell[2]), float(gamma), *extra, nx, ny, nslices, pv[0], pa[0],
                 prec, where]
        fn = {"rs": load().msx_potential_rs, "rec": load().msx_potential_rec,
              "rs_curve": load().msx_potential_rs_curve, "rec_curve": load().msx_potential_rec_curve}[kind]
        self._order(out, out_abs)
        _check(fn(*args))
        return (out, out_abs) if out_abs is not None else out

    def potential_rec(self, atoms, cell, nx, ny, nslices, dw=None, ads=None, gamma=np.pi / 2, nmult=1, mmult=1,
                      prec=F32, out=None, out_abs=None, absorptive=False):
        """Reciprocal-space (Doyle-Turner structure factor) potential, msx_potential_rec.
        cell = (a, b, c) of the unit cell; the grid spans (nmult*a, mmult*b).
        Returns V (and Vabs when absorptive=True), [nslices][nx][ny] in V*A."""
        return self._dt_potential("rec", atoms, cell, nx, ny, nslices, dw, ads, gamma, (int(nmult), int(mmult)),
                                  prec, out, out_abs, absorptive)

    def transmission(self, V, ax, by, kev, bw=2.0 / 3.0, out=None, Vabs=None):
        """exp(i sigma V / 1000), band-limited; V: (nlayer, nx, ny) real."""
        prec = _prec_of(V)
        _need(len(tuple(V.shape)) == 3, "V must be (nlayer, nx, ny)")
        nl, nx, ny = V.shape
        _need(out is None or tuple(out.shape) == (nl, nx, ny), f"out must be ({nl}, {nx}, {ny})")
        _need(Vabs is None or tuple(Vabs.shape) == (nl, nx, ny), f"Vabs must be ({nl}, {nx}, {ny})")
        pv = _buf(V, prec, complex_=False)
        if out is None:
            if _is_torch(V):
                import torch
                out = torch.empty((nl, nx, ny), dtype=torch.complex128 if prec == F64 else torch.complex64,
                                  device=V.device)
            else:
                out = np.zeros((nl, nx, ny), np.complex128 if prec == F64 else np.complex64)
        pt = _buf(out, prec, writable=True)
        bufs = [pv, pt]
        pa = None
        if Vabs is not None:
            pa = _buf(Vabs, prec, complex_=False)
            bufs.append(pa)
        where = _ptr_where(bufs)
        self._order(V, Vabs, out)
        _check(load().msx_transmission_abs(self.h, pv[0], None if pa is None else pa[0], nl, nx, ny, ax, by, kev, bw,
                                           pt[0], prec, where))
        return out

    @staticmethod
    def _stack(nx, ny, ax, by, kev, cz, layer, tilt=(0.0, 0.0), bw=2.0 / 3.0, gam=0.0):
        cz = np.ascontiguousarray(cz, np.float64)
        layer = np.ascontiguousarray(layer, np.int32)
        st = Stack(nx, ny, ax, by, kev, tilt[0], tilt[1], bw, len(cz),
                   cz.ctypes.data_as(C.POINTER(C.c_double)), len(layer),
                   layer.ctypes.data_as(C.POINTER(C.c_int32)), float(gam))
        return st, (cz, layer)

    def multislice(self, trans, cz, layer, psi, ax, by, kev, tilt=(0.0, 0.0), bw=2.0 / 3.0, psi0=None, gam=0.0):
        """Slice loop (mulslice.cpp:703-739) on nwave waves psi (nwave, nx, ny) in place.
        With psi0 (one image) every wave starts from psi0 and psi is output only.
        gam: angle between the cell axes (rad; 0 or pi/2 orthogonal): an oblique
        cell propagates with Musli's metric (Propagator.cpp:55-107)."""
        prec = _prec_of(trans)
        nl, nx, ny = _check_trans(trans)
        _check_layers(nl, cz, layer)
        _check_waves("psi", psi, nx, ny)
        _need(psi0 is None or _numel(psi0) == nx * ny, f"psi0 must be one ({nx}, {ny}) image")
        st, keep = self._stack(nx, ny, ax, by, kev, cz, layer, tilt, bw, gam)
        pt = _buf(trans, prec)
        pw = _buf(psi, prec, writable=True)
        bufs = [pt, pw]
        p0 = None
        if psi0 is not None:
            b0 = _buf(psi0, prec)
            bufs.append(b0)
            p0 = b0[0]
        where = _ptr_where(bufs)
        nwave = int(np.prod(psi.shape[:-2])) if len(psi.shape) > 2 else 1
        self._order(trans, psi, psi0)
        _check(load().msx_multislice(self.h, C.byref(st), pt[0], p0, pw[0], nwave, prec, where))
        return psi

    def multislice_series(self, trans, cz, layer, psi, ax, by, kev, at=(), beams=(), intensity=False,
                          tilt=(0.0, 0.0), bw=2.0 / 3.0, psi0=None, gam=0.0):
        """`multislice` (psi in place, same exit waves) recording a thickness
        series (msx_multislice_series; mulslice.cpp:682-737): returns
        (waves, beam_values, intensity) with waves (len(at), nwave, nx, ny) the
        exit waves after the slice counts `at`, beam_values (nslice, nwave,
        len(beams)) the k-space values at the (h, k) of `beams` over nx*ny
        after every slice, intensity (nslice, nwave) sum |psi|^2/(nx*ny) after
        every slice (None unless asked).  Outputs are numpy arrays for host
        psi, CUDA tensors for device psi."""
        prec = _prec_of(trans)
        nl, nx, ny = _check_trans(trans)
        _check_layers(nl, cz, layer)
        _check_waves("psi", psi, nx, ny)
        _need(psi0 is None or _numel(psi0) == nx * ny, f"psi0 must be one ({nx}, {ny}) image")
        at = np.ascontiguousarray(np.asarray(at, np.int32).reshape(-1))
        hk = np.ascontiguousarray(np.asarray(beams, np.int32).reshape(-1, 2))
        ns = len(layer)
        _need(np.all((at >= 1) & (at <= ns)) and np.all(np.diff(at) > 0),
              f"at must ascend strictly within 1..{ns}")
        _need(len(hk) <= 4096, "at most 4096 beams")
        nwave = int(np.prod(psi.shape[:-2])) if len(psi.shape) > 2 else 1
        st, keep = self._stack(nx, ny, ax, by, kev, cz, layer, tilt, bw, gam)
        pt = _buf(trans, prec)
        pw = _buf(psi, prec, writable=True)
        bufs = [pt, pw]
        p0 = None
        if psi0 is not None:
            b0 = _buf(psi0, prec)
            bufs.append(b0)
            p0 = b0[0]
        where = _ptr_where(bufs)
        cdt = np.complex64 if prec == F32 else np.complex128
        if where == DEVICE:
            import torch
            dev = psi.device
            tdt = torch.complex64 if prec == F32 else torch.complex128
            waves = torch.empty((len(at), nwave, nx, ny), dtype=tdt, device=dev)
            bv = torch.empty((ns, nwave, len(hk)), dtype=tdt, device=dev)
            inten = torch.empty((ns, nwave), dtype=torch.float64, device=dev) if intensity else None
            ptr = lambda x: x.data_ptr() if x is not None and x.numel() else None
        else:
            waves = np.empty((len(at), nwave, nx, ny), cdt)
            bv = np.empty((ns, nwave, len(hk)), cdt)
            inten = np.empty((ns, nwave), np.float64) if intensity else None
            ptr = lambda x: x.ctypes.data if x is not None and x.size else None
        se = Series(at.ctypes.data if len(at) else None, len(at), ptr(waves), hk.ctypes.data if len(hk) else None,
                    len(hk), ptr(bv), ptr(inten))
        self._order(trans, psi, psi0)
        _check(load().msx_multislice_series(self.h, C.byref(st), pt[0], p0, pw[0], nwave, prec, where, C.byref(se)))
        return waves, bv, inten

    def multislice_batch(self, jobs):
        """Uneven HRTEM jobs in shared launches (msx_multislice_batch).  jobs:
        list of dicts with trans (torch CUDA (nlayer, nx, ny) complex), psi (torch
        CUDA (nx, ny), incident in / exit out), ax, by, kev, cz, layer, and
        optional tilt, bw, psi0."""
        if not jobs:
            return
        prec = _prec_of(jobs[0]["trans"])
        keep, arr = [], (Job * len(jobs))()
        for k, j in enumerate(jobs):
            nl, nx, ny = _check_trans(j["trans"])
            _check_layers(nl, j["cz"], j["layer"])
            _need(_numel(j["psi"]) == nx * ny, f"job {k}: psi must be one ({nx}, {ny}) image")
            _need(j.get("psi0") is None or _numel(j["psi0"]) == nx * ny, f"job {k}: psi0 must be one ({nx}, {ny}) image")
            st, kp = self._stack(nx, ny, j["ax"], j["by"], j["kev"], j["cz"], j["layer"], j.get("tilt", (0.0, 0.0)),
                                 j.get("bw", 2.0 / 3.0))
            keep += [st, kp, j]
            for key in ("trans", "psi") + (("psi0",) if j.get("psi0") is not None else ()):
                t = j[key]
                if not (_is_torch(t) and t.is_cuda and t.is_contiguous() and _prec_of(t) == prec):
                    raise ValueError(f"job {k}: {key} must be a contiguous CUDA tensor of the batch precision")
            arr[k] = Job(C.addressof(st), j["trans"].data_ptr(),
                         j["psi0"].data_ptr() if j.get("psi0") is not None else None, j["psi"].data_ptr())
        self._order(jobs[0]["trans"])
        _check(load().msx_multislice_batch(self.h, arr, len(jobs), prec, DEVICE))

    def stem(self, trans, cz, layer, ax, by, kev, probe, positions, detectors, checkpoints,
             bw=2.0 / 3.0, batch=0, weight=1.0, out=None, tilt=(0.0, 0.0)):
        """STEM scan; returns out[ck][det][probe] (fraction of incident intensity).
        probe: dict(alpha_max, alpha_min=0, cs=0, df=0) (rad, Angstrom);
        detectors: [(k_inner, k_outer)] in 1/Angstrom; tilt: crystal tilt (rad)."""
        prec = _prec_of(trans)
        nl, nx, ny = _check_trans(trans)
        _check_layers(nl, cz, layer)
        st, keep = self._stack(nx, ny, ax, by, kev, cz, layer, tilt, bw)
        pt = _buf(trans, prec)
        _need(np.size(positions) % 2 == 0, "positions must be (nprobe, 2)")
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 2)
        det = np.ascontiguousarray(detectors, np.float64).reshape(-1, 2)
        ck = np.ascontiguousarray(checkpoints, np.int32)
        pr = Probe(float(probe["alpha_max"]), float(probe.get("alpha_min", 0.0)), float(probe.get("cs", 0.0)),
                   float(probe.get("df", 0.0)))
        dets = (Detector * len(det))(*[Detector(float(a), float(b)) for a, b in det])
        acc = 1
        if out is None:
            out = np.zeros((len(ck), len(det), len(pos)), np.float64)
            acc = 0
        _need(isinstance(out, np.ndarray) and out.dtype == np.float64 and out.flags.c_contiguous
              and out.flags.writeable, "out must be a writable C-contiguous float64 numpy array")
        _need(out.shape == (len(ck), len(det), len(pos)),
              f"out must be ({len(ck)}, {len(det)}, {len(pos)}) = (checkpoints, detectors, probes), got {out.shape}")
        self._order(trans)
        _check(load().msx_stem(self.h, C.byref(st), pt[0], C.byref(pr), C.c_void_p(pos.ctypes.data), len(pos),
                               dets, len(det), C.c_void_p(ck.ctypes.data), len(ck), float(weight), acc,
                               C.c_void_p(out.ctypes.data), int(batch), prec, pt[1]))
        return out

    def partial_coherence(self, trans, cz, layer, ax, by, kev, lens, acmin, acmax, sigmaf, dfdelt,
                          tilt=(0.0, 0.0), bw=2.0 / 3.0, batch=0):
        """Partial-coherence image (mulslice.cpp:539-676) of the stack `trans`:
        illumination angles acmin..acmax (rad), Gaussian defocus spread sigmaf
        sampled every dfdelt (Angstrom).  lens: dict(cs, df, aobj, dfa2=0,
        dfa2phi=0, dfa3=0, dfa3phi=0).  Returns (image float64 (nx, ny), number
        of angles, number of defocus steps); the image is a CUDA tensor when
        trans is one, else a numpy array."""
        prec = _prec_of(trans)
        nl, nx, ny = _check_trans(trans)
        _check_layers(nl, cz, layer)
        st, keep = self._stack(nx, ny, ax, by, kev, cz, layer, tilt, bw)
        pt = _buf(trans, prec)
        g = lambda k: float(lens.get(k, 0.0))  # noqa: E731
        ln = Lens(g("cs"), g("df"), g("aobj"), g("dfa2"), g("dfa2phi"), g("dfa3"), g("dfa3phi"))
        if pt[1] == DEVICE:
            import torch
            img = torch.empty((nx, ny), dtype=torch.float64, device=trans.device)
            pimg = C.c_void_p(img.data_ptr())
        else:
            img = np.zeros((nx, ny), np.float64)
            pimg = C.c_void_p(img.ctypes.data)
        na, nd = C.c_int32(0), C.c_int32(0)
        self._order(trans)
        _check(load().msx_partial_coherence(self.h, C.byref(st), pt[0], C.byref(ln), float(acmin), float(acmax),
                                            float(sigmaf), float(dfdelt), int(batch), pimg, C.byref(na),
                                            C.byref(nd), prec, pt[1]))
        return img, na.value, nd.value

    def ctf_image(self, psi, ax, by, kev, cs, df, aobj, want_wave=False):
        """Objective-lens image |ifft(fft(psi) exp(-i chi) A)|^2 of (nwave, nx, ny) waves."""
        prec = _prec_of(psi)
        _need(len(tuple(psi.shape)) >= 2, "psi must be (..., nx, ny)")
        nx, ny = psi.shape[-2], psi.shape[-1]
        nw = int(np.prod(psi.shape[:-2])) if len(psi.shape) > 2 else 1
        pw = _buf(psi, prec)
        if pw[1] == DEVICE:
            import torch
            inten = torch.empty(psi.shape, dtype=torch.float64 if prec == F64 else torch.float32, device=psi.device)
            img = torch.empty_like(psi) if want_wave else None
        else:
            inten = np.zeros(psi.shape, np.float64 if prec == F64 else np.float32)
            img = np.zeros(psi.shape, psi.dtype) if want_wave else None
        pi = _buf(inten, prec, complex_=False, writable=True)
        pimg = _buf(img, prec, writable=True)[0] if want_wave else None
        self._order(psi)
        _check(load().msx_ctf_image(self.h, pw[0], nw, nx, ny, ax, by, kev, cs, df, aobj, pimg, pi[0], prec,
                                    pw[1]))
        return (inten, img) if want_wave else inten


def unet_conv(ctx, mode, ks, out, w, x=None, bias=None, g=None, y=None, relu=False):
    """One convolution of the surrogate's kernels (msx_unet_conv) on torch CUDA
    float32 tensors: x [n, cin, h, w], w [cout, cin*ks*ks], g/y [n, cout, h, w];
    out is written (mode 0, 1) or accumulated into (mode 2)."""
    def ptr(t):
        if t is None:
            return None
        if not (_is_torch(t) and t.is_cuda and t.is_contiguous() and str(t.dtype) == "torch.float32"):
            raise ValueError("msx_unet_conv takes contiguous float32 CUDA tensors")
        return C.c_void_p(t.data_ptr())
    _need(mode in (0, 1, 2) and ks in (1, 3), "mode must be 0, 1 or 2 and ks 1 or 3")
    _need(len(tuple(w.shape)) == 2 and w.shape[1] % (ks * ks) == 0, "w must be (cout, cin*ks*ks)")
    ref = x if x is not None else g
    _need(ref is not None and len(tuple(ref.shape)) == 4, "x (or g) must be (n, c, h, w)")
    n, _, hh, ww = ref.shape
    cout = w.shape[0]
    cin = w.shape[1] // (ks * ks)
    shapes = {"x": (n, cin, hh, ww), "g": (n, cout, hh, ww), "y": (n, cout, hh, ww),
              "out": {0: (n, cout, hh, ww), 1: (n, cin, hh, ww), 2: (cout, cin * ks * ks)}[mode]}
    for nm, t in (("x", x), ("g", g), ("y", y), ("out", out)):
        _need(t is None or tuple(t.shape) == shapes[nm], f"{nm} must be {shapes[nm]}, got {None if t is None else tuple(t.shape)}")
    _need(bias is None or tuple(bias.shape) == (cout,), f"bias must be ({cout},)")
    _need(out is not None and (x is not None if mode in (0, 2) else g is not None) and (mode == 0 or g is not None),
          "mode 0 needs x and out, mode 1 g and out, mode 2 x, g and out")
    ctx._order(x, w, bias, g, y, out)
    _check(load().msx_unet_conv(ctx.h, mode, ks, ptr(x), ptr(w), ptr(bias), ptr(g), ptr(y), ptr(out), n, cin, cout,
                                hh, ww, int(relu)))
    return out


class UNet:
    """The frozen-phonon surrogate network (msx_unet_*), resident on the
    context's GPU.  Inputs are (n, h, w) float32 images for a one-channel
    network, (n, in_channels, h, w) otherwise (one image per frozen-phonon
    configuration); outputs and targets (n, h, w).  numpy (host) or torch CUDA
    tensors (device); h, w multiples of 2**(levels-1)."""

    def __init__(self, ctx, levels=4, base=32, seed=0, in_channels=1):
        self.ctx, self.levels, self.base, self.in_channels = ctx, levels, base, int(in_channels)
        h = C.c_void_p()
        _check(load().msx_unet_create_in(ctx.h, levels, base, self.in_channels, seed, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            load().msx_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nparam(self):
        return int(load().msx_unet_param_count(self.h))

    def params(self):
        p = np.empty(self.nparam, np.float32)
        _check(load().msx_unet_params(self.h, C.c_void_p(p.ctypes.data), HOST))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.nparam:
            raise ValueError(f"{p.size} parameters given, the network has {self.nparam}")
        _check(load().msx_unet_set_params(self.h, C.c_void_p(p.ctypes.data), HOST))

    def _in(self, x):
        """Input buffer and its (n, h, w)."""
        b = _buf(x, F32, complex_=False)
        shp = tuple(b[2].shape)
        if self.in_channels == 1 and len(shp) == 3:
            return b, shp
        if len(shp) != 4 or shp[1] != self.in_channels:
            raise ValueError(f"inputs must be (n, {self.in_channels}, h, w)"
                             + (" or (n, h, w)" if self.in_channels == 1 else "") + f", got {shp}")
        return b, (shp[0], shp[2], shp[3])

    @staticmethod
    def _img(x, shp, name="target"):
        b = _buf(x, F32, complex_=False)
        _need(tuple(b[2].shape) == shp, f"{name} must be {shp}, got {tuple(b[2].shape)}")
        return b

    def forward(self, x, out=None):
        bx, (n, hh, ww) = self._in(x)
        if out is None:
            out = np.empty((n, hh, ww), np.float32) if bx[1] == HOST else x.new_empty((n, hh, ww))
        _need(tuple(out.shape) == (n, hh, ww), f"out must be ({n}, {hh}, {ww})")
        bo = _buf(out, F32, complex_=False, writable=True)
        where = _ptr_where([bx, bo])
        self.ctx._order(x, out)
        _check(load().msx_unet_forward(self.h, bx[0], bo[0], n, hh, ww, where))
        return out

    def gradient(self, x, target):
        """(loss, packed gradient as numpy float32)."""
        bx, (n, hh, ww) = self._in(x)
        bt = self._img(target, (n, hh, ww))
        g = np.empty(self.nparam, np.float32)
        loss = C.c_double()
        where = _ptr_where([bx, bt])
        self.ctx._order(x, target)
        if where == DEVICE:
            import torch
            gd = torch.empty(self.nparam, dtype=torch.float32, device=x.device)
            _check(load().msx_unet_gradient(self.h, bx[0], bt[0], n, hh, ww, C.byref(loss),
                                            C.c_void_p(gd.data_ptr()), DEVICE))
            g = gd.cpu().numpy()
        else:
            _check(load().msx_unet_gradient(self.h, bx[0], bt[0], n, hh, ww, C.byref(loss),
                                            C.c_void_p(g.ctypes.data), HOST))
        return loss.value, g

    def train_step(self, x, target, lr):
        bx, (n, hh, ww) = self._in(x)
        bt = self._img(target, (n, hh, ww))
        loss = C.c_double()
        self.ctx._order(x, target)
        _check(load().msx_unet_train_step(self.h, bx[0], bt[0], n, hh, ww, float(lr), C.byref(loss),
                                          _ptr_where([bx, bt])))
        return loss.value


_default = None


def context():
    """Process-wide default context on device 0."""
    global _default
    with _lock:
        if _default is None:
            _default = Context(0)
        return _default


def parlay(sequence, nlayers=52, nsmax=1 << 20):
    """Layer indices of a Kirkland stacking sequence (msx_parlay): "2(abc)d"
    -> [0, 1, 2, 0, 1, 2, 3]; letters past nlayers are out of range.  Raises
    MsxError with parlay's code on a malformed sequence."""
    out = np.zeros(nsmax, np.int32)
    n = C.c_int32(0)
    _check(load().msx_parlay(sequence.encode(), C.c_void_p(out.ctypes.data), nsmax, nlayers - 1, C.byref(n)))
    return out[:n.value].copy()


def tiff_write(path, image, params=None, dx=0.0, dy=0.0, rmin=None, rmax=None, imin=None, imax=None):
    """Write a Kirkland floating-point TIFF (msx_tiff_write).  image: real
    (ny, nx) float array, or complex (ny, width) -> real and imaginary halves
    side by side (npix = 2).  The preview range defaults to the data range."""
    im = np.asarray(image)
    if np.iscomplexobj(im):
        pix = np.ascontiguousarray(np.concatenate([im.real, im.imag], axis=1), np.float32)
        npix = 2
    else:
        pix = np.ascontiguousarray(im, np.float32)
        npix = 1
    half = pix[:, : pix.shape[1] // npix]
    rmin = float(half.min()) if rmin is None else rmin
    rmax = float(half.max()) if rmax is None else rmax
    if npix == 2:
        ih = pix[:, pix.shape[1] // 2:]
        imin = float(ih.min()) if imin is None else imin
        imax = float(ih.max()) if imax is None else imax
    prm = np.zeros(512, np.float32) if params is None else np.ascontiguousarray(params, np.float32)
    _check(load().msx_tiff_write(os.fsencode(path), C.c_void_p(pix.ctypes.data), pix.shape[1], pix.shape[0], npix,
                                 C.c_void_p(prm.ctypes.data), len(prm), rmin, rmax, imin or 0.0, imax or 0.0, dx, dy))


def tiff_read(path):
    """Read a Kirkland floating-point TIFF (msx_tiff_read): returns (image,
    params); a complex file (npix = 2) comes back as a complex (ny, nx/2) array."""
    nx, ny, npix = C.c_int32(), C.c_int32(), C.c_int32()
    prm = np.zeros(512, np.float32)
    _check(load().msx_tiff_read(os.fsencode(path), None, 0, C.byref(nx), C.byref(ny), C.byref(npix), None))
    pix = np.zeros((ny.value, nx.value), np.float32)
    _check(load().msx_tiff_read(os.fsencode(path), C.c_void_p(pix.ctypes.data), pix.size, C.byref(nx), C.byref(ny),
                                C.byref(npix), C.c_void_p(prm.ctypes.data)))
    if npix.value == 2:
        h = nx.value // 2
        return pix[:, :h] + 1j * pix[:, h:], prm
    return pix, prm


def musli_stem_write(path, image, scan_x, scan_y, thickness, kev, gamma=np.pi / 2):
    """Musli .stem file (msx_musli_stem_write): image (stem_m, stem_n) doubles
    in scan order (x fastest), scan extents and thickness in Angstrom."""
    im = np.ascontiguousarray(image, np.float64)
    m, n = im.shape
    _check(load().msx_musli_stem_write(os.fsencode(path), C.c_void_p(im.ctypes.data), n, m, float(scan_x),
                                       float(scan_y), float(thickness), float(gamma), float(kev)))


def musli_wave_write(path, wave, a, b, thickness, kev, gamma=np.pi / 2):
    """Musli .wav file (msx_musli_wave_write): wave (m, n) complex, x fastest
    (Musli's layout; an engine wave [ix][iy] is wave.T)."""
    w = np.ascontiguousarray(wave, np.complex128)
    m, n = w.shape
    _check(load().msx_musli_wave_write(os.fsencode(path), C.c_void_p(w.ctypes.data), n, m, float(a), float(b),
                                       float(thickness), float(gamma), float(kev)))


def amorphous(cell, Z=6, density_g_cm3=2.0, min_dist=1.4, seed=0, mass=None, max_tries=200):
    """Amorphous specimen by random sequential addition (msx_amorphous) in the
    periodic box cell = (a, b, c) Angstrom: atoms of element Z at the given
    mass density (mass in u; carbon 12.011 by default), at least min_dist
    apart.  Returns an ATOM_DTYPE array (fewer atoms than the density asks
    only if max_tries draws per atom did not place them all)."""
    a, b, c = (float(v) for v in cell)
    m = mass if mass is not None else {6: 12.011}.get(int(Z))
    if m is None:
        raise ValueError("give the atomic mass for Z != 6")
    n = int(round(density_g_cm3 / m * 6.02214076e23 * 1e-24 * a * b * c))
    out = np.zeros(max(n, 1), ATOM_DTYPE)
    placed = C.c_int64(0)
    _check(load().msx_amorphous(a, b, c, int(Z), n, float(min_dist), int(seed) & (2 ** 64 - 1), n * int(max_tries),
                                C.c_void_p(out.ctypes.data), C.byref(placed)))
    return out[:placed.value]


def displace(atoms, u_rms, seed, config, ctx=None, out=None):
    """Frozen-phonon displacement (msx_displace).  atoms: structured ATOM_DTYPE
    numpy array (host, ctx optional) or a torch CUDA uint8 view of one (device,
    ctx required); u_rms: {Z: u} or a 104-array (Angstrom per axis)."""
    u = np.zeros(104)
    if isinstance(u_rms, dict):
        for z, v in u_rms.items():
            u[int(z)] = v
    else:
        u[:] = np.asarray(u_rms, np.float64)
    if _is_torch(atoms):
        if ctx is None:
            raise ValueError("device displacement needs a Context")
        _need(atoms.is_cuda and atoms.is_contiguous() and atoms.numel() % ATOM_DTYPE.itemsize == 0,
              "device atoms: a contiguous CUDA tensor of whole ATOM_DTYPE records")
        n = atoms.numel() // ATOM_DTYPE.itemsize
        if out is None:
            out = atoms.clone()
        _need(_is_torch(out) and out.is_cuda and out.is_contiguous() and out.numel() == atoms.numel(),
              "out must be a contiguous CUDA tensor of the atoms' size")
        ctx._order(atoms, out)
        _check(load().msx_displace(ctx.h, C.c_void_p(atoms.data_ptr()), n, C.c_void_p(u.ctypes.data), seed, config,
                                   C.c_void_p(out.data_ptr()), DEVICE))
        return out
    at = np.ascontiguousarray(atoms, dtype=ATOM_DTYPE)
    if out is None:
        out = np.empty_like(at)
    _need(isinstance(out, np.ndarray) and out.dtype == ATOM_DTYPE and out.shape == at.shape and out.flags.c_contiguous,
          "out must be a C-contiguous ATOM_DTYPE array of the atoms' length")
    _check(load().msx_displace(ctx.h if ctx else None, C.c_void_p(at.ctypes.data), len(at),
                               C.c_void_p(u.ctypes.data), seed, config, C.c_void_p(out.ctypes.data), HOST))
    return out


def create_pots_rec_dt(types, dw, npoints, dds, vol):
    """CreatePots_rec_DT_DW (Muslib/Curves.cpp:45-63): Doyle-Turner
    reciprocal-space curves (len(types), npoints) at s = i dds, times vol."""
    tp = np.ascontiguousarray(types, np.int32)
    d = np.ascontiguousarray(np.broadcast_to(dw, tp.shape), np.float64)
    out = np.zeros((len(tp), int(npoints)), np.float64)
    _check(load().msx_create_pots_rec_dt(len(tp), int(npoints), float(dds), C.c_void_p(tp.ctypes.data),
                                         C.c_void_p(d.ctypes.data), float(vol), C.c_void_p(out.ctypes.data)))
    return out


def create_pots_rs_dt(types, dw, npoints, dds):
    """CreatePots_rs_DT_DW (Muslib/Curves.cpp:64-81): Doyle-Turner
    real-space curves (len(types), npoints) at r = i dds."""
    tp = np.ascontiguousarray(types, np.int32)
    d = np.ascontiguousarray(np.broadcast_to(dw, tp.shape), np.float64)
    out = np.zeros((len(tp), int(npoints)), np.float64)
    _check(load().msx_create_pots_rs_dt(len(tp), int(npoints), float(dds), C.c_void_p(tp.ctypes.data),
                                        C.c_void_p(d.ctypes.data), C.c_void_p(out.ctypes.data)))
    return out


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    load().msx_philox4x32_10(C.c_void_p(c.ctypes.data), C.c_void_p(k.ctypes.data), C.c_void_p(o.ctypes.data))
    return o


def wavelength(kev):
    return load().msx_wavelength(kev)


def sigma(kev):
    return load().msx_sigma(kev)


def vzatom(Z, r, table=TABLE_LMULTISLICE, formula=FORMULA_KIRKLAND):
    return load().msx_vzatom(int(Z), float(r), int(table), int(formula))


def featom(Z, k2, table=TABLE_LMULTISLICE):
    return load().msx_featom(int(Z), float(k2), int(table))
