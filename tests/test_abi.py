"""The C-ABI library (libmsx.so) loads, exports every entry point that
include/msx.h declares, and its host-side functions agree with the oracle.
No GPU is needed (no compute call is made)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import multislice_amd as msx
from oracle import Oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = []
    for h in ("include/msx.h",):
        text = open(os.path.join(ROOT, h)).read()
        names += re.findall(r"MSX_API\s+[\w\s\*]+?\b(msx_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_expected_surface():
    names = declared_entry_points()
    for n in ["msx_create", "msx_destroy", "msx_fft2d", "msx_potential", "msx_transmission",
              "msx_multislice", "msx_stem", "msx_ctf_image", "msx_last_error"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(msx.LIB_PATH)
    missing = [n for n in declared_entry_points() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding knows every declared entry point
    assert set(declared_entry_points()) <= set(msx.EXPORTS)


def test_host_scalars_match_oracle():
    for kev in (60.0, 100.0, 200.0, 300.0):
        assert msx.wavelength(kev) == O.wavelength(kev)
        assert abs(msx.sigma(kev) - O.sigma(kev)) <= 1e-15 * O.sigma(kev)
    rng = np.random.default_rng(0)
    for Z in (1, 6, 14, 38, 79, 103):
        for r in np.concatenate([rng.random(6) * 4, [0.0, 1e-11]]):
            for table in (0, 1):
                for formula in (0, 1):
                    a = msx.vzatom(Z, r, table, formula)
                    b = O.vzatom(Z, r, table, formula)
                    assert abs(a - b) <= 1e-13 * abs(b)
        for k2 in rng.random(4) * 8:
            assert abs(msx.featom(Z, k2, 1) - O.featom(Z, k2, 1)) <= 1e-14 * O.featom(Z, k2, 1)


def test_out_of_range_Z_is_zero():
    assert msx.vzatom(0, 1.0) == 0.0 and msx.vzatom(104, 1.0) == 0.0
    assert msx.featom(-3, 1.0) == 0.0


def test_null_context_is_rejected_with_message():
    L = msx.load()
    rc = L.msx_fft2d(None, None, 1, 64, 64, 1, 0, 0)
    assert rc == -5
    assert b"ctx" in L.msx_last_error()


def test_create_reports_missing_device_or_succeeds():
    """Without a GPU msx_create fails with a HIP/argument error (never exits);
    with one it succeeds."""
    L = msx.load()
    h = C.c_void_p()
    rc = L.msx_create(0, C.byref(h))
    if rc == 0:
        L.msx_destroy(h)
    else:
        assert rc in (-1, -3)
        assert len(L.msx_last_error()) > 0


def test_structs_match_header_layout(tmp_path):
    """The ctypes structs of the binding have the sizes and field offsets the
    C compiler gives the header's structs."""
    import subprocess
    from multislice_amd.engine import Stack, Probe, Detector, Tuning, Job, Lens, Series
    assert msx.ATOM_DTYPE.itemsize == 32
    fields = {"msx_stack": (Stack, ["nx", "ny", "ax", "by", "kev", "ctilt_x", "ctilt_y", "bw", "nlayer", "cz",
                                     "nslice", "layer", "gam"]),
              "msx_tuning": (Tuning, ["row_wpb", "col_wpb", "batch", "ms_per_slice", "col_tile", "row_plan", "col_plan"]),
              "msx_probe": (Probe, ["alpha_max", "alpha_min", "cs", "df"]),
              "msx_detector": (Detector, ["k_inner", "k_outer"]),
              "msx_job": (Job, ["stack", "trans", "psi0", "psi"]),
              "msx_lens": (Lens, ["cs", "df", "aobj", "dfa2", "dfa2phi", "dfa3", "dfa3phi"]),
              "msx_series": (Series, ["at", "nat", "waves", "hk", "nbeams", "beams", "intensity"])}
    src = ["#include <stdio.h>", "#include <stddef.h>", "#include \"msx.h\"", "int main(void) {"]
    for st, (_, names) in fields.items():
        src.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for n in names:
            src.append(f'printf("{st}.{n} %zu\\n", offsetof({st}, {n}));')
    src.append("return 0; }")
    (tmp_path / "lay.c").write_text("\n".join(src))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["gcc", "-I" + os.path.join(root, "include"), str(tmp_path / "lay.c"), "-o", str(tmp_path / "lay")],
                   check=True)
    out = dict(line.split() for line in subprocess.run([str(tmp_path / "lay")], capture_output=True, text=True,
                                                      check=True).stdout.splitlines())
    for st, (cls, names) in fields.items():
        assert int(out[st]) == C.sizeof(cls), st
        for n in names:
            assert int(out[f"{st}.{n}"]) == getattr(cls, n).offset, (st, n)
