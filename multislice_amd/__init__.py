"""msx -- MI355X-native multislice TEM/STEM engine.

The compute path is libmsx.so (hand-written HIP kernels for gfx950 behind the
C ABI of include/msx.h).  This package is its Python host binding plus the
host-side mirror of the reference's LMultiSlice classes.
"""
from .engine import (  # noqa: F401
    F32, F64, HOST, DEVICE, TABLE_LMULTISLICE, TABLE_KIRKLAND, FORMULA_KIRKLAND, FORMULA_LMULTISLICE,
    SIZES, ATOM_DTYPE, Context, MsxError, context, load, wavelength, sigma, vzatom, featom, LIB_PATH, EXPORTS,
    PROF_ROW, PROF_COL, PROF_SPLAT, PROF_DET, displace, philox4x32_10, UNet, unet_conv, amorphous, parlay, tiff_read, tiff_write, musli_stem_write, musli_wave_write,
    create_pots_rec_dt, create_pots_rs_dt,
)
