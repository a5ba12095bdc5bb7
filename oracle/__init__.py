"""CPU oracle for the msx engine -- TEST INFRASTRUCTURE ONLY.

``oracle.lib()`` loads ``liboracle.so`` (the C++ restatement of the
reference's CPU multislice path, oracle/msx_oracle.cpp) and ``oracle.ref()``
loads the reference's own slicelib.cpp / FFTW 2.1.5 built by build_ref.sh
(present only where the reference tree was available at build time).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this package; the product (multislice_amd) never does.
"""
from .wrap import lib, ref, ref_fftw2, Oracle  # noqa: F401
