// msx_lms.hip -- out-of-line definitions of the C++ mirror's exported symbols
// (include/msx/lmultislice.hpp).  The reference's PotentialBuilder DLL
// (ImageSimulation/PotentialBuilder/PotentialBuilder.cpp:9-22) exports an
// empty class constructor, the variable nPotentialBuilder = 0 and the
// function fnPotentialBuilder() returning 42; libmsx.so exports the same
// three, once, so callers link them from the library.
#include "../../include/msx/lmultislice.hpp"

namespace msx {
namespace lms {

MSX_API int nPotentialBuilder = 0;

MSX_API int fnPotentialBuilder(void) { return 42; }

CPotentialBuilder::CPotentialBuilder(void) {}

}  // namespace lms
}  // namespace msx
