/* convGemm.h -- the reference's spelling of the header (src/convGemm.h,
 * installed as include/convGemm.h by src/CMakeLists.txt:40).  Everything is
 * declared in convgemm.h. */
#include "convgemm.h"
