// aslDataInc.h — data containers of ASL's API (src/aslDataInc.h).
#ifndef ASLDATAINC_H
#define ASLDATAINC_H
#include "acl/aclGenerators.h"
#include "data/aslBlocks.h"
#include "data/aslDataWithGhostNodes.h"
#include "utilities/aslUValue.h"

namespace asl {
/// initData(d, a) / initData(d, AVec) (src/aslDataInc.h:36-37): every node of
/// every component (one component per vector element) set to the value.
void initData(SPAbstractData d, double a);
void initData(SPAbstractData d, const AVec<double> &a);
}  // namespace asl
#endif
