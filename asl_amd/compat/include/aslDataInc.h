// aslDataInc.h — data containers of ASL's API (src/aslDataInc.h).
#ifndef ASLDATAINC_H
#define ASLDATAINC_H
#include "acl/aclGenerators.h"
#include "data/aslBlocks.h"
#include "data/aslDataWithGhostNodes.h"
#include "utilities/aslUValue.h"
#endif
