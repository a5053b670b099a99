#include "aclGenerators.h"
#include "aclTypes.h"
