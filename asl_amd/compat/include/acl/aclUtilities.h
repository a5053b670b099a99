#include "aclTypes.h"
