#include "aslDataInc.h"
