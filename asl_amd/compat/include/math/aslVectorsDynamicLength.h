#include "aslVectors.h"
