// aslSingleKernelNM.h — numerical method run by one kernel per execute()
// (src/num/aslSingleKernelNM.h:46-80).
#ifndef ASLSINGLEKERNELNM_H
#define ASLSINGLEKERNELNM_H
#include "aslNumMethod.h"
namespace asl {
class SingleKernelNM : public NumMethod {};
}  // namespace asl
#endif
