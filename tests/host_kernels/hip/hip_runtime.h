// Host build of the generated mechanism code for CPU tests: the device
// qualifiers become plain C++ and the math comes from the host libm.
#pragma once
#include <cmath>
#define __device__
#define __forceinline__ inline
using std::exp;
using std::fabs;
using std::fmax;
using std::log;
using std::log10;
using std::pow;
typedef int hipError_t;
typedef void* hipStream_t;
