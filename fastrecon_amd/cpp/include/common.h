// Common types of the fastrecon C++ API (reference ImageData/common.h:7-18).
#ifndef COMMON_H
#define COMMON_H

#include <complex>
#include <vector>

typedef struct
{
    unsigned x;
    unsigned y;
    unsigned z;
} ImageSize;

template<typename T>
using hostVector = std::vector<T>;

template<typename T>
using ComplexVector = hostVector<std::complex<T>>;

#endif // COMMON_H
