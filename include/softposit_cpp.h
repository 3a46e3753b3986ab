/*
 * softposit_cpp.h -- posit32 value type for the libRpack drop-in.
 *
 * The reference's API headers (libRpack/mpblas_posit.h:35, mplapack_posit.h:35) include
 * SoftPosit's softposit_cpp.h for the posit32 class.  This header provides a class with
 * the same layout (one uint32_t `value`, standard layout, 4 bytes) and the members the
 * reference code and drivers use: construction / assignment from double and int,
 * + - * / with posit32 or double operands, unary -, += -= *= /=, comparisons (posit
 * order = signed integer order of the patterns), toDouble(), sqrt().  Arithmetic calls
 * the library's scalar entry points (include/posit_gpu_amd.h), which round exactly as
 * the device kernels do.
 */
#ifndef PGA_SOFTPOSIT_CPP_H
#define PGA_SOFTPOSIT_CPP_H

#include <stdint.h>

#include "posit_gpu_amd.h"

struct posit32 {
    uint32_t value;

    posit32() : value(0) {}
    posit32(double x) : value(pga_p32_from_double(x)) {}
    posit32(float x) : value(pga_p32_from_double((double)x)) {}
    posit32(int x) : value(pga_p32_from_double((double)x)) {}
    posit32(long x) : value(pga_p32_from_double((double)x)) {}

    static posit32 from_bits(uint32_t bits) {
        posit32 p;
        p.value = bits;
        return p;
    }

    posit32 &operator=(double x) { value = pga_p32_from_double(x); return *this; }
    posit32 &operator=(int x) { value = pga_p32_from_double((double)x); return *this; }

    posit32 operator+(const posit32 &b) const { return from_bits(pga_p32_add(value, b.value)); }
    posit32 operator-(const posit32 &b) const { return from_bits(pga_p32_sub(value, b.value)); }
    posit32 operator*(const posit32 &b) const { return from_bits(pga_p32_mul(value, b.value)); }
    posit32 operator/(const posit32 &b) const { return from_bits(pga_p32_div(value, b.value)); }
    posit32 operator-() const { return from_bits(0u - value); }
    posit32 &operator+=(const posit32 &b) { value = pga_p32_add(value, b.value); return *this; }
    posit32 &operator-=(const posit32 &b) { value = pga_p32_sub(value, b.value); return *this; }
    posit32 &operator*=(const posit32 &b) { value = pga_p32_mul(value, b.value); return *this; }
    posit32 &operator/=(const posit32 &b) { value = pga_p32_div(value, b.value); return *this; }

    bool operator==(const posit32 &b) const { return value == b.value; }
    bool operator!=(const posit32 &b) const { return value != b.value; }
    bool operator<(const posit32 &b) const { return (int32_t)value < (int32_t)b.value; }
    bool operator<=(const posit32 &b) const { return (int32_t)value <= (int32_t)b.value; }
    bool operator>(const posit32 &b) const { return (int32_t)value > (int32_t)b.value; }
    bool operator>=(const posit32 &b) const { return (int32_t)value >= (int32_t)b.value; }

    double toDouble() const { return pga_p32_to_double(value); }
    bool isNaR() const { return value == 0x80000000u; }
    posit32 &sqrt() { value = pga_p32_sqrt(value); return *this; }
};

inline posit32 sqrt(posit32 a) { return posit32::from_bits(pga_p32_sqrt(a.value)); }
inline posit32 operator+(double a, const posit32 &b) { return posit32(a) + b; }
inline posit32 operator-(double a, const posit32 &b) { return posit32(a) - b; }
inline posit32 operator*(double a, const posit32 &b) { return posit32(a) * b; }
inline posit32 operator/(double a, const posit32 &b) { return posit32(a) / b; }

#endif
