// mueller.h -- ampl_to_mueller (outputs_mod.f90:473-549; Bohren & Huffman 3.16)
// for one bin (host + device), in the reference's order of operations.
//
// Amplitude a[8]: re/im of S(1,1), S(2,1), S(1,2), S(2,2) (the Fortran pair
// order r + 2c); M[16]: S11, S12, ..., S44 row by row.
#pragma once
#include <hip/hip_runtime.h>

namespace abt {

struct Cx { double r, i; };

__host__ __device__ __forceinline__ Cx mul_conj(Cx a, Cx b)  // a * conjg(b)
{
#pragma clang fp contract(off)
    return Cx{a.r * b.r + a.i * b.i, a.i * b.r - a.r * b.i};
}

__host__ __device__ __forceinline__ void mueller16(const double a[8], double M[16])
{
#pragma clang fp contract(off)
    const Cx s11{a[0], a[1]}, s21{a[2], a[3]}, s12{a[4], a[5]}, s22{a[6], a[7]};
    const Cx a11 = mul_conj(s11, s11), a12 = mul_conj(s12, s12);
    const Cx a21 = mul_conj(s21, s21), a22 = mul_conj(s22, s22);
    const Cx p1112 = mul_conj(s11, s12), p2221 = mul_conj(s22, s21);
    const Cx p1121 = mul_conj(s11, s21), p2212 = mul_conj(s22, s12);
    const Cx p1122 = mul_conj(s11, s22), p1221 = mul_conj(s12, s21), p2112 = mul_conj(s21, s12);
    const Cx p2111 = mul_conj(s21, s11), p2211 = mul_conj(s22, s11);
    M[0] = 0.5 * (((a11.r + a12.r) + a21.r) + a22.r);
    M[1] = 0.5 * (((a11.r - a12.r) + a21.r) - a22.r);
    M[2] = p1112.r + p2221.r;
    M[3] = p1112.i - p2221.i;
    M[4] = 0.5 * (((a11.r + a12.r) - a21.r) - a22.r);
    M[5] = 0.5 * (((a11.r - a12.r) - a21.r) + a22.r);
    M[6] = p1112.r - p2221.r;
    M[7] = p1112.i + p2221.i;
    M[8] = p1121.r + p2212.r;
    M[9] = p1121.r - p2212.r;
    M[10] = p1122.r + p1221.r;
    M[11] = p1122.i + p2112.i;
    M[12] = p2111.i + p2212.i;
    M[13] = p2111.i - p2212.i;
    M[14] = p2211.i - p1221.i;
    M[15] = p2211.r - p1221.r;
}

// S11 alone: 0.5 (|S11|^2 + |S12|^2 + |S21|^2 + |S22|^2), same order as mueller16
__host__ __device__ __forceinline__ double mueller_s11(const double a[8])
{
#pragma clang fp contract(off)
    const double a11 = a[0] * a[0] + a[1] * a[1], a21 = a[2] * a[2] + a[3] * a[3];
    const double a12 = a[4] * a[4] + a[5] * a[5], a22 = a[6] * a[6] + a[7] * a[7];
    return 0.5 * (((a11 + a12) + a21) + a22);
}

}  // namespace abt
