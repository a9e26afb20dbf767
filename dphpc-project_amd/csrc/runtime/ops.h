// Datatypes and reduction operators.
//
// Type sizes follow gpu_mpi/mpi_types_list.cuh:4-33 (MPI_LONG_DOUBLE is a double, the
// complex types are pairs of float/double).  Operator bodies follow
// gpu_mpi/operators.cu:73-151 -- inout[i] = op(in[i], inout[i]), with MPI_LXOR being value
// inequality as in the reference -- and the legality table operators.cu:153-200
// (MAXLOC/MINLOC have no legal type in this type list).  The library is compiled with
// -ffp-contract=off, so complex products and every other FP expression round exactly
// as written, matching the oracle.
#pragma once

#include "common.h"
#include "../../../include/gpumpi/mpi.h"

namespace gpumpi {

enum OpId : int { kOpMax, kOpMin, kOpSum, kOpProd, kOpLand, kOpBand, kOpLor, kOpBor,
                  kOpLxor, kOpBxor, kOpMaxloc, kOpMinloc };

__host__ __device__ __forceinline__ int type_size(int t) {
    switch (t) {
    case MPI_CHAR: case MPI_SIGNED_CHAR: case MPI_UNSIGNED_CHAR: case MPI_C_BOOL:
    case MPI_INT8_T: case MPI_UINT8_T: case MPI_BYTE: return 1;
    case MPI_SHORT: case MPI_UNSIGNED_SHORT: case MPI_INT16_T: case MPI_UINT16_T: return 2;
    case MPI_INT: case MPI_UNSIGNED: case MPI_FLOAT: case MPI_WCHAR: case MPI_INT32_T:
    case MPI_UINT32_T: return 4;
    case MPI_LONG: case MPI_LONG_LONG_INT: case MPI_LONG_LONG: case MPI_UNSIGNED_LONG:
    case MPI_UNSIGNED_LONG_LONG: case MPI_DOUBLE: case MPI_LONG_DOUBLE: case MPI_INT64_T:
    case MPI_UINT64_T: case MPI_C_COMPLEX: case MPI_C_FLOAT_COMPLEX: return 8;
    case MPI_C_DOUBLE_COMPLEX: case MPI_C_LONG_DOUBLE_COMPLEX: return 16;
    default: return -1;
    }
}

__host__ __device__ __forceinline__ bool type_is_integer(int t) {
    switch (t) {
    case MPI_SHORT: case MPI_INT: case MPI_LONG: case MPI_LONG_LONG_INT: case MPI_LONG_LONG:
    case MPI_SIGNED_CHAR: case MPI_UNSIGNED_CHAR: case MPI_UNSIGNED_SHORT: case MPI_UNSIGNED:
    case MPI_UNSIGNED_LONG: case MPI_UNSIGNED_LONG_LONG: case MPI_INT8_T: case MPI_INT16_T:
    case MPI_INT32_T: case MPI_INT64_T: case MPI_UINT8_T: case MPI_UINT16_T: case MPI_UINT32_T:
    case MPI_UINT64_T: return true;
    default: return false;
    }
}
__host__ __device__ __forceinline__ bool type_is_float(int t) {
    return t == MPI_FLOAT || t == MPI_DOUBLE || t == MPI_LONG_DOUBLE;
}
__host__ __device__ __forceinline__ bool type_is_complex(int t) {
    return t == MPI_C_COMPLEX || t == MPI_C_FLOAT_COMPLEX || t == MPI_C_DOUBLE_COMPLEX ||
           t == MPI_C_LONG_DOUBLE_COMPLEX;
}

__host__ __device__ __forceinline__ bool op_allowed(int t, int op) {
    switch (op) {
    case kOpMax: case kOpMin: return type_is_integer(t) || type_is_float(t);
    case kOpSum: case kOpProd:
        return type_is_integer(t) || type_is_float(t) || type_is_complex(t);
    case kOpLand: case kOpLor: case kOpLxor: return type_is_integer(t) || t == MPI_C_BOOL;
    case kOpBand: case kOpBor: case kOpBxor: return type_is_integer(t) || t == MPI_BYTE;
    default: return false;
    }
}

template <class T> struct cplx { T re, im; };

template <class T, int OP>
__device__ __forceinline__ T apply(T in, T io) {
    if constexpr (OP == kOpMax) return in > io ? in : io;
    else if constexpr (OP == kOpMin) return in > io ? io : in;
    else if constexpr (OP == kOpSum) return (T)(in + io);
    else if constexpr (OP == kOpProd) return (T)(in * io);
    else if constexpr (OP == kOpLand) return (T)(in && io);
    else if constexpr (OP == kOpLor) return (T)(in || io);
    else if constexpr (OP == kOpLxor) return (T)(in != io);
    else if constexpr (OP == kOpBand) return (T)(in & io);
    else if constexpr (OP == kOpBor) return (T)(in | io);
    else return (T)(in ^ io);
}

template <class T> struct is_cplx { static constexpr bool value = false; };
template <class T> struct is_cplx<cplx<T>> { static constexpr bool value = true; };

template <class T, int OP>
__device__ __forceinline__ cplx<T> apply_cplx(cplx<T> a, cplx<T> b) {
    cplx<T> r;
    if constexpr (OP == kOpSum) {
        r.re = a.re + b.re;
        r.im = a.im + b.im;
    } else {
        T rr = a.re * b.re, ii = a.im * b.im, ri = a.re * b.im, ir = a.im * b.re;
        r.re = rr - ii;
        r.im = ri + ir;
    }
    return r;
}

// inout = op(in, inout) for any legal storage type.
template <class T, int OP>
__device__ __forceinline__ T combine(T in, T io) {
    if constexpr (is_cplx<T>::value) {
        return apply_cplx<decltype(in.re), OP>(in, io);
    } else {
        return apply<T, OP>(in, io);
    }
}

// Dispatch a functor templated on <storage type, op> for a legal (datatype, op) pair.
// F::template run<T, OP>() is invoked; returns false if the pair is illegal.
#define GPUMPI_OPS_INT(F, T)                                                        \
    switch (op) {                                                                   \
    case kOpMax: F.template run<T, kOpMax>(); return true;                          \
    case kOpMin: F.template run<T, kOpMin>(); return true;                          \
    case kOpSum: F.template run<T, kOpSum>(); return true;                          \
    case kOpProd: F.template run<T, kOpProd>(); return true;                        \
    case kOpLand: F.template run<T, kOpLand>(); return true;                        \
    case kOpLor: F.template run<T, kOpLor>(); return true;                          \
    case kOpLxor: F.template run<T, kOpLxor>(); return true;                        \
    case kOpBand: F.template run<T, kOpBand>(); return true;                        \
    case kOpBor: F.template run<T, kOpBor>(); return true;                          \
    case kOpBxor: F.template run<T, kOpBxor>(); return true;                        \
    default: return false;                                                          \
    }

template <class F>
__device__ bool dispatch_op(int t, int op, F& f) {
    if (!op_allowed(t, op)) return false;
    switch (t) {
    case MPI_SHORT: case MPI_INT16_T: GPUMPI_OPS_INT(f, int16_t)
    case MPI_UNSIGNED_SHORT: case MPI_UINT16_T: GPUMPI_OPS_INT(f, uint16_t)
    case MPI_INT: case MPI_INT32_T: GPUMPI_OPS_INT(f, int32_t)
    case MPI_UNSIGNED: case MPI_UINT32_T: GPUMPI_OPS_INT(f, uint32_t)
    case MPI_LONG: case MPI_LONG_LONG_INT: case MPI_LONG_LONG: case MPI_INT64_T:
        GPUMPI_OPS_INT(f, int64_t)
    case MPI_UNSIGNED_LONG: case MPI_UNSIGNED_LONG_LONG: case MPI_UINT64_T:
        GPUMPI_OPS_INT(f, uint64_t)
    case MPI_SIGNED_CHAR: case MPI_INT8_T: GPUMPI_OPS_INT(f, int8_t)
    case MPI_UNSIGNED_CHAR: case MPI_UINT8_T: case MPI_BYTE: case MPI_C_BOOL:
        GPUMPI_OPS_INT(f, uint8_t)
    case MPI_FLOAT:
        switch (op) {
        case kOpMax: f.template run<float, kOpMax>(); return true;
        case kOpMin: f.template run<float, kOpMin>(); return true;
        case kOpSum: f.template run<float, kOpSum>(); return true;
        default: f.template run<float, kOpProd>(); return true;
        }
    case MPI_DOUBLE: case MPI_LONG_DOUBLE:
        switch (op) {
        case kOpMax: f.template run<double, kOpMax>(); return true;
        case kOpMin: f.template run<double, kOpMin>(); return true;
        case kOpSum: f.template run<double, kOpSum>(); return true;
        default: f.template run<double, kOpProd>(); return true;
        }
    case MPI_C_COMPLEX: case MPI_C_FLOAT_COMPLEX:
        if (op == kOpSum) f.template run<cplx<float>, kOpSum>();
        else f.template run<cplx<float>, kOpProd>();
        return true;
    case MPI_C_DOUBLE_COMPLEX: case MPI_C_LONG_DOUBLE_COMPLEX:
        if (op == kOpSum) f.template run<cplx<double>, kOpSum>();
        else f.template run<cplx<double>, kOpProd>();
        return true;
    default: return false;
    }
}
#undef GPUMPI_OPS_INT

}  // namespace gpumpi
