// Drop-in for include/hcorepp/common/Definitions.hpp of ecrc/hcorepp: the enums the
// TLR GEMM surface uses.  blas::Op / blas::Layout are defined here unless BLAS++'s
// blas.hh was included first (same enumerator values: 'N','T','C' and 'C','R').
#pragma once

#include <cstddef>
#include <cstdint>

#ifndef BLAS_HH
namespace blas {
enum class Op : char { NoTrans = 'N', Trans = 'T', ConjTrans = 'C' };
enum class Layout : char { ColMajor = 'C', RowMajor = 'R' };
}  // namespace blas
#endif

namespace hcorepp {
namespace common {
// reference include/hcorepp/common/Definitions.hpp:101-104
enum class CompressionType { LAPACK_GESVD = 0, LAPACK_GESDD = 1 };
}  // namespace common
}  // namespace hcorepp
