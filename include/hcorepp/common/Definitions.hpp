// Drop-in for include/hcorepp/common/Definitions.hpp of ecrc/hcorepp: the enums the
// TLR GEMM and Cholesky-flow surface uses.  blas::Op / Layout / Uplo / Side / Diag are
// defined here unless BLAS++'s blas.hh was included first (same enumerator values).
#pragma once

#include <cstddef>
#include <cstdint>

#ifndef BLAS_HH
namespace blas {
enum class Op : char { NoTrans = 'N', Trans = 'T', ConjTrans = 'C' };
enum class Layout : char { ColMajor = 'C', RowMajor = 'R' };
enum class Uplo : char { Upper = 'U', Lower = 'L', General = 'G' };
enum class Side : char { Left = 'L', Right = 'R' };
enum class Diag : char { NonUnit = 'N', Unit = 'U' };
}  // namespace blas
#endif

namespace hcorepp {
namespace common {
// reference include/hcorepp/common/Definitions.hpp:101-104
enum class CompressionType { LAPACK_GESVD = 0, LAPACK_GESDD = 1 };
}  // namespace common
}  // namespace hcorepp
