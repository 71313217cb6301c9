// Host-side helpers of the reference's native module that its ctypes binding
// sets argtypes on (CPET/utils/c_ops.py:26-77): sparse and dense
// matrix-vector products, a vector add, row sums and the k-weighted einsum
// reductions.  None of them is on a hot path (the reference's own drivers
// call only the field, potential and streamline kernels), so they run on the
// host, sequentially, in the element order of the reference's loops and in
// the reference's arithmetic types (CPET/utils/math_module.c:15-252).  They
// exist so that the reference's Math_ops() constructor binds this library
// with only the library path changed.
//
// Argument validation follows the rest of the C-ABI: a negative size or a
// NULL array that would be read or written returns CPET_ERR_ARG with a
// message in cpet_last_error(); 0 means success.  The reference functions
// return void; ctypes callers that declare restype None are unaffected.

#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "../../include/cpet_hip.h"

namespace cpet {
void set_error(const char* fmt, ...);
}

namespace {

// k in eV*A/e^2 as the reference's einsum kernels hold it: a float
// initialised from 14.3996451 (math_module.c:138,178).
constexpr float kCoulombF = 14.3996451f;

bool bad(bool cond, const char* fn, const char* what) {
  if (cond) cpet::set_error("%s: %s", fn, what);
  return cond;
}

}  // namespace

extern "C" {

// ret[r] = sum_{j in [indptr[r], indptr[r+1])} A[j] * B[indA[j]] for the
// indptrlen - 1 rows of a CSR matrix, FP64.  math_module.c:15-46 (its OpenMP
// region makes every thread walk all rows; the row results are these).
int sparse_dot(double* ret, const int* indptr, int indptrlen, const int* indA, int lenindA,
               const double* A, int lenA, const double* B, int size_array) {
  const char* fn = "sparse_dot";
  if (bad(indptrlen < 1 || lenindA < 0 || lenA < 0 || size_array < 0, fn, "negative size"))
    return CPET_ERR_ARG;
  if (indptrlen == 1) return CPET_OK;
  if (bad(!ret || !indptr, fn, "NULL array")) return CPET_ERR_ARG;
  for (int i = 1; i < indptrlen; ++i) {
    const int a = indptr[i - 1], b = indptr[i];
    if (bad(a < 0 || b < a || b > lenA || b > lenindA, fn, "indptr out of range"))
      return CPET_ERR_ARG;
    if (b > a && bad(!A || !indA || !B, fn, "NULL array")) return CPET_ERR_ARG;
    double sum = 0.0;
    for (int j = a; j < b; ++j) {
      const int c = indA[j];
      if (bad(c < 0 || c >= size_array, fn, "column index out of range")) return CPET_ERR_ARG;
      sum += A[j] * B[c];
    }
    ret[i - 1] = sum;
  }
  return CPET_OK;
}

// ret[i] = sum_j A[i*cols + j] * B[j], FP64.  math_module.c:49-68.
int dot(double* ret, const double* A, const double* B, int rows, int cols) {
  if (bad(rows < 0 || cols < 0, "dot", "negative size")) return CPET_ERR_ARG;
  if (rows == 0) return CPET_OK;
  if (bad(!ret || (cols > 0 && (!A || !B)), "dot", "NULL array")) return CPET_ERR_ARG;
  for (int i = 0; i < rows; ++i) {
    double sum = 0.0;
    for (int j = 0; j < cols; ++j) sum += A[(size_t)i * cols + j] * B[j];
    ret[i] = sum;
  }
  return CPET_OK;
}

// ret = A + B elementwise, float32.  math_module.c:124-132.
int vecaddn(float* ret, const float* A, const float* B, int lenA) {
  if (bad(lenA < 0, "vecaddn", "negative size")) return CPET_ERR_ARG;
  if (lenA == 0) return CPET_OK;
  if (bad(!ret || !A || !B, "vecaddn", "NULL array")) return CPET_ERR_ARG;
  for (int i = 0; i < lenA; ++i) ret[i] = A[i] + B[i];
  return CPET_OK;
}

// result[k] = sum_i (k_C * Q[i] * r_mag[k][i]) * R[k][i][:], float32
// accumulation in row order.  math_module.c:135-172.
int einsum_operation_batch(int batch, int rows, const float* r_mag, const float* Q,
                           const float* R, float* result) {
  const char* fn = "einsum_operation_batch";
  if (bad(batch < 0 || rows < 0, fn, "negative size")) return CPET_ERR_ARG;
  if (batch == 0) return CPET_OK;
  if (bad(!result || (rows > 0 && (!r_mag || !Q || !R)), fn, "NULL array")) return CPET_ERR_ARG;
  for (int k = 0; k < batch; ++k) {
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    for (int i = 0; i < rows; ++i) {
      const size_t ki = (size_t)k * rows + i;
      const float c = kCoulombF * Q[i] * r_mag[ki];
      s0 += c * R[3 * ki];
      s1 += c * R[3 * ki + 1];
      s2 += c * R[3 * ki + 2];
    }
    result[3 * (size_t)k] = s0;
    result[3 * (size_t)k + 1] = s1;
    result[3 * (size_t)k + 2] = s2;
  }
  return CPET_OK;
}

// The single-point form of einsum_operation_batch.  math_module.c:175-209.
int einsum_operation(int rows, const float* r_mag, const float* Q, const float* R,
                     float* result) {
  if (bad(rows < 0, "einsum_operation", "negative size")) return CPET_ERR_ARG;
  if (bad(!result, "einsum_operation", "NULL array")) return CPET_ERR_ARG;
  return einsum_operation_batch(1, rows, r_mag, Q, R, result);
}

// ret[k][i] = sum_j A[k][i][j], float32 accumulation.  math_module.c:212-232.
int einsum_ij_i_batch(int batch, int rows, int cols, const float* A, float* ret) {
  const char* fn = "einsum_ij_i_batch";
  if (bad(batch < 0 || rows < 0 || cols < 0, fn, "negative size")) return CPET_ERR_ARG;
  if ((size_t)batch * rows == 0) return CPET_OK;
  if (bad(!ret || (cols > 0 && !A), fn, "NULL array")) return CPET_ERR_ARG;
  for (size_t r = 0; r < (size_t)batch * rows; ++r) {
    float sum = 0.f;
    for (int j = 0; j < cols; ++j) sum += A[r * cols + j];
    ret[r] = sum;
  }
  return CPET_OK;
}

// ret[i] = sum_j A[i][j], float32 accumulation.  math_module.c:235-252.
int einsum_ij_i(int rows, int cols, const float* A, float* ret) {
  return einsum_ij_i_batch(1, rows, cols, A, ret);
}

}  // extern "C"
