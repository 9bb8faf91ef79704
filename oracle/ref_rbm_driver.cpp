// Thin C-ABI driver over the reference's own rbm_math.h (included unmodified
// from the reference tree at build time).  Test infrastructure only: it lets
// the tests compare the oracle's restated AVOS arithmetic and SpGEMM against
// the reference's compiled code.  Load with ctypes.PyDLL so OverflowError set
// by the reference's PyErr_Format surfaces as a Python exception.
#include <Python.h>
#include <numpy/ndarraytypes.h>
#include <stdexcept>
#include "rbm_math.h"

extern "C" {
unsigned long long ref_avos_product_ll(long long a, long long b) {
  return avos_product<npy_longlong, npy_ulonglong>(a, b);
}
long long ref_avos_sum_ll(long long a, long long b) {
  return avos_sum<npy_longlong, npy_ulonglong>(a, b);
}
unsigned int ref_avos_product_i(int a, int b) { return avos_product<npy_int, npy_uint>(a, b); }
int ref_avos_sum_i(int a, int b) { return avos_sum<npy_int, npy_uint>(a, b); }

void ref_matmat_pass1(int n_row, int n_col, const int *Ap, const int *Aj, const int *Bp,
                      const int *Bj, int *Cp) {
  rbm_matmat_pass1<int>(n_row, n_col, Ap, Aj, Bp, Bj, Cp);
}
void ref_matmat_pass2(int n_row, int n_col, const int *Ap, const int *Aj, const long long *Ax,
                      const int *Bp, const int *Bj, const long long *Bx, int *Cp, int *Cj,
                      long long *Cx) {
  rbm_matmat_pass2<int, npy_longlong, npy_ulonglong>(n_row, n_col, Ap, Aj, Ax, Bp, Bj, Bx, Cp,
                                                     Cj, Cx);
}
}
