// Small dense FP64 SVD used by the IRLBA restart.
#pragma once

namespace bpca {

int golub_reinsch_svd(int rows, int cols, double *A, double *S, double *V);
int bidiag_svd(int m, const double *B, double *Ub, double *Sb, double *Vbt);

}  // namespace bpca
