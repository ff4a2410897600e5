// Drop-in for kern_interp/ki_mat.h (reference ki_mat.h:19-101): owning
// column-major FP64 matrix with lda = height.  Host value type: its dense
// helpers (GEMM, LU, ID) are plain host arithmetic; the factor/solve path
// itself runs on the GPU through include/kigpu.h.
#ifndef KERN_INTERP_KI_MAT_H_
#define KERN_INTERP_KI_MAT_H_

#include <vector>

namespace kern_interp {

typedef int lapack_int;  // LAPACKE's index type in the reference headers

struct ki_Mat {
  double* mat;
  int lda_, height_, width_;

  ki_Mat();
  ~ki_Mat();
  ki_Mat(int h, int w);
  ki_Mat(const ki_Mat& o);
  ki_Mat& operator=(const ki_Mat& copy);
  ki_Mat(ki_Mat&& move);
  ki_Mat& operator=(ki_Mat&& move);

  double get(int i, int j) const;
  void set(int i, int j, double a);
  void addset(int i, int j, double a);
  void set_submatrix(const std::vector<int>& I_, const std::vector<int>& J_, const ki_Mat& A,
                     bool transpose_A = false);
  void set_submatrix(int row_s, int row_e, int col_s, int col_e, const ki_Mat& A,
                     bool transpose_A = false, bool timing = false);
  void set_submatrix(const std::vector<int>& I_, int col_s, int col_e, const ki_Mat& A,
                     bool transpose_A = false);
  void set_submatrix(int row_s, int row_e, const std::vector<int>& J_, const ki_Mat& A,
                     bool transpose_A = false);

  void transpose_into(ki_Mat* transpose) const;
  void eye(int n);
  ki_Mat transpose() const;

  int height() const;
  int width() const;

  ki_Mat& operator-=(const ki_Mat& o);
  ki_Mat& operator+=(const ki_Mat& o);
  ki_Mat& operator*=(double o);
  ki_Mat& operator/=(double o);

  ki_Mat operator-() const;
  ki_Mat operator-(const ki_Mat& o) const;
  ki_Mat operator+(const ki_Mat& o) const;
  ki_Mat operator*(const ki_Mat& o) const;
  ki_Mat operator*(double o) const;
  ki_Mat operator/(double o) const;

  ki_Mat operator()(const std::vector<int>& I_, const std::vector<int>& J_) const;
  ki_Mat operator()(const std::vector<int>& I_, int col_s, int col_e) const;
  ki_Mat operator()(int row_s, int row_e, const std::vector<int>& J_) const;
  ki_Mat operator()(int row_s, int row_e, int col_s, int col_e) const;

  double one_norm() const;
  double vec_two_norm() const;
  double frob_norm() const;
  double max_entry_magnitude() const;

  // LU with partial pivoting; piv is 1-based like LAPACK's dgetrf
  void LU_factorize(ki_Mat* K_LU, std::vector<lapack_int>* piv) const;
  void left_multiply_inverse(const ki_Mat& K, ki_Mat* U) const;
  void right_multiply_inverse(const ki_Mat& K, ki_Mat* L) const;
  void left_multiply_inverse(const ki_Mat& K, const std::vector<lapack_int>& piv,
                             ki_Mat* U) const;
  void right_multiply_inverse(const ki_Mat& K, const std::vector<lapack_int>& piv,
                              ki_Mat* L) const;

  // interpolative decomposition by Householder QR with column pivoting;
  // returns the number of skeleton columns (0: no compression)
  int id(std::vector<int>* p, ki_Mat* Z, double tol) const;

  // eigenvalues of a square matrix whose imaginary part is below 1e-14
  // (reference ki_mat.cpp:544-560, LAPACK dgeev there), by Hessenberg
  // reduction and Francis double-shift QR; the matrix is left unchanged
  std::vector<double> real_eigenvalues();
};

}  // namespace kern_interp

#endif  // KERN_INTERP_KI_MAT_H_
