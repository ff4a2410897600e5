// ki_Mat: host value type of the drop-in API (interface of the reference's
// kern_interp/ki_mat.h:19-101).  Dense helpers are written out here (no
// BLAS/LAPACK dependency); the factor/solve path runs on the GPU.
#include "kern_interp/ki_mat.h"

#include "runtime.h"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace kern_interp {

ki_Mat::ki_Mat() : mat(nullptr), lda_(0), height_(0), width_(0) {}

ki_Mat::~ki_Mat() { delete[] mat; }

ki_Mat::ki_Mat(int h, int w) : lda_(h), height_(h), width_(w) {
  mat = new double[(size_t)h * w]();
}

ki_Mat::ki_Mat(const ki_Mat& o) : lda_(o.lda_), height_(o.height_), width_(o.width_) {
  mat = new double[(size_t)height_ * width_];
  if (height_ * width_) std::memcpy(mat, o.mat, sizeof(double) * height_ * width_);
}

ki_Mat& ki_Mat::operator=(const ki_Mat& o) {
  if (this == &o) return *this;
  if (height_ * width_ != o.height_ * o.width_ || !mat) {
    delete[] mat;
    mat = new double[(size_t)o.height_ * o.width_];
  }
  lda_ = o.lda_;
  height_ = o.height_;
  width_ = o.width_;
  if (height_ * width_) std::memcpy(mat, o.mat, sizeof(double) * height_ * width_);
  return *this;
}

ki_Mat::ki_Mat(ki_Mat&& o) : mat(o.mat), lda_(o.lda_), height_(o.height_), width_(o.width_) {
  o.mat = nullptr;
  o.lda_ = o.height_ = o.width_ = 0;
}

ki_Mat& ki_Mat::operator=(ki_Mat&& o) {
  if (this == &o) return *this;
  delete[] mat;
  mat = o.mat;
  lda_ = o.lda_;
  height_ = o.height_;
  width_ = o.width_;
  o.mat = nullptr;
  o.lda_ = o.height_ = o.width_ = 0;
  return *this;
}

double ki_Mat::get(int i, int j) const {
  assert(i < height_ && j < width_ && mat);
  return mat[i + (size_t)lda_ * j];
}

void ki_Mat::set(int i, int j, double a) {
  assert(i < height_ && j < width_ && mat);
  mat[i + (size_t)lda_ * j] = a;
}

void ki_Mat::addset(int i, int j, double a) {
  assert(i < height_ && j < width_ && mat);
  mat[i + (size_t)lda_ * j] += a;
}

void ki_Mat::set_submatrix(const std::vector<int>& I_, const std::vector<int>& J_,
                           const ki_Mat& A, bool transpose_A) {
  for (size_t j = 0; j < J_.size(); ++j)
    for (size_t i = 0; i < I_.size(); ++i)
      set(I_[i], J_[j], transpose_A ? A.get(j, i) : A.get(i, j));
}

void ki_Mat::set_submatrix(int row_s, int row_e, int col_s, int col_e, const ki_Mat& A,
                           bool transpose_A, bool) {
  for (int j = 0; j < col_e - col_s; ++j)
    for (int i = 0; i < row_e - row_s; ++i)
      set(row_s + i, col_s + j, transpose_A ? A.get(j, i) : A.get(i, j));
}

void ki_Mat::set_submatrix(const std::vector<int>& I_, int col_s, int col_e, const ki_Mat& A,
                           bool transpose_A) {
  for (int j = 0; j < col_e - col_s; ++j)
    for (size_t i = 0; i < I_.size(); ++i)
      set(I_[i], col_s + j, transpose_A ? A.get(j, i) : A.get(i, j));
}

void ki_Mat::set_submatrix(int row_s, int row_e, const std::vector<int>& J_, const ki_Mat& A,
                           bool transpose_A) {
  for (size_t j = 0; j < J_.size(); ++j)
    for (int i = 0; i < row_e - row_s; ++i)
      set(row_s + i, J_[j], transpose_A ? A.get(j, i) : A.get(i, j));
}

void ki_Mat::transpose_into(ki_Mat* t) const {
  if (t->height_ != width_ || t->width_ != height_) *t = ki_Mat(width_, height_);
  for (int j = 0; j < width_; ++j)
    for (int i = 0; i < height_; ++i) t->mat[j + (size_t)t->lda_ * i] = get(i, j);
}

void ki_Mat::eye(int n) {
  if (height_ != n || width_ != n) *this = ki_Mat(n, n);
  for (int i = 0; i < n; ++i) set(i, i, 1.0);
}

ki_Mat ki_Mat::transpose() const {
  ki_Mat t(width_, height_);
  transpose_into(&t);
  return t;
}

int ki_Mat::height() const { return height_; }
int ki_Mat::width() const { return width_; }

ki_Mat& ki_Mat::operator-=(const ki_Mat& o) {
  assert(o.height_ == height_ && o.width_ == width_);
  for (int j = 0; j < width_; ++j)
    for (int i = 0; i < height_; ++i) mat[i + (size_t)lda_ * j] -= o.get(i, j);
  return *this;
}

ki_Mat& ki_Mat::operator+=(const ki_Mat& o) {
  assert(o.height_ == height_ && o.width_ == width_);
  for (int j = 0; j < width_; ++j)
    for (int i = 0; i < height_; ++i) mat[i + (size_t)lda_ * j] += o.get(i, j);
  return *this;
}

ki_Mat& ki_Mat::operator*=(double o) {
  for (size_t e = 0; e < (size_t)height_ * width_; ++e) mat[e] *= o;
  return *this;
}

ki_Mat& ki_Mat::operator/=(double o) {
  assert(std::abs(o) > 1e-8 && "Error: divide matrix by 0.");
  for (size_t e = 0; e < (size_t)height_ * width_; ++e) mat[e] /= o;
  return *this;
}

ki_Mat ki_Mat::operator-() const { return (*this) * -1.0; }

ki_Mat ki_Mat::operator-(const ki_Mat& o) const {
  ki_Mat r(*this);
  r -= o;
  return r;
}

ki_Mat ki_Mat::operator+(const ki_Mat& o) const {
  ki_Mat r(*this);
  r += o;
  return r;
}

ki_Mat ki_Mat::operator*(const ki_Mat& o) const {
  assert(width_ == o.height_);
  ki_Mat r(height_, o.width_);
  for (int j = 0; j < o.width_; ++j)
    for (int t = 0; t < width_; ++t) {
      const double b = o.mat[t + (size_t)o.lda_ * j];
      if (b == 0.0) continue;
      const double* a = mat + (size_t)lda_ * t;
      double* c = r.mat + (size_t)r.lda_ * j;
      for (int i = 0; i < height_; ++i) c[i] += a[i] * b;
    }
  return r;
}

ki_Mat ki_Mat::operator*(double o) const {
  ki_Mat r(*this);
  r *= o;
  return r;
}

ki_Mat ki_Mat::operator/(double o) const {
  ki_Mat r(*this);
  r /= o;
  return r;
}

ki_Mat ki_Mat::operator()(const std::vector<int>& I_, const std::vector<int>& J_) const {
  ki_Mat r(I_.size(), J_.size());
  for (size_t j = 0; j < J_.size(); ++j)
    for (size_t i = 0; i < I_.size(); ++i) r.mat[i + r.lda_ * j] = get(I_[i], J_[j]);
  return r;
}

ki_Mat ki_Mat::operator()(const std::vector<int>& I_, int col_s, int col_e) const {
  ki_Mat r(I_.size(), col_e - col_s);
  for (int j = 0; j < col_e - col_s; ++j)
    for (size_t i = 0; i < I_.size(); ++i) r.mat[i + r.lda_ * j] = get(I_[i], col_s + j);
  return r;
}

ki_Mat ki_Mat::operator()(int row_s, int row_e, const std::vector<int>& J_) const {
  ki_Mat r(row_e - row_s, J_.size());
  for (size_t j = 0; j < J_.size(); ++j)
    for (int i = 0; i < row_e - row_s; ++i) r.mat[i + r.lda_ * j] = get(row_s + i, J_[j]);
  return r;
}

ki_Mat ki_Mat::operator()(int row_s, int row_e, int col_s, int col_e) const {
  ki_Mat r(row_e - row_s, col_e - col_s);
  for (int j = 0; j < col_e - col_s; ++j)
    for (int i = 0; i < row_e - row_s; ++i) r.mat[i + r.lda_ * j] = get(row_s + i, col_s + j);
  return r;
}

double ki_Mat::frob_norm() const {
  double s = 0;
  for (int j = 0; j < width_; ++j)
    for (int i = 0; i < height_; ++i) s += get(i, j) * get(i, j);
  return std::sqrt(s);
}

double ki_Mat::max_entry_magnitude() const {
  double m = 0;
  for (int j = 0; j < width_; ++j)
    for (int i = 0; i < height_; ++i) m = std::max(m, std::abs(get(i, j)));
  return m;
}

double ki_Mat::one_norm() const {
  double top = 0;
  for (int j = 0; j < width_; ++j) {
    double s = 0;
    for (int i = 0; i < height_; ++i) s += std::fabs(get(i, j));
    top = std::max(top, s);
  }
  return top;
}

double ki_Mat::vec_two_norm() const {
  assert(width() == 1);
  return frob_norm();
}

void ki_Mat::LU_factorize(ki_Mat* K_LU, std::vector<lapack_int>* piv) const {
  *K_LU = *this;
  const int n = std::min(height_, width_);
  piv->assign(n, 0);
  ki_Mat& A = *K_LU;
  for (int c = 0; c < n; ++c) {
    int p = c;
    double best = std::fabs(A.get(c, c));
    for (int i = c + 1; i < height_; ++i)
      if (std::fabs(A.get(i, c)) > best) {
        best = std::fabs(A.get(i, c));
        p = i;
      }
    (*piv)[c] = p + 1;
    if (p != c)
      for (int j = 0; j < width_; ++j) std::swap(A.mat[c + (size_t)A.lda_ * j],
                                                 A.mat[p + (size_t)A.lda_ * j]);
    const double d = A.get(c, c);
    if (d == 0.0) continue;
    for (int i = c + 1; i < height_; ++i) A.mat[i + (size_t)A.lda_ * c] /= d;
    for (int j = c + 1; j < width_; ++j) {
      const double u = A.get(c, j);
      for (int i = c + 1; i < height_; ++i)
        A.mat[i + (size_t)A.lda_ * j] -= A.get(i, c) * u;
    }
  }
}

// in place: B <- op(LU)^{-1} B
static void lu_solve(const ki_Mat& LU, const std::vector<lapack_int>& piv, ki_Mat* B, bool trans) {
  const int n = LU.height();
  for (int col = 0; col < B->width(); ++col) {
    double* b = B->mat + (size_t)B->lda_ * col;
    if (!trans) {
      for (int c = 0; c < n; ++c) std::swap(b[c], b[piv[c] - 1]);
      for (int i = 0; i < n; ++i)
        for (int t = 0; t < i; ++t) b[i] -= LU.get(i, t) * b[t];
      for (int i = n - 1; i >= 0; --i) {
        for (int t = i + 1; t < n; ++t) b[i] -= LU.get(i, t) * b[t];
        b[i] /= LU.get(i, i);
      }
    } else {
      for (int i = 0; i < n; ++i) {
        for (int t = 0; t < i; ++t) b[i] -= LU.get(t, i) * b[t];
        b[i] /= LU.get(i, i);
      }
      for (int i = n - 1; i >= 0; --i)
        for (int t = i + 1; t < n; ++t) b[i] -= LU.get(t, i) * b[t];
      for (int c = n - 1; c >= 0; --c) std::swap(b[c], b[piv[c] - 1]);
    }
  }
}

void ki_Mat::left_multiply_inverse(const ki_Mat& K, ki_Mat* U) const {
  ki_Mat lu;
  std::vector<lapack_int> piv;
  LU_factorize(&lu, &piv);
  *U = K;
  lu_solve(lu, piv, U, false);
}

void ki_Mat::right_multiply_inverse(const ki_Mat& K, ki_Mat* L) const {
  ki_Mat lu;
  std::vector<lapack_int> piv;
  LU_factorize(&lu, &piv);
  ki_Mat kt = K.transpose();
  lu_solve(lu, piv, &kt, true);
  *L = kt.transpose();
}

void ki_Mat::left_multiply_inverse(const ki_Mat& K, const std::vector<lapack_int>& piv,
                                   ki_Mat* U) const {
  *U = K;
  lu_solve(*this, piv, U, false);
}

void ki_Mat::right_multiply_inverse(const ki_Mat& K, const std::vector<lapack_int>& piv,
                                    ki_Mat* L) const {
  ki_Mat kt = K.transpose();
  lu_solve(*this, piv, &kt, true);
  *L = kt.transpose();
}

// ki_mat.cpp:505-541: pivoted QR to the first |R_ii| < tol |R_00| (i >= 1),
// then Z = R11^{-1} R12; run on the GPU by the factorization's own routine
// (kigpu_interp_decomp).  Returns the number of skeleton columns, 0 when no
// |R_ii| falls below the threshold.
int ki_Mat::id(std::vector<int>* p, ki_Mat* Z, double tol) const {
  assert(height() != 0 && width() != 0);
  const int m = height_, n = width_;
  int k = 0;
  std::vector<int> perm(n);
  std::vector<double> t((size_t)n * n);
  detail::check(kigpu_interp_decomp(detail::context(), mat, m, n, tol, &k, perm.data(), t.data()),
                "kigpu_interp_decomp");
  if (k == 0) return 0;
  for (int j = 0; j < n; ++j) p->push_back(perm[j]);
  *Z = ki_Mat(k, n - k);
  for (int j = 0; j < n - k; ++j)
    for (int i = 0; i < k; ++i) Z->set(i, j, t[i + (size_t)j * k]);
  return k;
}


// Householder reduction to upper Hessenberg form, then Francis double-shift
// QR steps on the active block (Golub & Van Loan, Matrix Computations,
// algorithms 7.4.2 and 7.5.2), deflating at negligible subdiagonals; only the
// block is updated, as no Schur vectors are wanted.
std::vector<double> ki_Mat::real_eigenvalues() {
  assert(height_ == width_);
  const int n = width_;
  std::vector<double> H((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) H[i + (size_t)n * j] = get(i, j);
  auto h = [&](int i, int j) -> double& { return H[i + (size_t)n * j]; };
  // v^T-reflect rows r0..r0+m-1 over columns c0..c1, then columns over rows q0..q1
  auto reflect = [&](const double* v, int m, int r0, int c0, int c1, int q0, int q1) {
    double vv = 0;
    for (int a = 0; a < m; ++a) vv += v[a] * v[a];
    if (vv == 0) return;
    for (int j = c0; j <= c1; ++j) {
      double s = 0;
      for (int a = 0; a < m; ++a) s += v[a] * h(r0 + a, j);
      s *= 2 / vv;
      for (int a = 0; a < m; ++a) h(r0 + a, j) -= s * v[a];
    }
    for (int i = q0; i <= q1; ++i) {
      double s = 0;
      for (int a = 0; a < m; ++a) s += h(i, r0 + a) * v[a];
      s *= 2 / vv;
      for (int a = 0; a < m; ++a) h(i, r0 + a) -= s * v[a];
    }
  };
  auto house = [](double* v, int m) {  // v <- x + sign(x0) |x| e0
    double nx = 0;
    for (int a = 0; a < m; ++a) nx += v[a] * v[a];
    nx = std::sqrt(nx);
    v[0] += v[0] >= 0 ? nx : -nx;
  };
  std::vector<double> v(std::max(n, 3));
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    for (int a = 0; a < m; ++a) v[a] = h(k + 1 + a, k);
    house(v.data(), m);
    reflect(v.data(), m, k + 1, 0, n - 1, 0, n - 1);
  }
  std::vector<double> re;
  const double tiny = 1e-14, ulp = 2.220446049250313e-16;
  int hi = n - 1, iter = 0;
  while (hi >= 0) {
    int l = hi;
    while (l > 0 && std::fabs(h(l, l - 1)) > ulp * (std::fabs(h(l, l)) + std::fabs(h(l - 1, l - 1))))
      --l;
    if (l == hi) {
      re.push_back(h(hi, hi));
      --hi;
      iter = 0;
      continue;
    }
    if (l == hi - 1) {
      const double a = h(l, l), b = h(l, hi), c = h(hi, l), d = h(hi, hi);
      const double p = 0.5 * (a - d), disc = p * p + b * c;
      if (disc >= 0) {
        const double sq = std::sqrt(disc), m = 0.5 * (a + d);
        const double e1 = m + (p >= 0 ? sq : -sq);
        // the other root from the determinant, unless e1 is zero
        const double e2 = e1 != 0 ? (a * d - b * c) / e1 : m - (p >= 0 ? sq : -sq);
        re.push_back(e1);
        re.push_back(e2);
      } else if (std::sqrt(-disc) < tiny) {
        re.push_back(0.5 * (a + d));
        re.push_back(0.5 * (a + d));
      }
      hi -= 2;
      iter = 0;
      continue;
    }
    if (++iter > 60 * n) throw std::runtime_error("real_eigenvalues: QR iteration did not converge");
    double s = h(hi - 1, hi - 1) + h(hi, hi);
    double t = h(hi - 1, hi - 1) * h(hi, hi) - h(hi - 1, hi) * h(hi, hi - 1);
    if (iter % 11 == 10) {  // exceptional shift
      const double w = std::fabs(h(hi, hi - 1)) + std::fabs(h(hi - 1, hi - 2));
      s = 1.5 * w;
      t = w * w;
    }
    double x = h(l, l) * h(l, l) + h(l, l + 1) * h(l + 1, l) - s * h(l, l) + t;
    double y = h(l + 1, l) * (h(l, l) + h(l + 1, l + 1) - s);
    double z = h(l + 1, l) * h(l + 2, l + 1);
    for (int k = l; k <= hi - 2; ++k) {
      double w3[3] = {x, y, z};
      house(w3, 3);
      reflect(w3, 3, k, std::max(l, k - 1), hi, l, std::min(k + 3, hi));
      x = h(k + 1, k);
      y = h(k + 2, k);
      if (k < hi - 2) z = h(k + 3, k);
    }
    double w2[2] = {x, y};
    house(w2, 2);
    reflect(w2, 2, hi - 1, hi - 2, hi, l, hi);
  }
  return re;
}

}  // namespace kern_interp
