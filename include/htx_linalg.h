// htx_linalg.h -- fixed-size column vectors and matrices for the host classes
// of the drop-in (Equations::Euler1D/2D, Schemes::Lagrange/Hermite/SSPRK).
//
// The reference's class interfaces pass small fixed-size vectors and
// matrices by value (Vector4d, Matrix4d, Matrix<double, 9, 1>, ...,
// Equations_Euler2D.h:8-15, Schemes_Lagrange.h:8-18).  This header gives
// those names a plain value type so that code written against the
// reference's interfaces compiles unchanged: element access v(i), m(i, j),
// comma initialisation `v << a, b, c`, sums and differences, scalar scaling,
// matrix products, dot, sum, setZero, cwiseProduct / cwiseMax / cwiseMin /
// cwiseInverse, transpose.  Storage is column-major; no expression
// templates, no dynamic sizes.  C++11.
#pragma once
#include <cmath>
#include <cstddef>
#include <ostream>

namespace htx {
namespace la {

template <typename S, int R, int C>
class Matrix {
public:
    Matrix() {
        for (int k = 0; k < R * C; ++k) a_[k] = S(0);
    }

    static constexpr int RowsAtCompileTime = R;
    static constexpr int ColsAtCompileTime = C;
    int rows() const { return R; }
    int cols() const { return C; }
    int size() const { return R * C; }

    S& operator()(int i, int j) { return a_[j * R + i]; }
    const S& operator()(int i, int j) const { return a_[j * R + i]; }
    // linear (vector) access
    S& operator()(int k) { return a_[k]; }
    const S& operator()(int k) const { return a_[k]; }
    S& operator[](int k) { return a_[k]; }
    const S& operator[](int k) const { return a_[k]; }
    S* data() { return a_; }
    const S* data() const { return a_; }

    void setZero() { fill(S(0)); }
    void setOnes() { fill(S(1)); }
    void setConstant(S v) { fill(v); }
    static Matrix Zero() { return Matrix(); }
    static Matrix Ones() {
        Matrix m;
        m.setOnes();
        return m;
    }
    static Matrix Identity() {
        Matrix m;
        for (int k = 0; k < (R < C ? R : C); ++k) m(k, k) = S(1);
        return m;
    }

    // comma initialisation in row-major reading order: m << a, b, c, ...
    class Inserter {
    public:
        Inserter(Matrix& m, int k) : m_(m), k_(k) {}
        Inserter& operator,(S v) {
            put(v);
            return *this;
        }
        void put(S v) {
            if (k_ < R * C) m_(k_ / C, k_ % C) = v;
            ++k_;
        }

    private:
        Matrix& m_;
        int k_;
    };
    Inserter operator<<(S v) {
        Inserter in(*this, 0);
        in.put(v);
        return in;
    }

    Matrix& operator+=(const Matrix& o) {
        for (int k = 0; k < R * C; ++k) a_[k] += o.a_[k];
        return *this;
    }
    Matrix& operator-=(const Matrix& o) {
        for (int k = 0; k < R * C; ++k) a_[k] -= o.a_[k];
        return *this;
    }
    Matrix& operator*=(S s) {
        for (int k = 0; k < R * C; ++k) a_[k] *= s;
        return *this;
    }
    Matrix& operator/=(S s) {
        for (int k = 0; k < R * C; ++k) a_[k] /= s;
        return *this;
    }

    S sum() const {
        S s = a_[0];
        for (int k = 1; k < R * C; ++k) s += a_[k];
        return s;
    }
    S dot(const Matrix& o) const {
        S s = a_[0] * o.a_[0];
        for (int k = 1; k < R * C; ++k) s += a_[k] * o.a_[k];
        return s;
    }
    S maxCoeff() const {
        S m = a_[0];
        for (int k = 1; k < R * C; ++k) m = a_[k] > m ? a_[k] : m;
        return m;
    }
    S minCoeff() const {
        S m = a_[0];
        for (int k = 1; k < R * C; ++k) m = a_[k] < m ? a_[k] : m;
        return m;
    }
    Matrix cwiseProduct(const Matrix& o) const { return zip(o, [](S x, S y) { return x * y; }); }
    Matrix cwiseQuotient(const Matrix& o) const { return zip(o, [](S x, S y) { return x / y; }); }
    Matrix cwiseMax(const Matrix& o) const { return zip(o, [](S x, S y) { return std::fmax(x, y); }); }
    Matrix cwiseMin(const Matrix& o) const { return zip(o, [](S x, S y) { return std::fmin(x, y); }); }
    Matrix cwiseMax(S v) const { return map([v](S x) { return std::fmax(x, v); }); }
    Matrix cwiseMin(S v) const { return map([v](S x) { return std::fmin(x, v); }); }
    Matrix cwiseAbs() const { return map([](S x) { return std::fabs(x); }); }
    Matrix cwiseInverse() const { return map([](S x) { return S(1) / x; }); }
    Matrix<S, C, R> transpose() const {
        Matrix<S, C, R> t;
        for (int i = 0; i < R; ++i)
            for (int j = 0; j < C; ++j) t(j, i) = (*this)(i, j);
        return t;
    }

private:
    void fill(S v) {
        for (int k = 0; k < R * C; ++k) a_[k] = v;
    }
    template <typename F>
    Matrix map(F f) const {
        Matrix r;
        for (int k = 0; k < R * C; ++k) r.a_[k] = f(a_[k]);
        return r;
    }
    template <typename F>
    Matrix zip(const Matrix& o, F f) const {
        Matrix r;
        for (int k = 0; k < R * C; ++k) r.a_[k] = f(a_[k], o.a_[k]);
        return r;
    }
    S a_[R * C];
};

template <typename S, int R, int C>
Matrix<S, R, C> operator+(Matrix<S, R, C> a, const Matrix<S, R, C>& b) {
    return a += b;
}
template <typename S, int R, int C>
Matrix<S, R, C> operator-(Matrix<S, R, C> a, const Matrix<S, R, C>& b) {
    return a -= b;
}
template <typename S, int R, int C>
Matrix<S, R, C> operator-(const Matrix<S, R, C>& a) {
    Matrix<S, R, C> r;
    for (int k = 0; k < R * C; ++k) r(k) = -a(k);
    return r;
}
template <typename S, int R, int C>
Matrix<S, R, C> operator*(Matrix<S, R, C> a, S s) {
    return a *= s;
}
template <typename S, int R, int C>
Matrix<S, R, C> operator*(S s, const Matrix<S, R, C>& a) {
    Matrix<S, R, C> r;
    for (int k = 0; k < R * C; ++k) r(k) = s * a(k);
    return r;
}
template <typename S, int R, int C>
Matrix<S, R, C> operator/(Matrix<S, R, C> a, S s) {
    return a /= s;
}
// matrix product; each entry sums over k in ascending order
template <typename S, int R, int K, int C>
Matrix<S, R, C> operator*(const Matrix<S, R, K>& a, const Matrix<S, K, C>& b) {
    Matrix<S, R, C> r;
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < C; ++j) {
            S s = a(i, 0) * b(0, j);
            for (int k = 1; k < K; ++k) s += a(i, k) * b(k, j);
            r(i, j) = s;
        }
    return r;
}

template <typename S, int R, int C>
std::ostream& operator<<(std::ostream& os, const Matrix<S, R, C>& m) {
    for (int i = 0; i < R; ++i) {
        for (int j = 0; j < C; ++j) os << (j ? " " : "") << m(i, j);
        if (i + 1 < R) os << "\n";
    }
    return os;
}

}  // namespace la
}  // namespace htx

// the names the reference's headers bring into the global namespace
using htx::la::Matrix;
typedef Matrix<double, 2, 1> Vector2d;
typedef Matrix<double, 3, 1> Vector3d;
typedef Matrix<double, 4, 1> Vector4d;
typedef Matrix<double, 5, 1> Vector5d;
typedef Matrix<double, 6, 1> Vector6d;
typedef Matrix<double, 8, 1> Vector8d;
typedef Matrix<double, 9, 1> Vector9d;
typedef Matrix<double, 12, 1> Vector12d;
typedef Matrix<double, 2, 2> Matrix2d;
typedef Matrix<double, 3, 3> Matrix3d;
typedef Matrix<double, 4, 4> Matrix4d;
