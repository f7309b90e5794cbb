// clm_box.hpp -- host-side Box construction in the working precision T.
//
// Follows src/internals/Box.jl:191-270 (Box, _construct_box, _compute_nc_and_cell_size),
// src/internals/CellOperations.jl:337-479 (align_cell, cell_limits) and Box.jl:579-636
// (check_unit_cell).  Arithmetic is written in the evaluation order of the Julia /
// StaticArrays expressions (left-to-right sums, no fused multiply-add; this file must
// be compiled with -ffp-contract=off), because the device binning kernels consume these
// numbers and neighbour-list membership at the cutoff edge depends on their last bit.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace clm {

enum CellType { ORTHO = 0, TRICLINIC = 1, NONPERIODIC = 2 };
enum BoxError { BOX_OK = 0, BOX_BAD_LCELL = 1, BOX_BAD_CELL = 2, BOX_BAD_DIM = 3 };

// 3x3 column-major storage; element (i, j) is m[i + 3 * j].
template <typename T>
struct Mat3 {
    T a[9];
    T &operator()(int i, int j) { return a[i + 3 * j]; }
    T operator()(int i, int j) const { return a[i + 3 * j]; }
    static Mat3 eye() {
        Mat3 m;
        for (int k = 0; k < 9; ++k) m.a[k] = T(0);
        m.a[0] = m.a[4] = m.a[8] = T(1);
        return m;
    }
    static Mat3 zero() {
        Mat3 m;
        for (int k = 0; k < 9; ++k) m.a[k] = T(0);
        return m;
    }
};

template <typename T>
struct HostBox {
    int dim = 3, celltype = ORTHO, lcell = 1;
    int nc[3] = {1, 1, 1};
    T cutoff = 0, cutoff_sqr = 0;
    Mat3<T> input = Mat3<T>::zero(), aligned = Mat3<T>::zero(), rot = Mat3<T>::eye(), invrot = Mat3<T>::eye();
    T cs[3] = {0, 0, 0}, cbmin[3] = {0, 0, 0}, cbmax[3] = {0, 0, 0}, origin[3] = {0, 0, 0};
};

namespace detail {

template <typename T>
T sum_sq(const T *v, int n) {  // sum(abs2, v), sequential
    T s = v[0] * v[0];
    for (int k = 1; k < n; ++k) s = s + v[k] * v[k];
    return s;
}

template <typename T>
T dotn(const T *a, const T *b, int n) {
    T s = a[0] * b[0];
    for (int k = 1; k < n; ++k) s = s + a[k] * b[k];
    return s;
}

template <typename T>
void cross3(const T *a, const T *b, T *c) {
    T c0 = a[1] * b[2] - a[2] * b[1];
    T c1 = a[2] * b[0] - a[0] * b[2];
    T c2 = a[0] * b[1] - a[1] * b[0];
    c[0] = c0; c[1] = c1; c[2] = c2;
}

template <typename T>
Mat3<T> mul(const Mat3<T> &a, const Mat3<T> &b, int n) {
    Mat3<T> c = Mat3<T>::zero();
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            T s = a(i, 0) * b(0, j);
            for (int k = 1; k < n; ++k) s = s + a(i, k) * b(k, j);
            c(i, j) = s;
        }
    return c;
}

template <typename T>
void column(const Mat3<T> &m, int j, T *v) {
    v[0] = m(0, j); v[1] = m(1, j); v[2] = m(2, j);
}

// Inverse of the (rotation) matrix by cofactors; used only to map output vectors back.
template <typename T>
Mat3<T> inverse(const Mat3<T> &m, int n) {
    Mat3<T> r = Mat3<T>::eye();
    if (n == 3) {
        T c1[3], c2[3], c3[3], y0[3], y1[3], y2[3];
        column(m, 0, c1); column(m, 1, c2); column(m, 2, c3);
        cross3(c2, c3, y0);
        T d = dotn(c1, y0, 3);
        cross3(c3, c1, y1);
        cross3(c1, c2, y2);
        for (int k = 0; k < 3; ++k) {
            r(0, k) = y0[k] / d;
            r(1, k) = y1[k] / d;
            r(2, k) = y2[k] / d;
        }
    } else {
        T a = m(0, 0), b = m(0, 1), c = m(1, 0), e = m(1, 1);
        T d = a * e - b * c;
        r(0, 0) = e / d; r(0, 1) = -b / d;
        r(1, 0) = -c / d; r(1, 1) = a / d;
    }
    return r;
}

// CellOperations.jl:377-423
template <typename T>
void align3(const Mat3<T> &m0, Mat3<T> &mout, Mat3<T> &R) {
    static const int comb[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    T nrm2[3];
    for (int k = 0; k < 3; ++k) {
        T c[3];
        column(m0, k, c);
        nrm2[k] = sum_sq(c, 3);
    }
    int ia = 0;
    for (auto &cb : comb)
        if (nrm2[cb[0]] >= nrm2[cb[1]] && nrm2[cb[1]] >= nrm2[cb[2]]) { ia = cb[0]; break; }
    T a[3];
    column(m0, ia, a);
    T na = std::sqrt(sum_sq(a, 3));
    T a1[3] = {a[0] / na, a[1] / na, a[2] / na};
    T v[3] = {T(0), a1[2], -a1[1]};
    Mat3<T> R1 = Mat3<T>::eye();
    if (!(sum_sq(v, 3) == T(0))) {
        Mat3<T> vs = Mat3<T>::zero();
        vs(0, 1) = -v[2]; vs(0, 2) = v[1];
        vs(1, 0) = v[2];  vs(1, 2) = -v[0];
        vs(2, 0) = -v[1]; vs(2, 1) = v[0];
        Mat3<T> vs2 = mul(vs, vs, 3);
        T s = T(1) / (T(1) + a1[0]);
        Mat3<T> I = Mat3<T>::eye();
        for (int k = 0; k < 9; ++k) R1.a[k] = (I.a[k] + vs.a[k]) + vs2.a[k] * s;
    }
    Mat3<T> m = mul(R1, m0, 3);
    T x = m(0, 1), y = m(1, 1), z = m(2, 1);
    Mat3<T> R2 = Mat3<T>::eye();
    if (!((y * y + z * z) == T(0))) {
        T col[3] = {x, y, z};
        T b = std::sqrt(sum_sq(col, 3) - x * x);
        T sint = ((-z) * b) / (y * y + z * z);
        T cost = std::sqrt(T(1) - sint * sint);
        R2(1, 1) = cost; R2(1, 2) = -sint;
        R2(2, 1) = sint; R2(2, 2) = cost;
    }
    mout = mul(R2, m, 3);
    R = mul(R2, R1, 3);
}

// CellOperations.jl:353-375
template <typename T>
void align2(const Mat3<T> &m0, Mat3<T> &mout, Mat3<T> &R) {
    T a[2] = {m0(0, 0), m0(1, 0)};
    T b[2] = {m0(0, 1), m0(1, 1)};
    T na = std::sqrt(sum_sq(a, 2));
    T nb = std::sqrt(sum_sq(b, 2));
    if (nb > na) { a[0] = b[0]; a[1] = b[1]; na = nb; }
    R = Mat3<T>::eye();
    mout = m0;
    if (a[1] == T(0)) return;
    T sint = (-na) / ((a[0] * a[0]) / a[1] + a[1]);
    T cost = ((-a[0]) * sint) / a[1];
    R(0, 0) = cost; R(0, 1) = -sint;
    R(1, 0) = sint; R(1, 1) = cost;
    mout = mul(R, m0, 2);
}

// Box.jl:579-636
template <typename T>
bool unit_cell_ok(const Mat3<T> &m, T rc, int dim) {
    const T two_rc = T(2) * rc;
    if (dim == 3) {
        T a[3], b[3], c[3], bc[3], ab[3], ca[3];
        column(m, 0, a); column(m, 1, b); column(m, 2, c);
        cross3(b, c, bc);
        T n1 = std::sqrt(sum_sq(bc, 3));
        for (T &e : bc) e = e / n1;
        T aproj = dotn(a, bc, 3);
        cross3(a, b, ab);
        T n2 = std::sqrt(sum_sq(ab, 3));
        for (T &e : ab) e = e / n2;
        T cproj = dotn(c, ab, 3);
        cross3(c, a, ca);
        T n3 = std::sqrt(sum_sq(ca, 3));
        for (T &e : ca) e = e / n3;
        T bproj = dotn(b, ca, 3);
        if (aproj <= two_rc || bproj <= two_rc || cproj <= two_rc) return false;
        return !(std::isnan(aproj) || std::isnan(bproj) || std::isnan(cproj));
    }
    T a[2] = {m(0, 0), m(1, 0)}, b[2] = {m(0, 1), m(1, 1)};
    T na = std::sqrt(sum_sq(a, 2));
    T iu[2] = {a[0] / na, a[1] / na};
    T dbi = dotn(b, iu, 2);
    T bproj = std::sqrt(sum_sq(b, 2) - dbi * dbi);
    T nb = std::sqrt(sum_sq(b, 2));
    T ju[2] = {b[0] / nb, b[1] / nb};
    T daj = dotn(a, ju, 2);
    T aproj = std::sqrt(sum_sq(a, 2) - daj * daj);
    if (aproj <= two_rc || bproj <= two_rc) return false;
    return !(std::isnan(aproj) || std::isnan(bproj));
}

// CellOperations.jl:431-479: min/max over the cell vertices, starting from the origin vertex
template <typename T>
void limits_of_cell(const Mat3<T> &m, int dim, T *xmin, T *xmax) {
    T a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) { a[k] = m(k, 0); b[k] = m(k, 1); if (dim == 3) c[k] = m(k, 2); }
    T v[8][3];
    int nv = dim == 3 ? 8 : 4;
    for (int k = 0; k < 3; ++k) {
        v[0][k] = T(0); v[1][k] = a[k]; v[2][k] = a[k] + b[k]; v[3][k] = b[k];
        v[4][k] = a[k] + c[k]; v[5][k] = c[k]; v[6][k] = b[k] + c[k]; v[7][k] = (a[k] + b[k]) + c[k];
    }
    for (int k = 0; k < dim; ++k) { xmin[k] = v[0][k]; xmax[k] = v[0][k]; }
    for (int i = 1; i < nv; ++i)
        for (int k = 0; k < dim; ++k) {
            if (v[i][k] < xmin[k]) xmin[k] = v[i][k];
            if (v[i][k] > xmax[k]) xmax[k] = v[i][k];
        }
}

}  // namespace detail

// _construct_box (Box.jl:240-270) for a unit cell given column-major in T.
template <typename T>
int build_box(int dim, const T *cell, int celltype, T cutoff, int lcell, const T *origin, HostBox<T> &b) {
    if (dim != 2 && dim != 3) return BOX_BAD_DIM;
    if (lcell < 1) return BOX_BAD_LCELL;
    b = HostBox<T>();
    b.dim = dim; b.celltype = celltype; b.lcell = lcell; b.cutoff = cutoff;
    for (int j = 0; j < dim; ++j)
        for (int i = 0; i < dim; ++i) b.input(i, j) = cell[i + dim * j];
    if (celltype == TRICLINIC) {
        if (dim == 3) detail::align3(b.input, b.aligned, b.rot);
        else detail::align2(b.input, b.aligned, b.rot);
    } else {
        b.aligned = b.input;
        b.rot = Mat3<T>::eye();
    }
    if (!detail::unit_cell_ok(b.aligned, cutoff, dim)) return BOX_BAD_CELL;
    b.invrot = detail::inverse(b.rot, dim);
    T xmin[3] = {0, 0, 0}, xmax[3] = {0, 0, 0};
    detail::limits_of_cell(b.aligned, dim, xmin, xmax);
    const T step = cutoff / T(lcell);
    for (int k = 0; k < dim; ++k) {
        T q = (xmax[k] - xmin[k]) / step;
        long ncore;
        if (celltype == TRICLINIC) {
            ncore = (long)std::ceil(q);
            b.cs[k] = cutoff / T(lcell);
        } else {
            ncore = (long)std::floor(q);
            b.cs[k] = (xmax[k] - xmin[k]) / T(ncore);
        }
        b.nc[k] = int(ncore + 2 * lcell + 1);
        b.origin[k] = origin ? origin[k] : T(0);
        b.cbmin[k] = (xmin[k] + b.origin[k]) - T(lcell) * b.cs[k];
        b.cbmax[k] = (xmax[k] + b.origin[k]) + T(lcell) * b.cs[k];
    }
    for (int k = dim; k < 3; ++k) b.nc[k] = 1;
    b.cutoff_sqr = cutoff * cutoff;
    return BOX_OK;
}

}  // namespace clm
