// CVector.hpp -- the small fixed-size vector the reference's host API passes
// around (CVector<3,int>, CVector<3,T>, CVector<4,T>; src/libmath/CVector3.hpp,
// src/libmath/CVector4.hpp).  Only the members the solver, controller and
// manager surfaces use.
#pragma once

#include <cmath>
#include <ostream>

template <int N, typename T>
class CVector {
public:
    T data[N];

    CVector() {
        for (int i = 0; i < N; ++i) data[i] = T(0);
    }
    CVector(T a, T b, T c) {
        static_assert(N == 3, "three components");
        data[0] = a;
        data[1] = b;
        data[2] = c;
    }
    CVector(T a, T b, T c, T d) {
        static_assert(N == 4, "four components");
        data[0] = a;
        data[1] = b;
        data[2] = c;
        data[3] = d;
    }
    T &operator[](int i) { return data[i]; }
    const T &operator[](int i) const { return data[i]; }

    // product of the components: cells of a box (src/libmath/CVector3.hpp:177)
    T elements() const {
        T r = T(1);
        for (int i = 0; i < N; ++i) r *= data[i];
        return r;
    }
    T dotProd(const CVector &v) const {
        T r = T(0);
        for (int i = 0; i < N; ++i) r += data[i] * v.data[i];
        return r;
    }
    T length2() const {
        T r = data[0] * data[0];
        for (int i = 1; i < N; ++i) r = r + data[i] * data[i];
        return r;
    }
    // sqrt of the sum of squares in T (src/libmath/CVector3.hpp:185-188)
    T length() const { return (T)std::sqrt(length2()); }
    CVector operator*(T a) const {
        CVector r;
        for (int i = 0; i < N; ++i) r.data[i] = data[i] * a;
        return r;
    }
    CVector &operator*=(T a) {
        for (int i = 0; i < N; ++i) data[i] *= a;
        return *this;
    }
    bool operator==(const CVector &v) const {
        for (int i = 0; i < N; ++i)
            if (data[i] != v.data[i]) return false;
        return true;
    }
};

template <int N, typename T>
std::ostream &operator<<(std::ostream &os, const CVector<N, T> &v) {
    os << "[";
    for (int i = 0; i < N; ++i) os << (i ? ", " : "") << v.data[i];
    return os << "]";
}
