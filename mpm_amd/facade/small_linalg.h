// small_linalg.h -- the few fixed-size types the reference takes from Eigen
// (Vector3f, Vector3i, Matrix3f), as plain host PODs with Eigen-style
// element access: v(i), v.x(), m(i, j), setIdentity(), setZero().
#ifndef CHAINS_SMALL_LINALG_H_
#define CHAINS_SMALL_LINALG_H_

namespace chains {

template <class T>
struct Vec3 {
    T d[3] = {T(0), T(0), T(0)};
    Vec3() = default;
    Vec3(T x, T y, T z) { d[0] = x; d[1] = y; d[2] = z; }
    T& operator()(int i) { return d[i]; }
    const T& operator()(int i) const { return d[i]; }
    T& operator[](int i) { return d[i]; }
    const T& operator[](int i) const { return d[i]; }
    T& x() { return d[0]; }
    T& y() { return d[1]; }
    T& z() { return d[2]; }
    const T& x() const { return d[0]; }
    const T& y() const { return d[1]; }
    const T& z() const { return d[2]; }
};

using Vector3f = Vec3<float>;
using Vector3i = Vec3<int>;

struct Matrix3f {
    float d[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    float& operator()(int i, int j) { return d[i][j]; }
    const float& operator()(int i, int j) const { return d[i][j]; }
    void setZero() {
        for (auto& r : d)
            for (float& v : r) v = 0.0f;
    }
    void setIdentity() {
        setZero();
        d[0][0] = d[1][1] = d[2][2] = 1.0f;
    }
    static Matrix3f Identity() {
        Matrix3f m;
        m.setIdentity();
        return m;
    }
};

inline Vector3f operator*(float s, const Vector3f& v) { return Vector3f(s * v(0), s * v(1), s * v(2)); }

}  // namespace chains

#endif  // CHAINS_SMALL_LINALG_H_
