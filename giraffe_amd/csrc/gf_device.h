// Device-side 3x3 algebra for the Rodrigues-parameter kinematics (gfx950).
// Row-major 3x3 matrices in registers; templated on the arithmetic type.
#pragma once
#include <hip/hip_runtime.h>

#define GF_DEV __device__ __forceinline__

template <typename T> struct m3 { T a[9]; };
template <typename T> struct v3 { T x, y, z; };

template <typename T> GF_DEV v3<T> mk3(T x, T y, T z) { v3<T> r; r.x = x; r.y = y; r.z = z; return r; }
template <typename T> GF_DEV T dot3(const v3<T> &a, const v3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> GF_DEV v3<T> cross3(const v3<T> &a, const v3<T> &b)
{
    return mk3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename T> GF_DEV v3<T> add3(const v3<T> &a, const v3<T> &b) { return mk3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T> GF_DEV v3<T> sub3(const v3<T> &a, const v3<T> &b) { return mk3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T> GF_DEV v3<T> scl3(T s, const v3<T> &a) { return mk3<T>(s * a.x, s * a.y, s * a.z); }

template <typename T> GF_DEV v3<T> mv(const m3<T> &m, const v3<T> &v)
{
    return mk3<T>(m.a[0] * v.x + m.a[1] * v.y + m.a[2] * v.z,
                  m.a[3] * v.x + m.a[4] * v.y + m.a[5] * v.z,
                  m.a[6] * v.x + m.a[7] * v.y + m.a[8] * v.z);
}
template <typename T> GF_DEV v3<T> mtv(const m3<T> &m, const v3<T> &v)
{
    return mk3<T>(m.a[0] * v.x + m.a[3] * v.y + m.a[6] * v.z,
                  m.a[1] * v.x + m.a[4] * v.y + m.a[7] * v.z,
                  m.a[2] * v.x + m.a[5] * v.y + m.a[8] * v.z);
}
template <typename T> GF_DEV m3<T> mm3(const m3<T> &a, const m3<T> &b)
{
    m3<T> c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            c.a[3 * i + j] = a.a[3 * i] * b.a[j] + a.a[3 * i + 1] * b.a[3 + j] + a.a[3 * i + 2] * b.a[6 + j];
    return c;
}
// a^T b
template <typename T> GF_DEV m3<T> mtm3(const m3<T> &a, const m3<T> &b)
{
    m3<T> c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            c.a[3 * i + j] = a.a[i] * b.a[j] + a.a[3 + i] * b.a[3 + j] + a.a[6 + i] * b.a[6 + j];
    return c;
}
template <typename T> GF_DEV m3<T> tr3(const m3<T> &a)
{
    m3<T> c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) c.a[3 * i + j] = a.a[3 * j + i];
    return c;
}
template <typename T> GF_DEV m3<T> skew3(const v3<T> &v)
{
    m3<T> s;
    s.a[0] = 0;    s.a[1] = -v.z; s.a[2] = v.y;
    s.a[3] = v.z;  s.a[4] = 0;    s.a[5] = -v.x;
    s.a[6] = -v.y; s.a[7] = v.x;  s.a[8] = 0;
    return s;
}
template <typename T> GF_DEV m3<T> madd(const m3<T> &a, const m3<T> &b)
{
    m3<T> c;
#pragma unroll
    for (int i = 0; i < 9; i++) c.a[i] = a.a[i] + b.a[i];
    return c;
}
template <typename T> GF_DEV m3<T> msub(const m3<T> &a, const m3<T> &b)
{
    m3<T> c;
#pragma unroll
    for (int i = 0; i < 9; i++) c.a[i] = a.a[i] - b.a[i];
    return c;
}
template <typename T> GF_DEV m3<T> mscl(T s, const m3<T> &a)
{
    m3<T> c;
#pragma unroll
    for (int i = 0; i < 9; i++) c.a[i] = s * a.a[i];
    return c;
}
// c += s * (u outer w)
template <typename T> GF_DEV void add_outer(m3<T> &c, T s, const v3<T> &u, const v3<T> &w)
{
    T uu[3] = {u.x, u.y, u.z}, ww[3] = {w.x, w.y, w.z};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) c.a[3 * i + j] += s * uu[i] * ww[j];
}

// Rodrigues: Q = I + g (A + A^2/2), Xi = g (I + A/2), g = 4 / (4 + |a|^2).
template <typename T> GF_DEV void rodrigues(const v3<T> &a, m3<T> &Q, m3<T> &Xi, T &g)
{
    T aa2 = dot3(a, a);
    g = T(4) / (T(4) + aa2);
    m3<T> A = skew3(a);
    m3<T> A2 = mm3(A, A);
#pragma unroll
    for (int i = 0; i < 9; i++) {
        Q.a[i] = g * (A.a[i] + T(0.5) * A2.a[i]);
        Xi.a[i] = g * T(0.5) * A.a[i];
    }
    Q.a[0] += T(1); Q.a[4] += T(1); Q.a[8] += T(1);
    Xi.a[0] += g; Xi.a[4] += g; Xi.a[8] += g;
}

// V(x, t) with h3 = h5 = 0: (h8 t - h4 x*t) (x) x + h2 skew(t).
template <typename T> GF_DEV m3<T> opV(const v3<T> &x, const v3<T> &t)
{
    T h = T(4) / (T(4) + dot3(x, x));
    T h2 = T(0.5) * h, h4 = T(-0.25) * h * h, h8 = T(-0.5) * h * h;
    v3<T> xt = cross3(x, t);
    v3<T> w = sub3(scl3(h8, t), scl3(h4, xt));
    m3<T> r = mscl(h2, skew3(t));
    add_outer(r, T(1), w, x);
    return r;
}

// d_V(x, dx, t) with h3 = h5 = h7 = 0.
template <typename T> GF_DEV m3<T> opdV(const v3<T> &x, const v3<T> &dx, const v3<T> &t)
{
    T h = T(4) / (T(4) + dot3(x, x));
    T h4 = T(-0.25) * h * h, h6 = T(0.25) * h * h * h, h8 = T(-0.5) * h * h, h9 = T(0.5) * h * h * h;
    T xdx = dot3(x, dx);
    v3<T> xt = cross3(x, t), dxt = cross3(dx, t);
    v3<T> w1 = sub3(scl3(h9, t), scl3(h6, xt));
    v3<T> w2 = sub3(scl3(h8, t), scl3(h4, xt));
    v3<T> w3 = scl3(-h4, dxt);
    m3<T> r = mscl(h4 * xdx, skew3(t));
    add_outer(r, xdx, w1, x);
    add_outer(r, T(1), w2, dx);
    add_outer(r, T(1), w3, x);
    return r;
}

// Parameters of Q(a2) Q(a1): 4/(4 - a2.a1) (a2 + a1 + a2 x a1 / 2).
template <typename T> GF_DEV v3<T> compose(const v3<T> &a2, const v3<T> &a1)
{
    v3<T> c = cross3(a2, a1);
    T s = T(4) / (T(4) - dot3(a2, a1));
    return scl3(s, add3(add3(a2, a1), scl3(T(0.5), c)));
}
