// Complex arithmetic and the small-radix DFT butterflies of the gfx950 line
// engine. Everything here is __host__ __device__ so that the host planner
// and the kernels share one definition of each butterfly.
//
// Sign convention (crates/pencil-fft/src/lib.rs:391-438): the forward
// transform uses exp(-2 pi i jk/n). The kernels only ever run the forward
// butterflies; a positive-sign transform is conj(forward(conj(x))).
#pragma once

#include <hip/hip_runtime.h>

namespace pfh {

template <typename T>
struct cpx {
    T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T a, T b)
{
    cpx<T> r;
    r.x = a;
    r.y = b;
    return r;
}

// Non-temporal global accesses, for stages whose arrays are far larger than
// the caches (each element is read once and written once per stage; on
// MI355X a 512 MiB copy moves 6.2 TB/s with the hint against 5.3-5.7 TB/s
// without). Stages of small arrays keep ordinary accesses: their neighbouring
// tiles share cache lines and the next stage can find its input in the
// last-level cache. Complex values are accessed as two scalars so no
// alignment beyond the scalar's is assumed.
template <typename T>
__device__ __forceinline__ T ldg_nt(const T *p)
{
    return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ cpx<T> ldg_nt(const cpx<T> *p)
{
    const T *q = &p->x;
    return mk(__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1));
}
template <typename T>
__device__ __forceinline__ void stg_nt(T *p, T v)
{
    __builtin_nontemporal_store(v, p);
}
template <typename T>
__device__ __forceinline__ void stg_nt(cpx<T> *p, cpx<T> v)
{
    T *q = &p->x;
    __builtin_nontemporal_store(v.x, q);
    __builtin_nontemporal_store(v.y, q + 1);
}
template <bool NT, typename P>
__device__ __forceinline__ P gld(const P *p)
{
    if constexpr (NT) return ldg_nt(p);
    else return *p;
}
template <bool NT, typename P>
__device__ __forceinline__ void gst(P *p, P v)
{
    if constexpr (NT) stg_nt(p, v);
    else *p = v;
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> conjg(cpx<T> a) { return mk(a.x, -a.y); }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> scal(cpx<T> a, T s) { return mk(a.x * s, a.y * s); }
// multiply by -i
template <typename T>
__host__ __device__ __forceinline__ cpx<T> mul_mi(cpx<T> a) { return mk(a.y, -a.x); }
// multiply by +i
template <typename T>
__host__ __device__ __forceinline__ cpx<T> mul_pi(cpx<T> a) { return mk(-a.y, a.x); }

// cos/sin of 2 pi m / R used by the odd-prime butterflies (m = 1..(R-1)/2)
struct Roots {
    static constexpr double c3 = -0.5;
    static constexpr double s3 = 0.86602540378443864676372317075293618;
    static constexpr double c51 = 0.30901699437494742410229341718281906;
    static constexpr double c52 = -0.80901699437494742410229341718281906;
    static constexpr double s51 = 0.95105651629515357211643933337938214;
    static constexpr double s52 = 0.58778525229247312916870595463907277;
    static constexpr double r2 = 0.70710678118654752440084436210484904; // sqrt(1/2)
    static constexpr double c16 = 0.92387953251128675612818318939678829; // cos(pi/8)
    static constexpr double s16 = 0.38268343236508977172845998403039887; // sin(pi/8)
};

// Generic odd-prime butterfly constants: cos and sin of 2 pi m / R.
template <int R>
struct PrimeTable;
template <>
struct PrimeTable<7> {
    static __host__ __device__ double c(int m)
    {
        constexpr double t[4] = {1.0, 0.62348980185873353052500488400423981, -0.22252093395631440428890256449679476,
                                 -0.90096886790241912623610231950744505};
        return t[m];
    }
    static __host__ __device__ double s(int m)
    {
        constexpr double t[4] = {0.0, 0.78183148246802980870844452667405775, 0.97492791218182360701813168299393122,
                                 0.43388373911755812047576833284835875};
        return t[m];
    }
};
template <>
struct PrimeTable<11> {
    static __host__ __device__ double c(int m)
    {
        constexpr double t[6] = {1.0,
                                 0.84125353283118116886181164892859535,
                                 0.41541501300188642552927414923589580,
                                 -0.14231483827328514044379266862568197,
                                 -0.65486073394528506405692507247392354,
                                 -0.95949297361449738989036805707508320};
        return t[m];
    }
    static __host__ __device__ double s(int m)
    {
        constexpr double t[6] = {0.0,
                                 0.54064081745559758210763595432497169,
                                 0.90963199535451837140730764992353738,
                                 0.98982144188093273237609203778056162,
                                 0.75574957435425828377403584397126241,
                                 0.28173255684142969771141791702969015};
        return t[m];
    }
};
template <>
struct PrimeTable<13> {
    static __host__ __device__ double c(int m)
    {
        constexpr double t[7] = {1.0,
                                 0.88545602565320989587194927539206956,
                                 0.56806474673115580251180755912609802,
                                 0.12053668025532305334906768745254405,
                                 -0.35460488704253562596963789260002069,
                                 -0.74851074817110109863463059970135305,
                                 -0.97094181742605202715698227629378922};
        return t[m];
    }
    static __host__ __device__ double s(int m)
    {
        constexpr double t[7] = {0.0,
                                 0.46472317204376854565834833350187811,
                                 0.82298386589365639457961901765981568,
                                 0.99270887409805399280075125052775768,
                                 0.93501624268541482343679989590226025,
                                 0.66312265824079520237678188010227530,
                                 0.23931566428755776714177722772460734};
        return t[m];
    }
};

// Forward DFT of R values in place: v[k] = sum_j v[j] exp(-2 pi i jk / R).
template <typename T, int R>
struct Dft;

template <typename T>
struct Dft<T, 2> {
    __host__ __device__ static __forceinline__ void run(cpx<T> *v)
    {
        cpx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};

template <typename T>
struct Dft<T, 3> {
    __host__ __device__ static __forceinline__ void run(cpx<T> *v)
    {
        const T c = (T)Roots::c3, s = (T)Roots::s3;
        cpx<T> t1 = v[1] + v[2];
        cpx<T> t2 = v[0] + scal(t1, c);
        cpx<T> t3 = scal(mul_mi(v[1] - v[2]), s);
        v[0] = v[0] + t1;
        v[1] = t2 + t3;
        v[2] = t2 - t3;
    }
};

template <typename T>
struct Dft<T, 4> {
    __host__ __device__ static __forceinline__ void run(cpx<T> *v)
    {
        cpx<T> a = v[0] + v[2], b = v[0] - v[2];
        cpx<T> c = v[1] + v[3], d = mul_mi(v[1] - v[3]);
        v[0] = a + c;
        v[2] = a - c;
        v[1] = b + d;
        v[3] = b - d;
    }
};

template <typename T>
struct Dft<T, 5> {
    __host__ __device__ static __forceinline__ void run(cpx<T> *v)
    {
        const T c1 = (T)Roots::c51, c2 = (T)Roots::c52, s1 = (T)Roots::s51, s2 = (T)Roots::s52;
        cpx<T> a1 = v[1] + v[4], b1 = v[1] - v[4];
        cpx<T> a2 = v[2] + v[3], b2 = v[2] - v[3];
        cpx<T> r1 = v[0] + scal(a1, c1) + scal(a2, c2);
        cpx<T> r2 = v[0] + scal(a1, c2) + scal(a2, c1);
        cpx<T> i1 = mul_mi(scal(b1, s1) + scal(b2, s2));
        cpx<T> i2 = mul_mi(scal(b1, s2) - scal(b2, s1));
        v[0] = v[0] + a1 + a2;
        v[1] = r1 + i1;
        v[4] = r1 - i1;
        v[2] = r2 + i2;
        v[3] = r2 - i2;
    }
};

// Odd primes 7, 11, 13: pairwise symmetric form, O(R^2 / 2), with the
// symmetric sums / differences kept in the input slots (2R values live).
template <typename T, int R>
struct DftPrime {
    __host__ __device__ static __forceinline__ void run(cpx<T> *v)
    {
        constexpr int H = (R - 1) / 2;
        cpx<T> s0 = v[0];
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            const cpx<T> p = v[j] + v[R - j], q = v[j] - v[R - j];
            v[j] = p;     // a_j
            v[R - j] = q; // b_j
            s0 = s0 + p;
        }
        cpx<T> out[R];
        out[0] = s0;
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            cpx<T> re = v[0], im = mk((T)0, (T)0);
#pragma unroll
            for (int j = 1; j <= H; ++j) {
                const int e = (j * k) % R;
                const int ee = e <= H ? e : R - e;
                const T c = (T)PrimeTable<R>::c(ee);
                const T s = (T)(e <= H ? PrimeTable<R>::s(ee) : -PrimeTable<R>::s(ee));
                re = re + scal(v[j], c);
                im = im + scal(v[R - j], s);
            }
            // exp(-i t) = cos t - i sin t
            const cpx<T> mi = mul_mi(im);
            out[k] = re + mi;
            out[R - k] = re - mi;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = out[k];
    }
};
template <typename T>
struct Dft<T, 7> : DftPrime<T, 7> {};
template <typename T>
struct Dft<T, 11> : DftPrime<T, 11> {};
template <typename T>
struct Dft<T, 13> : DftPrime<T, 13> {};

template <typename T>
struct Dft<T, 8> {
    __host__ __device__ static __forceinline__ void run(cpx<T> *v)
    {
        const T r = (T)Roots::r2;
        cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
        Dft<T, 4>::run(e);
        Dft<T, 4>::run(o);
        // o[k] *= exp(-i pi k / 4)
        o[1] = scal(mk(o[1].x + o[1].y, o[1].y - o[1].x), r);
        o[2] = mul_mi(o[2]);
        o[3] = scal(mk(o[3].y - o[3].x, -o[3].x - o[3].y), r);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = e[k] + o[k];
            v[k + 4] = e[k] - o[k];
        }
    }
};

template <typename T>
struct Dft<T, 16> {
    __host__ __device__ static __forceinline__ void run(cpx<T> *v)
    {
        // j = r1 + 4 r2, k = k2 + 4 k1
        cpx<T> t[4][4];
#pragma unroll
        for (int r1 = 0; r1 < 4; ++r1) {
            cpx<T> c[4] = {v[r1], v[r1 + 4], v[r1 + 8], v[r1 + 12]};
            Dft<T, 4>::run(c);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) t[r1][k2] = c[k2];
        }
        const T c1 = (T)Roots::c16, s1 = (T)Roots::s16, r = (T)Roots::r2;
        // twiddle exp(-2 pi i r1 k2 / 16)
        t[1][1] = cmul(t[1][1], mk(c1, -s1));
        t[1][2] = cmul(t[1][2], mk(r, -r));
        t[1][3] = cmul(t[1][3], mk(s1, -c1));
        t[2][1] = cmul(t[2][1], mk(r, -r));
        t[2][2] = mul_mi(t[2][2]);
        t[2][3] = cmul(t[2][3], mk(-r, -r));
        t[3][1] = cmul(t[3][1], mk(s1, -c1));
        t[3][2] = cmul(t[3][2], mk(-r, -r));
        t[3][3] = cmul(t[3][3], mk(-c1, s1));
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            cpx<T> c[4] = {t[0][k2], t[1][k2], t[2][k2], t[3][k2]};
            Dft<T, 4>::run(c);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) v[k2 + 4 * k1] = c[k1];
        }
    }
};

// exp(-2 pi i e / 32) = (W32c(e), -W32s(e)), e = 0..21
__host__ __device__ inline double W32c(int e)
{
    constexpr double t[22] = {1.0,
                                 0.9807852804032304491262,
                                 0.9238795325112867561282,
                                 0.8314696123025452370788,
                                 0.7071067811865475244008,
                                 0.5555702330196022247428,
                                 0.3826834323650897717285,
                                 0.1950903220161282678483,
                                 8.478427660368899643959e-32,
                                 -0.1950903220161282678483,
                                 -0.3826834323650897717285,
                                 -0.5555702330196022247428,
                                 -0.7071067811865475244008,
                                 -0.8314696123025452370788,
                                 -0.9238795325112867561282,
                                 -0.9807852804032304491262,
                                 -1.0,
                                 -0.9807852804032304491262,
                                 -0.9238795325112867561282,
                                 -0.8314696123025452370788,
                                 -0.7071067811865475244008,
                                 -0.5555702330196022247428};
    return t[e];
}
__host__ __device__ inline double W32s(int e)
{
    constexpr double t[22] = {0.0,
                                 0.1950903220161282678483,
                                 0.3826834323650897717285,
                                 0.5555702330196022247428,
                                 0.7071067811865475244008,
                                 0.8314696123025452370788,
                                 0.9238795325112867561282,
                                 0.9807852804032304491262,
                                 1.0,
                                 0.9807852804032304491262,
                                 0.9238795325112867561282,
                                 0.8314696123025452370788,
                                 0.7071067811865475244008,
                                 0.5555702330196022247428,
                                 0.3826834323650897717285,
                                 0.1950903220161282678483,
                                 1.695685532073779928792e-31,
                                 -0.1950903220161282678483,
                                 -0.3826834323650897717285,
                                 -0.5555702330196022247428,
                                 -0.7071067811865475244008,
                                 -0.8314696123025452370788};
    return t[e];
}

template <typename T>
struct Dft<T, 32> {
    __host__ __device__ static __forceinline__ void run(cpx<T> *v)
    {
        // j = r1 + 4 r2, k = k2 + 8 k1
        cpx<T> t[4][8];
#pragma unroll
        for (int r1 = 0; r1 < 4; ++r1) {
            cpx<T> c[8];
#pragma unroll
            for (int r2 = 0; r2 < 8; ++r2) c[r2] = v[r1 + 4 * r2];
            Dft<T, 8>::run(c);
#pragma unroll
            for (int k2 = 0; k2 < 8; ++k2) {
                const int e = r1 * k2;
                t[r1][k2] = (e == 0) ? c[k2] : cmul(c[k2], mk((T)W32c(e), (T)-W32s(e)));
            }
        }
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) {
            cpx<T> c[4] = {t[0][k2], t[1][k2], t[2][k2], t[3][k2]};
            Dft<T, 4>::run(c);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) v[k2 + 8 * k1] = c[k1];
        }
    }
};

} // namespace pfh
