// dev_math.h -- complex helpers and in-register small DFTs for gfx950.
//
// Every DFT here is a forward transform X[k] = sum_n x[n] e^{-2 pi i nk/R}
// on an array held in VGPRs, natural order in and out. Radices 2, 3, 4, 5 are
// written out; composite radices (6, 8, 9, 10, 12, 15, 16) are one
// Cooley-Tukey step over them with compile-time twiddle constants, so a whole
// composite butterfly stays in registers. (The reference builds its radix
// 6/8/16 butterflies from the same 2/3/4 pieces: src/fftset_vec.c:536-845.)
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace fsa {

// ---------------------------------------------------------------- complex ops
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// ------------------------------------------------------- streaming HBM I/O
// Signal, output and kernel-spectrum accesses touch each element once, so
// they are issued as nontemporal (streaming) loads and stores: measured on
// MI355X, 3-6% faster than cached accesses for forward, inverse and
// convolution batches.
typedef float fsa_f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ldg(const float2* p)
{
    const fsa_f2v v = __builtin_nontemporal_load((const fsa_f2v*)p);
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float ldg(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stg(float2* p, float2 v) { __builtin_nontemporal_store(fsa_f2v{v.x, v.y}, (fsa_f2v*)p); }
__device__ __forceinline__ void stg(float* p, float v) { __builtin_nontemporal_store(v, p); }
// Element i of p with its byte offset formed in 32 bits (0 <= i and
// i * sizeof(T) < 2^32). For a p that is uniform across the wave (a
// twiddle or modulation table) the access is an SGPR base plus a 32-bit
// lane offset, where p + i costs 64-bit address arithmetic per lane. For a
// per-lane p (one transform's signal among several in a workgroup) plain
// p + i is cheaper: the compiler folds constant offsets into one 64-bit
// base.
template <class T>
__device__ __forceinline__ const T* off32(const T* p, int i)
{
    return (const T*)((const char*)p + (unsigned)i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ T* off32(T* p, int i)
{
    return (T*)((char*)p + (unsigned)i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ const T& at32(const T* p, int i)
{
    return *off32(p, i);
}
// p[uni + lane] for a wave-uniform p and uni (a twiddle table and a pass's
// compile-time or uniform offset) and a per-lane lane >= 0: the uniform part
// goes into the scalar base, so the twiddles of one butterfly share one
// 32-bit lane offset.
template <class T>
__device__ __forceinline__ const T& at2(const T* p, int uni, int lane)
{
    return *off32(p + uni, lane);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// -i * a and +i * a
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }
__device__ __forceinline__ float2 mul_pi(float2 a) { return make_float2(-a.y, a.x); }

// ------------------------------------------ register and address pinning
// Keeps a table pointer in SGPRs and its loads inside the current step (an
// empty asm statement, no memory access): left to itself the compiler hoists
// the twiddle loads of every chunk to the top and spills.
template <class T>
__device__ __forceinline__ const T* fresh(const T* p)
{
    // through the global address space, so loads stay global (not flat)
    typedef const T __attribute__((address_space(1)))* gptr;
    gptr q = (gptr)p;
    asm volatile("" : "+s"(q));
    return (const T*)q;
}

// Element lo of p + uni, with p + uni formed in SGPRs (uniform) and the
// per-lane part a 32-bit offset: the access is then an SGPR base plus one
// lane-offset VGPR shared by every access of the step. Without the asm the
// compiler folds the lane offset into a 64-bit VGPR base first and needs a
// VGPR pair per distinct uniform offset.
template <class T>
__device__ __forceinline__ T* ub(T* p, int uni, int lo)
{
    typedef T __attribute__((address_space(1)))* gptr;
    gptr q = (gptr)(p + uni);
    asm volatile("" : "+s"(q));
    return off32((T*)q, lo);
}

// Materialises v here: values parked in VGPRs for a later phase would
// otherwise have their last multiply sunk to the use, keeping both factors
// live across the whole kernel.
__device__ __forceinline__ void pin(float2& v)
{
    fsa_f2v t = {v.x, v.y};  // one VGPR pair, as the 64-bit loads and stores want it
    asm volatile("" : "+v"(t));
    v = make_float2(t.x, t.y);
}

// x, opaque to the compiler: the write-back addresses of an in-place pass
// are recomputed from it instead of keeping the 16 read addresses live
// across the butterfly.
__device__ __forceinline__ int fresh_int(int x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// ------------------------------------------------------- compile-time loops
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// --------------------------------------------- compile-time roots of unity
// cos/sin(2 pi p / R) by a Taylor series after shifting the angle into
// [-pi, pi); evaluated only in constant expressions, rounded once to float.
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ct_series_sin(double x)
{
    double t = x, s = x;
    for (int k = 1; k < 30; ++k) {
        t *= -x * x / double((2 * k) * (2 * k + 1));
        s += t;
    }
    return s;
}
constexpr double ct_series_cos(double x)
{
    double t = 1.0, s = 1.0;
    for (int k = 1; k < 30; ++k) {
        t *= -x * x / double((2 * k - 1) * (2 * k));
        s += t;
    }
    return s;
}
constexpr double ct_angle(long p, long r)
{
    long q = ((p % r) + r) % r;
    if (2 * q > r)
        q -= r;  // angle in (-pi, pi]
    return 2.0 * kPi * double(q) / double(r);
}

// W_R^P = exp(-2 pi i P / R)
template <int P, int R>
struct Root {
    static constexpr float re = float(ct_series_cos(ct_angle(P, R)));
    static constexpr float im = float(-ct_series_sin(ct_angle(P, R)));
};

// x * W_R^P with exact handling of the quarter turns
template <int P, int R>
__device__ __forceinline__ float2 rootmul(float2 x)
{
    constexpr int p = ((P % R) + R) % R;
    if constexpr (p == 0) {
        return x;
    } else if constexpr (4 * p == R) {
        return mul_mi(x);
    } else if constexpr (2 * p == R) {
        return make_float2(-x.x, -x.y);
    } else if constexpr (4 * p == 3 * R) {
        return mul_pi(x);
    } else if constexpr (8 * p == R) {
        constexpr float h = Root<1, 8>::re;  // sqrt(1/2)
        return make_float2(h * (x.x + x.y), h * (x.y - x.x));
    } else if constexpr (8 * p == 3 * R) {
        constexpr float h = Root<1, 8>::re;
        return make_float2(h * (x.y - x.x), -h * (x.x + x.y));
    } else {
        return cmul(x, make_float2(Root<p, R>::re, Root<p, R>::im));
    }
}

// --------------------------------------------------------------- butterflies
template <int R>
struct DFT;

template <>
struct DFT<1> {
    static __device__ __forceinline__ void run(float2*) {}
};

template <>
struct DFT<2> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <>
struct DFT<3> {
    static __device__ __forceinline__ void run(float2* v)
    {
        constexpr float s3 = 0.86602540378443864676f;  // sin(2 pi / 3)
        float2 t = cadd(v[1], v[2]);
        float2 d = csub(v[1], v[2]);
        float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
        float2 s = cscale(d, s3);
        v[0] = cadd(v[0], t);
        v[1] = make_float2(m.x + s.y, m.y - s.x);  // m - i s
        v[2] = make_float2(m.x - s.y, m.y + s.x);  // m + i s
    }
};

template <>
struct DFT<4> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        float2 c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, mul_mi(d));
        v[3] = csub(b, mul_mi(d));
    }
};

template <>
struct DFT<5> {
    static __device__ __forceinline__ void run(float2* v)
    {
        constexpr float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
        constexpr float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
        constexpr float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
        constexpr float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
        float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        float2 b1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
        float2 b2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(a1, mul_mi(b1));
        v[4] = csub(a1, mul_mi(b1));
        v[2] = cadd(a2, mul_mi(b2));
        v[3] = csub(a2, mul_mi(b2));
    }
};

// One Cooley-Tukey step in registers, R = A*B:
// X[k1 + A k2] = sum_n2 W_B^{n2 k2} W_R^{n2 k1} sum_n1 x[B n1 + n2] W_A^{n1 k1}
template <int A, int B>
struct DFTct {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 t[A * B];
        static_for<B>([&](auto N2) {
            constexpr int n2 = decltype(N2)::value;
            float2 c[A];
            static_for<A>([&](auto N1) { c[N1] = v[B * decltype(N1)::value + n2]; });
            DFT<A>::run(c);
            static_for<A>([&](auto K1) {
                constexpr int k1 = decltype(K1)::value;
                t[k1 * B + n2] = rootmul<n2 * k1, A * B>(c[k1]);
            });
        });
        static_for<A>([&](auto K1) {
            constexpr int k1 = decltype(K1)::value;
            float2 r[B];
            static_for<B>([&](auto N2) { r[N2] = t[k1 * B + decltype(N2)::value]; });
            DFT<B>::run(r);
            static_for<B>([&](auto K2) { v[k1 + A * decltype(K2)::value] = r[K2]; });
        });
    }
};

template <> struct DFT<6> : DFTct<2, 3> {};
template <> struct DFT<8> : DFTct<2, 4> {};
template <> struct DFT<9> : DFTct<3, 3> {};
template <> struct DFT<10> : DFTct<2, 5> {};
template <> struct DFT<12> : DFTct<4, 3> {};
template <> struct DFT<15> : DFTct<3, 5> {};
template <> struct DFT<16> : DFTct<4, 4> {};

// DFTct for a whole column held by one lane (rcol_impl.h): the same step,
// with each small DFT fenced off from the machine scheduler, which would
// otherwise interleave all of them and run out of VGPRs.
template <int A, int B>
struct DFTctSeq {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 t[A * B];
        static_for<B>([&](auto N2) {
            constexpr int n2 = decltype(N2)::value;
            float2 c[A];
            static_for<A>([&](auto N1) { c[N1] = v[B * decltype(N1)::value + n2]; });
            DFT<A>::run(c);
            static_for<A>([&](auto K1) {
                constexpr int k1 = decltype(K1)::value;
                t[k1 * B + n2] = rootmul<n2 * k1, A * B>(c[k1]);
            });
            __builtin_amdgcn_sched_barrier(0);
        });
        static_for<A>([&](auto K1) {
            constexpr int k1 = decltype(K1)::value;
            float2 r[B];
            static_for<B>([&](auto N2) { r[N2] = t[k1 * B + decltype(N2)::value]; });
            DFT<B>::run(r);
            static_for<B>([&](auto K2) { v[k1 + A * decltype(K2)::value] = r[K2]; });
            __builtin_amdgcn_sched_barrier(0);
        });
    }
};
template <> struct DFT<60> : DFTctSeq<4, 15> {};
template <> struct DFT<18> : DFTctSeq<2, 9> {};

}  // namespace fsa
