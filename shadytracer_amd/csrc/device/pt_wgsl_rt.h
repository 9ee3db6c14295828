// The runtime of the scene's WGSL code as translated to C++ (shadytracer_amd/js/wgsl.js): WGSL's vector, matrix and
// scalar types, its operators, constructors, conversions, swizzles and builtin functions (the WGSL specification's
// numeric builtins, componentwise on vectors), over a scalar WF -- float under hipRTC, double in the oracle's host
// compile of the same translated code. The language, not the scene's library: shapes, noise and the Hit are each
// side's own.
//
// Set before including: WG_FN (the function qualifiers), WF (the float type), WL(x) (a float literal of type WF).
#pragma once

template <class T> struct tv2 {
    T x, y;
    WG_FN tv2() : x(T(0)), y(T(0)) {}
    WG_FN tv2(T a, T b) : x(a), y(b) {}
    WG_FN explicit tv2(T s) : x(s), y(s) {}
    WG_FN T& operator[](int i) { return i == 0 ? x : y; }
    WG_FN T operator[](int i) const { return i == 0 ? x : y; }
};
template <class T> struct tv3 {
    T x, y, z;
    WG_FN tv3() : x(T(0)), y(T(0)), z(T(0)) {}
    WG_FN tv3(T a, T b, T c) : x(a), y(b), z(c) {}
    WG_FN explicit tv3(T s) : x(s), y(s), z(s) {}
    WG_FN T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    WG_FN T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> struct tv4 {
    T x, y, z, w;
    WG_FN tv4() : x(T(0)), y(T(0)), z(T(0)), w(T(0)) {}
    WG_FN tv4(T a, T b, T c, T d) : x(a), y(b), z(c), w(d) {}
    WG_FN explicit tv4(T s) : x(s), y(s), z(s), w(s) {}
    WG_FN T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
    WG_FN T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
};
typedef tv2<WF> wvec2;
typedef tv3<WF> wvec3;
typedef tv4<WF> wvec4;
typedef tv2<int> wivec2;
typedef tv3<int> wivec3;
typedef tv4<int> wivec4;
typedef tv2<unsigned int> wuvec2;
typedef tv3<unsigned int> wuvec3;
typedef tv4<unsigned int> wuvec4;
typedef tv2<bool> wbvec2;
typedef tv3<bool> wbvec3;
typedef tv4<bool> wbvec4;
template <class T, int N> struct warray {
    T v[N];
    WG_FN T& operator[](int i) { return v[i]; }
    WG_FN T operator[](int i) const { return v[i]; }
};

// ---- scalar conversions (WGSL: a float to an integer truncates toward zero, clamped to the integer's range) ------
template <class A> WG_FN WF wf32(A a) { return WF(a); }
WG_FN WF wf32(bool b) { return b ? WF(1) : WF(0); }
WG_FN WF wf32() { return WF(0); }
WG_FN unsigned int wu32() { return 0u; }
WG_FN int wi32() { return 0; }
WG_FN bool wbool() { return false; }
WG_FN unsigned int wu32(unsigned int a) { return a; }
WG_FN unsigned int wu32(int a) { return (unsigned int)a; }
WG_FN unsigned int wu32(bool a) { return a ? 1u : 0u; }
WG_FN unsigned int wu32(float a) { return !(a > 0.0f) ? 0u : (a >= 4294967040.0f ? 4294967040u : (unsigned int)a); }
WG_FN unsigned int wu32(double a) { return !(a > 0.0) ? 0u : (a >= 4294967040.0 ? 4294967040u : (unsigned int)a); }
WG_FN int wi32(int a) { return a; }
WG_FN int wi32(unsigned int a) { return (int)a; }
WG_FN int wi32(bool a) { return a ? 1 : 0; }
WG_FN int wi32(float a) { return a != a ? 0 : (a <= -2147483648.0f ? (-2147483647 - 1) : (a >= 2147483520.0f ? 2147483520 : (int)a)); }
WG_FN int wi32(double a) { return a != a ? 0 : (a <= -2147483648.0 ? (-2147483647 - 1) : (a >= 2147483520.0 ? 2147483520 : (int)a)); }
template <class A> WG_FN bool wbool(A a) { return a != A(0); }
template <class T> struct wscal { typedef T type; };
template <class T> struct wscal<tv2<T> > { typedef T type; };
template <class T> struct wscal<tv3<T> > { typedef T type; };
template <class T> struct wscal<tv4<T> > { typedef T type; };
template <class D, class S> struct wcv { static WG_FN D go(S s) { return D(s); } };
template <class S> struct wcv<WF, S> { static WG_FN WF go(S s) { return wf32(s); } };
template <class S> struct wcv<unsigned int, S> { static WG_FN unsigned int go(S s) { return wu32(s); } };
template <class S> struct wcv<int, S> { static WG_FN int go(S s) { return wi32(s); } };
template <class S> struct wcv<bool, S> { static WG_FN bool go(S s) { return wbool(s); } };
template <class D, class S> WG_FN D wsc(S s) { return wcv<D, S>::go(s); }

// ---- vector constructors: mk_<type>(...) as WGSL's: zero, a splat, components, or vectors and scalars in order ------
#define WG_MK(P, T)                                                                                                     \
    WG_FN tv2<T> mk_##P##vec2() { return tv2<T>(); }                                                                    \
    template <class A> WG_FN tv2<T> mk_##P##vec2(A s) { return tv2<T>(wsc<T>(s)); }                                     \
    template <class A> WG_FN tv2<T> mk_##P##vec2(tv2<A> v) { return tv2<T>(wsc<T>(v.x), wsc<T>(v.y)); }                 \
    template <class A, class B> WG_FN tv2<T> mk_##P##vec2(A a, B b) { return tv2<T>(wsc<T>(a), wsc<T>(b)); }           \
    WG_FN tv3<T> mk_##P##vec3() { return tv3<T>(); }                                                                    \
    template <class A> WG_FN tv3<T> mk_##P##vec3(A s) { return tv3<T>(wsc<T>(s)); }                                     \
    template <class A> WG_FN tv3<T> mk_##P##vec3(tv3<A> v) { return tv3<T>(wsc<T>(v.x), wsc<T>(v.y), wsc<T>(v.z)); }    \
    template <class A, class B> WG_FN tv3<T> mk_##P##vec3(tv2<A> a, B b) { return tv3<T>(wsc<T>(a.x), wsc<T>(a.y), wsc<T>(b)); } \
    template <class A, class B> WG_FN tv3<T> mk_##P##vec3(A a, tv2<B> b) { return tv3<T>(wsc<T>(a), wsc<T>(b.x), wsc<T>(b.y)); } \
    template <class A, class B, class C> WG_FN tv3<T> mk_##P##vec3(A a, B b, C c) { return tv3<T>(wsc<T>(a), wsc<T>(b), wsc<T>(c)); } \
    WG_FN tv4<T> mk_##P##vec4() { return tv4<T>(); }                                                                    \
    template <class A> WG_FN tv4<T> mk_##P##vec4(A s) { return tv4<T>(wsc<T>(s)); }                                     \
    template <class A> WG_FN tv4<T> mk_##P##vec4(tv4<A> v) { return tv4<T>(wsc<T>(v.x), wsc<T>(v.y), wsc<T>(v.z), wsc<T>(v.w)); } \
    template <class A, class B> WG_FN tv4<T> mk_##P##vec4(tv3<A> a, B b) { return tv4<T>(wsc<T>(a.x), wsc<T>(a.y), wsc<T>(a.z), wsc<T>(b)); } \
    template <class A, class B> WG_FN tv4<T> mk_##P##vec4(A a, tv3<B> b) { return tv4<T>(wsc<T>(a), wsc<T>(b.x), wsc<T>(b.y), wsc<T>(b.z)); } \
    template <class A, class B> WG_FN tv4<T> mk_##P##vec4(tv2<A> a, tv2<B> b) { return tv4<T>(wsc<T>(a.x), wsc<T>(a.y), wsc<T>(b.x), wsc<T>(b.y)); } \
    template <class A, class B, class C> WG_FN tv4<T> mk_##P##vec4(tv2<A> a, B b, C c) { return tv4<T>(wsc<T>(a.x), wsc<T>(a.y), wsc<T>(b), wsc<T>(c)); } \
    template <class A, class B, class C> WG_FN tv4<T> mk_##P##vec4(A a, tv2<B> b, C c) { return tv4<T>(wsc<T>(a), wsc<T>(b.x), wsc<T>(b.y), wsc<T>(c)); } \
    template <class A, class B, class C> WG_FN tv4<T> mk_##P##vec4(A a, B b, tv2<C> c) { return tv4<T>(wsc<T>(a), wsc<T>(b), wsc<T>(c.x), wsc<T>(c.y)); } \
    template <class A, class B, class C, class D> WG_FN tv4<T> mk_##P##vec4(A a, B b, C c, D d) { return tv4<T>(wsc<T>(a), wsc<T>(b), wsc<T>(c), wsc<T>(d)); }
WG_MK(w, WF)
WG_MK(wi, int)
WG_MK(wu, unsigned int)
WG_MK(wb, bool)
#undef WG_MK

// `var x: T = e`: e as T (an abstract literal or a vector of the same type)
template <class D, class S> struct wconv_ { static WG_FN D go(S s) { return wsc<D>(s); } };
template <class T> struct wconv_<T, T> { static WG_FN T go(T s) { return s; } };
template <class D, class S> WG_FN D wconv(S s) { return wconv_<D, S>::go(s); }

// ---- swizzles: wsw2(v, 0, 1) is v.xy ------------------------------------------------------------------------------
template <class V> WG_FN tv2<typename wscal<V>::type> wsw2(const V& v, int a, int b) { return tv2<typename wscal<V>::type>(v[a], v[b]); }
template <class V> WG_FN tv3<typename wscal<V>::type> wsw3(const V& v, int a, int b, int c) { return tv3<typename wscal<V>::type>(v[a], v[b], v[c]); }
template <class V> WG_FN tv4<typename wscal<V>::type> wsw4(const V& v, int a, int b, int c, int d) { return tv4<typename wscal<V>::type>(v[a], v[b], v[c], v[d]); }

// ---- componentwise operators ----------------------------------------------------------------------------------------
#define WG_MAP2(V, e) V(e(a.x, b.x), e(a.y, b.y))
#define WG_OPS(OP)                                                                                                      \
    template <class T> WG_FN tv2<T> operator OP(tv2<T> a, tv2<T> b) { return tv2<T>(a.x OP b.x, a.y OP b.y); }          \
    template <class T> WG_FN tv3<T> operator OP(tv3<T> a, tv3<T> b) { return tv3<T>(a.x OP b.x, a.y OP b.y, a.z OP b.z); } \
    template <class T> WG_FN tv4<T> operator OP(tv4<T> a, tv4<T> b) { return tv4<T>(a.x OP b.x, a.y OP b.y, a.z OP b.z, a.w OP b.w); } \
    template <class T, class S> WG_FN tv2<T> operator OP(tv2<T> a, S s) { const T b = T(s); return tv2<T>(a.x OP b, a.y OP b); } \
    template <class T, class S> WG_FN tv3<T> operator OP(tv3<T> a, S s) { const T b = T(s); return tv3<T>(a.x OP b, a.y OP b, a.z OP b); } \
    template <class T, class S> WG_FN tv4<T> operator OP(tv4<T> a, S s) { const T b = T(s); return tv4<T>(a.x OP b, a.y OP b, a.z OP b, a.w OP b); } \
    template <class T, class S> WG_FN tv2<T> operator OP(S s, tv2<T> a) { const T b = T(s); return tv2<T>(b OP a.x, b OP a.y); } \
    template <class T, class S> WG_FN tv3<T> operator OP(S s, tv3<T> a) { const T b = T(s); return tv3<T>(b OP a.x, b OP a.y, b OP a.z); } \
    template <class T, class S> WG_FN tv4<T> operator OP(S s, tv4<T> a) { const T b = T(s); return tv4<T>(b OP a.x, b OP a.y, b OP a.z, b OP a.w); } \
    template <class T, class S> WG_FN tv2<T>& operator OP##=(tv2<T>& a, S b) { a = a OP b; return a; }                 \
    template <class T, class S> WG_FN tv3<T>& operator OP##=(tv3<T>& a, S b) { a = a OP b; return a; }                 \
    template <class T, class S> WG_FN tv4<T>& operator OP##=(tv4<T>& a, S b) { a = a OP b; return a; }
WG_OPS(+)
WG_OPS(-)
WG_OPS(*)
WG_OPS(/)
WG_OPS(&)
WG_OPS(|)
WG_OPS(^)
WG_OPS(<<)
WG_OPS(>>)
#undef WG_OPS
#define WG_CMP(OP)                                                                                                      \
    template <class T> WG_FN wbvec2 operator OP(tv2<T> a, tv2<T> b) { return wbvec2(a.x OP b.x, a.y OP b.y); }           \
    template <class T> WG_FN wbvec3 operator OP(tv3<T> a, tv3<T> b) { return wbvec3(a.x OP b.x, a.y OP b.y, a.z OP b.z); } \
    template <class T> WG_FN wbvec4 operator OP(tv4<T> a, tv4<T> b) { return wbvec4(a.x OP b.x, a.y OP b.y, a.z OP b.z, a.w OP b.w); }
WG_CMP(<)
WG_CMP(>)
WG_CMP(<=)
WG_CMP(>=)
WG_CMP(==)
WG_CMP(!=)
#undef WG_CMP
template <class T> WG_FN tv2<T> operator-(tv2<T> a) { return tv2<T>(-a.x, -a.y); }
template <class T> WG_FN tv3<T> operator-(tv3<T> a) { return tv3<T>(-a.x, -a.y, -a.z); }
template <class T> WG_FN tv4<T> operator-(tv4<T> a) { return tv4<T>(-a.x, -a.y, -a.z, -a.w); }
template <class T> WG_FN tv2<T> operator~(tv2<T> a) { return tv2<T>(~a.x, ~a.y); }
template <class T> WG_FN tv3<T> operator~(tv3<T> a) { return tv3<T>(~a.x, ~a.y, ~a.z); }
template <class T> WG_FN tv4<T> operator~(tv4<T> a) { return tv4<T>(~a.x, ~a.y, ~a.z, ~a.w); }
WG_FN wbvec2 operator!(wbvec2 a) { return wbvec2(!a.x, !a.y); }
WG_FN wbvec3 operator!(wbvec3 a) { return wbvec3(!a.x, !a.y, !a.z); }
WG_FN wbvec4 operator!(wbvec4 a) { return wbvec4(!a.x, !a.y, !a.z, !a.w); }

// ---- a function of scalars, componentwise on vectors ----------------------------------------------------------------
#define WG_VEC1(name)                                                                                                   \
    template <class T> WG_FN tv2<T> name(tv2<T> a) { return tv2<T>(name(a.x), name(a.y)); }                            \
    template <class T> WG_FN tv3<T> name(tv3<T> a) { return tv3<T>(name(a.x), name(a.y), name(a.z)); }                  \
    template <class T> WG_FN tv4<T> name(tv4<T> a) { return tv4<T>(name(a.x), name(a.y), name(a.z), name(a.w)); }
#define WG_VEC2(name)                                                                                                   \
    template <class T> WG_FN tv2<T> name(tv2<T> a, tv2<T> b) { return tv2<T>(name(a.x, b.x), name(a.y, b.y)); }         \
    template <class T> WG_FN tv3<T> name(tv3<T> a, tv3<T> b) { return tv3<T>(name(a.x, b.x), name(a.y, b.y), name(a.z, b.z)); } \
    template <class T> WG_FN tv4<T> name(tv4<T> a, tv4<T> b) { return tv4<T>(name(a.x, b.x), name(a.y, b.y), name(a.z, b.z), name(a.w, b.w)); }
#define WG_VEC3(name)                                                                                                   \
    template <class T> WG_FN tv2<T> name(tv2<T> a, tv2<T> b, tv2<T> c) { return tv2<T>(name(a.x, b.x, c.x), name(a.y, b.y, c.y)); } \
    template <class T> WG_FN tv3<T> name(tv3<T> a, tv3<T> b, tv3<T> c) { return tv3<T>(name(a.x, b.x, c.x), name(a.y, b.y, c.y), name(a.z, b.z, c.z)); } \
    template <class T> WG_FN tv4<T> name(tv4<T> a, tv4<T> b, tv4<T> c) { return tv4<T>(name(a.x, b.x, c.x), name(a.y, b.y, c.y), name(a.z, b.z, c.z), name(a.w, b.w, c.w)); }

// the operands of a two- or three-argument builtin in one type: the usual promotions (an abstract int with a float is
// the float), a scalar splat to the other's vector (WGSL's mix, step, clamp with a scalar are not all legal, but an
// abstract literal beside a vector is)
template <class A, class B> struct wcommon { typedef decltype(A() + B()) type; };
template <class T> struct wcommon<tv2<T>, tv2<T> > { typedef tv2<T> type; };
template <class T> struct wcommon<tv3<T>, tv3<T> > { typedef tv3<T> type; };
template <class T> struct wcommon<tv4<T>, tv4<T> > { typedef tv4<T> type; };
template <class D, class S> struct wlift { static WG_FN D go(S s) { return D(s); } };
template <class T, class S> struct wlift<tv2<T>, S> { static WG_FN tv2<T> go(S s) { return tv2<T>(T(s)); } };
template <class T, class S> struct wlift<tv3<T>, S> { static WG_FN tv3<T> go(S s) { return tv3<T>(T(s)); } };
template <class T, class S> struct wlift<tv4<T>, S> { static WG_FN tv4<T> go(S s) { return tv4<T>(T(s)); } };
template <class T> struct wlift<tv2<T>, tv2<T> > { static WG_FN tv2<T> go(tv2<T> s) { return s; } };
template <class T> struct wlift<tv3<T>, tv3<T> > { static WG_FN tv3<T> go(tv3<T> s) { return s; } };
template <class T> struct wlift<tv4<T>, tv4<T> > { static WG_FN tv4<T> go(tv4<T> s) { return s; } };
template <class D, class S> WG_FN D wl(S s) { return wlift<D, S>::go(s); }

// scalar kernels
WG_FN WF wabs_(WF x) { return fabs(x); }
WG_FN int wabs_(int x) { return x < 0 ? -x : x; }
WG_FN unsigned int wabs_(unsigned int x) { return x; }
template <class T> WG_FN T wsign_(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }
template <class T> WG_FN T wmin_(T a, T b) { return b < a ? b : a; }
template <class T> WG_FN T wmax_(T a, T b) { return a < b ? b : a; }
WG_FN WF wmin_(WF a, WF b) { return fmin(a, b); }
WG_FN WF wmax_(WF a, WF b) { return fmax(a, b); }
template <class T> WG_FN T wclamp_(T x, T lo, T hi) { return wmin_(wmax_(x, lo), hi); }
WG_FN WF wmix_(WF a, WF b, WF t) { return a * (WF(1) - t) + b * t; }
WG_FN WF wstep_(WF e, WF x) { return x < e ? WF(0) : WF(1); }
WG_FN WF wsmooth_(WF e0, WF e1, WF x) { WF t = wclamp_((x - e0) / (e1 - e0), WF(0), WF(1)); return t * t * (WF(3) - WF(2) * t); }
WG_FN WF wfract_(WF x) { return x - floor(x); }
WG_FN WF wround_(WF x) { return rint(x); }   // half to even, as WGSL's round
WG_FN WF wisqrt_(WF x) { return WF(1) / sqrt(x); }
WG_FN WF wsat_(WF x) { return wclamp_(x, WF(0), WF(1)); }
WG_FN WF wfma_(WF a, WF b, WF c) { return fma(a, b, c); }
WG_FN WF watan2_(WF y, WF x) { return atan2(y, x); }
WG_FN WF wpow_(WF x, WF y) { return pow(x, y); }
WG_FN WF wmod_(WF a, WF b) { return a - b * trunc(a / b); }
WG_FN int wmod_(int a, int b) { return b == 0 ? a : a % b; }
WG_FN unsigned int wmod_(unsigned int a, unsigned int b) { return b == 0u ? a : a % b; }
#define WG_SC1(name, body) WG_FN WF name##_(WF x) { return body; }
WG_SC1(wsqrt, sqrt(x))
WG_SC1(wexp, exp(x))
WG_SC1(wexp2, exp2(x))
WG_SC1(wlog, log(x))
WG_SC1(wlog2, log2(x))
WG_SC1(wsin, sin(x))
WG_SC1(wcos, cos(x))
WG_SC1(wtan, tan(x))
WG_SC1(wasin, asin(x))
WG_SC1(wacos, acos(x))
WG_SC1(watan, atan(x))
WG_SC1(wsinh, sinh(x))
WG_SC1(wcosh, cosh(x))
WG_SC1(wtanh, tanh(x))
WG_SC1(wfloor, floor(x))
WG_SC1(wceil, ceil(x))
WG_SC1(wtrunc, trunc(x))
WG_SC1(wradians, x * WL(0.017453292519943295))
WG_SC1(wdegrees, x * WL(57.29577951308232))
#undef WG_SC1
#define WG_V1(k) WG_VEC1(k)
WG_V1(wabs_) WG_V1(wsign_) WG_V1(wfract_) WG_V1(wround_) WG_V1(wisqrt_) WG_V1(wsat_) WG_V1(wsqrt_) WG_V1(wexp_)
WG_V1(wexp2_) WG_V1(wlog_) WG_V1(wlog2_) WG_V1(wsin_) WG_V1(wcos_) WG_V1(wtan_) WG_V1(wasin_) WG_V1(wacos_)
WG_V1(watan_) WG_V1(wsinh_) WG_V1(wcosh_) WG_V1(wtanh_) WG_V1(wfloor_) WG_V1(wceil_) WG_V1(wtrunc_)
WG_V1(wradians_) WG_V1(wdegrees_)
#undef WG_V1
WG_VEC2(wmin_) WG_VEC2(wmax_) WG_VEC2(wstep_) WG_VEC2(watan2_) WG_VEC2(wpow_) WG_VEC2(wmod_)
WG_VEC3(wclamp_) WG_VEC3(wmix_) WG_VEC3(wsmooth_) WG_VEC3(wfma_)

// the WGSL names (w_<name>): arguments brought to one type, then the kernel
#define WG_F1(name, k) template <class A> WG_FN A name(A a) { return k(a); }
WG_F1(w_abs, wabs_) WG_F1(w_sign, wsign_) WG_F1(w_fract, wfract_) WG_F1(w_round, wround_) WG_F1(w_inverseSqrt, wisqrt_)
WG_F1(w_saturate, wsat_) WG_F1(w_sqrt, wsqrt_) WG_F1(w_exp, wexp_) WG_F1(w_exp2, wexp2_) WG_F1(w_log, wlog_)
WG_F1(w_log2, wlog2_) WG_F1(w_sin, wsin_) WG_F1(w_cos, wcos_) WG_F1(w_tan, wtan_) WG_F1(w_asin, wasin_)
WG_F1(w_acos, wacos_) WG_F1(w_atan, watan_) WG_F1(w_sinh, wsinh_) WG_F1(w_cosh, wcosh_) WG_F1(w_tanh, wtanh_)
WG_F1(w_floor, wfloor_) WG_F1(w_ceil, wceil_) WG_F1(w_trunc, wtrunc_) WG_F1(w_radians, wradians_)
WG_F1(w_degrees, wdegrees_)
#undef WG_F1
// a float function given an abstract int: WGSL converts the literal
WG_FN WF wf_(int a) { return WF(a); }
#define WG_FI(name) WG_FN WF name(int a) { return name(WF(a)); }
WG_FI(w_sqrt) WG_FI(w_exp) WG_FI(w_exp2) WG_FI(w_log) WG_FI(w_log2) WG_FI(w_sin) WG_FI(w_cos) WG_FI(w_tan) WG_FI(w_floor)
WG_FI(w_ceil) WG_FI(w_fract) WG_FI(w_round) WG_FI(w_inverseSqrt) WG_FI(w_saturate) WG_FI(w_radians) WG_FI(w_degrees)
#undef WG_FI
#define WG_F2(name, k) template <class A, class B> WG_FN typename wcommon<A, B>::type name(A a, B b) { typedef typename wcommon<A, B>::type C; return k(wl<C>(a), wl<C>(b)); }
WG_F2(w_min, wmin_) WG_F2(w_max, wmax_) WG_F2(w_step, wstep_) WG_F2(w_atan2, watan2_) WG_F2(w_pow, wpow_)
WG_F2(wmod, wmod_)
#undef WG_F2
template <class A, class B, class C> struct wcommon3 { typedef typename wcommon<typename wcommon<A, B>::type, C>::type type; };
#define WG_F3(name, k) template <class A, class B, class C> WG_FN typename wcommon3<A, B, C>::type name(A a, B b, C c) { typedef typename wcommon3<A, B, C>::type D; return k(wl<D>(a), wl<D>(b), wl<D>(c)); }
WG_F3(w_clamp, wclamp_) WG_F3(w_smoothstep, wsmooth_) WG_F3(w_fma, wfma_)
#undef WG_F3
// mix(a, b, t): t a scalar or a vector as a and b
template <class A, class B, class C> WG_FN typename wcommon<A, B>::type w_mix(A a, B b, C t) {
    typedef typename wcommon<A, B>::type D;
    return wmix_(wl<D>(a), wl<D>(b), wl<D>(t));
}
// select(f, t, cond): cond a bool, or a bool vector picking per component
template <class A, class B> WG_FN typename wcommon<A, B>::type w_select(A f, B t, bool c) { typedef typename wcommon<A, B>::type D; return c ? wl<D>(t) : wl<D>(f); }
template <class T> WG_FN tv2<T> w_select(tv2<T> f, tv2<T> t, wbvec2 c) { return tv2<T>(c.x ? t.x : f.x, c.y ? t.y : f.y); }
template <class T> WG_FN tv3<T> w_select(tv3<T> f, tv3<T> t, wbvec3 c) { return tv3<T>(c.x ? t.x : f.x, c.y ? t.y : f.y, c.z ? t.z : f.z); }
template <class T> WG_FN tv4<T> w_select(tv4<T> f, tv4<T> t, wbvec4 c) { return tv4<T>(c.x ? t.x : f.x, c.y ? t.y : f.y, c.z ? t.z : f.z, c.w ? t.w : f.w); }
WG_FN bool w_any(bool b) { return b; }
WG_FN bool w_all(bool b) { return b; }
WG_FN bool w_any(wbvec2 b) { return b.x || b.y; }
WG_FN bool w_any(wbvec3 b) { return b.x || b.y || b.z; }
WG_FN bool w_any(wbvec4 b) { return b.x || b.y || b.z || b.w; }
WG_FN bool w_all(wbvec2 b) { return b.x && b.y; }
WG_FN bool w_all(wbvec3 b) { return b.x && b.y && b.z; }
WG_FN bool w_all(wbvec4 b) { return b.x && b.y && b.z && b.w; }

// ---- geometry ---------------------------------------------------------------------------------------------------------
WG_FN WF w_dot(wvec2 a, wvec2 b) { return a.x * b.x + a.y * b.y; }
WG_FN WF w_dot(wvec3 a, wvec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
WG_FN WF w_dot(wvec4 a, wvec4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
WG_FN int w_dot(wivec3 a, wivec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
WG_FN wvec3 w_cross(wvec3 a, wvec3 b) { return wvec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
WG_FN WF w_length(WF a) { return fabs(a); }
template <class V> WG_FN WF w_length(V a) { return sqrt(w_dot(a, a)); }
template <class V> WG_FN WF w_distance(V a, V b) { return w_length(a - b); }
WG_FN WF w_distance(WF a, WF b) { return fabs(a - b); }
template <class V> WG_FN V w_normalize(V a) { return a * (WF(1) / sqrt(w_dot(a, a))); }
template <class V> WG_FN V w_reflect(V e, V n) { return e - n * (WF(2) * w_dot(n, e)); }
template <class V> WG_FN V w_faceForward(V n, V i, V r) { return w_dot(r, i) < WF(0) ? n : -n; }
template <class V> WG_FN V w_refract(V e, V n, WF eta) {
    const WF c = w_dot(n, e), k = WF(1) - eta * eta * (WF(1) - c * c);
    return k < WF(0) ? V() : e * eta - n * (eta * c + sqrt(k));
}

// ---- square matrices, column-major (m[i] is column i) --------------------------------------------------------------
struct wmat2 { wvec2 c[2]; WG_FN wvec2& operator[](int i) { return c[i]; } WG_FN wvec2 operator[](int i) const { return c[i]; } };
struct wmat3 { wvec3 c[3]; WG_FN wvec3& operator[](int i) { return c[i]; } WG_FN wvec3 operator[](int i) const { return c[i]; } };
struct wmat4 { wvec4 c[4]; WG_FN wvec4& operator[](int i) { return c[i]; } WG_FN wvec4 operator[](int i) const { return c[i]; } };
WG_FN wmat2 mk_wmat2() { wmat2 m; return m; }
WG_FN wmat3 mk_wmat3() { wmat3 m; return m; }
WG_FN wmat4 mk_wmat4() { wmat4 m; return m; }
WG_FN wmat2 mk_wmat2(wvec2 a, wvec2 b) { wmat2 m; m.c[0] = a; m.c[1] = b; return m; }
WG_FN wmat3 mk_wmat3(wvec3 a, wvec3 b, wvec3 c) { wmat3 m; m.c[0] = a; m.c[1] = b; m.c[2] = c; return m; }
WG_FN wmat4 mk_wmat4(wvec4 a, wvec4 b, wvec4 c, wvec4 d) { wmat4 m; m.c[0] = a; m.c[1] = b; m.c[2] = c; m.c[3] = d; return m; }
template <class A, class B, class C, class D> WG_FN wmat2 mk_wmat2(A a, B b, C c, D d) { return mk_wmat2(mk_wvec2(a, b), mk_wvec2(c, d)); }
template <class A, class B, class C, class D, class E, class F, class G, class H, class I>
WG_FN wmat3 mk_wmat3(A a, B b, C c, D d, E e, F f, G g, H h, I i) { return mk_wmat3(mk_wvec3(a, b, c), mk_wvec3(d, e, f), mk_wvec3(g, h, i)); }
WG_FN wvec2 operator*(wmat2 m, wvec2 v) { return m.c[0] * v.x + m.c[1] * v.y; }
WG_FN wvec3 operator*(wmat3 m, wvec3 v) { return m.c[0] * v.x + m.c[1] * v.y + m.c[2] * v.z; }
WG_FN wvec4 operator*(wmat4 m, wvec4 v) { return m.c[0] * v.x + m.c[1] * v.y + m.c[2] * v.z + m.c[3] * v.w; }
WG_FN wvec2 operator*(wvec2 v, wmat2 m) { return wvec2(w_dot(v, m.c[0]), w_dot(v, m.c[1])); }
WG_FN wvec3 operator*(wvec3 v, wmat3 m) { return wvec3(w_dot(v, m.c[0]), w_dot(v, m.c[1]), w_dot(v, m.c[2])); }
WG_FN wvec4 operator*(wvec4 v, wmat4 m) { return wvec4(w_dot(v, m.c[0]), w_dot(v, m.c[1]), w_dot(v, m.c[2]), w_dot(v, m.c[3])); }
WG_FN wmat2 operator*(wmat2 a, wmat2 b) { return mk_wmat2(a * b.c[0], a * b.c[1]); }
WG_FN wmat3 operator*(wmat3 a, wmat3 b) { return mk_wmat3(a * b.c[0], a * b.c[1], a * b.c[2]); }
WG_FN wmat4 operator*(wmat4 a, wmat4 b) { return mk_wmat4(a * b.c[0], a * b.c[1], a * b.c[2], a * b.c[3]); }
template <class S> WG_FN wmat3 operator*(wmat3 a, S s) { return mk_wmat3(a.c[0] * s, a.c[1] * s, a.c[2] * s); }
template <class S> WG_FN wmat3 operator*(S s, wmat3 a) { return mk_wmat3(a.c[0] * s, a.c[1] * s, a.c[2] * s); }
WG_FN wmat3 operator+(wmat3 a, wmat3 b) { return mk_wmat3(a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]); }
WG_FN wmat3 operator-(wmat3 a, wmat3 b) { return mk_wmat3(a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2]); }
WG_FN wmat3& operator*=(wmat3& a, wmat3 b) { a = a * b; return a; }
WG_FN wmat3 w_transpose(wmat3 m) {
    return mk_wmat3(wvec3(m.c[0].x, m.c[1].x, m.c[2].x), wvec3(m.c[0].y, m.c[1].y, m.c[2].y), wvec3(m.c[0].z, m.c[1].z, m.c[2].z));
}
WG_FN WF w_determinant(wmat3 m) { return w_dot(m.c[0], w_cross(m.c[1], m.c[2])); }
WG_FN WF w_determinant(wmat2 m) { return m.c[0].x * m.c[1].y - m.c[1].x * m.c[0].y; }

// ---- bit reinterpretation (bitcast<T>) ----------------------------------------------------------------------------------
WG_FN unsigned int wbitcast_unsigned_int(int a) { return (unsigned int)a; }
WG_FN int wbitcast_int(unsigned int a) { return (int)a; }
WG_FN wuvec3 wbitcast_wuvec3(wivec3 a) { return wuvec3((unsigned int)a.x, (unsigned int)a.y, (unsigned int)a.z); }
WG_FN wivec3 wbitcast_wivec3(wuvec3 a) { return wivec3((int)a.x, (int)a.y, (int)a.z); }
