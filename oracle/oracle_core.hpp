// Test infrastructure only: the oracle.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker.  The product path never calls it.
//
// CPU restatement of the ADECO deposit path (kaehlerr/ADECO, src/), written from the reference's
// semantics.  Templated on the coordinate type T and the grid type G:
//   T = G = float   "float-faithful": reproduces the reference's float arithmetic op for op,
//                   including its double temporaries (1.-x, /6., 0.16666666666666666*vol), so on
//                   the same inputs and in the same order its grids are bit-identical to the
//                   compiled reference (checked against oracle/_ref by the tests);
//   T = G = double  the FP64 accuracy oracle the HIP kernels are held to;
//   T = Counted<double>  the same FP64 arithmetic with every floating-point operation counted.
// Must be compiled with -ffp-contract=off (no fused multiply-add) so float results are
// reproducible; oracle/Makefile does so.
//
// Tables that have to be alike: the 16-case sub-tet table (kSubTets, kSubTetCount, kEdges) and the
// 6-tet cube table (kTetConn) are the reference's (AdTetIntersection.h:66-96,
// AdDepositTets2Grid.h:207-231), because the vertex order of every sub-tet fixes the float
// rounding that a bit-exact restatement must reproduce.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstddef>
#include <cstring>
#include <algorithm>
#include <limits>
#include <vector>
#include <stack>

namespace oracle {

// ---------------------------------------------------------------- op counting
struct OpCounts {
    uint64_t add = 0, mul = 0, div = 0, cmp = 0;
    uint64_t flops() const { return add + mul + div; }
};
extern thread_local OpCounts* g_ops;

template <class R>
struct Counted {
    R v;
    Counted() : v(0) {}
    Counted(R x) : v(x) {}  // NOLINT: implicit like a builtin
    explicit operator R() const { return v; }
    explicit operator int() const { return int(v); }
    static void a() { if (g_ops) ++g_ops->add; }
    static void m() { if (g_ops) ++g_ops->mul; }
    static void d() { if (g_ops) ++g_ops->div; }
    static void c() { if (g_ops) ++g_ops->cmp; }
    friend Counted operator+(Counted x, Counted y) { a(); return Counted(x.v + y.v); }
    friend Counted operator-(Counted x, Counted y) { a(); return Counted(x.v - y.v); }
    friend Counted operator*(Counted x, Counted y) { m(); return Counted(x.v * y.v); }
    friend Counted operator/(Counted x, Counted y) { d(); return Counted(x.v / y.v); }
    Counted operator-() const { return Counted(-v); }
    Counted& operator+=(Counted y) { a(); v += y.v; return *this; }
    Counted& operator-=(Counted y) { a(); v -= y.v; return *this; }
    Counted& operator*=(Counted y) { m(); v *= y.v; return *this; }
    friend bool operator<(Counted x, Counted y) { c(); return x.v < y.v; }
    friend bool operator>(Counted x, Counted y) { c(); return x.v > y.v; }
    friend bool operator<=(Counted x, Counted y) { c(); return x.v <= y.v; }
    friend bool operator>=(Counted x, Counted y) { c(); return x.v >= y.v; }
    friend bool operator==(Counted x, Counted y) { c(); return x.v == y.v; }
    friend bool operator!=(Counted x, Counted y) { c(); return x.v != y.v; }
};

template <class T> struct NumTraits { using D = double; static double dbl(T x) { return double(x); } };
template <> struct NumTraits<long double> { using D = long double; static long double dbl(long double x) { return x; } };
template <class R> struct NumTraits<Counted<R>> {
    using D = Counted<R>;
    static Counted<R> dbl(Counted<R> x) { return x; }
};

inline float fabs_t(float x) { return std::fabs(x); }
inline double fabs_t(double x) { return std::fabs(x); }
inline long double fabs_t(long double x) { return std::fabs(x); }
inline long double floor_t(long double x) { return std::floor(x); }
inline long double ceil_t(long double x) { return std::ceil(x); }
template <class R> Counted<R> fabs_t(Counted<R> x) { return Counted<R>(std::fabs(x.v)); }
inline float floor_t(float x) { return std::floor(x); }
inline double floor_t(double x) { return std::floor(x); }
template <class R> Counted<R> floor_t(Counted<R> x) { return Counted<R>(std::floor(x.v)); }
inline float ceil_t(float x) { return std::ceil(x); }
inline double ceil_t(double x) { return std::ceil(x); }
template <class R> Counted<R> ceil_t(Counted<R> x) { return Counted<R>(std::ceil(x.v)); }
template <class T> int to_int(T x) { return int(x); }

// "1. - x" and friends: the reference mixes a double literal with T; for T=float the difference is
// formed in double and rounded to float once.
template <class T> inline T one_minus(T x) { return T(1. - NumTraits<T>::dbl(x)); }
template <class T> inline T minus_one(T x) { return T(NumTraits<T>::dbl(x) - 1.); }
template <class R> inline Counted<R> one_minus(Counted<R> x) { return Counted<R>(1.) - x; }
template <class R> inline Counted<R> minus_one(Counted<R> x) { return x - Counted<R>(1.); }

// ---------------------------------------------------------------- tables
// AdTetIntersection.h:66-96
static const unsigned char kSubTets[16][3][4] = {
    {{255, 255, 255, 255}, {255, 255, 255, 255}, {255, 255, 255, 255}},
    {{0, 5, 6, 4}, {255, 255, 255, 255}, {255, 255, 255, 255}},
    {{7, 1, 8, 4}, {255, 255, 255, 255}, {255, 255, 255, 255}},
    {{0, 6, 8, 7}, {6, 0, 5, 7}, {1, 0, 8, 7}},
    {{2, 7, 9, 5}, {255, 255, 255, 255}, {255, 255, 255, 255}},
    {{6, 0, 9, 7}, {0, 6, 4, 7}, {0, 2, 9, 7}},
    {{5, 1, 9, 8}, {9, 1, 5, 2}, {1, 5, 4, 8}},
    {{6, 0, 9, 8}, {0, 2, 9, 8}, {0, 1, 2, 8}},
    {{8, 3, 9, 6}, {255, 255, 255, 255}, {255, 255, 255, 255}},
    {{0, 5, 9, 8}, {5, 0, 4, 8}, {3, 0, 9, 8}},
    {{1, 6, 9, 7}, {1, 9, 6, 3}, {6, 1, 4, 7}},
    {{0, 5, 9, 7}, {3, 0, 9, 7}, {1, 0, 3, 7}},
    {{6, 2, 8, 7}, {6, 8, 2, 3}, {2, 6, 5, 7}},
    {{4, 0, 8, 7}, {0, 3, 8, 7}, {0, 2, 3, 7}},
    {{1, 4, 6, 5}, {3, 1, 6, 5}, {2, 1, 3, 5}},
    {{0, 2, 3, 1}, {255, 255, 255, 255}, {255, 255, 255, 255}}};
static const int kSubTetCount[16] = {0, 1, 1, 3, 1, 3, 3, 3, 1, 3, 3, 3, 3, 3, 3, 1};
static const int kEdges[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

// AdDepositTets2Grid.h:207-231: the 6-tet table over the 8 cube slots
static const int kTetConn[6][4] = {{1, 0, 2, 4}, {4, 2, 1, 3}, {3, 5, 1, 4},
                                   {4, 5, 6, 3}, {3, 2, 6, 4}, {3, 7, 5, 6}};

// Cube-corner codes (dx | dy<<1 | dz<<2) of the 4 vertices of each of the 6 tets, per
// decomposition.  0 = REF_EFFECTIVE: the slots as sample_cell_vertices fills them
// (AdDepositTets2Grid.h:947-972; P0..P7 = 000,100,010,110,011,111,001,101 in (i,j,k) offsets);
// 1 = REF_INTENDED: the slots in vert[] order, index 4x+2y+z (AdDepositTets2Grid.h:188-192);
// 2 = KUHN: the six tets along the (0,0,0)-(1,1,1) diagonal.
enum Decomp { REF_EFFECTIVE = 0, REF_INTENDED = 1, KUHN = 2 };
inline void decomposition_corners(int decomp, int out[6][4]) {
    static const int slot_eff[8] = {0, 1, 2, 3, 6, 7, 4, 5};
    static const int slot_int[8] = {0, 4, 2, 6, 1, 5, 3, 7};
    if (decomp == KUHN) {
        static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        for (int t = 0; t < 6; ++t) {
            out[t][0] = 0;
            out[t][1] = 1 << perm[t][0];
            out[t][2] = (1 << perm[t][0]) | (1 << perm[t][1]);
            out[t][3] = 7;
        }
        return;
    }
    const int* slot = decomp == REF_INTENDED ? slot_int : slot_eff;
    for (int t = 0; t < 6; ++t)
        for (int v = 0; v < 4; ++v) out[t][v] = slot[kTetConn[t][v]];
}

// ---------------------------------------------------------------- clip (AdTetIntersection.h)
struct ClipStats {
    uint64_t faces = 0, edge_points = 0, subtet_vols = 0;
};

template <class T>
inline T vol6(const T* a, const T* b, const T* c, const T* d) {  // AdTetIntersection.h:183-203
    const T ady = a[1] - d[1], bdy = b[1] - d[1], cdy = c[1] - d[1];
    const T adz = a[2] - d[2], bdz = b[2] - d[2], cdz = c[2] - d[2];
    return (a[0] - d[0]) * (bdy * cdz - bdz * cdy) + (b[0] - d[0]) * (cdy * adz - cdz * ady) +
           (c[0] - d[0]) * (ady * bdz - adz * bdy);
}

template <class T>
inline int sgn(T a) { return int(T(a > T(0.f)) - T(a < T(0.f))); }  // AdTetIntersection.h:99-104

// AdTetIntersection.h:230-257
template <class T>
inline void edge_point(int face, int o, int dir, const T* a, const T* b, T* out) {
    T nom = face < 4 ? (dir == 0 ? a[o] : one_minus(a[o])) : (dir == 0 ? -a[o] : minus_one(a[o]));
    const T l[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    T den = face < 4 ? (dir == 0 ? -l[o] : l[o]) : (dir == 0 ? l[o] : -l[o]);
    const T tmp = den == T(0.) ? T(0.) : nom / den;
    const T d = (tmp <= T(0.)) ? T(0.) : ((tmp <= T(1.)) ? tmp : T(1.));
    out[0] = a[0] + d * l[0];
    out[1] = a[1] + d * l[1];
    out[2] = a[2] + d * l[2];
}

// AdTetIntersection.h:261-358
template <class T>
inline int process_face(int f, int axis, int dir, int& lookup, T pts[10][3], T& dvol, ClipStats* st) {
    int vo[4];
    for (int p = 0; p < 4; ++p) {
        T o;
        if (f < 4) o = dir == 0 ? pts[p][axis] : one_minus(pts[p][axis]);
        else o = dir == 0 ? -pts[p][axis] : minus_one(pts[p][axis]);
        vo[p] = (1 + sgn(o)) >> 1;
    }
    lookup = vo[0] | (vo[1] << 1) | (vo[2] << 2) | (vo[3] << 3);
    if (st) ++st->faces;
    const int n = kSubTetCount[lookup];
    if (n == 0) return 0;
    for (int e = 0; e < 6; ++e)
        if (vo[kEdges[e][0]] != vo[kEdges[e][1]]) {
            edge_point(f, axis, dir, pts[kEdges[e][0]], pts[kEdges[e][1]], pts[e + 4]);
            if (st) ++st->edge_points;
        }
    if (f > 3)
        for (int i = 0; i < n; ++i) {
            const unsigned char* c = kSubTets[lookup][i];
            dvol += vol6(pts[c[0]], pts[c[1]], pts[c[2]], pts[c[3]]);
            if (st) ++st->subtet_vols;
        }
    return n;
}

template <class T>
inline void take_subtet(int lookup, int k, const T src[10][3], T dst[10][3]) {
    const unsigned char* c = kSubTets[lookup][k];
    for (int v = 0; v < 4; ++v)
        for (int a = 0; a < 3; ++a) dst[v][a] = src[c[v]][a];
}

// |tet ∩ [0,1]^3| for a positively oriented tet (AdTetIntersection.h:383-476)
template <class T>
inline T clip_unit_cube(const T tet[4][3], ClipStats* st = nullptr) {
    T volume = T(0.);
    T dummy = T(0.);
    T p1[10][3];
    for (int v = 0; v < 4; ++v)
        for (int a = 0; a < 3; ++a) p1[v][a] = tet[v][a];
    int l1, l2, l3, l4, l56;
    const int n1 = process_face(0, 0, 0, l1, p1, dummy, st);
    for (int t1 = 0; t1 < n1; ++t1) {
        T p2[10][3];
        take_subtet(l1, t1, p1, p2);
        const int n2 = process_face(1, 0, 1, l2, p2, dummy, st);
        for (int t2 = 0; t2 < n2; ++t2) {
            T p3[10][3];
            take_subtet(l2, t2, p2, p3);
            const int n3 = process_face(2, 1, 0, l3, p3, dummy, st);
            for (int t3 = 0; t3 < n3; ++t3) {
                T p4[10][3];
                take_subtet(l3, t3, p3, p4);
                const int n4 = process_face(3, 1, 1, l4, p4, dummy, st);
                for (int t4 = 0; t4 < n4; ++t4) {
                    T p5[10][3];
                    take_subtet(l4, t4, p4, p5);
                    T dvol1 = T(0.);
                    process_face(4, 2, 0, l56, p5, dvol1, st);
                    T dvol2 = T(0.);
                    process_face(5, 2, 1, l56, p5, dvol2, st);
                    volume += vol6(p5[0], p5[1], p5[2], p5[3]) - dvol1 - dvol2;
                    if (st) ++st->subtet_vols;
                }
            }
        }
    }
    return T(0.16666666666666666 * NumTraits<T>::dbl(volume));
}

// ---------------------------------------------------------------- per-tet deposit
enum Intersection { NO_INTERSECTION = 0, TET_INSIDE_CELL, CELL_INSIDE_TET, PARTIAL };

template <class T>
struct Box {
    T mn[3], mx[3];
};

template <class T>
inline bool strict_overlap(const Box<T>& a, const Box<T>& b) {  // AdDepositTets2Grid.h:87-94
    return a.mn[0] < b.mx[0] && b.mn[0] < a.mx[0] && a.mn[1] < b.mx[1] && b.mn[1] < a.mx[1] &&
           a.mn[2] < b.mx[2] && b.mn[2] < a.mx[2];
}

template <class T>
inline bool inside(const Box<T>& b, const T* p) {  // AdDepositTets2Grid.h:77-82
    return (p[0] >= b.mn[0] && p[0] <= b.mx[0]) && (p[1] >= b.mn[1] && p[1] <= b.mx[1]) &&
           (p[2] >= b.mn[2] && p[2] <= b.mx[2]);
}

template <class T>
inline void cross3(const T* a, const T* b, T* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

// AdDepositTets2Grid.h:286-363, evaluated against the unit cell in vert[] corner order
template <class T>
inline Intersection intersection_test(const T s[4][3]) {
    T e[6][3];
    for (int a = 0; a < 3; ++a) {
        e[0][a] = s[1][a] - s[0][a];
        e[1][a] = s[2][a] - s[0][a];
        e[2][a] = s[3][a] - s[0][a];
        e[3][a] = s[3][a] - s[1][a];
        e[4][a] = s[2][a] - s[1][a];
    }
    T n[4][3];
    cross3(e[0], e[1], n[0]);
    cross3(e[1], e[2], n[1]);
    cross3(e[2], e[0], n[2]);
    cross3(e[3], e[4], n[3]);
    const T* off[4] = {s[0], s[0], s[0], s[1]};
    int combined = 0;
    for (int f = 0; f < 4; ++f) {
        int lookup = 0;
        for (int p = 0; p < 8; ++p) {
            const T cv[3] = {T(float((p >> 2) & 1)), T(float((p >> 1) & 1)), T(float(p & 1))};
            const T o = (off[f][0] - cv[0]) * n[f][0] + (off[f][1] - cv[1]) * n[f][1] +
                        (off[f][2] - cv[2]) * n[f][2];
            lookup |= ((1 + sgn(o)) >> 1) << p;
        }
        if (lookup == 0) return NO_INTERSECTION;
        combined += lookup;
    }
    if (combined == 1020) return CELL_INSIDE_TET;
    Box<T> unit;
    for (int a = 0; a < 3; ++a) { unit.mn[a] = T(0.f); unit.mx[a] = T(1.f); }
    if (inside(unit, s[0]) && inside(unit, s[1]) && inside(unit, s[2]) && inside(unit, s[3]))
        return TET_INSIDE_CELL;
    return PARTIAL;
}

struct DepositStats {
    uint64_t tets = 0, tets_active = 0, candidates = 0, aabb_pass = 0;
    uint64_t no = 0, tet_inside = 0, cell_inside = 0, partial = 0, nonzero = 0;
    ClipStats clip;
};

template <class T>
inline int clampi(int v, int lo, int hi) { return v <= lo ? lo : (v >= hi ? hi : v); }

// AdDepositTets2Grid.h:509-665 (grid index i + dims*(j + dims*k), accumulate with +=).
// per_axis_scale = false is the reference's scale = 1/dx on every axis (AdDepositTets2Grid.h:569);
// true scales each axis by its own 1/delta (an exact partition of a non-cubic float box).
template <class G, class T>
inline void deposit_tet(const T v[4][3], G tet_mass, const int dims[3], const Box<T>& gb, G* grid,
                        DepositStats* st, bool per_axis_scale = false) {
    if (st) ++st->tets;
    Box<T> tb;
    for (int a = 0; a < 3; ++a) tb.mn[a] = tb.mx[a] = v[0][a];
    for (int i = 1; i < 4; ++i)
        for (int a = 0; a < 3; ++a) {
            if (v[i][a] < tb.mn[a]) tb.mn[a] = v[i][a];
            else if (v[i][a] > tb.mx[a]) tb.mx[a] = v[i][a];
        }
    if (!strict_overlap(tb, gb)) return;
    Box<T> ov;
    for (int a = 0; a < 3; ++a) {
        ov.mn[a] = tb.mn[a] < gb.mn[a] ? gb.mn[a] : tb.mn[a];
        ov.mx[a] = tb.mx[a] < gb.mx[a] ? tb.mx[a] : gb.mx[a];
    }
    T delta[3];
    for (int a = 0; a < 3; ++a) delta[a] = (gb.mx[a] - gb.mn[a]) / T(float(dims[a]));
    const T cell_vol = delta[0] * delta[1] * delta[2];
    int cov[3][2];
    for (int a = 0; a < 3; ++a) {
        cov[a][0] = clampi<T>(to_int(floor_t((ov.mn[a] - gb.mn[a]) / delta[a])), 0, dims[a] - 1);
        cov[a][1] = clampi<T>(to_int(ceil_t((ov.mx[a] - gb.mn[a]) / delta[a])), 0, dims[a] - 1);
    }
    const T tet_vol = T(NumTraits<T>::dbl(vol6(v[0], v[1], v[2], v[3])) / 6.);
    if (tet_vol == T(0.f)) return;
    T src[4][3];
    const int order_pos[4] = {0, 1, 2, 3}, order_neg[4] = {3, 1, 2, 0};
    const int* ord = tet_vol < T(0.f) ? order_neg : order_pos;
    for (int i = 0; i < 4; ++i)
        for (int a = 0; a < 3; ++a) src[i][a] = v[ord[i]][a];
    if (st) ++st->tets_active;
    T scale[3];
    for (int a = 0; a < 3; ++a) scale[a] = T(1.f) / delta[per_axis_scale ? a : 0];
    const T abs_vol = fabs_t(tet_vol);
    Box<T> cb;
    for (int k = cov[2][0]; k <= cov[2][1]; ++k) {
        cb.mn[2] = gb.mn[2] + T(float(k)) * delta[2];
        cb.mx[2] = cb.mn[2] + delta[2];
        for (int j = cov[1][0]; j <= cov[1][1]; ++j) {
            cb.mn[1] = gb.mn[1] + T(float(j)) * delta[1];
            cb.mx[1] = cb.mn[1] + delta[1];
            for (int i = cov[0][0]; i <= cov[0][1]; ++i) {
                cb.mn[0] = gb.mn[0] + T(float(i)) * delta[0];
                cb.mx[0] = cb.mn[0] + delta[0];
                if (st) ++st->candidates;
                if (!strict_overlap(tb, cb)) continue;
                if (st) ++st->aabb_pass;
                T s[4][3];
                for (int q = 0; q < 4; ++q)
                    for (int a = 0; a < 3; ++a) s[q][a] = scale[a] * (src[q][a] - cb.mn[a]);
                const Intersection it = intersection_test(s);
                G dmass = G(0.f);
                if (it == NO_INTERSECTION) {
                    if (st) ++st->no;
                    continue;
                } else if (it == TET_INSIDE_CELL) {
                    if (st) ++st->tet_inside;
                    dmass = tet_mass;
                } else if (it == CELL_INSIDE_TET) {
                    if (st) ++st->cell_inside;
                    dmass = tet_mass * G(cell_vol / abs_vol);
                } else {
                    if (st) ++st->partial;
                    T dvol = clip_unit_cube(s, st ? &st->clip : nullptr);
                    if (dvol > T(0.f)) {
                        dvol *= cell_vol;
                        dmass = tet_mass * G(dvol / abs_vol);
                    }
                }
                if (dmass > G(0.f)) {
                    if (st) ++st->nonzero;
                    grid[size_t(i) + size_t(dims[0]) * (size_t(j) + size_t(dims[1]) * size_t(k))] += dmass;
                }
            }
        }
    }
}

// ---------------------------------------------------------------- meta-patch loop
// AdDepositTets2Grid.h:917-1053 and AdDeposit2Grid.cpp:67-125
template <class G, class T>
inline void deposit_patch_cell(int cell, const float* patch, int m, const Box<T>& gb, int dims,
                               G tet_mass, G* grid, DepositStats* st, bool per_axis_scale = false) {
    const int mm = m * m;
    const int k = cell / mm, rest = cell % mm, j = rest / m, i = rest % m;
    const int base = 3 * (i + (m + 1) * (j + k * (m + 1)));
    const int ax = 3, ay = 3 * (m + 1), az = 3 * (m + 1) * (m + 1);
    const int slot_off[8] = {0, ax, ay, ay + ax, ay + az, ay + az + ax, az, az + ax};
    float cp[8][3];
    for (int s = 0; s < 8; ++s)
        for (int a = 0; a < 3; ++a) cp[s][a] = patch[base + slot_off[s] + a];
    for (int s = 0; s < 8; ++s)
        if (double(std::fabs(cp[s][0] - 666.f) + std::fabs(cp[s][1] - 666.f) + std::fabs(cp[s][2] - 666.f)) < 1.E-03)
            return;
    const int gd[3] = {dims, dims, dims};
    for (int t = 0; t < 6; ++t) {
        T v[4][3];
        for (int q = 0; q < 4; ++q)
            for (int a = 0; a < 3; ++a) v[q][a] = T(cp[kTetConn[t][q]][a]);
        deposit_tet<G, T>(v, tet_mass, gd, gb, grid, st, per_axis_scale);
    }
}

template <class T>
inline Box<T> bbox6(const float* b) {
    Box<T> g;
    g.mn[0] = T(b[0]); g.mx[0] = T(b[1]);
    g.mn[1] = T(b[2]); g.mx[1] = T(b[3]);
    g.mn[2] = T(b[4]); g.mx[2] = T(b[5]);
    return g;
}

template <class T>
inline Box<T> bbox6d(const double* b) {
    Box<T> g;
    g.mn[0] = T(b[0]); g.mx[0] = T(b[1]);
    g.mn[1] = T(b[2]); g.mx[1] = T(b[3]);
    g.mn[2] = T(b[4]); g.mx[2] = T(b[5]);
    return g;
}

// ---------------------------------------------------------------- lattice tets
// Tet `t` of an (npx,npy,npz) SoA vertex lattice: cube c = t/6 (i fastest over (np-1)^3 cubes),
// local tet t%6 through the decomposition's corner table.
template <class T>
inline void lattice_tet(const float* x, const float* y, const float* z, const int np[3],
                        const int corners[6][4], int64_t t, T v[4][3]) {
    const int64_t c = t / 6;
    const int lt = int(t % 6);
    const int64_t ncx = np[0] - 1, ncy = np[1] - 1;
    const int64_t i = c % ncx, j = (c / ncx) % ncy, k = c / (ncx * ncy);
    for (int q = 0; q < 4; ++q) {
        const int code = corners[lt][q];
        const int64_t vi = (i + (code & 1)) + int64_t(np[0]) * ((j + ((code >> 1) & 1)) + int64_t(np[1]) * (k + ((code >> 2) & 1)));
        v[q][0] = T(x[vi]);
        v[q][1] = T(y[vi]);
        v[q][2] = T(z[vi]);
    }
}

}  // namespace oracle
