// Exact tetrahedron ∩ cell geometry for the ADECO mass deposit, written for CDNA4 (gfx950) and
// also compiled for the host so the CPU test suite can check the same code.
//
// Semantics follow the reference CPU deposit (AdDepositTets2Grid.h:509-665,
// AdTetIntersection.h:183-476): per tet, the covered cell range of its AABB, the strict AABB cull,
// the scaled cell-local tet s = (v - cell_min) / dx, a classification into
// no-intersection / tet-inside-cell / cell-inside-tet / partial, and for partial cells the exact
// volume |s ∩ [0,1]^3|.  The clip itself is this project's own formulation (the reference walks a
// 16-case lookup table through six faces with 10-point scratch arrays per level, which on a GPU
// means dynamically indexed private memory):
//   * the x- and y-faces are applied only when they actually cut the tet; each cut turns a tet
//     into one tet (1 vertex inside), into the tet minus the corner tet beyond the plane (3
//     inside: two signed pieces), or into a wedge of 3 tets (2 inside), written as a point
//     sequence whose pieces are windows of 4 consecutive points, so every piece is addressed with
//     selects, never a table lookup;
//   * the z-slab of each remaining sub-tet is integrated in closed form: the fraction of a tet
//     beyond a plane, from the signed distances d_i of its vertices, is
//       1 vertex beyond (a = d > 0, others b_j >= 0 on the near side):  a^3 / prod_j (a + b_j)
//       2 beyond (a, b) and 2 near (c, e):
//         [a^2 b^2 + ab(a+b)(c+e) + ce(a^2+ab+b^2)] / ((a+c)(a+e)(b+c)(b+e))
//       3 beyond: 1 - (the 1-vertex formula seen from the near vertex)
//     (all terms positive: no cancellation), so a sub-tet costs one determinant and a few
//     multiplies instead of two more table-driven clips.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ADECO_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define ADECO_HD inline
#endif

namespace adeco {

// Cube-corner codes (dx | dy<<1 | dz<<2) of the 4 vertices of each of the 6 tets of a lattice
// cube.  REF_EFFECTIVE is the decomposition the reference CPU/GPU deposit actually applies: the
// 6-tet table of AdDepositTets2Grid.h:213-218 over the vertex slots as sample_cell_vertices fills
// them (AdDepositTets2Grid.h:947-972).  REF_INTENDED is the same table over the vert[] slot order
// (AdDepositTets2Grid.h:188-192), which tiles the cube.  KUHN is the six tets around the
// (0,0,0)-(1,1,1) diagonal.
enum Decomposition : int { REF_EFFECTIVE = 0, REF_INTENDED = 1, KUHN = 2 };

struct CornerTable {
    unsigned char c[6][4];
};

// The same table packed into two 64-bit words (16 bits per tet, 3 bits per corner), so a kernel
// reads corner codes with shifts instead of indexing a private array.
struct PackedCorners {
    unsigned long long lo, hi;  // tets 0-3, tets 4-5
    ADECO_HD int code(int lt, int q) const {
        const unsigned long long w = lt < 4 ? (lo >> (16 * lt)) : (hi >> (16 * (lt - 4)));
        return int((w >> (3 * q)) & 7u);
    }
};

ADECO_HD CornerTable corner_table(int decomp) {
    CornerTable t;
    // slot -> corner code for the two reference slot orders
    const unsigned char eff[8] = {0, 1, 2, 3, 6, 7, 4, 5};
    const unsigned char intended[8] = {0, 4, 2, 6, 1, 5, 3, 7};
    const unsigned char conn[6][4] = {{1, 0, 2, 4}, {4, 2, 1, 3}, {3, 5, 1, 4},
                                      {4, 5, 6, 3}, {3, 2, 6, 4}, {3, 7, 5, 6}};
    const unsigned char perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (int i = 0; i < 6; ++i) {
        if (decomp == KUHN) {
            t.c[i][0] = 0;
            t.c[i][1] = (unsigned char)(1 << perm[i][0]);
            t.c[i][2] = (unsigned char)((1 << perm[i][0]) | (1 << perm[i][1]));
            t.c[i][3] = 7;
        } else {
            const unsigned char* s = decomp == REF_INTENDED ? intended : eff;
            for (int v = 0; v < 4; ++v) t.c[i][v] = s[conn[i][v]];
        }
    }
    return t;
}

inline PackedCorners pack_corners(const CornerTable& t) {
    PackedCorners p{0ull, 0ull};
    for (int i = 0; i < 6; ++i)
        for (int v = 0; v < 4; ++v) {
            const unsigned long long bits = (unsigned long long)(t.c[i][v] & 7u) << (3 * v);
            if (i < 4) p.lo |= bits << (16 * i);
            else p.hi |= bits << (16 * (i - 4));
        }
    return p;
}

template <class R>
struct V3 {
    R x, y, z;
};

template <class R>
ADECO_HD V3<R> mk(R x, R y, R z) {
    V3<R> v;
    v.x = x; v.y = y; v.z = z;
    return v;
}

template <class R>
ADECO_HD R comp(const V3<R>& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

template <class R>
ADECO_HD void set_comp(V3<R>& v, int a, R val) {
    if (a == 0) v.x = val;
    else if (a == 1) v.y = val;
    else v.z = val;
}

template <class R>
ADECO_HD V3<R> sel(bool p, const V3<R>& a, const V3<R>& b) {
    return mk(p ? a.x : b.x, p ? a.y : b.y, p ? a.z : b.z);
}

// 6 x signed volume in the reference's sign convention: (a-d) . ((b-d) x (c-d))
// (AdTetIntersection.h:183-203)
template <class R>
ADECO_HD R vol6(const V3<R>& a, const V3<R>& b, const V3<R>& c, const V3<R>& d) {
    const R ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
    const R adz = a.z - d.z, bdz = b.z - d.z, cdz = c.z - d.z;
    return (a.x - d.x) * (bdy * cdz - bdz * cdy) + (b.x - d.x) * (cdy * adz - cdz * ady) +
           (c.x - d.x) * (ady * bdz - adz * bdy);
}

template <class R>
ADECO_HD R absr(R x) { return x < R(0) ? -x : x; }

// a / b for well-scaled positive denominators: the hardware reciprocal refined by two Newton steps
// and one correction of the quotient (within an ulp or two of the IEEE quotient, at about half the
// instructions of the IEEE division sequence).
ADECO_HD double qdiv(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    r = fma(fma(-b, r, 1.0), r, r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
#else
    return a / b;
#endif
}
ADECO_HD float qdiv(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float r = __builtin_amdgcn_rcpf(b);
    const float q = a * r;
    return fmaf(fmaf(-b, q, a), r, q);
#else
    return a / b;
#endif
}

// Fraction of a tet's volume on the far side (d > 0) of a plane, from the 4 vertex distances.
// With a = the beyond magnitudes, c = the near ones (|d|), and D = the product over the vertex
// pairs on opposite sides of (|d_i| + |d_j|) (= |d_i - d_j|, no cancellation):
//   1 beyond (a):        a^3 / D
//   2 beyond (a, b):     [ (ab)^2 + ab (a+b)(c+e) + ce ((a+b)^2 - ab) ] / D
//   3 beyond (near c):   1 - c^3 / D
// The products of the beyond and of the near magnitudes and their sums select a, b, c, e without
// sorting, so one division and no branches serve every case.
template <class R>
ADECO_HD R beyond_fraction(R d0, R d1, R d2, R d3) {
    const bool b0 = d0 > R(0), b1 = d1 > R(0), b2 = d2 > R(0), b3 = d3 > R(0);
    const int n = int(b0) + int(b1) + int(b2) + int(b3);
    const R m0 = absr(d0), m1 = absr(d1), m2 = absr(d2), m3 = absr(d3);
    const R one = R(1), zero = R(0);
    const R den = ((b0 != b1) ? m0 + m1 : one) * ((b0 != b2) ? m0 + m2 : one) * ((b0 != b3) ? m0 + m3 : one) *
                  ((b1 != b2) ? m1 + m2 : one) * ((b1 != b3) ? m1 + m3 : one) * ((b2 != b3) ? m2 + m3 : one);
    const R p = (b0 ? m0 : one) * (b1 ? m1 : one) * (b2 ? m2 : one) * (b3 ? m3 : one);
    const R q = (b0 ? one : m0) * (b1 ? one : m1) * (b2 ? one : m2) * (b3 ? one : m3);
    const R sb = (b0 ? m0 : zero) + (b1 ? m1 : zero) + (b2 ? m2 : zero) + (b3 ? m3 : zero);
    const R sn = (b0 ? zero : m0) + (b1 ? zero : m1) + (b2 ? zero : m2) + (b3 ? zero : m3);
    const R num2 = p * p + p * sb * sn + q * (sb * sb - p);
    const R e = n == 1 ? p : q;
    const R num = n == 2 ? num2 : e * e * e;
    const R f = (n == 0 || n == 4) ? zero : qdiv(num, den);
    return n == 4 ? one : (n == 3 ? one - f : f);
}

// Both sides of a plane from one closed form: *above = fraction with d > 0, *below = the rest.  The
// side with fewer vertices is the one computed directly; the other is 1 minus it.
template <class R>
ADECO_HD void beyond_split(R d0, R d1, R d2, R d3, R* above, R* below) {
    const bool b0 = d0 > R(0), b1 = d1 > R(0), b2 = d2 > R(0), b3 = d3 > R(0);
    const int n = int(b0) + int(b1) + int(b2) + int(b3);
    const R m0 = absr(d0), m1 = absr(d1), m2 = absr(d2), m3 = absr(d3);
    const R one = R(1), zero = R(0);
    const R den = ((b0 != b1) ? m0 + m1 : one) * ((b0 != b2) ? m0 + m2 : one) * ((b0 != b3) ? m0 + m3 : one) *
                  ((b1 != b2) ? m1 + m2 : one) * ((b1 != b3) ? m1 + m3 : one) * ((b2 != b3) ? m2 + m3 : one);
    const R p = (b0 ? m0 : one) * (b1 ? m1 : one) * (b2 ? m2 : one) * (b3 ? m3 : one);
    const R q = (b0 ? one : m0) * (b1 ? one : m1) * (b2 ? one : m2) * (b3 ? one : m3);
    const R sb = (b0 ? m0 : zero) + (b1 ? m1 : zero) + (b2 ? m2 : zero) + (b3 ? m3 : zero);
    const R sn = (b0 ? zero : m0) + (b1 ? zero : m1) + (b2 ? zero : m2) + (b3 ? zero : m3);
    const R num2 = p * p + p * sb * sn + q * (sb * sb - p);
    const R e = n == 1 ? p : q;
    const R num = n == 2 ? num2 : e * e * e;
    const R f = (n == 0 || n == 4) ? zero : qdiv(num, den);
    // f is the d > 0 side for n = 1, 2 and the d <= 0 side for n = 3
    *above = n == 4 ? one : (n == 3 ? one - f : f);
    *below = n == 0 ? one : (n == 3 ? f : (n == 4 ? zero : one - f));
}

// |vol6| of the part of tet (a,b,c,d) inside the slab 0 <= z <= 1.
template <class R>
ADECO_HD R zslab_vol6(const V3<R>& a, const V3<R>& b, const V3<R>& c, const V3<R>& d) {
    const R w = absr(vol6(a, b, c, d));
    const R zmin = fmin(fmin(a.z, b.z), fmin(c.z, d.z));
    const R zmax = fmax(fmax(a.z, b.z), fmax(c.z, d.z));
    R f = R(1);
    if (zmin < R(0)) f -= beyond_fraction(-a.z, -b.z, -c.z, -d.z);
    if (zmax > R(1)) f -= beyond_fraction(a.z - R(1), b.z - R(1), c.z - R(1), d.z - R(1));
    return f > R(0) ? w * f : R(0);
}

// One clip stage: the part of tet t on the inside of plane (axis, value c, keep >= c when ge)
// written as a point sequence seq[0..5] with `count` sub-tets seq[k..k+3], k < count.
template <class R, int NS = 7>
struct Stage {
    V3<R> s[NS];
    int count;     // pieces: windows 0..count-1, except that piece 1 is window 3 when corner is set
    bool corner;   // 3 vertices inside: piece 0 = the input tet (+), piece 1 = the corner beyond (-)
};

template <class R>
ADECO_HD V3<R> cut_point(const V3<R>& in, R din, const V3<R>& out, R dout, int axis, R c) {
    const R t = qdiv(din, din - dout);  // din > 0 >= dout: t in (0, 1]
    V3<R> p = mk(in.x + t * (out.x - in.x), in.y + t * (out.y - in.y), in.z + t * (out.z - in.z));
    set_comp(p, axis, c);
    return p;
}

template <class R, int NS = 7>
ADECO_HD void clip_stage(const V3<R> t[4], int axis, R c, bool ge, Stage<R, NS>& st, bool use_signed = true) {
    R d[4];
    for (int i = 0; i < 4; ++i) {
        const R x = comp(t[i], axis);
        d[i] = ge ? (x - c) : (c - x);
    }
    const int m = int(d[0] > R(0)) | (int(d[1] > R(0)) << 1) | (int(d[2] > R(0)) << 2) | (int(d[3] > R(0)) << 3);
    const int n = (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1) + ((m >> 3) & 1);
    // every path writes all six slots, so the sequence stays in registers
    st.corner = false;
    if (n == 0 || n == 4) {
        st.s[0] = t[0]; st.s[1] = t[1]; st.s[2] = t[2]; st.s[3] = t[3]; st.s[4] = t[3]; st.s[5] = t[3];
        if (NS == 7) st.s[NS - 1] = t[3];
        st.count = n == 4 ? 1 : 0;
        return;
    }
    // canonical order, inside vertices first: a compare-and-swap network on the inside flag keeps
    // every access statically indexed (registers, no private-memory arrays)
    V3<R> p[4] = {t[0], t[1], t[2], t[3]};
    R dd[4] = {d[0], d[1], d[2], d[3]};
#define ADECO_CSWAP(a, b)                                        \
    {                                                            \
        const bool sw = !(dd[a] > R(0)) && (dd[b] > R(0));       \
        const V3<R> pa = p[a], pb = p[b];                        \
        const R da = dd[a], db = dd[b];                          \
        p[a] = sel(sw, pb, pa); p[b] = sel(sw, pa, pb);          \
        dd[a] = sw ? db : da; dd[b] = sw ? da : db;              \
    }
    ADECO_CSWAP(0, 1) ADECO_CSWAP(2, 3) ADECO_CSWAP(0, 2) ADECO_CSWAP(1, 3) ADECO_CSWAP(1, 2)
#undef ADECO_CSWAP
    if (n == 1) {
        st.s[0] = p[0];
        st.s[1] = cut_point(p[0], dd[0], p[1], dd[1], axis, c);
        st.s[2] = cut_point(p[0], dd[0], p[2], dd[2], axis, c);
        st.s[3] = cut_point(p[0], dd[0], p[3], dd[3], axis, c);
        st.s[4] = st.s[3];
        st.s[5] = st.s[3];
        if (NS == 7) st.s[NS - 1] = st.s[3];
        st.count = 1;
    } else if (n == 3 && NS == 7 && use_signed) {  // the tet (a,b,c,d) minus the corner (d, p_ad, p_bd, p_cd)
        st.s[0] = p[0]; st.s[1] = p[1]; st.s[2] = p[2]; st.s[3] = p[3];
        st.s[4] = cut_point(p[0], dd[0], p[3], dd[3], axis, c);
        st.s[5] = cut_point(p[1], dd[1], p[3], dd[3], axis, c);
        st.s[NS - 1] = cut_point(p[2], dd[2], p[3], dd[3], axis, c);
        st.count = 2;
        st.corner = true;
    } else if (n == 3) {  // prism (a,b,c) - (p_ad, p_bd, p_cd): 3 tets
        st.s[0] = p[0]; st.s[1] = p[1]; st.s[2] = p[2];
        st.s[3] = cut_point(p[0], dd[0], p[3], dd[3], axis, c);
        st.s[4] = cut_point(p[1], dd[1], p[3], dd[3], axis, c);
        st.s[5] = cut_point(p[2], dd[2], p[3], dd[3], axis, c);
        if (NS == 7) st.s[NS - 1] = st.s[5];
        st.count = 3;
    } else {  // wedge (a, p_ac, p_ad) - (b, p_bc, p_bd)
        st.s[0] = p[0];
        st.s[1] = cut_point(p[0], dd[0], p[2], dd[2], axis, c);
        st.s[2] = cut_point(p[0], dd[0], p[3], dd[3], axis, c);
        st.s[3] = p[1];
        st.s[4] = cut_point(p[1], dd[1], p[2], dd[2], axis, c);
        st.s[5] = cut_point(p[1], dd[1], p[3], dd[3], axis, c);
        if (NS == 7) st.s[NS - 1] = st.s[5];
        st.count = 3;
    }
}

// sub-tet k (0..2) = seq[k..k+3]; written out slot by slot so every index is a constant and the
// sequence never leaves registers
template <class R>
ADECO_HD R take(const Stage<R, 7>& st, int k, V3<R> out[4]) {
    const bool k0 = k == 0, k1 = k == 1 && !st.corner, k3 = k == 1 && st.corner;
    out[0] = sel(k0, st.s[0], sel(k1, st.s[1], sel(k3, st.s[3], st.s[2])));
    out[1] = sel(k0, st.s[1], sel(k1, st.s[2], sel(k3, st.s[4], st.s[3])));
    out[2] = sel(k0, st.s[2], sel(k1, st.s[3], sel(k3, st.s[5], st.s[4])));
    out[3] = sel(k0, st.s[3], sel(k1, st.s[4], sel(k3, st.s[6], st.s[5])));
    return k3 ? R(-1) : R(1);
}

template <class R>
ADECO_HD R take(const Stage<R, 6>& st, int k, V3<R> out[4]) {
    const bool k0 = k == 0, k1 = k == 1;
    out[0] = sel(k0, st.s[0], sel(k1, st.s[1], st.s[2]));
    out[1] = sel(k0, st.s[1], sel(k1, st.s[2], st.s[3]));
    out[2] = sel(k0, st.s[2], sel(k1, st.s[3], st.s[4]));
    out[3] = sel(k0, st.s[3], sel(k1, st.s[4], st.s[5]));
    return R(1);
}

// Moves the next piece of a stage into slots 0..3: pieces of a sequence are consecutive windows,
// so the next one is the sequence shifted by one, except after the first piece of a corner stage,
// whose second piece is window 3.  The loops over a stage's pieces read slots 0..3 only.
template <class R>
ADECO_HD void advance(Stage<R, 7>& st) {
    const bool c = st.corner;
    st.s[0] = sel(c, st.s[3], st.s[1]);
    st.s[1] = sel(c, st.s[4], st.s[2]);
    st.s[2] = sel(c, st.s[5], st.s[3]);
    st.s[3] = sel(c, st.s[6], st.s[4]);
    st.s[4] = st.s[5];
}

template <class R>
ADECO_HD void advance(Stage<R, 6>& st) {
    st.s[0] = st.s[1]; st.s[1] = st.s[2]; st.s[2] = st.s[3]; st.s[3] = st.s[4]; st.s[4] = st.s[5];
}

// the sign of piece k of a stage whose pieces are read from slots 0..3 (advance)
template <class R>
ADECO_HD R piece_sign(const Stage<R, 7>& st, int k) { return (k == 1 && st.corner) ? R(-1) : R(1); }
template <class R>
ADECO_HD R piece_sign(const Stage<R, 6>&, int) { return R(1); }

// Which of the faces x>=0, x<=1, y>=0, y<=1 cut the tet (bit f set), or -1 if the tet lies
// entirely outside one of them.
template <class R>
ADECO_HD int active_xy_faces(const V3<R> t[4]) {
    int act = 0;
    for (int f = 0; f < 4; ++f) {
        const int axis = f >> 1;
        const bool ge = (f & 1) == 0;
        const R c = ge ? R(0) : R(1);
        int nin = 0;
        for (int i = 0; i < 4; ++i) {
            const R x = comp(t[i], axis);
            nin += int(ge ? (x - c) > R(0) : (c - x) > R(0));
        }
        if (nin == 0) return -1;
        if (nin < 4) act |= 1 << f;
    }
    return act;
}

struct ClipCounters {
    uint32_t stages, subtets;
};

// Hands the pieces of one stage, with their signs, to the leaf.
// (unrolled over the at most 3 pieces, so that take() reads fixed slots)
template <class R, int NS, class Leaf>
ADECO_HD void stage_leaves(const Stage<R, NS>& st, R g, Leaf& leaf, ClipCounters* cnt) {
    if (cnt) cnt->subtets += uint32_t(st.count);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k < st.count) {
            V3<R> q[4];
            const R gk = g * take(st, k, q);
            leaf(q, gk);
        }
    }
}

// The clip tree of t over its cutting x/y faces; every leaf piece q (sign g) goes to leaf(q, g),
// which integrates its z extent in closed form.  use_signed: a cut that leaves 3 vertices inside
// gives the whole tet (+) and the corner beyond the plane (-) instead of 3 prism tets.  One
// instance serves both trees, so a kernel holds the clip code once.
// t is in the canonical frame of clip_pieces: the cutting faces are, by level, x >= 0, then x <= 1
// (two x faces) or y >= 0 (one x face: `xy`), then y >= 0, then y <= 1, so only the second level's
// face is chosen at run time.
template <class R, int NS, class Leaf>
ADECO_HD void clip_tree(const V3<R> t[4], int nact, bool xy, bool use_signed, Leaf& leaf, ClipCounters* cnt) {
    Stage<R, NS> s1, s2, s3, s4;
    V3<R> q1[4], q2[4], q3[4];
    clip_stage(t, 0, R(0), true, s1, use_signed);
    if (cnt) cnt->stages += 1;
    if (nact == 1) {
        stage_leaves(s1, R(1), leaf, cnt);
        return;
    }
    const int ax1 = xy ? 1 : 0;
    const R c1 = xy ? R(0) : R(1);
    for (int k1 = 0; k1 < s1.count; ++k1) {
        const R g1 = piece_sign(s1, k1);
        for (int q = 0; q < 4; ++q) q1[q] = s1.s[q];
        advance(s1);
        clip_stage(q1, ax1, c1, xy, s2, use_signed);
        if (cnt) cnt->stages += 1;
        if (nact == 2) {
            stage_leaves(s2, g1, leaf, cnt);
            continue;
        }
        for (int k2 = 0; k2 < s2.count; ++k2) {
            const R g2 = g1 * piece_sign(s2, k2);
            for (int q = 0; q < 4; ++q) q2[q] = s2.s[q];
            advance(s2);
            clip_stage(q2, 1, R(0), true, s3, use_signed);
            if (cnt) cnt->stages += 1;
            if (nact == 3) {
                stage_leaves(s3, g2, leaf, cnt);
                continue;
            }
            for (int k3 = 0; k3 < s3.count; ++k3) {
                const R g3 = g2 * piece_sign(s3, k3);
                for (int q = 0; q < 4; ++q) q3[q] = s3.s[q];
                advance(s3);
                clip_stage(q3, 1, R(1), false, s4, use_signed);
                if (cnt) cnt->stages += 1;
                stage_leaves(s4, g3, leaf, cnt);
            }
        }
    }
}

// The pieces of t inside the x/y square [0,1]^2 (any z) to the leaf: t itself when no x/y face
// cuts it, none when it lies outside one of them.  NS = 6 compiles the exact tree alone
// (use_signed is then ignored).  The tree runs in a canonical frame: x and y swapped when y has
// more cutting faces, and an axis with only its face at 1 cutting mirrored (u -> 1 - u), so that
// its cutting face becomes the one at 0; volumes are unchanged.
template <class R, int NS = 7, class Leaf>
ADECO_HD void clip_pieces(const V3<R> t[4], bool use_signed, Leaf& leaf, ClipCounters* cnt = nullptr) {
    const int act = active_xy_faces(t);
    if (act < 0) return;
    const int nx = (act & 1) + ((act >> 1) & 1), ny = ((act >> 2) & 1) + ((act >> 3) & 1);
    const int nact = nx + ny;
    if (nact == 0) {
        if (cnt) cnt->subtets += 1;
        leaf(t, R(1));
        return;
    }
    const bool swap = ny > nx;
    const int ax_ = swap ? (act >> 2) : act, ay_ = swap ? act : (act >> 2);  // bit 0: face at 0, bit 1: face at 1
    const bool mx = (ax_ & 3) == 2, my = (ay_ & 3) == 2;
    V3<R> c[4];
    for (int q = 0; q < 4; ++q) {
        const R u = swap ? t[q].y : t[q].x, v = swap ? t[q].x : t[q].y;
        c[q] = mk(mx ? R(1) - u : u, my ? R(1) - v : v, t[q].z);
    }
    clip_tree<R, NS>(c, nact, nx == 1 && ny == 1, use_signed, leaf, cnt);
}

// Leaf of the unit cell: signed sum of the pieces' z-slab volumes and the sum of their magnitudes.
template <class R>
struct ZslabSum {
    R vol, sabs;
    ADECO_HD void operator()(const V3<R> q[4], R g) {
        const R w = zslab_vol6(q[0], q[1], q[2], q[3]);
        vol += g * w;
        sabs += w;
    }
};

// Leaf of a column segment of ls unit cells along z ([c, c+1], c < ls; ls = 1 is the unit cell):
// piece q (sign g, |vol6| w) adds g w (1 - below(c) - above(c+1)) to acc[c * kStride] and
// w (1 - below(c) - above(c+1)) to sab[c * kStride] (of type S: the magnitudes only set the
// cancellation-noise threshold, so a kernel may keep them in FP32).  The fractions depend on the 4 heights only:
// sorted once (z0 <= z1 <= z2 <= z3), a plane h in (z0, z3) has 1, 2 or 3 heights below it, and
// the side with fewer vertices is computed directly as in beyond_fraction (D is fixed per interval):
//   (z0, z1]: below = a^3 / D1                 a = h - z0,  D1 = (z1-z0)(z2-z0)(z3-z0)
//   (z1, z2]: below = [(ab)^2 + ab(a+b)(c+e) + ce((a+b)^2 - ab)] / D2
//             a, b = h - z0, h - z1;  c, e = z2 - h, z3 - h;  D2 = (z2-z0)(z3-z0)(z2-z1)(z3-z1)
//   (z2, z3): above = e^3 / D3                 D3 = (z3-z0)(z3-z1)(z3-z2)
// Each plane between two cells is evaluated once for both.
template <class R, int kStride, class S = double>
struct ColumnLeaf {
    double* acc;
    S* sab;
    int ls;
    ADECO_HD void operator()(const V3<R> q[4], R g) {
        const R w = absr(vol6(q[0], q[1], q[2], q[3]));
        R z0 = q[0].z, z1 = q[1].z, z2 = q[2].z, z3 = q[3].z;
#define ADECO_MINMAX(u, v)      \
    {                           \
        const R lo_ = fmin(u, v); \
        v = fmax(u, v);         \
        u = lo_;                \
    }
        ADECO_MINMAX(z0, z1) ADECO_MINMAX(z2, z3) ADECO_MINMAX(z0, z2) ADECO_MINMAX(z1, z3) ADECO_MINMAX(z1, z2)
#undef ADECO_MINMAX
        const R top = R(ls);
        // cells [c0, c1] meet the piece's z range; comparisons with NaN are false, which keeps both
        // in [0, ls)
        const int c0 = z0 > R(0) ? int(fmin(floor(z0), top - R(1))) : 0;
        const int c1 = z3 < top ? int(fmax(ceil(z3) - R(1), R(0))) : ls - 1;
        const R d20 = z2 - z0, d30 = z3 - z0, d31 = z3 - z1;
        const R D1 = (z1 - z0) * d20 * d30, D2 = d20 * d30 * (z2 - z1) * d31, D3 = d30 * d31 * (z3 - z2);
        R below_prev = R(0);
        for (int h = c0; h <= c1 + 1; ++h) {
            const R zh = R(h);
            R above = R(0), below = R(1);
            if (z0 >= zh) {
                above = R(1);
                below = R(0);
            } else if (z3 > zh) {
                const R a = zh - z0, b = zh - z1, c = z2 - zh, e = z3 - zh;
                const bool one = zh <= z1, three = zh > z2;
                const R ab = a * b, apb = a + b;
                const R num2 = ab * ab + ab * apb * (c + e) + c * e * (apb * apb - ab);
                const R num = one ? a * a * a : (three ? e * e * e : num2);
                const R f = qdiv(num, one ? D1 : (three ? D3 : D2));
                below = three ? R(1) - f : f;
                above = three ? f : R(1) - f;
            }
            if (h > c0) {
                const R m = R(1) - below_prev - above;
                if (m > R(0)) {
                    acc[(h - 1) * kStride] += double(g * w * m);
                    sab[(h - 1) * kStride] += S(w * m);
                }
            }
            below_prev = below;
        }
    }
};

// 6 x |t ∩ [0,1]^3| by the signed (use_signed) or the exact tree; *sabs = the sum of the magnitudes
// of the terms it was summed from.
template <class R, int NS = 7>
ADECO_HD R clip_any_vol6(const V3<R> t[4], bool use_signed, R* sabs, ClipCounters* cnt = nullptr) {
    ZslabSum<R> acc{R(0), R(0)};
    clip_pieces<R, NS>(t, use_signed, acc, cnt);
    *sabs = acc.sabs;
    return acc.vol;
}

// Signed tree of a cell-local tet with its noise measure: returns 6 x |t ∩ [0,1]^3| and sets *sabs
// to the sum of the magnitudes of the terms it was summed from.
template <class R>
ADECO_HD R clip_signed_vol6(const V3<R> t[4], R* sabs, ClipCounters* cnt = nullptr) {
    return clip_any_vol6(t, true, sabs, cnt);
}

// Tree of non-overlapping pieces: every term is positive, no cancellation.
template <class R>
ADECO_HD R clip_exact_vol6(const V3<R> t[4], ClipCounters* cnt = nullptr) {
    R sabs;
    return clip_any_vol6<R, 6>(t, false, &sabs, cnt);
}

// A signed result below 1e-6 of the magnitudes it was summed from may be cancellation noise (a few
// eps of those magnitudes); such a cell is recomputed with the exact tree.
template <class R>
ADECO_HD bool signed_is_noisy(R v, R sabs) { return !(v >= R(1e-6) * sabs); }

// 6 x |t ∩ [0,1]^3| (t in cell units, any orientation): the signed tree (fewer pieces), recomputed
// with the exact tree where it may be noise.
// The loop keeps one instance of the tree in the caller.
template <class R>
ADECO_HD R clip_unit_cube_vol6(const V3<R> t[4], ClipCounters* cnt = nullptr) {
    bool use_signed = true;
    for (;;) {
        R sabs;
        const R v = clip_any_vol6(t, use_signed, &sabs, cnt);
        if (!use_signed || !signed_is_noisy(v, sabs)) return v;
        use_signed = false;
    }
}

// ---------------------------------------------------------------------------------------------
// Classification of a scaled cell-local tet against the unit cell, as the reference's
// intersection_test (AdDepositTets2Grid.h:286-363): per tet face, the cube corners strictly on
// the inner side; no corner inside one face -> no intersection; all corners inside every face ->
// the cell lies in the tet; all tet vertices in [0,1]^3 -> the tet lies in the cell.
enum CellClass : int { CLS_NONE = 0, CLS_TET_IN_CELL = 1, CLS_CELL_IN_TET = 2, CLS_PARTIAL = 3 };

template <class R>
ADECO_HD V3<R> cross(const V3<R>& a, const V3<R>& b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

template <class R>
ADECO_HD V3<R> sub(const V3<R>& a, const V3<R>& b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }

// Conservative classification of the positively oriented cell-local tet s against the box [0,r]
// (FP64): per tet face, the box corners' plane values at their extremes (from the normal's
// signs), with a relative margin of 1e-12 so that an uncertain box is PARTIAL, never NONE or
// CELL_IN_TET.
ADECO_HD int classify_margin(const V3<double> s[4], const double* r) {
    const V3<double> e01 = sub(s[1], s[0]), e02 = sub(s[2], s[0]), e03 = sub(s[3], s[0]);
    const V3<double> e13 = sub(s[3], s[1]), e12 = sub(s[2], s[1]);
    const V3<double> n[4] = {cross(e01, e02), cross(e02, e03), cross(e03, e01), cross(e13, e12)};
    bool all_in = true;
    for (int f = 0; f < 4; ++f) {
        const V3<double> o = f == 3 ? s[1] : s[0];
        const double v0 = o.x * n[f].x + o.y * n[f].y + o.z * n[f].z;
        // plane value (o - p).n at the cube corners p in {0, r}: extremes from the normal's signs
        const double px = n[f].x * r[0], py = n[f].y * r[1], pz = n[f].z * r[2];
        const double pos = fmax(px, 0.0) + fmax(py, 0.0) + fmax(pz, 0.0);
        const double neg = fmin(px, 0.0) + fmin(py, 0.0) + fmin(pz, 0.0);
        const double scl = (fabs(n[f].x) + fabs(n[f].y) + fabs(n[f].z)) *
                           (1.0 + fmax(fmax(fabs(o.x), fabs(o.y)), fabs(o.z)));
        const double tol = 1e-12 * scl;
        const double vmax = v0 - neg;
        const double vmin = v0 - pos;
        if (vmax < -tol) return CLS_NONE;
        all_in = all_in && (vmin > tol);
    }
    if (all_in) return CLS_CELL_IN_TET;
    bool tin = true;
    for (int q = 0; q < 4; ++q)
        tin = tin && s[q].x >= 0.0 && s[q].x <= r[0] && s[q].y >= 0.0 && s[q].y <= r[1] && s[q].z >= 0.0 &&
              s[q].z <= r[2];
    return tin ? CLS_TET_IN_CELL : CLS_PARTIAL;
}

// The reference's corner test against the unit cell (described above the CellClass values).
template <class R>
ADECO_HD int classify_unit(const V3<R> s[4]) {
    const V3<R> e01 = sub(s[1], s[0]), e02 = sub(s[2], s[0]), e03 = sub(s[3], s[0]);
    const V3<R> e13 = sub(s[3], s[1]), e12 = sub(s[2], s[1]);
    const V3<R> n[4] = {cross(e01, e02), cross(e02, e03), cross(e03, e01), cross(e13, e12)};
    bool all_in = true;
    for (int f = 0; f < 4; ++f) {
        const V3<R> o = f == 3 ? s[1] : s[0];
        int inside = 0;
        for (int p = 0; p < 8; ++p) {
            const R px = R((p >> 2) & 1), py = R((p >> 1) & 1), pz = R(p & 1);
            const R v = (o.x - px) * n[f].x + (o.y - py) * n[f].y + (o.z - pz) * n[f].z;
            inside += int(v > R(0));
        }
        if (inside == 0) return CLS_NONE;
        all_in = all_in && inside == 8;
    }
    if (all_in) return CLS_CELL_IN_TET;
    bool tin = true;
    for (int i = 0; i < 4; ++i)
        tin = tin && s[i].x >= R(0) && s[i].x <= R(1) && s[i].y >= R(0) && s[i].y <= R(1) &&
              s[i].z >= R(0) && s[i].z <= R(1);
    return tin ? CLS_TET_IN_CELL : CLS_PARTIAL;
}

}  // namespace adeco
