// Per-tet exact deposit onto one uniform grid, FP64 throughout (host + device).
//
// Follows deposit_tet_2_grid (AdDepositTets2Grid.h:509-665): tet AABB and its strict overlap with
// the grid, the covered cell range clamped to the grid (the cells its strict test can pass), the signed volume (zero ->
// skip; negative -> vertex order (v3, v1, v2, v0)), then per covered cell the strict AABB cull,
// the scaled tet s = scale_a * (src - cell_min) per axis, the classification and the mass
//   tet in cell  -> m
//   cell in tet  -> m * cell_vol / |V|
//   partial      -> m * (|s ∩ [0,1]^3| * cell_vol) / |V|      (when > 0)
// accumulated with `emit(cell_index, dmass)`.
#pragma once

#include "adeco_geom.hpp"
#include "kernels.hpp"

namespace adeco {

struct TetCounters {
    unsigned long long v[kNumStats];
};

template <class Emit>
ADECO_HD void deposit_tet_f64(const V3<double> v[4], const GridGeom& g, double tet_mass, Emit& emit,
                              TetCounters* cnt) {
    if (cnt) cnt->v[0]++;
    double tmn[3] = {v[0].x, v[0].y, v[0].z}, tmx[3] = {v[0].x, v[0].y, v[0].z};
    for (int i = 1; i < 4; ++i) {
        const double p[3] = {v[i].x, v[i].y, v[i].z};
        for (int a = 0; a < 3; ++a) {
            tmn[a] = p[a] < tmn[a] ? p[a] : tmn[a];
            tmx[a] = p[a] > tmx[a] ? p[a] : tmx[a];
        }
    }
    for (int a = 0; a < 3; ++a)
        if (!(tmn[a] < g.gmax[a] && g.gmin[a] < tmx[a])) return;
    int lo[3], hi[3];
    const int dm1 = g.dims - 1;
    for (int a = 0; a < 3; ++a) {
        const double omn = tmn[a] < g.gmin[a] ? g.gmin[a] : tmn[a];
        const double omx = tmx[a] < g.gmax[a] ? tmx[a] : g.gmax[a];
        // [floor, floor] widened by 1e-9 of a cell: the cells the strict overlap test can pass
        // (the reference's ceil adds one that never does)
        int l = int(floor((omn - g.gmin[a]) / g.delta[a] - 1e-9));
        int h = int(floor((omx - g.gmin[a]) / g.delta[a] + 1e-9));
        lo[a] = l <= 0 ? 0 : (l >= dm1 ? dm1 : l);
        hi[a] = h <= 0 ? 0 : (h >= dm1 ? dm1 : h);
    }
    const double vol = vol6(v[0], v[1], v[2], v[3]) / 6.0;
    if (vol == 0.0) return;
    V3<double> src[4];
    if (vol < 0.0) {
        src[0] = v[3]; src[1] = v[1]; src[2] = v[2]; src[3] = v[0];
    } else {
        src[0] = v[0]; src[1] = v[1]; src[2] = v[2]; src[3] = v[3];
    }
    if (cnt) cnt->v[1]++;
    const double inv_abs = 1.0 / (vol < 0.0 ? -vol : vol);
    const double m_cell = tet_mass * (g.cell_vol * inv_abs);
    for (int k = lo[2]; k <= hi[2]; ++k) {
        const double cz = g.gmin[2] + double(k) * g.delta[2];
        const double czx = cz + g.delta[2];
        for (int j = lo[1]; j <= hi[1]; ++j) {
            const double cy = g.gmin[1] + double(j) * g.delta[1];
            const double cyx = cy + g.delta[1];
            for (int i = lo[0]; i <= hi[0]; ++i) {
                const double cx = g.gmin[0] + double(i) * g.delta[0];
                const double cxx = cx + g.delta[0];
                if (cnt) cnt->v[2]++;
                if (!(tmn[0] < cxx && cx < tmx[0] && tmn[1] < cyx && cy < tmx[1] && tmn[2] < czx && cz < tmx[2]))
                    continue;
                if (cnt) cnt->v[3]++;
                V3<double> s[4];
                for (int q = 0; q < 4; ++q)
                    s[q] = mk(g.scale[0] * (src[q].x - cx), g.scale[1] * (src[q].y - cy), g.scale[2] * (src[q].z - cz));
                const int cls = classify_unit(s);
                double dm = 0.0;
                if (cls == CLS_NONE) {
                    if (cnt) cnt->v[4]++;
                    continue;
                } else if (cls == CLS_TET_IN_CELL) {
                    if (cnt) cnt->v[5]++;
                    dm = tet_mass;
                } else if (cls == CLS_CELL_IN_TET) {
                    if (cnt) cnt->v[6]++;
                    dm = m_cell;
                } else {
                    if (cnt) cnt->v[7]++;
                    ClipCounters cc = {0, 0};
                    const double w6 = clip_unit_cube_vol6(s, cnt ? &cc : nullptr);
                    if (cnt) { cnt->v[9] += cc.stages; cnt->v[10] += cc.subtets; }
                    dm = m_cell * (w6 * (1.0 / 6.0));
                }
                if (dm > 0.0) {
                    if (cnt) cnt->v[8]++;
                    emit(int64_t(i) + int64_t(g.dims) * (int64_t(j) + int64_t(g.dims) * int64_t(k)), dm);
                }
            }
        }
    }
}

ADECO_HD bool is_sentinel(float x, float y, float z) {  // AdDepositTets2Grid.h:1029-1038
    const float s = fabsf(x - 666.f) + fabsf(y - 666.f) + fabsf(z - 666.f);
    return double(s) < 1.E-03;
}

// The 4 vertices of lattice tet t (cube t/6, i fastest over (np-1)^3 cubes; local tet t%6).
ADECO_HD void lattice_tet_vertices(const LatticeView& lat, const PackedCorners& ct, int64_t t, V3<double> v[4]) {
    if (lat.soup) {
        const float* p = lat.soup + 12 * t;
        for (int q = 0; q < 4; ++q) v[q] = mk(double(p[3 * q]), double(p[3 * q + 1]), double(p[3 * q + 2]));
        return;
    }
    const int64_t ncx = lat.np[0] - 1, ncy = lat.np[1] - 1;
    int64_t c, i, j, k;
    int lt;
    if (t < (int64_t(1) << 32)) {
        // 32-bit index arithmetic (a 64-bit division is several times longer)
        const unsigned tt = unsigned(t), cc = tt / 6u;
        lt = int(tt - 6u * cc);
        const unsigned cj = cc / unsigned(ncx);
        const unsigned kk = cj / unsigned(ncy);
        c = cc;
        i = int64_t(cc - cj * unsigned(ncx));
        k = kk;
        j = int64_t(cj - kk * unsigned(ncy));
    } else {
        c = t / 6;
        lt = int(t - 6 * c);
        const int64_t cj = c / ncx;
        i = c - cj * ncx;
        k = cj / ncy;
        j = cj - k * ncy;
    }
    (void)c;
    for (int q = 0; q < 4; ++q) {
        const int code = ct.code(lt, q);
        const int64_t vi = (i + (code & 1)) +
                           int64_t(lat.np[0]) * ((j + ((code >> 1) & 1)) + int64_t(lat.np[1]) * (k + ((code >> 2) & 1)));
        v[q] = mk(double(lat.x[vi]), double(lat.y[vi]), double(lat.z[vi]));
    }
}

}  // namespace adeco
