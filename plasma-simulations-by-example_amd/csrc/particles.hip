// Species: particle storage in tile bins, fused push + gather + deposit + rebin, density
// finish, counting sort, diagnostics, device injection.
//
// Layout.  Particles are structure-of-arrays, grouped in per-tile bins (tile = TILE^3 cells,
// tiles numbered x-fastest).  Every push reads bins of buffer `cur` and appends each surviving
// particle to the bin of its new tile in the other buffer, so bins stay exact every step,
// dead particles vanish (the compaction of ch3/ver2/Species.cpp:36-46), and each workgroup
// sees only particles inside its own tile: its E tile lives in LDS and its deposits land in an
// LDS window of the tile plus a HALO-cell rim, written out once per tile (no global atomics
// except for particles that jump further than the rim).
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <mutex>

#include "picamd_internal.h"

namespace pic {

int compute_ef(pic_world *w, bool want32, bool want64);

constexpr int NSLOT_DIAG = 5;
constexpr int NDIAG_SLOTS = NSLOT_DIAG + 1;  // + particle count

// ---------------------------------------------------------------------------------------
// push parameters
// ---------------------------------------------------------------------------------------
struct PushParams {
    Grid g;
    int bc;
    int sphere_on;
    // PIC_MIXED (FP32, cell-relative units)
    float qm_dt;     // dt*q/m
    float dt_dh[3];  // dt/dh
    // frame velocity u: the record holds v - u; a step moves the offset by
    // fdh_hi + ((v - u) dt/dh + fdh_lo), u dt/dh split over two floats; a wall reverses
    // v, so the stored v - u becomes -(v - u) - 2u
    float fdh_hi[3], fdh_lo[3], fr2[3];
    int sph_c[3];    // sphere centre in cells: integer part
    float sph_f[3];  //                          fractional part
    float dhf[3];
    float r2f;
    // PIC_FP64 (physical units, the reference's arithmetic)
    double s64;  // dt*charge/mass, as ch2/Species.cpp:26
    double dt;
    double sph_c64[3];
    double sph_r2;
};

__device__ inline int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

// Groups the lanes of a wavefront by key (keys >= 0) with ballots only: for each lane, the
// lowest lane holding its key (its leader), its rank among those lanes and their number.
struct KeyGroup {
    int leader;
    uint32_t before, count;
};

__device__ inline KeyGroup wave_group(int key) {
    KeyGroup kg{-1, 0u, 0u};
    unsigned long long active = __ballot(key >= 0);
    while (active) {
        int leader = __ffsll((long long)active) - 1;
        int lk = __builtin_amdgcn_readlane(key, leader);
        unsigned long long m = __ballot(key == lk);
        if (key == lk) {
            kg.leader = leader;
            kg.before = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            kg.count = (uint32_t)__popcll(m);
        }
        active &= ~m;
    }
    return kg;
}

// Reserves counter[key] slots for the lanes of each distinct key: the leaders of all keys
// issue their atomics in one instruction (one round trip per wavefront, not one per key),
// then each lane reads its leader's base.  Returns the lane's slot (0 for key < 0).
__device__ inline uint32_t wave_group_base(const KeyGroup &kg, int key, uint32_t leader_base) {
    int src = kg.leader < 0 ? (int)__lane_id() : kg.leader;
    uint32_t base = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)leader_base);
    return key >= 0 ? base + kg.before : 0u;
}

__device__ inline uint32_t wave_reserve_global(int key, uint32_t *counter) {
    KeyGroup kg = wave_group(key);
    uint32_t base = 0;
    if (key >= 0 && kg.before == 0) base = atomicAdd(&counter[key], kg.count);
    return wave_group_base(kg, key, base);
}

__device__ inline unsigned long long wave_count_u32(int key, uint32_t *counter) {
    // histogram: one atomic per distinct key per wavefront
    unsigned long long active = __ballot(key >= 0);
    while (active) {
        int leader = __ffsll((long long)active) - 1;
        int lk = __builtin_amdgcn_readlane(key, leader);
        unsigned long long m = __ballot(key == lk);
        if ((int)__lane_id() == leader) atomicAdd(&counter[lk], (uint32_t)__popcll(m));
        active &= ~m;
    }
    return 0;
}

// ---------------------------------------------------------------------------------------
// particle state in registers
// ---------------------------------------------------------------------------------------
struct PMixed {
    int ci, cj, ck;
    float ox, oy, oz, vx, vy, vz, w;
    uint32_t id;
};

struct PF64 {
    double x, y, z, vx, vy, vz, w;
    uint32_t id;
};

// PIC_MIXED particles are 32-byte records {cell, ox, oy, oz | vx, vy, vz, w}: two 16-byte
// accesses per lane, so one wavefront instruction moves 1 KB of contiguous particles and a
// particle written to another tile's inflow segment is one contiguous 32-byte piece.  (Eight
// separate 4-byte arrays measured 13-20% slower in the push: twice the memory instructions
// per byte and eight partial-line writes per moved particle.)
__device__ inline void load_p(const PartView &v, long idx, PMixed &p) {
    const float4 *r = v.rec + 2 * idx;
    float4 a = r[0], b = r[1];
    int c = __float_as_int(a.x);
    p.ci = cell_i(c);
    p.cj = cell_j(c);
    p.ck = cell_k(c);
    p.ox = a.y;
    p.oy = a.z;
    p.oz = a.w;
    p.vx = b.x;
    p.vy = b.y;
    p.vz = b.z;
    p.w = b.w;
    p.id = v.id ? v.id[idx] : 0u;
}

__device__ inline void store_p(const PartView &v, long idx, const PMixed &p) {
    float4 *r = v.rec + 2 * idx;
    r[0] = make_float4(__int_as_float(pack_cell(p.ci, p.cj, p.ck)), p.ox, p.oy, p.oz);
    r[1] = make_float4(p.vx, p.vy, p.vz, p.w);
    if (v.id) v.id[idx] = p.id;
}

__device__ inline void load_p(const PartView &v, long idx, PF64 &p) {
    p.x = v.x[idx];
    p.y = v.y[idx];
    p.z = v.z[idx];
    p.vx = v.dvx[idx];
    p.vy = v.dvy[idx];
    p.vz = v.dvz[idx];
    p.w = v.dw[idx];
    p.id = v.id ? v.id[idx] : 0u;
}

__device__ inline void store_p(const PartView &v, long idx, const PF64 &p) {
    v.x[idx] = p.x;
    v.y[idx] = p.y;
    v.z[idx] = p.z;
    v.dvx[idx] = p.vx;
    v.dvy[idx] = p.vy;
    v.dvz[idx] = p.vz;
    v.dw[idx] = p.w;
    if (v.id) v.id[idx] = p.id;
}

// cell of an FP64 particle: i = (int)((x-x0)/dh), as World::XtoL + Field_::scatter
// (ch2/World.h:72-78, Field.h:167-171); clamped so that a position that rounds onto the
// upper wall never indexes past the last cell.
__device__ inline void cell_of(const Grid &g, const PF64 &p, int &i, int &j, int &k, double &di,
                               double &dj, double &dk) {
#pragma clang fp contract(off)
    double lx = (p.x - g.x0[0]) / g.dh[0];
    double ly = (p.y - g.x0[1]) / g.dh[1];
    double lz = (p.z - g.x0[2]) / g.dh[2];
    i = (int)lx;
    j = (int)ly;
    k = (int)lz;
    i = min(max(i, 0), g.nci - 1);
    j = min(max(j, 0), g.ncj - 1);
    k = min(max(k, 0), g.nck - 1);
    di = lx - i;
    dj = ly - j;
    dk = lz - k;
}

__device__ inline int tile_id(const Grid &g, int ci, int cj, int ck) {
    return (ci >> TILE_SHIFT) + g.ntx * ((cj >> TILE_SHIFT) + g.nty * (ck >> TILE_SHIFT));
}

__device__ inline int tile_of(const Grid &g, const PMixed &p) { return tile_id(g, p.ci, p.cj, p.ck); }

__device__ inline int tile_of(const Grid &g, const PF64 &p) {
    int i, j, k;
    double a, b, c;
    cell_of(g, p, i, j, k, a, b, c);
    return tile_id(g, i, j, k);
}

// ---- FP32 push in cell-relative coordinates -------------------------------------------
// Advances one axis: offset o in [0,1) of cell c moves by d cells.
__device__ inline void advance_axis(int &c, float &o, float d) {
    float n = o + d;
    float fl = floorf(n);
    n -= fl;
    int s = (int)fl;
    if (n >= 1.0f) {  // o+d-floor rounded up to 1
        n = 0.0f;
        s += 1;
    }
    c += s;
    o = n;
}

// Specular wall on one axis with nc cells (ch2/Species.cpp:32-36 in logical units):
// L < 0 -> -L, L >= nc -> 2nc - L, velocity reversed.  A particle that lands exactly on the
// upper wall (L == nc after reflection, where the reference scatters past the array) is
// clamped just inside it.
__device__ inline void reflect_axis(int &c, float &o, float &v, int nc, float u2) {
    if (c < 0) {
        c = -c - 1;
        o = 1.0f - o;
        if (o >= 1.0f) {
            o = 0.0f;
            c += 1;
        }
        v = -v - u2;
    } else if (c >= nc) {
        c = 2 * nc - c - 1;
        o = 1.0f - o;
        if (o >= 1.0f) {
            o = 0.0f;
            c += 1;
        }
        v = -v - u2;
    }
    if (c >= nc) {
        c = nc - 1;
        o = 0x1.fffffep-1f;
    }
    if (c < 0) {  // moved more than the box width in one step: keep memory-safe
        c = 0;
        o = 0.0f;
    }
}

// E at a particle from the LDS E tile: three FP32 planes (Ex, Ey, Ez) of TN^3 nodes,
// trilinear (CIC).  Planes instead of one float4 per node: 12 B/node of LDS, which lets
// eight push workgroups share a CU.
__device__ inline float3 gather_lds(const float *pl, int li, int lj, int lk, float ox, float oy, float oz) {
    int b = li + TN * (lj + TN * lk);
    float ax = 1.0f - ox, ay = 1.0f - oy, az = 1.0f - oz;
    float w000 = ax * ay * az, w100 = ox * ay * az, w010 = ax * oy * az, w110 = ox * oy * az;
    float w001 = ax * ay * oz, w101 = ox * ay * oz, w011 = ax * oy * oz, w111 = ox * oy * oz;
    float r[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float *e = pl + c * TNODES + b;
        r[c] = e[0] * w000 + e[1] * w100 + e[TN + 1] * w110 + e[TN] * w010 + e[TN * TN] * w001 +
               e[TN * TN + 1] * w101 + e[TN * TN + TN + 1] * w111 + e[TN * TN + TN] * w011;
    }
    return make_float3(r[0], r[1], r[2]);
}

__device__ inline float3 gather_global32(const Grid &g, const float4 *ef, int ci, int cj, int ck,
                                         float ox, float oy, float oz) {
    long sj = g.ni, sk = (long)g.ni * g.nj;
    long b = ci + sj * cj + sk * ck;
    float4 e000 = ef[b], e100 = ef[b + 1], e010 = ef[b + sj], e110 = ef[b + sj + 1];
    float4 e001 = ef[b + sk], e101 = ef[b + sk + 1], e011 = ef[b + sk + sj], e111 = ef[b + sk + sj + 1];
    float ax = 1.0f - ox, ay = 1.0f - oy, az = 1.0f - oz;
    float w000 = ax * ay * az, w100 = ox * ay * az, w010 = ax * oy * az, w110 = ox * oy * az;
    float w001 = ax * ay * oz, w101 = ox * ay * oz, w011 = ax * oy * oz, w111 = ox * oy * oz;
    float3 E;
    E.x = e000.x * w000 + e100.x * w100 + e110.x * w110 + e010.x * w010 + e001.x * w001 +
          e101.x * w101 + e111.x * w111 + e011.x * w011;
    E.y = e000.y * w000 + e100.y * w100 + e110.y * w110 + e010.y * w010 + e001.y * w001 +
          e101.y * w101 + e111.y * w111 + e011.y * w011;
    E.z = e000.z * w000 + e100.z * w100 + e110.z * w110 + e010.z * w010 + e001.z * w001 +
          e101.z * w101 + e111.z * w111 + e011.z * w011;
    return E;
}

// Leapfrog step of a PIC_MIXED particle given E: v += E*(dt q/m); x += v dt, then the wall
// treatment.  Returns false when the particle is absorbed.
__device__ inline bool push_mixed(const PushParams &P, PMixed &p, float3 E) {
    p.vx += E.x * P.qm_dt;
    p.vy += E.y * P.qm_dt;
    p.vz += E.z * P.qm_dt;
    advance_axis(p.ci, p.ox, P.fdh_hi[0] + (p.vx * P.dt_dh[0] + P.fdh_lo[0]));
    advance_axis(p.cj, p.oy, P.fdh_hi[1] + (p.vy * P.dt_dh[1] + P.fdh_lo[1]));
    advance_axis(p.ck, p.oz, P.fdh_hi[2] + (p.vz * P.dt_dh[2] + P.fdh_lo[2]));
    if (P.bc == PIC_BC_REFLECT) {
        reflect_axis(p.ci, p.ox, p.vx, P.g.nci, P.fr2[0]);
        reflect_axis(p.cj, p.oy, p.vy, P.g.ncj, P.fr2[1]);
        reflect_axis(p.ck, p.oz, p.vz, P.g.nck, P.fr2[2]);
        return true;
    }
    // absorbing: !inBounds (half-open [x0,xm)) || inSphere  (ch3/ver2/Species.cpp:28)
    if (p.ci < 0 || p.ci >= P.g.nci || p.cj < 0 || p.cj >= P.g.ncj || p.ck < 0 || p.ck >= P.g.nck)
        return false;
    if (P.sphere_on) {
        float rx = ((float)(p.ci - P.sph_c[0]) + (p.ox - P.sph_f[0])) * P.dhf[0];
        float ry = ((float)(p.cj - P.sph_c[1]) + (p.oy - P.sph_f[1])) * P.dhf[1];
        float rz = ((float)(p.ck - P.sph_c[2]) + (p.oz - P.sph_f[2])) * P.dhf[2];
        if (rx * rx + ry * ry + rz * rz <= P.r2f) return false;
    }
    return true;
}

// ---- FP64 push: the reference's arithmetic (ch2/Species.cpp:19-36, ch3/ver2/Species.cpp) --
__device__ inline double3 gather64_weights(const double *Ex, const double *Ey, const double *Ez,
                                           long b, long sj, long sk, double di, double dj,
                                           double dk) {
#pragma clang fp contract(off)
    // Field_::gather (ch2/Field.h:186-208): corners in the reference's order, each term
    // E*(w_i)*(w_j)*(w_k), summed left to right.
    const long off[8] = {0, 1, 1 + sj, sj, sk, 1 + sk, 1 + sj + sk, sj + sk};
    const double wi[8] = {1 - di, di, di, 1 - di, 1 - di, di, di, 1 - di};
    const double wj[8] = {1 - dj, 1 - dj, dj, dj, 1 - dj, 1 - dj, dj, dj};
    const double wk[8] = {1 - dk, 1 - dk, 1 - dk, 1 - dk, dk, dk, dk, dk};
    double3 r;
    r.x = Ex[b + off[0]] * wi[0] * wj[0] * wk[0];
    r.y = Ey[b + off[0]] * wi[0] * wj[0] * wk[0];
    r.z = Ez[b + off[0]] * wi[0] * wj[0] * wk[0];
#pragma unroll
    for (int n = 1; n < 8; n++) {
        r.x = r.x + Ex[b + off[n]] * wi[n] * wj[n] * wk[n];
        r.y = r.y + Ey[b + off[n]] * wi[n] * wj[n] * wk[n];
        r.z = r.z + Ez[b + off[n]] * wi[n] * wj[n] * wk[n];
    }
    return r;
}

__device__ inline double3 gather64_global(const double4 *ef, long b, long sj, long sk, double di,
                                          double dj, double dk) {
#pragma clang fp contract(off)
    const long off[8] = {0, 1, 1 + sj, sj, sk, 1 + sk, 1 + sj + sk, sj + sk};
    const double wi[8] = {1 - di, di, di, 1 - di, 1 - di, di, di, 1 - di};
    const double wj[8] = {1 - dj, 1 - dj, dj, dj, 1 - dj, 1 - dj, dj, dj};
    const double wk[8] = {1 - dk, 1 - dk, 1 - dk, 1 - dk, dk, dk, dk, dk};
    double4 e = ef[b];
    double3 r;
    r.x = e.x * wi[0] * wj[0] * wk[0];
    r.y = e.y * wi[0] * wj[0] * wk[0];
    r.z = e.z * wi[0] * wj[0] * wk[0];
#pragma unroll
    for (int n = 1; n < 8; n++) {
        e = ef[b + off[n]];
        r.x = r.x + e.x * wi[n] * wj[n] * wk[n];
        r.y = r.y + e.y * wi[n] * wj[n] * wk[n];
        r.z = r.z + e.z * wi[n] * wj[n] * wk[n];
    }
    return r;
}

__device__ inline bool push_f64(const PushParams &P, PF64 &p, double3 E) {
#pragma clang fp contract(off)
    const Grid &g = P.g;
    p.vx += E.x * P.s64;
    p.vy += E.y * P.s64;
    p.vz += E.z * P.s64;
    p.x += p.vx * P.dt;
    p.y += p.vy * P.dt;
    p.z += p.vz * P.dt;
    if (P.bc == PIC_BC_REFLECT) {
        double *xs[3] = {&p.x, &p.y, &p.z};
        double *vs[3] = {&p.vx, &p.vy, &p.vz};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double &x = *xs[c];
            if (x < g.x0[c]) {
                x = 2 * g.x0[c] - x;
                *vs[c] *= -1.0;
            } else if (x >= g.xm[c]) {
                x = 2 * g.xm[c] - x;
                *vs[c] *= -1.0;
            }
        }
        return true;
    }
    if (p.x < g.x0[0] || p.x >= g.xm[0] || p.y < g.x0[1] || p.y >= g.xm[1] || p.z < g.x0[2] ||
        p.z >= g.xm[2])
        return false;
    if (P.sphere_on) {
        double r0 = p.x - P.sph_c64[0], r1 = p.y - P.sph_c64[1], r2 = p.z - P.sph_c64[2];
        double m2 = (r0 * r0 + r1 * r1 + r2 * r2);
        if (m2 <= P.sph_r2) return false;
    }
    return true;
}

// ---- deposit helpers --------------------------------------------------------------------
__device__ inline void deposit_global64(const Grid &g, double *den, int i, int j, int k, double di,
                                        double dj, double dk, double w) {
#pragma clang fp contract(off)
    // Field_::scatter (ch2/Field.h:164-183)
    long sj = g.ni, sk = (long)g.ni * g.nj;
    long b = i + sj * j + sk * k;
    atomicAdd(&den[b], w * (1 - di) * (1 - dj) * (1 - dk));
    atomicAdd(&den[b + 1], w * (di) * (1 - dj) * (1 - dk));
    atomicAdd(&den[b + 1 + sj], w * (di) * (dj) * (1 - dk));
    atomicAdd(&den[b + sj], w * (1 - di) * (dj) * (1 - dk));
    atomicAdd(&den[b + sk], w * (1 - di) * (1 - dj) * (dk));
    atomicAdd(&den[b + 1 + sk], w * (di) * (1 - dj) * (dk));
    atomicAdd(&den[b + 1 + sj + sk], w * (di) * (dj) * (dk));
    atomicAdd(&den[b + sj + sk], w * (1 - di) * (dj) * (dk));
}

// CIC deposit into the LDS window: weights computed in the particle's arithmetic T, summed in
// FP64 (ds_add_f64) for both forms.  Cell-sorted particles make the lanes of a wavefront hit
// the same window nodes; under that contention ds_add_f32 measured 2.8x slower than
// ds_add_f64 on gfx950 (44.5 ms vs 16.0 ms for the whole 1e9-particle push).
// cell and CIC offsets of a particle for a global deposit
__device__ inline void dep_params(const Grid &g, const PMixed &p, int &i, int &j, int &k, double &di, double &dj,
                                  double &dk, double &w) {
    i = p.ci;
    j = p.cj;
    k = p.ck;
    di = p.ox;
    dj = p.oy;
    dk = p.oz;
    w = p.w;
}

__device__ inline void dep_params(const Grid &g, const PF64 &p, int &i, int &j, int &k, double &di, double &dj,
                                  double &dk, double &w) {
    cell_of(g, p, i, j, k, di, dj, dk);
    w = p.w;
}

// Global CIC deposit of one particle (outside the LDS windows).  Per-lane atomics: the
// particles that leave a tile's window in one step mostly land in distinct cells, and
// grouping equal cells across the wavefront costs more than the atomics it saves.
template <typename PT>
__device__ inline void deposit_global(const Grid &g, double *den, const PT &p) {
    int i, j, k;
    double di, dj, dk, w;
    dep_params(g, p, i, j, k, di, dj, dk, w);
    deposit_global64(g, den, i, j, k, di, dj, dk, w);
}

template <typename T, int WN = pic::WN>
__device__ inline void deposit_lds(double *sW, int wi, int wj, int wk, T ox, T oy, T oz, T w) {
    int b = wi + WN * (wj + WN * wk);
    T ax = T(1) - ox, ay = T(1) - oy, az = T(1) - oz;
    atomicAdd(&sW[b], (double)(w * ax * ay * az));
    atomicAdd(&sW[b + 1], (double)(w * ox * ay * az));
    atomicAdd(&sW[b + 1 + WN], (double)(w * ox * oy * az));
    atomicAdd(&sW[b + WN], (double)(w * ax * oy * az));
    atomicAdd(&sW[b + WN * WN], (double)(w * ax * ay * oz));
    atomicAdd(&sW[b + 1 + WN * WN], (double)(w * ox * ay * oz));
    atomicAdd(&sW[b + 1 + WN + WN * WN], (double)(w * ox * oy * oz));
    atomicAdd(&sW[b + WN + WN * WN], (double)(w * ax * oy * oz));
}

// ---------------------------------------------------------------------------------------
// per-step bin plan: capacities of the output bins from the input counts
// ---------------------------------------------------------------------------------------
// Tile t gets an own segment of exactly n[t] slots (its input count: every particle that
// stays fits) followed by an inflow segment for one step of arrivals: the larger of
// max(n/8, nb/4) (nb the largest face-neighbour population) and 3/2 of the arrivals the tile
// saw in the last push (the raw reservation count, including arrivals that found the
// segment full) plus the staged particles now in the tile, then + 64.  Sizing from observed
// arrivals lets converging flows (the wake behind the sphere) get the room they need, and
// counting staged particles drains the staging list instead of re-spilling it.  The inflow
// room is scaled down uniformly when the buffer is short.  Mode 1 (counting sort) sizes bins
// from a histogram with no inflow segment.  Five small launches over all tiles: counts,
// totals, capacities, block offsets, starts (a single-workgroup plan was latency-bound at
// 0.3 ms for 65k tiles).
constexpr int PLAN_THREADS = 256;
constexpr int PLAN_SCAN_THREADS = 1024;

// scratch of the plan (per species): per-tile need and slack, per-block sums, alpha
struct PlanTmp {
    uint32_t *need, *slack;
    uint32_t *staged;         // staged particles per current tile, PLAN_STAGED_SHARDS partial
                              // counts per tile (shard-major); zeroed again by k_plan_caps
    uint32_t *nchunk;         // push workgroups per tile
    unsigned long long *blk;  // [2 * nblocks]
    unsigned long long *blk2; // [nblocks] chunk block sums, then their offsets
    unsigned long long *total;// [0] push workgroups over all tiles
    double *alpha;            // [0] alpha, [1] 1 when the buffer is short
};

__device__ inline uint32_t tile_count(const PartView &v, int t) {
    BinSeg b = bin_of(v, t);
    return b.c0 + b.c1;
}

__device__ inline void plan_sizes(const Grid &g, const PartView &in, const uint32_t *hist, const uint32_t *staged,
                                  int mode, int t, uint32_t &n, uint32_t &s) {
    if (mode == 0) {
        n = tile_count(in, t);
        uint32_t nb = 0;
        int tx = t % g.ntx, ty = (t / g.ntx) % g.nty, tz = t / (g.ntx * g.nty);
        if (tx > 0) nb = max(nb, tile_count(in, t - 1));
        if (tx < g.ntx - 1) nb = max(nb, tile_count(in, t + 1));
        if (ty > 0) nb = max(nb, tile_count(in, t - g.ntx));
        if (ty < g.nty - 1) nb = max(nb, tile_count(in, t + g.ntx));
        if (tz > 0) nb = max(nb, tile_count(in, t - g.ntx * g.nty));
        if (tz < g.ntz - 1) nb = max(nb, tile_count(in, t + g.ntx * g.nty));
        uint32_t arrivals = in.bin_in[t];
        for (int k = 0; k < PLAN_STAGED_SHARDS; k++) arrivals += staged[(long)k * g.ntiles + t];
        s = max(max(n / 8, nb / 4), arrivals + arrivals / 2) + 64;
    } else {
        n = hist[t];
        s = 0;
    }
}

__device__ inline unsigned long long block_sum_u64(unsigned long long v, unsigned long long *sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    unsigned long long r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) r += sh[i];
    __syncthreads();
    return r;
}

// staged particles per current tile (the staging list of the input buffer).  The list is
// mostly a step's injection, all on the inlet face: the counts are split over
// PLAN_STAGED_SHARDS copies (by workgroup) so that fewer atomics queue on each address.
template <int PREC>
__global__ void __launch_bounds__(256) k_plan_staged(Grid g, PartView in, uint32_t *staged) {
    using PT = typename std::conditional<PREC == PIC_MIXED, PMixed, PF64>::type;
    uint32_t ns = min(*in.stage_count, (uint32_t)in.cap_stage);
    uint32_t *mine = staged + (long)(blockIdx.x % PLAN_STAGED_SHARDS) * g.ntiles;
    for (uint32_t q = blockIdx.x * 256 + threadIdx.x; q < ns; q += gridDim.x * 256) {
        PT p;
        load_p(in, in.cap_bins + q, p);
        atomicAdd(&mine[tile_of(g, p)], 1u);
    }
}

__global__ void __launch_bounds__(PLAN_THREADS)
    k_plan_count(Grid g, PartView in, const uint32_t *hist, int mode, PlanTmp T) {
    __shared__ unsigned long long sh[PLAN_THREADS / 64];
    int t = blockIdx.x * PLAN_THREADS + threadIdx.x;
    uint32_t n = 0, sl = 0;
    if (t < g.ntiles) {
        plan_sizes(g, in, hist, T.staged, mode, t, n, sl);
        T.need[t] = n;
        T.slack[t] = sl;
    }
    unsigned long long a = block_sum_u64(n, sh), b = block_sum_u64(sl, sh);
    if (threadIdx.x == 0) {
        T.blk[2 * blockIdx.x] = a;
        T.blk[2 * blockIdx.x + 1] = b;
    }
}

__global__ void __launch_bounds__(PLAN_SCAN_THREADS)
    k_plan_total(PlanTmp T, int nblk, unsigned long long capb, uint32_t *err) {
    __shared__ unsigned long long sh[PLAN_SCAN_THREADS / 64];
    unsigned long long a = 0, b = 0;
    for (int i = threadIdx.x; i < nblk; i += PLAN_SCAN_THREADS) {
        a += T.blk[2 * i];
        b += T.blk[2 * i + 1];
    }
    unsigned long long need = block_sum_u64(a, sh), slack = block_sum_u64(b, sh);
    if (threadIdx.x == 0) {
        double alpha = 1.0, short_buf = 0.0;
        if (need > capb) {
            alpha = 0.0;
            short_buf = 1.0;
            atomicOr(err, 1u);  // the caller grows the buffers
        } else if (need + slack > capb) {
            alpha = (double)(capb - need) / (double)slack;
        }
        T.alpha[0] = alpha;
        T.alpha[1] = short_buf;
    }
}

__global__ void __launch_bounds__(PLAN_THREADS) k_plan_caps(Grid g, PlanTmp T, PartView out) {
    __shared__ unsigned long long sh[PLAN_THREADS / 64];
    int t = blockIdx.x * PLAN_THREADS + threadIdx.x;
    uint32_t c = 0;
    if (t < g.ntiles) {
        uint32_t own = T.alpha[1] != 0.0 ? 0u : T.need[t];
        c = own + (uint32_t)(T.alpha[0] * (double)T.slack[t]);
        out.bin_cap[t] = c;
        out.bin_off[t] = own;
        for (int k = 0; k < PLAN_STAGED_SHARDS; k++) T.staged[(long)k * g.ntiles + t] = 0;  // next plan
    }
    unsigned long long a = block_sum_u64(c, sh);
    if (threadIdx.x == 0) T.blk[blockIdx.x] = a;
}

// exclusive scan of the nblk block sums in blk, in place (one workgroup, fixed order);
// total (nullable) receives their sum
__global__ void __launch_bounds__(PLAN_SCAN_THREADS)
    k_plan_scan(unsigned long long *blk, int nblk, unsigned long long *total) {
    __shared__ unsigned long long s_scan[PLAN_SCAN_THREADS];
    int per = (nblk + PLAN_SCAN_THREADS - 1) / PLAN_SCAN_THREADS;
    int b0 = threadIdx.x * per, b1 = min(nblk, b0 + per);
    unsigned long long local = 0;
    for (int b = b0; b < b1; b++) local += blk[b];
    s_scan[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < PLAN_SCAN_THREADS; off <<= 1) {
        unsigned long long v = threadIdx.x >= off ? s_scan[threadIdx.x - off] : 0;
        __syncthreads();
        s_scan[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long run = s_scan[threadIdx.x] - local;
    for (int b = b0; b < b1; b++) {
        unsigned long long v = blk[b];
        blk[b] = run;
        run += v;
    }
    if (total && threadIdx.x == PLAN_SCAN_THREADS - 1) *total = s_scan[threadIdx.x];
}

__global__ void __launch_bounds__(PLAN_THREADS) k_plan_starts(Grid g, PlanTmp T, PartView out) {
    __shared__ unsigned long long s_scan[PLAN_THREADS];
    int t = blockIdx.x * PLAN_THREADS + threadIdx.x;
    unsigned long long c = t < g.ntiles ? out.bin_cap[t] : 0;
    s_scan[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < PLAN_THREADS; off <<= 1) {
        unsigned long long v = threadIdx.x >= off ? s_scan[threadIdx.x - off] : 0;
        __syncthreads();
        s_scan[threadIdx.x] += v;
        __syncthreads();
    }
    if (t < g.ntiles) {
        out.bin_start[t] = (uint32_t)(T.blk[blockIdx.x] + s_scan[threadIdx.x] - c);
        out.bin_own[t] = 0;
        out.bin_in[t] = 0;
    }
    if (t == 0) *out.stage_count = 0;
}

// push workgroups per tile: one per `chunk` of its input particles, at least one
// and which instantiation pushes each tile: the wide-rim one when at least 16 of its
// particles, and 1/128 of them, deposited outside the window in the last push (force: 1 all
// tiles, -1 none); the far counts restart from 0
__global__ void __launch_bounds__(PLAN_THREADS)
    k_plan_chunks(Grid g, PlanTmp T, uint32_t *far, uint32_t *wide, int force, uint32_t chunk) {
    __shared__ unsigned long long sh[PLAN_THREADS / 64];
    int t = blockIdx.x * PLAN_THREADS + threadIdx.x;
    uint32_t c = 0;
    if (t < g.ntiles) {
        c = max(1u, (T.need[t] + chunk - 1) / chunk);
        T.nchunk[t] = c;
        uint32_t f = far[t];
        far[t] = 0;
        wide[t] = force > 0 ? 1u : force < 0 ? 0u : (f >= 16u && (unsigned long long)f * 128u >= T.need[t]) ? 1u : 0u;
    }
    unsigned long long a = block_sum_u64(c, sh);
    if (threadIdx.x == 0) T.blk2[blockIdx.x] = a;
}

// map[w] = tile | chunk << 32 | wide << 63 for push workgroup w, tiles in order (wide: the
// tile's flag from k_plan_chunks, read by each push workgroup with its entry); entries past
// the map's capacity (only when the particle count exceeds the bins, which the plan flags)
// are dropped
__global__ void __launch_bounds__(PLAN_THREADS)
    k_plan_map(Grid g, PlanTmp T, const uint32_t *wide, uint64_t *map, long cap) {
    __shared__ unsigned long long s_scan[PLAN_THREADS];
    int t = blockIdx.x * PLAN_THREADS + threadIdx.x;
    unsigned long long c = t < g.ntiles ? T.nchunk[t] : 0;
    s_scan[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < PLAN_THREADS; off <<= 1) {
        unsigned long long v = threadIdx.x >= off ? s_scan[threadIdx.x - off] : 0;
        __syncthreads();
        s_scan[threadIdx.x] += v;
        __syncthreads();
    }
    if (t < g.ntiles) {
        unsigned long long w0 = T.blk2[blockIdx.x] + s_scan[threadIdx.x] - c;
        const uint64_t wb = wide[t] ? (1ull << 63) : 0ull;
        for (unsigned long long k = 0; k < c && w0 + k < (unsigned long long)cap; k++)
            map[w0 + k] = (uint64_t)t | ((uint64_t)k << 32) | wb;
    }
}

// ---------------------------------------------------------------------------------------
// the fused push: gather (LDS E tile) + leapfrog + walls + CIC deposit (LDS window) + rebin
// ---------------------------------------------------------------------------------------
struct PushOut {
    const uint64_t *chunk_map;           // workgroup -> tile | chunk << 32 | wide-rim flag << 63
    const unsigned long long *total_groups;  // workgroups the map holds
    float *win32;
    double *win64;
    uint8_t *win_used;
    double *den_extra;
    double *diag_part;  // [NDIAG_SLOTS][gridDim.x] per-block partials, reduced on demand
    uint32_t *err;
    uint32_t *far;          // per tile: particles deposited outside the window this push
    uint32_t *far_in;       // per tile: 1 when marked far arrivals landed in its inflow segment
    uint32_t chunk;         // particles per push workgroup (the plan's chunk size)
    uint32_t stripe;        // map entries per XCD stripe
};

// slot of an arrival in the inflow segment of tile dtile, -1 when the segment is full
__device__ inline long inflow_slot(const PartView &out, int dtile, uint32_t pos) {
    uint32_t off = out.bin_off[dtile], cap = out.bin_cap[dtile];
    return pos < cap - off ? (long)out.bin_start[dtile] + off + pos : -1L;
}

// slots in the staging list for the lanes with `need` (one atomic per wavefront); -1, and a
// capacity error, when the list is full.  Call from wavefront-uniform control flow.
__device__ inline long staging_slot(bool need, const PartView &out, uint32_t *err) {
    const unsigned long long m = __ballot(need);
    if (!m) return -1L;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(out.stage_count, (uint32_t)__popcll(m));
    const uint32_t s = (uint32_t)__builtin_amdgcn_readlane((int)base, leader) +
                       __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
    if (!need) return -1L;
    if (s < out.cap_stage) return out.cap_bins + s;
    atomicOr(err, 2u);  // staging full: particle dropped, reported as a capacity error
    return -1L;
}

// RIM: the window's rim in cells, HALO or HALO_WIDE.  Both instantiations are launched over
// the same workgroups; each pushes the tiles the plan gave it (the staging list goes to the
// HALO one) and returns at once elsewhere, leaving that workgroup's partials to the other.
template <int PREC, int RIM>
__global__ void __launch_bounds__(PUSH_THREADS)
    k_push(PushParams P, PartView in, PartView out, const float4 *__restrict__ ef32,
           const double4 *__restrict__ ef64, PushOut O) {
    using PT = typename std::conditional<PREC == PIC_MIXED, PMixed, PF64>::type;
    constexpr bool WIDE = RIM != HALO;
    constexpr int WNR = TILE + 1 + 2 * RIM;  // window nodes per edge
    constexpr int WNODESR = WNR * WNR * WNR;
    const Grid &g = P.g;
    __shared__ uint32_t sOwn, sFar;
    __shared__ double sRed[NDIAG_SLOTS][PUSH_THREADS / 64];
    // LDS: E tile (3 float or 3 double planes), the FP64 deposit window, and the output bin
    // segments of the 27 tiles around this one (start, inflow offset, capacity)
    __shared__ __attribute__((aligned(16))) unsigned char
        sRaw[(PREC == PIC_MIXED ? TNODES * 12 : TNODES * 24)];
    __shared__ double sW[WNODESR];
    __shared__ uint32_t sNb[3][27];

    // per-thread diagnostic sums in the particle arithmetic (a thread sees a few dozen
    // particles of one tile), reduced in FP64 across the workgroup and the grid.  The weight
    // sum is FP64 in both forms: a run of equal FP32 weights rounds the same way at every
    // addition, and the real count (and, through the frame velocity, the momentum and KE:
    // sum w v = S + u W) would inherit that bias.
    using AT = typename std::conditional<PREC == PIC_MIXED, float, double>::type;
    AT acc[NDIAG_SLOTS] = {0, 0, 0, 0, 0, 0};
    double accw = 0;
    const int tid = threadIdx.x;
    auto tally = [&](const PT &p) {
        AT vx = p.vx, vy = p.vy, vz = p.vz, w = p.w;
        accw += (double)p.w;
        acc[1] += vx * w;
        acc[2] += vy * w;
        acc[3] += vz * w;
        acc[4] += w * (vx * vx + vy * vy + vz * vz);
        acc[5] += 1;
    };

    // Workgroups [0, PUSH_STAGE_WGS) push the staging list, the rest the binned tiles.  The
    // staging ones come first in dispatch order: their particles (a step's injection) gather
    // and deposit through global memory and take longer per particle, so started last they
    // would be the kernel's tail.
    const unsigned nbin = gridDim.x - PUSH_STAGE_WGS;  // workgroups for the binned tiles
    if (blockIdx.x >= PUSH_STAGE_WGS) {
        // ------------------------------------------------ binned tile: waves run independently
        // workgroup -> (tile, chunk) from the plan's map.  Workgroups are dispatched
        // round-robin over the 8 XCDs (XCD b % 8, PUSH_STAGE_WGS a multiple of 8); the
        // j-th binned one of an XCD takes entry (j / S * 8 + xcd) * S + j % S,
        // S = O.stripe: a bijection of [0, nbin) (nbin is a multiple of 8 S) that
        // deals out stripes of S entries round-robin
        const unsigned total = (unsigned)min(*O.total_groups, (unsigned long long)nbin);
        const unsigned b = blockIdx.x - PUSH_STAGE_WGS;
        const unsigned xcd = b % 8u, j = b / 8u;
        const unsigned wg = ((j / O.stripe) * 8u + xcd) * O.stripe + j % O.stripe;
        if (wg >= total) {
            if (tid < NDIAG_SLOTS) O.diag_part[tid * gridDim.x + blockIdx.x] = 0.0;
            return;
        }
        const uint64_t tc = O.chunk_map[wg];
        const int t = (int)(uint32_t)tc;
        const uint32_t chunk = (uint32_t)(tc >> 32) & 0x7fffffffu;
        if (((tc >> 63) != 0) != WIDE) return;  // the other instantiation's tile
        BinSeg bs = bin_of(in, t);
        const uint32_t cnt_tile = bs.c0 + bs.c1;
        const bool split = cnt_tile > O.chunk;  // several workgroups share tile t
        const uint32_t q_lo = chunk * O.chunk;
        const uint32_t q_hi = min(cnt_tile, q_lo + O.chunk);
        if (q_lo >= q_hi) {
            if (tid == 0) O.win_used[t] = 0;
            if (tid < NDIAG_SLOTS) O.diag_part[tid * gridDim.x + blockIdx.x] = 0.0;
            return;
        }
        int tx = t % g.ntx, ty = (t / g.ntx) % g.nty, tz = t / (g.ntx * g.nty);
        int c0x = tx * TILE, c0y = ty * TILE, c0z = tz * TILE;
        long sj = g.ni, sk = (long)g.ni * g.nj;
        for (int n = tid; n < TNODES; n += PUSH_THREADS) {
            int a = n % TN, b = (n / TN) % TN, c = n / (TN * TN);
            int gi = c0x + a, gj = c0y + b, gk = c0z + c;
            bool ok = gi < g.ni && gj < g.nj && gk < g.nk;
            long u = gi + sj * gj + sk * gk;
            if (PREC == PIC_MIXED) {
                float4 e = ok ? ef32[u] : make_float4(0.f, 0.f, 0.f, 0.f);
                float *pl = reinterpret_cast<float *>(sRaw);
                pl[n] = e.x;
                pl[TNODES + n] = e.y;
                pl[2 * TNODES + n] = e.z;
            } else {
                double4 e = ok ? ef64[u] : make_double4(0, 0, 0, 0);
                double *pl = reinterpret_cast<double *>(sRaw);
                pl[n] = e.x;
                pl[TNODES + n] = e.y;
                pl[2 * TNODES + n] = e.z;
            }
        }
        for (int n = tid; n < WNODESR; n += PUSH_THREADS) sW[n] = 0.0;
        if (tid < 27) {
            int nx = tx + tid % 3 - 1, ny = ty + (tid / 3) % 3 - 1, nz = tz + tid / 9 - 1;
            bool in = nx >= 0 && nx < g.ntx && ny >= 0 && ny < g.nty && nz >= 0 && nz < g.ntz;
            int nt = nx + g.ntx * (ny + g.nty * nz);
            sNb[0][tid] = in ? out.bin_start[nt] : 0u;
            sNb[1][tid] = in ? out.bin_off[nt] : 0u;
            sNb[2][tid] = in ? out.bin_cap[nt] : 0u;
        }
        if (tid == 0) {
            sOwn = 0;
            sFar = 0;
        }
        __syncthreads();
        const long own0 = out.bin_start[t];  // own segment of this tile in the output buffer
        const uint32_t own_cap = out.bin_off[t];

        // the next particle of each thread is loaded while the current one is processed
        PT pn;
        if (q_lo + tid < q_hi) load_p(in, seg_index(bs, q_lo + tid), pn);
        for (uint32_t base = q_lo; base < q_hi; base += PUSH_THREADS) {
            uint32_t q = base + tid;
            PT p = pn;
            if (q + PUSH_THREADS < q_hi) load_p(in, seg_index(bs, q + PUSH_THREADS), pn);
            bool alive = false;
            int dtile = -1, nbi = -1;   // destination tile, and its index among the 27 around t
            int wi = -1, wj = 0, wk = 0;  // deposit window cell; wi < 0: outside the window
            double di = 0, dj = 0, dk = 0;  // FP64 form: CIC offsets after the push
            if (q < q_hi) {
                int ci, cj, ck;  // cell after the push
                if constexpr (PREC == PIC_MIXED) {
                    int li = min(max(p.ci - c0x, 0), TILE - 1);
                    int lj = min(max(p.cj - c0y, 0), TILE - 1);
                    int lk = min(max(p.ck - c0z, 0), TILE - 1);
                    float3 E = gather_lds(reinterpret_cast<const float *>(sRaw), li, lj, lk, p.ox, p.oy, p.oz);
                    alive = push_mixed(P, p, E);
                    ci = p.ci;
                    cj = p.cj;
                    ck = p.ck;
                } else {
                    cell_of(g, p, ci, cj, ck, di, dj, dk);
                    int li = min(max(ci - c0x, 0), TILE - 1);
                    int lj = min(max(cj - c0y, 0), TILE - 1);
                    int lk = min(max(ck - c0z, 0), TILE - 1);
                    const double *pl = reinterpret_cast<const double *>(sRaw);
                    double3 E = gather64_weights(pl, pl + TNODES, pl + 2 * TNODES, li + TN * (lj + TN * lk), TN,
                                                 TN * TN, di, dj, dk);
                    alive = push_f64(P, p, E);
                    if (alive) cell_of(g, p, ci, cj, ck, di, dj, dk);
                }
                if (alive) {
                    tally(p);
                    int a = ci - c0x + RIM, b = cj - c0y + RIM, c = ck - c0z + RIM;
                    if (a >= 0 && a < TILE + 2 * RIM && b >= 0 && b < TILE + 2 * RIM && c >= 0 &&
                        c < TILE + 2 * RIM) {
                        wi = a;
                        wj = b;
                        wk = c;
                    }
                    dtile = tile_id(g, ci, cj, ck);
                    int dx = (ci >> TILE_SHIFT) - tx, dy = (cj >> TILE_SHIFT) - ty, dz = (ck >> TILE_SHIFT) - tz;
                    if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1)
                        nbi = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
                }
            }
            // Slots first, deposit second: the reservation atomics are in flight while the
            // deposit runs.  Stayers: one atomic per wavefront on the own segment's counter
            // (LDS, or global when the tile is split); arrivals elsewhere: one atomic per
            // distinct destination, all in one instruction.
            const bool stay = alive && dtile == t;
            const unsigned long long m = __ballot(stay);
            const int sleader = m ? __ffsll((long long)m) - 1 : 0;
            uint32_t sbase = 0;
            if (m && (int)__lane_id() == sleader)
                sbase = split ? atomicAdd(&out.bin_own[t], (uint32_t)__popcll(m))
                              : atomicAdd(&sOwn, (uint32_t)__popcll(m));
            const int key = (alive && !stay) ? dtile : -1;
            const KeyGroup kg = wave_group(key);
            uint32_t lbase = 0;
            if (key >= 0 && kg.before == 0) lbase = atomicAdd(&out.bin_in[key], kg.count);
            if (wi >= 0) {
                if constexpr (PREC == PIC_MIXED)
                    deposit_lds<float, WNR>(sW, wi, wj, wk, p.ox, p.oy, p.oz, p.w);
                else
                    deposit_lds<double, WNR>(sW, wi, wj, wk, di, dj, dk, p.w);
            }
            const unsigned long long fm = __ballot(alive && wi < 0);  // far deposits, counted
            if (fm && (int)__lane_id() == __ffsll((long long)fm) - 1) atomicAdd(&sFar, (uint32_t)__popcll(fm));
            long idx = -1;
            if (m) {
                uint32_t slot = (uint32_t)__builtin_amdgcn_readlane((int)sbase, sleader) +
                                __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                if (stay && slot < own_cap) idx = own0 + slot;
            }
            uint32_t pos = wave_group_base(kg, key, lbase);
            if (key >= 0) {
                if (nbi >= 0) {
                    uint32_t off = sNb[1][nbi], cap = sNb[2][nbi];
                    idx = pos < cap - off ? (long)sNb[0][nbi] + off + pos : -1L;
                } else {
                    idx = inflow_slot(out, key, pos);
                }
            }
            // no room (bins shrunk by a short buffer, inflow segment full): staging list
            const bool to_stage = alive && idx < 0;
            const long sidx = staging_slot(to_stage, out, O.err);
            if (to_stage) idx = sidx;
            // A particle that deposits outside the window (always an arrival: a stayer's cell
            // is in the tile) is marked in its slot's byte, and k_far_arrivals deposits it
            // through its destination tile's window after the push, in place of eight global
            // atomics per particle.
            if (alive && wi < 0) {
                if (idx >= 0) {
                    out.mark[idx] = 1;
                    if (!to_stage) O.far_in[dtile] = 1u;
                } else {
                    deposit_global(g, O.den_extra, p);  // dropped (staging full): deposit now
                }
            }
            if (idx >= 0) store_p(out, idx, p);
        }
        __syncthreads();
        if (tid == 0 && sFar) atomicAdd(&O.far[t], sFar);
        // window node n outside the tile's window slot (further chunks of a split tile: all of
        // them; wide rim: the nodes beyond HALO cells) goes to the far-deposit array
        auto add_far = [&](int n, double v) {
            if (v == 0.0) return;
            int gi = c0x - RIM + n % WNR, gj = c0y - RIM + (n / WNR) % WNR, gk = c0z - RIM + n / (WNR * WNR);
            if (gi >= 0 && gi < g.ni && gj >= 0 && gj < g.nj && gk >= 0 && gk < g.nk)
                atomicAdd(&O.den_extra[gi + sj * gj + sk * gk], v);
        };
        if (chunk == 0) {
            // flush the deposit window into the tile's window slot
            for (int n = tid; n < WNODESR; n += PUSH_THREADS) {
                int m = n;  // index in the slot's HALO-rim window
                if constexpr (WIDE) {
                    constexpr int D = RIM - HALO;
                    int a = n % WNR - D, b = (n / WNR) % WNR - D, c = n / (WNR * WNR) - D;
                    if (a < 0 || a >= WN || b < 0 || b >= WN || c < 0 || c >= WN) {
                        add_far(n, sW[n]);
                        continue;
                    }
                    m = a + WN * (b + WN * c);
                }
                if constexpr (PREC == PIC_MIXED)
                    O.win32[(long)t * WNODES + m] = (float)sW[n];
                else
                    O.win64[(long)t * WNODES + m] = sW[n];
            }
            if (tid == 0) {
                O.win_used[t] = 1;
                if (!split) out.bin_own[t] = sOwn;
            }
        } else {
            for (int n = tid; n < WNODESR; n += PUSH_THREADS) add_far(n, sW[n]);
        }
    } else {
        // ------------------------------------------------ unbinned staging list
        if constexpr (WIDE) return;  // pushed by the HALO instantiation
        uint32_t ns = min(*in.stage_count, (uint32_t)in.cap_stage);
        uint32_t nthreads = PUSH_STAGE_WGS * PUSH_THREADS;
        uint32_t first = blockIdx.x * PUSH_THREADS + tid;
        uint32_t trips = (ns + nthreads - 1) / nthreads;  // uniform per wavefront
        for (uint32_t it = 0; it < trips; it++) {
            uint32_t q = first + it * nthreads;
            PT p;
            int dt = -1;
            bool alive = false;
            if (q < ns) {
                load_p(in, in.cap_bins + q, p);
                if constexpr (PREC == PIC_MIXED) {
                    float3 E = gather_global32(g, ef32, p.ci, p.cj, p.ck, p.ox, p.oy, p.oz);
                    alive = push_mixed(P, p, E);
                    if (alive) dt = tile_of(g, p);
                } else {
                    int i, j, k;
                    double di, dj, dk;
                    cell_of(g, p, i, j, k, di, dj, dk);
                    long sj = g.ni, sk = (long)g.ni * g.nj;
                    double3 E = gather64_global(ef64, i + sj * j + sk * k, sj, sk, di, dj, dk);
                    alive = push_f64(P, p, E);
                    if (alive) {
                        cell_of(g, p, i, j, k, di, dj, dk);
                        dt = tile_id(g, i, j, k);
                    }
                }
                if (alive) tally(p);
            }
            if (alive) deposit_global(g, O.den_extra, p);
            uint32_t pos = wave_reserve_global(dt, out.bin_in);
            long idx = dt >= 0 ? inflow_slot(out, dt, pos) : -1L;
            const bool to_stage = dt >= 0 && idx < 0;
            const long sidx = staging_slot(to_stage, out, O.err);
            if (to_stage) idx = sidx;
            if (idx >= 0) store_p(out, idx, p);
        }
    }
    // per-block partials of getRealCount, getMomentum, getKE and the particle count
#pragma unroll
    for (int q = 0; q < NDIAG_SLOTS; q++) {
        double v = warp_sum(q == 0 ? accw : (double)acc[q]);
        if ((tid & 63) == 0) sRed[q][tid >> 6] = v;
    }
    __syncthreads();
    if (tid < NDIAG_SLOTS) {
        double s = 0;
        for (int i = 0; i < PUSH_THREADS / 64; i++) s += sRed[tid][i];
        O.diag_part[tid * gridDim.x + blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------------------
// deferred far deposits of the push
// ---------------------------------------------------------------------------------------
// Workgroup t < ntiles: when the push flagged tile t, the marked particles of its inflow
// segment (one byte per slot, cleared here) are deposited into an LDS window of the tile's
// (TILE+1)^3 nodes, in the push's arithmetic.  The window is then added into the tile's
// window slot (this workgroup owns it: no atomics), or written with the rest of the slot
// zeroed when the push left the slot unused, and the flag is cleared.  Workgroups past ntiles
// walk the staging list, whose marked particles (arrivals that found their segment full)
// deposit by global atomics.
template <int PREC>
__global__ void __launch_bounds__(256)
    k_far_arrivals(Grid g, PartView out, uint32_t *far_in, float *win32, double *win64, uint8_t *win_used,
                   double *den_extra) {
    using PT = typename std::conditional<PREC == PIC_MIXED, PMixed, PF64>::type;
    __shared__ double sD[TNODES];
    const int tid = threadIdx.x;
    const int t = blockIdx.x;
    if (t >= g.ntiles) {
        const uint32_t ns = min(*out.stage_count, (uint32_t)out.cap_stage);
        const uint32_t nth = (gridDim.x - g.ntiles) * 256u;
        for (uint32_t q = (t - g.ntiles) * 256u + tid; q < ns; q += nth) {
            const long idx = out.cap_bins + q;
            if (!out.mark[idx]) continue;
            out.mark[idx] = 0;
            PT p;
            load_p(out, idx, p);
            deposit_global(g, den_extra, p);
        }
        return;
    }
    if (far_in[t] == 0u) return;
    for (int n = tid; n < TNODES; n += 256) sD[n] = 0.0;
    __syncthreads();
    if (tid == 0) far_in[t] = 0u;
    const int tx = t % g.ntx, ty = (t / g.ntx) % g.nty, tz = t / (g.ntx * g.nty);
    const int c0x = tx * TILE, c0y = ty * TILE, c0z = tz * TILE;
    const BinSeg bs = bin_of(out, t);
    // the marks of the inflow segment [lo, hi), 16 bytes per load (the mark array is 256-byte
    // aligned and padded to 256 bytes; bytes of the neighbouring segments are skipped)
    const long lo = (long)bs.s1, hi = lo + bs.c1;
    for (long base = (lo & ~15L) + 16L * tid; base < hi; base += 16L * 256) {
        const uint4 m4 = *reinterpret_cast<const uint4 *>(out.mark + base);
        if ((m4.x | m4.y | m4.z | m4.w) == 0u) continue;
        const uint32_t words[4] = {m4.x, m4.y, m4.z, m4.w};
        for (int e = 0; e < 16; e++) {
            if (((words[e >> 2] >> (8 * (e & 3))) & 0xffu) == 0u) continue;
            const long idx = base + e;
            if (idx < lo || idx >= hi) continue;
            out.mark[idx] = 0;
            PT p;
            load_p(out, idx, p);
            if constexpr (PREC == PIC_MIXED) {
                deposit_lds<float, TN>(sD, p.ci - c0x, p.cj - c0y, p.ck - c0z, p.ox, p.oy, p.oz, p.w);
            } else {
                int ci, cj, ck;
                double di, dj, dk;
                cell_of(g, p, ci, cj, ck, di, dj, dk);
                deposit_lds<double, TN>(sD, ci - c0x, cj - c0y, ck - c0z, di, dj, dk, p.w);
            }
        }
    }
    __syncthreads();
    const bool used = win_used[t] != 0;
    if (used) {
        // add the tile's (TILE+1)^3 nodes into its slot
        for (int n = tid; n < TNODES; n += 256) {
            const double v = sD[n];
            if (v == 0.0) continue;
            const int a = n % TN + HALO, b = (n / TN) % TN + HALO, c = n / (TN * TN) + HALO;
            const long m = (long)t * WNODES + a + WN * (b + WN * c);
            if constexpr (PREC == PIC_MIXED)
                win32[m] = (float)((double)win32[m] + v);
            else
                win64[m] += v;
        }
    } else {
        // the push left the slot unused: write it whole
        for (int n = tid; n < WNODES; n += 256) {
            const int a = n % WN - HALO, b = (n / WN) % WN - HALO, c = n / (WN * WN) - HALO;
            const bool inner = a >= 0 && a < TN && b >= 0 && b < TN && c >= 0 && c < TN;
            const double v = inner ? sD[a + TN * (b + TN * c)] : 0.0;
            const long m = (long)t * WNODES + n;
            if constexpr (PREC == PIC_MIXED)
                win32[m] = (float)v;
            else
                win64[m] = v;
        }
    }
    if (!used && tid == 0) win_used[t] = 1;
}

// ---------------------------------------------------------------------------------------
// density finish: den = (sum of tile windows + global extras) / node_vol
// ---------------------------------------------------------------------------------------
template <typename WT>
__global__ void __launch_bounds__(256)
    k_finish_density(Grid g, const WT *__restrict__ win, const uint8_t *__restrict__ win_used,
                     double *den_extra, double *den) {
#pragma clang fp contract(off)
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= g.nn) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    double raw = den_extra[u];
    den_extra[u] = 0;  // ready for the next deposit
    // tiles whose window [a*TILE-HALO, a*TILE+TILE+HALO] holds node i
    int ax0 = max(0, floordiv(i - TILE - HALO + TILE - 1, TILE)), ax1 = min(g.ntx - 1, floordiv(i + HALO, TILE));
    int ay0 = max(0, floordiv(j - TILE - HALO + TILE - 1, TILE)), ay1 = min(g.nty - 1, floordiv(j + HALO, TILE));
    int az0 = max(0, floordiv(k - TILE - HALO + TILE - 1, TILE)), az1 = min(g.ntz - 1, floordiv(k + HALO, TILE));
    for (int c = az0; c <= az1; c++)
        for (int b = ay0; b <= ay1; b++)
            for (int a = ax0; a <= ax1; a++) {
                int t = a + g.ntx * (b + g.nty * c);
                if (!win_used[t]) continue;
                int li = i - a * TILE + HALO, lj = j - b * TILE + HALO, lk = k - c * TILE + HALO;
                raw += (double)win[(long)t * WNODES + li + WN * (lj + WN * lk)];
            }
    // den /= node_vol (ch2/Field.h:125-134; volumes as World::computeNodeVolumes)
    double V = g.dh[0] * g.dh[1] * g.dh[2];
    if (i == 0 || i == g.ni - 1) V *= 0.5;
    if (j == 0 || j == g.nj - 1) V *= 0.5;
    if (k == 0 || k == g.nk - 1) V *= 0.5;
    den[u] = (V != 0) ? raw / V : 0.0;
}

// deposit-only pass (computeNumberDensity without a preceding push), on the push's windows:
// each tile's particles go into an FP64 LDS window of the tile (HALO cells around it) that
// is written to the tile's window slot, which k_finish_density sums; particles outside the
// window (a bin edited from the host) and the staging list use global FP64 atomics.  The
// caller clears win_used: empty tiles leave their slot unused.
template <int PREC>
__global__ void __launch_bounds__(256)
    k_deposit_tiles(Grid g, PartView in, float *win32, double *win64, uint8_t *win_used, double *den_extra) {
    using PT = typename std::conditional<PREC == PIC_MIXED, PMixed, PF64>::type;
    __shared__ double sW[WNODES];
    const int b = blockIdx.x, tid = threadIdx.x;
    if (b < g.ntiles) {
        const BinSeg bs = bin_of(in, b);
        const uint32_t cnt = bs.c0 + bs.c1;
        if (cnt == 0) return;
        const int c0x = (b % g.ntx) * TILE, c0y = ((b / g.ntx) % g.nty) * TILE, c0z = (b / (g.ntx * g.nty)) * TILE;
        for (int n = tid; n < WNODES; n += blockDim.x) sW[n] = 0.0;
        __syncthreads();
        for (uint32_t q = tid; q < cnt; q += blockDim.x) {
            PT p;
            load_p(in, seg_index(bs, q), p);
            int ci, cj, ck;
            double di = 0, dj = 0, dk = 0;
            if constexpr (PREC == PIC_MIXED) {
                ci = p.ci;
                cj = p.cj;
                ck = p.ck;
            } else {
                cell_of(g, p, ci, cj, ck, di, dj, dk);
            }
            const int wi = ci - c0x + HALO, wj = cj - c0y + HALO, wk = ck - c0z + HALO;
            if (wi >= 0 && wi < TILE + 2 * HALO && wj >= 0 && wj < TILE + 2 * HALO && wk >= 0 && wk < TILE + 2 * HALO) {
                if constexpr (PREC == PIC_MIXED)
                    deposit_lds<float>(sW, wi, wj, wk, p.ox, p.oy, p.oz, p.w);
                else
                    deposit_lds<double>(sW, wi, wj, wk, di, dj, dk, p.w);
            } else {
                deposit_global(g, den_extra, p);
            }
        }
        __syncthreads();
        if (PREC == PIC_MIXED) {
            float *dst = win32 + (long)b * WNODES;
            for (int n = tid; n < WNODES; n += blockDim.x) dst[n] = (float)sW[n];
        } else {
            double *dst = win64 + (long)b * WNODES;
            for (int n = tid; n < WNODES; n += blockDim.x) dst[n] = sW[n];
        }
        if (tid == 0) win_used[b] = 1;
    } else {
        uint32_t ns = min(*in.stage_count, (uint32_t)in.cap_stage);
        for (uint32_t q = (b - g.ntiles) * blockDim.x + tid; q < ns; q += (gridDim.x - g.ntiles) * blockDim.x) {
            PT p;
            load_p(in, in.cap_bins + q, p);
            deposit_global(g, den_extra, p);
        }
    }
}

// diagnostics over all particles (when the fused ones are stale): per-block partials
template <int PREC>
__global__ void __launch_bounds__(256) k_diag_all(Grid g, PartView in, double *diag_part) {
    using PT = typename std::conditional<PREC == PIC_MIXED, PMixed, PF64>::type;
    __shared__ double sh[8];
    double acc[NDIAG_SLOTS] = {0, 0, 0, 0, 0, 0};
    auto add = [&](const PT &p) {
        double vx = p.vx, vy = p.vy, vz = p.vz, w = p.w;
        acc[0] += w;
        acc[1] += vx * w;
        acc[2] += vy * w;
        acc[3] += vz * w;
        acc[4] += w * (vx * vx + vy * vy + vz * vz);
        acc[5] += 1;
    };
    int b = blockIdx.x;
    if (b < g.ntiles) {
        BinSeg bs = bin_of(in, b);
        for (uint32_t q = threadIdx.x; q < bs.c0 + bs.c1; q += blockDim.x) {
            PT p;
            load_p(in, seg_index(bs, q), p);
            add(p);
        }
    } else {
        uint32_t ns = min(*in.stage_count, (uint32_t)in.cap_stage);
        for (uint32_t q = (b - g.ntiles) * blockDim.x + threadIdx.x; q < ns; q += (gridDim.x - g.ntiles) * blockDim.x) {
            PT p;
            load_p(in, in.cap_bins + q, p);
            add(p);
        }
    }
    for (int q = 0; q < NDIAG_SLOTS; q++) {
        double s = block_sum(acc[q], sh);
        if (threadIdx.x == 0) diag_part[q * gridDim.x + blockIdx.x] = s;
    }
}

// sums nslot rows of nb per-block partials in a fixed order (one workgroup)
__global__ void __launch_bounds__(1024) k_reduce_slots(const double *part, int nb, int nslot, double *out) {
    __shared__ double sh[16];
    for (int q = 0; q < nslot; q++) {
        double s = 0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(long)q * nb + b];
        s = block_sum(s, sh);
        if (threadIdx.x == 0) out[q] = s;
    }
}

// Scratch of pic_species_diag_noalloc: NDIAG_SLOTS rows of per-block partials for up to
// DIAG_NOALLOC_MAX_BLOCKS workgroups, then the NDIAG_SLOTS reduced sums (6.3 MB).  Each GPU
// holds its own copy for the life of the process.  Host code never passes `scratch` itself
// to a kernel: its device address comes from hipGetSymbolAddress, per device.
constexpr long DIAG_NOALLOC_MAX_BLOCKS = 131072;
constexpr long DIAG_NOALLOC_SUMS = NDIAG_SLOTS * DIAG_NOALLOC_MAX_BLOCKS;
static __device__ __align__(256) double scratch[DIAG_NOALLOC_SUMS + 8];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// device address of `scratch` on device dev, looked up once per device with dev current
static int scratch_on_device(int dev, double **out) {
    PIC_CHECK(dev >= 0 && dev < SCRATCH_MAX_DEVICES, PIC_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        PIC_HIP(hipGetDevice(&cur));
        if (cur != dev) PIC_HIP(hipSetDevice(dev));
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) PIC_HIP(hipSetDevice(cur));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        PIC_CHECK(p != nullptr, PIC_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return PIC_OK;
}

// ---------------------------------------------------------------------------------------
// counting sort: histogram -> plan (mode 1) -> scatter
// ---------------------------------------------------------------------------------------
// a block's share of the particles: its tile's two segments, or a slice of the staging list
struct Share {
    BinSeg bs;
    bool staging;
    long sbase;
    uint32_t cnt, first, step;
};

__device__ inline Share share_of(const Grid &g, const PartView &in) {
    Share s{};
    int b = blockIdx.x;
    if (b < g.ntiles) {
        s.bs = bin_of(in, b);
        s.staging = false;
        s.cnt = s.bs.c0 + s.bs.c1;
        s.first = threadIdx.x;
        s.step = blockDim.x;
    } else {
        s.staging = true;
        s.sbase = in.cap_bins;
        s.cnt = min(*in.stage_count, (uint32_t)in.cap_stage);
        s.first = (b - g.ntiles) * blockDim.x + threadIdx.x;
        s.step = (gridDim.x - g.ntiles) * blockDim.x;
    }
    return s;
}

__device__ inline long share_index(const Share &s, uint32_t q) {
    return s.staging ? s.sbase + q : seg_index(s.bs, q);
}

__device__ inline uint32_t share_trips(const Share &s) {
    uint32_t base = s.first - threadIdx.x;  // the same for every thread of the block
    return s.cnt > base ? (s.cnt - base + s.step - 1) / s.step : 0;
}

template <int PREC>
__global__ void __launch_bounds__(256) k_sort_hist(Grid g, PartView in, uint32_t *hist) {
    using PT = typename std::conditional<PREC == PIC_MIXED, PMixed, PF64>::type;
    Share s = share_of(g, in);
    uint32_t trips = share_trips(s);
    for (uint32_t it = 0; it < trips; it++) {
        uint32_t q = s.first + it * s.step;
        int key = -1;
        if (q < s.cnt) {
            PT p;
            load_p(in, share_index(s, q), p);
            key = tile_of(g, p);
        }
        wave_count_u32(key, hist);
    }
}

template <int PREC>
__global__ void __launch_bounds__(256) k_sort_scatter(Grid g, PartView in, PartView out, uint32_t *err) {
    using PT = typename std::conditional<PREC == PIC_MIXED, PMixed, PF64>::type;
    Share s = share_of(g, in);
    uint32_t trips = share_trips(s);
    for (uint32_t it = 0; it < trips; it++) {
        uint32_t q = s.first + it * s.step;
        int key = -1;
        PT p;
        if (q < s.cnt) {
            load_p(in, share_index(s, q), p);
            key = tile_of(g, p);
        }
        uint32_t pos = wave_reserve_global(key, out.bin_own);
        if (key >= 0) {
            if (pos < out.bin_off[key])
                store_p(out, (long)out.bin_start[key] + pos, p);
            else
                atomicOr(err, 4u);  // cannot happen: own segments are sized by the histogram
        }
    }
}

// ---------------------------------------------------------------------------------------
// uploads, downloads, injection
// ---------------------------------------------------------------------------------------
// Converts FP64 host-format particles (interleaved pos/vel) into the species format and
// appends them to the staging list.  rewind != 0 applies addParticle's half-step rewind
// v -= (q/m)*E(x)*(0.5*dt) with the current E (ch2/Species.cpp:55-71); positions were
// already bounds-checked on the host.
template <int PREC>
__global__ void __launch_bounds__(256)
    k_append(Grid g, long n, const double *pos, const double *vel, const double *mpw, const uint32_t *ids,
             uint32_t id0, int rewind, double qm, double half_dt, const float4 *ef32,
             const double4 *ef64, PartView dst, uint32_t stage0, double3 frame) {
#pragma clang fp contract(off)
    long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (q >= n) return;
    double x = pos[3 * q], y = pos[3 * q + 1], z = pos[3 * q + 2];
    double vx = vel[3 * q], vy = vel[3 * q + 1], vz = vel[3 * q + 2];
    double lx = (x - g.x0[0]) / g.dh[0], ly = (y - g.x0[1]) / g.dh[1], lz = (z - g.x0[2]) / g.dh[2];
    int i = min(max((int)lx, 0), g.nci - 1);
    int j = min(max((int)ly, 0), g.ncj - 1);
    int k = min(max((int)lz, 0), g.nck - 1);
    double di = lx - i, dj = ly - j, dk = lz - k;
    if (rewind) {
        long sj = g.ni, sk = (long)g.ni * g.nj;
        double3 E;
        if (PREC == PIC_FP64) {
            E = gather64_global(ef64, i + sj * j + sk * k, sj, sk, di, dj, dk);
        } else {
            float3 e = gather_global32(g, ef32, i, j, k, (float)di, (float)dj, (float)dk);
            E = make_double3(e.x, e.y, e.z);
        }
        vx -= (E.x * qm) * half_dt;
        vy -= (E.y * qm) * half_dt;
        vz -= (E.z * qm) * half_dt;
    }
    long idx = dst.cap_bins + stage0 + q;
    if (PREC == PIC_MIXED) {
        PMixed p;
        p.ci = i;
        p.cj = j;
        p.ck = k;
        float fo[3] = {(float)di, (float)dj, (float)dk};
        int *cc[3] = {&p.ci, &p.cj, &p.ck};
        int nc[3] = {g.nci, g.ncj, g.nck};
        for (int a = 0; a < 3; a++) {
            if (fo[a] >= 1.0f) {  // offset rounded up to one cell
                if (*cc[a] + 1 < nc[a]) {
                    fo[a] = 0.0f;
                    *cc[a] += 1;
                } else {
                    fo[a] = 0x1.fffffep-1f;
                }
            }
        }
        p.ox = fo[0];
        p.oy = fo[1];
        p.oz = fo[2];
        p.vx = (float)(vx - frame.x);
        p.vy = (float)(vy - frame.y);
        p.vz = (float)(vz - frame.z);
        p.w = (float)mpw[q];
        p.id = ids ? ids[q] : id0 + (uint32_t)q;
        store_p(dst, idx, p);
    } else {
        PF64 p{x, y, z, vx, vy, vz, mpw[q], ids ? ids[q] : id0 + (uint32_t)q};
        store_p(dst, idx, p);
    }
}

// Gathers bins + staging into contiguous FP64 arrays (device order).
template <int PREC>
__global__ void __launch_bounds__(256)
    k_gather_out(Grid g, PartView in, const uint32_t *offsets, double *pos, double *vel, double *mpw,
                 uint32_t *ids, double3 frame) {
    using PT = typename std::conditional<PREC == PIC_MIXED, PMixed, PF64>::type;
    Share s = share_of(g, in);
    uint32_t o = offsets[blockIdx.x < (unsigned)g.ntiles ? blockIdx.x : g.ntiles];
    for (uint32_t q = s.first; q < s.cnt; q += s.step) {
        PT p;
        load_p(in, share_index(s, q), p);
        long d = (long)o + q;
        if constexpr (PREC == PIC_MIXED) {
            pos[3 * d + 0] = g.x0[0] + ((double)p.ci + (double)p.ox) * g.dh[0];
            pos[3 * d + 1] = g.x0[1] + ((double)p.cj + (double)p.oy) * g.dh[1];
            pos[3 * d + 2] = g.x0[2] + ((double)p.ck + (double)p.oz) * g.dh[2];
        } else {
            pos[3 * d + 0] = p.x;
            pos[3 * d + 1] = p.y;
            pos[3 * d + 2] = p.z;
        }
        if constexpr (PREC == PIC_MIXED) {
            vel[3 * d + 0] = frame.x + (double)p.vx;
            vel[3 * d + 1] = frame.y + (double)p.vy;
            vel[3 * d + 2] = frame.z + (double)p.vz;
        } else {
            vel[3 * d + 0] = p.vx;
            vel[3 * d + 1] = p.vy;
            vel[3 * d + 2] = p.vz;
        }
        mpw[d] = p.w;
        if (ids) ids[d] = p.id;
    }
}

// PIC_MIXED records re-expressed in another frame: v - u_old -> v - u_new, du = u_old - u_new.
__global__ void __launch_bounds__(256) k_shift_frame(Grid g, PartView in, double3 du) {
    Share s = share_of(g, in);
    for (uint32_t q = s.first; q < s.cnt; q += s.step) {
        PMixed p;
        long idx = share_index(s, q);
        load_p(in, idx, p);
        p.vx = (float)((double)p.vx + du.x);
        p.vy = (float)((double)p.vy + du.y);
        p.vz = (float)((double)p.vz + du.z);
        store_p(in, idx, p);
    }
}

// Device beam / box source.  Particle q of this call uses Philox counter (q, ctr_lo, ctr_hi,
// stream) with key (seed); draws: 2 for the position on the face (or 3 in a box), then for a
// warm beam the isotropic direction (2) and the speed (9), as ch4/Species.cpp:149-174.
struct InjectParams {
    Grid g;
    long n;
    int box;         // 0: k=0 face (beam source), 1: box [x1,x2)
    double x1[3], x2[3];
    double v_drift;
    double vth;      // sqrt(2 k T / m) (0 for a cold beam)
    double w;
    double qm, half_dt;
    int rewind;
    uint32_t key0, key1, ctr_lo, ctr_hi;
    uint32_t id0;
    double frame[3];  // PIC_MIXED records hold v - frame
};

template <int PREC>
__global__ void __launch_bounds__(256)
    k_inject(InjectParams I, const float4 *ef32, const double4 *ef64, PartView dst, uint32_t stage0,
             long flat0, uint32_t *err) {
    // the rewind keeps addParticle's rounding (ch3/ver2/Species.cpp:75-79): no FMA
#pragma clang fp contract(off)
    long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (q >= I.n) return;
    const Grid &g = I.g;
    double u[16];
    for (int blk = 0; blk < 4; blk++) {
        U4 r = philox4x32_10(U4{(uint32_t)q, (uint32_t)(q >> 32) ^ I.ctr_lo, I.ctr_hi, (uint32_t)blk},
                             I.key0, I.key1);
        u[4 * blk + 0] = u01(r.x, r.y);
        u[4 * blk + 1] = u01(r.z, r.w);
        U4 r2 = philox4x32_10(U4{(uint32_t)q, (uint32_t)(q >> 32) ^ I.ctr_lo, I.ctr_hi, (uint32_t)blk + 4u},
                              I.key0, I.key1);
        u[4 * blk + 2] = u01(r2.x, r2.y);
        u[4 * blk + 3] = u01(r2.z, r2.w);
    }
    double x, y, z;
    if (I.box) {
        x = I.x1[0] + u[0] * (I.x2[0] - I.x1[0]);
        y = I.x1[1] + u[1] * (I.x2[1] - I.x1[1]);
        z = I.x1[2] + u[2] * (I.x2[2] - I.x1[2]);
    } else {
        x = g.x0[0] + u[0] * (g.dh[0] * (g.ni - 1));
        y = g.x0[1] + u[1] * (g.dh[1] * (g.nj - 1));
        z = g.x0[2];
    }
    double vx = 0, vy = 0, vz = I.v_drift;
    if (I.vth > 0) {
        double theta = 2 * PI_REF * u[3];
        double r = -1.0 + 2 * u[4];
        double a = sqrt(1 - r * r);
        double v1 = I.vth * (u[5] + u[6] + u[7] - 1.5);
        double v2 = I.vth * (u[8] + u[9] + u[10] - 1.5);
        double v3 = I.vth * (u[11] + u[12] + u[13] - 1.5);
        double vmag = 3 / sqrt(2.0 + 2 + 2) * sqrt(v1 * v1 + v2 * v2 + v3 * v3);
        vx = r * vmag;
        vy = cos(theta) * a * vmag;
        vz += sin(theta) * a * vmag;
    }
    // clamp into [x0,xm): a box upper bound equal to xm may round onto it
    double pp[3] = {x, y, z};
    int cc[3];
    double dd[3];
    int nc[3] = {g.nci, g.ncj, g.nck};
    for (int a = 0; a < 3; a++) {
        double l = (pp[a] - g.x0[a]) / g.dh[a];
        int c = (int)l;
        if (c < 0) c = 0;
        if (c >= nc[a]) c = nc[a] - 1;
        cc[a] = c;
        dd[a] = l - c;
        if (dd[a] >= 1.0) dd[a] = 0.9999999999999999;
    }
    if (I.rewind) {
        long sj = g.ni, sk = (long)g.ni * g.nj;
        double3 E;
        if (PREC == PIC_FP64) {
            E = gather64_global(ef64, cc[0] + sj * cc[1] + sk * cc[2], sj, sk, dd[0], dd[1], dd[2]);
        } else {
            float3 e = gather_global32(g, ef32, cc[0], cc[1], cc[2], (float)dd[0], (float)dd[1], (float)dd[2]);
            E = make_double3(e.x, e.y, e.z);
        }
        vx -= (E.x * I.qm) * I.half_dt;
        vy -= (E.y * I.qm) * I.half_dt;
        vz -= (E.z * I.qm) * I.half_dt;
    }
    long idx = flat0 >= 0 ? flat0 + q : dst.cap_bins + stage0 + q;
    if (PREC == PIC_MIXED) {
        PMixed p;
        p.ci = cc[0];
        p.cj = cc[1];
        p.ck = cc[2];
        float fo[3];
        for (int a = 0; a < 3; a++) {
            fo[a] = (float)dd[a];
            if (fo[a] >= 1.0f) fo[a] = 0x1.fffffep-1f;
        }
        p.ox = fo[0];
        p.oy = fo[1];
        p.oz = fo[2];
        p.vx = (float)(vx - I.frame[0]);
        p.vy = (float)(vy - I.frame[1]);
        p.vz = (float)(vz - I.frame[2]);
        p.w = (float)I.w;
        p.id = I.id0 + (uint32_t)q;
        store_p(dst, idx, p);
    } else {
        // keep the sampled position; one that rounded onto an upper wall moves just inside
        PF64 p{x, y, z, vx, vy, vz, I.w, I.id0 + (uint32_t)q};
        if (p.x >= g.xm[0]) p.x = nextafter(g.xm[0], g.x0[0]);
        if (p.y >= g.xm[1]) p.y = nextafter(g.xm[1], g.x0[1]);
        if (p.z >= g.xm[2]) p.z = nextafter(g.xm[2], g.x0[2]);
        store_p(dst, idx, p);
    }
}

__global__ void k_avg(long nn, double *ave, const double *inst, int samples) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= nn) return;
    // Field_::updateAverage (ch3/ver2/Field.h:214-220)
    ave[u] = (inst[u] + samples * ave[u]) / (samples + 1);
}

__global__ void k_fill_u8(long n, uint8_t *p, uint8_t v) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u < n) p[u] = v;
}

__global__ void k_fill_f64(long n, double *p, double v) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u < n) p[u] = v;
}

__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }

// Snapshot of a push, in stream order right after it: the live count (sum of the push's
// per-workgroup count partials), the staging count, and the capacity flags raised since the
// last snapshot (taken and cleared in one atomic, so none is lost to a later push).
__global__ void __launch_bounds__(256)
    k_push_status(const double *cnt_part, int nb, const uint32_t *stage_count, long cap_stage, SpStatus *st) {
    __shared__ double sh[4];
    double a = 0;
    for (int i = threadIdx.x; i < nb; i += 256) a += cnt_part[i];
    a = block_sum(a, sh);
    if (threadIdx.x == 0) {
        st->np = (long long)(a + 0.5);
        const uint32_t sc = *stage_count;
        st->stage = (long)sc < cap_stage ? sc : (uint32_t)cap_stage;
        st->err_snap = atomicExch(&st->err, 0u);
    }
}

inline unsigned nb256(long n) { return (unsigned)((n + 255) / 256); }

// =========================================================================================
// host side
// =========================================================================================
static size_t part_bytes(int prec, bool ids) { return (prec == PIC_MIXED ? 32 : 56) + (ids ? 4 : 0); }

static int alloc_buf(pic_species *s, PartBuf &b, long cap_bins, long cap_stage) {
    const Grid &g = s->w->g;
    long n = cap_bins + cap_stage;
    size_t esz = s->prec == PIC_MIXED ? 4 : 8;
    size_t bytes = 0;
    size_t arr = (size_t)n * esz;
    arr = (arr + 255) & ~(size_t)255;
    bytes += arr * 7;
    if (s->prec == PIC_MIXED) bytes += arr;  // 8 x 4 B per particle in one record array
    size_t idb = s->track_ids ? (((size_t)n * 4 + 255) & ~(size_t)255) : 0;
    bytes += idb;
    size_t binb = (((size_t)g.ntiles * 4 + 255) & ~(size_t)255);
    bytes += binb * 5 + 256;
    const size_t markb = ((size_t)n + 255) & ~(size_t)255;
    bytes += markb;
    void *mem = nullptr;
    PIC_HIP(hipMalloc(&mem, bytes));
    b = PartBuf{};
    b.mem = mem;
    b.cap_bins = cap_bins;
    b.cap_stage = cap_stage;
    char *p = (char *)mem;
    auto take = [&](size_t sz) {
        char *r = p;
        p += sz;
        return r;
    };
    if (s->prec == PIC_MIXED) {
        b.rec = (float4 *)take(arr * 8);  // 32-byte records
    } else {
        b.x = (double *)take(arr);
        b.y = (double *)take(arr);
        b.z = (double *)take(arr);
        b.dvx = (double *)take(arr);
        b.dvy = (double *)take(arr);
        b.dvz = (double *)take(arr);
        b.dw = (double *)take(arr);
    }
    if (s->track_ids) b.id = (uint32_t *)take(idb);
    b.bin_start = (uint32_t *)take(binb);
    b.bin_cap = (uint32_t *)take(binb);
    b.bin_off = (uint32_t *)take(binb);
    b.bin_own = (uint32_t *)take(binb);
    b.bin_in = (uint32_t *)take(binb);
    b.stage_count = (uint32_t *)take(256);
    b.mark = (uint8_t *)take(markb);
    // empty bins, empty staging, no marks
    PIC_HIP(hipMemsetAsync(b.bin_start, 0, binb * 5 + 256 + markb, s->w->stream));
    return PIC_OK;
}

static void free_buf(PartBuf &b) {
    if (b.mem) hipFree(b.mem);
    b = PartBuf{};
}

static long default_stage_cap(long cap_bins) { return std::max<long>(65536, cap_bins / 16); }

// Which tiles the three-cell-rim push takes: none (the default; -1), those flagged by their
// far deposits (PIC_PUSH_WIDE=auto; 0) or all (PIC_PUSH_WIDE=all; 1).  With the far deposits
// deferred to k_far_arrivals, the two-cell rim alone measured faster late in the sphere run
// (13.6 against 14.4-14.9 ms per push at steps 81-100, same early): the wide window's
// occupancy (4 workgroups per CU against 6) costs more than the far deposits it saves.
static int wide_policy() {
    const char *fw = getenv("PIC_PUSH_WIDE");
    if (!fw) return -1;
    if (!strcmp(fw, "all")) return 1;
    if (!strcmp(fw, "auto")) return 0;
    return -1;
}

static PlanTmp plan_tmp(pic_species *s) {
    const Grid &g = s->w->g;
    size_t nt = (size_t)g.ntiles;
    char *p = (char *)s->plan_mem;
    PlanTmp T;
    size_t nblk = (nt + PLAN_THREADS - 1) / PLAN_THREADS;
    T.need = (uint32_t *)p;
    T.slack = (uint32_t *)(p + 4 * nt);
    T.nchunk = (uint32_t *)(p + 8 * nt);
    T.staged = (uint32_t *)(p + 12 * nt);
    T.alpha = (double *)(p + (((12 + 4 * (size_t)PLAN_STAGED_SHARDS) * nt + 7) & ~(size_t)7));
    T.total = (unsigned long long *)((char *)T.alpha + 16);
    T.blk = T.total + 1;
    T.blk2 = T.blk + 2 * nblk;
    return T;
}

static size_t plan_bytes(const Grid &g) {
    size_t nblk = ((size_t)g.ntiles + PLAN_THREADS - 1) / PLAN_THREADS;
    return (((12 + 4 * (size_t)PLAN_STAGED_SHARDS) * (size_t)g.ntiles + 7) & ~(size_t)7) + 16 + 8 + 24 * nblk;
}

// push workgroups over the tiles: at most one per tile plus one per chunk of bin slots,
// padded to whole rounds of XCD stripes
static long push_bin_groups(const pic_species *s, long cap_bins) {
    const Grid &g = s->w->g;
    const long r = 8L * s->xcd_stripe;
    long n = (long)g.ntiles + cap_bins / s->push_chunk + 1;
    return (n + r - 1) / r * r;
}

// chunk map and per-workgroup diagnostic partials sized for the buffers' bin capacity
static int ensure_push_scratch(pic_species *s) {
    const Grid &g = s->w->g;
    long nb = push_bin_groups(s, s->buf[1 - s->cur].cap_bins);
    if (nb > s->chunk_cap) {
        if (s->chunk_map) PIC_HIP(hipFree(s->chunk_map));
        s->chunk_map = nullptr;
        PIC_HIP(hipMalloc(&s->chunk_map, sizeof(uint64_t) * (size_t)nb));
        s->chunk_cap = nb;
    }
    long nd = std::max<long>(nb + PUSH_STAGE_WGS, g.ntiles + STAGE_WGS);
    if (nd > s->diag_cap) {
        if (s->diag_part) PIC_HIP(hipFree(s->diag_part));
        s->diag_part = nullptr;
        PIC_HIP(hipMalloc(&s->diag_part, sizeof(double) * NDIAG_SLOTS * (size_t)nd));
        s->diag_cap = nd;
    }
    return PIC_OK;
}

// segment plan of out from the counts of in (mode 0) or from the histogram hist (mode 1)
static int launch_plan(pic_species *s, const PartView &in, const uint32_t *hist, int mode, const PartBuf &out,
                       uint32_t *err) {
    pic_world *w = s->w;
    const Grid &g = w->g;
    PlanTmp T = plan_tmp(s);
    int nblk = (g.ntiles + PLAN_THREADS - 1) / PLAN_THREADS;
    PartView vo = view_of(out);
    if (mode == 0) {
        if (s->prec == PIC_MIXED)
            k_plan_staged<PIC_MIXED><<<PLAN_STAGED_WGS, 256, 0, w->stream>>>(g, in, T.staged);
        else
            k_plan_staged<PIC_FP64><<<PLAN_STAGED_WGS, 256, 0, w->stream>>>(g, in, T.staged);
    }
    k_plan_count<<<nblk, PLAN_THREADS, 0, w->stream>>>(g, in, hist, mode, T);
    k_plan_total<<<1, PLAN_SCAN_THREADS, 0, w->stream>>>(T, nblk, (unsigned long long)out.cap_bins, err);
    k_plan_caps<<<nblk, PLAN_THREADS, 0, w->stream>>>(g, T, vo);
    k_plan_scan<<<1, PLAN_SCAN_THREADS, 0, w->stream>>>(T.blk, nblk, nullptr);
    k_plan_starts<<<nblk, PLAN_THREADS, 0, w->stream>>>(g, T, vo);
    if (mode == 0) {  // push workgroups per tile
        const int force = wide_policy();
        k_plan_chunks<<<nblk, PLAN_THREADS, 0, w->stream>>>(g, T, s->far_tile, s->far_tile + g.ntiles, force,
                                                            (uint32_t)s->push_chunk);
        k_plan_scan<<<1, PLAN_SCAN_THREADS, 0, w->stream>>>(T.blk2, nblk, T.total);
        k_plan_map<<<nblk, PLAN_THREADS, 0, w->stream>>>(g, T, s->far_tile + g.ntiles, s->chunk_map, s->chunk_cap);
    }
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

static int launch_plan_push(pic_species *s, uint32_t *err) {
    PartBuf &in = s->buf[s->cur], &out = s->buf[1 - s->cur];
    return launch_plan(s, view_of(in), nullptr, 0, out, err);
}

// capacity flags of this species' device operations (1: bins short, 2: particle dropped,
// 4: sort overflow)
static uint32_t *err_flag(pic_species *s) { return &s->st_dev->err; }

// Reads the snapshot of the last push once its copy has landed (wait: blocks on its event
// only, not on the stream).  Tightens np_upper to the pushed count plus what was appended
// since, and learns the staging count; PIC_ERR_CAPACITY when that push dropped particles.
static int poll_status(pic_species *s, bool wait) {
    if (!s->st_pending) return PIC_OK;
    if (wait) {
        PIC_HIP(hipEventSynchronize(s->st_ev));
    } else {
        hipError_t q = hipEventQuery(s->st_ev);
        if (q == hipErrorNotReady) return PIC_OK;
        if (q != hipSuccess) return hip_fail(q, "hipEventQuery", __FILE__, __LINE__);
    }
    s->st_pending = false;
    const SpStatus &st = *s->st_host;
    s->np_upper = (long)st.np + s->added_since_push;
    if (!s->stage_known) {
        s->stage_host = (long)st.stage + s->added_since_push;
        s->stage_known = true;
    }
    if (st.err_snap & 6u) {  // 2: staging full in a push or injection, 4: sort overflow
        s->np_exact = false;  // the snapshot counted the dropped particles
        set_error("the particle staging list overflowed in advance(): particles were dropped (capacity)");
        return PIC_ERR_CAPACITY;
    }
    return PIC_OK;
}

static int check_err(pic_species *s) {
    int rc = poll_status(s, true);
    if (rc) return rc;
    uint32_t e = 0;
    PIC_HIP(hipMemcpyAsync(&s->w->scal_host->counter, err_flag(s), 4, hipMemcpyDeviceToHost, s->w->stream));
    PIC_HIP(hipStreamSynchronize(s->w->stream));
    e = s->w->scal_host->counter;
    if (e) {
        k_set_u32<<<1, 1, 0, s->w->stream>>>(err_flag(s), 0u);
        set_error(e & 2 ? "particle staging list overflowed (capacity)"
                        : "particle buffers too small for the particle count (capacity)");
        return PIC_ERR_CAPACITY;
    }
    return PIC_OK;
}

static int do_sort(pic_species *s);

// grows both buffers so that cap_bins >= need (keeps the particles)
static int ensure_capacity(pic_species *s, long need_bins, long need_stage) {
    PartBuf &cur = s->buf[s->cur];
    const Grid &g = s->w->g;
    long want_bins = cur.cap_bins, want_stage = cur.cap_stage;
    long slack_min = (long)g.ntiles * 48;
    if (need_bins + need_bins / 4 + slack_min > want_bins) want_bins = need_bins + need_bins / 2 + slack_min;
    if (need_stage > want_stage) want_stage = std::max(need_stage + need_stage / 2, default_stage_cap(want_bins));
    if (want_bins == cur.cap_bins && want_stage == cur.cap_stage) return PIC_OK;
    PIC_CHECK(want_bins + want_stage < 4000000000L, PIC_ERR_CAPACITY, "more than 4e9 particle slots");
    // new buffers: sort from the current layout into the new one
    PartBuf nb, nb2;
    int rc = alloc_buf(s, nb, want_bins, want_stage);
    if (rc) return rc;
    rc = alloc_buf(s, nb2, want_bins, want_stage);
    if (rc) {
        free_buf(nb);
        return rc;
    }
    PartBuf old_other = s->buf[1 - s->cur];
    s->buf[1 - s->cur] = nb;  // sort target
    rc = do_sort(s);          // moves cur -> nb and flips cur
    // old current buffer is now the "other": replace it
    PartBuf old_cur = s->buf[1 - s->cur];
    s->buf[1 - s->cur] = nb2;
    PIC_HIP(hipStreamSynchronize(s->w->stream));
    free_buf(old_cur);
    free_buf(old_other);
    return rc;
}

template <int PREC>
static int sort_launch(pic_species *s) {
    pic_world *w = s->w;
    const Grid &g = w->g;
    PartBuf &in = s->buf[s->cur], &out = s->buf[1 - s->cur];
    PartView vin = view_of(in), vout = view_of(out);
    hipEvent_t ev;
    timer_begin(w, PIC_T_SORT, &ev);
    // histogram into out.bin_in (zeroed); the plan sizes the own segments from it (its
    // first kernel copies the counts) and its last kernel zeroes the counters
    PIC_HIP(hipMemsetAsync(out.bin_in, 0, sizeof(uint32_t) * g.ntiles, w->stream));
    unsigned grid = g.ntiles + STAGE_WGS * 4;
    k_sort_hist<PREC><<<grid, 256, 0, w->stream>>>(g, vin, out.bin_in);
    PIC_LAUNCH_CHECK();
    {
        int rc = launch_plan(s, vin, out.bin_in, 1, out, err_flag(s));
        if (rc) return rc;
    }
    k_sort_scatter<PREC><<<grid, 256, 0, w->stream>>>(g, vin, vout, err_flag(s));
    PIC_LAUNCH_CHECK();
    timer_end(w, PIC_T_SORT, ev);
    s->cur = 1 - s->cur;
    s->sorted = true;
    s->stage_added = 0;
    s->stage_host = 0;  // everything is in the bins now
    s->stage_known = true;
    return PIC_OK;
}

static int do_sort(pic_species *s) {
    return s->prec == PIC_MIXED ? sort_launch<PIC_MIXED>(s) : sort_launch<PIC_FP64>(s);
}

// host-side count of particles in the current buffer (bins + staging); syncs
static int count_particles(pic_species *s, std::vector<uint32_t> &cnt, uint32_t &stage) {
    const Grid &g = s->w->g;
    PartBuf &b = s->buf[s->cur];
    size_t nt = (size_t)g.ntiles;
    std::vector<uint32_t> cap(nt), off(nt), own(nt), inn(nt);
    cnt.resize(nt);
    PIC_HIP(hipMemcpyAsync(cap.data(), b.bin_cap, 4 * nt, hipMemcpyDeviceToHost, s->w->stream));
    PIC_HIP(hipMemcpyAsync(off.data(), b.bin_off, 4 * nt, hipMemcpyDeviceToHost, s->w->stream));
    PIC_HIP(hipMemcpyAsync(own.data(), b.bin_own, 4 * nt, hipMemcpyDeviceToHost, s->w->stream));
    PIC_HIP(hipMemcpyAsync(inn.data(), b.bin_in, 4 * nt, hipMemcpyDeviceToHost, s->w->stream));
    PIC_HIP(hipMemcpyAsync(&s->w->scal_host->iters, b.stage_count, 4, hipMemcpyDeviceToHost, s->w->stream));
    PIC_HIP(hipStreamSynchronize(s->w->stream));
    for (size_t t = 0; t < nt; t++) {
        uint32_t c0 = std::min(own[t], off[t]);
        uint32_t room = cap[t] > off[t] ? cap[t] - off[t] : 0u;
        cnt[t] = c0 + std::min(inn[t], room);
    }
    stage = std::min<uint32_t>((uint32_t)s->w->scal_host->iters, (uint32_t)b.cap_stage);
    return PIC_OK;
}

}  // namespace pic

using namespace pic;

extern "C" {

int pic_species_create(pic_world *w, double mass, double charge, double mpw0, int precision,
                       long capacity, int track_ids, pic_species **out) {
    PIC_CHECK(w && out, PIC_ERR_ARG, "null argument");
    PIC_CHECK(precision == PIC_MIXED || precision == PIC_FP64, PIC_ERR_ARG, "bad precision");
    PIC_CHECK(mass > 0, PIC_ERR_ARG, "mass must be positive");
    PIC_CHECK(w->extents_set, PIC_ERR_STATE, "setExtents must precede species creation");
    *out = nullptr;
    pic_species *s = new pic_species();
    s->w = w;
    s->mass = mass;
    s->charge = charge;
    s->mpw0 = mpw0;
    s->prec = precision;
    // launch shapes of the push for this grid and particle count (tuning table, else defaults)
    s->push_chunk = std::max<long>(256, (long)tuned("push", (double)w->g.nn, (double)std::max<long>(capacity, 1),
                                                      "chunk", (double)PUSH_CHUNK));
    s->xcd_stripe = std::max(1, (int)tuned("push", (double)w->g.nn, (double)std::max<long>(capacity, 1), "stripe",
                                           (double)PUSH_XCD_STRIPE));
    s->track_ids = track_ids != 0;
    const Grid &g = w->g;
    long cap = std::max<long>(capacity, 1024);
    long cap_bins = cap + (cap * 3) / 10 + (long)g.ntiles * 64;
    long cap_stage = default_stage_cap(cap_bins);
    int rc = PIC_OK;
    size_t wbytes = (size_t)g.ntiles * WNODES * (precision == PIC_MIXED ? 4 : 8);
    if ((rc = alloc_buf(s, s->buf[0], cap_bins, cap_stage))) goto fail;
    if ((rc = alloc_buf(s, s->buf[1], cap_bins, cap_stage))) goto fail;
#define TRY(x)                                                 \
    do {                                                       \
        hipError_t e_ = (x);                                   \
        if (e_ != hipSuccess) {                                \
            rc = hip_fail(e_, #x, __FILE__, __LINE__);         \
            goto fail;                                         \
        }                                                      \
    } while (0)
    TRY(hipMalloc(&s->st_dev, sizeof(SpStatus)));
    TRY(hipHostMalloc(&s->st_host, sizeof(SpStatus)));
    TRY(hipEventCreateWithFlags(&s->st_ev, hipEventDisableTiming));
    TRY(hipMalloc(&s->den, sizeof(double) * g.nn));
    TRY(hipMalloc(&s->den_extra, sizeof(double) * g.nn));
    TRY(hipMalloc(&s->den_ave, sizeof(double) * g.nn));
    TRY(hipMalloc(&s->win, wbytes));
    TRY(hipMalloc(&s->win_used, g.ntiles));
    TRY(hipMalloc(&s->far_tile, sizeof(uint32_t) * 3 * (size_t)g.ntiles));
    TRY(hipMalloc(&s->plan_mem, plan_bytes(g)));
    TRY(hipMalloc(&s->diag_dev, sizeof(double) * 8));
    if ((rc = ensure_push_scratch(s))) goto fail;
    TRY(hipHostMalloc(&s->diag_host, sizeof(double) * 8));
    TRY(hipMemsetAsync(s->st_dev, 0, sizeof(SpStatus), w->stream));
    TRY(hipMemsetAsync(s->den, 0, sizeof(double) * g.nn, w->stream));
    TRY(hipMemsetAsync(s->den_extra, 0, sizeof(double) * g.nn, w->stream));
    TRY(hipMemsetAsync(s->far_tile, 0, sizeof(uint32_t) * 3 * (size_t)g.ntiles, w->stream));
    TRY(hipMemsetAsync(s->den_ave, 0, sizeof(double) * g.nn, w->stream));
    TRY(hipMemsetAsync(s->win_used, 0, g.ntiles, w->stream));
    TRY(hipMemsetAsync(s->plan_mem, 0, plan_bytes(g), w->stream));  // staged counts start at 0
    TRY(hipMemsetAsync(s->diag_dev, 0, sizeof(double) * 8, w->stream));
    TRY(hipStreamSynchronize(w->stream));
#undef TRY
    *s->st_host = SpStatus{};
    s->den_state = 2;  // no particles: den = 0 is current
    s->sorted = true;
    *out = s;
    return PIC_OK;
fail:
    pic_species_destroy(s);
    return rc;
}

int pic_species_destroy(pic_species *s) {
    if (!s) return PIC_OK;
    if (s->w && s->w->stream) hipStreamSynchronize(s->w->stream);
    free_buf(s->buf[0]);
    free_buf(s->buf[1]);
    hipFree(s->st_dev);
    if (s->st_host) hipHostFree(s->st_host);
    if (s->st_ev) hipEventDestroy(s->st_ev);
    hipFree(s->den);
    hipFree(s->den_extra);
    hipFree(s->den_ave);
    hipFree(s->win);
    hipFree(s->win_used);
    hipFree(s->far_tile);
    hipFree(s->plan_mem);
    hipFree(s->diag_dev);
    hipFree(s->diag_part);
    hipFree(s->chunk_map);
    if (s->diag_host) hipHostFree(s->diag_host);
    delete s;
    return PIC_OK;
}

static int append_host(pic_species *s, long n, const double *pos, const double *vel, const double *mpw,
                       const uint32_t *ids, int rewind, double dt) {
    pic_world *w = s->w;
    if (n == 0) return PIC_OK;
    // staging must hold the new particles on top of what it already holds
    std::vector<uint32_t> cnt;
    uint32_t stage = 0;
    int rc = count_particles(s, cnt, stage);
    if (rc) return rc;
    long inbins = 0;
    for (auto c : cnt) inbins += c;
    rc = ensure_capacity(s, inbins + stage + n, stage + n);
    if (rc) return rc;
    rc = count_particles(s, cnt, stage);  // a regrow sorts staging into the bins
    if (rc) return rc;
    inbins = 0;
    for (auto c : cnt) inbins += c;
    if (s->buf[s->cur].cap_stage < stage + n) {
        // staging still short after growth: sort what is there into the bins first
        rc = do_sort(s);
        if (rc) return rc;
        inbins += stage;
        stage = 0;
    }
    PartBuf &b = s->buf[s->cur];
    double *d = nullptr;
    size_t bytes = sizeof(double) * (size_t)n * 7 + (ids ? 4 * (size_t)n : 0);
    PIC_HIP(hipMalloc(&d, bytes));
    double *dpos = d, *dvel = d + 3 * n, *dmpw = d + 6 * n;
    uint32_t *dids = ids ? (uint32_t *)(d + 7 * n) : nullptr;
    PIC_HIP(hipMemcpyAsync(dpos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, w->stream));
    PIC_HIP(hipMemcpyAsync(dvel, vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice, w->stream));
    PIC_HIP(hipMemcpyAsync(dmpw, mpw, sizeof(double) * n, hipMemcpyHostToDevice, w->stream));
    if (ids) PIC_HIP(hipMemcpyAsync(dids, ids, 4 * (size_t)n, hipMemcpyHostToDevice, w->stream));
    double qm = s->charge / s->mass;
    if (s->prec == PIC_MIXED)
        k_append<PIC_MIXED><<<nb256(n), 256, 0, w->stream>>>(w->g, n, dpos, dvel, dmpw, dids, s->next_id, rewind,
                                                             qm, 0.5 * dt, w->ef32, w->ef64, view_of(b), stage,
                                                             frame_of(s));
    else
        k_append<PIC_FP64><<<nb256(n), 256, 0, w->stream>>>(w->g, n, dpos, dvel, dmpw, dids, s->next_id, rewind,
                                                            qm, 0.5 * dt, w->ef32, w->ef64, view_of(b), stage,
                                                            frame_of(s));
    PIC_LAUNCH_CHECK();
    k_set_u32<<<1, 1, 0, w->stream>>>(b.stage_count, stage + (uint32_t)n);
    PIC_HIP(hipStreamSynchronize(w->stream));
    PIC_HIP(hipFree(d));
    if (!ids) s->next_id += (uint32_t)n;
    // count_particles synchronised the stream: any push snapshot has landed and is stale now
    s->st_pending = false;
    s->np_upper = inbins + stage + n;
    s->added_since_push = 0;
    s->stage_host = stage + n;
    s->stage_known = true;
    s->stage_added += n;
    s->sorted = false;
    s->den_state = 0;
    s->diag_valid = false;
    return PIC_OK;
}

int pic_species_add_particles(pic_species *s, long n, const double *pos, const double *vel,
                              const double *mpw, double dt, long *n_added) {
    PIC_CHECK(s && (n == 0 || (pos && vel && mpw)), PIC_ERR_ARG, "null argument");
    const Grid &g = s->w->g;
    // World::inBounds (ch2/World.h:59-63) filter on the host, keeping call order
    std::vector<double> P, V, M;
    P.reserve(3 * n);
    V.reserve(3 * n);
    M.reserve(n);
    for (long q = 0; q < n; q++) {
        const double *x = pos + 3 * q;
        bool in = true;
        for (int c = 0; c < 3; c++)
            if (x[c] < g.x0[c] || x[c] >= g.xm[c]) in = false;
        if (!in) continue;
        for (int c = 0; c < 3; c++) {
            P.push_back(x[c]);
            V.push_back(vel[3 * q + c]);
        }
        M.push_back(mpw[q]);
    }
    long k = (long)M.size();
    if (n_added) *n_added = k;
    return append_host(s, k, P.data(), V.data(), M.data(), nullptr, 1, dt);
}

int pic_species_set_particles(pic_species *s, long n, const double *pos, const double *vel,
                              const double *mpw, const uint32_t *ids) {
    PIC_CHECK(s && (n == 0 || (pos && vel && mpw)), PIC_ERR_ARG, "null argument");
    pic_world *w = s->w;
    // drop everything: empty bins and staging
    for (int b = 0; b < 2; b++) {
        PartBuf &pb = s->buf[b];
        PIC_HIP(hipMemsetAsync(pb.bin_own, 0, 4 * (size_t)w->g.ntiles, w->stream));
        PIC_HIP(hipMemsetAsync(pb.bin_in, 0, 4 * (size_t)w->g.ntiles, w->stream));
        PIC_HIP(hipMemsetAsync(pb.bin_cap, 0, 4 * (size_t)w->g.ntiles, w->stream));
        PIC_HIP(hipMemsetAsync(pb.bin_off, 0, 4 * (size_t)w->g.ntiles, w->stream));
        PIC_HIP(hipMemsetAsync(pb.stage_count, 0, 4, w->stream));
    }
    s->next_id = 0;
    s->np_upper = 0;
    s->st_pending = false;
    s->added_since_push = 0;
    s->stage_host = 0;
    s->stage_known = true;
    if (n == 0) {
        s->den_state = 0;
        s->diag_valid = false;
        return PIC_OK;
    }
    int rc = append_host(s, n, pos, vel, mpw, ids, 0, 0.0);
    if (rc) return rc;
    return do_sort(s);
}

int pic_species_get_np(pic_species *s, long *np) {
    PIC_CHECK(s && np, PIC_ERR_ARG, "null argument");
    // the last push's snapshot (waited for on its event) plus what was appended since is
    // exact unless that push dropped particles, which poll_status reports
    int rc0 = poll_status(s, true);
    if (rc0) return rc0;
    if (s->np_exact) {
        *np = s->np_upper;
        return PIC_OK;
    }
    std::vector<uint32_t> cnt;
    uint32_t stage = 0;
    int rc = count_particles(s, cnt, stage);
    if (rc) return rc;
    long tot = stage;
    for (auto c : cnt) tot += c;
    *np = tot;
    s->np_upper = tot;
    s->np_exact = true;
    s->added_since_push = 0;
    return check_err(s);
}

int pic_species_get_particles(pic_species *s, double *pos, double *vel, double *mpw, uint32_t *ids) {
    PIC_CHECK(s && pos && vel && mpw, PIC_ERR_ARG, "null argument");
    PIC_CHECK(!ids || s->track_ids, PIC_ERR_ARG, "species does not track ids");
    pic_world *w = s->w;
    const Grid &g = w->g;
    std::vector<uint32_t> cnt;
    uint32_t stage = 0;
    int rc = count_particles(s, cnt, stage);
    if (rc) return rc;
    std::vector<uint32_t> off(g.ntiles + 1);
    long run = 0;
    for (int t = 0; t < g.ntiles; t++) {
        off[t] = (uint32_t)run;
        run += cnt[t];
    }
    off[g.ntiles] = (uint32_t)run;
    long n = run + stage;
    if (n == 0) return PIC_OK;
    uint32_t *doff = nullptr;
    double *d = nullptr;
    PIC_HIP(hipMalloc(&doff, 4 * (size_t)(g.ntiles + 1)));
    PIC_HIP(hipMalloc(&d, sizeof(double) * 7 * (size_t)n + 4 * (size_t)n));
    PIC_HIP(hipMemcpyAsync(doff, off.data(), 4 * (size_t)(g.ntiles + 1), hipMemcpyHostToDevice, w->stream));
    double *dpos = d, *dvel = d + 3 * n, *dmpw = d + 6 * n;
    uint32_t *dids = ids ? (uint32_t *)(d + 7 * n) : nullptr;
    unsigned grid = g.ntiles + STAGE_WGS;
    PartView v = view_of(s->buf[s->cur]);
    if (s->prec == PIC_MIXED)
        k_gather_out<PIC_MIXED><<<grid, 256, 0, w->stream>>>(g, v, doff, dpos, dvel, dmpw, dids, frame_of(s));
    else
        k_gather_out<PIC_FP64><<<grid, 256, 0, w->stream>>>(g, v, doff, dpos, dvel, dmpw, dids, frame_of(s));
    PIC_LAUNCH_CHECK();
    PIC_HIP(hipMemcpyAsync(pos, dpos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, w->stream));
    PIC_HIP(hipMemcpyAsync(vel, dvel, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, w->stream));
    PIC_HIP(hipMemcpyAsync(mpw, dmpw, sizeof(double) * n, hipMemcpyDeviceToHost, w->stream));
    if (ids) PIC_HIP(hipMemcpyAsync(ids, dids, 4 * (size_t)n, hipMemcpyDeviceToHost, w->stream));
    PIC_HIP(hipStreamSynchronize(w->stream));
    PIC_HIP(hipFree(d));
    PIC_HIP(hipFree(doff));
    return PIC_OK;
}

int pic_species_sort(pic_species *s) {
    PIC_CHECK(s, PIC_ERR_ARG, "null argument");
    int rc = do_sort(s);
    if (rc) return rc;
    return check_err(s);
}

int pic_species_advance(pic_species *s, int bc, double dt) {
    PIC_CHECK(s, PIC_ERR_ARG, "null argument");
    PIC_CHECK(bc == PIC_BC_REFLECT || bc == PIC_BC_ABSORB, PIC_ERR_ARG, "bad boundary");
    pic_world *w = s->w;
    const Grid &g = w->g;
    // the previous push's snapshot: particles it dropped are reported here, one call later
    // (the wait is on that push's event, which the solve between two pushes has passed)
    int rc = poll_status(s, true);
    if (rc) return rc;
    // the push uses whatever E the world holds (computeEF or a host write), as the
    // reference's advance() reads world.ef as it stands
    // room for one more step: output bins need the current population plus slack
    long up = s->np_upper;
    PartBuf &in0 = s->buf[s->cur];
    if (up + up / 4 + (long)g.ntiles * 48 > in0.cap_bins) {
        rc = ensure_capacity(s, up, 0);
        if (rc) return rc;
    }
    // a large unbinned list goes through the sort rather than the slow global path
    if (!s->sorted && s->stage_added > 65536 && s->stage_added * 16 > up) {
        rc = do_sort(s);
        if (rc) return rc;
    }
    PartBuf &in = s->buf[s->cur], &out = s->buf[1 - s->cur];
    rc = ensure_push_scratch(s);
    if (rc) return rc;
    // far deposits that no density finish consumed (a push or deposit pass not followed by
    // computeNumberDensity) must not add to this push's density
    if (s->extra_dirty) {
        k_fill_f64<<<nb256(g.nn), 256, 0, w->stream>>>(g.nn, s->den_extra, 0.0);
        PIC_LAUNCH_CHECK();
        s->extra_dirty = false;
    }
    // the push phase: segment plan of the output buffer, then the fused push
    hipEvent_t ev;
    timer_begin(w, PIC_T_PUSH, &ev);
    rc = launch_plan_push(s, err_flag(s));
    if (rc) return rc;
    PushParams P{};
    P.g = g;
    P.bc = bc;
    P.sphere_on = (bc == PIC_BC_ABSORB) && w->sphere.on;
    P.qm_dt = (float)(dt * s->charge / s->mass);
    P.s64 = dt * s->charge / s->mass;
    P.dt = dt;
    for (int c = 0; c < 3; c++) {
        P.dt_dh[c] = (float)(dt / g.dh[c]);
        P.dhf[c] = (float)g.dh[c];
        double Ls = (w->sphere.c[c] - g.x0[c]) / g.dh[c];
        double fl = floor(Ls);
        P.sph_c[c] = (int)fl;
        P.sph_f[c] = (float)(Ls - fl);
        P.sph_c64[c] = w->sphere.c[c];
    }
    P.r2f = (float)w->sphere.r2;
    P.sph_r2 = w->sphere.r2;
    for (int c = 0; c < 3; c++) {
        double u = s->frame[c] * dt / g.dh[c];
        P.fdh_hi[c] = (float)u;
        P.fdh_lo[c] = (float)(u - (double)P.fdh_hi[c]);
        P.fr2[c] = (float)(2 * s->frame[c]);
    }
    PushOut O{};
    O.chunk_map = s->chunk_map;
    O.total_groups = plan_tmp(s).total;
    O.win32 = (float *)s->win;
    O.win64 = (double *)s->win;
    O.win_used = s->win_used;
    O.den_extra = s->den_extra;
    O.diag_part = s->diag_part;
    O.err = err_flag(s);
    O.far = s->far_tile;
    O.far_in = s->far_tile + 2 * (size_t)g.ntiles;
    O.chunk = (uint32_t)s->push_chunk;
    O.stripe = (uint32_t)s->xcd_stripe;
    // workgroups for the tiles (the plan's map decides which run) + the staging workgroups
    unsigned grid = (unsigned)(push_bin_groups(s, out.cap_bins) + PUSH_STAGE_WGS);
    s->diag_nb = (int)grid;
    hipEvent_t evk;
    timer_begin(w, PIC_T_PUSH_KERNEL, &evk);
    // the two-cell-rim push, then the wide-rim one for the tiles the plan flagged (each
    // returns at once on the other's workgroups)
    // then the deferred far deposits (tiles the push flagged, and the staging list).  The
    // wide-rim launch is left out when the plan gave it no tile.
    const unsigned far_grid = (unsigned)g.ntiles + STAGE_WGS;
    const bool wide = wide_policy() >= 0;
    if (s->prec == PIC_MIXED) {
        k_push<PIC_MIXED, HALO><<<grid, PUSH_THREADS, 0, w->stream>>>(P, view_of(in), view_of(out), w->ef32, w->ef64, O);
        if (wide)
            k_push<PIC_MIXED, HALO_WIDE><<<grid, PUSH_THREADS, 0, w->stream>>>(P, view_of(in), view_of(out), w->ef32,
                                                                               w->ef64, O);
        k_far_arrivals<PIC_MIXED><<<far_grid, 256, 0, w->stream>>>(g, view_of(out), O.far_in, O.win32, O.win64,
                                                                   O.win_used, s->den_extra);
    } else {
        k_push<PIC_FP64, HALO><<<grid, PUSH_THREADS, 0, w->stream>>>(P, view_of(in), view_of(out), w->ef32, w->ef64, O);
        if (wide)
            k_push<PIC_FP64, HALO_WIDE><<<grid, PUSH_THREADS, 0, w->stream>>>(P, view_of(in), view_of(out), w->ef32,
                                                                              w->ef64, O);
        k_far_arrivals<PIC_FP64><<<far_grid, 256, 0, w->stream>>>(g, view_of(out), O.far_in, O.win32, O.win64,
                                                                  O.win_used, s->den_extra);
    }
    PIC_LAUNCH_CHECK();
    timer_end(w, PIC_T_PUSH_KERNEL, evk);
    timer_end(w, PIC_T_PUSH, ev);
    // snapshot of this push, read by the host at its next call without a stream sync
    k_push_status<<<1, 256, 0, w->stream>>>(s->diag_part + 5 * (size_t)grid, (int)grid, out.stage_count,
                                           out.cap_stage, s->st_dev);
    PIC_LAUNCH_CHECK();
    PIC_HIP(hipMemcpyAsync(s->st_host, s->st_dev, sizeof(SpStatus), hipMemcpyDeviceToHost, w->stream));
    PIC_HIP(hipEventRecord(s->st_ev, w->stream));
    s->st_pending = true;
    s->added_since_push = 0;
    s->stage_known = false;
    s->np_exact = true;
    s->extra_dirty = true;
    s->cur = 1 - s->cur;
    s->den_state = 1;
    s->diag_valid = true;
    s->sorted = true;  // overflow, if any, is in staging and gets pushed next step
    s->stage_added = 0;
    return PIC_OK;
}

int pic_species_compute_number_density(pic_species *s) {
    PIC_CHECK(s, PIC_ERR_ARG, "null argument");
    pic_world *w = s->w;
    const Grid &g = w->g;
    // particles the last push dropped are reported here when its snapshot has landed
    int rc = poll_status(s, false);
    if (rc) return rc;
    if (s->den_state == 2) return PIC_OK;
    hipEvent_t ev;
    timer_begin(w, PIC_T_DEPOSIT, &ev);
    if (s->den_state == 0) {
        if (s->extra_dirty)  // far deposits of a push whose density was never finished
            k_fill_f64<<<nb256(g.nn), 256, 0, w->stream>>>(g.nn, s->den_extra, 0.0);
        k_fill_u8<<<nb256(g.ntiles), 256, 0, w->stream>>>(g.ntiles, s->win_used, 0);
        unsigned grid = g.ntiles + STAGE_WGS;
        PartView v = view_of(s->buf[s->cur]);
        if (s->prec == PIC_MIXED)
            k_deposit_tiles<PIC_MIXED><<<grid, 256, 0, w->stream>>>(g, v, (float *)s->win, (double *)s->win,
                                                                    s->win_used, s->den_extra);
        else
            k_deposit_tiles<PIC_FP64><<<grid, 256, 0, w->stream>>>(g, v, (float *)s->win, (double *)s->win,
                                                                   s->win_used, s->den_extra);
        PIC_LAUNCH_CHECK();
    }
    if (s->prec == PIC_MIXED)
        k_finish_density<float><<<nbk8(g.nn), 256, 0, w->stream>>>(g, (const float *)s->win, s->win_used,
                                                                    s->den_extra, s->den);
    else
        k_finish_density<double><<<nbk8(g.nn), 256, 0, w->stream>>>(g, (const double *)s->win, s->win_used,
                                                                     s->den_extra, s->den);
    PIC_LAUNCH_CHECK();
    timer_end(w, PIC_T_DEPOSIT, ev);
    s->den_state = 2;
    s->extra_dirty = false;  // the finish consumed and zeroed den_extra
    return PIC_OK;
}

int pic_species_get_den(pic_species *s, double *out) {
    PIC_CHECK(s && out, PIC_ERR_ARG, "null argument");
    PIC_HIP(hipMemcpyAsync(out, s->den, sizeof(double) * s->w->g.nn, hipMemcpyDeviceToHost, s->w->stream));
    PIC_HIP(hipStreamSynchronize(s->w->stream));
    return PIC_OK;
}

int pic_species_get_den_ave(pic_species *s, double *out) {
    PIC_CHECK(s && out, PIC_ERR_ARG, "null argument");
    PIC_HIP(hipMemcpyAsync(out, s->den_ave, sizeof(double) * s->w->g.nn, hipMemcpyDeviceToHost, s->w->stream));
    PIC_HIP(hipStreamSynchronize(s->w->stream));
    return PIC_OK;
}

int pic_species_update_averages(pic_species *s) {
    PIC_CHECK(s, PIC_ERR_ARG, "null argument");
    k_avg<<<nb256(s->w->g.nn), 256, 0, s->w->stream>>>(s->w->g.nn, s->den_ave, s->den, s->ave_samples);
    PIC_LAUNCH_CHECK();
    s->ave_samples++;
    return PIC_OK;
}

// getRealCount, getMomentum = mass*sum(v*mpw), getKE = 0.5*mass*sum(mpw*v^2) from the
// reduced sums W = sum(mpw), S = sum(mpw*v), Q = sum(mpw*|v|^2) of the stored velocities.
// PIC_MIXED stores v - u (u the species frame): sum(mpw*v) = S + u W and
// sum(mpw*|v|^2) = Q + 2 u.S + |u|^2 W.
static void finish_diag(const pic_species *s, const double *h, double out[5]) {
    const double *u = s->frame;
    bool moving = s->prec == PIC_MIXED && (u[0] != 0 || u[1] != 0 || u[2] != 0);
    double S[3] = {h[1], h[2], h[3]}, Q = h[4];
    if (moving) {
        Q += 2 * (u[0] * S[0] + u[1] * S[1] + u[2] * S[2]) + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * h[0];
        for (int c = 0; c < 3; c++) S[c] += u[c] * h[0];
    }
    out[0] = h[0];
    out[1] = S[0] * s->mass;
    out[2] = S[1] * s->mass;
    out[3] = S[2] * s->mass;
    out[4] = 0.5 * s->mass * Q;
}

int pic_species_diag(pic_species *s, double out[5]) {
    PIC_CHECK(s && out, PIC_ERR_ARG, "null argument");
    pic_world *w = s->w;
    if (!s->diag_valid) {
        unsigned grid = w->g.ntiles + STAGE_WGS;
        PartView v = view_of(s->buf[s->cur]);
        if (s->prec == PIC_MIXED)
            k_diag_all<PIC_MIXED><<<grid, 256, 0, w->stream>>>(w->g, v, s->diag_part);
        else
            k_diag_all<PIC_FP64><<<grid, 256, 0, w->stream>>>(w->g, v, s->diag_part);
        PIC_LAUNCH_CHECK();
        s->diag_valid = true;
        s->diag_nb = (int)grid;
    }
    k_reduce_slots<<<1, 1024, 0, w->stream>>>(s->diag_part, s->diag_nb, NDIAG_SLOTS, s->diag_dev);
    PIC_LAUNCH_CHECK();
    PIC_HIP(hipMemcpyAsync(s->diag_host, s->diag_dev, sizeof(double) * 5, hipMemcpyDeviceToHost, w->stream));
    PIC_HIP(hipStreamSynchronize(w->stream));
    finish_diag(s, s->diag_host, out);
    return PIC_OK;
}

int pic_species_set_frame_velocity(pic_species *s, const double u[3]) {
    PIC_CHECK(s && u, PIC_ERR_ARG, "null argument");
    for (int c = 0; c < 3; c++) PIC_CHECK(std::isfinite(u[c]), PIC_ERR_ARG, "frame velocity not finite");
    if (s->prec != PIC_MIXED) return PIC_OK;  // FP64 records hold v itself
    double du[3] = {s->frame[0] - u[0], s->frame[1] - u[1], s->frame[2] - u[2]};
    if (du[0] == 0 && du[1] == 0 && du[2] == 0) return PIC_OK;
    pic_world *w = s->w;
    if (s->np_upper > 0) {
        k_shift_frame<<<(unsigned)(w->g.ntiles + STAGE_WGS), 256, 0, w->stream>>>(
            w->g, view_of(s->buf[s->cur]), make_double3(du[0], du[1], du[2]));
        PIC_LAUNCH_CHECK();
    }
    for (int c = 0; c < 3; c++) s->frame[c] = u[c];
    s->diag_valid = false;  // partial sums of the old frame's velocities
    return PIC_OK;
}

int pic_species_set_launch(pic_species *s, long push_chunk, int xcd_stripe) {
    PIC_CHECK(s, PIC_ERR_ARG, "null argument");
    PIC_CHECK(push_chunk >= 256 && push_chunk <= (1L << 24), PIC_ERR_ARG, "push_chunk: 256 to 2^24 particles");
    PIC_CHECK(xcd_stripe >= 1 && xcd_stripe <= 4096, PIC_ERR_ARG, "xcd_stripe: 1 to 4096 entries");
    PIC_HIP(hipStreamSynchronize(s->w->stream));
    s->push_chunk = push_chunk;
    s->xcd_stripe = xcd_stripe;
    return PIC_OK;
}

// pic_species_diag's full reduction (k_diag_all, then the fixed-order k_reduce_slots) with
// its partials and sums in the static `scratch` of this device instead of the species'
// buffers: no allocation, no pinned memory.  Calls overlapping on one GPU share `scratch`.
int pic_species_diag_noalloc(pic_species *s, double out[5]) {
    PIC_CHECK(s && out, PIC_ERR_ARG, "null argument");
    pic_world *w = s->w;
    const long nb = (long)w->g.ntiles + STAGE_WGS;
    if (nb > DIAG_NOALLOC_MAX_BLOCKS) {
        char msg[192];
        snprintf(msg, sizeof msg,
                 "pic_species_diag_noalloc: %ld tiles exceed its static scratch (at most %ld tiles); "
                 "use pic_species_diag",
                 (long)w->g.ntiles, DIAG_NOALLOC_MAX_BLOCKS - STAGE_WGS);
        set_error(msg);
        return PIC_ERR_CAPACITY;
    }
    double *ptr = nullptr;
    int rc = scratch_on_device(w->device, &ptr);
    if (rc) return rc;
    PartView v = view_of(s->buf[s->cur]);
    if (s->prec == PIC_MIXED)
        k_diag_all<PIC_MIXED><<<(unsigned)nb, 256, 0, w->stream>>>(w->g, v, ptr);
    else
        k_diag_all<PIC_FP64><<<(unsigned)nb, 256, 0, w->stream>>>(w->g, v, ptr);
    PIC_LAUNCH_CHECK();
    k_reduce_slots<<<1, 1024, 0, w->stream>>>(ptr, (int)nb, NDIAG_SLOTS, ptr + DIAG_NOALLOC_SUMS);
    PIC_LAUNCH_CHECK();
    double h[NDIAG_SLOTS];
    PIC_HIP(hipMemcpyAsync(h, ptr + DIAG_NOALLOC_SUMS, sizeof h, hipMemcpyDeviceToHost, w->stream));
    PIC_HIP(hipStreamSynchronize(w->stream));
    finish_diag(s, h, out);
    return PIC_OK;
}

static int inject_common(pic_species *s, InjectParams &I, bool flat_ok) {
    (void)flat_ok;
    pic_world *w = s->w;
    for (int c = 0; c < 3; c++) I.frame[c] = s->frame[c];
    if (I.n <= 0) return PIC_OK;
    uint32_t stage = 0;
    // the staging count comes from the last push's snapshot when it has landed (the solve
    // between two pushes has passed it), else from a synchronous read
    int rc = poll_status(s, false);
    if (rc) return rc;
    if (s->stage_known) {
        stage = (uint32_t)s->stage_host;
    } else {
        PIC_HIP(hipMemcpyAsync(&w->scal_host->iters, s->buf[s->cur].stage_count, 4, hipMemcpyDeviceToHost,
                               w->stream));
        PIC_HIP(hipStreamSynchronize(w->stream));
        stage = (uint32_t)w->scal_host->iters;
    }
    long n = I.n;
    long need_total = s->np_upper + n;
    if (stage + n > s->buf[s->cur].cap_stage || need_total + need_total / 4 > s->buf[s->cur].cap_bins) {
        rc = ensure_capacity(s, need_total, stage + n);
        if (rc) return rc;
        PIC_HIP(hipMemcpyAsync(&w->scal_host->iters, s->buf[s->cur].stage_count, 4, hipMemcpyDeviceToHost,
                               w->stream));
        PIC_HIP(hipStreamSynchronize(w->stream));
        stage = (uint32_t)w->scal_host->iters;
        if (stage + n > s->buf[s->cur].cap_stage) {
            rc = do_sort(s);
            if (rc) return rc;
            stage = 0;
        }
    }
    PartBuf &b = s->buf[s->cur];
    I.id0 = s->next_id;
    hipEvent_t ev;
    timer_begin(w, PIC_T_INJECT, &ev);
    if (s->prec == PIC_MIXED)
        k_inject<PIC_MIXED><<<nb256(n), 256, 0, w->stream>>>(I, w->ef32, w->ef64, view_of(b), stage, -1, err_flag(s));
    else
        k_inject<PIC_FP64><<<nb256(n), 256, 0, w->stream>>>(I, w->ef32, w->ef64, view_of(b), stage, -1, err_flag(s));
    PIC_LAUNCH_CHECK();
    k_set_u32<<<1, 1, 0, w->stream>>>(b.stage_count, stage + (uint32_t)n);
    timer_end(w, PIC_T_INJECT, ev);
    s->next_id += (uint32_t)n;
    s->np_upper += n;
    s->added_since_push += n;
    s->stage_host = stage + n;
    s->stage_known = true;
    s->stage_added += n;
    s->sorted = false;
    s->den_state = 0;
    s->diag_valid = false;
    return PIC_OK;
}

int pic_species_inject_beam(pic_species *s, double v_drift, double den, double T, double dt, uint64_t seed,
                            uint64_t counter, long *n_added) {
    PIC_CHECK(s, PIC_ERR_ARG, "null argument");
    PIC_CHECK(s->mpw0 > 0, PIC_ERR_ARG, "species needs mpw0 > 0 for injection");
    pic_world *w = s->w;
    const Grid &g = w->g;
    // ColdBeamSource::sample (ch3/ver2/Source.cpp:6-18)
    double Lx = g.dh[0] * (g.ni - 1);
    double Ly = g.dh[1] * (g.nj - 1);
    double A = Lx * Ly;
    double num_real = den * v_drift * A * dt;
    U4 r = philox4x32_10(U4{0xffffffffu, (uint32_t)counter, (uint32_t)(counter >> 32), 0xfffffffeu},
                         (uint32_t)seed, (uint32_t)(seed >> 32));
    int num_sim = (int)(num_real / s->mpw0 + u01(r.x, r.y));
    if (n_added) *n_added = num_sim;
    InjectParams I{};
    I.g = g;
    I.n = num_sim;
    I.box = 0;
    I.v_drift = v_drift;
    I.vth = T > 0 ? sqrt(2 * KB * T / s->mass) : 0.0;
    I.w = s->mpw0;
    I.qm = s->charge / s->mass;
    I.half_dt = 0.5 * dt;
    I.rewind = 1;
    I.key0 = (uint32_t)seed;
    I.key1 = (uint32_t)(seed >> 32);
    I.ctr_lo = (uint32_t)counter;
    I.ctr_hi = (uint32_t)(counter >> 32);
    return inject_common(s, I, false);
}

int pic_species_load_box_device(pic_species *s, const double x1[3], const double x2[3], double num_den,
                                long num_mp, double v_drift, double T, uint64_t seed) {
    PIC_CHECK(s && x1 && x2 && num_mp >= 0, PIC_ERR_ARG, "bad argument");
    pic_world *w = s->w;
    const Grid &g = w->g;
    for (int c = 0; c < 3; c++)
        PIC_CHECK(x1[c] >= g.x0[c] && x2[c] <= g.xm[c] && x2[c] > x1[c], PIC_ERR_ARG, "box outside the domain");
    double box_vol = (x2[0] - x1[0]) * (x2[1] - x1[1]) * (x2[2] - x1[2]);
    double mpw = num_den * box_vol / (double)num_mp;
    // batches no larger than the staging list; each batch is sorted into the bins
    int rc = ensure_capacity(s, s->np_upper + num_mp, 0);
    if (rc) return rc;
    long done = 0;
    uint64_t counter = 0;
    while (done < num_mp) {
        long bs = std::min<long>(num_mp - done, s->buf[s->cur].cap_stage / 2);
        InjectParams I{};
        I.g = g;
        I.n = bs;
        I.box = 1;
        for (int c = 0; c < 3; c++) {
            I.x1[c] = x1[c];
            I.x2[c] = x2[c];
        }
        I.v_drift = v_drift;
        I.vth = T > 0 ? sqrt(2 * KB * T / s->mass) : 0.0;
        I.w = mpw;
        I.qm = s->charge / s->mass;
        I.rewind = 0;
        I.key0 = (uint32_t)seed;
        I.key1 = (uint32_t)(seed >> 32);
        I.ctr_lo = (uint32_t)counter;
        I.ctr_hi = 0x5eed0000u;
        // counter advances by the batch so that draws never repeat across batches
        rc = inject_common(s, I, true);
        if (rc) return rc;
        rc = do_sort(s);
        if (rc) return rc;
        done += bs;
        counter += 1;
    }
    return check_err(s);
}

}  // extern "C"
