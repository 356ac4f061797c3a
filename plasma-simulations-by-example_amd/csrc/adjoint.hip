// Differentiable flow past a sphere and its hand-written adjoint.
//
// The forward run is the ch3/ver2 loop (ch3/ver2/Main.cpp:50-70): inject (ColdBeamSource::
// sample with a thermal spread, ch3/ver2/Source.cpp:4-27 and ch4/Source.cpp:31-56, plus the
// addParticle rewind, ch3/ver2/Species.cpp:65-81), advance (gather + leapfrog,
// ch3/ver2/Species.cpp:7-48), computeNumberDensity (scatter / node_vol, :51-62),
// computeChargeDensity (World.cpp:47-55), the Newton solve of the Boltzmann-electron Poisson
// equation (PotentialSolver.cpp:225-296, here on the reduced symmetric system with the
// multigrid-preconditioned CG) and computeEF (:465-504), all in FP64 with the reference's
// operation order.  Three changes make the map from the parameters theta = (drift speed,
// sphere potential, Te) to the outputs differentiable; the reference has no counterpart:
//   - absorption is soft: after each push a particle's weight is multiplied by
//     m(x) = m_s(x) m_w(x), m_s = clamp((|x-c| - r)/delta, 0, 1) around the sphere and m_w the
//     product of clamp(distance/delta, 0, 1) over the five open walls.  A particle with
//     m = 0 is dead, as inSphere || !inBounds makes it in the reference (:28-32); the weight
//     falls to 0 continuously instead of jumping.  The inlet wall (z = z0) keeps the hard
//     test.  Slots are never compacted, so step n's particles are a prefix of step n+1's.
//   - each step injects a fixed number J of particles whose weight carries the flux:
//     w = n0 v_drift A dt / J (the reference rounds the count randomly, Source.cpp:18).
//   - random numbers are Philox4x32-10 keyed by (seed, slot, step), so a run is a
//     deterministic function of theta.
// Outputs: the ion current to the sphere, q sum_n sum_p w_p (1 - m_s(x_p)) / (K dt), and the
// number density after the last step.
//
// The adjoint runs the steps backwards from checkpoints of the particles and phi stored every
// M steps (each segment is recomputed, its states kept): the reverse leapfrog, the transposes
// of the CIC gather and scatter (and their derivatives in x), of the soft absorption, of
// computeEF, and per step one linear solve with the symmetric Newton Jacobian J_r of the
// reduced Poisson system (solve_linearized).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "picamd_internal.h"
#include "stencil.h"

namespace pic {

int compute_ef(pic_world *w, bool want32, bool want64);
int solve_linearized(pic_solver *s, const double *rhs, double *lam, double tol, int *iters);

// FP64 particle slots, structure of arrays
struct DState {
    double *x, *y, *z, *vx, *vy, *vz, *w;
};

struct DParams {
    Grid g;
    double c[3], r;   // sphere
    double inv_delta;  // 1 / soft absorbing width [1/m]
    double s64;        // dt*charge/mass (ch3/ver2/Species.cpp:23)
    double qm_half_dt; // addParticle's rewind factor: (q/m), then *(0.5 dt)
    double qm, half_dt;
    double dt;
};

// ---- CIC helpers (World::XtoL, Field_::gather / scatter, ch3/ver2/Field.h:167-211) --------
struct Cic {
    long b;  // node (i, j, k)
    double d[3];
};

__device__ inline Cic cic_of(const Grid &g, double x, double y, double z) {
#pragma clang fp contract(off)
    double X[3] = {x, y, z};
    int c[3];
    Cic r;
    const int nc[3] = {g.nci, g.ncj, g.nck};
    for (int a = 0; a < 3; a++) {
        double l = (X[a] - g.x0[a]) / g.dh[a];
        int i = (int)l;
        if (i < 0) i = 0;
        if (i > nc[a] - 1) i = nc[a] - 1;
        c[a] = i;
        r.d[a] = l - i;
    }
    r.b = c[0] + (long)g.ni * (c[1] + (long)g.nj * c[2]);
    return r;
}

// corner n of the reference's eight-term order: offsets and which weight (d or 1-d) per axis
__device__ inline long corner_off(const Grid &g, int n) {
    long sj = g.ni, sk = (long)g.ni * g.nj;
    const long off[8] = {0, 1, 1 + sj, sj, sk, 1 + sk, 1 + sj + sk, sj + sk};
    return off[n];
}
__device__ inline int corner_bits(int n) {
    // bit 0: +i, bit 1: +j, bit 2: +k
    const int b[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    return b[n];
}

// E at the particle, the reference's term order
__device__ inline double3 gather_e(const Grid &g, const double4 *ef, const Cic &q) {
#pragma clang fp contract(off)
    double3 r = make_double3(0, 0, 0);
    for (int n = 0; n < 8; n++) {
        int cb = corner_bits(n);
        double wi = (cb & 1) ? q.d[0] : 1 - q.d[0];
        double wj = (cb & 2) ? q.d[1] : 1 - q.d[1];
        double wk = (cb & 4) ? q.d[2] : 1 - q.d[2];
        double4 e = ef[q.b + corner_off(g, n)];
        if (n == 0) {
            r.x = e.x * wi * wj * wk;
            r.y = e.y * wi * wj * wk;
            r.z = e.z * wi * wj * wk;
        } else {
            r.x = r.x + e.x * wi * wj * wk;
            r.y = r.y + e.y * wi * wj * wk;
            r.z = r.z + e.z * wi * wj * wk;
        }
    }
    return r;
}

// scalar field value at the particle and its gradient in x (per metre)
__device__ inline double gather_s_grad(const Grid &g, const double *f, const Cic &q, double gr[3]) {
    double v = 0;
    gr[0] = gr[1] = gr[2] = 0;
    for (int n = 0; n < 8; n++) {
        int cb = corner_bits(n);
        double wi = (cb & 1) ? q.d[0] : 1 - q.d[0], si = (cb & 1) ? 1.0 : -1.0;
        double wj = (cb & 2) ? q.d[1] : 1 - q.d[1], sj = (cb & 2) ? 1.0 : -1.0;
        double wk = (cb & 4) ? q.d[2] : 1 - q.d[2], sk = (cb & 4) ? 1.0 : -1.0;
        double fv = f[q.b + corner_off(g, n)];
        v += fv * wi * wj * wk;
        gr[0] += fv * si * wj * wk;
        gr[1] += fv * wi * sj * wk;
        gr[2] += fv * wi * wj * sk;
    }
    gr[0] /= g.dh[0];
    gr[1] /= g.dh[1];
    gr[2] /= g.dh[2];
    return v;
}

// scatter (ch3/ver2/Field.h:167-186)
__device__ inline void scatter_atomic(const Grid &g, double *f, const Cic &q, double v) {
#pragma clang fp contract(off)
    for (int n = 0; n < 8; n++) {
        int cb = corner_bits(n);
        double wi = (cb & 1) ? q.d[0] : 1 - q.d[0];
        double wj = (cb & 2) ? q.d[1] : 1 - q.d[1];
        double wk = (cb & 4) ? q.d[2] : 1 - q.d[2];
        atomicAdd(&f[q.b + corner_off(g, n)], v * wi * wj * wk);
    }
}

// Sums of NV values per lane along runs of adjacent lanes with the same key (every lane of
// the wavefront calls it; a lane with nothing to add passes a key of its own): the slots of a
// segment are kept in cell order, so a wavefront's lanes share a few cells, and each run's
// first lane then issues the run's atomics in place of one set per lane.  Returns whether
// this lane heads its run.
template <int NV>
__device__ inline bool run_sums(int key, double (&v)[NV]) {
    const int lane = (int)__lane_id();
    const int kprev = __shfl_up(key, 1);
    const bool head = lane == 0 || kprev != key;
    const unsigned long long heads = __ballot(head);
    const unsigned long long above = heads & ~((2ULL << lane) - 1ULL);
    const int end = above ? __ffsll((long long)above) - 2 : 63;  // the run's last lane
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const bool same = lane + off <= end;
#pragma unroll
        for (int n = 0; n < NV; n++) {
            const double y = __shfl_down(v[n], off);
            if (same) v[n] += y;
        }
    }
    return head;
}

// scatter_atomic for a wavefront: act, the lane deposits v at q
__device__ inline void scatter_runs(const Grid &g, double *f, bool act, const Cic &q, double v) {
#pragma clang fp contract(off)
    double c[8];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        int cb = corner_bits(n);
        double wi = (cb & 1) ? q.d[0] : 1 - q.d[0];
        double wj = (cb & 2) ? q.d[1] : 1 - q.d[1];
        double wk = (cb & 4) ? q.d[2] : 1 - q.d[2];
        c[n] = v * wi * wj * wk;
    }
    const bool head = run_sums<8>(act ? (int)q.b : -1 - (int)__lane_id(), c);
    if (!act || !head) return;
#pragma unroll
    for (int n = 0; n < 8; n++) atomicAdd(&f[q.b + corner_off(g, n)], c[n]);
}

// ---- soft absorption ----------------------------------------------------------------------
__device__ inline double ramp(double dist, double inv, double &dr) {
    double t = dist * inv;
    if (t <= 0.0) {
        dr = 0.0;
        return 0.0;
    }
    if (t >= 1.0) {
        dr = 0.0;
        return 1.0;
    }
    dr = inv;
    return t;
}

// m_s and its gradient
__device__ inline double soft_sphere(const DParams &P, const double X[3], double gs[3]) {
    double rx = X[0] - P.c[0], ry = X[1] - P.c[1], rz = X[2] - P.c[2];
    double rr = sqrt(rx * rx + ry * ry + rz * rz);
    double dr;
    double m = ramp(rr - P.r, P.inv_delta, dr);
    if (dr != 0.0 && rr > 0.0) {
        gs[0] = dr * rx / rr;
        gs[1] = dr * ry / rr;
        gs[2] = dr * rz / rr;
    } else {
        gs[0] = gs[1] = gs[2] = 0.0;
    }
    return m;
}

// m_w (x-, x+, y-, y+, z+ walls) and its gradient; z below the inlet plane: 0 (hard)
__device__ inline double soft_walls(const DParams &P, const double X[3], double gw[3]) {
    const Grid &g = P.g;
    double ma[3], da[3];
    for (int a = 0; a < 3; a++) {
        double dlo = 0.0, dhi;
        double lo = a < 2 ? ramp(X[a] - g.x0[a], P.inv_delta, dlo) : (X[a] >= g.x0[a] ? 1.0 : 0.0);
        double hi = ramp(g.xm[a] - X[a], P.inv_delta, dhi);
        ma[a] = lo * hi;
        da[a] = dlo * hi - lo * dhi;
    }
    gw[0] = da[0] * ma[1] * ma[2];
    gw[1] = ma[0] * da[1] * ma[2];
    gw[2] = ma[0] * ma[1] * da[2];
    return ma[0] * ma[1] * ma[2];
}

// ---- draws ----------------------------------------------------------------------------------
// 16 uniforms for (slot q, step n, tag): Philox counter (q lo, q hi ^ tag<<24, n, block)
__device__ inline void draws16(uint32_t k0, uint32_t k1, long q, int n, uint32_t tag, double u[16]) {
    for (int b = 0; b < 8; b++) {
        U4 r = philox4x32_10(U4{(uint32_t)q, (uint32_t)(q >> 32) ^ (tag << 24), (uint32_t)n, (uint32_t)b}, k0, k1);
        u[2 * b] = u01(r.x, r.y);
        u[2 * b + 1] = u01(r.z, r.w);
    }
}

// isotropic thermal velocity of ch4/Species.cpp:149-174 from u[3..13] (the build's
// WarmBeamSource sampler, the same as k_inject's)
__device__ inline void thermal(double vth, const double u[16], double v[3]) {
#pragma clang fp contract(off)
    v[0] = v[1] = v[2] = 0.0;
    if (vth <= 0.0) return;
    double theta = 2 * PI_REF * u[3];
    double r = -1.0 + 2 * u[4];
    double a = sqrt(1 - r * r);
    double v1 = vth * (u[5] + u[6] + u[7] - 1.5);
    double v2 = vth * (u[8] + u[9] + u[10] - 1.5);
    double v3 = vth * (u[11] + u[12] + u[13] - 1.5);
    double vmag = 3 / sqrt(2.0 + 2 + 2) * sqrt(v1 * v1 + v2 * v2 + v3 * v3);
    v[0] = r * vmag;
    v[1] = cos(theta) * a * vmag;
    v[2] = sin(theta) * a * vmag;
}

// ---- forward kernels ------------------------------------------------------------------------
struct LoadParams {
    DParams P;
    uint32_t k0, k1;
    long n;        // particles
    long slot0;    // first slot
    int step;      // step index of the draws
    int inject;    // 0: box fill of the domain, 1: inlet face
    double v_drift, vth, w;
};

// initial fill (uniform over the domain, dead inside the sphere) or one step's injection on
// the k = 0 face with addParticle's rewind in the current E
__global__ void __launch_bounds__(256) k_dsph_load(LoadParams L, const double4 *ef, DState st) {
#pragma clang fp contract(off)
    long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (q >= L.n) return;
    const Grid &g = L.P.g;
    long p = L.slot0 + q;
    double u[16];
    draws16(L.k0, L.k1, p, L.step, L.inject ? 1u : 2u, u);
    double X[3];
    double w = L.w;
    if (L.inject) {
        X[0] = g.x0[0] + u[0] * (g.dh[0] * (g.ni - 1));
        X[1] = g.x0[1] + u[1] * (g.dh[1] * (g.nj - 1));
        X[2] = g.x0[2];
    } else {
        for (int a = 0; a < 3; a++) X[a] = g.x0[a] + u[a] * (g.xm[a] - g.x0[a]);
        double rx = X[0] - L.P.c[0], ry = X[1] - L.P.c[1], rz = X[2] - L.P.c[2];
        if (rx * rx + ry * ry + rz * rz <= L.P.r * L.P.r) w = 0.0;
    }
    double v[3];
    thermal(L.vth, u, v);
    v[2] = L.v_drift + v[2];
    if (L.inject) {
        Cic c = cic_of(g, X[0], X[1], X[2]);
        double3 E = gather_e(g, ef, c);
        v[0] -= (L.P.qm * E.x) * L.P.half_dt;
        v[1] -= (L.P.qm * E.y) * L.P.half_dt;
        v[2] -= (L.P.qm * E.z) * L.P.half_dt;
    }
    st.x[p] = X[0];
    st.y[p] = X[1];
    st.z[p] = X[2];
    st.vx[p] = v[0];
    st.vy[p] = v[1];
    st.vz[p] = v[2];
    st.w[p] = w;
}

// one push of slots [0, n): gather, leapfrog, soft absorption, deposit of the new weight;
// per-block partials of the weight the sphere took
__global__ void __launch_bounds__(256)
    k_dsph_push(DParams P, DState in, DState out, long n, const double4 *ef, double *den_raw, double *part_s) {
#pragma clang fp contract(off)
    __shared__ double sh[4];
    const Grid &g = P.g;
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double took = 0.0;
    bool dep = false;  // this lane deposits wd at cd
    Cic cd{};
    double wd = 0.0;
    if (p < n) {
        double w = in.w[p];
        if (w > 0.0) {
            double x = in.x[p], y = in.y[p], z = in.z[p];
            double vx = in.vx[p], vy = in.vy[p], vz = in.vz[p];
            Cic c = cic_of(g, x, y, z);
            double3 E = gather_e(g, ef, c);
            vx += E.x * P.s64;
            vy += E.y * P.s64;
            vz += E.z * P.s64;
            x += vx * P.dt;
            y += vy * P.dt;
            z += vz * P.dt;
            double X[3] = {x, y, z}, gs[3], gw[3];
            double ms = soft_sphere(P, X, gs);
            double mw = soft_walls(P, X, gw);
            took = w * (1.0 - ms);
            w = w * ms * mw;
            if (w > 0.0) {
                dep = true;
                cd = cic_of(g, x, y, z);
                wd = w;
            }
            out.x[p] = x;
            out.y[p] = y;
            out.z[p] = z;
            out.vx[p] = vx;
            out.vy[p] = vy;
            out.vz[p] = vz;
        }
        out.w[p] = w > 0.0 ? w : 0.0;
    }
    scatter_runs(g, den_raw, dep, cd, wd);
    double s = block_sum(took, sh);
    if (threadIdx.x == 0) part_s[blockIdx.x] = s;
}

// deposit of slots [0, n) where they stand (the initial fill's computeNumberDensity)
__global__ void __launch_bounds__(256) k_dsph_deposit(Grid g, DState st, long n, double *den_raw) {
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= n) return;
    double w = st.w[p];
    if (w > 0.0) scatter_atomic(g, den_raw, cic_of(g, st.x[p], st.y[p], st.z[p]), w);
}

// sort keys of slots [0, n): the base node of the particle's cell (live ones) or nn (others)
__global__ void __launch_bounds__(256) k_dsph_keys(Grid g, DState st, long n, unsigned *keys, int *vals) {
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= n) return;
    keys[p] = st.w[p] > 0.0 ? (unsigned)cic_of(g, st.x[p], st.y[p], st.z[p]).b : (unsigned)g.nn;
    vals[p] = (int)p;
}

// dst[t] = src[perm[t]] for t < n (a state into cell order)
__global__ void __launch_bounds__(256) k_dsph_gather(DState src, DState dst, const int *perm, long n) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    long p = perm[t];
    dst.x[t] = src.x[p];
    dst.y[t] = src.y[p];
    dst.z[t] = src.z[p];
    dst.vx[t] = src.vx[p];
    dst.vy[t] = src.vy[p];
    dst.vz[t] = src.vz[p];
    dst.w[t] = src.w[p];
}

// dst[perm[t]] = src[t] for t < n (adjoints back to the previous segment's slot order)
__global__ void __launch_bounds__(256) k_dsph_scatter(DState src, DState dst, const int *perm, long n) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    long p = perm[t];
    dst.x[p] = src.x[t];
    dst.y[p] = src.y[t];
    dst.z[p] = src.z[t];
    dst.vx[p] = src.vx[t];
    dst.vy[p] = src.vy[t];
    dst.vz[p] = src.vz[t];
    dst.w[p] = src.w[t];
}

// den = den_raw / node_vol (Species::computeNumberDensity), rho = q den
// (World::computeChargeDensity); den_out (nullable) keeps den
__global__ void __launch_bounds__(256)
    k_dsph_finish(long nn, const double *den_raw, const double *node_vol, double q, double *rho, double *den_out) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= nn) return;
    double v = node_vol[u];
    double d = v != 0.0 ? den_raw[u] / v : 0.0;
    rho[u] = 0.0 + d * q;
    if (den_out) den_out[u] = d;
}

// acc[slot] += fixed-order sum of nb partials (one workgroup)
__global__ void __launch_bounds__(1024) k_dsph_accum(const double *part, long nb, double *acc, int slot) {
    __shared__ double sh[16];
    double s = 0;
    for (long b = threadIdx.x; b < nb; b += blockDim.x) s += part[b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) acc[slot] += s;
}

__global__ void k_dsph_zero(long n, double *a) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u < n) a[u] = 0.0;
}

// ---- reverse kernels ------------------------------------------------------------------------
struct AdjPush {
    DParams P;
    long n;       // slots live before the push (incl. this step's injection)
    long n_old;   // slots that existed before this step's injection
    double sbar;  // dL/d(weight taken by the sphere)
    double winj_per_vd;  // d(injected weight)/d(v_drift)
};

// Reverse of one step for slots [0, n): given the adjoints of (x', v', w') after the push,
// writes those of (x, v, w) before it, adds the transpose of the gather into Ebar, and for
// the slots injected this step the drift-speed partials and the transpose of the rewind.
// fbar: dL/d(den) / node_vol after this step.
__global__ void __launch_bounds__(256)
    k_dsph_push_adj(AdjPush A, DState st, const double4 *ef, const double *fbar, DState adj, double *ebx,
                    double *eby, double *ebz, double *part_vd) {
#pragma clang fp contract(off)
    __shared__ double sh[4];
    const DParams &P = A.P;
    const Grid &g = P.g;
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double gvd = 0.0;
    bool dep = false;  // this lane adds eb3 (corner by corner) into Ebar at base node cb
    long cb0 = 0;
    double eb3[24];
#pragma unroll
    for (int q = 0; q < 24; q++) eb3[q] = 0.0;
    if (p < A.n) {
        double w = st.w[p];
        if (w > 0.0) {
            double x = st.x[p], y = st.y[p], z = st.z[p];
            double vx = st.vx[p], vy = st.vy[p], vz = st.vz[p];
            // forward again (the same operations as k_dsph_push)
            Cic c = cic_of(g, x, y, z);
            double3 E = gather_e(g, ef, c);
            double v1[3] = {vx + E.x * P.s64, vy + E.y * P.s64, vz + E.z * P.s64};
            double X1[3] = {x + v1[0] * P.dt, y + v1[1] * P.dt, z + v1[2] * P.dt};
            double gs[3], gw[3];
            double ms = soft_sphere(P, X1, gs);
            double mw = soft_walls(P, X1, gw);
            double w1 = w * ms * mw;
            // adjoints after the push
            double xb[3] = {adj.x[p], adj.y[p], adj.z[p]};
            double vb[3] = {adj.vx[p], adj.vy[p], adj.vz[p]};
            double wb1 = adj.w[p];
            if (w1 > 0.0) {
                // deposit of w1 at X1: den += w1 S(X1) / V
                double gf[3];
                double fv = gather_s_grad(g, fbar, cic_of(g, X1[0], X1[1], X1[2]), gf);
                wb1 += fv;
                for (int a = 0; a < 3; a++) xb[a] += w1 * gf[a];
            }
            // w1 = w ms mw;  took = w (1 - ms)
            double wb = wb1 * (ms * mw) + A.sbar * (1.0 - ms);
            for (int a = 0; a < 3; a++) xb[a] += wb1 * w * (gs[a] * mw + ms * gw[a]) - A.sbar * w * gs[a];
            // X1 = X + v1 dt;  v1 = v + s E(X)
            for (int a = 0; a < 3; a++) vb[a] += xb[a] * P.dt;
            double eb[3] = {vb[0] * P.s64, vb[1] * P.s64, vb[2] * P.s64};
            // dE/dX^T eb and E's transpose into the nodes
            double gx[3] = {0, 0, 0};
            for (int n = 0; n < 8; n++) {
                int cb = corner_bits(n);
                double wi = (cb & 1) ? c.d[0] : 1 - c.d[0], si = (cb & 1) ? 1.0 : -1.0;
                double wj = (cb & 2) ? c.d[1] : 1 - c.d[1], sj = (cb & 2) ? 1.0 : -1.0;
                double wk = (cb & 4) ? c.d[2] : 1 - c.d[2], sk = (cb & 4) ? 1.0 : -1.0;
                long u = c.b + corner_off(g, n);
                double4 e = ef[u];
                double ed = e.x * eb[0] + e.y * eb[1] + e.z * eb[2];
                gx[0] += ed * si * wj * wk;
                gx[1] += ed * wi * sj * wk;
                gx[2] += ed * wi * wj * sk;
                double W = wi * wj * wk;
                double e0 = eb[0] * W, e1 = eb[1] * W, e2 = eb[2] * W;
                if (p >= A.n_old) {
                    // injected this step: v = v_drift z - (q/m) E(X) (0.5 dt) + thermal
                    double r = P.qm * P.half_dt * W;
                    e0 -= vb[0] * r;
                    e1 -= vb[1] * r;
                    e2 -= vb[2] * r;
                }
                eb3[3 * n] = e0;
                eb3[3 * n + 1] = e1;
                eb3[3 * n + 2] = e2;
            }
            dep = true;
            cb0 = c.b;
            xb[0] += gx[0] / g.dh[0];
            xb[1] += gx[1] / g.dh[1];
            xb[2] += gx[2] / g.dh[2];
            if (p >= A.n_old) {
                // v_drift enters the velocity (dv_z/dv_drift = 1) and the weight
                gvd = vb[2] + wb * A.winj_per_vd;
                xb[0] = xb[1] = xb[2] = 0.0;
                vb[0] = vb[1] = vb[2] = 0.0;
                wb = 0.0;
            }
            adj.x[p] = xb[0];
            adj.y[p] = xb[1];
            adj.z[p] = xb[2];
            adj.vx[p] = vb[0];
            adj.vy[p] = vb[1];
            adj.vz[p] = vb[2];
            adj.w[p] = wb;
        } else {
            adj.x[p] = adj.y[p] = adj.z[p] = 0.0;
            adj.vx[p] = adj.vy[p] = adj.vz[p] = 0.0;
            adj.w[p] = 0.0;
        }
    }
    // E's transpose into the nodes, summed along runs of lanes with the same cell first
    const bool head = run_sums<24>(dep ? (int)cb0 : -1 - (int)__lane_id(), eb3);
    if (dep && head) {
#pragma unroll
        for (int n = 0; n < 8; n++) {
            long u = cb0 + corner_off(g, n);
            atomicAdd(&ebx[u], eb3[3 * n]);
            atomicAdd(&eby[u], eb3[3 * n + 1]);
            atomicAdd(&ebz[u], eb3[3 * n + 2]);
        }
    }
    double s = block_sum(gvd, sh);
    if (threadIdx.x == 0) part_vd[blockIdx.x] = s;
}

// partials of sum vbar_z over the initial particles (v = v_drift z + thermal)
__global__ void __launch_bounds__(256) k_dsph_init_vd(long n, const double *w0, const double *vbz, double *part) {
    __shared__ double sh[4];
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double s = (p < n && w0[p] > 0.0) ? vbz[p] : 0.0;
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// d E_a[iv] / d phi[iu] of computeEF (ch3/ver2/PotentialSolver.cpp:465-504) along one axis
__device__ inline double ef_coef(int iv, int iu, int n, double h) {
    double i2h = 1.0 / (2 * h);
    if (iv == 0) {
        if (iu == 0) return 3 * i2h;
        if (iu == 1) return -4 * i2h;
        if (iu == 2) return 1 * i2h;
        return 0.0;
    }
    if (iv == n - 1) {
        if (iu == n - 3) return -1 * i2h;
        if (iu == n - 2) return 4 * i2h;
        if (iu == n - 1) return -3 * i2h;
        return 0.0;
    }
    if (iu == iv + 1) return -i2h;
    if (iu == iv - 1) return i2h;
    return 0.0;
}

// phibar = EF^T Ebar
__global__ void __launch_bounds__(256)
    k_dsph_ef_t(Grid g, const double *ebx, const double *eby, const double *ebz, double *phibar) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= g.nn) return;
    int ijk[3];
    node_ijk(g, u, ijk[0], ijk[1], ijk[2]);
    const int n[3] = {g.ni, g.nj, g.nk};
    const long st[3] = {1, (long)g.ni, (long)g.ni * g.nj};
    const double *eb[3] = {ebx, eby, ebz};
    double s = 0.0;
    for (int a = 0; a < 3; a++) {
        for (int m = -2; m <= 2; m++) {
            int iv = ijk[a] + m;
            if (iv < 0 || iv >= n[a]) continue;
            double cf = ef_coef(iv, ijk[a], n[a], g.dh[a]);
            if (cf != 0.0) s += cf * eb[a][u + m * st[a]];
        }
    }
    phibar[u] = s;
}

// Folds phibar onto the unknowns of the reduced system: a face node equals its chain target
// (a regular node, or the sphere at phi_s, or the inlet at 0), a sphere node is phi_s.
// rhs (zeroed before) receives the regular part; part_ps the sphere's.
__global__ void __launch_bounds__(256)
    k_dsph_fold(Grid g, const uint8_t *oid, const double *phibar, double *rhs, double *part_ps) {
    __shared__ double sh[4];
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double ps = 0.0;
    if (u < g.nn) {
        int i, j, k;
        node_ijk(g, u, i, j, k);
        int t = node_type(g, oid, u, i, j, k);
        double pb = phibar[u];
        if (t == NT_REG) {
            atomicAdd(&rhs[u], pb);
        } else if (t == NT_DIR) {
            if (oid[u] == 1) ps = pb;
        } else {
            long v = chain_target(g, oid, u, i, j, k);
            if (oid[v] == 0) atomicAdd(&rhs[v], pb);
            else if (oid[v] == 1) ps = pb;
        }
    }
    double s = block_sum(ps, sh);
    if (threadIdx.x == 0) part_ps[blockIdx.x] = s;
}

struct SolveAdj {
    PcgParams Q;
    double q;  // species charge (rho = q den)
};

// From lam = J_r^-1 rhs: rhobar = -lam/eps0 on the regular nodes, fbar = q rhobar / node_vol,
// and the partials of -lam . dF/dTe and -lam . dF/dphi_s.  F on a regular node:
// A x + rho/eps0 - q n0 exp((x - phi0)/Te)/eps0, with phi_s in A x through sphere neighbours.
__global__ void __launch_bounds__(256)
    k_dsph_solve_adj(SolveAdj S, const uint8_t *oid, const double *phi, const double *lam, const double *node_vol,
                     double *fbar, double *part_te, double *part_ps) {
    __shared__ double sh[4];
    const Grid &g = S.Q.g;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double te = 0.0, ps = 0.0;
    if (u < g.nn) {
        int i, j, k;
        node_ijk(g, u, i, j, k);
        int t = node_type(g, oid, u, i, j, k);
        double fb = 0.0;
        if (t == NT_REG) {
            double l = lam[u];
            fb = S.q * (-l / EPS_0) / node_vol[u];
            double dphi = phi[u] - S.Q.phi0;
            double dF_dTe = QE * S.Q.n0 * dphi / (EPS_0 * S.Q.Te0 * S.Q.Te0) * exp(dphi / S.Q.Te0);
            te = -l * dF_dTe;
            double cs = 0.0;
            for (int d = 0; d < 6; d++)
                if (oid[u + nb_offset(g, d)] == 1) cs += nb_coef(S.Q, d);
            ps = -l * cs;
        }
        fbar[u] = fb;
    }
    double s = block_sum(te, sh);
    if (threadIdx.x == 0) part_te[blockIdx.x] = s;
    s = block_sum(ps, sh);
    if (threadIdx.x == 0) part_ps[blockIdx.x] = s;
}

// fbar (+)= dL/dden / node_vol (the loss's density term after the last step)
__global__ void __launch_bounds__(256)
    k_dsph_fbar_loss(long nn, const double *dden, const double *node_vol, double *fbar, int add) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= nn) return;
    double v = node_vol[u];
    double f = (dden && v != 0.0) ? dden[u] / v : 0.0;
    fbar[u] = add ? fbar[u] + f : f;
}

__global__ void __launch_bounds__(256) k_dsph_sumsq(long n, const double *a, double *part) {
    __shared__ double sh[4];
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double v = u < n ? a[u] * a[u] : 0.0;
    v = block_sum(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

}  // namespace pic

using namespace pic;

struct pic_adjoint {
    pic_adjoint_config cfg{};
    pic_world *w = nullptr;
    pic_solver *s = nullptr;
    Grid g{};
    DParams P{};
    long cap = 0;             // slots: n_init + K * n_inject
    int nseg = 0;             // segment states (max(M, 2))
    int nck = 0;              // checkpoints (ceil(K / M))
    std::vector<DState> seg, ck;
    std::vector<void *> mem;  // every allocation, freed by destroy
    double *phi_seg = nullptr;  // nseg x nn
    double *phi_ck = nullptr;   // nck x nn
    DState adj{};
    double *den_raw = nullptr, *den_k = nullptr, *fbar = nullptr, *phibar = nullptr, *rhs = nullptr, *lam = nullptr;
    double *phi_k = nullptr;  // phi after the last step
    double *ebx = nullptr, *eby = nullptr, *ebz = nullptr;
    double *part = nullptr, *part2 = nullptr;
    long part_cap = 0;
    double *acc = nullptr;       // device: [0] sphere weight, [1] dvd, [2] dphis, [3] dTe, [4] scratch
    double *acc_host = nullptr;  // pinned
    double theta[3] = {0, 0, 0};
    bool have_forward = false;
    double w_init = 0, w_inj_per_vd = 0, vth = 0;
    double stats[4] = {0, 0, 0, 0};  // forward ms, backward ms, Newton solves, adjoint CG its
    // Each checkpoint's state is put into cell order (stable radix sort by cell) before it is
    // stored: perms[ci] maps its new slots to the old ones, so that the backward pass can
    // carry the adjoints back across the checkpoint
    int *perms = nullptr;  // nck x cap
    unsigned *keys = nullptr, *keys2 = nullptr;
    int *vals = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int key_bits = 32;
};

namespace {

int alloc_d(pic_adjoint *a, double **p, size_t n) {
    void *m = nullptr;
    PIC_HIP(hipMalloc(&m, sizeof(double) * std::max<size_t>(n, 1)));
    a->mem.push_back(m);
    *p = (double *)m;
    return PIC_OK;
}

int alloc_state(pic_adjoint *a, DState &s) {
    double **f[7] = {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz, &s.w};
    for (auto p : f) {
        int rc = alloc_d(a, p, (size_t)a->cap);
        if (rc) return rc;
    }
    return PIC_OK;
}

int copy_state(pic_adjoint *a, const DState &src, const DState &dst, long n, hipStream_t st) {
    const double *s7[7] = {src.x, src.y, src.z, src.vx, src.vy, src.vz, src.w};
    double *d7[7] = {dst.x, dst.y, dst.z, dst.vx, dst.vy, dst.vz, dst.w};
    for (int i = 0; i < 7; i++)
        PIC_HIP(hipMemcpyAsync(d7[i], s7[i], sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    return PIC_OK;
}

inline long n_before(const pic_adjoint *a, int n) { return a->cfg.n_init + (long)n * a->cfg.n_inject; }

unsigned nbl(long n) { return (unsigned)std::max<long>((n + 255) / 256, 1); }

int sync_acc(pic_adjoint *a) {
    PIC_HIP(hipMemcpyAsync(a->acc_host, a->acc, sizeof(double) * 8, hipMemcpyDeviceToHost, a->w->stream));
    PIC_HIP(hipStreamSynchronize(a->w->stream));
    return PIC_OK;
}

int grow_part(pic_adjoint *a, long nb) {
    if (nb <= a->part_cap) return PIC_OK;
    long cap = std::max(nb, 2 * a->part_cap);
    int rc = alloc_d(a, &a->part, (size_t)cap);
    if (rc) return rc;
    rc = alloc_d(a, &a->part2, (size_t)cap);
    if (rc) return rc;
    a->part_cap = cap;
    return PIC_OK;
}

// the world's phi is the solution of the Poisson problem for the world's rho: Newton solve
int newton(pic_adjoint *a) {
    double res = 0;
    int its = 0;
    int rc = pic_solver_solve(a->s, &res, &its);
    if (rc == PIC_NOT_CONVERGED) {
        set_error("differentiable run: the Newton solve did not converge");
        return PIC_NOT_CONVERGED;
    }
    a->stats[2] += 1;
    return rc;
}

// deposit of slots [0, n) of st happened into den_raw: finish rho, solve, E
int field_step(pic_adjoint *a, bool keep_den) {
    pic_world *w = a->w;
    k_dsph_finish<<<nbl(a->g.nn), 256, 0, w->stream>>>(a->g.nn, a->den_raw, w->node_vol, a->cfg.charge, w->rho,
                                                      keep_den ? a->den_k : nullptr);
    PIC_LAUNCH_CHECK();
    int rc = newton(a);
    if (rc) return rc;
    return compute_ef(w, false, true);
}

int deposit_all(pic_adjoint *a, const DState &st, long n) {
    pic_world *w = a->w;
    k_dsph_zero<<<nbl(a->g.nn), 256, 0, w->stream>>>(a->g.nn, a->den_raw);
    k_dsph_deposit<<<nbl(n), 256, 0, w->stream>>>(a->g, st, n, a->den_raw);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

int load(pic_adjoint *a, const DState &st, long n, long slot0, int step, int inject) {
    if (n <= 0) return PIC_OK;
    LoadParams L;
    L.P = a->P;
    L.k0 = (uint32_t)a->cfg.seed;
    L.k1 = (uint32_t)(a->cfg.seed >> 32);
    L.n = n;
    L.slot0 = slot0;
    L.step = step;
    L.inject = inject;
    L.v_drift = a->theta[0];
    L.vth = a->vth;
    L.w = inject ? a->w_inj_per_vd * a->theta[0] : a->w_init;
    k_dsph_load<<<nbl(n), 256, 0, a->w->stream>>>(L, a->w->ef64, st);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

// slots [0, n) of src into cell order in dst (stable: slots of one cell keep their order);
// perm receives the old slot of each new one
int order_by_cell(pic_adjoint *a, const DState &src, const DState &dst, long n, int *perm) {
    pic_world *w = a->w;
    if (n <= 0) return PIC_OK;
    k_dsph_keys<<<nbl(n), 256, 0, w->stream>>>(a->g, src, n, a->keys, a->vals);
    PIC_LAUNCH_CHECK();
    size_t bytes = a->sort_tmp_bytes;
    PIC_HIP(hipcub::DeviceRadixSort::SortPairs(a->sort_tmp, bytes, a->keys, a->keys2, a->vals, perm, (int)n, 0,
                                               a->key_bits, w->stream));
    k_dsph_gather<<<nbl(n), 256, 0, w->stream>>>(src, dst, perm, n);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

// push slots [0, n) of in -> out, deposit, accumulate the sphere's weight into acc[0]
int push(pic_adjoint *a, const DState &in, const DState &out, long n) {
    pic_world *w = a->w;
    k_dsph_zero<<<nbl(a->g.nn), 256, 0, w->stream>>>(a->g.nn, a->den_raw);
    int rc = grow_part(a, nbl(n));
    if (rc) return rc;
    k_dsph_push<<<nbl(n), 256, 0, w->stream>>>(a->P, in, out, n, w->ef64, a->den_raw, a->part);
    k_dsph_accum<<<1, 1024, 0, w->stream>>>(a->part, nbl(n), a->acc, 0);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

int set_sphere_potential(pic_adjoint *a, double phi_s) {
    const pic_adjoint_config &c = a->cfg;
    return pic_world_add_sphere(a->w, c.sphere_c, c.sphere_r, phi_s);
}

// initial state of the run: fill, deposit, solve (continued in the sphere potential from a
// quarter of it, as a preloaded plasma needs), E
int initial_state(pic_adjoint *a, const DState &st) {
    pic_world *w = a->w;
    k_dsph_zero<<<nbl(a->g.nn), 256, 0, w->stream>>>(a->g.nn, w->phi);
    PIC_LAUNCH_CHECK();
    int rc = load(a, st, a->cfg.n_init, 0, 0, 0);
    if (rc) return rc;
    rc = deposit_all(a, st, a->cfg.n_init);
    if (rc) return rc;
    k_dsph_finish<<<nbl(a->g.nn), 256, 0, w->stream>>>(a->g.nn, a->den_raw, w->node_vol, a->cfg.charge, w->rho,
                                                      nullptr);
    PIC_LAUNCH_CHECK();
    for (int q = 1; q <= 4; q++) {
        rc = set_sphere_potential(a, a->theta[1] * (double)q / 4.0);
        if (rc) return rc;
        rc = newton(a);
        if (rc) return rc;
    }
    return compute_ef(w, false, true);
}

// one forward step n: inject into st_in, push into st_out, fields
int step(pic_adjoint *a, const DState &st_in, const DState &st_out, int n, bool keep_den) {
    long n0 = n_before(a, n);
    int rc = load(a, st_in, a->cfg.n_inject, n0, n, 1);
    if (rc) return rc;
    rc = push(a, st_in, st_out, n0 + a->cfg.n_inject);
    if (rc) return rc;
    return field_step(a, keep_den);
}

int reduce_into(pic_adjoint *a, const double *part, long nb, int slot) {
    k_dsph_accum<<<1, 1024, 0, a->w->stream>>>(part, nb, a->acc, slot);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

// reverse of the solve that produced the world's current phi, from phibar = dL/dphi:
// fbar = dL/dden / node_vol of the density that solve read, theta partials
int solve_adjoint_phibar(pic_adjoint *a) {
    pic_world *w = a->w;
    const Grid &g = a->g;
    long nn = g.nn;
    unsigned nb = nbl(nn);
    int rc = grow_part(a, nb);
    if (rc) return rc;
    k_dsph_zero<<<nb, 256, 0, w->stream>>>(nn, a->rhs);
    k_dsph_fold<<<nb, 256, 0, w->stream>>>(g, w->object_id, a->phibar, a->rhs, a->part);
    PIC_LAUNCH_CHECK();
    rc = reduce_into(a, a->part, nb, 2);
    if (rc) return rc;
    // tolerance relative to the right-hand side's RMS
    k_dsph_sumsq<<<nb, 256, 0, w->stream>>>(nn, a->rhs, a->part);
    a->acc_host[4] = 0;
    PIC_HIP(hipMemsetAsync(a->acc + 4, 0, sizeof(double), w->stream));
    rc = reduce_into(a, a->part, nb, 4);
    if (rc) return rc;
    rc = sync_acc(a);
    if (rc) return rc;
    double rms = sqrt(a->acc_host[4] / (double)nn);
    if (rms == 0.0) {
        k_dsph_zero<<<nb, 256, 0, w->stream>>>(nn, a->fbar);
        PIC_LAUNCH_CHECK();
        return PIC_OK;
    }
    int its = 0;
    rc = solve_linearized(a->s, a->rhs, a->lam, rms * a->cfg.adj_rel_tol, &its);
    if (rc) return rc;
    a->stats[3] += its;
    SolveAdj S;
    S.Q.g = g;
    S.Q.idx = 1.0 / g.dh[0];
    S.Q.idy = 1.0 / g.dh[1];
    S.Q.idz = 1.0 / g.dh[2];
    S.Q.idx2 = S.Q.idx * S.Q.idx;
    S.Q.idy2 = S.Q.idy * S.Q.idy;
    S.Q.idz2 = S.Q.idz * S.Q.idz;
    S.Q.phi0 = a->s->phi0;
    S.Q.Te0 = a->s->Te0;
    S.Q.n0 = a->s->n0;
    S.Q.tol = 0;
    S.q = a->cfg.charge;
    k_dsph_solve_adj<<<nb, 256, 0, w->stream>>>(S, w->object_id, w->phi, a->lam, w->node_vol, a->fbar, a->part,
                                              a->part2);
    PIC_LAUNCH_CHECK();
    rc = reduce_into(a, a->part, nb, 3);
    if (rc) return rc;
    return reduce_into(a, a->part2, nb, 2);
}

// the same from Ebar = dL/dE of computeEF's output: phibar = EF^T Ebar
int solve_adjoint(pic_adjoint *a) {
    pic_world *w = a->w;
    k_dsph_ef_t<<<nbl(a->g.nn), 256, 0, w->stream>>>(a->g, a->ebx, a->eby, a->ebz, a->phibar);
    PIC_LAUNCH_CHECK();
    return solve_adjoint_phibar(a);
}

}  // namespace

extern "C" {

int pic_adjoint_create(const pic_adjoint_config *cfg, pic_adjoint **out) {
    PIC_CHECK(cfg && out, PIC_ERR_ARG, "null argument");
    const pic_adjoint_config &c = *cfg;
    PIC_CHECK(c.ni >= 5 && c.nj >= 5 && c.nk >= 5, PIC_ERR_ARG, "grid too small");
    PIC_CHECK(c.steps >= 1 && c.checkpoint_every >= 1, PIC_ERR_ARG, "steps and checkpoint interval must be >= 1");
    PIC_CHECK(c.n_init >= 0 && c.n_inject >= 0 && c.n_init + (long)c.steps * c.n_inject > 0, PIC_ERR_ARG,
              "no particles");
    PIC_CHECK(c.soft_width > 0 && c.dt > 0 && c.mass > 0 && c.charge != 0, PIC_ERR_ARG, "bad physical parameter");
    pic_adjoint *a = new pic_adjoint();
    a->cfg = c;
    if (a->cfg.cg_tol <= 0) a->cfg.cg_tol = 1e-6;
    if (a->cfg.nr_tol <= 0) a->cfg.nr_tol = 1e-8;
    if (a->cfg.adj_rel_tol <= 0) a->cfg.adj_rel_tol = 1e-8;
    auto fail = [&](int rc) {
        pic_adjoint_destroy(a);
        return rc;
    };
    int rc = pic_world_create(c.ni, c.nj, c.nk, &a->w);
    if (rc) {
        delete a;
        return rc;
    }
    rc = pic_world_set_extents(a->w, c.x0, c.xm);
    if (rc) return fail(rc);
    rc = pic_world_add_sphere(a->w, c.sphere_c, c.sphere_r, -1.0);
    if (rc) return fail(rc);
    rc = pic_world_add_inlet(a->w);
    if (rc) return fail(rc);
    rc = pic_solver_create(a->w, PIC_SOLVER_MG, 20000, a->cfg.cg_tol, &a->s);
    if (rc) return fail(rc);
    a->s->nr_tol = a->cfg.nr_tol;
    a->g = a->w->g;
    DParams &P = a->P;
    P.g = a->g;
    for (int i = 0; i < 3; i++) P.c[i] = c.sphere_c[i];
    P.r = c.sphere_r;
    P.inv_delta = 1.0 / (c.soft_width * a->g.dh[0]);
    P.dt = c.dt;
    P.s64 = c.dt * c.charge / c.mass;
    P.qm = c.charge / c.mass;
    P.half_dt = 0.5 * c.dt;
    double L[3];
    for (int i = 0; i < 3; i++) L[i] = c.xm[i] - c.x0[i];
    double vol = L[0] * L[1] * L[2];
    a->w_init = c.n_init > 0 ? c.n0 * vol / (double)c.n_init : 0.0;
    double A = (a->g.dh[0] * (c.ni - 1)) * (a->g.dh[1] * (c.nj - 1));
    a->w_inj_per_vd = c.n_inject > 0 ? c.n0 * A * c.dt / (double)c.n_inject : 0.0;
    a->vth = c.T_ion > 0 ? sqrt(2 * KB * c.T_ion / c.mass) : 0.0;
    a->cap = c.n_init + (long)c.steps * c.n_inject;
    a->nseg = std::max(c.checkpoint_every, 2);
    a->nck = (c.steps + c.checkpoint_every - 1) / c.checkpoint_every;
    a->seg.resize(a->nseg);
    a->ck.resize(a->nck);
    for (auto &s : a->seg)
        if ((rc = alloc_state(a, s))) return fail(rc);
    for (auto &s : a->ck)
        if ((rc = alloc_state(a, s))) return fail(rc);
    if ((rc = alloc_state(a, a->adj))) return fail(rc);
    long nn = a->g.nn;
    double **fields[] = {&a->den_raw, &a->den_k, &a->fbar, &a->phibar, &a->rhs, &a->lam, &a->ebx, &a->eby, &a->ebz,
                         &a->phi_k};
    for (auto f : fields)
        if ((rc = alloc_d(a, f, (size_t)nn))) return fail(rc);
    if ((rc = alloc_d(a, &a->phi_seg, (size_t)nn * a->nseg))) return fail(rc);
    if ((rc = alloc_d(a, &a->phi_ck, (size_t)nn * a->nck))) return fail(rc);
    if ((rc = alloc_d(a, &a->acc, 8))) return fail(rc);
    if ((rc = grow_part(a, nbl(std::max(a->cap, nn))))) return fail(rc);
    {
        // the cell-order sort of up to cap slots: keys below nn + 1
        PIC_CHECK(a->cap < (1L << 31) && nn < (1L << 31), PIC_ERR_ARG, "too many slots or nodes for the cell sort");
        void *m[4] = {nullptr, nullptr, nullptr, nullptr};
        const size_t capn = (size_t)std::max<long>(a->cap, 1);
        size_t sz[3] = {sizeof(int) * capn * a->nck, sizeof(unsigned) * capn, sizeof(unsigned) * capn};
        for (int i = 0; i < 3; i++) {
            hipError_t he = hipMalloc(&m[i], sz[i]);
            if (he != hipSuccess) return fail(hip_fail(he, "hipMalloc (cell sort)", __FILE__, __LINE__));
            a->mem.push_back(m[i]);
        }
        a->perms = (int *)m[0];
        a->keys = (unsigned *)m[1];
        a->keys2 = (unsigned *)m[2];
        hipError_t he = hipMalloc(&m[3], sizeof(int) * capn);
        if (he != hipSuccess) return fail(hip_fail(he, "hipMalloc (cell sort)", __FILE__, __LINE__));
        a->mem.push_back(m[3]);
        a->vals = (int *)m[3];
        a->key_bits = 1;
        while ((1L << a->key_bits) <= nn) a->key_bits++;
        size_t bytes = 0;
        he = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, a->keys, a->keys2, a->vals, a->perms, (int)capn, 0,
                                                a->key_bits, a->w->stream);
        if (he != hipSuccess) return fail(hip_fail(he, "hipcub radix sort (size)", __FILE__, __LINE__));
        he = hipMalloc(&a->sort_tmp, std::max<size_t>(bytes, 1));
        if (he != hipSuccess) return fail(hip_fail(he, "hipMalloc (cell sort)", __FILE__, __LINE__));
        a->mem.push_back(a->sort_tmp);
        a->sort_tmp_bytes = bytes;
    }
    hipError_t e = hipHostMalloc((void **)&a->acc_host, sizeof(double) * 8, hipHostMallocDefault);
    if (e != hipSuccess) return fail(hip_fail(e, "hipHostMalloc", __FILE__, __LINE__));
    *out = a;
    return PIC_OK;
}

int pic_adjoint_destroy(pic_adjoint *a) {
    if (!a) return PIC_OK;
    if (a->w) hipStreamSynchronize(a->w->stream);
    for (void *m : a->mem) hipFree(m);
    if (a->acc_host) hipHostFree(a->acc_host);
    if (a->s) pic_solver_destroy(a->s);
    if (a->w) pic_world_destroy(a->w);
    delete a;
    return PIC_OK;
}

int pic_adjoint_forward(pic_adjoint *a, const double theta[3], double *current, double *den_out, double *phi_out) {
    PIC_CHECK(a && theta, PIC_ERR_ARG, "null argument");
    PIC_CHECK(theta[0] > 0 && theta[2] > 0, PIC_ERR_ARG, "drift speed and Te must be positive");
    auto t0 = std::chrono::steady_clock::now();
    pic_world *w = a->w;
    const pic_adjoint_config &c = a->cfg;
    for (int i = 0; i < 3; i++) a->theta[i] = theta[i];
    a->have_forward = false;
    a->stats[2] = 0;
    int rc = pic_solver_set_reference_values(a->s, 0.0, theta[2], c.n0);
    if (rc) return rc;
    PIC_HIP(hipMemsetAsync(a->acc, 0, sizeof(double) * 8, w->stream));
    DState A = a->seg[0], B = a->seg[1];
    rc = initial_state(a, A);
    if (rc) return rc;
    for (int n = 0; n < c.steps; n++) {
        if (n % c.checkpoint_every == 0) {
            int ci = n / c.checkpoint_every;
            // the checkpoint holds the state in cell order, and the run goes on from it
            rc = order_by_cell(a, A, a->ck[ci], n_before(a, n), a->perms + (size_t)ci * a->cap);
            if (rc) return rc;
            rc = copy_state(a, a->ck[ci], A, n_before(a, n), w->stream);
            if (rc) return rc;
            PIC_HIP(hipMemcpyAsync(a->phi_ck + (size_t)ci * a->g.nn, w->phi, sizeof(double) * a->g.nn,
                                   hipMemcpyDeviceToDevice, w->stream));
        }
        rc = step(a, A, B, n, n == c.steps - 1);
        if (rc) return rc;
        std::swap(A, B);
    }
    PIC_HIP(hipMemcpyAsync(a->phi_k, w->phi, sizeof(double) * a->g.nn, hipMemcpyDeviceToDevice, w->stream));
    if (den_out)
        PIC_HIP(hipMemcpyAsync(den_out, a->den_k, sizeof(double) * a->g.nn, hipMemcpyDeviceToDevice, w->stream));
    if (phi_out)
        PIC_HIP(hipMemcpyAsync(phi_out, w->phi, sizeof(double) * a->g.nn, hipMemcpyDeviceToDevice, w->stream));
    rc = sync_acc(a);
    if (rc) return rc;
    if (current) *current = c.charge * a->acc_host[0] / ((double)c.steps * c.dt);
    a->have_forward = true;
    a->stats[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PIC_OK;
}

int pic_adjoint_backward(pic_adjoint *a, double dcurrent, const double *dden, const double *dphi, double grad[3]) {
    PIC_CHECK(a && grad, PIC_ERR_ARG, "null argument");
    PIC_CHECK(a->have_forward, PIC_ERR_STATE, "pic_adjoint_backward needs a pic_adjoint_forward first");
    auto t0 = std::chrono::steady_clock::now();
    pic_world *w = a->w;
    const pic_adjoint_config &c = a->cfg;
    const Grid &g = a->g;
    long nn = g.nn;
    a->stats[3] = 0;
    // acc[1..3]: gradient; acc[0] keeps the forward's sphere weight
    PIC_HIP(hipMemsetAsync(a->acc + 1, 0, sizeof(double) * 4, w->stream));
    DState adj = a->adj;
    double *a7[7] = {adj.x, adj.y, adj.z, adj.vx, adj.vy, adj.vz, adj.w};
    for (double *p : a7) PIC_HIP(hipMemsetAsync(p, 0, sizeof(double) * a->cap, w->stream));
    int rc;
    if (dphi) {
        // the last step's solve: phi_K = solve(rho_K) feeds the loss directly
        PIC_HIP(hipMemcpyAsync(w->phi, a->phi_k, sizeof(double) * nn, hipMemcpyDeviceToDevice, w->stream));
        rc = set_sphere_potential(a, a->theta[1]);
        if (rc) return rc;
        PIC_HIP(hipMemcpyAsync(a->phibar, dphi, sizeof(double) * nn, hipMemcpyDeviceToDevice, w->stream));
        rc = solve_adjoint_phibar(a);
        if (rc) return rc;
    }
    k_dsph_fbar_loss<<<nbl(nn), 256, 0, w->stream>>>(nn, dden, w->node_vol, a->fbar, dphi ? 1 : 0);
    PIC_LAUNCH_CHECK();
    const double sbar = dcurrent * c.charge / ((double)c.steps * c.dt);
    const int M = c.checkpoint_every;
    for (int ci = a->nck - 1; ci >= 0; ci--) {
        int n_first = ci * M, n_last = std::min(c.steps, n_first + M) - 1;
        // recompute the segment from its checkpoint, keeping each step's state and phi
        rc = copy_state(a, a->ck[ci], a->seg[0], n_before(a, n_first), w->stream);
        if (rc) return rc;
        PIC_HIP(hipMemcpyAsync(w->phi, a->phi_ck + (size_t)ci * nn, sizeof(double) * nn, hipMemcpyDeviceToDevice,
                               w->stream));
        rc = set_sphere_potential(a, a->theta[1]);
        if (rc) return rc;
        rc = compute_ef(w, false, true);
        if (rc) return rc;
        for (int n = n_first; n <= n_last; n++) {
            int j = n - n_first;
            PIC_HIP(hipMemcpyAsync(a->phi_seg + (size_t)j * nn, w->phi, sizeof(double) * nn, hipMemcpyDeviceToDevice,
                                   w->stream));
            if (n < n_last) {
                rc = step(a, a->seg[j], a->seg[j + 1], n, false);
            } else {
                rc = load(a, a->seg[j], c.n_inject, n_before(a, n), n, 1);
            }
            if (rc) return rc;
        }
        // reverse through the segment
        for (int n = n_last; n >= n_first; n--) {
            int j = n - n_first;
            PIC_HIP(hipMemcpyAsync(w->phi, a->phi_seg + (size_t)j * nn, sizeof(double) * nn, hipMemcpyDeviceToDevice,
                                   w->stream));
            rc = compute_ef(w, false, true);
            if (rc) return rc;
            k_dsph_zero<<<nbl(nn), 256, 0, w->stream>>>(nn, a->ebx);
            k_dsph_zero<<<nbl(nn), 256, 0, w->stream>>>(nn, a->eby);
            k_dsph_zero<<<nbl(nn), 256, 0, w->stream>>>(nn, a->ebz);
            AdjPush A;
            A.P = a->P;
            A.n = n_before(a, n) + c.n_inject;
            A.n_old = n_before(a, n);
            A.sbar = sbar;
            A.winj_per_vd = a->w_inj_per_vd;
            rc = grow_part(a, nbl(A.n));
            if (rc) return rc;
            k_dsph_push_adj<<<nbl(A.n), 256, 0, w->stream>>>(A, a->seg[j], w->ef64, a->fbar, adj, a->ebx, a->eby,
                                                            a->ebz, a->part);
            PIC_LAUNCH_CHECK();
            rc = reduce_into(a, a->part, nbl(A.n), 1);
            if (rc) return rc;
            // the solve that produced phi_n (the initial one for n = 0) at phi_n
            rc = solve_adjoint(a);
            if (rc) return rc;
        }
        if (ci > 0) {
            // the adjoints of checkpoint ci's slots back to the order the previous segment ended in
            long nb = n_before(a, n_first);
            if (nb > 0) {
                k_dsph_scatter<<<nbl(nb), 256, 0, w->stream>>>(adj, a->seg[0], a->perms + (size_t)ci * a->cap, nb);
                PIC_LAUNCH_CHECK();
                std::swap(a->adj, a->seg[0]);
                adj = a->adj;
            }
        }
    }
    // initial particles: v = v_drift z + thermal
    if (c.n_init > 0) {
        k_dsph_init_vd<<<nbl(c.n_init), 256, 0, w->stream>>>(c.n_init, a->ck[0].w, adj.vz, a->part);
        PIC_LAUNCH_CHECK();
        rc = reduce_into(a, a->part, nbl(c.n_init), 1);
        if (rc) return rc;
    }
    rc = sync_acc(a);
    if (rc) return rc;
    grad[0] = a->acc_host[1];
    grad[1] = a->acc_host[2];
    grad[2] = a->acc_host[3];
    a->stats[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PIC_OK;
}

int pic_adjoint_stats(pic_adjoint *a, double out[4]) {
    PIC_CHECK(a && out, PIC_ERR_ARG, "null argument");
    for (int i = 0; i < 4; i++) out[i] = a->stats[i];
    return PIC_OK;
}

}  // extern "C"
