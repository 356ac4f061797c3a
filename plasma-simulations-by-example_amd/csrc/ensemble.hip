// Batched ensemble of independent grounded boxes (the ch2 program, ch2/Main.cpp:14-75, run for
// many boxes of uneven size at once).
//
// Every box is a ch2 world: specular walls (ch2/Species.cpp:32-36), linear SOR w = 1.4 with
// the residual every 25 iterations (ch2/PotentialSolver.cpp:11-67), computeEF
// (:70-110).  The boxes share nothing, so one launch per phase covers all of them through
// a box-descriptor table:
//   - k_ens_push: one workgroup per chunk of a (box, species) particle range: gather from
//     the box's E, leapfrog, reflection, CIC deposit into an LDS window of nodes around the
//     chunk's cells (the particles are counting-sorted by cell), flushed with one global
//     atomic per non-zero node; deposits that fall outside the window go straight to global
//     atomics;
//   - k_ens_finish: den = raw / node_vol per species (clearing raw), rho = sum q_s den_s;
//   - SOR: one launch per two iterations over all boxes (k_ens_sor_block: blocks of 8^3
//     nodes, four half-sweeps from an LDS window, into a second phi buffer), chunks of 25
//     iterations and a per-box residual check (k_ens_resid, k_ens_check), converged boxes
//     skipping at once.  When every box's phi fits in LDS, one workgroup per box runs its
//     whole solve there instead (k_ens_sor); with PIC_ENS_SOR_BLOCKS=0 those boxes do so and
//     the others get one launch per half-sweep (k_ens_sweep);
//   - k_ens_ef: computeEF of every box.
// The step is captured in three hipGraphs (push + density + the first SOR iteration, a chunk
// of 25 SOR iterations, computeEF) and replayed; the host reads the count of converged boxes
// between chunks.  A counting sort of every range runs between steps every sort_every steps.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "picamd_internal.h"

namespace pic {

// one grounded box
struct EnsBox {
    int ni, nj, nk, nci, ncj, nck;
    long node0, nn;
    double x0[3], xm[3], dh[3];
};

// one species' particles in one box
struct EnsRange {
    int box, sp;
    long p0, n;
    long cell0;      // first slot of this range's cell counters in the sort's histogram
    double s64;      // dt*q/m (FP64 form)
    float qm_dt;     // dt*q/m (mixed form)
    float dt_dh[3];  // dt/dh (mixed form: displacement in cells)
};

struct EnsChunk {
    int range;
    uint32_t first, count;
};

// particle storage of the mixed form: {cell bits i | j<<10 | k<<20, ox, oy, oz} {vx, vy, vz, w}
struct EnsParts {
    float4 *rec;  // 2 per particle (mixed)
    double *x, *y, *z, *vx, *vy, *vz, *w;  // FP64 form
    uint32_t *id;
};

// ---- shared helpers -------------------------------------------------------------------------
__device__ inline long ens_cell_node(const EnsBox &b, int i, int j, int k) {
    return b.node0 + i + (long)b.ni * (j + (long)b.nj * k);
}

__device__ inline double ens_node_vol(const EnsBox &b, int i, int j, int k) {
    // World::computeNodeVolumes (ch2/World.cpp:58-69)
    double V = b.dh[0] * b.dh[1] * b.dh[2];
    if (i == 0 || i == b.ni - 1) V *= 0.5;
    if (j == 0 || j == b.nj - 1) V *= 0.5;
    if (k == 0 || k == b.nk - 1) V *= 0.5;
    return V;
}

__device__ inline void ens_node_ijk(const EnsBox &b, long v, int &i, int &j, int &k) {
    unsigned uu = (unsigned)v, ni = (unsigned)b.ni, nj = (unsigned)b.nj;
    unsigned r = uu / ni;
    i = (int)(uu - r * ni);
    unsigned kk = r / nj;
    j = (int)(r - kk * nj);
    k = (int)kk;
}

// CIC corner n in the reference's term order (Field.h:164-208): bit 0 +i, bit 1 +j, bit 2 +k
__device__ inline int ens_corner(int n) {
    const int b[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    return b[n];
}

// ---- push + deposit -------------------------------------------------------------------------
struct EnsPushArgs {
    const EnsBox *box;
    const EnsRange *rng;
    const EnsChunk *chunk;
    EnsParts P;
    const float4 *ef32;
    const double4 *ef64;
    double *raw;   // [nsp][total nodes]
    long total_nodes;
    int wmax;      // LDS window capacity in nodes (0: no window)
    double dt;
};

template <int PREC>
__global__ void __launch_bounds__(256) k_ens_push(EnsPushArgs A) {
#pragma clang fp contract(off)
    extern __shared__ double sW[];
    __shared__ long s_w0;
    __shared__ int s_wn;
    const EnsChunk ch = A.chunk[blockIdx.x];
    const EnsRange R = A.rng[ch.range];
    const EnsBox B = A.box[R.box];
    double *raw = A.raw + (long)R.sp * A.total_nodes;
    const long sj = B.ni, sk = (long)B.ni * B.nj;
    // the window: nodes [w0, w0 + wn) from the chunk's first particle's cell to its last
    // particle's cell plus one plane and one row (the CIC reach)
    if (threadIdx.x == 0) {
        long first = R.p0 + ch.first, last = R.p0 + ch.first + ch.count - 1;
        long n0, n1;
        if constexpr (PREC == PIC_MIXED) {
            int c0 = __float_as_int(A.P.rec[2 * first].x), c1 = __float_as_int(A.P.rec[2 * last].x);
            n0 = cell_i(c0) + sj * cell_j(c0) + sk * cell_k(c0);
            n1 = cell_i(c1) + sj * cell_j(c1) + sk * cell_k(c1);
        } else {
            auto cnode = [&](long p) {
                int c[3];
                double X[3] = {A.P.x[p], A.P.y[p], A.P.z[p]};
                const int nc[3] = {B.nci, B.ncj, B.nck};
                for (int a = 0; a < 3; a++) {
                    double l = (X[a] - B.x0[a]) / B.dh[a];
                    int i = (int)l;
                    c[a] = i < 0 ? 0 : (i > nc[a] - 1 ? nc[a] - 1 : i);
                }
                return c[0] + sj * c[1] + sk * c[2];
            };
            n0 = cnode(first);
            n1 = cnode(last);
        }
        long lo = min(n0, n1), hi = max(n0, n1) + sk + sj + 2;
        long span = hi - lo;
        s_w0 = lo;
        s_wn = A.wmax > 0 ? (int)min(span, (long)A.wmax) : 0;
    }
    __syncthreads();
    const long w0 = s_w0;
    const int wn = s_wn;
    for (int q = threadIdx.x; q < wn; q += blockDim.x) sW[q] = 0.0;
    __syncthreads();
    // CIC deposit of a lane's particle at base node b (act: the lane has one).  The
    // particles are sorted by cell, so in a dense box most lanes of a wavefront share a few
    // cells and their atomics would queue on the same 8 addresses.  The 8 weights are first
    // summed along each run of adjacent lanes with the same b (shuffles bounded by the run's
    // last lane; between two sorts a cell can recur further on, in a run of its own), and the
    // run's first lane deposits the sums.
    const int lane = threadIdx.x & 63;
    auto deposit = [&](bool act, long b, double d0, double d1, double d2, double w) {
        double v[8];
#pragma unroll
        for (int n = 0; n < 8; n++) {
            int cb = ens_corner(n);
            double wi = (cb & 1) ? d0 : 1 - d0;
            double wj = (cb & 2) ? d1 : 1 - d1;
            double wk = (cb & 4) ? d2 : 1 - d2;
            v[n] = w * wi * wj * wk;
        }
        const int key = act ? (int)b : -1 - lane;  // idle lanes: runs of their own
        const int kprev = __shfl_up(key, 1);
        const bool head = lane == 0 || kprev != key;
        const unsigned long long heads = __ballot(head);
        const unsigned long long above = heads & ~((2ULL << lane) - 1ULL);
        const int end = above ? __ffsll((long long)above) - 2 : 63;  // the run's last lane
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const bool same = lane + off <= end;
#pragma unroll
            for (int n = 0; n < 8; n++) {
                const double y = __shfl_down(v[n], off);
                if (same) v[n] += y;
            }
        }
        if (!act || !head) return;
#pragma unroll
        for (int n = 0; n < 8; n++) {
            int cb = ens_corner(n);
            long u = b + ((cb & 1) ? 1 : 0) + ((cb & 2) ? sj : 0) + ((cb & 4) ? sk : 0);
            long q = u - w0;
            if (q >= 0 && q < wn)
                atomicAdd(&sW[q], v[n]);
            else
                atomicAdd(&raw[B.node0 + u], v[n]);
        }
    };
    // a trip count the same for every lane, so that the whole wavefront takes part in the
    // deposit's shuffles
    const uint32_t trips = (ch.count + blockDim.x - 1) / blockDim.x;
    for (uint32_t it = 0; it < trips; it++) {
        const uint32_t t = threadIdx.x + it * blockDim.x;
        const bool act = t < ch.count;
        long p = R.p0 + ch.first + (act ? t : 0);
        if constexpr (PREC == PIC_MIXED) {
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f), v = a;
            if (act) {
                a = A.P.rec[2 * p];
                v = A.P.rec[2 * p + 1];
            }
            int c = __float_as_int(a.x);
            int ci = cell_i(c), cj = cell_j(c), ck = cell_k(c);
            float o[3] = {a.y, a.z, a.w};
            // gather (FP32 E, the reference's term order)
            long b = ci + sj * cj + sk * ck;
            float ex = 0.f, ey = 0.f, ez = 0.f;
            for (int n = 0; n < 8; n++) {
                int cb = ens_corner(n);
                float wi = (cb & 1) ? o[0] : 1.f - o[0];
                float wj = (cb & 2) ? o[1] : 1.f - o[1];
                float wk = (cb & 4) ? o[2] : 1.f - o[2];
                long u = B.node0 + b + ((cb & 1) ? 1 : 0) + ((cb & 2) ? sj : 0) + ((cb & 4) ? sk : 0);
                float4 e = A.ef32[u];
                float W = wi * wj * wk;
                ex += e.x * W;
                ey += e.y * W;
                ez += e.z * W;
            }
            float vel[3] = {v.x + ex * R.qm_dt, v.y + ey * R.qm_dt, v.z + ez * R.qm_dt};
            int cc[3] = {ci, cj, ck};
            const int nc[3] = {B.nci, B.ncj, B.nck};
            for (int d = 0; d < 3; d++) {
                float s = o[d] + vel[d] * R.dt_dh[d];
                float fl = floorf(s);
                int c2 = cc[d] + (int)fl;
                float o2 = s - fl;
                if (o2 >= 1.0f) {  // s just below an integer rounded up
                    o2 = 0.0f;
                    c2 += 1;
                }
                // specular walls (ch2/Species.cpp:32-36) in cell units: p -> -p, p -> 2 nc - p
                if (c2 < 0) {
                    if (o2 > 0.0f) {
                        c2 = -c2 - 1;
                        o2 = 1.0f - o2;
                    } else {
                        c2 = -c2;
                    }
                    vel[d] = -vel[d];
                } else if (c2 >= nc[d]) {
                    if (o2 > 0.0f) {
                        c2 = 2 * nc[d] - c2 - 1;
                        o2 = 1.0f - o2;
                    } else {
                        c2 = 2 * nc[d] - c2;
                    }
                    vel[d] = -vel[d];
                    if (c2 >= nc[d]) {  // landed on the wall: one ulp inside
                        c2 = nc[d] - 1;
                        o2 = 0x1.fffffep-1f;
                    }
                }
                if (c2 < 0) {  // a move of more than the box: keep it inside
                    c2 = 0;
                    o2 = 0.0f;
                }
                cc[d] = c2;
                o[d] = o2;
            }
            if (act) {
                A.P.rec[2 * p] = make_float4(__int_as_float(pack_cell(cc[0], cc[1], cc[2])), o[0], o[1], o[2]);
                A.P.rec[2 * p + 1] = make_float4(vel[0], vel[1], vel[2], v.w);
            }
            deposit(act, cc[0] + sj * cc[1] + sk * cc[2], o[0], o[1], o[2], v.w);
        } else {
            double X[3] = {A.P.x[p], A.P.y[p], A.P.z[p]};
            double V[3] = {A.P.vx[p], A.P.vy[p], A.P.vz[p]};
            const int nc[3] = {B.nci, B.ncj, B.nck};
            auto cell = [&](const double *Xs, int *c, double *d) {
                for (int a = 0; a < 3; a++) {
                    double l = (Xs[a] - B.x0[a]) / B.dh[a];  // World::XtoL
                    int i = (int)l;
                    if (i < 0) i = 0;
                    if (i > nc[a] - 1) i = nc[a] - 1;
                    c[a] = i;
                    d[a] = l - i;
                }
            };
            int c[3];
            double d[3];
            cell(X, c, d);
            long b = B.node0 + c[0] + sj * c[1] + sk * c[2];
            double E[3];
            for (int n = 0; n < 8; n++) {
                int cb = ens_corner(n);
                double wi = (cb & 1) ? d[0] : 1 - d[0];
                double wj = (cb & 2) ? d[1] : 1 - d[1];
                double wk = (cb & 4) ? d[2] : 1 - d[2];
                double4 e = A.ef64[b + ((cb & 1) ? 1 : 0) + ((cb & 2) ? sj : 0) + ((cb & 4) ? sk : 0)];
                if (n == 0) {
                    E[0] = e.x * wi * wj * wk;
                    E[1] = e.y * wi * wj * wk;
                    E[2] = e.z * wi * wj * wk;
                } else {
                    E[0] = E[0] + e.x * wi * wj * wk;
                    E[1] = E[1] + e.y * wi * wj * wk;
                    E[2] = E[2] + e.z * wi * wj * wk;
                }
            }
            // ch2/Species.cpp:23-36
            for (int a = 0; a < 3; a++) V[a] += E[a] * R.s64;
            for (int a = 0; a < 3; a++) X[a] += V[a] * A.dt;
            for (int a = 0; a < 3; a++) {
                if (X[a] < B.x0[a]) {
                    X[a] = 2 * B.x0[a] - X[a];
                    V[a] *= -1.0;
                } else if (X[a] >= B.xm[a]) {
                    X[a] = 2 * B.xm[a] - X[a];
                    V[a] *= -1.0;
                }
            }
            if (act) {
                A.P.x[p] = X[0];
                A.P.y[p] = X[1];
                A.P.z[p] = X[2];
                A.P.vx[p] = V[0];
                A.P.vy[p] = V[1];
                A.P.vz[p] = V[2];
            }
            cell(X, c, d);
            deposit(act, c[0] + sj * c[1] + sk * c[2], d[0], d[1], d[2], A.P.w[p]);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < wn; q += blockDim.x) {
        double v = sW[q];
        if (v != 0.0) atomicAdd(&raw[B.node0 + w0 + q], v);
    }
}

// ---- node kernels over all boxes: one workgroup of 256 belongs to one box ------------------
struct EnsNodeArgs {
    const EnsBox *box;
    const int *blk_box;   // box of each workgroup
    const long *blk0;     // first workgroup of each box
    int nsp;
    long total_nodes;
};

// den_s = raw_s / node_vol (raw_s cleared for the next deposit), rho = sum_s q_s den_s
// (Species::computeNumberDensity + World::computeChargeDensity, ch2/Species.cpp:41-52,
// ch2/World.cpp:46-54)
__global__ void __launch_bounds__(256)
    k_ens_finish(EnsNodeArgs N, const double *charges, double *raw, double *den, double *rho) {
    int bi = N.blk_box[blockIdx.x];
    const EnsBox B = N.box[bi];
    long v = (long)(blockIdx.x - N.blk0[bi]) * 256 + threadIdx.x;
    if (v >= B.nn) return;
    int i, j, k;
    ens_node_ijk(B, v, i, j, k);
    double V = ens_node_vol(B, i, j, k);
    long u = B.node0 + v;
    double r = 0.0;
    for (int s = 0; s < N.nsp; s++) {
        long us = (long)s * N.total_nodes + u;
        double d = raw[us] / V;
        raw[us] = 0.0;
        den[us] = d;
        if (charges[s] != 0.0) r = r + charges[s] * d;
    }
    rho[u] = r;
}

// computeEF (ch2/PotentialSolver.cpp:70-110)
__global__ void __launch_bounds__(256)
    k_ens_ef(EnsNodeArgs N, const double *phi, float4 *ef32, double4 *ef64) {
    int bi = N.blk_box[blockIdx.x];
    const EnsBox B = N.box[bi];
    long v = (long)(blockIdx.x - N.blk0[bi]) * 256 + threadIdx.x;
    if (v >= B.nn) return;
    int i, j, k;
    ens_node_ijk(B, v, i, j, k);
    long u = B.node0 + v;
    long si = 1, sj = B.ni, sk = (long)B.ni * B.nj;
    double dx = B.dh[0], dy = B.dh[1], dz = B.dh[2];
    double ex, ey, ez;
    if (i == 0)
        ex = -(-3 * phi[u] + 4 * phi[u + si] - phi[u + 2 * si]) / (2 * dx);
    else if (i == B.ni - 1)
        ex = -(phi[u - 2 * si] - 4 * phi[u - si] + 3 * phi[u]) / (2 * dx);
    else
        ex = -(phi[u + si] - phi[u - si]) / (2 * dx);
    if (j == 0)
        ey = -(-3 * phi[u] + 4 * phi[u + sj] - phi[u + 2 * sj]) / (2 * dy);
    else if (j == B.nj - 1)
        ey = -(phi[u - 2 * sj] - 4 * phi[u - sj] + 3 * phi[u]) / (2 * dy);
    else
        ey = -(phi[u + sj] - phi[u - sj]) / (2 * dy);
    if (k == 0)
        ez = -(-3 * phi[u] + 4 * phi[u + sk] - phi[u + 2 * sk]) / (2 * dz);
    else if (k == B.nk - 1)
        ez = -(phi[u - 2 * sk] - 4 * phi[u - sk] + 3 * phi[u]) / (2 * dz);
    else
        ez = -(phi[u + sk] - phi[u - sk]) / (2 * dz);
    ef32[u] = make_float4((float)ex, (float)ey, (float)ez, 0.f);
    if (ef64) ef64[u] = make_double4(ex, ey, ez, 0.0);
}

// ---- SOR: one workgroup per box -------------------------------------------------------------
struct EnsSorArgs {
    const EnsBox *box;
    const int *boxes;  // the boxes of this launch
    double *phi;
    const double *rho;
    double w, tol;
    int max_it;
    int *iters;        // per box: iterations run (negative: not converged)
    double *res;       // per box: last residual norm
};

// The red-black form of ch2/PotentialSolver.cpp:11-67: interior nodes, w = 1.4, the residual
// L2 = sqrt(sum R^2 / (ni nj nk)) every 25 iterations from iteration 0, stop when L2 < tol.
// LDS: phi of the whole box in LDS (boxes that fit); otherwise phi in global memory.
template <bool LDS>
__global__ void __launch_bounds__(1024) k_ens_sor(EnsSorArgs S) {
    extern __shared__ double sphi[];
    __shared__ double sh[16];
    __shared__ int s_conv;
    const int bi = S.boxes[blockIdx.x];
    const EnsBox B = S.box[bi];
    double *gphi = S.phi + B.node0;
    const double *rho = S.rho + B.node0;
    double *phi = LDS ? sphi : gphi;
    const long nn = B.nn, sj = B.ni, sk = (long)B.ni * B.nj;
    if (LDS) {
        for (long u = threadIdx.x; u < nn; u += blockDim.x) sphi[u] = gphi[u];
    }
    if (threadIdx.x == 0) s_conv = 0;
    __syncthreads();
    const double idx2 = 1.0 / (B.dh[0] * B.dh[0]), idy2 = 1.0 / (B.dh[1] * B.dh[1]),
                 idz2 = 1.0 / (B.dh[2] * B.dh[2]);
    const double denom = 2 * idx2 + 2 * idy2 + 2 * idz2;
    const unsigned mi = (unsigned)(B.ni - 2), mj = (unsigned)(B.nj - 2);
    const unsigned half = (mi + 1) / 2, rows = mj * (unsigned)(B.nk - 2);
    const unsigned per_color = half * rows;
    double L2 = 0;
    int it;
    for (it = 0; it < S.max_it; it++) {
        for (int color = 0; color < 2; color++) {
            for (unsigned q = threadIdx.x; q < per_color; q += blockDim.x) {
                unsigned r = q / half, s = q - r * half;
                unsigned j = 1 + r % mj, k = 1 + r / mj;
                unsigned i = 1 + 2 * s + (((1 + j + k) & 1) != (unsigned)color ? 1u : 0u);
                if (i > mi) continue;
                long u = i + sj * j + sk * k;
                double phi_new = (rho[u] / EPS_0 + idx2 * (phi[u - 1] + phi[u + 1]) + idy2 * (phi[u - sj] + phi[u + sj]) +
                                  idz2 * (phi[u - sk] + phi[u + sk])) /
                                 denom;
                phi[u] = phi[u] + S.w * (phi_new - phi[u]);
            }
            if (!LDS) __threadfence_block();
            __syncthreads();
        }
        if (it % 25 == 0) {
            double sum = 0;
            for (unsigned q = threadIdx.x; q < mi * rows; q += blockDim.x) {
                unsigned r = q / mi;
                unsigned i = 1 + (q - r * mi);
                unsigned j = 1 + r % mj, k = 1 + r / mj;
                long u = i + sj * j + sk * k;
                double R = -phi[u] * denom + rho[u] / EPS_0 + idx2 * (phi[u - 1] + phi[u + 1]) +
                           idy2 * (phi[u - sj] + phi[u + sj]) + idz2 * (phi[u - sk] + phi[u + sk]);
                sum += R * R;
            }
            double tot = block_sum(sum, sh);
            if (threadIdx.x == 0) {
                L2 = sqrt(tot / (double)(B.ni * B.nj * B.nk));
                s_conv = L2 < S.tol;
                sh[0] = L2;
            }
            __syncthreads();
            L2 = sh[0];
            int conv = s_conv;
            __syncthreads();
            if (conv) break;
        }
    }
    if (LDS) {
        for (long u = threadIdx.x; u < nn; u += blockDim.x) gphi[u] = sphi[u];
    }
    if (threadIdx.x == 0) {
        S.iters[bi] = s_conv ? it + 1 : -it;
        S.res[bi] = L2;
    }
}

// ---- SOR of the boxes that do not fit in LDS: one launch per half-sweep over all of them ----
// The same iteration as k_ens_sor, spread over the whole GPU: every box gets its own run of
// workgroups (a workgroup never spans two boxes), a box that has converged skips the launch
// at once (its workgroups return), and the residual of iteration 0, 25, 50, ... is reduced per
// box in a fixed order.
struct EnsSweepArgs {
    const EnsBox *box;
    const int *blk_box;  // box of each sweep workgroup
    const long *blk0;    // first sweep workgroup of each box
    double *phi;
    const double *rho;
    const int *done;
    double w;
    double *copy = nullptr;  // k_ens_resid: also copies the interior of phi here
};

__global__ void __launch_bounds__(256) k_ens_sweep(EnsSweepArgs S, int color) {
#pragma clang fp contract(off)
    const int bi = S.blk_box[blockIdx.x];
    if (S.done[bi]) return;
    const EnsBox B = S.box[bi];
    const unsigned mi = (unsigned)(B.ni - 2), mj = (unsigned)(B.nj - 2);
    const unsigned half = (mi + 1) / 2, rows = mj * (unsigned)(B.nk - 2);
    const unsigned q = (unsigned)(blockIdx.x - S.blk0[bi]) * 256u + threadIdx.x;
    if (q >= half * rows) return;
    unsigned r = q / half, sl = q - r * half;
    unsigned j = 1 + r % mj, k = 1 + r / mj;
    unsigned i = 1 + 2 * sl + (((1 + j + k) & 1) != (unsigned)color ? 1u : 0u);
    if (i > mi) return;
    const long sj = B.ni, sk = (long)B.ni * B.nj;
    double *phi = S.phi + B.node0;
    const double *rho = S.rho + B.node0;
    const double idx2 = 1.0 / (B.dh[0] * B.dh[0]), idy2 = 1.0 / (B.dh[1] * B.dh[1]),
                 idz2 = 1.0 / (B.dh[2] * B.dh[2]);
    const double denom = 2 * idx2 + 2 * idy2 + 2 * idz2;
    long u = i + sj * j + sk * k;
    double phi_new = (rho[u] / EPS_0 + idx2 * (phi[u - 1] + phi[u + 1]) + idy2 * (phi[u - sj] + phi[u + sj]) +
                      idz2 * (phi[u - sk] + phi[u + sk])) /
                     denom;
    phi[u] = phi[u] + S.w * (phi_new - phi[u]);
}

// ---- the same SOR, two iterations per launch --------------------------------------------------
// A launch of k_ens_sweep costs about 13 us however few nodes it updates, and the largest box
// needs about 1400 iterations, most of them after the smaller boxes have converged.  Here a
// workgroup takes a block of ENS_SB^3 interior nodes of one box and its window with a T-node
// rim in LDS, runs T half-sweeps there (red first), each on the nodes the later ones still
// need (the block grown by T - 1 - h nodes), and writes the block's nodes to the other phi
// buffer.  Every node gets the same updates, from the same values, as from T launches of
// k_ens_sweep; the rim's nodes are recomputed by each block that needs them.  T = 4 (two
// iterations) with 8^3 blocks measured best: larger blocks leave the largest box too few
// workgroups, and a wider rim loads windows many times the block's nodes.
constexpr int ENS_SB = 8;
constexpr int ENS_T = 4;

struct EnsBlockArgs {
    const EnsBox *box;
    const int4 *blk;     // per workgroup: box, first interior i, j, k of its block
    const double *src;   // phi before this launch
    double *dst;         // phi after it (the other buffer)
    const double *rho;
    const int *done;
    double w;
};

template <int T>
__global__ void __launch_bounds__(256) k_ens_sor_block(EnsBlockArgs S) {
#pragma clang fp contract(off)
    extern __shared__ double sp[];
    constexpr int W = ENS_SB + 2 * T;
    const int4 bk = S.blk[blockIdx.x];
    if (S.done[bk.x]) return;
    const EnsBox B = S.box[bk.x];
    const long sj = B.ni, sk = (long)B.ni * B.nj;
    const double *src = S.src + B.node0;
    double *dst = S.dst + B.node0;
    const double *rho = S.rho + B.node0;
    const int i0 = bk.y - T, j0 = bk.z - T, k0 = bk.w - T;  // grid index of window node 0
    const double idx2 = 1.0 / (B.dh[0] * B.dh[0]), idy2 = 1.0 / (B.dh[1] * B.dh[1]),
                 idz2 = 1.0 / (B.dh[2] * B.dh[2]);
    const double denom = 2 * idx2 + 2 * idy2 + 2 * idz2;
    // the window: every load of a thread issued before the first LDS store
    constexpr int NW = W * W * W, PER = (NW + 255) / 256;
    double v[PER];
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const int n = threadIdx.x + 256 * t;
        const int a = n % W, b = (n / W) % W, c = n / (W * W);
        const int gi = i0 + a, gj = j0 + b, gk = k0 + c;
        const bool in = n < NW && gi >= 0 && gi < B.ni && gj >= 0 && gj < B.nj && gk >= 0 && gk < B.nk;
        v[t] = in ? src[gi + sj * gj + sk * gk] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const int n = threadIdx.x + 256 * t;
        if (n < NW) sp[n] = v[t];
    }
    __syncthreads();
    // one node of a half-sweep: the q-th of colour `color` in the region [ia, ia + mi) x
    // [ja, ja + mj) x [ka, ...), its rows colour-strided (half = (mi + 1) / 2 slots per row)
    auto update = [&](int q, int color, int ia, int ja, int ka, int mi, int mj, int half) {
        const int r = q / half, s = q - r * half;
        const int gj = ja + r % mj, gk = ka + r / mj;
        const int gi = ia + 2 * s + (((ia + gj + gk) & 1) != color ? 1 : 0);
        if (gi >= ia + mi) return;
        const int l = (gi - i0) + W * ((gj - j0) + W * (gk - k0));
        const double phi_new = (rho[gi + sj * gj + sk * gk] / EPS_0 + idx2 * (sp[l - 1] + sp[l + 1]) +
                                idy2 * (sp[l - W] + sp[l + W]) + idz2 * (sp[l - W * W] + sp[l + W * W])) /
                               denom;
        sp[l] = sp[l] + S.w * (phi_new - sp[l]);
    };
    // a block whose grown regions all lie inside the box's interior: sizes known at compile
    // time (the index divisions become multiplications)
    const bool full = bk.y - (T - 1) >= 1 && bk.y + ENS_SB + (T - 1) <= B.ni - 1 && bk.z - (T - 1) >= 1 &&
                      bk.z + ENS_SB + (T - 1) <= B.nj - 1 && bk.w - (T - 1) >= 1 &&
                      bk.w + ENS_SB + (T - 1) <= B.nk - 1;
#pragma unroll
    for (int h = 0; h < T; h++) {
        const int color = h & 1, e = T - 1 - h;
        if (full) {
            const int m = ENS_SB + 2 * e, half = (m + 1) / 2;
            for (int q = threadIdx.x; q < half * m * m; q += blockDim.x)
                update(q, color, bk.y - e, bk.z - e, bk.w - e, m, m, half);
        } else {
            // interior nodes of the block grown by e: [a, b) per axis
            const int ia = max(bk.y - e, 1), ib = min(bk.y + ENS_SB + e, B.ni - 1);
            const int ja = max(bk.z - e, 1), jb = min(bk.z + ENS_SB + e, B.nj - 1);
            const int ka = max(bk.w - e, 1), kb = min(bk.w + ENS_SB + e, B.nk - 1);
            const int mi = ib - ia, mj = jb - ja, mk = kb - ka;
            const int half = (mi + 1) / 2;
            const int cnt = (mi > 0 && mj > 0 && mk > 0) ? half * mj * mk : 0;
            for (int q = threadIdx.x; q < cnt; q += blockDim.x) update(q, color, ia, ja, ka, mi, mj, half);
        }
        __syncthreads();
    }
    const int mi = min(bk.y + ENS_SB, B.ni - 1) - bk.y, mj = min(bk.z + ENS_SB, B.nj - 1) - bk.z,
              mk = min(bk.w + ENS_SB, B.nk - 1) - bk.w;
    for (int q = threadIdx.x; q < mi * mj * mk; q += blockDim.x) {
        const int a = q % mi, r = q / mi, b = r % mj, c = r / mj;
        const int gi = bk.y + a, gj = bk.z + b, gk = bk.w + c;
        dst[gi + sj * gj + sk * gk] = sp[(gi - i0) + W * ((gj - j0) + W * (gk - k0))];
    }
}

// LDS bytes of k_ens_sor_block<T>: its window of phi
inline size_t sor_block_lds(int T) {
    const size_t w = ENS_SB + 2 * T;
    return sizeof(double) * w * w * w;
}

// per-workgroup partial sums of R^2 over the interior nodes (ch2/PotentialSolver.cpp:48-58)
__global__ void __launch_bounds__(256) k_ens_resid(EnsSweepArgs S, double *part) {
    __shared__ double sh[4];
    const int bi = S.blk_box[blockIdx.x];
    if (S.done[bi]) {
        if (threadIdx.x == 0) part[blockIdx.x] = 0.0;
        return;
    }
    const EnsBox B = S.box[bi];
    const unsigned mi = (unsigned)(B.ni - 2), mj = (unsigned)(B.nj - 2);
    const unsigned rows = mj * (unsigned)(B.nk - 2);
    const unsigned q = (unsigned)(blockIdx.x - S.blk0[bi]) * 256u + threadIdx.x;
    double R2 = 0.0;
    if (q < mi * rows) {
        unsigned r = q / mi;
        unsigned i = 1 + (q - r * mi);
        unsigned j = 1 + r % mj, k = 1 + r / mj;
        const long sj = B.ni, sk = (long)B.ni * B.nj;
        const double *phi = S.phi + B.node0;
        const double *rho = S.rho + B.node0;
        const double idx2 = 1.0 / (B.dh[0] * B.dh[0]), idy2 = 1.0 / (B.dh[1] * B.dh[1]),
                     idz2 = 1.0 / (B.dh[2] * B.dh[2]);
        const double denom = 2 * idx2 + 2 * idy2 + 2 * idz2;
        long u = i + sj * j + sk * k;
        double R = -phi[u] * denom + rho[u] / EPS_0 + idx2 * (phi[u - 1] + phi[u + 1]) +
                   idy2 * (phi[u - sj] + phi[u + sj]) + idz2 * (phi[u - sk] + phi[u + sk]);
        R2 = R * R;
        if (S.copy) S.copy[B.node0 + u] = phi[u];
    }
    double t = block_sum(R2, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

struct EnsCheckArgs {
    const EnsBox *box;
    const int *large;   // the boxes of this path
    const long *rblk0;  // first residual workgroup of each box
    const long *rblkn;  // residual workgroups of each box
    const double *part;
    double tol;
    int *box_it, *done, *iters, *ndone;
    double *res;
};

// one workgroup per box: L2 = sqrt(sum R^2 / (ni nj nk)) after `its` more iterations; the box
// stops when L2 < tol (its iteration count then is the reference's it + 1)
__global__ void __launch_bounds__(256) k_ens_check(EnsCheckArgs C, int its) {
    __shared__ double sh[4];
    const int bi = C.large[blockIdx.x];
    if (C.done[bi]) return;
    const EnsBox B = C.box[bi];
    double s = 0.0;
    for (long q = threadIdx.x; q < C.rblkn[bi]; q += 256) s += C.part[C.rblk0[bi] + q];
    double tot = block_sum(s, sh);
    if (threadIdx.x == 0) {
        double L2 = sqrt(tot / (double)(B.ni * B.nj * B.nk));
        int it = C.box_it[bi] + its;
        C.box_it[bi] = it;
        C.res[bi] = L2;
        if (L2 < C.tol) {
            C.done[bi] = 1;
            C.iters[bi] = it;
            atomicAdd(C.ndone, 1);
        } else {
            C.iters[bi] = -it;
        }
    }
}

__global__ void k_ens_sor_reset(int n, const int *large, int *box_it, int *done, int *ndone) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) {
        box_it[large[q]] = 0;
        done[large[q]] = 0;
    }
    if (q == 0) *ndone = 0;
}

// ---- counting sort of every range by cell ---------------------------------------------------
struct EnsSortArgs {
    const EnsBox *box;
    const EnsRange *rng;
    int nrange;
    EnsParts in, out;
    uint32_t *cnt;  // [total cells over ranges], then exclusive offsets
};

template <int PREC>
__device__ inline long ens_cell_of(const EnsSortArgs &A, const EnsRange &R, long p) {
    const EnsBox &B = A.box[R.box];
    if constexpr (PREC == PIC_MIXED) {
        int c = __float_as_int(A.in.rec[2 * p].x);
        return cell_i(c) + (long)B.nci * (cell_j(c) + (long)B.ncj * cell_k(c));
    } else {
        double X[3] = {A.in.x[p], A.in.y[p], A.in.z[p]};
        const int nc[3] = {B.nci, B.ncj, B.nck};
        int c[3];
        for (int a = 0; a < 3; a++) {
            double l = (X[a] - B.x0[a]) / B.dh[a];
            int i = (int)l;
            c[a] = i < 0 ? 0 : (i > nc[a] - 1 ? nc[a] - 1 : i);
        }
        return c[0] + (long)B.nci * (c[1] + (long)B.ncj * c[2]);
    }
}

// grid-stride over all particles; range of particle p by binary search over range starts
__device__ inline int ens_range_of(const EnsRange *rng, int nrange, long p) {
    int lo = 0, hi = nrange - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (rng[mid].p0 <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int PREC>
__global__ void __launch_bounds__(256) k_ens_hist(EnsSortArgs A, long np) {
    for (long p = blockIdx.x * 256L + threadIdx.x; p < np; p += (long)gridDim.x * 256) {
        const EnsRange R = A.rng[ens_range_of(A.rng, A.nrange, p)];
        atomicAdd(&A.cnt[R.cell0 + ens_cell_of<PREC>(A, R, p)], 1u);
    }
}

template <int PREC>
__global__ void __launch_bounds__(256) k_ens_scatter(EnsSortArgs A, long np, int ids) {
    for (long p = blockIdx.x * 256L + threadIdx.x; p < np; p += (long)gridDim.x * 256) {
        const EnsRange R = A.rng[ens_range_of(A.rng, A.nrange, p)];
        long d = atomicAdd(&A.cnt[R.cell0 + ens_cell_of<PREC>(A, R, p)], 1u);
        if constexpr (PREC == PIC_MIXED) {
            A.out.rec[2 * d] = A.in.rec[2 * p];
            A.out.rec[2 * d + 1] = A.in.rec[2 * p + 1];
        } else {
            A.out.x[d] = A.in.x[p];
            A.out.y[d] = A.in.y[p];
            A.out.z[d] = A.in.z[p];
            A.out.vx[d] = A.in.vx[p];
            A.out.vy[d] = A.in.vy[p];
            A.out.vz[d] = A.in.vz[p];
            A.out.w[d] = A.in.w[p];
        }
        if (ids) A.out.id[d] = A.in.id[p];
    }
}

// uploads of host particles (FP64, interleaved) into a range: mixed form converts to
// cell + FP32 offset
__global__ void __launch_bounds__(256)
    k_ens_upload(EnsBox B, long p0, long n, const double *pos, const double *vel, const double *mpw, int mixed,
                 EnsParts P, uint32_t id0, int ids) {
#pragma clang fp contract(off)
    long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= n) return;
    long p = p0 + q;
    double X[3] = {pos[3 * q], pos[3 * q + 1], pos[3 * q + 2]};
    if (mixed) {
        const int nc[3] = {B.nci, B.ncj, B.nck};
        int c[3];
        float o[3];
        for (int a = 0; a < 3; a++) {
            double l = (X[a] - B.x0[a]) / B.dh[a];
            int i = (int)l;
            i = i < 0 ? 0 : (i > nc[a] - 1 ? nc[a] - 1 : i);
            float f = (float)(l - i);
            if (f >= 1.0f) f = 0x1.fffffep-1f;
            if (f < 0.0f) f = 0.0f;
            c[a] = i;
            o[a] = f;
        }
        P.rec[2 * p] = make_float4(__int_as_float(pack_cell(c[0], c[1], c[2])), o[0], o[1], o[2]);
        P.rec[2 * p + 1] = make_float4((float)vel[3 * q], (float)vel[3 * q + 1], (float)vel[3 * q + 2], (float)mpw[q]);
    } else {
        P.x[p] = X[0];
        P.y[p] = X[1];
        P.z[p] = X[2];
        P.vx[p] = vel[3 * q];
        P.vy[p] = vel[3 * q + 1];
        P.vz[p] = vel[3 * q + 2];
        P.w[p] = mpw[q];
    }
    if (ids) P.id[p] = id0 + (uint32_t)q;
}

__global__ void __launch_bounds__(256)
    k_ens_download(EnsBox B, long p0, long n, int mixed, EnsParts P, double *pos, double *vel, double *mpw,
                   uint32_t *ids) {
    long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= n) return;
    long p = p0 + q;
    if (mixed) {
        float4 a = P.rec[2 * p], v = P.rec[2 * p + 1];
        int c = __float_as_int(a.x);
        pos[3 * q] = B.x0[0] + ((double)cell_i(c) + (double)a.y) * B.dh[0];
        pos[3 * q + 1] = B.x0[1] + ((double)cell_j(c) + (double)a.z) * B.dh[1];
        pos[3 * q + 2] = B.x0[2] + ((double)cell_k(c) + (double)a.w) * B.dh[2];
        vel[3 * q] = v.x;
        vel[3 * q + 1] = v.y;
        vel[3 * q + 2] = v.z;
        mpw[q] = v.w;
    } else {
        pos[3 * q] = P.x[p];
        pos[3 * q + 1] = P.y[p];
        pos[3 * q + 2] = P.z[p];
        vel[3 * q] = P.vx[p];
        vel[3 * q + 1] = P.vy[p];
        vel[3 * q + 2] = P.vz[p];
        mpw[q] = P.w[p];
    }
    if (ids) ids[q] = P.id[p];
}

// device load of loadParticlesBox (ch2/Species.cpp:74-97): uniform in [x1, x2), v = 0,
// Philox draws keyed by (seed, range, particle)
__global__ void __launch_bounds__(256)
    k_ens_load(EnsBox B, long p0, long n, double x10, double x11, double x12, double x20, double x21, double x22,
               double w, uint32_t k0, uint32_t k1, uint32_t tag, int mixed, EnsParts P, int ids) {
#pragma clang fp contract(off)
    long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= n) return;
    U4 r = philox4x32_10(U4{(uint32_t)q, (uint32_t)(q >> 32), tag, 0u}, k0, k1);
    U4 r2 = philox4x32_10(U4{(uint32_t)q, (uint32_t)(q >> 32), tag, 1u}, k0, k1);
    double X[3] = {x10 + u01(r.x, r.y) * (x20 - x10), x11 + u01(r.z, r.w) * (x21 - x11),
                   x12 + u01(r2.x, r2.y) * (x22 - x12)};
    long p = p0 + q;
    if (mixed) {
        const int nc[3] = {B.nci, B.ncj, B.nck};
        int c[3];
        float o[3];
        for (int a = 0; a < 3; a++) {
            double l = (X[a] - B.x0[a]) / B.dh[a];
            int i = (int)l;
            i = i < 0 ? 0 : (i > nc[a] - 1 ? nc[a] - 1 : i);
            float f = (float)(l - i);
            if (f >= 1.0f) f = 0x1.fffffep-1f;
            c[a] = i;
            o[a] = f;
        }
        P.rec[2 * p] = make_float4(__int_as_float(pack_cell(c[0], c[1], c[2])), o[0], o[1], o[2]);
        P.rec[2 * p + 1] = make_float4(0.f, 0.f, 0.f, (float)w);
    } else {
        P.x[p] = X[0];
        P.y[p] = X[1];
        P.z[p] = X[2];
        P.vx[p] = P.vy[p] = P.vz[p] = 0.0;
        P.w[p] = w;
    }
    if (ids) P.id[p] = (uint32_t)q;
}

// per-range diagnostics partials: sum w, sum w v (3), sum w v^2
__global__ void __launch_bounds__(256)
    k_ens_diag(const EnsRange *rng, const EnsChunk *chunk, EnsParts P, int mixed, double *part) {
    __shared__ double sh[4];
    const EnsChunk ch = chunk[blockIdx.x];
    const EnsRange R = rng[ch.range];
    double acc[5] = {0, 0, 0, 0, 0};
    for (uint32_t t = threadIdx.x; t < ch.count; t += 256) {
        long p = R.p0 + ch.first + t;
        double vx, vy, vz, w;
        if (mixed) {
            float4 v = P.rec[2 * p + 1];
            vx = v.x;
            vy = v.y;
            vz = v.z;
            w = v.w;
        } else {
            vx = P.vx[p];
            vy = P.vy[p];
            vz = P.vz[p];
            w = P.w[p];
        }
        acc[0] += w;
        acc[1] += w * vx;
        acc[2] += w * vy;
        acc[3] += w * vz;
        acc[4] += w * (vx * vx + vy * vy + vz * vz);
    }
    for (int q = 0; q < 5; q++) {
        double s = block_sum(acc[q], sh);
        if (threadIdx.x == 0) part[5 * (long)blockIdx.x + q] = s;
    }
}

}  // namespace pic

using namespace pic;

struct pic_ensemble {
    int device = 0;
    hipStream_t stream = nullptr;
    int nbox = 0, nsp = 0, prec = PIC_MIXED;
    bool track_ids = false;
    std::vector<EnsBox> hbox;
    std::vector<EnsRange> hrng;
    std::vector<EnsChunk> hchunk;
    std::vector<double> mass, charge;
    std::vector<uint32_t> next_id;
    long total_nodes = 0, np = 0, total_cells = 0;
    EnsBox *box = nullptr;
    EnsRange *rng = nullptr;
    EnsChunk *chunk = nullptr;
    int *blk_box = nullptr;
    long *blk0 = nullptr;
    long nblk_nodes = 0;
    EnsParts P{}, Q{};  // particles, sort target
    void *pmem[2] = {nullptr, nullptr};
    double *raw = nullptr, *den = nullptr, *rho = nullptr, *phi = nullptr, *charges_dev = nullptr;
    float4 *ef32 = nullptr;
    double4 *ef64 = nullptr;
    int *iters = nullptr;
    double *res = nullptr;
    int *sor_small = nullptr, *sor_large = nullptr;  // box lists
    int n_small = 0, n_large = 0;
    // the boxes whose phi is not held in LDS: sweep and residual workgroup tables, per-box
    // iteration counters and flags, the count of converged boxes (and its pinned mirror)
    int *sw_blk_box = nullptr, *rs_blk_box = nullptr;
    long *sw_blk0 = nullptr, *rs_blk0 = nullptr, *rs_blkn = nullptr;
    long n_sw_blk = 0, n_rs_blk = 0;
    // the same boxes in blocks of ENS_SB^3 interior nodes (k_ens_sor_block) and phi's second
    // buffer; sor_blocks false: one launch of k_ens_sweep per half-sweep (PIC_ENS_SOR_BLOCKS=0)
    int4 *sblk = nullptr;
    long n_sblk = 0;
    double *phi2 = nullptr;
    bool sor_blocks = true;
    double *rs_part = nullptr;
    int *box_it = nullptr, *done = nullptr, *ndone = nullptr, *ndone_host = nullptr;
    hipGraph_t graph_c = nullptr, graph_e = nullptr;
    hipGraphExec_t exec_c = nullptr, exec_e = nullptr;
    long last_sor_its = 0;  // iterations issued to the large boxes in the last step
    size_t sor_lds = 0;  // bytes of LDS for the small-box SOR
    uint32_t *cnt = nullptr, *off = nullptr;  // sort: per-cell counts, then their exclusive scan
    void *scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    double *diag_part = nullptr;
    // tuning
    int chunk_particles = 4096;
    int window_nodes = 6144;
    int sor_threads = 1024;
    int sort_every = 20;
    bool use_graph = true;
    // graph of one step
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double g_dt = -1, g_tol = -1;
    int g_max_it = -1;
    int graph_nodes = 0;
    long steps_done = 0, steps_since_sort = 0;
    bool sorted = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_step_ms = 0;
};

namespace {

int ens_alloc(void **p, size_t bytes) {
    PIC_HIP(hipMalloc(p, std::max<size_t>(bytes, 16)));
    return PIC_OK;
}

// chunk table for the current chunk size
int build_chunks(pic_ensemble *e) {
    e->hchunk.clear();
    for (int r = 0; r < (int)e->hrng.size(); r++) {
        const EnsRange &R = e->hrng[r];
        for (long f = 0; f < R.n; f += e->chunk_particles) {
            EnsChunk c;
            c.range = r;
            c.first = (uint32_t)f;
            c.count = (uint32_t)std::min<long>(e->chunk_particles, R.n - f);
            e->hchunk.push_back(c);
        }
    }
    if (e->chunk) hipFree(e->chunk);
    e->chunk = nullptr;
    int rc = ens_alloc((void **)&e->chunk, sizeof(EnsChunk) * std::max<size_t>(e->hchunk.size(), 1));
    if (rc) return rc;
    PIC_HIP(hipMemcpy(e->chunk, e->hchunk.data(), sizeof(EnsChunk) * e->hchunk.size(), hipMemcpyHostToDevice));
    if (e->diag_part) hipFree(e->diag_part);
    e->diag_part = nullptr;
    return ens_alloc((void **)&e->diag_part, sizeof(double) * 5 * std::max<size_t>(e->hchunk.size(), 1));
}

int drop_graph(pic_ensemble *e) {
    hipGraphExec_t *x[3] = {&e->exec, &e->exec_c, &e->exec_e};
    hipGraph_t *g[3] = {&e->graph, &e->graph_c, &e->graph_e};
    for (int i = 0; i < 3; i++) {
        if (*x[i]) hipGraphExecDestroy(*x[i]);
        if (*g[i]) hipGraphDestroy(*g[i]);
        *x[i] = nullptr;
        *g[i] = nullptr;
    }
    return PIC_OK;
}

EnsPushArgs push_args(pic_ensemble *e, double dt) {
    EnsPushArgs A;
    A.box = e->box;
    A.rng = e->rng;
    A.chunk = e->chunk;
    A.P = e->P;
    A.ef32 = e->ef32;
    A.ef64 = e->ef64;
    A.raw = e->raw;
    A.total_nodes = e->total_nodes;
    A.wmax = e->window_nodes;
    A.dt = dt;
    return A;
}

EnsNodeArgs node_args(pic_ensemble *e) {
    return EnsNodeArgs{e->box, e->blk_box, e->blk0, e->nsp, e->total_nodes};
}

int launch_push(pic_ensemble *e, double dt) {
    if (e->hchunk.empty()) return PIC_OK;
    EnsPushArgs A = push_args(e, dt);
    size_t lds = sizeof(double) * (size_t)std::max(e->window_nodes, 1);
    if (e->prec == PIC_MIXED)
        k_ens_push<PIC_MIXED><<<(unsigned)e->hchunk.size(), 256, lds, e->stream>>>(A);
    else
        k_ens_push<PIC_FP64><<<(unsigned)e->hchunk.size(), 256, lds, e->stream>>>(A);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

int launch_finish(pic_ensemble *e) {
    k_ens_finish<<<(unsigned)e->nblk_nodes, 256, 0, e->stream>>>(node_args(e), e->charges_dev, e->raw, e->den, e->rho);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

// the boxes held in LDS: one workgroup each runs the whole solve
int launch_sor_small(pic_ensemble *e, int max_it, double tol) {
    if (!e->n_small) return PIC_OK;
    EnsSorArgs S;
    S.box = e->box;
    S.phi = e->phi;
    S.rho = e->rho;
    S.w = 1.4;  // ch2/PotentialSolver.cpp:40
    S.tol = tol;
    S.max_it = max_it;
    S.iters = e->iters;
    S.res = e->res;
    S.boxes = e->sor_small;
    k_ens_sor<true><<<e->n_small, e->sor_threads, e->sor_lds, e->stream>>>(S);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

EnsSweepArgs sweep_args(pic_ensemble *e, bool resid) {
    EnsSweepArgs S;
    S.box = e->box;
    S.blk_box = resid ? e->rs_blk_box : e->sw_blk_box;
    S.blk0 = resid ? e->rs_blk0 : e->sw_blk0;
    S.phi = e->phi;
    S.rho = e->rho;
    S.done = e->done;
    S.w = 1.4;  // ch2/PotentialSolver.cpp:40
    return S;
}

// `its` iterations of every large box not yet converged, then its residual check
int launch_sor_chunk(pic_ensemble *e, int its, double tol) {
    if (!e->n_large) return PIC_OK;
    EnsSweepArgs SW = sweep_args(e, false), RS = sweep_args(e, true);
    if (e->sor_blocks) {
        // one launch per two iterations (and one for an odd last), alternating the phi
        // buffers; the residual pass reads the last one and, after an odd count of launches,
        // copies it back into phi
        EnsBlockArgs A{e->box, e->sblk, e->phi, e->phi2, e->rho, e->done, 1.4};  // ch2/PotentialSolver.cpp:40
        int launches = 0;
        for (int it = 0; it < its;) {
            if (its - it >= 2) {
                k_ens_sor_block<ENS_T><<<(unsigned)e->n_sblk, 256, sor_block_lds(ENS_T), e->stream>>>(A);
                it += ENS_T / 2;
            } else {  // an odd last iteration
                k_ens_sor_block<2><<<(unsigned)e->n_sblk, 256, sor_block_lds(2), e->stream>>>(A);
                it += 1;
            }
            launches++;
            double *next = const_cast<double *>(A.src);  // both are phi buffers of this ensemble
            A.src = A.dst;
            A.dst = next;
        }
        RS.phi = const_cast<double *>(A.src);
        RS.copy = (launches & 1) ? e->phi : nullptr;
    } else {
        for (int it = 0; it < its; it++)
            for (int color = 0; color < 2; color++)
                k_ens_sweep<<<(unsigned)e->n_sw_blk, 256, 0, e->stream>>>(SW, color);
    }
    k_ens_resid<<<(unsigned)e->n_rs_blk, 256, 0, e->stream>>>(RS, e->rs_part);
    EnsCheckArgs C{e->box, e->sor_large, e->rs_blk0, e->rs_blkn, e->rs_part, tol,
                   e->box_it, e->done, e->iters, e->ndone, e->res};
    k_ens_check<<<(unsigned)e->n_large, 256, 0, e->stream>>>(C, its);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

int launch_sor_start(pic_ensemble *e, int max_it, double tol) {
    int rc = launch_sor_small(e, max_it, tol);
    if (rc) return rc;
    if (!e->n_large) return PIC_OK;
    k_ens_sor_reset<<<(e->n_large + 255) / 256, 256, 0, e->stream>>>(e->n_large, e->sor_large, e->box_it, e->done,
                                                                    e->ndone);
    PIC_LAUNCH_CHECK();
    // iteration 0 and the reference's first residual check (it % 25 == 0 at it = 0)
    return launch_sor_chunk(e, 1, tol);
}

int launch_ef(pic_ensemble *e) {
    k_ens_ef<<<(unsigned)e->nblk_nodes, 256, 0, e->stream>>>(node_args(e), e->phi, e->ef32,
                                                           e->prec == PIC_FP64 ? e->ef64 : nullptr);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

int do_sort(pic_ensemble *e) {
    if (e->np == 0) return PIC_OK;
    EnsSortArgs A;
    A.box = e->box;
    A.rng = e->rng;
    A.nrange = (int)e->hrng.size();
    A.in = e->P;
    A.out = e->Q;
    A.cnt = e->cnt;
    PIC_HIP(hipMemsetAsync(e->cnt, 0, sizeof(uint32_t) * (e->total_cells + 1), e->stream));
    unsigned grid = (unsigned)std::min<long>((e->np + 255) / 256, 8192);
    if (e->prec == PIC_MIXED)
        k_ens_hist<PIC_MIXED><<<grid, 256, 0, e->stream>>>(A, e->np);
    else
        k_ens_hist<PIC_FP64><<<grid, 256, 0, e->stream>>>(A, e->np);
    PIC_LAUNCH_CHECK();
    size_t bytes = e->scan_tmp_bytes;
    PIC_HIP(hipcub::DeviceScan::ExclusiveSum(e->scan_tmp, bytes, e->cnt, e->off, (int)(e->total_cells + 1),
                                             e->stream));
    A.cnt = e->off;
    if (e->prec == PIC_MIXED)
        k_ens_scatter<PIC_MIXED><<<grid, 256, 0, e->stream>>>(A, e->np, e->track_ids ? 1 : 0);
    else
        k_ens_scatter<PIC_FP64><<<grid, 256, 0, e->stream>>>(A, e->np, e->track_ids ? 1 : 0);
    PIC_LAUNCH_CHECK();
    // back into the primary buffer, so the captured step keeps its pointers
    if (e->prec == PIC_MIXED) {
        PIC_HIP(hipMemcpyAsync(e->P.rec, e->Q.rec, 32 * (size_t)e->np, hipMemcpyDeviceToDevice, e->stream));
    } else {
        double *d[7] = {e->P.x, e->P.y, e->P.z, e->P.vx, e->P.vy, e->P.vz, e->P.w};
        double *s[7] = {e->Q.x, e->Q.y, e->Q.z, e->Q.vx, e->Q.vy, e->Q.vz, e->Q.w};
        for (int i = 0; i < 7; i++)
            PIC_HIP(hipMemcpyAsync(d[i], s[i], 8 * (size_t)e->np, hipMemcpyDeviceToDevice, e->stream));
    }
    if (e->track_ids)
        PIC_HIP(hipMemcpyAsync(e->P.id, e->Q.id, 4 * (size_t)e->np, hipMemcpyDeviceToDevice, e->stream));
    e->sorted = true;
    e->steps_since_sort = 0;
    return PIC_OK;
}

// The step in three parts: (1) push, density and rho, the SOR of the LDS boxes and the first
// iteration of the others; (2) chunks of 25 iterations of the boxes not yet converged,
// issued until all have converged or max_it iterations have run; (3) computeEF.
// CHUNK: the reference checks the residual every 25 iterations (ch2/PotentialSolver.cpp:46).
constexpr int SOR_CHUNK = 25;

int enqueue_first(pic_ensemble *e, double dt, int max_it, double tol) {
    int rc = launch_push(e, dt);
    if (rc) return rc;
    rc = launch_finish(e);
    if (rc) return rc;
    return launch_sor_start(e, max_it, tol);
}

int capture(pic_ensemble *e, int part, double dt, int max_it, double tol, hipGraph_t *g, hipGraphExec_t *x,
            int *nodes) {
    PIC_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = part == 0 ? enqueue_first(e, dt, max_it, tol) : part == 1 ? launch_sor_chunk(e, SOR_CHUNK, tol) : launch_ef(e);
    hipGraph_t gg = nullptr;
    hipError_t ee = hipStreamEndCapture(e->stream, &gg);
    if (rc) {
        if (gg) hipGraphDestroy(gg);
        return rc;
    }
    if (ee != hipSuccess) return hip_fail(ee, "hipStreamEndCapture", __FILE__, __LINE__);
    *g = gg;
    size_t nn = 0;
    PIC_HIP(hipGraphGetNodes(gg, nullptr, &nn));
    *nodes += (int)nn;
    PIC_HIP(hipGraphInstantiate(x, gg, nullptr, nullptr, 0));
    return PIC_OK;
}

int ensure_graph(pic_ensemble *e, double dt, int max_it, double tol) {
    if (e->exec && e->g_dt == dt && e->g_max_it == max_it && e->g_tol == tol) return PIC_OK;
    drop_graph(e);
    e->graph_nodes = 0;
    int rc = capture(e, 0, dt, max_it, tol, &e->graph, &e->exec, &e->graph_nodes);
    if (!rc && e->n_large) rc = capture(e, 1, dt, max_it, tol, &e->graph_c, &e->exec_c, &e->graph_nodes);
    if (!rc) rc = capture(e, 2, dt, max_it, tol, &e->graph_e, &e->exec_e, &e->graph_nodes);
    if (rc) {
        drop_graph(e);
        return rc;
    }
    e->g_dt = dt;
    e->g_max_it = max_it;
    e->g_tol = tol;
    return PIC_OK;
}

// one whole step, from the graphs or eagerly
int run_step(pic_ensemble *e, double dt, int max_it, double tol) {
    int rc;
    if (e->use_graph) {
        PIC_HIP(hipGraphLaunch(e->exec, e->stream));
    } else if ((rc = enqueue_first(e, dt, max_it, tol))) {
        return rc;
    }
    long its = 1;
    if (e->n_large) {
        // chunks between host reads of the converged count: 1, 2, 4, ... (at most 8)
        int burst = 1;
        for (;;) {
            PIC_HIP(hipMemcpyAsync(e->ndone_host, e->ndone, sizeof(int), hipMemcpyDeviceToHost, e->stream));
            PIC_HIP(hipStreamSynchronize(e->stream));
            if (*e->ndone_host >= e->n_large || its >= max_it) break;
            for (int b = 0; b < burst && its < max_it; b++) {
                if (e->use_graph) {
                    PIC_HIP(hipGraphLaunch(e->exec_c, e->stream));
                } else if ((rc = launch_sor_chunk(e, SOR_CHUNK, tol))) {
                    return rc;
                }
                its += SOR_CHUNK;
            }
            burst = std::min(2 * burst, 8);
        }
    }
    e->last_sor_its = its;
    if (e->use_graph) {
        PIC_HIP(hipGraphLaunch(e->exec_e, e->stream));
        return PIC_OK;
    }
    return launch_ef(e);
}

int box_range_check(pic_ensemble *e, int b, int s) {
    PIC_CHECK(b >= 0 && b < e->nbox && s >= 0 && s < e->nsp, PIC_ERR_ARG, "box or species index out of range");
    return PIC_OK;
}

}  // namespace

extern "C" {

// the blocked SOR of the boxes outside LDS (default), or one launch per half-sweep
static bool sor_blocks_env() {
    const char *t = getenv("PIC_ENS_SOR_BLOCKS");
    return !(t && atoi(t) == 0);
}

int pic_ensemble_create(int nbox, const pic_box *boxes, int nsp, const double *mass, const double *charge,
                        const long *np, int precision, int track_ids, pic_ensemble **out) {
    PIC_CHECK(out && boxes && mass && charge && np && nbox > 0 && nsp > 0, PIC_ERR_ARG, "null or empty argument");
    PIC_CHECK(precision == PIC_MIXED || precision == PIC_FP64, PIC_ERR_ARG, "bad precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device");
        return PIC_ERR_NODEVICE;
    }
    pic_ensemble *e = new pic_ensemble();
    auto fail = [&](int rc) {
        pic_ensemble_destroy(e);
        return rc;
    };
    hipGetDevice(&e->device);
    e->nbox = nbox;
    e->nsp = nsp;
    e->prec = precision;
    e->track_ids = track_ids != 0;
    e->mass.assign(mass, mass + nsp);
    e->charge.assign(charge, charge + nsp);
    e->next_id.assign((size_t)nbox * nsp, 0);
    long node0 = 0, cell0 = 0, p0 = 0;
    size_t max_small = 0;
    size_t lds_cap = 65536;
    {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) == hipSuccess && v > 0)
            lds_cap = (size_t)v;
    }
    std::vector<int> small, large;
    for (int b = 0; b < nbox; b++) {
        const pic_box &x = boxes[b];
        if (x.ni < 3 || x.nj < 3 || x.nk < 3 || x.ni > MAX_CELLS_I || x.nj > MAX_CELLS_J || x.nk > MAX_CELLS_K)
            return fail((pic::set_error("box size out of range (3 to 1024 nodes per axis)"), PIC_ERR_ARG));
        EnsBox B{};
        B.ni = x.ni;
        B.nj = x.nj;
        B.nk = x.nk;
        B.nci = x.ni - 1;
        B.ncj = x.nj - 1;
        B.nck = x.nk - 1;
        B.node0 = node0;
        B.nn = (long)x.ni * x.nj * x.nk;
        for (int a = 0; a < 3; a++) {
            B.x0[a] = x.x0[a];
            B.xm[a] = x.xm[a];
        }
        // World::setExtents (ch2/World.cpp:21-37)
        B.dh[0] = (x.xm[0] - x.x0[0]) / (x.ni - 1);
        B.dh[1] = (x.xm[1] - x.x0[1]) / (x.nj - 1);
        B.dh[2] = (x.xm[2] - x.x0[2]) / (x.nk - 1);
        node0 += B.nn;
        e->hbox.push_back(B);
        if ((size_t)B.nn * sizeof(double) + 1024 <= lds_cap) {
            small.push_back(b);
            max_small = std::max(max_small, (size_t)B.nn * sizeof(double));
        } else {
            large.push_back(b);
        }
        for (int s = 0; s < nsp; s++) {
            EnsRange R{};
            R.box = b;
            R.sp = s;
            R.p0 = p0;
            R.n = np[(long)b * nsp + s];
            if (R.n < 0) return fail((pic::set_error("negative particle count"), PIC_ERR_ARG));
            R.cell0 = cell0;
            cell0 += (long)B.nci * B.ncj * B.nck;
            p0 += R.n;
            e->hrng.push_back(R);
        }
    }
    e->total_nodes = node0;
    e->total_cells = cell0;
    e->np = p0;
    // With boxes outside LDS, the blocked launches run anyway until the largest has
    // converged, and they are latency-bound: the boxes that fit in LDS join those launches
    // rather than queueing a single-workgroup solve of their own before them.
    if (!large.empty() && !small.empty() && sor_blocks_env()) {
        large.insert(large.end(), small.begin(), small.end());
        std::sort(large.begin(), large.end());
        small.clear();
        max_small = 0;
    }
    e->n_small = (int)small.size();
    e->n_large = (int)large.size();
    e->sor_lds = max_small;
    int rc;
    PIC_HIP(hipStreamCreate(&e->stream));
    PIC_HIP(hipEventCreate(&e->ev0));
    PIC_HIP(hipEventCreate(&e->ev1));
    if ((rc = ens_alloc((void **)&e->box, sizeof(EnsBox) * nbox))) return fail(rc);
    if ((rc = ens_alloc((void **)&e->rng, sizeof(EnsRange) * e->hrng.size()))) return fail(rc);
    PIC_HIP(hipMemcpy(e->box, e->hbox.data(), sizeof(EnsBox) * nbox, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(e->rng, e->hrng.data(), sizeof(EnsRange) * e->hrng.size(), hipMemcpyHostToDevice));
    // node workgroups: box of each
    std::vector<int> bb;
    std::vector<long> b0;
    for (int b = 0; b < nbox; b++) {
        b0.push_back((long)bb.size());
        long nb = (e->hbox[b].nn + 255) / 256;
        for (long q = 0; q < nb; q++) bb.push_back(b);
    }
    e->nblk_nodes = (long)bb.size();
    if ((rc = ens_alloc((void **)&e->blk_box, sizeof(int) * bb.size()))) return fail(rc);
    if ((rc = ens_alloc((void **)&e->blk0, sizeof(long) * b0.size()))) return fail(rc);
    PIC_HIP(hipMemcpy(e->blk_box, bb.data(), sizeof(int) * bb.size(), hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(e->blk0, b0.data(), sizeof(long) * b0.size(), hipMemcpyHostToDevice));
    {
        // sweep and residual workgroups of the boxes outside LDS, box by box
        std::vector<int> swb, rsb;
        std::vector<long> sw0((size_t)nbox, 0), rs0((size_t)nbox, 0), rsn((size_t)nbox, 0);
        for (int b : large) {
            const EnsBox &B = e->hbox[b];
            long mi = B.ni - 2, rows = (long)(B.nj - 2) * (B.nk - 2);
            long per_color = (mi + 1) / 2 * rows;
            sw0[b] = (long)swb.size();
            for (long q = 0; q < (per_color + 255) / 256; q++) swb.push_back(b);
            rs0[b] = (long)rsb.size();
            rsn[b] = (mi * rows + 255) / 256;
            for (long q = 0; q < rsn[b]; q++) rsb.push_back(b);
        }
        e->n_sw_blk = (long)swb.size();
        e->n_rs_blk = (long)rsb.size();
        int **ip[2] = {&e->sw_blk_box, &e->rs_blk_box};
        std::vector<int> *iv[2] = {&swb, &rsb};
        for (int k = 0; k < 2; k++) {
            if ((rc = ens_alloc((void **)ip[k], sizeof(int) * std::max<size_t>(iv[k]->size(), 1)))) return fail(rc);
            if (!iv[k]->empty())
                PIC_HIP(hipMemcpy(*ip[k], iv[k]->data(), sizeof(int) * iv[k]->size(), hipMemcpyHostToDevice));
        }
        long **lp[3] = {&e->sw_blk0, &e->rs_blk0, &e->rs_blkn};
        std::vector<long> *lv[3] = {&sw0, &rs0, &rsn};
        for (int k = 0; k < 3; k++) {
            if ((rc = ens_alloc((void **)lp[k], sizeof(long) * (size_t)nbox))) return fail(rc);
            PIC_HIP(hipMemcpy(*lp[k], lv[k]->data(), sizeof(long) * nbox, hipMemcpyHostToDevice));
        }
        if ((rc = ens_alloc((void **)&e->rs_part, sizeof(double) * std::max<long>(e->n_rs_blk, 1)))) return fail(rc);
        std::vector<int4> sb;
        for (int b : large) {
            const EnsBox &B = e->hbox[b];
            for (int z = 1; z < B.nk - 1; z += ENS_SB)
                for (int y = 1; y < B.nj - 1; y += ENS_SB)
                    for (int x = 1; x < B.ni - 1; x += ENS_SB) sb.push_back(make_int4(b, x, y, z));
        }
        e->n_sblk = (long)sb.size();
        if ((rc = ens_alloc((void **)&e->sblk, sizeof(int4) * std::max<size_t>(sb.size(), 1)))) return fail(rc);
        if (!sb.empty()) PIC_HIP(hipMemcpy(e->sblk, sb.data(), sizeof(int4) * sb.size(), hipMemcpyHostToDevice));
        e->sor_blocks = sor_blocks_env();
        if (sor_block_lds(ENS_T) > lds_cap) e->sor_blocks = false;
        int **fl[3] = {&e->box_it, &e->done, &e->ndone};
        for (int k = 0; k < 3; k++) {
            if ((rc = ens_alloc((void **)fl[k], sizeof(int) * (size_t)(nbox + 1)))) return fail(rc);
            PIC_HIP(hipMemset(*fl[k], 0, sizeof(int) * (size_t)(nbox + 1)));
        }
        hipError_t he = hipHostMalloc((void **)&e->ndone_host, sizeof(int), hipHostMallocDefault);
        if (he != hipSuccess) return fail(hip_fail(he, "hipHostMalloc", __FILE__, __LINE__));
    }
    if ((rc = ens_alloc((void **)&e->sor_small, sizeof(int) * std::max<size_t>(small.size(), 1)))) return fail(rc);
    if ((rc = ens_alloc((void **)&e->sor_large, sizeof(int) * std::max<size_t>(large.size(), 1)))) return fail(rc);
    if (!small.empty()) PIC_HIP(hipMemcpy(e->sor_small, small.data(), sizeof(int) * small.size(), hipMemcpyHostToDevice));
    if (!large.empty()) PIC_HIP(hipMemcpy(e->sor_large, large.data(), sizeof(int) * large.size(), hipMemcpyHostToDevice));
    // fields
    long tn = e->total_nodes;
    if ((rc = ens_alloc((void **)&e->raw, sizeof(double) * tn * nsp))) return fail(rc);
    if ((rc = ens_alloc((void **)&e->den, sizeof(double) * tn * nsp))) return fail(rc);
    if ((rc = ens_alloc((void **)&e->rho, sizeof(double) * tn))) return fail(rc);
    if ((rc = ens_alloc((void **)&e->phi, sizeof(double) * tn))) return fail(rc);
    if ((rc = ens_alloc((void **)&e->ef32, sizeof(float4) * tn))) return fail(rc);
    if (precision == PIC_FP64 && (rc = ens_alloc((void **)&e->ef64, sizeof(double4) * tn))) return fail(rc);
    if ((rc = ens_alloc((void **)&e->charges_dev, sizeof(double) * nsp))) return fail(rc);
    if ((rc = ens_alloc((void **)&e->iters, sizeof(int) * nbox))) return fail(rc);
    if ((rc = ens_alloc((void **)&e->res, sizeof(double) * nbox))) return fail(rc);
    PIC_HIP(hipMemcpy(e->charges_dev, charge, sizeof(double) * nsp, hipMemcpyHostToDevice));
    PIC_HIP(hipMemset(e->raw, 0, sizeof(double) * tn * nsp));
    PIC_HIP(hipMemset(e->den, 0, sizeof(double) * tn * nsp));
    PIC_HIP(hipMemset(e->rho, 0, sizeof(double) * tn));
    PIC_HIP(hipMemset(e->phi, 0, sizeof(double) * tn));
    // phi's second buffer: the walls are 0 in both and no kernel writes them
    if ((rc = ens_alloc((void **)&e->phi2, sizeof(double) * tn))) return fail(rc);
    PIC_HIP(hipMemset(e->phi2, 0, sizeof(double) * tn));
    PIC_HIP(hipMemset(e->ef32, 0, sizeof(float4) * tn));
    if (e->ef64) PIC_HIP(hipMemset(e->ef64, 0, sizeof(double4) * tn));
    // particles: two buffers (the second is the sort's target)
    size_t per = precision == PIC_MIXED ? 32 : 56;
    size_t idb = e->track_ids ? 4 : 0;
    for (int k = 0; k < 2; k++) {
        if ((rc = ens_alloc(&e->pmem[k], (per + idb) * (size_t)std::max<long>(e->np, 1)))) return fail(rc);
        EnsParts &P = k == 0 ? e->P : e->Q;
        char *m = (char *)e->pmem[k];
        long n = std::max<long>(e->np, 1);
        if (precision == PIC_MIXED) {
            P.rec = (float4 *)m;
            m += 32 * (size_t)n;
        } else {
            double **f[7] = {&P.x, &P.y, &P.z, &P.vx, &P.vy, &P.vz, &P.w};
            for (auto q : f) {
                *q = (double *)m;
                m += 8 * (size_t)n;
            }
        }
        P.id = e->track_ids ? (uint32_t *)m : nullptr;
    }
    if ((rc = ens_alloc((void **)&e->cnt, sizeof(uint32_t) * (e->total_cells + 1)))) return fail(rc);
    if ((rc = ens_alloc((void **)&e->off, sizeof(uint32_t) * (e->total_cells + 1)))) return fail(rc);
    {
        size_t bytes = 0;
        PIC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, e->cnt, e->off, (int)(e->total_cells + 1)));
        e->scan_tmp_bytes = bytes;
        if ((rc = ens_alloc(&e->scan_tmp, bytes))) return fail(rc);
    }
    // launch shapes for this size from the tuning table (else the defaults); per-range dt
    // factors are set by pic_ensemble_step
    {
        double nd = (double)e->total_nodes, pd = (double)std::max<long>(e->np, 1);
        e->chunk_particles = std::max(256, (int)tuned("ensemble", nd, pd, "chunk", e->chunk_particles));
        e->window_nodes = std::min(8192, std::max(0, (int)tuned("ensemble", nd, pd, "window", e->window_nodes)));
        int st = (int)tuned("ensemble", nd, pd, "sor_threads", e->sor_threads);
        e->sor_threads = (st == 256 || st == 512 || st == 1024) ? st : 1024;
        e->sort_every = std::max(0, (int)tuned("ensemble", nd, pd, "sort_every", e->sort_every));
    }
    if ((rc = build_chunks(e))) return fail(rc);
    *out = e;
    return PIC_OK;
}

int pic_ensemble_destroy(pic_ensemble *e) {
    if (!e) return PIC_OK;
    if (e->stream) hipStreamSynchronize(e->stream);
    drop_graph(e);
    void *ptrs[] = {e->box, e->rng, e->chunk, e->blk_box, e->blk0, e->pmem[0], e->pmem[1], e->raw, e->den, e->rho,
                    e->phi, e->charges_dev, e->ef32, e->ef64, e->iters, e->res, e->sor_small, e->sor_large, e->cnt,
                    e->off, e->scan_tmp, e->diag_part, e->sw_blk_box, e->rs_blk_box, e->sw_blk0, e->rs_blk0,
                    e->rs_blkn, e->rs_part, e->box_it, e->done, e->ndone, e->sblk, e->phi2};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (e->ndone_host) hipHostFree(e->ndone_host);
    if (e->ev0) hipEventDestroy(e->ev0);
    if (e->ev1) hipEventDestroy(e->ev1);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return PIC_OK;
}

int pic_ensemble_set_particles(pic_ensemble *e, int b, int s, long n, const double *pos, const double *vel,
                               const double *mpw) {
    PIC_CHECK(e, PIC_ERR_ARG, "null argument");
    int rc = box_range_check(e, b, s);
    if (rc) return rc;
    const EnsRange &R = e->hrng[(size_t)b * e->nsp + s];
    PIC_CHECK(n == R.n, PIC_ERR_ARG, "particle count differs from the count given at creation");
    if (n == 0) return PIC_OK;
    PIC_CHECK(pos && vel && mpw, PIC_ERR_ARG, "null argument");
    double *d = nullptr;
    PIC_HIP(hipMalloc(&d, sizeof(double) * 7 * (size_t)n));
    PIC_HIP(hipMemcpyAsync(d, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, e->stream));
    PIC_HIP(hipMemcpyAsync(d + 3 * n, vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice, e->stream));
    PIC_HIP(hipMemcpyAsync(d + 6 * n, mpw, sizeof(double) * n, hipMemcpyHostToDevice, e->stream));
    k_ens_upload<<<(unsigned)((n + 255) / 256), 256, 0, e->stream>>>(e->hbox[b], R.p0, n, d, d + 3 * n, d + 6 * n,
                                                                   e->prec == PIC_MIXED, e->P, 0u,
                                                                   e->track_ids ? 1 : 0);
    PIC_LAUNCH_CHECK();
    PIC_HIP(hipStreamSynchronize(e->stream));
    PIC_HIP(hipFree(d));
    e->sorted = false;
    return PIC_OK;
}

int pic_ensemble_load_box(pic_ensemble *e, int b, int s, const double x1[3], const double x2[3], double num_den,
                          uint64_t seed) {
    PIC_CHECK(e && x1 && x2, PIC_ERR_ARG, "null argument");
    int rc = box_range_check(e, b, s);
    if (rc) return rc;
    const EnsRange &R = e->hrng[(size_t)b * e->nsp + s];
    if (R.n == 0) return PIC_OK;
    // loadParticlesBox (ch2/Species.cpp:74-97): mpw = num_den * V / num_mp
    double V = (x2[0] - x1[0]) * (x2[1] - x1[1]) * (x2[2] - x1[2]);
    double w = num_den * V / (double)R.n;
    uint32_t tag = (uint32_t)((size_t)b * e->nsp + s);
    k_ens_load<<<(unsigned)((R.n + 255) / 256), 256, 0, e->stream>>>(e->hbox[b], R.p0, R.n, x1[0], x1[1], x1[2], x2[0],
                                                                    x2[1], x2[2], w, (uint32_t)seed,
                                                                    (uint32_t)(seed >> 32), tag, e->prec == PIC_MIXED,
                                                                    e->P, e->track_ids ? 1 : 0);
    PIC_LAUNCH_CHECK();
    e->sorted = false;
    return PIC_OK;
}

int pic_ensemble_get_particles(pic_ensemble *e, int b, int s, double *pos, double *vel, double *mpw, uint32_t *ids) {
    PIC_CHECK(e && pos && vel && mpw, PIC_ERR_ARG, "null argument");
    PIC_CHECK(!ids || e->track_ids, PIC_ERR_ARG, "ensemble does not track ids");
    int rc = box_range_check(e, b, s);
    if (rc) return rc;
    const EnsRange &R = e->hrng[(size_t)b * e->nsp + s];
    if (R.n == 0) return PIC_OK;
    long n = R.n;
    double *d = nullptr;
    PIC_HIP(hipMalloc(&d, sizeof(double) * 7 * (size_t)n + 4 * (size_t)n));
    uint32_t *did = (uint32_t *)(d + 7 * n);
    k_ens_download<<<(unsigned)((n + 255) / 256), 256, 0, e->stream>>>(e->hbox[b], R.p0, n, e->prec == PIC_MIXED,
                                                                     e->P, d, d + 3 * n, d + 6 * n,
                                                                     ids ? did : nullptr);
    PIC_LAUNCH_CHECK();
    PIC_HIP(hipMemcpyAsync(pos, d, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, e->stream));
    PIC_HIP(hipMemcpyAsync(vel, d + 3 * n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, e->stream));
    PIC_HIP(hipMemcpyAsync(mpw, d + 6 * n, sizeof(double) * n, hipMemcpyDeviceToHost, e->stream));
    if (ids) PIC_HIP(hipMemcpyAsync(ids, did, 4 * (size_t)n, hipMemcpyDeviceToHost, e->stream));
    PIC_HIP(hipStreamSynchronize(e->stream));
    PIC_HIP(hipFree(d));
    return PIC_OK;
}

int pic_ensemble_sort(pic_ensemble *e) {
    PIC_CHECK(e, PIC_ERR_ARG, "null argument");
    return do_sort(e);
}

int pic_ensemble_set_tuning(pic_ensemble *e, int chunk_particles, int window_nodes, int sor_threads,
                            int sort_every, int use_graph) {
    PIC_CHECK(e, PIC_ERR_ARG, "null argument");
    PIC_CHECK(chunk_particles >= 256 && chunk_particles <= (1 << 20), PIC_ERR_ARG, "chunk_particles out of range");
    PIC_CHECK(window_nodes >= 0 && (size_t)window_nodes * 8 <= 65536, PIC_ERR_ARG,
              "window_nodes: 0 to 8192 (64 KB of LDS)");
    PIC_CHECK(sor_threads == 256 || sor_threads == 512 || sor_threads == 1024, PIC_ERR_ARG,
              "sor_threads: 256, 512 or 1024");
    PIC_CHECK(sort_every >= 0, PIC_ERR_ARG, "negative sort interval");
    PIC_HIP(hipStreamSynchronize(e->stream));
    e->window_nodes = window_nodes;
    e->sor_threads = sor_threads;
    e->sort_every = sort_every;
    e->use_graph = use_graph != 0;
    if (chunk_particles != e->chunk_particles) {
        e->chunk_particles = chunk_particles;
        int rc = build_chunks(e);
        if (rc) return rc;
    }
    return drop_graph(e);
}

// Initial potential of every box (PotentialSolver::solve before the loop, ch2/Main.cpp:38-42):
// deposit, rho, SOR, computeEF, without moving the particles.
int pic_ensemble_solve_initial(pic_ensemble *e, int max_it, double tol) {
    PIC_CHECK(e, PIC_ERR_ARG, "null argument");
    // a push with dt = 0 moves nothing (v += E*0, x += v*0) and deposits
    for (auto &R : e->hrng) {
        R.s64 = 0.0;
        R.qm_dt = 0.0f;
        R.dt_dh[0] = R.dt_dh[1] = R.dt_dh[2] = 0.0f;
    }
    PIC_HIP(hipMemcpyAsync(e->rng, e->hrng.data(), sizeof(EnsRange) * e->hrng.size(), hipMemcpyHostToDevice,
                           e->stream));
    drop_graph(e);
    bool g = e->use_graph;
    e->use_graph = false;
    int rc = run_step(e, 0.0, max_it, tol);
    e->use_graph = g;
    if (rc) return rc;
    e->g_dt = -1;
    PIC_HIP(hipStreamSynchronize(e->stream));
    return PIC_OK;
}

// nsteps steps of the ch2 loop for every box: advance (all species) -> computeNumberDensity ->
// computeChargeDensity -> solve -> computeEF, replayed from one hipGraph; a counting sort
// every sort_every steps
int pic_ensemble_step(pic_ensemble *e, int nsteps, double dt, int max_it, double tol) {
    PIC_CHECK(e, PIC_ERR_ARG, "null argument");
    PIC_CHECK(dt > 0 && nsteps >= 0 && max_it >= 1, PIC_ERR_ARG, "bad step arguments");
    if (e->g_dt != dt) {
        for (auto &R : e->hrng) {
            const EnsBox &B = e->hbox[R.box];
            R.s64 = dt * e->charge[R.sp] / e->mass[R.sp];
            R.qm_dt = (float)(dt * e->charge[R.sp] / e->mass[R.sp]);
            for (int a = 0; a < 3; a++) R.dt_dh[a] = (float)(dt / B.dh[a]);
        }
        PIC_HIP(hipMemcpyAsync(e->rng, e->hrng.data(), sizeof(EnsRange) * e->hrng.size(), hipMemcpyHostToDevice,
                               e->stream));
    }
    int rc = PIC_OK;
    if (e->use_graph) {
        rc = ensure_graph(e, dt, max_it, tol);
        if (rc) return rc;
    } else {
        e->g_dt = dt;  // the dt the range table holds
    }
    PIC_HIP(hipEventRecord(e->ev0, e->stream));
    for (int s = 0; s < nsteps; s++) {
        if (e->sort_every > 0 && (!e->sorted || e->steps_since_sort >= e->sort_every)) {
            rc = do_sort(e);
            if (rc) return rc;
        }
        rc = run_step(e, dt, max_it, tol);
        if (rc) return rc;
        e->steps_done++;
        e->steps_since_sort++;
    }
    PIC_HIP(hipEventRecord(e->ev1, e->stream));
    PIC_HIP(hipEventSynchronize(e->ev1));
    float ms = 0;
    PIC_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    e->last_step_ms = nsteps > 0 ? ms / nsteps : 0.0;
    return PIC_OK;
}

// which: 0 phi, 1 rho, 2 ef (3 per node), 10 + s den of species s
int pic_ensemble_get_field(pic_ensemble *e, int b, int which, double *out) {
    PIC_CHECK(e && out, PIC_ERR_ARG, "null argument");
    PIC_CHECK(b >= 0 && b < e->nbox, PIC_ERR_ARG, "box index out of range");
    const EnsBox &B = e->hbox[b];
    PIC_HIP(hipStreamSynchronize(e->stream));
    if (which == 0) {
        PIC_HIP(hipMemcpy(out, e->phi + B.node0, sizeof(double) * B.nn, hipMemcpyDeviceToHost));
    } else if (which == 1) {
        PIC_HIP(hipMemcpy(out, e->rho + B.node0, sizeof(double) * B.nn, hipMemcpyDeviceToHost));
    } else if (which == 2) {
        if (e->ef64) {
            std::vector<double4> t((size_t)B.nn);
            PIC_HIP(hipMemcpy(t.data(), e->ef64 + B.node0, sizeof(double4) * B.nn, hipMemcpyDeviceToHost));
            for (long u = 0; u < B.nn; u++) {
                out[3 * u] = t[u].x;
                out[3 * u + 1] = t[u].y;
                out[3 * u + 2] = t[u].z;
            }
        } else {
            std::vector<float4> t((size_t)B.nn);
            PIC_HIP(hipMemcpy(t.data(), e->ef32 + B.node0, sizeof(float4) * B.nn, hipMemcpyDeviceToHost));
            for (long u = 0; u < B.nn; u++) {
                out[3 * u] = t[u].x;
                out[3 * u + 1] = t[u].y;
                out[3 * u + 2] = t[u].z;
            }
        }
    } else if (which >= 10 && which < 10 + e->nsp) {
        PIC_HIP(hipMemcpy(out, e->den + (long)(which - 10) * e->total_nodes + B.node0, sizeof(double) * B.nn,
                          hipMemcpyDeviceToHost));
    } else {
        set_error("unknown ensemble field");
        return PIC_ERR_ARG;
    }
    return PIC_OK;
}

// per box: SOR iterations of the last solve (negative: not converged) and its residual
int pic_ensemble_solver_status(pic_ensemble *e, int *iters, double *residual) {
    PIC_CHECK(e, PIC_ERR_ARG, "null argument");
    PIC_HIP(hipStreamSynchronize(e->stream));
    if (iters) PIC_HIP(hipMemcpy(iters, e->iters, sizeof(int) * e->nbox, hipMemcpyDeviceToHost));
    if (residual) PIC_HIP(hipMemcpy(residual, e->res, sizeof(double) * e->nbox, hipMemcpyDeviceToHost));
    return PIC_OK;
}

// out[5 * (b * nsp + s) + q]: sum w, sum w v (3), sum w |v|^2 of each range
int pic_ensemble_diag(pic_ensemble *e, double *out) {
    PIC_CHECK(e && out, PIC_ERR_ARG, "null argument");
    size_t nr = e->hrng.size();
    std::fill(out, out + 5 * nr, 0.0);
    if (e->hchunk.empty()) return PIC_OK;
    k_ens_diag<<<(unsigned)e->hchunk.size(), 256, 0, e->stream>>>(e->rng, e->chunk, e->P, e->prec == PIC_MIXED,
                                                                e->diag_part);
    PIC_LAUNCH_CHECK();
    std::vector<double> part(5 * e->hchunk.size());
    PIC_HIP(hipMemcpyAsync(part.data(), e->diag_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                           e->stream));
    PIC_HIP(hipStreamSynchronize(e->stream));
    for (size_t c = 0; c < e->hchunk.size(); c++)
        for (int q = 0; q < 5; q++) out[5 * e->hchunk[c].range + q] += part[5 * c + q];
    return PIC_OK;
}

// out: [0] ms per step of the last pic_ensemble_step (graph replays + sorts), [1] kernels
// in the step's graph, [2] particles, [3] nodes, [4] chunks (push workgroups)
int pic_ensemble_stats(pic_ensemble *e, double out[5]) {
    PIC_CHECK(e && out, PIC_ERR_ARG, "null argument");
    out[0] = e->last_step_ms;
    out[1] = e->graph_nodes;
    out[2] = (double)e->np;
    out[3] = (double)e->total_nodes;
    out[4] = (double)e->hchunk.size();
    return PIC_OK;
}

}  // extern "C"
