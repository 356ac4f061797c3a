// PotentialSolver: linear SOR (ch2), nonlinear Boltzmann SOR (ch3 solveGS), Newton-PCG
// (ch3 solveNRPCG + solvePCGLinear + solveGSLinear), solveQN and computeEF, on the device.
//
// Ordering.  The reference sweeps lexicographically (i outer, k inner).  On the GPU the
// sweeps are red-black: same stencil, same relaxation factor, same residual definition and
// stopping test, a different iterate path.  Converged solutions agree to the tolerance.
// All arithmetic is FP64; dots and norms are FP64 sums with a fixed reduction tree.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>

#include "picamd_internal.h"
#include "stencil.h"

#pragma clang fp contract(off)

namespace pic {

int compute_ef(pic_world *w, bool want32, bool want64);
int solve_qn(pic_world *w, double phi0, double Te0, double n0);

struct SorParams {
    double idx2, idy2, idz2;
    double denom;  // 2*idx2 + 2*idy2 + 2*idz2
    double w;      // relaxation factor
    double phi0, Te0, n0;
    double tol;
    int max_it;
};


// ---- one node update -------------------------------------------------------------------
// linear SOR, interior nodes only (ch2/PotentialSolver.cpp:32-42)
__device__ inline double sor_lin_new(const double *phi, double rho, long u, long sj, long sk,
                                     const SorParams &S) {
    double phi_new = (rho / EPS_0 + S.idx2 * (phi[u - 1] + phi[u + 1]) + S.idy2 * (phi[u - sj] + phi[u + sj]) +
                      S.idz2 * (phi[u - sk] + phi[u + sk])) /
                     S.denom;
    return phi[u] + S.w * (phi_new - phi[u]);
}

__device__ inline double sor_lin_res(const double *phi, double rho, long u, long sj, long sk,
                                     const SorParams &S) {
    return -phi[u] * S.denom + rho / EPS_0 + S.idx2 * (phi[u - 1] + phi[u + 1]) +
           S.idy2 * (phi[u - sj] + phi[u + sj]) + S.idz2 * (phi[u - sk] + phi[u + sk]);
}

// nonlinear Boltzmann SOR (ch3/ver2/PotentialSolver.cpp:349-383); returns the new value
__device__ inline double sor_boltz_new(const Grid &g, const double *phi, double rho, long u, int i, int j,
                                       int k, const SorParams &S) {
    long sj = g.ni, sk = (long)g.ni * g.nj;
    if (i == 0) return phi[u + 1];
    if (i == g.ni - 1) return phi[u - 1];
    if (j == 0) return phi[u + sj];
    if (j == g.nj - 1) return phi[u - sj];
    if (k == 0) return phi[u + sk];
    if (k == g.nk - 1) return phi[u - sk];
    double ne = S.n0 * exp((phi[u] - S.phi0) / S.Te0);
    double phi_new = ((rho - QE * ne) / EPS_0 + S.idx2 * (phi[u - 1] + phi[u + 1]) +
                      S.idy2 * (phi[u - sj] + phi[u + sj]) + S.idz2 * (phi[u - sk] + phi[u + sk])) /
                     S.denom;
    return phi[u] + S.w * (phi_new - phi[u]);
}

// residual of solveGS (ch3/ver2/PotentialSolver.cpp:387-419)
__device__ inline double sor_boltz_res(const Grid &g, const double *phi, double rho, long u, int i, int j,
                                       int k, const SorParams &S) {
    long sj = g.ni, sk = (long)g.ni * g.nj;
    if (i == 0) return phi[u] - phi[u + 1];
    if (i == g.ni - 1) return phi[u] - phi[u - 1];
    if (j == 0) return phi[u] - phi[u + sj];
    if (j == g.nj - 1) return phi[u] - phi[u - sj];
    if (k == 0) return phi[u] - phi[u + sk];
    if (k == g.nk - 1) return phi[u] - phi[u - sk];
    double ne = S.n0 * exp((phi[u] - S.phi0) / S.Te0);
    return -phi[u] * S.denom + (rho - QE * ne) / EPS_0 + S.idx2 * (phi[u - 1] + phi[u + 1]) +
           S.idy2 * (phi[u - sj] + phi[u + sj]) + S.idz2 * (phi[u - sk] + phi[u + sk]);
}

// ---- single-workgroup solver with phi resident in LDS (small grids) -------------------
// The whole iteration (sweeps, residual every 25 iterations, stopping test) runs in one
// launch: no host round trip and no launch per sweep.
constexpr int LDS_THREADS = 1024;

template <bool BOLTZ>
__global__ void __launch_bounds__(LDS_THREADS)
    k_sor_lds(Grid g, double *phi, const double *__restrict__ rho, const uint8_t *__restrict__ oid,
              SorParams S, DevScalars *out) {
    extern __shared__ double sphi[];
    __shared__ double sh[LDS_THREADS / 64];
    __shared__ int s_conv;
    long nn = g.nn;
    long sj = g.ni, sk = (long)g.ni * g.nj;
    for (long u = threadIdx.x; u < nn; u += LDS_THREADS) sphi[u] = phi[u];
    if (threadIdx.x == 0) s_conv = 0;
    __syncthreads();
    double L2 = 0;
    int it;
    for (it = 0; it < S.max_it; it++) {
        for (int color = 0; color < 2; color++) {
            for (long u = threadIdx.x; u < nn; u += LDS_THREADS) {
                int i, j, k;
                node_ijk(g, u, i, j, k);
                if (((i + j + k) & 1) != color) continue;
                if (BOLTZ) {
                    if (oid[u] > 0) continue;
                    sphi[u] = sor_boltz_new(g, sphi, rho[u], u, i, j, k, S);
                } else {
                    if (i == 0 || j == 0 || k == 0 || i == g.ni - 1 || j == g.nj - 1 || k == g.nk - 1) continue;
                    sphi[u] = sor_lin_new(sphi, rho[u], u, sj, sk, S);
                }
            }
            __syncthreads();
        }
        if (it % 25 == 0) {
            double sum = 0;
            for (long u = threadIdx.x; u < nn; u += LDS_THREADS) {
                int i, j, k;
                node_ijk(g, u, i, j, k);
                double R;
                if (BOLTZ) {
                    if (oid[u] > 0) continue;
                    R = sor_boltz_res(g, sphi, rho[u], u, i, j, k, S);
                } else {
                    if (i == 0 || j == 0 || k == 0 || i == g.ni - 1 || j == g.nj - 1 || k == g.nk - 1) continue;
                    R = sor_lin_res(sphi, rho[u], u, sj, sk, S);
                }
                sum += R * R;
            }
            double tot = block_sum(sum, sh);
            if (threadIdx.x == 0) {
                L2 = sqrt(tot / (double)(g.ni * g.nj * g.nk));
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
    for (long u = threadIdx.x; u < nn; u += LDS_THREADS) phi[u] = sphi[u];
    if (threadIdx.x == 0) {
        out->red[0] = L2;
        out->done = s_conv;
        out->iters = s_conv ? it + 1 : it;
    }
}

// ---- multi-workgroup red-black sweeps (large grids) -------------------------------------
template <bool BOLTZ>
__global__ void __launch_bounds__(256)
    k_sor_sweep(Grid g, double *phi, const double *__restrict__ rho, const uint8_t *__restrict__ oid,
                SorParams S, int color, const DevScalars *sc) {
    if (sc->done) return;
    // thread -> pair of nodes along i; the one of this colour is updated
    long half = (g.ni + 1) / 2;
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long rows = (long)g.nj * g.nk;
    if (t >= half * rows) return;
    long row = t / half;
    int j = (int)(row % g.nj), k = (int)(row / g.nj);
    int i = (int)(t % half) * 2;
    if (((i + j + k) & 1) != color) i += 1;
    if (i >= g.ni) return;
    long u = i + g.ni * (long)j + (long)g.ni * g.nj * k;
    long sj = g.ni, sk = (long)g.ni * g.nj;
    if (BOLTZ) {
        if (oid[u] > 0) return;
        phi[u] = sor_boltz_new(g, phi, rho[u], u, i, j, k, S);
    } else {
        if (i == 0 || j == 0 || k == 0 || i == g.ni - 1 || j == g.nj - 1 || k == g.nk - 1) return;
        phi[u] = sor_lin_new(phi, rho[u], u, sj, sk, S);
    }
}

template <bool BOLTZ>
__global__ void __launch_bounds__(256)
    k_sor_residual(Grid g, const double *phi, const double *__restrict__ rho, const uint8_t *__restrict__ oid,
                   SorParams S, double *partials, const DevScalars *sc) {
    __shared__ double sh[4];
    if (sc->done) return;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double R2 = 0;
    if (u < g.nn) {
        int i, j, k;
        node_ijk(g, u, i, j, k);
        long sj = g.ni, sk = (long)g.ni * g.nj;
        if (BOLTZ) {
            if (!(oid[u] > 0)) {
                double R = sor_boltz_res(g, phi, rho[u], u, i, j, k, S);
                R2 = R * R;
            }
        } else if (!(i == 0 || j == 0 || k == 0 || i == g.ni - 1 || j == g.nj - 1 || k == g.nk - 1)) {
            double R = sor_lin_res(phi, rho[u], u, sj, sk, S);
            R2 = R * R;
        }
    }
    double s = block_sum(R2, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// one workgroup: sums nb partials per slot in a fixed order

__global__ void __launch_bounds__(1024)
    k_sor_check(const double *partials, int nb, double denom_nodes, double tol, int it, DevScalars *sc) {
    __shared__ double sh[16];
    if (sc->done) return;
    double tot = fixed_sum(partials, nb, sh);
    if (threadIdx.x == 0) {
        double L2 = sqrt(tot / denom_nodes);
        sc->red[0] = L2;
        if (L2 < tol) {
            sc->done = 1;
            sc->iters = it + 1;
        }
    }
}

// ---- Newton-PCG, matrix-free -----------------------------------------------------------
// Row of A (buildMatrix, ch3/ver2/PotentialSolver.cpp:146-201) applied to v in the slot
// order of the reference matrix; diag is the row's diagonal (A's, or J = A - diag(P)).




// A's diagonal
__device__ inline double a_diag(const PcgParams &Q, int type, int i, int j) {
    if (type == NT_DIR) return 1.0;
    if (type == NT_REG) return -2.0 * (Q.idx2 + Q.idy2 + Q.idz2);
    const Grid &g = Q.g;
    if (i == 0 || i == g.ni - 1) return Q.idx;
    if (j == 0 || j == g.nj - 1) return Q.idy;
    return Q.idz;
}

__device__ inline double row_apply(const PcgParams &Q, int type, long u, int i, int j, int k, double diag,
                                   const double *v) {
    const Grid &g = Q.g;
    long sj = g.ni, sk = (long)g.ni * g.nj;
    if (type == NT_DIR) return diag * v[u];
    if (type == NT_REG) {
        double s = 0;
        s += Q.idz2 * v[u - sk];
        s += Q.idy2 * v[u - sj];
        s += Q.idx2 * v[u - 1];
        s += diag * v[u];
        s += Q.idx2 * v[u + 1];
        s += Q.idy2 * v[u + sj];
        s += Q.idz2 * v[u + sk];
        return s;
    }
    double s = 0;
    s += diag * v[u];
    if (i == 0) s += -Q.idx * v[u + 1];
    else if (i == g.ni - 1) s += -Q.idx * v[u - 1];
    else if (j == 0) s += -Q.idy * v[u + sj];
    else if (j == g.nj - 1) s += -Q.idy * v[u - sj];
    else if (k == 0) s += -Q.idz * v[u + sk];
    else s += -Q.idz * v[u - sk];
    return s;
}

// F = A x - b, minus the Boltzmann term on regular nodes, and the Jacobian diagonal
// jd = A(u,u) - P(u), P = d(ne term)/dphi (ch3/ver2/PotentialSolver.cpp:240-266)
__global__ void __launch_bounds__(256)
    k_nr_F(PcgParams Q, const double *x, const double *rho, const uint8_t *oid, double *F, double *jd) {
    long u = xcd_block() * blockDim.x + threadIdx.x;
    const Grid &g = Q.g;
    if (u >= g.nn) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    int t = node_type(g, oid, u, i, j, k);
    double ad = a_diag(Q, t, i, j);
    double Ax = row_apply(Q, t, u, i, j, k, ad, x);
    double b;
    if (t == NT_NEU) b = 0;
    else if (t == NT_DIR) b = x[u];
    else b = -rho[u] / EPS_0;
    double f = Ax - b;
    double p = 0;
    if (t == NT_REG) {
        f -= QE * Q.n0 * exp((x[u] - Q.phi0) / Q.Te0) / EPS_0;
        p = Q.n0 * QE / (EPS_0 * Q.Te0) * exp((x[u] - Q.phi0) / Q.Te0);
    }
    F[u] = f;
    jd[u] = ad - p;  // Matrix::diagSubtract (:60-64)
}

// ---- PCG (solvePCGLinear, :299-331) in two vector passes per iteration -----------------
// Iteration k of the reference:  z = J d; a = g.s; b = d.z; y += (a/b) d; g += (a/b) z;
// s = M g; a' = g.s; d = (a'/a) d - s; stop when RMS(g) < tol.  Here:
//   pass A:  d_new = mu d_old - M g (recomputed at the 7 stencil points), y += lam d_old
//            (the previous iteration's deferred update), z = J d_new, partials d_new.z
//   pass B:  g += (a/b) z, partials g.(M g) and g.g
// with M g = (1.0/jd) g as Matrix::invDiagonal.  d ping-pongs between two buffers so that
// pass A reads only old values.  The last deferred y update is applied by k_pcg_finish.

// The streaming PCG passes (init, pass B) run on a fixed grid of at most PCG_GRID
// workgroups, each walking a contiguous chunk of whole 256-node strides and writing one
// partial, so the scalar step reduces at most PCG_GRID values.  Workgroups are dispatched
// round-robin over the 8 XCDs; with a multiple of 8 workgroups, XCD x gets chunks
// [x G/8, (x+1) G/8), a slab of consecutive k planes.  The stencil pass A keeps one node per
// thread (its gathers need the latency hiding of many resident wavefronts); its per-block
// partials are first folded to FOLD_GRID values by k_fold.
constexpr unsigned PCG_GRID = 2048;
constexpr unsigned FOLD_GRID = 128;

// out[b] = sum of the b-th of FOLD_GRID contiguous slices of in[0, n), fixed order
__global__ void __launch_bounds__(256) k_fold(const double *in, long n, double *out, const DevScalars *sc) {
    __shared__ double sh[4];
    if (sc->done) return;
    long per = (n + gridDim.x - 1) / gridDim.x;
    long a = blockIdx.x * per, b = a + per < n ? a + per : n;
    double s = 0;
    for (long i = a + threadIdx.x; i < b; i += 256) s += in[i];
    double r = block_sum(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

struct NodeChunk {
    long u0, u1;
};

__device__ inline NodeChunk pcg_chunk(long nn) {
    unsigned G = gridDim.x, b = blockIdx.x;
    unsigned c = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
    long per = (nn + G - 1) / G;
    per = (per + 255) / 256 * 256;
    long u0 = (long)c * per;
    long u1 = u0 + per < nn ? u0 + per : nn;
    return NodeChunk{u0, u1};
}

// g = J y - F; d_0 = 0; partials of g.s  (s = M g)
__global__ void __launch_bounds__(256)
    k_pcg_init(PcgParams Q, const uint8_t *oid, const double *jd, const double *y, const double *F, double *gv,
               double *d0, double *partials) {
    __shared__ double sh[4];
    const Grid &g = Q.g;
    NodeChunk ch = pcg_chunk(g.nn);
    double gs = 0;
    for (long u = ch.u0 + threadIdx.x; u < ch.u1; u += 256) {
        int i, j, k;
        node_ijk(g, u, i, j, k);
        int t = node_type(g, oid, u, i, j, k);
        double gg = row_apply(Q, t, u, i, j, k, jd[u], y) - F[u];
        double s = (1.0 / jd[u]) * gg;
        gv[u] = gg;
        d0[u] = 0;
        gs += gg * s;
    }
    double r = block_sum(gs, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__device__ inline double dnew_at(const double *d_old, const double *gv, const double *jd, long v, double mu) {
    return mu * d_old[v] - (1.0 / jd[v]) * gv[v];
}

__global__ void __launch_bounds__(256)
    k_pcg_a(PcgParams Q, const uint8_t *oid, const double *jd, const double *gv, const double *d_old,
            double *d_new, double *y, double *z, double *partials, const DevScalars *sc) {
    __shared__ double sh[4];
    if (sc->done) return;
    const Grid &g = Q.g;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    double mu = sc->red[6], lam = sc->red[7];
    double dz = 0;
    if (u < g.nn) {
        int i, j, k;
        node_ijk(g, u, i, j, k);
        int t = node_type(g, oid, u, i, j, k);
        long sj = g.ni, sk = (long)g.ni * g.nj;
        y[u] = y[u] + lam * d_old[u];
        double dc = dnew_at(d_old, gv, jd, u, mu);
        double zz;
        if (t == NT_DIR) {
            zz = jd[u] * dc;
        } else if (t == NT_REG) {
            zz = 0;
            zz += Q.idz2 * dnew_at(d_old, gv, jd, u - sk, mu);
            zz += Q.idy2 * dnew_at(d_old, gv, jd, u - sj, mu);
            zz += Q.idx2 * dnew_at(d_old, gv, jd, u - 1, mu);
            zz += jd[u] * dc;
            zz += Q.idx2 * dnew_at(d_old, gv, jd, u + 1, mu);
            zz += Q.idy2 * dnew_at(d_old, gv, jd, u + sj, mu);
            zz += Q.idz2 * dnew_at(d_old, gv, jd, u + sk, mu);
        } else {
            long nb;
            double a;
            if (i == 0) { nb = u + 1; a = -Q.idx; }
            else if (i == g.ni - 1) { nb = u - 1; a = -Q.idx; }
            else if (j == 0) { nb = u + sj; a = -Q.idy; }
            else if (j == g.nj - 1) { nb = u - sj; a = -Q.idy; }
            else if (k == 0) { nb = u + sk; a = -Q.idz; }
            else { nb = u - sk; a = -Q.idz; }
            zz = 0;
            zz += jd[u] * dc;
            zz += a * dnew_at(d_old, gv, jd, nb, mu);
        }
        d_new[u] = dc;
        z[u] = zz;
        dz = dc * zz;
    }
    double r = block_sum(dz, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ void __launch_bounds__(256)
    k_pcg_b(PcgParams Q, const double *jd, double *gv, const double *z, double *partials, int nb,
            const DevScalars *sc) {
    __shared__ double sh[4];
    if (sc->done) return;
    NodeChunk ch = pcg_chunk(Q.g.nn);
    double alpha = sc->red[0], beta = sc->red[1];
    double gs = 0, g2 = 0;
    for (long u = ch.u0 + threadIdx.x; u < ch.u1; u += 256) {
        double gg = gv[u] + (alpha / beta) * z[u];
        gv[u] = gg;
        double s = (1.0 / jd[u]) * gg;
        gs += gg * s;
        g2 += gg * gg;
    }
    double r1 = block_sum(gs, sh);
    double r2 = block_sum(g2, sh);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = r1;
        partials[nb + blockIdx.x] = r2;
    }
}

// the deferred y update of the last completed iteration
template <typename TD>
__global__ void __launch_bounds__(256) k_pcg_finish(long nn, const TD *d, double *y, const DevScalars *sc) {
    long u = xcd_block() * blockDim.x + threadIdx.x;
    double lam = sc->red[7];
    if (u < nn) y[u] = y[u] + lam * (double)d[u];
}

// scalar steps (one workgroup): red[0] alpha, red[1] beta, red[2] g.g, red[3] alpha',
// red[4] RMS(g), red[6] mu for the next pass A, red[7] lam (pending y update)
// mode 0: after init;  mode 1: after pass A;  mode 2: after pass B (convergence test)
__global__ void __launch_bounds__(1024)
    k_pcg_scalar(const double *partials, int nb, int mode, double nu, double tol, int it, DevScalars *sc) {
    __shared__ double sh[16];
    if (mode != 0 && sc->done) return;
    double s0 = fixed_sum(partials, nb, sh);
    double s1 = mode == 2 ? fixed_sum(partials + nb, nb, sh) : 0.0;
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        sc->red[0] = s0;
        sc->red[6] = 0.0;  // d_0 = -s_0
        sc->red[7] = 0.0;
        sc->iters = 0;
    } else if (mode == 1) {
        sc->red[1] = s0;
    } else {
        double alpha = sc->red[0], beta = sc->red[1];
        sc->red[7] = alpha / beta;  // y += (alpha/beta) d, applied by the next pass A
        sc->red[3] = s0;
        sc->red[2] = s1;
        sc->red[6] = s0 / alpha;   // d = (alpha'/alpha) d - s
        sc->red[0] = s0;
        double l2 = sqrt(s1 / nu);  // vec::norm is RMS (:114-121)
        sc->red[4] = l2;
        sc->iters = it + 1;
        if (l2 < tol) sc->done = 1;
    }
}

// red-black Gauss-Seidel (w = 1) on J y = F: the fallback solveGSLinear (:433-461)
__global__ void __launch_bounds__(256)
    k_gsl_sweep(PcgParams Q, const uint8_t *oid, const double *jdv, double *y, const double *F, int color,
                const DevScalars *sc) {
    if (sc->done) return;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const Grid &g = Q.g;
    if (u >= g.nn) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    if (((i + j + k) & 1) != color) return;
    int t = node_type(g, oid, u, i, j, k);
    double jd = jdv[u];
    double S = row_apply(Q, t, u, i, j, k, jd, y) - jd * y[u];
    double phi_new = (F[u] - S) / jd;
    y[u] = y[u] + 1. * (phi_new - y[u]);
}

__global__ void __launch_bounds__(256)
    k_gsl_residual(PcgParams Q, const uint8_t *oid, const double *jdv, const double *y, const double *F,
                   double *partials, const DevScalars *sc) {
    __shared__ double sh[4];
    if (sc->done) return;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const Grid &g = Q.g;
    double r2 = 0;
    if (u < g.nn) {
        int i, j, k;
        node_ijk(g, u, i, j, k);
        int t = node_type(g, oid, u, i, j, k);
        double R = row_apply(Q, t, u, i, j, k, jdv[u], y) - F[u];
        r2 = R * R;
    }
    double r = block_sum(r2, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// y = 0 on Dirichlet nodes; x -= y; partials of y.y  (:273-283)
__global__ void __launch_bounds__(256)
    k_nr_update(PcgParams Q, const uint8_t *oid, double *x, double *y, double *partials) {
    __shared__ double sh[4];
    long u = xcd_block() * blockDim.x + threadIdx.x;
    const Grid &g = Q.g;
    double y2 = 0;
    if (u < g.nn) {
        double yy = y[u];
        if (oid[u] > 0) {
            yy = 0;
            y[u] = 0;
        }
        x[u] = x[u] - yy;
        y2 = yy * yy;
    }
    double r = block_sum(y2, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ void k_reset_flags(DevScalars *sc) {
    sc->done = 0;
    sc->iters = 0;
}

__global__ void __launch_bounds__(1024) k_sum_to(const double *partials, int nb, double *out) {
    __shared__ double sh[16];
    double s = fixed_sum(partials, nb, sh);
    if (threadIdx.x == 0) *out = s;
}

__global__ void k_fill(long n, double *p, double v) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u < n) p[u] = v;
}

// ---- Newton + PCG on the Neumann-reduced system (PIC_SOLVER_PCG_SYM) -------------------
// Face nodes are not unknowns: each equals the node its copy chain reaches (the priority
// i, j, k of ch3/ver2/PotentialSolver.cpp:359-370, which is what the reference's Neumann
// rows enforce at convergence).  For a regular node every Neumann neighbour reaches the
// node itself, so eliminating them adds their coefficient to the diagonal: the reduced
// Jacobian is symmetric and negative definite.  Vectors are zero off the regular nodes.

// x on face nodes := x at their chain target
__global__ void __launch_bounds__(256) k_neumann_copy(Grid g, const uint8_t *oid, double *x) {
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u >= g.nn) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    if (node_type(g, oid, u, i, j, k) != NT_NEU) return;
    x[u] = x[chain_target(g, oid, u, i, j, k)];
}

// F on regular nodes (the reference's residual, ch3/ver2/PotentialSolver.cpp:240-257), the
// reduced diagonal jr = A(u,u) - P(u) + sum of Neumann-neighbour coefficients, and a mask:
// bits 0-5 regular neighbours in direction d, bit 6 the node itself is regular.
// CgStart (nullable): the start of a correction solve from y = 0, written here instead of
// by a pass of its own: g = J 0 - F = -F (FP64 and FP32).  The first CG pass takes d = 0
// and y = 0 without reading them (k_mgpcg_a<true>).  idiag32: the multigrid's level-0
// reciprocal diagonal in FP32, (float)(1 / jr).
struct CgStart {
    double *g;
    float *g32;
    float *idiag32;
};

__global__ void __launch_bounds__(256)
    k_nrs_F(PcgParams Q, const double *x, const double *rho, const uint8_t *oid, double *F, double *jr,
            uint8_t *mask, double *boltz, CgStart st) {
    long u = xcd_block() * blockDim.x + threadIdx.x;
    const Grid &g = Q.g;
    if (u >= g.nn) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    int t = node_type(g, oid, u, i, j, k);
    if (t != NT_REG) {
        F[u] = 0;
        jr[u] = 1;
        mask[u] = 0;
        if (boltz) boltz[u] = 0;
        if (st.g) {
            st.g[u] = 0;
            st.g32[u] = 0.0f;
            st.idiag32[u] = 1.0f;
        }
        return;
    }
    double ad = a_diag(Q, t, i, j);
    double Ax = row_apply(Q, t, u, i, j, k, ad, x);
    double f = Ax - (-rho[u] / EPS_0);
    f -= QE * Q.n0 * exp((x[u] - Q.phi0) / Q.Te0) / EPS_0;
    double p = Q.n0 * QE / (EPS_0 * Q.Te0) * exp((x[u] - Q.phi0) / Q.Te0);
    double diag = ad - p;
    uint8_t m = 1u << 6;
    const int di[6] = {0, 0, -1, 1, 0, 0}, dj[6] = {0, -1, 0, 0, 1, 0}, dk[6] = {-1, 0, 0, 0, 0, 1};
    for (int d = 0; d < 6; d++) {
        long v = u + nb_offset(g, d);
        int tv = node_type(g, oid, v, i + di[d], j + dj[d], k + dk[d]);
        if (tv == NT_REG) m |= (uint8_t)(1u << d);
        else if (tv == NT_NEU) diag += nb_coef(Q, d);
    }
    F[u] = f;
    jr[u] = diag;
    mask[u] = m;
    if (boltz) boltz[u] = p;
    if (st.g) {
        st.g[u] = -f;
        st.g32[u] = (float)(-f);
        st.idiag32[u] = (float)(1.0 / diag);
    }
}


__global__ void __launch_bounds__(256)
    k_pcgs_init(PcgParams Q, const uint8_t *mask, const double *jr, const double *y, const double *F, double *gv,
                double *d0, double *partials) {
    __shared__ double sh[4];
    NodeChunk ch = pcg_chunk(Q.g.nn);
    double gs = 0;
    for (long u = ch.u0 + threadIdx.x; u < ch.u1; u += 256) {
        uint8_t m = mask[u];
        double gg = 0;
        if (m & 64) {
            gg = jr_apply(Q, u, m, jr[u], y) - F[u];
            gs += gg * ((1.0 / jr[u]) * gg);
        }
        gv[u] = gg;
        d0[u] = 0;
    }
    double r = block_sum(gs, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ void __launch_bounds__(256)
    k_pcgs_a(PcgParams Q, const uint8_t *mask, const double *jr, const double *gv, const double *d_old,
             double *d_new, double *y, double *z, double *partials, const DevScalars *sc) {
    __shared__ double sh[4];
    if (sc->done) return;
    const Grid &g = Q.g;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    double mu = sc->red[6], lam = sc->red[7];
    double dz = 0;
    if (u < g.nn) {
        uint8_t m = mask[u];
        if (m & 64) {
            y[u] = y[u] + lam * d_old[u];
            double dc = dnew_at(d_old, gv, jr, u, mu);
            double zz = 0;
            for (int d = 0; d < 3; d++)
                if (m & (1u << d)) {
                    long v = u + nb_offset(g, d);
                    zz += nb_coef(Q, d) * dnew_at(d_old, gv, jr, v, mu);
                }
            zz += jr[u] * dc;
            for (int d = 3; d < 6; d++)
                if (m & (1u << d)) {
                    long v = u + nb_offset(g, d);
                    zz += nb_coef(Q, d) * dnew_at(d_old, gv, jr, v, mu);
                }
            d_new[u] = dc;
            z[u] = zz;
            dz = dc * zz;
        } else {
            d_new[u] = 0;
            z[u] = 0;
        }
    }
    double r = block_sum(dz, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// ---- CG preconditioned by one multigrid V-cycle (PIC_SOLVER_MG) -------------------------
// The iteration of solvePCGLinear (:299-331) with s = V(g) in place of s = M g:
//   init:    g = J y - F = -F (y = 0; k_nrs_F writes g, its FP32 copy, y = 0 and d_0 = 0);
//            s = V(g);  a = g.s
//   pass A:  d = mu d_old - s (at the 7 stencil points), y += lam d_old, z = J d, partials d.z
//   pass B:  g += (a/b) z, partials g.g (convergence: RMS(g) < tol)
//   s = V(g), with the partials of g.s from its last sweep:  mu = a'/a
// Scalars: red[0] a, red[1] b, red[4] RMS(g), red[6] mu, red[7] lam = a/b.
// The direction d is stored in FP32 (s, of which it is a running combination, is FP32 from
// the V-cycle already).  z = J d is formed from the stored, rounded d, and y advances by the
// same stored d, so the FP64 recurrence g = J y - F stays exact; the rounding only perturbs
// the direction, as a slightly different preconditioner would.

// FIRST: the first pass of a solve from y = 0, d = 0 (mu = lam = 0), neither of them read
template <bool FIRST>
__global__ void __launch_bounds__(256)
    k_mgpcg_a(PcgParams Q, const uint8_t *mask, const double *jr, const float *sv, const float *d_old,
              float *d_new, double *y, double *z, double *partials, const DevScalars *sc) {
    // d_old and s are zero off the regular nodes and every regular node is interior, so the
    // six neighbours are read without waiting for the mask (non-regular ones add 0)
    __shared__ double sh[4];
    if (sc->done) return;
    const Grid &g = Q.g;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    double mu = sc->red[6], lam = sc->red[7];
    double dz = 0;
    if (u < g.nn) {
        int i, j, k;
        node_ijk(g, u, i, j, k);
        bool in = i > 0 && i < g.ni - 1 && j > 0 && j < g.nj - 1 && k > 0 && k < g.nk - 1;
        double dc = 0, zz = 0;
        if (in) {
            long sj = g.ni, sk = (long)g.ni * g.nj;
            uint8_t m = mask[u];
            auto dold = [&](long v) { return FIRST ? 0.0 : (double)d_old[v]; };
            // the new direction at node v as stored (one fma, then FP32): every thread that
            // forms it gets the same value, so z below is J applied to the stored d
            auto dnew = [&](long v) { return (double)(float)fma(mu, dold(v), -(double)sv[v]); };
            double jru = jr[u], dou = dold(u), yu = FIRST ? 0.0 : y[u];
            double n0 = dnew(u - sk);
            double n1 = dnew(u - sj);
            double n2 = dnew(u - 1);
            double n3 = dnew(u + 1);
            double n4 = dnew(u + sj);
            double n5 = dnew(u + sk);
            if (m & 64) {
                y[u] = yu + lam * dou;
                dc = dnew(u);
                // slot order of the reference rows: -z, -y, -x, centre, +x, +y, +z
                zz = Q.idz2 * n0;
                zz += Q.idy2 * n1;
                zz += Q.idx2 * n2;
                zz += jru * dc;
                zz += Q.idx2 * n3;
                zz += Q.idy2 * n4;
                zz += Q.idz2 * n5;
            }
        }
        d_new[u] = (float)dc;
        z[u] = zz;
        dz = dc * zz;
    }
    double r = block_sum(dz, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ void __launch_bounds__(256)
    k_mgpcg_b(long nn, double *gv, float *g32, const double *z, double *partials, const DevScalars *sc) {
    __shared__ double sh[4];
    if (sc->done) return;
    NodeChunk ch = pcg_chunk(nn);
    double lam = sc->red[7];
    double g2 = 0;
    for (long u = ch.u0 + threadIdx.x; u < ch.u1; u += 256) {
        double gg = gv[u] + lam * z[u];
        gv[u] = gg;
        g32[u] = (float)gg;
        g2 += gg * gg;
    }
    double r = block_sum(g2, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// mode 0: a = g.s after init;  mode 1: b = d.z, lam = a/b;  mode 2: RMS(g), convergence;
// mode 3: a' = g.s, mu = a'/a
__global__ void __launch_bounds__(1024)
    k_mgpcg_scalar(const double *partials, int nb, int mode, double nu, double tol, int it, DevScalars *sc) {
    __shared__ double sh[16];
    if (mode != 0 && sc->done) return;
    double s0 = fixed_sum(partials, nb, sh);
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        sc->red[0] = s0;
        sc->red[6] = 0.0;
        sc->red[7] = 0.0;
        sc->iters = 0;
    } else if (mode == 1) {
        sc->red[1] = s0;
        sc->red[7] = sc->red[0] / s0;
    } else if (mode == 2) {
        double l2 = sqrt(s0 / nu);
        sc->red[4] = l2;
        sc->iters = it + 1;
        if (l2 < tol) sc->done = 1;
    } else {
        sc->red[6] = s0 / sc->red[0];
        sc->red[0] = s0;
    }
}

// red-black Gauss-Seidel (w = 1) on the reduced system: the solveGSLinear fallback
__global__ void __launch_bounds__(256)
    k_gsls_sweep(PcgParams Q, const uint8_t *mask, const double *jr, double *y, const double *F, int color,
                 const DevScalars *sc) {
    if (sc->done) return;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const Grid &g = Q.g;
    if (u >= g.nn) return;
    uint8_t m = mask[u];
    if (!(m & 64)) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    if (((i + j + k) & 1) != color) return;
    double S = jr_apply(Q, u, m, jr[u], y) - jr[u] * y[u];
    double phi_new = (F[u] - S) / jr[u];
    y[u] = y[u] + 1. * (phi_new - y[u]);
}

__global__ void __launch_bounds__(256)
    k_gsls_residual(PcgParams Q, const uint8_t *mask, const double *jr, const double *y, const double *F,
                    double *partials, const DevScalars *sc) {
    __shared__ double sh[4];
    if (sc->done) return;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double r2 = 0;
    if (u < Q.g.nn) {
        uint8_t m = mask[u];
        if (m & 64) {
            double R = jr_apply(Q, u, m, jr[u], y) - F[u];
            r2 = R * R;
        }
    }
    double r = block_sum(r2, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// y on face nodes := y at their chain target, 0 on Dirichlet nodes; x -= y; partials of y.y.
// d_pend (nullable): the CG's last deferred update y += lam d is applied here, on the values
// read, instead of by k_pcg_finish (lam = sc->red[7]; the same operations, so the same y)
__global__ void __launch_bounds__(256)
    k_nrs_update(PcgParams Q, const uint8_t *oid, double *x, const double *y, const float *d_pend,
                 const DevScalars *sc, double *partials) {
    __shared__ double sh[4];
    long u = xcd_block() * blockDim.x + threadIdx.x;
    const Grid &g = Q.g;
    const double lam = d_pend ? sc->red[7] : 0.0;
    auto yv = [&](long v) { return d_pend ? y[v] + lam * (double)d_pend[v] : y[v]; };
    double y2 = 0;
    if (u < g.nn) {
        int i, j, k;
        node_ijk(g, u, i, j, k);
        int t = node_type(g, oid, u, i, j, k);
        double yy = 0;
        if (t == NT_REG) {
            yy = yv(u);
        } else if (t == NT_NEU) {
            // a chain that ends on the sphere or the inlet (a Dirichlet node, which the solve
            // never updates and whose y is never written) leaves the face node as it is
            long v = chain_target(g, oid, u, i, j, k);
            if (oid[v] == 0) yy = yv(v);
        }
        x[u] = x[u] - yy;
        y2 = yy * yy;
    }
    double r = block_sum(y2, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}



static SorParams sor_params(const pic_solver *s) {
    const Grid &g = s->w->g;
    SorParams S;
    S.idx2 = 1.0 / (g.dh[0] * g.dh[0]);
    S.idy2 = 1.0 / (g.dh[1] * g.dh[1]);
    S.idz2 = 1.0 / (g.dh[2] * g.dh[2]);
    S.denom = 2 * S.idx2 + 2 * S.idy2 + 2 * S.idz2;
    S.w = 1.4;
    S.phi0 = s->phi0;
    S.Te0 = s->Te0;
    S.n0 = s->n0;
    S.tol = s->tol;
    S.max_it = s->max_it;
    return S;
}

static size_t max_lds_bytes(int dev) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 65536;
    return (size_t)v;
}

static int read_scalars(pic_world *w) {
    PIC_HIP(hipMemcpyAsync(w->scal_host, w->scal, sizeof(DevScalars), hipMemcpyDeviceToHost, w->stream));
    PIC_HIP(hipStreamSynchronize(w->stream));
    return PIC_OK;
}

// SOR (linear or Boltzmann): returns PIC_OK / PIC_NOT_CONVERGED
template <bool BOLTZ>
static int solve_sor(pic_solver *s, double *residual, int *iters) {
    pic_world *w = s->w;
    const Grid &g = w->g;
    SorParams S = sor_params(s);
    k_reset_flags<<<1, 1, 0, w->stream>>>(w->scal);
    size_t lds = (size_t)g.nn * sizeof(double);
    size_t lim = max_lds_bytes(w->device);
    bool done_lds = false;
    if (lds + 1024 <= lim) {
        static bool attr_set[2] = {false, false};
        if (!attr_set[BOLTZ]) {
            hipFuncSetAttribute((const void *)k_sor_lds<BOLTZ>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(lim - 1024));
            attr_set[BOLTZ] = true;
        }
        k_sor_lds<BOLTZ><<<1, LDS_THREADS, lds, w->stream>>>(g, w->phi, w->rho, w->object_id, S, w->scal);
        if (hipGetLastError() == hipSuccess) done_lds = true;
    }
    if (!done_lds) {
        long half = (g.ni + 1) / 2;
        unsigned nb_sweep = nbk(half * (long)g.nj * g.nk);
        unsigned nb_res = nbk(g.nn);
        int rc = grow_partials(w, (int)nb_res);
        if (rc) return rc;
        const int CHUNK = 100;
        for (int it0 = 0; it0 < S.max_it; it0 += CHUNK) {
            int it1 = std::min(S.max_it, it0 + CHUNK);
            for (int it = it0; it < it1; it++) {
                k_sor_sweep<BOLTZ><<<nb_sweep, 256, 0, w->stream>>>(g, w->phi, w->rho, w->object_id, S, 0, w->scal);
                k_sor_sweep<BOLTZ><<<nb_sweep, 256, 0, w->stream>>>(g, w->phi, w->rho, w->object_id, S, 1, w->scal);
                if (it % 25 == 0) {
                    k_sor_residual<BOLTZ><<<nb_res, 256, 0, w->stream>>>(g, w->phi, w->rho, w->object_id, S,
                                                                         w->partials, w->scal);
                    k_sor_check<<<1, 1024, 0, w->stream>>>(w->partials, (int)nb_res,
                                                          (double)(g.ni * g.nj * g.nk), S.tol, it, w->scal);
                }
            }
            PIC_LAUNCH_CHECK();
            rc = read_scalars(w);
            if (rc) return rc;
            if (w->scal_host->done) break;
        }
        if (!w->scal_host->done) w->scal_host->iters = S.max_it;
    } else {
        int rc = read_scalars(w);
        if (rc) return rc;
    }
    if (residual) *residual = w->scal_host->red[0];
    if (iters) *iters = w->scal_host->iters;
    w->ef_valid = false;
    return w->scal_host->done ? PIC_OK : PIC_NOT_CONVERGED;
}

static PcgParams pcg_params(const pic_solver *s) {
    const Grid &g = s->w->g;
    PcgParams Q;
    Q.g = g;
    Q.idx = 1.0 / g.dh[0];
    Q.idy = 1.0 / g.dh[1];
    Q.idz = 1.0 / g.dh[2];
    Q.idx2 = Q.idx * Q.idx;
    Q.idy2 = Q.idy * Q.idy;
    Q.idz2 = Q.idz * Q.idz;
    Q.phi0 = s->phi0;
    Q.Te0 = s->Te0;
    Q.n0 = s->n0;
    Q.tol = s->tol;
    return Q;
}

// solvePCGLinear on J y = F; returns converged and adds iterations to *its
static int pcg_linear(pic_solver *s, const PcgParams &Q, bool *converged, int *its) {
    pic_world *w = s->w;
    long nn = Q.g.nn;
    unsigned nb = nbk(nn);
    unsigned nb8 = nbk8(nn);  // stencil kernels: XCD-remapped grid, one partial per workgroup
    unsigned G = nb < PCG_GRID ? nb : PCG_GRID;  // streaming-pass workgroups = partials
    int rc = grow_partials(w, (int)nb8);
    if (rc) return rc;
    double *folded = w->partials + 4 * (size_t)w->partials_cap;
    double nu = (double)(int)nn;
    double *dbuf[2] = {s->d, s->z2};
    k_reset_flags<<<1, 1, 0, w->stream>>>(w->scal);
    k_pcg_init<<<G, 256, 0, w->stream>>>(Q, w->object_id, s->P, s->y, s->F, s->g, dbuf[0], w->partials);
    k_pcg_scalar<<<1, 1024, 0, w->stream>>>(w->partials, (int)G, 0, nu, Q.tol, 0, w->scal);
    const int CHUNK = 32;
    bool conv = false;
    for (int it0 = 0; it0 < s->max_it; it0 += CHUNK) {
        int it1 = std::min(s->max_it, it0 + CHUNK);
        for (int it = it0; it < it1; it++) {
            k_pcg_a<<<nb8, 256, 0, w->stream>>>(Q, w->object_id, s->P, s->g, dbuf[it & 1], dbuf[(it + 1) & 1], s->y,
                                              s->z, w->partials, w->scal);
            k_fold<<<FOLD_GRID, 256, 0, w->stream>>>(w->partials, nb8, folded, w->scal);
            k_pcg_scalar<<<1, 1024, 0, w->stream>>>(folded, (int)FOLD_GRID, 1, nu, Q.tol, it, w->scal);
            k_pcg_b<<<G, 256, 0, w->stream>>>(Q, s->P, s->g, s->z, w->partials, (int)G, w->scal);
            k_pcg_scalar<<<1, 1024, 0, w->stream>>>(w->partials, (int)G, 2, nu, Q.tol, it, w->scal);
        }
        PIC_LAUNCH_CHECK();
        rc = read_scalars(w);
        if (rc) return rc;
        if (w->scal_host->done) {
            conv = true;
            break;
        }
    }
    // d of the last completed iteration sits in dbuf[iters & 1]
    int done_its = w->scal_host->iters;
    k_pcg_finish<double><<<nb8, 256, 0, w->stream>>>(nn, dbuf[done_its & 1], s->y, w->scal);
    PIC_LAUNCH_CHECK();
    *its += done_its;
    *converged = conv;
    return PIC_OK;
}

static int gs_linear(pic_solver *s, const PcgParams &Q) {
    pic_world *w = s->w;
    long nn = Q.g.nn;
    unsigned nb = nbk(nn);
    k_reset_flags<<<1, 1, 0, w->stream>>>(w->scal);
    const int CHUNK = 100;
    for (int it0 = 0; it0 < s->max_it; it0 += CHUNK) {
        int it1 = std::min(s->max_it, it0 + CHUNK);
        for (int it = it0; it < it1; it++) {
            k_gsl_sweep<<<nb, 256, 0, w->stream>>>(Q, w->object_id, s->P, s->y, s->F, 0, w->scal);
            k_gsl_sweep<<<nb, 256, 0, w->stream>>>(Q, w->object_id, s->P, s->y, s->F, 1, w->scal);
            if (it % 25 == 0) {
                k_gsl_residual<<<nb, 256, 0, w->stream>>>(Q, w->object_id, s->P, s->y, s->F, w->partials, w->scal);
                k_sor_check<<<1, 1024, 0, w->stream>>>(w->partials, (int)nb, (double)(int)nn, Q.tol, it, w->scal);
            }
        }
        PIC_LAUNCH_CHECK();
        int rc = read_scalars(w);
        if (rc) return rc;
        if (w->scal_host->done) break;
    }
    return PIC_OK;
}

static int ensure_pcg_vectors(pic_solver *s) {
    if (s->y) return PIC_OK;
    long nn = s->w->g.nn;
    double **v[7] = {&s->y, &s->g, &s->d, &s->z, &s->F, &s->P, &s->z2};
    for (auto p : v) PIC_HIP(hipMalloc(p, sizeof(double) * nn));
    // y = 0 everywhere once: the multigrid path writes y on regular nodes only, and the
    // Newton update reads the face nodes' chain targets, which may be Dirichlet nodes (as
    // the reference zeroes y there, ch3/ver2/PotentialSolver.cpp:273-275)
    k_fill<<<nbk(nn), 256, 0, s->w->stream>>>(nn, s->y, 0.0);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

static int solve_nrpcg(pic_solver *s, double *residual, int *iters) {
    pic_world *w = s->w;
    int rc = ensure_pcg_vectors(s);
    if (rc) return rc;
    PcgParams Q = pcg_params(s);
    long nn = w->g.nn;
    unsigned nb = nbk(nn);
    unsigned nb8 = nbk8(nn);  // stencil kernels: XCD-remapped grid, one partial per workgroup
    rc = grow_partials(w, (int)nb8);
    if (rc) return rc;
    // y starts at zero for each solve and is warm-started across Newton steps (:232)
    k_fill<<<nb, 256, 0, w->stream>>>(nn, s->y, 0.0);
    const int NR_MAX_IT = 20;
    const double NR_TOL = 1e-3;
    // PIC_VERBOSE=1: one progress line per Newton iteration on stderr
    const char *venv = getenv("PIC_VERBOSE");
    bool verbose = venv && venv[0] == '1';
    auto t0 = std::chrono::steady_clock::now();
    bool converged = false;
    double norm = 0;
    int pcg_its = 0;
    for (int it = 0; it < NR_MAX_IT; it++) {
        k_nr_F<<<nb8, 256, 0, w->stream>>>(Q, w->phi, w->rho, w->object_id, s->F, s->P);
        PIC_LAUNCH_CHECK();
        bool pc = false;
        rc = pcg_linear(s, Q, &pc, &pcg_its);
        if (rc) return rc;
        s->info_newton = it + 1;
        if (!pc) {
            rc = gs_linear(s, Q);
            if (rc) return rc;
            s->info_fallbacks++;
        }
        k_nr_update<<<nb8, 256, 0, w->stream>>>(Q, w->object_id, w->phi, s->y, w->partials);
        k_sum_to<<<1, 1024, 0, w->stream>>>(w->partials, (int)nb8, &w->scal->red[5]);
        PIC_LAUNCH_CHECK();
        rc = read_scalars(w);
        if (rc) return rc;
        norm = sqrt(w->scal_host->red[5] / (double)(int)nn);
        if (verbose)
            fprintf(stderr, "[picamd] NR it %d: PCG %s after %d its total, |y|rms %.3e, %.1f ms\n", it,
                    pc ? "converged" : "fell back to GS", pcg_its, norm,
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        if (norm < NR_TOL) {
            converged = true;
            break;
        }
    }
    if (residual) *residual = norm;
    if (iters) *iters = pcg_its;
    w->ef_valid = false;
    return converged ? PIC_OK : PIC_NOT_CONVERGED;
}

static int pcgs_linear(pic_solver *s, const PcgParams &Q, bool *converged, int *its) {
    pic_world *w = s->w;
    long nn = Q.g.nn;
    unsigned nb = nbk(nn);
    unsigned nb8 = nbk8(nn);  // stencil kernels: XCD-remapped grid, one partial per workgroup
    unsigned G = nb < PCG_GRID ? nb : PCG_GRID;  // streaming-pass workgroups = partials
    int rc = grow_partials(w, (int)nb8);
    if (rc) return rc;
    double *folded = w->partials + 4 * (size_t)w->partials_cap;
    double nu = (double)(int)nn;
    double *dbuf[2] = {s->d, s->z2};
    k_reset_flags<<<1, 1, 0, w->stream>>>(w->scal);
    k_pcgs_init<<<G, 256, 0, w->stream>>>(Q, s->mask, s->P, s->y, s->F, s->g, dbuf[0], w->partials);
    k_pcg_scalar<<<1, 1024, 0, w->stream>>>(w->partials, (int)G, 0, nu, Q.tol, 0, w->scal);
    const int CHUNK = 32;
    bool conv = false;
    for (int it0 = 0; it0 < s->max_it; it0 += CHUNK) {
        int it1 = std::min(s->max_it, it0 + CHUNK);
        for (int it = it0; it < it1; it++) {
            k_pcgs_a<<<nb8, 256, 0, w->stream>>>(Q, s->mask, s->P, s->g, dbuf[it & 1], dbuf[(it + 1) & 1], s->y,
                                               s->z, w->partials, w->scal);
            k_fold<<<FOLD_GRID, 256, 0, w->stream>>>(w->partials, nb8, folded, w->scal);
            k_pcg_scalar<<<1, 1024, 0, w->stream>>>(folded, (int)FOLD_GRID, 1, nu, Q.tol, it, w->scal);
            k_pcg_b<<<G, 256, 0, w->stream>>>(Q, s->P, s->g, s->z, w->partials, (int)G, w->scal);
            k_pcg_scalar<<<1, 1024, 0, w->stream>>>(w->partials, (int)G, 2, nu, Q.tol, it, w->scal);
        }
        PIC_LAUNCH_CHECK();
        int rc = read_scalars(w);
        if (rc) return rc;
        if (w->scal_host->done) {
            conv = true;
            break;
        }
    }
    int done_its = w->scal_host->iters;
    k_pcg_finish<double><<<nb8, 256, 0, w->stream>>>(nn, dbuf[done_its & 1], s->y, w->scal);
    PIC_LAUNCH_CHECK();
    *its += done_its;
    *converged = conv;
    return PIC_OK;
}

static int gsls_linear(pic_solver *s, const PcgParams &Q) {
    pic_world *w = s->w;
    long nn = Q.g.nn;
    unsigned nb = nbk(nn);
    k_reset_flags<<<1, 1, 0, w->stream>>>(w->scal);
    const int CHUNK = 100;
    for (int it0 = 0; it0 < s->max_it; it0 += CHUNK) {
        int it1 = std::min(s->max_it, it0 + CHUNK);
        for (int it = it0; it < it1; it++) {
            k_gsls_sweep<<<nb, 256, 0, w->stream>>>(Q, s->mask, s->P, s->y, s->F, 0, w->scal);
            k_gsls_sweep<<<nb, 256, 0, w->stream>>>(Q, s->mask, s->P, s->y, s->F, 1, w->scal);
            if (it % 25 == 0) {
                k_gsls_residual<<<nb, 256, 0, w->stream>>>(Q, s->mask, s->P, s->y, s->F, w->partials, w->scal);
                k_sor_check<<<1, 1024, 0, w->stream>>>(w->partials, (int)nb, (double)(int)nn, Q.tol, it, w->scal);
            }
        }
        PIC_LAUNCH_CHECK();
        int rc = read_scalars(w);
        if (rc) return rc;
        if (w->scal_host->done) break;
    }
    return PIC_OK;
}

// Newton iteration of solveNRPCG (:225-296) on the Neumann-reduced system
int mg_build(pic_solver *s, MgHierarchy **out);
int mg_setup(MgHierarchy *H, const PcgParams &Q0, const double *boltz0, unsigned geom_stamp, hipStream_t st);
int mg_vcycle(MgHierarchy *H, const double *b0, const float *b0f, float **x0, double *dot_partials, hipStream_t st,
              const DevScalars *sc);
int mg_dot_partials(const MgHierarchy *H);
float *mg_level0_idiag(MgHierarchy *H);
void mg_destroy(MgHierarchy *H);

// solvePCGLinear on the reduced system with one V-cycle as the preconditioner
// *d_pend: when converged, the direction whose update y += lam d the caller's k_nrs_update
// applies (null when y is final here)
static int mgpcg_linear(pic_solver *s, const PcgParams &Q, bool *converged, int *its, const float **d_pend) {
    pic_world *w = s->w;
    long nn = Q.g.nn;
    unsigned nb = nbk(nn);
    unsigned nb8 = nbk8(nn);  // stencil kernels: XCD-remapped grid, one partial per workgroup
    unsigned G = nb < PCG_GRID ? nb : PCG_GRID;
    int rc = grow_partials(w, (int)nb8);
    if (rc) return rc;
    double *folded = w->partials + 4 * (size_t)w->partials_cap;
    double nu = (double)(int)nn;
    // the FP32 directions, in the storage of the FP64 vectors the Jacobi-preconditioned CG uses
    float *dbuf[2] = {reinterpret_cast<float *>(s->d), reinterpret_cast<float *>(s->z2)};
    hipStream_t st = w->stream;
    k_reset_flags<<<1, 1, 0, st>>>(w->scal);
    // s = V(g) in FP32; the V-cycle's last sweep also writes the partials of g.s
    float *sv = nullptr;
    const long ndot = mg_dot_partials(s->mg);
    rc = mg_vcycle(s->mg, s->g, s->g32, &sv, w->partials, st, w->scal);
    if (rc) return rc;
    k_fold<<<FOLD_GRID, 256, 0, st>>>(w->partials, ndot, folded, w->scal);
    k_mgpcg_scalar<<<1, 1024, 0, st>>>(folded, (int)FOLD_GRID, 0, nu, Q.tol, 0, w->scal);
    // a few iterations are queued between host checks of the convergence flag; kernels of
    // iterations past convergence return at once
    const int CHUNK = 2;
    bool conv = false;
    for (int it0 = 0; it0 < s->max_it; it0 += CHUNK) {
        int it1 = std::min(s->max_it, it0 + CHUNK);
        for (int it = it0; it < it1; it++) {
            if (it == 0)
                k_mgpcg_a<true><<<nb8, 256, 0, st>>>(Q, s->mask, s->P, sv, dbuf[0], dbuf[1], s->y, s->z, w->partials,
                                                     w->scal);
            else
                k_mgpcg_a<false><<<nb8, 256, 0, st>>>(Q, s->mask, s->P, sv, dbuf[it & 1], dbuf[(it + 1) & 1], s->y,
                                                      s->z, w->partials, w->scal);
            k_fold<<<FOLD_GRID, 256, 0, st>>>(w->partials, nb8, folded, w->scal);
            k_mgpcg_scalar<<<1, 1024, 0, st>>>(folded, (int)FOLD_GRID, 1, nu, Q.tol, it, w->scal);
            k_mgpcg_b<<<G, 256, 0, st>>>(nn, s->g, s->g32, s->z, w->partials, w->scal);
            k_mgpcg_scalar<<<1, 1024, 0, st>>>(w->partials, (int)G, 2, nu, Q.tol, it, w->scal);
            rc = mg_vcycle(s->mg, s->g, s->g32, &sv, w->partials, st, w->scal);
            if (rc) return rc;
            k_fold<<<FOLD_GRID, 256, 0, st>>>(w->partials, ndot, folded, w->scal);
            k_mgpcg_scalar<<<1, 1024, 0, st>>>(folded, (int)FOLD_GRID, 3, nu, Q.tol, it, w->scal);
        }
        PIC_LAUNCH_CHECK();
        rc = read_scalars(w);
        if (rc) return rc;
        if (w->scal_host->done) {
            conv = true;
            break;
        }
    }
    int done_its = w->scal_host->iters;
    *d_pend = nullptr;
    if (done_its > 0 && conv)
        *d_pend = dbuf[done_its & 1];
    else if (done_its > 0)  // the Gauss-Seidel fallback continues from y
        k_pcg_finish<float><<<nb8, 256, 0, st>>>(nn, dbuf[done_its & 1], s->y, w->scal);
    else  // converged at the start (or no iteration allowed): the correction is 0
        k_fill<<<nbk(nn), 256, 0, st>>>(nn, s->y, 0.0);
    PIC_LAUNCH_CHECK();
    *its += done_its;
    *converged = conv;
    return PIC_OK;
}

static int ensure_mg(pic_solver *s) {
    if (!s->boltz) PIC_HIP(hipMalloc(&s->boltz, (size_t)s->w->g.nn * sizeof(double)));
    if (!s->g32) PIC_HIP(hipMalloc(&s->g32, (size_t)s->w->g.nn * sizeof(float)));
    return mg_build(s, &s->mg);
}

static int solve_nrpcg_sym(pic_solver *s, double *residual, int *iters) {
    pic_world *w = s->w;
    int rc = ensure_pcg_vectors(s);
    if (rc) return rc;
    if (!s->mask) PIC_HIP(hipMalloc(&s->mask, w->g.nn));
    const bool use_mg = s->type == PIC_SOLVER_MG;
    if (use_mg) {
        rc = ensure_mg(s);
        if (rc) return rc;
    }
    PcgParams Q = pcg_params(s);
    long nn = w->g.nn;
    unsigned nb = nbk(nn);
    unsigned nb8 = nbk8(nn);  // stencil kernels: XCD-remapped grid, one partial per workgroup
    rc = grow_partials(w, (int)nb8);
    if (rc) return rc;
    const char *venv = getenv("PIC_VERBOSE");
    bool verbose = venv && venv[0] == '1';
    auto t0 = std::chrono::steady_clock::now();
    k_neumann_copy<<<nb8, 256, 0, w->stream>>>(w->g, w->object_id, w->phi);
    if (!use_mg) k_fill<<<nb, 256, 0, w->stream>>>(nn, s->y, 0.0);
    const int NR_MAX_IT = 20;
    const double NR_TOL = s->nr_tol;  // 1e-3 as the reference (:229) unless a caller tightened it
    bool converged = false;
    double norm = 0;
    int pcg_its = 0;
    for (int it = 0; it < NR_MAX_IT; it++) {
        // Each correction solve starts from y = 0, whose residual is the (small) Newton
        // residual F, where the reference's warm start from the previous correction (kept
        // by the PCG path) begins about as far from the new correction as that one is.
        // With the multigrid preconditioner, k_nrs_F also writes that start.
        CgStart cs{nullptr, nullptr, nullptr};
        if (use_mg) cs = CgStart{s->g, s->g32, mg_level0_idiag(s->mg)};
        k_nrs_F<<<nb8, 256, 0, w->stream>>>(Q, w->phi, w->rho, w->object_id, s->F, s->P, s->mask, s->boltz, cs);
        if (!use_mg && it > 0) k_fill<<<nb, 256, 0, w->stream>>>(nn, s->y, 0.0);
        PIC_LAUNCH_CHECK();
        bool pc = false;
        const float *d_pend = nullptr;
        if (use_mg) {
            rc = mg_setup(s->mg, Q, s->boltz, w->geom_stamp, w->stream);
            if (rc) return rc;
            rc = mgpcg_linear(s, Q, &pc, &pcg_its, &d_pend);
        } else {
            rc = pcgs_linear(s, Q, &pc, &pcg_its);
        }
        if (rc) return rc;
        s->info_newton = it + 1;
        if (!pc) {
            rc = gsls_linear(s, Q);
            if (rc) return rc;
            s->info_fallbacks++;
        }
        k_nrs_update<<<nb8, 256, 0, w->stream>>>(Q, w->object_id, w->phi, s->y, d_pend, w->scal, w->partials);
        k_sum_to<<<1, 1024, 0, w->stream>>>(w->partials, (int)nb8, &w->scal->red[5]);
        PIC_LAUNCH_CHECK();
        rc = read_scalars(w);
        if (rc) return rc;
        norm = sqrt(w->scal_host->red[5] / (double)(int)nn);
        if (verbose)
            fprintf(stderr, "[picamd] NR(%s) it %d: PCG %s after %d its total, |y|rms %.3e, %.1f ms\n",
                    use_mg ? "mg" : "sym", it,
                    pc ? "converged" : "fell back to GS", pcg_its, norm,
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        if (norm < NR_TOL) {
            converged = true;
            break;
        }
    }
    if (residual) *residual = norm;
    if (iters) *iters = pcg_its;
    w->ef_valid = false;
    return converged ? PIC_OK : PIC_NOT_CONVERGED;
}

// Right-hand side of a linearised solve: F = rhs on the regular nodes (0 elsewhere) and the
// CG start from y = 0, g = -F (FP64 and the V-cycle's FP32 copy)
__global__ void __launch_bounds__(256)
    k_lin_rhs(long nn, const uint8_t *mask, const double *rhs, double *F, double *gv, float *g32) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= nn) return;
    double f = (mask[u] & 64) ? rhs[u] : 0.0;
    F[u] = f;
    gv[u] = -f;
    if (g32) g32[u] = (float)(-f);
}

// Solves J lam = rhs on the regular nodes of the reduced system, J the Newton Jacobian at the
// world's current phi (A_r - diag(P), symmetric negative definite), by the solver's own CG
// (multigrid-preconditioned for PIC_SOLVER_MG, Jacobi for PIC_SOLVER_PCG_SYM) until
// RMS(J lam - rhs) < tol.  lam is 0 off the regular nodes.  This is the linear solve of the
// adjoint of solve(): J is symmetric, so J^T lam = rhs is this system.
int solve_linearized(pic_solver *s, const double *rhs, double *lam, double tol, int *iters) {
    pic_world *w = s->w;
    PIC_CHECK(s->type == PIC_SOLVER_MG || s->type == PIC_SOLVER_PCG_SYM, PIC_ERR_STATE,
              "linearised solves need the reduced symmetric system (PIC_SOLVER_MG or PIC_SOLVER_PCG_SYM)");
    int rc = ensure_pcg_vectors(s);
    if (rc) return rc;
    if (!s->mask) PIC_HIP(hipMalloc(&s->mask, w->g.nn));
    const bool use_mg = s->type == PIC_SOLVER_MG;
    if (use_mg) {
        rc = ensure_mg(s);
        if (rc) return rc;
    }
    PcgParams Q = pcg_params(s);
    Q.tol = tol;
    long nn = w->g.nn;
    unsigned nb8 = nbk8(nn);
    rc = grow_partials(w, (int)nb8);
    if (rc) return rc;
    CgStart cs{nullptr, nullptr, nullptr};
    if (use_mg) cs = CgStart{s->g, s->g32, mg_level0_idiag(s->mg)};
    k_nrs_F<<<nb8, 256, 0, w->stream>>>(Q, w->phi, w->rho, w->object_id, s->F, s->P, s->mask, s->boltz, cs);
    k_lin_rhs<<<nbk(nn), 256, 0, w->stream>>>(nn, s->mask, rhs, s->F, s->g, use_mg ? s->g32 : nullptr);
    PIC_LAUNCH_CHECK();
    bool conv = false;
    int its = 0;
    if (use_mg) {
        rc = mg_setup(s->mg, Q, s->boltz, w->geom_stamp, w->stream);
        if (rc) return rc;
        const float *d_pend = nullptr;
        rc = mgpcg_linear(s, Q, &conv, &its, &d_pend);
        if (rc) return rc;
        if (d_pend) k_pcg_finish<float><<<nb8, 256, 0, w->stream>>>(nn, d_pend, s->y, w->scal);
    } else {
        k_fill<<<nbk(nn), 256, 0, w->stream>>>(nn, s->y, 0.0);
        rc = pcgs_linear(s, Q, &conv, &its);
        if (rc) return rc;
    }
    PIC_LAUNCH_CHECK();
    PIC_HIP(hipMemcpyAsync(lam, s->y, sizeof(double) * nn, hipMemcpyDeviceToDevice, w->stream));
    if (iters) *iters = its;
    if (!conv) {
        set_error("linearised solve did not reach its tolerance within max_it CG iterations");
        return PIC_NOT_CONVERGED;
    }
    return PIC_OK;
}

}  // namespace pic

using namespace pic;

extern "C" {

int pic_solver_create(pic_world *w, int type, int max_it, double tol, pic_solver **out) {
    PIC_CHECK(w && out, PIC_ERR_ARG, "null argument");
    PIC_CHECK(type >= PIC_SOLVER_GS && type <= PIC_SOLVER_MG, PIC_ERR_ARG, "bad solver type");
    PIC_CHECK(max_it >= 0, PIC_ERR_ARG, "negative max_it");
    PIC_CHECK(w->extents_set, PIC_ERR_STATE, "setExtents must precede the solver");
    pic_solver *s = new pic_solver();
    s->w = w;
    s->type = type;
    s->max_it = max_it;
    s->tol = tol;
    *out = s;
    return PIC_OK;
}

int pic_solver_destroy(pic_solver *s) {
    if (!s) return PIC_OK;
    if (s->w && s->w->stream) hipStreamSynchronize(s->w->stream);
    hipFree(s->y);
    hipFree(s->g);
    hipFree(s->d);
    hipFree(s->z);
    hipFree(s->F);
    hipFree(s->P);
    hipFree(s->z2);
    hipFree(s->mask);
    mg_destroy(s->mg);
    hipFree(s->boltz);
    hipFree(s->g32);
    delete s;
    return PIC_OK;
}

int pic_solver_set_reference_values(pic_solver *s, double phi0, double Te0, double n0) {
    PIC_CHECK(s, PIC_ERR_ARG, "null argument");
    PIC_CHECK(Te0 > 0 && n0 > 0, PIC_ERR_ARG, "Te0 and n0 must be positive");
    s->phi0 = phi0;
    s->Te0 = Te0;
    s->n0 = n0;
    return PIC_OK;
}

int pic_solver_solve(pic_solver *s, double *residual, int *iters) {
    PIC_CHECK(s, PIC_ERR_ARG, "null argument");
    pic_world *w = s->w;
    hipEvent_t ev;
    timer_begin(w, PIC_T_SOLVE, &ev);
    int rc;
    int its_local = 0;
    if (!iters) iters = &its_local;
    s->info_newton = 0;
    s->info_fallbacks = 0;
    switch (s->type) {
    case PIC_SOLVER_SOR_LINEAR: rc = solve_sor<false>(s, residual, iters); break;
    case PIC_SOLVER_GS: rc = solve_sor<true>(s, residual, iters); break;
    case PIC_SOLVER_PCG: rc = solve_nrpcg(s, residual, iters); break;
    case PIC_SOLVER_PCG_SYM:
    case PIC_SOLVER_MG: rc = solve_nrpcg_sym(s, residual, iters); break;
    default:
        rc = solve_qn(w, s->phi0, s->Te0, s->n0);
        if (residual) *residual = 0;
        if (iters) *iters = 1;
        break;
    }
    timer_end(w, PIC_T_SOLVE, ev);
    s->info_iters = *iters;
    s->info_converged = rc == PIC_OK ? 1 : 0;
    return rc;
}

int pic_solver_set_mg(pic_solver *s, int sweeps, int gal_from, long fuse_max) {
    PIC_CHECK(s, PIC_ERR_ARG, "null argument");
    PIC_CHECK(sweeps >= 0 && sweeps <= 8 && gal_from >= 0 && gal_from <= 16 && fuse_max >= 0, PIC_ERR_ARG,
              "multigrid shape out of range");
    PIC_HIP(hipStreamSynchronize(s->w->stream));
    s->mg_sweeps = sweeps;
    s->mg_gal_from = gal_from;
    s->mg_fuse_max = fuse_max;
    if (s->mg) {  // rebuilt with the new shape by the next solve
        mg_destroy(s->mg);
        s->mg = nullptr;
    }
    return PIC_OK;
}

int pic_solver_get_info(pic_solver *s, int info[4]) {
    PIC_CHECK(s && info, PIC_ERR_ARG, "null argument");
    info[0] = s->info_newton;
    info[1] = s->info_iters;
    info[2] = s->info_fallbacks;
    info[3] = s->info_converged;
    return PIC_OK;
}

int pic_solver_solve_qn(pic_solver *s) {
    PIC_CHECK(s, PIC_ERR_ARG, "null argument");
    return solve_qn(s->w, s->phi0, s->Te0, s->n0);
}

int pic_solver_compute_ef(pic_solver *s) {
    PIC_CHECK(s, PIC_ERR_ARG, "null argument");
    int rc = compute_ef(s->w, true, true);
    if (rc) return rc;
    s->w->ef_valid = true;
    return PIC_OK;
}

}  // extern "C"
