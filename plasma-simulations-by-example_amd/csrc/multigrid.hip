// Geometric multigrid V-cycle as the preconditioner of CG on the Neumann-reduced Newton
// system (PIC_SOLVER_MG).
//
// The Newton iteration, its residual F, its stopping tests (RMS(g) < tol for the linear
// solve, RMS(y) < 1e-3 for Newton) and the discrete solution are those of solveNRPCG +
// solvePCGLinear (ch3/ver2/PotentialSolver.cpp:225-331) on the reduced system of
// PIC_SOLVER_PCG_SYM; only the preconditioner differs.  The reference's Jacobi preconditioner
// (Matrix::invDiagonal, :52-57) needs O(n) CG iterations on an n-node-wide grid (about 2000
// per Newton step at 257x257x513); one V-cycle per iteration makes the count nearly
// independent of the grid.
//
// Hierarchy.  Vertex-centred coarsening by 2 in every direction while every (n-1) is even
// and every n >= 5 (257x257x513 -> 129x129x257 -> ... -> 3x3x5).  Coarse nodes take their
// object id and Boltzmann derivative P from the fine node at the same position (injection).
// Level 1 discretises the reduced 7-point operator at its own spacing, with Neumann face
// nodes eliminated into the diagonal exactly as on the finest level; levels from
// MG_GALERKIN_FROM on carry the Galerkin product R A P of the level above (27 points, built
// once per geometry) with the injected P on its diagonal.
//
// Transfers.  Prolongation is trilinear interpolation of the coarse correction after its
// face nodes are filled from their copy-chain targets (E), written to regular fine nodes
// only.  Restriction is its transpose over 8 (full weighting, then the face values folded
// into their chain targets, E^T), so that the V-cycle is a symmetric operator.
//
// Smoothing.  One red-black Gauss-Seidel sweep before (red, black) and one after (black,
// red) the coarse correction on the 7-point levels, two damped-Jacobi sweeps on each side on
// the Galerkin levels: the post-smoother is the adjoint of the pre-smoother, as CG requires
// of its preconditioner.  The coarsest level is solved by sweeps.
//
// Every kernel here returns at once when the solve has converged (DevScalars::done), so
// the host can queue several iterations between convergence checks.
#include <algorithm>
#include <vector>

#include "stencil.h"

#pragma clang fp contract(off)

namespace pic {

struct MgLevel {
    Grid g;                  // ni, nj, nk, nn, dh
    PcgParams Q;             // coefficients at this spacing
    uint8_t *oid = nullptr;  // object ids (level 0: the world's)
    uint8_t *mask = nullptr; // bits 0-5 regular neighbours, bit 6 regular (level 0: the solver's)
    float *idiag = nullptr;  // reciprocal of the reduced diagonal
    float *boltz = nullptr;  // Boltzmann derivative P (levels >= 1)
    float *x = nullptr, *r = nullptr;
    float *b = nullptr;      // right-hand side (levels >= 1; level 0 reads the CG residual)
    float *a27 = nullptr;    // Galerkin levels: the 27 coefficients of R J P (plane-major)
    float *x2 = nullptr;     // Galerkin levels: the second Jacobi buffer
    bool gal = false;
    bool owned = false;
};

struct MgDev;

#ifndef MG_NU
#define MG_NU 1
#endif
// red-black sweep pairs per side of the V-cycle on the levels outside the fused coarse cycle
constexpr int MG_SWEEPS = MG_NU;

// Levels >= MG_GALERKIN_FROM carry the Galerkin operator R J P of the level above (27
// coefficients per node) and are smoothed by damped Jacobi; the levels above keep the
// rediscretised 7-point operator and red-black Gauss-Seidel.  The rediscretised coarse
// operators of the Neumann-reduced system cost about half the convergence rate of the
// preconditioned CG (0.30 against 0.17-0.20 per iteration in a CPU model of this cycle);
// level 1 keeps the cheap 7-point form, which gives up little of that.  Measured on the
// 257x257x513 sphere flow: 14.6 -> 9.5 CG iterations and 29.6 -> 21.6 ms of Newton solve
// per time step.
#ifndef MG_GAL_FROM
#define MG_GAL_FROM 2
#endif
#ifndef MG_JAC_W
#define MG_JAC_W 0.9f
#endif
constexpr int MG_GALERKIN_FROM = MG_GAL_FROM;
constexpr float MG_JACOBI_W = MG_JAC_W;

struct MgHierarchy {
    std::vector<MgLevel> L;
    // launch shapes (MG_NU, MG_GAL_FROM, MG_FUSE_MAX unless the tuning table or the solver
    // sets others): smoothing sweep pairs on the 7-point levels, first Galerkin level, largest
    // level of the single-workgroup coarse cycle
    int sweeps = MG_NU;
    int gal_from = MG_GAL_FROM;
    long fuse_max = 2000;
    unsigned rap_stamp = ~0u;  // world geometry the Galerkin operators were built for
    int coarse_sweeps = 0;
    int fuse_l0 = -1;       // first level of the single-workgroup coarse cycle (-1: none)
    MgDev *dev = nullptr;   // device copy of the level descriptors
};

// the six neighbour coefficients of a level in slot order (-z, -y, -x, +x, +y, +z)
struct Coef6 {
    float c[6];
};

static Coef6 coef6(const PcgParams &Q) {
    Coef6 k;
    k.c[0] = (float)Q.idz2;
    k.c[1] = (float)Q.idy2;
    k.c[2] = (float)Q.idx2;
    k.c[3] = (float)Q.idx2;
    k.c[4] = (float)Q.idy2;
    k.c[5] = (float)Q.idz2;
    return k;
}

// ---- kernels -----------------------------------------------------------------------------
// coarse object id and P by injection from the fine node at the same position
template <typename TP>
__global__ void __launch_bounds__(256)
    k_mg_inject(Grid gf, Grid gc, const uint8_t *oid_f, const TP *P_f, uint8_t *oid_c, float *P_c) {
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u >= gc.nn) return;
    int i, j, k;
    node_ijk(gc, u, i, j, k);
    long f = 2L * i + (long)gf.ni * (2L * j + (long)gf.nj * 2L * k);
    oid_c[u] = oid_f[f];
    P_c[u] = (float)P_f[f];
}

// mask and reduced diagonal of a coarse level (the operator part of k_nrs_F)
__global__ void __launch_bounds__(256)
    k_mg_ops(PcgParams Q, const uint8_t *oid, const float *P, uint8_t *mask, float *idiag) {
    long u = xcd_block() * blockDim.x + threadIdx.x;
    const Grid &g = Q.g;
    if (u >= g.nn) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    if (node_type(g, oid, u, i, j, k) != NT_REG) {
        mask[u] = 0;
        idiag[u] = 1;
        return;
    }
    double dg = -2.0 * (Q.idx2 + Q.idy2 + Q.idz2) - (double)P[u];
    uint8_t m = 1u << 6;
    const int di[6] = {0, 0, -1, 1, 0, 0}, dj[6] = {0, -1, 0, 0, 1, 0}, dk[6] = {-1, 0, 0, 0, 0, 1};
    for (int d = 0; d < 6; d++) {
        long v = u + nb_offset(g, d);
        int tv = node_type(g, oid, v, i + di[d], j + dj[d], k + dk[d]);
        if (tv == NT_REG) m |= (uint8_t)(1u << d);
        else if (tv == NT_NEU) dg += nb_coef(Q, d);
    }
    mask[u] = m;
    idiag[u] = (float)(1.0 / dg);
}

// ---- per-node steps (shared by the per-level kernels and the fused coarse cycle) ----------
// Iterates x, r are zero off the regular nodes (smoothers and prolongation write regular
// nodes only, and the first half-sweep of a V-cycle zeroes the rest), and every regular node
// is interior.  So the six neighbours of an interior node are loaded unconditionally and
// summed with their coefficients: the non-regular ones add 0, as in the reduced operator,
// and the loads do not wait for the mask.

__device__ inline bool interior(const Grid &g, int i, int j, int k) {
    return i > 0 && i < g.ni - 1 && j > 0 && j < g.nj - 1 && k > 0 && k < g.nk - 1;
}

__device__ inline float nb_sum(const Grid &g, const Coef6 &C, const float *x, long u) {
    long sj = g.ni, sk = (long)g.ni * g.nj;
    return C.c[0] * x[u - sk] + C.c[1] * x[u - sj] + C.c[2] * x[u - 1] + C.c[3] * x[u + 1] +
           C.c[4] * x[u + sj] + C.c[5] * x[u + sk];
}

// The first sweep of a level from x = 0, red then black, in one pass: a red regular node
// takes b/diag; a black regular node the Gauss-Seidel update from those red values, computed
// from b/diag at its neighbours (b, like x, is zero off the regular nodes); other nodes 0.
template <typename TB>
__device__ inline void mg_pre_node(const Grid &g, const Coef6 &C, const uint8_t *mask, const float *idiag, float *x,
                                   const TB *b, long u) {
    int i, j, k;
    node_ijk(g, u, i, j, k);
    float v = 0.0f;
    if (interior(g, i, j, k)) {
        uint8_t m = mask[u];
        float bu = (float)b[u], du = idiag[u];
        if (((i + j + k) & 1) == 0) {
            v = bu * du;
        } else {
            long sj = g.ni, sk = (long)g.ni * g.nj;
            float s = C.c[0] * ((float)b[u - sk] * idiag[u - sk]) + C.c[1] * ((float)b[u - sj] * idiag[u - sj]) +
                      C.c[2] * ((float)b[u - 1] * idiag[u - 1]) + C.c[3] * ((float)b[u + 1] * idiag[u + 1]) +
                      C.c[4] * ((float)b[u + sj] * idiag[u + sj]) + C.c[5] * ((float)b[u + sk] * idiag[u + sk]);
            v = (bu - s) * du;
        }
        if (!(m & 64)) v = 0.0f;
    }
    x[u] = v;
}

// Gauss-Seidel update of the t-th node of one colour (x-fastest, every second node of a row)
template <typename TB>
__device__ inline void mg_smooth_node(const Grid &g, const Coef6 &C, const uint8_t *mask, const float *idiag,
                                      float *x, const TB *b, int color, long t) {
    unsigned half = (unsigned)(g.ni + 1) / 2;
    unsigned row = (unsigned)t / half, ih = (unsigned)t - row * half;
    unsigned kk = row / (unsigned)g.nj;
    int j = (int)(row - kk * (unsigned)g.nj), k = (int)kk;
    int i = (int)(2 * ih) + ((j + k + color) & 1);
    if (!interior(g, i, j, k)) return;
    long u = i + (long)g.ni * row;
    uint8_t m = mask[u];
    float s = nb_sum(g, C, x, u);
    float bb = (float)b[u], du = idiag[u];
    if (m & 64) x[u] = (bb - s) * du;
}

__device__ inline long smooth_count(const Grid &g) { return (long)((g.ni + 1) / 2) * g.nj * g.nk; }

// r = b - J x on regular nodes, 0 elsewhere
template <typename TB>
__device__ inline void mg_residual_node(const Grid &g, const Coef6 &C, const uint8_t *mask, const float *idiag,
                                        const float *x, const TB *b, float *r, long u) {
    int i, j, k;
    node_ijk(g, u, i, j, k);
    float v = 0.0f;
    if (interior(g, i, j, k)) {
        uint8_t m = mask[u];
        float s = x[u] / idiag[u] + nb_sum(g, C, x, u);
        float bb = (float)b[u];
        if (m & 64) v = bb - s;
    }
    r[u] = v;
}

// full weighting of the fine residual onto coarse node u (faces included)
__device__ inline void mg_restrict_node(const Grid &gf, const Grid &gc, const float *rf, float *rc, long u) {
    int I, J, K;
    node_ijk(gc, u, I, J, K);
    long sj = gf.ni, sk = (long)gf.ni * gf.nj;
    float s = 0;
    for (int c = -1; c <= 1; c++) {
        int k = 2 * K + c;
        if (k < 0 || k >= gf.nk) continue;
        float wc = c == 0 ? 0.5f : 0.25f;
        for (int b = -1; b <= 1; b++) {
            int j = 2 * J + b;
            if (j < 0 || j >= gf.nj) continue;
            float wb = wc * (b == 0 ? 0.5f : 0.25f);
            for (int a = -1; a <= 1; a++) {
                int i = 2 * I + a;
                if (i < 0 || i >= gf.ni) continue;
                s += wb * (a == 0 ? 0.5f : 0.25f) * rf[i + sj * j + sk * k];
            }
        }
    }
    rc[u] = s;
}

// E^T: a regular coarse node adds the values of the face nodes whose copy chain ends on it
// (at most 7, fixed order); non-regular nodes become 0
__device__ inline void mg_fold_node(const Grid &g, const uint8_t *oid, const uint8_t *mask, const float *in,
                                    float *out, long u) {
    if (!(mask[u] & 64)) {
        out[u] = 0.0f;
        return;
    }
    int i, j, k;
    node_ijk(g, u, i, j, k);
    float s = in[u];
    int da0 = i == 1 ? -1 : 0, da1 = i == g.ni - 2 ? 1 : 0;
    int db0 = j == 1 ? -1 : 0, db1 = j == g.nj - 2 ? 1 : 0;
    int dc0 = k == 1 ? -1 : 0, dc1 = k == g.nk - 2 ? 1 : 0;
    if (da0 | da1 | db0 | db1 | dc0 | dc1) {
        long sj = g.ni, sk = (long)g.ni * g.nj;
        for (int c = dc0; c <= dc1; c++)
            for (int b = db0; b <= db1; b++)
                for (int a = da0; a <= da1; a++) {
                    if (a == 0 && b == 0 && c == 0) continue;
                    long f = u + a + sj * b + sk * c;
                    if (chain_target(g, oid, f, i + a, j + b, k + c) == u) s += in[f];
                }
    }
    out[u] = s;
}

// E: a face node takes the value of its chain target (Dirichlet nodes keep 0)
__device__ inline void mg_extend_node(const Grid &g, const uint8_t *oid, float *x, long u) {
    int i, j, k;
    node_ijk(g, u, i, j, k);
    if (node_type(g, oid, u, i, j, k) != NT_NEU) return;
    x[u] = x[chain_target(g, oid, u, i, j, k)];
}

// x_f += trilinear interpolation of x_c at the t-th red node of a row (colour-strided like
// mg_smooth_node).  Only red nodes: the post-smoothing black half-sweep that follows
// overwrites every black value without reading it, and the red half-sweep after it reads
// only black values, so a correction written to black nodes would never be read.
__device__ inline void mg_prolong_node(const Grid &gf, const Grid &gc, const uint8_t *mask_f, const float *xc,
                                       float *xf, long t) {
    unsigned half = (unsigned)(gf.ni + 1) / 2;
    unsigned row = (unsigned)t / half, ih = (unsigned)t - row * half;
    unsigned kk = row / (unsigned)gf.nj;
    int j = (int)(row - kk * (unsigned)gf.nj), k = (int)kk;
    int i = (int)(2 * ih) + ((j + k) & 1);
    if (i >= gf.ni) return;
    long u = i + (long)gf.ni * row;
    // the mask, the coarse values and x are loaded together (an odd fine index has its
    // upper coarse neighbour inside the coarse grid), and the mask decides the store
    int I = i >> 1, J = j >> 1, K = k >> 1;
    int ai = i & 1, aj = j & 1, ak = k & 1;
    long sj = gc.ni, sk = (long)gc.ni * gc.nj;
    long b = I + sj * J + sk * K;
    uint8_t m = mask_f[u];
    float xu = xf[u];
    float s = 0;
    for (int c = 0; c <= ak; c++)
        for (int bb = 0; bb <= aj; bb++)
            for (int a = 0; a <= ai; a++) s += xc[b + a + sj * bb + sk * c];
    // weights: 1, 1/2, 1/4 or 1/8 for 1, 2, 4 or 8 coarse nodes
    if (m & 64) xf[u] = xu + s * (1.0f / (float)(1 << (ai + aj + ak)));
}

// ---- Galerkin levels ------------------------------------------------------------------------
// a27[q nn + u], q = (di+1) + 3 (dj+1) + 9 (dk+1) (q = 13 the centre): the Laplacian part of
// R J P at regular node u; the Boltzmann derivative stays on the diagonal, injected, in
// idiag = 1 / (a27[13] - P).  Zero off the regular nodes, whose 26 neighbours all exist.
__device__ inline long off27(const Grid &g, int q) {
    int di = q % 3 - 1, dj = (q / 3) % 3 - 1, dk = q / 9 - 1;
    return di + (long)g.ni * (dj + (long)g.nj * dk);
}

// (A x)(u) less the centre term
__device__ inline float a27_offdiag(const Grid &g, const float *a27, const float *x, long u) {
    float s = 0;
#pragma unroll
    for (int q = 0; q < 27; q++)
        if (q != 13) s += a27[q * g.nn + u] * x[u + off27(g, q)];
    return s;
}

// first damped-Jacobi sweep from x = 0: x_out = w b / diag at the regular nodes; the others
// become 0 in both buffers (x_oth may hold face values of the last extension, or a node the
// geometry has since made Dirichlet), so that prolongation reads 0 there
template <typename TB>
__device__ inline void jac_first_node(const Grid &g, const uint8_t *mask, const float *idiag, const TB *b,
                                      float *x_out, float *x_oth, long u) {
    if (mask[u] & 64) {
        x_out[u] = MG_JACOBI_W * (float)b[u] * idiag[u];
    } else {
        x_out[u] = 0.0f;
        x_oth[u] = 0.0f;
    }
}

// damped-Jacobi sweep: x_out = x_in + w (b - A x_in) / diag at the regular nodes (the
// others stay 0 in both buffers)
template <typename TB>
__device__ inline void jac_node(const Grid &g, const float *a27, const uint8_t *mask, const float *idiag,
                                const float *x_in, const TB *b, float *x_out, long u) {
    if (!(mask[u] & 64)) return;
    float s = a27_offdiag(g, a27, x_in, u);
    float du = idiag[u], xu = x_in[u];
    x_out[u] = xu + MG_JACOBI_W * (((float)b[u] - s) * du - xu);
}

// r = b - A x at the regular nodes, 0 elsewhere
template <typename TB>
__device__ inline void res27_node(const Grid &g, const float *a27, const uint8_t *mask, const float *idiag,
                                  const float *x, const TB *b, float *r, long u) {
    float v = 0.0f;
    if (mask[u] & 64) v = (float)b[u] - (a27_offdiag(g, a27, x, u) + x[u] / idiag[u]);
    r[u] = v;
}

// trilinear interpolation of x_c at fine node (i, j, k): 1, 2, 4 or 8 coarse nodes (an odd
// fine index has its upper coarse neighbour inside the coarse grid)
__device__ inline float prolong_at(const Grid &gc, const float *xc, int i, int j, int k) {
    int ai = i & 1, aj = j & 1, ak = k & 1;
    long sj = gc.ni, sk = (long)gc.ni * gc.nj;
    long b = (i >> 1) + sj * (j >> 1) + sk * (k >> 1);
    float s = 0;
    for (int c = 0; c <= ak; c++)
        for (int bb = 0; bb <= aj; bb++)
            for (int a = 0; a <= ai; a++) s += xc[b + a + sj * bb + sk * c];
    return s * (1.0f / (float)(1 << (ai + aj + ak)));
}

// x_f += trilinear interpolation of x_c at every regular node (the Jacobi sweeps that follow
// read both colours)
__device__ inline void prolong_all_node(const Grid &gf, const Grid &gc, const uint8_t *mask_f, const float *xc,
                                        float *xf, long u) {
    int i, j, k;
    node_ijk(gf, u, i, j, k);
    uint8_t m = mask_f[u];
    float xu = xf[u];
    float c = prolong_at(gc, xc, i, j, k);
    if (m & 64) xf[u] = xu + c;
}

// the regular coarse node whose value node K takes under the extension E (K itself, or the
// end of its copy chain); -1 for a Dirichlet node or a chain that ends on one
__device__ inline long ext_target(const Grid &g, const uint8_t *oid, long K, int i, int j, int k) {
    int t = node_type(g, oid, K, i, j, k);
    if (t == NT_REG) return K;
    if (t == NT_DIR) return -1;
    long c = chain_target(g, oid, K, i, j, k);
    int ci, cj, ck;
    node_ijk(g, c, ci, cj, ck);
    return node_type(g, oid, c, ci, cj, ck) == NT_REG ? c : -1;
}

// The operator (Laplacian part) of the level above a Galerkin level: the rediscretised
// 7-point one (a27 == nullptr: neighbour coefficients C, a Neumann neighbour folded into the
// centre as k_mg_ops does) or its own Galerkin operator.
struct FineOp {
    Grid g;
    const uint8_t *oid;
    Coef6 C;
    float dc;  // -2 (idx2 + idy2 + idz2)
    const float *a27;
};

// row I of (1/8) P^T A_f P, P = trilinear interpolation after E on the coarse level, over the
// regular nodes of both levels (the restriction of the cycle is (1/8) P^T: full weighting,
// then the fold E^T)
__device__ inline void rap_row(const FineOp &F, const Grid &gc, const uint8_t *oid_c, long I, float *out,
                               long stride) {
    int Ii, Ij, Ik;
    node_ijk(gc, I, Ii, Ij, Ik);
    double acc[27];
#pragma unroll
    for (int q = 0; q < 27; q++) acc[q] = 0.0;
    if (node_type(gc, oid_c, I, Ii, Ij, Ik) == NT_REG) {
        // a regular node with a non-regular coarse neighbour takes values through copy chains
        // of nodes up to one step away, which reach fine nodes up to three steps from 2I
        bool near = false;
        for (int q = 0; q < 27 && !near; q++) {
            int di = q % 3 - 1, dj = (q / 3) % 3 - 1, dk = q / 9 - 1;
            long K = I + di + (long)gc.ni * (dj + (long)gc.nj * dk);
            near = node_type(gc, oid_c, K, Ii + di, Ij + dj, Ik + dk) != NT_REG;
        }
        const int R = near ? 3 : 1;
        const Grid &gf = F.g;
        auto parent_target = [&](int pi, int pj, int pk) -> long {
            long K = pi + (long)gc.ni * (pj + (long)gc.nj * pk);
            return near ? ext_target(gc, oid_c, K, pi, pj, pk) : K;
        };
        for (int fk = max(0, 2 * Ik - R); fk <= min(gf.nk - 1, 2 * Ik + R); fk++)
            for (int fj = max(0, 2 * Ij - R); fj <= min(gf.nj - 1, 2 * Ij + R); fj++)
                for (int fi = max(0, 2 * Ii - R); fi <= min(gf.ni - 1, 2 * Ii + R); fi++) {
                    long f = fi + (long)gf.ni * (fj + (long)gf.nj * fk);
                    if (node_type(gf, F.oid, f, fi, fj, fk) != NT_REG) continue;
                    // weight of I in the interpolation at f
                    int ai = fi & 1, aj = fj & 1, ak = fk & 1;
                    double w = 1.0 / (double)(1 << (ai + aj + ak));
                    double pf = 0;
                    for (int c = 0; c <= ak; c++)
                        for (int b = 0; b <= aj; b++)
                            for (int a = 0; a <= ai; a++)
                                if (parent_target((fi >> 1) + a, (fj >> 1) + b, (fk >> 1) + c) == I) pf += w;
                    if (pf == 0) continue;
                    // the fine row at f
                    const int nq = F.a27 ? 27 : 7;
                    for (int e = 0; e < nq; e++) {
                        int gi, gj, gk;
                        double av;
                        if (F.a27) {
                            gi = fi + e % 3 - 1;
                            gj = fj + (e / 3) % 3 - 1;
                            gk = fk + e / 9 - 1;
                            av = F.a27[e * gf.nn + f];
                        } else {
                            // e = 0: centre; 1..6: the six neighbours (-z, -y, -x, +x, +y, +z)
                            const int di[7] = {0, 0, 0, -1, 1, 0, 0}, dj[7] = {0, 0, -1, 0, 0, 1, 0},
                                      dk[7] = {0, -1, 0, 0, 0, 0, 1};
                            gi = fi + di[e];
                            gj = fj + dj[e];
                            gk = fk + dk[e];
                            long gg = gi + (long)gf.ni * (gj + (long)gf.nj * gk);
                            if (e == 0) {
                                double dgv = F.dc;
                                for (int d = 0; d < 6; d++) {
                                    long v = f + nb_offset(gf, d);
                                    const int ddi[6] = {0, 0, -1, 1, 0, 0}, ddj[6] = {0, -1, 0, 0, 1, 0},
                                              ddk[6] = {-1, 0, 0, 0, 0, 1};
                                    if (node_type(gf, F.oid, v, fi + ddi[d], fj + ddj[d], fk + ddk[d]) == NT_NEU)
                                        dgv += F.C.c[d];
                                }
                                av = dgv;
                            } else {
                                av = node_type(gf, F.oid, gg, gi, gj, gk) == NT_REG ? F.C.c[e - 1] : 0.0;
                            }
                        }
                        if (av == 0) continue;
                        // interpolation weights of g onto coarse nodes J near I
                        int bi = gi & 1, bj = gj & 1, bk = gk & 1;
                        double wg = 1.0 / (double)(1 << (bi + bj + bk));
                        for (int c = 0; c <= bk; c++)
                            for (int b = 0; b <= bj; b++)
                                for (int a = 0; a <= bi; a++) {
                                    int pi = (gi >> 1) + a, pj = (gj >> 1) + b, pk = (gk >> 1) + c;
                                    long J = parent_target(pi, pj, pk);
                                    if (J < 0) continue;
                                    int Ji, Jj, Jk;
                                    node_ijk(gc, J, Ji, Jj, Jk);
                                    int dI = Ji - Ii, dJ = Jj - Ij, dK = Jk - Ik;
                                    if (dI < -1 || dI > 1 || dJ < -1 || dJ > 1 || dK < -1 || dK > 1) continue;
                                    acc[(dI + 1) + 3 * (dJ + 1) + 9 * (dK + 1)] += pf * av * wg;
                                }
                    }
                }
    }
    for (int q = 0; q < 27; q++) out[q * stride] = (float)(acc[q] / 8.0);
}

// ---- per-level kernels (one node per thread, XCD-remapped workgroups) ---------------------
template <typename TB>
__global__ void __launch_bounds__(256)
    k_mg_pre(Grid g, Coef6 C, const uint8_t *mask, const float *idiag, float *x, const TB *b, const DevScalars *sc) {
    if (sc->done) return;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u < g.nn) mg_pre_node(g, C, mask, idiag, x, b, u);
}

template <typename TB>
__global__ void __launch_bounds__(256)
    k_mg_smooth(Grid g, Coef6 C, const uint8_t *mask, const float *idiag, float *x, const TB *b, int color,
                const DevScalars *sc) {
    if (sc->done) return;
    long t = xcd_block() * blockDim.x + threadIdx.x;
    if (t < smooth_count(g)) mg_smooth_node(g, C, mask, idiag, x, b, color, t);
}

// The last half-sweep of a V-cycle (red, finest level) fused with the CG product g.s:
// thread t updates the t-th red node of its row and adds g*s for that node and for the
// black node beside it (final since the previous half-sweep), so every node is counted
// once.  One partial per workgroup.
__global__ void __launch_bounds__(256)
    k_mg_smooth_dot(Grid g, Coef6 C, const uint8_t *mask, const float *idiag, float *x, const double *b,
                    double *partials, const DevScalars *sc) {
    __shared__ double sh[4];
    if (sc->done) return;
    long t = xcd_block() * blockDim.x + threadIdx.x;
    double acc = 0;
    if (t < smooth_count(g)) {
        unsigned half = (unsigned)(g.ni + 1) / 2;
        unsigned row = (unsigned)t / half, ih = (unsigned)t - row * half;
        unsigned kk = row / (unsigned)g.nj;
        int j = (int)(row - kk * (unsigned)g.nj), k = (int)kk;
        int par = (j + k) & 1;
        int i = (int)(2 * ih) + par, ib = (int)(2 * ih) + 1 - par;
        if (i < g.ni && interior(g, i, j, k)) {
            long u = i + (long)g.ni * row;
            uint8_t m = mask[u];
            float s = nb_sum(g, C, x, u);
            double bb = b[u];
            float du = idiag[u];
            if (m & 64) {
                float xn = ((float)bb - s) * du;
                x[u] = xn;
                acc += bb * (double)xn;
            }
        }
        if (ib < g.ni) {
            long ub = ib + (long)g.ni * row;
            acc += b[ub] * (double)x[ub];
        }
    }
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

template <typename TB>
__global__ void __launch_bounds__(256)
    k_mg_residual(Grid g, Coef6 C, const uint8_t *mask, const float *idiag, const float *x, const TB *b, float *r,
                  const DevScalars *sc) {
    if (sc->done) return;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u < g.nn) mg_residual_node(g, C, mask, idiag, x, b, r, u);
}

__global__ void __launch_bounds__(256)
    k_mg_restrict(Grid gf, Grid gc, const float *rf, float *rc, const DevScalars *sc) {
    if (sc->done) return;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u < gc.nn) mg_restrict_node(gf, gc, rf, rc, u);
}

__global__ void __launch_bounds__(256)
    k_mg_fold(Grid g, const uint8_t *oid, const uint8_t *mask, const float *in, float *out, const DevScalars *sc) {
    if (sc->done) return;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u < g.nn) mg_fold_node(g, oid, mask, in, out, u);
}

__global__ void __launch_bounds__(256) k_mg_extend(Grid g, const uint8_t *oid, float *x, const DevScalars *sc) {
    if (sc->done) return;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u < g.nn) mg_extend_node(g, oid, x, u);
}

__global__ void __launch_bounds__(256)
    k_mg_prolong(Grid gf, Grid gc, const uint8_t *mask_f, const float *xc, float *xf, const DevScalars *sc) {
    if (sc->done) return;
    long t = xcd_block() * blockDim.x + threadIdx.x;
    if (t < smooth_count(gf)) mg_prolong_node(gf, gc, mask_f, xc, xf, t);
}

// Galerkin levels
__global__ void __launch_bounds__(256) k_mg_rap(FineOp F, Grid gc, const uint8_t *oid_c, float *a27) {
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u < gc.nn) rap_row(F, gc, oid_c, u, a27 + u, gc.nn);
}

template <typename TB>
__global__ void __launch_bounds__(256) k_mg_jac_first(Grid g, const uint8_t *mask, const float *idiag, const TB *b,
                                                                                 float *x_out, float *x_oth, const DevScalars *sc) {
    if (sc->done) return;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u < g.nn) jac_first_node(g, mask, idiag, b, x_out, x_oth, u);
}

template <typename TB>
__global__ void __launch_bounds__(256)
    k_mg_jac(Grid g, const float *a27, const uint8_t *mask, const float *idiag, const float *x_in, const TB *b,
             float *x_out, const DevScalars *sc) {
    if (sc->done) return;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u < g.nn) jac_node(g, a27, mask, idiag, x_in, b, x_out, u);
}

template <typename TB>
__global__ void __launch_bounds__(256)
    k_mg_res27(Grid g, const float *a27, const uint8_t *mask, const float *idiag, const float *x, const TB *b,
               float *r, const DevScalars *sc) {
    if (sc->done) return;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u < g.nn) res27_node(g, a27, mask, idiag, x, b, r, u);
}

__global__ void __launch_bounds__(256)
    k_mg_prolong_all(Grid gf, Grid gc, const uint8_t *mask_f, const float *xc, float *xf, const DevScalars *sc) {
    if (sc->done) return;
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u < gf.nn) prolong_all_node(gf, gc, mask_f, xc, xf, u);
}

// reciprocal diagonal of a Galerkin level: 1 / (a27 centre - P) (the mask from k_mg_ops)
__global__ void __launch_bounds__(256)
    k_mg_gal_diag(Grid g, const float *a27, const uint8_t *mask, const float *P, float *idiag) {
    long u = xcd_block() * blockDim.x + threadIdx.x;
    if (u >= g.nn) return;
    idiag[u] = (mask[u] & 64) ? (float)(1.0 / ((double)a27[13 * g.nn + u] - (double)P[u])) : 1.0f;
}

// ---- the coarse levels in one workgroup -------------------------------------------------
// Levels l0..nl-1 (each at most MG_FUSE_NODES nodes) run the down leg, the coarsest solve and
// the up leg of the V-cycle inside one launch of 1024 threads, a workgroup barrier between
// steps; each step is the same per-node code as the per-level kernels.  Replaces some 40
// launches of a few microseconds each per V-cycle (the coarsest solve's sweeps among them).
struct MgDev {
    Grid g;
    Coef6 C;
    const uint8_t *oid, *mask;
    const float *idiag;
    float *x, *r, *b;
    const float *a27;  // Galerkin levels (else nullptr)
    float *x2;
};

// Levels of at most MG_FUSE_MAX nodes (9x9x17 and below at 257x257x513).  With the 27-point
// Galerkin sweeps, 17x17x33 (9537 nodes) runs faster as per-level kernels of 40 workgroups
// than inside the single workgroup (about 1 ms per time step).
#ifndef MG_FUSE_MAX
#define MG_FUSE_MAX 2000
#endif
constexpr long MG_FUSE_NODES = MG_FUSE_MAX;  // about 2 nodes per thread and step
// the largest level a tuning table or a solver may hand to the single workgroup
constexpr long MG_FUSE_NODES_MAX = 100000;
constexpr int MG_FUSE_THREADS = 1024;

__global__ void __launch_bounds__(MG_FUSE_THREADS)
    k_mg_coarse(const MgDev *L, int l0, int nl, int sweeps, const DevScalars *sc) {
    if (sc->done) return;
    const long T = MG_FUSE_THREADS;
    // n damped-Jacobi sweeps of a Galerkin level: from x = 0 (into x2, then alternating) or
    // from x; an even n ends in x
    auto jacobi = [&](const MgDev &f, int n, bool from_zero) {
        for (int s = 0; s < n; s++) {
            if (s == 0 && from_zero) {
                for (long u = threadIdx.x; u < f.g.nn; u += T) jac_first_node(f.g, f.mask, f.idiag, f.b, f.x2, f.x, u);
            } else {
                const bool in2 = (s & 1) != 0;  // odd sweeps read x2 (both starts)
                const float *xi = in2 ? f.x2 : f.x;
                float *xo = in2 ? f.x : f.x2;
                for (long u = threadIdx.x; u < f.g.nn; u += T) jac_node(f.g, f.a27, f.mask, f.idiag, xi, f.b, xo, u);
            }
            __syncthreads();
        }
    };
    for (int l = l0; l < nl - 1; l++) {
        const MgDev &f = L[l], &c = L[l + 1];
        if (f.a27) {
            jacobi(f, 2, true);
            for (long u = threadIdx.x; u < f.g.nn; u += T) res27_node(f.g, f.a27, f.mask, f.idiag, f.x, f.b, f.r, u);
        } else {
            for (long u = threadIdx.x; u < f.g.nn; u += T) mg_pre_node(f.g, f.C, f.mask, f.idiag, f.x, f.b, u);
            __syncthreads();
            for (long u = threadIdx.x; u < f.g.nn; u += T)
                mg_residual_node(f.g, f.C, f.mask, f.idiag, f.x, f.b, f.r, u);
        }
        __syncthreads();
        for (long u = threadIdx.x; u < c.g.nn; u += T) mg_restrict_node(f.g, c.g, f.r, c.r, u);
        __syncthreads();
        for (long u = threadIdx.x; u < c.g.nn; u += T) mg_fold_node(c.g, c.oid, c.mask, c.r, c.b, u);
        __syncthreads();
    }
    {
        const MgDev &c = L[nl - 1];
        if (c.a27) {
            jacobi(c, 2 * sweeps, true);
        } else {
            auto half_sweep = [&](int color) {
                for (long t = threadIdx.x; t < smooth_count(c.g); t += T)
                    mg_smooth_node(c.g, c.C, c.mask, c.idiag, c.x, c.b, color, t);
                __syncthreads();
            };
            for (long u = threadIdx.x; u < c.g.nn; u += T) mg_pre_node(c.g, c.C, c.mask, c.idiag, c.x, c.b, u);
            __syncthreads();
            // (red, black) pairs, then as many (black, red) pairs: a symmetric coarsest solve
            for (int it = 1; it < sweeps; it++) {
                half_sweep(0);
                half_sweep(1);
            }
            for (int it = 0; it < sweeps; it++) {
                half_sweep(1);
                half_sweep(0);
            }
        }
    }
    for (int l = nl - 2; l >= l0; l--) {
        const MgDev &f = L[l], &c = L[l + 1];
        for (long u = threadIdx.x; u < c.g.nn; u += T) mg_extend_node(c.g, c.oid, c.x, u);
        __syncthreads();
        if (f.a27) {
            for (long u = threadIdx.x; u < f.g.nn; u += T) prolong_all_node(f.g, c.g, f.mask, c.x, f.x, u);
            __syncthreads();
            jacobi(f, 2, false);
        } else {
            for (long t = threadIdx.x; t < smooth_count(f.g); t += T)
                mg_prolong_node(f.g, c.g, f.mask, c.x, f.x, t);
            __syncthreads();
            for (long t = threadIdx.x; t < smooth_count(f.g); t += T)
                mg_smooth_node(f.g, f.C, f.mask, f.idiag, f.x, f.b, 1, t);
            __syncthreads();
            for (long t = threadIdx.x; t < smooth_count(f.g); t += T)
                mg_smooth_node(f.g, f.C, f.mask, f.idiag, f.x, f.b, 0, t);
            __syncthreads();
        }
    }
}

// ---- host --------------------------------------------------------------------------------
static PcgParams level_params(const PcgParams &Q0, const Grid &g) {
    PcgParams Q = Q0;
    Q.g = g;
    Q.idx = 1.0 / g.dh[0];
    Q.idy = 1.0 / g.dh[1];
    Q.idz = 1.0 / g.dh[2];
    Q.idx2 = Q.idx * Q.idx;
    Q.idy2 = Q.idy * Q.idy;
    Q.idz2 = Q.idz * Q.idz;
    return Q;
}

static unsigned smooth_blocks(const Grid &g) {
    long half = (g.ni + 1) / 2;
    return nbk8(half * (long)g.nj * g.nk);
}

void mg_destroy(MgHierarchy *H) {
    if (!H) return;
    hipFree(H->dev);
    for (auto &l : H->L) {
        hipFree(l.idiag);
        hipFree(l.x);
        hipFree(l.r);
        hipFree(l.a27);
        hipFree(l.x2);
        if (!l.owned) continue;
        hipFree(l.oid);
        hipFree(l.mask);
        hipFree(l.boltz);
        hipFree(l.b);
    }
    delete H;
}

// Builds the level list (once per solver).  Level 0 borrows the world's object ids and the
// solver's mask; its right-hand side is the CG residual (FP64) and its solution s->sv.
int mg_build(pic_solver *s, MgHierarchy **out) {
    if (*out) return PIC_OK;
    pic_world *w = s->w;
    MgHierarchy *H = new MgHierarchy();
    H->fuse_max = MG_FUSE_NODES;
    double nodes = (double)w->g.nn;
    H->sweeps = s->mg_sweeps > 0 ? s->mg_sweeps : (int)tuned("mg", nodes, 0.0, "sweeps", MG_SWEEPS);
    H->gal_from = s->mg_gal_from > 0 ? s->mg_gal_from : (int)tuned("mg", nodes, 0.0, "gal_from", MG_GALERKIN_FROM);
    H->fuse_max = s->mg_fuse_max > 0 ? s->mg_fuse_max : (long)tuned("mg", nodes, 0.0, "fuse_max", (double)MG_FUSE_NODES);
    H->sweeps = std::max(1, H->sweeps);
    H->gal_from = std::max(1, H->gal_from);
    MgLevel l0;
    l0.g = w->g;
    l0.oid = w->object_id;
    l0.mask = s->mask;
    H->L.push_back(l0);
    for (;;) {
        const Grid &f = H->L.back().g;
        if ((f.ni - 1) % 2 || (f.nj - 1) % 2 || (f.nk - 1) % 2 || f.ni < 5 || f.nj < 5 || f.nk < 5) break;
        MgLevel c;
        c.g = f;
        c.g.ni = (f.ni - 1) / 2 + 1;
        c.g.nj = (f.nj - 1) / 2 + 1;
        c.g.nk = (f.nk - 1) / 2 + 1;
        c.g.nn = (long)c.g.ni * c.g.nj * c.g.nk;
        for (int a = 0; a < 3; a++) c.g.dh[a] = 2 * f.dh[a];
        c.owned = true;
        c.gal = (int)H->L.size() >= H->gal_from;
        H->L.push_back(c);
    }
    hipError_t e = hipSuccess;
    for (size_t l = 0; l < H->L.size() && e == hipSuccess; l++) {
        MgLevel &c = H->L[l];
        size_t n = (size_t)c.g.nn;
        if (e == hipSuccess) e = hipMalloc(&c.idiag, n * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&c.x, n * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&c.r, n * sizeof(float));
        if (c.gal) {
            // x2 is read (times a zero coefficient) at nodes no sweep writes: it starts at 0
            if (e == hipSuccess) e = hipMalloc(&c.a27, 27 * n * sizeof(float));
            if (e == hipSuccess) e = hipMalloc(&c.x2, n * sizeof(float));
            if (e == hipSuccess) e = hipMemset(c.x2, 0, n * sizeof(float));
            if (e == hipSuccess) e = hipMemset(c.x, 0, n * sizeof(float));
        }
        if (!c.owned) continue;
        if (e == hipSuccess) e = hipMalloc(&c.oid, n);
        if (e == hipSuccess) e = hipMalloc(&c.mask, n);
        if (e == hipSuccess) e = hipMalloc(&c.boltz, n * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&c.b, n * sizeof(float));
    }
    if (e != hipSuccess) {
        mg_destroy(H);
        return hip_fail(e, "hipMalloc (multigrid level)", __FILE__, __LINE__);
    }
    const Grid &cg = H->L.back().g;
    // symmetric sweep pairs on the coarsest level: about its width, enough for the few
    // unknowns it has (3 at 3x3x5, 64 at 6x6x6)
    H->coarse_sweeps = std::max(2, std::max(cg.ni, std::max(cg.nj, cg.nk)));
    // levels small enough for one workgroup (level 0's right-hand side is the FP64 CG
    // residual, so the fused cycle starts at level 1 or deeper)
    const int nl = (int)H->L.size();
    for (int l = 1; l < nl; l++)
        if (H->L[l].g.nn <= H->fuse_max && H->L[l].g.nn <= MG_FUSE_NODES_MAX) {
            H->fuse_l0 = l;
            break;
        }
    if (H->fuse_l0 > 0) {
        std::vector<MgDev> hd(nl);
        for (int l = 0; l < nl; l++) {
            const MgLevel &v = H->L[l];
            MgDev &d = hd[l];
            d.g = v.g;
            for (int a = 0; a < 6; a++) {
                double id = 1.0 / v.g.dh[a < 3 ? 2 - a : a - 3];  // slots -z, -y, -x, +x, +y, +z
                d.C.c[a] = (float)(id * id);
            }
            d.oid = v.oid;
            d.mask = v.mask;
            d.idiag = v.idiag;
            d.x = v.x;
            d.r = v.r;
            d.b = v.b;
            d.a27 = v.a27;
            d.x2 = v.x2;
        }
        e = hipMalloc(&H->dev, sizeof(MgDev) * nl);
        if (e == hipSuccess) e = hipMemcpy(H->dev, hd.data(), sizeof(MgDev) * nl, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            mg_destroy(H);
            return hip_fail(e, "multigrid level descriptors", __FILE__, __LINE__);
        }
    }
    *out = H;
    return PIC_OK;
}

// Level operators for the current Newton step: level 0's mask, P and FP32 reciprocal
// diagonal come fresh from k_nrs_F.  The Laplacian parts of the Galerkin operators depend on
// the geometry only and are rebuilt when its stamp changes; their diagonals take each step's
// injected P.
int mg_setup(MgHierarchy *H, const PcgParams &Q0, const double *boltz0, unsigned geom_stamp, hipStream_t st) {
    const bool rap = H->rap_stamp != geom_stamp;
    for (size_t l = 0; l < H->L.size(); l++) {
        MgLevel &c = H->L[l];
        c.Q = level_params(Q0, c.g);
        if (l == 0) continue;
        const MgLevel &f = H->L[l - 1];
        if (l == 1)
            k_mg_inject<double><<<nbk8(c.g.nn), 256, 0, st>>>(f.g, c.g, f.oid, boltz0, c.oid, c.boltz);
        else
            k_mg_inject<float><<<nbk8(c.g.nn), 256, 0, st>>>(f.g, c.g, f.oid, f.boltz, c.oid, c.boltz);
        k_mg_ops<<<nbk8(c.g.nn), 256, 0, st>>>(c.Q, c.oid, c.boltz, c.mask, c.idiag);
        if (!c.gal) continue;
        if (rap) {
            FineOp F{f.g, f.oid, coef6(f.Q), (float)(-2.0 * (f.Q.idx2 + f.Q.idy2 + f.Q.idz2)), f.gal ? f.a27 : nullptr};
            k_mg_rap<<<nbk8(c.g.nn), 256, 0, st>>>(F, c.g, c.oid, c.a27);
        }
        k_mg_gal_diag<<<nbk8(c.g.nn), 256, 0, st>>>(c.g, c.a27, c.mask, c.boltz, c.idiag);
    }
    PIC_LAUNCH_CHECK();
    H->rap_stamp = geom_stamp;
    return PIC_OK;
}

// pre-smoothing, residual and restriction of level l (right-hand side b of type TB)
template <typename TB>
static void down(MgHierarchy *H, int l, const TB *b, hipStream_t st, const DevScalars *sc) {
    MgLevel &f = H->L[l], &c = H->L[l + 1];
    Coef6 C = coef6(f.Q);
    if (f.gal) {  // two damped-Jacobi sweeps from 0 (into x2, then x)
        k_mg_jac_first<TB><<<nbk8(f.g.nn), 256, 0, st>>>(f.g, f.mask, f.idiag, b, f.x2, f.x, sc);
        k_mg_jac<TB><<<nbk8(f.g.nn), 256, 0, st>>>(f.g, f.a27, f.mask, f.idiag, f.x2, b, f.x, sc);
        k_mg_res27<TB><<<nbk8(f.g.nn), 256, 0, st>>>(f.g, f.a27, f.mask, f.idiag, f.x, b, f.r, sc);
        k_mg_restrict<<<nbk8(c.g.nn), 256, 0, st>>>(f.g, c.g, f.r, c.r, sc);
        k_mg_fold<<<nbk8(c.g.nn), 256, 0, st>>>(c.g, c.oid, c.mask, c.r, c.b, sc);
        return;
    }
    k_mg_pre<TB><<<nbk8(f.g.nn), 256, 0, st>>>(f.g, C, f.mask, f.idiag, f.x, b, sc);
    for (int it = 1; it < H->sweeps; it++) {  // further (red, black) pairs
        k_mg_smooth<TB><<<smooth_blocks(f.g), 256, 0, st>>>(f.g, C, f.mask, f.idiag, f.x, b, 0, sc);
        k_mg_smooth<TB><<<smooth_blocks(f.g), 256, 0, st>>>(f.g, C, f.mask, f.idiag, f.x, b, 1, sc);
    }
    k_mg_residual<TB><<<nbk8(f.g.nn), 256, 0, st>>>(f.g, C, f.mask, f.idiag, f.x, b, f.r, sc);
    k_mg_restrict<<<nbk8(c.g.nn), 256, 0, st>>>(f.g, c.g, f.r, c.r, sc);
    k_mg_fold<<<nbk8(c.g.nn), 256, 0, st>>>(c.g, c.oid, c.mask, c.r, c.b, sc);
}

// the last (red) half-sweep of the finest level, with the product g.s when asked for
static void last_sweep(MgLevel &f, const double *b0, double *dot_partials, hipStream_t st, const DevScalars *sc) {
    Coef6 C = coef6(f.Q);
    if (dot_partials)
        k_mg_smooth_dot<<<smooth_blocks(f.g), 256, 0, st>>>(f.g, C, f.mask, f.idiag, f.x, b0, dot_partials, sc);
    else
        k_mg_smooth<double><<<smooth_blocks(f.g), 256, 0, st>>>(f.g, C, f.mask, f.idiag, f.x, b0, 0, sc);
}

// prolongation of level l+1 into level l, then post-smoothing (black, red); on level 0 the
// red half-sweep is last_sweep
template <typename TB>
static void up(MgHierarchy *H, int l, const TB *b, hipStream_t st, const DevScalars *sc) {
    MgLevel &f = H->L[l], &c = H->L[l + 1];
    Coef6 C = coef6(f.Q);
    k_mg_extend<<<nbk8(c.g.nn), 256, 0, st>>>(c.g, c.oid, c.x, sc);
    if (f.gal) {  // x += P x_c at every regular node, then two damped-Jacobi sweeps
        k_mg_prolong_all<<<nbk8(f.g.nn), 256, 0, st>>>(f.g, c.g, f.mask, c.x, f.x, sc);
        k_mg_jac<TB><<<nbk8(f.g.nn), 256, 0, st>>>(f.g, f.a27, f.mask, f.idiag, f.x, b, f.x2, sc);
        k_mg_jac<TB><<<nbk8(f.g.nn), 256, 0, st>>>(f.g, f.a27, f.mask, f.idiag, f.x2, b, f.x, sc);
        return;
    }
    k_mg_prolong<<<smooth_blocks(f.g), 256, 0, st>>>(f.g, c.g, f.mask, c.x, f.x, sc);
    k_mg_smooth<TB><<<smooth_blocks(f.g), 256, 0, st>>>(f.g, C, f.mask, f.idiag, f.x, b, 1, sc);
    for (int it = 1; it < H->sweeps; it++) {  // the pre-smoothing pairs in reverse: (black, red)
        k_mg_smooth<TB><<<smooth_blocks(f.g), 256, 0, st>>>(f.g, C, f.mask, f.idiag, f.x, b, 0, sc);
        k_mg_smooth<TB><<<smooth_blocks(f.g), 256, 0, st>>>(f.g, C, f.mask, f.idiag, f.x, b, 1, sc);
    }
    if (l > 0) k_mg_smooth<TB><<<smooth_blocks(f.g), 256, 0, st>>>(f.g, C, f.mask, f.idiag, f.x, b, 0, sc);
}

// x0 = V(b0): one V-cycle from a zero initial guess on every level; x0 is level 0's
// solution (FP32), b0 the CG residual (FP64) and b0f its FP32 copy.  The sweeps of level 0
// read b0f (half the bytes, and the same value: they compute in FP32); the last half-sweep
// reads b0, for its FP64 partials of b0.x0 (mg_dot_partials of them) when asked for.
int mg_vcycle(MgHierarchy *H, const double *b0, const float *b0f, float **x0, double *dot_partials, hipStream_t st,
              const DevScalars *sc) {
    const int nl = (int)H->L.size();
    *x0 = H->L[0].x;
    MgLevel &f0 = H->L[0];
    if (nl == 1) {  // no coarsening possible: the preconditioner is one symmetric sweep pair
        Coef6 C = coef6(f0.Q);
        k_mg_pre<float><<<nbk8(f0.g.nn), 256, 0, st>>>(f0.g, C, f0.mask, f0.idiag, f0.x, b0f, sc);
        k_mg_smooth<float><<<smooth_blocks(f0.g), 256, 0, st>>>(f0.g, C, f0.mask, f0.idiag, f0.x, b0f, 1, sc);
        last_sweep(f0, b0, dot_partials, st, sc);
        PIC_LAUNCH_CHECK();
        return PIC_OK;
    }
    down<float>(H, 0, b0f, st, sc);
    if (H->fuse_l0 > 0) {
        const int l0 = H->fuse_l0;
        for (int l = 1; l < l0; l++) down<float>(H, l, H->L[l].b, st, sc);
        k_mg_coarse<<<1, MG_FUSE_THREADS, 0, st>>>(H->dev, l0, nl, H->coarse_sweeps, sc);
        for (int l = l0 - 1; l >= 1; l--) up<float>(H, l, H->L[l].b, st, sc);
        up<float>(H, 0, b0f, st, sc);
        last_sweep(f0, b0, dot_partials, st, sc);
        PIC_LAUNCH_CHECK();
        return PIC_OK;
    }
    for (int l = 1; l < nl - 1; l++) down<float>(H, l, H->L[l].b, st, sc);
    {
        MgLevel &c = H->L[nl - 1];
        Coef6 C = coef6(c.Q);
        if (c.gal) {  // 2 coarse_sweeps damped-Jacobi sweeps from 0, ending in x
            k_mg_jac_first<float><<<nbk8(c.g.nn), 256, 0, st>>>(c.g, c.mask, c.idiag, c.b, c.x2, c.x, sc);
            for (int it = 1; it < 2 * H->coarse_sweeps; it++) {
                const bool in2 = (it & 1) != 0;
                k_mg_jac<float><<<nbk8(c.g.nn), 256, 0, st>>>(c.g, c.a27, c.mask, c.idiag, in2 ? c.x2 : c.x, c.b,
                                                               in2 ? c.x : c.x2, sc);
            }
        } else {
        k_mg_pre<float><<<nbk8(c.g.nn), 256, 0, st>>>(c.g, C, c.mask, c.idiag, c.x, c.b, sc);
        // symmetric sweeps: (red, black) pairs, then as many (black, red) pairs
        for (int it = 1; it < H->coarse_sweeps; it++) {
            k_mg_smooth<float><<<smooth_blocks(c.g), 256, 0, st>>>(c.g, C, c.mask, c.idiag, c.x, c.b, 0, sc);
            k_mg_smooth<float><<<smooth_blocks(c.g), 256, 0, st>>>(c.g, C, c.mask, c.idiag, c.x, c.b, 1, sc);
        }
        for (int it = 0; it < H->coarse_sweeps; it++) {
            k_mg_smooth<float><<<smooth_blocks(c.g), 256, 0, st>>>(c.g, C, c.mask, c.idiag, c.x, c.b, 1, sc);
            k_mg_smooth<float><<<smooth_blocks(c.g), 256, 0, st>>>(c.g, C, c.mask, c.idiag, c.x, c.b, 0, sc);
        }
        }
    }
    for (int l = nl - 2; l >= 1; l--) up<float>(H, l, H->L[l].b, st, sc);
    up<float>(H, 0, b0f, st, sc);
    last_sweep(f0, b0, dot_partials, st, sc);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

// number of partials the fused product of mg_vcycle writes
int mg_dot_partials(const MgHierarchy *H) { return (int)smooth_blocks(H->L[0].g); }

int mg_levels(const MgHierarchy *H) { return H ? (int)H->L.size() : 0; }

// level 0's FP32 reciprocal diagonal (written by the Newton residual kernel)
float *mg_level0_idiag(MgHierarchy *H) { return H->L[0].idiag; }

}  // namespace pic
