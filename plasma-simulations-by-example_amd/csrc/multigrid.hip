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
// object id and Boltzmann derivative P from the fine node at the same position (injection);
// each level discretises the reduced 7-point operator at its own spacing, with Neumann face
// nodes eliminated into the diagonal exactly as on the finest level.
//
// Transfers.  Prolongation is trilinear interpolation of the coarse correction after its
// face nodes are filled from their copy-chain targets (E), written to regular fine nodes
// only.  Restriction is its transpose over 8 (full weighting, then the face values folded
// into their chain targets, E^T), so that the V-cycle is a symmetric operator.
//
// Smoothing.  One red-black Gauss-Seidel sweep before (red, black) and one after (black,
// red) the coarse correction: the post-smoother is the adjoint of the pre-smoother, as CG
// requires of its preconditioner.  The coarsest level is solved by sweeps.
//
// Every kernel here returns at once when the solve has converged (DevScalars::done), so
// the host can queue several iterations between convergence checks.
#include <algorithm>
#include <vector>

#include "stencil.h"

#pragma clang fp contract(off)

namespace pic {

struct MgLevel {
    Grid g;                   // ni, nj, nk, nn, dh
    PcgParams Q;              // coefficients at this spacing
    uint8_t *oid = nullptr;   // object ids (level 0: the world's)
    uint8_t *mask = nullptr;  // bits 0-5 regular neighbours, bit 6 regular (level 0: the solver's)
    double *diag = nullptr;   // reduced diagonal (level 0: the solver's jr)
    double *boltz = nullptr;  // Boltzmann derivative P (level 0: the solver's)
    double *x = nullptr, *b = nullptr, *r = nullptr;
    bool owned = false;
};

struct MgHierarchy {
    std::vector<MgLevel> L;
    int coarse_sweeps = 0;
};

// ---- kernels -----------------------------------------------------------------------------
// coarse object id and P by injection from the fine node at the same position
__global__ void __launch_bounds__(256)
    k_mg_inject(Grid gf, Grid gc, const uint8_t *oid_f, const double *P_f, uint8_t *oid_c, double *P_c) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= gc.nn) return;
    int i, j, k;
    node_ijk(gc, u, i, j, k);
    long f = 2L * i + (long)gf.ni * (2L * j + (long)gf.nj * 2L * k);
    oid_c[u] = oid_f[f];
    P_c[u] = P_f[f];
}

// mask and reduced diagonal of a coarse level (the operator part of k_nrs_F)
__global__ void __launch_bounds__(256)
    k_mg_ops(PcgParams Q, const uint8_t *oid, const double *P, uint8_t *mask, double *diag) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const Grid &g = Q.g;
    if (u >= g.nn) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    if (node_type(g, oid, u, i, j, k) != NT_REG) {
        mask[u] = 0;
        diag[u] = 1;
        return;
    }
    double dg = -2.0 * (Q.idx2 + Q.idy2 + Q.idz2) - P[u];
    uint8_t m = 1u << 6;
    const int di[6] = {0, 0, -1, 1, 0, 0}, dj[6] = {0, -1, 0, 0, 1, 0}, dk[6] = {-1, 0, 0, 0, 0, 1};
    for (int d = 0; d < 6; d++) {
        long v = u + nb_offset(g, d);
        int tv = node_type(g, oid, v, i + di[d], j + dj[d], k + dk[d]);
        if (tv == NT_REG) m |= (uint8_t)(1u << d);
        else if (tv == NT_NEU) dg += nb_coef(Q, d);
    }
    mask[u] = m;
    diag[u] = dg;
}

// first half-sweep from x = 0: red regular nodes x = b/diag, every other node 0
__global__ void __launch_bounds__(256)
    k_mg_smooth_first(Grid g, const uint8_t *mask, const double *diag, double *x, const double *b,
                      const DevScalars *sc) {
    if (sc->done) return;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= g.nn) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    bool red = ((i + j + k) & 1) == 0;
    x[u] = (red && (mask[u] & 64)) ? b[u] / diag[u] : 0.0;
}

// Gauss-Seidel half-sweep over the regular nodes of one colour
__global__ void __launch_bounds__(256)
    k_mg_smooth(PcgParams Q, const uint8_t *mask, const double *diag, double *x, const double *b, int color,
                const DevScalars *sc) {
    if (sc->done) return;
    const Grid &g = Q.g;
    // thread t handles the t-th node of this colour: x-fastest, every second node of a row
    long half = (g.ni + 1) / 2;
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long row = t / half;
    if (row >= (long)g.nj * g.nk) return;
    int j = (int)(row % g.nj), k = (int)(row / g.nj);
    int i = (int)(2 * (t % half)) + ((j + k + color) & 1);
    if (i >= g.ni) return;
    long u = i + (long)g.ni * row;
    uint8_t m = mask[u];
    if (!(m & 64)) return;
    double s = 0;
    for (int d = 0; d < 6; d++)
        if (m & (1u << d)) s += nb_coef(Q, d) * x[u + nb_offset(g, d)];
    x[u] = (b[u] - s) / diag[u];
}

// r = b - J x on regular nodes, 0 elsewhere
__global__ void __launch_bounds__(256)
    k_mg_residual(PcgParams Q, const uint8_t *mask, const double *diag, const double *x, const double *b,
                  double *r, const DevScalars *sc) {
    if (sc->done) return;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= Q.g.nn) return;
    uint8_t m = mask[u];
    r[u] = (m & 64) ? b[u] - jr_apply(Q, u, m, diag[u], x) : 0.0;
}

// full weighting of the fine residual onto every coarse node (faces included)
__global__ void __launch_bounds__(256)
    k_mg_restrict(Grid gf, Grid gc, const double *rf, double *rc, const DevScalars *sc) {
    if (sc->done) return;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= gc.nn) return;
    int I, J, K;
    node_ijk(gc, u, I, J, K);
    long sj = gf.ni, sk = (long)gf.ni * gf.nj;
    double s = 0;
    for (int c = -1; c <= 1; c++) {
        int k = 2 * K + c;
        if (k < 0 || k >= gf.nk) continue;
        double wc = c == 0 ? 0.5 : 0.25;
        for (int b = -1; b <= 1; b++) {
            int j = 2 * J + b;
            if (j < 0 || j >= gf.nj) continue;
            double wb = wc * (b == 0 ? 0.5 : 0.25);
            for (int a = -1; a <= 1; a++) {
                int i = 2 * I + a;
                if (i < 0 || i >= gf.ni) continue;
                s += wb * (a == 0 ? 0.5 : 0.25) * rf[i + sj * j + sk * k];
            }
        }
    }
    rc[u] = s;
}

// E^T: every regular coarse node adds the values of the face nodes whose copy chain ends on
// it (at most 7, fixed order); non-regular nodes become 0
__global__ void __launch_bounds__(256)
    k_mg_fold(Grid g, const uint8_t *oid, const uint8_t *mask, const double *in, double *out,
              const DevScalars *sc) {
    if (sc->done) return;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= g.nn) return;
    if (!(mask[u] & 64)) {
        out[u] = 0.0;
        return;
    }
    int i, j, k;
    node_ijk(g, u, i, j, k);
    double s = in[u];
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

// E: face nodes take the value of their chain target (Dirichlet nodes keep 0)
__global__ void __launch_bounds__(256) k_mg_extend(Grid g, const uint8_t *oid, double *x, const DevScalars *sc) {
    if (sc->done) return;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= g.nn) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    if (node_type(g, oid, u, i, j, k) != NT_NEU) return;
    x[u] = x[chain_target(g, oid, u, i, j, k)];
}

// x_f += trilinear interpolation of x_c, on regular fine nodes
__global__ void __launch_bounds__(256)
    k_mg_prolong(Grid gf, Grid gc, const uint8_t *mask_f, const double *xc, double *xf, const DevScalars *sc) {
    if (sc->done) return;
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= gf.nn) return;
    if (!(mask_f[u] & 64)) return;
    int i, j, k;
    node_ijk(gf, u, i, j, k);
    int I = i >> 1, J = j >> 1, K = k >> 1;
    int ai = i & 1, aj = j & 1, ak = k & 1;
    long sj = gc.ni, sk = (long)gc.ni * gc.nj;
    long b = I + sj * J + sk * K;
    double s = 0;
    for (int c = 0; c <= ak; c++)
        for (int bb = 0; bb <= aj; bb++)
            for (int a = 0; a <= ai; a++) s += xc[b + a + sj * bb + sk * c];
    // weights: 1, 1/2, 1/4 or 1/8 for 1, 2, 4 or 8 coarse nodes
    xf[u] += s * (1.0 / (double)(1 << (ai + aj + ak)));
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
    return nbk(half * (long)g.nj * g.nk);
}

void mg_destroy(MgHierarchy *H) {
    if (!H) return;
    for (auto &l : H->L) {
        if (!l.owned) continue;
        hipFree(l.oid);
        hipFree(l.mask);
        hipFree(l.diag);
        hipFree(l.boltz);
        hipFree(l.x);
        hipFree(l.b);
        hipFree(l.r);
    }
    delete H;
}

// Builds the level list (once per solver); level 0 borrows the solver's arrays.
int mg_build(pic_solver *s, MgHierarchy **out) {
    if (*out) return PIC_OK;
    pic_world *w = s->w;
    MgHierarchy *H = new MgHierarchy();
    MgLevel l0;
    l0.g = w->g;
    l0.oid = w->object_id;
    l0.mask = s->mask;
    l0.diag = s->P;
    l0.boltz = s->boltz;
    l0.x = s->sv;
    l0.b = s->g;
    l0.r = s->rv;
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
        size_t n = (size_t)c.g.nn;
        hipError_t e = hipSuccess;
        if (e == hipSuccess) e = hipMalloc(&c.oid, n);
        if (e == hipSuccess) e = hipMalloc(&c.mask, n);
        if (e == hipSuccess) e = hipMalloc(&c.diag, n * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&c.boltz, n * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&c.x, n * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&c.b, n * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&c.r, n * sizeof(double));
        H->L.push_back(c);
        if (e != hipSuccess) {
            mg_destroy(H);
            return hip_fail(e, "hipMalloc (multigrid level)", __FILE__, __LINE__);
        }
    }
    const Grid &cg = H->L.back().g;
    H->coarse_sweeps = std::max(8, cg.ni + cg.nj + cg.nk);
    *out = H;
    return PIC_OK;
}

// Coarse-level operators for the current Newton step (level 0's mask, jr and P are fresh
// from k_nrs_F).
int mg_setup(MgHierarchy *H, const PcgParams &Q0, hipStream_t st) {
    for (size_t l = 0; l < H->L.size(); l++) {
        MgLevel &c = H->L[l];
        c.Q = level_params(Q0, c.g);
        if (l == 0) continue;
        const MgLevel &f = H->L[l - 1];
        k_mg_inject<<<nbk(c.g.nn), 256, 0, st>>>(f.g, c.g, f.oid, f.boltz, c.oid, c.boltz);
        k_mg_ops<<<nbk(c.g.nn), 256, 0, st>>>(c.Q, c.oid, c.boltz, c.mask, c.diag);
    }
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

// x_0 = V(b_0): one V-cycle from a zero initial guess on every level
int mg_vcycle(MgHierarchy *H, hipStream_t st, const DevScalars *sc) {
    const int nl = (int)H->L.size();
    for (int l = 0; l < nl - 1; l++) {
        MgLevel &f = H->L[l], &c = H->L[l + 1];
        k_mg_smooth_first<<<nbk(f.g.nn), 256, 0, st>>>(f.g, f.mask, f.diag, f.x, f.b, sc);
        k_mg_smooth<<<smooth_blocks(f.g), 256, 0, st>>>(f.Q, f.mask, f.diag, f.x, f.b, 1, sc);
        k_mg_residual<<<nbk(f.g.nn), 256, 0, st>>>(f.Q, f.mask, f.diag, f.x, f.b, f.r, sc);
        k_mg_restrict<<<nbk(c.g.nn), 256, 0, st>>>(f.g, c.g, f.r, c.r, sc);
        k_mg_fold<<<nbk(c.g.nn), 256, 0, st>>>(c.g, c.oid, c.mask, c.r, c.b, sc);
    }
    {
        MgLevel &c = H->L[nl - 1];
        k_mg_smooth_first<<<nbk(c.g.nn), 256, 0, st>>>(c.g, c.mask, c.diag, c.x, c.b, sc);
        k_mg_smooth<<<smooth_blocks(c.g), 256, 0, st>>>(c.Q, c.mask, c.diag, c.x, c.b, 1, sc);
        // symmetric sweeps: (red, black) pairs, then as many (black, red) pairs
        for (int it = 1; it < H->coarse_sweeps; it++) {
            k_mg_smooth<<<smooth_blocks(c.g), 256, 0, st>>>(c.Q, c.mask, c.diag, c.x, c.b, 0, sc);
            k_mg_smooth<<<smooth_blocks(c.g), 256, 0, st>>>(c.Q, c.mask, c.diag, c.x, c.b, 1, sc);
        }
        for (int it = 0; it < H->coarse_sweeps; it++) {
            k_mg_smooth<<<smooth_blocks(c.g), 256, 0, st>>>(c.Q, c.mask, c.diag, c.x, c.b, 1, sc);
            k_mg_smooth<<<smooth_blocks(c.g), 256, 0, st>>>(c.Q, c.mask, c.diag, c.x, c.b, 0, sc);
        }
    }
    for (int l = nl - 2; l >= 0; l--) {
        MgLevel &f = H->L[l], &c = H->L[l + 1];
        k_mg_extend<<<nbk(c.g.nn), 256, 0, st>>>(c.g, c.oid, c.x, sc);
        k_mg_prolong<<<nbk(f.g.nn), 256, 0, st>>>(f.g, c.g, f.mask, c.x, f.x, sc);
        k_mg_smooth<<<smooth_blocks(f.g), 256, 0, st>>>(f.Q, f.mask, f.diag, f.x, f.b, 1, sc);
        k_mg_smooth<<<smooth_blocks(f.g), 256, 0, st>>>(f.Q, f.mask, f.diag, f.x, f.b, 0, sc);
    }
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

int mg_levels(const MgHierarchy *H) { return H ? (int)H->L.size() : 0; }

}  // namespace pic
