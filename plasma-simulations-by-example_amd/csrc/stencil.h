// Node-stencil helpers shared by the Poisson solvers (solver.hip) and the multigrid
// preconditioner (multigrid.hip): node classification, the Neumann copy chains of the
// reduced system, the reduced 7-point operator and fixed-order reductions.
#ifndef PICAMD_STENCIL_H
#define PICAMD_STENCIL_H

#include "picamd_internal.h"

namespace pic {

// Node classes: regular (an unknown), Neumann face (copies the node its chain reaches),
// Dirichlet (object_id > 0: sphere, inlet).
enum { NT_REG = 0, NT_NEU = 1, NT_DIR = 2 };

struct PcgParams {
    Grid g;
    double idx, idy, idz, idx2, idy2, idz2;
    double phi0, Te0, n0;
    double tol;
};

__device__ inline int node_type(const Grid &g, const uint8_t *oid, long u, int i, int j, int k) {
    if (oid[u] > 0) return NT_DIR;
    if (i == 0 || i == g.ni - 1 || j == 0 || j == g.nj - 1 || k == 0 || k == g.nk - 1) return NT_NEU;
    return NT_REG;
}

// Face nodes are not unknowns of the reduced system: each equals the node its copy chain
// reaches (the priority i, j, k of ch3/ver2/PotentialSolver.cpp:359-370, which is what the
// reference's Neumann rows enforce at convergence).
__device__ inline long chain_target(const Grid &g, const uint8_t *oid, long u, int i, int j, int k) {
    long sj = g.ni, sk = (long)g.ni * g.nj;
    for (int hop = 0; hop < 3; hop++) {
        if (oid[u] > 0) return u;
        if (i == 0) { i = 1; u += 1; }
        else if (i == g.ni - 1) { i = g.ni - 2; u -= 1; }
        else if (j == 0) { j = 1; u += sj; }
        else if (j == g.nj - 1) { j = g.nj - 2; u -= sj; }
        else if (k == 0) { k = 1; u += sk; }
        else if (k == g.nk - 1) { k = g.nk - 2; u -= sk; }
        else return u;
    }
    return u;
}

// direction d: 0 -z, 1 -y, 2 -x, 3 +x, 4 +y, 5 +z (the reference's slot order around the centre)
__device__ inline long nb_offset(const Grid &g, int d) {
    long sj = g.ni, sk = (long)g.ni * g.nj;
    const long o[6] = {-sk, -sj, -1, 1, sj, sk};
    return o[d];
}

__device__ inline double nb_coef(const PcgParams &Q, int d) {
    const double c[6] = {Q.idz2, Q.idy2, Q.idx2, Q.idx2, Q.idy2, Q.idz2};
    return c[d];
}

// reduced operator row: regular neighbours (mask bits 0-5) in slot order, diagonal jru
__device__ inline double jr_apply(const PcgParams &Q, long u, uint8_t m, double jru, const double *v) {
    const Grid &g = Q.g;
    double s = 0;
    for (int d = 0; d < 3; d++)
        if (m & (1u << d)) s += nb_coef(Q, d) * v[u + nb_offset(g, d)];
    s += jru * v[u];
    for (int d = 3; d < 6; d++)
        if (m & (1u << d)) s += nb_coef(Q, d) * v[u + nb_offset(g, d)];
    return s;
}

// sum of nb per-block partials by one workgroup, fixed order
__device__ inline double fixed_sum(const double *p, int nb, double *sh) {
    double s = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += p[b];
    return block_sum(s, sh);
}


}  // namespace pic
#endif
