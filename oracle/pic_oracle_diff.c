/*
 * pic_oracle_diff.c -- FP64 CPU restatement of the differentiable flow past a sphere.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the GPU library's pic_adjoint_* entry points
 * (csrc/adjoint.hip).  Finite differences of orc_dsph_run give the FP64 CPU reference
 * gradient the GPU adjoint is compared with.  The product library never calls this.
 *
 * The run is the ch3/ver2 loop (ch3/ver2/Main.cpp:50-70) built from the oracle's restated
 * reference functions -- gather (Field.h:189-211), scatter (:167-186), node volumes
 * (World.cpp:59-70), addSphere / addInlet (:88-116), computeEF (PotentialSolver.cpp:
 * 465-504), the addParticle rewind (Species.cpp:65-81), the leapfrog (Species.cpp:14-26) --
 * with the build's three changes that make it differentiable (soft absorption layers, a
 * fixed injection count carrying the flux in its weight, counter-based random numbers),
 * and the Newton solve on the Neumann-reduced symmetric system (the fixed point of
 * solveNRPCG, :225-296) by Jacobi-preconditioned CG, converged to tight tolerances.
 *
 * Pinning: orc_solve_nr_sym against orc_solve_nrpcg (itself bit-exact against the compiled
 * reference) at the reference's tolerances and against the uniform-density Boltzmann closed
 * form; orc_philox4x32_10 against the published known-answer vectors of Random123;
 * orc_dsph_run's weight bookkeeping (fill + injected = live + taken by the sphere + taken
 * by the walls) and its agreement with a run whose soft layers are thinner.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pic_oracle.h"

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), restated from the paper */
void orc_philox4x32_10(const uint32_t ctr_in[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)M0 * c[0];
        uint64_t p1 = (uint64_t)M1 * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
        k0 += W0;
        k1 += W1;
    }
    memcpy(out, c, sizeof c);
}

static double u01(uint32_t a, uint32_t b) {
    uint64_t v = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
    return (double)(v & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
}

static void draws16(uint64_t seed, long q, int n, uint32_t tag, double u[16]) {
    for (int b = 0; b < 8; b++) {
        uint32_t c[4] = {(uint32_t)q, (uint32_t)((uint64_t)q >> 32) ^ (tag << 24), (uint32_t)n, (uint32_t)b};
        uint32_t r[4];
        orc_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32), r);
        u[2 * b] = u01(r[0], r[1]);
        u[2 * b + 1] = u01(r[2], r[3]);
    }
}

/* the isotropic thermal velocity of ch4/Species.cpp:149-174, from u[3..13] */
static void thermal(double vth, const double u[16], double v[3]) {
    v[0] = v[1] = v[2] = 0.0;
    if (vth <= 0.0) return;
    double theta = 2 * ORC_PI * u[3];
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

/* ---- Newton on the Neumann-reduced system ----------------------------------------------- */
enum { REG = 0, NEU = 1, DIR = 2 };

static int node_type(const orc_grid *g, const int *oid, long u, int i, int j, int k) {
    if (oid[u] > 0) return DIR;
    if (i == 0 || i == g->ni - 1 || j == 0 || j == g->nj - 1 || k == 0 || k == g->nk - 1) return NEU;
    return REG;
}

/* the node a face node copies: priority i, j, k (ch3/ver2/PotentialSolver.cpp:359-370) */
static long chain_target(const orc_grid *g, const int *oid, long u, int i, int j, int k) {
    long sj = g->ni, sk = (long)g->ni * g->nj;
    for (int hop = 0; hop < 3; hop++) {
        if (oid[u] > 0) return u;
        if (i == 0) { i = 1; u += 1; }
        else if (i == g->ni - 1) { i = g->ni - 2; u -= 1; }
        else if (j == 0) { j = 1; u += sj; }
        else if (j == g->nj - 1) { j = g->nj - 2; u -= sj; }
        else if (k == 0) { k = 1; u += sk; }
        else if (k == g->nk - 1) { k = g->nk - 2; u -= sk; }
        else return u;
    }
    return u;
}

typedef struct {
    const orc_grid *g;
    const int *oid;
    double c[6];  /* neighbour coefficients: -z, -y, -x, +x, +y, +z */
    long off[6];
    long nn;
    unsigned char *mask;  /* bits 0-5 regular neighbours, bit 6 regular */
    double *jr;           /* reduced Jacobian diagonal */
} Reduced;

static double apply_row(const Reduced *R, long u, const double *v) {
    unsigned char m = R->mask[u];
    double s = 0;
    for (int d = 0; d < 3; d++)
        if (m & (1u << d)) s += R->c[d] * v[u + R->off[d]];
    s += R->jr[u] * v[u];
    for (int d = 3; d < 6; d++)
        if (m & (1u << d)) s += R->c[d] * v[u + R->off[d]];
    return s;
}

static double dotr(const Reduced *R, const double *a, const double *b) {
    double s = 0;
    for (long u = 0; u < R->nn; u++)
        if (R->mask[u] & 64) s += a[u] * b[u];
    return s;
}

/* CG with the Jacobi preconditioner on J y = F (regular nodes), y from 0, until
 * RMS(J y - F) over all nodes < tol.  Returns iterations (negative if not converged). */
static int cg_reduced(const Reduced *R, const double *F, double *y, double tol, int max_it, double *work) {
    long nn = R->nn;
    double *g = work, *s = work + nn, *d = work + 2 * nn, *z = work + 3 * nn;
    for (long u = 0; u < nn; u++) {
        y[u] = 0;
        g[u] = (R->mask[u] & 64) ? -F[u] : 0.0;
        s[u] = (R->mask[u] & 64) ? g[u] / R->jr[u] : 0.0;
        d[u] = -s[u];
        z[u] = 0;
    }
    double a = dotr(R, g, s);
    for (int it = 0; it < max_it; it++) {
        double gg = dotr(R, g, g);
        if (sqrt(gg / (double)nn) < tol) return it;
        for (long u = 0; u < nn; u++) z[u] = (R->mask[u] & 64) ? apply_row(R, u, d) : 0.0;
        double b = dotr(R, d, z);
        double lam = a / b;
        for (long u = 0; u < nn; u++) {
            if (!(R->mask[u] & 64)) continue;
            y[u] += lam * d[u];
            g[u] += lam * z[u];
            s[u] = g[u] / R->jr[u];
        }
        double a2 = dotr(R, g, s);
        double mu = a2 / a;
        a = a2;
        for (long u = 0; u < nn; u++)
            if (R->mask[u] & 64) d[u] = mu * d[u] - s[u];
    }
    double gg = dotr(R, g, g);
    return sqrt(gg / (double)nn) < tol ? max_it : -max_it;
}

int orc_solve_nr_sym(const orc_grid *g, double *phi, const double *rho, const int *oid, double phi0,
                     double Te0, double n0, double cg_tol, double nr_tol, int max_cg, double *stats) {
    long nn = (long)g->ni * g->nj * g->nk;
    long sj = g->ni, sk = (long)g->ni * g->nj;
    double idx2 = 1.0 / (g->dh[0] * g->dh[0]), idy2 = 1.0 / (g->dh[1] * g->dh[1]),
           idz2 = 1.0 / (g->dh[2] * g->dh[2]);
    Reduced R;
    R.g = g;
    R.oid = oid;
    R.nn = nn;
    double c[6] = {idz2, idy2, idx2, idx2, idy2, idz2};
    long off[6] = {-sk, -sj, -1, 1, sj, sk};
    memcpy(R.c, c, sizeof c);
    memcpy(R.off, off, sizeof off);
    R.mask = calloc((size_t)nn, 1);
    R.jr = calloc((size_t)nn, sizeof(double));
    double *F = calloc((size_t)nn, sizeof(double));
    double *y = calloc((size_t)nn, sizeof(double));
    double *work = calloc((size_t)nn * 4, sizeof(double));
    int di[6] = {0, 0, -1, 1, 0, 0}, dj[6] = {0, -1, 0, 0, 1, 0}, dk[6] = {-1, 0, 0, 0, 0, 1};
    /* face nodes start as copies of their chain targets */
    for (int k = 0; k < g->nk; k++)
        for (int j = 0; j < g->nj; j++)
            for (int i = 0; i < g->ni; i++) {
                long u = i + sj * j + sk * k;
                if (node_type(g, oid, u, i, j, k) == NEU) phi[u] = phi[chain_target(g, oid, u, i, j, k)];
            }
    int converged = 0, newton = 0, cg_total = 0;
    double norm = 0;
    for (int it = 0; it < 50; it++) {
        for (int k = 0; k < g->nk; k++)
            for (int j = 0; j < g->nj; j++)
                for (int i = 0; i < g->ni; i++) {
                    long u = i + sj * j + sk * k;
                    R.mask[u] = 0;
                    R.jr[u] = 1;
                    F[u] = 0;
                    if (node_type(g, oid, u, i, j, k) != REG) continue;
                    double ad = -2.0 * (idx2 + idy2 + idz2);
                    double s = 0;
                    s += idz2 * phi[u - sk];
                    s += idy2 * phi[u - sj];
                    s += idx2 * phi[u - 1];
                    s += ad * phi[u];
                    s += idx2 * phi[u + 1];
                    s += idy2 * phi[u + sj];
                    s += idz2 * phi[u + sk];
                    double f = s - (-rho[u] / ORC_EPS_0);
                    f -= ORC_QE * n0 * exp((phi[u] - phi0) / Te0) / ORC_EPS_0;
                    double p = n0 * ORC_QE / (ORC_EPS_0 * Te0) * exp((phi[u] - phi0) / Te0);
                    double diag = ad - p;
                    unsigned char m = 64;
                    for (int d = 0; d < 6; d++) {
                        long v = u + off[d];
                        int tv = node_type(g, oid, v, i + di[d], j + dj[d], k + dk[d]);
                        if (tv == REG) m |= (unsigned char)(1u << d);
                        else if (tv == NEU) diag += c[d];
                    }
                    F[u] = f;
                    R.jr[u] = diag;
                    R.mask[u] = m;
                }
        int its = cg_reduced(&R, F, y, cg_tol, max_cg, work);
        newton++;
        cg_total += its < 0 ? -its : its;
        if (its < 0) break;
        double y2 = 0;
        for (int k = 0; k < g->nk; k++)
            for (int j = 0; j < g->nj; j++)
                for (int i = 0; i < g->ni; i++) {
                    long u = i + sj * j + sk * k;
                    int t = node_type(g, oid, u, i, j, k);
                    double yy = 0;
                    if (t == REG) yy = y[u];
                    else if (t == NEU) {
                        long v = chain_target(g, oid, u, i, j, k);
                        if (oid[v] == 0) yy = y[v];
                    }
                    phi[u] = phi[u] - yy;
                    y2 += yy * yy;
                }
        norm = sqrt(y2 / (double)nn);
        if (norm < nr_tol) {
            converged = 1;
            break;
        }
    }
    if (stats) {
        stats[0] = newton;
        stats[1] = cg_total;
        stats[2] = 0;
        stats[3] = norm;
    }
    free(R.mask);
    free(R.jr);
    free(F);
    free(y);
    free(work);
    return converged;
}

/* ---- the differentiable run ---------------------------------------------------------------- */
static double ramp(double dist, double inv) {
    double t = dist * inv;
    if (t <= 0.0) return 0.0;
    if (t >= 1.0) return 1.0;
    return t;
}

static double soft_sphere(const orc_dsph_config *c, double inv, const double X[3]) {
    double rx = X[0] - c->sphere_c[0], ry = X[1] - c->sphere_c[1], rz = X[2] - c->sphere_c[2];
    return ramp(sqrt(rx * rx + ry * ry + rz * rz) - c->sphere_r, inv);
}

static double soft_walls(const orc_grid *g, double inv, const double X[3]) {
    double m = 1.0;
    for (int a = 0; a < 3; a++) {
        double lo = a < 2 ? ramp(X[a] - g->x0[a], inv) : (X[a] >= g->x0[a] ? 1.0 : 0.0);
        double hi = ramp(g->xm[a] - X[a], inv);
        m *= lo * hi;
    }
    return m;
}

/* CIC weights of X clamped into the grid, as the GPU's cic_of */
static void cell_of(const orc_grid *g, const double X[3], double lc[3]) {
    int nc[3] = {g->ni - 1, g->nj - 1, g->nk - 1};
    for (int a = 0; a < 3; a++) {
        double l = (X[a] - g->x0[a]) / g->dh[a];
        int i = (int)l;
        if (i < 0) i = 0;
        if (i > nc[a] - 1) i = nc[a] - 1;
        lc[a] = (double)i + (l - i);  /* the same cell and offset as the GPU */
        if ((int)lc[a] != i) lc[a] = (double)i + 0.9999999999999999;
    }
}

static void finish(long nn, const double *raw, const double *nv, double q, double *rho, double *den) {
    for (long u = 0; u < nn; u++) {
        double d = nv[u] != 0.0 ? raw[u] / nv[u] : 0.0;
        rho[u] = 0.0 + d * q;
        if (den) den[u] = d;
    }
}

int orc_dsph_run(const orc_dsph_config *c, const double theta[3], double *current, double *den_out,
                 double *phi_out, double *stats) {
    orc_grid G;
    orc_grid_init(&G, c->ni, c->nj, c->nk, c->x0, c->xm);
    const orc_grid *g = &G;
    long nn = (long)c->ni * c->nj * c->nk;
    long cap = c->n_init + (long)c->steps * c->n_inject;
    double *pos = calloc((size_t)cap * 3, sizeof(double)), *vel = calloc((size_t)cap * 3, sizeof(double));
    double *w = calloc((size_t)cap, sizeof(double));
    double *nv = calloc((size_t)nn, sizeof(double)), *phi = calloc((size_t)nn, sizeof(double));
    double *rho = calloc((size_t)nn, sizeof(double)), *raw = calloc((size_t)nn, sizeof(double));
    double *ef = calloc((size_t)nn * 3, sizeof(double));
    int *oid = calloc((size_t)nn, sizeof(int));
    orc_node_volumes(g, nv);
    const double inv = 1.0 / (c->soft_width * g->dh[0]);
    const double s64 = c->dt * c->charge / c->mass;
    const double qm = c->charge / c->mass, half_dt = 0.5 * c->dt;
    const double vol = (c->xm[0] - c->x0[0]) * (c->xm[1] - c->x0[1]) * (c->xm[2] - c->x0[2]);
    const double w_init = c->n_init > 0 ? c->n0 * vol / (double)c->n_init : 0.0;
    const double A = (g->dh[0] * (c->ni - 1)) * (g->dh[1] * (c->nj - 1));
    const double w_inj = c->n_inject > 0 ? c->n0 * A * c->dt / (double)c->n_inject * theta[0] : 0.0;
    const double vth = c->T_ion > 0 ? sqrt(2 * ORC_K * c->T_ion / c->mass) : 0.0;
    int ok = 1, solves = 0, cg = 0;
    double st[4];
    /* initial fill: uniform over the domain, dead inside the sphere */
    for (long p = 0; p < c->n_init; p++) {
        double u[16], v[3];
        draws16(c->seed, p, 0, 2u, u);
        for (int a = 0; a < 3; a++) pos[3 * p + a] = g->x0[a] + u[a] * (g->xm[a] - g->x0[a]);
        double rx = pos[3 * p] - c->sphere_c[0], ry = pos[3 * p + 1] - c->sphere_c[1],
               rz = pos[3 * p + 2] - c->sphere_c[2];
        w[p] = rx * rx + ry * ry + rz * rz <= c->sphere_r * c->sphere_r ? 0.0 : w_init;
        thermal(vth, u, v);
        vel[3 * p] = v[0];
        vel[3 * p + 1] = v[1];
        vel[3 * p + 2] = theta[0] + v[2];
    }
    memset(raw, 0, sizeof(double) * (size_t)nn);
    for (long p = 0; p < c->n_init; p++) {
        if (w[p] <= 0.0) continue;
        double lc[3];
        cell_of(g, &pos[3 * p], lc);
        orc_scatter(g, raw, lc, w[p]);
    }
    finish(nn, raw, nv, c->charge, rho, NULL);
    /* the initial solve, continued in the sphere potential from a quarter of it */
    for (int q = 1; q <= 4 && ok; q++) {
        orc_add_sphere(g, oid, phi, c->sphere_c, c->sphere_r, theta[1] * (double)q / 4.0);
        orc_add_inlet(g, oid, phi);
        ok = orc_solve_nr_sym(g, phi, rho, oid, 0.0, theta[2], c->n0, c->cg_tol, c->nr_tol, 20000, st);
        solves++;
        cg += (int)st[1];
    }
    orc_compute_ef(g, phi, ef);
    double took = 0.0;
    long np = c->n_init;
    for (int n = 0; n < c->steps && ok; n++) {
        /* injection on the k = 0 face with the addParticle rewind */
        for (long q = 0; q < c->n_inject; q++) {
            long p = np + q;
            double u[16], v[3], lc[3];
            draws16(c->seed, p, n, 1u, u);
            pos[3 * p] = g->x0[0] + u[0] * (g->dh[0] * (c->ni - 1));
            pos[3 * p + 1] = g->x0[1] + u[1] * (g->dh[1] * (c->nj - 1));
            pos[3 * p + 2] = g->x0[2];
            thermal(vth, u, v);
            v[2] = theta[0] + v[2];
            cell_of(g, &pos[3 * p], lc);
            double E[3];
            orc_gather3(g, ef, lc, E);
            for (int a = 0; a < 3; a++) v[a] -= (qm * E[a]) * half_dt;
            memcpy(&vel[3 * p], v, sizeof v);
            w[p] = w_inj;
        }
        np += c->n_inject;
        /* push, soft absorption, deposit */
        memset(raw, 0, sizeof(double) * (size_t)nn);
        for (long p = 0; p < np; p++) {
            if (w[p] <= 0.0) continue;
            double lc[3], E[3];
            cell_of(g, &pos[3 * p], lc);
            orc_gather3(g, ef, lc, E);
            for (int a = 0; a < 3; a++) vel[3 * p + a] += E[a] * s64;
            for (int a = 0; a < 3; a++) pos[3 * p + a] += vel[3 * p + a] * c->dt;
            double ms = soft_sphere(c, inv, &pos[3 * p]);
            double mw = soft_walls(g, inv, &pos[3 * p]);
            took += w[p] * (1.0 - ms);
            w[p] = w[p] * ms * mw;
            if (w[p] > 0.0) {
                cell_of(g, &pos[3 * p], lc);
                orc_scatter(g, raw, lc, w[p]);
            } else {
                w[p] = 0.0;
            }
        }
        finish(nn, raw, nv, c->charge, rho, n == c->steps - 1 ? den_out : NULL);
        ok = orc_solve_nr_sym(g, phi, rho, oid, 0.0, theta[2], c->n0, c->cg_tol, c->nr_tol, 20000, st);
        solves++;
        cg += (int)st[1];
        orc_compute_ef(g, phi, ef);
    }
    if (current) *current = c->charge * took / ((double)c->steps * c->dt);
    if (phi_out) memcpy(phi_out, phi, sizeof(double) * (size_t)nn);
    if (stats) {
        double live = 0;
        for (long p = 0; p < np; p++) live += w[p];
        stats[0] = solves;
        stats[1] = cg;
        stats[2] = live;
        stats[3] = took;
    }
    free(pos);
    free(vel);
    free(w);
    free(nv);
    free(phi);
    free(rho);
    free(raw);
    free(ef);
    free(oid);
    return ok;
}
