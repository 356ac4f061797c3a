/*
 * pic_oracle.c -- FP64 CPU restatement of the reference ES-PIC inner loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or shipped).
 * Each function names the reference function it restates; loop order and floating-point
 * operation order follow the reference so results are bit-exact against oracle/_ref.
 * Node arrays: flat, U = k*ni*nj + j*ni + i (ch3/ver2/Field.h:161).
 */
#include "pic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define U3(g, i, j, k) ((long)(k) * (g)->ni * (g)->nj + (long)(j) * (g)->ni + (i))

void orc_grid_init(orc_grid *g, int ni, int nj, int nk, const double *x0, const double *xm) {
    /* World::setExtents, ch2/World.cpp:21-37: dh = (xm-x0)/(nn-1), xc = 0.5*(x0+xm) */
    int nn[3] = {ni, nj, nk};
    g->ni = ni;
    g->nj = nj;
    g->nk = nk;
    for (int c = 0; c < 3; c++) {
        g->x0[c] = x0[c];
        g->xm[c] = xm[c];
        g->dh[c] = (xm[c] - x0[c]) / (nn[c] - 1);
        g->xc[c] = 0.5 * (x0[c] + xm[c]);
    }
}

void orc_node_volumes(const orc_grid *g, double *nv) {
    /* ch2/World.cpp:58-69 */
    for (int i = 0; i < g->ni; i++)
        for (int j = 0; j < g->nj; j++)
            for (int k = 0; k < g->nk; k++) {
                double V = g->dh[0] * g->dh[1] * g->dh[2];
                if (i == 0 || i == g->ni - 1) V *= 0.5;
                if (j == 0 || j == g->nj - 1) V *= 0.5;
                if (k == 0 || k == g->nk - 1) V *= 0.5;
                nv[U3(g, i, j, k)] = V;
            }
}

void orc_xtol(const orc_grid *g, const double *x, double *lc) {
    /* ch2/World.h:72-78 */
    for (int c = 0; c < 3; c++) lc[c] = (x[c] - g->x0[c]) / g->dh[c];
}

int orc_in_bounds(const orc_grid *g, const double *x) {
    /* ch2/World.h:59-63: half-open [x0, xm) per axis */
    for (int c = 0; c < 3; c++)
        if (x[c] < g->x0[c] || x[c] >= g->xm[c]) return 0;
    return 1;
}

int orc_in_sphere(const double *x, const double *c, double r2) {
    /* ch3/ver2/World.cpp:119-126 */
    double r0 = x[0] - c[0], r1 = x[1] - c[1], r2_ = x[2] - c[2];
    double m2 = (r0 * r0 + r1 * r1 + r2_ * r2_);
    return m2 <= r2;
}

/* CIC corner order of ch2/Field.h:172-182 / :198-205 */
static const int CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                 {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

void orc_gather3(const orc_grid *g, const double *ef3, const double *lc, double *out) {
    /* ch2/Field.h:186-208 with T = double3: per component
     * val = sum_c  E_c * w0 * w1 * w2, evaluated left to right */
    int i = (int)lc[0];
    double di = lc[0] - i;
    int j = (int)lc[1];
    double dj = lc[1] - j;
    int k = (int)lc[2];
    double dk = lc[2] - k;
    for (int c = 0; c < 3; c++) {
        double val = 0;
        for (int n = 0; n < 8; n++) {
            int a = CORNER[n][0], b = CORNER[n][1], e = CORNER[n][2];
            double wi = a ? di : (1 - di);
            double wj = b ? dj : (1 - dj);
            double wk = e ? dk : (1 - dk);
            double term = ef3[3 * U3(g, i + a, j + b, k + e) + c] * wi * wj * wk;
            val = (n == 0) ? term : val + term;
        }
        out[c] = val;
    }
}

double orc_gather1(const orc_grid *g, const double *f, const double *lc) {
    int i = (int)lc[0];
    double di = lc[0] - i;
    int j = (int)lc[1];
    double dj = lc[1] - j;
    int k = (int)lc[2];
    double dk = lc[2] - k;
    double val = 0;
    for (int n = 0; n < 8; n++) {
        int a = CORNER[n][0], b = CORNER[n][1], e = CORNER[n][2];
        double wi = a ? di : (1 - di);
        double wj = b ? dj : (1 - dj);
        double wk = e ? dk : (1 - dk);
        double term = f[U3(g, i + a, j + b, k + e)] * wi * wj * wk;
        val = (n == 0) ? term : val + term;
    }
    return val;
}

void orc_scatter(const orc_grid *g, double *f, const double *lc, double v) {
    /* ch2/Field.h:164-183: f[node] += value*w_i*w_j*w_k, same corner order */
    int i = (int)lc[0];
    double di = lc[0] - i;
    int j = (int)lc[1];
    double dj = lc[1] - j;
    int k = (int)lc[2];
    double dk = lc[2] - k;
    for (int n = 0; n < 8; n++) {
        int a = CORNER[n][0], b = CORNER[n][1], e = CORNER[n][2];
        double wi = a ? di : (1 - di);
        double wj = b ? dj : (1 - dj);
        double wk = e ? dk : (1 - dk);
        f[U3(g, i + a, j + b, k + e)] += v * wi * wj * wk;
    }
}

static void push_one(const orc_grid *g, double *x, double *v, const double *ef3, double dt,
                     double charge, double mass) {
    /* ch2/Species.cpp:19-29: lc = XtoL(x); E = gather; v += E*(dt*q/m); x += v*dt */
    double lc[3], E[3];
    orc_xtol(g, x, lc);
    orc_gather3(g, ef3, lc, E);
    double s = dt * charge / mass;
    for (int c = 0; c < 3; c++) v[c] += E[c] * s;
    for (int c = 0; c < 3; c++) x[c] += v[c] * dt;
}

void orc_push_reflect(const orc_grid *g, long np, double *pos, double *vel, const double *ef3,
                      double dt, double charge, double mass) {
    for (long p = 0; p < np; p++) {
        double *x = pos + 3 * p, *v = vel + 3 * p;
        push_one(g, x, v, ef3, dt, charge, mass);
        /* ch2/Species.cpp:32-36: specular reflection on each axis */
        for (int c = 0; c < 3; c++) {
            if (x[c] < g->x0[c]) {
                x[c] = 2 * g->x0[c] - x[c];
                v[c] *= -1.0;
            } else if (x[c] >= g->xm[c]) {
                x[c] = 2 * g->xm[c] - x[c];
                v[c] *= -1.0;
            }
        }
    }
}

long orc_push_absorb(const orc_grid *g, long np, double *pos, double *vel, double *mpw, long *ids,
                     const double *ef3, double dt, double charge, double mass,
                     const double *sphere_c, double sphere_r2) {
    /* ch3/ver2/Species.cpp:13-33 */
    for (long p = 0; p < np; p++) {
        double *x = pos + 3 * p, *v = vel + 3 * p;
        push_one(g, x, v, ef3, dt, charge, mass);
        if (orc_in_sphere(x, sphere_c, sphere_r2) || !orc_in_bounds(g, x)) mpw[p] = 0;
    }
    /* ch3/ver2/Species.cpp:36-46: dead particles replaced by the last live entry */
    long n = np;
    for (long p = 0; p < n; p++) {
        if (mpw[p] > 0) continue;
        memcpy(pos + 3 * p, pos + 3 * (n - 1), 3 * sizeof(double));
        memcpy(vel + 3 * p, vel + 3 * (n - 1), 3 * sizeof(double));
        mpw[p] = mpw[n - 1];
        if (ids) ids[p] = ids[n - 1];
        n--;
        p--;
    }
    return n;
}

void orc_number_density(const orc_grid *g, long np, const double *pos, const double *mpw,
                        const double *nv, double *den) {
    /* ch2/Species.cpp:41-52 and Field.h:125-134 (/= writes 0 where the divisor is 0) */
    long nn = (long)g->ni * g->nj * g->nk;
    memset(den, 0, nn * sizeof(double));
    for (long p = 0; p < np; p++) {
        double lc[3];
        orc_xtol(g, pos + 3 * p, lc);
        orc_scatter(g, den, lc, mpw[p]);
    }
    for (long u = 0; u < nn; u++) {
        if (nv[u] != 0)
            den[u] /= nv[u];
        else
            den[u] = 0;
    }
}

void orc_charge_density(long nn, int nsp, const double *charges, const double *const *dens,
                        double *rho) {
    /* ch2/World.cpp:46-54: rho = 0; rho += (den*q) for charged species */
    for (long u = 0; u < nn; u++) rho[u] = 0;
    for (int s = 0; s < nsp; s++) {
        if (charges[s] == 0) continue;
        for (long u = 0; u < nn; u++) rho[u] += dens[s][u] * charges[s];
    }
}

int orc_add_particle(const orc_grid *g, const double *ef3, double dt, double charge, double mass,
                     const double *x, const double *v, double w, long *np, double *pos,
                     double *vel, double *mpw) {
    /* ch2/Species.cpp:55-71: reject out of bounds; v -= (q/m)*E*(0.5*dt) */
    if (!orc_in_bounds(g, x)) return 0;
    double lc[3], E[3];
    orc_xtol(g, x, lc);
    orc_gather3(g, ef3, lc, E);
    long p = *np;
    for (int c = 0; c < 3; c++) {
        pos[3 * p + c] = x[c];
        vel[3 * p + c] = v[c] - (E[c] * (charge / mass)) * (0.5 * dt);
    }
    mpw[p] = w;
    *np = p + 1;
    return 1;
}

void orc_load_box(const orc_grid *g, orc_mt19937 *rng, const double *x1, const double *x2,
                  double num_den, int num_mp, const double *ef3, double dt, double charge,
                  double mass, long *np, double *pos, double *vel, double *mpw) {
    /* ch2/Species.cpp:74-97 */
    double box_vol = (x2[0] - x1[0]) * (x2[1] - x1[1]) * (x2[2] - x1[2]);
    double num_real = num_den * box_vol;
    double w = num_real / num_mp;
    for (int p = 0; p < num_mp; p++) {
        double x[3], v[3] = {0, 0, 0};
        x[0] = x1[0] + orc_mt_uniform(rng) * (x2[0] - x1[0]);
        x[1] = x1[1] + orc_mt_uniform(rng) * (x2[1] - x1[1]);
        x[2] = x1[2] + orc_mt_uniform(rng) * (x2[2] - x1[2]);
        orc_add_particle(g, ef3, dt, charge, mass, x, v, w, np, pos, vel, mpw);
    }
}

void orc_load_box_qs(const orc_grid *g, const double *x1, const double *x2, double num_den,
                     const int *num_mp, const double *ef3, double dt, double charge, double mass,
                     long *np, double *pos, double *vel, double *mpw) {
    /* ch2/Species.cpp:101-141 */
    double box_vol = (x2[0] - x1[0]) * (x2[1] - x1[1]) * (x2[2] - x1[2]);
    int num_mp_tot = (num_mp[0] - 1) * (num_mp[1] - 1) * (num_mp[2] - 1);
    double num_real = num_den * box_vol;
    double w0 = num_real / num_mp_tot;
    double di = (x2[0] - x1[0]) / (num_mp[0] - 1);
    double dj = (x2[1] - x1[1]) / (num_mp[1] - 1);
    double dk = (x2[2] - x1[2]) / (num_mp[2] - 1);
    for (int i = 0; i < num_mp[0]; i++)
        for (int j = 0; j < num_mp[1]; j++)
            for (int k = 0; k < num_mp[2]; k++) {
                double x[3], v[3] = {0, 0, 0};
                x[0] = x1[0] + i * di;
                x[1] = x1[1] + j * dj;
                x[2] = x1[2] + k * dk;
                if (x[0] == x2[0]) x[0] -= 1e-4 * di;
                if (x[1] == x2[1]) x[1] -= 1e-4 * dj;
                if (x[2] == x2[2]) x[2] -= 1e-4 * dk;
                double w = 1;
                if (i == 0 || i == num_mp[0] - 1) w *= 0.5;
                if (j == 0 || j == num_mp[1] - 1) w *= 0.5;
                if (k == 0 || k == num_mp[2] - 1) w *= 0.5;
                orc_add_particle(g, ef3, dt, charge, mass, x, v, w0 * w, np, pos, vel, mpw);
            }
}

int orc_cold_beam_sample(const orc_grid *g, orc_mt19937 *rng, double v_drift, double den,
                         double mpw0, const double *ef3, double dt, double charge, double mass,
                         long cap, long *np, double *pos, double *vel, double *mpw) {
    /* ch3/ver2/Source.cpp:4-27 */
    double Lx = g->dh[0] * (g->ni - 1);
    double Ly = g->dh[1] * (g->nj - 1);
    double A = Lx * Ly;
    double num_real = den * v_drift * A * dt;
    int num_sim = (int)(num_real / mpw0 + orc_mt_uniform(rng));
    int added = 0;
    for (int i = 0; i < num_sim; i++) {
        double x[3], v[3] = {0, 0, v_drift};
        x[0] = g->x0[0] + orc_mt_uniform(rng) * Lx;
        x[1] = g->x0[1] + orc_mt_uniform(rng) * Ly;
        x[2] = g->x0[2];
        if (*np >= cap) return -1;
        added += orc_add_particle(g, ef3, dt, charge, mass, x, v, mpw0, np, pos, vel, mpw);
    }
    return added;
}

static double sample_vth(orc_mt19937 *rng, double T, double mass) {
    /* ch4/Species.cpp:149-158: speed from three sums of three uniforms */
    double v_th = sqrt(2 * ORC_K * T / mass);
    double a, b, c;
    a = orc_mt_uniform(rng); b = orc_mt_uniform(rng); c = orc_mt_uniform(rng);
    double v1 = v_th * (a + b + c - 1.5);
    a = orc_mt_uniform(rng); b = orc_mt_uniform(rng); c = orc_mt_uniform(rng);
    double v2 = v_th * (a + b + c - 1.5);
    a = orc_mt_uniform(rng); b = orc_mt_uniform(rng); c = orc_mt_uniform(rng);
    double v3 = v_th * (a + b + c - 1.5);
    return 3 / sqrt(2 + 2 + 2) * sqrt(v1 * v1 + v2 * v2 + v3 * v3);
}

int orc_warm_beam_sample(const orc_grid *g, orc_mt19937 *rng, double v_drift, double den,
                         double T, double mpw0, const double *ef3, double dt, double charge,
                         double mass, long cap, long *np, double *pos, double *vel, double *mpw) {
    /* ch4/Source.cpp:31-56 and ch4/Species.cpp:161-174 */
    double Lx = g->dh[0] * (g->ni - 1);
    double Ly = g->dh[1] * (g->nj - 1);
    double A = Lx * Ly;
    double num_real = den * v_drift * A * dt;
    int num_sim = (int)(num_real / mpw0 + orc_mt_uniform(rng));
    int added = 0;
    for (int i = 0; i < num_sim; i++) {
        double x[3], v[3];
        x[0] = g->x0[0] + orc_mt_uniform(rng) * Lx;
        x[1] = g->x0[1] + orc_mt_uniform(rng) * Ly;
        x[2] = g->x0[2];
        double theta = 2 * ORC_PI * orc_mt_uniform(rng);
        double r = -1.0 + 2 * orc_mt_uniform(rng);
        double a = sqrt(1 - r * r);
        double d[3] = {r, cos(theta) * a, sin(theta) * a};
        double vth = sample_vth(rng, T, mass);
        for (int c = 0; c < 3; c++) v[c] = d[c] * vth;
        v[2] += v_drift;
        if (*np >= cap) return -1;
        added += orc_add_particle(g, ef3, dt, charge, mass, x, v, mpw0, np, pos, vel, mpw);
    }
    return added;
}

int orc_solve_sor_linear(const orc_grid *g, double *phi, const double *rho, int max_it, double tol,
                         double *l2_out, int *its) {
    /* ch2/PotentialSolver.cpp:11-67: lexicographic i,j,k (k fastest) GS with w=1.4 on
     * interior nodes; boundary nodes are never written.  Residual every 25 iterations. */
    double idx2 = 1.0 / (g->dh[0] * g->dh[0]);
    double idy2 = 1.0 / (g->dh[1] * g->dh[1]);
    double idz2 = 1.0 / (g->dh[2] * g->dh[2]);
    long si = 1, sj = g->ni, sk = (long)g->ni * g->nj;
    double L2 = 0;
    int converged = 0;
    unsigned it;
    for (it = 0; it < (unsigned)max_it; it++) {
        for (int i = 1; i < g->ni - 1; i++)
            for (int j = 1; j < g->nj - 1; j++)
                for (int k = 1; k < g->nk - 1; k++) {
                    long u = U3(g, i, j, k);
                    double phi_new = (rho[u] / ORC_EPS_0 + idx2 * (phi[u - si] + phi[u + si]) +
                                      idy2 * (phi[u - sj] + phi[u + sj]) +
                                      idz2 * (phi[u - sk] + phi[u + sk])) /
                                     (2 * idx2 + 2 * idy2 + 2 * idz2);
                    phi[u] = phi[u] + 1.4 * (phi_new - phi[u]);
                }
        if (it % 25 == 0) {
            double sum = 0;
            for (int i = 1; i < g->ni - 1; i++)
                for (int j = 1; j < g->nj - 1; j++)
                    for (int k = 1; k < g->nk - 1; k++) {
                        long u = U3(g, i, j, k);
                        double R = -phi[u] * (2 * idx2 + 2 * idy2 + 2 * idz2) + rho[u] / ORC_EPS_0 +
                                   idx2 * (phi[u - si] + phi[u + si]) +
                                   idy2 * (phi[u - sj] + phi[u + sj]) +
                                   idz2 * (phi[u - sk] + phi[u + sk]);
                        sum += R * R;
                    }
            L2 = sqrt(sum / (g->ni * g->nj * g->nk));
            if (L2 < tol) {
                converged = 1;
                break;
            }
        }
    }
    if (l2_out) *l2_out = L2;
    if (its) *its = (int)(converged ? it + 1 : it);
    return converged;
}

int orc_solve_gs_boltzmann(const orc_grid *g, double *phi, const double *rho, const int *object_id,
                           double phi0, double Te0, double n0, int max_it, double tol, double *l2_out,
                           int *its) {
    /* ch3/ver2/PotentialSolver.cpp:334-430 */
    double idx2 = 1.0 / (g->dh[0] * g->dh[0]);
    double idy2 = 1.0 / (g->dh[1] * g->dh[1]);
    double idz2 = 1.0 / (g->dh[2] * g->dh[2]);
    int ni = g->ni, nj = g->nj, nk = g->nk;
    long si = 1, sj = ni, sk = (long)ni * nj;
    double L2 = 0;
    int converged = 0;
    unsigned it;
    for (it = 0; it < (unsigned)max_it; it++) {
        for (int i = 0; i < ni; i++)
            for (int j = 0; j < nj; j++)
                for (int k = 0; k < nk; k++) {
                    long u = U3(g, i, j, k);
                    if (object_id[u] > 0) continue;
                    if (i == 0)
                        phi[u] = phi[u + si];
                    else if (i == ni - 1)
                        phi[u] = phi[u - si];
                    else if (j == 0)
                        phi[u] = phi[u + sj];
                    else if (j == nj - 1)
                        phi[u] = phi[u - sj];
                    else if (k == 0)
                        phi[u] = phi[u + sk];
                    else if (k == nk - 1)
                        phi[u] = phi[u - sk];
                    else {
                        double ne = n0 * exp((phi[u] - phi0) / Te0);
                        double phi_new = ((rho[u] - ORC_QE * ne) / ORC_EPS_0 +
                                          idx2 * (phi[u - si] + phi[u + si]) +
                                          idy2 * (phi[u - sj] + phi[u + sj]) +
                                          idz2 * (phi[u - sk] + phi[u + sk])) /
                                         (2 * idx2 + 2 * idy2 + 2 * idz2);
                        phi[u] = phi[u] + 1.4 * (phi_new - phi[u]);
                    }
                }
        if (it % 25 == 0) {
            double sum = 0;
            for (int i = 0; i < ni; i++)
                for (int j = 0; j < nj; j++)
                    for (int k = 0; k < nk; k++) {
                        long u = U3(g, i, j, k);
                        if (object_id[u] > 0) continue;
                        double R = 0;
                        if (i == 0)
                            R = phi[u] - phi[u + si];
                        else if (i == ni - 1)
                            R = phi[u] - phi[u - si];
                        else if (j == 0)
                            R = phi[u] - phi[u + sj];
                        else if (j == nj - 1)
                            R = phi[u] - phi[u - sj];
                        else if (k == 0)
                            R = phi[u] - phi[u + sk];
                        else if (k == nk - 1)
                            R = phi[u] - phi[u - sk];
                        else {
                            double ne = n0 * exp((phi[u] - phi0) / Te0);
                            R = -phi[u] * (2 * idx2 + 2 * idy2 + 2 * idz2) +
                                (rho[u] - ORC_QE * ne) / ORC_EPS_0 +
                                idx2 * (phi[u - si] + phi[u + si]) +
                                idy2 * (phi[u - sj] + phi[u + sj]) +
                                idz2 * (phi[u - sk] + phi[u + sk]);
                        }
                        sum += R * R;
                    }
            L2 = sqrt(sum / (ni * nj * nk));
            if (L2 < tol) {
                converged = 1;
                break;
            }
        }
    }
    if (l2_out) *l2_out = L2;
    if (its) *its = (int)(converged ? it + 1 : it);
    return converged;
}

void orc_solve_qn(const orc_grid *g, double *phi, const double *rho, const int *object_id,
                  double phi0, double Te0, double n0) {
    /* ch3/ver2/PotentialSolver.cpp:204-222 */
    double rho0 = n0 * ORC_QE;
    double rho_ratio_min = 1e-6;
    long nn = (long)g->ni * g->nj * g->nk;
    for (long u = 0; u < nn; u++) {
        if (object_id[u] > 0) continue;
        double rho_ratio = rho[u] / rho0;
        if (rho_ratio < rho_ratio_min) rho_ratio = rho_ratio_min;
        phi[u] = phi0 + Te0 * log(rho_ratio);
    }
}

/* ---- matrix-free restatement of the 7-slot Matrix of ch3/ver2/PotentialSolver.cpp ---- */
enum { NT_REG = 0, NT_NEUMANN = 1, NT_DIRICHLET = 2 };

typedef struct {
    int nslot;
    long col[7];
    double a[7]; /* coefficient of A (before diagSubtract) */
    int diag;    /* slot of the diagonal */
} orc_row;

/* buildMatrix (ch3/ver2/PotentialSolver.cpp:146-201): slot insertion order is kept so that
 * Matrix::operator* / multRow sums in the same order. */
static int build_row(const orc_grid *g, const int *object_id, int i, int j, int k, orc_row *r) {
    double idx = 1.0 / g->dh[0], idy = 1.0 / g->dh[1], idz = 1.0 / g->dh[2];
    double idx2 = idx * idx, idy2 = idy * idy, idz2 = idz * idz;
    int ni = g->ni, nj = g->nj, nk = g->nk;
    long u = U3(g, i, j, k);
    long sj = ni, sk = (long)ni * nj;
    if (object_id[u] > 0) {
        r->nslot = 1; r->col[0] = u; r->a[0] = 1; r->diag = 0;
        return NT_DIRICHLET;
    }
    r->diag = 0;
    r->nslot = 2;
    r->col[0] = u;
    if (i == 0) { r->a[0] = idx; r->col[1] = u + 1; r->a[1] = -idx; }
    else if (i == ni - 1) { r->a[0] = idx; r->col[1] = u - 1; r->a[1] = -idx; }
    else if (j == 0) { r->a[0] = idy; r->col[1] = u + sj; r->a[1] = -idy; }
    else if (j == nj - 1) { r->a[0] = idy; r->col[1] = u - sj; r->a[1] = -idy; }
    else if (k == 0) { r->a[0] = idz; r->col[1] = u + sk; r->a[1] = -idz; }
    else if (k == nk - 1) { r->a[0] = idz; r->col[1] = u - sk; r->a[1] = -idz; }
    else {
        r->nslot = 7;
        r->col[0] = u - sk; r->a[0] = idz2;
        r->col[1] = u - sj; r->a[1] = idy2;
        r->col[2] = u - 1;  r->a[2] = idx2;
        r->col[3] = u;      r->a[3] = -2.0 * (idx2 + idy2 + idz2);
        r->col[4] = u + 1;  r->a[4] = idx2;
        r->col[5] = u + sj; r->a[5] = idy2;
        r->col[6] = u + sk; r->a[6] = idz2;
        r->diag = 3;
        return NT_REG;
    }
    return NT_NEUMANN;
}

typedef struct {
    long nu;
    orc_row *rows;
    int *type;
} orc_mat;

static void mat_build(const orc_grid *g, const int *object_id, orc_mat *A) {
    A->nu = (long)g->ni * g->nj * g->nk;
    A->rows = (orc_row *)malloc(A->nu * sizeof(orc_row));
    A->type = (int *)malloc(A->nu * sizeof(int));
    for (int k = 0; k < g->nk; k++)
        for (int j = 0; j < g->nj; j++)
            for (int i = 0; i < g->ni; i++) {
                long u = U3(g, i, j, k);
                A->type[u] = build_row(g, object_id, i, j, k, &A->rows[u]);
            }
}

static void mat_free(orc_mat *A) {
    free(A->rows);
    free(A->type);
}

/* Matrix::operator* (:24-35) with the diagonal replaced by diag[u] (diagSubtract, :60-64) */
static void mat_mul(const orc_mat *A, const double *diag, const double *v, double *r) {
    for (long u = 0; u < A->nu; u++) {
        const orc_row *row = &A->rows[u];
        double s = 0;
        for (int q = 0; q < row->nslot; q++) {
            double a = (q == row->diag) ? diag[u] : row->a[q];
            s += a * v[row->col[q]];
        }
        r[u] = s;
    }
}

static double vdot(long n, const double *a, const double *b) {
    /* vec::dot (:104-111) */
    double d = 0;
    for (long j = 0; j < n; j++) d += a[j] * b[j];
    return d;
}

static double vnorm(long n, const double *v) {
    /* vec::norm (:114-121): RMS, sqrt(sum/n) with n an int */
    double s = 0;
    int nu = (int)n;
    for (int j = 0; j < nu; j++) s += v[j] * v[j];
    return sqrt(s / nu);
}

/* solvePCGLinear (:299-331) on J = A - diag(P) (diag passed in) */
static int pcg_linear(const orc_mat *A, const double *diag, double *x, const double *b, int max_it,
                      double tol, int *its) {
    long nu = A->nu;
    double *g = (double *)malloc(nu * sizeof(double));
    double *s = (double *)malloc(nu * sizeof(double));
    double *d = (double *)malloc(nu * sizeof(double));
    double *z = (double *)malloc(nu * sizeof(double));
    double *M = (double *)malloc(nu * sizeof(double));
    int converged = 0;
    double l2 = 0;
    for (long u = 0; u < nu; u++) M[u] = 1.0 / diag[u];
    mat_mul(A, diag, x, g);
    for (long u = 0; u < nu; u++) g[u] = g[u] - b[u];
    for (long u = 0; u < nu; u++) s[u] = M[u] * g[u];
    for (long u = 0; u < nu; u++) d[u] = -1 * s[u];
    unsigned it;
    for (it = 0; it < (unsigned)max_it; it++) {
        mat_mul(A, diag, d, z);
        double alpha = vdot(nu, g, s);
        double beta = vdot(nu, d, z);
        for (long u = 0; u < nu; u++) x[u] = x[u] + (alpha / beta) * d[u];
        for (long u = 0; u < nu; u++) g[u] = g[u] + (alpha / beta) * z[u];
        for (long u = 0; u < nu; u++) s[u] = M[u] * g[u];
        beta = alpha;
        alpha = vdot(nu, g, s);
        for (long u = 0; u < nu; u++) d[u] = (alpha / beta) * d[u] - s[u];
        l2 = vnorm(nu, g);
        if (l2 < tol) {
            converged = 1;
            break;
        }
    }
    if (its) *its = (int)(converged ? it + 1 : it);
    free(g); free(s); free(d); free(z); free(M);
    return converged;
}

/* solveGSLinear (:433-461), w = 1.0 */
static int gs_linear(const orc_mat *A, const double *diag, double *x, const double *b, int max_it,
                     double tol) {
    long nu = A->nu;
    double *R = (double *)malloc(nu * sizeof(double));
    int converged = 0;
    for (unsigned it = 0; it < (unsigned)max_it; it++) {
        for (long u = 0; u < nu; u++) {
            const orc_row *row = &A->rows[u];
            double sum = 0;
            for (int q = 0; q < row->nslot; q++) {
                double a = (q == row->diag) ? diag[u] : row->a[q];
                sum += a * x[row->col[q]];
            }
            double S = sum - diag[u] * x[u];
            double phi_new = (b[u] - S) / diag[u];
            x[u] = x[u] + 1. * (phi_new - x[u]);
        }
        if (it % 25 == 0) {
            mat_mul(A, diag, x, R);
            for (long u = 0; u < nu; u++) R[u] = R[u] - b[u];
            double L2 = vnorm(nu, R);
            if (L2 < tol) {
                converged = 1;
                break;
            }
        }
    }
    free(R);
    return converged;
}

int orc_solve_nrpcg(const orc_grid *g, double *phi, const double *rho, const int *object_id,
                    double phi0, double Te0, double n0, int max_it, double tol, double *stats) {
    /* solveNRPCG (:225-296) */
    const int NR_MAX_IT = 20;
    const double NR_TOL = 1e-3;
    orc_mat A;
    mat_build(g, object_id, &A);
    long nu = A.nu;
    double *P = (double *)calloc(nu, sizeof(double));
    double *y = (double *)calloc(nu, sizeof(double));
    double *x = (double *)malloc(nu * sizeof(double));
    double *b = (double *)malloc(nu * sizeof(double));
    double *F = (double *)malloc(nu * sizeof(double));
    double *Adiag = (double *)malloc(nu * sizeof(double));
    double *Jdiag = (double *)malloc(nu * sizeof(double));
    for (long u = 0; u < nu; u++) {
        x[u] = phi[u];
        b[u] = rho[u];
        Adiag[u] = A.rows[u].a[A.rows[u].diag];
    }
    for (long u = 0; u < nu; u++) {
        if (A.type[u] == NT_NEUMANN) b[u] = 0;
        else if (A.type[u] == NT_DIRICHLET) b[u] = x[u];
        else b[u] = -b[u] / ORC_EPS_0;
    }
    double norm = 0;
    int converged = 0, nr_it = 0, pcg_total = 0, gs_fallbacks = 0;
    for (int it = 0; it < NR_MAX_IT; it++) {
        nr_it = it + 1;
        mat_mul(&A, Adiag, x, F);
        for (long n = 0; n < nu; n++) F[n] = F[n] - b[n];
        for (long n = 0; n < nu; n++)
            if (A.type[n] == NT_REG) F[n] -= ORC_QE * n0 * exp((x[n] - phi0) / Te0) / ORC_EPS_0;
        for (long n = 0; n < nu; n++)
            if (A.type[n] == NT_REG) P[n] = n0 * ORC_QE / (ORC_EPS_0 * Te0) * exp((x[n] - phi0) / Te0);
        for (long u = 0; u < nu; u++) Jdiag[u] = Adiag[u] - P[u];
        int pits = 0;
        if (!pcg_linear(&A, Jdiag, y, F, max_it, tol, &pits)) {
            gs_fallbacks++;
            gs_linear(&A, Jdiag, y, F, max_it, tol);
        }
        pcg_total += pits;
        for (long u = 0; u < nu; u++)
            if (A.type[u] == NT_DIRICHLET) y[u] = 0;
        for (long u = 0; u < nu; u++) x[u] = x[u] - y[u];
        norm = vnorm(nu, y);
        if (norm < NR_TOL) {
            converged = 1;
            break;
        }
    }
    for (long u = 0; u < nu; u++) phi[u] = x[u];
    if (stats) {
        stats[0] = nr_it;
        stats[1] = pcg_total;
        stats[2] = gs_fallbacks;
        stats[3] = norm;
    }
    free(P); free(y); free(x); free(b); free(F); free(Adiag); free(Jdiag);
    mat_free(&A);
    return converged;
}

void orc_compute_ef(const orc_grid *g, const double *phi, double *ef3) {
    /* ch3/ver2/PotentialSolver.cpp:465-504: central in the interior, second-order one-sided
     * on the walls */
    int ni = g->ni, nj = g->nj, nk = g->nk;
    double dx = g->dh[0], dy = g->dh[1], dz = g->dh[2];
    long si = 1, sj = ni, sk = (long)ni * nj;
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++)
            for (int k = 0; k < nk; k++) {
                long u = U3(g, i, j, k);
                double *e = ef3 + 3 * u;
                if (i == 0)
                    e[0] = -(-3 * phi[u] + 4 * phi[u + si] - phi[u + 2 * si]) / (2 * dx);
                else if (i == ni - 1)
                    e[0] = -(phi[u - 2 * si] - 4 * phi[u - si] + 3 * phi[u]) / (2 * dx);
                else
                    e[0] = -(phi[u + si] - phi[u - si]) / (2 * dx);
                if (j == 0)
                    e[1] = -(-3 * phi[u] + 4 * phi[u + sj] - phi[u + 2 * sj]) / (2 * dy);
                else if (j == nj - 1)
                    e[1] = -(phi[u - 2 * sj] - 4 * phi[u - sj] + 3 * phi[u]) / (2 * dy);
                else
                    e[1] = -(phi[u + sj] - phi[u - sj]) / (2 * dy);
                if (k == 0)
                    e[2] = -(-3 * phi[u] + 4 * phi[u + sk] - phi[u + 2 * sk]) / (2 * dz);
                else if (k == nk - 1)
                    e[2] = -(phi[u - 2 * sk] - 4 * phi[u - sk] + 3 * phi[u]) / (2 * dz);
                else
                    e[2] = -(phi[u + sk] - phi[u - sk]) / (2 * dz);
            }
}

void orc_add_sphere(const orc_grid *g, int *object_id, double *phi, const double *c, double r,
                    double phi_s) {
    /* ch3/ver2/World.cpp:88-106: x = x0 + dh*(i,j,k); mark nodes with |x-c|^2 <= r^2 */
    double r2 = r * r;
    for (int i = 0; i < g->ni; i++)
        for (int j = 0; j < g->nj; j++)
            for (int k = 0; k < g->nk; k++) {
                double x[3] = {g->x0[0] + g->dh[0] * (double)i, g->x0[1] + g->dh[1] * (double)j,
                               g->x0[2] + g->dh[2] * (double)k};
                if (orc_in_sphere(x, c, r2)) {
                    long u = U3(g, i, j, k);
                    object_id[u] = 1;
                    phi[u] = phi_s;
                }
            }
}

void orc_add_inlet(const orc_grid *g, int *object_id, double *phi) {
    /* ch3/ver2/World.cpp:109-116 */
    for (int i = 0; i < g->ni; i++)
        for (int j = 0; j < g->nj; j++) {
            long u = U3(g, i, j, 0);
            object_id[u] = 2;
            phi[u] = 0;
        }
}

double orc_get_pe(const orc_grid *g, const double *ef3, const double *nv) {
    /* ch2/World.cpp:72-84, loop order i,j,k */
    double pe = 0;
    for (int i = 0; i < g->ni; i++)
        for (int j = 0; j < g->nj; j++)
            for (int k = 0; k < g->nk; k++) {
                long u = U3(g, i, j, k);
                const double *e = ef3 + 3 * u;
                double ef2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
                pe += ef2 * nv[u];
            }
    return 0.5 * ORC_EPS_0 * pe;
}

void orc_species_diag(long np, const double *vel, const double *mpw, double mass, double *out) {
    /* ch2/Species.cpp:144-168 */
    double cnt = 0, m0 = 0, m1 = 0, m2 = 0, ke = 0;
    for (long p = 0; p < np; p++) cnt += mpw[p];
    for (long p = 0; p < np; p++) {
        m0 += vel[3 * p + 0] * mpw[p];
        m1 += vel[3 * p + 1] * mpw[p];
        m2 += vel[3 * p + 2] * mpw[p];
    }
    for (long p = 0; p < np; p++) {
        const double *v = vel + 3 * p;
        double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        ke += mpw[p] * v2;
    }
    out[0] = cnt;
    out[1] = m0 * mass;
    out[2] = m1 * mass;
    out[3] = m2 * mass;
    out[4] = 0.5 * mass * ke;
}

void orc_update_average(long nn, double *ave, const double *inst, int *samples) {
    /* ch3/ver2/Field.h:214-220 */
    int s = *samples;
    for (long u = 0; u < nn; u++) ave[u] = (inst[u] + s * ave[u]) / (s + 1);
    *samples = s + 1;
}

/* ---- std::mt19937 (Matsumoto & Nishimura 1998, the parameters fixed by the C++ standard)
 *      and libstdc++'s generate_canonical<double,53> used by uniform_real_distribution ---- */
void orc_mt_seed(orc_mt19937 *r, uint32_t seed) {
    r->mt[0] = seed;
    for (int i = 1; i < 624; i++)
        r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
    r->idx = 624;
}

uint32_t orc_mt_next(orc_mt19937 *r) {
    if (r->idx >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (r->mt[i] & 0x80000000u) | (r->mt[(i + 1) % 624] & 0x7fffffffu);
            r->mt[i] = r->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        r->idx = 0;
    }
    uint32_t y = r->mt[r->idx++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

double orc_mt_uniform(orc_mt19937 *r) {
    /* two 32-bit draws: (x1 + x2*2^32) / 2^64, clamped below 1 */
    double sum = (double)orc_mt_next(r);
    double hi = (double)orc_mt_next(r);
    sum += hi * 4294967296.0;
    double ret = sum / 18446744073709551616.0;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}
