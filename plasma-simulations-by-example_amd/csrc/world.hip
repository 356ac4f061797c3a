// World: node fields on the device, geometry, computeEF, solveQN, charge density, timers.
#include <cmath>
#include <cstring>
#include <mutex>

#include "picamd_internal.h"

// The FP64 node kernels here restate the reference's expressions operation for operation;
// contraction into FMA would change their rounding, so it is off for this file.
#pragma clang fp contract(off)

namespace pic {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    g_last_error = buf;
    return PIC_ERR_HIP;
}

int grow_partials(pic_world *w, int nblocks) {
    if (nblocks <= w->partials_cap) return PIC_OK;
    if (w->partials) PIC_HIP(hipFree(w->partials));
    int cap = nblocks < 4096 ? 4096 : nblocks;
    PIC_HIP(hipMalloc(&w->partials, sizeof(double) * 8 * (size_t)cap));
    w->partials_cap = cap;
    return PIC_OK;
}

void timer_begin(pic_world *w, int phase, hipEvent_t *a) {
    (void)phase;
    *a = nullptr;
    if (!w->timing) return;
    if (w->free_events.empty()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        w->free_events.push_back(e);
    }
    *a = w->free_events.back();
    w->free_events.pop_back();
    hipEventRecord(*a, w->stream);
}

void timer_end(pic_world *w, int phase, hipEvent_t a) {
    if (!w->timing || !a) return;
    hipEvent_t b;
    if (w->free_events.empty()) {
        if (hipEventCreate(&b) != hipSuccess) return;
    } else {
        b = w->free_events.back();
        w->free_events.pop_back();
    }
    hipEventRecord(b, w->stream);
    w->pending.push_back(TimerRec{phase, a, b});
}

static void resolve_timers(pic_world *w) {
    if (w->pending.empty()) return;
    hipStreamSynchronize(w->stream);
    for (auto &r : w->pending) {
        float ms = 0;
        hipEventElapsedTime(&ms, r.a, r.b);
        w->ms[r.phase] += ms;
        w->calls[r.phase] += 1;
        w->free_events.push_back(r.a);
        w->free_events.push_back(r.b);
    }
    w->pending.clear();
}

// ---- kernels ---------------------------------------------------------------------------

__global__ void k_node_volumes(Grid g, double *nv) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= g.nn) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    // World::computeNodeVolumes (ch2/World.cpp:58-69)
    double V = g.dh[0] * g.dh[1] * g.dh[2];
    if (i == 0 || i == g.ni - 1) V *= 0.5;
    if (j == 0 || j == g.nj - 1) V *= 0.5;
    if (k == 0 || k == g.nk - 1) V *= 0.5;
    nv[u] = V;
}

__global__ void k_add_sphere(Grid g, uint8_t *oid, double *phi, double cx, double cy, double cz,
                             double r2, double phi_s) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= g.nn) return;
    int i, j, k;
    node_ijk(g, u, i, j, k);
    // World::pos(i,j,k) = x0 + dh*(i,j,k); World::inSphere (ch3/ver2/World.cpp:88-126)
    double x = g.x0[0] + g.dh[0] * (double)i;
    double y = g.x0[1] + g.dh[1] * (double)j;
    double z = g.x0[2] + g.dh[2] * (double)k;
    double r0 = x - cx, r1 = y - cy, rz = z - cz;
    double m2 = (r0 * r0 + r1 * r1 + rz * rz);
    if (m2 <= r2) {
        oid[u] = 1;
        phi[u] = phi_s;
    }
}

__global__ void k_add_inlet(Grid g, uint8_t *oid, double *phi) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= (long)g.ni * g.nj) return;
    oid[u] = 2;
    phi[u] = 0;
}

// computeEF (ch3/ver2/PotentialSolver.cpp:465-504): central differences inside, second-order
// one-sided on the walls.  Writes the FP32 copy used by PIC_MIXED pushes and/or the FP64
// copy, and accumulates getPE's sum |E|^2 node_vol in FP64 (World::getPE,
// ch3/ver2/World.cpp:73-85) so that diagnostics cost no extra pass.
template <bool W32, bool W64>
__global__ void __launch_bounds__(256) k_compute_ef(Grid g, const double *__restrict__ phi,
                                                    float4 *ef32, double4 *ef64, double *pe_acc) {
    __shared__ double sh[8];
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double pe = 0;
    if (u < g.nn) {
        int ni = g.ni, nj = g.nj, nk = g.nk;
        int i, j, k;
        node_ijk(g, u, i, j, k);
        long si = 1, sj = ni, sk = (long)ni * nj;
        double dx = g.dh[0], dy = g.dh[1], dz = g.dh[2];
        double ex, ey, ez;
        if (i == 0)
            ex = -(-3 * phi[u] + 4 * phi[u + si] - phi[u + 2 * si]) / (2 * dx);
        else if (i == ni - 1)
            ex = -(phi[u - 2 * si] - 4 * phi[u - si] + 3 * phi[u]) / (2 * dx);
        else
            ex = -(phi[u + si] - phi[u - si]) / (2 * dx);
        if (j == 0)
            ey = -(-3 * phi[u] + 4 * phi[u + sj] - phi[u + 2 * sj]) / (2 * dy);
        else if (j == nj - 1)
            ey = -(phi[u - 2 * sj] - 4 * phi[u - sj] + 3 * phi[u]) / (2 * dy);
        else
            ey = -(phi[u + sj] - phi[u - sj]) / (2 * dy);
        if (k == 0)
            ez = -(-3 * phi[u] + 4 * phi[u + sk] - phi[u + 2 * sk]) / (2 * dz);
        else if (k == nk - 1)
            ez = -(phi[u - 2 * sk] - 4 * phi[u - sk] + 3 * phi[u]) / (2 * dz);
        else
            ez = -(phi[u + sk] - phi[u - sk]) / (2 * dz);
        if (W32) ef32[u] = make_float4((float)ex, (float)ey, (float)ez, 0.f);
        if (W64) ef64[u] = make_double4(ex, ey, ez, 0.0);
        double V = g.dh[0] * g.dh[1] * g.dh[2];
        if (i == 0 || i == ni - 1) V *= 0.5;
        if (j == 0 || j == nj - 1) V *= 0.5;
        if (k == 0 || k == nk - 1) V *= 0.5;
        pe = (ex * ex + ey * ey + ez * ez) * V;
    }
    double s = block_sum(pe, sh);
    if (threadIdx.x == 0) pe_acc[blockIdx.x] = s;  // per-block partial, summed by get_pe
}

// World::getPE over a stored E (after a host write of ef)
__global__ void __launch_bounds__(256) k_pe_from_ef(Grid g, const double4 *ef64, double *pe_acc) {
    __shared__ double sh[8];
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double pe = 0;
    if (u < g.nn) {
        int i, j, k;
        node_ijk(g, u, i, j, k);
        double V = g.dh[0] * g.dh[1] * g.dh[2];
        if (i == 0 || i == g.ni - 1) V *= 0.5;
        if (j == 0 || j == g.nj - 1) V *= 0.5;
        if (k == 0 || k == g.nk - 1) V *= 0.5;
        double4 e = ef64[u];
        pe = (e.x * e.x + e.y * e.y + e.z * e.z) * V;
    }
    double s = block_sum(pe, sh);
    if (threadIdx.x == 0) pe_acc[blockIdx.x] = s;
}

__global__ void __launch_bounds__(1024) k_sum_partials(const double *part, long nb, double *out) {
    __shared__ double sh[16];
    double s = 0;
    for (long b = threadIdx.x; b < nb; b += blockDim.x) s += part[b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *out = s;
}

// solveQN (ch3/ver2/PotentialSolver.cpp:204-222)
__global__ void k_solve_qn(Grid g, double *phi, const double *rho, const uint8_t *oid, double phi0,
                           double Te0, double rho0) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= g.nn) return;
    if (oid[u] > 0) return;
    double rho_ratio = rho[u] / rho0;
    if (rho_ratio < 1e-6) rho_ratio = 1e-6;
    phi[u] = phi0 + Te0 * log(rho_ratio);
}

// World::computeChargeDensity (ch3/ver2/World.cpp:47-55): rho = 0; rho += den_s*q_s
struct DenList {
    const double *den[16];
    double q[16];
    int n;
};

__global__ void k_charge_density(long nn, DenList L, double *rho) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= nn) return;
    double r = 0;
    for (int s = 0; s < L.n; s++) r += L.den[s][u] * L.q[s];
    rho[u] = r;
}

__global__ void k_unpack_ef(long nn, const double4 *ef64, double *out) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= nn) return;
    double4 e = ef64[u];
    out[3 * u + 0] = e.x;
    out[3 * u + 1] = e.y;
    out[3 * u + 2] = e.z;
}

__global__ void k_pack_ef(long nn, const double *in, double4 *ef64, float4 *ef32) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= nn) return;
    double ex = in[3 * u], ey = in[3 * u + 1], ez = in[3 * u + 2];
    ef64[u] = make_double4(ex, ey, ez, 0.0);
    ef32[u] = make_float4((float)ex, (float)ey, (float)ez, 0.f);
}

__global__ void k_oid_to_double(long nn, const uint8_t *oid, double *out) {
    long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (u >= nn) return;
    out[u] = (double)oid[u];
}

inline unsigned nblocks(long n, int t) { return (unsigned)((n + t - 1) / t); }

int compute_ef(pic_world *w, bool want32, bool want64) {
    const Grid &g = w->g;
    PIC_CHECK(g.ni >= 3 && g.nj >= 3 && g.nk >= 3, PIC_ERR_ARG,
              "computeEF needs at least 3 nodes per axis");
    hipEvent_t ev;
    timer_begin(w, PIC_T_EF, &ev);
    unsigned nb = nbk8(g.nn);  // one PE partial per workgroup (nbk8: the partials' count)
    if (want32 && want64)
        k_compute_ef<true, true><<<nb, 256, 0, w->stream>>>(g, w->phi, w->ef32, w->ef64, w->pe_part);
    else if (want32)
        k_compute_ef<true, false><<<nb, 256, 0, w->stream>>>(g, w->phi, w->ef32, w->ef64, w->pe_part);
    else
        k_compute_ef<false, true><<<nb, 256, 0, w->stream>>>(g, w->phi, w->ef32, w->ef64, w->pe_part);
    PIC_LAUNCH_CHECK();
    timer_end(w, PIC_T_EF, ev);
    w->pe_valid = true;
    return PIC_OK;
}

int solve_qn(pic_world *w, double phi0, double Te0, double n0) {
    double rho0 = n0 * QE;
    k_solve_qn<<<nblocks(w->g.nn, 256), 256, 0, w->stream>>>(w->g, w->phi, w->rho, w->object_id,
                                                             phi0, Te0, rho0);
    PIC_LAUNCH_CHECK();
    w->ef_valid = false;
    return PIC_OK;
}

}  // namespace pic

using namespace pic;

// =========================================================================================
// C ABI
// =========================================================================================
extern "C" {

int pic_abi_version(void) { return PICAMD_ABI_VERSION; }

const char *pic_last_error(void) { return g_last_error.c_str(); }

int pic_device_count(int *n) {
    PIC_CHECK(n, PIC_ERR_ARG, "null argument");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return PIC_OK;
}

int pic_world_create(int ni, int nj, int nk, pic_world **out) {
    PIC_CHECK(out, PIC_ERR_ARG, "null out");
    *out = nullptr;
    PIC_CHECK(ni >= 2 && nj >= 2 && nk >= 2, PIC_ERR_ARG, "need at least 2 nodes per axis");
    PIC_CHECK(ni - 1 <= MAX_CELLS_I && nj - 1 <= MAX_CELLS_J && nk - 1 <= MAX_CELLS_K, PIC_ERR_ARG,
              "grid exceeds the packed cell index range (1024x1024x2048 cells)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device visible");
        return PIC_ERR_NODEVICE;
    }
    pic_world *w = new pic_world();
    hipGetDevice(&w->device);
    Grid &g = w->g;
    g.ni = ni;
    g.nj = nj;
    g.nk = nk;
    g.nci = ni - 1;
    g.ncj = nj - 1;
    g.nck = nk - 1;
    g.ntx = (g.nci + TILE - 1) / TILE;
    g.nty = (g.ncj + TILE - 1) / TILE;
    g.ntz = (g.nck + TILE - 1) / TILE;
    g.ntiles = g.ntx * g.nty * g.ntz;
    g.nn = (long)ni * nj * nk;
    int rc = PIC_OK;
#define TRY(x)                          \
    do {                                \
        hipError_t e_ = (x);            \
        if (e_ != hipSuccess) {         \
            rc = hip_fail(e_, #x, __FILE__, __LINE__); \
            goto fail;                  \
        }                               \
    } while (0)
    TRY(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    TRY(hipMalloc(&w->phi, sizeof(double) * g.nn));
    TRY(hipMalloc(&w->rho, sizeof(double) * g.nn));
    TRY(hipMalloc(&w->node_vol, sizeof(double) * g.nn));
    TRY(hipMalloc(&w->ef32, sizeof(float4) * g.nn));
    TRY(hipMalloc(&w->ef64, sizeof(double4) * g.nn));
    TRY(hipMalloc(&w->object_id, g.nn));
    TRY(hipMalloc(&w->pe_part, sizeof(double) * nbk8(g.nn)));
    TRY(hipMemsetAsync(w->pe_part, 0, sizeof(double) * nbk8(g.nn), w->stream));
    TRY(hipMalloc(&w->scal, sizeof(DevScalars)));
    TRY(hipHostMalloc(&w->scal_host, sizeof(DevScalars)));
    TRY(hipMemsetAsync(w->phi, 0, sizeof(double) * g.nn, w->stream));
    TRY(hipMemsetAsync(w->rho, 0, sizeof(double) * g.nn, w->stream));
    TRY(hipMemsetAsync(w->node_vol, 0, sizeof(double) * g.nn, w->stream));
    TRY(hipMemsetAsync(w->ef32, 0, sizeof(float4) * g.nn, w->stream));
    TRY(hipMemsetAsync(w->ef64, 0, sizeof(double4) * g.nn, w->stream));
    TRY(hipMemsetAsync(w->object_id, 0, g.nn, w->stream));
    TRY(hipMemsetAsync(w->scal, 0, sizeof(DevScalars), w->stream));
    TRY(hipStreamSynchronize(w->stream));
#undef TRY
    w->ef_valid = true;  // phi = 0 and ef = 0 are consistent
    *out = w;
    return PIC_OK;
fail:
    pic_world_destroy(w);
    return rc;
}

int pic_world_destroy(pic_world *w) {
    if (!w) return PIC_OK;
    if (w->stream) hipStreamSynchronize(w->stream);
    for (auto &r : w->pending) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    for (auto e : w->free_events) hipEventDestroy(e);
    hipFree(w->phi);
    hipFree(w->rho);
    hipFree(w->node_vol);
    hipFree(w->ef32);
    hipFree(w->ef64);
    hipFree(w->object_id);
    hipFree(w->scal);
    hipFree(w->partials);
    hipFree(w->pe_part);
    if (w->scal_host) hipHostFree(w->scal_host);
    if (w->stream) hipStreamDestroy(w->stream);
    delete w;
    return PIC_OK;
}

int pic_world_set_extents(pic_world *w, const double x0[3], const double xm[3]) {
    PIC_CHECK(w && x0 && xm, PIC_ERR_ARG, "null argument");
    Grid &g = w->g;
    int nn[3] = {g.ni, g.nj, g.nk};
    for (int c = 0; c < 3; c++) {
        PIC_CHECK(xm[c] > x0[c], PIC_ERR_ARG, "xm must exceed x0 on every axis");
        g.x0[c] = x0[c];
        g.xm[c] = xm[c];
        g.dh[c] = (xm[c] - x0[c]) / (nn[c] - 1);  // ch2/World.cpp:30-31
    }
    w->extents_set = true;
    k_node_volumes<<<nblocks(g.nn, 256), 256, 0, w->stream>>>(g, w->node_vol);
    PIC_LAUNCH_CHECK();
    return PIC_OK;
}

int pic_world_add_sphere(pic_world *w, const double c[3], double r, double phi_s) {
    PIC_CHECK(w && c, PIC_ERR_ARG, "null argument");
    PIC_CHECK(w->extents_set, PIC_ERR_STATE, "setExtents must precede addSphere");
    // the same sphere again (a new potential only) leaves object_id as it was
    const bool same = w->sphere.on && w->sphere.c[0] == c[0] && w->sphere.c[1] == c[1] &&
                      w->sphere.c[2] == c[2] && w->sphere.r2 == r * r;
    w->sphere.on = 1;
    for (int q = 0; q < 3; q++) w->sphere.c[q] = c[q];
    w->sphere.r2 = r * r;
    k_add_sphere<<<nblocks(w->g.nn, 256), 256, 0, w->stream>>>(w->g, w->object_id, w->phi, c[0], c[1],
                                                               c[2], r * r, phi_s);
    PIC_LAUNCH_CHECK();
    w->ef_valid = false;
    if (!same) w->geom_stamp++;
    return PIC_OK;
}

int pic_world_add_inlet(pic_world *w) {
    PIC_CHECK(w, PIC_ERR_ARG, "null argument");
    long n = (long)w->g.ni * w->g.nj;
    k_add_inlet<<<nblocks(n, 256), 256, 0, w->stream>>>(w->g, w->object_id, w->phi);
    PIC_LAUNCH_CHECK();
    w->ef_valid = false;
    w->geom_stamp++;
    return PIC_OK;
}

int pic_world_get_field(pic_world *w, int which, double *out) {
    PIC_CHECK(w && out, PIC_ERR_ARG, "null argument");
    long nn = w->g.nn;
    double *tmp = nullptr;
    switch (which) {
    case PIC_FIELD_PHI:
        PIC_HIP(hipMemcpyAsync(out, w->phi, sizeof(double) * nn, hipMemcpyDeviceToHost, w->stream));
        break;
    case PIC_FIELD_RHO:
        PIC_HIP(hipMemcpyAsync(out, w->rho, sizeof(double) * nn, hipMemcpyDeviceToHost, w->stream));
        break;
    case PIC_FIELD_NODE_VOL:
        PIC_HIP(hipMemcpyAsync(out, w->node_vol, sizeof(double) * nn, hipMemcpyDeviceToHost, w->stream));
        break;
    case PIC_FIELD_EF:
        PIC_HIP(hipMalloc(&tmp, sizeof(double) * 3 * nn));
        k_unpack_ef<<<nblocks(nn, 256), 256, 0, w->stream>>>(nn, w->ef64, tmp);
        PIC_HIP(hipMemcpyAsync(out, tmp, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, w->stream));
        break;
    case PIC_FIELD_OBJECT_ID:
        PIC_HIP(hipMalloc(&tmp, sizeof(double) * nn));
        k_oid_to_double<<<nblocks(nn, 256), 256, 0, w->stream>>>(nn, w->object_id, tmp);
        PIC_HIP(hipMemcpyAsync(out, tmp, sizeof(double) * nn, hipMemcpyDeviceToHost, w->stream));
        break;
    default:
        set_error("unknown field");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(w->stream));
    if (tmp) PIC_HIP(hipFree(tmp));
    return PIC_OK;
}

int pic_world_set_field(pic_world *w, int which, const double *in) {
    PIC_CHECK(w && in, PIC_ERR_ARG, "null argument");
    long nn = w->g.nn;
    double *tmp = nullptr;
    switch (which) {
    case PIC_FIELD_PHI:
        PIC_HIP(hipMemcpyAsync(w->phi, in, sizeof(double) * nn, hipMemcpyHostToDevice, w->stream));
        w->ef_valid = false;
        break;
    case PIC_FIELD_RHO:
        PIC_HIP(hipMemcpyAsync(w->rho, in, sizeof(double) * nn, hipMemcpyHostToDevice, w->stream));
        break;
    case PIC_FIELD_EF:
        PIC_HIP(hipMalloc(&tmp, sizeof(double) * 3 * nn));
        PIC_HIP(hipMemcpyAsync(tmp, in, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, w->stream));
        k_pack_ef<<<nblocks(nn, 256), 256, 0, w->stream>>>(nn, tmp, w->ef64, w->ef32);
        w->ef_valid = true;
        w->pe_valid = false;
        break;
    default:
        set_error("field is not writable");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(w->stream));
    if (tmp) PIC_HIP(hipFree(tmp));
    return PIC_OK;
}

int pic_world_compute_charge_density(pic_world *w, pic_species *const *sp, int nsp) {
    PIC_CHECK(w && (sp || nsp == 0), PIC_ERR_ARG, "null argument");
    DenList L{};
    for (int s = 0; s < nsp; s++) {
        PIC_CHECK(sp[s] && sp[s]->w == w, PIC_ERR_ARG, "species belongs to another world");
        if (sp[s]->charge == 0) continue;  // neutrals skipped, ch3/ver2/World.cpp:52
        PIC_CHECK(L.n < 16, PIC_ERR_ARG, "at most 16 charged species");
        PIC_CHECK(sp[s]->den_state == 2, PIC_ERR_STATE,
                  "computeNumberDensity must run before computeChargeDensity");
        L.den[L.n] = sp[s]->den;
        L.q[L.n] = sp[s]->charge;
        L.n++;
    }
    hipEvent_t ev;
    timer_begin(w, PIC_T_RHO, &ev);
    k_charge_density<<<nblocks(w->g.nn, 256), 256, 0, w->stream>>>(w->g.nn, L, w->rho);
    PIC_LAUNCH_CHECK();
    timer_end(w, PIC_T_RHO, ev);
    return PIC_OK;
}

int pic_world_get_pe(pic_world *w, double *pe) {
    PIC_CHECK(w && pe, PIC_ERR_ARG, "null argument");
    if (!w->pe_valid) {
        // E was written from the host: sum |E|^2 V over the stored field
        k_pe_from_ef<<<nbk8(w->g.nn), 256, 0, w->stream>>>(w->g, w->ef64, w->pe_part);
        PIC_LAUNCH_CHECK();
        w->pe_valid = true;
    }
    k_sum_partials<<<1, 1024, 0, w->stream>>>(w->pe_part, (long)nbk8(w->g.nn), &w->scal->pe);
    PIC_LAUNCH_CHECK();
    PIC_HIP(hipMemcpyAsync(&w->scal_host->pe, &w->scal->pe, sizeof(double), hipMemcpyDeviceToHost,
                           w->stream));
    PIC_HIP(hipStreamSynchronize(w->stream));
    *pe = 0.5 * EPS_0 * w->scal_host->pe;
    return PIC_OK;
}

int pic_world_sync(pic_world *w) {
    PIC_CHECK(w, PIC_ERR_ARG, "null argument");
    PIC_HIP(hipStreamSynchronize(w->stream));
    return PIC_OK;
}

int pic_world_stream(pic_world *w, void **stream) {
    PIC_CHECK(w && stream, PIC_ERR_ARG, "null argument");
    *stream = (void *)w->stream;
    return PIC_OK;
}

int pic_world_set_timing(pic_world *w, int enable) {
    PIC_CHECK(w, PIC_ERR_ARG, "null argument");
    w->timing = enable != 0;
    return PIC_OK;
}

int pic_world_get_timers(pic_world *w, double *ms, long *calls, int reset) {
    PIC_CHECK(w, PIC_ERR_ARG, "null argument");
    resolve_timers(w);
    for (int p = 0; p < PIC_T_COUNT; p++) {
        if (ms) ms[p] = w->ms[p];
        if (calls) calls[p] = w->calls[p];
        if (reset) {
            w->ms[p] = 0;
            w->calls[p] = 0;
        }
    }
    return PIC_OK;
}

}  // extern "C"
