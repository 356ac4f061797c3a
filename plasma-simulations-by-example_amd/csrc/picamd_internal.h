// Internal definitions of libpicamd (gfx950).  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "picamd.h"

namespace pic {

// ---- physical constants (ch2/World.h:11-20) -------------------------------------------
constexpr double EPS_0 = 8.85418782e-12;
constexpr double QE = 1.602176565e-19;
constexpr double KB = 1.380648e-23;
constexpr double PI_REF = 3.141592653;  // the reference's Const::PI

// ---- tiling ---------------------------------------------------------------------------
// Particles live in per-tile bins; a tile is TILE^3 cells.  The push of one tile is one
// workgroup: its E tile (TILE+1)^3 nodes sits in LDS, and deposits of the pushed
// particles land in an LDS window of the tile plus HALO cells on every side.  A particle
// that ends more than HALO cells outside its tile deposits with global atomics instead; in
// the sphere flow ions accelerated by the sheath outrun a one-cell rim (1.3% of them after
// seven steps, and growing), a two-cell rim keeps nearly all deposits in LDS.
constexpr int TILE_SHIFT = 3;
constexpr int TILE = 1 << TILE_SHIFT;
constexpr int TN = TILE + 1;              // E-tile nodes per edge
constexpr int TNODES = TN * TN * TN;      // 729
constexpr int HALO = 2;
// Tiles whose particles left the two-cell rim often in the last push (ions the sphere
// accelerates move several cells a step) are pushed by a second instantiation with a
// HALO_WIDE-cell rim; the part of its window beyond HALO cells goes out by global atomics
// once per node.  Three cells measured faster than four late in the sphere run (17.9 against
// 18.7 ms per push at steps 81-100; five: 19.6): 4 workgroups per CU against 3.
constexpr int HALO_WIDE = 3;
constexpr int WN = TILE + 1 + 2 * HALO;   // window nodes per edge
constexpr int WNODES = WN * WN * WN;      // 2197
constexpr int PUSH_THREADS = 256;
constexpr int NDEST = 27;                 // destination tiles reachable in one step (3x3x3)
constexpr int STAGE_WGS = 64;             // workgroups that walk the unbinned list (sort, diag)
// The push gives each tile one workgroup per PUSH_CHUNK of its particles (the default; the
// tuning table may set another per grid and particle count): flows that pile
// ions into a few tiles (the focused wake behind the sphere reaches 40x the mean tile
// population within 100 steps) would otherwise leave one workgroup running long after the
// rest.  PUSH_STAGE_WGS workgroups push the unbinned list.
constexpr int PUSH_CHUNK = 32768;
constexpr int PUSH_STAGE_WGS = 256;
// The staged histogram of the push plan walks the staging list, which holds a step's
// injection (1.75e6 particles per step in the sphere configurations): a wide grid keeps its
// per-thread chain of loads short, and PLAN_STAGED_SHARDS partial counts per tile spread the
// atomics of particles that share a tile.
constexpr int PLAN_STAGED_WGS = 1024;
constexpr int PLAN_STAGED_SHARDS = 16;
// Push workgroups are handed to the 8 XCDs in stripes of PUSH_XCD_STRIPE (default; tunable)
// consecutive
// (tile, chunk) entries, round-robin: each XCD's L2 serves runs of neighbouring tiles, and
// every XCD gets a share of each region of the domain.  (Contiguous eighths per XCD left
// the XCDs that own the wake running at twice the others' time.)
constexpr int PUSH_XCD_STRIPE = 32;
static_assert(PUSH_STAGE_WGS % 8 == 0, "binned push workgroups keep their XCD (blockIdx % 8)");

// packed cell index of PIC_MIXED particles: i | j<<10 | k<<20
constexpr int CELL_BITS_I = 10, CELL_BITS_J = 10, CELL_BITS_K = 11;
constexpr int MAX_CELLS_I = 1 << CELL_BITS_I, MAX_CELLS_J = 1 << CELL_BITS_J,
              MAX_CELLS_K = 1 << CELL_BITS_K;

__host__ __device__ inline int pack_cell(int i, int j, int k) { return i | (j << 10) | (k << 20); }
__host__ __device__ inline int cell_i(int c) { return c & 1023; }
__host__ __device__ inline int cell_j(int c) { return (c >> 10) & 1023; }
__host__ __device__ inline int cell_k(int c) { return (c >> 20) & 2047; }

struct Grid {
    int ni, nj, nk;     // nodes
    int nci, ncj, nck;  // cells
    int ntx, nty, ntz;  // tiles
    int ntiles;
    long nn;
    double x0[3], xm[3], dh[3];
};

struct Sphere {
    int on;
    double c[3];
    double r2;
};

// ---- particle storage ---------------------------------------------------------------
// One buffer: [0, cap_bins) holds tile bins, [cap_bins, cap_bins+cap_stage) the unbinned
// staging list (host uploads, injected particles, bin overflow).
struct PartBuf {
    void *mem = nullptr;
    long cap_bins = 0, cap_stage = 0;
    // PIC_MIXED: 32-byte records {int cell (bit pattern), ox, oy, oz | vx, vy, vz, w}
    float4 *rec = nullptr;
    // PIC_FP64
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *dvx = nullptr, *dvy = nullptr, *dvz = nullptr, *dw = nullptr;
    uint32_t *id = nullptr;
    // tile bins.  Bin t holds two segments: [start, start + own) and
    // [start + off, start + off + min(in, cap - off)).  A push writes the particles that stay
    // in tile t to the first (sized to t's input count, so it never overflows) and appends
    // arrivals from other tiles to the second.
    uint32_t *bin_start = nullptr, *bin_cap = nullptr, *bin_off = nullptr;
    uint32_t *bin_own = nullptr, *bin_in = nullptr;
    uint32_t *stage_count = nullptr;  // device counter of the staging list
    // one byte per slot: 1 for a particle whose deposit the push left to k_far_arrivals (set
    // by the push, cleared by k_far_arrivals right after it: 0 everywhere else)
    uint8_t *mark = nullptr;
};

// Device view of a buffer, passed by value to kernels.
struct PartView {
    float4 *rec;
    double *x, *y, *z, *dvx, *dvy, *dvz, *dw;
    uint32_t *id;
    uint32_t *bin_start, *bin_cap, *bin_off, *bin_own, *bin_in, *stage_count;
    long cap_bins, cap_stage;
    uint8_t *mark;
};

inline PartView view_of(const PartBuf &b) {
    return PartView{b.rec, b.x, b.y, b.z,
                    b.dvx, b.dvy, b.dvz, b.dw, b.id, b.bin_start, b.bin_cap, b.bin_off,
                    b.bin_own, b.bin_in, b.stage_count, b.cap_bins, b.cap_stage, b.mark};
}

// the two segments of tile t: q in [0, c0+c1) lives at q < c0 ? s0+q : s1+(q-c0)
struct BinSeg {
    uint32_t s0, c0, s1, c1;
};

__device__ inline BinSeg bin_of(const PartView &v, int t) {
    uint32_t cap = v.bin_cap[t], off = v.bin_off[t], st = v.bin_start[t];
    uint32_t c0 = v.bin_own[t];
    c0 = c0 < off ? c0 : off;
    uint32_t room = cap > off ? cap - off : 0u;
    uint32_t c1 = v.bin_in[t];
    c1 = c1 < room ? c1 : room;
    return BinSeg{st, c0, st + off, c1};
}

__device__ inline long seg_index(const BinSeg &b, uint32_t q) {
    return q < b.c0 ? (long)b.s0 + q : (long)b.s1 + (q - b.c0);
}

// ---- device scalars of a world ---------------------------------------------------------
struct DevScalars {
    double pe;            // getPE accumulator
    double red[8];        // general reductions
    unsigned int done;    // solver convergence flag
    unsigned int counter; // last-block counter
    int iters;
    int pad;
};

// per-species device status word (see pic_species::st_dev)
struct SpStatus {
    uint32_t err;       // capacity flags raised since the last snapshot: 1 bins short, 2 particle dropped
    uint32_t err_snap;  // flags taken (and cleared) by the last snapshot
    uint32_t stage;     // staging count after the last push
    uint32_t pad;
    long long np;       // live particles after the last push
};

struct TimerRec {
    int phase;
    hipEvent_t a, b;
};

// (i, j, k) of flat node u = k*ni*nj + j*ni + i; 32-bit unsigned arithmetic (grids hold
// fewer than 2^32 nodes), which divides in a few instructions where 64-bit division takes
// dozens
__device__ inline void node_ijk(const Grid &g, long u, int &i, int &j, int &k) {
    unsigned uu = (unsigned)u, ni = (unsigned)g.ni, nj = (unsigned)g.nj;
    unsigned r = uu / ni;
    i = (int)(uu - r * ni);
    unsigned kk = r / nj;
    j = (int)(r - kk * nj);
    k = (int)kk;
}

// Workgroup index for node-stencil kernels.  Workgroups are dispatched round-robin over the
// 8 XCDs, each with its own L2; with this remap XCD x runs workgroups [x G/8, (x+1) G/8),
// a contiguous slab of k planes, so the +-j and +-k neighbours a workgroup reads were
// fetched into the same L2 by the workgroups next to it.  Launch with nbk8() workgroups.
__device__ inline long xcd_block() {
    unsigned G = gridDim.x, b = blockIdx.x;
    return (G % 8) ? (long)b : (long)((b % 8) * (G / 8) + b / 8);
}

// workgroups of 256 threads over n nodes, and that count padded to a multiple of 8 (for
// xcd_block)
inline unsigned nbk(long n) { return (unsigned)((n + 255) / 256); }
inline unsigned nbk8(long n) { return (nbk(n) + 7) / 8 * 8; }

}  // namespace pic

struct pic_world {
    int device = 0;
    hipStream_t stream = nullptr;
    pic::Grid g{};
    bool extents_set = false;
    double *phi = nullptr, *rho = nullptr, *node_vol = nullptr;
    float4 *ef32 = nullptr;
    double4 *ef64 = nullptr;
    uint8_t *object_id = nullptr;
    pic::Sphere sphere{};
    pic::DevScalars *scal = nullptr;  // device
    pic::DevScalars *scal_host = nullptr;  // pinned mirror
    double *partials = nullptr;       // per-block reduction partials
    int partials_cap = 0;
    double *pe_part = nullptr;        // per-block |E|^2 V partials of the last computeEF
    bool ef_valid = false;            // ef32/ef64 hold computeEF of the current phi
    unsigned geom_stamp = 0;          // bumped whenever object_id changes (sphere, inlet)
    bool pe_valid = false;
    // timers
    bool timing = false;
    std::vector<pic::TimerRec> pending;
    std::vector<hipEvent_t> free_events;
    double ms[PIC_T_COUNT] = {0};
    long calls[PIC_T_COUNT] = {0};
};

struct pic_species {
    pic_world *w = nullptr;
    double mass = 0, charge = 0, mpw0 = 0;
    int prec = PIC_MIXED;
    bool track_ids = false;
    pic::PartBuf buf[2];
    int cur = 0;                  // buffer holding the current particles
    long np_upper = 0;            // host-side upper bound of the particle count
    uint32_t next_id = 0;
    // Status of the last push, written in stream order after it (k_push_status) and copied
    // to pinned memory behind st_ev: the live count, the staging count and the capacity
    // flags that push raised.  The host reads it one call later without a stream sync.
    pic::SpStatus *st_dev = nullptr;
    pic::SpStatus *st_host = nullptr;  // pinned mirror
    hipEvent_t st_ev = nullptr;
    bool st_pending = false;      // a snapshot is in flight and not yet read
    long added_since_push = 0;    // particles appended since the snapshotted push
    bool stage_known = true;      // stage_host is the exact staging count
    long stage_host = 0;
    bool np_exact = true;         // np_upper is the exact count (no host edit since a count)
    bool extra_dirty = false;     // den_extra holds deposits that no density finish consumed
    double *den = nullptr, *den_extra = nullptr, *den_ave = nullptr;
    int ave_samples = 0;
    void *win = nullptr;          // per-tile deposit windows (float or double)
    uint8_t *win_used = nullptr;
    uint32_t *far_tile = nullptr;  // per tile: far deposits of the last push, then [ntiles..) the
                                   // plan's wide-rim flags, then [2 ntiles..) the push's
                                   // far-arrival flags (cleared by k_far_arrivals)
    void *plan_mem = nullptr;     // scratch of the segment plan (per-tile counts, block sums)
    uint64_t *chunk_map = nullptr;  // push workgroup -> (tile, chunk), written by the plan
    long chunk_cap = 0;           // entries of chunk_map
    long diag_cap = 0;            // workgroups diag_part has room for
    int diag_nb = 0;              // workgroups whose partials diag_part holds
    double *diag_dev = nullptr;   // [6] reduced diagnostics
    double *diag_part = nullptr;  // per-block partials of the last push (or diag pass)
    double *diag_host = nullptr;  // pinned
    bool diag_valid = false;      // diag_dev matches the current particles
    int den_state = 0;            // 0 stale, 1 windows hold the last push deposit, 2 den final
    bool sorted = false;          // particles are all in bins (staging empty)
    long stage_added = 0;         // particles appended to staging since the last push or sort
    // launch shapes of the push (from the tuning table when it has an entry for this size):
    // particles per push workgroup, map entries per XCD stripe
    long push_chunk = pic::PUSH_CHUNK;
    int xcd_stripe = pic::PUSH_XCD_STRIPE;
    // PIC_MIXED: the records hold v - frame in FP32 (FP64 species store v itself)
    double frame[3] = {0, 0, 0};
};

inline double3 frame_of(const pic_species *s) { return make_double3(s->frame[0], s->frame[1], s->frame[2]); }

namespace pic {
struct MgHierarchy;
}

struct pic_solver {
    pic_world *w = nullptr;
    int type = PIC_SOLVER_GS;
    int max_it = 10000;
    double tol = 1e-4;
    double phi0 = 0, n0 = 1e12, Te0 = 1.5;  // defaults of ch3/ver2/PotentialSolver.h:93-95
    // NR-PCG vectors
    // (P holds the Jacobian diagonal jd = A(u,u) - P(u); d and z2 are the ping-pong pair of d)
    double *y = nullptr, *g = nullptr, *d = nullptr, *z = nullptr, *F = nullptr, *P = nullptr, *z2 = nullptr;
    uint8_t *mask = nullptr;  // PCG_SYM: regular-neighbour bits per node
    // PIC_SOLVER_MG: Boltzmann derivative of the Newton step, multigrid levels
    double *boltz = nullptr;
    float *g32 = nullptr;  // FP32 copy of the CG residual, the V-cycle's right-hand side
    pic::MgHierarchy *mg = nullptr;
    // what the last solve() did (pic_solver_get_info)
    int info_newton = 0, info_iters = 0, info_fallbacks = 0, info_converged = 0;
    // Newton stopping test of the reduced solvers: RMS(y) < nr_tol (the reference's NR_TOL,
    // ch3/ver2/PotentialSolver.cpp:229; tightened by the differentiable run)
    double nr_tol = 1e-3;
    // multigrid launch shapes set through pic_solver_set_mg (0: the tuning table's, else the
    // built-in defaults)
    int mg_sweeps = 0, mg_gal_from = 0;
    long mg_fuse_max = 0;
};

// ---- error handling ------------------------------------------------------------------
namespace pic {
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);
int grow_partials(pic_world *w, int nblocks);

// the tuning table's value of `key` for the entry of `kind` nearest to (nodes, particles);
// dflt without one (tuning.hip)
double tuned(const char *kind, double nodes, double particles, const char *key, double dflt);

// phase timing helpers (no-ops unless timing is enabled)
void timer_begin(pic_world *w, int phase, hipEvent_t *a);
void timer_end(pic_world *w, int phase, hipEvent_t a);
}  // namespace pic

#define PIC_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) return pic::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define PIC_CHECK(cond, code, msg)       \
    do {                                 \
        if (!(cond)) {                   \
            pic::set_error(msg);         \
            return (code);               \
        }                                \
    } while (0)

#define PIC_LAUNCH_CHECK() PIC_HIP(hipGetLastError())

// ---- device helpers ------------------------------------------------------------------
namespace pic {

__device__ inline double warp_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block-wide FP64 sum; result valid in thread 0.  sh must hold >= blockDim/64 doubles.
__device__ inline double block_sum(double v, double *sh) {
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    v = warp_sum(v);
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    int nw = (blockDim.x + 63) >> 6;
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < nw; i++) r += sh[i];
    __syncthreads();
    return r;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11): counter-based, so every particle's draws depend only on (seed, counter).
struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)M0 * ctr.x;
        uint64_t p1 = (uint64_t)M1 * ctr.z;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        U4 n;
        n.x = hi1 ^ ctr.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ ctr.w ^ k1;
        n.w = lo0;
        ctr = n;
        k0 += W0;
        k1 += W1;
    }
    return ctr;
}

// uniform double in [0,1) from two 32-bit words (53 random bits)
__host__ __device__ inline double u01(uint32_t a, uint32_t b) {
    uint64_t v = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
    return (double)(v & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
}

}  // namespace pic
