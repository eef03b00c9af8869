// ctx.hpp -- the device-resident context behind sphgpu_ctx.
//
// Particle state lives in HBM as structure-of-arrays, permuted into octree
// key order at every tree build so that neighbour gathers touch contiguous
// memory.  `id[k]` is the reference particle index (SPHParticle::id) of the
// particle stored at sorted position k; host exchange goes through it.
//
// FP64 master arrays hold everything the integrator and the host see.  The
// pair loops read packed per-particle records (FP32, or FP64 in the FP64 pair
// form) refreshed from the masters after each integrator stage and updated
// in place by PreInteraction.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/sphgpu.h"
#include "device_math.hpp"

namespace sphgpu {

// scalar FP64 fields kept per particle (sorted order)
enum SField {
    S_MASS, S_DENS, S_PRES, S_ENE, S_ENE_P, S_DENE, S_SML, S_SOUND, S_BALSARA,
    S_ALPHA, S_GRADH, S_PHI, S_NSCALAR
};

// Pair records: per particle, 16-byte (FP32) or 32-byte (FP64) 4-vectors in
// key order, so one neighbour costs a few wide loads:
//   R_POSH   x, y, z, h          R_VELM   vx, vy, vz, m
//   R_THERMO dens, pres, sound, u R_AUX   gradh, balsara, alpha, 0
//   R_G0..3  GSPH gradients: (gd0 gd1 gd2 gp0) (gp1 gp2 gv0 gv1)
//            (gv2 gv3 gv4 gv5) (gv6 gv7 gv8 0), gv index = 3*component + axis
enum RecField { R_POSH, R_VELM, R_THERMO, R_AUX, R_G0, R_G1, R_G2, R_G3, R_NREC };

// GSPH gradient rows in FP64 (grad_d[3], grad_p[3], grad_v[9])
constexpr int kNGrad = 15;

template <typename T> struct V4T;
template <> struct V4T<float> { typedef float4 type; };
template <> struct V4T<double> { typedef double4 type; };

template <typename T> struct RecView {
    typedef typename V4T<T>::type V4;
    const V4 *r[R_NREC];
};

// Node record of the gravity walk: centre of mass, edge^2, mass, member
// range [start, end), subtree end (skip) and signed level (negative: leaf);
// 64 bytes, one wave-uniform load per visited node.
struct alignas(16) GravNode {
    double cx, cy, cz, l2;
    double mass, pad;
    int start, end, skip, level;
};

struct Tree {
    // per sorted particle
    unsigned long long *key = nullptr, *key_alt = nullptr;
    int *perm = nullptr, *perm_alt = nullptr;
    unsigned char *kdig = nullptr;   // first differing level with the previous key
    unsigned char *lam = nullptr;    // leaf level of the particle's leaf
    int *leaf_end = nullptr;         // end of the leaf that starts here (if any)
    int *ncount = nullptr;           // nodes starting at each position (+ scan)
    int *nbase = nullptr;            // exclusive scan of ncount, size n+1
    // per node, pre-order
    int *start = nullptr, *end = nullptr, *skip = nullptr;
    int *level = nullptr;            // root = 1; negative marks a leaf: -level
    double *center = nullptr;        // [3][cap]
    double *edge = nullptr;
    double *ks = nullptr;            // kernel_size (max h in subtree)
    double *mass = nullptr;
    double *com = nullptr;           // [3][cap]
    // FP32 walk records for the neighbour search: (cx, cy, cz, half extent x),
    // (half extents y, z), all rounded up, and (start, end, skip, leaf flag |
    // kernel_size rounded up); a leaf's box is its members' bounding box
    float4 *nbox = nullptr;
    float2 *nhalf = nullptr;
    int4 *nlink = nullptr;
    // children of each node stored contiguously (breadth-first walks):
    // nchild[k] = (offset into child, count); leaves have count 0
    int2 *nchild = nullptr;
    int *child = nullptr;
    unsigned int *ccount = nullptr;   // child counts, scanned into offsets
    // hierarchy of block reductions over sorted particles (64 and 4096)
    double *bmax1 = nullptr, *bmax2 = nullptr;
    double *bsum1 = nullptr, *bsum2 = nullptr;  // [4][nb] : m, mx, my, mz
    // gravity walk records (one per node, refreshed by each gravity call)
    GravNode *gnode = nullptr;
    int gnode_cap = 0;
    // root geometry (device): center[3], edge
    double *root = nullptr;
    int node_cap = 0;
    int n_nodes = 0;
    bool valid = false;
    // radix sort scratch
    unsigned int *hist = nullptr;
    int hist_blocks = 0;
    // node count and the partial-sort check, read together by the host
    int *sync_words = nullptr;         // [0] node count, [1] prefix run > leaf
    int *host_words = nullptr;         // pinned host copy of sync_words
    hipEvent_t words_ready = nullptr;  // recorded after that copy
    // partial sorts order the keys on their top sorted_digits digits; a
    // failed check sorts that build fully (full_once) and adds 4 digits to
    // the later builds' partial sorts (full_sort once they cover the key)
    int sorted_digits = 8;
    bool full_once = false;
    bool full_sort = false;
    long long partial_sorts = 0, full_sorts = 0;
    unsigned int *scan_tmp = nullptr;  // tile sums of the multi-block scans
};

// Neighbour-list layout: particles in blocks of 64 (one wavefront), slot-major
// inside a block.  A wavefront's slot-q loads are one 256-byte segment, and
// the list block a wavefront writes during the search is contiguous
// (cap * 256 bytes), which keeps its scattered appends on few pages.
__device__ __host__ inline size_t list_slot(int i, int q, int cap)
{
    return ((size_t)(i >> 6) * (size_t)cap + (size_t)q) * 64 + (size_t)(i & 63);
}

struct Lists {
    int *idx = nullptr;   // see list_slot; ceil(n/64) * cap * 64 entries
    int *cnt = nullptr;   // [n]
    int *cnt_a = nullptr; // [n] force lists built from the pre-interaction:
                          // entries [0, cnt_a) have r < h_i, the rest r < h_j only
    int cap = 0;
};

struct Timer {
    hipEvent_t a, b;
    std::string name;
};

struct MLP;  // learned Riemann guess (mlp.hip)

} // namespace sphgpu

struct sphgpu_ctx {
    sphgpu_params prm;
    int n = 0;
    int dim = 3;
    hipStream_t stream = nullptr;
    std::string err;
    sphgpu::Box box;

    // FP64 masters, sorted order
    double *pos[3] = {}, *vel[3] = {}, *vel_p[3] = {}, *acc[3] = {};
    double *sc[sphgpu::S_NSCALAR] = {};
    double *grad[sphgpu::kNGrad] = {};
    int *id = nullptr, *neighbor = nullptr;
    // permutation double-buffers for the persistent fields
    double *alt[16] = {};
    int *id_alt = nullptr;
    // pair records: FP32 always (they also feed the search prefilter), FP64
    // in FP64 pair mode
    float4 *rec32[sphgpu::R_NREC] = {};
    double4 *rec64[sphgpu::R_NREC] = {};
    // per-particle scratch
    double *hs = nullptr;        // search radius of the pre-interaction
    // alpha before the time step's pre-interaction (time-dependent AV): a
    // repeated pre-interaction of the step starts its alpha update from it
    double *alpha_snap = nullptr;
    bool alpha_snap_on = false;
    // every particle has the same mass (checked when masses are uploaded;
    // SPH steps never change them): the density sums then skip the mass loads
    bool mass_uniform = false;
    double mass_u = 0.0;
    double *stage = nullptr;     // host-exchange staging, [NFIELDS][n]
    double *red = nullptr;       // reduction partials
    int red_blocks = 0;

    // device scalars: [0]=dt [1]=time [2]=h_per_v_sig [3]=dt_force_min
    double *dscal = nullptr;
    double *hscal = nullptr;     // pinned host mirror
    unsigned long long *dmin = nullptr;  // order-preserving bits for atomicMin
    // counters, see sphgpu_stats layout
    long long *dstats = nullptr;
    int *dflag = nullptr;        // [0] list overflow, [1] node overflow
    int *hflag = nullptr;        // pinned host copy of dflag
    int *hflag_dev = nullptr;    // its device address (written by k_flist_order)
    hipEvent_t flag_ready = nullptr;  // recorded after that copy

    sphgpu::Tree tree;
    sphgpu::Lists lpre, lforce;
    bool sorted_valid = false;   // state is in current tree order
    bool mirror_dirty = true;    // pair records lag the FP64 masters
    bool pre_first = true;       // PreInteraction::m_first
    bool have_pre = false;
    // force pair lists derived from the last pre-interaction (no is_ij
    // search): 0 none, 1 built but not yet checked, 2 valid
    int flist = 0;
    long long flist_searches = 0;  // FluidForce calls that searched instead
    // the pre-interaction search's overflow flag has not been read yet (the
    // time step reads it together with the force-list flag, one copy)
    bool pre_search_pending = false;
    // Search radius of the time step's pre-interaction: the reference's
    // kappa h_guess (kappa = 1.2) scaled by pre_rscale.  Inside the time
    // step's deferred path the radius is (1 + pre_eps) h_guess; the
    // pre-interaction flags any particle whose Newton iterates reach it, and
    // that step's pre-interaction then runs again at the reference's radius.
    // pre_eps doubles after such a step (up to kappa - 1, the reference's
    // radius) and halves again after 64 steps without one.
    double pre_rscale = 1.0;
    double pre_eps = 0.05;
    int pre_quiet = 0;
    // set by the time-step loop around its tree build: fields that the step
    // recomputes before reading them are not permuted
    bool in_step = false;

    // timing
    int timing = 0;  // 0 off, 1 every kernel group, 2 pair kernels only
    bool timing_open = false;  // time_begin recorded an event for this group
    std::vector<sphgpu::Timer> pending;
    std::map<std::string, std::pair<double, long long>> ktimes;
    std::vector<hipEvent_t> event_pool;

    // learned guess
    sphgpu::MLP *mlp = nullptr;
    float *rlog = nullptr;
    long long rlog_cap = 0;
    int rlog_stride = 1;
    unsigned long long *rlog_count = nullptr;
    bool rlog_on = false;
};

namespace sphgpu {

// error helpers --------------------------------------------------------------
int fail(sphgpu_ctx *c, int code, const std::string &msg);
#define HIPCHK(c, expr)                                                        \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess)                                                  \
            return ::sphgpu::fail((c), SPHGPU_E_HIP,                           \
                std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

// timing around a launch
void time_begin(sphgpu_ctx *c, const char *name);
void time_end(sphgpu_ctx *c);
void collect_times(sphgpu_ctx *c);

inline int grid_for(long long n, int block) { return (int)((n + block - 1) / block); }

// tree.hip
int tree_alloc(sphgpu_ctx *c);
void tree_free(sphgpu_ctx *c);
int tree_build(sphgpu_ctx *c);
int tree_set_kernel(sphgpu_ctx *c);
int tree_dump(sphgpu_ctx *c, int64_t *n_nodes, int32_t *level, int32_t *count, int32_t *first,
              int32_t *is_leaf, double *center, double *edge, double *mass, double *com,
              double *kernel_size, int32_t *members);

// neighbor.hip
int lists_alloc(sphgpu_ctx *c, Lists &l, int cap);
/* doubles l's capacity up to the reference's nn * 1000; SPHGPU_E_NEIGHBOR_OVERFLOW at that size */
int lists_grow(sphgpu_ctx *c, Lists &l);
void lists_free(Lists &l);
// is_ij = 0: radius hs[i] (r2 < hs^2);  is_ij = 1: sqrt(r2) < max(h_i, h_j), r2 != 0
// defer: launch only; the overflow flag is read later (pre_resolve)
// flag_cleared: the overflow flag was reset by an earlier launch of the step
int search(sphgpu_ctx *c, Lists &l, int is_ij, const double *radius, bool defer = false,
           bool flag_cleared = false, double rscale = 1.0);
// settles a pending force list (flist == 1) from the device flag: 0 = use it
int flist_check(sphgpu_ctx *c);

// hydro.hip
int pre_interaction(sphgpu_ctx *c);
// the same with its search's overflow check left to pre_resolve (with the
// time-dependent viscosity, a repeated run takes alpha from alpha_snap)
int pre_interaction_deferred(sphgpu_ctx *c);
// reads a deferred search's overflow flag (and the force-list flag in the
// same copy); after an overflow the lists grow and the pre-interaction runs
// again from the same search radii
int pre_resolve(sphgpu_ctx *c);
int fluid_force(sphgpu_ctx *c);
int riemann_batch(sphgpu_ctx *c, int solver, int precision, int n, const double *left,
                  const double *right, double *pstar, double *vstar, int32_t *iters,
                  int32_t *learned_used);
// the same without allocating: staging in the library's static device area
int riemann_batch_static(sphgpu_ctx *c, int solver, int precision, int n, const double *left,
                         const double *right, double *pstar, double *vstar, int32_t *iters,
                         int32_t *learned_used);

// gravity.hip
int gravity(sphgpu_ctx *c);

// integrate.hip
int predict(sphgpu_ctx *c);
int correct(sphgpu_ctx *c);
int timestep(sphgpu_ctx *c);
int energy_momentum(sphgpu_ctx *c, double out[7]);
int refresh_mirrors(sphgpu_ctx *c);
int init_fields(sphgpu_ctx *c);
int advance_time(sphgpu_ctx *c);

// mlp.hip
int mlp_create(sphgpu_ctx *c);
void mlp_destroy(sphgpu_ctx *c);
const float *mlp_device_params(sphgpu_ctx *c);

template <typename T> RecView<T> rec_view(sphgpu_ctx *c);

} // namespace sphgpu
