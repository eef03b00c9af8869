// ctx.hpp -- the device-resident context behind sphgpu_ctx.
//
// Particle state lives in HBM as structure-of-arrays, permuted into octree
// key order at every tree build so that neighbour gathers touch contiguous
// memory.  `id[k]` is the reference particle index (SPHParticle::id) of the
// particle stored at sorted position k; host exchange goes through it.
//
// FP64 master arrays hold everything the integrator and the host see.  FP32
// mirrors of the per-pair inputs feed the FP32 pair loops; in FP64 pair mode
// the pair loops read the masters directly.
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

// FP32 mirror fields used by pair loops
enum MField {
    M_X, M_Y, M_Z, M_VX, M_VY, M_VZ, M_MASS, M_ENE, M_SOUND, M_SML, M_DENS,
    M_PRES, M_GRADH, M_BALSARA, M_ALPHA,
    M_GD0, M_GD1, M_GD2, M_GP0, M_GP1, M_GP2,
    M_GV0, M_GV1, M_GV2, M_GV3, M_GV4, M_GV5, M_GV6, M_GV7, M_GV8,
    M_NFIELD
};

// GSPH gradient rows in FP64 (grad_d[3], grad_p[3], grad_v[9])
constexpr int kNGrad = 15;

// Pointers a pair kernel reads for neighbour j, in precision T.
template <typename T> struct PairView {
    const T *x[3];
    const T *v[3];
    const T *mass, *ene, *sound, *sml, *dens, *pres, *gradh, *balsara, *alpha;
    const T *gd[3], *gp[3], *gv[9];
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
    // hierarchy of block reductions over sorted particles (64 and 4096)
    double *bmax1 = nullptr, *bmax2 = nullptr;
    double *bsum1 = nullptr, *bsum2 = nullptr;  // [4][nb] : m, mx, my, mz
    // root geometry (device): center[3], edge
    double *root = nullptr;
    int node_cap = 0;
    int n_nodes = 0;
    bool valid = false;
    // radix sort scratch
    unsigned int *hist = nullptr;
    int hist_blocks = 0;
};

struct Lists {
    int *idx = nullptr;   // [cap][n] sorted indices (slot-major for coalescing)
    int *cnt = nullptr;   // [n]
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
    // FP32 mirrors
    float *mf[sphgpu::M_NFIELD] = {};
    // per-particle scratch
    double *hs = nullptr;        // search radius of the pre-interaction
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

    sphgpu::Tree tree;
    sphgpu::Lists lpre, lforce;
    bool sorted_valid = false;   // state is in current tree order
    bool mirror_dirty = true;    // FP32 mirrors lag the FP64 masters
    bool pre_first = true;       // PreInteraction::m_first
    bool have_pre = false;

    // timing
    bool timing = false;
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
void lists_free(Lists &l);
// is_ij = 0: radius hs[i] (r2 < hs^2);  is_ij = 1: sqrt(r2) < max(h_i, h_j), r2 != 0
int search(sphgpu_ctx *c, Lists &l, int is_ij, const double *radius);

// hydro.hip
int pre_interaction(sphgpu_ctx *c);
int fluid_force(sphgpu_ctx *c);
int riemann_batch(sphgpu_ctx *c, int solver, int precision, int n, const double *left,
                  const double *right, double *pstar, double *vstar, int32_t *iters);

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

template <typename T> PairView<T> pair_view(sphgpu_ctx *c);

} // namespace sphgpu
