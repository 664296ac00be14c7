// Internal structures shared by the vulcan_amd HIP sources.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "../../include/vulcan_amd.h"

#define VK_NI_MAX VK_NI_LIMIT  // largest network the kernels are compiled for
#define VK_NATOM_MAX 8
#define VK_JAC_LIGHT 8   // term lists up to this length are evaluated in the dense pass

#define VK_CHECK(expr)                                                   \
    do {                                                                 \
        hipError_t e_ = (expr);                                          \
        if (e_ != hipSuccess) return -(int)e_;                           \
    } while (0)


// one term of a Jacobian entry: coef * dv[slot] (one 16-byte load)
// one Jacobian term packed in 32 bits: dv slot << 8 | net stoichiometric
// coefficient (a small integer, signed 8 bits); the coefficient converts
// back to the same double, so products and sums are unchanged
constexpr int VK_JAC_SLOT_MAX = 1 << 23;
__host__ __device__ inline int vk_jac_pack(int slot, int coef) { return (slot << 8) | (coef & 0xFF); }
__host__ __device__ inline int vk_jac_slot(int w) { return w >> 8; }
__host__ __device__ inline double vk_jac_coef(int w) { return (double)((int)((unsigned)w << 24) >> 24); }

struct vk_plan {
    int ni, npairs, nr, natom, nchem, nslot, njnz, njac;
    int factor_variant;
    int *reac, *prod, *kind, *rev;
    double *par, *nasa9, *atom_mat, *ms, *alpha;
    int *chem_ptr, *chem_pair, *chem_sign;
    int2 *chem_ent;            // (chem_pair, chem_sign) interleaved
    int *slot_pair, *slot_sp;
    int *jac_map, *jac_ptr, *jac_slot;
    int njent;                 // entries with chemical terms plus the diagonal
    int4 *jac_ent;             // (dense index, first term, end of terms, row), longest term lists first
    int *jac_emap;             // dense index -> (first term << 4 | term count) of a short list, -1 zero, -2 long list
    int njheavy;               // entries with more than VK_JAC_LIGHT terms (the first ones of jac_ent)
    int *jac_tp;               // packed terms (vk_jac_pack) in jac_coef / jac_slot order
    double *jac_coef;
    std::vector<void *> allocs;
};

// Block cyclic reduction schedule of a set of columns: at each level every
// other active layer of a column is eliminated, the rest survive (from an
// odd count the even positions, both ends included, are eliminated, so
// floor(n/2) layers survive and a column of nz layers needs floor(log2 nz)
// levels); one layer per column remains on top.
struct vk_bcr_sched {
    int nlev = 0;
    std::vector<int> elim_off, surv_off;   // nlev+1 each, into the entry arrays
    int nsurv = 0, ntop = 0;
    int4 *d_elim = nullptr;   // (layer, left, right or -1, column)
    int4 *d_surv = nullptr;   // (layer, left or -1, right or -1, column)
    int *d_sflag = nullptr;   // per surviving entry: bit 0 new left coupling, bit 1 new right coupling
    int2 *d_top = nullptr;    // (layer, column)
    // one column whose last level (not level 0) eliminates the layer right
    // of the top survivor (two active layers): that level's sweeps and the
    // top solve run as one launch
    bool tail = false;
};

// workspace of the block cyclic reduction: A, B, C per layer (n x n each),
// P, Q per surviving (level, layer) entry, XB = A^-1 B and XC = A^-1 C per
// eliminated layer (the back-substitution's operators)
struct vk_bcr_ws {
    double *A = nullptr, *B = nullptr, *C = nullptr, *P = nullptr, *Q = nullptr, *XB = nullptr, *XC = nullptr;
};

struct vk_columns {
    int ncols, L, max_nz;
    std::vector<int> off;      // host copy, ncols+1
    int *d_off;                // ncols+1
    int *d_layer_col;          // L
    // chain order of the blocked chain factorisations: workgroup b runs
    // chain order[b] (2c: top half of column c, 2c+1: bottom half).
    // [0]: one workgroup per CU, longest chains first (a freed CU takes the
    // longest chain left); [1]: two per CU, where workgroups b and b + nCU
    // share a CU (first nCU workgroups one per CU), pairing the longest
    // chains with the shortest.  The order moves chains between CUs only;
    // each chain's arithmetic is the same wherever it runs.
    int *d_chain_order[2];     // 2 * ncols each
    vk_bcr_sched bcr;
};

// per-layer partial results of the finalize kernel
enum { VKP_DELTA = 0, VKP_NEG = 1, VKP_SMALL = 2, VKP_NEGY = 3, VKP_ATOM = 4 };

// photochemistry radiative transfer of a set of columns (vk_photo.hip)
int vk_launch_photo(vk_photo *p, const vk_plan *plan, const vk_columns *cols, const double *y, const double *ymix,
                    const double *dz, double *k, vk_colstate *st, int force, int ini_frq, int final_frq,
                    double yconv_min, hipStream_t s);

struct vk_integrator {
    const vk_plan *plan;
    const vk_columns *cols;
    int history;               // ring length for conv (0 = no history)
    int tcap;                  // capacity of t_time per column
    double *D, *Dinv, *Lsub, *Usup;     // lhs blocks, inverse Schur complements, couplings
    double *b1, *x1, *b2, *x2, *work;   // right-hand sides and solutions of the two stages
    double *ztmp;                       // forward-sweep scratch of the twisted solve
    double *yk2, *part;
    double *ysum;                       // layer totals for the diffusion terms
    vk_bcr_ws bcr;                      // block cyclic reduction workspace (allocated when selected)
    int *info;
    double *yhist;             // [history][L][ni]
    double *thist;             // [ncols][tcap]
    int *flags;                // [ncols] per-attempt decision
    double *colscr;            // [ncols][16] per-column scratch
    hipEvent_t ev0, ev1;
    int timing;                // time the factor kernel of the last attempt per call
    double factor_ms;          // summed duration of the timed factor launches
    long long factor_launches; // number of timed factor launches
    int use_graph;             // replay one captured loop iteration per attempt
    hipGraphExec_t gexec;
    hipStream_t gstream;
    std::vector<char> gkey;    // arguments baked into gexec
    vk_photo *photo;           // radiative transfer refreshed in the loop (may be null)
    int photo_ini_frq, photo_final_frq;
    std::vector<void *> allocs;
};

// Block-inversion launchers of the factor variants (nb = ceil(ni / 16),
// one translation unit each); 0 or a VK_E* code.
int vk_launch_factor_register(int nb, int ni, int ncols, const int *off, const double *D, const double *Lsub,
                              const double *Usup, double *Dinv, int *info, const vk_colstate *st, hipStream_t s);
int vk_launch_factor_blocked1(int nb, int ni, int ncols, const int *off, const int *order, const double *D,
                              const double *Lsub, const double *Usup, double *Dinv, int *info, const vk_colstate *st,
                              hipStream_t s);
int vk_launch_factor_blocked2(int nb, int ni, int ncols, const int *off, const int *order, const double *D,
                              const double *Lsub, const double *Usup, double *Dinv, int *info, const vk_colstate *st,
                              hipStream_t s);
int vk_launch_factor_wave(int nb, int ni, int ncols, const int *off, const double *D, const double *Lsub,
                          const double *Usup, double *Dinv, int *info, const vk_colstate *st, hipStream_t s);
int vk_launch_bcr_factor(int nb, int ni, const vk_columns *cols, const double *D, const double *Lsub,
                         const double *Usup, double *Dinv, int *info, const vk_colstate *st, const vk_bcr_ws *ws,
                         hipStream_t s);
// acc (nullable): also add the solution into acc, row by row as it is written
int vk_launch_bcr_solve(int ni, const vk_columns *cols, const double *Dinv, const double *Lsub, const double *Usup,
                        const double *in, double *out, double *tmp, const vk_colstate *st, const vk_bcr_ws *ws,
                        double *acc, hipStream_t s);
