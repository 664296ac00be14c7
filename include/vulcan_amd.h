/*
 * vulcan_amd: MI355X (gfx950) kernels for VULCAN's Ros2 kinetics-plus-
 * diffusion step.  Plain C ABI for ctypes / any FFI.
 *
 * Conventions
 *  - Every array argument named d_* is a device pointer owned by the caller
 *    (for example a PyTorch-ROCm tensor's data_ptr()).  The library owns only
 *    the immutable network tables (vk_plan), the column layout (vk_columns)
 *    and the integrator workspace (vk_integrator).
 *  - Layers of all columns are stacked: column c owns layers
 *    [off[c], off[c] + nz[c]) of every per-layer array.  Per-layer arrays
 *    are [L] or [L][ni] (species fastest), rate arrays are [L][nr] with
 *    forward k at 2p and reverse k at 2p+1 for reaction pair p (VULCAN's
 *    k[2p+1], k[2p+2]).  Interface quantities of a column (Kzz, Dzz, dzi, Ti,
 *    Hpi: nz-1 values) sit in the first nz-1 layer slots of that column.
 *  - Jacobian / lhs blocks are [L][ni][ni] row-major: entry (i, m) is
 *    d f_i / d y_m of layer l; Lsub[l] (Usup[l]) is the diagonal of the
 *    coupling block to layer l-1 (l+1).
 *  - Kernels are asynchronous on the given stream (hipStream_t passed as
 *    void*; NULL = default stream).  Return value: 0 = ok, < 0 = negated HIP
 *    error or VK_E* code below.  Numerical failure is never an error: it is
 *    reported through the state (negative y, delta, element loss), as in the
 *    reference.
 *  - A plan, columns object or integrator may be used by one host thread at a
 *    time.
 */
#ifndef VULCAN_AMD_H
#define VULCAN_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define VK_EINVAL   (-1001)   /* bad argument / shape            */
#define VK_ENOMEM   (-1002)   /* device allocation failed        */
#define VK_ETOOBIG  (-1003)   /* ni beyond the compiled maximum  */
#define VK_ENOTCHECKED (-1004) /* query of the bounds-checked build on the normal one */

#define VK_NI_LIMIT 160           /* largest ni the kernels are compiled for */
#define VK_SPECIES_WORDS 5        /* ceil(VK_NI_LIMIT / 32): one bit per species */
#define VK_CONV_IGNORE_WORDS VK_SPECIES_WORDS

/* rate-expression kinds (par[p][0..7] = A n E Ainf ninf Einf Fc J0) */
#define VK_KIND_ARRH   0  /* op.py:166               */
#define VK_KIND_LIND   1  /* op.py:176-182           */
#define VK_KIND_K0     2  /* op.py:185-186           */
#define VK_KIND_TROE   3  /* op.py:203-207 form      */
#define VK_KIND_PHOTO  4  /* J0 exp(-tau), op.py:2786 */
#define VK_KIND_CH3OH  5  /* op.py:197-207           */

typedef struct vk_plan vk_plan;
typedef struct vk_columns vk_columns;
typedef struct vk_integrator vk_integrator;
typedef struct vk_photo vk_photo;

/* Host description of a network (copied into device tables by
 * vk_plan_create).  Terms: species index, -1 = M, -2 = empty slot. */
typedef struct {
    int ni;                 /* species                                   */
    int npairs;             /* reversible pairs (nr = 2*npairs rate rows) */
    int natom;              /* monitored elements                        */
    const int *reac;        /* [npairs][4] terms as written              */
    const int *prod;        /* [npairs][4]                               */
    const int *kind;        /* [npairs]                                  */
    const double *par;      /* [npairs][8]                               */
    const int *rev;         /* [npairs] 1: k_r = k_f / K_eq, 0: k_r = 0  */
    const double *nasa9;    /* [ni][20] low [0:10], high [10:20]         */
    const double *atom_mat; /* [ni][natom] element counts                */
    const double *ms;       /* [ni] molar mass (g/mol)                   */
    const double *alpha;    /* [ni] thermal diffusion factor             */
    /* derived tables (see vulcan_amd/network.py: build_tables) */
    const int *chem_ptr;    /* [ni+1]   */
    const int *chem_pair;   /* [nchem]  */
    const int *chem_sign;   /* [nchem]  */
    int nchem;
    const int *slot_pair;   /* [nslot]  */
    const int *slot_sp;     /* [nslot]  */
    int nslot;
    const int *jac_map;     /* [ni*ni]  */
    const int *jac_ptr;     /* [njnz+1] */
    const int *jac_slot;    /* [njac]   */
    const double *jac_coef; /* [njac]   */
    int njnz;
    int njac;
} vk_network_desc;

/* Device atmosphere arrays of a set of columns (all [L] unless noted). */
typedef struct {
    const double *Tco;
    const double *M;        /* third body = total number density n_0   */
    const double *Kzz;      /* interfaces                              */
    const double *Dzz;      /* [L][ni] interfaces                      */
    double *dzi;            /* interfaces (updated by update_mu_dz)    */
    const double *Ti;       /* interfaces                              */
    double *Hpi;            /* interfaces (updated)                    */
    double *g;              /* layers (updated)                        */
    double *Hp;             /* layers (updated)                        */
    const double *tau;      /* photolysis optical depth, may be NULL   */
    /* Transport terms and boundary conditions beyond the default
     * configuration; a NULL array or a zero switch turns each off. */
    const double *vz;       /* interfaces: vertical velocity, upwind advection (use_vz; op.py:1545-1570) */
    const double *vm;       /* [L][ni] layers: molecular-diffusion velocity; selects the upwind
                               molecular terms of diffdf_vm / lhs_jac_tot_vm (use_vm_mol; op.py:1599-1694,
                               2044-2120)                                                               */
    double *top_flux;       /* [ncols][ni] top flux, cm^-2 s^-1 (store.py:138); rewritten for the
                               escaping species every update_frq steps (op.py:986-999)                  */
    const double *bot_flux; /* [ncols][ni] bottom flux (store.py:139)                                   */
    const double *bot_vdep; /* [ncols][ni] deposition velocity, cm s^-1 (store.py:140)                  */
    const double *fix_bot;  /* [ncols][ni] fixed bottom mixing ratio, < 0: free (use_fix_sp_bot;
                               op.py:2945-2946); NULL: none                                             */
    int use_topflux;        /* diffdf adds top_flux / dzi[-1] (op.py:1589-1592)                         */
    int use_botflux;        /* diffdf adds (bot_flux - y bot_vdep) / dzi[0], the lhs the deposition
                               (op.py:1593-1596, 2029); delta skips layer 0 (op.py:2953)               */
    /* diff_esc (vulcan_cfg.py:101): bit (i % 32) of word i / 32 set = species i
     * escapes at the diffusion limit, capped at max_flux (vulcan_cfg.py:102) */
    unsigned int esc[VK_SPECIES_WORDS];
    double max_flux;
} vk_atm;

/* Column geometry inputs for update_mu_dz (per column / per layer). */
typedef struct {
    const double *lnp;      /* [L] log(pico[l]/pico[l+1])              */
    double *zco;            /* [L + ncols] interface heights (nz+1 per column, offset off[c]+c) */
    double *dz;             /* [L]                                     */
    double *mu;             /* [L]                                     */
    const double *gs;       /* [ncols]                                 */
    const double *Rp;       /* [ncols]                                 */
    const int *pref;        /* [ncols]                                 */
} vk_geom;

/* Numerical knobs (vulcan_cfg.py:124-159; names as there). */
typedef struct {
    double dt_min, dt_max, dt_var_min, dt_var_max;
    double rtol, atol, mtol, mtol_conv, pos_cut, nega_cut, loss_eps;
    double trun_min, runtime, st_factor, yconv_cri, slope_cri, yconv_min, flux_cri;
    int count_min, count_max, conv_step, update_frq;
    int use_moldiff;
    int use_conv;           /* evaluate the steady-state test (needs history)  */
    int use_update_mu;      /* run update_mu_dz every update_frq steps         */
    int refine;             /* refinement steps per linear solve (1 = default) */
    /* conver_ignore (vulcan_cfg.py:153, op.py:1045-1046): bit (i % 32) of
     * word i / 32 set = species i is left out of conv()'s longdy */
    unsigned int conv_ignore[VK_CONV_IGNORE_WORDS];
} vk_params;

/* Per-column integrator state (device resident; mirrors var/para fields). */
typedef struct {
    double t, dt, delta, longdy, longdydt, small_y, nega_y;
    double aflux_change;  /* actinic-flux change tested by conv() (op.py:1060) */
    int count, attempts, nega_count, loss_count, delta_count, end_case;
    int retry;      /* 1: the next attempt retries the current step       */
    int done;       /* 1: stop() returned true, column no longer steps     */
} vk_colstate;

/* ---- plan / layout -------------------------------------------------------*/
int vk_plan_create(const vk_network_desc *net, vk_plan **out);   /* make_chem_funs.py output + op.ReadRate tables */
int vk_plan_destroy(vk_plan *plan);
int vk_plan_info(const vk_plan *plan, int *ni, int *nr, int *max_ni);
int vk_columns_create(int ncols, const int *nz, vk_columns **out);
int vk_columns_destroy(vk_columns *cols);
int vk_columns_layers(const vk_columns *cols, int *nlayers);

/* ---- reference functions, one kernel each ------------------------------- */
/* op.ReadRate.read_rate + rev_rate (op.py:63-309), photolysis k = J (op.py:2786) */
int vk_rates(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm,
             double *d_k, void *stream);
/* chem_funs.chemdf(y, M, k) (chem_funs.py:931-1000) */
int vk_chemdf(const vk_plan *plan, const vk_columns *cols, const double *d_y,
              const double *d_M, const double *d_k, double *d_dydt, void *stream);
/* ODESolver.diffdf / diffdf_no_mol (op.py:1496-1597 / 1438-1494) */
int vk_diffdf(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm,
              const double *d_y, int use_moldiff, double *d_out, void *stream);
/* chemdf + diffdf fused; with d_k1 != NULL also - 2/(r dt_c) k1 (op.py:2892, 2918-2928) */
int vk_rhs(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm,
           const double *d_y, const double *d_k, int use_moldiff,
           const double *d_k1, const double *d_dt, double *d_out, void *stream);
/* chem_funs.symjac diagonal blocks (chem_funs.py:3222-7453); neg_symjac = -J */
int vk_jac(const vk_plan *plan, const vk_columns *cols, const double *d_y,
           const double *d_M, const double *d_k, double *d_J, void *stream);
/* ODESolver.lhs_jac_tot / lhs_jac_no_mol (op.py:1973-2042 / 2122-2166) as blocks */
int vk_lhs(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm,
           const double *d_y, const double *d_k, int use_moldiff,
           const double *d_dt, double *d_D, double *d_Lsub, double *d_Usup, void *stream);
/* replaces store_bandM + LAPACK gbtrf inside scipy.linalg.solve_banded
 * (op.py:2833-2858, 2914): twisted block-Thomas factorisation (a lower
 * chain from layer 0 and an upper chain from layer nz-1 meet at layer
 * nz/2); d_Dinv[l] receives the inverse of layer l's Schur complement,
 * stored transposed (d_Dinv[l][m][i] = S_l^-1[i][m]); d_D is not modified;
 * d_info[c] = 1 + layer of a singular block, or 0 */
int vk_block_factor(const vk_plan *plan, const vk_columns *cols, const double *d_D,
                    const double *d_Lsub, const double *d_Usup, double *d_Dinv,
                    int *d_info, void *stream);
/* Block-inversion kernel of the factorisation (no counterpart in the
 * reference, whose banded LU has one code path): VK_FACTOR_REGISTER keeps
 * the block in registers and runs one barrier-synchronised Gauss-Jordan
 * step per pivot; VK_FACTOR_BLOCKED groups the pivots in panels of 16,
 * factors each panel in one wave and applies the rank-16 Schur update on
 * FP64 MFMA (compiled for one or for two resident workgroups per CU);
 * VK_FACTOR_WAVE holds the block in the registers of two waves and runs
 * one two-wave barrier per pivot step (small workgroups: several chains per
 * CU).  VK_FACTOR_BCR replaces the chain over layers by log2(nz) levels of
 * block cyclic reduction (independent inversions per level, dense
 * Schur-complement updates on FP64 MFMA): far shorter dependence chains for
 * few columns; it keeps its own workspace, so it runs inside the
 * integrator and vk_block_solve_system, while vk_block_factor /
 * vk_block_solve reject it.  All give the solution to rounding; vk_tune_factor times every
 * variant on a synthetic system of the columns' shape (`reps` launches
 * each: one factorisation and two solves, milliseconds per repetition into
 * ms[VK_FACTOR_NVARIANTS], -1 for a variant that does not apply: the
 * two-wave kernel above ni = 112, cyclic reduction above 64 columns) and
 * keeps the
 * fastest in the plan. */
#define VK_FACTOR_REGISTER     0
#define VK_FACTOR_BLOCKED      1   /* register budget for one workgroup per CU      */
#define VK_FACTOR_BLOCKED_OCC2 2   /* register budget for two workgroups per CU     */
#define VK_FACTOR_WAVE         3   /* two waves per block, unblocked (ni <= 112)    */
#define VK_FACTOR_BCR          4   /* block cyclic reduction across layers          */
#define VK_FACTOR_NVARIANTS    5
int vk_plan_set_factor_variant(vk_plan *plan, int variant);
int vk_plan_factor_variant(const vk_plan *plan);
int vk_tune_factor(vk_plan *plan, const vk_columns *cols, int reps, float *ms, void *stream);
/* x = lhs^-1 b for the block system (scipy.linalg.solve_banded at op.py:2914,
 * 2929, with `refine` steps of double-double iterative refinement), any
 * factor variant; allocates its workspace on the stream.  info as for
 * vk_block_factor. */
int vk_block_solve_system(const vk_plan *plan, const vk_columns *cols, const double *d_D,
                          const double *d_Lsub, const double *d_Usup, const double *d_b, double *d_x,
                          int refine, int *d_info, void *stream);
/* vk_block_solve_system that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace (inverse blocks, refinement vectors and, for
 * cyclic reduction, its coupling and operator blocks) comes from one static
 * device array of VK_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends; a system that needs more returns VK_ETOOBIG (for cyclic reduction
 * about one nz = 150, ni = 93 column; for the chain variants about 1400
 * layers at ni = 93).  Every call on a GPU shares that GPU's array, so calls
 * that overlap on one GPU (other streams, other host threads) race; calls
 * ordered on one stream do not.  Nothing in the scratch needs to start at
 * zero.  Results are those of vk_block_solve_system. */
#define VK_NOALLOC_SCRATCH_BYTES (96u << 20)
int vk_block_solve_system_noalloc(const vk_plan *plan, const vk_columns *cols, const double *d_D,
                                  const double *d_Lsub, const double *d_Usup, const double *d_b,
                                  double *d_x, int refine, int *d_info, void *stream);
/* Self-test of the FP64 MFMA operand layout the blocked inversion relies
 * on: out[i*16+j] = sum_k A[i][k] B[k][j] with A[i][k] = i + k/4 + 1,
 * B[k][j] = 1 - k/2 + j/8 (16 x 4 x 16), computed by one MFMA. */
int vk_selftest_mfma_f64(double *d_out, void *stream);
/* Self-test of the wave-wide maximum used by the blocked inversion's pivot
 * search (32-bit keys): out[l] = max over the 64 lanes of in[], for every
 * lane l. */
int vk_selftest_wave_max_u32(const unsigned *d_in, unsigned *d_out, void *stream);
/* Self-test of the reciprocal of the blocked inversion's pivot step (the
 * hardware estimate refined by two Newton-Raphson steps): d_out[i] =
 * that reciprocal of d_in[i], for i < n. */
int vk_selftest_recip_f64(int n, const double *d_in, double *d_out, void *stream);

/* replaces the gbtrs half of solve_banded (op.py:2914, 2929): rhs <- lhs^-1 rhs */
int vk_block_solve(const vk_plan *plan, const vk_columns *cols, const double *d_Dinv,
                   const double *d_Lsub, const double *d_Usup, double *d_rhs, void *stream);
/* r = b - lhs x accumulated in double-double: the residual of one step of
 * iterative refinement, which brings the block solve to the accuracy of
 * LAPACK's banded LU or better on ill-conditioned blocks (no counterpart in
 * the reference; used inside the Ros2 attempt for both solves) */
int vk_block_residual(const vk_plan *plan, const vk_columns *cols, const double *d_D,
                      const double *d_Lsub, const double *d_Usup, const double *d_x,
                      const double *d_b, double *d_r, void *stream);

/* ---- integrator (Ros2.solver / one_step / step_size, Integration loop) --- */
int vk_integrator_create(const vk_plan *plan, const vk_columns *cols, int history,
                         vk_integrator **out);
int vk_integrator_destroy(vk_integrator *it);
/* Ros2.solver (op.py:2860-3007): one attempt for every column that is not
 * done; y is replaced by sol, ymix by sol/sum(sol), st.delta set.  The
 * attempt does not clip or accept. */
int vk_ros2_solver(vk_integrator *it, const vk_atm *atm, const vk_params *prm,
                   const double *d_k, double *d_y, double *d_ymix,
                   vk_colstate *d_st, void *stream);
/* Loop body of Integration.__call__ (op.py:813-923) including Ros2.one_step
 * (op.py:3091-3103), clip/loss/step_ok/step_reject (op.py:2447-2534),
 * update_mu_dz (op.py:944-984), renormalisation (909-914), save_step,
 * step_size (3105-3125) and stop/conv (1018-1085).  Runs n_attempts
 * solver attempts (accepted or not) on every column that is not done.
 * d_y_prev and d_atom_* are per-column state arrays ([L][ni],
 * [ncols][natom]). */
int vk_integrate(vk_integrator *it, const vk_atm *atm, const vk_geom *geom,
                 const vk_params *prm, const double *d_k, double *d_y,
                 double *d_ymix, double *d_y_prev, double *d_atom_ini,
                 double *d_atom_loss, double *d_atom_loss_prev,
                 vk_colstate *d_st, int n_attempts, void *stream);
/* Integration.stop() with conv() (op.py:1018-1085) for every column that is
 * not done: sets st.longdy / st.longdydt, and end_case / done when the
 * column converged (1), ran past runtime (2) or past count_max (3).
 * History as vk_integrate keeps it: y_time[q] (the state after step q+1)
 * at d_yhist + (q % H) * L * ni for q >= max(count - conv_step, 0) and
 * t_time[q] at d_thist[c * tcap + q]; d_y must be y_time[count-1].
 * With prm->use_conv = 0 only the runtime and count limits are tested. */
int vk_conv(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm,
            const vk_params *prm, const double *d_y, const double *d_ymix,
            const double *d_yhist, int H, const double *d_thist, int tcap,
            vk_colstate *d_st, void *stream);
/* Host copies of the history vk_integrate keeps for conv() (save_step's
 * y_time / t_time, op.py:1087-1105): with write = 1 loads t_time[0:count]
 * of every column (h_t: [ncols][count]) and y_time[q] for q0 <= q < count
 * (h_y: [count - q0][L][ni]) into the integrator, e.g. to resume a run at
 * para.count > 0; with write = 0 reads them back.  count - q0 may not exceed
 * the integrator's history length; either host pointer may be NULL.
 * Synchronises the device. */
int vk_integrator_history(vk_integrator *it, int write, int count, int q0,
                          double *h_t, double *h_y);
/* update_mu_dz (op.py:944-984) for every column that is not done */
int vk_update_mu_dz(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm,
                    const vk_geom *geom, const double *d_ymix, const vk_colstate *d_st,
                    int force, void *stream);
/* update_phi_esc (op.py:986-999): top_flux[c][i] = max(-Dzz[-1][i] y[-1][i]
 * (1/Hp[-1] - ms[i] g[-1] / (Navo kb T[-1])), -max_flux) for the species of
 * atm->esc, every column (no-op when none is marked) */
int vk_update_phi_esc(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm,
                      const double *d_y, void *stream);
/* element sums: atom_sum[c][a] = sum_{l,i} atom_mat[i][a] y[l][i] (op.py:2472-2487, build_atm.py:283-291) */
int vk_atom_sums(const vk_plan *plan, const vk_columns *cols, const double *d_y,
                 double *d_atom_sum, void *stream);
/* With timing on, vk_integrate records HIP events around the factor kernel
 * of its last attempt (on the call's stream) and, before returning, waits
 * for them and adds the duration to the totals that vk_integrator_stats
 * reports (summed milliseconds, number of timed launches). */
int vk_integrator_set_timing(vk_integrator *it, int on);
int vk_integrator_stats(const vk_integrator *it, double *factor_ms, long long *factor_launches);
/* With graphs on (and timing off), vk_integrate captures one loop iteration
 * (attempt, decision, bookkeeping, conv: about 25 kernels) into a HIP graph
 * on a default-priority stream owned by the integrator and replays it once
 * per attempt, ordered after and before the caller's stream with events.
 * The graph is re-captured whenever the scalar arguments or device pointers
 * change.  No counterpart in the reference (its loop is per-step Python);
 * results are bitwise those of the eager launches. */
int vk_integrator_set_graph(vk_integrator *it, int on);

/* ---- photochemistry radiative transfer ----------------------------------
 * Tables of ODESolver.make_bins_read_cross (op.py:495-781) and
 * Atm.read_sflux (build_atm.py:617-657), prepared on the host: wavelength
 * bins, absorption cross sections of the photolysed species (with per-layer
 * temperature-dependent tables where vulcan_cfg.T_cross_sp asks for them),
 * Rayleigh cross sections of the scattering species, and per
 * photodissociation branch its cross section and the rate column it sets
 * (VULCAN's pho_rate_index - 1: the reference indexes k by the id printed
 * in the network file, op.py:241-244).  Per-layer tables are [n][L][nbin] over the stacked
 * layers of the columns. */
typedef struct {
    int nbin;                  /* wavelength bins                              */
    int nabs;                  /* absorbing (photolysed) species               */
    int nscat;                 /* Rayleigh-scattering species (scat_sp)        */
    int nbranch;               /* photodissociation branches                   */
    int nT, nTJ;               /* temperature-dependent absorption / branch tables */
    int din12;                 /* first bin of width dbin2 (-1: none)          */
    const int *abs_sp;         /* [nabs] species index                         */
    const int *abs_T;          /* [nabs] index into cross_T or -1              */
    const int *scat_sp;        /* [nscat] species index                        */
    const int *br_col;         /* [nbranch] rate column: k[l][col] = J f_diurnal */
    const int *br_T;           /* [nbranch] index into cross_J_T or -1         */
    const double *bins;        /* [nbin] nm                                    */
    const double *sflux_top;   /* [nbin] stellar flux at the top               */
    const double *cross;       /* [nabs][nbin] cm^2                            */
    const double *cross_T;     /* [nT][L][nbin]                                */
    const double *cross_scat;  /* [nscat][nbin]                                */
    const double *cross_J;     /* [nbranch][nbin]                              */
    const double *cross_J_T;   /* [nTJ][L][nbin]                               */
    double dbin1, dbin2;       /* bin widths below / from the switch (nm)      */
    double cos_sl;             /* cos(sl_angle)                                */
    double edd, f_diurnal, flux_atol, hc;
} vk_photo_desc;

#define VK_PHOTO_TAU     0     /* [L + ncols][nbin] optical depth (levels)    */
#define VK_PHOTO_SFLUX   1     /* [L + ncols][nbin] direct beam (levels)      */
#define VK_PHOTO_DFLUX_U 2     /* [L + ncols][nbin] diffuse up flux (levels)  */
#define VK_PHOTO_DFLUX_D 3     /* [L + ncols][nbin] diffuse down flux (levels) */
#define VK_PHOTO_AFLUX   4     /* [L][nbin] actinic flux (layers)             */

/* Uploads the tables and allocates the radiative-transfer state of the
 * columns (column c's levels at rows off[c] + c .. off[c] + c + nz). */
int vk_photo_create(const vk_plan *plan, const vk_columns *cols, const vk_photo_desc *desc, vk_photo **out);
int vk_photo_destroy(vk_photo *photo);
/* ODESolver.compute_tau + compute_flux + compute_J (op.py:2580-2787) for
 * every column whose update is due: with force, all of them; otherwise the
 * columns not done whose st.count is a multiple of the update frequency in
 * force and that have not been updated at this count, the frequency going
 * from ini_frq to final_frq for good once longdy < 10 yconv_min and
 * longdydt < 1e-6 (op.py:818-830).  Writes the photolysis rates
 * k[l][col] = J f_diurnal and st.aflux_change (the largest relative change
 * of the actinic flux above flux_atol, op.py:2737), which conv() tests. */
int vk_photo_update(vk_photo *photo, const vk_plan *plan, const vk_columns *cols, const double *d_y,
                    const double *d_ymix, const double *d_dz, double *d_k, vk_colstate *d_st, int force,
                    int ini_frq, int final_frq, double yconv_min, void *stream);
/* Copies one of the VK_PHOTO_* arrays to host memory (synchronises). */
int vk_photo_read(const vk_photo *photo, int what, double *h_out);
/* Runs vk_photo_update at the start of every vk_integrate attempt (the top
 * of the loop, op.py:813-830): the photolysis rows of the d_k that
 * vk_integrate is given are rewritten in place.  NULL detaches. */
int vk_integrator_set_photo(vk_integrator *it, vk_photo *photo, int ini_frq, int final_frq);

/* ---- warm-start MLP (no counterpart in the reference: it only seeds the
 * plain Ros2 integration, which still decides convergence) ---------------
 * BF16 tensors are passed as raw 16-bit arrays.  GEMM: C[m][n] = sum_k
 * A(m,k) B(k,n) (+ bias[n]) (+ C if accumulate) with A(m,k) at
 * A[m*sam + k*sak] and B(k,n) at B[k*sbk + n*sbn]; FP32 accumulation on
 * BF16 MFMA; with d_Yact != NULL also Yact = SiLU(C) as BF16. */
int vk_mlp_gemm(int M, int N, int K, const void *d_A, long sam, long sak,
                const void *d_B, long sbk, long sbn, float *d_C, long ldc,
                const float *d_bias, int accumulate, void *d_Yact, long ldy, void *stream);
/* dZ = dY * SiLU'(Z) (FP32 and BF16 copies) */
int vk_mlp_silu_bwd(long n, const float *d_Z, const float *d_dY, float *d_dZ, void *d_dZb,
                    void *stream);
/* masked MSE: dY = 2 w (Y - T) inv_count, loss partial sums per 256 entries */
int vk_mlp_mse(long n, const float *d_Y, const float *d_T, const float *d_W, float inv_count,
               float *d_dY, void *d_dYb, float *d_loss_part, void *stream);
/* out[n] = sum_m X[m][n] (bias gradients) */
int vk_mlp_colsum(int M, int N, const float *d_X, float *d_out, void *stream);
/* FP32 -> BF16 (round to nearest even) */
int vk_mlp_to_bf16(long n, const float *d_x, void *d_y, void *stream);
/* one Adam step (bias-corrected, decoupled weight decay wd) on FP32 master
 * weights; refreshes the BF16 copy d_wb when not NULL */
int vk_mlp_adam(long n, float *d_w, const float *d_g, float *d_m, float *d_v, float lr, float b1,
                float b2, float eps, int step, float wd, void *d_wb, void *stream);

/* ---- debug aid -----------------------------------------------------------*/
/* No counterpart in the reference (a debug aid for this library).  In the
 * bounds-checked build (libvulcan_amd_checked.so, the chemistry kernels
 * compiled with VK_CHECKED) every table index the rate, right-hand-side and
 * assembly kernels gather through is checked against its array; *site_out
 * receives 0 if none was out of range since the last reset, else the site
 * of the first (enum VkCheckSite in vk_chem.hip); reset != 0 clears it.
 * The normal build returns VK_ENOTCHECKED. */
int vk_debug_checks(int *site_out, int reset);

#ifdef __cplusplus
}
#endif
#endif
