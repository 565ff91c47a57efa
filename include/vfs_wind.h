/*
 * vfs_wind.h -- C ABI of the MI355X (gfx950) fractional-step core.
 *
 * The library keeps one single-rank PETSc-DMDA block per context: every field
 * is a device-resident "global" array of mz*my*mx points with i fastest
 * (init.c:157-160), vector fields interleaved {x,y,z} (Cmpnts,
 * variables.h:44-46).  Cells 1..m-2 are interior, 0 and m-1 ghost cells; in a
 * DA-periodic direction (ii/jj/kk_periodic) a stencil index outside [0,m)
 * wraps modulo m, as a one-rank periodic DA ghost exchange does.
 *
 * A context plays the role of the reference's UserCtx: the entry points take
 * it where the reference takes UserCtx* and act on the arrays it owns, so a
 * host driver uploads its DMDA arrays (VecGetArray) once, calls the same
 * sequence Flow_Solver calls (solvers.c:313-759) and downloads the results.
 * All entry points return 0 on success and a nonzero VFS_E* code on failure;
 * vfs_last_error() then describes the failure.  A context is not thread safe;
 * all work is ordered on the context's own HIP stream.
 */
#ifndef VFS_WIND_H
#define VFS_WIND_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VFS_OK = 0,
    VFS_EINVAL = 1,   /* bad argument or unsupported configuration */
    VFS_EHIP = 2,     /* HIP runtime error */
    VFS_ENOMEM = 3,   /* device allocation failed */
    VFS_ENOCONV = 4   /* iterative solver did not reach its tolerance */
};

/* Every option a hot-path kernel reads; the reference keeps these in
 * globals (main.c:11-355, variables.h:908-1348) and in UserCtx. */
typedef struct {
    int mx, my, mz;            /* DA sizes: IM+1, JM+1, KM+1 (init.c:157)       */
    int per_i, per_j, per_k;   /* -ii_periodic, -jj_periodic, -kk_periodic      */
    int bctype[6];             /* bcs.dat codes: i-, i+, j-, j+, k-, k+         */
    int les;                   /* -les: 0, 1 (constant) or 2 (dynamic Smagorinsky) */
    int second_order;          /* -second_order                                */
    int laplacian;             /* -laplacian                                   */
    double ren, dt, st;        /* -ren, -dt, Strouhal (init.c:59-72)           */
    double cs2;                /* value of lCs (Cs^2) used by les==1           */
    double inlet_flux;         /* -flux; < 0 when unset                        */
    double poisson_threshold;  /* poisson_threshold, main.c:124 (0.1)          */
    double flux_in_sum;        /* FluxInSum of InflowFlux (outlet balance)     */
    double max_cs;             /* -max_cs: upper clip of the dynamic lCs (0.5) */
    int dynamic_freq;          /* -dynamic_freq: steps between dynamic updates (1) */
    int i_homo_filter, j_homo_filter, k_homo_filter;  /* -i/j/k_homo_filter: average
                                  LM, MM over homogeneous directions (les.c:799-958) */
    int inletprofile;          /* -inlet (main.c:24): profile of the k- inlet (1)  */
    int pseudo_periodic;       /* -pseudo (main.c:289): the k- inlet takes the
                                  velocity of plane (6 mz) / 10 (bcs.c:29-95)    */
} vfs_params;

/* Device-resident arrays of a context. */
typedef enum {
    VFS_COOR = 0,   /* node coordinates (Cmpnts)                         */
    VFS_NVERT,      /* blanking (0 = fluid)                              */
    VFS_CSI, VFS_ETA, VFS_ZET, VFS_AJ,           /* cell metrics            */
    VFS_ICSI, VFS_IETA, VFS_IZET, VFS_IAJ,       /* i-face metrics          */
    VFS_JCSI, VFS_JETA, VFS_JZET, VFS_JAJ,       /* j-face metrics          */
    VFS_KCSI, VFS_KETA, VFS_KZET, VFS_KAJ,       /* k-face metrics          */
    VFS_CENT,       /* cell centres                                      */
    VFS_UCONT,      /* contravariant face fluxes (lUcont)                */
    VFS_UCAT,       /* Cartesian cell velocity (Ucat)                    */
    VFS_UCAT_OLD,   /* lUcat_old, the initial field used by slip walls   */
    VFS_NU_T,       /* eddy viscosity (lNu_t)                            */
    VFS_P,          /* pressure                                          */
    VFS_PHI,        /* pressure correction                               */
    VFS_DP,         /* contravariant pressure gradient (dP)              */
    VFS_RHS,        /* momentum right-hand side (Rhs)                    */
    VFS_B,          /* Poisson right-hand side                           */
    VFS_FEUL,       /* turbine body force on faces (lF_eul)              */
    VFS_UCONT_O,    /* Ucont at the start of the step (Ucont_o)          */
    VFS_RHS_O,      /* Formfunction_2 of Ucont_o (RHS_o, solvers.c:627)   */
    VFS_CS,         /* Smagorinsky coefficient lCs (= Cs^2) per point     */
    VFS_NFIELDS
} vfs_field;

typedef struct vfs_ctx vfs_ctx;

/* ---- context ---------------------------------------------------------- */
const char *vfs_last_error(void);
int vfs_ctx_create(const vfs_params *params, vfs_ctx **out);
int vfs_ctx_destroy(vfs_ctx *ctx);
int vfs_set_params(vfs_ctx *ctx, const vfs_params *params);
int vfs_get_params(const vfs_ctx *ctx, vfs_params *params);
size_t vfs_field_size(const vfs_ctx *ctx, vfs_field f);      /* doubles */
int vfs_upload(vfs_ctx *ctx, vfs_field f, const double *host);
int vfs_download(vfs_ctx *ctx, vfs_field f, double *host);
/* Device address of a field.  The context treats the field as written when
 * the pointer is handed out; a host that keeps the pointer and writes through
 * it after a later library call must call vfs_field_written before the next
 * call that reads the field, or results computed from its old contents
 * (Ucat, the Poisson operator) are reused. */
double *vfs_field_device_ptr(vfs_ctx *ctx, vfs_field f);
/* Declares that f was written outside the library (through a device pointer):
 * drops every result the context derived from it. */
int vfs_field_written(vfs_ctx *ctx, vfs_field f);
int vfs_sync(vfs_ctx *ctx);

/* ---- setup (metrics.c, bcs.c) ------------------------------------------ */
/* FormMetrics (metrics.c:12-1355): all metric fields from VFS_COOR. */
int vfs_form_metrics(vfs_ctx *ctx);
/* Calc_Inlet_Area's mean_k_area (bcs.c:137-199); stored in the context. */
int vfs_mean_k_area(vfs_ctx *ctx, double *area);

/* ---- momentum (rhs.c, momentum.c, les.c) ------------------------------- */
/* Contra2Cart (rhs.c:35 -> Contra2Cart_2, rhs.c:65-749). */
int vfs_contra2cart(vfs_ctx *ctx);
/* FormBCS (bcs.c:1684-2767) for codes 1 (wall), 10 (slip), 4 (outlet), 5 (uniform inlet,
   whose fluxes vfs_inflow_flux sets) and periodic directions. */
int vfs_form_bcs(vfs_ctx *ctx, int outflow_scale);
/* InflowFlux (bcs.c:247-626) for the inlet (code 5) on the k- face with -inlet
 * 0 and 3 (the constant Flux_in of main.c:13), 1 and 2 (uniform), 10 (power
 * law), 12 (pipe), 14 (tabulated pipe profile), 15 (round jet), 21
 * (Poiseuille) and 100 or -pseudo 1 (the inflow plane, vfs_inflow_set_plane /
 * vfs_pseudo_periodic_plane), with the held bulk flux of -inlet 14, 100 and
 * -pseudo when -flux is given: inlet fluxes and ghost velocities; FluxInSum
 * returned and kept for FormBCS.  Profile 11 (time-seeded noise) and the
 * profiles that leave the inlet velocity unset (13, 16, 20, 101) are refused. */
int vfs_inflow_flux(vfs_ctx *ctx, double *flux_in);
/* Compute_Smagorinsky_Constant_1 (les.c:75-1141) from VFS_UCAT into lCs:
 * 0 while ti < 2 on a fresh start (tistart == 0 and no restart), params.cs2
 * (0.01) for les == 1, the dynamic model for les == 2 (3-D Simpson test
 * filter, homogeneous averaging per the *_homo_filter flags, clipped to
 * [0, max_cs]).  Until it is first called lCs holds params.cs2. */
int vfs_smagorinsky_constant(vfs_ctx *ctx, int ti, int tistart, int rstart_flg);
/* Compute_eddy_viscosity_LES (les.c:1143-1361) with the current lCs. */
int vfs_eddy_viscosity(vfs_ctx *ctx);
/* Calc_k_Flux (rhs2.c:1549-1633): mean_k_flux, stored in the context. */
int vfs_calc_k_flux(vfs_ctx *ctx, double *mean_k_flux);
/* Pressure_Gradient (momentum.c:203-439): VFS_P -> VFS_DP. */
int vfs_pressure_gradient(vfs_ctx *ctx);
/* Formfunction_2 (momentum.c:454-2013): VFS_RHS += scale * F(U). */
int vfs_formfunction2(vfs_ctx *ctx, double scale);
/* RungeKutta (rhs.c:1586-1708), -imp 0. */
int vfs_runge_kutta(vfs_ctx *ctx);
/* Calc_Minimum_dt (momentum.c:110-201): the CFL-1 time step. */
int vfs_min_dt(vfs_ctx *ctx, double *dt_cfl1);

/* ---- pressure (poisson.c, poisson_hypre.c, main.c) --------------------- */
typedef struct {
    double tol;          /* relative residual ||b-A phi||/||b|| (1e-8)       */
    int restart;         /* FGMRES restart length (20)                       */
    int maxit;           /* total iteration cap                              */
    int levels;          /* most multigrid levels (6); fewer if a grid stops halving */
    int pre, post;       /* smoothing sweeps (2, 3)                          */
    int coarse_sweeps;   /* coarsest-level sweeps when it is too big for a dense solve */
    int mg_fp32;         /* 1: FP32 V-cycle inside the FP64 FGMRES           */
    int use_guess;       /* 1: start from VFS_PHI; 0: from zero              */
    int amg;             /* 1: algebraic multigrid preconditioner (the role of
                            HYPRE BoomerAMG, poisson_hypre.c:34-97) in place
                            of the geometric V-cycle; FP64, mg_fp32 ignored   */
    double amg_thresh;   /* -amg_thresh: strength threshold (0.75)           */
    int deep_level;      /* first V-cycle level that smooths deep_pre and
                            deep_post sweeps instead of pre and post;
                            0 (the default): every level takes pre and post */
    int deep_pre, deep_post;
} vfs_solve_opts;

typedef struct {
    int iterations;      /* FGMRES iterations (= V-cycles)                   */
    double relres;       /* final true relative residual                     */
    double ms;           /* wall time of the solve                           */
    int fallback;        /* learned stage rejected (see vfs_poisson_solve_learned) */
} vfs_solve_stats;

/* PoissonRHS2_hypre + Remove_Nullspace_Scale (poisson_hypre.c:118-407): VFS_UCONT -> VFS_B. */
int vfs_poisson_rhs(vfs_ctx *ctx);
/* y = A x with the PoissonLHSNew operator (poisson.c:977-1688); device arrays. */
int vfs_poisson_apply(vfs_ctx *ctx, const double *x_dev, double *y_dev);
/* Solve A phi = b (PoissonSolver_Hypre's role, poisson_hypre.c:506-636):
 * VFS_B -> VFS_PHI, plain-mean null-space removal (poisson_hypre.c:608). */
int vfs_poisson_solve(vfs_ctx *ctx, const vfs_solve_opts *opts, vfs_solve_stats *stats);
/* UpdatePressure with Convert_Phi2_Phi's ghost handling (poisson.c:98-186, 3137-3297). */
int vfs_update_pressure(vfs_ctx *ctx);
/* Projection (poisson.c:2700-3066), including its closing Contra2Cart. */
int vfs_projection(vfs_ctx *ctx);
/* Divergence (main.c:936-1074): max |div U| * aj over fluid cells. */
int vfs_divergence(vfs_ctx *ctx, double *maxdiv);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops): vfs_divergence on
 * device arrays of the DA layout described by params (ucont: 3 doubles per
 * point, aj and nvert: 1), enqueued on `stream` (a hipStream_t; NULL = the
 * null stream) with the result written to the device double *maxdiv and no
 * host synchronisation.  The per-block partials live in a fixed 8 KiB device
 * array per GPU (1024 doubles), which calls on one GPU share: calls that
 * overlap on one GPU race.  Grids of 2^31 points or more are refused. */
int vfs_divergence_noalloc(const vfs_params *params, const double *ucont, const double *aj, const double *nvert,
                           double *maxdiv, void *stream);

/* ---- implicit momentum, -imp 4 (implicitsolver.c, momentum.c) ---------- */
typedef struct {
    double rtol;         /* -imp_tol: SNES and GMRES relative tolerance (1e-4)  */
    int restart;         /* GMRES restart length (30)                         */
    int max_newton;      /* Newton iteration cap (50)                         */
    int max_linear;      /* GMRES iteration cap per Newton step (1000)        */
    int use_feul;        /* add VFS_FEUL (rotor / nacelle / IB delta models)  */
} vfs_newton_opts;

typedef struct {
    int newton_iterations;
    int linear_iterations;
    int function_evals;
    double fnorm0, fnorm;  /* ||R|| at the start and the end                 */
    double ms;
} vfs_newton_stats;

/* FormFunction_SNES (momentum.c:2237-2336), time_coeff() == 1:
 * r = -(u - Ucont_o)/dt + 0.5 F(u) + 0.5 RHS_o - dP (+ F_eul); device arrays of
 * 3 doubles per point.  Leaves VFS_UCONT/VFS_UCAT at the wall-zeroed u; VFS_RHS
 * is scratch (unspecified afterwards). */
int vfs_formfunction_snes(vfs_ctx *ctx, const double *u_dev, double *r_dev, int use_feul);
/* Ucont_o = Ucont and RHS_o = Formfunction_2(Ucont) (solvers.c:627-628). */
int vfs_store_old_state(vfs_ctx *ctx);
/* Implicit_MatrixFree (implicitsolver.c:4203-4446): solve R(Ucont) = 0 with
 * matrix-free Newton-Krylov, then the post-solve boundary fix and FormBCS. */
int vfs_implicit_matrixfree(vfs_ctx *ctx, const vfs_newton_opts *opts, vfs_newton_stats *stats);

/* ---- one step ------------------------------------------------------------ */
typedef struct {
    int poisson_iterations;
    double poisson_relres;
    double maxdiv;
    double ms_momentum, ms_poisson, ms_total;
    int newton_iterations, linear_iterations;  /* implicit step only */
} vfs_step_stats;

/* The time-step index the step drivers pass to Compute_Smagorinsky_Constant_1
 * (the reference's globals ti, tistart, rstart_flg).  Each step calls it when
 * ti % dynamic_freq == 0 or ti == tistart, then advances ti by one.  A new
 * context starts at ti = tistart = 0 with rstart_flg = 1 (the first-steps
 * zero coefficient off). */
int vfs_set_time_index(vfs_ctx *ctx, int ti, int tistart, int rstart_flg);
/* Flow_Solver (solvers.c:313-759) for -imp 0: k-flux, LES, Pressure_Gradient,
 * RungeKutta, Poisson, UpdatePressure, Projection, Divergence. */
int vfs_fractional_step(vfs_ctx *ctx, const vfs_solve_opts *opts, vfs_step_stats *stats);
/* The same step with the -imp 4 momentum solve (solvers.c:626-630). */
int vfs_fractional_step_cn(vfs_ctx *ctx, const vfs_newton_opts *nopts, const vfs_solve_opts *opts,
                           vfs_step_stats *stats);

/* ---- actuator line (rotor_model.c) ---------------------------------------
 * One turbine with line elements (rotor_model 3 or 6).  The caller passes
 * what ACL_read_ucd / airfoil_ACL and the Turbine.inp reader (main.c:2040-2101)
 * produce: blade nodes in grid units with the hub offset applied, the rotor
 * axis and hub centre, and per foil type the radial twist/chord table and the
 * CL/CD polars (angles in degrees, lengths divided by reflength_wt). */
typedef struct {
    double nx, ny, nz;      /* rotor axis; normalised here (main.c:2051-2054)  */
    double xc, yc, zc;      /* hub centre                                       */
    double angvel;          /* angular velocity (-fixturbineangvel: angvel_fixed) */
    double pitch;           /* pitch[0], degrees                                */
    double r_nacelle;       /* r_nacelle / reflength_wt                         */
    double refangle;        /* refangle_AL in degrees (rotor_model 6)           */
    int rotor_model;        /* 3 or 6                                           */
    int deltafunc;          /* -deltafunc: 0 dfunc_2h, 7 dfunc_exp, else s4h    */
    double halfwidth;       /* -halfwidth_dfunc (deltafunc 7)                   */
    /* Calc_F_lagr_ACL's optional corrections (rotor_model.c:2488-2598), all off
     * when zero; the reference's defaults (main.c:161, 224-238) are num_blade 3,
     * a_shen 0.125, b_shen 21, c_shen 0.1, correction_shen1 1, c0_cl 1. */
    int shen, prandtl, shen1;       /* -Shen_AL, -Prandtl_AL, -Shen1_AL tip-loss models */
    double r_rotor;                 /* fsi.r_rotor / reflength_wt (R of the models)      */
    double tsr;                     /* ibm.Tipspeedratio (vfs_alm_turbine_angvel sets it) */
    int num_blade;                  /* -num_blade (B of the models)                      */
    double a_shen, b_shen, c_shen;  /* -a_shen -b_shen -c_shen                            */
    double correction_shen1;        /* -correction_ALShen1                                */
    int correction3d_cl;            /* -correction3D_CL (inboard of 0.75 R)               */
    double c0_cl;                   /* -c0_CL                                             */
    int smoothforce;                /* -smoothforce_AL: smoothforce + 1 passes            */
} vfs_turbine;

/* Calc_turbineangvel's inputs (rotor_model.c:5519-5601; main.c globals and the
 * Turbine.inp control columns). */
typedef struct {
    int mode;                       /* 0 -fixturbineangvel, 1 -FixTipSpeedRatio, 2 -turbinetorquecontrol */
    double angvel_fixed;            /* mode 0: fsi.angvel_fixed                          */
    double tsr;                     /* mode 1: ibm.Tipspeedratio                         */
    double r_rotor;                 /* fsi.r_rotor (dimensional)                         */
    double reflength_wt, refvel_wt, refvel_cfd, rho_air;
    double J_rotation, CP_max, TSR_max, X_control;          /* mode 2                 */
    int fixturbinegeneratortorque;  /* mode 2: keep torque_generator                    */
    double torque_generator;        /* mode 2 with fixturbinegeneratortorque            */
} vfs_turbine_control;

typedef struct {
    int n_ac; const double *r_ac, *twist, *chord;   /* FOILnn                      */
    int n_cl; const double *ang_cl, *cl;            /* CLnn                        */
    int n_cd; const double *ang_cd, *cd;            /* CDnn                        */
} vfs_foil;

typedef enum {
    VFS_ALM_NODES = 0,  /* 3 per node: blade node positions                    */
    VFS_ALM_NODE_VEL,   /* 3 per node: node velocities (rotor_Rot)             */
    VFS_ALM_CENT,       /* 3 per element: element centres                      */
    VFS_ALM_REF_CENT,   /* 3 per element: reference-line centres (model 6)     */
    VFS_ALM_DA,         /* element lengths                                      */
    VFS_ALM_TWIST, VFS_ALM_CHORD,
    VFS_ALM_U_LAGR,     /* 3 per element (Calc_U_lagr)                          */
    VFS_ALM_F_LAGR,     /* 3 per element (Calc_F_lagr_ACL)                      */
    VFS_ALM_AOA, VFS_ALM_CL, VFS_ALM_CD,
    VFS_ALM_BOX,        /* 6 per element: i_min, i_max, j_min, j_max, k_min, k_max */
    VFS_ALM_REF_BOX     /* the same for the reference line                      */
} vfs_alm_array;

typedef struct vfs_alm vfs_alm;

/* ACL_read_ucd + airfoil_ACL state (rotor_model.c:659-1318, 2017-2137). */
int vfs_alm_create(vfs_ctx *ctx, const vfs_turbine *t, int num_blade, int nv_blade, const double *nodes, int nfoil,
                   const vfs_foil *foils, vfs_alm **out);
int vfs_alm_destroy(vfs_alm *alm);
int vfs_alm_counts(const vfs_alm *alm, int *n_elements, int *n_nodes);
/* rotor_Rot, dimension 1 (rotor_model.c:4741-4867), and refAL_Rot for model 6 (rotor_model.c:5034-5088). */
int vfs_alm_rotate(vfs_alm *alm, double dt);
/* Pre_process (rotor_model.c:1849-2011) of the blade (and reference) elements. */
int vfs_alm_preprocess(vfs_alm *alm);
/* Calc_U_lagr (rotor_model.c:2937-3146) from VFS_UCAT. */
int vfs_alm_u_lagr(vfs_alm *alm);
/* Calc_F_lagr_ACL (rotor_model.c:2370-2681) with the corrections vfs_turbine selects
 * (tip-loss models, 3D lift correction, force smoothing) and the nacelle zeroing. */
int vfs_alm_f_lagr(vfs_alm *alm);
/* Calc_forces_ACL (rotor_model.c:4148-4180): out[8] = {A_Sum, F_Sum, P_Sum, U_Sum / A_Sum,
 * Fx_Sum, Fy_Sum, Fz_Sum, M_Sum} of the current F_lagr and U_lagr; keeps Torque_aero =
 * -M_Sum for vfs_alm_turbine_angvel. */
int vfs_alm_loads(vfs_alm *alm, double *out);
/* Calc_turbineangvel (rotor_model.c:5519-5601) for time step ti with the reference
 * velocity uref (vfs_alm_uref): updates the angular velocity the next rotation uses and
 * the tip speed ratio of the tip-loss models; returns both (either pointer may be null).
 * A restart from the turbine-rotation file is not read here (set vfs_turbine.angvel). */
int vfs_alm_turbine_angvel(vfs_alm *alm, const vfs_turbine_control *tc, double dt, double uref, int ti, int tistart,
                           double *angvel, double *tsr);
/* VecSet(F_eul, 0) and Calc_F_eul (rotor_model.c:3668-3960) into VFS_FEUL. */
int vfs_alm_f_eul(vfs_alm *alm);
/* The rotor_model 6 block of Flow_Solver (solvers.c:535-563) at the current angular
 * velocity: rotate, Pre_process, U_lagr, F_lagr, F_eul (Uref_ACL, Calc_turbineangvel and
 * Calc_forces_ACL are the separate calls above, made where the host needs them). */
int vfs_alm_step(vfs_alm *alm, double dt);
/* Copy an element or node array to the host. */
int vfs_alm_get(vfs_alm *alm, vfs_alm_array what, double *host);
/* Uref_ACL (rotor_model.c:4373-4396): reference disk elements (Urefdata000) and its mean axial velocity. */
int vfs_alm_set_disk(vfs_alm *alm, int n, const double *cent, const double *dA);
int vfs_alm_uref(vfs_alm *alm, double *uref);

/* ---- learned Poisson stage (no counterpart in the reference) ------------
 * A 3D convolutional U-Net (BF16 on the matrix cores, FP32 accumulation and
 * master weights) on the interior cells of the context's grid, trained on
 * (Poisson right-hand side, solution) pairs the solver produces.  It gives
 * vfs_poisson_solve_learned its initial guess and one preconditioner stage. */
typedef struct vfs_unet vfs_unet;
/* base channels (16..64, multiple of 16), levels 2..5; interior cell counts
 * divisible by 2^(levels-1). */
int vfs_unet_create(vfs_ctx *ctx, int base_channels, int levels, unsigned long long seed, vfs_unet **out);
int vfs_unet_destroy(vfs_unet *net);
/* phi ~ out_scale * rms(b) * net(b / rms(b)); out_scale from the training data. */
int vfs_unet_set_output_scale(vfs_unet *net, double out_scale);
/* y = net(b) on DA-layout device arrays (ghost points of y set to 0). */
int vfs_unet_forward(vfs_unet *net, const double *b_dev, double *y_dev);
/* forward + mean-squared loss against phi + backward; gradients kept. */
int vfs_unet_gradients(vfs_unet *net, const double *b_dev, const double *phi_dev, double *loss);
/* vfs_unet_gradients + one Adam step with learning rate lr. */
int vfs_unet_train_step(vfs_unet *net, const double *b_dev, const double *phi_dev, double lr, double *loss);
long long vfs_unet_param_count(vfs_unet *net);
/* which = 0: FP32 master weights, 1: last gradients; flattened layer by layer. */
int vfs_unet_get_params(vfs_unet *net, int which, float *host);
int vfs_unet_set_params(vfs_unet *net, const float *host);
/* vfs_poisson_solve with the network: x0 = guess + alpha * net(b - A guess)
 * (alpha by least squares on the residual), then FGMRES with the V-cycle
 * (environment VFS_LEARNED_DIRECTION=1: the first direction preconditioned by
 * the network); when after three iterations the residual has not dropped
 * below half of its starting value the solve restarts from the guess with the
 * plain V-cycle (stats->fallback = 1). */
int vfs_poisson_solve_learned(vfs_ctx *ctx, vfs_unet *net, const vfs_solve_opts *opts, vfs_solve_stats *stats);

/* ---- inflow planes (-inlet 100, -pseudo 1) ----------------------------- */
/* The reference's user->ucont_plane: my x mx Cmpnts of Cartesian velocity
 * that InflowFlux copies into the k = 0 ghost cells for -inlet 100 and the
 * pseudo-periodic inlet (bcs.c:379-390).  vfs_inflow_set_plane uploads a
 * plane (read_inflow_section's result, rhs2.c:766-862); vfs_inflow_get_plane
 * downloads the current one (for save_inflow_section, rhs2.c:716-764). */
int vfs_inflow_set_plane(vfs_ctx *ctx, const double *host);
int vfs_inflow_get_plane(vfs_ctx *ctx, double *host);
/* pseudo_periodic_BC (bcs.c:29-95): the plane := Ucat of plane k = (6 mz) / 10,
 * on the device.  The step drivers call it first when -pseudo is set
 * (solvers.c:355-358). */
int vfs_pseudo_periodic_plane(vfs_ctx *ctx);
/* save_inflow_section (rhs2.c:716-764), host side: appends the plane to
 * dir/inflow_%06d_dt=%g.dat (the file of time index ti; the file is started
 * afresh when ti - save_inflow_minus == tistart or is 1 modulo the period). */
int vfs_inflow_write_section(const char *dir, int ti, int tistart, int save_inflow_period, int save_inflow_minus,
                             double dt, int mx, int my, const double *plane);
/* read_inflow_section (rhs2.c:766-862), host side: a reader keeps the open
 * file of path_inflow/inflow/inflow_%06d_dt=%g.dat between calls and returns
 * the plane of time index ti with the reference's recycling (-recycle period),
 * file period (-save_inflow_period) and offset (-save_inflow_minus) rules;
 * the first call must be for ti = tistart, later calls for ti + 1, ti + 2, ...
 * As in the reference, the plane is zero at ti = 1 of a fresh start, and
 * values past the end of a file (which a restart not at 1 modulo the file
 * period reaches, since files are switched only after tistart + 90) are zero. */
typedef struct vfs_inflow_reader vfs_inflow_reader;
int vfs_inflow_reader_open(const char *path_inflow, double dt_inflow, int save_inflow_period, int save_inflow_minus,
                           int recycle_period, int tistart, int mx, int my, vfs_inflow_reader **out);
int vfs_inflow_reader_next(vfs_inflow_reader *r, int ti, double *plane);
int vfs_inflow_reader_close(vfs_inflow_reader *r);

/* ---- file formats (host side) ----------------------------------------- */
/* PETSc binary Vec files as VecView writes them and VecLoadIntoVector reads
 * them (PETSc 3.1, the reference's PetscViewerBinaryOpen calls in
 * main.c:376-934): big-endian int32 class cookie 1211214, int32 length n,
 * n big-endian doubles; a DMDA vector in natural (= one-rank global) order.
 * vfs_vec_read fails unless the file holds exactly n values. */
int vfs_vec_write(const char *path, const double *host, long long n);
int vfs_vec_read(const char *path, double *host, long long n);
int vfs_vec_length(const char *path, long long *n);
/* Ucont_P_Binary_Output (main.c:684-770, -qcr 0, no averaging): writes
 * dir/vfield%06d_%1d.dat (Ucont), ufield (Ucat), pfield (P) and nvfield
 * (Nvert) for time index ti and block number block. */
int vfs_checkpoint_write(vfs_ctx *ctx, const char *dir, int ti, int block);
/* Ucont_P_Binary_Input (main.c:376-426, no averaging): loads Ucont, P and Ucat,
 * then Ucont_o = Ucont and Ucat_old = Ucat; the blanking is zeroed unless
 * immersed != 0, in which case nvfield is loaded into VFS_NVERT (the
 * immersed-boundary search that recomputes it is not part of this library). */
int vfs_checkpoint_read(vfs_ctx *ctx, const char *dir, int ti, int block, int immersed);
/* Grid files of MG_Initial (init.c:258-394): grid.dat as text (block count,
 * then per block "IM JM KM" and all x, all y, all z values in (k, j, i)
 * order), grid.dat binary (-binary_input: the same as native int / double),
 * xyz.dat (-xyz_input: "IM JM KM", then IM lines whose first column is x,
 * JM lines whose second is y, KM lines whose third is z). */
enum { VFS_GRID_TEXT = 0, VFS_GRID_BINARY = 1, VFS_GRID_XYZ = 2 };
int vfs_grid_dims(const char *path, int format, int block, int *IM, int *JM, int *KM);
/* Node coordinates of one block into a DA-layout coordinate array (mx = IM+1,
 * my = JM+1, mz = KM+1; nodes at indices 0..IM-1 etc., the last plane zero),
 * each value divided by cl and multiplied by L_dim as the reference does. */
int vfs_grid_read(const char *path, int format, int block, double cl, double L_dim, int mx, int my, int mz,
                  double *coor);

/* ---- algebraic multigrid, host side (diagnostics) ---------------------- */
/* The setup of the AMG preconditioner (the role of HYPRE BoomerAMG,
 * poisson_hypre.c:34-97: PMIS coarsening, FF1 interpolation, strength
 * threshold theta = -amg_thresh, Galerkin coarse operators) on a CSR matrix
 * given by the host; the solver builds the same hierarchy from its own
 * operator (vfs_solve_opts.amg).  Levels are added while the last has more
 * than max_coarse rows (512 when max_coarse <= 0) and fewer than max_levels
 * exist. */
typedef struct vfs_amg_host vfs_amg_host;
int vfs_amg_host_build(int n, const int *ptr, const int *col, const double *val, double theta, int max_levels,
                       int max_coarse, vfs_amg_host **out);
int vfs_amg_host_levels(const vfs_amg_host *h, int *nlevels);
/* which = 0: level's operator, 1: its interpolation to the next level; with
 * null arrays only the sizes are returned. */
int vfs_amg_host_matrix(const vfs_amg_host *h, int level, int which, int *rows, int *cols, long long *nnz, int *ptr,
                        int *col, double *val);
/* C/F splitting of a level: 1 = C-point, -1 = F-point. */
int vfs_amg_host_splitting(const vfs_amg_host *h, int level, int *cf);
/* The multicolour Gauss-Seidel smoother's colouring of a level (the GPU form
 * of HYPRE's default hybrid Gauss-Seidel relaxation): rows grouped by colour
 * (n entries) and colour offsets (ncolours + 1 entries); with null arrays
 * only *ncolours is returned. */
int vfs_amg_host_colouring(const vfs_amg_host *h, int level, int *ncolours, int *rows, int *offsets);
int vfs_amg_host_free(vfs_amg_host *h);

/* ---- measurement (no counterpart in the reference) -------------------- */
/* Enable (1) or disable (0) HIP-event timing of the hot kernels' launches on
 * the context stream and clear earlier records.  For benchmarks only. */
int vfs_profile(vfs_ctx *ctx, int enable);
/* The same, recording only the kernel ids whose bit (1u << id) is set in
 * kernel_mask (0 disables); nested regions of unselected ids record nothing.
 * Fewer event pairs on the stream perturb the timed run less. */
int vfs_profile_select(vfs_ctx *ctx, unsigned int kernel_mask);
/* Total time and launch count recorded for kernel id 0..13 (finest level for
 * the smoother): face_fluxes, cell_rhs, line_rhs_fp32, line_rhs_fp64,
 * poisson_apply_fp64, contra2cart, pressure_gradient, eddy_viscosity,
 * line_solve_fp32, line_solve_fp64, cell_fp, formfunction2 (one whole
 * Formfunction_2 call: face pass, its periodic fix and both cell passes),
 * line_rhs_compact_fp32 (the line right-hand side of the compact FP32
 * smoother, which a level takes when its cross couplings are weak) and
 * line_relax_compact_fp32 (the compact smoother's one-pass form: right-hand
 * side and line solve of one colour in one launch). */
int vfs_profile_read(vfs_ctx *ctx, int kernel, double *total_ms, long long *count, const char **name);

#ifdef __cplusplus
}
#endif
#endif
