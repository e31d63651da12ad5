/*
 * maos_tomo.h -- C ABI of the MI355X tomography solver.
 *
 * Two layers, both plain C (no torch or HIP types in any signature):
 *
 *  1. Host setup (FP64, C++ inside): builds the reconstruction geometry and
 *     operators exactly as MAOS's CPU setup does before it calls
 *     gpu_setup_recon():  ploc/xloc grids, saloc + subaperture area, the GP
 *     average-gradient operator, the NEA weighting W (saneai), the TT
 *     projector PTT, the Laplacian regulariser L2 and the FDPCG block
 *     preconditioner.  Result: an mt_recon_t, a POD view of plain arrays that
 *     plays the role of MAOS's RECON_T on this path.
 *
 *  2. GPU solve (HIP, gfx950): consumes an mt_recon_t (from (1) or from a
 *     caller that built the same arrays itself) and runs the FDPCG-PCG
 *     tomography solve.  These calls replace the reference's C entry points
 *     gpu_setup_recon / gpu_update_recon_cn2 / gpu_recon_reset / gpu_tomo
 *     (cuda/gpu.h:49-56, cuda/recon.cu:477-582) and the C++ operator entry
 *     points cutomo_grid::R/L (cuda/tomo.cu:607-677), cufdpcg_t::P
 *     (cuda/fdpcg.cu:168-226) and gpu_pcg (cuda/pcg.cu:87-242).
 *
 * Every function returns an int status (MT_OK == 0, negative on error) and
 * never aborts the process; mt_last_error() returns a message for the last
 * failure on the calling thread.
 *
 * Vector layouts (identical to MAOS dcell layouts):
 *   gradient vector : per WFS w (in order, skipped WFS have 0 entries),
 *                     2*nsa[w] doubles: all x slopes then all y slopes.
 *   layer vector    : per layer l, xnx[l]*xny[l] values, index ix+iy*xnx[l]
 *                     (the square xloc of tomo.square=1).
 */
#ifndef MAOS_TOMO_H
#define MAOS_TOMO_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* status codes                                                        */
/* ------------------------------------------------------------------ */
#define MT_OK          0
#define MT_EINVAL     -1   /* invalid argument or dimension            */
#define MT_ENOMEM     -2   /* host or device allocation failed         */
#define MT_EHIP       -3   /* HIP runtime error                        */
#define MT_ENOTFINITE -4   /* non-finite input (NaN/Inf gradients)     */
#define MT_ESTATE     -5   /* call not valid in the current state      */
#define MT_EFULL      -6   /* admission refused: over memory budget    */
#define MT_NOTCONV     1   /* positive flag: solve ran but residual > 1 */

const char *mt_last_error(void);

/* ------------------------------------------------------------------ */
/* 1. host setup                                                       */
/* ------------------------------------------------------------------ */

/* One wavefront-sensor type (MAOS powfs.* keys on this path). */
typedef struct {
    double dsa;       /* powfs.dsa: subaperture size (m)                    */
    double dx;        /* powfs.dx: OPD sampling inside a subaperture (m)    */
    double saat;      /* powfs.saat: relative-area threshold                */
    double nearecon;  /* powfs.nearecon: NEA for reconstruction (mas)       */
    double hs;        /* powfs.hs: guide star height (m); INFINITY for NGS  */
    int    trs;       /* powfs.trs: remove tip/tilt from this WFS' slopes   */
    int    lo;        /* powfs.lo: low-order WFS (not in tomography)        */
    int    skip;      /* powfs.skip                                         */
    int    dtrat;     /* powfs.dtrat                                        */
} mt_powfs_t;

/* One wavefront sensor (MAOS wfs.* keys). */
typedef struct {
    double thetax, thetay;  /* direction (radian)            */
    int    powfs;           /* index into mt_parms_t.powfs    */
} mt_wfs_t;

/* The .conf parameters the tomography path reads (PARMS_T subset). */
typedef struct {
    double aper_d, aper_din;       /* aper.d = [d din]                       */
    int    npsr;                   /* number of reconstructed layers         */
    const double *ht;              /* atmr.ht (m)                            */
    const double *wt;              /* atmr.wt                                */
    const int    *os;              /* atmr.os                                */
    double atmr_dx;                /* atmr.dx; 0 = smallest high-order dsa   */
    double r0z;                    /* atmr.r0z (m at 500 nm, zenith)         */
    double L0;                     /* atmr.L0 (m)                            */
    double za;                     /* sim.za (radian)                        */
    int    tomo_pos;               /* tomo.pos                               */
    int    tomo_guard;             /* tomo.guard                             */
    int    tomo_cone;              /* tomo.cone                              */
    double tomo_cxxscale;          /* tomo.cxxscale                          */
    int    tomo_precond;           /* tomo.precond: 0 none, 1 FDPCG          */
    int    recon_split;            /* recon.split                            */
    int    tomo_splitlrt;          /* tomo.splitlrt                          */
    int    npowfs;
    const mt_powfs_t *powfs;
    int    nwfs;
    const mt_wfs_t *wfs;
    int    ndir;                   /* extra beam directions (evl + fit)      */
    const double *dir_thetax;      /* radian                                 */
    const double *dir_thetay;      /* radian                                 */
    const double *dir_hs;          /* m; INFINITY for science                */
    double tomoscale;              /* TOMOSCALE (MAOS default 1e-12)         */
} mt_parms_t;

/*
 * POD view of the reconstruction operators (the RECON_T fields that
 * gpu_setup_recon reads, cuda/recon.cu:477-488, cuda/tomo.cu:174-312).
 * All pointers are owned by whoever produced the struct.
 */
typedef struct {
    int    npsr;
    const long   *xnx, *xny;       /* layer grid sizes (recon->xmap)         */
    const double *xox, *xoy, *xdx; /* layer grid origin / spacing (m)        */
    const double *ht;              /* layer heights (m)                      */
    long   pnx, pny;               /* ploc square grid (recon->pmap)         */
    double pox, poy, pdx;
    int    nwfs;
    const double *thetax, *thetay, *hs;   /* per WFS                         */
    const int    *skip;            /* per WFS: 1 = not in tomography         */
    const long   *nsa;             /* per WFS subaperture count              */
    const double *dsa;             /* per WFS subaperture size               */
    const double * const *saloc;   /* per WFS: locx[nsa] then locy[nsa]      */
    const double * const *saa;     /* per WFS: normalized subaperture area   */
    /* GP per WFS in MAOS dsp (CSC) layout: 2*nsa rows x pnx*pny columns      */
    const long   * const *gp_p;    /* ncol+1                                 */
    const long   * const *gp_i;    /* row indices                            */
    const double * const *gp_x;    /* values                                 */
    int    pos;                    /* tomo.pos (GP stencil is (pos+1)^2)     */
    const double * const *neai;    /* per WFS 3*nsa: [xx yy xy] per subap    */
    const int    *trs;             /* per WFS: RHS removes TT                */
    const double * const *ptt;     /* per WFS (trs): 2 x 2nsa column-major   */
    int    ptt_lhs;                /* 1: LHS also removes TT (!split||lrt)   */
    const double *l2c;             /* per layer: (0.25*laplacian_coef)^2
                                      * cxxscale * TOMOSCALE                 */
    /* FDPCG */
    int    precond;                /* 0 none, 1 FDPCG                         */
    int    fd_bs;                  /* block size = sum os^2                   */
    long   fd_nb;                  /* number of blocks (full spectrum)        */
    long   fd_nbhalf;              /* blocks of the half spectrum             */
    int    fd_scale;               /* 1: apply 1/sqrt(N) per FFT (needscale)  */
    const double *fd_Minv;         /* fd_nb blocks, bs*bs complex (re,im),
                                      column-major; first fd_nbhalf blocks are
                                      the half-spectrum ones                  */
    const long   *fd_perm;         /* fd_nb*bs, full spectrum, FFT order      */
    const long   *fd_permhf;       /* fd_nbhalf*bs, half spectrum; negative =
                                      conjugate of element -v                 */
    double tomoscale;
} mt_recon_t;

typedef struct mt_setup mt_setup_t;

/* Build geometry + operators + FDPCG (setup_recon_prep, setup_recon_tomo,
 * fdpcg_prepare: maos/setup_recon_prep.c:1043, maos/setup_recon.c:1212,
 * maos/fdpcg.c:331). */
int  mt_setup_create(const mt_parms_t *parms, mt_setup_t **out);
/* Re-weight layers / r0 and rebuild L2 and FDPCG (setup_recon_tomo_update,
 * maos/setup_recon.c:747). */
int  mt_setup_update_cn2(mt_setup_t *s, const double *wt, double r0z);
/* Fill a POD view; pointers stay valid until the next update or free. */
int  mt_setup_recon(const mt_setup_t *s, mt_recon_t *out);
void mt_setup_free(mt_setup_t *s);
/* Replace the GP operators with MAOS-saved ones (load.GP,
 * maos/setup_recon_prep.c:485-505): a cell (mt_bin_read of MAOS's GP.bin)
 * of one sparse 2 nsa x nploc matrix per WFS or per powfs; WFS of one
 * powfs must carry the same matrix.  Call before gpu_tomo_setup. */
typedef struct mt_bin mt_bin_t;
int  mt_setup_load_gp(mt_setup_t *s, const mt_bin_t *gp);
/* Pre-inversion FDPCG blocks (Mhat, same layout as fd_Minv), for checking. */
int  mt_setup_fdpcg_mhat(const mt_setup_t *s, const double **mhat);
/* Recon-grid sizes: nx (sum of layer sizes), ng (sum of gradient sizes). */
int  mt_setup_sizes(const mt_setup_t *s, long *nx, long *ng);

/* Lower-level FDPCG construction from explicit inputs: the pieces that
 * fdpcg_prepare (maos/fdpcg.c:331-580) reads from RECON_T / POWFS_T.
 * mt_setup_create calls the same code.  Outputs are allocated here and
 * released with mt_fdpcg_free. */
typedef struct {
    int    nps;                    /* layers                                  */
    const long   *nx, *ny, *os;    /* layer grid sizes and oversampling       */
    int    pos;                    /* tomo.pos                                */
    double dxp;                    /* ploc spacing                            */
    const double *ht;              /* layer heights                           */
    const double *l2cf;            /* laplacian_coef * sqrt(cxxscale*TOMOSCALE) */
    long   nsa;                    /* subapertures of the high-order powfs    */
    double dsa;
    const double *salocx, *salocy, *saa;
    int    nwfs;                   /* WFS of the high-order powfs             */
    const double *thetax, *thetay; /* radian                                  */
    const double *neam;            /* mean NEA / sqrt(TOMOSCALE) (recon->neam) */
} mt_fdpcg_in_t;

typedef struct {
    int    bs;
    long   nb, nbhalf;
    int    scale;
    double *Mhat;                  /* nb*bs*bs complex, before inversion      */
    double *Minv;                  /* nb*bs*bs complex                        */
    long   *perm;                  /* nb*bs                                   */
    long   *permhf;                /* nbhalf*bs                               */
} mt_fdpcg_out_t;

int  mt_fdpcg_prepare(const mt_fdpcg_in_t *in, mt_fdpcg_out_t *out);
void mt_fdpcg_free(mt_fdpcg_out_t *out);

/* ------------------------------------------------------------------ */
/* 2. GPU solve                                                        */
/* ------------------------------------------------------------------ */

typedef struct gpu_tomo gpu_tomo_t;

/* Upload operators and allocate device state for `nbatch` independent
 * systems sharing this geometry (nbatch = 1 is the reference's one solve).
 * Replaces gpu_setup_recon (cuda/recon.cu:477). */
int  gpu_tomo_setup(const mt_recon_t *recon, int device, int nbatch,
                    int maxit, int warm, gpu_tomo_t **out);
/* Re-upload L2 coefficients and FDPCG blocks after a Cn2 update (both
 * copies of the blocks: the LDS column kernel's and the register column
 * kernel's), in place, so batches and captured graphs stay valid.  Waits for
 * the plan's own streams only; no solve of this plan may be in flight on a
 * caller's stream.  Replaces gpu_update_recon_cn2 (cuda/recon.cu:526,
 * curecon_t::update_cn2 cuda/recon.cu:231-250). */
int  gpu_tomo_update_cn2(gpu_tomo_t *g, const mt_recon_t *recon);
/* Zero the warm-start state. Replaces gpu_recon_reset (cuda/recon.cu:538). */
int  gpu_tomo_reset(gpu_tomo_t *g);
void gpu_tomo_free(gpu_tomo_t *g);

/* Full solve with host buffers (double, MAOS dcell layout), one system in
 * slot isys of the plan: grad -> RHS -> PCG (maxit, warm start from the
 * slot's resident solution) -> x.  cgres receives sqrt(r'r / b'b).
 * Replaces gpu_tomo (cuda/recon.cu:564). */
int  gpu_tomo_solve(gpu_tomo_t *g, int isys, const double *grad,
                    double *x, double *cgres);

/* The same for nsys <= nbatch systems in one batched solve: grad is nsys
 * stacked gradient vectors (ng each), x nsys stacked layer vectors (nx
 * each, in: ignored, the warm start is the resident solution of each slot),
 * cgres nsys residuals.  Returns MT_NOTCONV if any residual exceeds 1.
 * Batch form of gpu_tomo (cuda/recon.cu:564) for independent loops. */
int  gpu_tomo_solve_batch(gpu_tomo_t *g, int nsys, const double *grad,
                          double *x, double *cgres);

/* Batched solve with device buffers already resident in HBM (FP32):
 * d_grad = nbatch stacked gradient vectors, d_x = nbatch stacked layer
 * vectors (in: warm start when warm=1; out: solution).  Enqueued on
 * `stream` (a hipStream_t, 0 = the plan's own stream); no host sync.
 * mode: 0 = one launch per kernel (plain), 1 = replay captured hipGraphs:
 * the nbatch systems as k flights over consecutive system ranges (k from
 * gpu_tomo_tune_set / gpu_tomo_autotune, default 1), each graph on a stream
 * of the plan, forked from and joined back to `stream` by events, so their
 * kernels can run concurrently (when the streams' hardware queues differ);
 * results do not depend on k. */
int  gpu_tomo_solve_dev(gpu_tomo_t *g, const float *d_grad, float *d_x,
                        int nbatch, int mode, void *stream);
/* gpu_tomo_solve_dev (mode 0, plain launches) that allocates nothing on the
 * device or the host at call time, for callers that cannot allocate: code
 * under hipGraph stream capture, tight serving loops, harnesses that forbid
 * hipMalloc.  d_grad (nsys stacked FP32 gradient vectors in the active-WFS
 * layout of gpu_tomo_solve_dev) is read, d_x (nsys stacked layer vectors:
 * warm start in when the plan has warm=1, solution out) is written in place,
 * and d_res (device, nsys doubles, may be NULL) receives the residuals.  The
 * solve's other state (r, z, p, A p, the WFS fields, both FDPCG spectra,
 * partial sums, scalars) lives in one library-owned device array of
 * 16 MiB per GPU, fixed at load (the 30 m / 6 LGS / 6 layer system needs
 * 2.9 MB); a plan whose one-system state exceeds it gets MT_EINVAL.  The
 * systems are solved one after another with the plan's one-system launch
 * tables (built by gpu_tomo_setup and by every call that changes the launch
 * shape), with the same bits as gpu_tomo_solve_dev.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other
 * threads, other plans) race and give wrong results; the array is cleared
 * at the start of each system.  Profiling must be off (MT_ESTATE).
 * No counterpart in the reference: MAOS allocates its CG scratch per
 * reconstructor (CGTMP_T, cuda/pcg.h:24-37); this is for callers that
 * cannot allocate. */
int  gpu_tomo_solve_dev_noalloc(gpu_tomo_t *g, const float *d_grad, float *d_x,
                                double *d_res, int nsys, void *stream);
/* Tomography half of MVM assembly: columns [col0, col0 + ncol) of the
 * matrix T with x = T g, each the solve of a unit gradient vector (cold
 * start, maxit iterations, nbatch columns per graph solve), into out
 * (host, nx x ncol doubles, column-major: column k = solution of e_{col0+k}
 * in gpu_tomo_solve's layer layout).  Columns of WFS the solve does not use
 * are zero.  The gradients and solutions of the plan's slots (the warm-start
 * state) are kept: the next gpu_tomo_solve starts where it would have.
 * MAOS assembles its MVM by solving tomography and fitting for every
 * gradient column, in batches (mvm_direct, cuda/mvm_direct.cu:132); this is
 * the tomography part. */
int  gpu_tomo_mvm_tomo(gpu_tomo_t *g, long col0, long ncol, double *out);
/* Residual sqrt(r'r/b'b) of each system after the last solve (host copy). */
int  gpu_tomo_residuals(gpu_tomo_t *g, double *cgres, int nbatch);

/* Operator applies for parity checks (host double buffers, one system):
 *   L : xout = beta*xout + alpha*(HX'G'WGHX + L2'L2) xin  (cutomo_grid::L)
 *   R : xout = alpha * HX'G'W(I-TT*PTT) g                 (cutomo_grid::R)
 *   P : xout = FDPCG^-1 xin                                (cufdpcg_t::P)
 */
int  gpu_tomo_L(gpu_tomo_t *g, const double *xin, double *xout,
                double alpha, double beta);
int  gpu_tomo_R(gpu_tomo_t *g, const double *grad, double *xout, double alpha);
int  gpu_tomo_P(gpu_tomo_t *g, const double *xin, double *xout);

/* Precision of the solve.  Operators (HX, GP, W, L2), x, p, z and the
 * inverse FDPCG transform are FP32 in every mode, dot products / alpha /
 * beta FP64 in modes 0-2; the modes differ in the residual recurrence
 * (r, A*p, b), the forward FDPCG spectrum and the reductions:
 *   0  r, A*p, b FP32; forward spectrum FP32
 *   1  r, A*p, b FP64; forward spectrum FP64   (default)
 *   2  r, A*p, b FP64; forward spectrum FP32
 *   3  as 0, and the dot products, alpha and beta in FP32 too (FP32
 *      everywhere; for the precision study)
 *   4  everything FP64: vectors, operator coefficients, the full-spectrum
 *      FDPCG transforms and the reductions (the reference's CUDA_DOUBLE
 *      build, cuda/common.h:39-45), on a separate FP64 plan made when the
 *      mode is selected; it serves gpu_tomo_solve, gpu_tomo_solve_batch,
 *      gpu_tomo_L/R/P, gpu_tomo_get_x and gpu_tomo_residuals, and the FP32
 *      device-buffer, group and MVM entry points refuse the plan (MT_EINVAL)
 *      while it is active.  Entering and leaving mode 4 carries the
 *      warm-start solutions over (rounded to FP32 on leaving).
 * A 30-iteration FDPCG solve of the 30 m system drifts ~2e-3 (relative L2)
 * from the FP64 solve with mode 0 and stays ~1e-5 with mode 1. */
int  gpu_tomo_precision(gpu_tomo_t *g, int mode);

/* Host gradients of system slot isys, and the solution / residual of slot
 * isys after a solve (for use with groups). */
int  gpu_tomo_set_grad(gpu_tomo_t *g, int isys, const double *grad);
int  gpu_tomo_get_x(gpu_tomo_t *g, int isys, double *x, double *cgres);

/* Uneven batch: system ranges [s0[i], s0[i]+nsys[i]) of n plans, of any
 * geometry (aperture, WFS and layer counts), solved in lockstep with one
 * launch per kernel per iteration covering all of them; every workgroup
 * finds its plan and block in a table.  Plans must share the device,
 * precision mode, maxit, warm flag and preconditioner.  A member whose
 * launch shape or precision changes later (gpu_tomo_tune_set,
 * gpu_tomo_precision, gpu_tomo_autotune) makes the group rebuild its tables
 * and graph at its next solve.  Inputs are the plans' gradient
 * slots (gpu_tomo_set_grad or gpu_tomo_solve_dev's buffers); results stay in
 * the plans (gpu_tomo_get_x).  mode: 0 = launches, 1 = replay a captured
 * hipGraph.  stream 0: the group's own stream, synchronised on return.
 * No reference counterpart: MAOS solves one system per gpu_tomo call
 * (cuda/recon.cu:564); this batches many independent loops. */
typedef struct gpu_tomo_group gpu_tomo_group_t;
int  gpu_tomo_group_create(int n, gpu_tomo_t *const *plans, const int *s0,
                           const int *nsys, gpu_tomo_group_t **out);
int  gpu_tomo_group_solve(gpu_tomo_group_t *grp, int mode, void *stream);
/* Autotuner for a group: times whole graph solves of the group under the
 * column kernel (registers where a plan allows / LDS stages), the layer
 * tile rows (64 / 32 / 16) and the systems per workgroup of the WFS and layer
 * passes (1, 2, 4), by coordinate descent from the built-in heuristic, and
 * sets the fastest on every member plan (as gpu_tomo_tune_set would, so a
 * plan's results alone and in the group stay the same bits).  choice[6]
 * (may be NULL) receives it in gpu_tomo_tune_set's order.  Overwrites the
 * solutions of the group's slots.  No reference counterpart. */
int  gpu_tomo_group_autotune(gpu_tomo_group_t *grp, int *choice);
void gpu_tomo_group_free(gpu_tomo_group_t *grp);

/* Request loop with continuous batching.  Each AO loop (add_loop) owns a
 * plan with one warm-start slot, admitted only while the plans' device
 * memory stays within mem_budget_gb (else MT_EFULL; remove_loop frees it).
 * submit() queues one time step's gradients of a loop and returns at once;
 * a worker thread launches the pending requests of loops that are not on
 * the GPU as one batch (one launch per kernel per iteration for all of
 * them; use_graph: replayed from a hipGraph cached per set of loops) on one
 * of max_inflight streams, at most max_batch requests per batch, so requests
 * that arrive while others run are admitted as soon as a stream and their
 * loop are free, and batches overlap on the device.  Each loop warm-starts
 * from its own previous solution; all loops of a server use one maxit.
 * poll() returns completed requests (tag,
 * residual, submit-to-completion latency); get_x() the loop's latest
 * solution.  No reference counterpart (MAOS solves one loop per process,
 * cuda/recon.cu:564); this serves many loops on one GPU. */
typedef struct gpu_tomo_server gpu_tomo_server_t;
int  gpu_tomo_server_create(int device, double mem_budget_gb, int max_inflight,
                            int max_batch, int use_graph,
                            gpu_tomo_server_t **out);
int  gpu_tomo_server_add_loop(gpu_tomo_server_t *s, const mt_recon_t *recon,
                              int maxit, int *loop);
int  gpu_tomo_server_remove_loop(gpu_tomo_server_t *s, int loop);
int  gpu_tomo_server_submit(gpu_tomo_server_t *s, int loop, const double *grad,
                            long tag);
int  gpu_tomo_server_poll(gpu_tomo_server_t *s, int maxn, int *loop, long *tag,
                          double *cgres, double *latency_us, int *n);
int  gpu_tomo_server_drain(gpu_tomo_server_t *s);
int  gpu_tomo_server_get_x(gpu_tomo_server_t *s, int loop, double *x,
                           double *cgres);
int  gpu_tomo_server_usage(gpu_tomo_server_t *s, double *used_gb,
                           double *budget_gb);
/* New turbulence profile of one served loop (recon from
 * mt_setup_update_cn2 of that loop's setup): the loop is held (its requests
 * wait, every other loop keeps being served), the flight that carries it, if
 * any, is waited for, and its plan's regulariser and FDPCG blocks are
 * replaced as gpu_tomo_update_cn2 does; its warm-start solution is kept.
 * The profile-update path of gpu_update_recon_cn2 (cuda/recon.cu:526) for a
 * loop of the request loop, which has no counterpart in MAOS. */
int  gpu_tomo_server_update_cn2(gpu_tomo_server_t *s, int loop,
                                const mt_recon_t *recon);
/* Autotuner for a server: gpu_tomo_group_autotune over the admitted loops
 * (one system each, as a batch carries them), the result set on every
 * loop's plan and the cached batch graphs captured again before their next
 * launch.  Each loop's warm-start solution is kept.  Call it while the
 * server is idle (nothing submitted and not yet completed: after drain());
 * MT_ESTATE otherwise, MT_EINVAL with no loops.  A loop added later keeps
 * the default launch shape.  choice[6] (may be NULL) receives the choice in
 * gpu_tomo_tune_set's order.  No reference counterpart. */
int  gpu_tomo_server_autotune(gpu_tomo_server_t *s, int *choice);
void gpu_tomo_server_free(gpu_tomo_server_t *s);

/* Autotuner: times the solve of nsys systems under each layer tile shape
 * (16 x 64, 16 x 32 or 16 x 16 nodes, where they fit), each column kernel (in
 * registers, where every FDPCG block couples one frequency across the
 * layers, at most 8, and the column length is 2^a x P with P in 2, 3, 4, 5,
 * 7; or in LDS stages, always available) and each choice of systems per
 * workgroup (1, 2, 4) of the WFS, layer and LDS column kernels, and keeps
 * the fastest for this plan; then times whole graph solves of the nsys
 * systems as 1, 2 and 4 flights (gpu_tomo_solve_dev mode 1).  choice[6]
 * receives wfs, layer, col (0 with the register column kernel), the layer
 * tile rows (16, 32 or 64), the column kernel (1 registers, 2 LDS stages) and
 * the flights.  Overwrites the gradients and solutions of slots [0, nsys).  Call it before the plan joins a group.
 * The tile shape and the column kernel set the order of sums, so they can
 * change results in the last bits; systems per workgroup never do.  No
 * reference counterpart. */
int  gpu_tomo_autotune(gpu_tomo_t *g, int nsys, int *choice);

/* Set the launch shape directly, e.g. to replay an autotuned choice without
 * timing: choice[6] = systems per workgroup of the WFS, layer and LDS
 * column kernels (each 1, 2 or 4, or 0 for the built-in heuristic), the
 * layer tile rows (16, 32, 64, or 0 for 64 where it fits), the column kernel
 * (1 registers where it applies, 2 LDS stages, 0 as 1; 2 also keeps the row
 * FFTs in LDS stages, which otherwise run in registers for single-plan
 * batches with layers of 64 or 96 points) and the flights of a
 * graph solve (1, 2, 4, or 0 for 1).  Values above a
 * kernel's maximum (WFS 2 when tomo.pos = 2, column 2 in precision modes 1
 * and 2) are clamped when the launch is built.  Call it before the plan
 * joins a group.  No reference counterpart. */
int  gpu_tomo_tune_set(gpu_tomo_t *g, const int *choice);

/* Plan facts: matched-sampling ray path (1) or generic ray tables (0), FFT
 * sizes, FDPCG block size, largest FFT radix. */
int  gpu_tomo_info(const gpu_tomo_t *g, int *matched, int *fnx, int *fny,
                   int *bs, int *rmax);

/* Sizes of one system's vectors. */
int  gpu_tomo_sizes(const gpu_tomo_t *g, long *nx, long *ng);

/* Average device duration (ms) per launch of each kernel class over the
 * last gpu_tomo_solve_dev call made with profiling enabled (HIP events on the
 * stream the kernels run on), and the number of launches of that class.
 * names/ms/calls are parallel arrays of length *n (in: capacity).
 * enable: 0 off, 1 on; n > 1 also launches each class whose kernels only
 * read their inputs (TT, WFS, layer and column passes) n times back to back
 * inside one event pair, which takes the event overhead out of the
 * per-launch average (the solve's results are unchanged). */
int  gpu_tomo_profile(gpu_tomo_t *g, int enable);
int  gpu_tomo_profile_read(gpu_tomo_t *g, const char **names, double *ms,
                           double *calls, int *n);

/* ------------------------------------------------------------------ */
/* 3. DM fitting (the solve after tomography)                          */
/* ------------------------------------------------------------------ */
/*
 * The fitting step of MAOS's minimum-variance reconstructor
 * (recon.alg=0, fit.alg=1 CG): the DM commands a minimise the aperture-
 * weighted piston-removed residual between the tomography estimate x seen
 * through the fitting directions and the DMs,
 *
 *   (sum_f wt_f HA_f' W HA_f + NW NW' + S + tikcr*4/nact I) a
 *       = sum_f wt_f HA_f' W HX_f x,                W = W0 - W1 W1',
 *
 * as FitR / FitL (maos/recon_utils.c:483-560, with the assembled matrix's
 * Tikhonov term, maos/setup_recon_fit.c:270-283) and the GPU's grid form
 * cufit_grid (cuda/fit.cu:47-230, gpu.fit=2: fit.square=1, tomo.square=1).
 * HA_f: bilinear interpolation from each DM's square actuator grid to the
 * fitting grid floc along fitting direction f (mkh, lib/mkh.c:47-135, the
 * linear influence function dm.iac=0); HX_f: the same from the tomography
 * layers; W0, W1: the bilinear-influence aperture weighting on floc
 * (mkw_circular, lib/mkw.c:363-439); NW: the piston and upper-DM tip/tilt
 * low-rank terms (setup_fit_lrt, maos/setup_recon_fit.c:104-171); S: the
 * slaving regulariser of weakly coupled actuators (slaving,
 * lib/slaving.c:63-227, with the coupling of genactcpl, :24-55).
 * Not supported (MT_EINVAL): dm.iac != 0 (cubic influence functions) and
 * fit.actinterp=1 (HA replaced by HA*actinterp); the reference's grid path
 * (gpu.fit=2) does not apply actinterp either.
 */
typedef struct {
    double ht;       /* dm.ht (m)                                          */
    double dx;       /* dm.dx (m); dm.ar = 1                               */
    double offset;   /* dm.offset (fraction of dx)                         */
    double guard;    /* dm.guard (actuator rings)                          */
    double iac;      /* dm.iac; must be 0                                  */
} mt_dm_t;

typedef struct {
    int    ndm;
    const mt_dm_t *dm;
    int    nfit;                   /* fit directions                        */
    const double *thetax, *thetay; /* fit.thetax/thetay (radian)            */
    const double *wt;              /* fit.wt (normalised to sum |wt| = 1)   */
    const double *hs;              /* fit.hs (m); INFINITY for science      */
    int    pos;                    /* fit.pos; <= 0: tomo.pos               */
    int    maxit;                  /* fit.maxit; 0: 10 (warm) / 100 (cold) as
                                      maos/setup_parms.c:2381-2385          */
    int    warm;                   /* recon.warm_restart                    */
    int    lrt_piston, lrt_tt;     /* fit.lrt_piston, fit.lrt_tt            */
    int    actslave;               /* fit.actslave                          */
    double actthres;               /* fit.actthres                          */
    double tikcr;                  /* fit.tikcr                             */
    int    actinterp;              /* fit.actinterp; must be 0              */
    int    annular_W;              /* dbg.annular_W                         */
} mt_fit_parms_t;

/* POD view of the fitting operators (RECON_T->fit fields on this path).
 * Grids are square maps (fit.square=1); vectors are stacked per DM,
 * index ix + iy*anx[dm]. */
typedef struct {
    int    ndm;
    const long   *anx, *any;       /* actuator grids (recon->amap)           */
    const double *aox, *aoy, *adx, *aht;
    long   fnx, fny;               /* fitting grid floc (recon->fmap)        */
    double fox, foy, fdx;
    int    nfit;
    const double *thetax, *thetay, *wt, *hs;
    const double *W0;              /* 9 entries per floc point: W0(i, j) for
                                      j at offsets (dx, dy) in {-1,0,1}^2,
                                      index (dy+1)*3 + dx+1                  */
    const double *W1;              /* per floc point                         */
    int    nnw;                    /* low-rank columns (NW)                  */
    const double *NW;              /* nact_total x nnw, column-major         */
    /* S in CSR over all actuators (stacked DMs; block diagonal) or NULL     */
    const long   *sl_p, *sl_i;
    const double *sl_x;
    double tikdiag;                /* tikcr * 4 / nact_total                 */
    const double *actcpl;          /* per actuator coupling (genactcpl)      */
    int    maxit, warm;
} mt_fit_recon_t;

typedef struct mt_fit mt_fit_t;

/* Fitting geometry and operators for the tomography geometry of s
 * (setup_recon_floc/aloc, setup_recon_fit: maos/setup_recon_prep.c:112-165,
 * 271-360, maos/setup_recon_fit.c:404-453). */
int  mt_fit_create(const mt_setup_t *s, const mt_fit_parms_t *p, mt_fit_t **out);
int  mt_fit_recon(const mt_fit_t *f, mt_fit_recon_t *out);
/* Actuator count (sum over DMs) and floc points. */
int  mt_fit_sizes(const mt_fit_t *f, long *nact, long *nfloc);
void mt_fit_free(mt_fit_t *f);

typedef struct gpu_fit gpu_fit_t;

/* Upload the fitting operators for nbatch independent systems; the
 * tomography layers x come in the layout of recon (the tomography's
 * mt_recon_t).  Replaces the fitting half of gpu_setup_recon
 * (curecon_t::update -> cufit_grid, cuda/recon.cu:104-125, cuda/fit.cu:47-117). */
int  gpu_fit_setup(const mt_recon_t *recon, const mt_fit_recon_t *fit, int device,
                   int nbatch, gpu_fit_t **out);
/* Fit of one system (slot isys, host doubles): opdr (tomography layers, as
 * gpu_tomo_solve returns them) -> RHS -> CG (maxit iterations, warm start
 * from the slot's previous DM commands when warm) -> dmfit (stacked DM
 * vectors); cgres = sqrt(r'r / b'b).  Replaces curecon_t::fit
 * (cuda/recon.cu:277-288). */
int  gpu_fit_solve(gpu_fit_t *g, int isys, const double *opdr, double *dmfit,
                   double *cgres);
/* nsys <= nbatch systems with device buffers (FP32 layer vectors in, FP64
 * DM vectors out), enqueued on stream (0: the plan's own); no host sync.
 * Five launches per CG iteration, shared by the nsys systems.  d_dmfit
 * holds the warm start when the plan has warm = 1. */
int  gpu_fit_solve_dev(gpu_fit_t *g, const float *d_opdr, double *d_dmfit,
                       int nsys, void *stream);
/* Operators for parity checks (host doubles, one system):
 *   R : out = alpha * FitR x              (cufit_grid::R)
 *   L : out = beta * out + alpha * FitL a (cufit_grid::L) */
int  gpu_fit_R(gpu_fit_t *g, const double *opdr, double *out, double alpha);
int  gpu_fit_L(gpu_fit_t *g, const double *a, double *out, double alpha, double beta);
/* Residuals of the last solve. */
int  gpu_fit_residuals(gpu_fit_t *g, double *cgres, int nsys);
/* MVM assembly: columns [col0, col0 + ncol) of the control matrix
 * M = F T (DM commands per gradient), each the cold tomography solve of a
 * unit gradient vector (t's maxit, nbatch columns per graph solve) followed
 * by the cold fit of that solution (f's maxit) on the device, into out
 * (host, nact x ncol doubles, column-major).  Columns of WFS the
 * tomography does not use are zero.  t's warm-start state is kept.  The
 * tomography-then-fitting assembly of mvm_direct (cuda/mvm_direct.cu:
 * 107-134; per column, as cuda/mvm_trans.cu:98-148 distributes them). */
int  gpu_mvm_assemble(gpu_tomo_t *t, gpu_fit_t *f, long col0, long ncol,
                      double *out);
void gpu_fit_free(gpu_fit_t *g);

/* ------------------------------------------------------------------ */
/* MAOS .conf files                                                    */
/* ------------------------------------------------------------------ */
/* The key store MAOS fills from its .conf files and command-line
 * overrides (open_config, lib/readcfg.c:236-456) and the value parsers it
 * reads keys with (readcfg_str / readcfg_dmat / readcfg_strarr,
 * lib/readcfg.c:523-640; readstr_numarr, sys/readstr.c:193-420), so that a
 * MAOS configuration can set up this library (maos_amd.conf maps the keys
 * of this path to mt_parms_t / mt_fit_parms_t as setup_parms does).
 * mt_conf_read takes a file name ending in .conf (found as given, then in
 * the directories of mt_conf_add_path, most recent first) or key=value text
 * (lines; what MAOS accepts on its command line).  A key read again at the
 * same or a higher priority replaces the earlier value; key += [..]
 * appends to an array.  Every get marks the key as used; mt_conf_unused
 * lists the keys never read (MAOS's close_config refuses them). */
typedef struct mt_conf mt_conf_t;
int  mt_conf_new(mt_conf_t **out);
void mt_conf_free(mt_conf_t *c);
int  mt_conf_add_path(mt_conf_t *c, const char *dir);
int  mt_conf_read(mt_conf_t *c, const char *file_or_text, const char *prefix,
                  int priority);
int  mt_conf_has(const mt_conf_t *c, const char *key);
/* The value with paired quotes removed (empty: n = 0); buf may be NULL to
 * query the length n. */
int  mt_conf_str(mt_conf_t *c, const char *key, char *buf, long len, long *n);
/* String array element i (i < 0: the count into n). */
int  mt_conf_strarr(mt_conf_t *c, const char *key, long i, char *buf, long len,
                    long *n);
/* Numeric array (column-major, shape nrow x ncol; a value naming a .bin or
 * .bin.gz file is that file's double array); maxn = 0 queries the count. */
int  mt_conf_dbl(mt_conf_t *c, const char *key, double *out, long maxn,
                 long *n, long *nrow, long *ncol);
/* Keys never read, space separated, and their count. */
int  mt_conf_unused(const mt_conf_t *c, char *buf, long len, long *n);

/* ------------------------------------------------------------------ */
/* MAOS .bin files                                                     */
/* ------------------------------------------------------------------ */
/* Double arrays and cells of them in MAOS's .bin layout (sys/bin.c:513-811,
 * math/cell.c:227-270; .bin.gz through zlib), so that gradients saved by MAOS
 * (gradlastcl, the input of gpu_tomo) and solutions (opdr) pass through
 * unchanged.  A node is a double matrix (nx rows, ny columns, column-major)
 * or a cell of nodes (an empty element is NULL); each may carry the header
 * text MAOS stores before it (key=value lines); the other element types
 * MAOS writes (complex, single precision, integer and sparse matrices) are
 * read and written too.  No reference counterpart as a C API: MAOS does this
 * I/O itself with readbin / writebin (math/cell.c). */

int  mt_bin_read(const char *path, mt_bin_t **out);       /* .bin or .bin.gz */
int  mt_bin_write(const char *path, const mt_bin_t *b);   /* gzip if the path ends in .gz */
void mt_bin_free(mt_bin_t *b);
int  mt_bin_info(const mt_bin_t *b, int *is_cell, long *nx, long *ny);
/* Element types (sys/bin.h:34-51): double, cell, complex double, single
 * float, single complex, integer (64- or 32-bit on file), sparse double,
 * sparse complex (64- or 32-bit indices on file; CSC). */
#define MT_BIN_DOUBLE   0
#define MT_BIN_CELL     1
#define MT_BIN_COMPLEX  2
#define MT_BIN_FLOAT    3
#define MT_BIN_FCOMPLEX 4
#define MT_BIN_INT      5
#define MT_BIN_SPARSE   6
#define MT_BIN_CSPARSE  7
int  mt_bin_type(const mt_bin_t *b);
/* Dense values as doubles: nx*ny (real) or 2*nx*ny (complex: re, im pairs;
 * single precision widened); NULL for cells, integers and sparse. */
const double   *mt_bin_data(const mt_bin_t *b);
const long     *mt_bin_idata(const mt_bin_t *b);           /* integer arrays */
/* A sparse matrix in CSC form: p (ny+1 column pointers), i (row indices),
 * x (values; complex: re, im pairs), nzmax = p[ny]. */
int  mt_bin_sparse(const mt_bin_t *b, const long **p, const long **i,
                   const double **x, long *nzmax);
const mt_bin_t *mt_bin_child(const mt_bin_t *b, long i);   /* NULL: empty or not a cell */
const char     *mt_bin_header(const mt_bin_t *b);
int  mt_bin_new_dbl(long nx, long ny, const double *data, const char *header, mt_bin_t **out);
int  mt_bin_new_cell(long nx, long ny, const char *header, mt_bin_t **out);
/* A dense node of another type (MT_BIN_COMPLEX / FLOAT / FCOMPLEX / INT;
 * data as mt_bin_data returns it; integers given as doubles). */
int  mt_bin_new_typed(int type, long nx, long ny, const double *data,
                      const char *header, mt_bin_t **out);
/* A sparse matrix (CSC; cplx: x holds re, im pairs). */
int  mt_bin_new_sparse(long nx, long ny, const long *p, const long *i,
                       const double *x, int cplx, const char *header,
                       mt_bin_t **out);
/* Put child at element i of cell (column-major); the cell takes ownership
 * (and frees any previous element). */
int  mt_bin_set(mt_bin_t *cell, long i, mt_bin_t *child);

/* The solve's vectors in MAOS's layouts: gradients from an nwfs x 1 cell of
 * 2*nsa x 1 arrays, x slopes then y slopes (gradlastcl,
 * maos/recon_utils.c:699; the elements of skipped WFS are ignored) into the
 * gradient vector of gpu_tomo_solve; and a solution x into an npsr x 1 cell
 * of xnx x xny arrays, as the GPU path holds opdr (cuda/recon.cu:72). */
int  mt_bin_grad(const mt_bin_t *b, const mt_recon_t *r, double *grad);
int  mt_bin_opdr(const mt_recon_t *r, const double *x, mt_bin_t **out);

#ifdef __cplusplus
}
#endif
#endif /* MAOS_TOMO_H */
