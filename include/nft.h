/* nft.h -- C ABI of the MI355X (gfx950) edge-based NFT time step.
 *
 * Drop-in boundary for the hot path of lippsLFroude_implgwopks3D_unstr3.f (the reference):
 * MPDATA advection over median-dual edges and the GCR(k) solve for the anelastic pressure
 * with implicit gravity waves. The reference's Fortran-77 subroutines keep their names and
 * argument lists in src_amd/fortran/nft_hip.f90 and call the functions below through
 * iso_c_binding; Python reaches them through ctypes (src_amd/nft.py).
 *
 * Conventions
 *  - All arrays are host arrays in the reference's column-major SoA layouts; integer node
 *    ids are 1-based as in the reference. Leading dimensions (the reference's compile-time
 *    npoinx / nedgex / nfacesx) are passed explicitly so COMMON arrays can be passed as is.
 *  - The mesh, the profiles and the time-step state live on the GPU after nft_init; the
 *    per-subroutine entry points copy their arguments in and out, the step driver does not.
 *  - Return codes: 0 ok; <0 argument or HIP error (text in nft_last_error()); >0 the GCR
 *    solve stopped before reaching its tolerance (iteration limit or stagnation).
 *  - One context per process thread; calls on one context are not reentrant.
 */
#ifndef NFT_H
#define NFT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nft_ctx nft_ctx;

/* Mesh as read by the reference's `input` (:2343-2501, COMMON /n/ /xyz/ /volume/ /edg/
 * /wge/ /bound/ /bpoints/ /terms/). col_start (optional, ncol > 0) lists 0-based node
 * ranges of vertical columns, bottom to top, for the vertical line preconditioner. */
typedef struct {
  int npoin, nedge, nbface, nboun;
  int ld_poin, ld_edge, ld_face;  /* leading dimensions (>= counts) */
  const int32_t *iedge;   /* iedge(ld_edge,2) */
  const double *sn;       /* sn(ld_edge,3) */
  const int32_t *kbface;  /* kbface(ld_face,2) */
  const double *sb;       /* sb(ld_face,3) */
  const int32_t *ibpoin;  /* ibpoin(ld_face,2): node, flag */
  const double *sbb;      /* sbb(ld_face,3) */
  const double *anp;      /* anp(ld_face,3) */
  const double *coord;    /* coord(ld_poin,3) */
  const double *vol;      /* vol(ld_poin) */
  int ncol;
  const int32_t *col_start; /* [ncol+1] */
} nft_mesh;

/* Doubly periodic box (the reference's iper = 2): the coincident node sets that `input`
 * reads after the mesh (:2569-2583; COMMON /coin/ icoin(nfacesx*2,2), /per2/
 * icoin2(nfacesx,2), /per4/ icoin4(nfacesx,4)), 1-based, with their leading dimensions.
 * Every node may belong to one set at most. */
typedef struct {
  int ncoin, ncoin2, ncoin4;
  int ld_coin, ld_coin2, ld_coin4;
  const int32_t *icoin;   /* icoin(ld_coin,2): pairs of the x sides */
  const int32_t *icoin2;  /* icoin2(ld_coin2,2): pairs of the y sides */
  const int32_t *icoin4;  /* icoin4(ld_coin4,4): quadruples of the corner columns */
  int vol_closed;         /* nonzero: vol already holds the sums over the sets (the reference's
                             vol after mpsolver :99-103); 0: nft_init_periodic forms them */
} nft_periodic;

/* Run parameters: the reference's inputinc.d / input_mg_v3.d knobs and the constants
 * hard-coded in its mpsolver (Cartesian, lipps=0, implgw=1, icorio=0; itraj below). */
typedef struct {
  double dt;                 /* delt -> timestp */
  double u0, v0, w0;         /* u0inf v0inf w0inf */
  double strat;              /* N^2/g */
  double theta0;             /* tetbar */
  int lord;                  /* GCR(k), 1..16 */
  int itr;                   /* max outer GCR cycles */
  int exit_mode;             /* 0: max|r| <= eps0/dt (reference); 1: max|r| <= eps0*max|b| */
  double eps0;
  int precon;                /* 0 none (iprc=0), 1 point relaxation (reference), 2 vertical line
                                (two-point-flux tridiagonal), 3 vertical line with the in-column
                                band of laplc (pentadiagonal, probed at nft_init; columns of one
                                length, at most 400 levels) */
  int sm_flag;               /* point relaxation: 1 Jacobi, 2 Richardson, 3 mixed */
  int itrp1;                 /* preconditioner sweeps */
  double omega;
  double dxabl, dxabr, dyabd, dyabu, dzabd, dzabu;  /* absorber depths */
  double towxl, towxr, towyd, towyu, towzd, towzu;  /* absorber time scales */
  int irl;
  double dom[6];             /* box used by beta_timestep_unstr (reference: input :2255-2260) */
  int precision;             /* 0: FP64 everywhere; 1: mixed (FP32 MPDATA + preconditioner);
                                2: as 1, and laplc inside the GCR iterations with FP32 edge
                                metrics and FP32 gradient temporaries (FP64 sums, residual,
                                dots and updates); each solve runs in passes that restart
                                from the FP64 residual of its pressure and ends only when
                                that residual meets the tolerance */
  int itraj;                 /* advector at n+1/2: 0 extrapolation (reference default, :427-438);
                                1 velprd's first-order prediction (:4448-4531, one more GCR
                                solve per step) */
} nft_params;

/* GCR statistics, as /itero/ of the reference (:1953-1963) plus timing. */
typedef struct {
  int niter;       /* preconditioned iterations of the last solve (nlc) */
  int nitsm;       /* running sum of niter */
  int icount;      /* number of solves */
  int converged;   /* last solve met its tolerance */
  double eer, eem; /* final/initial preconditioned residual ratios (L2^2, max) */
  double res_max;  /* max|r| at exit of the last solve */
  double tol;      /* absolute tolerance the last solve used */
} nft_gcr_stats;

typedef struct {
  int steps;
  int gcr_solves;
  int gcr_iters;          /* sum of niter over the solves */
  int gcr_unconverged;    /* solves that stopped early */
  double ms_total;        /* device time of the steps (hipEvent) */
  double ms_gcr;          /* device time inside gcrk calls */
  double ms_mpdata;       /* device time inside the four mpdatam calls */
} nft_run_stats;

/* ---- life cycle ---------------------------------------------------------------------- */
/* Replaces `input`'s hand-off to mpsolver's set-up (rhprof, thprof, absorbers,
 * beta_timestep_unstr(1), compute_d_unstr(1), coefw; :140-170, :250-252): uploads the
 * mesh, builds the node-sorted incidence and the column factors. */
int nft_init(const nft_mesh *mesh, const nft_params *params, nft_ctx **out);
/* nft_init for the doubly periodic box (iper = 2): mpsolver's set-up with vol summed over
 * the coincident sets first (:99-103, unless per->vol_closed says it is); afterwards every entry point runs the reference's
 * periodic branches: derivative without the boundary-point terms and summed over the sets
 * (cyclicr/cyclicr2/cyclicr4, :4015-4019, :4377-4446), bcv with the flag-11 slip only
 * (:1603-1612), courant without its boundary points (:2049), and nft_mpdatam and the step
 * with mpdatamper (:314-319, :3617-3941). FP64 only (precision 0); the multigrid branch, the
 * learned preconditioner and nft_laplc_t are not available. */
int nft_init_periodic(const nft_mesh *mesh, const nft_params *params, const nft_periodic *per, nft_ctx **out);
int nft_finalize(nft_ctx *ctx);
const char *nft_last_error(void);
int nft_device_name(char *buf, int len);

/* ---- reference subroutines (host arrays in/out) ------------------------------------- */
/* derivative(unk,dpdxa,idx)                     reference :3943-4022 */
int nft_derivative(nft_ctx *ctx, const double *unk, double *dpdxa, int idx);
/* laplc(press,alap,impfl)                       reference :1288-1353 */
int nft_laplc(nft_ctx *ctx, const double *press, double *alap, int impfl);
/* out = L^T y, the transpose of laplc (no reference counterpart: trains the learned
 * preconditioner) */
int nft_laplc_t(nft_ctx *ctx, const double *y, double *out, int impfl);
/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): laplc
 * (:1288-1353) on DEVICE arrays press and alap (npoin doubles each, on the context's GPU,
 * ready when the call is made), with laplc's intermediate held in a fixed device array of
 * the library instead of the context's work arrays. The call allocates nothing; it
 * enqueues two kernels (four in the periodic box, whose set sums live in the context) on
 * the context's stream and waits for them. Limit: meshes of at most
 * 1,048,576 nodes (the array holds 24 MiB on each GPU the library runs on, until the
 * process ends); larger meshes are rejected with an error. Every call on one GPU shares that
 * array, so calls that overlap on one GPU (from several contexts or threads) race. Results
 * equal nft_laplc's bit for bit. */
int nft_laplc_noalloc(nft_ctx *ctx, const double *press, double *alap, int impfl);
/* no reference counterpart: the FP32 operator that precision 2 applies inside the GCR
 * iterations (edge metrics and scaled gradient in FP32, FP64 sums), on host arrays; for
 * checking it against its CPU restatement. Precision 2 contexts only. */
int nft_laplc32(nft_ctx *ctx, const double *press, double *alap, int impfl);
/* prfr(press,ustr,vstr,wstr,uout,vout,wout,impfl) reference :1536-1580 */
int nft_prfr(nft_ctx *ctx, const double *press, const double *ustr, const double *vstr,
             const double *wstr, double *uout, double *vout, double *wout, int impfl);
/* bcv(uout,vout,wout,flag)                      reference :1583-1643 */
int nft_bcv(nft_ctx *ctx, double *u, double *v, double *w, double flag);
/* div(ua,va,wa,rhs,iflg), without the iflg=1 print  reference :2096-2146 */
int nft_div(nft_ctx *ctx, const double *ua, const double *va, const double *wa, double *rhs,
            int iflg);
/* precon(rhs,p,r,iprc,elas,impfl), fine-grid smoothing  reference :482-612 */
int nft_precon(nft_ctx *ctx, const double *rhs, double *p, int impfl);
/* gcrk(p,u1,v1,w1,u2,v2,w2,impfl)               reference :1776-1972 */
int nft_gcrk(nft_ctx *ctx, double *p, const double *u1, const double *v1, const double *w1,
             double *u2, double *v2, double *w2, int impfl, nft_gcr_stats *stats);
/* mpdatam(uuu,unk,iflag); uuu(ld,3)             reference :3029-3311 */
int nft_mpdatam(nft_ctx *ctx, const double *uuu, int ld, double *unk, int iflag);

/* ---- device-resident time stepping (mpsolver, reference :58-463) -------------------- */
/* mpsolver initialisation: potential-flow projection, coefw, advector (:177-266) */
int nft_state_init(nft_ctx *ctx);
/* nsteps iterations of the DO ITIME loop (:285-439), state kept on the GPU */
int nft_run_steps(nft_ctx *ctx, int nsteps, nft_run_stats *stats);
/* state blocks: "unk"(4), "frc"(4), "uuu"(3), "unkdash"(3), "prpot"(1); [k][npoin] */
int nft_get_state(nft_ctx *ctx, const char *name, double *host, int ld);
int nft_set_state(nft_ctx *ctx, const char *name, const double *host, int ld);
/* set-up fields: "alpha","d1","coefw","rh0","th0","ue","line_a","line_inv","line_c";
 * "courant": the per-node Courant numbers of the last nft_diagnostics call */
int nft_get_field(nft_ctx *ctx, const char *name, double *host);
int nft_get_gcr_stats(nft_ctx *ctx, nft_gcr_stats *stats);
/* Incidence used by the gradient/divergence sweeps (no reference counterpart; results are
 * identical either way): 1 = columnar (per-column slot template for the interior levels,
 * available when the mesh is extruded with identical interior levels; the default then),
 * 0 = node-sorted SELL-64. columnar < 0 only reports availability. */
int nft_set_incidence(nft_ctx *ctx, int columnar, int *available);

/* ---- multigrid branch of precon (mg_flag=1, reference :618-748) --------------------
 * After the fine sweeps, a three-level V-cycle: the fine residual restricted to coarse
 * level 1 (restriction_unstr :757-822), point relaxation there (laplc_c :1356-1446), its
 * residual restricted to level 2 and relaxed (laplc_cc :1450-1535), the corrections
 * prolonged back (prolongation_unstr :828-940) with relaxation on level 1 and itrp2 fine
 * sweeps. The coarse meshes are the reference's mesh3d_c.d / mesh3d_cc.d (COMMON /edg_c/
 * /wge_c/ ... as read; the reference reads the level-2 sn y/z columns from level 1's array,
 * :2467-2468, so pass its COMMON arrays as they are to reproduce it), match1 the 1-based
 * fine node of each level-1 node and match2 the level-1 node of each level-2 node
 * (match1.d / match2.d). Coarse levels use the point step of sm_flag with d2/d3 and
 * betaunstr2/3 (compute_d_unstr, beta_timestep_unstr :953-1284) computed here. */
typedef struct {
  int mg_flag;                        /* 1: run the V-cycle inside precon */
  int itrp2, itrp_c1, itrp_c2, itrp_cc;  /* sweeps as in input_mg_v3.d */
  int transfer;  /* 0: restriction/prolongation index the fine array by the neighbour
                    counter, as the reference does (:817, :913); 1: by the neighbour list */
  double omega_c;  /* relaxation factor of the coarse-level point steps; <= 0: omega, as
                      the reference (no counterpart otherwise: its coarse Jacobi with the
                      d2/d3 diagonal estimates diverges for omega >= 0.7) */
} nft_multigrid;
/* Coarse levels and V-cycle settings. With m1 == NULL only mg_flag and the sweep counts
 * change (the levels of an earlier call stay). A call with meshes allocates the levels'
 * device storage anew; storage of earlier levels is released by nft_finalize. */
int nft_set_multigrid(nft_ctx *ctx, const nft_mesh *m1, const int32_t *match1,
                      const nft_mesh *m2, const int32_t *match2, const nft_multigrid *mg);
/* per-level set-up values: "d" (d2 / d3), "beta" (1 value), "rh0", "pstep"; *count = length */
int nft_multigrid_field(nft_ctx *ctx, int level, const char *name, double *host, int *count);
/* the V-cycle's pieces on host arrays: op 0 restriction_unstr(in fine, out level),
 * op 1 prolongation_unstr(in level, out finer), op 2 laplc on the level (laplc_c/_cc) */
int nft_multigrid_op(nft_ctx *ctx, int op, int level, const double *in, double *out, int impfl);

/* ---- diagnostics (the reference prints them after every step, :456-459) ---------------
 * div(uuu,rhs,1) statistics of rhs*timestp (:2148-2168), courant(uuu,cour,iflag)
 * (:1974-2094) and stats(unk) (:2847-2912), computed on the GPU with fused reductions. */
typedef struct {
  double div[4];        /* dvmx, dvmn, dvav, dvsd */
  double eer, eem;      /* of the last GCR solve (/itero/) */
  int niter, nitav;     /* niter, nitsm/icount */
  int crnic;            /* 1-based node of the last Courant-number maximum */
  double cour[4];       /* crnmx, crnmn, crnav, crnsd */
  double u[4], v[4], w[4], t[4];  /* umx, umn, uav, usd ... tmx, tmn, tav, tsd */
  int64_t step;         /* time steps done when recorded */
} nft_diag;
/* the three diagnostics of the current state (courant with iflag 1, 2 or 3) */
int nft_diagnostics(nft_ctx *ctx, int courant_flag, nft_diag *out);
/* courant(uuu,cour,iflag) (:1974-2094) of the advector uuu(ld,3) given on the host: the
 * per-node Courant numbers into cour (may be null) and their statistics into out->cour,
 * out->crnic; the device-resident time-step state is not touched */
int nft_courant(nft_ctx *ctx, const double *uuu, int ld, double *cour, int iflag, nft_diag *out);
/* stats(unk) (:2847-2912) of unk(ld,4) given on the host: out->u, v, w, t; the
 * device-resident time-step state is not touched */
int nft_stats(nft_ctx *ctx, const double *unk, int ld, nft_diag *out);
/* record them on the GPU every `every` steps of nft_run_steps (0: off), keeping the last
 * `capacity` records; nft_get_diagnostics copies the records made since the previous call
 * (oldest first, at most max) and sets *count to their number */
int nft_set_diagnostics(nft_ctx *ctx, int every, int capacity, int courant_flag);
int nft_get_diagnostics(nft_ctx *ctx, int *count, nft_diag *out, int max);

/* ---- files ---------------------------------------------------------------------------
 * Full-precision state file, replacing the reference's f14.6 text restart
 * (Restart_Out_old / Restart_In_old :4779-4890, called at :472 and :277): the step-boundary
 * state unk, frc, uuu, unkdash, prpot plus the step and GCR counters; resuming from it
 * continues the run bit for bit. */
int nft_save_state(nft_ctx *ctx, const char *path);
int nft_load_state(nft_ctx *ctx, const char *path);
/* time steps since nft_state_init (or as loaded with a state file) */
int nft_steps_done(nft_ctx *ctx, int64_t *steps);
/* Binary mesh file with the content of the reference's mesh3d.d (`input` :2343-2501) and
 * the column table: host-only, no GPU needed. nft_mesh_load allocates the arrays (leading
 * dimensions = counts; kbface/sb and ibpoin/sbb/anp share one, max(nbface, nboun), as in
 * the reference's COMMON layout); release them with nft_mesh_release. */
int nft_mesh_save(const char *path, const nft_mesh *mesh);
int nft_mesh_load(const char *path, nft_mesh **out);
int nft_mesh_release(nft_mesh *mesh);

/* ---- measurement (no reference counterpart: the reference only times the whole run) -- */
/* Kernel ids for the timer: 1 laplc gradient half, 2 laplc divergence half, 3 line solve,
 * 4 GCR update, 5 GCR direction, 6..9 MPDATA donor/gradient/FCT/limit, 10 the coarse-level
 * part of the multigrid V-cycle (a chain of kernels; bytes_per_launch 0), 11 the learned
 * correction's network forward pass (bytes_per_launch holds its MFMA FLOPs), 12 MPDATA's
 * antidiffusive-flux sweep (each edge once, between the gradient and FCT sweeps). While a kernel id
 * is selected, every launch of that kernel is bracketed by HIP events on the context's
 * stream. bytes_per_launch is the kernel's algorithmic traffic: every array it reads or
 * writes counted once at its stored width (incidence slots actually used, edge metrics,
 * node arrays; for laplc's divergence also the vectors its fused dots read, averaged over
 * the timed launches). */
int nft_kernel_timer(nft_ctx *ctx, int kernel);
int nft_kernel_timer_read(nft_ctx *ctx, int *launches, double *total_ms, double *bytes_per_launch);
/* plain device copy of one node array (8*npoin bytes read + written) `reps` times, mean ms
 * per copy: the reference point for hardware byte counters */
int nft_copy_calibration(nft_ctx *ctx, int reps, double *ms);
/* copy of `bytes` between two device buffers allocated for the call (2 GiB is far larger
 * than the last-level cache), mean ms per copy: the practical HBM streaming rate beside the
 * published peak */
int nft_stream_calibration(nft_ctx *ctx, int64_t bytes, int reps, double *ms);
/* precision 2 only (no reference counterpart): a pass of a solve ends when the recursive
 * residual's max-norm has dropped by `drop` (0 < drop < 1, default 1e-3) from the FP64
 * residual the pass started from; smaller values mean fewer, longer passes. drop <= 0
 * only reads the current value into *current (may be null otherwise). */
int nft_set_mixed_drop(nft_ctx *ctx, double drop, double *current);

/* ---- learned preconditioner (no reference counterpart) ------------------------------
 * A column MLP with two horizontal graph convolutions (bf16 MFMA, FP32 accumulation)
 * adds a learned correction to the vertical line preconditioner inside the flexible GCR;
 * when an iteration stalls (rl2 > stall_ratio * rl2_prev, or no decrease) the solve drops
 * the correction and continues with the line preconditioner alone. Needs uniform columns
 * and a vertical line preconditioner (precon >= 2: the two-point-flux tridiagonal or the
 * in-column band of laplc), with any precision (0, 1 or 2). Parameters: per layer W (K x N, row-major) then b (N); layers
 * Kp->H, H->H, 2H->H, 2H->H, H->H, H->Kp with Kp = levels rounded up to a multiple of 8. */
int nft_net_init(nft_ctx *ctx, int hidden, uint64_t seed, int max_batch);
int nft_net_info(nft_ctx *ctx, int *kp, int *levels, int *ncol, int64_t *nparams);
int nft_net_params(nft_ctx *ctx, float *host, int64_t count, int set);
/* the whole training state, so training can continue in another context of the same
 * network shape (e.g. on another mesh's pairs): weights (layout of nft_net_params), Adam's
 * first and second moments in the same layout, then the Adam step count; count >= 3 *
 * nparams + 1 floats */
int nft_net_state(nft_ctx *ctx, float *host, int64_t count, int set);
/* forward (and with T, loss + backward) on a host batch [nsamples*ncol][Kp] */
int nft_net_eval(nft_ctx *ctx, const float *X, const float *T, int nsamples, float *Y, float *grad,
                 double *loss);
/* store every stride-th pair (r, q = M_line^-1 r) of the GCR iterations that follow, with
 * s = max|r| and |r|^2, until max_pairs are held, and for the error objective the iteration's
 * pressure, turned into x_k - x_final when its solve converges (pairs of a solve that does not
 * converge are dropped); device memory: 24*npoin bytes per pair */
int nft_dump_pairs(nft_ctx *ctx, int max_pairs, int stride);
/* pair store: count, and copies of pairs [first, first+n_get): r and q as n_get x npoin */
int nft_pairs(nft_ctx *ctx, int *count, int first, int n_get, double *r, double *q, double *s,
              double *r2);
/* Training objective (physics-informed): with z = q + s*N(r/s), the mean over pairs of
 * |L z - r|^2 / |r|^2, L = laplc; its gradient is s * L^T(2(Lz - r)/(B|r|^2)) per sample.
 * nft_net_pair_loss evaluates it on pairs [first, first+count); nft_net_train runs Adam. */
int nft_net_pair_loss(nft_ctx *ctx, int first, int count, double *loss, float *grad);
/* Adam over pairs [0, use_pairs) (all when use_pairs <= 0) */
int nft_net_train(nft_ctx *ctx, int steps, int batch, double lr, uint64_t seed, int use_pairs,
                  double *loss_first, double *loss_last);
/* training objective: 0 residual |Lz - r|^2/|r|^2, 1 angle 1 - <r,Lz>^2/(|r|^2 |Lz|^2),
 * 2 error |z - d|^2/|d|^2 with d = x_k - x_final (= L^-1 r to the solve's tolerance);
 * linear = 1: identity activations and biases held at zero (the correction is linear in r) */
int nft_net_options(nft_ctx *ctx, int objective, int linear);
/* kind 0: line preconditioner only; 1: learned correction + line fallback. kind < 0 only
 * reads the counters: fallbacks so far, GCR iterations that used the learned correction. */
int nft_set_learned(nft_ctx *ctx, int kind, double stall_ratio, int *fallbacks, int *learned_iters);

#ifdef __cplusplus
}
#endif
#endif /* NFT_H */
