/* CPU oracle for the edge-based NFT step: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference Fortran
 * (lippsLFroude_implgwopks3D_unstr3.f) for the Cartesian path, closed or doubly periodic
 * (isfera=0, iper=0 or 2, lipps=0, itraj=0 or 1, implgw=1, icorio=0). Each function keeps the
 * reference's loop order and operation order so that, compiled without FP contraction,
 * it reproduces the reference's rounding. Only tests, the smoke check and the CPU leg
 * of the benchmark use it, as what the HIP path's results are compared with.
 *
 * The doubly periodic variant (iper = 2) restates the reference's periodic branches:
 * cyclicr/cyclicr2/cyclicr4, mpdatamper and the periodic forms of derivative, bcv, courant
 * and mpsolver's set-up.
 *
 * Parity pinning (who checks each function):
 *   derivative, laplc, prfr, div, bcv, precon(point), gcrk, mpdatam, compute_d,
 *   absorbers, step driver  -> golden vectors written by the flang-compiled reference
 *                              (tests/golden) plus closed forms in tests/test_oracle.py
 *   periodic variant (closures, mpdatamper, the periodic derivative/bcv/courant, the
 *   closed vol in d1 and beta)
 *                           -> golden vectors of the reference built from the same
 *                              source and run with iper=2 on a generated periodic mesh
 *   multigrid branch of precon (restriction, prolongation, coarse laplc, d2/d3, the
 *   V-cycle)                -> golden vectors of the reference run with mg_flag=1 on
 *                              generated nested meshes (tests/golden)
 *   line preconditioner     -> no reference counterpart: pinned by a dense solve of the
 *                              tridiagonal system and by the closed form of the coefficients;
 *                              the probed band (precon 3) by the dense matrix of laplc
 *                              (unit-vector columns) and a dense solve of the band
 */
#ifndef NFT_ORACLE_H
#define NFT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_ctx oc_ctx;

typedef struct {
  int npoin, nedge, nbface, nboun;
  const int32_t *iedge;   /* [2][nedge], 1-based */
  const double *sn;       /* [3][nedge] */
  const int32_t *kbface;  /* [2][nbface] */
  const double *sb;       /* [3][nbface] */
  const int32_t *ibpoin;  /* [2][nboun] (node, flag) */
  const double *sbb;      /* [3][nboun] */
  const double *anp;      /* [3][nboun] */
  const double *coord;    /* [3][npoin] */
  const double *vol;      /* [npoin] */
  int ncol;               /* vertical columns (0: none) */
  const int32_t *col_start; /* [ncol+1], 0-based node ranges */
  /* doubly periodic box (iper = 2): the coincident node sets read after the mesh
   * (F:2569-2583), 1-based: pairs of the x sides, pairs of the y sides, quadruples of the
   * corner columns */
  int ncoin, ncoin2, ncoin4;
  const int32_t *icoin;   /* [2][ncoin] */
  const int32_t *icoin2;  /* [2][ncoin2] */
  const int32_t *icoin4;  /* [4][ncoin4] */
} oc_mesh;

typedef struct {
  double dt;              /* timestp (uniform) */
  double u0, v0, w0;      /* free stream */
  double strat;           /* N^2/g */
  double theta0;          /* tetbar */
  int lord, itr, exit_mode; double eps0;
  int precon;             /* 0 none (iprc=0), 1 point relaxation, 2 vertical line
                           * (two-point-flux tridiagonal), 3 vertical line with the
                           * in-column band of laplc itself (pentadiagonal, probed) */
  int sm_flag;            /* point relaxation: 1 Jacobi, 2 Richardson, 3 mixed */
  int itrp1; double omega;
  /* sponge (irl=1 when any tow is nonzero) */
  double dxabl, dxabr, dyabd, dyabu, dzabd, dzabu;
  double towxl, towxr, towyd, towyu, towzd, towzu;
  int irl;
  /* domain box used by beta_timestep_unstr */
  double dom[6];
  /* inner products: 0 = the reference's sequential sums; 1 = the HIP path's fixed
   * reduction tree (same lane/block/grid assignment), so the two can be compared bit for
   * bit and the effect of summation order measured on its own */
  int dot_order;
  /* advector at n+1/2: 0 extrapolation 1.5 rho u^n - 0.5 u^(n-1) (F:427-438); 1 the
   * first-order prediction velprd (F:4448-4531), which adds a GCR solve per step */
  int itraj;
  /* 0 FP64; 1 the mixed form: MPDATA as oc_mpdatam32 and, with the band line (precon 3),
   * the line solve as oc_line_solve32; everything else FP64; 2 as 1, with the FP32
   * operator (oc_laplc32) inside gcrk and the solve in passes from the FP64 residual */
  int precision;
  double mixed_drop;   /* precision 2: recursive-residual drop that ends a pass (0: 1e-3) */
  /* 0 the closed box; 2 the doubly periodic box (the reference's iper = 2, FP64 only): vol
   * summed over the coincident sets at set-up (F:99-103), derivative without the
   * boundary-point terms and summed over the sets (F:4007-4019), bcv on flag 11 only
   * (F:1603-1612), courant without its boundary-point loop (F:2049), mpdatamper in place
   * of mpdatam (F:314-319, F:3617-3941) */
  int iper;
} oc_params;

typedef struct { int niter, nitsm, icount; double eer, eem; } oc_stats;

oc_ctx *oc_create(const oc_mesh *m, const oc_params *p);
void oc_destroy(oc_ctx *c);

/* profile arrays owned by the context (read/write access for tests) */
double *oc_field(oc_ctx *c, const char *name);
double oc_beta(oc_ctx *c);
void oc_get_stats(oc_ctx *c, oc_stats *s);
/* maximum outer GCR cycles of later oc_gcrk calls */
void oc_set_itr(oc_ctx *c, int itr);

/* reference subroutines */
void oc_derivative(oc_ctx *c, const double *unk, double *dpdxa, int idx);
void oc_laplc(oc_ctx *c, const double *press, double *alap, int impfl);
/* the FP32 operator of the mixed form (precision 2): laplc with the edge metrics rounded to
 * FP32 and the scaled gradient rounded to FP32 before the divergence (no reference
 * counterpart) */
void oc_laplc32(oc_ctx *c, const double *press, double *alap, int impfl);
/* the band line solve of the mixed forms (precon 3): residual, LU factors and substitution
 * in FP32; -1 when there is no band */
int oc_line_solve32(oc_ctx *c, const double *rhs, double *z);
/* mpdatam of the mixed forms: FP32 arithmetic on FP32 roundings of the advector, metrics,
 * closures, vol and rh0, FP32 increments subtracted from the FP64 field */
void oc_mpdatam32(oc_ctx *c, const double *uuu, double *unk, int iflag);
void oc_prfr(oc_ctx *c, const double *press, const double *ustr, const double *vstr,
             const double *wstr, double *uout, double *vout, double *wout, int impfl);
void oc_bcv(oc_ctx *c, double *u, double *v, double *w, double flag);
void oc_div(oc_ctx *c, const double *ua, const double *va, const double *wa, double *rhs, int iflg);
void oc_precon(oc_ctx *c, const double *rhs, double *p, double *r, int impfl);
void oc_gcrk(oc_ctx *c, double *p, const double *u1, const double *v1, const double *w1,
             double *u2, double *v2, double *w2, int impfl);
void oc_mpdatam(oc_ctx *c, const double *uuu /*[3][npoin]*/, double *unk, int iflag);
void oc_line_solve(oc_ctx *c, const double *rhs, double *z);

/* precon's multigrid branch (mg_flag=1, F:618-748). Level 1 = coarse mesh 1 with match1
 * (1-based fine node of each level-1 node), level 2 = coarse mesh 2 with match2 (1-based
 * level-1 node). transfer 0: restriction/prolongation as the reference indexes them (the
 * neighbour counter used as the node id, F:817, F:913); 1: the neighbour lists.
 * Returns -1 when a match is out of range or a coarse mesh is larger than its parent.
 * m1 == NULL turns the branch off. */
int oc_set_multigrid(oc_ctx *c, const oc_mesh *m1, const int32_t *match1, const oc_mesh *m2,
                     const int32_t *match2, int itrp2, int itrp_c1, int itrp_c2, int itrp_cc,
                     int transfer);
/* relaxation factor of the coarse-level point steps (no reference counterpart: the
 * reference uses omega on every level, which is also the default; <= 0 restores it) */
void oc_set_multigrid_omega(oc_ctx *c, double omega_c);
/* restriction_unstr(.,.,level) and prolongation_unstr(.,.,level) */
void oc_mg_restrict(oc_ctx *c, int level, const double *src, double *out);
void oc_mg_prolong(oc_ctx *c, int level, const double *coarse, double *out);
/* context of a level (0 = fine): its oc_laplc is laplc_c / laplc_cc, its "d1" field is
 * d2 / d3 and its oc_beta is betaunstr2 / betaunstr3 */
oc_ctx *oc_level(oc_ctx *c, int level);

/* NFT driver: state arrays are [npoin] blocks: unk[4], frc[4], uuu[3], unkdash[3], prpot */
void oc_nft_init(oc_ctx *c);
void oc_nft_step(oc_ctx *c);
/* per-step diagnostics: div(uuu,rhs,1), courant(uuu,cour,flag), stats(unk) -> out[25] */
void oc_diagnostics(oc_ctx *c, int flag, double *out);
double *oc_state(oc_ctx *c, const char *name);

#ifdef __cplusplus
}
#endif
#endif
