/* CPU oracle for the edge-based NFT step: TEST INFRASTRUCTURE ONLY (see nft_oracle.h).
 *
 * Every function below restates one reference subroutine of
 * lippsLFroude_implgwopks3D_unstr3.f ("F:n" = line n of that file) in its own loop order
 * and operation order. Compile with -ffp-contract=off so products and sums round as
 * the reference's scalar code does.
 */
#include "nft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double G = 9.80616; /* F:94 */

struct oc_ctx {
  oc_mesh m;
  oc_params p;
  int n;
  /* profiles and setup fields */
  double *rh0, *th0, *the, *ue, *ve, *coefw, *alpha, *d1, *dthedz;
  double beta1;
  /* line preconditioner: Thomas factors and operator rows */
  double *la, *linv, *lc;
  double *llo, *ldg, *lup;
  /* precon 3: in-column band of laplc (impfl 1) by rows, bm2[i] = L[i][i-2] ... bp2[i] =
   * L[i][i+2], and its LU factors (unit lower l1, l2; upper 1/d, u1, u2); band_L = column
   * length (0: not built) */
  double *bm2, *bm1, *b0, *bp1, *bp2, *fl1, *fl2, *fu1, *fu2, *fdinv;
  int band_L, band_colors;
  int line_L;   /* uniform column length <= 64 (wavefront scans), else 0 */
  int line_R;   /* rows per lane of the coarsened scans for uniform 64 < L <= 512, else 0 (Thomas) */
  /* scratch */
  double *s[16];
  /* gcrk work */
  double *gx, *gax, *gr, *gqr, *gar, *gb[4];
  /* nft state */
  double *unk, *frc, *uuu, *unkdash, *prpot, *unk1, *uuu0;
  oc_stats st;
  /* launch geometry of the HIP path (nft_init) for dot_order = 1 */
  int nchunk, nblocks, xcd_per, ggrid, vgrid;
  double *part;
  /* neighbour lists in neigh_definition's order (F:2726-2767): interior edges, then the
   * boundary-face edges, each in input order */
  int32_t *nb_ptr, *nb_idx;
  /* multigrid branch of precon (F:618-748): coarse levels 1 and 2, each a context of its
   * own mesh; match[l] = 0-based node of level l-1 (0 = fine) for each node of level l,
   * imatch[l] the inverse over level l-1 (-1 where no coarse node sits) */
  struct oc_ctx *lv[2];
  int32_t *lv_match[2], *lv_imatch[2];
  int mg_on, mg_itrp2, mg_c1, mg_c2, mg_cc, mg_transfer;
  double mg_omega;   /* relaxation factor of the coarse levels (the reference: omega) */
  double *mg_w[12];
  /* velprd work (itraj = 1): uuu1[3], ff[3], pprpot */
  double *vp[7];
  /* periodic box: vol summed over the coincident sets (c->m.vol points here), and the
   * set sums of one closure */
  double *volc, *cuc;
};

static double amax1(double a, double b) { return a > b ? a : b; }
static double amin1(double a, double b) { return a < b ? a : b; }

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* cyclicr, cyclicr2, cyclicr4 (F:4377-4446), called in that order: every node of a
 * coincident set takes the sum of the set's values, each list's sums formed before any is
 * stored */
static void cyclic_all(oc_ctx *c, double *u) {
  const oc_mesh *m = &c->m;
  double *uc = c->cuc;
  for (int ic = 0; ic < m->ncoin; ++ic) uc[ic] = u[m->icoin[ic] - 1] + u[m->icoin[m->ncoin + ic] - 1];
  for (int ic = 0; ic < m->ncoin; ++ic) { u[m->icoin[ic] - 1] = uc[ic]; u[m->icoin[m->ncoin + ic] - 1] = uc[ic]; }
  for (int ic = 0; ic < m->ncoin2; ++ic) uc[ic] = u[m->icoin2[ic] - 1] + u[m->icoin2[m->ncoin2 + ic] - 1];
  for (int ic = 0; ic < m->ncoin2; ++ic) { u[m->icoin2[ic] - 1] = uc[ic]; u[m->icoin2[m->ncoin2 + ic] - 1] = uc[ic]; }
  const int q = m->ncoin4;
  for (int ic = 0; ic < q; ++ic) {
    const int32_t *s4 = m->icoin4;
    uc[ic] = u[s4[ic] - 1] + u[s4[q + ic] - 1] + u[s4[2 * q + ic] - 1] + u[s4[3 * q + ic] - 1];
  }
  for (int ic = 0; ic < q; ++ic)
    for (int k = 0; k < 4; ++k) u[m->icoin4[k * q + ic] - 1] = uc[ic];
}

static void build_line(oc_ctx *c);
static void build_band(oc_ctx *c);

/* compute_d_unstr (F:1112-1284) over the context's own mesh with density rh0_u */
static void compute_d(oc_ctx *c, const double *rh0u, double *d) {
  const oc_mesh *m = &c->m;
  int n = c->n;
  double *dix = c->s[0], *diy = c->s[1], *diz = c->s[2];
  for (int i = 0; i < n; ++i) { dix[i] = 0.; diy[i] = 0.; diz[i] = 0.; }
  for (int pass = 0; pass < 2; ++pass) {
    int ne = pass == 0 ? m->nedge : m->nbface;
    const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
    const double *S = pass == 0 ? m->sn : m->sb;
    for (int e = 0; e < ne; ++e) {
      int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
      double sx = S[e], sy = S[ne + e], sz = S[2 * (size_t)ne + e];
      dix[i1] = dix[i1] + sx * sx * rh0u[i2] / m->vol[i2];
      dix[i2] = dix[i2] + sx * sx * rh0u[i1] / m->vol[i1];
      diy[i1] = diy[i1] + sy * sy * rh0u[i2] / m->vol[i2];
      diy[i2] = diy[i2] + sy * sy * rh0u[i1] / m->vol[i1];
      diz[i1] = diz[i1] + sz * sz * rh0u[i2] / m->vol[i2];
      diz[i2] = diz[i2] + sz * sz * rh0u[i1] / m->vol[i1];
    }
  }
  for (int i = 0; i < n; ++i)
    d[i] = -(0.5 * c->p.dt * (dix[i] + diy[i] + diz[i]) / (4 * m->vol[i] * rh0u[i]));
}

oc_ctx *oc_create(const oc_mesh *m, const oc_params *p) {
  oc_ctx *c = (oc_ctx *)calloc(1, sizeof(oc_ctx));
  c->m = *m;
  c->p = *p;
  int n = m->npoin;
  c->n = n;
  if (p->iper == 2) {
    /* the periodic branches exist in FP64 only, and need the coincident sets */
    if (p->precision != 0 || (m->ncoin && !m->icoin) || (m->ncoin2 && !m->icoin2) || (m->ncoin4 && !m->icoin4)) {
      free(c);
      return NULL;
    }
    int mc = m->ncoin > m->ncoin2 ? m->ncoin : m->ncoin2;
    c->cuc = zalloc((size_t)(mc > m->ncoin4 ? mc : m->ncoin4));
    /* mpsolver F:99-103: vol summed over the coincident sets before anything reads it */
    c->volc = zalloc(n);
    memcpy(c->volc, m->vol, sizeof(double) * (size_t)n);
    cyclic_all(c, c->volc);
    c->m.vol = c->volc;
  } else {
    c->m.ncoin = c->m.ncoin2 = c->m.ncoin4 = 0;
  }
  c->rh0 = zalloc(n); c->th0 = zalloc(n); c->the = zalloc(n); c->ue = zalloc(n);
  c->ve = zalloc(n); c->coefw = zalloc(n); c->alpha = zalloc(n); c->d1 = zalloc(n);
  c->dthedz = zalloc(n);
  c->la = zalloc(n); c->linv = zalloc(n); c->lc = zalloc(n);
  c->llo = zalloc(n); c->ldg = zalloc(n); c->lup = zalloc(n);
  for (int i = 0; i < 16; ++i) c->s[i] = zalloc(n);
  int lord = p->lord > 0 ? p->lord : 1;
  c->gx = zalloc((size_t)lord * n); c->gax = zalloc((size_t)lord * n);
  c->gr = zalloc(n); c->gqr = zalloc(n); c->gar = zalloc(n);
  for (int i = 0; i < 4; ++i) c->gb[i] = zalloc(n);
  c->unk = zalloc(4 * (size_t)n); c->frc = zalloc(4 * (size_t)n); c->uuu = zalloc(3 * (size_t)n);
  c->unkdash = zalloc(3 * (size_t)n); c->prpot = zalloc(n); c->unk1 = zalloc(4 * (size_t)n);
  c->uuu0 = zalloc(3 * (size_t)n);
  if (p->itraj == 1)
    for (int i = 0; i < 7; ++i) c->vp[i] = zalloc(n);

  /* rhprof/thprof with lipps=0 (F:4549-4579, F:4532-4547); metric factors are 1 in the
   * Cartesian case (F:149-157) */
  for (int i = 0; i < n; ++i) {
    c->rh0[i] = 1.0;
    c->th0[i] = p->theta0;
    c->ue[i] = p->u0;
    c->ve[i] = p->v0;
    /* F:187-190 */
    c->the[i] = p->theta0 * (1. + p->strat * m->coord[2 * (size_t)n + i]);
    c->dthedz[i] = p->theta0 * p->strat * 1.0;
  }
  /* absorbers (F:2915-3026) */
  {
    double *alpha = c->alpha;
    const double *X = m->coord, *Y = m->coord + n, *Z = m->coord + 2 * (size_t)n;
    double eps = 1.e-07;
    if (p->irl == 1) {
      double abxmax = -1.e+07, abxmin = 1.e+07, abymax = -1.e+07, abymin = 1.e+07;
      double abzmax = -1.e+07, abzmin = 1.e+07;
      for (int i = 0; i < n; ++i) {
        abxmax = amax1(X[i], abxmax); abxmin = amin1(X[i], abxmin);
        abymax = amax1(Y[i], abymax); abymin = amin1(Y[i], abymin);
        abzmax = amax1(Z[i], abzmax); abzmin = amin1(Z[i], abzmin);
      }
      double r = 1.0;
      double xabr = abxmax * r - p->dxabr, xabl = abxmin * r + p->dxabl;
      double yabu = abymax * r - p->dyabu, yabd = abymin * r + p->dyabd;
      double zabu = abzmax - p->dzabu, zabd = abzmin + p->dzabd;
      double tolil = p->towxl != 0. ? 1. / p->towxl : 0.;
      double tolir = p->towxr != 0. ? 1. / p->towxr : 0.;
      double toliu = p->towyu != 0. ? 1. / p->towyu : 0.;
      double tolid = p->towyd != 0. ? 1. / p->towyd : 0.;
      double toliuz = p->towzu != 0. ? 1. / p->towzu : 0.;
      double tolidz = p->towzd != 0. ? 1. / p->towzd : 0.;
      for (int i = 0; i < n; ++i) {
        double relx, rely, relz, r0, r1, r2, relb;
        r0 = X[i] * r; r1 = 0.; r2 = 0.; relb = 0.;
        if (p->dxabl != 0.) r1 = amax1(0., xabl - r0) / p->dxabl;
        if (p->dxabr != 0.) r2 = amax1(0., r0 - xabr) / p->dxabr;
        if ((tolil + tolir) != 0.)
          relb = ((tolil * r1) * (tolil * r1) + (tolir * r2) * (tolir * r2)) / (tolil * r1 + tolir * r2 + eps);
        relx = relb;
        r1 = 0.; r2 = 0.; relb = 0.; r0 = Y[i] * r;
        if (p->dyabd != 0.) r1 = amax1(0., -r0 + yabd) / p->dyabd;
        if (p->dyabu != 0.) r2 = amax1(0., r0 - yabu) / p->dyabu;
        if ((tolid + toliu) != 0.)
          relb = ((tolid * r1) * (tolid * r1) + (toliu * r2) * (toliu * r2)) / (tolid * r1 + toliu * r2 + eps);
        rely = relb;
        r1 = 0.; r2 = 0.; relb = 0.; r0 = Z[i];
        if (p->dzabd != 0.) r1 = amax1(0., -r0 + zabd) / p->dzabd;
        if (p->dzabu != 0.) r2 = amax1(0., r0 - zabu) / p->dzabu;
        if ((tolidz + toliuz) != 0.)
          relb = ((tolidz * r1) * (tolidz * r1) + (toliuz * r2) * (toliuz * r2)) / (tolidz * r1 + toliuz * r2 + eps);
        relz = relb;
        double alpha1 = (rely * rely + relx * relx) / (rely + relx + eps);
        alpha[i] = (alpha1 * alpha1 + relz * relz) / (alpha1 + relz + eps);
      }
    }
  }
  /* beta_timestep_unstr(1) (F:953-1109) */
  {
    double ddx = 10.e+06, hill = 1500.0;
    const double *X = m->coord, *Y = m->coord + n, *Z = m->coord + 2 * (size_t)n;
    const double *d = p->dom;
    for (int e = 0; e < m->nedge; ++e) {
      int i1 = m->iedge[e] - 1, i2 = m->iedge[m->nedge + e] - 1;
      if (X[i1] != d[0] && X[i1] != d[1] && Y[i1] != d[2] && Y[i1] != d[3] && Z[i1] != d[4] &&
          Z[i1] != d[5] && X[i2] != d[0] && X[i2] != d[1] && Y[i2] != d[2] && Y[i2] != d[3] &&
          Z[i2] != d[4] && Z[i2] != d[5] && Z[i1] > hill && Z[i2] > hill) {
        double hx = fabs(m->sn[e]), hy = fabs(m->sn[m->nedge + e]), hz = fabs(m->sn[2 * (size_t)m->nedge + e]);
        double h = amax1(amax1(hx, hy), hz);
        ddx = amin1(amin1(ddx, c->m.vol[i1] / h), c->m.vol[i2] / h);
      }
    }
    c->beta1 = 0.5 * ddx * ddx * 1.0;
  }
  /* compute_d_unstr(1) (F:1112-1284) */
  compute_d(c, c->rh0, c->d1);
  /* neigh_definition (F:2726-2767) */
  {
    c->nb_ptr = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? m->nedge : m->nbface;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      for (int e = 0; e < ne; ++e) { c->nb_ptr[ie[e]]++; c->nb_ptr[ie[ne + e]]++; }
    }
    for (int i = 0; i < n; ++i) c->nb_ptr[i + 1] += c->nb_ptr[i];
    c->nb_idx = (int32_t *)calloc((size_t)c->nb_ptr[n] + 1, sizeof(int32_t));
    int32_t *fill = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? m->nedge : m->nbface;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        c->nb_idx[c->nb_ptr[i1] + fill[i1]++] = i2;
        c->nb_idx[c->nb_ptr[i2] + fill[i2]++] = i1;
      }
    }
    free(fill);
  }
  /* coefw (F:250-252); set here so laplc/precon can be called stand-alone */
  for (int i = 0; i < n; ++i) {
    double gth0 = G / c->th0[i];
    double cw = 1. + 0.25 * (p->dt * p->dt) * gth0 * c->dthedz[i];
    c->coefw[i] = 1. / cw;
  }
  build_line(c);
  if (p->precon == 3) build_band(c);
  /* launch geometry of the HIP path (src_amd/csrc/nft_capi.hip, nft_init) */
  c->nchunk = (n + 63) / 64;
  c->nblocks = (c->nchunk + 3) / 4;
  if (c->nblocks >= 64) { c->xcd_per = (c->nblocks + 7) / 8; c->ggrid = 8 * c->xcd_per; }
  else { c->xcd_per = 0; c->ggrid = c->nblocks; }
  c->vgrid = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
  c->part = zalloc((size_t)(c->ggrid > c->vgrid ? c->ggrid : c->vgrid));
  return c;
}

/* ---- the HIP path's reduction tree (dot_order = 1) --------------------------------
 * wave: lane l += lane l+off for off = 32,16,...,1 (shfl_down), result in lane 0;
 * block: ((w0 + w1) + w2) + w3 over its four wavefronts;
 * grid: 1024 threads (16 one-wavefront workgroups); thread t sums partials t, t+1024, ...
 * in order, then the wavefront trees and the 16 wavefront results in order. Inactive
 * lanes hold 0. */
static double tree64(double *v) {
  for (int off = 32; off > 0; off >>= 1)
    for (int l = 0; l < off; ++l) v[l] = v[l] + v[l + off];
  return v[0];
}

static double block_sum(const double *t256) {
  double w[64], r[4];
  for (int q = 0; q < 4; ++q) {
    for (int l = 0; l < 64; ++l) w[l] = t256[q * 64 + l];
    r[q] = tree64(w);
  }
  double a = r[0];
  for (int j = 1; j < 4; ++j) a = a + r[j];
  return a;
}

static double grid_sum(const double *part, int nblk) {
  double t[1024], w[64];
  for (int th = 0; th < 1024; ++th) {
    double acc = 0.0;
    for (int b = th; b < nblk; b += 1024) acc = acc + part[b];
    t[th] = acc;
  }
  double a = 0.0;
  for (int q = 0; q < 16; ++q) {
    for (int l = 0; l < 64; ++l) w[l] = t[q * 64 + l];
    double r = tree64(w);
    a = q == 0 ? r : a + r;
  }
  return a;
}

/* sum over nodes of a[i]*b[i] (a NULL: b[i]*b[i]) as the fused dots of the gather kernels */
static double dsum_gather(oc_ctx *c, const double *a, const double *b) {
  double t[256];
  for (int pb = 0; pb < c->ggrid; ++pb) {
    int lb = c->xcd_per > 0 ? (pb & 7) * c->xcd_per + (pb >> 3) : pb;
    for (int th = 0; th < 256; ++th) {
      double v = 0.0;
      if (lb < c->nblocks) {
        int chunk = lb * 4 + th / 64;
        int i = chunk * 64 + th % 64;
        if (chunk < c->nchunk && i < c->n) v = (a ? a[i] : b[i]) * b[i];
      }
      t[th] = v;
    }
    c->part[pb] = block_sum(t);
  }
  return grid_sum(c->part, c->ggrid);
}

/* sum of a[k]*b[k] as the grid-stride vector kernels accumulate it */
static double dsum_stride(oc_ctx *c, const double *a, const double *b) {
  double t[256];
  const int stride = c->vgrid * 256;
  for (int pb = 0; pb < c->vgrid; ++pb) {
    for (int th = 0; th < 256; ++th) {
      double v = 0.0;
      for (int k = pb * 256 + th; k < c->n; k += stride) v = v + a[k] * b[k];
      t[th] = v;
    }
    c->part[pb] = block_sum(t);
  }
  return grid_sum(c->part, c->vgrid);
}

void oc_destroy(oc_ctx *c) {
  if (!c) return;
  double *all[] = {c->rh0, c->th0, c->the, c->ue, c->ve, c->coefw, c->alpha, c->d1, c->dthedz,
                   c->la, c->linv, c->lc, c->gx, c->gax, c->gr, c->gqr, c->gar, c->unk, c->frc,
                   c->uuu, c->unkdash, c->prpot, c->unk1, c->uuu0, c->part, c->llo, c->ldg, c->lup};
  for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); ++i) free(all[i]);
  for (int i = 0; i < 16; ++i) free(c->s[i]);
  for (int i = 0; i < 4; ++i) free(c->gb[i]);
  free(c->nb_ptr); free(c->nb_idx);
  for (int i = 0; i < 7; ++i) free(c->vp[i]);
  {
    double *band[] = {c->bm2, c->bm1, c->b0, c->bp1, c->bp2, c->fl1, c->fl2, c->fu1, c->fu2, c->fdinv};
    for (size_t i = 0; i < sizeof(band) / sizeof(band[0]); ++i) free(band[i]);
  }
  for (int l = 0; l < 2; ++l) { oc_destroy(c->lv[l]); free(c->lv_match[l]); free(c->lv_imatch[l]); }
  for (int i = 0; i < 12; ++i) free(c->mg_w[i]);
  free(c->volc); free(c->cuc);
  free(c);
}

double *oc_field(oc_ctx *c, const char *name) {
  if (!strcmp(name, "vol")) return (double *)c->m.vol;   /* closed over the coincident sets in the periodic box */
  if (!strcmp(name, "rh0")) return c->rh0;
  if (!strcmp(name, "th0")) return c->th0;
  if (!strcmp(name, "the")) return c->the;
  if (!strcmp(name, "ue")) return c->ue;
  if (!strcmp(name, "ve")) return c->ve;
  if (!strcmp(name, "coefw")) return c->coefw;
  if (!strcmp(name, "alpha")) return c->alpha;
  if (!strcmp(name, "d1")) return c->d1;
  if (!strcmp(name, "dthedz")) return c->dthedz;
  if (!strcmp(name, "line_a")) return c->la;
  if (!strcmp(name, "line_inv")) return c->linv;
  if (!strcmp(name, "line_c")) return c->lc;
  if (!strcmp(name, "line_lo")) return c->llo;
  if (!strcmp(name, "line_dg")) return c->ldg;
  if (!strcmp(name, "line_up")) return c->lup;
  if (!strcmp(name, "band_m2")) return c->bm2;
  if (!strcmp(name, "band_m1")) return c->bm1;
  if (!strcmp(name, "band_0")) return c->b0;
  if (!strcmp(name, "band_p1")) return c->bp1;
  if (!strcmp(name, "band_p2")) return c->bp2;
  return NULL;
}

double oc_beta(oc_ctx *c) { return c->beta1; }
void oc_get_stats(oc_ctx *c, oc_stats *s) { *s = c->st; }
void oc_set_itr(oc_ctx *c, int itr) { c->p.itr = itr; }

/* derivative(unk,dpdxa,idx), F:3943-4022, Cartesian branches (closed and periodic); f32: the edge metrics sn, sb
 * rounded to FP32 first (the FP32 operator of the mixed form; sbb stays FP64) */
static void derivative_m(oc_ctx *c, const double *unk, double *dpdxa, int idx, int f32) {
  const oc_mesh *m = &c->m;
  int n = c->n, d = idx - 1;
  for (int i = 0; i < n; ++i) dpdxa[i] = 0.;
  for (int e = 0; e < m->nedge; ++e) {
    int i1 = m->iedge[e] - 1, i2 = m->iedge[m->nedge + e] - 1;
    double sx = m->sn[(size_t)d * m->nedge + e];
    if (f32) sx = (double)(float)sx;
    double u = (unk[i1] + unk[i2]) * 0.5;
    dpdxa[i1] = dpdxa[i1] + sx * u;
    dpdxa[i2] = dpdxa[i2] - sx * u;
  }
  for (int e = 0; e < m->nbface; ++e) {
    int i1 = m->kbface[e] - 1, i2 = m->kbface[m->nbface + e] - 1;
    double sx = m->sb[(size_t)d * m->nbface + e];
    if (f32) sx = (double)(float)sx;
    double u = (unk[i1] + unk[i2]) * 0.5;
    dpdxa[i1] = dpdxa[i1] + sx * u;
    dpdxa[i2] = dpdxa[i2] - sx * u;
  }
  if (c->p.iper == 2) {
    /* periodic branch (F:4015-4019): no boundary-point terms, sums over the sets */
    cyclic_all(c, dpdxa);
    return;
  }
  for (int b = 0; b < m->nboun; ++b) {
    int i1 = m->ibpoin[b] - 1;
    double sx = m->sbb[(size_t)d * m->nboun + b];
    double u = unk[i1];
    dpdxa[i1] = dpdxa[i1] + sx * u;
  }
}

void oc_derivative(oc_ctx *c, const double *unk, double *dpdxa, int idx) { derivative_m(c, unk, dpdxa, idx, 0); }

/* bcv(uout,vout,wout,flag), F:1583-1643, Cartesian branches: the periodic box
 * (F:1603-1612) keeps only the flag-11 slip projection */
void oc_bcv(oc_ctx *c, double *u, double *v, double *w, double flag) {
  const oc_mesh *m = &c->m;
  int nb = m->nboun;
  const int per = c->p.iper == 2;
  for (int b = 0; b < nb; ++b) {
    int ii = m->ibpoin[b] - 1, f = m->ibpoin[nb + b];
    double ax = m->anp[b], ay = m->anp[nb + b], az = m->anp[2 * (size_t)nb + b];
    if (per) {
      if (f == 11) {
        double vn = u[ii] * ax + v[ii] * ay + w[ii] * az;
        u[ii] = u[ii] - ax * vn;
        v[ii] = v[ii] - ay * vn;
        w[ii] = w[ii] - az * vn;
      }
      continue;
    }
    if (f == 11 || f == 5 || f == 6) {
      double vn = u[ii] * ax + v[ii] * ay + w[ii] * az;
      u[ii] = u[ii] - ax * vn;
      v[ii] = v[ii] - ay * vn;
      w[ii] = w[ii] - az * vn;
    }
    if (f == 7 || f == 8) u[ii] = c->ue[ii] * flag;
  }
}

/* laplc(press,alap,impfl), F:1288-1353; f32: the mixed form's FP32 operator (edge metrics
 * rounded to FP32, the scaled gradient rounded to FP32 before the divergence, FP64 sums) */
static void laplc_m(oc_ctx *c, const double *press, double *alap, int impfl, int f32) {
  int n = c->n;
  double *gx = c->s[3], *gy = c->s[4], *gz = c->s[5];
  double dt = c->p.dt;
  for (int i = 0; i < n; ++i) alap[i] = 0.;
  derivative_m(c, press, gx, 1, f32);
  derivative_m(c, press, gy, 2, f32);
  derivative_m(c, press, gz, 3, f32);
  for (int i = 0; i < n; ++i) {
    double v = c->m.vol[i];
    double fz = impfl ? c->coefw[i] : 1.0; /* coefw*impfl+(1-impfl) */
    gx[i] = dt * gx[i] / v;
    gy[i] = dt * gy[i] / v;
    gz[i] = dt * gz[i] / v;
    gx[i] = 0.5 * gx[i];
    gy[i] = 0.5 * gy[i];
    gz[i] = 0.5 * gz[i] * fz;
  }
  oc_bcv(c, gx, gy, gz, 0.);
  for (int i = 0; i < n; ++i) {
    gx[i] = gx[i] * c->rh0[i];
    gy[i] = gy[i] * c->rh0[i];
    gz[i] = gz[i] * c->rh0[i];
    if (f32) { gx[i] = (double)(float)gx[i]; gy[i] = (double)(float)gy[i]; gz[i] = (double)(float)gz[i]; }
  }
  double *ax = c->s[6], *ay = c->s[7], *az = c->s[8];
  derivative_m(c, gx, ax, 1, f32);
  derivative_m(c, gy, ay, 2, f32);
  derivative_m(c, gz, az, 3, f32);
  for (int i = 0; i < n; ++i) alap[i] = (ax[i] + ay[i] + az[i]) / (c->m.vol[i] * c->rh0[i]);
}

void oc_laplc(oc_ctx *c, const double *press, double *alap, int impfl) { laplc_m(c, press, alap, impfl, 0); }
void oc_laplc32(oc_ctx *c, const double *press, double *alap, int impfl) { laplc_m(c, press, alap, impfl, 1); }

/* prfr(press,ustr,vstr,wstr,uout,vout,wout,impfl), F:1536-1580 */
void oc_prfr(oc_ctx *c, const double *press, const double *ustr, const double *vstr,
             const double *wstr, double *uout, double *vout, double *wout, int impfl) {
  int n = c->n;
  double *gx = c->s[9], *gy = c->s[10], *gz = c->s[11];
  double dt = c->p.dt;
  oc_derivative(c, press, gx, 1);
  oc_derivative(c, press, gy, 2);
  oc_derivative(c, press, gz, 3);
  for (int i = 0; i < n; ++i) {
    double v = c->m.vol[i];
    gx[i] = gx[i] / v;
    gy[i] = gy[i] / v;
    gz[i] = gz[i] / v;
  }
  for (int i = 0; i < n; ++i) {
    double fz = impfl ? c->coefw[i] : 1.0;
    uout[i] = ustr[i] - 0.5 * dt * gx[i];
    vout[i] = vstr[i] - 0.5 * dt * gy[i];
    wout[i] = wstr[i] - 0.5 * dt * gz[i] * fz;
  }
  oc_bcv(c, uout, vout, wout, 1.);
}

/* div(ua,va,wa,rhs,iflg), F:2096-2146 (diagnostic print of iflg=1 omitted) */
void oc_div(oc_ctx *c, const double *ua, const double *va, const double *wa, double *rhs, int iflg) {
  int n = c->n;
  double *ub = c->s[12], *vb = c->s[13], *wb = c->s[14];
  for (int i = 0; i < n; ++i) {
    if (iflg == -1) { ub[i] = ua[i] * c->rh0[i]; vb[i] = va[i] * c->rh0[i]; wb[i] = wa[i] * c->rh0[i]; }
    else { ub[i] = ua[i]; vb[i] = va[i]; wb[i] = wa[i]; }
  }
  double *dx = c->s[3], *dy = c->s[4], *dz = c->s[5];
  oc_derivative(c, ub, dx, 1);
  oc_derivative(c, vb, dy, 2);
  oc_derivative(c, wb, dz, 3);
  double fl = (double)iflg;
  for (int i = 0; i < n; ++i) {
    double v = c->m.vol[i];
    rhs[i] = fl * (dx[i] / v + dy[i] / v + dz[i] / v) / c->rh0[i];
  }
}

/* ---- vertical line preconditioner (no reference counterpart) ----------------------
 * Two-point-flux edge Laplacian: for edge (i,j) with dual facet S and d = x_j - x_i,
 *   c_ij = 0.5*dt * 0.5*(rho_i+rho_j) * (Sx^2 + Sy^2 + cw*Sz^2) / max(S.d, 1e-3|S||d|)
 *          / (vol_i * rho_i),   cw = 0.5*(coefw_i+coefw_j)
 * M_ii = -sum_j c_ij over all edges of i; M_{i,i+-1} = c_ij over the vertical edges that
 * join consecutive nodes of a column. Per column the tridiagonal M is factorised once
 * (Thomas): a_k = M_{k,k-1}, cp_k = M_{k,k+1} * inv_k, inv_k = 1/(M_kk - a_k cp_{k-1}). */
static void build_line(oc_ctx *c) {
  const oc_mesh *m = &c->m;
  int n = c->n;
  double *diag = c->s[0], *lo = c->s[1], *up = c->s[2];
  for (int i = 0; i < n; ++i) { diag[i] = 0.; lo[i] = 0.; up[i] = 0.; }
  if (m->ncol <= 0) return;
  const double *X = m->coord, *Y = m->coord + n, *Z = m->coord + 2 * (size_t)n;
  double hdt = 0.5 * c->p.dt;
  /* column position of every node: (col, level) */
  int *colof = (int *)malloc(sizeof(int) * (size_t)n);
  int *levof = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; ++i) { colof[i] = -1; levof[i] = -1; }
  for (int k = 0; k < m->ncol; ++k)
    for (int i = m->col_start[k]; i < m->col_start[k + 1]; ++i) { colof[i] = k; levof[i] = i - m->col_start[k]; }
  for (int pass = 0; pass < 2; ++pass) {
    int ne = pass == 0 ? m->nedge : m->nbface;
    const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
    const double *S = pass == 0 ? m->sn : m->sb;
    for (int e = 0; e < ne; ++e) {
      int i = ie[e] - 1, j = ie[ne + e] - 1;
      double sx = S[e], sy = S[ne + e], sz = S[2 * (size_t)ne + e];
      double dx = X[j] - X[i], dy = Y[j] - Y[i], dz = Z[j] - Z[i];
      double sd = sx * dx + sy * dy + sz * dz;
      double sl = sqrt(sx * sx + sy * sy + sz * sz), dl = sqrt(dx * dx + dy * dy + dz * dz);
      double den = sd > 1e-3 * sl * dl ? sd : 1e-3 * sl * dl;
      double cw = 0.5 * (c->coefw[i] + c->coefw[j]);
      double rb = 0.5 * (c->rh0[i] + c->rh0[j]);
      double k0 = hdt * rb * (sx * sx + sy * sy + cw * sz * sz) / den;
      double ci = k0 / (m->vol[i] * c->rh0[i]);
      double cj = k0 / (m->vol[j] * c->rh0[j]);
      diag[i] -= ci;
      diag[j] -= cj;
      if (colof[i] >= 0 && colof[i] == colof[j]) {
        if (levof[j] == levof[i] + 1) { up[i] = ci; lo[j] = cj; }
        else if (levof[i] == levof[j] + 1) { lo[i] = ci; up[j] = cj; }
      }
    }
  }
  for (int k = 0; k < m->ncol; ++k) {
    int s0 = m->col_start[k], s1 = m->col_start[k + 1];
    double cprev = 0.;
    for (int i = s0; i < s1; ++i) {
      double a = (i == s0) ? 0. : lo[i];
      double den = diag[i] - a * cprev;
      double inv = 1. / den;
      double cp = (i == s1 - 1) ? 0. : up[i] * inv;
      c->la[i] = a; c->linv[i] = inv; c->lc[i] = cp;
      cprev = cp;
      c->llo[i] = a; c->ldg[i] = diag[i]; c->lup[i] = (i == s1 - 1) ? 0. : up[i];
    }
  }
  free(colof);
  free(levof);
  int L0 = m->col_start[1] - m->col_start[0];
  int uniform = L0 <= 64;
  for (int k = 0; k < m->ncol && uniform; ++k)
    if (m->col_start[k + 1] - m->col_start[k] != L0 || m->col_start[k] != k * L0) uniform = 0;
  c->line_L = uniform ? L0 : 0;
  c->line_R = 0;
  if (!uniform && m->ncol > 0) {
    int longcols = L0 > 64 && L0 <= 512;
    for (int k = 0; k < m->ncol && longcols; ++k)
      if (m->col_start[k + 1] - m->col_start[k] != L0 || m->col_start[k] != k * L0) longcols = 0;
    if (longcols) c->line_R = (L0 + 63) / 64;
  }
}

/* One column of L <= 64 rows in a 64-lane frame (rows past L carry the zero map), solved
 * with the fixed Thomas factors as two scans of affine maps, in the operation order of the
 * HIP path's wavefront solver (k_line_scan):
 *   forward  (beta, alpha) = (d*inv, -(a*inv)); for s = 1..32, lanes l >= s:
 *            beta = beta + alpha*beta[l-s]; alpha = alpha*alpha[l-s]   (old values)
 *   backward (beta, gamma) = (y, -c');         for s = 1..32, lanes l+s < 64:
 *            beta = beta + gamma*beta[l+s]; gamma = gamma*gamma[l+s]
 * A column's result does not depend on its lane offset in a wavefront: alpha = 0 at its
 * first row and gamma = 0 at its last row turn every cross-column term into an exact 0. */
/* precon 3: the in-column band of laplc(., impfl=1) probed column-color by column-color.
 * laplc couples nodes at most two edges apart, so with probe nodes on columns that are
 * pairwise at least three column hops apart (greedy colouring in column order) and on
 * levels five apart, entry (row, probe) of laplc(indicator) is the only nonzero term of
 * each row within two levels of a probe. The band is factorised per column (LU without
 * pivoting, band width 2). Uniform columns only (band_L = 0 otherwise). */
static int cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

static void build_band(oc_ctx *c) {
  const oc_mesh *m = &c->m;
  int n = c->n, nc = m->ncol;
  c->band_L = 0;
  if (nc <= 0 || n % nc) return;
  int L = n / nc;
  for (int k = 0; k <= nc; ++k)
    if (m->col_start[k] != k * L) return;
  /* column adjacency (CSR, unique neighbours) */
  int32_t *deg = (int32_t *)calloc((size_t)nc + 1, sizeof(int32_t));
  for (int pass = 0; pass < 2; ++pass) {
    int ne = pass == 0 ? m->nedge : m->nbface;
    const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
    for (int e = 0; e < ne; ++e) {
      int a = (ie[e] - 1) / L, b = (ie[ne + e] - 1) / L;
      if (a != b) { deg[a + 1]++; deg[b + 1]++; }
    }
  }
  for (int k = 0; k < nc; ++k) deg[k + 1] += deg[k];
  int32_t *adj = (int32_t *)malloc(sizeof(int32_t) * ((size_t)deg[nc] + 1));
  int32_t *fill = (int32_t *)calloc((size_t)nc, sizeof(int32_t));
  for (int pass = 0; pass < 2; ++pass) {
    int ne = pass == 0 ? m->nedge : m->nbface;
    const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
    for (int e = 0; e < ne; ++e) {
      int a = (ie[e] - 1) / L, b = (ie[ne + e] - 1) / L;
      if (a != b) { adj[deg[a] + fill[a]++] = b; adj[deg[b] + fill[b]++] = a; }
    }
  }
  /* unique neighbours per column (each column pair is joined by many edges) */
  {
    int32_t *uptr = (int32_t *)calloc((size_t)nc + 1, sizeof(int32_t));
    int32_t *uadj = (int32_t *)malloc(sizeof(int32_t) * ((size_t)deg[nc] + 1));
    int32_t w = 0;
    for (int k = 0; k < nc; ++k) {
      int32_t a0 = deg[k], a1 = deg[k + 1];
      /* sort the row, then copy without repeats */
      qsort(adj + a0, (size_t)(a1 - a0), sizeof(int32_t), cmp_i32);
      for (int32_t q = a0; q < a1; ++q)
        if (q == a0 || adj[q] != adj[q - 1]) uadj[w++] = adj[q];
      uptr[k + 1] = w;
    }
    free(deg); free(adj);
    deg = uptr; adj = uadj;
  }
  /* periodic box: twin columns (those holding the nodes of one coincident set) */
  int32_t *tptr = NULL, *tadj = NULL;
  if (c->p.iper == 2) {
    tptr = (int32_t *)calloc((size_t)nc + 1, sizeof(int32_t));
    const int32_t *lists[3] = {m->icoin, m->icoin2, m->icoin4};
    const int cnt[3] = {m->ncoin, m->ncoin2, m->ncoin4}, w[3] = {2, 2, 4};
    for (int pass = 0; pass < 2; ++pass) {
      int32_t *tf = pass ? (int32_t *)calloc((size_t)nc, sizeof(int32_t)) : NULL;
      for (int t = 0; t < 3; ++t)
        for (int ic = 0; ic < cnt[t]; ++ic)
          for (int a = 0; a < w[t]; ++a)
            for (int b = 0; b < w[t]; ++b) {
              int ca = (lists[t][a * cnt[t] + ic] - 1) / L, cb = (lists[t][b * cnt[t] + ic] - 1) / L;
              if (ca == cb) continue;
              if (pass == 0) tptr[ca + 1]++;
              else tadj[tptr[ca] + tf[ca]++] = cb;
            }
      if (pass == 0) {
        for (int k = 0; k < nc; ++k) tptr[k + 1] += tptr[k];
        tadj = (int32_t *)malloc(sizeof(int32_t) * ((size_t)tptr[nc] + 1));
      }
      free(tf);
    }
  }
  /* greedy colouring: a column takes the smallest colour absent among the columns its
   * probe reaches, within two column hops; in the periodic box the closures join twin
   * columns as well, so the reach is twin, hop, twin, hop, twin from the column */
  int *color = (int *)malloc(sizeof(int) * (size_t)nc);
  int *mark = (int *)calloc(4096, sizeof(int));
  int32_t *reach = tptr ? (int32_t *)malloc(sizeof(int32_t) * (size_t)nc) : NULL;
  int *seen = tptr ? (int *)calloc((size_t)nc, sizeof(int)) : NULL;
  int ncolors = 0;
  for (int k = 0; k < nc; ++k) color[k] = -1;
  for (int k = 0; k < nc; ++k) {
    if (tptr) {
      int cnt = 1;
      reach[0] = k;
      seen[k] = k + 1;
      for (int step = 0; step < 5; ++step) {
        const int32_t *ptr = step % 2 == 0 ? tptr : deg, *ad = step % 2 == 0 ? tadj : adj;
        int c0 = cnt;
        for (int q = 0; q < c0; ++q)
          for (int e = ptr[reach[q]]; e < ptr[reach[q] + 1]; ++e)
            if (seen[ad[e]] != k + 1) { seen[ad[e]] = k + 1; reach[cnt++] = ad[e]; }
      }
      for (int q = 1; q < cnt; ++q)
        if (color[reach[q]] >= 0 && color[reach[q]] < 4096) mark[color[reach[q]]] = k + 1;
    } else {
      for (int q = deg[k]; q < deg[k + 1]; ++q) {
        int a = adj[q];
        if (color[a] >= 0 && color[a] < 4096) mark[color[a]] = k + 1;
        for (int q2 = deg[a]; q2 < deg[a + 1]; ++q2) {
          int b = adj[q2];
          if (color[b] >= 0 && color[b] < 4096) mark[color[b]] = k + 1;
        }
      }
    }
    int col = 0;
    while (col < 4095 && mark[col] == k + 1) col++;
    color[k] = col;
    if (col + 1 > ncolors) ncolors = col + 1;
  }
  double **bands[] = {&c->bm2, &c->bm1, &c->b0, &c->bp1, &c->bp2, &c->fl1, &c->fl2, &c->fu1, &c->fu2, &c->fdinv};
  for (size_t q = 0; q < sizeof(bands) / sizeof(bands[0]); ++q) *bands[q] = zalloc(n);
  double *v = c->s[9], *Lv = c->s[10];
  for (int col = 0; col < ncolors; ++col) {
    for (int r = 0; r < 5; ++r) {
      for (int i = 0; i < n; ++i) v[i] = 0.;
      for (int k = 0; k < nc; ++k)
        if (color[k] == col)
          for (int l = r; l < L; l += 5) v[k * L + l] = 1.;
      oc_laplc(c, v, Lv, 1);
      for (int k = 0; k < nc; ++k) {
        if (color[k] != col) continue;
        for (int l = r; l < L; l += 5) {
          int i = k * L + l;
          c->b0[i] = Lv[i];
          if (l + 1 < L) c->bm1[i + 1] = Lv[i + 1];
          if (l + 2 < L) c->bm2[i + 2] = Lv[i + 2];
          if (l >= 1) c->bp1[i - 1] = Lv[i - 1];
          if (l >= 2) c->bp2[i - 2] = Lv[i - 2];
        }
      }
    }
  }
  /* LU per column: row k = (l2, l1, 1) x U, U row k = (d, u1, u2) */
  for (int k = 0; k < nc; ++k) {
    for (int l = 0; l < L; ++l) {
      int i = k * L + l;
      double l2 = 0., l1 = 0., t1 = c->bm1[i], d = c->b0[i];
      if (l >= 2) { l2 = c->bm2[i] * c->fdinv[i - 2]; t1 = t1 - l2 * c->fu1[i - 2]; d = d - l2 * c->fu2[i - 2]; }
      if (l >= 1) { l1 = t1 * c->fdinv[i - 1]; d = d - l1 * c->fu1[i - 1]; }
      double u1 = l + 1 < L ? c->bp1[i] : 0.;
      if (l >= 1 && l + 1 < L) u1 = u1 - l1 * c->fu2[i - 1];
      c->fl1[i] = l1; c->fl2[i] = l2;
      c->fu1[i] = u1;
      c->fu2[i] = l + 2 < L ? c->bp2[i] : 0.;
      c->fdinv[i] = 1. / d;
    }
  }
  c->band_L = L;
  c->band_colors = ncolors;
  free(deg); free(adj); free(fill); free(color); free(mark);
  free(tptr); free(tadj); free(reach); free(seen);
}

static void scan_column(const double *la, const double *linv, const double *lc, const double *rhs, int L,
                        double *z) {
  double al[64], be[64], ga[64], na[64], nb[64];
  for (int l = 0; l < 64; ++l) {
    if (l < L) { al[l] = -(la[l] * linv[l]); be[l] = rhs[l] * linv[l]; ga[l] = -lc[l]; }
    else { al[l] = 0.; be[l] = 0.; ga[l] = 0.; }
  }
  for (int s = 1; s < 64; s <<= 1) {
    for (int l = 0; l < 64; ++l) {
      if (l >= s) { nb[l] = be[l] + al[l] * be[l - s]; na[l] = al[l] * al[l - s]; }
      else { nb[l] = be[l]; na[l] = al[l]; }
    }
    for (int l = 0; l < 64; ++l) { be[l] = nb[l]; al[l] = na[l]; }
  }
  for (int s = 1; s < 64; s <<= 1) {
    for (int l = 0; l < 64; ++l) {
      if (l + s < 64) { nb[l] = be[l] + ga[l] * be[l + s]; na[l] = ga[l] * ga[l + s]; }
      else { nb[l] = be[l]; na[l] = ga[l]; }
    }
    for (int l = 0; l < 64; ++l) { be[l] = nb[l]; ga[l] = na[l]; }
  }
  for (int l = 0; l < L; ++l) z[l] = be[l];
}

/* One column of 64 < L <= 64*R rows, lane l holding rows l*R .. l*R+R-1 (rows >= L the
 * zero map), in the operation order of the HIP path's coarsened scan (k_line_scan_r):
 * each lane composes its rows' maps in order into (A, B) from (1, 0); the 64 lane maps
 * are scanned as in scan_column; lane l replays its rows from y_in = B of lane l-1 (0 for
 * lane 0); then the same backwards with gamma = -c' from the last row down. */
static void scan_column_r(const double *la, const double *linv, const double *lc, const double *rhs, int L,
                          int R, double *z) {
  double al[512], be[512], ga[512];
  double A[64], B[64], nA[64], nB[64], yin[64];
  for (int q = 0; q < 64 * R; ++q) {
    if (q < L) { al[q] = -(la[q] * linv[q]); be[q] = rhs[q] * linv[q]; ga[q] = -lc[q]; }
    else { al[q] = 0.; be[q] = 0.; ga[q] = 0.; }
  }
  for (int l = 0; l < 64; ++l) {
    double a = 1., b = 0.;
    for (int r = 0; r < R; ++r) { b = be[l * R + r] + al[l * R + r] * b; a = al[l * R + r] * a; }
    A[l] = a; B[l] = b;
  }
  for (int s = 1; s < 64; s <<= 1) {
    for (int l = 0; l < 64; ++l) {
      if (l >= s) { nB[l] = B[l] + A[l] * B[l - s]; nA[l] = A[l] * A[l - s]; }
      else { nB[l] = B[l]; nA[l] = A[l]; }
    }
    for (int l = 0; l < 64; ++l) { A[l] = nA[l]; B[l] = nB[l]; }
  }
  for (int l = 0; l < 64; ++l) yin[l] = l > 0 ? B[l - 1] : 0.;
  for (int l = 0; l < 64; ++l) {
    double y = yin[l];
    for (int r = 0; r < R; ++r) { y = be[l * R + r] + al[l * R + r] * y; be[l * R + r] = y; }
  }
  for (int l = 0; l < 64; ++l) {
    double g = 1., b = 0.;
    for (int r = R - 1; r >= 0; --r) { b = be[l * R + r] + ga[l * R + r] * b; g = ga[l * R + r] * g; }
    A[l] = g; B[l] = b;
  }
  for (int s = 1; s < 64; s <<= 1) {
    for (int l = 0; l < 64; ++l) {
      if (l + s < 64) { nB[l] = B[l] + A[l] * B[l + s]; nA[l] = A[l] * A[l + s]; }
      else { nB[l] = B[l]; nA[l] = A[l]; }
    }
    for (int l = 0; l < 64; ++l) { A[l] = nA[l]; B[l] = nB[l]; }
  }
  for (int l = 0; l < 64; ++l) yin[l] = l < 63 ? B[l + 1] : 0.;
  for (int l = 0; l < 64; ++l) {
    double zz = yin[l];
    for (int r = R - 1; r >= 0; --r) {
      zz = be[l * R + r] + ga[l * R + r] * zz;
      if (l * R + r < L) z[l * R + r] = zz;
    }
  }
}

/* z = M^-1 rhs, column by column: wavefront scans for uniform columns of at most 64
 * levels, coarsened scans up to 512 levels, otherwise a forward then backward Thomas
 * sweep */
/* precon 3, one column of L rows: forward y_k = r_k - l1_k y_(k-1) - l2_k y_(k-2), then
 * z_k = (y_k - u1_k z_(k+1) - u2_k z_(k+2)) * (1/d_k), terms past the column's ends
 * dropped; the HIP path's k_line_penta runs the same sequence per column. */
static void band_column(const oc_ctx *c, int s0, int L, const double *rhs, double *z) {
  double y1 = 0., y2 = 0.;
  for (int k = 0; k < L; ++k) {
    int i = s0 + k;
    double y = rhs[i];
    if (k >= 1) y = y - c->fl1[i] * y1;
    if (k >= 2) y = y - c->fl2[i] * y2;
    z[i] = y;
    y2 = y1; y1 = y;
  }
  double z1 = 0., z2 = 0.;
  for (int k = L - 1; k >= 0; --k) {
    int i = s0 + k;
    double t = z[i];
    if (k + 1 < L) t = t - c->fu1[i] * z1;
    if (k + 2 < L) t = t - c->fu2[i] * z2;
    t = t * c->fdinv[i];
    z[i] = t;
    z2 = z1; z1 = t;
  }
}

/* the same substitution in FP32 (the mixed forms): residual and LU factors rounded to FP32,
 * every operation in FP32 */
static void band_column_f32(const oc_ctx *c, int s0, int L, const double *rhs, double *z) {
  float *w = (float *)malloc(sizeof(float) * (size_t)L);
  float y1 = 0.f, y2 = 0.f;
  for (int k = 0; k < L; ++k) {
    int i = s0 + k;
    float y = (float)rhs[i];
    if (k >= 1) y = y - (float)c->fl1[i] * y1;
    if (k >= 2) y = y - (float)c->fl2[i] * y2;
    w[k] = y;
    y2 = y1; y1 = y;
  }
  float z1 = 0.f, z2 = 0.f;
  for (int k = L - 1; k >= 0; --k) {
    int i = s0 + k;
    float t = w[k];
    if (k + 1 < L) t = t - (float)c->fu1[i] * z1;
    if (k + 2 < L) t = t - (float)c->fu2[i] * z2;
    t = t * (float)c->fdinv[i];
    w[k] = t;
    z2 = z1; z1 = t;
  }
  for (int k = 0; k < L; ++k) z[s0 + k] = (double)w[k];
  free(w);
}

/* band line solve of the mixed forms (precon 3 only); returns -1 without a band */
int oc_line_solve32(oc_ctx *c, const double *rhs, double *z) {
  if (c->band_L <= 0) return -1;
  for (int k = 0; k < c->m.ncol; ++k) band_column_f32(c, k * c->band_L, c->band_L, rhs, z);
  return 0;
}

void oc_line_solve(oc_ctx *c, const double *rhs, double *z) {
  const oc_mesh *m = &c->m;
  if (c->band_L > 0) {
    for (int k = 0; k < m->ncol; ++k) band_column(c, k * c->band_L, c->band_L, rhs, z);
    return;
  }
  if (c->line_L > 0) {
    int L = c->line_L;
    for (int k = 0; k < m->ncol; ++k) {
      int s0 = k * L;
      scan_column(c->la + s0, c->linv + s0, c->lc + s0, rhs + s0, L, z + s0);
    }
    return;
  }
  if (c->line_R > 0) {
    int L = m->col_start[1], R = c->line_R;
    for (int k = 0; k < m->ncol; ++k) {
      int s0 = k * L;
      scan_column_r(c->la + s0, c->linv + s0, c->lc + s0, rhs + s0, L, R, z + s0);
    }
    return;
  }
  for (int k = 0; k < m->ncol; ++k) {
    int s0 = m->col_start[k], s1 = m->col_start[k + 1];
    double dprev = 0.;
    for (int i = s0; i < s1; ++i) {
      double d = (rhs[i] - c->la[i] * dprev) * c->linv[i];
      z[i] = d;
      dprev = d;
    }
    for (int i = s1 - 2; i >= s0; --i) z[i] = z[i] - c->lc[i] * z[i + 1];
  }
}

/* point-relaxation step per node (precon F:587-599 on the fine mesh, F:632-653 on the
 * coarse ones): d and beta of the level */
static void pstep_fill(int sm_flag, const double *d, double beta, int n, double *pstep) {
  for (int i = 0; i < n; ++i) {
    if (sm_flag == 1) pstep[i] = (-1.) / d[i];
    else if (sm_flag == 2) pstep[i] = beta;
    else pstep[i] = beta / (1 - d[i] * beta);
  }
}

/* one fine-mesh smoothing sweep of precon (F:607-612, F:740-745): r = L p (skipped on the
 * first sweep from p = 0, where it is zero), then the point step p += omega*pstep*(r-rhs),
 * or with precon=2 the line step p += omega * M^-1 (rhs - r) */
static void fine_sweep(oc_ctx *c, const double *rhs, double *p, double *r, int impfl, int first) {
  int n = c->n;
  double om = c->p.omega;
  if (!first) oc_laplc(c, p, r, impfl);
  if (c->p.precon == 1) {
    const double *pstep = c->s[15];
    for (int i = 0; i < n; ++i) p[i] = p[i] + om * pstep[i] * (r[i] - rhs[i]);
  } else {
    double *res = c->s[15], *z = c->s[2];
    for (int i = 0; i < n; ++i) res[i] = rhs[i] - r[i];
    if (c->p.precision >= 1 && c->band_L > 0) oc_line_solve32(c, res, z);
    else oc_line_solve(c, res, z);
    for (int i = 0; i < n; ++i) p[i] = p[i] + om * z[i];
  }
}

/* restriction_unstr(scalar,scalar_c,flg), F:757-822: the mean over the fine node's
 * neighbours of the edge averages. transfer 0 indexes the fine array by the neighbour
 * counter itself, as F:817 does (scalar(inei), inei = 1..count); transfer 1 by the
 * neighbour list. f = the finer level's context. */
static void mg_restrict(const oc_ctx *f, const int32_t *match, int nc, const double *src, double *out,
                        int transfer) {
  for (int ic = 0; ic < nc; ++ic) {
    int ip = match[ic];
    int cnt = f->nb_ptr[ip + 1] - f->nb_ptr[ip];
    double sc = 0.;
    for (int q = 0; q < cnt; ++q) {
      int j = transfer ? f->nb_idx[f->nb_ptr[ip] + q] : q;
      sc = sc + (src[ip] + src[j]) / 2;
    }
    out[ic] = sc / cnt;
  }
}

/* prolongation_unstr(scalar_c,scalar,flg), F:828-940: injection at the matched nodes, the
 * mean of the matched neighbours elsewhere (0 where there is none). A node counts as
 * matched when its injected value differs from the reference's sentinel -99999; the
 * neighbours are the counter values 1..count with transfer 0 (F:913-914) and the
 * neighbour list with transfer 1. */
static void mg_prolong(const oc_ctx *f, const int32_t *imatch, const double *coarse, double *out,
                       int transfer) {
  const double sentinel = -99999.;
  int n = f->n;
  double *aux = f->s[14];
  for (int ip = 0; ip < n; ++ip) { out[ip] = 0.; aux[ip] = sentinel; }
  for (int ip = 0; ip < n; ++ip)
    if (imatch[ip] >= 0) { out[ip] = coarse[imatch[ip]]; aux[ip] = coarse[imatch[ip]]; }
  for (int ip = 0; ip < n; ++ip) {
    if (aux[ip] != sentinel) continue;
    int cnt = f->nb_ptr[ip + 1] - f->nb_ptr[ip];
    double sc = out[ip], filled = 0.;
    for (int q = 0; q < cnt; ++q) {
      int j = transfer ? f->nb_idx[f->nb_ptr[ip] + q] : q;
      if (aux[j] != sentinel) { sc = sc + out[j]; filled = filled + 1; }
    }
    out[ip] = filled != 0. ? sc / filled : sc;
  }
}

/* the V-cycle of precon's multigrid branch, F:618-748 */
static void mg_cycle(oc_ctx *c, const double *rhs, double *p, double *r, int impfl) {
  oc_ctx *c1 = c->lv[0], *c2 = c->lv[1];
  int n = c->n, n1 = c1->n, n2 = c2->n, tr = c->mg_transfer;
  double om = c->mg_omega;
  double *rrhs = c->mg_w[0], *e = c->mg_w[1];
  double *p_c = c->mg_w[2], *r_c = c->mg_w[3], *rrhs_c = c->mg_w[4], *rout_c = c->mg_w[5];
  double *e_c = c->mg_w[6], *ps_c = c->mg_w[7];
  double *p_cc = c->mg_w[8], *r_cc = c->mg_w[9], *rrhs_cc = c->mg_w[10], *ps_cc = c->mg_w[11];
  pstep_fill(c->p.sm_flag, c1->d1, c1->beta1, n1, ps_c);
  pstep_fill(c->p.sm_flag, c2->d1, c2->beta1, n2, ps_cc);
  for (int i = 0; i < n1; ++i) { p_c[i] = 0.; r_c[i] = 0.; e_c[i] = 0.; }
  for (int i = 0; i < n2; ++i) { p_cc[i] = 0.; r_cc[i] = 0.; }
  /* fine residual, restricted */
  oc_laplc(c, p, r, impfl);
  for (int i = 0; i < n; ++i) rrhs[i] = rhs[i] - r[i];
  mg_restrict(c, c->lv_match[0], n1, rrhs, rrhs_c, tr);
  /* coarse 1: relax, residual, restrict */
  for (int i = 0; i < n1; ++i) p_c[i] = -(om * ps_c[i] * rrhs_c[i]);
  for (int it = 1; it <= c->mg_c1; ++it) {
    oc_laplc(c1, p_c, r_c, impfl);
    for (int i = 0; i < n1; ++i) p_c[i] = p_c[i] + om * ps_c[i] * (r_c[i] - rrhs_c[i]);
  }
  oc_laplc(c1, p_c, r_c, impfl);
  for (int i = 0; i < n1; ++i) rout_c[i] = rrhs_c[i] - r_c[i];
  mg_restrict(c1, c->lv_match[1], n2, rout_c, rrhs_cc, tr);
  /* coarse 2: relax */
  for (int i = 0; i < n2; ++i) p_cc[i] = -(rrhs_cc[i] * ps_cc[i] * om);
  for (int it = 1; it <= c->mg_cc; ++it) {
    oc_laplc(c2, p_cc, r_cc, impfl);
    for (int i = 0; i < n2; ++i) p_cc[i] = p_cc[i] + om * ps_cc[i] * (r_cc[i] - rrhs_cc[i]);
  }
  /* back up: prolong, correct, relax */
  mg_prolong(c1, c->lv_imatch[1], p_cc, e_c, tr);
  for (int i = 0; i < n1; ++i) p_c[i] = p_c[i] + e_c[i];
  for (int it = 1; it <= c->mg_c2; ++it) {
    oc_laplc(c1, p_c, r_c, impfl);
    for (int i = 0; i < n1; ++i) p_c[i] = p_c[i] + om * ps_c[i] * (r_c[i] - rrhs_c[i]);
  }
  mg_prolong(c, c->lv_imatch[0], p_c, e, tr);
  for (int i = 0; i < n; ++i) p[i] = p[i] + e[i];
  for (int it = 1; it <= c->mg_itrp2; ++it) fine_sweep(c, rhs, p, r, impfl, 0);
}

/* precon(rhs,p,r,iprc,elas,impfl), F:482-748: fine-grid smoothing, then the multigrid
 * V-cycle when oc_set_multigrid has turned it on (mg_flag=1).
 * precon=2 replaces the point step pstep*(r-rhs) by the line solve M^-1 (rhs-r) on the
 * fine mesh; the coarse meshes keep the point step. */
void oc_precon(oc_ctx *c, const double *rhs, double *p, double *r, int impfl) {
  int n = c->n;
  if (c->p.precon == 0) { for (int i = 0; i < n; ++i) p[i] = rhs[i]; return; }
  for (int i = 0; i < n; ++i) { p[i] = 0.; r[i] = 0.; }
  if (c->p.precon == 1) pstep_fill(c->p.sm_flag, c->d1, c->beta1, n, c->s[15]);
  for (int it = 1; it <= c->p.itrp1; ++it) fine_sweep(c, rhs, p, r, impfl, it == 1);
  if (c->mg_on) mg_cycle(c, rhs, p, r, impfl);
}

int oc_set_multigrid(oc_ctx *c, const oc_mesh *m1, const int32_t *match1, const oc_mesh *m2,
                     const int32_t *match2, int itrp2, int itrp_c1, int itrp_c2, int itrp_cc,
                     int transfer) {
  if (!m1) { c->mg_on = 0; return 0; }
  if (c->p.iper == 2) return -1;   /* the coarse levels' periodic closures are not restated */
  const oc_mesh *ms[2] = {m1, m2};
  const int32_t *mt[2] = {match1, match2};
  int nf[2] = {c->n, m1->npoin};
  for (int l = 0; l < 2; ++l) {
    if (ms[l]->npoin > nf[l]) return -1;  /* coarse arrays are indexed inside the fine ones */
    for (int i = 0; i < ms[l]->npoin; ++i)
      if (mt[l][i] < 1 || mt[l][i] > nf[l]) return -1;
  }
  for (int l = 0; l < 2; ++l) {
    oc_destroy(c->lv[l]); free(c->lv_match[l]); free(c->lv_imatch[l]);
    c->lv[l] = oc_create(ms[l], &c->p);
    int nc = ms[l]->npoin;
    c->lv_match[l] = (int32_t *)malloc(sizeof(int32_t) * (size_t)nc);
    c->lv_imatch[l] = (int32_t *)malloc(sizeof(int32_t) * (size_t)nf[l]);
    for (int i = 0; i < nf[l]; ++i) c->lv_imatch[l][i] = -1;
    for (int i = 0; i < nc; ++i) {
      c->lv_match[l][i] = mt[l][i] - 1;
      c->lv_imatch[l][mt[l][i] - 1] = i;
    }
  }
  oc_ctx *c1 = c->lv[0], *c2 = c->lv[1];
  /* laplc_c / laplc_cc (F:1356-1535): rh0 of the matched fine node; coefw, timestp and ue
   * indexed by the coarse node id in the fine arrays, as the reference does */
  for (int i = 0; i < c1->n; ++i) {
    c1->rh0[i] = c->rh0[c->lv_match[0][i]];
    c1->coefw[i] = c->coefw[i];
    c1->ue[i] = c->ue[i];
  }
  for (int i = 0; i < c2->n; ++i) {
    c2->rh0[i] = c->rh0[c->lv_match[0][c->lv_match[1][i]]];
    c2->coefw[i] = c->coefw[i];
    c2->ue[i] = c->ue[i];
  }
  c->mg_transfer = transfer;
  /* compute_d_unstr(2), (3) (F:1181-1199): density restricted from the fine mesh */
  double *r1 = (double *)calloc((size_t)c1->n, sizeof(double));
  double *r2 = (double *)calloc((size_t)c2->n, sizeof(double));
  mg_restrict(c, c->lv_match[0], c1->n, c->rh0, r1, transfer);
  compute_d(c1, r1, c1->d1);
  mg_restrict(c1, c->lv_match[1], c2->n, r1, r2, transfer);
  compute_d(c2, r2, c2->d1);
  free(r1); free(r2);
  for (int i = 0; i < 12; ++i) {
    free(c->mg_w[i]);
    c->mg_w[i] = zalloc((size_t)c->n);
  }
  c->mg_itrp2 = itrp2; c->mg_c1 = itrp_c1; c->mg_c2 = itrp_c2; c->mg_cc = itrp_cc;
  c->mg_omega = c->p.omega;
  c->mg_on = 1;
  return 0;
}

void oc_set_multigrid_omega(oc_ctx *c, double omega_c) { c->mg_omega = omega_c > 0 ? omega_c : c->p.omega; }

void oc_mg_restrict(oc_ctx *c, int level, const double *src, double *out) {
  const oc_ctx *f = level == 1 ? c : c->lv[0];
  mg_restrict(f, c->lv_match[level - 1], c->lv[level - 1]->n, src, out, c->mg_transfer);
}

void oc_mg_prolong(oc_ctx *c, int level, const double *coarse, double *out) {
  const oc_ctx *f = level == 1 ? c : c->lv[0];
  mg_prolong(f, c->lv_imatch[level - 1], coarse, out, c->mg_transfer);
}

oc_ctx *oc_level(oc_ctx *c, int level) { return level == 0 ? c : c->lv[level - 1]; }

/* gcrk(p,u1,v1,w1,u2,v2,w2,impfl), F:1776-1972 (ielast=0).
 * exit_mode 0: max|r| <= eps0/dt (reference, F:1805, F:1895);
 * exit_mode 1: max|r| <= eps0 * max|b|, b = the residual of the zero initial guess
 *              (the right-hand side -div(rho*u*)/(vol*rho) after the velocity bc). */
void oc_gcrk(oc_ctx *c, double *p, const double *u1, const double *v1, const double *w1,
             double *u2, double *v2, double *w2, int impfl) {
  int n = c->n, lord = c->p.lord;
  double *r = c->gr, *qr = c->gqr, *ar = c->gar, *x = c->gx, *ax = c->gax;
  double ax2[64], axar[64], del[64];
  double epa = 1.e-38;
  int nlc = 0;
  double eps = c->p.eps0 / c->p.dt;
  /* precision 2: the iterations apply the FP32 operator and the solve runs in passes,
   * each from the FP64 residual of the current pressure, ending when the recursive
   * residual has dropped by mixed_drop (or met the tolerance); a pass after the first
   * stops the solve when that FP64 residual meets the tolerance */
  const int op32 = c->p.precision == 2;
  const double drop = c->p.mixed_drop > 0 ? c->p.mixed_drop : 1.e-3;
  for (int k = 0; k < n; ++k) { r[k] = 0.; ar[k] = 0.; qr[k] = 0.; }
  for (size_t k = 0; k < (size_t)lord * n; ++k) { x[k] = 0.; ax[k] = 0.; }
  double bmx = 0.;
  if (c->p.exit_mode == 1) {
    double *z = c->gb[3];
    for (int k = 0; k < n; ++k) z[k] = 0.;
    oc_prfr(c, z, u1, v1, w1, c->gb[0], c->gb[1], c->gb[2], impfl);
    oc_div(c, c->gb[0], c->gb[1], c->gb[2], z, -1);
    for (int k = 0; k < n; ++k) bmx = amax1(bmx, fabs(z[k]));
    eps = c->p.eps0 * bmx;
  }
  const int dev = c->p.dot_order == 1;
  double eer0 = 0., eem0 = -1.e15;
  for (int pass = 0;; ++pass) {
    /* u2, v2, w2: prfr of the incoming pressure; later passes use scratch arrays */
    double *pu = pass ? c->gb[0] : u2, *pv = pass ? c->gb[1] : v2, *pw = pass ? c->gb[2] : w2;
    oc_prfr(c, p, u1, v1, w1, pu, pv, pw, impfl);
    oc_div(c, pu, pv, pw, r, -1);
    double rmx = 0.;
    for (int k = 0; k < n; ++k) rmx = amax1(rmx, fabs(r[k]));
    if (pass > 0 && (rmx <= eps || !(rmx == rmx) || pass >= 24)) break;
    oc_precon(c, r, qr, ar, impfl);
    double rl20 = 0.;
    if (pass == 0) {
      for (int k = 0; k < n; ++k) {
        eer0 = eer0 + qr[k] * qr[k];
        eem0 = amax1(eem0, fabs(qr[k]));
      }
      if (dev) eer0 = dsum_stride(c, qr, qr);
      eer0 = amax1(eer0, epa);
      eem0 = amax1(eem0, epa);
    }
    for (int k = 0; k < n; ++k) rl20 = rl20 + r[k] * r[k];
    if (dev) rl20 = dsum_gather(c, NULL, r);
    rl20 = amax1(rl20, epa);
    for (int k = 0; k < n; ++k) x[k] = qr[k];
    if (op32) oc_laplc32(c, x, ax, impfl);
    else oc_laplc(c, x, ax, impfl);
    const double eps_pass = op32 ? amax1(eps, drop * rmx) : eps;
    const int itmn = pass == 0 ? 3 : 1;
    int conv = 0;
    for (int it = 1; it <= c->p.itr; ++it) {
      for (int i = 1; i <= lord; ++i) {
        double *xi = x + (size_t)(i - 1) * n, *axi = ax + (size_t)(i - 1) * n;
        double rax = 0.;
        ax2[i - 1] = 0.;
        if (!dev) {
          for (int k = 0; k < n; ++k) {
            rax = rax + r[k] * axi[k];
            ax2[i - 1] = ax2[i - 1] + axi[k] * axi[k];
          }
        } else if (it == 1 && i == 1) {   /* fused into the divergence kernel of laplc */
          rax = dsum_gather(c, r, axi);
          ax2[i - 1] = dsum_gather(c, NULL, axi);
        } else {                          /* fused into the direction kernel */
          rax = dsum_stride(c, r, axi);
          ax2[i - 1] = dsum_stride(c, axi, axi);
        }
        ax2[i - 1] = amax1(epa, ax2[i - 1]);
        double beta = -rax / ax2[i - 1];
        double dvmx = -1.e15, rl2 = 0.;
        for (int k = 0; k < n; ++k) {
          p[k] = p[k] + beta * xi[k];
          r[k] = r[k] + beta * axi[k];
          dvmx = amax1(dvmx, fabs(r[k]));
          rl2 = rl2 + r[k] * r[k];
        }
        if (dev) rl2 = dsum_stride(c, r, r);
        if (dvmx <= eps_pass && it >= itmn) { conv = 1; goto pass_end; }
        if (rl2 >= rl20) goto pass_end;
        rl20 = amax1(rl2, epa);
        oc_precon(c, r, qr, ar, impfl);
        if (op32) oc_laplc32(c, qr, ar, impfl);
        else oc_laplc(c, qr, ar, impfl);
        nlc = nlc + 1;
        for (int ii = 1; ii <= i; ++ii) {
          double *axii = ax + (size_t)(ii - 1) * n;
          axar[ii - 1] = 0.;
          if (dev) axar[ii - 1] = dsum_gather(c, axii, ar);
          else
            for (int k = 0; k < n; ++k) axar[ii - 1] = axar[ii - 1] + axii[k] * ar[k];
          del[ii - 1] = -axar[ii - 1] / ax2[ii - 1];
        }
        if (i < lord) {
          double *xn = x + (size_t)i * n, *axn = ax + (size_t)i * n;
          for (int k = 0; k < n; ++k) { xn[k] = qr[k]; axn[k] = ar[k]; }
          for (int ii = 1; ii <= i; ++ii) {
            double *xii = x + (size_t)(ii - 1) * n, *axii = ax + (size_t)(ii - 1) * n;
            for (int k = 0; k < n; ++k) {
              xn[k] = xn[k] + del[ii - 1] * xii[k];
              axn[k] = axn[k] + del[ii - 1] * axii[k];
            }
          }
        } else {
          for (int k = 0; k < n; ++k) {
            x[k] = qr[k] + del[0] * x[k];
            ax[k] = ar[k] + del[0] * ax[k];
          }
          for (int ii = 2; ii <= i; ++ii) {
            double *xii = x + (size_t)(ii - 1) * n, *axii = ax + (size_t)(ii - 1) * n;
            for (int k = 0; k < n; ++k) {
              x[k] = x[k] + del[ii - 1] * xii[k];
              xii[k] = 0.;
              ax[k] = ax[k] + del[ii - 1] * axii[k];
              axii[k] = 0.;
            }
          }
        }
      }
    }
  pass_end:
    if (!op32 || !conv) break;
  }
  double eer = 0., eem = -1.e15;
  for (int k = 0; k < n; ++k) { eer = eer + qr[k] * qr[k]; eem = amax1(eem, fabs(qr[k])); }
  if (dev) eer = dsum_stride(c, qr, qr);
  c->st.eer = eer / eer0;
  c->st.eem = eem / eem0;
  c->st.niter = nlc;
  c->st.nitsm = c->st.nitsm + nlc;
  c->st.icount = c->st.icount + 1;
}

/* mpdatam(uuu,unk,iflag), F:3029-3311: donor cell + one antidiffusive pass (iord=2)
 * with the FCT limiter; bounds from the donor-cell solution only. */
static void mpdatamper(oc_ctx *c, const double *uuu, double *unk);

void oc_mpdatam(oc_ctx *c, const double *uuu, double *unk, int iflag) {
  if (c->p.iper == 2) {   /* mpsolver F:314-319 */
    mpdatamper(c, uuu, unk);
    return;
  }
  const oc_mesh *m = &c->m;
  int n = c->n, E = m->nedge, B = m->nbface, Bp = m->nboun;
  const double *U = uuu, *V = uuu + n, *W = uuu + 2 * (size_t)n;
  double dt = c->p.dt;
  double *aun = zalloc(E), *aunb = zalloc(B), *aunbb1 = zalloc(Bp);
  double *divv = c->s[0], *arh2 = c->s[1], *gx = c->s[2], *gy = c->s[3], *gz = c->s[4];
  double *mx = c->s[5], *mn = c->s[6], *rin = c->s[7], *rout = c->s[8], *cp = c->s[9], *cn = c->s[10];
  double afree[5] = {c->p.u0, c->p.v0, c->p.w0, 0., 1.};
  double tol = 1.e-10;
  for (int i = 0; i < n; ++i) { divv[i] = 0.; mx[i] = -1.e15; mn[i] = 1.e15; }
  for (int ipass = 1; ipass <= 2; ++ipass) {
    for (int i = 0; i < n; ++i) arh2[i] = 0.;
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      const double *S = pass == 0 ? m->sn : m->sb;
      double *a = pass == 0 ? aun : aunb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        double sx = S[e], sy = S[ne + e], sz = S[2 * (size_t)ne + e];
        double fl;
        if (ipass == 1) {
          a[e] = (U[i1] + U[i2]) * 0.5 * sx + (V[i1] + V[i2]) * 0.5 * sy + (W[i1] + W[i2]) * 0.5 * sz;
          fl = unk[i1] * amax1(0., a[e]) + unk[i2] * amin1(0., a[e]);
        } else {
          fl = a[e];
        }
        arh2[i1] = arh2[i1] + fl * (1. + 1.) * 0.5;
        arh2[i2] = arh2[i2] - fl * (1. + 1.) * 0.5;
        divv[i1] = divv[i1] + a[e];
        divv[i2] = divv[i2] - a[e];
      }
    }
    for (int b = 0; b < Bp; ++b) {
      int i1 = m->ibpoin[b] - 1, f = m->ibpoin[Bp + b];
      double fl1;
      if (ipass == 1) {
        aunbb1[b] = -divv[i1];
        double af = iflag == 1 ? c->ue[i1] : afree[iflag - 1];
        fl1 = af * amin1(0., aunbb1[b]) + unk[i1] * amax1(0., aunbb1[b]);
      } else {
        fl1 = 0.;
      }
      if (f == 11 || f == 5 || f == 6) { fl1 = 0.; aunbb1[b] = 0.; }
      arh2[i1] = arh2[i1] + fl1 * 1.;
    }
    for (int i = 0; i < n; ++i) unk[i] = unk[i] - arh2[i] * dt / m->vol[i] / c->rh0[i];
    if (ipass == 2) break;

    oc_derivative(c, unk, gx, 1);
    oc_derivative(c, unk, gy, 2);
    oc_derivative(c, unk, gz, 3);
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      double *a = pass == 0 ? aun : aunb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        double un = fabs(a[e]) * (unk[i2] - unk[i1]) * 0.5;
        double volt = 0.5 * (m->vol[i1] * c->rh0[i1] + m->vol[i2] * c->rh0[i2]);
        double dpdx = 0.5 * (gx[i1] + gx[i2]) / volt;
        double dpdy = 0.5 * (gy[i1] + gy[i2]) / volt;
        double dpdz = 0.5 * (gz[i1] + gz[i2]) / volt;
        double term = (U[i1] + U[i2]) * 0.5 * dpdx + (V[i1] + V[i2]) * 0.5 * dpdy + (W[i1] + W[i2]) * 0.5 * dpdz;
        a[e] = un - a[e] * term * dt * 0.5;
      }
    }
    for (int b = 0; b < Bp; ++b) aunbb1[b] = 0.;
    /* FCT bounds from the donor-cell solution (F:3228-3247) */
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        double p1 = unk[i1], p2 = unk[i2];
        mx[i1] = amax1(amax1(mx[i1], p1), p2);
        mn[i1] = amin1(amin1(mn[i1], p1), p2);
        mx[i2] = amax1(amax1(mx[i2], p2), p1);
        mn[i2] = amin1(amin1(mn[i2], p2), p1);
      }
    }
    for (int i = 0; i < n; ++i) { rin[i] = 0.; rout[i] = 0.; cp[i] = 0.; cn[i] = 0.; }
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      double *a = pass == 0 ? aun : aunb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        double pp = amax1(0., a[e]), pn = -amin1(0., a[e]);
        rin[i1] = rin[i1] + pn;
        rout[i1] = rout[i1] + pp;
        rin[i2] = rin[i2] + pp;
        rout[i2] = rout[i2] + pn;
      }
    }
    for (int i = 0; i < n; ++i) {
      cp[i] = (mx[i] - unk[i]) * m->vol[i] * c->rh0[i] / (rin[i] * dt + tol);
      cn[i] = (unk[i] - mn[i]) * m->vol[i] * c->rh0[i] / (rout[i] * dt + tol);
    }
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      double *a = pass == 0 ? aun : aunb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        if (a[e] > 0.) a[e] = a[e] * amin1(amin1(1., cp[i2]), cn[i1]);
        else a[e] = a[e] * amin1(amin1(1., cn[i2]), cp[i1]);
      }
    }
  }
  free(aun); free(aunb); free(aunbb1);
}

/* mpdatamper(uuu,unk,iflag), F:3617-3941: the two passes of mpdatam with the periodic
 * closures. No boundary-point fluxes (the boundary points add 0, F:3733-3739); the
 * increments are summed over the coincident sets before the update (F:3741-3743); the
 * FCT bounds of each set are merged (F:3831-3879: a pair takes the extremes of both
 * nodes; a quadruple takes the maxima of all four, and its minima are formed, as the
 * reference writes them, from the four nodes' maxima and values); the in/out flux sums
 * are summed over the sets (F:3897-3902). iflag selects nothing here. */
static double max8(const double *v) { double r = v[0]; for (int k = 1; k < 8; ++k) r = amax1(r, v[k]); return r; }
static double min8(const double *v) { double r = v[0]; for (int k = 1; k < 8; ++k) r = amin1(r, v[k]); return r; }

static void mpdatamper(oc_ctx *c, const double *uuu, double *unk) {
  const oc_mesh *m = &c->m;
  int n = c->n, E = m->nedge, B = m->nbface, Bp = m->nboun;
  const double *U = uuu, *V = uuu + n, *W = uuu + 2 * (size_t)n;
  double dt = c->p.dt;
  double *aun = zalloc(E), *aunb = zalloc(B);
  double *divv = c->s[0], *arh2 = c->s[1], *gx = c->s[2], *gy = c->s[3], *gz = c->s[4];
  double *mx = c->s[5], *mn = c->s[6], *rin = c->s[7], *rout = c->s[8], *cp = c->s[9], *cn = c->s[10];
  double tol = 1.e-10;
  for (int i = 0; i < n; ++i) { divv[i] = 0.; mx[i] = -1.e15; mn[i] = 1.e15; }
  for (int ipass = 1; ipass <= 2; ++ipass) {
    for (int i = 0; i < n; ++i) arh2[i] = 0.;
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      const double *S = pass == 0 ? m->sn : m->sb;
      double *a = pass == 0 ? aun : aunb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        double sx = S[e], sy = S[ne + e], sz = S[2 * (size_t)ne + e];
        double fl;
        if (ipass == 1) {
          a[e] = (U[i1] + U[i2]) * 0.5 * sx + (V[i1] + V[i2]) * 0.5 * sy + (W[i1] + W[i2]) * 0.5 * sz;
          fl = unk[i1] * amax1(0., a[e]) + unk[i2] * amin1(0., a[e]);
        } else {
          fl = a[e];
        }
        arh2[i1] = arh2[i1] + fl * (1. + 1.) * 0.5;
        arh2[i2] = arh2[i2] - fl * (1. + 1.) * 0.5;
        divv[i1] = divv[i1] + a[e];
        divv[i2] = divv[i2] - a[e];
      }
    }
    for (int b = 0; b < Bp; ++b) {
      int i1 = m->ibpoin[b] - 1;
      arh2[i1] = arh2[i1] + 0. * 1.;
    }
    cyclic_all(c, arh2);
    for (int i = 0; i < n; ++i) unk[i] = unk[i] - arh2[i] * dt / m->vol[i] / c->rh0[i];
    if (ipass == 2) break;

    oc_derivative(c, unk, gx, 1);
    oc_derivative(c, unk, gy, 2);
    oc_derivative(c, unk, gz, 3);
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      double *a = pass == 0 ? aun : aunb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        double un = fabs(a[e]) * (unk[i2] - unk[i1]) * 0.5;
        double volt = 0.5 * (m->vol[i1] * c->rh0[i1] + m->vol[i2] * c->rh0[i2]);
        double dpdx = 0.5 * (gx[i1] + gx[i2]) / volt;
        double dpdy = 0.5 * (gy[i1] + gy[i2]) / volt;
        double dpdz = 0.5 * (gz[i1] + gz[i2]) / volt;
        double term = (U[i1] + U[i2]) * 0.5 * dpdx + (V[i1] + V[i2]) * 0.5 * dpdy + (W[i1] + W[i2]) * 0.5 * dpdz;
        a[e] = un - a[e] * term * dt * 0.5;
      }
    }
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        double p1 = unk[i1], p2 = unk[i2];
        mx[i1] = amax1(amax1(mx[i1], p1), p2);
        mn[i1] = amin1(amin1(mn[i1], p1), p2);
        mx[i2] = amax1(amax1(mx[i2], p2), p1);
        mn[i2] = amin1(amin1(mn[i2], p2), p1);
      }
    }
    for (int list = 0; list < 2; ++list) {
      int nl = list == 0 ? m->ncoin : m->ncoin2;
      const int32_t *pr = list == 0 ? m->icoin : m->icoin2;
      for (int ic = 0; ic < nl; ++ic) {
        int i1 = pr[ic] - 1, i2 = pr[nl + ic] - 1;
        double p1 = unk[i1], p2 = unk[i2];
        mx[i1] = amax1(amax1(amax1(mx[i1], p1), p2), mx[i2]);
        mn[i1] = amin1(amin1(amin1(mn[i1], p1), p2), mn[i2]);
        mx[i2] = amax1(amax1(amax1(mx[i2], p1), p2), mx[i1]);
        mn[i2] = amin1(amin1(amin1(mn[i2], p1), p2), mn[i1]);
      }
    }
    for (int ic = 0; ic < m->ncoin4; ++ic) {
      int q = m->ncoin4, ii[4];
      for (int k = 0; k < 4; ++k) ii[k] = m->icoin4[k * q + ic] - 1;
      double v[8];
      for (int k = 0; k < 4; ++k) v[4 + k] = unk[ii[k]];
      for (int k = 0; k < 4; ++k) {
        for (int j = 0; j < 4; ++j) v[j] = mx[ii[j]];
        mx[ii[k]] = max8(v);
        v[k] = mx[ii[k]];
        mn[ii[k]] = min8(v);
      }
    }
    for (int i = 0; i < n; ++i) { rin[i] = 0.; rout[i] = 0.; cp[i] = 0.; cn[i] = 0.; }
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      double *a = pass == 0 ? aun : aunb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        double pp = amax1(0., a[e]), pn = -amin1(0., a[e]);
        rin[i1] = rin[i1] + pn;
        rout[i1] = rout[i1] + pp;
        rin[i2] = rin[i2] + pp;
        rout[i2] = rout[i2] + pn;
      }
    }
    cyclic_all(c, rin);
    cyclic_all(c, rout);
    for (int i = 0; i < n; ++i) {
      cp[i] = (mx[i] - unk[i]) * m->vol[i] * c->rh0[i] / (rin[i] * dt + tol);
      cn[i] = (unk[i] - mn[i]) * m->vol[i] * c->rh0[i] / (rout[i] * dt + tol);
    }
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      double *a = pass == 0 ? aun : aunb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        if (a[e] > 0.) a[e] = a[e] * amin1(amin1(1., cp[i2]), cn[i1]);
        else a[e] = a[e] * amin1(amin1(1., cn[i2]), cp[i1]);
      }
    }
  }
  free(aun); free(aunb);
}

/* mpdatam of the mixed forms: the loops above in FP32 with the HIP path's operands — the
 * advector, edge metrics, boundary closures, vol and rh0 rounded to FP32, the FP64 field
 * rounded to FP32 where it is read, every product and sum in FP32, and each pass's FP32
 * increment subtracted from the FP64 field (no reference counterpart). */
static float f32min(float a, float b) { return a < b ? a : b; }
static float f32max(float a, float b) { return a > b ? a : b; }

void oc_mpdatam32(oc_ctx *c, const double *uuu, double *unk, int iflag) {
  const oc_mesh *m = &c->m;
  int n = c->n, E = m->nedge, B = m->nbface, Bp = m->nboun;
  float *U = (float *)malloc(sizeof(float) * 3 * (size_t)n), *V = U + n, *W = U + 2 * (size_t)n;
  for (size_t k = 0; k < 3 * (size_t)n; ++k) U[k] = (float)uuu[k];
  float *vr = (float *)malloc(sizeof(float) * 12 * (size_t)n);
  float *vl = vr + n, *rh = vr + 2 * (size_t)n, *divv = vr + 3 * (size_t)n, *arh2 = vr + 4 * (size_t)n;
  float *gx = vr + 5 * (size_t)n, *gy = vr + 6 * (size_t)n, *gz = vr + 7 * (size_t)n;
  float *mx = vr + 8 * (size_t)n, *mn = vr + 9 * (size_t)n, *rin = vr + 10 * (size_t)n, *rout = vr + 11 * (size_t)n;
  float *cp = (float *)malloc(sizeof(float) * 2 * (size_t)n), *cn = cp + n;
  float *aun = (float *)malloc(sizeof(float) * ((size_t)E + B)), *aunb = aun + E;
  const float dt = (float)c->p.dt, tol = (float)1.e-10;
  const double afree[5] = {c->p.u0, c->p.v0, c->p.w0, 0., 1.};
  for (int i = 0; i < n; ++i) {
    vl[i] = (float)m->vol[i]; rh[i] = (float)c->rh0[i]; vr[i] = vl[i] * rh[i];
    divv[i] = 0.f; mx[i] = (float)-1.e15; mn[i] = (float)1.e15;
  }
  for (int ipass = 1; ipass <= 2; ++ipass) {
    for (int i = 0; i < n; ++i) arh2[i] = 0.f;
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      const double *S = pass == 0 ? m->sn : m->sb;
      float *a = pass == 0 ? aun : aunb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        float sx = (float)S[e], sy = (float)S[ne + e], sz = (float)S[2 * (size_t)ne + e];
        float fl;
        if (ipass == 1) {
          a[e] = (U[i1] + U[i2]) * 0.5f * sx + (V[i1] + V[i2]) * 0.5f * sy + (W[i1] + W[i2]) * 0.5f * sz;
          fl = (float)unk[i1] * f32max(0.f, a[e]) + (float)unk[i2] * f32min(0.f, a[e]);
          divv[i1] = divv[i1] + a[e];
          divv[i2] = divv[i2] - a[e];
        } else {
          fl = a[e];
        }
        arh2[i1] = arh2[i1] + fl;
        arh2[i2] = arh2[i2] - fl;
      }
    }
    if (ipass == 1) {
      for (int b = 0; b < Bp; ++b) {
        int i1 = m->ibpoin[b] - 1, f = m->ibpoin[Bp + b];
        float aunbb1 = -divv[i1];
        float af = iflag == 1 ? (float)c->ue[i1] : (float)afree[iflag - 1];
        float fl1 = (f == 11 || f == 5 || f == 6) ? 0.f
                                                  : af * f32min(0.f, aunbb1) + (float)unk[i1] * f32max(0.f, aunbb1);
        arh2[i1] = arh2[i1] + fl1 * 1.f;
      }
    } else {
      for (int b = 0; b < Bp; ++b) { int i1 = m->ibpoin[b] - 1; arh2[i1] = arh2[i1] + 0.f * 1.f; }
    }
    for (int i = 0; i < n; ++i) {
      float inc = arh2[i] * dt / vl[i] / rh[i];
      unk[i] = unk[i] - (double)inc;
    }
    if (ipass == 2) break;

    /* gradients of the donor-cell field */
    for (int i = 0; i < n; ++i) { gx[i] = 0.f; gy[i] = 0.f; gz[i] = 0.f; }
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      const double *S = pass == 0 ? m->sn : m->sb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        float sx = (float)S[e], sy = (float)S[ne + e], sz = (float)S[2 * (size_t)ne + e];
        float u = ((float)unk[i1] + (float)unk[i2]) * 0.5f;
        float tx = sx * u, ty = sy * u, tz = sz * u;
        gx[i1] = gx[i1] + tx; gy[i1] = gy[i1] + ty; gz[i1] = gz[i1] + tz;
        gx[i2] = gx[i2] - tx; gy[i2] = gy[i2] - ty; gz[i2] = gz[i2] - tz;
      }
    }
    for (int b = 0; b < Bp; ++b) {
      int i1 = m->ibpoin[b] - 1;
      float pi = (float)unk[i1];
      gx[i1] = gx[i1] + (float)m->sbb[b] * pi;
      gy[i1] = gy[i1] + (float)m->sbb[Bp + b] * pi;
      gz[i1] = gz[i1] + (float)m->sbb[2 * (size_t)Bp + b] * pi;
    }
    /* antidiffusive velocities, FCT bounds, in/out sums */
    for (int i = 0; i < n; ++i) { rin[i] = 0.f; rout[i] = 0.f; }
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      float *a = pass == 0 ? aun : aunb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        float p1 = (float)unk[i1], p2 = (float)unk[i2];
        float un = fabsf(a[e]) * (p2 - p1) * 0.5f;
        float volt = 0.5f * (vr[i1] + vr[i2]);
        float dpdx = 0.5f * (gx[i1] + gx[i2]) / volt;
        float dpdy = 0.5f * (gy[i1] + gy[i2]) / volt;
        float dpdz = 0.5f * (gz[i1] + gz[i2]) / volt;
        float term = (U[i1] + U[i2]) * 0.5f * dpdx + (V[i1] + V[i2]) * 0.5f * dpdy + (W[i1] + W[i2]) * 0.5f * dpdz;
        float A = un - a[e] * term * dt * 0.5f;
        a[e] = A;
        mx[i1] = f32max(f32max(mx[i1], p1), p2);
        mn[i1] = f32min(f32min(mn[i1], p1), p2);
        mx[i2] = f32max(f32max(mx[i2], p2), p1);
        mn[i2] = f32min(f32min(mn[i2], p2), p1);
        float pp = f32max(0.f, A), pn = -f32min(0.f, A);
        rin[i1] = rin[i1] + pn; rout[i1] = rout[i1] + pp;
        rin[i2] = rin[i2] + pp; rout[i2] = rout[i2] + pn;
      }
    }
    for (int i = 0; i < n; ++i) {
      float pi = (float)unk[i];
      cp[i] = (mx[i] - pi) * vl[i] * rh[i] / (rin[i] * dt + tol);
      cn[i] = (pi - mn[i]) * vl[i] * rh[i] / (rout[i] * dt + tol);
    }
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? E : B;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      float *a = pass == 0 ? aun : aunb;
      for (int e = 0; e < ne; ++e) {
        int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
        if (a[e] > 0.f) a[e] = a[e] * f32min(f32min(1.f, cp[i2]), cn[i1]);
        else a[e] = a[e] * f32min(f32min(1.f, cn[i2]), cp[i1]);
      }
    }
  }
  free(U); free(vr); free(cp); free(aun);
}

/* ---- per-step diagnostics (F:456-459), in the reference's loop order -------------------
 * out[25]: div(uuu,rhs,1) statistics of rhs*timestp (F:2148-2164: dvmx, dvmn, dvav, dvsd),
 * courant(uuu,cour,flag) (F:1974-2094: crnmx, crnmn, crnav, crnsd, ic), stats(unk)
 * (F:2847-2912: umx, umn, uav, usd, then v, w, t). */
void oc_diagnostics(oc_ctx *c, int flag, double *out) {
  const oc_mesh *m = &c->m;
  const int n = c->n;
  const double ts = c->p.dt;
  const double *u = c->uuu, *v = c->uuu + n, *w = c->uuu + 2 * (size_t)n;
  double *rhs = (double *)malloc(sizeof(double) * (size_t)n);
  double *cour = (double *)malloc(sizeof(double) * (size_t)n);
  oc_div(c, u, v, w, rhs, 1);
  double mx = -1.e15, mn = 1.e15, av = 0., sd = 0.;
  for (int ip = 0; ip < n; ++ip) {
    mx = fmax(mx, rhs[ip] * ts);
    mn = fmin(mn, rhs[ip] * ts);
    av = av + rhs[ip] * ts;
  }
  av = av / (double)n;
  for (int ip = 0; ip < n; ++ip) { double d = rhs[ip] * ts - av; sd = sd + d * d; }
  sd = sqrt(sd / n);
  out[0] = mx; out[1] = mn; out[2] = av; out[3] = sd;
  /* courant: the edge loops scatter into both nodes, then the boundary points */
  for (int ip = 0; ip < n; ++ip) cour[ip] = 0.;
  for (int pass = 0; pass < 2; ++pass) {
    int ne = pass == 0 ? m->nedge : m->nbface;
    const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
    const double *S = pass == 0 ? m->sn : m->sb;
    for (int e = 0; e < ne; ++e) {
      int i1 = ie[e] - 1, i2 = ie[ne + e] - 1;
      double sx = S[e], sy = S[ne + e], sz = S[2 * (size_t)ne + e];
      double rhav = 0.5 * (c->rh0[i1] + c->rh0[i2]);
      double aun = (u[i1] + u[i2]) * 0.5 * sx + (v[i1] + v[i2]) * 0.5 * sy + (w[i1] + w[i2]) * 0.5 * sz;
      aun = aun / rhav;
      if (flag == 1) { cour[i1] = cour[i1] + fabs(aun); cour[i2] = cour[i2] + fabs(aun); }
      if (flag == 2) { cour[i1] = cour[i1] + aun; cour[i2] = cour[i2] - aun; }
      if (flag == 3) { cour[i1] = cour[i1] + fmax(0., aun); cour[i2] = cour[i2] - fmin(0., aun); }
    }
  }
  /* the periodic box skips the boundary points (F:2049) */
  for (int b = 0; b < (c->p.iper == 2 ? 0 : m->nboun); ++b) {
    int i1 = m->ibpoin[b] - 1;
    double sx = m->sbb[b], sy = m->sbb[m->nboun + b], sz = m->sbb[2 * (size_t)m->nboun + b];
    double aun = u[i1] * sx + v[i1] * sy + w[i1] * sz;
    aun = aun / c->rh0[i1];
    if (flag == 1) cour[i1] = cour[i1] + fabs(aun);
    if (flag == 2) cour[i1] = cour[i1] + aun;
    if (flag == 3) cour[i1] = cour[i1] + fmax(0., aun);
  }
  for (int ip = 0; ip < n; ++ip) cour[ip] = cour[ip] * ts / m->vol[ip] / c->rh0[ip];
  mx = -1.e9; mn = 1.e9; av = 0.; sd = 0.;
  int ic = 0;
  for (int ip = 0; ip < n; ++ip) {
    av = av + cour[ip];
    mx = fmax(mx, cour[ip]);
    mn = fmin(mn, cour[ip]);
    if (mx == cour[ip]) ic = ip + 1;
  }
  av = av / n;
  for (int ip = 0; ip < n; ++ip) sd = sd + (cour[ip] - av) * (cour[ip] - av);
  sd = sqrt(sd / n);
  out[4] = mx; out[5] = mn; out[6] = av; out[7] = sd; out[8] = ic;
  /* stats(unk): rh0-weighted means and deviations; tsd divides by npoin (F:2905) */
  double rav = 0., fav[4] = {0, 0, 0, 0}, fmx[4], fmn[4], fsd[4] = {0, 0, 0, 0};
  for (int k = 0; k < 4; ++k) { fmx[k] = -1.e10; fmn[k] = 1.e10; }
  for (int ip = 0; ip < n; ++ip) {
    rav = rav + c->rh0[ip];
    for (int k = 0; k < 4; ++k) {
      double x = c->unk[(size_t)k * n + ip];
      fav[k] = fav[k] + c->rh0[ip] * x;
      fmx[k] = fmax(fmx[k], x);
      fmn[k] = fmin(fmn[k], x);
    }
  }
  for (int k = 0; k < 4; ++k) fav[k] = fav[k] / rav;
  for (int ip = 0; ip < n; ++ip)
    for (int k = 0; k < 4; ++k) {
      double d = c->unk[(size_t)k * n + ip] - fav[k];
      fsd[k] = fsd[k] + c->rh0[ip] * (d * d);
    }
  for (int k = 0; k < 4; ++k) {
    out[9 + 4 * k] = fmx[k]; out[10 + 4 * k] = fmn[k]; out[11 + 4 * k] = fav[k];
    out[12 + 4 * k] = sqrt(fsd[k] / (k < 3 ? rav : (double)n));
  }
  free(rhs);
  free(cour);
}

/* ---- NFT driver: mpsolver initialisation (F:177-266) and one DO ITIME step
 * (F:285-439); the diagnostics above are separate calls. */
double *oc_state(oc_ctx *c, const char *name) {
  if (!strcmp(name, "unk")) return c->unk;
  if (!strcmp(name, "frc")) return c->frc;
  if (!strcmp(name, "uuu")) return c->uuu;
  if (!strcmp(name, "unkdash")) return c->unkdash;
  if (!strcmp(name, "prpot")) return c->prpot;
  return NULL;
}

/* velprd(unk,frc,uuu,uuu1,unkdash), F:4448-4531 (Cartesian: the metric factors are 1):
 * a first-order prediction of u at n+1 from u^n + dt*frc advected by uuu, projected by a
 * GCR solve (impfl=0) and prfr; uuu becomes the mean of uuu and the prediction times rho
 * (the advector at n+1/2), unkdash the prediction. */
static void oc_velprd(oc_ctx *c) {
  int n = c->n;
  double dt = c->p.dt;
  double *unk = c->unk, *frc = c->frc, *uuu = c->uuu, *ud = c->unkdash;
  double *u1 = c->vp[0], *ff1 = c->vp[3], *ff2 = c->vp[4], *ff3 = c->vp[5], *pp = c->vp[6];
  double *v1 = c->vp[1], *w1 = c->vp[2];
  double *uo[3] = {u1, v1, w1};
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) ud[(size_t)k * n + i] = unk[(size_t)k * n + i] + dt * frc[(size_t)k * n + i];
  for (int k = 0; k < 3; ++k) {
    const double *f = ud + (size_t)k * n;
    oc_derivative(c, f, ff1, 1);
    oc_derivative(c, f, ff2, 2);
    oc_derivative(c, f, ff3, 3);
    for (int i = 0; i < n; ++i) {
      double v = c->m.vol[i];
      ff1[i] = ff1[i] / v;
      ff2[i] = ff2[i] / v;
      ff3[i] = ff3[i] / v;
    }
    for (int i = 0; i < n; ++i)
      uo[k][i] = f[i] - dt * (uuu[i] * ff1[i] + uuu[n + i] * ff2[i] + uuu[2 * n + i] * ff3[i]);
  }
  for (int i = 0; i < n; ++i) { pp[i] = 0.; ff1[i] = 0.; ff2[i] = 0.; ff3[i] = 0.; }
  oc_gcrk(c, pp, u1, v1, w1, ff1, ff2, ff3, 0);
  oc_prfr(c, pp, u1, v1, w1, ff1, ff2, ff3, 0);
  double *ff[3] = {ff1, ff2, ff3};
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      uuu[(size_t)k * n + i] = (uuu[(size_t)k * n + i] + ff[k][i]) * 0.5 * c->rh0[i];
      ud[(size_t)k * n + i] = ff[k][i];
    }
}

void oc_nft_init(oc_ctx *c) {
  int n = c->n;
  double *unk = c->unk, *frc = c->frc, *uuu = c->uuu, *ud = c->unkdash, *u1 = c->unk1;
  for (int i = 0; i < 4 * n; ++i) frc[i] = 0.;
  for (int i = 0; i < n; ++i) {
    unk[i] = c->ue[i];
    unk[n + i] = c->ve[i];
    unk[2 * n + i] = 0.2 * frc[i];
    unk[3 * n + i] = 0. + 0.001 * frc[i];
    u1[i] = unk[i];
    u1[n + i] = unk[n + i];
    u1[2 * n + i] = unk[2 * n + i];
    u1[3 * n + i] = 0.;
    c->prpot[i] = 0.;
  }
  memset(&c->st, 0, sizeof(c->st));
  oc_gcrk(c, c->prpot, u1, u1 + n, u1 + 2 * n, unk, unk + n, unk + 2 * n, 0);
  oc_prfr(c, c->prpot, u1, u1 + n, u1 + 2 * n, unk, unk + n, unk + 2 * n, 0);
  for (int i = 0; i < n; ++i) {
    uuu[i] = unk[i]; uuu[n + i] = unk[n + i]; uuu[2 * n + i] = unk[2 * n + i];
  }
  for (int i = 0; i < n; ++i) {
    c->prpot[i] = 0.;
    frc[i] = 0.; frc[n + i] = 0.; frc[2 * n + i] = 0.; frc[3 * n + i] = 0.;
  }
  if (c->p.itraj == 1) { oc_velprd(c); return; }
  for (int i = 0; i < n; ++i) {
    uuu[i] = uuu[i] * c->rh0[i];
    uuu[n + i] = uuu[n + i] * c->rh0[i];
    uuu[2 * n + i] = uuu[2 * n + i] * c->rh0[i];
    ud[i] = unk[i]; ud[n + i] = unk[n + i]; ud[2 * n + i] = unk[2 * n + i];
  }
}

void oc_nft_step(oc_ctx *c) {
  int n = c->n;
  double dt = c->p.dt;
  double *unk = c->unk, *frc = c->frc, *uuu = c->uuu, *ud = c->unkdash, *u1 = c->unk1;
  double *uuu0 = c->uuu0;
  const double *al = c->alpha, *ue = c->ue, *ve = c->ve, *th0 = c->th0, *cw = c->coefw;
  const int traj = c->p.itraj == 1;
  for (int i = 0; i < n; ++i) {
    if (!traj) { uuu0[i] = uuu[i]; uuu0[n + i] = uuu[n + i]; uuu0[2 * n + i] = uuu[2 * n + i]; }
    for (int k = 0; k < 4; ++k) unk[k * n + i] = unk[k * n + i] + 0.5 * frc[k * n + i] * dt;
  }
  for (int k = 0; k < 4; ++k) {
    if (c->p.precision >= 1) oc_mpdatam32(c, uuu, unk + (size_t)k * n, k + 1);
    else oc_mpdatam(c, uuu, unk + (size_t)k * n, k + 1);
  }
  for (int mtri = 1; mtri <= 2; ++mtri) {
    for (int i = 0; i < n; ++i) {
      double frm1 = 0. - al[i] * (ud[i] - ue[i]);
      double frm2 = 0. - al[i] * (ud[n + i] - ve[i]);
      double frm3 = (0. - al[i] * ud[2 * n + i]) + G * unk[3 * n + i] / th0[i];
      frc[i] = unk[i] + 0.5 * dt * frm1;
      frc[n + i] = unk[n + i] + 0.5 * dt * frm2;
      frc[2 * n + i] = (unk[2 * n + i] + 0.5 * dt * frm3) * cw[i];
      u1[i] = frc[i];
      u1[n + i] = frc[n + i];
      u1[2 * n + i] = frc[2 * n + i];
    }
    oc_gcrk(c, c->prpot, u1, u1 + n, u1 + 2 * n, uuu, uuu + n, uuu + 2 * n, 1);
    oc_prfr(c, c->prpot, u1, u1 + n, u1 + 2 * n, uuu, uuu + n, uuu + 2 * n, 1);
    for (int i = 0; i < n; ++i) {
      ud[i] = uuu[i];
      ud[n + i] = uuu[n + i];
      ud[2 * n + i] = uuu[2 * n + i];
      if (mtri == 2) {
        frc[i] = (ud[i] - frc[i]) * 2. / dt;
        frc[n + i] = (ud[n + i] - frc[n + i]) * 2. / dt;
        double frc33 = frc[2 * n + i] / cw[i] - (G * unk[3 * n + i] / th0[i]) * 0.5 * dt;
        frc[2 * n + i] = (ud[2 * n + i] - frc33) * 2. / dt;
        unk[i] = ud[i];
        unk[n + i] = ud[n + i];
        unk[2 * n + i] = ud[2 * n + i];
        frc[3 * n + i] = -(unk[2 * n + i] * c->dthedz[i]);
        unk[3 * n + i] = unk[3 * n + i] + 0.5 * dt * frc[3 * n + i];
      }
    }
  }
  for (int i = 0; i < n; ++i) {
    double frm1 = 0. - al[i] * (unk[i] - ue[i]);
    double frm2 = 0. - al[i] * (unk[n + i] - ve[i]);
    double frm3 = 0. - al[i] * unk[2 * n + i];
    double frm4 = -(al[i] * unk[3 * n + i]);
    frc[i] = frc[i] + frm1;
    frc[n + i] = frc[n + i] + frm2;
    frc[2 * n + i] = frc[2 * n + i] + frm3;
    frc[3 * n + i] = frc[3 * n + i] + frm4;
    if (traj) continue;
    double rh = c->rh0[i];
    uuu[i] = 1.5 * uuu[i] * rh - 0.5 * uuu0[i];
    uuu[n + i] = 1.5 * uuu[n + i] * rh - 0.5 * uuu0[n + i];
    uuu[2 * n + i] = 1.5 * uuu[2 * n + i] * rh - 0.5 * uuu0[2 * n + i];
    double rhoi = 1. / rh;
    uuu0[i] = uuu0[i] * rhoi;
    uuu0[n + i] = uuu0[n + i] * rhoi;
    uuu0[2 * n + i] = uuu0[2 * n + i] * rhoi;
    ud[i] = 2. * unk[i] - uuu0[i];
    ud[n + i] = 2. * unk[n + i] - uuu0[n + i];
    ud[2 * n + i] = 2. * unk[2 * n + i] - uuu0[2 * n + i];
  }
  if (traj) oc_velprd(c);
}
