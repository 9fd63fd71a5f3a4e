// C ABI of the gfx950 NFT step (declarations and contracts in include/nft.h).
//
// nft_init builds the device-resident mesh (node-sorted SELL-64 incidence, edge metrics,
// boundary entries grouped per node), computes the once-per-run set-up fields on the host
// in the reference's operation order, and allocates every work array, so the time-step
// path allocates nothing. The reference subroutines map onto kernels as
//   derivative -> k_grad<GRAD_RAW>             prfr  -> k_grad<GRAD_PRFR>
//   laplc      -> k_grad<GRAD_LAPLC> + k_div<DIV_LAPLC>
//   div        -> k_div<DIV_DIV>               bcv   -> fused into the epilogues (k_bcv alone)
//   precon     -> k_line (vertical Thomas) | k_point (+ laplc for later sweeps)
//   gcrk       -> the loop below: k_gcr_update, precon, laplc with fused <ax_ii,ar>, k_gcr_dir
//   mpdatam    -> k_mp_donor, k_mp_grad, k_mp_fct, k_mp_limit (all fields of a step at once)
// Only the GCR exit test needs the host: two doubles per inner iteration are read back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/nft.h"
#include "nft_kernels.h"
#include "nft_net.h"

using namespace nft;

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string &m) {
  g_err = m;
  return code;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return set_err(-2, std::string("HIP error: ") + hipGetErrorString(e_) + \
                                                 " at " #expr);                               \
  } while (0)

static const double kG = 9.80616;  // reference :94

}  // namespace

// ---- fixed scratch of the allocation-free entry point (nft_laplc_noalloc) ---------------
// laplc's intermediate, the scaled gradient ((x, y) pairs for every node, then z), for
// meshes of at most kScratchNodes nodes: 3 x 8 B x 1,048,576 = 24 MiB, held by every GPU
// the library runs on until the process ends. In host code `scratch` names a host-side
// stand-in, not device memory: kernels get the per-device address that
// hipGetSymbolAddress returns (g_scratch_addr), looked up once per device by nft_init.
// The array is shared by every call on one GPU and keeps the last call's contents.
static constexpr int kScratchNodes = 1 << 20;
static constexpr int kScratchDevices = 64;
static __device__ __align__(256) double scratch[3 * (size_t)kScratchNodes];
static double *g_scratch_addr[kScratchDevices] = {};

namespace {

// the device address of `scratch` on the current device (looked up on first use per device)
int scratch_address(double **out) {
  *out = nullptr;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) return set_err(-2, "scratch: device ordinal out of range");
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(-2, std::string("scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_scratch_addr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_addr[dev];
  return 0;
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
};

}  // namespace

struct nft_ctx {
  nft_params P;
  int n = 0, nedge = 0, nbface = 0, nboun = 0, etot = 0;
  int ncol = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {};
  std::vector<void *> allocs;

  Incidence inc;
  // every edge once, in the role of its first node i1 (MPDATA's antidiffusive-flux sweep);
  // same chunks and launch shape as inc
  Incidence inc_own{};
  double own_col_bytes = 0;
  // column patches of the LDS-staged MPDATA bounds and limiter sweeps (columns of 64 levels,
  // closed box): the incidence with LDS-relative neighbours, the patch tables, and the
  // largest staged set (columns); patch_on = 0 keeps the sweeps on memory gathers
  Incidence inc_lds{};
  MpPatches patches{};
  int patch_smax = 0, patch_on = 0;
  int gather_grid = 0;   // launched blocks of gather kernels
  int vec_grid = 0;      // fixed grid of grid-stride vector kernels
  int nodes_grid = 0;    // one thread per node
  EdgeS<double> S;
  EdgeS<float> Sf;

  // node fields (device)
  double *vol, *rh0, *th0, *ue, *ve, *coefw, *alpha, *d1, *dthedz, *pstep;
  NodeCoef nc{};
  float *volf, *rh0f;
  double *vr = nullptr;       // vol*rh0 (MPDATA)
  float *vrf = nullptr;       // the same product of the FP32 copies
  int32_t *col_start;
  double *la, *linv, *lc;
  float *laf, *linvf, *lcf;
  int line_L = 0;             // uniform column length <= 64 (wavefront scans), else 0
  // precon 3: in-column band of laplc (host rows bm2..bp2, factors interleaved on the device)
  std::vector<double> h_band[5];
  BandFactors<double> band{};
  BandFactors<float> bandf{};   // FP32 copies (precision >= 1)
  int band_L = 0, band_cpb = 0, band_colors = 0;
  int line_R = 0;             // rows per lane of the coarsened scan for 64 < L <= 512, else 0 (Thomas)
  int col_L = 0;              // uniform column length with nodes numbered column by column (0: not)
  int col_on = 0;             // gathers use the columnar incidence (inc.col_L > 0)
  int64_t steps_done = 0;     // time steps since nft_state_init (or as loaded with the state)
  // diagnostics recorded by nft_run_steps: device ring of records, host side of each record
  int diag_every = 0, diag_cap = 0, diag_flag = 1;
  double *diag_ring = nullptr, *diag_now = nullptr;
  std::vector<nft_diag> diag_meta;   // eer, eem, niter, nitav, step of each pending record
  std::vector<int> diag_slot;
  double col_bytes = 0;       // bytes of the columnar tables
  int n_flag78 = 0;           // boundary entries with flag 7 or 8 (bcv reads ue there)
  double *zbuf;
  float *zbuff;

  // host copies of set-up fields
  std::vector<double> h_alpha, h_d1, h_coefw, h_la, h_linv, h_lc;
  std::vector<std::pair<int, int>> h_hedges;   // column pairs joined by same-level edges
  double beta1 = 0;

  // state
  double *unk[4], *frc[4], *uuu[3], *uuu0[3], *ud[3], *prpot;
  float *uuuf[3];

  // work
  double *t[8];          // prfr/div/laplc temporaries
  double2 *txy;          // laplc: (x, y) of the scaled gradient (z in t[2])
  F3 *t3f = nullptr;        // the FP32 operator's scaled gradient, packed (precision 2)
  double op32_drop = 1.0e-3;   // recursive-residual drop that ends a pass (precision 2)
  double *gx[NFT_MAXD], *gax[NFT_MAXD];
  double *r, *qr, *ar, *pw, *rw;
  double *mp_psi1[4];
  void *mp_g[12], *mp_c[8];
  void *mp_flux = nullptr;   // antidiffusive velocity per edge: two (field, field) pair arrays
  double *partials;
  // second partials buffer: with fold, the direction kernel writes here while it reads the
  // divergence's partials, and the next update reads them here while it writes its own
  double *partials2 = nullptr;
  int fold = 0;        // finalize folded into the consuming kernels (small grids, lord <= NFT_FOLD_MAXD)
  ExitFold xfold{};    // set around a device-loop iteration whose exit test the band solve takes
  unsigned *counter;   // [0]: arrivals at the end of k_reduce_final (zero between launches)
  double *fin_wres;    // k_reduce_final's wavefront results: NFT_FIN_WG x (NFT_MAXD + 1)
  double *scal;
  double *h_scal;  // pinned
  // device-controlled GCR loop: exit flag handed to the loop kernels (null outside it) and
  // the same decision mirrored into mapped host memory for the host's poll
  const double *skip = nullptr;
  double *h_flag = nullptr, *d_hflag = nullptr;
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
  // io staging for the per-subroutine entry points
  double *io[8];
  // this device's address of the fixed scratch array (nft_laplc_noalloc); null if the
  // lookup failed at nft_init (the entry point then reports it)
  double *scratch_dev = nullptr;
  std::string scratch_err;

  nft_gcr_stats st{};

  // doubly periodic box (iper = 2): the coincident sets on the device (inc.per, inc.cslot
  // and inc.cval point into them) and, on the host, the sets as 0-based member lists
  int per = 0;
  PerSets ps{};
  std::vector<int4> h_sets;

  // multigrid branch of precon (mg_flag=1): coarse levels 1 and 2 on meshes of their own
  struct MgLevel {
    int n = 0, nodes_grid = 0, gather_grid = 0, col_on = 0;
    Incidence inc{};
    EdgeS<double> S{};
    NodeCoef nc{};
    double *vol = nullptr, *rho = nullptr, *pstep = nullptr;
    int32_t *match = nullptr;    // [n] 0-based node of the next finer level
    int32_t *imatch = nullptr;   // [n of the finer level] node of this level, -1 where none
    double *p = nullptr, *r = nullptr, *rrhs = nullptr, *rout = nullptr, *e = nullptr;
    std::vector<double> h_d;     // d2 / d3 of compute_d_unstr
    double beta = 0;             // betaunstr2 / betaunstr3
  } mg[2];
  int mg_ready = 0, mg_on = 0, mg_itrp2 = 0, mg_c1 = 0, mg_c2 = 0, mg_cc = 0, mg_transfer = 0;
  double mg_omega = 0;   // relaxation factor of the coarse levels
  double *mg_rrhs = nullptr, *mg_e = nullptr;

  // velprd (itraj = 1): its pressure and projected prediction (host and device pointer lists)
  double *vp_p = nullptr, *vp_ff[3] = {nullptr, nullptr, nullptr};
  double **vp_ffd = nullptr;

  // learned preconditioner (network, training pairs, inference state)
  struct Net {
    int ready = 0, H = 128, L = 0, Kp = 0, ncol = 0, max_rows = 0;
    int32_t *nptr = nullptr, *nidx = nullptr;
    int K[6] = {}, N[6] = {};
    float *w[12] = {}, *g[12] = {}, *m[12] = {}, *v[12] = {};
    uint16_t *wb[6] = {}, *wt[6] = {};
    float *X = nullptr, *A1 = nullptr, *C2 = nullptr, *C3 = nullptr, *A4 = nullptr, *A5 = nullptr, *Y = nullptr;
    float *T = nullptr, *dY = nullptr, *dZ5 = nullptr, *dZ4 = nullptr, *dC3 = nullptr, *dZ3 = nullptr;
    float *dC2 = nullptr, *dZ2 = nullptr, *dZ1 = nullptr, *P = nullptr, *lpart = nullptr;
    size_t P_count = 0;
    int step = 0;
    // training pairs: residual r and its line-preconditioned q, with s = max|r| and |r|^2;
    // pD: the pressure x_k of the iteration, replaced once its solve has converged by
    // x_k - x_final (= L^-1 r up to the final residual: the error objective's target)
    double *pR = nullptr, *pQ = nullptr, *pD = nullptr;
    std::vector<double> ps, pr2;
    std::vector<int> pimp;   // impfl of the solve each pair came from
    int npairs = 0, max_pairs = 0, stride = 1, seen = 0;
    // training work arrays (node length) and loss partials
    double *tz = nullptr, *tLz = nullptr, *ty = nullptr, *tg = nullptr, *tpart = nullptr;
    int dumping = 0;
    int objective = 0;     // training objective: 0 residual, 1 angle, 2 error
    int linear = 0;        // identity activations and zero biases: a fixed linear map
    int learned_iters = 0; // GCR iterations that used the learned correction
    int kind = 0;          // 0 off, 1 learned correction on top of the line preconditioner
    double stall = 0.99;   // fall back to the line preconditioner when rl2/rl2_prev exceeds this
    int active = 0;        // per solve: learned correction in use
    int fallbacks = 0;
  } net;

  // kernel timer
  int timer_kernel = 0;
  std::vector<hipEvent_t> tev;   // pairs
  int tev_used = 0;
  int t_launches = 0;
  double t_ms = 0;
  std::vector<double> tex;   // per recorded pair: bytes of the launch beyond kernel_bytes()
  double t_extra = 0;        // their sum over the folded pairs
};

namespace {

template <typename T>
int dalloc(nft_ctx *c, T **p, size_t n) {
  void *q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
  c->allocs.push_back(q);
  *p = (T *)q;
  return 0;
}

template <typename T>
int upload(nft_ctx *c, T **p, const T *h, size_t n) {
  int rc = dalloc(c, p, n);
  if (rc) return rc;
  if (n) HIPCHK(hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

inline int blocks_for(long n) { return (int)((n + NFT_BLOCK - 1) / NFT_BLOCK); }

// ---- set-up fields on the host, operation for operation as the reference -------------
// beta_timestep_unstr (:953-1109) over one mesh: half the square of the smallest
// vol/max|S| over interior edges off the box faces and above the hard-coded 1500 m
double beta_host(const nft_mesh *m, const double *d) {
  const int ldp = m->ld_poin;
  const double *X = m->coord, *Y = m->coord + ldp, *Z = m->coord + 2 * (size_t)ldp;
  double ddx = 10.e+06, hill = 1500.0;
  const int E = m->nedge, lde = m->ld_edge;
  for (int e = 0; e < E; ++e) {
    int i1 = m->iedge[e] - 1, i2 = m->iedge[lde + e] - 1;
    if (X[i1] != d[0] && X[i1] != d[1] && Y[i1] != d[2] && Y[i1] != d[3] && Z[i1] != d[4] &&
        Z[i1] != d[5] && X[i2] != d[0] && X[i2] != d[1] && Y[i2] != d[2] && Y[i2] != d[3] &&
        Z[i2] != d[4] && Z[i2] != d[5] && Z[i1] > hill && Z[i2] > hill) {
      double hx = std::fabs(m->sn[e]), hy = std::fabs(m->sn[lde + e]), hz = std::fabs(m->sn[2 * (size_t)lde + e]);
      double h = std::max(std::max(hx, hy), hz);
      ddx = std::min(std::min(ddx, m->vol[i1] / h), m->vol[i2] / h);
    }
  }
  return 0.5 * ddx * ddx * 1.0;
}

// compute_d_unstr (:1112-1284) over one mesh with density rh0u
void compute_d_host(const nft_mesh *m, const double *rh0, double dt, std::vector<double> &out) {
  const int n = m->npoin;
  std::vector<double> dix(n, 0.), diy(n, 0.), diz(n, 0.);
  for (int pass = 0; pass < 2; ++pass) {
    int ne = pass == 0 ? m->nedge : m->nbface;
    int ld = pass == 0 ? m->ld_edge : m->ld_face;
    const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
    const double *S = pass == 0 ? m->sn : m->sb;
    for (int e = 0; e < ne; ++e) {
      int i1 = ie[e] - 1, i2 = ie[ld + e] - 1;
      double sx = S[e], sy = S[ld + e], sz = S[2 * (size_t)ld + e];
      dix[i1] = dix[i1] + sx * sx * rh0[i2] / m->vol[i2];
      dix[i2] = dix[i2] + sx * sx * rh0[i1] / m->vol[i1];
      diy[i1] = diy[i1] + sy * sy * rh0[i2] / m->vol[i2];
      diy[i2] = diy[i2] + sy * sy * rh0[i1] / m->vol[i1];
      diz[i1] = diz[i1] + sz * sz * rh0[i2] / m->vol[i2];
      diz[i2] = diz[i2] + sz * sz * rh0[i1] / m->vol[i1];
    }
  }
  out.assign(n, 0.);
  for (int i = 0; i < n; ++i) out[i] = -(0.5 * dt * (dix[i] + diy[i] + diz[i]) / (4 * m->vol[i] * rh0[i]));
}

// point-relaxation step of precon (:587-599 fine, :632-653 coarse) from d and beta
void pstep_host(int sm_flag, const std::vector<double> &d, double beta, std::vector<double> &pstep) {
  pstep.assign(d.size(), 0.);
  for (size_t i = 0; i < d.size(); ++i) {
    if (sm_flag == 1) pstep[i] = (-1.) / d[i];
    else if (sm_flag == 2) pstep[i] = beta;
    else pstep[i] = beta / (1 - d[i] * beta);
  }
}

void setup_fields(nft_ctx *c, const nft_mesh *m, std::vector<double> &rh0, std::vector<double> &th0,
                  std::vector<double> &ue, std::vector<double> &ve, std::vector<double> &dthedz,
                  std::vector<double> &pstep) {
  const nft_params &p = c->P;
  const int n = m->npoin, ldp = m->ld_poin;
  const double *X = m->coord, *Y = m->coord + ldp, *Z = m->coord + 2 * (size_t)ldp;
  rh0.assign(n, 1.0); th0.assign(n, p.theta0); ue.assign(n, p.u0); ve.assign(n, p.v0);
  dthedz.assign(n, 0.0);
  for (int i = 0; i < n; ++i) dthedz[i] = p.theta0 * p.strat * 1.0;  // :192
  // absorbers (:2915-3026)
  c->h_alpha.assign(n, 0.0);
  if (p.irl == 1) {
    auto amax1 = [](double a, double b) { return a > b ? a : b; };
    auto amin1 = [](double a, double b) { return a < b ? a : b; };
    double eps = 1.e-07;
    double abxmax = -1.e+07, abxmin = 1.e+07, abymax = -1.e+07, abymin = 1.e+07, abzmax = -1.e+07,
           abzmin = 1.e+07;
    for (int i = 0; i < n; ++i) {
      abxmax = amax1(X[i], abxmax); abxmin = amin1(X[i], abxmin);
      abymax = amax1(Y[i], abymax); abymin = amin1(Y[i], abymin);
      abzmax = amax1(Z[i], abzmax); abzmin = amin1(Z[i], abzmin);
    }
    double r = 1.0;
    double xabr = abxmax * r - p.dxabr, xabl = abxmin * r + p.dxabl;
    double yabu = abymax * r - p.dyabu, yabd = abymin * r + p.dyabd;
    double zabu = abzmax - p.dzabu, zabd = abzmin + p.dzabd;
    double tolil = p.towxl != 0. ? 1. / p.towxl : 0., tolir = p.towxr != 0. ? 1. / p.towxr : 0.;
    double toliu = p.towyu != 0. ? 1. / p.towyu : 0., tolid = p.towyd != 0. ? 1. / p.towyd : 0.;
    double toliuz = p.towzu != 0. ? 1. / p.towzu : 0., tolidz = p.towzd != 0. ? 1. / p.towzd : 0.;
    for (int i = 0; i < n; ++i) {
      double relx, rely, relz, r0, r1, r2, relb;
      r0 = X[i] * r; r1 = 0.; r2 = 0.; relb = 0.;
      if (p.dxabl != 0.) r1 = amax1(0., xabl - r0) / p.dxabl;
      if (p.dxabr != 0.) r2 = amax1(0., r0 - xabr) / p.dxabr;
      if ((tolil + tolir) != 0.)
        relb = ((tolil * r1) * (tolil * r1) + (tolir * r2) * (tolir * r2)) / (tolil * r1 + tolir * r2 + eps);
      relx = relb;
      r1 = 0.; r2 = 0.; relb = 0.; r0 = Y[i] * r;
      if (p.dyabd != 0.) r1 = amax1(0., -r0 + yabd) / p.dyabd;
      if (p.dyabu != 0.) r2 = amax1(0., r0 - yabu) / p.dyabu;
      if ((tolid + toliu) != 0.)
        relb = ((tolid * r1) * (tolid * r1) + (toliu * r2) * (toliu * r2)) / (tolid * r1 + toliu * r2 + eps);
      rely = relb;
      r1 = 0.; r2 = 0.; relb = 0.; r0 = Z[i];
      if (p.dzabd != 0.) r1 = amax1(0., -r0 + zabd) / p.dzabd;
      if (p.dzabu != 0.) r2 = amax1(0., r0 - zabu) / p.dzabu;
      if ((tolidz + toliuz) != 0.)
        relb = ((tolidz * r1) * (tolidz * r1) + (toliuz * r2) * (toliuz * r2)) / (tolidz * r1 + toliuz * r2 + eps);
      relz = relb;
      double alpha1 = (rely * rely + relx * relx) / (rely + relx + eps);
      c->h_alpha[i] = (alpha1 * alpha1 + relz * relz) / (alpha1 + relz + eps);
    }
  }
  // beta_timestep_unstr(1) (:953-1109)
  c->beta1 = beta_host(m, p.dom);
  // compute_d_unstr(1) (:1112-1284)
  compute_d_host(m, rh0.data(), p.dt, c->h_d1);
  // coefw (:250-252)
  c->h_coefw.assign(n, 0.);
  for (int i = 0; i < n; ++i) {
    double gth0 = kG / th0[i];
    double cw = 1. + 0.25 * (p.dt * p.dt) * gth0 * dthedz[i];
    c->h_coefw[i] = 1. / cw;
  }
  // point relaxation step (precon :587-599)
  pstep_host(p.sm_flag, c->h_d1, c->beta1, pstep);
  // vertical line preconditioner (no reference counterpart): two-point-flux edge Laplacian
  // c_ij = 0.5*dt * 0.5(rho_i+rho_j) (Sx^2+Sy^2+cw Sz^2) / max(S.d, 1e-3|S||d|) / (vol_i rho_i),
  // M_ii = -sum_j c_ij, M_{i,i+-1} = c_ij along the column; Thomas factors per column.
  c->h_la.assign(n, 0.); c->h_linv.assign(n, 0.); c->h_lc.assign(n, 0.);
  if (m->ncol > 0) {
    std::vector<double> diag(n, 0.), lo(n, 0.), up(n, 0.);
    std::vector<int> colof(n, -1), levof(n, -1);
    for (int k = 0; k < m->ncol; ++k)
      for (int i = m->col_start[k]; i < m->col_start[k + 1]; ++i) { colof[i] = k; levof[i] = i - m->col_start[k]; }
    double hdt = 0.5 * p.dt;
    for (int pass = 0; pass < 2; ++pass) {
      int ne = pass == 0 ? m->nedge : m->nbface;
      int ld = pass == 0 ? m->ld_edge : m->ld_face;
      const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
      const double *S = pass == 0 ? m->sn : m->sb;
      for (int e = 0; e < ne; ++e) {
        int i = ie[e] - 1, j = ie[ld + e] - 1;
        double sx = S[e], sy = S[ld + e], sz = S[2 * (size_t)ld + e];
        double dx = X[j] - X[i], dy = Y[j] - Y[i], dz = Z[j] - Z[i];
        double sd = sx * dx + sy * dy + sz * dz;
        double sl = std::sqrt(sx * sx + sy * sy + sz * sz), dl = std::sqrt(dx * dx + dy * dy + dz * dz);
        double den = sd > 1e-3 * sl * dl ? sd : 1e-3 * sl * dl;
        double cw = 0.5 * (c->h_coefw[i] + c->h_coefw[j]);
        double rb = 0.5 * (rh0[i] + rh0[j]);
        double k0 = hdt * rb * (sx * sx + sy * sy + cw * sz * sz) / den;
        double ci = k0 / (m->vol[i] * rh0[i]);
        double cj = k0 / (m->vol[j] * rh0[j]);
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
        c->h_la[i] = a; c->h_linv[i] = inv; c->h_lc[i] = cp;
        cprev = cp;
      }
    }
    int L0 = m->col_start[1] - m->col_start[0];
    bool uniform = true;
    for (int k = 0; k < m->ncol && uniform; ++k)
      if (m->col_start[k + 1] - m->col_start[k] != L0 || m->col_start[k] != k * L0) uniform = false;
    c->col_L = uniform ? L0 : 0;
    c->line_L = uniform && L0 <= 64 ? L0 : 0;
    c->line_R = uniform && L0 > 64 && L0 <= 64 * 8 ? (L0 + 63) / 64 : 0;
  }
}

// ---- kernel timer ------------------------------------------------------------------------
enum { KT_LGRAD = 1, KT_LDIV = 2, KT_LINE = 3, KT_UPD = 4, KT_DIR = 5, KT_MP0 = 6, KT_MG = 10, KT_NET = 11, KT_MPF = 12 };

inline bool timing(nft_ctx *c, int k) { return c->timer_kernel == k && c->tev_used + 2 <= (int)c->tev.size(); }
inline void t_begin(nft_ctx *c, int k) {
  if (timing(c, k)) (void)hipEventRecord(c->tev[c->tev_used], c->stream);
}
// extra: algorithmic bytes of this launch beyond kernel_bytes() (the fused dots' inputs)
inline void t_end(nft_ctx *c, int k, double extra = 0.0) {
  if (timing(c, k)) {
    (void)hipEventRecord(c->tev[c->tev_used + 1], c->stream);
    c->tex[c->tev_used / 2] = extra;
    c->tev_used += 2;
  }
}
// fold the first `upto` recorded events (finished pairs) into the totals and move the rest
// to the front
void t_harvest_prefix(nft_ctx *c, int upto) {
  upto = std::min(upto, c->tev_used) & ~1;
  for (int i = 0; i + 1 < upto; i += 2) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->tev[i], c->tev[i + 1]) == hipSuccess) {
      c->t_ms += ms;
      c->t_launches++;
      c->t_extra += c->tex[i / 2];
    }
  }
  std::rotate(c->tev.begin(), c->tev.begin() + upto, c->tev.begin() + c->tev_used);
  std::rotate(c->tex.begin(), c->tex.begin() + upto / 2, c->tex.begin() + c->tev_used / 2);
  c->tev_used -= upto;
}
// after a stream synchronisation: fold finished event pairs into the totals
void t_harvest(nft_ctx *c) { t_harvest_prefix(c, c->tev_used); }

// ---- device operators -------------------------------------------------------------------
// periodic box: the set kernel of the three gathered sums that the next k_grad / k_div
// launch takes at the coincident nodes
void per_sums3(nft_ctx *c, const double *q0, const double *q1, const double *q2, const double2 *qxy,
               const double *scale) {
  PerSums3 a{q0, q1, q2, qxy, scale, c->skip};
  hipLaunchKernelGGL(k_per_sums3, dim3((c->ps.nset + NFT_SETS_PER_BLOCK - 1) / NFT_SETS_PER_BLOCK), dim3(NFT_BLOCK), 0,
                     c->stream, c->inc, c->ps, a);
}

int grad_launch(nft_ctx *c, int mode, const GradArgs &a) {
  dim3 g(c->gather_grid), b(NFT_BLOCK);
  if (c->per && mode != GRAD_ZERO_PRFR) per_sums3(c, a.f, a.f, a.f, nullptr, nullptr);
  if (mode == GRAD_LAPLC) t_begin(c, KT_LGRAD);
  if (c->col_on) {
    switch (mode) {
      case GRAD_RAW: hipLaunchKernelGGL((k_grad<GRAD_RAW, true>), g, b, 0, c->stream, c->inc, c->S, a); break;
      case GRAD_PRFR: hipLaunchKernelGGL((k_grad<GRAD_PRFR, true>), g, b, 0, c->stream, c->inc, c->S, a); break;
      case GRAD_LAPLC: hipLaunchKernelGGL((k_grad<GRAD_LAPLC, true>), g, b, 0, c->stream, c->inc, c->S, a); break;
      case GRAD_ADV: hipLaunchKernelGGL((k_grad<GRAD_ADV, true>), g, b, 0, c->stream, c->inc, c->S, a); break;
      default: hipLaunchKernelGGL((k_grad<GRAD_ZERO_PRFR, true>), g, b, 0, c->stream, c->inc, c->S, a); break;
    }
  } else {
    switch (mode) {
      case GRAD_RAW: hipLaunchKernelGGL((k_grad<GRAD_RAW, false>), g, b, 0, c->stream, c->inc, c->S, a); break;
      case GRAD_PRFR: hipLaunchKernelGGL((k_grad<GRAD_PRFR, false>), g, b, 0, c->stream, c->inc, c->S, a); break;
      case GRAD_LAPLC: hipLaunchKernelGGL((k_grad<GRAD_LAPLC, false>), g, b, 0, c->stream, c->inc, c->S, a); break;
      case GRAD_ADV: hipLaunchKernelGGL((k_grad<GRAD_ADV, false>), g, b, 0, c->stream, c->inc, c->S, a); break;
      default: hipLaunchKernelGGL((k_grad<GRAD_ZERO_PRFR, false>), g, b, 0, c->stream, c->inc, c->S, a); break;
    }
  }
  if (mode == GRAD_LAPLC) t_end(c, KT_LGRAD);
  HIPCHK(hipGetLastError());
  return 0;
}

GradArgs grad_args(nft_ctx *c) {
  GradArgs a{};
  a.nc = c->nc; a.dt = c->P.dt; a.skip = c->skip;
  return a;
}

DivArgs div_args(nft_ctx *c) {
  DivArgs a{};
  a.nc = c->nc; a.partials = c->partials; a.counter = c->counter;
  a.scal = c->scal; a.epa = 1.e-38; a.skip = c->skip;
  return a;
}

template <int MODE, int ND, int NMAX, typename TM = double>
void div_go(nft_ctx *c, const DivArgs &a) {
  if (c->col_on)
    hipLaunchKernelGGL((k_div<MODE, ND, NMAX, true, TM>), dim3(c->gather_grid), dim3(NFT_BLOCK), 0, c->stream,
                       c->inc, c->S, a);
  else
    hipLaunchKernelGGL((k_div<MODE, ND, NMAX, false, TM>), dim3(c->gather_grid), dim3(NFT_BLOCK), 0, c->stream,
                       c->inc, c->S, a);
}

template <int ND, typename TM>
bool div_laplc_nd(nft_ctx *c, const DivArgs &a, int nd) {
  if constexpr (ND > NFT_MAXD) {
    return false;
  } else {
    if (nd == ND) { div_go<DIV_LAPLC, ND, 0, TM>(c, a); return true; }
    return div_laplc_nd<ND + 1, TM>(c, a, nd);
  }
}

template <int K>
bool final_go(nft_ctx *c, int nblk, const FinalOp &op) {
  if constexpr (K > NFT_MAXD + 1) {
    return false;
  } else {
    if (op.K == K) {
      // small grids: one wavefront replays the 16 wavefronts' sums (same result, no counter)
      if (nblk <= NFT_FOLD_BLOCKS)
        hipLaunchKernelGGL(k_reduce_small<K>, dim3(1), dim3(64), 0, c->stream, c->partials, nblk, op, c->scal);
      else
        hipLaunchKernelGGL(k_reduce_final<K>, dim3(NFT_FIN_WG), dim3(64), 0, c->stream, c->partials, nblk, op,
                           c->scal, c->fin_wres, c->counter);
      return true;
    }
    return final_go<K + 1>(c, nblk, op);
  }
}

int final_launch(nft_ctx *c, int nblk, const FinalOp &op0) {
  FinalOp op = op0;
  op.skip = c->skip;
  if (!final_go<1>(c, nblk, op)) return set_err(-1, "reduction: too many values per block");
  HIPCHK(hipGetLastError());
  return 0;
}

// f32: the FP32 operator's divergence (FP32 metrics and inputs, DIV_LAPLC only)
int div_launch(nft_ctx *c, int mode, const DivArgs &a, bool f32 = false) {
  const int nd = a.dots.nd, nmax = a.dots.nmax ? 1 : 0;
  if (c->per) {
    if (f32) return set_err(-1, "periodic box: the FP32 operator is not available");
    if (mode == DIV_LAPLC) per_sums3(c, nullptr, nullptr, a.q2, a.qxy, nullptr);
    else per_sums3(c, a.q0, a.q1, a.q2, nullptr, a.iflg == -1 && !c->nc.rh0_one ? c->rh0 : nullptr);
  }
  if (mode == DIV_LAPLC) {
    if (nmax) return set_err(-1, "laplc: max reduction not supported");
    t_begin(c, KT_LDIV);
    bool ok = f32 ? div_laplc_nd<0, float>(c, a, nd) : div_laplc_nd<0, double>(c, a, nd);
    // the fused dots read one node array each, except a dot of the output with itself
    int reads = 0;
    for (int k = 0; k < nd; ++k) reads += a.dots.a[k] ? 1 : 0;
    t_end(c, KT_LDIV, 8.0 * c->n * reads);
    if (!ok) return set_err(-1, "laplc: too many fused dots");
  } else {
    if (nd == 0 && nmax == 0) div_go<DIV_DIV, 0, 0>(c, a);
    else if (nd == 1 && nmax == 1) div_go<DIV_DIV, 1, 1>(c, a);
    else if (nd == 1 && nmax == 0) div_go<DIV_DIV, 1, 0>(c, a);
    else return set_err(-1, "div: unsupported fused reduction");
  }
  HIPCHK(hipGetLastError());
  if (nd + nmax > 0 && !a.dots.fold) {
    FinalOp op{};
    op.kind = FK_DIV; op.K = nd + nmax; op.ismax = nmax ? (1u << nd) : 0u; op.dots = a.dots; op.epa = a.epa;
    return final_launch(c, c->gather_grid, op);
  }
  return 0;
}

// laplc(press, alap, impfl) with optional fused dots on alap
int laplc_dev(nft_ctx *c, const double *press, double *alap, int impfl, const DotList *dots) {
  GradArgs g = grad_args(c);
  g.f = press; g.oxy = c->txy; g.o2 = c->t[2]; g.impfl = impfl;
  int rc = grad_launch(c, GRAD_LAPLC, g);
  if (rc) return rc;
  DivArgs d = div_args(c);
  d.qxy = c->txy; d.q2 = c->t[2]; d.out = alap;
  if (dots) d.dots = *dots;
  return div_launch(c, DIV_LAPLC, d);
}

// laplc with FP32 edge metrics and FP32 gradient temporaries, FP64 sums and output: the
// operator of the Krylov iterations with precision 2
int laplc32_dev(nft_ctx *c, const double *press, double *alap, int impfl, const DotList *dots) {
  if (!c->t3f || !c->inc.s3f) return set_err(-1, "laplc32: FP32 operator not set up (precision 2 only)");
  GradArgs g = grad_args(c);
  g.f = press; g.o3f = c->t3f; g.impfl = impfl;
  dim3 gg(c->gather_grid), b(NFT_BLOCK);
  t_begin(c, KT_LGRAD);
  if (c->col_on)
    hipLaunchKernelGGL((k_grad<GRAD_LAPLC, true, float>), gg, b, 0, c->stream, c->inc, c->S, g);
  else
    hipLaunchKernelGGL((k_grad<GRAD_LAPLC, false, float>), gg, b, 0, c->stream, c->inc, c->S, g);
  t_end(c, KT_LGRAD);
  HIPCHK(hipGetLastError());
  DivArgs d = div_args(c);
  d.q3f = c->t3f; d.out = alap;
  if (dots) d.dots = *dots;
  return div_launch(c, DIV_LAPLC, d, true);
}

// the matrix-vector product of the GCR iterations
int krylov_laplc(nft_ctx *c, const double *press, double *alap, int impfl, const DotList *dots) {
  return c->P.precision == 2 ? laplc32_dev(c, press, alap, impfl, dots) : laplc_dev(c, press, alap, impfl, dots);
}

// out = L^T y (transpose of laplc_dev, two gathers; uses t[0..2])
int laplt_dev(nft_ctx *c, const double *y, double *out, int impfl) {
  LaplTArgs a{};
  a.y = y; a.g0 = c->t[0]; a.g1 = c->t[1]; a.g2 = c->t[2]; a.out = out;
  a.nc = c->nc; a.dt = c->P.dt; a.impfl = impfl;
  dim3 g(c->gather_grid), b(NFT_BLOCK);
  hipLaunchKernelGGL(k_laplT_1, g, b, 0, c->stream, c->inc, c->S, a);
  hipLaunchKernelGGL(k_laplT_2, g, b, 0, c->stream, c->inc, c->S, a);
  HIPCHK(hipGetLastError());
  return 0;
}

int prfr_dev(nft_ctx *c, const double *p, const double *us, const double *vs, const double *ws,
             double *uo, double *vo, double *wo, int impfl) {
  GradArgs g = grad_args(c);
  g.f = p; g.s0 = us; g.s1 = vs; g.s2 = ws; g.o0 = uo; g.o1 = vo; g.o2 = wo; g.impfl = impfl;
  return grad_launch(c, p ? GRAD_PRFR : GRAD_ZERO_PRFR, g);
}

int div_dev(nft_ctx *c, const double *ua, const double *va, const double *wa, double *rhs, int iflg,
            const DotList *dots) {
  DivArgs d = div_args(c);
  d.q0 = ua; d.q1 = va; d.q2 = wa; d.out = rhs; d.iflg = iflg;
  if (dots) d.dots = *dots;
  return div_launch(c, DIV_DIV, d);
}

// coarsened line scan for uniform columns of 64 < L <= 64*R rows (R = line_R)
template <typename T>
void line_scan_r_go(nft_ctx *c, int cg, const T *la, const T *linv, const T *lc, const double *res,
                    const double *pin, double *p) {
  const int L = c->col_L;
  const double om = c->P.omega;
  switch (c->line_R) {
    case 2: hipLaunchKernelGGL((k_line_scan_r<T, 2>), dim3(cg), dim3(NFT_BLOCK), 0, c->stream, c->ncol, L, la, linv, lc, res, pin, p, om, c->skip); break;
    case 3: hipLaunchKernelGGL((k_line_scan_r<T, 3>), dim3(cg), dim3(NFT_BLOCK), 0, c->stream, c->ncol, L, la, linv, lc, res, pin, p, om, c->skip); break;
    case 4: hipLaunchKernelGGL((k_line_scan_r<T, 4>), dim3(cg), dim3(NFT_BLOCK), 0, c->stream, c->ncol, L, la, linv, lc, res, pin, p, om, c->skip); break;
    case 5: hipLaunchKernelGGL((k_line_scan_r<T, 5>), dim3(cg), dim3(NFT_BLOCK), 0, c->stream, c->ncol, L, la, linv, lc, res, pin, p, om, c->skip); break;
    case 6: hipLaunchKernelGGL((k_line_scan_r<T, 6>), dim3(cg), dim3(NFT_BLOCK), 0, c->stream, c->ncol, L, la, linv, lc, res, pin, p, om, c->skip); break;
    case 7: hipLaunchKernelGGL((k_line_scan_r<T, 7>), dim3(cg), dim3(NFT_BLOCK), 0, c->stream, c->ncol, L, la, linv, lc, res, pin, p, om, c->skip); break;
    default: hipLaunchKernelGGL((k_line_scan_r<T, 8>), dim3(cg), dim3(NFT_BLOCK), 0, c->stream, c->ncol, L, la, linv, lc, res, pin, p, om, c->skip); break;
  }
}

// one fine-mesh smoothing sweep of precon (:607-612, :740-745): r = L p (not on the first
// sweep from p = 0, where it is zero), then the point step p += omega*pstep*(r-rhs), or
// with the line preconditioner p += omega * M^-1 (rhs - r)
int fine_sweep(nft_ctx *c, const double *rhs, double *p, int impfl, bool first) {
  const nft_params &P = c->P;
  const int n = c->n;
  if (!first) {
    int rc = laplc_dev(c, p, c->rw, impfl, nullptr);
    if (rc) return rc;
  }
  if (P.precon == 1) {
    hipLaunchKernelGGL(k_point, dim3(c->nodes_grid), dim3(NFT_BLOCK), 0, c->stream, n, c->pstep,
                       c->rw, rhs, p, P.omega, first ? 1 : 0, c->skip);
  } else {
    const double *res = rhs;
    if (!first) {
      hipLaunchKernelGGL(k_sub, dim3(c->nodes_grid), dim3(NFT_BLOCK), 0, c->stream, n, rhs, c->rw, c->t[3],
                         c->skip);
      res = c->t[3];
    }
    t_begin(c, KT_LINE);
    const double *pin = first ? nullptr : p;
    if (P.precon == 3) {
      const int nb = (c->ncol + c->band_cpb - 1) / c->band_cpb;
      // the exit test of the update before it, folded in on small grids (first sweep only)
      const ExitFold xf = first ? c->xfold : ExitFold{};
      if (P.precision >= 1) {
        const size_t shm = (size_t)c->band_cpb * (c->band_L + 1) * sizeof(float);
        if (xf.part)
          hipLaunchKernelGGL((k_line_penta<float, true>), dim3(nb), dim3(NFT_WAVE), shm, c->stream, c->ncol,
                             c->band_L, c->band_cpb, c->bandf, res, pin, p, P.omega, c->skip, xf);
        else
          hipLaunchKernelGGL((k_line_penta<float, false>), dim3(nb), dim3(NFT_WAVE), shm, c->stream, c->ncol,
                             c->band_L, c->band_cpb, c->bandf, res, pin, p, P.omega, c->skip, xf);
      } else {
        const size_t shm = (size_t)c->band_cpb * (c->band_L + 1) * sizeof(double);
        if (xf.part)
          hipLaunchKernelGGL((k_line_penta<double, true>), dim3(nb), dim3(NFT_WAVE), shm, c->stream, c->ncol,
                             c->band_L, c->band_cpb, c->band, res, pin, p, P.omega, c->skip, xf);
        else
          hipLaunchKernelGGL((k_line_penta<double, false>), dim3(nb), dim3(NFT_WAVE), shm, c->stream, c->ncol,
                             c->band_L, c->band_cpb, c->band, res, pin, p, P.omega, c->skip, xf);
      }
    } else if (c->line_L > 0) {
      int cpw = 64 / c->line_L;
      long waves = (c->ncol + cpw - 1) / cpw;
      int cg = (int)((waves + NFT_WPB - 1) / NFT_WPB);
      if (P.precision >= 1)
        hipLaunchKernelGGL(k_line_scan<float>, dim3(cg), dim3(NFT_BLOCK), 0, c->stream, c->ncol, c->line_L,
                           c->laf, c->linvf, c->lcf, res, pin, p, P.omega, c->skip);
      else
        hipLaunchKernelGGL(k_line_scan<double>, dim3(cg), dim3(NFT_BLOCK), 0, c->stream, c->ncol, c->line_L,
                           c->la, c->linv, c->lc, res, pin, p, P.omega, c->skip);
    } else if (c->line_R > 0) {
      int cg = (int)((c->ncol + NFT_WPB - 1) / NFT_WPB);
      if (P.precision >= 1) line_scan_r_go<float>(c, cg, c->laf, c->linvf, c->lcf, res, pin, p);
      else line_scan_r_go<double>(c, cg, c->la, c->linv, c->lc, res, pin, p);
    } else {
      int cg = blocks_for(c->ncol);
      if (P.precision >= 1)
        hipLaunchKernelGGL(k_line<float>, dim3(cg), dim3(NFT_BLOCK), 0, c->stream, c->ncol, c->col_start,
                           c->laf, c->linvf, c->lcf, res, pin, p, P.omega, c->zbuff, c->skip);
      else
        hipLaunchKernelGGL(k_line<double>, dim3(cg), dim3(NFT_BLOCK), 0, c->stream, c->ncol, c->col_start,
                           c->la, c->linv, c->lc, res, pin, p, P.omega, c->zbuf, c->skip);
    }
    t_end(c, KT_LINE);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// laplc_c / laplc_cc (:1356-1535): laplc over a coarse level's mesh; the gradient
// temporaries are the fine mesh's (coarse levels are smaller)
int laplc_level(nft_ctx *c, const nft_ctx::MgLevel &L, const double *press, double *alap, int impfl) {
  GradArgs g{};
  g.nc = L.nc; g.dt = c->P.dt; g.skip = c->skip;
  g.f = press; g.oxy = c->txy; g.o2 = c->t[2]; g.impfl = impfl;
  DivArgs d{};
  d.nc = L.nc; d.partials = c->partials; d.counter = c->counter; d.scal = c->scal; d.epa = 1.e-38;
  d.skip = c->skip; d.qxy = c->txy; d.q2 = c->t[2]; d.out = alap;
  dim3 gg(L.gather_grid), b(NFT_BLOCK);
  if (L.col_on) {
    hipLaunchKernelGGL((k_grad<GRAD_LAPLC, true>), gg, b, 0, c->stream, L.inc, L.S, g);
    hipLaunchKernelGGL((k_div<DIV_LAPLC, 0, 0, true>), gg, b, 0, c->stream, L.inc, L.S, d);
  } else {
    hipLaunchKernelGGL((k_grad<GRAD_LAPLC, false>), gg, b, 0, c->stream, L.inc, L.S, g);
    hipLaunchKernelGGL((k_div<DIV_LAPLC, 0, 0, false>), gg, b, 0, c->stream, L.inc, L.S, d);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// point relaxation on a coarse level (:676-681, :703-708, :723-728): it sweeps of
// r = L p, p += omega*pstep*(r - rrhs)
int relax_level(nft_ctx *c, nft_ctx::MgLevel &L, int sweeps, int impfl) {
  for (int it = 1; it <= sweeps; ++it) {
    int rc = laplc_level(c, L, L.p, L.r, impfl);
    if (rc) return rc;
    hipLaunchKernelGGL(k_point, dim3(L.nodes_grid), dim3(NFT_BLOCK), 0, c->stream, L.n, L.pstep, L.r, L.rrhs,
                       L.p, c->mg_omega, 0, c->skip);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// the V-cycle of precon's multigrid branch (:618-748), after the fine smoothing
int mg_cycle_dev(nft_ctx *c, const double *rhs, double *p, int impfl) {
  auto &A = c->mg[0], &B = c->mg[1];
  const int n = c->n, tr = c->mg_transfer;
  const double om = c->mg_omega;
  const dim3 b(NFT_BLOCK);
  int rc;
  // fine residual, restricted to level 1
  if ((rc = laplc_dev(c, p, c->rw, impfl, nullptr))) return rc;
  t_begin(c, KT_MG);
  hipLaunchKernelGGL(k_sub, dim3(c->nodes_grid), b, 0, c->stream, n, rhs, c->rw, c->mg_rrhs, c->skip);
  hipLaunchKernelGGL(k_mg_restrict, dim3(A.nodes_grid), b, 0, c->stream, A.n, A.match, c->inc, c->mg_rrhs,
                     A.rrhs, tr, c->skip);
  // level 1: p = -(omega*pstep*rrhs), relax, residual, restrict to level 2
  hipLaunchKernelGGL(k_point, dim3(A.nodes_grid), b, 0, c->stream, A.n, A.pstep, A.r, A.rrhs, A.p, om, 1,
                     c->skip);
  if ((rc = relax_level(c, A, c->mg_c1, impfl))) return rc;
  if ((rc = laplc_level(c, A, A.p, A.r, impfl))) return rc;
  hipLaunchKernelGGL(k_sub, dim3(A.nodes_grid), b, 0, c->stream, A.n, A.rrhs, A.r, A.rout, c->skip);
  hipLaunchKernelGGL(k_mg_restrict, dim3(B.nodes_grid), b, 0, c->stream, B.n, B.match, A.inc, A.rout, B.rrhs,
                     tr, c->skip);
  // level 2: p = -(rrhs*pstep*omega), relax
  hipLaunchKernelGGL(k_mg_start_cc, dim3(B.nodes_grid), b, 0, c->stream, B.n, B.pstep, B.rrhs, B.p, om, c->skip);
  if ((rc = relax_level(c, B, c->mg_cc, impfl))) return rc;
  // back to level 1: correct, relax
  hipLaunchKernelGGL(k_mg_prolong, dim3(A.nodes_grid), b, 0, c->stream, A.n, B.imatch, A.inc, B.p, A.e, tr,
                     c->skip);
  hipLaunchKernelGGL(k_mg_add, dim3(A.nodes_grid), b, 0, c->stream, A.n, A.p, A.e, c->skip);
  if ((rc = relax_level(c, A, c->mg_c2, impfl))) return rc;
  // back to the fine mesh: correct, relax
  hipLaunchKernelGGL(k_mg_prolong, dim3(c->nodes_grid), b, 0, c->stream, n, A.imatch, c->inc, A.p, c->mg_e, tr,
                     c->skip);
  hipLaunchKernelGGL(k_mg_add, dim3(c->nodes_grid), b, 0, c->stream, n, p, c->mg_e, c->skip);
  t_end(c, KT_MG);
  HIPCHK(hipGetLastError());
  for (int it = 1; it <= c->mg_itrp2; ++it)
    if ((rc = fine_sweep(c, rhs, p, impfl, false))) return rc;
  return 0;
}

// precon(rhs,p,r,iprc,elas,impfl) (:482-748): itrp1 fine sweeps (point relaxation or the
// vertical line preconditioner), then the V-cycle when the multigrid branch is on
int precon_dev(nft_ctx *c, const double *rhs, double *p, int impfl) {
  const nft_params &P = c->P;
  if (P.precon == 0) {
    hipLaunchKernelGGL(k_copy, dim3(c->nodes_grid), dim3(NFT_BLOCK), 0, c->stream, c->n, rhs, p, c->skip);
    HIPCHK(hipGetLastError());
    return 0;
  }
  for (int it = 1; it <= P.itrp1; ++it) {
    int rc = fine_sweep(c, rhs, p, impfl, it == 1);
    if (rc) return rc;
  }
  if (c->mg_on) return mg_cycle_dev(c, rhs, p, impfl);
  return 0;
}


// ---- learned preconditioner ------------------------------------------------------------
constexpr int NET_SPLIT_ROWS = 1024;

int net_gemm(nft_ctx *c, int M, int N, int K, const float *A, int lda, const uint16_t *Bt, const float *bias,
             int relu, const float *mask, int ldm, float *C, int ldc) {
  if (K % NET_TK == 0) {
    dim3 grid((M + 63) / 64, (N + 63) / 64);
    hipLaunchKernelGGL(k_net_gemm_t, grid, dim3(256), 0, c->stream, M, N, K, A, lda, Bt, bias, relu, mask, ldm, C,
                       ldc);
  } else {
    dim3 grid((M + 31) / 32, (N + 31) / 32);
    hipLaunchKernelGGL(k_net_gemm, grid, dim3(64), 0, c->stream, M, N, K, A, lda, Bt, bias, relu, mask, ldm, C, ldc);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// gradient of weight l from activations A [M][K] and output gradient D [M][N]
int net_wgrad(nft_ctx *c, int l, int M, const float *A, int lda, const float *D, int ldd) {
  auto &n = c->net;
  const int K = n.K[l], N = n.N[l];
  const int nsplit = (M + NET_SPLIT_ROWS - 1) / NET_SPLIT_ROWS;
  hipLaunchKernelGGL(k_net_wgrad, dim3((K + 31) / 32, (N + 31) / 32, nsplit), dim3(64), 0, c->stream, M, K, N, A,
                     lda, D, ldd, NET_SPLIT_ROWS, n.P);
  hipLaunchKernelGGL(k_net_reduce_splits, dim3((K * N + 255) / 256), dim3(256), 0, c->stream, nsplit, K * N, n.P,
                     n.g[2 * l]);
  hipLaunchKernelGGL(k_net_colsum, dim3((N + 127) / 128, nsplit), dim3(128), 0, c->stream, M, N, D, ldd,
                     NET_SPLIT_ROWS, n.P);
  hipLaunchKernelGGL(k_net_reduce_splits, dim3((N + 255) / 256), dim3(256), 0, c->stream, nsplit, N, n.P,
                     n.g[2 * l + 1]);
  HIPCHK(hipGetLastError());
  return 0;
}

// forward pass on M = nsamples*ncol rows of n.X -> n.Y (activations kept for backward)
int net_forward(nft_ctx *c, int M) {
  auto &n = c->net;
  const int H = n.H, H2 = 2 * H;
  int rc;
  int tb = 256, gb = (M * H + tb - 1) / tb;
  const int act = n.linear ? 0 : 1;
  if ((rc = net_gemm(c, M, H, n.Kp, n.X, n.Kp, n.wt[0], n.w[1], act, nullptr, 0, n.A1, H))) return rc;
  if ((rc = net_gemm(c, M, H, H, n.A1, H, n.wt[1], n.w[3], act, nullptr, 0, n.C2, H2))) return rc;
  hipLaunchKernelGGL(k_net_agg, dim3(gb), dim3(tb), 0, c->stream, M, n.ncol, H, n.nptr, n.nidx, n.C2, H2, n.C2 + H, H2);
  if ((rc = net_gemm(c, M, H, H2, n.C2, H2, n.wt[2], n.w[5], act, nullptr, 0, n.C3, H2))) return rc;
  hipLaunchKernelGGL(k_net_agg, dim3(gb), dim3(tb), 0, c->stream, M, n.ncol, H, n.nptr, n.nidx, n.C3, H2, n.C3 + H, H2);
  if ((rc = net_gemm(c, M, H, H2, n.C3, H2, n.wt[3], n.w[7], act, nullptr, 0, n.A4, H))) return rc;
  if ((rc = net_gemm(c, M, H, H, n.A4, H, n.wt[4], n.w[9], act, nullptr, 0, n.A5, H))) return rc;
  if ((rc = net_gemm(c, M, n.Kp, H, n.A5, H, n.wt[5], n.w[11], 0, nullptr, 0, n.Y, n.Kp))) return rc;
  HIPCHK(hipGetLastError());
  return 0;
}

// gradients of every parameter into n.g from the output gradient n.dY
int net_backprop(nft_ctx *c, int M) {
  auto &n = c->net;
  const int H = n.H, H2 = 2 * H, Kp = n.Kp;
  int rc;
  int tb = 256, gb = (M * H + tb - 1) / tb;
  // layer 6: Y = A5 W6 + b6
  if ((rc = net_wgrad(c, 5, M, n.A5, H, n.dY, Kp))) return rc;
  if ((rc = net_gemm(c, M, H, Kp, n.dY, Kp, n.wb[5], nullptr, 0, n.linear ? nullptr : n.A5, H, n.dZ5, H))) return rc;
  // layer 5: A5 = relu(A4 W5 + b5)
  if ((rc = net_wgrad(c, 4, M, n.A4, H, n.dZ5, H))) return rc;
  if ((rc = net_gemm(c, M, H, H, n.dZ5, H, n.wb[4], nullptr, 0, n.linear ? nullptr : n.A4, H, n.dZ4, H))) return rc;
  // layer 4: A4 = relu([A3|Agg A3] W4 + b4)
  if ((rc = net_wgrad(c, 3, M, n.C3, H2, n.dZ4, H))) return rc;
  if ((rc = net_gemm(c, M, H2, H, n.dZ4, H, n.wb[3], nullptr, 0, nullptr, 0, n.dC3, H2))) return rc;
  hipLaunchKernelGGL(k_net_agg_back, dim3(gb), dim3(tb), 0, c->stream, M, n.ncol, H, n.nptr, n.nidx, n.dC3, H2, n.linear ? nullptr : n.C3,
                     H2, n.dZ3, H);
  // layer 3: A3 = relu([A2|Agg A2] W3 + b3)
  if ((rc = net_wgrad(c, 2, M, n.C2, H2, n.dZ3, H))) return rc;
  if ((rc = net_gemm(c, M, H2, H, n.dZ3, H, n.wb[2], nullptr, 0, nullptr, 0, n.dC2, H2))) return rc;
  hipLaunchKernelGGL(k_net_agg_back, dim3(gb), dim3(tb), 0, c->stream, M, n.ncol, H, n.nptr, n.nidx, n.dC2, H2, n.linear ? nullptr : n.C2,
                     H2, n.dZ2, H);
  // layer 2: A2 = relu(A1 W2 + b2)
  if ((rc = net_wgrad(c, 1, M, n.A1, H, n.dZ2, H))) return rc;
  if ((rc = net_gemm(c, M, H, H, n.dZ2, H, n.wb[1], nullptr, 0, n.linear ? nullptr : n.A1, H, n.dZ1, H))) return rc;
  // layer 1: A1 = relu(X W1 + b1)
  if ((rc = net_wgrad(c, 0, M, n.X, Kp, n.dZ1, H))) return rc;
  HIPCHK(hipGetLastError());
  return 0;
}

// loss (MSE over the first L of Kp outputs against n.T) and gradients of every parameter
int net_backward(nft_ctx *c, int M, double *loss) {
  auto &n = c->net;
  const int nl = (M * n.Kp + 255) / 256;
  const float inv = 1.0f / (float)((size_t)M * n.L);
  hipLaunchKernelGGL(k_net_loss, dim3(nl), dim3(256), 0, c->stream, M, n.L, n.Kp, n.Y, n.T, inv, n.dY, n.lpart);
  int rc = net_backprop(c, M);
  if (rc) return rc;
  if (loss) {
    std::vector<float> part(nl);
    HIPCHK(hipMemcpyAsync(part.data(), n.lpart, nl * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double a = 0;
    for (int i = 0; i < nl; ++i) a += part[i];
    *loss = a * inv;
  }
  return 0;
}

int net_alloc(nft_ctx *c, int H, int max_rows, uint64_t seed) {
  auto &n = c->net;
  int rc;
  n.H = H;
  n.L = c->col_L;
  n.Kp = (n.L + 7) / 8 * 8;
  n.ncol = c->ncol;
  n.max_rows = max_rows;
  const int Kp = n.Kp, H2 = 2 * H;
  const int Ks[6] = {Kp, H, H2, H2, H, H}, Ns[6] = {H, H, H, H, H, Kp};
  // He initialisation from a fixed-seed 64-bit generator (Box-Muller), output layer x0.1
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
  auto next = [&]() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; };
  auto unif = [&]() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  for (int l = 0; l < 6; ++l) {
    n.K[l] = Ks[l]; n.N[l] = Ns[l];
    size_t cnt = (size_t)Ks[l] * Ns[l];
    std::vector<float> w(cnt), z(cnt, 0.f), zb(Ns[l], 0.f);
    double std_ = std::sqrt(2.0 / Ks[l]) * (l == 5 ? 0.1 : 1.0);
    for (size_t i = 0; i < cnt; ++i) {
      double u1 = unif(), u2 = unif();
      w[i] = (float)(std_ * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
    // rows of W1 beyond L see only zero padding; keep them zero
    if (l == 0)
      for (int k = n.L; k < Kp; ++k)
        for (int j = 0; j < Ns[l]; ++j) w[(size_t)k * Ns[l] + j] = 0.f;
    if ((rc = upload(c, &n.w[2 * l], w.data(), cnt)) || (rc = upload(c, &n.g[2 * l], z.data(), cnt)) ||
        (rc = upload(c, &n.m[2 * l], z.data(), cnt)) || (rc = upload(c, &n.v[2 * l], z.data(), cnt)) ||
        (rc = upload(c, &n.w[2 * l + 1], zb.data(), Ns[l])) || (rc = upload(c, &n.g[2 * l + 1], zb.data(), Ns[l])) ||
        (rc = upload(c, &n.m[2 * l + 1], zb.data(), Ns[l])) || (rc = upload(c, &n.v[2 * l + 1], zb.data(), Ns[l])) ||
        (rc = dalloc(c, &n.wb[l], cnt)) || (rc = dalloc(c, &n.wt[l], cnt)))
      return rc;
    hipLaunchKernelGGL(k_net_to_bf16, dim3((cnt + 255) / 256), dim3(256), 0, c->stream, Ks[l], Ns[l], n.w[2 * l],
                       n.wb[l], n.wt[l]);
  }
  const size_t R = max_rows;
  if ((rc = dalloc(c, &n.X, R * Kp)) || (rc = dalloc(c, &n.A1, R * H)) || (rc = dalloc(c, &n.C2, R * H2)) ||
      (rc = dalloc(c, &n.C3, R * H2)) || (rc = dalloc(c, &n.A4, R * H)) || (rc = dalloc(c, &n.A5, R * H)) ||
      (rc = dalloc(c, &n.Y, R * Kp)) || (rc = dalloc(c, &n.T, R * Kp)) || (rc = dalloc(c, &n.dY, R * Kp)) ||
      (rc = dalloc(c, &n.dZ5, R * H)) || (rc = dalloc(c, &n.dZ4, R * H)) || (rc = dalloc(c, &n.dC3, R * H2)) ||
      (rc = dalloc(c, &n.dZ3, R * H)) || (rc = dalloc(c, &n.dC2, R * H2)) || (rc = dalloc(c, &n.dZ2, R * H)) ||
      (rc = dalloc(c, &n.dZ1, R * H)) || (rc = dalloc(c, &n.lpart, (R * Kp + 255) / 256 + 1)))
    return rc;
  n.P_count = (size_t)((R + NET_SPLIT_ROWS - 1) / NET_SPLIT_ROWS) * H2 * std::max(H, Kp);
  if ((rc = dalloc(c, &n.P, n.P_count))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  n.ready = 1;
  return 0;
}

// learned correction on top of a line-preconditioned qr: qr += s * N(r / s), s = scal[sidx]
int net_apply_precon(nft_ctx *c, const double *r, double *qr, int sidx) {
  auto &n = c->net;
  const int M = n.ncol;
  hipLaunchKernelGGL(k_net_input, dim3((M * n.Kp + 255) / 256), dim3(256), 0, c->stream, n.ncol, n.L, n.Kp, r,
                     c->scal, sidx, n.X);
  t_begin(c, KT_NET);
  int rc = net_forward(c, M);
  t_end(c, KT_NET);
  if (rc) return rc;
  hipLaunchKernelGGL(k_net_output, dim3((M * n.L + 255) / 256), dim3(256), 0, c->stream, n.ncol, n.L, n.Kp, n.Y,
                     c->scal, sidx, qr);
  HIPCHK(hipGetLastError());
  return 0;
}

// keep every stride-th (r, line-preconditioned q) of the solves while the store has room
int net_store_pair(nft_ctx *c, double s, double r2, int impfl, const double *p) {
  auto &n = c->net;
  if (n.seen++ % n.stride != 0 || n.npairs >= n.max_pairs || !(s > 0) || !(r2 > 0)) return 0;
  const size_t N = c->n;
  HIPCHK(hipMemcpyAsync(n.pR + N * n.npairs, c->r, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(n.pQ + N * n.npairs, c->qr, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(n.pD + N * n.npairs, p, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  n.ps.push_back(s);
  n.pr2.push_back(r2);
  n.pimp.push_back(impfl);
  n.npairs++;
  return 0;
}

int read_scal(nft_ctx *c, int first, int count) {
  HIPCHK(hipMemcpyAsync(c->h_scal + first, c->scal + first, count * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  t_harvest(c);
  return 0;
}

int norms_dev(nft_ctx *c, const double *a, int sum_slot, int max_slot) {
  hipLaunchKernelGGL(k_norms, dim3(c->vec_grid), dim3(NFT_BLOCK), 0, c->stream, c->n, a, (double *)nullptr,
                     c->scal, sum_slot, max_slot, c->partials, c->counter);
  HIPCHK(hipGetLastError());
  FinalOp op{};
  op.kind = FK_NORMS; op.K = 2; op.ismax = 2u; op.sum_slot = sum_slot; op.max_slot = max_slot;
  return final_launch(c, c->vec_grid, op);
}

// One GCR inner iteration (reference :1886-1936) as launched by both loops: update with
// its reductions, preconditioner, laplc of the preconditioned residual with the fused
// <ax_ii, ar>, new direction with <r, ax>, <ax, ax>. `ctl` is the finalize of the update
// (device exit test when ctl.ctl is set).
// fold_slot >= 0: the previous direction's finalize is folded into this update (small grids);
// exit_fold: the exit test is left to the band solve that follows (c->xfold)
static int gcr_iteration_update(nft_ctx *c, double *p, int i, const FinalOp &ctl, int fold_slot,
                                bool exit_fold = false) {
  UpdFold fo{};
  if (fold_slot >= 0) {
    fo.part = c->partials2; fo.nblk = c->vec_grid; fo.slot = fold_slot; fo.epa = ctl.epa;
  }
  t_begin(c, KT_UPD);
  if (fo.part)
    hipLaunchKernelGGL(k_gcr_update<true>, dim3(c->vec_grid), dim3(NFT_BLOCK), 0, c->stream, c->n, p, c->r,
                       c->gx[i - 1], c->gax[i - 1], c->scal, c->partials, c->skip, fo);
  else
    hipLaunchKernelGGL(k_gcr_update<false>, dim3(c->vec_grid), dim3(NFT_BLOCK), 0, c->stream, c->n, p, c->r,
                       c->gx[i - 1], c->gax[i - 1], c->scal, c->partials, c->skip, fo);
  t_end(c, KT_UPD);
  HIPCHK(hipGetLastError());
  if (exit_fold) return 0;
  return final_launch(c, c->vec_grid, ctl);
}

// With c->fold (grids of at most NFT_FOLD_BLOCKS blocks) the finalize of the divergence's
// dots runs inside the direction kernel and the direction's inside the next update, so an
// iteration launches two grid-level sums fewer; the sums and their order are unchanged.
static int gcr_iteration_rest(nft_ctx *c, int i, int impfl, double epa) {
  const int lord = c->P.lord;
  const bool fold = c->fold != 0;
  int rc;
  if ((rc = precon_dev(c, c->r, c->qr, impfl))) return rc;
  {
    DotList dl{};
    dl.nd = i;
    for (int ii = 0; ii < i; ++ii) dl.a[ii] = c->gax[ii];
    dl.fin = FIN_DEL;
    dl.fold = fold ? 1 : 0;
    if ((rc = krylov_laplc(c, c->qr, c->ar, impfl, &dl))) return rc;
  }
  DirArgs d{};
  d.n = c->n; d.nx = i;
  int slot = (i < lord) ? i : 0;
  d.xn = c->gx[slot]; d.axn = c->gax[slot]; d.slot = slot;
  for (int ii = 0; ii < i; ++ii) { d.x[ii] = c->gx[ii]; d.ax[ii] = c->gax[ii]; }
  d.qr = c->qr; d.ar = c->ar; d.r = c->r;
  d.scal = c->scal; d.partials = fold ? c->partials2 : c->partials; d.counter = c->counter; d.epa = epa;
  d.skip = c->skip;
  if (fold) { d.fold_part = c->partials; d.fold_nblk = c->gather_grid; }
  t_begin(c, KT_DIR);
  if (fold) hipLaunchKernelGGL(k_gcr_dir<true>, dim3(c->vec_grid), dim3(NFT_BLOCK), 0, c->stream, d);
  else hipLaunchKernelGGL(k_gcr_dir<false>, dim3(c->vec_grid), dim3(NFT_BLOCK), 0, c->stream, d);
  t_end(c, KT_DIR);
  HIPCHK(hipGetLastError());
  if (fold) return 0;   // the next update folds this finalize in
  FinalOp op{};
  op.kind = FK_DIR; op.K = 2; op.ismax = 0u; op.slot = slot; op.epa = epa;
  return final_launch(c, c->vec_grid, op);
}

// The inner loop with the exit test on the device: every iteration is enqueued without a
// host round trip; the host reads the mapped exit flag one iteration behind (the GPU
// always has the next iteration queued) and stops enqueueing once it is set. Iterations
// enqueued past the exit return at once. Same arithmetic and exit rule as the host loop.
static int gcr_loop_device(nft_ctx *c, double *p, int impfl, double eps, int itmn, double rl20, int *nlc,
                           int *converged, double *dvmx) {
  const nft_params &P = c->P;
  const int lord = P.lord;
  const double epa = 1.e-38;
  int rc = 0;
  *c->h_flag = 0.0;
  hipLaunchKernelGGL(k_gcr_ctl_begin, dim3(1), dim3(64), 0, c->stream, c->scal, rl20);
  HIPCHK(hipGetLastError());
  c->skip = c->scal + SC_DONE;
  hipEvent_t ev[2] = {c->poll_ev[0], c->poll_ev[1]};
  // timer events in use before each enqueued iteration, counted from the loop's start
  std::vector<long> ev_mark;
  long ev_off = 0;   // events already folded into the totals
  int enq = 0;
  bool stop = false;
  int prev_slot = -1;   // slot of the direction whose finalize the next update folds in
  // small grids with the band line (one sweep, no multigrid): the exit test runs inside the
  // band solve of the same iteration; the previous sum r^2 alternates between two slots
  const bool xfold = c->fold && P.precon == 3 && !c->mg_on;
  for (int it = 1; it <= P.itr && !stop && !rc; ++it) {
    for (int i = 1; i <= lord && !stop; ++i) {
      ev_mark.push_back(ev_off + c->tev_used);
      FinalOp op{};
      op.kind = FK_UPDATE; op.K = 2; op.ismax = 1u;
      op.ctl = 1; op.it = it; op.itmn = itmn; op.eps = eps; op.epa = epa; op.hflag = c->d_hflag;
      if ((rc = gcr_iteration_update(c, p, i, op, c->fold ? prev_slot : -1, xfold))) break;
      if (xfold) {
        c->xfold = ExitFold{c->partials, c->vec_grid, it, itmn, enq & 1, eps, epa, c->scal, c->d_hflag};
      }
      rc = gcr_iteration_rest(c, i, impfl, epa);
      c->xfold = ExitFold{};
      if (rc) break;
      prev_slot = i < lord ? i : 0;
      if (hipEventRecord(ev[enq & 1], c->stream) != hipSuccess) { rc = set_err(-2, "hipEventRecord"); break; }
      ++enq;
      if (enq > 1) {
        if (hipEventSynchronize(ev[(enq - 2) & 1]) != hipSuccess) { rc = set_err(-2, "hipEventSynchronize"); break; }
        if (*(volatile double *)c->h_flag != 0.0) {
          stop = true;
        } else if (c->timer_kernel && c->tev_used > (int)c->tev.size() / 2) {
          // iterations before enq-1 all passed and have finished: their timings are final
          long upto = ev_mark[enq - 1] - ev_off;
          t_harvest_prefix(c, (int)upto);
          ev_off += upto;
        }
      }
    }
  }
  c->skip = nullptr;
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(c->h_scal + SC_DVMX, c->scal + SC_DVMX, sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(c->h_scal + SC_DONE, c->scal + SC_DONE, 3 * sizeof(double), hipMemcpyDeviceToHost));
  *nlc = (int)c->h_scal[SC_NLC];
  *converged = c->h_scal[SC_DONE] == 1.0;
  *dvmx = c->h_scal[SC_DVMX];
  // launches past the exit did no work: drop their timer events before they are read
  if (*nlc < enq) {
    long keep = ev_mark[*nlc] - ev_off;
    if (c->timer_kernel == KT_UPD) keep += 2;   // the update of the deciding iteration ran
    if (keep >= 0 && keep < c->tev_used) c->tev_used = (int)keep;
  }
  t_harvest(c);
  return 0;
}

// The host-controlled inner loop (learned correction, pair dumps): the exit test after
// every update is taken on the host, which may drop the learned correction.
static int gcr_loop_host(nft_ctx *c, double *p, int impfl, double eps, int itmn, double rl20, int *nlc,
                         int *converged, double *dvmx, const double **qr_last) {
  const nft_params &P = c->P;
  const int lord = P.lord;
  const double epa = 1.e-38;
  int rc;
  for (int it = 1; it <= P.itr; ++it) {
    for (int i = 1; i <= lord; ++i) {
      t_begin(c, KT_UPD);
      hipLaunchKernelGGL(k_gcr_update<false>, dim3(c->vec_grid), dim3(NFT_BLOCK), 0, c->stream, c->n, p, c->r,
                         c->gx[i - 1], c->gax[i - 1], c->scal, c->partials, (const double *)nullptr, UpdFold{});
      t_end(c, KT_UPD);
      {
        FinalOp op{};
        op.kind = FK_UPDATE; op.K = 2; op.ismax = 1u;
        if ((rc = final_launch(c, c->vec_grid, op))) return rc;
      }
      HIPCHK(hipGetLastError());
      if ((rc = read_scal(c, SC_DVMX, 2))) return rc;
      *dvmx = c->h_scal[SC_DVMX];
      double rl2 = c->h_scal[SC_RL2];
      if (std::isnan(rl2)) return 0;   // NaN residual: stop, reported as not converged
      if (*dvmx <= eps && it >= itmn) { *converged = 1; return 0; }
      if (rl2 >= rl20) {
        // stagnation: with the learned correction active, drop it and continue with the
        // line preconditioner alone instead of ending the solve
        if (!c->net.active) return 0;
        c->net.active = 0;
        c->net.fallbacks++;
      } else if (c->net.active && rl2 > c->net.stall * rl20) {
        c->net.active = 0;
        c->net.fallbacks++;
      }
      rl20 = std::max(rl2, epa);
      if ((rc = precon_dev(c, c->r, c->qr, impfl))) return rc;
      if (c->net.dumping && (rc = net_store_pair(c, *dvmx, rl2, impfl, p))) return rc;
      if (c->net.active) {
        if ((rc = net_apply_precon(c, c->r, c->qr, SC_DVMX))) return rc;
        c->net.learned_iters++;
      }
      *qr_last = c->qr;
      {
        DotList dl{};
        dl.nd = i;
        for (int ii = 0; ii < i; ++ii) dl.a[ii] = c->gax[ii];
        dl.fin = FIN_DEL;
        if ((rc = krylov_laplc(c, c->qr, c->ar, impfl, &dl))) return rc;
      }
      *nlc = *nlc + 1;
      {
        DirArgs d{};
        d.n = c->n; d.nx = i;
        int slot = (i < lord) ? i : 0;
        d.xn = c->gx[slot]; d.axn = c->gax[slot]; d.slot = slot;
        for (int ii = 0; ii < i; ++ii) { d.x[ii] = c->gx[ii]; d.ax[ii] = c->gax[ii]; }
        d.qr = c->qr; d.ar = c->ar; d.r = c->r;
        d.scal = c->scal; d.partials = c->partials; d.counter = c->counter; d.epa = epa;
        t_begin(c, KT_DIR);
        hipLaunchKernelGGL(k_gcr_dir<false>, dim3(c->vec_grid), dim3(NFT_BLOCK), 0, c->stream, d);
        t_end(c, KT_DIR);
        HIPCHK(hipGetLastError());
        FinalOp op{};
        op.kind = FK_DIR; op.K = 2; op.ismax = 0u; op.slot = slot; op.epa = epa;
        if ((rc = final_launch(c, c->vec_grid, op))) return rc;
      }
    }
  }
  return 0;
}

// gcrk (reference :1776-1972) on device arrays. Returns 0 when converged, 1 otherwise.
//
// With the FP32 operator (precision 2) the iterations apply laplc with FP32 metrics and
// temporaries while the residual, its updates, the dots and the orthogonalisation stay
// FP64. The recursive residual then drifts from the true one by the operator's rounding,
// so a solve runs in passes: each pass starts from the true FP64 residual of the current
// pressure (prfr + div, as the reference's start) and iterates until the recursive
// residual has dropped by op32_drop (nft_set_mixed_drop) or passed the tolerance; the
// next pass checks the tolerance on the true residual and stops there, or restarts GCR
// from it.
#ifndef NFT_OP32_PASSES
#define NFT_OP32_PASSES 24
#endif
int gcrk_dev(nft_ctx *c, double *p, const double *u1, const double *v1, const double *w1, double *u2,
             double *v2, double *w2, int impfl) {
  const nft_params &P = c->P;
  const double epa = 1.e-38;
  const bool op32 = P.precision == 2;
  int rc;
  double eps = P.eps0 / P.dt;
  if (P.exit_mode == 1) {
    // b = residual of the zero initial guess
    if ((rc = prfr_dev(c, nullptr, u1, v1, w1, c->t[4], c->t[5], c->t[6], impfl))) return rc;
    if ((rc = div_dev(c, c->t[4], c->t[5], c->t[6], c->t[7], -1, nullptr))) return rc;
    if ((rc = norms_dev(c, c->t[7], SC_TMP, SC_BMX))) return rc;
  }
  double *x0 = c->gx[0], *ax0 = c->gax[0];
  const double *qr_last = x0;
  int nlc = 0, converged = 0;
  double dvmx = 0, eer0 = epa, eem0 = epa;
  const int pair0 = c->net.npairs;   // pairs stored from here on belong to this solve
  c->net.active = c->net.kind == 1 && c->net.ready;
  for (int pass = 0;; ++pass) {
    // u2, v2, w2 are prfr of the incoming pressure (outputs of gcrk); later passes use
    // scratch arrays for theirs
    double *pu = pass ? c->t[4] : u2, *pv = pass ? c->t[5] : v2, *pw = pass ? c->t[6] : w2;
    if ((rc = prfr_dev(c, p, u1, v1, w1, pu, pv, pw, impfl))) return rc;
    {
      DotList dl{};
      dl.nd = 1; dl.a[0] = nullptr; dl.nmax = 1; dl.fin = FIN_NORMS;
      if ((rc = div_dev(c, pu, pv, pw, c->r, -1, &dl))) return rc;
    }
    if (pass > 0) {
      // the true residual of the pressure the previous pass reached
      if ((rc = read_scal(c, SC_R2, 2))) return rc;
      dvmx = c->h_scal[SC_RMX];
      if (dvmx <= eps) { converged = 1; break; }
      if (!(dvmx == dvmx) || pass >= NFT_OP32_PASSES) break;
    }
    if ((rc = precon_dev(c, c->r, x0, impfl))) return rc;
    if (c->net.active && (rc = net_apply_precon(c, c->r, x0, SC_RMX))) return rc;
    if ((rc = norms_dev(c, x0, SC_EER, SC_EEM))) return rc;
    {
      DotList dl{};
      dl.nd = 2; dl.a[0] = c->r; dl.a[1] = nullptr; dl.fin = FIN_BETA; dl.slot = 0;
      if ((rc = krylov_laplc(c, x0, ax0, impfl, &dl))) return rc;
    }
    if ((rc = read_scal(c, 0, SC_TOTAL))) return rc;
    if (pass == 0) {
      if (P.exit_mode == 1) eps = P.eps0 * c->h_scal[SC_BMX];
      eer0 = std::max(c->h_scal[SC_EER], epa);
      eem0 = std::max(c->h_scal[SC_EEM], epa);
    }
    const double rl20 = std::max(c->h_scal[SC_R2], epa);
    // a pass of the FP32 operator ends early, at a fixed drop of the recursive residual
    const double eps_pass = op32 ? std::max(eps, c->op32_drop * c->h_scal[SC_RMX]) : eps;
    const int itmn = pass == 0 ? 3 : 1;
    int nlc_pass = 0, conv_pass = 0;
    qr_last = x0;
    if (!c->net.active && !c->net.dumping) {
      if ((rc = gcr_loop_device(c, p, impfl, eps_pass, itmn, rl20, &nlc_pass, &conv_pass, &dvmx))) return rc;
      if (nlc_pass > 0) qr_last = c->qr;
    } else {
      if ((rc = gcr_loop_host(c, p, impfl, eps_pass, itmn, rl20, &nlc_pass, &conv_pass, &dvmx, &qr_last))) return rc;
    }
    nlc += nlc_pass;
    if (!op32) { converged = conv_pass; break; }
    if (!conv_pass) break;   // stagnation, NaN or the iteration cap: not converged
  }
  if (c->net.npairs > pair0) {
    // error targets of this solve's pairs: x_k - x_final; a solve that did not converge
    // gives no targets, and its pairs are dropped
    auto &n = c->net;
    if (converged) {
      for (int k = pair0; k < n.npairs; ++k) {
        double *d = n.pD + (size_t)c->n * k;
        hipLaunchKernelGGL(k_sub, dim3(c->nodes_grid), dim3(NFT_BLOCK), 0, c->stream, c->n, d, p, d,
                           (const double *)nullptr);
      }
      HIPCHK(hipGetLastError());
    } else {
      n.npairs = pair0;
      n.ps.resize(pair0); n.pr2.resize(pair0); n.pimp.resize(pair0);
    }
  }
  if ((rc = norms_dev(c, qr_last, SC_EER, SC_EEM))) return rc;
  if ((rc = read_scal(c, SC_EER, 2))) return rc;
  c->st.eer = c->h_scal[SC_EER] / eer0;
  c->st.eem = c->h_scal[SC_EEM] / eem0;
  c->st.niter = nlc;
  c->st.nitsm += nlc;
  c->st.icount += 1;
  c->st.converged = converged;
  c->st.res_max = dvmx;
  c->st.tol = eps;
  return converged ? 0 : 1;
}

// LDS bytes of the patch sweeps: the staged columns' donor-cell values (bounds) or (cp, cn)
// pairs (limiter)
template <typename T, int NF>
size_t patch_lds_bytes(const nft_ctx *c, bool limit) {
  return (size_t)c->patch_smax * NFT_WAVE * NF * sizeof(T) * (limit ? 2 : 1);
}
// the donor-cell sweep's: advector and fields
template <typename T, int NF>
size_t patch_lds_bytes_donor(const nft_ctx *c) {
  return (size_t)c->patch_smax * NFT_WAVE * (3 + NF) * sizeof(T);
}

template <typename T, int NF>
int mpdata_dev(nft_ctx *c, const T *U, const T *V, const T *W, double *const *fields, const int *iflags) {
  MpArgs<T, NF> a{};
  a.U = U; a.V = V; a.W = W;
  for (int f = 0; f < NF; ++f) {
    a.psi[f] = fields[f];
    a.psi1[f] = c->mp_psi1[f];
    a.gxy[f] = (typename MpArgs<T, NF>::T2 *)c->mp_g[3 * f];   // 2n values
    a.gz[f] = (T *)c->mp_g[3 * f + 2];
    a.cpn[f] = (typename MpArgs<T, NF>::T2 *)c->mp_c[2 * f];   // 2n values
    if (f % 2 == 0)
      a.aflux[f / 2] = (typename MpArgs<T, NF>::T2 *)c->mp_flux + (size_t)(f / 2) * c->etot;
    int fl = iflags[f];
    double afree[5] = {c->P.u0, c->P.v0, c->P.w0, 0., 1.};
    a.use_ue[f] = fl == 1;
    a.afree[f] = (T)afree[fl - 1];
  }
  if (sizeof(T) == 8) { a.vol = (const T *)c->vol; a.rh0 = (const T *)c->rh0; a.vr = (const T *)c->vr; }
  else { a.vol = (const T *)c->volf; a.rh0 = (const T *)c->rh0f; a.vr = (const T *)c->vrf; }
  a.ue = c->ue;
  a.dt = (T)c->P.dt;
  EdgeS<T> S;
  if (sizeof(T) == 8) S = *(EdgeS<T> *)&c->S;
  else S = *(EdgeS<T> *)&c->Sf;
  dim3 g(c->gather_grid), b(NFT_BLOCK);
  // periodic box (mpdatamper): each sweep's set kernel first
  const dim3 gs(blocks_for(c->ps.nset));
  const bool per = c->per != 0;
  if constexpr (sizeof(T) == 8) {
    if (per) hipLaunchKernelGGL(k_per_mp_donor<NF>, gs, b, 0, c->stream, c->inc, c->ps, a);
  }
  // columns of 64 levels in a closed box: the gathering sweeps over LDS-staged column patches
  const bool lds = c->patch_on && c->col_on && !per;
  const dim3 gp(c->patches.xcd_per > 0 ? 8 * c->patches.xcd_per : c->patches.npatch), bp(NFT_WAVE * NFT_PATCH_COLS);
  const size_t shd = patch_lds_bytes_donor<T, NF>(c);
  const size_t shf = patch_lds_bytes<T, NF>(c, false), shl = patch_lds_bytes<T, NF>(c, true);
  t_begin(c, KT_MP0);
  if (lds && NFT_MP_DONOR_LDS)
    hipLaunchKernelGGL((k_mp_donor_lds<T, NF>), gp, bp, shd, c->stream, c->inc_lds, S, c->patches, a);
  else if (c->col_on) hipLaunchKernelGGL((k_mp_donor<T, NF, true>), g, b, 0, c->stream, c->inc, S, a);
  else hipLaunchKernelGGL((k_mp_donor<T, NF, false>), g, b, 0, c->stream, c->inc, S, a);
  t_end(c, KT_MP0);
  // pass 2 reads the donor-cell fields
  if constexpr (sizeof(T) == 8) {
    if (per) hipLaunchKernelGGL(k_per_mp_grad<NF>, gs, b, 0, c->stream, c->inc, c->ps, a);
  }
  t_begin(c, KT_MP0 + 1);
  if (lds && NFT_MP_GRAD_LDS)
    hipLaunchKernelGGL((k_mp_grad_lds<T, NF>), gp, bp, shf, c->stream, c->inc_lds, S, c->patches, a);
  else if (c->col_on) hipLaunchKernelGGL((k_mp_grad<T, NF, true>), g, b, 0, c->stream, c->inc, S, a);
  else hipLaunchKernelGGL((k_mp_grad<T, NF, false>), g, b, 0, c->stream, c->inc, S, a);
  t_end(c, KT_MP0 + 1);
  if constexpr (sizeof(T) == 8) {
    if (per) hipLaunchKernelGGL(k_per_mp_fct<NF>, gs, b, 0, c->stream, c->inc, c->ps, a);
  }
  // every edge's antidiffusive velocity once, over the owned incidence (same chunks and grid)
  Incidence go = c->inc_own;
  go.sxy = c->inc.sxy; go.szz = c->inc.szz;
  t_begin(c, KT_MPF);
  if (c->col_on && go.col_L > 0) hipLaunchKernelGGL((k_mp_flux<T, NF, true>), g, b, 0, c->stream, go, S, a);
  else hipLaunchKernelGGL((k_mp_flux<T, NF, false>), g, b, 0, c->stream, go, S, a);
  t_end(c, KT_MPF);
  t_begin(c, KT_MP0 + 2);
  if (lds)
    hipLaunchKernelGGL((k_mp_fct_lds<T, NF>), gp, bp, shf, c->stream, c->inc_lds,
                       c->patches, a);
  else if (c->col_on) hipLaunchKernelGGL((k_mp_fct<T, NF, true>), g, b, 0, c->stream, c->inc, a);
  else hipLaunchKernelGGL((k_mp_fct<T, NF, false>), g, b, 0, c->stream, c->inc, a);
  t_end(c, KT_MP0 + 2);
  double *const *out = (double *const *)c->io[7];
  if constexpr (sizeof(T) == 8) {
    if (per) hipLaunchKernelGGL(k_per_mp_limit<NF>, gs, b, 0, c->stream, c->inc, c->ps, a);
  }
  t_begin(c, KT_MP0 + 3);
  if (lds)
    hipLaunchKernelGGL((k_mp_limit_lds<T, NF>), gp, bp, shl, c->stream, c->inc_lds,
                       c->patches, a, out);
  else if (c->col_on) hipLaunchKernelGGL((k_mp_limit<T, NF, true>), g, b, 0, c->stream, c->inc, a, out);
  else hipLaunchKernelGGL((k_mp_limit<T, NF, false>), g, b, 0, c->stream, c->inc, a, out);
  t_end(c, KT_MP0 + 3);
  HIPCHK(hipGetLastError());
  return 0;
}

StepArgs step_args(nft_ctx *c) {
  StepArgs a{};
  a.n = c->n;
  for (int k = 0; k < 4; ++k) { a.unk[k] = c->unk[k]; a.frc[k] = c->frc[k]; }
  for (int k = 0; k < 3; ++k) { a.uuu[k] = c->uuu[k]; a.uuu0[k] = c->uuu0[k]; a.ud[k] = c->ud[k]; }
  a.alpha = c->alpha; a.ue = c->ue; a.ve = c->ve; a.th0 = c->th0; a.coefw = c->coefw;
  a.rh0 = c->rh0; a.dthedz = c->dthedz;
  a.dt = c->P.dt; a.g = kG;
  a.itraj = c->P.itraj;
  for (int k = 0; k < 3; ++k) a.uuuf[k] = c->P.precision >= 1 ? c->uuuf[k] : nullptr;
  return a;
}

// velprd(unk,frc,uuu,uuu1,unkdash) (:4448-4531, Cartesian): predicts u at n+1 by advecting
// unk + dt*frc with uuu (gradient and advection fused in k_grad<GRAD_ADV>), projects it
// with a GCR solve (impfl = 0) and prfr, and sets uuu = (uuu + prediction)*0.5*rh0,
// unkdash = prediction. uuu1 lives in uuu0 (unused with itraj = 1).
int velprd_dev(nft_ctx *c, nft_run_stats *rs) {
  StepArgs a = step_args(c);
  dim3 g(c->nodes_grid), b(NFT_BLOCK);
  int rc;
  hipLaunchKernelGGL(k_vp_start, g, b, 0, c->stream, a);
  for (int k = 0; k < 3; ++k) {
    GradArgs ga = grad_args(c);
    ga.f = c->ud[k]; ga.s0 = c->uuu[0]; ga.s1 = c->uuu[1]; ga.s2 = c->uuu[2]; ga.o0 = c->uuu0[k];
    if ((rc = grad_launch(c, GRAD_ADV, ga))) return rc;
  }
  hipLaunchKernelGGL(k_fill, g, b, 0, c->stream, c->n, c->vp_p, 0.0);
  float ms = 0;
  HIPCHK(hipEventRecord(c->ev[0], c->stream));
  rc = gcrk_dev(c, c->vp_p, c->uuu0[0], c->uuu0[1], c->uuu0[2], c->vp_ff[0], c->vp_ff[1], c->vp_ff[2], 0);
  if (rc < 0) return rc;
  if (rs) { rs->gcr_solves++; rs->gcr_iters += c->st.niter; rs->gcr_unconverged += rc; }
  if ((rc = prfr_dev(c, c->vp_p, c->uuu0[0], c->uuu0[1], c->uuu0[2], c->vp_ff[0], c->vp_ff[1], c->vp_ff[2], 0)))
    return rc;
  HIPCHK(hipEventRecord(c->ev[1], c->stream));
  if (rs) {
    HIPCHK(hipEventSynchronize(c->ev[1]));
    t_harvest(c);
    HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    rs->ms_gcr += ms;
  }
  hipLaunchKernelGGL(k_vp_end, g, b, 0, c->stream, a, (const double *const *)c->vp_ffd);
  HIPCHK(hipGetLastError());
  return 0;
}

int set_out_fields(nft_ctx *c, double *const *fields, int nf) {
  HIPCHK(hipMemcpyAsync(c->io[7], fields, nf * sizeof(double *), hipMemcpyHostToDevice, c->stream));
  return 0;
}

// fields per launch of the step's MPDATA sweeps (4: u, v, w, theta' in one launch each;
// 2 and 1 re-read the geometry per group and measured slower on the production mesh:
// 4.7 and 6.3 against 3.7 ms per step, results identical)
#ifndef NFT_MP_FIELDS
#define NFT_MP_FIELDS 4
#endif

int mpdata_step(nft_ctx *c) {
  int fl[4] = {1, 2, 3, 4};
  constexpr int G = NFT_MP_FIELDS;
  static_assert(G == 1 || G == 2 || G == 4, "fields per MPDATA launch: 1, 2 or 4");
  for (int f0 = 0; f0 < 4; f0 += G) {
    int rc = set_out_fields(c, c->unk + f0, G);
    if (!rc)
      rc = c->P.precision >= 1 ? mpdata_dev<float, G>(c, c->uuuf[0], c->uuuf[1], c->uuuf[2], c->unk + f0, fl + f0)
                               : mpdata_dev<double, G>(c, c->uuu[0], c->uuu[1], c->uuu[2], c->unk + f0, fl + f0);
    if (rc) return rc;
  }
  return 0;
}

// ---- diagnostics ----------------------------------------------------------------------
constexpr int kDiagLen = 25;   // div 4, cour 4, crnic, u 4, v 4, w 4, t 4

template <int NF>
int diag_stats(nft_ctx *c, const DiagSpec &d) {
  hipLaunchKernelGGL(k_diag_mom<NF>, dim3(c->vec_grid), dim3(NFT_BLOCK), 0, c->stream, c->n, d, c->partials);
  FinalOp op{};
  op.kind = FK_DIAG_MOM; op.K = 3 * NF + 1; op.diag = d;
  for (int f = 0; f < NF; ++f) op.ismax |= (1u << (3 * f)) | (1u << (3 * f + 1));
  int rc = final_launch(c, c->vec_grid, op);
  if (rc) return rc;
  hipLaunchKernelGGL(k_diag_var<NF>, dim3(c->vec_grid), dim3(NFT_BLOCK), 0, c->stream, c->n, d, c->partials);
  FinalOp op2{};
  op2.kind = FK_DIAG_VAR; op2.K = NF + 1; op2.ismax = 1u << NF; op2.diag = d;
  return final_launch(c, c->vec_grid, op2);
}

// the reference's per-step diagnostics into rec[kDiagLen] (device), of the advector uuu
// and the fields unk given as device arrays; parts: 1 div, 2 courant (numbers in t[5]),
// 4 stats (sums of these)
int diag_dev_of(nft_ctx *c, int flag, double *rec, double *const *uuu, double *const *unk, int parts) {
  int rc;
  if (parts & 1) {
    if ((rc = div_dev(c, uuu[0], uuu[1], uuu[2], c->t[4], 1, nullptr))) return rc;
    DiagSpec d{};
    d.nf = 1; d.x[0] = c->t[4]; d.scale = c->P.dt; d.n = (double)c->n;
    d.init_mx = -1.e15; d.init_mn = 1.e15; d.out = rec;
    if ((rc = diag_stats<1>(c, d))) return rc;
  }
  if (parts & 2) {
    dim3 g(c->gather_grid), b(NFT_BLOCK);
#define NFT_COUR(F, COL)                                                                                        \
    hipLaunchKernelGGL((k_courant<F, COL>), g, b, 0, c->stream, c->inc, c->S, uuu[0], uuu[1], uuu[2], c->nc, \
                       c->P.dt, c->t[5])
    if (c->col_on) {
      if (flag == 2) NFT_COUR(2, true); else if (flag == 3) NFT_COUR(3, true); else NFT_COUR(1, true);
    } else {
      if (flag == 2) NFT_COUR(2, false); else if (flag == 3) NFT_COUR(3, false); else NFT_COUR(1, false);
    }
#undef NFT_COUR
    HIPCHK(hipGetLastError());
    DiagSpec cd{};
    cd.nf = 1; cd.x[0] = c->t[5]; cd.scale = 1.0; cd.n = (double)c->n;
    cd.init_mx = -1.e9; cd.init_mn = 1.e9; cd.argmax = 1; cd.out = rec + 4;
    if ((rc = diag_stats<1>(c, cd))) return rc;
  }
  if (parts & 4) {
    DiagSpec sd{};
    sd.nf = 4; sd.w = c->rh0; sd.scale = 1.0; sd.n = (double)c->n;
    sd.init_mx = -1.e10; sd.init_mn = 1.e10; sd.out = rec + 9;
    for (int f = 0; f < 4; ++f) { sd.x[f] = unk[f]; sd.mean_by_w[f] = 1; sd.sd_by_w[f] = f < 3 ? 1 : 0; }
    if ((rc = diag_stats<4>(c, sd))) return rc;
  }
  return 0;
}

// the three diagnostics of the device-resident state
int diag_dev(nft_ctx *c, int flag, double *rec) { return diag_dev_of(c, flag, rec, c->uuu, c->unk, 7); }

void diag_unpack(const double *r, nft_diag &d) {
  for (int k = 0; k < 4; ++k) { d.div[k] = r[k]; d.cour[k] = r[4 + k]; }
  d.crnic = (int)r[8];
  for (int k = 0; k < 4; ++k) { d.u[k] = r[9 + k]; d.v[k] = r[13 + k]; d.w[k] = r[17 + k]; d.t[k] = r[21 + k]; }
}

void diag_meta_fill(nft_ctx *c, nft_diag &d) {
  d.eer = c->st.eer; d.eem = c->st.eem; d.niter = c->st.niter;
  d.nitav = c->st.nitsm / std::max(c->st.icount, 1);
  d.step = c->steps_done;
}

int one_step(nft_ctx *c, nft_run_stats *rs) {
  StepArgs a = step_args(c);
  dim3 g(c->nodes_grid), b(NFT_BLOCK);
  int rc;
  hipLaunchKernelGGL(k_step_pre, g, b, 0, c->stream, a);
  HIPCHK(hipEventRecord(c->ev[2], c->stream));
  if ((rc = mpdata_step(c))) return rc;
  HIPCHK(hipEventRecord(c->ev[3], c->stream));
  float ms = 0;
  for (int mtri = 1; mtri <= 2; ++mtri) {
    hipLaunchKernelGGL(k_step_rhs, g, b, 0, c->stream, a);
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    rc = gcrk_dev(c, c->prpot, c->frc[0], c->frc[1], c->frc[2], c->uuu[0], c->uuu[1], c->uuu[2], 1);
    if (rc < 0) return rc;
    if (rs) { rs->gcr_solves++; rs->gcr_iters += c->st.niter; rs->gcr_unconverged += rc; }
    if ((rc = prfr_dev(c, c->prpot, c->frc[0], c->frc[1], c->frc[2], c->uuu[0], c->uuu[1], c->uuu[2], 1)))
      return rc;
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    HIPCHK(hipEventSynchronize(c->ev[1]));
    t_harvest(c);
    if (rs) { HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1])); rs->ms_gcr += ms; }
    hipLaunchKernelGGL(k_step_post, g, b, 0, c->stream, a, mtri == 2 ? 1 : 0);
  }
  hipLaunchKernelGGL(k_step_end, g, b, 0, c->stream, a);
  HIPCHK(hipGetLastError());
  if (rs) { HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3])); rs->ms_mpdata += ms; }
  if (c->P.itraj == 1) return velprd_dev(c, rs);
  return 0;
}

int copy_in(nft_ctx *c, double *dst, const double *src) {
  HIPCHK(hipMemcpyAsync(dst, src, c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return 0;
}
int copy_out(nft_ctx *c, double *dst, const double *src) {
  HIPCHK(hipMemcpyAsync(dst, src, c->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return 0;
}
int sync(nft_ctx *c) {
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// precon 3: the in-column band of laplc(., impfl=1), probed column-colour by colour.
// laplc couples nodes at most two edges apart, so with probe nodes on columns pairwise at
// least three column hops apart (greedy colouring in column order) and on levels five
// apart, entry (row, probe) of laplc(indicator) is the only nonzero term of each row within
// two levels of a probe. Each column's band is factorised (LU, no pivoting, band width 2)
// on the host in the oracle's operation order, and the factors are stored interleaved for
// k_line_penta.
int build_band(nft_ctx *c, const nft_mesh *m) {
  const int n = c->n, nc = m->ncol;
  if (nc <= 0 || n % nc) return set_err(-1, "nft_init: precon 3 needs columns of one length");
  const int L = n / nc;
  for (int k = 0; k <= nc; ++k)
    if (m->col_start[k] != k * L) return set_err(-1, "nft_init: precon 3 needs columns of one length");
  int cpb = L <= 96 ? 64 : L <= 200 ? 32 : L <= 400 ? 16 : 0;
  if (!cpb) return set_err(-1, "nft_init: precon 3 supports columns of at most 400 levels");
  // few columns: fewer per block, so that the one-wavefront blocks spread over the CUs
  while (cpb > 1 && (nc + cpb - 1) / cpb < NFT_PENTA_MINBLK) cpb /= 2;
  // column adjacency (interior edges, then boundary-face edges)
  std::vector<int32_t> deg(nc + 1, 0);
  for (int pass = 0; pass < 2; ++pass) {
    const int ne = pass == 0 ? m->nedge : m->nbface, ld = pass == 0 ? m->ld_edge : m->ld_face;
    const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
    for (int e = 0; e < ne; ++e) {
      int a = (ie[e] - 1) / L, b = (ie[ld + e] - 1) / L;
      if (a != b) { deg[a + 1]++; deg[b + 1]++; }
    }
  }
  for (int k = 0; k < nc; ++k) deg[k + 1] += deg[k];
  std::vector<int32_t> adj(deg[nc] + 1), fill(nc, 0);
  for (int pass = 0; pass < 2; ++pass) {
    const int ne = pass == 0 ? m->nedge : m->nbface, ld = pass == 0 ? m->ld_edge : m->ld_face;
    const int32_t *ie = pass == 0 ? m->iedge : m->kbface;
    for (int e = 0; e < ne; ++e) {
      int a = (ie[e] - 1) / L, b = (ie[ld + e] - 1) / L;
      if (a != b) { adj[deg[a] + fill[a]++] = b; adj[deg[b] + fill[b]++] = a; }
    }
  }
  // unique neighbours per column (each column pair is joined by many edges)
  {
    std::vector<int32_t> uptr(nc + 1, 0), uadj;
    uadj.reserve(adj.size() / 8 + 1);
    for (int k = 0; k < nc; ++k) {
      std::vector<int32_t> row(adj.begin() + deg[k], adj.begin() + deg[k + 1]);
      std::sort(row.begin(), row.end());
      row.erase(std::unique(row.begin(), row.end()), row.end());
      uadj.insert(uadj.end(), row.begin(), row.end());
      uptr[k + 1] = (int32_t)uadj.size();
    }
    deg.swap(uptr);
    adj.swap(uadj);
  }
  // periodic box: twin columns (those holding the nodes of one coincident set)
  std::vector<int32_t> tptr, tadj;
  if (c->per) {
    std::vector<std::vector<int32_t>> tw(nc);
    for (const int4 &q : c->h_sets) {
      const int cnt = q.z >= 0 ? 4 : 2;
      const int mem[4] = {q.x, q.y, q.z, q.w};
      for (int a = 0; a < cnt; ++a)
        for (int b = 0; b < cnt; ++b)
          if (mem[a] / L != mem[b] / L) tw[mem[a] / L].push_back(mem[b] / L);
    }
    tptr.assign(nc + 1, 0);
    for (int k = 0; k < nc; ++k) {
      std::sort(tw[k].begin(), tw[k].end());
      tw[k].erase(std::unique(tw[k].begin(), tw[k].end()), tw[k].end());
      tadj.insert(tadj.end(), tw[k].begin(), tw[k].end());
      tptr[k + 1] = (int32_t)tadj.size();
    }
  }
  // greedy colouring: a column takes the smallest colour absent among the columns its
  // probe reaches, within two column hops; in the periodic box the closures join twin
  // columns as well, so the reach is twin, hop, twin, hop, twin from the column
  std::vector<int> color(nc, -1), mark(4096, 0);
  std::vector<int32_t> reach;
  std::vector<int> seen(c->per ? nc : 0, 0);
  int ncolors = 0;
  for (int k = 0; k < nc; ++k) {
    if (c->per) {
      reach.assign(1, k);
      seen[k] = k + 1;
      for (int step = 0; step < 5; ++step) {
        const std::vector<int32_t> &ptr = step % 2 == 0 ? tptr : deg, &ad = step % 2 == 0 ? tadj : adj;
        const size_t c0 = reach.size();
        for (size_t q = 0; q < c0; ++q)
          for (int e = ptr[reach[q]]; e < ptr[reach[q] + 1]; ++e)
            if (seen[ad[e]] != k + 1) { seen[ad[e]] = k + 1; reach.push_back(ad[e]); }
      }
      for (size_t q = 1; q < reach.size(); ++q)
        if (color[reach[q]] >= 0 && color[reach[q]] < 4096) mark[color[reach[q]]] = k + 1;
    } else {
      for (int q = deg[k]; q < deg[k + 1]; ++q) {
        const int a = adj[q];
        if (color[a] >= 0 && color[a] < 4096) mark[color[a]] = k + 1;
        for (int q2 = deg[a]; q2 < deg[a + 1]; ++q2) {
          const int b = adj[q2];
          if (color[b] >= 0 && color[b] < 4096) mark[color[b]] = k + 1;
        }
      }
    }
    int col = 0;
    while (col < 4095 && mark[col] == k + 1) col++;
    color[k] = col;
    ncolors = std::max(ncolors, col + 1);
  }
  std::vector<double> &bm2 = c->h_band[0], &bm1 = c->h_band[1], &b0 = c->h_band[2], &bp1 = c->h_band[3],
                      &bp2 = c->h_band[4];
  for (auto &b : c->h_band) b.assign(n, 0.);
  {
    // probes built, applied and read out on the device; the rows come back once
    int32_t *d_color;
    int rc = upload(c, &d_color, color.data(), nc);
    if (rc) return rc;
    double *band[5] = {c->io[2], c->io[3], c->io[4], c->io[5], c->io[6]};
    const dim3 g(c->nodes_grid), b(NFT_BLOCK);
    for (int q = 0; q < 5; ++q) hipLaunchKernelGGL(k_fill, g, b, 0, c->stream, n, band[q], 0.0);
    for (int col = 0; col < ncolors; ++col) {
      for (int r = 0; r < 5; ++r) {
        hipLaunchKernelGGL(k_band_probe, g, b, 0, c->stream, n, L, d_color, col, r, c->io[0]);
        if ((rc = laplc_dev(c, c->io[0], c->io[1], 1, nullptr))) return rc;
        hipLaunchKernelGGL(k_band_extract, g, b, 0, c->stream, n, L, d_color, col, r, c->io[1], band[0], band[1],
                           band[2], band[3], band[4]);
      }
    }
    HIPCHK(hipGetLastError());
    for (int q = 0; q < 5; ++q)
      HIPCHK(hipMemcpyAsync(c->h_band[q].data(), band[q], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  // LU per column: row k = (l2, l1, 1) x U, U row k = (d, u1, u2)
  std::vector<double> fl1(n), fl2(n), fu1(n), fu2(n), fdinv(n);
  for (int k = 0; k < nc; ++k) {
    for (int l = 0; l < L; ++l) {
      const size_t i = (size_t)k * L + l;
      double l2 = 0., l1 = 0., t1 = bm1[i], d = b0[i];
      if (l >= 2) { l2 = bm2[i] * fdinv[i - 2]; t1 = t1 - l2 * fu1[i - 2]; d = d - l2 * fu2[i - 2]; }
      if (l >= 1) { l1 = t1 * fdinv[i - 1]; d = d - l1 * fu1[i - 1]; }
      double u1 = l + 1 < L ? bp1[i] : 0.;
      if (l >= 1 && l + 1 < L) u1 = u1 - l1 * fu2[i - 1];
      fl1[i] = l1; fl2[i] = l2;
      fu1[i] = u1;
      fu2[i] = l + 2 < L ? bp2[i] : 0.;
      fdinv[i] = 1. / d;
    }
  }
  // interleave: block cb of cpb columns, entry (cb*L + l)*cpb + lane
  const int nb = (nc + cpb - 1) / cpb;
  const size_t ni = (size_t)nb * L * cpb;
  std::vector<double> *src[5] = {&fl1, &fl2, &fu1, &fu2, &fdinv};
  double *dst[5];
  float *dstf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int q = 0; q < 5; ++q) {
    std::vector<double> il(ni, 0.);
    for (int k = 0; k < nc; ++k) {
      const int cb = k / cpb, lane = k % cpb;
      for (int l = 0; l < L; ++l) il[((size_t)cb * L + l) * cpb + lane] = (*src[q])[(size_t)k * L + l];
    }
    int rc = upload(c, &dst[q], il.data(), ni);
    if (rc) return rc;
    if (c->P.precision >= 1) {
      std::vector<float> ilf(ni);
      for (size_t t = 0; t < ni; ++t) ilf[t] = (float)il[t];
      if ((rc = upload(c, &dstf[q], ilf.data(), ni))) return rc;
    }
  }
  c->band = BandFactors<double>{dst[0], dst[1], dst[2], dst[3], dst[4]};
  c->bandf = BandFactors<float>{dstf[0], dstf[1], dstf[2], dstf[3], dstf[4]};
  c->band_L = L; c->band_cpb = cpb; c->band_colors = ncolors;
  return 0;
}

// connectivity and volume checks of a mesh before it reaches any kernel
int validate_mesh(const nft_mesh *m, const char *who) {
  const std::string w(who);
  if (m->npoin <= 0 || m->nedge < 0 || m->nbface < 0 || m->nboun < 0) return set_err(-1, w + ": bad mesh sizes");
  if (m->ld_poin < m->npoin || m->ld_edge < m->nedge || m->ld_face < std::max(m->nbface, m->nboun))
    return set_err(-1, w + ": leading dimension smaller than count");
  const int n = m->npoin;
  auto bad_node = [&](int32_t v) { return v < 1 || v > n; };
  for (int e = 0; e < m->nedge; ++e)
    if (bad_node(m->iedge[e]) || bad_node(m->iedge[m->ld_edge + e]) || m->iedge[e] == m->iedge[m->ld_edge + e])
      return set_err(-1, w + ": iedge entry out of range or degenerate at edge " + std::to_string(e + 1));
  for (int e = 0; e < m->nbface; ++e)
    if (bad_node(m->kbface[e]) || bad_node(m->kbface[m->ld_face + e]) || m->kbface[e] == m->kbface[m->ld_face + e])
      return set_err(-1, w + ": kbface entry out of range at " + std::to_string(e + 1));
  for (int b = 0; b < m->nboun; ++b)
    if (bad_node(m->ibpoin[b])) return set_err(-1, w + ": ibpoin node out of range at " + std::to_string(b + 1));
  if (m->ncol > 0) {
    if (!m->col_start || m->col_start[0] != 0 || m->col_start[m->ncol] != n)
      return set_err(-1, w + ": col_start must cover 0..npoin");
    for (int k = 0; k < m->ncol; ++k)
      if (m->col_start[k + 1] <= m->col_start[k]) return set_err(-1, w + ": empty column");
  }
  for (int i = 0; i < n; ++i)
    if (!(m->vol[i] > 0)) return set_err(-1, w + ": non-positive volume at node " + std::to_string(i + 1));
  return 0;
}

// Node-sorted slots of one incidence on the host: SELL-64 (each node's slots in the
// global edge order, interior edges then boundary-face edges) and, when the mesh is extruded
// with identical interior levels, the columnar form of the same slots. both: every edge in
// the roles of both its nodes (the gathers); otherwise only in the role of its first node
// i1 (the owned incidence: each edge appears once). Every chunk gets at least one slot and
// every column at least one bottom/top slot pair, so a lane whose node has no slots still
// reads a valid entry (which it then drops).
struct HostSlots {
  int nchunk = 0;
  std::vector<int32_t> deg, chunk_ptr, nbr, eid;
  int colL = 0;
  std::vector<int32_t> tptr, xptr;
  std::vector<int2> tab;
  int xoff = 0;
  double col_bytes = 0;
};

int host_slots(int n, int ncol, const int32_t *col_start, const std::vector<int32_t> &i1,
               const std::vector<int32_t> &i2, bool both, HostSlots &H) {
  const long etot = (long)i1.size();
  std::vector<int32_t> &deg = H.deg, &chunk_ptr = H.chunk_ptr, &nbr = H.nbr, &eid = H.eid;
  deg.assign(n, 0);
  for (long e = 0; e < etot; ++e) {
    deg[i1[e]]++;
    if (both) deg[i2[e]]++;
  }
  const int nchunk = (n + NFT_WAVE - 1) / NFT_WAVE;
  H.nchunk = nchunk;
  chunk_ptr.assign(nchunk + 1, 0);
  for (int ch = 0; ch < nchunk; ++ch) {
    int w = 1;
    for (int r = 0; r < NFT_WAVE; ++r) {
      int i = ch * NFT_WAVE + r;
      if (i < n) w = std::max(w, deg[i]);
    }
    if ((long)chunk_ptr[ch] + (long)NFT_WAVE * w > 2000000000L) return set_err(-1, "incidence too large");
    chunk_ptr[ch + 1] = chunk_ptr[ch] + NFT_WAVE * w;
  }
  const long nslot = chunk_ptr[nchunk];
  nbr.assign(nslot, 0); eid.assign(nslot, 0);
  std::vector<int32_t> fill(n, 0);
  for (int i = 0; i < n; ++i) {
    int ch = i / NFT_WAVE, r = i % NFT_WAVE;
    int w = (chunk_ptr[ch + 1] - chunk_ptr[ch]) / NFT_WAVE;
    for (int s = 0; s < w; ++s) { nbr[chunk_ptr[ch] + s * NFT_WAVE + r] = i; eid[chunk_ptr[ch] + s * NFT_WAVE + r] = 0; }
  }
  auto put = [&](int i, int j, int32_t code) {
    int ch = i / NFT_WAVE, r = i % NFT_WAVE;
    long slot = chunk_ptr[ch] + (long)fill[i] * NFT_WAVE + r;
    nbr[slot] = j; eid[slot] = code; fill[i]++;
  };
  for (long e = 0; e < etot; ++e) {
    put(i1[e], i2[e], (int32_t)e);
    if (both) put(i2[e], i1[e], ~(int32_t)e);
  }
  // columnar form: template of level 1 per column, checked against every interior level
  std::vector<int32_t> &tptr = H.tptr, &xptr = H.xptr;
  std::vector<int32_t> tnb, tco, xnb, xco;
  H.colL = 0;
  if (ncol > 0 && n % ncol == 0 && n / ncol >= 3) {
    const int Lc = n / ncol;
    bool ok = true;
    for (int k = 0; k <= ncol && ok; ++k) ok = col_start[k] == k * Lc;
    auto sl = [&](int i, int q) { int ch = i / NFT_WAVE, r = i % NFT_WAVE; return chunk_ptr[ch] + (long)q * NFT_WAVE + r; };
    tptr.assign(ncol + 1, 0);
    xptr.assign(ncol + 1, 0);
    for (int cc = 0; cc < ncol && ok; ++cc) {
      const int i1n = cc * Lc + 1, w = deg[i1n];
      for (int q = 0; q < w; ++q) {
        long s1 = sl(i1n, q);
        int e = eid[s1];
        // edge code base: ((edge - k + 1) << 1) | (i2 role), edge - k + 1 >= 0
        int ea = e >= 0 ? e : ~e;
        if (ea >= (1 << 29)) { ok = false; break; }
        tnb.push_back(nbr[s1] - 1);
        tco.push_back((ea << 1) | (e < 0 ? 1 : 0));
      }
      tptr[cc + 1] = tptr[cc] + w;
      for (int k = 1; k < Lc - 1 && ok; ++k) {
        const int i = cc * Lc + k;
        if (deg[i] != w) { ok = false; break; }
        for (int q = 0; q < w; ++q) {
          long sk = sl(i, q);
          int b = tco[tptr[cc] + q];
          int ea = (b >> 1) - 1 + k;
          int e = (b & 1) ? ~ea : ea;
          if (nbr[sk] != tnb[tptr[cc] + q] + k || eid[sk] != e) { ok = false; break; }
        }
      }
      const int ib = cc * Lc, it = cc * Lc + Lc - 1;
      const int xw = std::max(1, std::max(deg[ib], deg[it]));
      for (int q = 0; q < xw; ++q)
        for (int t = 0; t < 2; ++t) {
          int i = t ? it : ib;
          bool has = q < deg[i];
          xnb.push_back(has ? nbr[sl(i, q)] : i);
          xco.push_back(has ? eid[sl(i, q)] : 0);
        }
      xptr[cc + 1] = xptr[cc] + 2 * xw;
    }
    if (ok) {
      H.colL = Lc;
      H.tab.resize(tnb.size() + xnb.size());
      for (size_t q = 0; q < tnb.size(); ++q) H.tab[q] = make_int2(tnb[q], tco[q]);
      for (size_t q = 0; q < xnb.size(); ++q) H.tab[tnb.size() + q] = make_int2(xnb[q], xco[q]);
      H.xoff = (int)tnb.size();
      H.col_bytes = 4.0 * (tptr.size() + xptr.size()) + 8.0 * H.tab.size();
    }
  }
  return 0;
}

// the slot arrays of H on the device, into g (sizes, SELL-64 arrays, columnar tables and the
// launch shape; boundary tables and metrics are the caller's)
int upload_slots(nft_ctx *c, int n, const HostSlots &H, Incidence &g) {
  int rc;
  int32_t *d_chunk, *d_deg, *d_nbr, *d_eid;
  if ((rc = upload(c, &d_chunk, H.chunk_ptr.data(), H.chunk_ptr.size())) ||
      (rc = upload(c, &d_deg, H.deg.data(), H.deg.size())) || (rc = upload(c, &d_nbr, H.nbr.data(), H.nbr.size())) ||
      (rc = upload(c, &d_eid, H.eid.data(), H.eid.size())))
    return rc;
  g.n = n; g.nchunk = H.nchunk; g.chunk_ptr = d_chunk; g.deg = d_deg; g.nbr = d_nbr; g.eid = d_eid;
  g.col_L = 0;
  if (H.colL > 0) {
    int32_t *d_tptr, *d_xptr;
    int2 *d_tab;
    if ((rc = upload(c, &d_tptr, H.tptr.data(), H.tptr.size())) ||
        (rc = upload(c, &d_xptr, H.xptr.data(), H.xptr.size())) || (rc = upload(c, &d_tab, H.tab.data(), H.tab.size())))
      return rc;
    g.col_L = H.colL;
    g.tptr = d_tptr; g.xptr = d_xptr; g.tab = d_tab; g.xoff = H.xoff;
  }
  g.nblocks = (H.nchunk + NFT_WPB - 1) / NFT_WPB;
  g.xcd_per = g.nblocks >= 64 ? (g.nblocks + 7) / 8 : 0;
  return 0;
}

// Column patches for the LDS-staged MPDATA bounds and limiter sweeps (columns of L =
// NFT_WAVE levels, columnar slots H): each patch grows from the lowest unassigned column,
// adding the unassigned neighbour with most links into the patch (then the lowest id) while
// the staged set (its columns and every column their slots reach) stays within
// NFT_PATCH_STAGE, up to NFT_PATCH_COLS columns. Staged lists hold the patch's columns
// first, then the others in increasing order; tabl is H.tab with each neighbour rewritten
// as (staged index * L + level), relative to the patch of the column that owns the entry.
// Returns the largest staged set, or 0 when a column alone exceeds the limit.
int build_patches(int ncol, int L, const HostSlots &H, std::vector<int32_t> &ptr, std::vector<int32_t> &col,
                  std::vector<uint8_t> &nown, std::vector<int2> &tabl) {
  std::vector<std::vector<int32_t>> nb(ncol);
  for (int c = 0; c < ncol; ++c) {
    std::vector<int32_t> &v = nb[c];
    for (int t = H.tptr[c]; t < H.tptr[c + 1]; ++t) v.push_back((H.tab[t].x + 1) / L);
    for (int t = H.xptr[c]; t < H.xptr[c + 1]; ++t) v.push_back(H.tab[H.xoff + t].x / L);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    v.erase(std::remove(v.begin(), v.end(), c), v.end());
  }
  std::vector<int32_t> owner(ncol, -1), mark(ncol, -1), sidx(ncol, -1);
  ptr.assign(1, 0);
  col.clear();
  nown.clear();
  int smax = 0;
  for (int seed = 0; seed < ncol; ++seed) {
    if (owner[seed] >= 0) continue;
    const int p = (int)nown.size();
    std::vector<int32_t> own{seed}, S{seed};
    mark[seed] = p;
    for (int z : nb[seed])
      if (mark[z] != p) { mark[z] = p; S.push_back(z); }
    if ((int)S.size() > NFT_PATCH_STAGE) return 0;
    owner[seed] = p;
    while ((int)own.size() < NFT_PATCH_COLS) {
      int best = -1, bl = -1;
      for (int y : S) {
        if (owner[y] >= 0) continue;
        int inc = 0, links = 0;
        for (int z : nb[y]) {
          inc += mark[z] != p;
          links += owner[z] == p;
        }
        if ((int)S.size() + inc > NFT_PATCH_STAGE) continue;
        if (links > bl || (links == bl && y < best)) { best = y; bl = links; }
      }
      if (best < 0) break;
      owner[best] = p;
      own.push_back(best);
      for (int z : nb[best])
        if (mark[z] != p) { mark[z] = p; S.push_back(z); }
    }
    std::vector<int32_t> rest;
    for (int y : S)
      if (owner[y] != p) rest.push_back(y);
    std::sort(rest.begin(), rest.end());
    col.insert(col.end(), own.begin(), own.end());
    col.insert(col.end(), rest.begin(), rest.end());
    nown.push_back((uint8_t)own.size());
    ptr.push_back((int32_t)col.size());
    smax = std::max(smax, (int)S.size());
  }
  tabl = H.tab;
  const int np = (int)nown.size();
  for (int p = 0; p < np; ++p) {
    for (int k = ptr[p]; k < ptr[p + 1]; ++k) sidx[col[k]] = k - ptr[p];
    for (int k = ptr[p]; k < ptr[p] + nown[p]; ++k) {
      const int c = col[k];
      for (int t = H.tptr[c]; t < H.tptr[c + 1]; ++t) {
        const int cj = (H.tab[t].x + 1) / L;
        tabl[t].x = H.tab[t].x - cj * L + sidx[cj] * L;
      }
      for (int t = H.xoff + H.xptr[c]; t < H.xoff + H.xptr[c + 1]; ++t) {
        const int cj = H.tab[t].x / L;
        tabl[t].x = sidx[cj] * L + (H.tab[t].x - cj * L);
      }
    }
    for (int k = ptr[p]; k < ptr[p + 1]; ++k) sidx[col[k]] = -1;
  }
  return smax;
}

// Device geometry of one mesh: the node-sorted SELL-64 incidence (each node's slots in the
// global edge order: interior edges, then boundary-face edges), its columnar form when the
// mesh is extruded with identical interior levels, the per-node boundary-point tables and
// the edge metrics. Host copies of the edge lists and metrics stay in G for the callers.
struct Geometry {
  Incidence inc{};
  EdgeS<double> S{};
  int gather_grid = 0;
  int col_on = 0;
  double col_bytes = 0;
  int n_flag78 = 0;   // boundary entries with flag 7 or 8 (where bcv reads ue)
  // the owned incidence (every edge once, at its first node), built when want_own is set
  int want_own = 0;
  Incidence own{};
  double own_col_bytes = 0;
  // column patches (want_patches, columns of NFT_WAVE levels): see build_patches
  int want_patches = 0;
  Incidence inc_lds{};
  MpPatches patches{};
  int patch_smax = 0;
  std::vector<int32_t> i1, i2;
  std::vector<double> sx, sy, sz;
};

int build_geometry(nft_ctx *c, const nft_mesh *m, Geometry &G) {
  const int n = m->npoin;
  const long etot = (long)m->nedge + m->nbface;
  int rc = 0;
  G.i1.assign(etot, 0); G.i2.assign(etot, 0);
  G.sx.assign(etot, 0.); G.sy.assign(etot, 0.); G.sz.assign(etot, 0.);
  std::vector<int32_t> &i1 = G.i1, &i2 = G.i2;
  std::vector<double> &sx = G.sx, &sy = G.sy, &sz = G.sz;
  for (int e = 0; e < m->nedge; ++e) {
    i1[e] = m->iedge[e] - 1; i2[e] = m->iedge[m->ld_edge + e] - 1;
    sx[e] = m->sn[e]; sy[e] = m->sn[m->ld_edge + e]; sz[e] = m->sn[2 * (size_t)m->ld_edge + e];
  }
  for (int e = 0; e < m->nbface; ++e) {
    long k = m->nedge + e;
    i1[k] = m->kbface[e] - 1; i2[k] = m->kbface[m->ld_face + e] - 1;
    sx[k] = m->sb[e]; sy[k] = m->sb[m->ld_face + e]; sz[k] = m->sb[2 * (size_t)m->ld_face + e];
  }
  HostSlots H;
  if ((rc = host_slots(n, m->ncol, m->col_start, i1, i2, true, H))) return rc;

  // boundary points grouped per node, in entry order
  std::vector<int32_t> bp_ptr(n + 1, 0);
  for (int b = 0; b < m->nboun; ++b) bp_ptr[m->ibpoin[b]]++;
  for (int i = 0; i < n; ++i) bp_ptr[i + 1] += bp_ptr[i];
  const int nbp = m->nboun;
  std::vector<int32_t> bfill(bp_ptr.begin(), bp_ptr.end() - 1), bflag(nbp);
  std::vector<double> bsbb(3 * (size_t)nbp), banp(3 * (size_t)nbp);
  for (int b = 0; b < nbp; ++b) {
    int i = m->ibpoin[b] - 1;
    int q = bfill[i]++;
    bflag[q] = m->ibpoin[m->ld_face + b];
    if (bflag[q] == 7 || bflag[q] == 8) G.n_flag78++;
    for (int d = 0; d < 3; ++d) {
      bsbb[(size_t)d * nbp + q] = m->sbb[(size_t)d * m->ld_face + b];
      banp[(size_t)d * nbp + q] = m->anp[(size_t)d * m->ld_face + b];
    }
  }
  int32_t *d_bp_ptr, *d_bflag;
  double *d_bsbb, *d_banp;
  if ((rc = upload(c, &d_bp_ptr, bp_ptr.data(), bp_ptr.size())) ||
      (rc = upload(c, &d_bflag, bflag.data(), bflag.size())) || (rc = upload(c, &d_bsbb, bsbb.data(), bsbb.size())) ||
      (rc = upload(c, &d_banp, banp.data(), banp.size())))
    return rc;
  Incidence &g = G.inc;
  if ((rc = upload_slots(c, n, H, g))) return rc;
  if (g.col_L > 0) {
    G.col_on = 1;
    G.col_bytes = H.col_bytes;
  }
  g.bp_ptr = d_bp_ptr; g.bp_flag = d_bflag; g.bp_sbb = d_bsbb; g.bp_anp = d_banp; g.nbp = nbp;
  G.gather_grid = g.xcd_per > 0 ? 8 * g.xcd_per : g.nblocks;
  if (G.want_patches && H.colL == NFT_WAVE) {
    std::vector<int32_t> pptr, pcol;
    std::vector<uint8_t> pown;
    std::vector<int2> tabl;
    const int smax = build_patches(m->ncol, NFT_WAVE, H, pptr, pcol, pown, tabl);
    if (smax > 0) {
      int32_t *d_pptr, *d_pcol;
      uint8_t *d_pown;
      int2 *d_tabl;
      if ((rc = upload(c, &d_pptr, pptr.data(), pptr.size())) || (rc = upload(c, &d_pcol, pcol.data(), pcol.size())) ||
          (rc = upload(c, &d_pown, pown.data(), pown.size())) || (rc = upload(c, &d_tabl, tabl.data(), tabl.size())))
        return rc;
      const int np = (int)pown.size();
      G.patches = MpPatches{np, np >= 64 ? (np + 7) / 8 : 0, smax * NFT_WAVE, d_pptr, d_pcol, d_pown};
      G.inc_lds = g;
      G.inc_lds.tab = d_tabl;
      G.patch_smax = smax;
    }
  }
  if (G.want_own) {
    // each edge once, in the role of its first node: MPDATA's antidiffusive-flux sweep
    HostSlots Ho;
    if ((rc = host_slots(n, m->ncol, m->col_start, i1, i2, false, Ho)) || (rc = upload_slots(c, n, Ho, G.own)))
      return rc;
    G.own.nbp = 0;
    G.own_col_bytes = Ho.col_bytes;
  }
  double *dsx, *dsy, *dsz;
  if ((rc = upload(c, &dsx, sx.data(), etot)) || (rc = upload(c, &dsy, sy.data(), etot)) ||
      (rc = upload(c, &dsz, sz.data(), etot)))
    return rc;
  G.S = EdgeS<double>{dsx, dsy, dsz};
  {
    std::vector<double2> sxy(etot);
    for (long e = 0; e < etot; ++e) sxy[e] = make_double2(sx[e], sy[e]);
    double2 *d_sxy;
    if ((rc = upload(c, &d_sxy, sxy.data(), etot))) return rc;
    g.sxy = d_sxy;
    g.szz = dsz;
  }
  if (G.patch_smax > 0) { G.inc_lds.sxy = g.sxy; G.inc_lds.szz = g.szz; }
  return 0;
}

}  // namespace

// =====================================================================================
extern "C" {

const char *nft_last_error(void) { return g_err.c_str(); }

int nft_device_name(char *buf, int len) {
  hipDeviceProp_t prop;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return 0;
}

}  // extern "C"

namespace {

// the coincident sets of a periodic box: checked (1-based nodes in range, every node in at
// most one set) and flattened into 0-based member lists in the reference's order (icoin,
// icoin2, icoin4)
int periodic_sets(const nft_periodic *pp, int n, std::vector<int4> &sets) {
  if (pp->ncoin < 0 || pp->ncoin2 < 0 || pp->ncoin4 < 0 || pp->ld_coin < pp->ncoin || pp->ld_coin2 < pp->ncoin2 ||
      pp->ld_coin4 < pp->ncoin4)
    return set_err(-1, "nft_init_periodic: bad coincident-set counts or leading dimensions");
  if ((pp->ncoin && !pp->icoin) || (pp->ncoin2 && !pp->icoin2) || (pp->ncoin4 && !pp->icoin4))
    return set_err(-1, "nft_init_periodic: null coincident-set array");
  std::vector<char> seen(n, 0);
  sets.clear();
  const int32_t *lists[3] = {pp->icoin, pp->icoin2, pp->icoin4};
  const int cnt[3] = {pp->ncoin, pp->ncoin2, pp->ncoin4}, ld[3] = {pp->ld_coin, pp->ld_coin2, pp->ld_coin4};
  for (int t = 0; t < 3; ++t)
    for (int q = 0; q < cnt[t]; ++q) {
      int mem[4] = {-1, -1, -1, -1};
      for (int k = 0; k < (t == 2 ? 4 : 2); ++k) {
        const int32_t v = lists[t][(size_t)k * ld[t] + q];
        if (v < 1 || v > n) return set_err(-1, "nft_init_periodic: coincident node out of range");
        if (seen[v - 1]) return set_err(-1, "nft_init_periodic: node " + std::to_string(v) + " is in more than one coincident set");
        seen[v - 1] = 1;
        mem[k] = v - 1;
      }
      sets.push_back(make_int4(mem[0], mem[1], mem[2], mem[3]));
    }
  if (sets.empty()) return set_err(-1, "nft_init_periodic: no coincident sets");
  return 0;
}

// the patch sweeps' launch attribute beyond the 64 KB default LDS
template <typename T, int NF>
int patch_lds_attr(const nft_ctx *c, int optin) {
  const size_t bf = patch_lds_bytes<T, NF>(c, false), bl = patch_lds_bytes<T, NF>(c, true);
  const size_t bd = patch_lds_bytes_donor<T, NF>(c);
  if ((long)bl > optin || (long)bd > optin) return -1;
  if (hipFuncSetAttribute((const void *)k_mp_donor_lds<T, NF>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bd) !=
          hipSuccess ||
      hipFuncSetAttribute((const void *)k_mp_grad_lds<T, NF>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bf) !=
          hipSuccess ||
      hipFuncSetAttribute((const void *)k_mp_fct_lds<T, NF>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bf) !=
          hipSuccess ||
      hipFuncSetAttribute((const void *)k_mp_limit_lds<T, NF>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bl) !=
          hipSuccess)
    return -1;
  return 0;
}

// 0 when every instance of the patch sweeps can take its LDS on this device
int patch_lds_setup(const nft_ctx *c) {
  int dev = 0, optin = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess || optin <= 0) {
    (void)hipGetLastError();
    return -1;
  }
  if (patch_lds_attr<double, 4>(c, optin) || patch_lds_attr<float, 4>(c, optin) || patch_lds_attr<double, 1>(c, optin) ||
      patch_lds_attr<float, 1>(c, optin)) {
    (void)hipGetLastError();
    return -1;
  }
  return 0;
}

int init_impl(const nft_mesh *m, const nft_params *P, const nft_periodic *pp, nft_ctx **out) {
  if (!m || !P || !out) return set_err(-1, "nft_init: null argument");
  *out = nullptr;
  if (P->lord < 1 || P->lord > NFT_MAXD) return set_err(-1, "nft_init: lord out of range 1..18");
  if (P->precon >= 2 && m->ncol <= 0) return set_err(-1, "nft_init: line preconditioner needs columns");
  if (P->precon < 0 || P->precon > 3) return set_err(-1, "nft_init: unknown preconditioner");
  if (!(P->dt > 0)) return set_err(-1, "nft_init: dt must be positive");
  if (P->itraj != 0 && P->itraj != 1) return set_err(-1, "nft_init: itraj must be 0 or 1");
  if (P->precision < 0 || P->precision > 2) return set_err(-1, "nft_init: precision must be 0, 1 or 2");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(-3, "nft_init: no HIP device");
  const int n = m->npoin;
  const long etot = (long)m->nedge + m->nbface;
  {
    int rc0 = validate_mesh(m, "nft_init");
    if (rc0) return rc0;
  }
  // periodic box: vol summed over the coincident sets before anything reads it (mpsolver
  // :99-103, cyclicr then cyclicr2 then cyclicr4; the sets are disjoint)
  std::vector<int4> psets;
  std::vector<double> volc;
  nft_mesh mper{};
  if (pp) {
    if (P->precision != 0) return set_err(-1, "nft_init_periodic: the periodic box runs in FP64 (precision 0)");
    int rc0 = periodic_sets(pp, n, psets);
    if (rc0) return rc0;
    volc.assign(m->vol, m->vol + n);
    if (!pp->vol_closed) {
      for (const int4 &q : psets) {
        double u = volc[q.x] + volc[q.y];
        if (q.z >= 0) u = (u + volc[q.z]) + volc[q.w];
        volc[q.x] = u; volc[q.y] = u;
        if (q.z >= 0) { volc[q.z] = u; volc[q.w] = u; }
      }
    }
    mper = *m;
    mper.vol = volc.data();
    m = &mper;
  }
  nft_ctx *c = new nft_ctx();
  c->P = *P;
  c->n = n; c->nedge = m->nedge; c->nbface = m->nbface; c->nboun = m->nboun; c->etot = (int)etot;
  c->ncol = m->ncol;
  int rc = 0;
  auto fail = [&](int code) { nft_finalize(c); return code; };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(set_err(-2, "stream"));
  if (scratch_address(&c->scratch_dev)) c->scratch_err = g_err;
  for (auto &e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(set_err(-2, "event"));

  // ---- geometry: incidence, columnar form, boundary tables, edge metrics
  Geometry geo;
  geo.want_own = 1;
  geo.want_patches = pp ? 0 : 1;
  if ((rc = build_geometry(c, m, geo))) return fail(rc);
  c->inc = geo.inc; c->S = geo.S; c->gather_grid = geo.gather_grid;
  c->inc_own = geo.own; c->own_col_bytes = geo.own_col_bytes;
  c->inc_lds = geo.inc_lds; c->patches = geo.patches; c->patch_smax = geo.patch_smax;
  if (c->patch_smax > 0) c->patch_on = patch_lds_setup(c) == 0 ? 1 : 0;
  if (pp) {
    // node -> slot of the set sums (16 values per node: the most a sweep keeps, four
    // fields' MPDATA bounds and in/out sums)
    std::vector<int32_t> cslot(n, -1);
    int ncn = 0;
    for (const int4 &q : psets)
      for (int k = 0; k < (q.z >= 0 ? 4 : 2); ++k) cslot[k == 0 ? q.x : k == 1 ? q.y : k == 2 ? q.z : q.w] = ncn++;
    int4 *d_sets;
    int32_t *d_cslot;
    double *d_cval;
    if ((rc = upload(c, &d_sets, psets.data(), psets.size())) || (rc = upload(c, &d_cslot, cslot.data(), n)) ||
        (rc = dalloc(c, &d_cval, (size_t)16 * ncn)))
      return fail(rc);
    c->per = 1;
    c->h_sets = psets;
    c->ps = PerSets{(int)psets.size(), d_sets, d_cslot, d_cval};
    c->inc.per = 1; c->inc.cslot = d_cslot; c->inc.cval = d_cval;
  }
  c->col_on = geo.col_on; c->col_bytes = geo.col_bytes; c->n_flag78 = geo.n_flag78;
  c->vec_grid = std::min(blocks_for(n), 4096);
  c->nodes_grid = blocks_for(n);
  const std::vector<int32_t> &i1 = geo.i1, &i2 = geo.i2;
  const std::vector<double> &sx = geo.sx, &sy = geo.sy, &sz = geo.sz;
  if (m->ncol > 0) {
    // horizontal column adjacency (edges between different columns at one level)
    std::vector<int> colof(n), levof(n);
    for (int k = 0; k < m->ncol; ++k)
      for (int i = m->col_start[k]; i < m->col_start[k + 1]; ++i) { colof[i] = k; levof[i] = i - m->col_start[k]; }
    for (long e = 0; e < etot; ++e) {
      int a = colof[i1[e]], b = colof[i2[e]];
      if (a != b && levof[i1[e]] == levof[i2[e]]) c->h_hedges.push_back({std::min(a, b), std::max(a, b)});
    }
    std::sort(c->h_hedges.begin(), c->h_hedges.end());
    c->h_hedges.erase(std::unique(c->h_hedges.begin(), c->h_hedges.end()), c->h_hedges.end());
  }
  // ---- set-up fields
  std::vector<double> rh0, th0, ue, ve, dthedz, pstep;
  setup_fields(c, m, rh0, th0, ue, ve, dthedz, pstep);
  std::vector<double> volv(m->vol, m->vol + n);
  if ((rc = upload(c, &c->vol, volv.data(), n)) || (rc = upload(c, &c->rh0, rh0.data(), n)) ||
      (rc = upload(c, &c->th0, th0.data(), n)) || (rc = upload(c, &c->ue, ue.data(), n)) ||
      (rc = upload(c, &c->ve, ve.data(), n)) || (rc = upload(c, &c->coefw, c->h_coefw.data(), n)) ||
      (rc = upload(c, &c->alpha, c->h_alpha.data(), n)) || (rc = upload(c, &c->d1, c->h_d1.data(), n)) ||
      (rc = upload(c, &c->dthedz, dthedz.data(), n)) || (rc = upload(c, &c->pstep, pstep.data(), n)) ||
      (rc = upload(c, &c->la, c->h_la.data(), n)) || (rc = upload(c, &c->linv, c->h_linv.data(), n)) ||
      (rc = upload(c, &c->lc, c->h_lc.data(), n)))
    return fail(rc);
  if (m->ncol > 0) {
    if ((rc = upload(c, &c->col_start, m->col_start, m->ncol + 1))) return fail(rc);
  }
  // exact shortcuts for uniform node coefficients
  c->nc.vol = c->vol; c->nc.rh0 = c->rh0; c->nc.coefw = c->coefw; c->nc.ue = c->ue;
  c->nc.rh0_one = std::all_of(rh0.begin(), rh0.end(), [](double v) { return v == 1.0; }) ? 1 : 0;
  c->nc.coefw0 = c->h_coefw[0];
  c->nc.coefw_uniform =
      std::all_of(c->h_coefw.begin(), c->h_coefw.end(), [&](double v) { return v == c->nc.coefw0; }) ? 1 : 0;
  if ((rc = dalloc(c, &c->zbuf, n))) return fail(rc);
  {
    std::vector<double> vr(n);
    for (int i = 0; i < n; ++i) vr[i] = volv[i] * rh0[i];
    if ((rc = upload(c, &c->vr, vr.data(), n))) return fail(rc);
  }
  // FP32 copies for the mixed-precision form
  if (P->precision >= 1) {
    auto tof = [](const std::vector<double> &v) { std::vector<float> f(v.size()); for (size_t i = 0; i < v.size(); ++i) f[i] = (float)v[i]; return f; };
    std::vector<float> f;
    float *fx, *fy, *fz;
    if ((rc = upload(c, &fx, (f = tof(sx)).data(), etot)) || (rc = upload(c, &fy, (f = tof(sy)).data(), etot)) ||
        (rc = upload(c, &fz, (f = tof(sz)).data(), etot)))
      return fail(rc);
    c->Sf = EdgeS<float>{fx, fy, fz};
    {
      std::vector<float> vrf(n);
      for (int i = 0; i < n; ++i) {
        const float a = (float)volv[i], b = (float)rh0[i];
        vrf[i] = a * b;
      }
      if ((rc = upload(c, &c->vrf, vrf.data(), n))) return fail(rc);
    }
    if ((rc = upload(c, &c->volf, (f = tof(volv)).data(), n)) || (rc = upload(c, &c->rh0f, (f = tof(rh0)).data(), n)) ||
        (rc = upload(c, &c->laf, (f = tof(c->h_la)).data(), n)) || (rc = upload(c, &c->linvf, (f = tof(c->h_linv)).data(), n)) ||
        (rc = upload(c, &c->lcf, (f = tof(c->h_lc)).data(), n)) || (rc = dalloc(c, &c->zbuff, n)))
      return fail(rc);
    for (int k = 0; k < 3; ++k)
      if ((rc = dalloc(c, &c->uuuf[k], n))) return fail(rc);
    if (P->precision == 2) {
      // the FP32 operator: (Sx, Sy, Sz) rounded to FP32 and packed per edge
      std::vector<F3> s3f(etot);
      for (long e = 0; e < etot; ++e) s3f[e] = F3{(float)sx[e], (float)sy[e], (float)sz[e]};
      F3 *d_s3f;
      if ((rc = upload(c, &d_s3f, s3f.data(), etot)) || (rc = dalloc(c, &c->t3f, n))) return fail(rc);
      c->inc.s3f = d_s3f;
    }
  }

  // ---- state and work arrays (allocated once)
  for (int k = 0; k < 4; ++k)
    if ((rc = dalloc(c, &c->unk[k], n)) || (rc = dalloc(c, &c->frc[k], n))) return fail(rc);
  for (int k = 0; k < 3; ++k)
    if ((rc = dalloc(c, &c->uuu[k], n)) || (rc = dalloc(c, &c->uuu0[k], n)) || (rc = dalloc(c, &c->ud[k], n)))
      return fail(rc);
  if ((rc = dalloc(c, &c->prpot, n))) return fail(rc);
  for (int k = 0; k < 8; ++k)
    if ((rc = dalloc(c, &c->t[k], n))) return fail(rc);
  if ((rc = dalloc(c, &c->txy, n))) return fail(rc);
  for (int k = 0; k < P->lord; ++k)
    if ((rc = dalloc(c, &c->gx[k], n)) || (rc = dalloc(c, &c->gax[k], n))) return fail(rc);
  if ((rc = dalloc(c, &c->r, n)) || (rc = dalloc(c, &c->qr, n)) || (rc = dalloc(c, &c->ar, n)) ||
      (rc = dalloc(c, &c->pw, n)) || (rc = dalloc(c, &c->rw, n)))
    return fail(rc);
  const size_t tsz = P->precision >= 1 ? sizeof(float) : sizeof(double);
  for (int k = 0; k < 4; ++k)
    if ((rc = dalloc(c, &c->mp_psi1[k], n))) return fail(rc);
  // MPDATA work arrays per field: gradient (x, y) pairs (2n) and z (n), the (cp, cn) pairs
  // (2n); slots 3f+1 and 2f+1 are not used by the sweeps (the single-field mpdatam entry
  // keeps its FP32 advector in slots 9..11)
  for (int k = 0; k < 12; ++k) {
    void *q;
    if (hipMalloc(&q, (k % 3 == 0 ? 2 : 1) * n * tsz) != hipSuccess) return fail(set_err(-2, "alloc"));
    c->allocs.push_back(q); c->mp_g[k] = q;
  }
  for (int k = 0; k < 8; ++k) {
    void *q;
    if (hipMalloc(&q, (k % 2 == 0 ? 2 : 1) * n * tsz) != hipSuccess) return fail(set_err(-2, "alloc"));
    c->allocs.push_back(q); c->mp_c[k] = q;
  }
  {
    // two arrays of (field, field) pairs over all edges, sized for FP64
    void *q;
    if (hipMalloc(&q, (size_t)std::max(c->etot, 1) * 2 * sizeof(double2)) != hipSuccess)
      return fail(set_err(-2, "alloc"));
    c->allocs.push_back(q); c->mp_flux = q;
  }
  for (int k = 0; k < 7; ++k)
    if ((rc = dalloc(c, &c->io[k], n))) return fail(rc);
  if (P->itraj == 1) {
    if ((rc = dalloc(c, &c->vp_p, n)) || (rc = dalloc(c, &c->vp_ff[0], n)) || (rc = dalloc(c, &c->vp_ff[1], n)) ||
        (rc = dalloc(c, &c->vp_ff[2], n)) || (rc = dalloc(c, &c->vp_ffd, 3)))
      return fail(rc);
    if (hipMemcpy(c->vp_ffd, c->vp_ff, 3 * sizeof(double *), hipMemcpyHostToDevice) != hipSuccess)
      return fail(set_err(-2, "memcpy"));
  }
  if ((rc = dalloc(c, &c->io[7], 8))) return fail(rc);  // device array of field pointers
  size_t npart = (size_t)std::max(c->gather_grid, c->vec_grid) * (NFT_MAXD + 1);
  // small grids fold the GCR finalizes into the consuming kernels (same sums, same order)
  c->fold = c->gather_grid <= NFT_FOLD_BLOCKS && c->vec_grid <= NFT_FOLD_BLOCKS && P->lord <= NFT_FOLD_MAXD;
  if ((rc = dalloc(c, &c->partials, npart)) || (rc = dalloc(c, &c->partials2, npart)) ||
      (rc = dalloc(c, &c->counter, 4)) || (rc = dalloc(c, &c->scal, SC_TOTAL)) ||
      (rc = dalloc(c, &c->fin_wres, (size_t)NFT_FIN_WG * (NFT_MAXD + 1))))
    return fail(rc);
  if (hipMemset(c->counter, 0, 4 * sizeof(unsigned)) != hipSuccess) return fail(set_err(-2, "memset"));
  if (hipMemset(c->scal, 0, SC_TOTAL * sizeof(double)) != hipSuccess) return fail(set_err(-2, "memset"));
  if (hipHostMalloc(&c->h_scal, SC_TOTAL * sizeof(double)) != hipSuccess) return fail(set_err(-2, "host alloc"));
  if (hipHostMalloc(&c->h_flag, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
    return fail(set_err(-2, "host alloc"));
  *c->h_flag = 0.0;
  if (hipHostGetDevicePointer((void **)&c->d_hflag, c->h_flag, 0) != hipSuccess)
    return fail(set_err(-2, "hipHostGetDevicePointer"));
  if (hipEventCreateWithFlags(&c->poll_ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->poll_ev[1], hipEventDisableTiming) != hipSuccess)
    return fail(set_err(-2, "event"));
  for (int k = 0; k < 4; ++k)
    if (hipMemset(c->frc[k], 0, n * sizeof(double)) != hipSuccess) return fail(set_err(-2, "memset"));
  if (P->precon == 3 && (rc = build_band(c, m))) return fail(rc);
  if (hipDeviceSynchronize() != hipSuccess) return fail(set_err(-2, "sync"));
  *out = c;
  return 0;
}

}  // namespace

extern "C" {

int nft_init(const nft_mesh *m, const nft_params *P, nft_ctx **out) { return init_impl(m, P, nullptr, out); }

int nft_init_periodic(const nft_mesh *m, const nft_params *P, const nft_periodic *per, nft_ctx **out) {
  if (!per) return set_err(-1, "nft_init_periodic: null argument");
  return init_impl(m, P, per, out);
}

int nft_finalize(nft_ctx *c) {
  if (!c) return 0;
  if (c->stream) hipStreamSynchronize(c->stream);
  for (void *p : c->allocs) hipFree(p);
  if (c->h_scal) hipHostFree(c->h_scal);
  if (c->h_flag) hipHostFree(c->h_flag);
  for (auto &e : c->poll_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->ev)
    if (e) hipEventDestroy(e);
  for (auto &e : c->tev)
    if (e) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int nft_get_field(nft_ctx *c, const char *name, double *host) {
  if (!c || !name || !host) return set_err(-1, "nft_get_field: null argument");
  const std::vector<double> *v = nullptr;
  std::string s(name);
  if (s == "alpha") v = &c->h_alpha;
  else if (s == "d1") v = &c->h_d1;
  else if (s == "coefw") v = &c->h_coefw;
  else if (s == "line_a") v = &c->h_la;
  else if (s == "line_inv") v = &c->h_linv;
  else if (s == "line_c") v = &c->h_lc;
  else if (s == "band_m2" || s == "band_m1" || s == "band_0" || s == "band_p1" || s == "band_p2") {
    if (!c->band_L) return set_err(-1, "nft_get_field: no band (precon 3 only)");
    const char *names[5] = {"band_m2", "band_m1", "band_0", "band_p1", "band_p2"};
    for (int q = 0; q < 5; ++q)
      if (s == names[q]) v = &c->h_band[q];
  }
  if (v) { std::memcpy(host, v->data(), c->n * sizeof(double)); return 0; }
  double *d = nullptr;
  if (s == "rh0") d = c->rh0;
  else if (s == "th0") d = c->th0;
  else if (s == "ue") d = c->ue;
  else if (s == "beta") { host[0] = c->beta1; return 0; }
  else if (s == "courant") d = c->t[5];   // Courant numbers of the last nft_diagnostics call
  if (!d) return set_err(-1, "nft_get_field: unknown field " + s);
  HIPCHK(hipMemcpy(host, d, c->n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int nft_derivative(nft_ctx *c, const double *unk, double *dpdxa, int idx) {
  if (!c || !unk || !dpdxa || idx < 1 || idx > 3) return set_err(-1, "nft_derivative: bad argument");
  int rc;
  if ((rc = copy_in(c, c->io[0], unk))) return rc;
  GradArgs g = grad_args(c);
  g.f = c->io[0]; g.o0 = c->io[1]; g.o1 = c->io[2]; g.o2 = c->io[3];
  if ((rc = grad_launch(c, GRAD_RAW, g))) return rc;
  if ((rc = copy_out(c, dpdxa, c->io[idx]))) return rc;
  return sync(c);
}

int nft_laplc(nft_ctx *c, const double *press, double *alap, int impfl) {
  if (!c || !press || !alap) return set_err(-1, "nft_laplc: null argument");
  int rc;
  if ((rc = copy_in(c, c->io[0], press))) return rc;
  if ((rc = laplc_dev(c, c->io[0], c->io[1], impfl, nullptr))) return rc;
  if ((rc = copy_out(c, alap, c->io[1]))) return rc;
  return sync(c);
}

int nft_laplc_noalloc(nft_ctx *c, const double *press, double *alap, int impfl) {
  if (!c || !press || !alap || (impfl != 0 && impfl != 1)) return set_err(-1, "nft_laplc_noalloc: bad argument");
  if (c->n > kScratchNodes)
    return set_err(-1, "nft_laplc_noalloc: the mesh has " + std::to_string(c->n) + " nodes, more than the " +
                           std::to_string(kScratchNodes) + " its fixed scratch array holds");
  if (!c->scratch_dev) return set_err(-2, "nft_laplc_noalloc: " + c->scratch_err);
  // k_grad writes the scaled gradient of every node before k_div reads any of it, so the
  // scratch needs no clearing; the kernels are launched directly (no timer events)
  GradArgs g = grad_args(c);
  g.f = press; g.oxy = reinterpret_cast<double2 *>(c->scratch_dev); g.o2 = c->scratch_dev + 2 * (size_t)c->n;
  g.impfl = impfl;
  dim3 gg(c->gather_grid), b(NFT_BLOCK);
  if (c->per) per_sums3(c, press, press, press, nullptr, nullptr);
  if (c->col_on)
    hipLaunchKernelGGL((k_grad<GRAD_LAPLC, true>), gg, b, 0, c->stream, c->inc, c->S, g);
  else
    hipLaunchKernelGGL((k_grad<GRAD_LAPLC, false>), gg, b, 0, c->stream, c->inc, c->S, g);
  HIPCHK(hipGetLastError());
  DivArgs d = div_args(c);
  d.qxy = g.oxy; d.q2 = g.o2; d.out = alap;
  if (c->per) per_sums3(c, nullptr, nullptr, g.o2, g.oxy, nullptr);
  div_go<DIV_LAPLC, 0, 0>(c, d);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int nft_laplc32(nft_ctx *c, const double *press, double *alap, int impfl) {
  if (!c || !press || !alap) return set_err(-1, "nft_laplc32: null argument");
  if (c->P.precision != 2) return set_err(-1, "nft_laplc32: the FP32 operator exists with precision 2 only");
  int rc;
  if ((rc = copy_in(c, c->io[0], press))) return rc;
  if ((rc = laplc32_dev(c, c->io[0], c->io[1], impfl, nullptr))) return rc;
  if ((rc = copy_out(c, alap, c->io[1]))) return rc;
  return sync(c);
}

int nft_laplc_t(nft_ctx *c, const double *y, double *out, int impfl) {
  if (!c || !y || !out) return set_err(-1, "nft_laplc_t: null argument");
  if (c->per) return set_err(-1, "nft_laplc_t: not available for the periodic box");
  int rc;
  if ((rc = copy_in(c, c->io[0], y))) return rc;
  if ((rc = laplt_dev(c, c->io[0], c->io[1], impfl))) return rc;
  if ((rc = copy_out(c, out, c->io[1]))) return rc;
  return sync(c);
}

int nft_prfr(nft_ctx *c, const double *press, const double *ustr, const double *vstr, const double *wstr,
             double *uout, double *vout, double *wout, int impfl) {
  if (!c || !press || !ustr || !vstr || !wstr || !uout || !vout || !wout) return set_err(-1, "nft_prfr: null argument");
  int rc;
  if ((rc = copy_in(c, c->io[0], press)) || (rc = copy_in(c, c->io[1], ustr)) || (rc = copy_in(c, c->io[2], vstr)) ||
      (rc = copy_in(c, c->io[3], wstr)))
    return rc;
  if ((rc = prfr_dev(c, c->io[0], c->io[1], c->io[2], c->io[3], c->io[4], c->io[5], c->io[6], impfl))) return rc;
  if ((rc = copy_out(c, uout, c->io[4])) || (rc = copy_out(c, vout, c->io[5])) || (rc = copy_out(c, wout, c->io[6])))
    return rc;
  return sync(c);
}

int nft_bcv(nft_ctx *c, double *u, double *v, double *w, double flag) {
  if (!c || !u || !v || !w) return set_err(-1, "nft_bcv: null argument");
  int rc;
  if ((rc = copy_in(c, c->io[0], u)) || (rc = copy_in(c, c->io[1], v)) || (rc = copy_in(c, c->io[2], w))) return rc;
  hipLaunchKernelGGL(k_bcv, dim3(c->nodes_grid), dim3(NFT_BLOCK), 0, c->stream, c->inc, c->io[0], c->io[1],
                     c->io[2], c->ue, flag);
  HIPCHK(hipGetLastError());
  if ((rc = copy_out(c, u, c->io[0])) || (rc = copy_out(c, v, c->io[1])) || (rc = copy_out(c, w, c->io[2]))) return rc;
  return sync(c);
}

int nft_div(nft_ctx *c, const double *ua, const double *va, const double *wa, double *rhs, int iflg) {
  if (!c || !ua || !va || !wa || !rhs) return set_err(-1, "nft_div: null argument");
  int rc;
  if ((rc = copy_in(c, c->io[0], ua)) || (rc = copy_in(c, c->io[1], va)) || (rc = copy_in(c, c->io[2], wa))) return rc;
  if ((rc = div_dev(c, c->io[0], c->io[1], c->io[2], c->io[3], iflg, nullptr))) return rc;
  if ((rc = copy_out(c, rhs, c->io[3]))) return rc;
  return sync(c);
}

int nft_precon(nft_ctx *c, const double *rhs, double *p, int impfl) {
  if (!c || !rhs || !p) return set_err(-1, "nft_precon: null argument");
  int rc;
  if ((rc = copy_in(c, c->io[0], rhs))) return rc;
  if ((rc = precon_dev(c, c->io[0], c->io[1], impfl))) return rc;
  if ((rc = copy_out(c, p, c->io[1]))) return rc;
  return sync(c);
}

int nft_gcrk(nft_ctx *c, double *p, const double *u1, const double *v1, const double *w1, double *u2, double *v2,
             double *w2, int impfl, nft_gcr_stats *stats) {
  if (!c || !p || !u1 || !v1 || !w1 || !u2 || !v2 || !w2) return set_err(-1, "nft_gcrk: null argument");
  int rc;
  if ((rc = copy_in(c, c->io[0], p)) || (rc = copy_in(c, c->io[1], u1)) || (rc = copy_in(c, c->io[2], v1)) ||
      (rc = copy_in(c, c->io[3], w1)))
    return rc;
  int res = gcrk_dev(c, c->io[0], c->io[1], c->io[2], c->io[3], c->io[4], c->io[5], c->io[6], impfl);
  if (res < 0) return res;
  if ((rc = copy_out(c, p, c->io[0])) || (rc = copy_out(c, u2, c->io[4])) || (rc = copy_out(c, v2, c->io[5])) ||
      (rc = copy_out(c, w2, c->io[6])))
    return rc;
  if ((rc = sync(c))) return rc;
  if (stats) *stats = c->st;
  return res;
}

int nft_mpdatam(nft_ctx *c, const double *uuu, int ld, double *unk, int iflag) {
  if (!c || !uuu || !unk || iflag < 1 || iflag > 4 || ld < c->n) return set_err(-1, "nft_mpdatam: bad argument");
  int rc;
  if ((rc = copy_in(c, c->io[0], uuu)) || (rc = copy_in(c, c->io[1], uuu + ld)) ||
      (rc = copy_in(c, c->io[2], uuu + 2 * (size_t)ld)) || (rc = copy_in(c, c->io[3], unk)))
    return rc;
  double *fields[1] = {c->io[3]};
  int fl[1] = {iflag};
  if ((rc = set_out_fields(c, fields, 1))) return rc;
  if (c->P.precision >= 1) {
    float *uf[3] = {(float *)c->mp_g[9], (float *)c->mp_g[10], (float *)c->mp_g[11]};
    // the NF=1 call uses only the first 3 gradient slots, so slots 9..11 are free for uuu in FP32
    for (int k = 0; k < 3; ++k)
      hipLaunchKernelGGL(k_to_t<float>, dim3(c->nodes_grid), dim3(NFT_BLOCK), 0, c->stream, c->n, c->io[k], uf[k]);
    rc = mpdata_dev<float, 1>(c, uf[0], uf[1], uf[2], fields, fl);
  } else {
    rc = mpdata_dev<double, 1>(c, c->io[0], c->io[1], c->io[2], fields, fl);
  }
  if (rc) return rc;
  if ((rc = copy_out(c, unk, c->io[3]))) return rc;
  return sync(c);
}

int nft_state_init(nft_ctx *c) {
  if (!c) return set_err(-1, "nft_state_init: null context");
  c->steps_done = 0;
  const int n = c->n;
  dim3 g(c->nodes_grid), b(NFT_BLOCK);
  int rc;
  // unk = (ue, ve, 0, 0); unk1 (kept in frc[0..2]) = (ue, ve, 0); prpot = 0   (:177-206)
  HIPCHK(hipMemcpyAsync(c->unk[0], c->ue, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->unk[1], c->ve, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_fill, g, b, 0, c->stream, n, c->unk[2], 0.0);
  hipLaunchKernelGGL(k_fill, g, b, 0, c->stream, n, c->unk[3], 0.0);
  hipLaunchKernelGGL(k_fill, g, b, 0, c->stream, n, c->prpot, 0.0);
  HIPCHK(hipMemcpyAsync(c->frc[0], c->ue, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->frc[1], c->ve, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_fill, g, b, 0, c->stream, n, c->frc[2], 0.0);
  c->st = nft_gcr_stats{};
  rc = gcrk_dev(c, c->prpot, c->frc[0], c->frc[1], c->frc[2], c->unk[0], c->unk[1], c->unk[2], 0);
  if (rc < 0) return rc;
  if ((rc = prfr_dev(c, c->prpot, c->frc[0], c->frc[1], c->frc[2], c->unk[0], c->unk[1], c->unk[2], 0))) return rc;
  hipLaunchKernelGGL(k_fill, g, b, 0, c->stream, n, c->prpot, 0.0);
  StepArgs a = step_args(c);
  hipLaunchKernelGGL(k_init_post, g, b, 0, c->stream, a);
  HIPCHK(hipGetLastError());
  if (c->P.itraj == 1 && (rc = velprd_dev(c, nullptr)) < 0) return rc;
  return sync(c);
}

int nft_run_steps(nft_ctx *c, int nsteps, nft_run_stats *stats) {
  if (!c || nsteps < 0) return set_err(-1, "nft_run_steps: bad argument");
  nft_run_stats rs{};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int s = 0; s < nsteps; ++s) {
    int rc = one_step(c, &rs);
    if (rc) { hipEventDestroy(e0); hipEventDestroy(e1); return rc; }
    rs.steps++;
    c->steps_done++;
    if (c->diag_every > 0 && c->steps_done % c->diag_every == 0) {
      const int slot = (int)(c->diag_meta.size() % c->diag_cap);
      rc = diag_dev(c, c->diag_flag, c->diag_ring + (size_t)slot * kDiagLen);
      if (rc) { hipEventDestroy(e0); hipEventDestroy(e1); return rc; }
      nft_diag m{};
      diag_meta_fill(c, m);
      c->diag_meta.push_back(m);
      c->diag_slot.push_back(slot);
    }
  }
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  rs.ms_total = ms;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (stats) *stats = rs;
  return 0;
}

static int state_ptrs(nft_ctx *c, const char *name, double **p, int *k) {
  std::string s(name ? name : "");
  if (s == "unk") { for (int i = 0; i < 4; ++i) p[i] = c->unk[i]; *k = 4; }
  else if (s == "frc") { for (int i = 0; i < 4; ++i) p[i] = c->frc[i]; *k = 4; }
  else if (s == "uuu") { for (int i = 0; i < 3; ++i) p[i] = c->uuu[i]; *k = 3; }
  else if (s == "unkdash") { for (int i = 0; i < 3; ++i) p[i] = c->ud[i]; *k = 3; }
  else if (s == "prpot") { p[0] = c->prpot; *k = 1; }
  else return set_err(-1, "unknown state block " + s);
  return 0;
}

int nft_get_state(nft_ctx *c, const char *name, double *host, int ld) {
  if (!c || !host || ld < c->n) return set_err(-1, "nft_get_state: bad argument");
  double *p[4]; int k;
  int rc = state_ptrs(c, name, p, &k);
  if (rc) return rc;
  for (int i = 0; i < k; ++i)
    if ((rc = copy_out(c, host + (size_t)i * ld, p[i]))) return rc;
  return sync(c);
}

int nft_set_state(nft_ctx *c, const char *name, const double *host, int ld) {
  if (!c || !host || ld < c->n) return set_err(-1, "nft_set_state: bad argument");
  double *p[4]; int k;
  int rc = state_ptrs(c, name, p, &k);
  if (rc) return rc;
  for (int i = 0; i < k; ++i)
    if ((rc = copy_in(c, p[i], host + (size_t)i * ld))) return rc;
  if (std::string(name) == "uuu" && c->P.precision >= 1)
    for (int i = 0; i < 3; ++i)
      hipLaunchKernelGGL(k_to_t<float>, dim3(c->nodes_grid), dim3(NFT_BLOCK), 0, c->stream, c->n, c->uuu[i], c->uuuf[i]);
  return sync(c);
}

int nft_get_gcr_stats(nft_ctx *c, nft_gcr_stats *s) {
  if (!c || !s) return set_err(-1, "nft_get_gcr_stats: bad argument");
  *s = c->st;
  return 0;
}

int nft_kernel_timer(nft_ctx *c, int kernel) {
  if (!c || kernel < 0 || kernel > 12) return set_err(-1, "nft_kernel_timer: bad argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  t_harvest(c);
  if (c->tev.empty()) {
    c->tev.resize(1024);
    for (auto &e : c->tev) HIPCHK(hipEventCreate(&e));
    c->tex.assign(c->tev.size() / 2, 0.0);
  }
  c->timer_kernel = kernel;
  c->tev_used = 0;
  c->t_launches = 0;
  c->t_ms = 0;
  c->t_extra = 0;
  return 0;
}

// algorithmic bytes of one launch: each array read or written counted once at its width
static double kernel_bytes(nft_ctx *c, int k) {
  const double N = c->n, E = c->etot, NB = c->inc.nbp;
  const double chunks = 4 * (c->inc.nchunk + 1) + 4 * N;   // chunk pointers and degrees
  const double rest = chunks + 4 * (N + 1) + NB * (4 + 48);
  // the sweeps read the columnar tables instead of the nbr/eid slot streams when in use
  const double slots = c->col_on ? c->col_bytes : 2 * E * 8;
  const double geomg = 24 * E + slots + rest;   // FP64 metrics
  const double geomf = 12 * E + slots + rest;   // FP32 metrics
  const double geom = geomg;
  const bool mixed = c->P.precision >= 1, op32 = c->P.precision == 2;
  const double rho_b = c->nc.rh0_one ? 0.0 : 8 * N, cw_b = c->nc.coefw_uniform ? 0.0 : 8 * N;
  const double ts = mixed ? 4 : 8;
  switch (k) {
    // the FP32 operator (precision 2) reads 12-byte metrics and writes / gathers FP32 q; rh0
    // and coefw are read only when they are not uniform (1.0 / one value), ue only at the
    // boundary entries of flags 7 and 8
    case KT_LGRAD: return (op32 ? geomg - 12 * E : geomg) + 8 * N /*p*/ + 8 * N /*vol*/ + rho_b + cw_b +
                          8.0 * c->n_flag78 /*ue*/ + 3 * (op32 ? 4 : 8) * N /*q*/;
    // the divergence reads only sbb of the boundary entries (not their flags or anp)
    case KT_LDIV: return (op32 ? geomg - 12 * E : geomg) - NB * (4 + 24) + 3 * (op32 ? 4 : 8) * N /*q*/ +
                         8 * N /*vol*/ + rho_b + 8 * N /*out*/;
    case KT_LINE:
      if (c->P.precon == 3) return 5 * ts * N /*factors*/ + 8 * N /*rhs*/ + 8 * N /*pin*/ + 8 * N /*p*/;
      return c->line_L > 0 ? 3 * ts * N /*rows*/ + 8 * N /*rhs*/ + 8 * N /*p*/
                                       : 4 * (c->ncol + 1) + 3 * ts * N /*factors*/ + 8 * N + 8 * N + 2 * ts * N /*z*/;
    case KT_UPD: return 6 * 8 * N;  /* p,r read+write; x, ax read */
    case KT_DIR: return (3 + 2.0 * c->P.lord) * 8 * N;  /* qr, ar, r + up to lord x/ax; writes 2 */
    case KT_MP0: return (mixed ? geomf : geom) + 3 * ts * N + 4 * 8 * N + 4 * 8 * N + 2 * ts * N + 8 * N;
    case KT_MP0 + 1: return (mixed ? geomf : geom) + 4 * 8 * N + 12 * ts * N;
    // the flux sweep walks the owned incidence (each edge once) and reads the metrics, the
    // advector, vr, the donor-cell fields and their gradients; it writes 4 values per edge
    case KT_MPF: return (c->col_on && c->inc_own.col_L > 0 ? c->own_col_bytes : 8 * E) + chunks + (mixed ? 12 : 24) * E +
                        3 * ts * N + ts * N + 4 * 8 * N + 12 * ts * N + 4 * ts * E;
    // bounds and ratios: slots, donor-cell fields, the edge fluxes, vol and rh0; (cp, cn) out
    case KT_MP0 + 2: return slots + chunks + 4 * 8 * N + 4 * ts * E + 2 * ts * N + 8 * ts * N;
    // limited update: slots, boundary-entry ranges, the edge fluxes, (cp, cn), donor-cell
    // fields, vol and rh0; the fields out
    case KT_MP0 + 3: return slots + chunks + 4 * (N + 1) + 4 * ts * E + 8 * ts * N + 4 * 8 * N + 2 * ts * N + 4 * 8 * N;
    case KT_NET: {  // FLOPs of one forward pass over every column (the six MFMA GEMMs)
      const auto &nt = c->net;
      double fl = 0;
      for (int l = 0; l < 6; ++l) fl += 2.0 * nt.ncol * nt.K[l] * nt.N[l];
      return fl;
    }
  }
  return 0;
}

int nft_kernel_timer_read(nft_ctx *c, int *launches, double *total_ms, double *bytes_per_launch) {
  if (!c) return set_err(-1, "nft_kernel_timer_read: null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  t_harvest(c);
  if (launches) *launches = c->t_launches;
  if (total_ms) *total_ms = c->t_ms;
  if (bytes_per_launch)
    *bytes_per_launch = kernel_bytes(c, c->timer_kernel) + (c->t_launches ? c->t_extra / c->t_launches : 0.0);
  return 0;
}


// ---- learned preconditioner ------------------------------------------------------------
int nft_net_init(nft_ctx *c, int hidden, uint64_t seed, int max_batch) {
  if (!c || hidden < 32 || hidden % 32 || max_batch < 1) return set_err(-1, "nft_net_init: bad argument");
  if (c->ncol <= 0 || c->col_L == 0) return set_err(-1, "nft_net_init: needs uniform columns numbered column by column");
  if (c->per) return set_err(-1, "nft_net_init: the learned preconditioner is not available for the periodic box");
  if (c->net.ready) return set_err(-1, "nft_net_init: already initialised");
  // horizontal column graph from the edges joining different columns at one level
  std::vector<std::vector<int>> nb(c->ncol);
  for (const auto &e : c->h_hedges) {
    int a = e.first, b = e.second;
    nb[a].push_back(b);
    nb[b].push_back(a);
  }
  std::vector<int32_t> ptr(c->ncol + 1, 0), idx;
  for (int k = 0; k < c->ncol; ++k) {
    std::sort(nb[k].begin(), nb[k].end());
    nb[k].erase(std::unique(nb[k].begin(), nb[k].end()), nb[k].end());
    if (nb[k].empty()) return set_err(-1, "nft_net_init: a column without horizontal neighbours");
    ptr[k + 1] = ptr[k] + (int)nb[k].size();
    idx.insert(idx.end(), nb[k].begin(), nb[k].end());
  }
  int rc;
  if ((rc = upload(c, &c->net.nptr, ptr.data(), ptr.size())) || (rc = upload(c, &c->net.nidx, idx.data(), idx.size())))
    return rc;
  return net_alloc(c, hidden, max_batch * c->ncol, seed);
}

int nft_net_info(nft_ctx *c, int *kp, int *levels, int *ncol, int64_t *nparams) {
  if (!c || !c->net.ready) return set_err(-1, "nft_net_info: network not initialised");
  auto &n = c->net;
  int64_t cnt = 0;
  for (int l = 0; l < 6; ++l) cnt += (int64_t)n.K[l] * n.N[l] + n.N[l];
  if (kp) *kp = n.Kp;
  if (levels) *levels = n.L;
  if (ncol) *ncol = n.ncol;
  if (nparams) *nparams = cnt;
  return 0;
}

// flat parameter vector: for each layer W (K x N row-major) then b (N)
int nft_net_params(nft_ctx *c, float *host, int64_t count, int set) {
  if (!c || !c->net.ready || !host) return set_err(-1, "nft_net_params: bad argument");
  auto &n = c->net;
  int64_t off = 0;
  for (int l = 0; l < 6; ++l) {
    size_t kw = (size_t)n.K[l] * n.N[l], kb = n.N[l];
    if (off + (int64_t)(kw + kb) > count) return set_err(-1, "nft_net_params: buffer too small");
    if (set) {
      HIPCHK(hipMemcpy(n.w[2 * l], host + off, kw * sizeof(float), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(n.w[2 * l + 1], host + off + kw, kb * sizeof(float), hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_net_to_bf16, dim3((kw + 255) / 256), dim3(256), 0, c->stream, n.K[l], n.N[l], n.w[2 * l],
                         n.wb[l], n.wt[l]);
    } else {
      HIPCHK(hipMemcpy(host + off, n.w[2 * l], kw * sizeof(float), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(host + off + kw, n.w[2 * l + 1], kb * sizeof(float), hipMemcpyDeviceToHost));
    }
    off += kw + kb;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// forward (and, with T, backward) on a host batch X/T [nsamples*ncol][Kp]; outputs Y and
// the flat gradient vector (same layout as nft_net_params) when given
int nft_net_eval(nft_ctx *c, const float *X, const float *T, int nsamples, float *Y, float *grad, double *loss) {
  if (!c || !c->net.ready || !X || nsamples < 1) return set_err(-1, "nft_net_eval: bad argument");
  auto &n = c->net;
  const int M = nsamples * n.ncol;
  if (M > n.max_rows) return set_err(-1, "nft_net_eval: batch larger than max_batch");
  HIPCHK(hipMemcpyAsync(n.X, X, (size_t)M * n.Kp * sizeof(float), hipMemcpyHostToDevice, c->stream));
  int rc;
  if ((rc = net_forward(c, M))) return rc;
  if (T) {
    HIPCHK(hipMemcpyAsync(n.T, T, (size_t)M * n.Kp * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if ((rc = net_backward(c, M, loss))) return rc;
  }
  if (Y) HIPCHK(hipMemcpyAsync(Y, n.Y, (size_t)M * n.Kp * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  if (grad) {
    int64_t off = 0;
    for (int l = 0; l < 6; ++l) {
      size_t kw = (size_t)n.K[l] * n.N[l], kb = n.N[l];
      HIPCHK(hipMemcpyAsync(grad + off, n.g[2 * l], kw * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipMemcpyAsync(grad + off + kw, n.g[2 * l + 1], kb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      off += kw + kb;
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// store every stride-th (r, M_line^-1 r) pair of the GCR solves that follow, up to max_pairs
int nft_dump_pairs(nft_ctx *c, int max_pairs, int stride) {
  if (!c || !c->net.ready || max_pairs < 0 || stride < 1) return set_err(-1, "nft_dump_pairs: bad argument");
  if (max_pairs > 0 && c->P.precon < 2) return set_err(-1, "nft_dump_pairs: needs a line preconditioner");
  auto &n = c->net;
  const size_t N = c->n;
  int rc;
  if (max_pairs > n.max_pairs) {
    if ((rc = dalloc(c, &n.pR, N * max_pairs)) || (rc = dalloc(c, &n.pQ, N * max_pairs)) ||
        (rc = dalloc(c, &n.pD, N * max_pairs)))
      return rc;
    n.max_pairs = max_pairs;
    n.npairs = 0;
    n.ps.clear();
    n.pr2.clear();
    n.pimp.clear();
  }
  n.stride = stride;
  n.seen = 0;
  n.dumping = max_pairs > 0;
  return 0;
}

// pair store: count, and copies of pairs [first, first+n_get) (r, q: n_get x npoin; s, r2)
int nft_pairs(nft_ctx *c, int *count, int first, int n_get, double *r, double *q, double *s, double *r2) {
  if (!c || !c->net.ready) return set_err(-1, "nft_pairs: network not initialised");
  auto &n = c->net;
  const size_t N = c->n;
  if (n_get > 0) {
    if (first < 0 || first + n_get > n.npairs || !r || !q || !s || !r2) return set_err(-1, "nft_pairs: range");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(r, n.pR + N * first, N * n_get * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(q, n.pQ + N * first, N * n_get * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < n_get; ++k) { s[k] = n.ps[first + k]; r2[k] = n.pr2[first + k]; }
  }
  if (count) *count = n.npairs;
  return 0;
}

// Physics-informed loss of one minibatch. For sample b with pair (r, q, s, a = |r|^2),
// z_b = q + s N(r/s), v = L z_b, b = <v,v>, c = <r,v>:
//   objective 0 (residual): l_b = |v - r|^2 / a = (b - 2c + a)/a,  dl/dv = 2(v - r)/a;
//   objective 1 (angle):    l_b = 1 - c^2/(a b),                 dl/dv = (2c/(ab)) ((c/b) v - r);
// loss = mean over the B samples, dloss/dY_b = s L^T(dl_b/dv) / B.
// Forward of the network on the B samples must already be in n.Y.
static int net_physics_grad(nft_ctx *c, const int *idx, int B, double *loss) {
  auto &n = c->net;
  const int N = c->n;
  const size_t per = (size_t)n.ncol * n.Kp;
  const int nb = (N + 255) / 256;
  int rc;
  double total = 0;
  std::vector<double> part(2 * (size_t)nb);
  for (int bi = 0; bi < B; ++bi) {
    const int k = idx[bi];
    const double s = n.ps[k], a = n.pr2[k];
    const double *r = n.pR + (size_t)N * k, *q = n.pQ + (size_t)N * k;
    hipLaunchKernelGGL(k_net_z, dim3(nb), dim3(256), 0, c->stream, n.ncol, n.L, n.Kp, q, n.Y + per * bi, s, n.tz);
    if (n.objective == 2) {
      // error objective: l = |z - d|^2/|d|^2 with d = x_k - x_final, dl/dz = 2(z - d)/|d|^2
      const double *d = n.pD + (size_t)N * k;
      hipLaunchKernelGGL(k_net_axpby, dim3(nb), dim3(256), 0, c->stream, N, 1.0, n.tz, -1.0, d, n.tLz);
      hipLaunchKernelGGL(k_net_dots, dim3(nb), dim3(256), 0, c->stream, N, n.tLz, d, n.tpart);
      HIPCHK(hipMemcpyAsync(part.data(), n.tpart, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      hipLaunchKernelGGL(k_net_dots, dim3(nb), dim3(256), 0, c->stream, N, d, d, n.tpart + 2 * nb);
      std::vector<double> part2(2 * (size_t)nb);
      HIPCHK(hipMemcpyAsync(part2.data(), n.tpart + 2 * nb, 2 * nb * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      double ee = 0, dd = 0;
      for (int i = 0; i < nb; ++i) { ee += part[2 * i]; dd += part2[2 * i]; }
      if (!(dd > 0)) return set_err(-1, "net training: a pair with a zero error target");
      total += ee / dd;
      hipLaunchKernelGGL(k_net_axpby, dim3(nb), dim3(256), 0, c->stream, N, 2.0 / (dd * B), n.tLz, 0.0, d, n.ty);
      hipLaunchKernelGGL(k_net_dY, dim3((per + 255) / 256), dim3(256), 0, c->stream, n.ncol, n.L, n.Kp, n.ty, s,
                         n.dY + per * bi);
      HIPCHK(hipGetLastError());
      continue;
    }
    if ((rc = laplc_dev(c, n.tz, n.tLz, n.pimp[k], nullptr))) return rc;
    hipLaunchKernelGGL(k_net_dots, dim3(nb), dim3(256), 0, c->stream, N, n.tLz, r, n.tpart);
    HIPCHK(hipMemcpyAsync(part.data(), n.tpart, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double bb = 0, cc = 0;
    for (int i = 0; i < nb; ++i) { bb += part[2 * i]; cc += part[2 * i + 1]; }
    double alpha, beta, l;
    if (n.objective == 1 && bb > 0) {
      l = 1.0 - cc * cc / (a * bb);
      alpha = 2.0 * cc * cc / (a * bb * bb);
      beta = -2.0 * cc / (a * bb);
    } else {
      l = (bb - 2.0 * cc + a) / a;
      alpha = 2.0 / a;
      beta = -2.0 / a;
    }
    total += l;
    hipLaunchKernelGGL(k_net_axpby, dim3(nb), dim3(256), 0, c->stream, N, alpha / B, n.tLz, beta / B, r, n.ty);
    if ((rc = laplt_dev(c, n.ty, n.tg, n.pimp[k]))) return rc;
    hipLaunchKernelGGL(k_net_dY, dim3((per + 255) / 256), dim3(256), 0, c->stream, n.ncol, n.L, n.Kp, n.tg, s,
                       n.dY + per * bi);
    HIPCHK(hipGetLastError());
  }
  if (loss) *loss = total / B;
  return 0;
}

static int net_train_alloc(nft_ctx *c) {
  auto &n = c->net;
  if (n.tz) return 0;
  const size_t N = c->n;
  int rc;
  if ((rc = dalloc(c, &n.tz, N)) || (rc = dalloc(c, &n.tLz, N)) || (rc = dalloc(c, &n.ty, N)) ||
      (rc = dalloc(c, &n.tg, N)) || (rc = dalloc(c, &n.tpart, 4 * ((N + 255) / 256))))
    return rc;
  return 0;
}

// Forward on the stored pairs idx[0..B) (network inputs r/s)
static int net_forward_pairs(nft_ctx *c, const int *idx, int B) {
  auto &n = c->net;
  const size_t per = (size_t)n.ncol * n.Kp;
  for (int b = 0; b < B; ++b) {
    const int k = idx[b];
    hipLaunchKernelGGL(k_net_input_s, dim3((per + 255) / 256), dim3(256), 0, c->stream, n.ncol, n.L, n.Kp,
                       n.pR + (size_t)c->n * k, n.ps[k], n.X + per * b);
  }
  HIPCHK(hipGetLastError());
  return net_forward(c, B * n.ncol);
}

// physics-informed loss over the stored pairs [first, first+count) with the current
// network (the zero-network value is |L q - r|^2/|r|^2); with grad (count <= max_batch),
// also the flat parameter gradient of that loss (layout of nft_net_params)
int nft_net_pair_loss(nft_ctx *c, int first, int count, double *loss, float *grad) {
  if (!c || !c->net.ready || !loss || count < 1) return set_err(-1, "nft_net_pair_loss: bad argument");
  auto &n = c->net;
  if (first < 0 || first + count > n.npairs) return set_err(-1, "nft_net_pair_loss: range");
  const int maxb = n.max_rows / n.ncol;
  if (grad && count > maxb) return set_err(-1, "nft_net_pair_loss: gradient needs count <= max_batch");
  int rc;
  if ((rc = net_train_alloc(c))) return rc;
  double acc = 0;
  for (int b0 = 0; b0 < count; b0 += maxb) {
    int B = std::min(maxb, count - b0);
    std::vector<int> idx(B);
    for (int b = 0; b < B; ++b) idx[b] = first + b0 + b;
    double l = 0;
    if ((rc = net_forward_pairs(c, idx.data(), B)) || (rc = net_physics_grad(c, idx.data(), B, &l))) return rc;
    acc += l * B;
    if (grad && (rc = net_backprop(c, B * n.ncol))) return rc;
  }
  *loss = acc / count;
  if (grad) {
    int64_t off = 0;
    for (int l = 0; l < 6; ++l) {
      size_t kw = (size_t)n.K[l] * n.N[l], kb = n.N[l];
      HIPCHK(hipMemcpyAsync(grad + off, n.g[2 * l], kw * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipMemcpyAsync(grad + off + kw, n.g[2 * l + 1], kb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      off += kw + kb;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

// Adam on the stored pairs [0, use_pairs) (all when use_pairs <= 0) with the physics-informed
// loss: `steps` minibatches of `batch` pairs drawn with a fixed-seed generator; losses of the
// first and last minibatch returned
// the whole training state: weights (nft_net_params' layout), Adam's first and second
// moments in the same layout, then the step count
int nft_net_state(nft_ctx *c, float *host, int64_t count, int set) {
  if (!c || !c->net.ready || !host) return set_err(-1, "nft_net_state: bad argument");
  auto &n = c->net;
  int64_t P = 0;
  for (int l = 0; l < 6; ++l) P += (int64_t)n.K[l] * n.N[l] + n.N[l];
  if (count < 3 * P + 1) return set_err(-1, "nft_net_state: buffer too small (3 * nparams + 1 floats)");
  int rc;
  if ((rc = nft_net_params(c, host, P, set))) return rc;
  int64_t off = P;
  for (int part = 0; part < 2; ++part) {
    float **arr = part == 0 ? n.m : n.v;
    for (int l = 0; l < 6; ++l) {
      for (int wb = 0; wb < 2; ++wb) {
        size_t k = wb == 0 ? (size_t)n.K[l] * n.N[l] : (size_t)n.N[l];
        if (set) HIPCHK(hipMemcpy(arr[2 * l + wb], host + off, k * sizeof(float), hipMemcpyHostToDevice));
        else HIPCHK(hipMemcpy(host + off, arr[2 * l + wb], k * sizeof(float), hipMemcpyDeviceToHost));
        off += (int64_t)k;
      }
    }
  }
  if (set) {
    if (!(host[off] >= 0.f) || host[off] > 1e9f) return set_err(-1, "nft_net_state: bad step count");
    n.step = (int)host[off];
  } else {
    host[off] = (float)n.step;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int nft_net_train(nft_ctx *c, int steps, int batch, double lr, uint64_t seed, int use_pairs, double *loss_first,
                  double *loss_last) {
  if (!c || !c->net.ready || steps < 1 || batch < 1) return set_err(-1, "nft_net_train: bad argument");
  auto &n = c->net;
  const int np = use_pairs > 0 ? std::min(use_pairs, n.npairs) : n.npairs;
  if (np < 1) return set_err(-1, "nft_net_train: no training pairs");
  if (batch * n.ncol > n.max_rows) return set_err(-1, "nft_net_train: batch larger than max_batch");
  const int M = batch * n.ncol;
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 7;
  auto next = [&]() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; };
  int rc;
  if ((rc = net_train_alloc(c))) return rc;
  std::vector<int> idx(batch);
  for (int it = 0; it < steps; ++it) {
    for (int b = 0; b < batch; ++b) idx[b] = (int)(next() % (uint64_t)np);
    if ((rc = net_forward_pairs(c, idx.data(), batch))) return rc;
    double loss = 0;
    bool want = (it == 0 && loss_first) || (it == steps - 1 && loss_last);
    if ((rc = net_physics_grad(c, idx.data(), batch, want ? &loss : nullptr))) return rc;
    if ((rc = net_backprop(c, M))) return rc;
    if (it == 0 && loss_first) *loss_first = loss;
    if (it == steps - 1 && loss_last) *loss_last = loss;
    n.step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)n.step), c2 = 1.0f - std::pow(b2, (float)n.step);
    for (int l = 0; l < 6; ++l) {
      size_t kw = (size_t)n.K[l] * n.N[l], kb = n.N[l];
      hipLaunchKernelGGL(k_net_adam, dim3((kw + 255) / 256), dim3(256), 0, c->stream, (int)kw, n.K[l], n.N[l],
                         n.w[2 * l], n.g[2 * l], n.m[2 * l], n.v[2 * l], (float)lr, b1, b2, c1, c2, 1e-8f, n.wb[l],
                         n.wt[l]);
      if (!n.linear)
        hipLaunchKernelGGL(k_net_adam, dim3((kb + 255) / 256), dim3(256), 0, c->stream, (int)kb, 1, (int)kb,
                           n.w[2 * l + 1], n.g[2 * l + 1], n.m[2 * l + 1], n.v[2 * l + 1], (float)lr, b1, b2, c1,
                           c2, 1e-8f, (uint16_t *)nullptr, (uint16_t *)nullptr);
    }
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// device copy of one node array (read 8N + write 8N bytes), for calibrating byte counters
int nft_copy_calibration(nft_ctx *c, int reps, double *ms) {
  if (!c || reps < 1) return set_err(-1, "nft_copy_calibration: bad argument");
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy, dim3(c->nodes_grid), dim3(NFT_BLOCK), 0, c->stream, c->n, c->io[0], c->io[1],
                       (const double *)nullptr);
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t / reps;
  return 0;
}

// streaming reference: a copy of `bytes` between two fresh device buffers, 16 bytes per
// lane per load, grid-stride; mean ms per copy over `reps` after one untimed copy
int nft_stream_calibration(nft_ctx *c, int64_t bytes, int reps, double *ms) {
  if (!c || reps < 1 || bytes < 16) return set_err(-1, "nft_stream_calibration: bad argument");
  const size_t n2 = (size_t)bytes / 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n2 * 16) != hipSuccess) return set_err(-2, "nft_stream_calibration: alloc");
  if (hipMalloc(&b, n2 * 16) != hipSuccess) { (void)hipFree(a); return set_err(-2, "nft_stream_calibration: alloc"); }
  int rc = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float t = 0;
  const int grid = 8 * 256 * 4;   // a few wavefronts per SIMD on every CU, grid-stride
  if (hipMemsetAsync(a, 0, n2 * 16, c->stream) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = set_err(-2, "nft_stream_calibration: setup");
  } else {
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(NFT_BLOCK), 0, c->stream, (const double2 *)a, (double2 *)b, n2);
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(NFT_BLOCK), 0, c->stream, (const double2 *)a, (double2 *)b, n2);
    (void)hipEventRecord(e1, c->stream);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)
      rc = set_err(-2, "nft_stream_calibration: timing");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(a);
  (void)hipFree(b);
  if (!rc && ms) *ms = t / reps;
  return rc;
}

int nft_set_mixed_drop(nft_ctx *c, double drop, double *current) {
  if (!c) return set_err(-1, "nft_set_mixed_drop: null context");
  if (drop > 0) {
    if (!(drop < 1)) return set_err(-1, "nft_set_mixed_drop: drop must lie in (0, 1)");
    c->op32_drop = drop;
  }
  if (current) *current = c->op32_drop;
  return 0;
}

// training objective (0 residual, 1 angle, 2 error) and activation (linear = 1: identity activations
// with the biases held at zero, so the correction is a fixed linear map of r)
int nft_net_options(nft_ctx *c, int objective, int linear) {
  if (!c || !c->net.ready || objective < 0 || objective > 2 || linear < 0 || linear > 1)
    return set_err(-1, "nft_net_options: bad argument");
  auto &n = c->net;
  n.objective = objective;
  n.linear = linear;
  if (linear)
    for (int l = 0; l < 6; ++l) HIPCHK(hipMemsetAsync(n.w[2 * l + 1], 0, n.N[l] * sizeof(float), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int nft_diagnostics(nft_ctx *c, int courant_flag, nft_diag *out) {
  if (!c || !out || courant_flag < 1 || courant_flag > 3) return set_err(-1, "nft_diagnostics: bad argument");
  int rc;
  if (!c->diag_now && (rc = dalloc(c, &c->diag_now, kDiagLen))) return rc;
  if ((rc = diag_dev(c, courant_flag, c->diag_now))) return rc;
  double r[kDiagLen];
  HIPCHK(hipMemcpyAsync(r, c->diag_now, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *out = nft_diag{};
  diag_unpack(r, *out);
  diag_meta_fill(c, *out);
  return 0;
}

// courant(uuu,cour,iflag) and stats(unk) on host arrays: the arguments are staged in the
// io buffers, so the device-resident time-step state is left as it is
static int diag_host(nft_ctx *c, int flag, const double *const *cols, int ncols, int parts, double *cour,
                     nft_diag *out) {
  int rc;
  if (!c->diag_now && (rc = dalloc(c, &c->diag_now, kDiagLen))) return rc;
  for (int k = 0; k < ncols; ++k)
    if ((rc = copy_in(c, c->io[k], cols[k]))) return rc;
  HIPCHK(hipMemsetAsync(c->diag_now, 0, kDiagLen * sizeof(double), c->stream));
  if ((rc = diag_dev_of(c, flag, c->diag_now, c->io, c->io, parts))) return rc;
  double r[kDiagLen];
  HIPCHK(hipMemcpyAsync(r, c->diag_now, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  if (cour && (rc = copy_out(c, cour, c->t[5]))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  *out = nft_diag{};
  diag_unpack(r, *out);
  diag_meta_fill(c, *out);
  return 0;
}

int nft_courant(nft_ctx *c, const double *uuu, int ld, double *cour, int iflag, nft_diag *out) {
  if (!c || !uuu || !out || ld < c->n || iflag < 1 || iflag > 3) return set_err(-1, "nft_courant: bad argument");
  const double *cols[3] = {uuu, uuu + (size_t)ld, uuu + 2 * (size_t)ld};
  return diag_host(c, iflag, cols, 3, 2, cour, out);
}

int nft_stats(nft_ctx *c, const double *unk, int ld, nft_diag *out) {
  if (!c || !unk || !out || ld < c->n) return set_err(-1, "nft_stats: bad argument");
  const double *cols[4] = {unk, unk + (size_t)ld, unk + 2 * (size_t)ld, unk + 3 * (size_t)ld};
  return diag_host(c, 1, cols, 4, 4, nullptr, out);
}

int nft_set_diagnostics(nft_ctx *c, int every, int capacity, int courant_flag) {
  if (!c || every < 0 || (every > 0 && capacity < 1) || courant_flag < 1 || courant_flag > 3)
    return set_err(-1, "nft_set_diagnostics: bad argument");
  if (every > 0 && capacity > c->diag_cap) {
    int rc;
    if ((rc = dalloc(c, &c->diag_ring, (size_t)capacity * kDiagLen))) return rc;
    c->diag_cap = capacity;
  }
  c->diag_every = every;
  c->diag_flag = courant_flag;
  c->diag_meta.clear();
  c->diag_slot.clear();
  return 0;
}

int nft_get_diagnostics(nft_ctx *c, int *count, nft_diag *out, int max) {
  if (!c || !count || (max > 0 && !out)) return set_err(-1, "nft_get_diagnostics: bad argument");
  const int total = (int)c->diag_meta.size();
  const int first = std::max(0, total - c->diag_cap);   // older records were overwritten
  const int n = std::min(total - first, std::max(max, 0));
  *count = n;
  if (n > 0) {
    std::vector<double> ring((size_t)c->diag_cap * kDiagLen);
    HIPCHK(hipMemcpyAsync(ring.data(), c->diag_ring, ring.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < n; ++k) {
      out[k] = c->diag_meta[first + k];
      diag_unpack(ring.data() + (size_t)c->diag_slot[first + k] * kDiagLen, out[k]);
    }
  }
  c->diag_meta.clear();
  c->diag_slot.clear();
  return 0;
}

// gradient / divergence gathers: 1 columnar incidence (when the mesh has one), 0 SELL-64
int nft_set_incidence(nft_ctx *c, int columnar, int *available) {
  if (!c) return set_err(-1, "nft_set_incidence: null context");
  if (available) *available = c->inc.col_L > 0;
  if (columnar < 0) return 0;
  if (columnar && c->inc.col_L == 0) return set_err(-1, "nft_set_incidence: mesh has no columnar form");
  c->col_on = columnar ? 1 : 0;
  return 0;
}

int nft_set_learned(nft_ctx *c, int kind, double stall_ratio, int *fallbacks, int *learned_iters) {
  if (!c) return set_err(-1, "nft_set_learned: null context");
  if (kind == 1 && !c->net.ready) return set_err(-1, "nft_set_learned: network not initialised");
  if (kind == 1 && c->P.precon < 2) return set_err(-1, "nft_set_learned: needs a line preconditioner");
  if (kind >= 0) { c->net.kind = kind; c->net.stall = stall_ratio; }
  if (fallbacks) *fallbacks = c->net.fallbacks;
  if (learned_iters) *learned_iters = c->net.learned_iters;
  return 0;
}


// ---- full-precision state and mesh files ----------------------------------------------
// State file (replaces the reference's f14.6 text restart, Restart_Out_old/Restart_In_old
// :4779-4890, which loses digits): "NFTSTAT1", int64 npoin, steps, nitsm, icount, then
// the step-boundary state as little-endian float64 blocks unk(4), frc(4), uuu(3),
// unkdash(3), prpot(1), each [k][npoin]. Resuming from it continues the run bit for bit.
// ---- multigrid branch of precon ---------------------------------------------------------
int nft_set_multigrid(nft_ctx *c, const nft_mesh *m1, const int32_t *match1, const nft_mesh *m2,
                      const int32_t *match2, const nft_multigrid *mgp) {
  if (!c || !mgp) return set_err(-1, "nft_set_multigrid: null argument");
  if (c->per) return set_err(-1, "nft_set_multigrid: the coarse levels have no periodic closures here");
  if (mgp->itrp2 < 0 || mgp->itrp_c1 < 0 || mgp->itrp_c2 < 0 || mgp->itrp_cc < 0)
    return set_err(-1, "nft_set_multigrid: negative sweep count");
  if (mgp->transfer != MG_AS_REFERENCE && mgp->transfer != MG_NEIGHBOURS)
    return set_err(-1, "nft_set_multigrid: transfer must be 0 or 1");
  if (!m1 != !m2 || !m1 != !match1 || !m2 != !match2)
    return set_err(-1, "nft_set_multigrid: give both coarse meshes and both matches, or none");
  if (!m1) {
    if (mgp->mg_flag && !c->mg_ready) return set_err(-1, "nft_set_multigrid: no coarse levels set");
    if (mgp->mg_flag && mgp->transfer != c->mg_transfer)
      return set_err(-1, "nft_set_multigrid: changing the transfer needs the coarse meshes again");
  } else {
    if (c->P.precon == 0) return set_err(-1, "nft_set_multigrid: precon 0 has no multigrid branch");
    const nft_mesh *ms[2] = {m1, m2};
    const int32_t *mt[2] = {match1, match2};
    const int nfine[2] = {c->n, m1->npoin};
    for (int l = 0; l < 2; ++l) {
      int rc = validate_mesh(ms[l], "nft_set_multigrid");
      if (rc) return rc;
      // the coarse levels index the fine mesh's coefw and ue by their own node ids, as the
      // reference does (laplc_c, :1387-1392)
      if (ms[l]->npoin > nfine[l]) return set_err(-1, "nft_set_multigrid: coarse mesh larger than its parent");
      std::vector<char> seen(nfine[l], 0);
      for (int i = 0; i < ms[l]->npoin; ++i) {
        int32_t f = mt[l][i];
        if (f < 1 || f > nfine[l]) return set_err(-1, "nft_set_multigrid: match entry out of range at " + std::to_string(i + 1));
        if (seen[f - 1]) return set_err(-1, "nft_set_multigrid: two coarse nodes match fine node " + std::to_string(f));
        seen[f - 1] = 1;
      }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->mg_on = 0; c->mg_ready = 0;
    int rc;
    std::vector<double> rh0(c->n);
    HIPCHK(hipMemcpy(rh0.data(), c->rh0, c->n * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int32_t> chain(m1->npoin);   // fine node of each level-1 node
    for (int l = 0; l < 2; ++l) {
      auto &L = c->mg[l];
      const nft_mesh *cm = ms[l];
      const int nc = cm->npoin, nf = nfine[l];
      L = nft_ctx::MgLevel{};
      L.n = nc;
      Geometry G;
      if ((rc = build_geometry(c, cm, G))) return rc;
      L.inc = G.inc; L.S = G.S; L.gather_grid = G.gather_grid; L.col_on = G.col_on;
      L.nodes_grid = blocks_for(nc);
      std::vector<int32_t> match(nc), imatch(nf, -1);
      for (int i = 0; i < nc; ++i) { match[i] = mt[l][i] - 1; imatch[match[i]] = i; }
      // density of the matched fine node (laplc_c :1419-1426, laplc_cc :1512-1518)
      std::vector<double> rho(nc);
      for (int i = 0; i < nc; ++i) {
        int f = l == 0 ? match[i] : chain[match[i]];
        if (l == 0) chain[i] = f;
        rho[i] = rh0[f];
      }
      std::vector<double> vol(cm->vol, cm->vol + nc);
      if ((rc = upload(c, &L.match, match.data(), nc)) || (rc = upload(c, &L.imatch, imatch.data(), nf)) ||
          (rc = upload(c, &L.rho, rho.data(), nc)) || (rc = upload(c, &L.vol, vol.data(), nc)) ||
          (rc = dalloc(c, &L.pstep, nc)) || (rc = dalloc(c, &L.p, nc)) || (rc = dalloc(c, &L.r, nc)) ||
          (rc = dalloc(c, &L.rrhs, nc)) || (rc = dalloc(c, &L.rout, nc)) || (rc = dalloc(c, &L.e, nc)))
        return rc;
      L.nc = c->nc;
      L.nc.vol = L.vol;
      L.nc.rh0 = L.rho;
      L.nc.rh0_one = std::all_of(rho.begin(), rho.end(), [](double v) { return v == 1.0; }) ? 1 : 0;
      // coefw and ue stay the fine arrays (indexed by the coarse node id)
      L.beta = beta_host(cm, c->P.dom);
    }
    // compute_d_unstr(2), (3) (:1181-1199) with the density restricted from the fine mesh
    c->mg_transfer = mgp->transfer;
    {
      auto &A = c->mg[0], &B = c->mg[1];
      const dim3 b(NFT_BLOCK);
      hipLaunchKernelGGL(k_mg_restrict, dim3(A.nodes_grid), b, 0, c->stream, A.n, A.match, c->inc, c->rh0, A.r,
                         mgp->transfer, nullptr);
      hipLaunchKernelGGL(k_mg_restrict, dim3(B.nodes_grid), b, 0, c->stream, B.n, B.match, A.inc, A.r, B.r,
                         mgp->transfer, nullptr);
      HIPCHK(hipGetLastError());
      std::vector<double> r1(A.n), r2(B.n);
      HIPCHK(hipMemcpyAsync(r1.data(), A.r, A.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipMemcpyAsync(r2.data(), B.r, B.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      compute_d_host(m1, r1.data(), c->P.dt, A.h_d);
      compute_d_host(m2, r2.data(), c->P.dt, B.h_d);
      for (auto *L : {&A, &B}) {
        std::vector<double> ps;
        pstep_host(c->P.sm_flag, L->h_d, L->beta, ps);
        HIPCHK(hipMemcpy(L->pstep, ps.data(), ps.size() * sizeof(double), hipMemcpyHostToDevice));
      }
    }
    if (!c->mg_rrhs) {
      if ((rc = dalloc(c, &c->mg_rrhs, c->n)) || (rc = dalloc(c, &c->mg_e, c->n))) return rc;
    }
    c->mg_ready = 1;
  }
  c->mg_itrp2 = mgp->itrp2; c->mg_c1 = mgp->itrp_c1; c->mg_c2 = mgp->itrp_c2; c->mg_cc = mgp->itrp_cc;
  c->mg_omega = mgp->omega_c > 0 ? mgp->omega_c : c->P.omega;
  c->mg_on = mgp->mg_flag ? 1 : 0;
  return 0;
}

int nft_multigrid_field(nft_ctx *c, int level, const char *name, double *host, int *count) {
  if (!c || !name || level < 1 || level > 2) return set_err(-1, "nft_multigrid_field: bad argument");
  if (!c->mg_ready) return set_err(-1, "nft_multigrid_field: no coarse levels set");
  const auto &L = c->mg[level - 1];
  const std::string s(name);
  if (count) *count = L.n;
  if (s == "d") {
    if (host) std::copy(L.h_d.begin(), L.h_d.end(), host);
  } else if (s == "beta") {
    if (count) *count = 1;
    if (host) host[0] = L.beta;
  } else if (s == "rh0" || s == "pstep") {
    if (host) {
      HIPCHK(hipStreamSynchronize(c->stream));
      HIPCHK(hipMemcpy(host, s == "rh0" ? L.rho : L.pstep, L.n * sizeof(double), hipMemcpyDeviceToHost));
    }
  } else {
    return set_err(-1, "nft_multigrid_field: unknown field " + s);
  }
  return 0;
}

// restriction_unstr / prolongation_unstr / laplc_c / laplc_cc on host arrays
int nft_multigrid_op(nft_ctx *c, int op, int level, const double *in, double *out, int impfl) {
  if (!c || !in || !out || level < 1 || level > 2) return set_err(-1, "nft_multigrid_op: bad argument");
  if (!c->mg_ready) return set_err(-1, "nft_multigrid_op: no coarse levels set");
  auto &L = c->mg[level - 1];
  const int nfine = level == 1 ? c->n : c->mg[0].n;
  const Incidence &fine = level == 1 ? c->inc : c->mg[0].inc;
  const dim3 b(NFT_BLOCK);
  double *din = c->io[0], *dout = c->io[1];
  int nin, nout;
  switch (op) {
    case 0: nin = nfine; nout = L.n; break;    // restriction
    case 1: nin = L.n; nout = nfine; break;    // prolongation
    case 2: nin = L.n; nout = L.n; break;      // laplc on the level
    default: return set_err(-1, "nft_multigrid_op: op must be 0, 1 or 2");
  }
  HIPCHK(hipMemcpyAsync(din, in, nin * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (op == 0)
    hipLaunchKernelGGL(k_mg_restrict, dim3(L.nodes_grid), b, 0, c->stream, L.n, L.match, fine, din, dout,
                       c->mg_transfer, nullptr);
  else if (op == 1)
    hipLaunchKernelGGL(k_mg_prolong, dim3(blocks_for(nfine)), b, 0, c->stream, nfine, L.imatch, fine, din, dout,
                       c->mg_transfer, nullptr);
  else {
    int rc = laplc_level(c, L, din, dout, impfl);
    if (rc) return rc;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static const char kStateMagic[8] = {'N', 'F', 'T', 'S', 'T', 'A', 'T', '1'};
static const char *kStateBlocks[5] = {"unk", "frc", "uuu", "unkdash", "prpot"};
static const int kStateK[5] = {4, 4, 3, 3, 1};

int nft_save_state(nft_ctx *c, const char *path) {
  if (!c || !path) return set_err(-1, "nft_save_state: bad argument");
  const size_t n = c->n;
  std::vector<double> buf(4 * n);
  FILE *f = std::fopen(path, "wb");
  if (!f) return set_err(-1, std::string("nft_save_state: cannot open ") + path);
  int64_t hdr[4] = {(int64_t)n, c->steps_done, (int64_t)c->st.nitsm, (int64_t)c->st.icount};
  bool ok = std::fwrite(kStateMagic, 1, 8, f) == 8 && std::fwrite(hdr, sizeof(int64_t), 4, f) == 4;
  for (int b = 0; b < 5 && ok; ++b) {
    int rc = nft_get_state(c, kStateBlocks[b], buf.data(), (int)n);
    if (rc) { std::fclose(f); return rc; }
    ok = std::fwrite(buf.data(), sizeof(double), kStateK[b] * n, f) == kStateK[b] * n;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : set_err(-1, std::string("nft_save_state: write failed for ") + path);
}

int nft_load_state(nft_ctx *c, const char *path) {
  if (!c || !path) return set_err(-1, "nft_load_state: bad argument");
  const size_t n = c->n;
  FILE *f = std::fopen(path, "rb");
  if (!f) return set_err(-1, std::string("nft_load_state: cannot open ") + path);
  char magic[8];
  int64_t hdr[4];
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kStateMagic, 8) != 0 ||
      std::fread(hdr, sizeof(int64_t), 4, f) != 4) {
    std::fclose(f);
    return set_err(-1, std::string("nft_load_state: not a state file: ") + path);
  }
  if (hdr[0] != (int64_t)n) {
    std::fclose(f);
    return set_err(-1, "nft_load_state: state is for " + std::to_string(hdr[0]) + " nodes, mesh has " +
                           std::to_string(n));
  }
  std::vector<double> buf(4 * n);
  for (int b = 0; b < 5; ++b) {
    if (std::fread(buf.data(), sizeof(double), kStateK[b] * n, f) != kStateK[b] * n) {
      std::fclose(f);
      return set_err(-1, std::string("nft_load_state: truncated file ") + path);
    }
    int rc = nft_set_state(c, kStateBlocks[b], buf.data(), (int)n);
    if (rc) { std::fclose(f); return rc; }
  }
  std::fclose(f);
  c->steps_done = hdr[1];
  c->st.nitsm = (int)hdr[2];
  c->st.icount = (int)hdr[3];
  return 0;
}

int nft_steps_done(nft_ctx *c, int64_t *steps) {
  if (!c || !steps) return set_err(-1, "nft_steps_done: bad argument");
  *steps = c->steps_done;
  return 0;
}

// Mesh file: binary form of the reference's mesh3d.d (:2343-2501), with the column table:
// "NFTMESH1", int64 npoin, nedge, nbface, nboun, ncol, then iedge(nedge,2) int32,
// sn(nedge,3) float64, kbface(nbface,2) int32, sb(nbface,3), ibpoin(nboun,2) int32,
// sbb(nboun,3), anp(nboun,3), coord(npoin,3), vol(npoin), col_start(ncol+1) int32, all
// column-major with leading dimension = count and 1-based node ids as in the reference.
// Reading it is a few bulk reads instead of parsing ~10^8 text numbers.
static const char kMeshMagic[8] = {'N', 'F', 'T', 'M', 'E', 'S', 'H', '1'};

}  // extern "C"

namespace {
struct OwnedMesh {
  nft_mesh m{};
  std::vector<int32_t> iedge, kbface, ibpoin, col;
  std::vector<double> sn, sb, sbb, anp, coord, vol;
};

template <typename T>
bool write_cols(FILE *f, const T *a, int ld, int count, int k) {
  for (int j = 0; j < k; ++j)
    if (count > 0 && std::fwrite(a + (size_t)j * ld, sizeof(T), count, f) != (size_t)count) return false;
  return true;
}
}  // namespace

extern "C" {

int nft_mesh_save(const char *path, const nft_mesh *m) {
  if (!path || !m || m->npoin <= 0 || m->nedge < 0 || m->nbface < 0 || m->nboun < 0)
    return set_err(-1, "nft_mesh_save: bad argument");
  FILE *f = std::fopen(path, "wb");
  if (!f) return set_err(-1, std::string("nft_mesh_save: cannot open ") + path);
  const int ncol = m->ncol > 0 && m->col_start ? m->ncol : 0;
  int64_t hdr[5] = {m->npoin, m->nedge, m->nbface, m->nboun, ncol};
  bool ok = std::fwrite(kMeshMagic, 1, 8, f) == 8 && std::fwrite(hdr, sizeof(int64_t), 5, f) == 5 &&
            write_cols(f, m->iedge, m->ld_edge, m->nedge, 2) && write_cols(f, m->sn, m->ld_edge, m->nedge, 3) &&
            write_cols(f, m->kbface, m->ld_face, m->nbface, 2) && write_cols(f, m->sb, m->ld_face, m->nbface, 3) &&
            write_cols(f, m->ibpoin, m->ld_face, m->nboun, 2) && write_cols(f, m->sbb, m->ld_face, m->nboun, 3) &&
            write_cols(f, m->anp, m->ld_face, m->nboun, 3) && write_cols(f, m->coord, m->ld_poin, m->npoin, 3) &&
            write_cols(f, m->vol, m->ld_poin, m->npoin, 1) &&
            (ncol == 0 || write_cols(f, m->col_start, ncol + 1, ncol + 1, 1));
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : set_err(-1, std::string("nft_mesh_save: write failed for ") + path);
}

int nft_mesh_load(const char *path, nft_mesh **out) {
  if (!path || !out) return set_err(-1, "nft_mesh_load: bad argument");
  *out = nullptr;
  FILE *f = std::fopen(path, "rb");
  if (!f) return set_err(-1, std::string("nft_mesh_load: cannot open ") + path);
  char magic[8];
  int64_t h[5];
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kMeshMagic, 8) != 0 ||
      std::fread(h, sizeof(int64_t), 5, f) != 5 || h[0] <= 0 || h[0] > 2000000000LL || h[1] < 0 ||
      h[1] > 2000000000LL || h[2] < 0 || h[2] > 2000000000LL || h[3] < 0 || h[3] > 2000000000LL || h[4] < 0 ||
      h[4] > h[0]) {
    std::fclose(f);
    return set_err(-1, std::string("nft_mesh_load: not a mesh file: ") + path);
  }
  auto *o = new OwnedMesh();
  const size_t N = h[0], E = h[1], B = h[2], P = h[3], C = h[4];
  auto rd = [&](auto &v, size_t cnt) {
    v.resize(cnt);
    return cnt == 0 || std::fread(v.data(), sizeof(v[0]), cnt, f) == cnt;
  };
  bool ok = rd(o->iedge, 2 * E) && rd(o->sn, 3 * E) && rd(o->kbface, 2 * B) && rd(o->sb, 3 * B) &&
            rd(o->ibpoin, 2 * P) && rd(o->sbb, 3 * P) && rd(o->anp, 3 * P) && rd(o->coord, 3 * N) &&
            rd(o->vol, N) && (C == 0 || rd(o->col, C + 1));
  std::fclose(f);
  if (!ok) {
    delete o;
    return set_err(-1, std::string("nft_mesh_load: truncated file ") + path);
  }
  nft_mesh &m = o->m;
  m.npoin = (int)N; m.nedge = (int)E; m.nbface = (int)B; m.nboun = (int)P;
  m.ld_poin = (int)N; m.ld_edge = (int)E; m.ld_face = (int)std::max(B, P);
  // the reference's layout shares one leading dimension for face and boundary arrays
  if (B != P) {
    const size_t ld = m.ld_face;
    auto widen = [&](auto &v, size_t cnt, int k) {
      std::remove_reference_t<decltype(v)> w(ld * k, 0);
      for (int j = 0; j < k; ++j) std::copy(v.begin() + j * cnt, v.begin() + (j + 1) * cnt, w.begin() + j * ld);
      v.swap(w);
    };
    widen(o->kbface, B, 2); widen(o->sb, B, 3);
    widen(o->ibpoin, P, 2); widen(o->sbb, P, 3); widen(o->anp, P, 3);
  }
  m.iedge = o->iedge.data(); m.sn = o->sn.data(); m.kbface = o->kbface.data(); m.sb = o->sb.data();
  m.ibpoin = o->ibpoin.data(); m.sbb = o->sbb.data(); m.anp = o->anp.data(); m.coord = o->coord.data();
  m.vol = o->vol.data(); m.ncol = (int)C; m.col_start = C ? o->col.data() : nullptr;
  *out = &o->m;
  return 0;
}

int nft_mesh_release(nft_mesh *m) {
  // m is the first member of the OwnedMesh nft_mesh_load allocated
  if (m) delete reinterpret_cast<OwnedMesh *>(m);
  return 0;
}

}  // extern "C"
