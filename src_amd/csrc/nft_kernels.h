// Device kernels of the edge-based NFT step for gfx950 (one translation unit with
// nft_capi.hip).
//
// Layout in HBM
//  * Node-sorted incidence in SELL-64: nodes are taken 64 at a time (one wavefront, one
//    node per lane); chunk c owns slots [chunk_ptr[c], chunk_ptr[c+1]) laid out slot-major
//    (slot m of lane r at chunk_ptr[c] + 64*m + r), so a wavefront reading slot m issues
//    one coalesced load. A slot holds the neighbour node and the signed edge index:
//    e >= 0 when the node is the edge's first node i1, ~e when it is i2. The slots of a
//    node are in increasing global edge order (interior edges, then boundary semi-edges),
//    which is exactly the order in which the reference's edge loops scatter into that
//    node, so every gathered sum rounds as the reference's scatter does.
//  * Edge metrics SoA: sx[e], sy[e], sz[e] over interior edges followed by semi-edges.
//  * Boundary points grouped per node (CSR bp_ptr) in the reference's entry order:
//    flag, sbb(3), anp(3).
//  * Fields are node arrays [npoin]; columns of a generated mesh are level-contiguous, so
//    the 64 lanes of a wavefront usually walk one column and their neighbour gathers and
//    edge-metric loads are near-contiguous.
// All arithmetic is written operand-for-operand like the reference Fortran and compiled
// without FP contraction, so every non-reduction result equals the CPU oracle bit for bit.
// Reductions (GCR inner products) use a fixed tree: per-wave shuffles, per-block shared
// memory, then the last block to finish sums the per-block partials in block order, so
// results are identical from run to run.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NFT_WAVE 64
#define NFT_BLOCK 256
#define NFT_WPB (NFT_BLOCK / NFT_WAVE)
#define NFT_MAXD 18
// the grid-level finalize of the reductions: 1024 virtual threads in 16 wavefronts
#define NFT_FIN_THREADS 1024
#define NFT_FIN_WG (NFT_FIN_THREADS / 64)
// GCR orders up to this one, on grids of at most NFT_FOLD_BLOCKS blocks, have their
// finalize folded into the consuming kernel (at most 1,024: one partial per virtual thread)
#define NFT_FOLD_MAXD 4
#ifndef NFT_FOLD_BLOCKS
#define NFT_FOLD_BLOCKS 256
#endif

namespace nft {

// three packed FP32 values (12 bytes, one dwordx3 load)
struct __align__(4) F3 {
  float x, y, z;
};

struct Incidence {
  int n, nchunk;
  const int32_t *chunk_ptr;  // [nchunk+1]
  const int32_t *deg;        // [n]
  const int32_t *nbr;        // slots
  const int32_t *eid;        // slots
  const int32_t *bp_ptr;     // [n+1]
  const int32_t *bp_flag;    // [nbp]
  const double *bp_sbb;      // [3][nbp]
  const double *bp_anp;      // [3][nbp]
  int nbp;
  int xcd_per;               // logical blocks per XCD slice (remap), 0 = identity
  int nblocks;               // logical blocks
  // Columnar form of the same slots (extruded meshes, node i = c*col_L + k): the interior
  // levels 0 < k < col_L-1 of column c share one slot template t in [tptr[c], tptr[c+1]):
  // neighbour tab[t].x + k, edge (tab[t].y >> 1) - 1 + k, in the i2 role when tab[t].y & 1;
  // the bottom and top nodes keep explicit (neighbour, signed edge) slots at
  // tab[xoff + xptr[c] + 2*m + (0 bottom, 1 top)]. Built only when it reproduces nbr/eid
  // exactly; col_L = 0 otherwise.
  int col_L;
  const int32_t *tptr, *xptr;
  // (neighbour, code) pairs: the templates (tptr) then, from xoff, the explicit slots of
  // the bottom and top nodes (xptr), so every lane reads its slot with one 8-byte load
  const int2 *tab;
  int xoff;
  // edge metrics for the gradient/divergence sweeps as a 16-byte (Sx, Sy) pair and Sz:
  // two aligned loads per slot instead of three
  const double2 *sxy;
  const double *szz;
  // FP32 roundings of the same metrics for the FP32 operator of the mixed form, packed as
  // (Sx, Sy, Sz) so a slot reads them with one 12-byte load (null otherwise)
  const F3 *s3f;
  // Doubly periodic box (the reference's iper = 2): per = 1 drops the boundary-point terms
  // of the gathers and keeps only bcv's flag-11 slip. A node with cslot[i] >= 0 belongs to
  // a coincident set; its gathered sums are replaced by cval[cslot[i]*K + k], the sums over
  // its set (cyclicr/cyclicr2/cyclicr4) that a set kernel (k_per_*) left there just before.
  int per;
  const int32_t *cslot;
  const double *cval;
};

// One slot's edge metric, widened to FP64 (TM = float reads the FP32 copies)
template <typename TM>
__device__ __forceinline__ void load_metric(const Incidence &g, int e, double &sx, double &sy, double &sz) {
  if constexpr (sizeof(TM) == 4) {
    const F3 s = g.s3f[e];
    sx = s.x; sy = s.y; sz = s.z;
  } else {
    const double2 s = g.sxy[e];
    sx = s.x; sy = s.y; sz = g.szz[e];
  }
}

// Per-lane view of the columnar form.
struct ColLane {
  int k;       // level
  int base;    // tab index of slot 0
  int stride;  // 1 (template) or 2 (explicit bottom/top slots)
  int kk;      // level offset added to template entries (0 for explicit slots)
};

__device__ __forceinline__ void col_setup(const Incidence &g, int i, ColLane &C) {
  int c = i / g.col_L;
  C.k = i - c * g.col_L;
  const bool inner = C.k > 0 && C.k < g.col_L - 1;
  C.base = inner ? g.tptr[c] : g.xoff + g.xptr[c] + (C.k == 0 ? 0 : 1);
  C.stride = inner ? 1 : 2;
  C.kk = inner ? C.k : -1;
}

template <typename T>
struct EdgeS {
  const T *sx, *sy, *sz;
};

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous range of chunks
// so neighbouring columns share that XCD's L2.
__device__ __forceinline__ int logical_block(const Incidence &g) {
  int b = blockIdx.x;
  if (g.xcd_per <= 0) return b;
  return (b & 7) * g.xcd_per + (b >> 3);
}

struct Lane {
  int i;      // node id, may be >= n for tail lanes
  int base;   // slot base of the chunk
  int width;  // slots per lane in this chunk
  int deg;    // slots used by this node
  bool valid;
  int lane;
};

__device__ __forceinline__ bool lane_setup(const Incidence &g, Lane &L) {
  int lb = logical_block(g);
  if (lb >= g.nblocks) return false;
  int chunk = lb * NFT_WPB + (threadIdx.x >> 6);
  if (chunk >= g.nchunk) return false;
  L.lane = threadIdx.x & 63;
  L.i = chunk * NFT_WAVE + L.lane;
  // chunk values are the same in every lane of the wavefront: keep them in scalar registers
  L.base = __builtin_amdgcn_readfirstlane(g.chunk_ptr[chunk]);
  L.width = __builtin_amdgcn_readfirstlane((g.chunk_ptr[chunk + 1] - L.base) >> 6);
  L.valid = L.i < g.n;
  L.deg = L.valid ? g.deg[L.i] : 0;
  return true;
}

// Slot m of this lane (SELL-64 or columnar) in two steps: the raw 8-byte entry, then its
// decoding into (neighbour, signed edge), so a batch's entries can be loaded one batch
// ahead. Loads are unconditional: slots past the node's degree read its last slot again
// and lanes past the mesh read slot 0 of their chunk (valid addresses in both cases); the
// callers drop those slots when they accumulate.
template <bool COL>
__device__ __forceinline__ int2 slot_raw(const Incidence &g, const Lane &L, const ColLane &C, int m) {
  const int mm = L.deg > 0 ? min(m, L.deg - 1) : 0;
  if (COL) return g.tab[C.base + mm * C.stride];
  const int slot = L.base + mm * NFT_WAVE + L.lane;
  return make_int2(g.nbr[slot], g.eid[slot]);
}

template <bool COL>
__device__ __forceinline__ void slot_decode(const ColLane &C, int2 v, int &j, int &e) {
  if (COL) {
    const bool tmpl = C.kk >= 0;
    j = v.x + (tmpl ? C.kk : 0);
    const int ea = (v.y >> 1) - 1 + C.kk;
    const int et = (v.y & 1) ? ~ea : ea;
    e = tmpl ? et : v.y;
  } else {
    j = v.x;
    e = v.y;
  }
}

// ------------------------------------------------------------------------------------
// Deterministic reductions
// ------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// Reduces K values over the block (sum, or max where ismax bit k is set); result valid in
// thread 0. Uses sh[K*NFT_WPB].
template <int K>
__device__ __forceinline__ void block_reduce(double (&v)[K], unsigned ismax, double *sh) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = ((ismax >> k) & 1) ? wave_max(v[k]) : wave_sum(v[k]);
  if (l == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k * NFT_WPB + w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double a = sh[k * NFT_WPB];
      for (int j = 1; j < NFT_WPB; ++j) a = ((ismax >> k) & 1) ? fmax(a, sh[k * NFT_WPB + j]) : a + sh[k * NFT_WPB + j];
      v[k] = a;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void st_agent(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_agent(const double *p) {
  return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Writes this block's K partial results (block tree, result of thread 0) to
// partials[blockIdx.x*K + k]. The grid-level sum is k_reduce_final, launched next on the
// same stream: the kernel boundary makes every partial visible without a fence per block.
template <int K>
__device__ __forceinline__ void block_partial(double (&v)[K], unsigned ismax, double *partials, double *sh) {
  block_reduce<K>(v, ismax, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partials[(size_t)blockIdx.x * K + k] = v[k];
  }
}

// The grid-level sum of k_reduce_final, evaluated inside the consuming kernel for grids of
// at most 1,024 blocks, where every virtual thread t = 64*w + l of the finalize holds one
// block's K partials: wavefront 0 of each block replays the 16 virtual wavefronts (the same
// 64-lane trees over the same values, init where t >= nblk) and adds their results in order
// w = 0..15, so the result equals k_reduce_final's bit for bit; every thread of the block
// reads it from sh. Sums only (the GCR dots). K <= KMAX values per block.
// virtual-wavefront replay of the finalize's sum over nblk <= NFT_FOLD_BLOCKS blocks, in
// wavefront 0 (all 64 lanes); value k is a max where bit k of ismax is set; result in lane 0
template <int KMAX>
__device__ __forceinline__ void fold_wave(const double *partials, int nblk, int K, unsigned ismax,
                                          double (&tot)[KMAX]) {
  // virtual wavefronts that can hold blocks; their loads are all issued before the first tree
  constexpr int NW = (NFT_FOLD_BLOCKS + 63) / 64;
  static_assert(NW <= NFT_FIN_WG, "fold: at most one partial per virtual thread");
  const int l = threadIdx.x & 63;
  double pv[NW][KMAX];
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const int t = 64 * w + l;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) pv[w][k] = k < K && t < nblk ? partials[(size_t)t * K + k] : 0.0;
  }
#pragma unroll
  for (int w = 0; w < NFT_FIN_WG; ++w) {
    double acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) acc[k] = ((ismax >> k) & 1) ? -1.0e300 : 0.0;
    if (w < NW && 64 * w < nblk) {
      // wave-uniform branch: all 64 lanes take part in the trees; lanes past nblk add
      // nothing, as the finalize's threads past nblk
      const int t = 64 * w + l;
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        const double v = pv[w < NW ? w : 0][k];
        const bool mx = (ismax >> k) & 1;
        acc[k] = t < nblk ? (mx ? fmax(acc[k], v) : acc[k] + v) : acc[k];
        acc[k] = mx ? wave_max(acc[k]) : wave_sum(acc[k]);
      }
    }
    // a virtual wavefront without blocks reduces 64 initial values: +0.0 or -1e300
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
      tot[k] = w == 0 ? acc[k] : (((ismax >> k) & 1) ? fmax(tot[k], acc[k]) : tot[k] + acc[k]);
  }
}

// the same inside a consuming block: every thread reads the sums from sh (sums only)
template <int KMAX>
__device__ __forceinline__ void fold_reduce(const double *partials, int nblk, int K, double (&res)[KMAX],
                                            double *sh) {
  if (threadIdx.x < 64) {
    double tot[KMAX];
    fold_wave<KMAX>(partials, nblk, K, 0u, tot);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < KMAX; ++k) sh[k] = tot[k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KMAX; ++k) res[k] = sh[k];
  __syncthreads();
}

// ------------------------------------------------------------------------------------
// Boundary conditions on one node (bcv, reference :1583-1643): entries in order.
// u at flags 7/8 becomes ue*flag; ue is read only there.
// ------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void bcv_node(const Incidence &g, int i, T &u, T &v, T &w, const double *ue,
                                         double flag) {
  int b0 = g.bp_ptr[i], b1 = g.bp_ptr[i + 1];
  for (int b = b0; b < b1; ++b) {
    int f = g.bp_flag[b];
    if (g.per && f != 11) continue;   // periodic box (:1603-1612): the flag-11 slip only
    if (f == 11 || f == 5 || f == 6) {
      T ax = (T)g.bp_anp[b], ay = (T)g.bp_anp[g.nbp + b], az = (T)g.bp_anp[2 * g.nbp + b];
      T vn = u * ax + v * ay + w * az;
      u = u - ax * vn;
      v = v - ay * vn;
      w = w - az * vn;
    }
    if (f == 7 || f == 8) u = (T)(ue[i] * flag);
  }
}

// Node coefficients with exact shortcuts: when every rh0 is 1.0 the reference's products
// and quotients by rh0 are identities and rh0 is not read; a uniform coefw is passed by value.
struct NodeCoef {
  const double *vol, *rh0, *coefw, *ue;
  int rh0_one;
  int coefw_uniform;
  double coefw0;
};

__device__ __forceinline__ double rho(const NodeCoef &c, int i) { return c.rh0_one ? 1.0 : c.rh0[i]; }
__device__ __forceinline__ double cw(const NodeCoef &c, int i) { return c.coefw_uniform ? c.coefw0 : c.coefw[i]; }

// Kernels of the GCR loop take the device exit flag (scal + SC_DONE, or null outside the
// loop) and return at once when the solve has ended: iterations enqueued ahead of the
// host's check then do no work. The flag is uniform, so whole blocks leave together.
__device__ __forceinline__ bool skipped(const double *flag) { return flag && *flag != 0.0; }

// Slots are consumed NFT_BATCH_G (gradient) or NFT_BATCH_D (divergence) at a time: the batch's indices, then all its gathers, are
// issued before any use, so a wavefront keeps several independent loads in flight; the
// sums still run slot by slot in the reference's edge order.
#ifndef NFT_MP_GRAD_LDS
#define NFT_MP_GRAD_LDS 1   // MPDATA's gradient sweep over LDS-staged column patches too
#endif
#ifndef NFT_MP_DONOR_LDS
#define NFT_MP_DONOR_LDS 1   // MPDATA's donor-cell sweep over LDS-staged column patches too
#endif
#ifndef NFT_BATCH_G
#define NFT_BATCH_G 2   // gradient sweeps (2 measured best on gfx950, 1..8 tried)
#endif
#ifndef NFT_BATCH_D
#define NFT_BATCH_D 2   // divergence sweeps
#endif
// the same for the FP32 operator's sweeps (half-size metric and gradient loads)
#ifndef NFT_BATCH_G32
#define NFT_BATCH_G32 2
#endif
#ifndef NFT_BATCH_D32
#define NFT_BATCH_D32 2
#endif

// ------------------------------------------------------------------------------------
// Gradient of one field (derivative x3) with per-mode epilogues.
// ------------------------------------------------------------------------------------
enum GradMode {
  GRAD_RAW = 0,     // out0..2 = vol-integrated gradient (derivative idx 1..3)
  GRAD_PRFR = 1,    // out = str - 0.5*dt*g/vol (w: *fz), then bcv(...,1.)       (prfr)
  GRAD_LAPLC = 2,   // out = rh0 * bcv0(0.5*(dt*g/vol), w: *fz) -> (oxy, o2)        (laplc 1st half)
  GRAD_ZERO_PRFR = 3,  // pressure identically zero: out = bcv(str, 1.)
  GRAD_ADV = 4      // out0 = f - dt*((s0*gx + s1*gy) + s2*gz), g = gradient/vol      (velprd)
};

struct GradArgs {
  const double *f;
  double *o0, *o1, *o2;
  double2 *oxy;                // GRAD_LAPLC: (x, y) components; z in o2
  F3 *o3f;                     // GRAD_LAPLC with FP32 metrics: the same rounded to FP32, packed
  const double *s0, *s1, *s2;  // ustr, vstr, wstr (PRFR)
  NodeCoef nc;
  double dt;
  int impfl;
  const double *skip;
};

// TM = float (GRAD_LAPLC only): FP32 metrics, FP64 sums, FP32 outputs (the mixed form's
// FP32 operator)
template <int MODE, bool COL, typename TM = double>
__global__ void __launch_bounds__(NFT_BLOCK) k_grad(Incidence g, EdgeS<double> S, GradArgs a) {
  static_assert(sizeof(TM) == 8 || MODE == GRAD_LAPLC, "FP32 metrics: laplc only");
  constexpr int BG = sizeof(TM) == 4 ? NFT_BATCH_G32 : NFT_BATCH_G;
  if (skipped(a.skip)) return;
  Lane L;
  if (!lane_setup(g, L)) return;
  const int i = L.i;
  ColLane C{};
  if (COL) col_setup(g, L.valid ? i : g.n - 1, C);  // lanes past the mesh: a valid column, no slots used
  double gx = 0., gy = 0., gz = 0.;
  if (MODE != GRAD_ZERO_PRFR) {
    double fi = L.valid ? a.f[i] : 0.;
    // the next batch's slot entries are loaded while this batch's gathers are in flight
    int2 rn[BG];
#pragma unroll
    for (int u = 0; u < BG; ++u) rn[u] = slot_raw<COL>(g, L, C, u);
    for (int m0 = 0; m0 < L.width; m0 += BG) {
      int jj[BG], ee[BG];
#pragma unroll
      for (int u = 0; u < BG; ++u) slot_decode<COL>(C, rn[u], jj[u], ee[u]);
      if (m0 + BG < L.width) {
#pragma unroll
        for (int u = 0; u < BG; ++u) rn[u] = slot_raw<COL>(g, L, C, m0 + BG + u);
      }
      double fj[BG], sx[BG], sy[BG], sz[BG];
#pragma unroll
      for (int u = 0; u < BG; ++u) {
        int e = ee[u] >= 0 ? ee[u] : ~ee[u];
        fj[u] = a.f[jj[u]];
        load_metric<TM>(g, e, sx[u], sy[u], sz[u]);
      }
#pragma unroll
      for (int u = 0; u < BG; ++u) {
        // selects, not branches: slots past the degree leave the sums exactly as they are
        const bool ok = m0 + u < L.deg;
        const double v = (fi + fj[u]) * 0.5;
        const double tx = sx[u] * v, ty = sy[u] * v, tz = sz[u] * v;
        const bool pos = ee[u] >= 0;
        const double nx = pos ? gx + tx : gx - tx, ny = pos ? gy + ty : gy - ty, nz = pos ? gz + tz : gz - tz;
        gx = ok ? nx : gx; gy = ok ? ny : gy; gz = ok ? nz : gz;
      }
    }
    if (L.valid && !g.per) {
      for (int b = g.bp_ptr[i]; b < g.bp_ptr[i + 1]; ++b) {
        gx = gx + g.bp_sbb[b] * fi;
        gy = gy + g.bp_sbb[g.nbp + b] * fi;
        gz = gz + g.bp_sbb[2 * g.nbp + b] * fi;
      }
    }
    if (L.valid && g.per) {
      const int cs = g.cslot[i];
      if (cs >= 0) { gx = g.cval[3 * cs]; gy = g.cval[3 * cs + 1]; gz = g.cval[3 * cs + 2]; }
    }
  }
  if (!L.valid) return;
  if (MODE == GRAD_RAW) {
    a.o0[i] = gx; a.o1[i] = gy; a.o2[i] = gz;
  } else if (MODE == GRAD_ADV) {
    // velprd (lippsLFroude_implgwopks3D_unstr3.f:4469-4481): advection of f by s
    double vl = a.nc.vol[i];
    gx = gx / vl; gy = gy / vl; gz = gz / vl;
    a.o0[i] = a.f[i] - a.dt * (a.s0[i] * gx + a.s1[i] * gy + a.s2[i] * gz);
  } else if (MODE == GRAD_PRFR || MODE == GRAD_ZERO_PRFR) {
    double u, v, w;
    if (MODE == GRAD_PRFR) {
      double vl = a.nc.vol[i];
      gx = gx / vl; gy = gy / vl; gz = gz / vl;
      double fz = a.impfl ? cw(a.nc, i) : 1.0;
      double h = 0.5 * a.dt;
      u = a.s0[i] - h * gx;
      v = a.s1[i] - h * gy;
      w = a.s2[i] - h * gz * fz;
    } else {
      u = a.s0[i]; v = a.s1[i]; w = a.s2[i];
    }
    bcv_node<double>(g, i, u, v, w, a.nc.ue, 1.0);
    a.o0[i] = u; a.o1[i] = v; a.o2[i] = w;
  } else {  // GRAD_LAPLC
    double vl = a.nc.vol[i];
    double fz = a.impfl ? cw(a.nc, i) : 1.0;
    gx = a.dt * gx / vl; gy = a.dt * gy / vl; gz = a.dt * gz / vl;
    gx = 0.5 * gx; gy = 0.5 * gy; gz = 0.5 * gz * fz;
    bcv_node<double>(g, i, gx, gy, gz, a.nc.ue, 0.0);
    if (!a.nc.rh0_one) {
      double r = a.nc.rh0[i];
      gx = gx * r; gy = gy * r; gz = gz * r;
    }
    if constexpr (sizeof(TM) == 4) {
      a.o3f[i] = F3{(float)gx, (float)gy, (float)gz};
    } else {
      a.oxy[i] = make_double2(gx, gy);
      a.o2[i] = gz;
    }
  }
}

// ------------------------------------------------------------------------------------
// Divergence of three fields (derivative(q1,.,1)+derivative(q2,.,2)+derivative(q3,.,3))
// with ND fused inner products of the result (and optionally max|out|).
// ------------------------------------------------------------------------------------
enum DivMode {
  DIV_LAPLC = 0,   // out = ((ax+ay)+az)/(vol*rh0), inputs (qxy, q2)              (laplc 2nd half)
  DIV_DIV = 1,     // out = iflg*((ax/vol+ay/vol)+az/vol)/rh0, inputs scaled by rh0 if iflg=-1
};

struct DotList {
  int nd;                         // number of dot products a_k . out
  const double *a[NFT_MAXD];      // nullptr means `out` itself
  int fin;                        // finalize mode (see gcr_finalize)
  int slot;                       // GCR slot for FIN_BETA
  int nmax;                       // nonzero: also reduce max|out| (entry nd)
  int fold;                       // nonzero: the consuming kernel does the finalize
};

enum Finalize { FIN_NONE = 0, FIN_BETA = 1, FIN_DEL = 2, FIN_NORMS = 3 };

// scalar slots in the device scalar block
enum Scal {
  SC_BETA = 0, SC_RAX = 1, SC_DVMX = 2, SC_RL2 = 3, SC_BMX = 4, SC_EER = 5, SC_EEM = 6,
  SC_R2 = 7, SC_RMX = 8, SC_TMP = 9,
  SC_DONE = 10,  // GCR exit decided on the device: 0 running, 1 converged, 2 stagnated
  SC_NLC = 11,   // inner iterations that passed the exit test
  SC_RL20 = 12,  // previous sum r^2 (stagnation test); 12 and 13 alternate when the exit test is folded
  SC_AX2 = 16,   // [16, 16+NFT_MAXD)
  SC_DEL = 40,   // [40, 40+NFT_MAXD)
  SC_TOTAL = 64
};

struct DivArgs {
  const double *q0, *q1, *q2;
  const double2 *qxy;             // DIV_LAPLC: (x, y) components; z in q2
  const F3 *q3f;                  // DIV_LAPLC with FP32 metrics: FP32 inputs, packed
  double *out;
  NodeCoef nc;
  int iflg;
  DotList dots;
  double *partials;
  unsigned *counter;
  double *scal;
  double epa;
  const double *skip;
};

// a[idx] for a run-time idx as a select chain, so a stays in registers (an indexed access
// would place the array in scratch memory)
template <int K>
__device__ __forceinline__ double pick(const double (&a)[K], int idx) {
  double r = 0.;
#pragma unroll
  for (int k = 0; k < K; ++k) r = k == idx ? a[k] : r;
  return r;
}

template <int K>
__device__ __forceinline__ void gcr_finalize(const DotList &d, const double (&red)[K], double *scal,
                                             double epa) {
  if (d.fin == FIN_BETA) {
    // red[0] = <r, ax_s>, red[1] = <ax_s, ax_s>  (reference :1878-1885)
    double ax2 = fmax(epa, pick(red, 1));
    scal[SC_AX2 + d.slot] = ax2;
    scal[SC_RAX] = red[0];
    scal[SC_BETA] = -red[0] / ax2;
  } else if (d.fin == FIN_DEL) {
    // red[ii] = <ax_ii, ar>, del_ii = -axar/ax2_ii  (reference :1906-1912)
    for (int ii = 0; ii < d.nd; ++ii) scal[SC_DEL + ii] = -pick(red, ii) / scal[SC_AX2 + ii];
  } else if (d.fin == FIN_NORMS) {
    scal[SC_R2] = red[0];
    if (d.nmax) scal[SC_RMX] = pick(red, d.nd);
  }
}

// TM = float (DIV_LAPLC only): FP32 metrics and inputs, FP64 sums and output
template <int MODE, int ND, int NMAX, bool COL, typename TM = double>
__global__ void __launch_bounds__(NFT_BLOCK) k_div(Incidence g, EdgeS<double> S, DivArgs a) {
  static_assert(sizeof(TM) == 8 || MODE == DIV_LAPLC, "FP32 metrics: laplc only");
  constexpr bool F32 = sizeof(TM) == 4;
  constexpr int BD = F32 ? NFT_BATCH_D32 : NFT_BATCH_D;
  constexpr int K = ND + NMAX;
  __shared__ double sh[(K > 0 ? K : 1) * NFT_WPB];
  if (skipped(a.skip)) return;
  Lane L;
  bool active = lane_setup(g, L);
  if (K == 0 && !active) return;
  // lanes of blocks past the end still take part in the block reduction
  double out = 0.;
  if (active) {
    const int i = L.i;
    ColLane C{};
    if (COL) col_setup(g, L.valid ? i : g.n - 1, C);
    const bool scale = MODE == DIV_DIV && a.iflg == -1 && !a.nc.rh0_one;
    double ax = 0., ay = 0., az = 0.;
    double qi0 = 0., qi1 = 0., qi2 = 0.;
    if (L.valid) {
      if (F32) { const F3 v = a.q3f[i]; qi0 = v.x; qi1 = v.y; qi2 = v.z; }
      else {
        if (MODE == DIV_LAPLC) { const double2 v = a.qxy[i]; qi0 = v.x; qi1 = v.y; }
        else { qi0 = a.q0[i]; qi1 = a.q1[i]; }
        qi2 = a.q2[i];
      }
      if (scale) { double r0 = a.nc.rh0[i]; qi0 = qi0 * r0; qi1 = qi1 * r0; qi2 = qi2 * r0; }
    }
    int2 rn[BD];
#pragma unroll
    for (int u = 0; u < BD; ++u) rn[u] = slot_raw<COL>(g, L, C, u);
    for (int m0 = 0; m0 < L.width; m0 += BD) {
      int jj[BD], ee[BD];
#pragma unroll
      for (int u = 0; u < BD; ++u) slot_decode<COL>(C, rn[u], jj[u], ee[u]);
      if (m0 + BD < L.width) {
#pragma unroll
        for (int u = 0; u < BD; ++u) rn[u] = slot_raw<COL>(g, L, C, m0 + BD + u);
      }
      double q0[BD], q1[BD], q2[BD], sx[BD], sy[BD], sz[BD];
#pragma unroll
      for (int u = 0; u < BD; ++u) {
        int e = ee[u] >= 0 ? ee[u] : ~ee[u];
        int j = jj[u];
        if (F32) { const F3 v = a.q3f[j]; q0[u] = v.x; q1[u] = v.y; q2[u] = v.z; }
        else {
          if (MODE == DIV_LAPLC) { const double2 v = a.qxy[j]; q0[u] = v.x; q1[u] = v.y; }
          else { q0[u] = a.q0[j]; q1[u] = a.q1[j]; }
          q2[u] = a.q2[j];
        }
        load_metric<TM>(g, e, sx[u], sy[u], sz[u]);
        if (scale) { double rj = a.nc.rh0[j]; q0[u] = q0[u] * rj; q1[u] = q1[u] * rj; q2[u] = q2[u] * rj; }
      }
#pragma unroll
      for (int u = 0; u < BD; ++u) {
        const bool ok = m0 + u < L.deg;
        const double u0 = (qi0 + q0[u]) * 0.5, u1 = (qi1 + q1[u]) * 0.5, u2 = (qi2 + q2[u]) * 0.5;
        const double tx = sx[u] * u0, ty = sy[u] * u1, tz = sz[u] * u2;
        const bool pos = ee[u] >= 0;
        const double nx = pos ? ax + tx : ax - tx, ny = pos ? ay + ty : ay - ty, nz = pos ? az + tz : az - tz;
        ax = ok ? nx : ax; ay = ok ? ny : ay; az = ok ? nz : az;
      }
    }
    if (L.valid) {
      for (int b = g.bp_ptr[i]; b < (g.per ? 0 : g.bp_ptr[i + 1]); ++b) {
        ax = ax + g.bp_sbb[b] * qi0;
        ay = ay + g.bp_sbb[g.nbp + b] * qi1;
        az = az + g.bp_sbb[2 * g.nbp + b] * qi2;
      }
      if (g.per) {
        const int cs = g.cslot[i];
        if (cs >= 0) { ax = g.cval[3 * cs]; ay = g.cval[3 * cs + 1]; az = g.cval[3 * cs + 2]; }
      }
      if (MODE == DIV_LAPLC) {
        out = a.nc.rh0_one ? (ax + ay + az) / a.nc.vol[i] : (ax + ay + az) / (a.nc.vol[i] * a.nc.rh0[i]);
      } else {
        double vl = a.nc.vol[i];
        double t = (double)a.iflg * (ax / vl + ay / vl + az / vl);
        out = a.nc.rh0_one ? t : t / a.nc.rh0[i];
      }
      a.out[i] = out;
    }
    if (!L.valid) out = 0.;
  }
  if constexpr (K > 0) {
    // fused reductions over the block grid (every block participates)
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.;
    if (active && L.valid) {
#pragma unroll
      for (int k = 0; k < ND; ++k) v[k] = (a.dots.a[k] ? a.dots.a[k][L.i] : out) * out;
      if (NMAX) v[ND] = fabs(out);
    }
    block_partial<K>(v, NMAX ? (1u << ND) : 0u, a.partials, sh);
  }
}

// ------------------------------------------------------------------------------------
// Transpose of laplc (used to train the learned preconditioner). With yh = y/(vol*rho):
//   stage 1 (divergence^T and the per-node maps): for each node j and component c
//     w_j = sum_{e at j} s_je S_e 0.5 (yh_j - yh_i) + sbb_j yh_j,
//     gt_j = D_j P_j^T (rho_j w_j), D_j = diag(0.5 dt/vol_j (1, 1, fz)), P_j^T = the node's
//     boundary maps (slip projections, x-zeroing at flags 7/8) in reverse entry order;
//   stage 2 (gradient^T): out_j = sum_c [ sum_{e at j} s_je S_e^c 0.5 (gt_j^c - gt_i^c) + sbb_j^c gt_j^c ].
// ------------------------------------------------------------------------------------
struct LaplTArgs {
  const double *y;
  double *g0, *g1, *g2;
  double *out;
  NodeCoef nc;
  double dt;
  int impfl;
};

__global__ void __launch_bounds__(NFT_BLOCK) k_laplT_1(Incidence g, EdgeS<double> S, LaplTArgs a) {
  Lane L;
  if (!lane_setup(g, L) || !L.valid) return;
  const int i = L.i;
  const double yi = a.y[i] / (a.nc.vol[i] * rho(a.nc, i));
  double wx = 0., wy = 0., wz = 0.;
  for (int m = 0; m < L.deg; ++m) {
    int slot = L.base + m * NFT_WAVE + L.lane;
    int j = g.nbr[slot], e = g.eid[slot];
    double sg = e >= 0 ? 0.5 : -0.5;
    int ee = e >= 0 ? e : ~e;
    double d = yi - a.y[j] / (a.nc.vol[j] * rho(a.nc, j));
    wx += sg * S.sx[ee] * d; wy += sg * S.sy[ee] * d; wz += sg * S.sz[ee] * d;
  }
  int b0 = g.bp_ptr[i], b1 = g.bp_ptr[i + 1];
  for (int b = b0; b < b1; ++b) {
    wx += g.bp_sbb[b] * yi; wy += g.bp_sbb[g.nbp + b] * yi; wz += g.bp_sbb[2 * g.nbp + b] * yi;
  }
  double r = rho(a.nc, i);
  wx *= r; wy *= r; wz *= r;
  for (int b = b1 - 1; b >= b0; --b) {
    int f = g.bp_flag[b];
    if (f == 11 || f == 5 || f == 6) {
      double nx = g.bp_anp[b], ny = g.bp_anp[g.nbp + b], nz = g.bp_anp[2 * g.nbp + b];
      double vn = wx * nx + wy * ny + wz * nz;
      wx -= nx * vn; wy -= ny * vn; wz -= nz * vn;
    }
    if (f == 7 || f == 8) wx = 0.;
  }
  double h = 0.5 * a.dt / a.nc.vol[i];
  double fz = a.impfl ? cw(a.nc, i) : 1.0;
  a.g0[i] = h * wx; a.g1[i] = h * wy; a.g2[i] = h * wz * fz;
}

__global__ void __launch_bounds__(NFT_BLOCK) k_laplT_2(Incidence g, EdgeS<double> S, LaplTArgs a) {
  Lane L;
  if (!lane_setup(g, L) || !L.valid) return;
  const int i = L.i;
  const double gx = a.g0[i], gy = a.g1[i], gz = a.g2[i];
  double acc = 0.;
  for (int m = 0; m < L.deg; ++m) {
    int slot = L.base + m * NFT_WAVE + L.lane;
    int j = g.nbr[slot], e = g.eid[slot];
    double sg = e >= 0 ? 0.5 : -0.5;
    int ee = e >= 0 ? e : ~e;
    acc += sg * (S.sx[ee] * (gx - a.g0[j]) + S.sy[ee] * (gy - a.g1[j]) + S.sz[ee] * (gz - a.g2[j]));
  }
  for (int b = g.bp_ptr[i]; b < g.bp_ptr[i + 1]; ++b)
    acc += g.bp_sbb[b] * gx + g.bp_sbb[g.nbp + b] * gy + g.bp_sbb[2 * g.nbp + b] * gz;
  a.out[i] = acc;
}

// ------------------------------------------------------------------------------------
// Vector kernels of GCR (grid-stride, fixed grid -> deterministic partials)
// ------------------------------------------------------------------------------------
// p += beta*x; r += beta*ax; reduce max|r| and sum r^2   (reference :1886-1893)
// With fold (grids of at most 1,024 blocks) the direction's finalize (reference :1878-1885)
// is folded in: beta = -<r,ax>/max(epa, <ax,ax>) from the direction kernel's partials, and
// block 0 stores the scalars k_reduce_final would have stored.
struct UpdFold {
  const double *part;   // the direction kernel's partials (null: beta from scal)
  int nblk, slot;
  double epa;
};

template <bool FOLD>
__global__ void __launch_bounds__(NFT_BLOCK) k_gcr_update(int n, double *p, double *r, const double *x,
                                                          const double *ax, double *scal,
                                                          double *partials, const double *skip, UpdFold fo) {
  __shared__ double sh[2 * NFT_WPB];
  if (skipped(skip)) return;
  double beta;
  if constexpr (FOLD) {
    double red[2];
    fold_reduce<2>(fo.part, fo.nblk, 2, red, sh);
    const double ax2 = fmax(fo.epa, red[1]);
    beta = -red[0] / ax2;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      scal[SC_AX2 + fo.slot] = ax2;
      scal[SC_RAX] = red[0];
      scal[SC_BETA] = beta;
    }
  } else {
    beta = scal[SC_BETA];
  }
  double v[2] = {-1.e15, 0.};
  const int stride = gridDim.x * NFT_BLOCK;
  for (int k0 = blockIdx.x * NFT_BLOCK + threadIdx.x; k0 < n; k0 += 4 * stride) {
    double pk[4], rk[4], xk[4], ak[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      int k = k0 + u * stride;
      if (k < n) { pk[u] = p[k]; rk[u] = r[k]; xk[u] = x[k]; ak[u] = ax[k]; }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      int k = k0 + u * stride;
      if (k < n) {
        double pn = pk[u] + beta * xk[u];
        double rn = rk[u] + beta * ak[u];
        p[k] = pn;
        r[k] = rn;
        v[0] = fmax(v[0], fabs(rn));
        v[1] = v[1] + rn * rn;
      }
    }
  }
  block_partial<2>(v, 1u, partials, sh);
}

struct DirArgs {
  int n, nx;                 // number of previous directions
  double *xn, *axn;          // destination slot (may alias x[0]/ax[0] when collapsing)
  const double *x[NFT_MAXD], *ax[NFT_MAXD];
  const double *qr, *ar, *r;
  double *scal, *partials;
  unsigned *counter;
  int slot;                  // slot index of xn
  double epa;
  const double *skip;
  // folded finalize of the divergence's dots <ax_ii, ar> (reference :1906-1912): their
  // partials over fold_nblk blocks (null: del_ii from scal)
  const double *fold_part;
  int fold_nblk;
};

// x_new = qr + sum del_ii x_ii ; ax_new = ar + sum del_ii ax_ii   (reference :1913-1936)
// then <r, ax_new>, <ax_new, ax_new> -> beta for the next inner iteration
template <bool FOLD>
__global__ void __launch_bounds__(NFT_BLOCK) k_gcr_dir(DirArgs a) {
  __shared__ double sh[(NFT_FOLD_MAXD > 2 * NFT_WPB ? NFT_FOLD_MAXD : 2 * NFT_WPB)];
  if (skipped(a.skip)) return;
  double del[NFT_MAXD];
  if constexpr (FOLD) {
    double red[NFT_FOLD_MAXD];
    fold_reduce<NFT_FOLD_MAXD>(a.fold_part, a.fold_nblk, a.nx, red, sh);
#pragma unroll
    for (int ii = 0; ii < NFT_MAXD; ++ii)
      del[ii] = ii < a.nx && ii < NFT_FOLD_MAXD ? -red[ii < NFT_FOLD_MAXD ? ii : 0] / a.scal[SC_AX2 + ii] : 0.;
  } else {
#pragma unroll
    for (int ii = 0; ii < NFT_MAXD; ++ii) del[ii] = ii < a.nx ? a.scal[SC_DEL + ii] : 0.;
  }
  double v[2] = {0., 0.};
  for (int k = blockIdx.x * NFT_BLOCK + threadIdx.x; k < a.n; k += gridDim.x * NFT_BLOCK) {
    double xv = a.qr[k], av = a.ar[k];
#pragma unroll
    for (int ii = 0; ii < NFT_MAXD; ++ii) {
      if (ii < a.nx) {
        xv = xv + del[ii] * a.x[ii][k];
        av = av + del[ii] * a.ax[ii][k];
      }
    }
    a.xn[k] = xv;
    a.axn[k] = av;
    v[0] = v[0] + a.r[k] * av;
    v[1] = v[1] + av * av;
  }
  block_partial<2>(v, 0u, a.partials, sh);
}

// sum a^2 and max|a| (and optionally copy a into dst)
__global__ void __launch_bounds__(NFT_BLOCK) k_norms(int n, const double *a, double *dst, double *scal,
                                                     int sum_slot, int max_slot, double *partials,
                                                     unsigned *counter) {
  __shared__ double sh[2 * NFT_WPB];
  double v[2] = {0., -1.e15};
  for (int k = blockIdx.x * NFT_BLOCK + threadIdx.x; k < n; k += gridDim.x * NFT_BLOCK) {
    double x = a[k];
    if (dst) dst[k] = x;
    v[0] = v[0] + x * x;
    v[1] = fmax(v[1], fabs(x));
  }
  block_partial<2>(v, 2u, partials, sh);
}

// Grid-level end of every fused reduction: one block of 1024 threads sums the per-block
// partials in a fixed order (thread t: blocks t, t+1024, ...; wavefront trees; the 16
// wavefront results in order) and applies the solver's scalar update.
enum FinalKind { FK_DIV = 0, FK_UPDATE = 1, FK_DIR = 2, FK_NORMS = 3, FK_DIAG_MOM = 4, FK_DIAG_VAR = 5 };

// Diagnostic statistics of up to 4 node fields (the reference's div(...,1), courant and
// stats print blocks): per field max, min, mean and standard deviation, written to out[].
#define NFT_DIAG_MAXF 4
struct DiagSpec {
  int nf;                            // fields
  const double *x[NFT_DIAG_MAXF];    // field values
  const double *w;                   // weight (null: 1 per node)
  double scale;                      // each value is x*scale (1: unscaled)
  double n;                          // node count (as the reference's float(npoin))
  double init_mx, init_mn;           // the reference's starting values of its max/min loops
  int mean_by_w[NFT_DIAG_MAXF];      // mean denominator: sum of weights (1) or n (0)
  int sd_by_w[NFT_DIAG_MAXF];        // sd denominator: sum of weights (1) or n (0)
  int argmax;                        // also the 1-based index of the last maximum of field 0
  double *out;                       // [4*nf] (+1 for argmax): max, min, mean, sd per field
};

struct FinalOp {
  int kind;
  int K;            // values per block
  unsigned ismax;   // bit k: value k is a max
  DotList dots;     // FK_DIV
  int slot;         // FK_DIR: slot of the new direction
  int sum_slot, max_slot;  // FK_NORMS
  double epa;
  const double *skip;
  // FK_UPDATE with ctl: the exit test of the reference's inner loop (:1893-1895), decided
  // here: converged when max|r| <= eps at outer iteration it >= itmn, stagnated when
  // sum r^2 >= its previous value; otherwise the previous value becomes max(sum r^2, epa)
  // and the pass is counted. The decision is also stored to *hflag (mapped host memory).
  int ctl, it, itmn;
  double eps;
  double *hflag;
  DiagSpec diag;   // FK_DIAG_MOM / FK_DIAG_VAR
};

// the solver's scalar update from a grid-level result acc (k_reduce_final, k_reduce_small)
template <int K>
__device__ __forceinline__ void final_apply(const FinalOp &op, double (&acc)[K], double *scal) {
  if (op.kind == FK_DIAG_MOM) {
    // acc: per field (max, min, sum of w*x), then the sum of weights
    const DiagSpec &d = op.diag;
    const double W = d.w ? pick(acc, 3 * d.nf) : d.n;
    for (int f = 0; f < d.nf; ++f) {
      d.out[4 * f + 0] = fmax(d.init_mx, pick(acc, 3 * f + 0));
      d.out[4 * f + 1] = fmin(d.init_mn, -pick(acc, 3 * f + 1));
      d.out[4 * f + 2] = pick(acc, 3 * f + 2) / (d.mean_by_w[f] ? W : d.n);
      d.out[4 * f + 3] = W;   // kept for the second pass (overwritten by the sd)
    }
    return;
  }
  if (op.kind == FK_DIAG_VAR) {
    // acc: per field sum of w*(x-mean)^2, then (argmax) the largest index attaining the max
    const DiagSpec &d = op.diag;
    for (int f = 0; f < d.nf; ++f) {
      const double W = d.out[4 * f + 3];
      d.out[4 * f + 3] = sqrt(pick(acc, f) / (d.sd_by_w[f] ? W : d.n));
    }
    if (d.argmax) d.out[4 * d.nf] = pick(acc, d.nf);
    return;
  }
  if (op.kind == FK_DIV) {
    gcr_finalize(op.dots, acc, scal, op.epa);
  } else if (op.kind == FK_UPDATE) {
    scal[SC_DVMX] = acc[0];
    scal[SC_RL2] = pick(acc, 1);
    if (op.ctl) {
      double done = 0.0;
      if (!(pick(acc, 1) == pick(acc, 1))) done = 2.0;   // NaN residual: stop, reported as not converged
      else if (acc[0] <= op.eps && op.it >= op.itmn) done = 1.0;
      else if (pick(acc, 1) >= scal[SC_RL20]) done = 2.0;
      else {
        scal[SC_RL20] = fmax(pick(acc, 1), op.epa);
        scal[SC_NLC] = scal[SC_NLC] + 1.0;
      }
      if (done != 0.0) {
        scal[SC_DONE] = done;
        if (op.hflag) *op.hflag = done;
      }
    }
  } else if (op.kind == FK_DIR) {
    double ax2 = fmax(op.epa, pick(acc, 1));
    scal[SC_AX2 + op.slot] = ax2;
    scal[SC_RAX] = acc[0];
    scal[SC_BETA] = -acc[0] / ax2;
  } else {
    scal[op.sum_slot] = acc[0];
    scal[op.max_slot] = pick(acc, 1);
  }
}

// The grid-level sum runs as NFT_FIN_WG one-wavefront workgroups, so the partials are
// fetched by 16 CUs at once (one CU alone is bound by the misses it can keep in flight).
// Summation order: virtual thread t = 64*w + l (workgroup w, lane l) sums blocks t, t+1024,
// ... in that order; wavefront tree; then the 16 wavefront results in order w = 0..15,
// summed by the workgroup that arrives last at the counter.
template <int K>
__global__ void __launch_bounds__(64) k_reduce_final(const double *partials, int nblk, FinalOp op, double *scal,
                                                     double *wres, unsigned *arrived) {
  constexpr int U = K <= 4 ? 16 : (64 / K > 0 ? 64 / K : 1);
  if (skipped(op.skip)) return;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = ((op.ismax >> k) & 1) ? -1.0e300 : 0.0;
  // U blocks' K values are loaded before they are added
  for (int b0 = blockIdx.x * 64 + threadIdx.x; b0 < nblk; b0 += U * NFT_FIN_THREADS) {
    double pv[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u * NFT_FIN_THREADS;
#pragma unroll
      for (int k = 0; k < K; ++k) pv[u][k] = b < nblk ? partials[(size_t)b * K + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (b0 + u * NFT_FIN_THREADS < nblk) {
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = ((op.ismax >> k) & 1) ? fmax(acc[k], pv[u][k]) : acc[k] + pv[u][k];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = ((op.ismax >> k) & 1) ? wave_max(acc[k]) : wave_sum(acc[k]);
  if (threadIdx.x != 0) return;
  // publish this wavefront's result, then count the arrival (release); the last arrival
  // (acquire) reads all 16 results and resets the counter for the next launch
#pragma unroll
  for (int k = 0; k < K; ++k)
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(wres) + blockIdx.x * K + k,
                       (unsigned long long)__double_as_longlong(acc[k]), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  const unsigned prev = __hip_atomic_fetch_add(arrived, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
  if (prev != NFT_FIN_WG - 1) return;
  const unsigned long long *wr = reinterpret_cast<const unsigned long long *>(wres);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double a = __longlong_as_double(
        (long long)__hip_atomic_load(wr + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    for (int j = 1; j < NFT_FIN_WG; ++j) {
      const double x = __longlong_as_double(
          (long long)__hip_atomic_load(wr + j * K + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      a = ((op.ismax >> k) & 1) ? fmax(a, x) : a + x;
    }
    acc[k] = a;
  }
  __hip_atomic_store(arrived, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  final_apply<K>(op, acc, scal);
}

// The same grid-level sum for at most NFT_FOLD_BLOCKS blocks as one wavefront that replays
// the 16 virtual wavefronts (fold_wave): no arrival counter, same result bit for bit.
template <int K>
__global__ void __launch_bounds__(64) k_reduce_small(const double *partials, int nblk, FinalOp op, double *scal) {
  if (skipped(op.skip)) return;
  double acc[K];
  fold_wave<K>(partials, nblk, K, op.ismax, acc);
  if (threadIdx.x != 0) return;
  final_apply<K>(op, acc, scal);
}

// ------------------------------------------------------------------------------------
// Diagnostics (the reference prints them after every step, :456-459)
// ------------------------------------------------------------------------------------
// First pass: per field max, min and sum of w*x (x = value*scale), and the sum of weights.
// Sums run over the grid-stride order of the vector kernels (block partials, then
// k_reduce_final); max and min are exact in any order.
template <int NF>
__device__ __forceinline__ unsigned mins_mask() {
  unsigned m = 0;
  for (int f = 0; f < NF; ++f) m |= 1u << (3 * f + 1);
  return m;
}

template <int NF>
__global__ void __launch_bounds__(NFT_BLOCK) k_diag_mom(int n, DiagSpec d, double *partials) {
  constexpr int K = 3 * NF + 1;
  __shared__ double sh[K * NFT_WPB];
  double v[K];
#pragma unroll
  for (int f = 0; f < NF; ++f) { v[3 * f] = -1.0e300; v[3 * f + 1] = 1.0e300; v[3 * f + 2] = 0.0; }
  v[3 * NF] = 0.0;
  for (int i = blockIdx.x * NFT_BLOCK + threadIdx.x; i < n; i += gridDim.x * NFT_BLOCK) {
    const double w = d.w ? d.w[i] : 1.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double x = d.scale == 1.0 ? d.x[f][i] : d.x[f][i] * d.scale;
      v[3 * f] = fmax(v[3 * f], x);
      v[3 * f + 1] = fmin(v[3 * f + 1], x);
      v[3 * f + 2] = v[3 * f + 2] + (d.w ? w * x : x);
    }
    v[3 * NF] = v[3 * NF] + w;
  }
  unsigned ismax = 0;
#pragma unroll
  for (int f = 0; f < NF; ++f) ismax |= 1u << (3 * f);
  // the min is reduced as the max of -x
#pragma unroll
  for (int f = 0; f < NF; ++f) v[3 * f + 1] = -v[3 * f + 1];
  block_partial<K>(v, ismax | mins_mask<NF>(), partials, sh);
}

// Second pass: per field sum of w*(x-mean)^2 and, for the argmax field, the largest
// 1-based index whose value equals the max (the reference's running `if (crnmx.eq.cour(ip))
// ic=ip` ends there).
template <int NF>
__global__ void __launch_bounds__(NFT_BLOCK) k_diag_var(int n, DiagSpec d, double *partials) {
  constexpr int K = NF + 1;
  __shared__ double sh[K * NFT_WPB];
  double v[K];
  double mean[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) { v[f] = 0.0; mean[f] = d.out[4 * f + 2]; }
  v[NF] = 0.0;
  const double mx = d.out[0];
  for (int i = blockIdx.x * NFT_BLOCK + threadIdx.x; i < n; i += gridDim.x * NFT_BLOCK) {
    const double w = d.w ? d.w[i] : 1.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double x = d.scale == 1.0 ? d.x[f][i] : d.x[f][i] * d.scale;
      const double dv = x - mean[f];
      const double q = dv * dv;
      v[f] = v[f] + (d.w ? w * q : q);
    }
    if (d.argmax) {
      const double x0 = d.scale == 1.0 ? d.x[0][i] : d.x[0][i] * d.scale;
      if (x0 == mx) v[NF] = fmax(v[NF], (double)(i + 1));
    }
  }
  block_partial<K>(v, 1u << NF, partials, sh);
}

// courant(uuu,cour,iflag) (:1974-2094): per node the sum over its edges, in the reference's
// edge order, of |a_e| (iflag 1), +-a_e (2) or the signed upwind parts (3), with
// a_e = ((u_i1+u_i2)*0.5*Sx + (v_i1+v_i2)*0.5*Sy + (w_i1+w_i2)*0.5*Sz) / (0.5*(rho_i1+rho_i2)),
// then the boundary points' (u*sbbx + v*sbby + w*sbbz)/rho, then *dt/vol/rho.
template <int FLAG, bool COL>
__global__ void __launch_bounds__(NFT_BLOCK) k_courant(Incidence g, EdgeS<double> S, const double *u,
                                                       const double *v, const double *w, NodeCoef nc, double dt,
                                                       double *cour) {
  Lane L;
  if (!lane_setup(g, L)) return;
  const int i = L.i;
  ColLane C{};
  if (COL) col_setup(g, L.valid ? i : g.n - 1, C);
  if (!L.valid) return;
  const double ui = u[i], vi = v[i], wi = w[i], ri = rho(nc, i);
  double acc = 0.;
  for (int m = 0; m < L.deg; ++m) {
    int j, e;
    slot_decode<COL>(C, slot_raw<COL>(g, L, C, m), j, e);
    const int ea = e >= 0 ? e : ~e;
    double a = (ui + u[j]) * 0.5 * S.sx[ea] + (vi + v[j]) * 0.5 * S.sy[ea] + (wi + w[j]) * 0.5 * S.sz[ea];
    if (!nc.rh0_one) a = a / (0.5 * (ri + rho(nc, j)));
    if (FLAG == 1) acc = acc + fabs(a);
    else if (FLAG == 2) acc = e >= 0 ? acc + a : acc - a;
    else acc = e >= 0 ? acc + fmax(0., a) : acc - fmin(0., a);
  }
  // the periodic box skips the boundary points (:2049)
  for (int b = g.bp_ptr[i]; b < (g.per ? 0 : g.bp_ptr[i + 1]); ++b) {
    double a = ui * g.bp_sbb[b] + vi * g.bp_sbb[g.nbp + b] + wi * g.bp_sbb[2 * g.nbp + b];
    if (!nc.rh0_one) a = a / ri;
    if (FLAG == 1) acc = acc + fabs(a);
    else if (FLAG == 2) acc = acc + a;
    else acc = acc + fmax(0., a);
  }
  double c = acc * dt / nc.vol[i];
  if (!nc.rh0_one) c = c / ri;
  cour[i] = c;
}

// ------------------------------------------------------------------------------------
// Preconditioners
// ------------------------------------------------------------------------------------
// vertical line, general columns: forward/backward Thomas sweep per column, then
// p = p + omega*z (it=1: p = 0 + omega*M^-1 rhs; it>1: rhs - L p is passed in res).
template <typename T>
__global__ void __launch_bounds__(NFT_BLOCK) k_line(int ncol, const int32_t *col_start, const T *la,
                                                    const T *linv, const T *lc, const double *res,
                                                    const double *pin, double *p, double omega,
                                                    T *zbuf, const double *skip) {
  if (skipped(skip)) return;
  int c = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (c >= ncol) return;
  int s0 = col_start[c], s1 = col_start[c + 1];
  T dprev = (T)0;
  for (int i = s0; i < s1; ++i) {
    T d = ((T)res[i] - la[i] * dprev) * linv[i];
    zbuf[i] = d;
    dprev = d;
  }
  T znext = (T)0;
  for (int i = s1 - 1; i >= s0; --i) {
    T z = (i == s1 - 1) ? zbuf[i] : zbuf[i] - lc[i] * znext;
    znext = z;
    double p0 = pin ? pin[i] : 0.;
    p[i] = p0 + omega * (double)z;
  }
}

// vertical line, columns of equal length L <= 64: a wavefront holds 64/L columns, one
// level per lane. The factors of the Thomas algorithm are fixed (a: sub-diagonal, inv:
// 1/pivot, c': scaled super-diagonal), so the solve is two first-order linear recurrences
//   y_k = d_k*inv_k + alpha_k*y_(k-1),  alpha_k = -(a_k*inv_k)      (forward)
//   z_k = y_k + gamma_k*z_(k+1),        gamma_k = -c'_k               (backward)
// each evaluated as an inclusive scan of affine maps across the wavefront in 6 steps
// (distance 1..32): (beta, alpha) <- (beta + alpha*beta', alpha*alpha') with the map s
// lanes away. alpha = 0 at a column's first row and gamma = 0 at its last row (a = 0,
// c' = 0 there), which drops every term from a neighbouring column exactly; lanes past
// the columns carry the zero map.
template <typename T>
__device__ __forceinline__ T lane_get(T v, int src) { return __shfl(v, src, 64); }

template <typename T>
__global__ void __launch_bounds__(NFT_BLOCK) k_line_scan(int ncol, int L, const T *la, const T *linv,
                                                         const T *lc, const double *res, const double *pin,
                                                         double *p, double omega, const double *skip) {
  if (skipped(skip)) return;
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * NFT_BLOCK + threadIdx.x) >> 6;
  const int cpw = 64 / L;
  const int col = wave * cpw + lane / L;
  const bool real = lane < cpw * L && col < ncol;
  const int i = col * L + lane % L;
  T al = (T)0, be = (T)0, ga = (T)0;
  if (real) {
    T iv = linv[i];
    al = -(la[i] * iv);
    be = (T)res[i] * iv;
    ga = -lc[i];
  }
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    T ap = lane_get(al, lane >= s ? lane - s : lane), bp = lane_get(be, lane >= s ? lane - s : lane);
    if (lane >= s) {
      be = be + al * bp;
      al = al * ap;
    }
  }
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    T gp = lane_get(ga, lane + s < 64 ? lane + s : lane), bp = lane_get(be, lane + s < 64 ? lane + s : lane);
    if (lane + s < 64) {
      be = be + ga * bp;
      ga = ga * gp;
    }
  }
  if (real) {
    double p0 = pin ? pin[i] : 0.;
    p[i] = p0 + omega * (double)be;
  }
}

// vertical line, columns of equal length 64 < L <= 64*R: one wavefront per column, lane l
// holding rows l*R .. l*R+R-1 (rows >= L carry the zero map). Same recurrences as
// k_line_scan, evaluated as a thread-coarsened scan:
//   forward:  each lane composes its rows' maps in order into (A, B) (A = 1, B = 0 to
//             start; B = beta + alpha*B, A = alpha*A), the wavefront scans the lane maps
//             as k_line_scan does, lane l takes y_in = B of lane l-1 (0 for lane 0) and
//             replays its rows: y = beta + alpha*y;
//   backward: the same from the last row down with gamma = -c' and the forward results.
template <typename T, int R>
__global__ void __launch_bounds__(NFT_BLOCK) k_line_scan_r(int ncol, int L, const T *la, const T *linv,
                                                           const T *lc, const double *res, const double *pin,
                                                           double *p, double omega, const double *skip) {
  if (skipped(skip)) return;
  const int lane = threadIdx.x & 63;
  const int col = (blockIdx.x * NFT_BLOCK + threadIdx.x) >> 6;
  if (col >= ncol) return;
  const size_t c0 = (size_t)col * L;
  T al[R], be[R], ga[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int row = lane * R + r;
    al[r] = (T)0; be[r] = (T)0; ga[r] = (T)0;
    if (row < L) {
      const size_t i = c0 + row;
      T iv = linv[i];
      al[r] = -(la[i] * iv);
      be[r] = (T)res[i] * iv;
      ga[r] = -lc[i];
    }
  }
  // forward
  T A = (T)1, B = (T)0;
#pragma unroll
  for (int r = 0; r < R; ++r) { B = be[r] + al[r] * B; A = al[r] * A; }
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    T ap = lane_get(A, lane >= s ? lane - s : lane), bp = lane_get(B, lane >= s ? lane - s : lane);
    if (lane >= s) {
      B = B + A * bp;
      A = A * ap;
    }
  }
  T y = lane_get(B, lane > 0 ? lane - 1 : 0);
  if (lane == 0) y = (T)0;
#pragma unroll
  for (int r = 0; r < R; ++r) { y = be[r] + al[r] * y; be[r] = y; }
  // backward (be now holds the forward results)
  T G = (T)1;
  B = (T)0;
#pragma unroll
  for (int r = R - 1; r >= 0; --r) { B = be[r] + ga[r] * B; G = ga[r] * G; }
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    T gp = lane_get(G, lane + s < 64 ? lane + s : lane), bp = lane_get(B, lane + s < 64 ? lane + s : lane);
    if (lane + s < 64) {
      B = B + G * bp;
      G = G * gp;
    }
  }
  T z = lane_get(B, lane < 63 ? lane + 1 : 63);
  if (lane == 63) z = (T)0;
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    z = be[r] + ga[r] * z;
    const int row = lane * R + r;
    if (row < L) {
      const size_t i = c0 + row;
      double p0 = pin ? pin[i] : 0.;
      p[i] = p0 + omega * (double)z;
    }
  }
}

// point relaxation: p = p + omega*pstep*(r - rhs)   (reference :607-612)
__global__ void __launch_bounds__(NFT_BLOCK) k_point(int n, const double *pstep, const double *r,
                                                     const double *rhs, double *p, double omega,
                                                     int first, const double *skip) {
  if (skipped(skip)) return;
  int k = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (k >= n) return;
  double pk = first ? 0. : p[k];
  double rk = first ? 0. : r[k];
  p[k] = pk + omega * pstep[k] * (rk - rhs[k]);
}

__global__ void k_sub(int n, const double *a, const double *b, double *c, const double *skip) {
  if (skipped(skip)) return;
  int k = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (k < n) c[k] = a[k] - b[k];
}

// ------------------------------------------------------------------------------------
// Vertical line preconditioner with the in-column band of laplc itself (precon 3): per
// column the banded LU factors (unit lower l1, l2; upper 1/d, u1, u2), stored interleaved
// as f[(block*L + k)*cpb + c] so lane c of a block reads them coalesced level by level.
// One wavefront per block of cpb consecutive columns: the residual block (contiguous in
// memory) is staged in LDS with coalesced loads, each lane runs its column's forward and
// backward substitution sequentially from LDS (rows padded to L+1 doubles), and the result
// is written back coalesced as p = pin + omega*z. Same operation sequence as the oracle.
// ------------------------------------------------------------------------------------
// band probing (set-up of precon 3): v = indicator of the probe nodes (columns of colour
// col, levels = r mod 5), then each probe's column rows within two levels of it take
// their entry of laplc(v)
__global__ void k_band_probe(int n, int L, const int32_t *color, int col, int r, double *v) {
  const int i = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (i >= n) return;
  v[i] = (color[i / L] == col && (i % L) % 5 == r) ? 1.0 : 0.0;
}

__global__ void k_band_extract(int n, int L, const int32_t *color, int col, int r, const double *Lv, double *bm2,
                               double *bm1, double *b0, double *bp1, double *bp2) {
  const int i = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (i >= n) return;
  const int l = i % L;
  if (color[i / L] != col || l % 5 != r) return;
  b0[i] = Lv[i];
  if (l + 1 < L) bm1[i + 1] = Lv[i + 1];
  if (l + 2 < L) bm2[i + 2] = Lv[i + 2];
  if (l >= 1) bp1[i - 1] = Lv[i - 1];
  if (l >= 2) bp2[i - 2] = Lv[i - 2];
}

#ifndef NFT_PENTA_CHUNK
#define NFT_PENTA_CHUNK 16
#endif

#ifndef NFT_PENTA_MINBLK
// band solve: columns per block halved until the grid has at least this many one-wavefront
// blocks (2,000 columns: 4 per block, 500 blocks; measured 15.0 -> 13.2 us against 64 per
// block, while 2 and 1 per block were slower)
#define NFT_PENTA_MINBLK 256
#endif

#ifndef NFT_PENTA_STAGE
#define NFT_PENTA_STAGE 16  // loads in flight per lane while staging the band solve's columns (8 and 16 measured)
#endif

template <typename T>
struct BandFactors {
  const T *l1, *l2, *u1, *u2, *dinv;
};

// The exit test of GCR's inner loop (reference :1893-1897) folded into the band solve that
// follows the residual update, for small grids: every block replays the finalize's sum of
// the update's partials (max|r|, sum r^2; fold_wave) and takes the same decision as
// final_apply's FK_UPDATE; block 0 stores the scalars, the exit flag and its host copy. The
// previous sum r^2 alternates between two slots by iteration parity (read par, written
// par ^ 1), so no block reads a value another block of the same launch has updated.
struct ExitFold {
  const double *part;   // the update's partials (null: no fold)
  int nblk, it, itmn, par;
  double eps, epa;
  double *scal, *hflag;
};

// T = double (FP64 path, same operation sequence as the oracle) or float (mixed precision:
// factors, staged residual and substitution in FP32, the update of p in FP64)
template <typename T, bool XF = false>
__global__ void __launch_bounds__(NFT_WAVE) k_line_penta(int ncol, int L, int cpb, BandFactors<T> f,
                                                         const double *res, const double *pin, double *p,
                                                         double omega, const double *skip, ExitFold xf) {
  if (skipped(skip)) return;
  if constexpr (XF) {
    double acc[2];
    fold_wave<2>(xf.part, xf.nblk, 2, 1u, acc);   // the block is one wavefront
    const double dvmx = __shfl(acc[0], 0, 64), rl2 = __shfl(acc[1], 0, 64);
    double done = 0.0;
    if (!(rl2 == rl2)) done = 2.0;   // NaN residual: stop, reported as not converged
    else if (dvmx <= xf.eps && xf.it >= xf.itmn) done = 1.0;
    else if (rl2 >= xf.scal[SC_RL20 + xf.par]) done = 2.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      xf.scal[SC_DVMX] = dvmx;
      xf.scal[SC_RL2] = rl2;
      if (done != 0.0) {
        xf.scal[SC_DONE] = done;
        if (xf.hflag) *xf.hflag = done;
      } else {
        xf.scal[SC_RL20 + (xf.par ^ 1)] = fmax(rl2, xf.epa);
        xf.scal[SC_NLC] = xf.scal[SC_NLC] + 1.0;
      }
    }
    if (done != 0.0) return;
  }
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *lds = (T *)lds_raw;
  const int lane = threadIdx.x, cb = blockIdx.x;
  const int c0 = cb * cpb, ncb = min(cpb, ncol - c0);
  const int RL = L + 1;
  const size_t node0 = (size_t)c0 * L;
  const int tot = ncb * L;
  // staging: NFT_PENTA_STAGE loads in flight per lane before their LDS stores (one wavefront
  // per block leaves little else to hide the latency)
  int idx = lane;
  for (; idx + (NFT_PENTA_STAGE - 1) * NFT_WAVE < tot; idx += NFT_PENTA_STAGE * NFT_WAVE) {
    double v[NFT_PENTA_STAGE];
#pragma unroll
    for (int u = 0; u < NFT_PENTA_STAGE; ++u) v[u] = res[node0 + idx + u * NFT_WAVE];
#pragma unroll
    for (int u = 0; u < NFT_PENTA_STAGE; ++u) {
      const int q = idx + u * NFT_WAVE;
      lds[(q / L) * RL + q % L] = (T)v[u];
    }
  }
  for (; idx < tot; idx += NFT_WAVE) lds[(idx / L) * RL + idx % L] = (T)res[node0 + idx];
  __syncthreads();
  if (lane < ncb) {
    // the factors of NFT_PENTA_CHUNK levels are loaded together ahead of their steps, so
    // the substitution chain waits on memory once per chunk, not once per level
    T *z = lds + lane * RL;
    const size_t fb = (size_t)cb * L * cpb + lane;
    T y1 = 0, y2 = 0;
    for (int k0 = 0; k0 < L; k0 += NFT_PENTA_CHUNK) {
      T a1[NFT_PENTA_CHUNK], a2[NFT_PENTA_CHUNK];
#pragma unroll
      for (int j = 0; j < NFT_PENTA_CHUNK; ++j) {
        const int k = min(k0 + j, L - 1);
        a1[j] = f.l1[fb + (size_t)k * cpb];
        a2[j] = f.l2[fb + (size_t)k * cpb];
      }
#pragma unroll
      for (int j = 0; j < NFT_PENTA_CHUNK; ++j) {
        const int k = k0 + j;
        if (k < L) {
          T y = z[k];
          if (k >= 1) y = y - a1[j] * y1;
          if (k >= 2) y = y - a2[j] * y2;
          z[k] = y;
          y2 = y1; y1 = y;
        }
      }
    }
    T z1 = 0, z2 = 0;
    for (int k0 = L - 1; k0 >= 0; k0 -= NFT_PENTA_CHUNK) {
      T b1[NFT_PENTA_CHUNK], b2[NFT_PENTA_CHUNK], bd[NFT_PENTA_CHUNK];
#pragma unroll
      for (int j = 0; j < NFT_PENTA_CHUNK; ++j) {
        const int k = max(k0 - j, 0);
        b1[j] = f.u1[fb + (size_t)k * cpb];
        b2[j] = f.u2[fb + (size_t)k * cpb];
        bd[j] = f.dinv[fb + (size_t)k * cpb];
      }
#pragma unroll
      for (int j = 0; j < NFT_PENTA_CHUNK; ++j) {
        const int k = k0 - j;
        if (k >= 0) {
          T t = z[k];
          if (k + 1 < L) t = t - b1[j] * z1;
          if (k + 2 < L) t = t - b2[j] * z2;
          t = t * bd[j];
          z[k] = t;
          z2 = z1; z1 = t;
        }
      }
    }
  }
  __syncthreads();
  idx = lane;
  for (; idx + (NFT_PENTA_STAGE - 1) * NFT_WAVE < tot; idx += NFT_PENTA_STAGE * NFT_WAVE) {
    double p0[NFT_PENTA_STAGE];
#pragma unroll
    for (int u = 0; u < NFT_PENTA_STAGE; ++u) p0[u] = pin ? pin[node0 + idx + u * NFT_WAVE] : 0.;
#pragma unroll
    for (int u = 0; u < NFT_PENTA_STAGE; ++u) {
      const int q = idx + u * NFT_WAVE;
      p[node0 + q] = p0[u] + omega * (double)lds[(q / L) * RL + q % L];
    }
  }
  for (; idx < tot; idx += NFT_WAVE) {
    const double p0 = pin ? pin[node0 + idx] : 0.;
    p[node0 + idx] = p0 + omega * (double)lds[(idx / L) * RL + idx % L];
  }
}

// plain streaming copy (bandwidth reference)
__global__ void __launch_bounds__(NFT_BLOCK) k_stream_copy(const double2 *a, double2 *b, size_t n) {
  for (size_t k = (size_t)blockIdx.x * NFT_BLOCK + threadIdx.x; k < n; k += (size_t)gridDim.x * NFT_BLOCK) b[k] = a[k];
}

__global__ void k_copy(int n, const double *a, double *b, const double *skip) {
  if (skipped(skip)) return;
  int k = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (k < n) b[k] = a[k];
}

// ------------------------------------------------------------------------------------
// Multigrid branch of precon (lippsLFroude_implgwopks3D_unstr3.f:618-748): grid transfers
// and the coarse-level point steps. One thread per node; the neighbours of a fine node
// are its incidence slots, which hold them in neigh_definition's order (interior edges,
// then boundary-face edges, each in input order).
// ------------------------------------------------------------------------------------
enum MgTransfer {
  MG_AS_REFERENCE = 0,  // the neighbour counter q = 1..count used as a node id (:817, :913)
  MG_NEIGHBOURS = 1,    // the neighbour list
};

__device__ __forceinline__ int mg_neighbour(const Incidence &f, int ip, int q, int transfer) {
  if (transfer == MG_AS_REFERENCE) return q;
  const int ch = ip / NFT_WAVE, r = ip % NFT_WAVE;
  return f.nbr[f.chunk_ptr[ch] + (long)q * NFT_WAVE + r];
}

// restriction_unstr (:757-822): out[ic] = (sum_q (f[ip] + f[j_q]) / 2) / count, ip = match[ic]
__global__ void __launch_bounds__(NFT_BLOCK) k_mg_restrict(int nc, const int32_t *match, Incidence fine,
                                                           const double *f, double *out, int transfer,
                                                           const double *skip) {
  if (skipped(skip)) return;
  const int ic = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (ic >= nc) return;
  const int ip = match[ic];
  const int cnt = fine.deg[ip];
  const double fi = f[ip];
  double sc = 0.;
  for (int q = 0; q < cnt; ++q) sc = sc + (fi + f[mg_neighbour(fine, ip, q, transfer)]) / 2.0;
  out[ic] = sc / (double)cnt;
}

// prolongation_unstr (:828-940): injection at the fine nodes that carry a coarse node, the
// mean of such neighbours elsewhere (0 where there is none). A node counts as carrying a
// value when it differs from the reference's marker -99999.
__global__ void __launch_bounds__(NFT_BLOCK) k_mg_prolong(int nf, const int32_t *imatch, Incidence fine,
                                                          const double *coarse, double *out, int transfer,
                                                          const double *skip) {
  if (skipped(skip)) return;
  const int ip = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (ip >= nf) return;
  const double marker = -99999.;
  const int ic = imatch[ip];
  const double v0 = ic >= 0 ? coarse[ic] : 0.;
  if (ic >= 0 && v0 != marker) { out[ip] = v0; return; }
  const int cnt = fine.deg[ip];
  double sc = v0, filled = 0.;
  for (int q = 0; q < cnt; ++q) {
    const int jc = imatch[mg_neighbour(fine, ip, q, transfer)];
    if (jc >= 0) {
      const double vj = coarse[jc];
      if (vj != marker) { sc = sc + vj; filled = filled + 1.; }
    }
  }
  out[ip] = filled != 0. ? sc / filled : sc;
}

// start of the coarsest relaxation (:701-703): p = -(rrhs * pstep * omega), in that order
__global__ void k_mg_start_cc(int n, const double *pstep, const double *rrhs, double *p, double omega,
                              const double *skip) {
  if (skipped(skip)) return;
  const int k = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (k < n) p[k] = -(rrhs[k] * pstep[k] * omega);
}

// coarse-grid correction (:716-718, :733-735): a += b
__global__ void k_mg_add(int n, double *a, const double *b, const double *skip) {
  if (skipped(skip)) return;
  const int k = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (k < n) a[k] = a[k] + b[k];
}

// start of a device-controlled GCR loop: flag cleared, pass count 0, previous sum r^2
__global__ void k_gcr_ctl_begin(double *scal, double rl20) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    scal[SC_DONE] = 0.0;
    scal[SC_NLC] = 0.0;
    scal[SC_RL20] = rl20;
  }
}

__global__ void k_fill(int n, double *a, double v) {
  int k = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (k < n) a[k] = v;
}

// bcv on stand-alone arrays
__global__ void k_bcv(Incidence g, double *u, double *v, double *w, const double *ue, double flag) {
  int i = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (i >= g.n) return;
  double a = u[i], b = v[i], c = w[i];
  bcv_node<double>(g, i, a, b, c, ue, flag);
  u[i] = a; v[i] = b; w[i] = c;
}

// ------------------------------------------------------------------------------------
// MPDATA (reference mpdatam :3029-3311), NF fields advected by one momentum field.
// T is the arithmetic type of fluxes, antidiffusive velocities and limiter; the fields
// themselves stay FP64 and are updated by increments computed in T.
// ------------------------------------------------------------------------------------
// (x, y) pair of T (double2 / float2)
template <typename T> struct Pair;
template <> struct Pair<double> { typedef double2 type; };
template <> struct Pair<float> { typedef float2 type; };

template <typename T, int NF>
struct MpArgs {
  typedef typename Pair<T>::type T2;
  const T *U, *V, *W;              // advecting momentum uuu
  const double *psi[NF];           // fields in (pass 1) / donor-cell fields (pass 2)
  double *psi1[NF];                // donor-cell result / final result
  T2 *gxy[NF];                     // gradients of the donor-cell fields: (x, y) pairs ...
  T *gz[NF];                       // ... and z
  T2 *cpn[NF];                     // limiter ratios (cp, cn) per node
  T2 *aflux[(NF + 1) / 2];         // antidiffusive velocity per edge, fields in pairs
  const T *vol, *rh0;              // T copies
  const T *vr;                     // vol*rh0, the same product formed once per node
  const double *ue;
  T afree[NF];                     // inflow value per field (0 => use ue)
  int use_ue[NF];
  T dt;
};

template <typename T>
__device__ __forceinline__ T edge_aun(T U1, T U2, T V1, T V2, T W1, T W2, T sx, T sy, T sz) {
  return (U1 + U2) * (T)0.5 * sx + (V1 + V2) * (T)0.5 * sy + (W1 + W2) * (T)0.5 * sz;
}

template <typename T>
__device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T>
__device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }

// Edge metric of the MPDATA sweeps: FP64 from the (Sx, Sy) pair copy, FP32 from its own
// arrays (the same values either way)
template <typename T>
__device__ __forceinline__ void mp_metric(const Incidence &g, const EdgeS<T> &S, int e, T &sx, T &sy, T &sz) {
  if constexpr (sizeof(T) == 8) {
    const double2 v = g.sxy[e];
    sx = v.x; sy = v.y; sz = g.szz[e];
  } else {
    sx = S.sx[e]; sy = S.sy[e]; sz = S.sz[e];
  }
}

// The MPDATA sweeps gather per node over its slots in the reference's edge order. The
// edge quantities are written with the edge's (i1, i2) roles where they are not symmetric
// (p2 - p1, the limiter's ratios); sums and products of the two ends are symmetric in IEEE
// arithmetic, so both end nodes of an edge obtain the reference's bits. The antidiffusive
// velocity, the costly one (three divisions per field), is evaluated once per edge by its
// first node and stored; the bounds and limiter sweeps read it. Each node's own
// values are read once, before its slot loop. As in the gradient and divergence sweeps,
// every lane of a wavefront walks the chunk's full slot width with clamped (always valid)
// slot entries, the next batch's entries are loaded while the current batch's gathers are
// in flight, and slots past a node's degree are dropped by selects, not branches, so the
// wavefront never diverges around its loads.
#ifndef NFT_BATCH_MP
#define NFT_BATCH_MP 2   // slots per gather batch of the MPDATA sweeps (1 and 2 measured: 2 best)
#endif
#ifndef NFT_BATCH_MPL
#define NFT_BATCH_MPL 2  // the same for the bounds and limiter sweeps (stored fluxes, few registers)
#endif

// the per-slot gathers common to the sweeps: neighbour j, signed edge e, metric (sx, sy, sz)
template <typename T, bool COL, int B>
__device__ __forceinline__ void mp_batch_slots(const Incidence &g, const EdgeS<T> &S, const ColLane &C,
                                               const int2 (&rn)[B], int (&jj)[B], int (&ee)[B], T (&sx)[B],
                                               T (&sy)[B], T (&sz)[B]) {
#pragma unroll
  for (int u = 0; u < B; ++u) {
    slot_decode<COL>(C, rn[u], jj[u], ee[u]);
    const int e = ee[u] >= 0 ? ee[u] : ~ee[u];
    mp_metric<T>(g, S, e, sx[u], sy[u], sz[u]);
  }
}

// pass 1: donor-cell fluxes, boundary closure and update (reference :3094-3174)
// Per-node body of the donor-cell sweep; rd(j, U, V, W, psi) reads the advector and the
// fields at neighbour j (from memory, or from the patch's LDS copy)
template <typename T, int NF, bool COL, typename RD>
__device__ __forceinline__ void mp_donor_node(const Incidence &g, const EdgeS<T> &S, const Lane &L, const ColLane &C,
                                              const MpArgs<T, NF> &a, RD rd) {
  constexpr int B = NFT_BATCH_MP;
  const int i = L.i;
  T Ui = a.U[i], Vi = a.V[i], Wi = a.W[i];
  T pi[NF], arh2[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) { pi[f] = (T)a.psi[f][i]; arh2[f] = (T)0; }
  T divv = (T)0;
  int2 rn[B];
#pragma unroll
  for (int u = 0; u < B; ++u) rn[u] = slot_raw<COL>(g, L, C, u);
  for (int m0 = 0; m0 < L.width; m0 += B) {
    int jj[B], ee[B];
    T sx[B], sy[B], sz[B];
    mp_batch_slots<T, COL, B>(g, S, C, rn, jj, ee, sx, sy, sz);
    if (m0 + B < L.width) {
#pragma unroll
      for (int u = 0; u < B; ++u) rn[u] = slot_raw<COL>(g, L, C, m0 + B + u);
    }
    T Uj[B], Vj[B], Wj[B], pj[B][NF];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      rd(jj[u], Uj[u], Vj[u], Wj[u], pj[u]);
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const bool ok = m0 + u < L.deg;
      const bool first = ee[u] >= 0;
      T aun = edge_aun<T>(Ui, Uj[u], Vi, Vj[u], Wi, Wj[u], sx[u], sy[u], sz[u]);
      T ap = tmax<T>((T)0, aun), an = tmin<T>((T)0, aun);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        T fl = first ? pi[f] * ap + pj[u][f] * an : pj[u][f] * ap + pi[f] * an;
        T nv = first ? arh2[f] + fl : arh2[f] - fl;
        arh2[f] = ok ? nv : arh2[f];
      }
      T nd = first ? divv + aun : divv - aun;
      divv = ok ? nd : divv;
    }
  }
  for (int b = g.bp_ptr[i]; b < g.bp_ptr[i + 1]; ++b) {
    int fl_ = g.bp_flag[b];
    T aunbb1 = -divv;
    // the periodic box (mpdatamper :3733-3739) has no boundary-point flux
    bool slip = g.per || (fl_ == 11 || fl_ == 5 || fl_ == 6);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      T af = a.use_ue[f] ? (T)a.ue[i] : a.afree[f];
      T fl1 = slip ? (T)0 : af * tmin<T>((T)0, aunbb1) + pi[f] * tmax<T>((T)0, aunbb1);
      arh2[f] = arh2[f] + fl1 * (T)1;
    }
  }
  if (g.per) {
    const int cs = g.cslot[i];
#pragma unroll
    for (int f = 0; f < NF; ++f) arh2[f] = cs >= 0 ? (T)g.cval[NF * cs + f] : arh2[f];
  }
  T vl = a.vol[i], rh = a.rh0[i];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    T inc = arh2[f] * a.dt / vl / rh;
    a.psi1[f][i] = a.psi[f][i] - (double)inc;
  }
}

template <typename T, int NF, bool COL>
__global__ void __launch_bounds__(NFT_BLOCK) k_mp_donor(Incidence g, EdgeS<T> S, MpArgs<T, NF> a) {
  Lane L;
  if (!lane_setup(g, L)) return;
  ColLane C{};
  if (COL) col_setup(g, L.valid ? L.i : g.n - 1, C);
  if (!L.valid) return;
  mp_donor_node<T, NF, COL>(g, S, L, C, a, [&](int j, T &U, T &V, T &W, T (&p)[NF]) {
    U = a.U[j]; V = a.V[j]; W = a.W[j];
#pragma unroll
    for (int f = 0; f < NF; ++f) p[f] = (T)a.psi[f][j];
  });
}

// gradients of the donor-cell fields (derivative x3 per field, reference :3182-3184)
// Per-node body of the gradient sweep of the donor-cell fields; rd(j, f) reads field f at
// neighbour j (from memory, or from the patch's LDS copy)
template <typename T, int NF, bool COL, typename RD>
__device__ __forceinline__ void mp_grad_node(const Incidence &g, const EdgeS<T> &S, const Lane &L, const ColLane &C,
                                             const MpArgs<T, NF> &a, RD rd) {
  constexpr int B = NFT_BATCH_MP;
  const int i = L.i;
  T pi[NF], gx[NF], gy[NF], gz[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) { pi[f] = (T)a.psi1[f][i]; gx[f] = gy[f] = gz[f] = (T)0; }
  int2 rn[B];
#pragma unroll
  for (int u = 0; u < B; ++u) rn[u] = slot_raw<COL>(g, L, C, u);
  for (int m0 = 0; m0 < L.width; m0 += B) {
    int jj[B], ee[B];
    T sx[B], sy[B], sz[B];
    mp_batch_slots<T, COL, B>(g, S, C, rn, jj, ee, sx, sy, sz);
    if (m0 + B < L.width) {
#pragma unroll
      for (int u = 0; u < B; ++u) rn[u] = slot_raw<COL>(g, L, C, m0 + B + u);
    }
    T pj[B][NF];
#pragma unroll
    for (int u = 0; u < B; ++u) {
#pragma unroll
      for (int f = 0; f < NF; ++f) pj[u][f] = rd(jj[u], f);
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const bool ok = m0 + u < L.deg;
      const bool first = ee[u] >= 0;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        T uu = (pi[f] + pj[u][f]) * (T)0.5;
        T tx = sx[u] * uu, ty = sy[u] * uu, tz = sz[u] * uu;
        T nx = first ? gx[f] + tx : gx[f] - tx;
        T ny = first ? gy[f] + ty : gy[f] - ty;
        T nz = first ? gz[f] + tz : gz[f] - tz;
        gx[f] = ok ? nx : gx[f]; gy[f] = ok ? ny : gy[f]; gz[f] = ok ? nz : gz[f];
      }
    }
  }
  for (int b = g.bp_ptr[i]; b < (g.per ? 0 : g.bp_ptr[i + 1]); ++b) {
    T bx = (T)g.bp_sbb[b], by = (T)g.bp_sbb[g.nbp + b], bz = (T)g.bp_sbb[2 * g.nbp + b];
#pragma unroll
    for (int f = 0; f < NF; ++f) { gx[f] = gx[f] + bx * pi[f]; gy[f] = gy[f] + by * pi[f]; gz[f] = gz[f] + bz * pi[f]; }
  }
  if (g.per) {
    const int cs = g.cslot[i];
    if (cs >= 0) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        gx[f] = (T)g.cval[3 * NF * cs + 3 * f]; gy[f] = (T)g.cval[3 * NF * cs + 3 * f + 1];
        gz[f] = (T)g.cval[3 * NF * cs + 3 * f + 2];
      }
    }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    typename MpArgs<T, NF>::T2 v;
    v.x = gx[f]; v.y = gy[f];
    a.gxy[f][i] = v;
    a.gz[f][i] = gz[f];
  }
}

template <typename T, int NF, bool COL>
__global__ void __launch_bounds__(NFT_BLOCK) k_mp_grad(Incidence g, EdgeS<T> S, MpArgs<T, NF> a) {
  Lane L;
  if (!lane_setup(g, L)) return;
  ColLane C{};
  if (COL) col_setup(g, L.valid ? L.i : g.n - 1, C);
  if (!L.valid) return;
  mp_grad_node<T, NF, COL>(g, S, L, C, a, [&](int j, int f) { return (T)a.psi1[f][j]; });
}

// antidiffusive velocity of one edge for one field (reference :3187-3204) from the two end
// nodes' donor-cell values p1, p2 (edge roles i1, i2) and gradient sums gsx = gx_i1 + gx_i2, ...
template <typename T>
__device__ __forceinline__ T anti_flux(T p1, T p2, T gsx, T gsy, T gsz, T aun, T volt, T U12, T V12, T W12, T dt) {
  T un = fabs(aun) * (p2 - p1) * (T)0.5;
  T dpdx = (T)0.5 * gsx / volt;
  T dpdy = (T)0.5 * gsy / volt;
  T dpdz = (T)0.5 * gsz / volt;
  T term = U12 * (T)0.5 * dpdx + V12 * (T)0.5 * dpdy + W12 * (T)0.5 * dpdz;
  return un - aun * term * dt * (T)0.5;
}

// Antidiffusive velocities of the edges (reference :3187-3221), each evaluated once: a node
// walks the owned incidence (only the slots where it is the edge's first node i1), so every
// edge is taken exactly once, and stores the edge's value of every field at the edge index:
// aflux[0][e] = (field 0, field 1), aflux[1][e] = (field 2, field 3). The bounds and limiter
// sweeps read it from there in both end nodes' slot loops.
template <typename T, int NF, bool COL>
__global__ void __launch_bounds__(NFT_BLOCK) k_mp_flux(Incidence g, EdgeS<T> S, MpArgs<T, NF> a) {
  constexpr int B = NFT_BATCH_MP;
  typedef typename MpArgs<T, NF>::T2 T2;
  Lane L;
  if (!lane_setup(g, L)) return;
  ColLane C{};
  if (COL) col_setup(g, L.valid ? L.i : g.n - 1, C);
  if (!L.valid) return;
  const int i = L.i;
  T Ui = a.U[i], Vi = a.V[i], Wi = a.W[i];
  T vri = a.vr[i];
  T pi[NF], gxi[NF], gyi[NF], gzi[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    pi[f] = (T)a.psi1[f][i];
    const T2 g2 = a.gxy[f][i];
    gxi[f] = g2.x; gyi[f] = g2.y; gzi[f] = a.gz[f][i];
  }
  int2 rn[B];
#pragma unroll
  for (int u = 0; u < B; ++u) rn[u] = slot_raw<COL>(g, L, C, u);
  for (int m0 = 0; m0 < L.width; m0 += B) {
    int jj[B], ee[B];
    T sx[B], sy[B], sz[B];
    mp_batch_slots<T, COL, B>(g, S, C, rn, jj, ee, sx, sy, sz);
    if (m0 + B < L.width) {
#pragma unroll
      for (int u = 0; u < B; ++u) rn[u] = slot_raw<COL>(g, L, C, m0 + B + u);
    }
    T Uj[B], Vj[B], Wj[B], vrj[B], pj[B][NF], gjx[B][NF], gjy[B][NF], gjz[B][NF];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int j = jj[u];
      Uj[u] = a.U[j]; Vj[u] = a.V[j]; Wj[u] = a.W[j]; vrj[u] = a.vr[j];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        pj[u][f] = (T)a.psi1[f][j];
        const T2 gj = a.gxy[f][j];
        gjx[u][f] = gj.x; gjy[u][f] = gj.y; gjz[u][f] = a.gz[f][j];
      }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      // every slot of the owned incidence has this node as i1
      T aun = edge_aun<T>(Ui, Uj[u], Vi, Vj[u], Wi, Wj[u], sx[u], sy[u], sz[u]);
      T volt = (T)0.5 * (vri + vrj[u]);
      T U12 = Ui + Uj[u], V12 = Vi + Vj[u], W12 = Wi + Wj[u];
      T A[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f)
        A[f] = anti_flux<T>(pi[f], pj[u][f], gxi[f] + gjx[u][f], gyi[f] + gjy[u][f], gzi[f] + gjz[u][f], aun, volt,
                            U12, V12, W12, a.dt);
      if (m0 + u < L.deg) {
#pragma unroll
        for (int q = 0; q < (NF + 1) / 2; ++q) {
          T2 v;
          v.x = A[2 * q];
          v.y = 2 * q + 1 < NF ? A[2 * q + 1] : (T)0;
          a.aflux[q][ee[u]] = v;
        }
      }
    }
  }
}

// the antidiffusive velocities of one slot's edge, all fields, from the flux sweep's store
template <typename T, int NF>
__device__ __forceinline__ void mp_load_flux(const MpArgs<T, NF> &a, int e, T (&A)[NF]) {
#pragma unroll
  for (int q = 0; q < (NF + 1) / 2; ++q) {
    const typename MpArgs<T, NF>::T2 v = a.aflux[q][e];
    A[2 * q] = v.x;
    if (2 * q + 1 < NF) A[2 * q + 1] = v.y;
  }
}

// FCT bounds from the donor-cell solution, in/out fluxes and ratios (reference :3228-3280)
// Per-node body of the bounds sweep; rd(j, f) reads the donor-cell value of field f at
// neighbour j (from memory, or from the patch's LDS copy, where j is the LDS index)
template <typename T, int NF, bool COL, typename RD>
__device__ __forceinline__ void mp_fct_node(const Incidence &g, const Lane &L, const ColLane &C, const MpArgs<T, NF> &a,
                                            RD rd) {
  constexpr int B = NFT_BATCH_MPL;
  typedef typename MpArgs<T, NF>::T2 T2;
  const int i = L.i;
  T mx[NF], mn[NF], rin[NF], rout[NF], pi[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    mx[f] = (T)-1.e15; mn[f] = (T)1.e15; rin[f] = (T)0; rout[f] = (T)0;
    pi[f] = (T)a.psi1[f][i];
  }
  int2 rn[B];
#pragma unroll
  for (int u = 0; u < B; ++u) rn[u] = slot_raw<COL>(g, L, C, u);
  for (int m0 = 0; m0 < L.width; m0 += B) {
    int jj[B], ee[B];
#pragma unroll
    for (int u = 0; u < B; ++u) slot_decode<COL>(C, rn[u], jj[u], ee[u]);
    if (m0 + B < L.width) {
#pragma unroll
      for (int u = 0; u < B; ++u) rn[u] = slot_raw<COL>(g, L, C, m0 + B + u);
    }
    T pj[B][NF], Af[B][NF];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      mp_load_flux<T, NF>(a, ee[u] >= 0 ? ee[u] : ~ee[u], Af[u]);
#pragma unroll
      for (int f = 0; f < NF; ++f) pj[u][f] = rd(jj[u], f);
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const bool ok = m0 + u < L.deg;
      const bool first = ee[u] >= 0;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const T p = pj[u][f], A = Af[u][f];
        T nmx = tmax<T>(tmax<T>(mx[f], pi[f]), p);
        T nmn = tmin<T>(tmin<T>(mn[f], pi[f]), p);
        T pp = tmax<T>((T)0, A), pn = -tmin<T>((T)0, A);
        T nin = rin[f] + (first ? pn : pp);
        T nout = rout[f] + (first ? pp : pn);
        mx[f] = ok ? nmx : mx[f]; mn[f] = ok ? nmn : mn[f];
        rin[f] = ok ? nin : rin[f]; rout[f] = ok ? nout : rout[f];
      }
    }
  }
  if (g.per) {
    // bounds merged and in/out sums closed over the coincident set (mpdatamper :3831-3902)
    const int cs = g.cslot[i];
    if (cs >= 0) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double *q = g.cval + 4 * NF * (size_t)cs + 4 * f;
        mx[f] = (T)q[0]; mn[f] = (T)q[1]; rin[f] = (T)q[2]; rout[f] = (T)q[3];
      }
    }
  }
  T vl = a.vol[i], rh = a.rh0[i];
  const T tol = (T)1.e-10;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    T2 r;
    r.x = (mx[f] - pi[f]) * vl * rh / (rin[f] * a.dt + tol);
    r.y = (pi[f] - mn[f]) * vl * rh / (rout[f] * a.dt + tol);
    a.cpn[f][i] = r;
  }
}


template <typename T, int NF, bool COL>
__global__ void __launch_bounds__(NFT_BLOCK) k_mp_fct(Incidence g, MpArgs<T, NF> a) {
  Lane L;
  if (!lane_setup(g, L)) return;
  ColLane C{};
  if (COL) col_setup(g, L.valid ? L.i : g.n - 1, C);
  if (!L.valid) return;
  mp_fct_node<T, NF, COL>(g, L, C, a, [&](int j, int f) { return (T)a.psi1[f][j]; });
}

// limited antidiffusive fluxes and the second update (reference :3284-3302, :3098-3174)
// Per-node body of the limiter sweep; rd(j, f) reads the (cp, cn) pair of field f at
// neighbour j (from memory, or from the patch's LDS copy)
template <typename T, int NF, bool COL, typename RD>
__device__ __forceinline__ void mp_limit_node(const Incidence &g, const Lane &L, const ColLane &C,
                                              const MpArgs<T, NF> &a, double *const *out, RD rd) {
  constexpr int B = NFT_BATCH_MPL;
  typedef typename MpArgs<T, NF>::T2 T2;
  const int i = L.i;
  T arh2[NF], cpi[NF], cni[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    arh2[f] = (T)0;
    const T2 c2 = a.cpn[f][i];
    cpi[f] = c2.x; cni[f] = c2.y;
  }
  int2 rn[B];
#pragma unroll
  for (int u = 0; u < B; ++u) rn[u] = slot_raw<COL>(g, L, C, u);
  for (int m0 = 0; m0 < L.width; m0 += B) {
    int jj[B], ee[B];
#pragma unroll
    for (int u = 0; u < B; ++u) slot_decode<COL>(C, rn[u], jj[u], ee[u]);
    if (m0 + B < L.width) {
#pragma unroll
      for (int u = 0; u < B; ++u) rn[u] = slot_raw<COL>(g, L, C, m0 + B + u);
    }
    T Af[B][NF], cpj[B][NF], cnj[B][NF];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      mp_load_flux<T, NF>(a, ee[u] >= 0 ? ee[u] : ~ee[u], Af[u]);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const T2 cj = rd(jj[u], f);
        cpj[u][f] = cj.x; cnj[u][f] = cj.y;
      }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const bool ok = m0 + u < L.deg;
      const bool first = ee[u] >= 0;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const T A = Af[u][f];
        // edge roles: cp/cn of i2 and i1
        const T cp2 = first ? cpj[u][f] : cpi[f], cn1 = first ? cni[f] : cnj[u][f];
        const T cn2 = first ? cnj[u][f] : cni[f], cp1 = first ? cpi[f] : cpj[u][f];
        T lim = A > (T)0 ? tmin<T>(tmin<T>((T)1, cp2), cn1) : tmin<T>(tmin<T>((T)1, cn2), cp1);
        T fl = A * lim;
        T nv = first ? arh2[f] + fl : arh2[f] - fl;
        arh2[f] = ok ? nv : arh2[f];
      }
    }
  }
  for (int b = g.bp_ptr[i]; b < g.bp_ptr[i + 1]; ++b) {
#pragma unroll
    for (int f = 0; f < NF; ++f) arh2[f] = arh2[f] + (T)0 * (T)1;
  }
  if (g.per) {
    const int cs = g.cslot[i];
#pragma unroll
    for (int f = 0; f < NF; ++f) arh2[f] = cs >= 0 ? (T)g.cval[NF * cs + f] : arh2[f];
  }
  T vl = a.vol[i], rh = a.rh0[i];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    T inc = arh2[f] * a.dt / vl / rh;
    out[f][i] = a.psi1[f][i] - (double)inc;
  }
}

template <typename T, int NF, bool COL>
__global__ void __launch_bounds__(NFT_BLOCK, 7) k_mp_limit(Incidence g, MpArgs<T, NF> a, double *const *out) {
  Lane L;
  if (!lane_setup(g, L)) return;
  ColLane C{};
  if (COL) col_setup(g, L.valid ? L.i : g.n - 1, C);
  if (!L.valid) return;
  mp_limit_node<T, NF, COL>(g, L, C, a, out, [&](int j, int f) { return a.cpn[f][j]; });
}

// ------------------------------------------------------------------------------------
// LDS-staged MPDATA sweeps over column patches (columns of NFT_WAVE levels): donor cell,
// gradients, bounds and limiter
// ------------------------------------------------------------------------------------
// A patch is up to NFT_PATCH_COLS columns, one wavefront each, in one workgroup. Its
// columns and every column their slots reach (col[ptr[p]] .. col[ptr[p+1]-1], the patch's
// own nown[p] columns first) are copied to LDS with coalesced loads, and the slot loops read
// neighbour values from that copy: the incidence passed in has tab entries rewritten to
// LDS-relative neighbours (staged index * NFT_WAVE + level, decoded exactly as the node
// ids), so each node's slot order and arithmetic are unchanged.
#ifndef NFT_PATCH_COLS
#define NFT_PATCH_COLS 8
#endif
#ifndef NFT_PATCH_STAGE
#define NFT_PATCH_STAGE 20   // staged columns per patch at most (at most 80 KB of LDS at FP64: two patches per CU)
#endif
// staged nodes per thread at most
#define NFT_PATCH_SU ((NFT_PATCH_STAGE * NFT_WAVE + NFT_WAVE * NFT_PATCH_COLS - 1) / (NFT_WAVE * NFT_PATCH_COLS))

struct MpPatches {
  int npatch;
  int xcd_per;               // patches per XCD slice (block remap), 0 = identity
  int stride;                // LDS values per field: the largest staged set * NFT_WAVE
  const int32_t *ptr;        // [npatch+1] into col
  const int32_t *col;        // staged columns
  const uint8_t *nown;       // [npatch] columns owned by the patch
};

__device__ __forceinline__ int patch_of_block(const MpPatches &P) {
  const int b = blockIdx.x;
  return P.xcd_per > 0 ? (b & 7) * P.xcd_per + (b >> 3) : b;
}

// the lane of node (column c, level = lane) in a whole-column wavefront
__device__ __forceinline__ void patch_lane(const Incidence &g, int c, Lane &L) {
  L.lane = threadIdx.x & 63;
  L.i = c * NFT_WAVE + L.lane;
  L.base = __builtin_amdgcn_readfirstlane(g.chunk_ptr[c]);
  L.width = __builtin_amdgcn_readfirstlane((g.chunk_ptr[c + 1] - L.base) >> 6);
  L.valid = true;
  L.deg = g.deg[L.i];
}

template <typename T, int NF>
__global__ void __launch_bounds__(NFT_WAVE *NFT_PATCH_COLS) k_mp_fct_lds(Incidence g, MpPatches P, MpArgs<T, NF> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *sp = (T *)lds_raw;   // [field][staged column * NFT_WAVE + level]: lanes read consecutive words
  const int p = patch_of_block(P);
  if (p >= P.npatch) return;
  const int s0 = P.ptr[p], ns = P.ptr[p + 1] - s0;
  {
    // every staging load in flight before the first LDS store (clamped, always valid)
    T v[NFT_PATCH_SU][NF];
#pragma unroll
    for (int r = 0; r < NFT_PATCH_SU; ++r) {
      const int q = min((int)threadIdx.x + r * NFT_WAVE * NFT_PATCH_COLS, ns * NFT_WAVE - 1);
      const int node = P.col[s0 + (q >> 6)] * NFT_WAVE + (q & (NFT_WAVE - 1));
#pragma unroll
      for (int f = 0; f < NF; ++f) v[r][f] = (T)a.psi1[f][node];
    }
#pragma unroll
    for (int r = 0; r < NFT_PATCH_SU; ++r) {
      const int q = threadIdx.x + r * NFT_WAVE * NFT_PATCH_COLS;
      if (q < ns * NFT_WAVE) {
#pragma unroll
        for (int f = 0; f < NF; ++f) sp[f * P.stride + q] = v[r][f];
      }
    }
  }
  __syncthreads();
  const int w = threadIdx.x >> 6;
  if (w >= P.nown[p]) return;
  Lane L;
  patch_lane(g, P.col[s0 + w], L);
  ColLane C{};
  col_setup(g, L.i, C);
  mp_fct_node<T, NF, true>(g, L, C, a, [&](int j, int f) { return sp[f * P.stride + j]; });
}

template <typename T, int NF>
__global__ void __launch_bounds__(NFT_WAVE *NFT_PATCH_COLS) k_mp_limit_lds(Incidence g, MpPatches P, MpArgs<T, NF> a,
                                                                         double *const *out) {
  typedef typename MpArgs<T, NF>::T2 T2;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T2 *sc = (T2 *)lds_raw;   // [field][staged column * NFT_WAVE + level]
  const int p = patch_of_block(P);
  if (p >= P.npatch) return;
  const int s0 = P.ptr[p], ns = P.ptr[p + 1] - s0;
  {
    T vx[NFT_PATCH_SU][NF], vy[NFT_PATCH_SU][NF];   // (cp, cn) as scalars: kept in registers
#pragma unroll
    for (int r = 0; r < NFT_PATCH_SU; ++r) {
      const int q = min((int)threadIdx.x + r * NFT_WAVE * NFT_PATCH_COLS, ns * NFT_WAVE - 1);
      const int node = P.col[s0 + (q >> 6)] * NFT_WAVE + (q & (NFT_WAVE - 1));
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const T2 v = a.cpn[f][node];
        vx[r][f] = v.x; vy[r][f] = v.y;
      }
    }
#pragma unroll
    for (int r = 0; r < NFT_PATCH_SU; ++r) {
      const int q = threadIdx.x + r * NFT_WAVE * NFT_PATCH_COLS;
      if (q < ns * NFT_WAVE) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          T2 v;
          v.x = vx[r][f]; v.y = vy[r][f];
          sc[f * P.stride + q] = v;
        }
      }
    }
  }
  __syncthreads();
  const int w = threadIdx.x >> 6;
  if (w >= P.nown[p]) return;
  Lane L;
  patch_lane(g, P.col[s0 + w], L);
  ColLane C{};
  col_setup(g, L.i, C);
  mp_limit_node<T, NF, true>(g, L, C, a, out, [&](int j, int f) { return sc[f * P.stride + j]; });
}

// the gradient sweep of the donor-cell fields over the same patches
template <typename T, int NF>
__global__ void __launch_bounds__(NFT_WAVE *NFT_PATCH_COLS) k_mp_grad_lds(Incidence g, EdgeS<T> S, MpPatches P,
                                                                        MpArgs<T, NF> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *sp = (T *)lds_raw;   // [field][staged column * NFT_WAVE + level]
  const int p = patch_of_block(P);
  if (p >= P.npatch) return;
  const int s0 = P.ptr[p], ns = P.ptr[p + 1] - s0;
  {
    T v[NFT_PATCH_SU][NF];
#pragma unroll
    for (int r = 0; r < NFT_PATCH_SU; ++r) {
      const int q = min((int)threadIdx.x + r * NFT_WAVE * NFT_PATCH_COLS, ns * NFT_WAVE - 1);
      const int node = P.col[s0 + (q >> 6)] * NFT_WAVE + (q & (NFT_WAVE - 1));
#pragma unroll
      for (int f = 0; f < NF; ++f) v[r][f] = (T)a.psi1[f][node];
    }
#pragma unroll
    for (int r = 0; r < NFT_PATCH_SU; ++r) {
      const int q = threadIdx.x + r * NFT_WAVE * NFT_PATCH_COLS;
      if (q < ns * NFT_WAVE) {
#pragma unroll
        for (int f = 0; f < NF; ++f) sp[f * P.stride + q] = v[r][f];
      }
    }
  }
  __syncthreads();
  const int w = threadIdx.x >> 6;
  if (w >= P.nown[p]) return;
  Lane L;
  patch_lane(g, P.col[s0 + w], L);
  ColLane C{};
  col_setup(g, L.i, C);
  mp_grad_node<T, NF, true>(g, S, L, C, a, [&](int j, int f) { return sp[f * P.stride + j]; });
}

// the donor-cell sweep over the same patches: advector and fields staged in LDS
template <typename T, int NF>
__global__ void __launch_bounds__(NFT_WAVE *NFT_PATCH_COLS) k_mp_donor_lds(Incidence g, EdgeS<T> S, MpPatches P,
                                                                         MpArgs<T, NF> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *sv = (T *)lds_raw;   // [U, V, W, field 0 .. NF-1][staged column * NFT_WAVE + level]
  const int p = patch_of_block(P);
  if (p >= P.npatch) return;
  const int s0 = P.ptr[p], ns = P.ptr[p + 1] - s0;
  {
    T v[NFT_PATCH_SU][3 + NF];
#pragma unroll
    for (int r = 0; r < NFT_PATCH_SU; ++r) {
      const int q = min((int)threadIdx.x + r * NFT_WAVE * NFT_PATCH_COLS, ns * NFT_WAVE - 1);
      const int node = P.col[s0 + (q >> 6)] * NFT_WAVE + (q & (NFT_WAVE - 1));
      v[r][0] = a.U[node]; v[r][1] = a.V[node]; v[r][2] = a.W[node];
#pragma unroll
      for (int f = 0; f < NF; ++f) v[r][3 + f] = (T)a.psi[f][node];
    }
#pragma unroll
    for (int r = 0; r < NFT_PATCH_SU; ++r) {
      const int q = threadIdx.x + r * NFT_WAVE * NFT_PATCH_COLS;
      if (q < ns * NFT_WAVE) {
#pragma unroll
        for (int k = 0; k < 3 + NF; ++k) sv[k * P.stride + q] = v[r][k];
      }
    }
  }
  __syncthreads();
  const int w = threadIdx.x >> 6;
  if (w >= P.nown[p]) return;
  Lane L;
  patch_lane(g, P.col[s0 + w], L);
  ColLane C{};
  col_setup(g, L.i, C);
  mp_donor_node<T, NF, true>(g, S, L, C, a, [&](int j, T &U, T &V, T &W, T (&pj)[NF]) {
    U = sv[j]; V = sv[P.stride + j]; W = sv[2 * P.stride + j];
#pragma unroll
    for (int f = 0; f < NF; ++f) pj[f] = sv[(3 + f) * P.stride + j];
  });
}

// ------------------------------------------------------------------------------------
// Periodic box (the reference's iper = 2): set kernels. Per coincident set (cyclicr pairs,
// cyclicr2 pairs, cyclicr4 quadruples; the sets are disjoint) one thread, or one per member
// in k_per_sums3, the kernel of every laplc/div/prfr, forms each member's gathered sums exactly as the gather kernel that follows would (the member's
// SELL slots in the reference's edge order, the same operations), sums them over the set
// in the set's order (a + b; ((a + b) + c) + d) and stores the result for every member in
// cval; the gather kernel then takes these sums at the set's nodes. The sets hold the two
// or four copies of a node on opposite sides of the box, so the work is a thin layer of
// side columns.
// ------------------------------------------------------------------------------------
struct PerSets {
  int nset;
  const int4 *sets;       // 0-based member nodes in the set's order; z = w = -1 for pairs
  const int32_t *cslot;   // node -> slot in cval, -1 outside the sets
  double *cval;
};

__device__ __forceinline__ int set_size(const int4 &s) { return s.z >= 0 ? 4 : 2; }
__device__ __forceinline__ int set_member(const int4 &s, int q) {
  return q == 0 ? s.x : q == 1 ? s.y : q == 2 ? s.z : s.w;
}

// slot m of node j in the SELL form
__device__ __forceinline__ void sell_slot(const Incidence &g, int j, int m, int &nb, int &e) {
  const int slot = g.chunk_ptr[j >> 6] + m * NFT_WAVE + (j & 63);
  nb = g.nbr[slot];
  e = g.eid[slot];
}

// v[q][k] of the set's members -> the set's sums, stored for every member
template <int K>
__device__ __forceinline__ void per_close_store(const PerSets &p, const int4 &s, double (&v)[4][K]) {
  const int cnt = set_size(s);
  double c[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    c[k] = v[0][k] + v[1][k];
    if (cnt == 4) c[k] = (c[k] + v[2][k]) + v[3][k];
  }
  for (int q = 0; q < cnt; ++q) {
    double *d = p.cval + (size_t)K * p.cslot[set_member(s, q)];
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = c[k];
  }
}

// Three gathered sums per node, sum over slots of +-S_c * ((q_c,i + q_c,j) * 0.5): the
// gradient of one field (q0 = q1 = q2, k_grad) or the divergence of three (k_div; packed
// (x, y) inputs when qxy is set, each input times rh0 when scale is set).
struct PerSums3 {
  const double *q0, *q1, *q2;
  const double2 *qxy;
  const double *scale;
  const double *skip;
};

// four threads per set, one per member (pairs leave two idle): each gathers its member's
// sums, then every member thread closes the set's sums from LDS and stores its own copy
#define NFT_SETS_PER_BLOCK (NFT_BLOCK / 4)
__global__ void __launch_bounds__(NFT_BLOCK) k_per_sums3(Incidence g, PerSets p, PerSums3 a) {
  __shared__ double sh[3][NFT_BLOCK];
  if (skipped(a.skip)) return;   // uniform: every thread leaves before the barrier
  const int t = blockIdx.x * NFT_SETS_PER_BLOCK + (threadIdx.x >> 2), q = threadIdx.x & 3;
  const bool on = t < p.nset;
  const int4 s = on ? p.sets[t] : make_int4(-1, -1, -1, -1);
  const int cnt = on ? set_size(s) : 0;
  double ax = 0., ay = 0., az = 0.;
  if (q < cnt) {
    const int i = set_member(s, q);
    double qi0, qi1, qi2;
    if (a.qxy) { const double2 x = a.qxy[i]; qi0 = x.x; qi1 = x.y; }
    else { qi0 = a.q0[i]; qi1 = a.q1[i]; }
    qi2 = a.q2[i];
    if (a.scale) { const double r0 = a.scale[i]; qi0 = qi0 * r0; qi1 = qi1 * r0; qi2 = qi2 * r0; }
    const int deg = g.deg[i];
    for (int m = 0; m < deg; ++m) {
      int j, e;
      sell_slot(g, i, m, j, e);
      const int ea = e >= 0 ? e : ~e;
      double q0, q1, q2;
      if (a.qxy) { const double2 x = a.qxy[j]; q0 = x.x; q1 = x.y; }
      else { q0 = a.q0[j]; q1 = a.q1[j]; }
      q2 = a.q2[j];
      if (a.scale) { const double rj = a.scale[j]; q0 = q0 * rj; q1 = q1 * rj; q2 = q2 * rj; }
      const double2 sxy = g.sxy[ea];
      const double sz = g.szz[ea];
      const double u0 = (qi0 + q0) * 0.5, u1 = (qi1 + q1) * 0.5, u2 = (qi2 + q2) * 0.5;
      const double tx = sxy.x * u0, ty = sxy.y * u1, tz = sz * u2;
      ax = e >= 0 ? ax + tx : ax - tx;
      ay = e >= 0 ? ay + ty : ay - ty;
      az = e >= 0 ? az + tz : az - tz;
    }
  }
  sh[0][threadIdx.x] = ax; sh[1][threadIdx.x] = ay; sh[2][threadIdx.x] = az;
  __syncthreads();
  if (q < cnt) {
    const int b = threadIdx.x & ~3;
    double* d = p.cval + (size_t)3 * p.cslot[set_member(s, q)];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double c = sh[k][b] + sh[k][b + 1];
      if (cnt == 4) c = (c + sh[k][b + 2]) + sh[k][b + 3];
      d[k] = c;
    }
  }
}

// mpdatamper pass 1 (:3666-3743): donor-cell increments of NF fields, the boundary points'
// zero fluxes added as the reference adds them
template <int NF>
__global__ void __launch_bounds__(NFT_BLOCK) k_per_mp_donor(Incidence g, PerSets p, MpArgs<double, NF> a) {
  const int t = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (t >= p.nset) return;
  const int4 s = p.sets[t];
  double v[4][NF];
  for (int q = 0; q < set_size(s); ++q) {
    const int i = set_member(s, q);
    const double Ui = a.U[i], Vi = a.V[i], Wi = a.W[i];
    double pi[NF], arh2[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) { pi[f] = a.psi[f][i]; arh2[f] = 0.; }
    for (int m = 0; m < g.deg[i]; ++m) {
      int j, e;
      sell_slot(g, i, m, j, e);
      const int ea = e >= 0 ? e : ~e;
      const double2 sxy = g.sxy[ea];
      const bool first = e >= 0;
      const double aun = edge_aun<double>(Ui, a.U[j], Vi, a.V[j], Wi, a.W[j], sxy.x, sxy.y, g.szz[ea]);
      const double ap = tmax<double>(0., aun), an = tmin<double>(0., aun);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double pj = a.psi[f][j];
        const double fl = first ? pi[f] * ap + pj * an : pj * ap + pi[f] * an;
        arh2[f] = first ? arh2[f] + fl : arh2[f] - fl;
      }
    }
    for (int b = g.bp_ptr[i]; b < g.bp_ptr[i + 1]; ++b) {
#pragma unroll
      for (int f = 0; f < NF; ++f) arh2[f] = arh2[f] + 0. * 1.;
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) v[q][f] = arh2[f];
  }
  per_close_store<NF>(p, s, v);
}

// the periodic derivative x3 of the NF donor-cell fields (:3749-3751)
template <int NF>
__global__ void __launch_bounds__(NFT_BLOCK) k_per_mp_grad(Incidence g, PerSets p, MpArgs<double, NF> a) {
  const int t = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (t >= p.nset) return;
  const int4 s = p.sets[t];
  double v[4][3 * NF];
  for (int q = 0; q < set_size(s); ++q) {
    const int i = set_member(s, q);
    double pi[NF], gx[NF], gy[NF], gz[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) { pi[f] = a.psi1[f][i]; gx[f] = gy[f] = gz[f] = 0.; }
    for (int m = 0; m < g.deg[i]; ++m) {
      int j, e;
      sell_slot(g, i, m, j, e);
      const int ea = e >= 0 ? e : ~e;
      const double2 sxy = g.sxy[ea];
      const double sz = g.szz[ea];
      const bool first = e >= 0;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double uu = (pi[f] + a.psi1[f][j]) * 0.5;
        const double tx = sxy.x * uu, ty = sxy.y * uu, tz = sz * uu;
        gx[f] = first ? gx[f] + tx : gx[f] - tx;
        gy[f] = first ? gy[f] + ty : gy[f] - ty;
        gz[f] = first ? gz[f] + tz : gz[f] - tz;
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) { v[q][3 * f] = gx[f]; v[q][3 * f + 1] = gy[f]; v[q][3 * f + 2] = gz[f]; }
  }
  per_close_store<3 * NF>(p, s, v);
}

// FCT bounds and in/out sums of the set's members (:3806-3902): each member's bounds from
// its edges, merged over the set (a pair: the extremes of both nodes and both values; a
// quadruple: node by node the maximum of the four maxima and values, and as minimum the
// least of the four maxima and values, as the reference writes it); the in/out sums summed
// over the set.
template <int NF>
__global__ void __launch_bounds__(NFT_BLOCK) k_per_mp_fct(Incidence g, PerSets p, MpArgs<double, NF> a) {
  const int t = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (t >= p.nset) return;
  const int4 s = p.sets[t];
  const int cnt = set_size(s);
  double mx[4][NF], mn[4][NF], rio[4][2 * NF], pv[4][NF];
  for (int q = 0; q < cnt; ++q) {
    const int i = set_member(s, q);
    const double Ui = a.U[i], Vi = a.V[i], Wi = a.W[i], vri = a.vr[i];
    double pi[NF], gxi[NF], gyi[NF], gzi[NF], rin[NF], rout[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      pi[f] = a.psi1[f][i];
      pv[q][f] = pi[f];
      const double2 g2 = a.gxy[f][i];
      gxi[f] = g2.x; gyi[f] = g2.y; gzi[f] = a.gz[f][i];
      mx[q][f] = -1.e15; mn[q][f] = 1.e15; rin[f] = 0.; rout[f] = 0.;
    }
    for (int m = 0; m < g.deg[i]; ++m) {
      int j, e;
      sell_slot(g, i, m, j, e);
      const int ea = e >= 0 ? e : ~e;
      const double2 sxy = g.sxy[ea];
      const bool first = e >= 0;
      const double Uj = a.U[j], Vj = a.V[j], Wj = a.W[j];
      const double aun = edge_aun<double>(Ui, Uj, Vi, Vj, Wi, Wj, sxy.x, sxy.y, g.szz[ea]);
      const double volt = 0.5 * (vri + a.vr[j]);
      const double U12 = Ui + Uj, V12 = Vi + Vj, W12 = Wi + Wj;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double pj = a.psi1[f][j];
        const double2 gj = a.gxy[f][j];
        const double A = anti_flux<double>(first ? pi[f] : pj, first ? pj : pi[f], gxi[f] + gj.x, gyi[f] + gj.y,
                                           gzi[f] + a.gz[f][j], aun, volt, U12, V12, W12, a.dt);
        mx[q][f] = tmax<double>(tmax<double>(mx[q][f], pi[f]), pj);
        mn[q][f] = tmin<double>(tmin<double>(mn[q][f], pi[f]), pj);
        const double pp = tmax<double>(0., A), pn = -tmin<double>(0., A);
        rin[f] = rin[f] + (first ? pn : pp);
        rout[f] = rout[f] + (first ? pp : pn);
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) { rio[q][2 * f] = rin[f]; rio[q][2 * f + 1] = rout[f]; }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    if (cnt == 2) {
      const double p1 = pv[0][f], p2 = pv[1][f];
      mx[0][f] = tmax<double>(tmax<double>(tmax<double>(mx[0][f], p1), p2), mx[1][f]);
      mn[0][f] = tmin<double>(tmin<double>(tmin<double>(mn[0][f], p1), p2), mn[1][f]);
      mx[1][f] = tmax<double>(tmax<double>(tmax<double>(mx[1][f], p1), p2), mx[0][f]);
      mn[1][f] = tmin<double>(tmin<double>(tmin<double>(mn[1][f], p1), p2), mn[0][f]);
    } else {
      for (int k = 0; k < 4; ++k) {
        double hi = mx[0][f], lo;
        for (int j = 1; j < 4; ++j) hi = tmax<double>(hi, mx[j][f]);
        for (int j = 0; j < 4; ++j) hi = tmax<double>(hi, pv[j][f]);
        mx[k][f] = hi;
        lo = mx[0][f];
        for (int j = 1; j < 4; ++j) lo = tmin<double>(lo, mx[j][f]);
        for (int j = 0; j < 4; ++j) lo = tmin<double>(lo, pv[j][f]);
        mn[k][f] = lo;
      }
    }
  }
  double c[2 * NF];
#pragma unroll
  for (int k = 0; k < 2 * NF; ++k) {
    c[k] = rio[0][k] + rio[1][k];
    if (cnt == 4) c[k] = (c[k] + rio[2][k]) + rio[3][k];
  }
  for (int q = 0; q < cnt; ++q) {
    double *d = p.cval + (size_t)4 * NF * p.cslot[set_member(s, q)];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      d[4 * f] = mx[q][f]; d[4 * f + 1] = mn[q][f]; d[4 * f + 2] = c[2 * f]; d[4 * f + 3] = c[2 * f + 1];
    }
  }
}

// mpdatamper pass 2 (:3666-3743 with ipass = 2): limited antidiffusive increments
template <int NF>
__global__ void __launch_bounds__(NFT_BLOCK) k_per_mp_limit(Incidence g, PerSets p, MpArgs<double, NF> a) {
  const int t = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (t >= p.nset) return;
  const int4 s = p.sets[t];
  double v[4][NF];
  for (int q = 0; q < set_size(s); ++q) {
    const int i = set_member(s, q);
    const double Ui = a.U[i], Vi = a.V[i], Wi = a.W[i], vri = a.vr[i];
    double pi[NF], gxi[NF], gyi[NF], gzi[NF], cpi[NF], cni[NF], arh2[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      pi[f] = a.psi1[f][i];
      const double2 g2 = a.gxy[f][i], c2 = a.cpn[f][i];
      gxi[f] = g2.x; gyi[f] = g2.y; gzi[f] = a.gz[f][i];
      cpi[f] = c2.x; cni[f] = c2.y;
      arh2[f] = 0.;
    }
    for (int m = 0; m < g.deg[i]; ++m) {
      int j, e;
      sell_slot(g, i, m, j, e);
      const int ea = e >= 0 ? e : ~e;
      const double2 sxy = g.sxy[ea];
      const bool first = e >= 0;
      const double Uj = a.U[j], Vj = a.V[j], Wj = a.W[j];
      const double aun = edge_aun<double>(Ui, Uj, Vi, Vj, Wi, Wj, sxy.x, sxy.y, g.szz[ea]);
      const double volt = 0.5 * (vri + a.vr[j]);
      const double U12 = Ui + Uj, V12 = Vi + Vj, W12 = Wi + Wj;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double pj = a.psi1[f][j];
        const double2 gj = a.gxy[f][j], cj = a.cpn[f][j];
        const double A = anti_flux<double>(first ? pi[f] : pj, first ? pj : pi[f], gxi[f] + gj.x, gyi[f] + gj.y,
                                           gzi[f] + a.gz[f][j], aun, volt, U12, V12, W12, a.dt);
        const double cp2 = first ? cj.x : cpi[f], cn1 = first ? cni[f] : cj.y;
        const double cn2 = first ? cj.y : cni[f], cp1 = first ? cpi[f] : cj.x;
        const double lim = A > 0. ? tmin<double>(tmin<double>(1., cp2), cn1) : tmin<double>(tmin<double>(1., cn2), cp1);
        const double fl = A * lim;
        arh2[f] = first ? arh2[f] + fl : arh2[f] - fl;
      }
    }
    for (int b = g.bp_ptr[i]; b < g.bp_ptr[i + 1]; ++b) {
#pragma unroll
      for (int f = 0; f < NF; ++f) arh2[f] = arh2[f] + 0. * 1.;
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) v[q][f] = arh2[f];
  }
  per_close_store<NF>(p, s, v);
}

// ------------------------------------------------------------------------------------
// Pointwise loops of the time step (reference mpsolver)
// ------------------------------------------------------------------------------------
struct StepArgs {
  int n;
  int itraj;       // 1: the advector comes from velprd, so no extrapolation and no uuu0
  double *unk[4], *frc[4], *uuu[3], *uuu0[3], *ud[3];
  const double *alpha, *ue, *ve, *th0, *coefw, *rh0, *dthedz;
  double dt, g;
  float *uuuf[3];  // FP32 copy of the advector (mixed precision), may be null
};

// :290-301  uuu0 = uuu; unk += 0.5*frc*dt
__global__ void k_step_pre(StepArgs a) {
  int i = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (i >= a.n) return;
  if (!a.itraj)
    for (int k = 0; k < 3; ++k) a.uuu0[k][i] = a.uuu[k][i];
  for (int k = 0; k < 4; ++k) a.unk[k][i] = a.unk[k][i] + 0.5 * a.frc[k][i] * a.dt;
}

// :324-359  explicit part of the trapezoidal step; frc(1..3) <- u*, which is also unk1
__global__ void k_step_rhs(StepArgs a) {
  int i = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (i >= a.n) return;
  double al = a.alpha[i];
  double frm1 = 0. - al * (a.ud[0][i] - a.ue[i]);
  double frm2 = 0. - al * (a.ud[1][i] - a.ve[i]);
  double frm3 = (0. - al * a.ud[2][i]) + a.g * a.unk[3][i] / a.th0[i];
  double h = 0.5 * a.dt;
  a.frc[0][i] = a.unk[0][i] + h * frm1;
  a.frc[1][i] = a.unk[1][i] + h * frm2;
  a.frc[2][i] = (a.unk[2][i] + h * frm3) * a.coefw[i];
}

// :367-383  after the implicit solve (last: forcing and theta update)
__global__ void k_step_post(StepArgs a, int last) {
  int i = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (i >= a.n) return;
  double d0 = a.uuu[0][i], d1 = a.uuu[1][i], d2 = a.uuu[2][i];
  a.ud[0][i] = d0; a.ud[1][i] = d1; a.ud[2][i] = d2;
  if (last) {
    double dt = a.dt;
    a.frc[0][i] = (d0 - a.frc[0][i]) * 2. / dt;
    a.frc[1][i] = (d1 - a.frc[1][i]) * 2. / dt;
    double frc33 = a.frc[2][i] / a.coefw[i] - (a.g * a.unk[3][i] / a.th0[i]) * 0.5 * dt;
    a.frc[2][i] = (d2 - frc33) * 2. / dt;
    a.unk[0][i] = d0; a.unk[1][i] = d1; a.unk[2][i] = d2;
    double f4 = -(d2 * a.dthedz[i]);
    a.frc[3][i] = f4;
    a.unk[3][i] = a.unk[3][i] + 0.5 * dt * f4;
  }
}

// :387-439  explicit forcing for the next step and advector extrapolation
__global__ void k_step_end(StepArgs a) {
  int i = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (i >= a.n) return;
  double al = a.alpha[i];
  double u0 = a.unk[0][i], u1 = a.unk[1][i], u2 = a.unk[2][i], u3 = a.unk[3][i];
  a.frc[0][i] = a.frc[0][i] + (0. - al * (u0 - a.ue[i]));
  a.frc[1][i] = a.frc[1][i] + (0. - al * (u1 - a.ve[i]));
  a.frc[2][i] = a.frc[2][i] + (0. - al * u2);
  a.frc[3][i] = a.frc[3][i] + (-(al * u3));
  if (a.itraj) return;
  double rh = a.rh0[i];
  double rhoi = 1. / rh;
  for (int k = 0; k < 3; ++k) {
    double uu = a.uuu[k][i], u0k = a.uuu0[k][i];
    double nu = 1.5 * uu * rh - 0.5 * u0k;
    a.uuu[k][i] = nu;
    if (a.uuuf[0]) a.uuuf[k][i] = (float)nu;
    double o = u0k * rhoi;
    a.uuu0[k][i] = o;
    a.ud[k][i] = 2. * a.unk[k][i] - o;
  }
}

// mpsolver initialisation after the potential-flow projection (:229-266)
__global__ void k_init_post(StepArgs a) {
  int i = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (i >= a.n) return;
  double rh = a.rh0[i];
  for (int k = 0; k < 3; ++k) {
    double u = a.unk[k][i];
    if (a.itraj) {   // velprd follows with uuu = u (:229-231, :266)
      a.uuu[k][i] = u;
      continue;
    }
    double m = u * rh;
    a.uuu[k][i] = m;
    if (a.uuuf[0]) a.uuuf[k][i] = (float)m;
    a.ud[k][i] = u;
  }
  for (int k = 0; k < 4; ++k) a.frc[k][i] = 0.;
}

// velprd (:4464-4467): unkdash = unk + dt*frc, the start of the prediction
__global__ void k_vp_start(StepArgs a) {
  int i = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (i >= a.n) return;
  for (int k = 0; k < 3; ++k) a.ud[k][i] = a.unk[k][i] + a.dt * a.frc[k][i];
}

// velprd (:4520-4529): advector at n+1/2 and the prediction; p = 0 for the next call
__global__ void k_vp_end(StepArgs a, const double *const *ff) {
  int i = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (i >= a.n) return;
  double rh = a.rh0[i];
  for (int k = 0; k < 3; ++k) {
    double f = ff[k][i];
    double m = (a.uuu[k][i] + f) * 0.5 * rh;
    a.uuu[k][i] = m;
    if (a.uuuf[0]) a.uuuf[k][i] = (float)m;
    a.ud[k][i] = f;
  }
}

template <typename T>
__global__ void k_to_t(int n, const double *a, T *b) {
  int k = blockIdx.x * NFT_BLOCK + threadIdx.x;
  if (k < n) b[k] = (T)a[k];
}

}  // namespace nft
