// Terrain-following extruded tetrahedral mesh with median-dual edge metrics.
//
// The reference reads its mesh from `mesh3d.d` (lippsLFroude_implgwopks3D_unstr3.f:2343-2501)
// and ships no generator, so this file builds one from a 2-D triangulation of
// jittered points (the triangulation itself is an input):
//   * columns c = 0..nh-1, levels k = 0..L-1, node id n = c*L + k (level fastest),
//     z = zs(c) + zeta_k * (H - zs(c)) / H (Gal-Chen terrain following);
//   * every prism (triangle a<b<c by split rank, layer k; the rank is the column id unless
//     given) is split into 3 tets with the diagonal of each quad face leaving its
//     lowest-ranked node, which makes the split conforming:
//     {a0,b0,c0,c1}, {a0,b0,b1,c1}, {a0,a1,b1,c1};
//   * median-dual metrics: per tet edge (p,q) with the other nodes (r,s) in positive
//     orientation the dual facet area vector is S_pq = ((r+s)-(p+q)) x (s-r) / 24, and
//     every node receives a quarter of the tet volume;
//   * the boundary closure of node i on a boundary part is sbb = sum(A_f)/3 over the
//     boundary triangles f of that part touching i (outward area vectors), and anp is its
//     unit vector. With this closure, sum(+-S) + sum(sbb) = 0 holds at every node, which
//     is what the edge Green-Gauss gradient (reference `derivative`, :3943-4022) needs to
//     keep a constant field constant.
// Edges lying on the boundary surface go to the semi-edge list (`kbface`/`sb`), all other
// edges to `iedge`/`sn`, as the reference's input expects (:1212, :1235). Each list keeps
// the raw order: per column c the vertical edges, then for every horizontal pair (c,c2>c)
// its horizontal edges and its diagonal edges, each family level-contiguous.
// Flags: bottom/top 11 (slip), y=ymin 5 and y=ymax 6 (slip), x=xmin 7 and x=xmax 8
// (u prescribed). One boundary-point entry per (node, part), ordered by part
// (bottom, top, south, north, west, east) and then node.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>
#include <string>
#include <random>

namespace {

struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

enum Part { BOTTOM = 0, TOP = 1, SOUTH = 2, NORTH = 3, WEST = 4, EAST = 5, NPART = 6 };
const int kPartFlag[NPART] = {11, 11, 5, 6, 7, 8};

struct Mesh {
  int nh = 0, L = 0;
  int64_t npoin = 0;
  std::vector<double> coord;  // [3][npoin]
  std::vector<double> vol;
  std::vector<int32_t> iedge;   // [2][nedge], 1-based
  std::vector<double> sn;       // [3][nedge]
  std::vector<int32_t> kbface;  // [2][nbface]
  std::vector<double> sb;       // [3][nbface]
  std::vector<int32_t> ibpoin;  // [2][nboun]
  std::vector<double> sbb;      // [3][nboun]
  std::vector<double> anp;      // [3][nboun]
  int64_t nedge = 0, nbface = 0, nboun = 0;
  std::string err;
};

}  // namespace

extern "C" {

// Opaque handle for a generated mesh.
struct nft_genmesh;

// Builds the mesh. xy: [2][nh] column positions (x block then y block); tri: [3][nt]
// 0-based column ids (any orientation); zs: terrain height per column; zeta: L level
// heights in [0,H] (zeta[0]=0, zeta[L-1]=H); box = {xmin,xmax,ymin,ymax}; rank: optional
// [nh] distinct split ranks (NULL: the column id). Prisms are split from their
// lowest-ranked column, so a renumbering of the columns that passes the old ids as ranks
// yields the same tetrahedra.
// Returns NULL on failure (message in *errbuf when given).
nft_genmesh* nft_meshgen_build(int nh, const double* xy, int nt, const int32_t* tri,
                               const double* zs, int L, const double* zeta, double H,
                               const double* box, const int32_t* rank, char* errbuf, int errlen) {
  auto fail = [&](const char* m) -> nft_genmesh* {
    if (errbuf && errlen > 0) { std::snprintf(errbuf, errlen, "%s", m); }
    return nullptr;
  };
  if (nh < 3 || nt < 1 || L < 2) return fail("meshgen: need nh>=3, nt>=1, L>=2");
  if ((int64_t)nh * L > 2000000000LL) return fail("meshgen: too many nodes for int32 ids");
  Mesh* m = new Mesh();
  m->nh = nh; m->L = L; m->npoin = (int64_t)nh * L;
  const int64_t N = m->npoin;
  const double xmin = box[0], xmax = box[1], ymin = box[2], ymax = box[3];
  auto rk = [&](int c) -> int64_t { return rank ? (int64_t)rank[c] : (int64_t)c; };

  // ---- node coordinates
  m->coord.assign(3 * N, 0.0);
  m->vol.assign(N, 0.0);
  std::vector<V3> P(N);
  for (int c = 0; c < nh; ++c) {
    double x = xy[c], y = xy[nh + c], s = zs[c];
    for (int k = 0; k < L; ++k) {
      double z = s + zeta[k] * (H - s) / H;
      int64_t n = (int64_t)c * L + k;
      P[n] = {x, y, z};
      m->coord[n] = x; m->coord[N + n] = y; m->coord[2 * N + n] = z;
    }
  }

  // ---- unique horizontal pairs (c1<c2), per-column sorted adjacency
  std::vector<std::vector<std::pair<int, int>>> adj(nh);  // (c2, count of triangles)
  auto add_pair = [&](int a, int b) {
    if (a > b) std::swap(a, b);
    for (auto& pr : adj[a]) if (pr.first == b) { pr.second++; return; }
    adj[a].push_back({b, 1});
  };
  for (int t = 0; t < nt; ++t) {
    int a = tri[t], b = tri[nt + t], c = tri[2 * nt + t];
    if (a < 0 || b < 0 || c < 0 || a >= nh || b >= nh || c >= nh || a == b || b == c || a == c) {
      delete m; return fail("meshgen: bad triangle");
    }
    add_pair(a, b); add_pair(b, c); add_pair(a, c);
  }
  std::vector<int64_t> hstart(nh + 1, 0);
  for (int c = 0; c < nh; ++c) {
    std::sort(adj[c].begin(), adj[c].end());
    hstart[c + 1] = hstart[c] + (int64_t)adj[c].size();
  }
  const int64_t nhe = hstart[nh];
  std::vector<int> hc1(nhe), hc2(nhe);
  std::vector<char> hull(nhe, 0);
  std::vector<char> colbnd(nh, 0);
  for (int c = 0; c < nh; ++c)
    for (size_t j = 0; j < adj[c].size(); ++j) {
      int64_t h = hstart[c] + j;
      hc1[h] = c; hc2[h] = adj[c][j].first;
      if (adj[c][j].second == 1) { hull[h] = 1; colbnd[c] = 1; colbnd[adj[c][j].first] = 1; }
      else if (adj[c][j].second != 2) { delete m; return fail("meshgen: non-manifold triangulation"); }
    }
  auto hid = [&](int a, int b) -> int64_t {
    if (a > b) std::swap(a, b);
    const auto& v = adj[a];
    auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(b, -1));
    return hstart[a] + (it - v.begin());
  };

  // ---- raw edge numbering: per column, V(c) then for each h with c1=c: H(h), D(h)
  std::vector<int64_t> vbase(nh), hbase(nhe), dbase(nhe);
  int64_t nraw = 0;
  for (int c = 0; c < nh; ++c) {
    vbase[c] = nraw; nraw += L - 1;
    for (int64_t h = hstart[c]; h < hstart[c + 1]; ++h) {
      hbase[h] = nraw; nraw += L;
      dbase[h] = nraw; nraw += L - 1;
    }
  }
  if (nraw > 2000000000LL) { delete m; return fail("meshgen: too many edges for int32 ids"); }
  // raw edge -> endpoints and boundary-surface flag
  std::vector<int32_t> r1(nraw), r2(nraw);
  std::vector<char> rsurf(nraw, 0);
  for (int c = 0; c < nh; ++c) {
    for (int k = 0; k < L - 1; ++k) {
      int64_t e = vbase[c] + k;
      r1[e] = (int32_t)((int64_t)c * L + k); r2[e] = r1[e] + 1; rsurf[e] = colbnd[c];
    }
    for (int64_t h = hstart[c]; h < hstart[c + 1]; ++h) {
      int c2 = hc2[h];
      for (int k = 0; k < L; ++k) {
        int64_t e = hbase[h] + k;
        r1[e] = (int32_t)((int64_t)c * L + k); r2[e] = (int32_t)((int64_t)c2 * L + k);
        rsurf[e] = (k == 0 || k == L - 1 || hull[h]) ? 1 : 0;
      }
      // the quad's diagonal rises from the lower-ranked column; endpoints stored id-ascending
      const bool up = rk(c) < rk(c2);
      for (int k = 0; k < L - 1; ++k) {
        int64_t e = dbase[h] + k;
        r1[e] = (int32_t)((int64_t)c * L + k + (up ? 0 : 1)); r2[e] = (int32_t)((int64_t)c2 * L + k + (up ? 1 : 0));
        rsurf[e] = hull[h];
      }
    }
  }
  std::vector<double> rS(3 * nraw, 0.0);

  // ---- tets: accumulate dual facets and volumes
  auto rawid = [&](int64_t p, int64_t q, const int64_t hh[3], const int cc[3]) -> int64_t {
    // p<q node ids in the same prism
    int cp = (int)(p / L), kp = (int)(p % L), cq = (int)(q / L), kq = (int)(q % L);
    if (cp == cq) return vbase[cp] + std::min(kp, kq);
    int64_t h = -1;
    for (int j = 0; j < 3; ++j) {
      int a = cc[j], b = cc[(j + 1) % 3];
      if ((a == cp && b == cq) || (a == cq && b == cp)) { h = hh[j]; break; }
    }
    if (kp == kq) return hbase[h] + kp;
    return dbase[h] + std::min(kp, kq);
  };
  double minvol = 1e300;
  for (int t = 0; t < nt; ++t) {
    int cc[3] = {tri[t], tri[nt + t], tri[2 * nt + t]};
    std::sort(cc, cc + 3, [&](int x, int y) { return rk(x) < rk(y); });
    const int a = cc[0], b = cc[1], c = cc[2];
    int64_t hh[3] = {hid(a, b), hid(b, c), hid(c, a)};
    for (int k = 0; k < L - 1; ++k) {
      int64_t a0 = (int64_t)a * L + k, b0 = (int64_t)b * L + k, c0 = (int64_t)c * L + k;
      int64_t a1 = a0 + 1, b1 = b0 + 1, c1 = c0 + 1;
      int64_t tets[3][4] = {{a0, b0, c0, c1}, {a0, b0, b1, c1}, {a0, a1, b1, c1}};
      for (auto& T : tets) {
        V3 X[4] = {P[T[0]], P[T[1]], P[T[2]], P[T[3]]};
        double v6 = dot(X[1] - X[0], cross(X[2] - X[0], X[3] - X[0]));
        double v = std::fabs(v6) / 6.0;
        minvol = std::min(minvol, v);
        for (int j = 0; j < 4; ++j) m->vol[T[j]] += 0.25 * v;
        static const int ep[6][4] = {{0, 1, 2, 3}, {0, 2, 1, 3}, {0, 3, 1, 2},
                                     {1, 2, 0, 3}, {1, 3, 0, 2}, {2, 3, 0, 1}};
        for (auto& q : ep) {
          int ip = q[0], iq = q[1], ir = q[2], is = q[3];
          V3 p = X[ip], qq = X[iq], r = X[ir], s = X[is];
          if (dot(qq - p, cross(r - p, s - p)) < 0) { std::swap(ir, is); std::swap(r, s); }
          V3 S = (1.0 / 24.0) * cross((r + s) - (p + qq), s - r);
          int64_t np = T[ip], nq = T[iq];
          double sg = 1.0;
          if (np > nq) { std::swap(np, nq); sg = -1.0; }
          int64_t e = rawid(np, nq, hh, cc);
          rS[e] += sg * S.x; rS[nraw + e] += sg * S.y; rS[2 * nraw + e] += sg * S.z;
        }
      }
    }
  }
  if (!(minvol > 0)) { delete m; return fail("meshgen: degenerate tetrahedron"); }

  // ---- split into interior edges and boundary semi-edges (raw order kept)
  for (int64_t e = 0; e < nraw; ++e) (rsurf[e] ? m->nbface : m->nedge)++;
  const int64_t E = m->nedge, B = m->nbface;
  m->iedge.resize(2 * E); m->sn.resize(3 * E);
  m->kbface.resize(2 * B); m->sb.resize(3 * B);
  int64_t ie = 0, ib = 0;
  for (int64_t e = 0; e < nraw; ++e) {
    if (!rsurf[e]) {
      m->iedge[ie] = r1[e] + 1; m->iedge[E + ie] = r2[e] + 1;
      for (int d = 0; d < 3; ++d) m->sn[d * E + ie] = rS[d * nraw + e];
      ++ie;
    } else {
      m->kbface[ib] = r1[e] + 1; m->kbface[B + ib] = r2[e] + 1;
      for (int d = 0; d < 3; ++d) m->sb[d * B + ib] = rS[d * nraw + e];
      ++ib;
    }
  }

  // ---- boundary faces -> per (part,node) closure vectors
  std::vector<std::vector<V3>> acc(NPART, std::vector<V3>());
  std::vector<std::vector<char>> has(NPART);
  for (int p = 0; p < NPART; ++p) { acc[p].assign(N, V3{0, 0, 0}); has[p].assign(N, 0); }
  auto addface = [&](int part, int64_t i, int64_t j, int64_t k, V3 outward) {
    V3 A = 0.5 * cross(P[j] - P[i], P[k] - P[i]);
    if (dot(A, outward) < 0) A = -1.0 * A;
    V3 A3 = (1.0 / 3.0) * A;
    for (int64_t n : {i, j, k}) { acc[part][n] = acc[part][n] + A3; has[part][n] = 1; }
  };
  for (int t = 0; t < nt; ++t) {
    int a = tri[t], b = tri[nt + t], c = tri[2 * nt + t];
    addface(BOTTOM, (int64_t)a * L, (int64_t)b * L, (int64_t)c * L, V3{0, 0, -1});
    addface(TOP, (int64_t)a * L + L - 1, (int64_t)b * L + L - 1, (int64_t)c * L + L - 1, V3{0, 0, 1});
  }
  for (int64_t h = 0; h < nhe; ++h) {
    if (!hull[h]) continue;
    int c1 = hc1[h], c2 = hc2[h];
    double x1 = xy[c1], x2 = xy[c2], y1 = xy[nh + c1], y2 = xy[nh + c2];
    int part; V3 out;
    if (x1 == xmin && x2 == xmin) { part = WEST; out = {-1, 0, 0}; }
    else if (x1 == xmax && x2 == xmax) { part = EAST; out = {1, 0, 0}; }
    else if (y1 == ymin && y2 == ymin) { part = SOUTH; out = {0, -1, 0}; }
    else if (y1 == ymax && y2 == ymax) { part = NORTH; out = {0, 1, 0}; }
    else { delete m; return fail("meshgen: hull edge not on the box"); }
    // the side quad is split along the prism's diagonal, which rises from the lower-ranked column
    const int lo = rk(c1) < rk(c2) ? c1 : c2, hi = lo == c1 ? c2 : c1;
    for (int k = 0; k < L - 1; ++k) {
      int64_t p0 = (int64_t)lo * L + k, q0 = (int64_t)hi * L + k;
      addface(part, p0, q0, q0 + 1, out);
      addface(part, p0, q0 + 1, p0 + 1, out);
    }
  }
  for (int p = 0; p < NPART; ++p)
    for (int64_t n = 0; n < N; ++n) if (has[p][n]) m->nboun++;
  const int64_t Bp = m->nboun;
  m->ibpoin.resize(2 * Bp); m->sbb.resize(3 * Bp); m->anp.resize(3 * Bp);
  int64_t q = 0;
  for (int p = 0; p < NPART; ++p)
    for (int64_t n = 0; n < N; ++n) {
      if (!has[p][n]) continue;
      V3 s = acc[p][n];
      double len = std::sqrt(dot(s, s));
      m->ibpoin[q] = (int32_t)(n + 1); m->ibpoin[Bp + q] = kPartFlag[p];
      m->sbb[q] = s.x; m->sbb[Bp + q] = s.y; m->sbb[2 * Bp + q] = s.z;
      m->anp[q] = s.x / len; m->anp[Bp + q] = s.y / len; m->anp[2 * Bp + q] = s.z / len;
      ++q;
    }
  return reinterpret_cast<nft_genmesh*>(m);
}

// sizes[0..6] = npoin, nedge, nbface, nboun, nh, L, (unused 0)
void nft_meshgen_sizes(const nft_genmesh* h, int64_t* sizes) {
  const Mesh* m = reinterpret_cast<const Mesh*>(h);
  sizes[0] = m->npoin; sizes[1] = m->nedge; sizes[2] = m->nbface; sizes[3] = m->nboun;
  sizes[4] = m->nh; sizes[5] = m->L; sizes[6] = 0;
}

// Copies the arrays out in the reference's column-major SoA layouts (1-based ids).
void nft_meshgen_export(const nft_genmesh* h, double* coord, double* vol, int32_t* iedge,
                        double* sn, int32_t* kbface, double* sb, int32_t* ibpoin, double* sbb,
                        double* anp) {
  const Mesh* m = reinterpret_cast<const Mesh*>(h);
  auto cp = [](auto* dst, const auto& v) { if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0])); };
  cp(coord, m->coord); cp(vol, m->vol); cp(iedge, m->iedge); cp(sn, m->sn);
  cp(kbface, m->kbface); cp(sb, m->sb); cp(ibpoin, m->ibpoin); cp(sbb, m->sbb); cp(anp, m->anp);
}

// Writes the mesh in the reference's list-directed text format (Appendix-A layout of
// `input`, :2343-2501): header, points, edges, edge metrics, semi-edges, semi-edge
// metrics, boundary points, sbb, anp. Reals are printed with 17 significant digits so a
// list-directed read returns the same doubles.
int nft_meshgen_write_text(const nft_genmesh* h, const char* path) {
  const Mesh* m = reinterpret_cast<const Mesh*>(h);
  FILE* f = std::fopen(path, "w");
  if (!f) return -1;
  const int64_t N = m->npoin, E = m->nedge, B = m->nbface, Bp = m->nboun;
  std::fprintf(f, "%lld %lld %lld %lld %lld\n", (long long)N, (long long)E, (long long)B,
               (long long)Bp, (long long)Bp);
  for (int64_t i = 0; i < N; ++i)
    std::fprintf(f, "%lld %.17g %.17g %.17g %.17g\n", (long long)(i + 1), m->coord[i],
                 m->coord[N + i], m->coord[2 * N + i], m->vol[i]);
  for (int64_t e = 0; e < E; ++e)
    std::fprintf(f, "%lld %d %d\n", (long long)(e + 1), m->iedge[e], m->iedge[E + e]);
  for (int64_t e = 0; e < E; ++e)
    std::fprintf(f, "%lld %.17g %.17g %.17g\n", (long long)(e + 1), m->sn[e], m->sn[E + e], m->sn[2 * E + e]);
  for (int64_t e = 0; e < B; ++e)
    std::fprintf(f, "%lld %d %d\n", (long long)(e + 1), m->kbface[e], m->kbface[B + e]);
  for (int64_t e = 0; e < B; ++e)
    std::fprintf(f, "%lld %.17g %.17g %.17g\n", (long long)(e + 1), m->sb[e], m->sb[B + e], m->sb[2 * B + e]);
  for (int64_t e = 0; e < Bp; ++e)
    std::fprintf(f, "%lld %d %d\n", (long long)(e + 1), m->ibpoin[e], m->ibpoin[Bp + e]);
  for (int64_t e = 0; e < Bp; ++e)
    std::fprintf(f, "%lld %.17g %.17g %.17g\n", (long long)(e + 1), m->sbb[e], m->sbb[Bp + e], m->sbb[2 * Bp + e]);
  for (int64_t e = 0; e < Bp; ++e)
    std::fprintf(f, "%lld %.17g %.17g %.17g\n", (long long)(e + 1), m->anp[e], m->anp[Bp + e], m->anp[2 * Bp + e]);
  int rc = std::ferror(f) ? -2 : 0;
  std::fclose(f);
  return rc;
}

void nft_meshgen_free(nft_genmesh* h) { delete reinterpret_cast<Mesh*>(h); }

// Jittered nx x ny point set on [0,lx] x [0,ly] (column id c = iy*nx + ix), written as
// [2][nx*ny]. Interior points move by U(-jitter,jitter) times the grid spacing on each
// axis; points on a side move only along that side; corners stay. The stream is
// std::mt19937_64(seed), two draws per point (x then y), u = (r >> 11) * 2^-53 mapped to
// [-1,1), so the set is reproducible from (nx, ny, lx, ly, jitter, seed).
void nft_meshgen_points(int nx, int ny, double lx, double ly, double jitter, uint64_t seed,
                        double* xy) {
  std::mt19937_64 rng(seed);
  const double dx = lx / (nx - 1), dy = ly / (ny - 1);
  const int64_t nh = (int64_t)nx * ny;
  for (int iy = 0; iy < ny; ++iy)
    for (int ix = 0; ix < nx; ++ix) {
      double ux = (double)(rng() >> 11) * 0x1.0p-53 * 2.0 - 1.0;
      double uy = (double)(rng() >> 11) * 0x1.0p-53 * 2.0 - 1.0;
      double x = ix * dx, y = iy * dy;
      if (ix == 0) x = 0.0;
      if (ix == nx - 1) x = lx;
      if (iy == 0) y = 0.0;
      if (iy == ny - 1) y = ly;
      bool bx = (ix == 0 || ix == nx - 1), by = (iy == 0 || iy == ny - 1);
      if (!bx) x += jitter * dx * ux;
      if (!by) y += jitter * dy * uy;
      int64_t c = (int64_t)iy * nx + ix;
      xy[c] = x; xy[nh + c] = y;
    }
}

}  // extern "C"
