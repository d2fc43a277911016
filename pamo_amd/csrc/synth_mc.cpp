// Seeded SDF mesh generator for benchmarks (host C++, OpenMP).
//
// Field: f(p) = min( smin-fold over spheres, min over thin slabs ), where the
// smooth-min fold runs over the spheres with signed distance < cutoff in
// ascending sphere order, starting from cutoff (polynomial smin, width k).
// The surface is f = 0 on an n^3 lattice over [0,1]^3.
//
// Polygonisation: marching cubes by loop tracing.  On every cube face the
// crossing points are joined by segments; a face with four crossings (a
// saddle) joins the diagonal whose value product is larger (ties: the
// positive diagonal), a rule both cubes sharing the face evaluate
// identically.  Segments are oriented so that the surface normal points to
// f > 0, chained into loops per cube and fan-triangulated.  Each crossing
// point lies on exactly one loop per cube, so the result is a closed,
// consistently oriented 2-manifold.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Field {
  std::vector<double> sph;   // cx cy cz r
  std::vector<double> slab;  // 13 per slab
  int ns = 0, nb = 0;
  double k = 0.02, cutoff = 0.04;
  // bins
  int G = 64;
  std::vector<int> bin_off, bin_list;   // candidate spheres per bin
  std::vector<int> sbin_off, sbin_list; // candidate slabs per bin
  std::vector<int8_t> bin_state;        // 0 active, -1 inside, +1 outside

  static double smin(double a, double b, double k) {
    double h = std::max(k - std::fabs(a - b), 0.0) / k;
    return std::min(a, b) - h * h * k * 0.25;
  }
  double box_sdf(const double* s, double x, double y, double z) const {
    double px = x - s[0], py = y - s[1], pz = z - s[2];
    const double* n = s + 3;
    const double* t = s + 6;
    double b[3] = {n[1] * t[2] - n[2] * t[1], n[2] * t[0] - n[0] * t[2], n[0] * t[1] - n[1] * t[0]};
    double q0 = std::fabs(px * n[0] + py * n[1] + pz * n[2]) - s[9];
    double q1 = std::fabs(px * t[0] + py * t[1] + pz * t[2]) - s[10];
    double q2 = std::fabs(px * b[0] + py * b[1] + pz * b[2]) - s[11];
    double m0 = std::max(q0, 0.0), m1 = std::max(q1, 0.0), m2 = std::max(q2, 0.0);
    return std::sqrt(m0 * m0 + m1 * m1 + m2 * m2) + std::min(std::max(q0, std::max(q1, q2)), 0.0);
  }
  double slab_radius(const double* s) const {   // bounding-sphere radius of a slab box
    return std::sqrt(s[9] * s[9] + s[10] * s[10] + s[11] * s[11]);
  }

  void build_bins() {
    int nbins = G * G * G;
    std::vector<std::vector<int>> per(nbins), sper(nbins);
    bin_state.assign(nbins, 1);
    double bs = 1.0 / G;
    std::vector<double> hi_min(nbins, 1e30);   // min over spheres of the max distance (containment)
    for (int i = 0; i < ns; i++) {
      const double* s = &sph[4 * i];
      double reach = s[3] + cutoff;
      int lo[3], hi[3];
      for (int a = 0; a < 3; a++) {
        lo[a] = std::max(0, (int)std::floor((s[a] - reach) / bs));
        hi[a] = std::min(G - 1, (int)std::floor((s[a] + reach) / bs));
      }
      for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++)
          for (int x = lo[0]; x <= hi[0]; x++) {
            double bl[3] = {x * bs, y * bs, z * bs};
            double dmin2 = 0, dmax2 = 0;
            for (int a = 0; a < 3; a++) {
              double c = s[a], l = bl[a], h = bl[a] + bs;
              double dn = c < l ? l - c : (c > h ? c - h : 0.0);
              double df = std::max(std::fabs(c - l), std::fabs(c - h));
              dmin2 += dn * dn;
              dmax2 += df * df;
            }
            double lo_d = std::sqrt(dmin2) - s[3], hi_d = std::sqrt(dmax2) - s[3];
            if (lo_d < cutoff) {
              int b = (z * G + y) * G + x;
              per[b].push_back(i);
              hi_min[b] = std::min(hi_min[b], hi_d);
            }
          }
    }
    for (int j = 0; j < nb; j++) {
      const double* s = &slab[13 * j];
      double reach = slab_radius(s) + 2.0 * bs;
      int lo[3], hi[3];
      for (int a = 0; a < 3; a++) {
        lo[a] = std::max(0, (int)std::floor((s[a] - reach) / bs));
        hi[a] = std::min(G - 1, (int)std::floor((s[a] + reach) / bs));
      }
      for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++)
          for (int x = lo[0]; x <= hi[0]; x++) sper[(z * G + y) * G + x].push_back(j);
    }
    bin_off.assign(nbins + 1, 0);
    sbin_off.assign(nbins + 1, 0);
    for (int b = 0; b < nbins; b++) {
      bin_off[b + 1] = bin_off[b] + (int)per[b].size();
      sbin_off[b + 1] = sbin_off[b] + (int)sper[b].size();
      if (!sper[b].empty()) bin_state[b] = 0;
      else if (per[b].empty()) bin_state[b] = 1;
      else if (hi_min[b] < -k) bin_state[b] = -1;   // f <= min d < -k over the whole bin
      else bin_state[b] = 0;
    }
    bin_list.resize(bin_off[nbins]);
    sbin_list.resize(sbin_off[nbins]);
    for (int b = 0; b < nbins; b++) {
      std::copy(per[b].begin(), per[b].end(), bin_list.begin() + bin_off[b]);
      std::copy(sper[b].begin(), sper[b].end(), sbin_list.begin() + sbin_off[b]);
    }
  }

  double eval(double x, double y, double z) const {
    int bx = std::min(G - 1, std::max(0, (int)(x * G)));
    int by = std::min(G - 1, std::max(0, (int)(y * G)));
    int bz = std::min(G - 1, std::max(0, (int)(z * G)));
    int b = (bz * G + by) * G + bx;
    double f;
    if (bin_state[b] < 0) f = -1.0;
    else if (bin_state[b] > 0 && sbin_off[b] == sbin_off[b + 1]) f = cutoff;
    else {
      f = cutoff;
      for (int q = bin_off[b]; q < bin_off[b + 1]; q++) {
        const double* s = &sph[4 * bin_list[q]];
        double dx = x - s[0], dy = y - s[1], dz = z - s[2];
        double d = std::sqrt(dx * dx + dy * dy + dz * dz) - s[3];
        if (d < cutoff) f = smin(f, d, k);
      }
      for (int q = sbin_off[b]; q < sbin_off[b + 1]; q++) f = std::min(f, box_sdf(&slab[13 * sbin_list[q]], x, y, z));
    }
    if (f == 0.0) f = 1e-300;   // keep the iso-value off the lattice values
    return f;
  }
};

// cube tables: corner c = dx + 2dy + 4dz; edges 0-3 along x, 4-7 along y, 8-11 along z
int EC[12][2];
int FACE_C[6][4], FACE_E[6][4], FACE_N[6][3];
int FMASK[12];   // faces containing each edge (bitmask)

int edge_between(int a, int b) {
  for (int e = 0; e < 12; e++)
    if ((EC[e][0] == a && EC[e][1] == b) || (EC[e][0] == b && EC[e][1] == a)) return e;
  return -1;
}

void init_tables() {
  for (int y = 0; y < 2; y++)
    for (int z = 0; z < 2; z++) { int e = y + 2 * z; EC[e][0] = 2 * y + 4 * z; EC[e][1] = 1 + 2 * y + 4 * z; }
  for (int x = 0; x < 2; x++)
    for (int z = 0; z < 2; z++) { int e = 4 + x + 2 * z; EC[e][0] = x + 4 * z; EC[e][1] = x + 2 + 4 * z; }
  for (int x = 0; x < 2; x++)
    for (int y = 0; y < 2; y++) { int e = 8 + x + 2 * y; EC[e][0] = x + 2 * y; EC[e][1] = x + 2 * y + 4; }
  int f = 0;
  for (int a = 0; a < 3; a++)
    for (int s = 0; s < 2; s++, f++) {
      int b = (a + 1) % 3, c = (a + 2) % 3;
      int cyc[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
      for (int q = 0; q < 4; q++) {
        int p[3];
        p[a] = s; p[b] = cyc[q][0]; p[c] = cyc[q][1];
        FACE_C[f][q] = p[0] + 2 * p[1] + 4 * p[2];
      }
      for (int q = 0; q < 4; q++) FACE_E[f][q] = edge_between(FACE_C[f][q], FACE_C[f][(q + 1) % 4]);
      FACE_N[f][0] = FACE_N[f][1] = FACE_N[f][2] = 0;
      FACE_N[f][a] = s ? 1 : -1;
    }
  for (int e = 0; e < 12; e++) FMASK[e] = 0;
  for (int f2 = 0; f2 < 6; f2++)
    for (int q = 0; q < 4; q++) FMASK[FACE_E[f2][q]] |= 1 << f2;
}

inline void cpos(int c, long long p[3]) { p[0] = c & 1; p[1] = (c >> 1) & 1; p[2] = (c >> 2) & 1; }
inline long long det3(const long long a[3], const long long b[3], const long long c[3]) {
  return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
}
// directed segment eA -> eB on face f given nu (any positive multiple)
inline void orient_seg(int f, int eA, int eB, const long long nu[3], int& from, int& to) {
  long long ma[3], mb[3], pa[3], pb[3], qa[3], qb[3];
  cpos(EC[eA][0], pa); cpos(EC[eA][1], qa); cpos(EC[eB][0], pb); cpos(EC[eB][1], qb);
  for (int k = 0; k < 3; k++) { ma[k] = pa[k] + qa[k]; mb[k] = pb[k] + qb[k]; }
  long long d[3] = {mb[0] - ma[0], mb[1] - ma[1], mb[2] - ma[2]};
  long long n[3] = {FACE_N[f][0], FACE_N[f][1], FACE_N[f][2]};
  if (det3(d, nu, n) > 0) { from = eA; to = eB; }
  else { from = eB; to = eA; }
}

// next[e] for the crossing edges of one cube; returns false on inconsistency
bool cube_segments(const double val[8], int next[12]) {
  for (int e = 0; e < 12; e++) next[e] = -1;
  for (int f = 0; f < 6; f++) {
    const int* C = FACE_C[f];
    bool pos[4];
    for (int q = 0; q < 4; q++) pos[q] = val[C[q]] > 0;
    int ncross = 0;
    for (int q = 0; q < 4; q++) ncross += pos[q] != pos[(q + 1) % 4];
    if (ncross == 0) continue;
    long long P[4][3];
    for (int q = 0; q < 4; q++) cpos(C[q], P[q]);
    auto add = [&](int eA, int eB, const long long nu[3]) {
      int a, b;
      orient_seg(f, eA, eB, nu, a, b);
      if (next[a] != -1) return false;
      next[a] = b;
      return true;
    };
    if (ncross == 2) {
      int es[2], m = 0;
      for (int q = 0; q < 4; q++) if (pos[q] != pos[(q + 1) % 4]) es[m++] = FACE_E[f][q];
      long long sp[3] = {0, 0, 0}, sn[3] = {0, 0, 0};
      int np = 0, nn = 0;
      for (int q = 0; q < 4; q++) {
        for (int k = 0; k < 3; k++) (pos[q] ? sp : sn)[k] += P[q][k];
        (pos[q] ? np : nn)++;
      }
      long long nu[3];
      for (int k = 0; k < 3; k++) nu[k] = sp[k] * (12 / np) - sn[k] * (12 / nn);
      if (!add(es[0], es[1], nu)) return false;
    } else {
      double p02 = val[C[0]] * val[C[2]], p13 = val[C[1]] * val[C[3]];
      bool conn02 = p02 > p13 || (p02 == p13 && pos[0]);
      int cuts[2] = {conn02 ? 1 : 0, conn02 ? 3 : 2};
      for (int ci = 0; ci < 2; ci++) {
        int q = cuts[ci];
        int eA = FACE_E[f][(q + 3) % 4], eB = FACE_E[f][q];
        long long sum[3] = {0, 0, 0};
        for (int r = 0; r < 4; r++) for (int k = 0; k < 3; k++) sum[k] += P[r][k];
        long long nu[3];
        for (int k = 0; k < 3; k++) nu[k] = pos[q] ? (4 * P[q][k] - sum[k]) : (sum[k] - 4 * P[q][k]);
        if (!add(eA, eB, nu)) return false;
      }
    }
  }
  return true;
}

struct Mesh {
  std::vector<float> V;
  std::vector<int> F;
};
Mesh g_mesh;

}  // namespace

extern "C" {

// Returns 0 on success (counts[0] = V, counts[1] = F), <0 on error.
long long synth_sdf_mesh(const double* spheres, int ns, const double* slabs, int nb, double k, int n, int threads,
                         long long* counts, void*) {
  init_tables();
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  Field fd;
  fd.sph.assign(spheres, spheres + 4 * (size_t)ns);
  if (nb) fd.slab.assign(slabs, slabs + 13 * (size_t)nb);
  fd.ns = ns; fd.nb = nb; fd.k = k; fd.cutoff = 2.0 * k;
  fd.G = std::max(8, std::min(128, n / 8));
  fd.build_bins();
  const int N = n, P1 = N + 1;
  const double h = 1.0 / N;
  std::vector<double> sl[2];
  sl[0].resize((size_t)P1 * P1);
  sl[1].resize((size_t)P1 * P1);
  // vertex ids: x-edges per slice [j][i] (i < N), y-edges per slice [j][i] (j < N), z-edges per slab
  std::vector<int> xe[2], ye[2], ze;
  for (int s = 0; s < 2; s++) { xe[s].assign((size_t)P1 * N, -1); ye[s].assign((size_t)N * P1, -1); }
  ze.assign((size_t)P1 * P1, -1);
  Mesh& M = g_mesh;
  M.V.clear(); M.F.clear();

  // the outermost lattice layer is forced outside so the surface closes
  // where the field reaches the domain boundary
  auto eval_slice = [&](int kz, std::vector<double>& out) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < P1; j++)
      for (int i = 0; i < P1; i++) {
        double f = fd.eval(i * h, j * h, kz * h);
        if (kz == 0 || kz == N || j == 0 || j == N || i == 0 || i == N) f = std::max(f, h);
        out[(size_t)j * P1 + i] = f;
      }
  };
  auto tpos = [](double a, double b) {
    double t = a / (a - b);
    return std::min(0.999, std::max(0.001, t));
  };
  // assign vertex ids in row order for one family of edges
  auto make_vertices = [&](int rows, int cols, auto&& crosses, auto&& place, std::vector<int>& ids) {
    std::vector<int> rc(rows + 1, 0);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; r++) {
      int c = 0;
      for (int q = 0; q < cols; q++) c += crosses(r, q);
      rc[r + 1] = c;
    }
    for (int r = 0; r < rows; r++) rc[r + 1] += rc[r];
    size_t base = M.V.size() / 3;
    M.V.resize(3 * (base + rc[rows]));
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; r++) {
      int id = (int)base + rc[r];
      for (int q = 0; q < cols; q++) {
        if (crosses(r, q)) {
          float p[3];
          place(r, q, p);
          M.V[3 * (size_t)id] = p[0]; M.V[3 * (size_t)id + 1] = p[1]; M.V[3 * (size_t)id + 2] = p[2];
          ids[(size_t)r * cols + q] = id++;
        } else ids[(size_t)r * cols + q] = -1;
      }
    }
  };
  auto slice_vertices = [&](int kz, int s) {
    const std::vector<double>& f = sl[s];
    make_vertices(P1, N, [&](int j, int i) { return (f[(size_t)j * P1 + i] > 0) != (f[(size_t)j * P1 + i + 1] > 0); },
                  [&](int j, int i, float* p) {
                    double t = tpos(f[(size_t)j * P1 + i], f[(size_t)j * P1 + i + 1]);
                    p[0] = (float)((i + t) * h); p[1] = (float)(j * h); p[2] = (float)(kz * h);
                  }, xe[s]);
    make_vertices(N, P1, [&](int j, int i) { return (f[(size_t)j * P1 + i] > 0) != (f[(size_t)(j + 1) * P1 + i] > 0); },
                  [&](int j, int i, float* p) {
                    double t = tpos(f[(size_t)j * P1 + i], f[(size_t)(j + 1) * P1 + i]);
                    p[0] = (float)(i * h); p[1] = (float)((j + t) * h); p[2] = (float)(kz * h);
                  }, ye[s]);
  };
  int cur = 0;
  eval_slice(0, sl[0]);
  slice_vertices(0, 0);
  bool bad = false;
  for (int kz = 0; kz < N; kz++) {
    int nx = cur ^ 1;
    eval_slice(kz + 1, sl[nx]);
    const std::vector<double>& f0 = sl[cur];
    const std::vector<double>& f1 = sl[nx];
    make_vertices(P1, P1, [&](int j, int i) { return (f0[(size_t)j * P1 + i] > 0) != (f1[(size_t)j * P1 + i] > 0); },
                  [&](int j, int i, float* p) {
                    double t = tpos(f0[(size_t)j * P1 + i], f1[(size_t)j * P1 + i]);
                    p[0] = (float)(i * h); p[1] = (float)(j * h); p[2] = (float)((kz + t) * h);
                  }, ze);
    slice_vertices(kz + 1, nx);
    std::vector<std::vector<int>> rows(N);
    std::vector<std::vector<float>> extra(N);
    int rbad = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : rbad)
    for (int j = 0; j < N; j++) {
      std::vector<int>& out = rows[j];
      for (int i = 0; i < N; i++) {
        double v[8];
        for (int c = 0; c < 8; c++) {
          int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
          v[c] = (dz ? f1 : f0)[(size_t)(j + dy) * P1 + i + dx];
        }
        bool any_pos = false, any_neg = false;
        for (int c = 0; c < 8; c++) { if (v[c] > 0) any_pos = true; else any_neg = true; }
        if (!any_pos || !any_neg) continue;
        int next[12];
        if (!cube_segments(v, next)) { rbad++; continue; }
        int vid[12];
        for (int e = 0; e < 12; e++) {
          vid[e] = -1;
          if (next[e] < 0) continue;
          if (e < 4) { int y = e & 1, z = e >> 1; vid[e] = xe[z ? nx : cur][(size_t)(j + y) * N + i]; }
          else if (e < 8) { int x = (e - 4) & 1, z = (e - 4) >> 1; vid[e] = ye[z ? nx : cur][(size_t)j * P1 + i + x]; }
          else { int x = (e - 8) & 1, y = (e - 8) >> 1; vid[e] = ze[(size_t)(j + y) * P1 + i + x]; }
        }
        bool seen[12] = {false};
        for (int e0 = 0; e0 < 12; e0++) {
          if (next[e0] < 0 || seen[e0]) continue;
          int loop[12], m = 0, e = e0;
          while (!seen[e] && m < 12) { seen[e] = true; loop[m++] = e; e = next[e]; }
          if (e != e0 || m < 3) { rbad++; continue; }
          // fan apex whose triangles never lie flat in a cube face (a flat
          // triangle on a saddle face would be emitted again, reversed, by
          // the neighbouring cube); otherwise a centre vertex
          int apex = -1;
          for (int r = 0; r < m && apex < 0; r++) {
            bool flat = false;
            for (int q = 1; q + 1 < m && !flat; q++)
              flat = (FMASK[loop[r]] & FMASK[loop[(r + q) % m]] & FMASK[loop[(r + q + 1) % m]]) != 0;
            if (!flat) apex = r;
          }
          if (apex >= 0) {
            for (int q = 1; q + 1 < m; q++) {
              out.push_back(vid[loop[apex]]);
              out.push_back(vid[loop[(apex + q) % m]]);
              out.push_back(vid[loop[(apex + q + 1) % m]]);
            }
          } else {
            double c[3] = {0, 0, 0};
            for (int q = 0; q < m; q++)
              for (int a = 0; a < 3; a++) c[a] += M.V[3 * (size_t)vid[loop[q]] + a];
            int local = (int)(extra[j].size() / 3);
            for (int a = 0; a < 3; a++) extra[j].push_back((float)(c[a] / m));
            for (int q = 0; q < m; q++) {
              out.push_back(-1 - local); out.push_back(vid[loop[q]]); out.push_back(vid[loop[(q + 1) % m]]);
            }
          }
        }
      }
    }
    if (rbad) bad = true;
    for (int j = 0; j < N; j++) {
      int base = (int)(M.V.size() / 3);
      M.V.insert(M.V.end(), extra[j].begin(), extra[j].end());
      for (int& x : rows[j]) if (x < 0) x = base + (-1 - x);
      M.F.insert(M.F.end(), rows[j].begin(), rows[j].end());
    }
    cur = nx;
  }
  counts[0] = (long long)(M.V.size() / 3);
  counts[1] = (long long)(M.F.size() / 3);
  for (int x : M.F) if (x < 0) return -2;
  return bad ? -3 : 0;
}

void synth_fetch(float* verts, int* faces) {
  memcpy(verts, g_mesh.V.data(), g_mesh.V.size() * 4);
  memcpy(faces, g_mesh.F.data(), g_mesh.F.size() * 4);
  g_mesh.V.clear(); g_mesh.V.shrink_to_fit();
  g_mesh.F.clear(); g_mesh.F.shrink_to_fit();
}

}  // extern "C"
