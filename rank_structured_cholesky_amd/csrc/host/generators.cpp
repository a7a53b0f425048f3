// Synthetic SPD test systems (seeded, deterministic): finite-difference
// Poisson operators and trilinear-hex linear elasticity, plus uniform
// right-hand sides.  The reference ships FiniteDifference::generateLaplacian
// (src/rschol/geometry/FiniteDifference.cpp:54-88) and IsoHexElement
// (src/rschol/mesh/IsoHexElement.cpp) as guides; these generators are this
// library's own and are the inputs of the benchmark configurations.
#include <cmath>
#include <random>
#include <stdexcept>

#include "rsc_host.h"

namespace rsc {

namespace {

inline double unit53(std::mt19937_64& g) {
  return (double)(g() >> 11) * (1.0 / 9007199254740992.0);
}

inline double normal(std::mt19937_64& g) {
  double u1 = unit53(g), u2 = unit53(g);
  return std::sqrt(-2.0 * std::log(1.0 - u1)) *
         std::cos(2.0 * 3.14159265358979323846 * u2);
}

}  // namespace

Csc poisson2d(int N) {
  if (N < 1) throw std::invalid_argument("poisson2d: N < 1");
  Csc A;
  A.n = N * N;
  A.p.assign(A.n + 1, 0);
  A.i.reserve(5 * (size_t)A.n);
  A.x.reserve(5 * (size_t)A.n);
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      int c = x + N * y;
      auto put = [&](int r, double v) {
        A.i.push_back(r);
        A.x.push_back(v);
      };
      if (y > 0) put(c - N, -1.0);
      if (x > 0) put(c - 1, -1.0);
      put(c, 4.0);
      if (x + 1 < N) put(c + 1, -1.0);
      if (y + 1 < N) put(c + N, -1.0);
      A.p[c + 1] = (int)A.i.size();
    }
  return A;
}

Csc poisson3d_lognormal(int N, double sigma, uint64_t seed) {
  if (N < 1) throw std::invalid_argument("poisson3d: N < 1");
  const int64_t n = (int64_t)N * N * N;
  if (n > (1LL << 30)) throw std::invalid_argument("poisson3d: too large");
  std::mt19937_64 g(seed);
  const int64_t NF = (int64_t)(N + 1) * N * N;
  // kx[(z*N + y)*(N+1) + x], ky[(z*(N+1) + y)*N + x], kz[(z*N + y)*N + x]
  std::vector<double> kx(NF), ky(NF), kz(NF);
  for (auto& v : kx) v = std::exp(sigma * normal(g));
  for (auto& v : ky) v = std::exp(sigma * normal(g));
  for (auto& v : kz) v = std::exp(sigma * normal(g));
  auto KX = [&](int x, int y, int z) { return kx[((int64_t)z * N + y) * (N + 1) + x]; };
  auto KY = [&](int x, int y, int z) { return ky[((int64_t)z * (N + 1) + y) * N + x]; };
  auto KZ = [&](int x, int y, int z) { return kz[((int64_t)z * N + y) * N + x]; };
  Csc A;
  A.n = (int)n;
  A.p.assign(n + 1, 0);
  A.i.reserve(7 * (size_t)n);
  A.x.reserve(7 * (size_t)n);
  const int64_t NN = (int64_t)N * N;
  for (int z = 0; z < N; ++z)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < N; ++x) {
        int64_t c = x + (int64_t)N * y + NN * z;
        double kxm = KX(x, y, z), kxp = KX(x + 1, y, z);
        double kym = KY(x, y, z), kyp = KY(x, y + 1, z);
        double kzm = KZ(x, y, z), kzp = KZ(x, y, z + 1);
        auto put = [&](int64_t r, double v) {
          A.i.push_back((int)r);
          A.x.push_back(v);
        };
        if (z > 0) put(c - NN, -kzm);
        if (y > 0) put(c - N, -kym);
        if (x > 0) put(c - 1, -kxm);
        put(c, kxm + kxp + kym + kyp + kzm + kzp);
        if (x + 1 < N) put(c + 1, -kxp);
        if (y + 1 < N) put(c + N, -kyp);
        if (z + 1 < N) put(c + NN, -kzp);
        A.p[c + 1] = (int)A.i.size();
      }
  return A;
}

// Trilinear hexahedral (Q1) elasticity on an NE^3 unit-spaced element grid,
// E_e = exp(sigma * N(0,1)) per element, Poisson ratio nu, nodes on the x=0
// face clamped (removed).  Unknowns: 3 per free node, node order x fastest.
Csc elasticity3d(int NE, double sigma, double nu, uint64_t seed) {
  if (NE < 1) throw std::invalid_argument("elasticity3d: NE < 1");
  const int NN = NE + 1;
  std::mt19937_64 g(seed);
  std::vector<double> E((size_t)NE * NE * NE);
  for (auto& v : E) v = std::exp(sigma * normal(g));
  // reference element stiffness for E = 1 (unit cube, 2x2x2 Gauss)
  double K0[24][24] = {};
  {
    double lam = nu / ((1 + nu) * (1 - 2 * nu));
    double mu = 1.0 / (2 * (1 + nu));
    double D[6][6] = {};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) D[a][b] = lam;
    for (int a = 0; a < 3; ++a) D[a][a] += 2 * mu;
    for (int a = 3; a < 6; ++a) D[a][a] = mu;
    const double gp = 0.5 / std::sqrt(3.0);
    const double pts[2] = {0.5 - gp, 0.5 + gp};
    for (int qa = 0; qa < 2; ++qa)
      for (int qb = 0; qb < 2; ++qb)
        for (int qc = 0; qc < 2; ++qc) {
          double xi = pts[qa], et = pts[qb], ze = pts[qc];
          double dN[8][3];
          for (int k = 0; k < 8; ++k) {
            int ix = k & 1, iy = (k >> 1) & 1, iz = (k >> 2) & 1;
            double fx = ix ? xi : 1 - xi, fy = iy ? et : 1 - et,
                   fz = iz ? ze : 1 - ze;
            double sx = ix ? 1 : -1, sy = iy ? 1 : -1, sz = iz ? 1 : -1;
            dN[k][0] = sx * fy * fz;
            dN[k][1] = fx * sy * fz;
            dN[k][2] = fx * fy * sz;
          }
          double B[6][24] = {};
          for (int k = 0; k < 8; ++k) {
            B[0][3 * k + 0] = dN[k][0];
            B[1][3 * k + 1] = dN[k][1];
            B[2][3 * k + 2] = dN[k][2];
            B[3][3 * k + 0] = dN[k][1];
            B[3][3 * k + 1] = dN[k][0];
            B[4][3 * k + 1] = dN[k][2];
            B[4][3 * k + 2] = dN[k][1];
            B[5][3 * k + 0] = dN[k][2];
            B[5][3 * k + 2] = dN[k][0];
          }
          double w = 1.0 / 8.0;
          for (int r = 0; r < 24; ++r)
            for (int c = 0; c < 24; ++c) {
              double s = 0;
              for (int a = 0; a < 6; ++a)
                for (int b = 0; b < 6; ++b) s += B[a][r] * D[a][b] * B[b][c];
              K0[r][c] += w * s;
            }
        }
  }
  // free-node numbering: clamp nodes with x == 0
  const int64_t nnodes = (int64_t)NN * NN * NN;
  std::vector<int> dof(nnodes, -1);
  int nfree = 0;
  for (int z = 0; z < NN; ++z)
    for (int y = 0; y < NN; ++y)
      for (int x = 0; x < NN; ++x)
        if (x > 0) dof[x + (int64_t)NN * (y + (int64_t)NN * z)] = nfree++;
  const int n = 3 * nfree;
  // each free node couples to <= 27 nodes; assemble per column with a dense
  // 3x3 block per neighbour
  Csc A;
  A.n = n;
  A.p.assign(n + 1, 0);
  A.i.reserve((size_t)n * 81);
  A.x.reserve((size_t)n * 81);
  std::vector<double> blk(27 * 9);
  for (int z = 0; z < NN; ++z)
    for (int y = 0; y < NN; ++y)
      for (int x = 1; x < NN; ++x) {
        int64_t node = x + (int64_t)NN * (y + (int64_t)NN * z);
        std::fill(blk.begin(), blk.end(), 0.0);
        // elements touching this node
        for (int ez = z - 1; ez <= z; ++ez)
          for (int ey = y - 1; ey <= y; ++ey)
            for (int ex = x - 1; ex <= x; ++ex) {
              if (ex < 0 || ey < 0 || ez < 0 || ex >= NE || ey >= NE || ez >= NE)
                continue;
              double Ee = E[ex + (size_t)NE * (ey + (size_t)NE * ez)];
              int lc = (x - ex) | ((y - ey) << 1) | ((z - ez) << 2);
              for (int k = 0; k < 8; ++k) {
                int nx = ex + (k & 1), ny = ey + ((k >> 1) & 1),
                    nz = ez + ((k >> 2) & 1);
                int nb = (nx - x + 1) + 3 * ((ny - y + 1) + 3 * (nz - z + 1));
                for (int a = 0; a < 3; ++a)
                  for (int b = 0; b < 3; ++b)
                    blk[nb * 9 + a * 3 + b] += Ee * K0[3 * k + a][3 * lc + b];
              }
            }
        for (int b = 0; b < 3; ++b) {
          int col = 3 * dof[node] + b;
          for (int nz = z - 1; nz <= z + 1; ++nz)
            for (int ny = y - 1; ny <= y + 1; ++ny)
              for (int nx = x - 1; nx <= x + 1; ++nx) {
                if (nx < 1 || ny < 0 || nz < 0 || nx >= NN || ny >= NN || nz >= NN)
                  continue;
                int nb = (nx - x + 1) + 3 * ((ny - y + 1) + 3 * (nz - z + 1));
                int d = dof[nx + (int64_t)NN * (ny + (int64_t)NN * nz)];
                for (int a = 0; a < 3; ++a) {
                  A.i.push_back(3 * d + a);
                  A.x.push_back(blk[nb * 9 + a * 3 + b]);
                }
              }
          A.p[col + 1] = (int)A.i.size();
        }
      }
  return A;
}

std::vector<double> grid_positions(int nx, int ny, int nz) {
  std::vector<double> p(3 * (size_t)nx * ny * nz);
  size_t k = 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        p[k++] = x;
        p[k++] = y;
        p[k++] = z;
      }
  return p;
}

std::vector<double> uniform_rhs(int64_t count, uint64_t seed) {
  std::mt19937_64 g(seed);
  std::vector<double> b(count);
  for (auto& v : b) v = 2.0 * unit53(g) - 1.0;
  return b;
}

}  // namespace rsc
