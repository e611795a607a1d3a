// Host-side AddressSanitizer check of the library's host half: geometry,
// upwind schedule and LDS slot plan, through the C ABI, on a jittered cube of
// 6-tet hexes with its own face matching, at the LDS sizes the sweep uses and
// at tight ones that force staging, late and overflow reads.  No device calls.
// Build and run: make -C lionbolt_amd/csrc asan-host
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <map>
#include <random>
#include <vector>

#include "../../include/lionbolt.h"

namespace {

struct Mesh {
  int64_t NE = 0, NK = 0;
  std::vector<double> rg;
  std::vector<int64_t> node, nbr, mat;
};

// n^3 hexes of side 1, each split into 6 tets around the main diagonal; interior
// nodes jittered by up to 0.2.  Neighbours by matching sorted face triples.
Mesh cube(int n, unsigned seed) {
  Mesh m;
  const int p = n + 1;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(-0.2, 0.2);
  for (int z = 0; z < p; ++z)
    for (int y = 0; y < p; ++y)
      for (int x = 0; x < p; ++x) {
        const bool in = x > 0 && y > 0 && z > 0 && x < n && y < n && z < n;
        m.rg.push_back(x + (in ? u(rng) : 0.0));
        m.rg.push_back(y + (in ? u(rng) : 0.0));
        m.rg.push_back(z + (in ? u(rng) : 0.0));
      }
  m.NK = (int64_t)p * p * p;
  auto id = [&](int x, int y, int z) { return (int64_t)(z * p + y) * p + x; };
  static const int path[6][2] = {{1, 3}, {1, 5}, {2, 3}, {2, 6}, {4, 5}, {4, 6}};
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        int64_t c[8];
        for (int b = 0; b < 8; ++b) c[b] = id(x + (b & 1), y + ((b >> 1) & 1), z + ((b >> 2) & 1));
        for (auto& q : path) {
          int64_t t[4] = {c[0], c[q[0]], c[q[1]], c[7]};
          double J[3][3];
          for (int r = 0; r < 3; ++r)
            for (int d = 0; d < 3; ++d) J[r][d] = m.rg[t[r + 1] * 3 + d] - m.rg[t[0] * 3 + d];
          const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                             J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
          if (det < 0) std::swap(t[1], t[2]);   // positive orientation
          m.node.insert(m.node.end(), {t[0], t[1], t[2], t[3]});
          m.mat.push_back(z * 2 < n ? 0 : 1);
        }
      }
  m.NE = (int64_t)m.mat.size();
  m.nbr.assign(4 * m.NE, -1);
  static const int faces[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  std::map<std::array<int64_t, 3>, int64_t> open;
  for (int64_t e = 0; e < m.NE; ++e)
    for (int f = 0; f < 4; ++f) {
      std::array<int64_t, 3> t{m.node[e * 4 + faces[f][0]], m.node[e * 4 + faces[f][1]], m.node[e * 4 + faces[f][2]]};
      std::sort(t.begin(), t.end());
      auto it = open.find(t);
      if (it == open.end()) {
        open.emplace(t, e * 4 + f);
      } else {
        m.nbr[e * 4 + f] = it->second / 4;
        m.nbr[it->second] = e;
        open.erase(it);
      }
    }
  return m;
}

}  // namespace

int main() {
  int bad = 0;
  for (int n : {4, 9}) {
    Mesh m = cube(n, 11u + n);
    // directions on a spiral over the sphere, equal weights
    const int NI = 48;
    std::vector<double> k(3 * NI), w(NI, 4.0 * M_PI / NI);
    for (int i = 0; i < NI; ++i) {
      const double mu = 1.0 - (2.0 * i + 1.0) / NI, s = std::sqrt(1.0 - mu * mu), a = 2.399963229728653 * i;
      k[3 * i] = s * std::cos(a);
      k[3 * i + 1] = s * std::sin(a);
      k[3 * i + 2] = mu;
    }
    lb_mesh_desc md{m.NE, m.NK, m.rg.data(), m.node.data(), m.nbr.data(), m.mat.data(), 0};
    lb_quad_desc qd{NI, k.data(), w.data(), 1};
    std::vector<int32_t> sl(m.NE * NI), nlev(NI);
    std::vector<double> nrm(12 * m.NE);
    int rc = lb_schedule_host(&md, &qd, sl.data(), nlev.data(), nrm.data());
    if (rc) {
      std::printf("n=%d schedule rc=%d %s\n", n, rc, lb_last_error());
      return 1;
    }
    const int maxlev = *std::max_element(nlev.begin(), nlev.end());
    const int cases[][2] = {{1, 160 * 1024}, {1, 96 * 32}, {1, 128 * 32}, {2, 100 * 64},
                            {3, 160 * 1024}, {4, 48 * 128}, {8, 160 * 1024}, {24, 160 * 1024}, {1, 32}};
    for (auto& c : cases) {
      int64_t ns = 0;
      rc = lb_trace_plan_host(&md, &qd, c[0], c[1], nullptr, nullptr, nullptr, nullptr, &ns, nullptr, nullptr);
      if (rc) {
        std::printf("n=%d G=%d lds=%d size query rc=%d %s\n", n, c[0], c[1], rc, lb_last_error());
        ++bad;
        continue;
      }
      std::vector<int32_t> us(NI * m.NE * 4), ws(NI * m.NE), so(NI * (maxlev + 2)), st(2 * std::max<int64_t>(ns, 1));
      int32_t sizes[2] = {0, 0};
      int64_t stats[4] = {0, 0, 0, 0};
      rc = lb_trace_plan_host(&md, &qd, c[0], c[1], us.data(), ws.data(), so.data(), st.data(), &ns, sizes, stats);
      std::printf("n=%d G=%d lds=%d rc=%d ring=%d stage=%d staged=%lld reads: ring %lld staged %lld late %lld "
                  "overflow %lld\n", n, c[0], c[1], rc, sizes[0], sizes[1], (long long)ns, (long long)stats[0],
                  (long long)stats[1], (long long)stats[2], (long long)stats[3]);
      bad += rc != 0;
    }
    // malformed input: a neighbour index past the last element
    m.nbr[5] = m.NE;
    rc = lb_schedule_host(&md, &qd, sl.data(), nlev.data(), nrm.data());
    std::printf("n=%d bad neighbour rc=%d (%s)\n", n, rc, lb_last_error());
    bad += rc != LB_E_ARG;
  }
  std::printf(bad ? "FAILED\n" : "ok\n");
  return bad != 0;
}
