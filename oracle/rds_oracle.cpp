// CPU oracle for the PyRDSim step loop -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, in plain C++, the algorithm of the reference's
// time-step loop so that the GPU engine can be checked against it.  Only the
// test suite, the smoke check and the benchmark's CPU-baseline leg load it;
// the engine never calls it.
//
// What each part follows (all citations into pyrdsim/pyrdsim.pyx):
//   random stream        :71-108   (std::mt19937; a fresh
//                                   normal_distribution / uniform_real_
//                                   distribution object per draw)
//   force laws           :1037-1154 (f = dU/dr, force on i1 is +f*xhat_12)
//   cell grid            :1415-1445
//   neighbour list       :1448-1587 (Verlet list on the rs/ratio grid,
//                                   type gating, historic accumulation via
//                                   the shallow dict copy at :1513)
//   distances            :1590-1652 (minimum image with libc round)
//   drift                :1941-2067 (interactions in file order, then bonds,
//                                   then angles)
//   angles               :1322-1407 (straight-angle series; the flag of the
//                                   last angle defined read by every angle;
//                                   legs join the bonded list)
//   reactions            :2070-2222 (rates first, then execution; p = dt*k)
//   integrator           :2225-2250 (two separate FP64 additions)
//   boundaries           :2253-2292 (Python-style % ; single reflection)
//   step order           :2296-2322
//   initial drift        :2398-2401
//
// Besides the reference-faithful neighbour list there is an "exact" mode:
// every pair within the largest cutoff, rebuilt every step, no type gating.
// The GPU engine implements the exact semantics.
//
// Parity pins for every function here live in tests/test_oracle_*.py.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <set>
#include <string>
#include <vector>
#include <algorithm>

namespace {

enum { K_LJ = 1, K_TANH = 2, K_YUKAWA = 3, K_HARMONIC = 4, K_PWLINEAR = 5, K_WALL = 6 };
enum { RX_LINEAR = 1, RX_CATALYZED = 2, RX_INHIBITED = 3 };
enum { RNG_MT = 0, RNG_INJECTED = 1, RNG_PHILOX = 2 };
enum { NL_FAITHFUL = 0, NL_EXACT = 1 };

thread_local std::string g_err;

}  // namespace

extern "C" {

// Interaction in file order.  For walls kind = 6, p0 = particle_type,
// p1 = direction (pyx:1144-1145).
typedef struct {
  int32_t kind, p0, p1, n;
  double u0, r0, rc, a, dr;
} orc_interaction;

typedef struct {
  int32_t kind, reactant, product, catalyst, inhibitor, inhibitor2;
  double rate, r0, rate_decrease;
} orc_reaction;

typedef struct {
  int32_t i1, i2;
  double r0, u0;
} orc_bond;

typedef struct {
  int32_t n, n_dim, n_species, n_inter, n_rx, n_bonds;
  double L[3];
  double periodic[3];
  double dt;
  double include_random_force;
  const double* force_prefac;  // [n_species]
  const double* noise_prefac;  // [n_species]
  const orc_interaction* inter;
  const orc_reaction* rx;
  const orc_bond* bonds;
  const int32_t* molecule;  // [n] molecule number (IC column 0)
  const int32_t* partners;  // [n_species*n_species] interaction_partners
  double rs;
  int32_t ratio, nl_freq, nl_mode, rng_mode;
  int32_t reaction_prob;  // 0: p = dt*k (reference), 1: p = 1-exp(-dt*k)
  int32_t pad;
  uint64_t seed;
} orc_spec;

const char* orc_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11), restated from the published round function.
// ---------------------------------------------------------------------------
void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

}  // extern "C"

namespace {

inline double u53(uint32_t lo, uint32_t hi) {
  uint64_t w = ((uint64_t)hi << 32) | lo;
  return (double)(w >> 11) * 0x1.0p-53;
}

// The engine's noise definition: counter = (particle index within the system,
// step lo, step hi, stream), key = (seed lo, seed hi); Box-Muller in FP32 on
// an odd 24-bit uniform in (0,1) and a 24-bit angle in [0,1).  Stream 0 gives
// the x and y normals; stream 2 gives the z normal (its first output) in three
// dimensions; stream 1 is the reaction uniforms.
inline void philox_normals(uint64_t seed, int64_t step, int32_t i, double* n2,
                           uint32_t stream = 0u) {
  uint32_t ctr[4] = {(uint32_t)i, (uint32_t)step, (uint32_t)((uint64_t)step >> 32), stream};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  orc_philox4x32_10(ctr, key, w);
  float ua = (float)((w[0] >> 9) * 2u + 1u) * 0x1.0p-24f;
  float ub = (float)(w[1] >> 8) * 0x1.0p-24f;
  float rr = std::sqrt(-2.0f * std::log(ua));
  double th = M_PI * (double)(2.0f * ub);
  n2[0] = (double)(rr * (float)std::cos(th));
  n2[1] = (double)(rr * (float)std::sin(th));
}

inline void philox_uniforms(uint64_t seed, int64_t step, int32_t i, double* u) {
  uint32_t ctr[4] = {(uint32_t)i, (uint32_t)step, (uint32_t)((uint64_t)step >> 32), 1u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  orc_philox4x32_10(ctr, key, w);
  u[0] = u53(w[0], w[1]);
  u[1] = u53(w[2], w[3]);
}

// Cython's % on doubles (Python semantics, result takes the divisor's sign).
inline double py_mod(double a, double b) {
  double r = std::fmod(a, b);
  if (r != 0.0 && ((r < 0.0) != (b < 0.0))) r += b;
  return r;
}

struct Pair {
  int32_t j;
};

struct Oracle {
  orc_spec s;
  std::vector<orc_interaction> inter;
  std::vector<orc_reaction> rx;
  std::vector<orc_bond> bonds;
  std::vector<int32_t> molecule;
  std::vector<int32_t> partners;
  std::vector<double> fpre, npre;
  int n, d;
  std::vector<double> x, a;
  std::vector<int32_t> types;
  // faithful neighbour list: bonded_set[i] accumulates candidates forever
  // (pyx:1513 shallow copy), nl[i] is the distance-filtered list.
  std::vector<std::set<int32_t>> bonded_set;
  std::vector<std::set<int32_t>> bonded_static;  // bonds only (exact mode)
  std::vector<std::vector<int32_t>> nl;
  // distances for listed pairs, keyed like d_xx / xx_hat (sparse)
  std::vector<std::vector<double>> nl_d, nl_hat;  // per i, parallel to nl[i]
  // last computed distance/unit vector for each bond (stale when unlisted)
  std::vector<double> bond_d, bond_hat;
  // angles (pyx:1322-1407): participants, theta0 [rad], k0 = u0 [per rad^2]
  struct Angle {
    int32_t i1, i2, i3;
    double theta0, k0;
  };
  std::vector<Angle> angles;
  double angle_threshold = 0.1;  // angle_pi_approximation_threshold (pyx:354)
  // angle_is_pi as the angle closures read it: the enclosing variable at call
  // time, i.e. the value set for the last angle defined (pyx:1334-1338, 1378)
  double angle_is_pi = -1.0;
  bool angle_error = false;
  // grid
  std::vector<double> cl;
  std::vector<int32_t> ncell, pbc_nc;
  std::vector<int32_t> cell_id;
  // rng
  std::mt19937 gen;
  const double* inj_normals = nullptr;
  const double* inj_u1 = nullptr;
  const double* inj_u2 = nullptr;
  int64_t inj_steps = 0, inj_base = 1;
  int64_t cur_step = 0;
  // reaction diagnostics
  int64_t events = 0;
  std::vector<double> rates, total;
  double rs_exact = 0.0;

  double normal(int64_t step, int32_t j, int k) {
    if (s.rng_mode == RNG_MT) {
      std::normal_distribution<double> dist(0.0, 1.0);
      return dist(gen);
    } else if (s.rng_mode == RNG_INJECTED) {
      int64_t o = step - inj_base;
      if (o < 0 || o >= inj_steps) throw std::string("injected normals exhausted");
      return inj_normals[(o * n + j) * d + k];
    } else {
      double nn[2];
      philox_normals(s.seed, step, j, nn, k < 2 ? 0u : 2u);
      return nn[k & 1];
    }
  }
  double uniform(int64_t step, int32_t j, int which) {
    if (s.rng_mode == RNG_MT) {
      std::uniform_real_distribution<double> dist(0.0, 1.0);
      return dist(gen);
    } else if (s.rng_mode == RNG_INJECTED) {
      int64_t o = step - inj_base;
      if (o < 0 || o >= inj_steps) throw std::string("injected uniforms exhausted");
      return (which == 0 ? inj_u1 : inj_u2)[o * n + j];
    } else {
      double u[2];
      philox_uniforms(s.seed, step, j, u);
      return u[which];
    }
  }

  // pyx:1415-1445
  void setup_cells() {
    cl.assign(d, 0.0);
    ncell.assign(d, 0);
    pbc_nc.assign(d, 0);
    double r = (double)s.ratio;
    for (int k = 0; k < d; ++k) {
      int l = 1;
      // With rs = 0 (no interaction range and no rs given) the reference's
      // loop never ends; one cell is the only meaningful grid then.
      if (s.rs > 0.0)
        while (s.L[k] / (double)(l + 1) >= s.rs / r) ++l;
      cl[k] = s.L[k] / (double)l;
      ncell[k] = l;
      pbc_nc[k] = l - s.ratio;
    }
    cell_id.assign((size_t)n * d, 0);
  }

  // pyx:1303-1320: bonded pairs seed the neighbour list
  void setup_bonded() {
    bonded_set.assign(n, {});
    for (auto& b : bonds) {
      int i1 = std::min(b.i1, b.i2), i2 = std::max(b.i1, b.i2);
      bonded_set[i1].insert(i2);
    }
    bonded_static = bonded_set;
    bond_d.assign(bonds.size(), 0.0);
    bond_hat.assign(bonds.size() * d, 0.0);
  }

  void pair_vec(int i1, int i2, double* dx, double* dist) {
    double d2 = 0.0;
    for (int k = 0; k < d; ++k) {
      double v = x[(size_t)i2 * d + k] - x[(size_t)i1 * d + k];
      if (s.periodic[k] > 0) v = v - s.L[k] * std::round(v / s.L[k]);
      d2 += v * v;
      dx[k] = v;
    }
    *dist = std::sqrt(d2);
  }

  // pyx:1590-1652 (and the bond distances that live in d_xx)
  void calc_distances() {
    nl_d.assign(n, {});
    nl_hat.assign(n, {});
    double dx[3], dist;
    for (int i = 0; i < n; ++i) {
      auto& L = nl[i];
      nl_d[i].resize(L.size());
      nl_hat[i].resize(L.size() * d);
      for (size_t q = 0; q < L.size(); ++q) {
        pair_vec(i, L[q], dx, &dist);
        nl_d[i][q] = dist;
        for (int k = 0; k < d; ++k) nl_hat[i][q * d + k] = dx[k] / dist;
      }
    }
    if (s.nl_mode == NL_EXACT) {
      for (size_t b = 0; b < bonds.size(); ++b) {
        pair_vec(bonds[b].i1, bonds[b].i2, dx, &dist);
        bond_d[b] = dist;
        for (int k = 0; k < d; ++k) bond_hat[b * d + k] = dx[k] / dist;
      }
      return;
    }
    // refresh bond distances for bonds whose pair is listed
    for (size_t b = 0; b < bonds.size(); ++b) {
      int i1 = bonds[b].i1, i2 = bonds[b].i2;
      int lo = std::min(i1, i2), hi = std::max(i1, i2);
      auto& L = nl[lo];
      auto it = std::lower_bound(L.begin(), L.end(), hi);
      if (it != L.end() && *it == hi) {
        size_t q = it - L.begin();
        bond_d[b] = nl_d[lo][q];
        for (int k = 0; k < d; ++k) {
          double h = nl_hat[lo][q * d + k];
          bond_hat[b * d + k] = (i1 == lo) ? h : -h;
        }
      }
    }
  }

  // pyx:1448-1587 (faithful) or every pair within rs_exact (exact)
  void create_neighbor_list() {
    if (s.nl_mode == NL_EXACT && !(rs_exact > 0.0) && bonds.empty()) {
      // no pair interaction, reaction radius or bond: nothing to list
      nl.assign(n, {});
      calc_distances();
      return;
    }
    if (s.nl_mode == NL_EXACT) {
      nl.assign(n, {});
      double dx[3], dist;
      for (int i = 0; i < n; ++i) {
        std::vector<int32_t> row(bonded_static[i].begin(), bonded_static[i].end());
        for (int j = i + 1; j < n; ++j) {
          if (molecule[i] == molecule[j]) continue;
          pair_vec(i, j, dx, &dist);
          if (!(dist > rs_exact)) row.push_back(j);
        }
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        // bonded pairs obey the same distance filter as in pyx:1578-1586
        std::vector<int32_t> keep;
        for (int j : row) {
          pair_vec(i, j, dx, &dist);
          if (!(dist > rs_exact)) keep.push_back(j);
        }
        nl[i].swap(keep);
      }
      calc_distances();
      return;
    }
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < d; ++k)
        cell_id[(size_t)i * d + k] = (int32_t)std::floor(x[(size_t)i * d + k] / cl[k]);
    for (int i = 0; i < n; ++i) {
      int ti = types[i];
      for (int j = i + 1; j < n; ++j) {
        if (molecule[i] == molecule[j]) continue;
        int tj = types[j];
        if (!partners[(size_t)tj * s.n_species + ti]) continue;
        bool nb = true;
        for (int k = 0; k < d; ++k) {
          int dc = std::abs(cell_id[(size_t)j * d + k] - cell_id[(size_t)i * d + k]);
          if (dc <= s.ratio) continue;
          if (s.periodic[k] > 0) {
            if (dc < pbc_nc[k]) { nb = false; break; }
          } else {
            nb = false;
            break;
          }
        }
        if (nb) bonded_set[i].insert(j);  // mutates the shared set (pyx:1513)
      }
    }
    nl.assign(n, {});
    for (int i = 0; i < n; ++i) nl[i].assign(bonded_set[i].begin(), bonded_set[i].end());
    calc_distances();
    // drop pairs farther than rs (pyx:1578-1586)
    for (int i = 0; i < n; ++i) {
      std::vector<int32_t> keep;
      std::vector<double> kd, kh;
      for (size_t q = 0; q < nl[i].size(); ++q) {
        if (nl_d[i][q] > s.rs) continue;
        keep.push_back(nl[i][q]);
        kd.push_back(nl_d[i][q]);
        for (int k = 0; k < d; ++k) kh.push_back(nl_hat[i][q * d + k]);
      }
      nl[i].swap(keep);
      nl_d[i].swap(kd);
      nl_hat[i].swap(kh);
    }
  }

  static double pair_force(const orc_interaction& it, double r) {
    switch (it.kind) {
      case K_LJ: {
        double r6 = std::pow(it.r0 / r, (double)it.n);
        return -2. * it.n * it.u0 / r * r6 * (r6 - 1.);
      }
      case K_TANH: {
        double ainv = 1. / it.a;
        double t = std::tanh((r - it.r0) * ainv);
        return it.u0 * (1. - t * t) * ainv / 2.;
      }
      case K_YUKAWA:
        return -it.u0 * (1. + it.r0 / r) * std::exp(-r / it.r0) / (r * it.r0);
      case K_HARMONIC:
        return it.u0 * r;
      case K_PWLINEAR:
        return (r < it.r0 - it.dr) ? 0.0 : it.u0 / it.dr;
    }
    return 0.0;
  }

  // pyx:1941-2067
  void update_drifts() {
    std::fill(a.begin(), a.end(), 0.0);
    for (auto& it : inter) {
      if (it.kind == K_WALL) {
        double ainv = 1. / it.a;
        for (int i = 0; i < n; ++i) {
          if (types[i] != it.p0) continue;
          double t = std::tanh((x[(size_t)i * d + it.p1] - it.r0) * ainv);
          double f = it.u0 * (1. - t * t) * ainv / 2.;
          a[(size_t)i * d + it.p1] += -f;
        }
        continue;
      }
      for (int i1 = 0; i1 < n; ++i1) {
        auto& L = nl[i1];
        for (size_t q = 0; q < L.size(); ++q) {
          int i2 = L[q];
          if (molecule[i1] == molecule[i2]) continue;
          int p1 = std::min(types[i1], types[i2]), p2 = std::max(types[i1], types[i2]);
          if (!(p1 == it.p0 && p2 == it.p1)) continue;
          double r = nl_d[i1][q];
          if (r > it.rc) continue;
          double f = pair_force(it, r);
          for (int k = 0; k < d; ++k) {
            double h = nl_hat[i1][q * d + k];
            a[(size_t)i1 * d + k] += f * h;
            a[(size_t)i2 * d + k] += -f * h;
          }
        }
      }
    }
    for (size_t b = 0; b < bonds.size(); ++b) {
      double f = bonds[b].u0 * (bond_d[b] - bonds[b].r0);
      for (int k = 0; k < d; ++k) {
        double h = bond_hat[b * d + k];
        a[(size_t)bonds[b].i1 * d + k] += f * h;
        a[(size_t)bonds[b].i2 * d + k] += -f * h;
      }
    }
    // pyx:2052-2065 with the closure of pyx:1340-1391
    for (const Angle& g : angles) {
      double v1[3], v2[3], d1, d2;
      pair_vec(g.i2, g.i1, v1, &d1);  // x[i1] - x[i2], minimum image
      pair_vec(g.i2, g.i3, v2, &d2);  // x[i3] - x[i2]
      double u1[3], u2[3], h = 0.0;
      for (int k = 0; k < d; ++k) {
        u1[k] = v1[k] / d1;
        u2[k] = v2[k] / d2;
        h += u1[k] * u2[k];
      }
      double theta;
      if (h < -1.) {
        if (h < -1.0000001) {
          angle_error = true;
          char buf[160];
          snprintf(buf, sizeof buf,
                   "Inner product of units vectors is smaller than -1. Value = %.17g", h);
          g_err = buf;
        }
        h = -1.;
        theta = M_PI;
      } else {
        theta = std::acos(h);
      }
      double prefac;
      if (angle_is_pi > 0.) {
        if (theta < g.theta0 - angle_threshold) {
          prefac = g.k0 * (theta - g.theta0) / std::sqrt(1. - h * h);
        } else {
          double t = theta - g.theta0;
          prefac = -g.k0 / (1. - t * t / 6. + (t * t * t * t) / 120.);
        }
      } else {
        prefac = g.k0 * (theta - g.theta0) / std::sqrt(1. - h * h);
      }
      for (int k = 0; k < d; ++k) {
        double f = prefac / d1 * (u2[k] - h * u1[k]);
        a[(size_t)g.i1 * d + k] += f;
        a[(size_t)g.i2 * d + k] -= f;
        f = prefac / d2 * (u1[k] - h * u2[k]);
        a[(size_t)g.i3 * d + k] += f;
        a[(size_t)g.i2 * d + k] -= f;
      }
    }
  }

  // pyx:2070-2222
  void reaction_step(int64_t step) {
    int R = (int)rx.size();
    rates.assign((size_t)R * n, 0.0);
    total.assign(n, 0.0);
    for (int r = 0; r < R; ++r) {
      const orc_reaction& q = rx[r];
      double* rr = &rates[(size_t)r * n];
      if (q.kind == RX_LINEAR) {
        for (int j = 0; j < n; ++j)
          if (types[j] == q.reactant) { rr[j] = q.rate; total[j] += q.rate; }
      } else if (q.kind == RX_CATALYZED) {
        for (int j = 0; j < n; ++j) {
          auto& L = nl[j];
          if (types[j] == q.reactant)
            for (size_t p = 0; p < L.size(); ++p)
              if (types[L[p]] == q.catalyst && nl_d[j][p] < q.r0) {
                rr[j] += q.rate; total[j] += q.rate;
              }
          if (types[j] == q.catalyst)
            for (size_t p = 0; p < L.size(); ++p)
              if (types[L[p]] == q.reactant && nl_d[j][p] < q.r0) {
                rr[L[p]] += q.rate; total[L[p]] += q.rate;
              }
        }
      } else if (q.kind == RX_INHIBITED) {
        for (int j = 0; j < n; ++j) rr[j] = (types[j] == q.reactant) ? q.rate : 0.0;
        for (int j = 0; j < n; ++j) {
          auto& L = nl[j];
          int tj = types[j];
          if (tj == q.reactant)
            for (size_t p = 0; p < L.size(); ++p) {
              int tk = types[L[p]];
              if ((tk == q.inhibitor || tk == q.inhibitor2) && nl_d[j][p] < q.r0)
                rr[j] -= q.rate_decrease;
            }
          if (tj == q.inhibitor || tj == q.inhibitor2)
            for (size_t p = 0; p < L.size(); ++p)
              if (types[L[p]] == q.reactant && nl_d[j][p] < q.r0)
                rr[L[p]] -= q.rate_decrease;
        }
        for (int j = 0; j < n; ++j) {
          if (rr[j] < 0) rr[j] = 0.0;
          total[j] += rr[j];
        }
      }
    }
    for (int i = 0; i < n; ++i) {
      double k = total[i];
      double p = (s.reaction_prob == 0) ? s.dt * k : -std::expm1(-s.dt * k);
      if (p > uniform(step, i, 0)) {
        double u = uniform(step, i, 1);
        double acc = 0.0;
        for (int r = 0; r < R; ++r) {
          acc += rates[(size_t)r * n + i] / k;
          if (acc >= u) {
            types[i] = rx[r].product;
            ++events;
            break;
          }
        }
      }
    }
  }

  // pyx:2225-2250
  void update_positions(int64_t step) {
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < d; ++k) {
        size_t o = (size_t)j * d + k;
        int t = types[j];
        x[o] += fpre[t] * a[o];
        x[o] += s.include_random_force * npre[t] * normal(step, j, k);
      }
  }

  // pyx:2253-2292
  bool apply_bc() {
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < d; ++k) {
        size_t o = (size_t)j * d + k;
        if (s.periodic[k] < 0) {
          if (x[o] > s.L[k]) x[o] -= 2 * (x[o] - s.L[k]);
          if (x[o] < 0.) x[o] -= 2 * x[o];
        } else {
          x[o] = py_mod(x[o], s.L[k]);
        }
        if (x[o] < 0 || x[o] > s.L[k]) {
          char buf[256];
          snprintf(buf, sizeof buf,
                   "x[%d,%d] = %.17g out of bounds. Maybe the timestep is too "
                   "large for the given the interaction parameters?",
                   j, k, x[o]);
          g_err = buf;
          return false;
        }
      }
    return true;
  }

  // pyx:2296-2322
  bool step(int64_t stepno) {
    update_positions(stepno);
    if (!apply_bc()) return false;
    if (s.nl_mode == NL_EXACT || stepno % s.nl_freq == 0)
      create_neighbor_list();
    else
      calc_distances();
    reaction_step(stepno);
    update_drifts();
    return !angle_error;
  }
};

}  // namespace

extern "C" {

void* orc_create(const orc_spec* spec, const double* x0, const int32_t* types0) {
  try {
    Oracle* o = new Oracle();
    o->s = *spec;
    o->n = spec->n;
    o->d = spec->n_dim;
    o->inter.assign(spec->inter, spec->inter + spec->n_inter);
    o->rx.assign(spec->rx, spec->rx + spec->n_rx);
    o->bonds.assign(spec->bonds, spec->bonds + spec->n_bonds);
    o->molecule.assign(spec->molecule, spec->molecule + spec->n);
    o->partners.assign(spec->partners, spec->partners + (size_t)spec->n_species * spec->n_species);
    o->fpre.assign(spec->force_prefac, spec->force_prefac + spec->n_species);
    o->npre.assign(spec->noise_prefac, spec->noise_prefac + spec->n_species);
    o->x.assign(x0, x0 + (size_t)spec->n * spec->n_dim);
    o->a.assign((size_t)spec->n * spec->n_dim, 0.0);
    o->types.assign(types0, types0 + spec->n);
    o->s.inter = nullptr; o->s.rx = nullptr; o->s.bonds = nullptr;
    o->s.molecule = nullptr; o->s.partners = nullptr;
    o->s.force_prefac = nullptr; o->s.noise_prefac = nullptr;
    // rs of the exact mode: the largest pair cutoff or reaction radius
    double rmax = 0.0;
    for (auto& it : o->inter)
      if (it.kind != K_WALL) rmax = std::max(rmax, it.rc);
    for (auto& q : o->rx)
      if (q.kind == RX_CATALYZED || q.kind == RX_INHIBITED) rmax = std::max(rmax, q.r0);
    o->rs_exact = rmax;
    if (spec->rng_mode == RNG_MT) {
      // pyx:104-108: mt19937(seed) with the seed already reduced mod 2^63-1
      o->gen = std::mt19937((unsigned long)spec->seed);
    }
    o->setup_cells();
    o->setup_bonded();
    return o;
  } catch (...) {
    g_err = "orc_create failed";
    return nullptr;
  }
}

void orc_destroy(void* h) { delete (Oracle*)h; }

int orc_set_injected(void* h, const double* normals, const double* u1, const double* u2,
                     int64_t n_steps, int64_t step_base) {
  Oracle* o = (Oracle*)h;
  o->inj_normals = normals;
  o->inj_u1 = u1;
  o->inj_u2 = u2;
  o->inj_steps = n_steps;
  o->inj_base = step_base;
  return 0;
}

// Initial neighbour list and drift, as simulate() does before its loop
// (pyx:2398-2401).
int orc_init(void* h) {
  Oracle* o = (Oracle*)h;
  o->create_neighbor_list();
  o->update_drifts();
  return 0;
}

// Angles (pyx:1322-1407), in definition order: participants, theta0 [rad],
// k0 = u0 [per rad^2], and the pi-approximation threshold (pyx:354).  Their
// legs join the bonded neighbour list (pyx:1405-1406).  Call before orc_init.
int orc_set_angles(void* h, int64_t n, const int32_t* i1, const int32_t* i2,
                   const int32_t* i3, const double* theta0, const double* k0,
                   double threshold) {
  Oracle* o = (Oracle*)h;
  o->angles.clear();
  for (int64_t q = 0; q < n; ++q) {
    o->angles.push_back(Oracle::Angle{i1[q], i2[q], i3[q], theta0[q], k0[q]});
    auto leg = [&](int32_t p, int32_t r) {
      int32_t lo = std::min(p, r), hi = std::max(p, r);
      o->bonded_set[lo].insert(hi);
      o->bonded_static[lo].insert(hi);
    };
    leg(i1[q], i2[q]);
    leg(i2[q], i3[q]);
  }
  o->angle_threshold = threshold;
  // np.isclose(theta0, pi, atol=1e-5): |a - b| <= atol + rtol |b|, rtol = 1e-5
  if (n > 0)
    o->angle_is_pi = std::fabs(theta0[n - 1] - M_PI) <= 1e-5 + 1e-5 * M_PI ? 1.0 : -1.0;
  return 0;
}

// Run integration steps first_step .. first_step+n_steps-1 (the reference's
// step argument l+1).  Returns 0, or -3 on an out-of-bounds coordinate
// (message in orc_last_error) or an angle whose cosine is below -1.0000001
// (pyx:1364-1370), -1 on an exhausted injected stream.
int orc_run(void* h, int64_t first_step, int64_t n_steps) {
  Oracle* o = (Oracle*)h;
  try {
    for (int64_t q = 0; q < n_steps; ++q) {
      if (!o->step(first_step + q)) return -3;
      o->cur_step = first_step + q;
    }
  } catch (const std::string& e) {
    g_err = e;
    return -1;
  }
  return 0;
}

int orc_get_state(void* h, double* x, int32_t* types, double* drift) {
  Oracle* o = (Oracle*)h;
  if (x) std::memcpy(x, o->x.data(), o->x.size() * sizeof(double));
  if (types) std::memcpy(types, o->types.data(), o->types.size() * sizeof(int32_t));
  if (drift) std::memcpy(drift, o->a.data(), o->a.size() * sizeof(double));
  return 0;
}

// Replaces positions, types and the drift of the last step (each optional);
// the neighbour list is kept as it is until its next refresh.
int orc_set_state(void* h, const double* x, const int32_t* types, const double* drift) {
  Oracle* o = (Oracle*)h;
  if (x) std::memcpy(o->x.data(), x, o->x.size() * sizeof(double));
  if (types) std::memcpy(o->types.data(), types, o->types.size() * sizeof(int32_t));
  if (drift) std::memcpy(o->a.data(), drift, o->a.size() * sizeof(double));
  return 0;
}

// Per-particle reaction rates of the last reaction step, [n_rx][n] then
// total [n]; returns the number of reaction events so far.
int64_t orc_reaction_info(void* h, double* rates, double* total) {
  Oracle* o = (Oracle*)h;
  if (rates && !o->rates.empty()) std::memcpy(rates, o->rates.data(), o->rates.size() * sizeof(double));
  if (total && !o->total.empty()) std::memcpy(total, o->total.data(), o->total.size() * sizeof(double));
  return o->events;
}

// Number of listed pairs (i<j) in the current neighbour list.
int64_t orc_nl_size(void* h) {
  Oracle* o = (Oracle*)h;
  int64_t c = 0;
  for (auto& L : o->nl) c += (int64_t)L.size();
  return c;
}

// Dump the current neighbour list as (i, j) pairs; returns the count.
int64_t orc_nl_pairs(void* h, int32_t* out, int64_t cap) {
  Oracle* o = (Oracle*)h;
  int64_t c = 0;
  for (int i = 0; i < o->n; ++i)
    for (int j : o->nl[i]) {
      if (c < cap) { out[2 * c] = i; out[2 * c + 1] = j; }
      ++c;
    }
  return c;
}

// Raw draws of the reference stream: kinds[q] = 0 normal, 1 uniform.  A
// fresh distribution object per draw (pyx:71-85).
int orc_mt_sequence(uint64_t seed, int64_t n, const int8_t* kinds, double* out) {
  std::mt19937 gen((unsigned long)seed);
  for (int64_t q = 0; q < n; ++q) {
    if (kinds[q] == 0) {
      std::normal_distribution<double> dist(0.0, 1.0);
      out[q] = dist(gen);
    } else {
      std::uniform_real_distribution<double> dist(0.0, 1.0);
      out[q] = dist(gen);
    }
  }
  return 0;
}

int orc_mt_raw(uint64_t seed, int64_t n, uint32_t* out) {
  std::mt19937 gen((unsigned long)seed);
  for (int64_t q = 0; q < n; ++q) out[q] = (uint32_t)gen();
  return 0;
}

// The engine's Philox noise / reaction uniforms for particles [0, n) at
// one step (see philox_normals / philox_uniforms above).
int orc_philox_streams(uint64_t seed, int64_t step, int32_t n, double* normals, double* u) {
  for (int32_t i = 0; i < n; ++i) {
    if (normals) philox_normals(seed, step, i, normals + 2 * (size_t)i);
    if (u) philox_uniforms(seed, step, i, u + 2 * (size_t)i);
  }
  return 0;
}

// Pair force law f(r) = dU/dr for one interaction (pyx:1068-1130), and the
// wall law (pyx:1134-1136) when kind == 6 (r is then the coordinate).
double orc_force_law(const orc_interaction* it, double r) {
  if (it->kind == K_WALL) {
    double ainv = 1. / it->a;
    double t = std::tanh((r - it->r0) * ainv);
    return it->u0 * (1. - t * t) * ainv / 2.;
  }
  return Oracle::pair_force(*it, r);
}

}  // extern "C"
