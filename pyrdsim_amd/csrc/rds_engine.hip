// MI355X (gfx950) engine for PyRDSim's overdamped-Langevin time-step loop.
//
// What one step does (the reference's integration_step, pyrdsim.pyx:2296-2322):
//   x_{n+1}   = BC( x_n + fpre*F(x_n, tau_n) + npre*xi )     update_positions,
//                                                            apply_boundary_conditions
//   tau_{n+1} = R( x_{n+1}, tau_n )                          reaction_step
//   F_{n+1}   = F( x_{n+1}, tau_{n+1} )                      update_drifts
//
// How it runs here.  Particles live in HBM sorted by cell ("slots"), FP64
// positions as double2, int32 type / pid / molecule columns.  Many independent
// systems ("replicas") are concatenated: replica r owns slots
// [pbase_r, pbase_r + N_r) and cells [cbase_r, cbase_r + lx_r*ly_r).  Per step:
//   k_bonds    (molecular models only) harmonic bond and angle forces per slot
//   k_force    one workgroup of R wavefronts per tile of R cell rows by W
//              cells (home particles dealt 64 at a time): the (R+2) x (W+2) cells are
//              staged in LDS with their periodic images already shifted; each
//              lane owns one home particle and walks its three stencil row
//              windows in LDS; the force is consumed at once by the
//              Euler-Maruyama update, the boundary conditions and the binning
//              of the new position (per-cell histogram, one atomic per run of
//              lanes, and the rank within the cell)
//   k_cellscan cell starts from the histogram in one pass, in chunks of <= 2048
//              cells (a replica spreads over many workgroups): each chunk
//              publishes its count and sums its replica's earlier chunks';
//              histogram cleared for the next step, step counter advanced
//   k_scatter  counting-sort reorder of the new state into the other buffer;
//              with reactions it also decides the particles whose rates need
//              no neighbours and lists the others
//   k_react_walk (models with catalysed / inhibited reactions) the listed
//              particles count their neighbours within r0 and decide, unless
//              the particle count of their stencil already shows that no count
//              could make them react
// Three dimensions run the same sequence with k3_tile (tiles of 2 x 2 cell
// rows by W cells, nine row windows per lane), k3_scatter and k3_react.
// Positions, the minimum-image displacement and the integrator are FP64; the
// pair-force arithmetic is FP32 (RDS_MIXED) or FP64 (RDS_FP64).
//
// Random numbers: Philox4x32-10 (Salmon et al., SC'11) keyed by the replica
// seed, counter (pid, step, stream); Box-Muller in FP32 for the two normals of
// a particle, two 53-bit uniforms for the reactions; or injected buffers.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "rds.h"

#define MAXT RDS_MAX_TYPES
#define MAXTERMS RDS_MAX_TERMS
#define MAXWALLS RDS_MAX_WALLS
#define MAXRX RDS_MAX_REACTIONS
#define MAXW 60        // cells per tile strip
#define HMAX 1024      // home particles per tile with a direct home-number -> cell table

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(RDS_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// device tables
// ---------------------------------------------------------------------------

struct DTerm {
  int kind, n, pw, odd;  // pw: number of factors of the LJ power, odd: n odd
  double u0, r0, rc2, a_inv, rin, c_lj, r0sq, slope;
  float fu0, fr0, frc2, fa_inv, frin, fc_lj, fr0sq, fslope;
};

struct DWall {
  int type, dir, pad0, pad1;
  double u0, r0, a_inv;
};

struct DRx {
  int kind, reactant, product, catalyst, inh, inh2, pad0, pad1;
  double r0, rate_decrease;
  double r0sq_lo, r0sq_hi;  // r0^2 (1 -+ 1e-14): outside this band d2 decides d < r0
};

struct DModel {
  int nt, nterms, nwalls, nrx;
  int nbr_rx;      // any catalysed / inhibited reaction
  unsigned walk_types;  // bit t: type t is the reactant of a catalysed / inhibited reaction
  int pad1, pad2;
  double rc2max;   // largest pair cutoff squared
  int pfirst[MAXT * MAXT];
  int pcount[MAXT * MAXT];
  DTerm terms[MAXTERMS];
  DWall walls[MAXWALLS];
  DRx rx[MAXRX];
};

struct DRep {
  double L[3], cl[3];     // box and cell lengths per axis (z: three dimensions only)
  double icl[3];          // 1 / cl, rounded
  int lx, ly, cbase, pbase;
  int n, model, px, py;   // periodic axes (boundary conditions)
  unsigned seed_lo, seed_hi;
  int sx, sy;             // periodic images in the cell stencil
  int lz, pz, sz;         // the z axis (lz = 1 in two dimensions)
  int ncell;              // lx * ly * lz
  double fpre[MAXT], npre[MAXT], rate[MAXRX];
  double dt;
  double rl2;             // list mode: (largest pair cutoff + skin)^2
  float Lh[2], Ll[2];     // box lengths as float pairs (compensated positions)
  float fph[MAXT], fpl[MAXT], nph[MAXT], npl[MAXT];  // prefactors as float pairs
};

struct DStatus {
  unsigned long long err_key;  // (step << 32) | (pid*2 + axis), min over errors
  double err_val;              // the coordinate of the particle err_key names
  int bounds, stream;
  int err_lock;                // guards err_key / err_val together
  int err_lost;                // a failing lane could not take the lock
  unsigned long long events;
  int integrity;               // bit 0: scatter index out of range, bit 1: histogram total,
                               // bit 2: cell-start look-back gave up waiting
  int rebuild;                 // list mode: set when a particle moved more than skin/2
  unsigned long long rebuilds, overflow, entries;
  unsigned done;               // blocks of the list build that finished (reset by the last)
  int angle;                   // an angle's cosine fell below -1.0000001
  int cl_changed;              // cluster labels: a label changed in this pass
  double angle_val;            // that cosine
};

#define LIST_OVER 0xFFFF  // list length marking a particle that uses the stencil scan

// Everything a kernel of one step needs.  Passed by value.
struct StepArgs {
  const DRep* rep;
  const DModel* models;
  const int4* tiles;  // (replica, cy, cx0, w)
  int n_tiles;
  int n_total;
  const int* cstart;  // [C_total + 1]
  // current sorted state
  const double2* pos;
  const int* type;
  const int* pid;
  const int* mol;
  // next-step outputs of k_force
  double2* upos;
  int* ucell;
  int* urank;
  int* count;
  // bonds
  const double2* fbond;
  // force-only / count outputs (caller order)
  double2* fout;
  int* iout;
  double rprobe2;
  // random numbers
  const long long* step;
  const double* inj_n;
  const double* inj_u1;
  const double* inj_u2;
  long long inj_first, inj_count;
  int rng, prob_mode;
  double include_rf;
  DStatus* status;
  // reaction output
  int* type_out;
  // staged particles per LDS round (dynamic shared memory)
  int cap;
  const int* rep_of;  // replica of each pid
  const int4* sblocks;  // (replica, first slot, end slot) blocks of <= 256 slots
  // list mode
  uint2* nl4;           // [K/4][n_total] staged indices of each slot's partners,
                        // four 16-bit entries per chunk, the last chunk padded
                        // with its first entry
  unsigned short* ncnt; // [n_total] list length, or LIST_OVER
  int K;                // multiple of 4
  double2* xref;        // position of each slot at the last rebuild
  const int* scell;     // cell of each slot at the last rebuild (null: cell mode)
  double hskin2;        // (skin / 2)^2
  int cpos;             // positions stored as float pairs (compensated FP32)
  const int* boff;      // bond CSR by pid (null without bonds)
  const int* bpart;
  // cell mode: slots whose reaction rates need a neighbour count, listed by
  // slot block (sblocks): block b's in wl[256 b ..) as (slot, cell, type,
  // pid), their number in wl_n[b]; the cell is the new one in cell mode and
  // that of the last rebuild in list mode
  int4* wl;
  int* wl_n;
  const int4* wblocks;  // (slot block, replica, model) of replicas whose model has a walk
  // three dimensions: positions {x, y, z, 0}, next positions, bond forces
  const double4* pos4;
  double4* upos4;
  const double4* fbond4;
};

// ---------------------------------------------------------------------------
// random numbers
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint4 philox4x32_10(uint4 c, unsigned k0, unsigned k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // each 32 x 32 -> 64-bit product as one multiply (both halves at once)
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c.x;
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c.z;
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
    const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Two standard normals of particle `pid` (index within its replica) at
// reference step `step`.
// (Stream 0; the third normal of a three-dimensional particle comes from
// stream 2, the same way.)
__device__ __forceinline__ void philox_normals(const DRep& R, int pid, long long step,
                                               float& n0, float& n1, unsigned stream = 0u) {
  uint4 w = philox4x32_10(make_uint4((unsigned)pid, (unsigned)step,
                                     (unsigned)((unsigned long long)step >> 32), stream),
                          R.seed_lo, R.seed_hi);
  float ua = (float)((w.x >> 9) * 2u + 1u) * 0x1.0p-24f;  // odd / 2^24, in (0,1)
  float ub = (float)(w.y >> 8) * 0x1.0p-24f;              // [0,1)
  // -2 ln(ua) = -2 ln(2) log2(ua); ua >= 2^-24 is a normal float, so the
  // hardware log2 and square root (about 1 ulp each) need no range handling
  float rr = __builtin_amdgcn_sqrtf(-1.38629436f * __builtin_amdgcn_logf(ua));
  float s, c;
  sincospif(2.0f * ub, &s, &c);
  n0 = rr * c;
  n1 = rr * s;
}

// The two reaction uniforms u1 (decision) and u2 (choice), 53-bit.
__device__ __forceinline__ void philox_uniforms(const DRep& R, int pid, long long step,
                                                double& u1, double& u2) {
  uint4 w = philox4x32_10(make_uint4((unsigned)pid, (unsigned)step,
                                     (unsigned)((unsigned long long)step >> 32), 1u),
                          R.seed_lo, R.seed_hi);
  unsigned long long a = ((unsigned long long)w.y << 32) | w.x;
  unsigned long long b = ((unsigned long long)w.w << 32) | w.z;
  u1 = (double)(a >> 11) * 0x1.0p-53;
  u2 = (double)(b >> 11) * 0x1.0p-53;
}

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

// Cell coordinate floor(x / cl) (the reference's x // cl, pyrdsim.pyx:1497),
// clamped into [0, l): a periodic wrap may return exactly L, and a coordinate
// that left the box (reported separately) must still index a valid cell.
// The reference's minimum image dx - L*round(dx/L) (pyrdsim.pyx:1628-1640).
// For |dx| < L/4 the rounded quotient is +-0 and the result is dx (up to the
// sign of a zero, which no distance sees), so the division is skipped there.
__device__ __forceinline__ double min_image(double dx, double L) {
  return fabs(dx) < 0.25 * L ? dx : dx - L * round(dx / L);
}

// Compensated FP32 positions (RDS_COMPENSATED): each coordinate is the
// unevaluated sum h + l of two floats, |l| <= ulp(h)/2, stored as
// float4 {xh, yh, xl, yl} in the 16-byte slot of a double2.  Its value is
// exactly (double)h + (double)l.  The step updates it with float arithmetic
// only (Knuth's two-sum / Dekker's fast-two-sum); everything else reads the
// exact FP64 value.
struct DF {
  float h, l;
};
__device__ __forceinline__ DF two_sum(float a, float b) {
#pragma clang fp contract(off)
  const float s = a + b;
  const float bb = s - a;
  return DF{s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DF fast_two_sum(float a, float b) {  // |a| >= |b|
#pragma clang fp contract(off)
  const float s = a + b;
  return DF{s, b - (s - a)};
}
__device__ __forceinline__ DF df_add(DF x, float y) {
#pragma clang fp contract(off)
  const DF t = two_sum(x.h, y);
  return fast_two_sum(t.h, t.l + x.l);
}
__device__ __forceinline__ DF df_add(DF x, DF y) {
#pragma clang fp contract(off)
  const DF t = two_sum(x.h, y.h);
  return fast_two_sum(t.h, t.l + (x.l + y.l));
}
// (a.h + a.l)(b.h + b.l) to about 2^-44 relative: the exact product of the
// heads (fma two-product) plus the cross terms
__device__ __forceinline__ DF df_mul(DF a, DF b) {
#pragma clang fp contract(off)
  const float p = a.h * b.h;
  const float e = __builtin_fmaf(a.h, b.h, -p);
  return fast_two_sum(p, e + (a.h * b.l + a.l * b.h));
}
__device__ __forceinline__ DF df_of(double x) {
  const float h = (float)x;
  return DF{h, (float)(x - (double)h)};
}
__device__ __forceinline__ bool df_lt(DF a, DF b) { return a.h < b.h || (a.h == b.h && a.l < b.l); }
__device__ __forceinline__ double df_val(DF a) { return (double)a.h + (double)a.l; }
__device__ __forceinline__ float4 df_split(double x, double y) {
  const float xh = (float)x, yh = (float)y;
  return make_float4(xh, yh, (float)(x - (double)xh), (float)(y - (double)yh));
}
// position of slot s as FP64, either storage
__device__ __forceinline__ double2 rd_pos(const double2* __restrict__ p, int s, int cpos) {
  if (!cpos) return p[s];
  const float4 v = reinterpret_cast<const float4*>(p)[s];
  return make_double2((double)v.x + (double)v.z, (double)v.y + (double)v.w);
}

// Records a coordinate that left the box: the smallest key (step, pid, axis)
// over all failures and the coordinate of the particle it names, written
// together under a lock (failures are rare, so the lock costs nothing on the
// normal path).  The lane that takes the lock does its update in the same
// loop iteration, so lanes of one wave cannot wait on each other; the spin is
// bounded, and a lane that never gets the lock sets err_lost instead.
__device__ __noinline__ void record_bounds(DStatus* st, unsigned long long key, double val) {
  atomicExch(&st->bounds, 1);
  if (key >= atomicAdd(&st->err_key, 0ull)) return;
  for (int spin = 0; spin < (1 << 20); ++spin) {
    if (atomicCAS(&st->err_lock, 0, 1) == 0) {
      if (key < atomicAdd(&st->err_key, 0ull)) {
        atomicExch(&st->err_key, key);
        atomicExch(reinterpret_cast<unsigned long long*>(&st->err_val),
                   (unsigned long long)__double_as_longlong(val));
      }
      __threadfence();
      atomicExch(&st->err_lock, 0);
      return;
    }
  }
  atomicExch(&st->err_lost, 1);
}

// q / d for 0 <= q < 2^20 and 1 <= d < 2^10 from inv_d = 1.0f / d: exact,
// since (q + 1/2) / d lies at least 1 / (2d) from an integer and the float
// product errs by less than that
__device__ __forceinline__ int div_small(int q, float inv_d) {
  return (int)(((float)q + 0.5f) * inv_d);
}

__device__ __forceinline__ int cell_coord(double x, double cl, int l) {
  double q = floor(x / cl);
  if (!(q >= 0.0)) return 0;
  if (q >= (double)l) return l - 1;
  return (int)q;
}

// The same cell coordinate from a multiplication by inv = 1/cl (rounded):
// x*inv is within a few ulp of x/cl, so its floor is floor(x/cl) except
// where the quotient lies within that of an integer; those rare cases take
// the division.
__device__ __forceinline__ int cell_coord_fast(double x, double cl, double inv, int l) {
  const double t = x * inv;
  const double q = floor(t);
  if (t - q > 1e-9 && t - q < 1.0 - 1e-9) {
    if (!(q >= 0.0)) return 0;
    if (q >= (double)l) return l - 1;
    return (int)q;
  }
  return cell_coord(x, cl, l);
}

// Block -> tile map that gives each XCD a contiguous run of tiles (workgroups
// are dealt to the 8 XCDs round-robin; neighbouring tiles share halo rows, which
// should then meet in the same L2).
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int X = 8;
  int x = b % X, k = b / X;
  int q = n / X, r = n % X;
  return x * q + min(x, r) + k;
}

template <typename Real>
struct Real2 {
  Real x, y;
};

// ---------------------------------------------------------------------------
// tiles: R cell rows x W cells.  The staged region is rows cy0-1 .. cy0+R and
// cells cx0-1 .. cx0+W of each row, periodic images shifted by -L / +L.  Each
// staged row is at most 3 contiguous slot segments (left image, in-range run,
// right image).  A tile is served by tile_waves(R) wavefronts (R up to 4, four
// for taller tiles), which take its home particles 64 at a time.
// ---------------------------------------------------------------------------

constexpr int tile_waves(int R) { return R < 4 ? R : 4; }

template <int R>
struct TileSh {
  int cs[R + 2][MAXW + 2];   // first slot of staged cell k (k = 0..W+1)
  int off[R + 2][MAXW + 3];  // staged offset of cell k within its row (prefix)
  int row_base[R + 3];
  int seg_slot[R + 2][3], seg_n[R + 2][3], seg_off[R + 2][3];
  double seg_sx[R + 2][3];
  double row_sy[R + 2];
  int hp[R + 1];             // home particles of the home rows before row q
  int chp[R * MAXW + 1];     // home particles of the home cells before cell i = q * W + k
  unsigned short hkey[HMAX];  // home number -> its cell (q << 6 | k), when <= HMAX of them
  int total, nbatch;
};

// Exclusive prefix sum of a[0..n) in place, by the 64 lanes of the calling
// wave (n <= 64 * PER); returns the total to every lane.
template <int PER>
__device__ __forceinline__ int wave_exclusive_scan(int* a, int n, int lane) {
  int v[PER];
  int sum = 0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int c = lane * PER + i;
    v[i] = c < n ? a[c] : 0;
    sum += v[i];
  }
  int inc = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  int run = inc - sum;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int c = lane * PER + i;
    if (c < n) a[c] = run;
    run += v[i];
  }
  return __shfl(inc, 63, 64);
}

template <int R>
__device__ void tile_setup(const DRep& Rp, const int* __restrict__ cstart, int cy0, int cx0,
                           int W, int nrows, TileSh<R>& sh) {
  const int t = threadIdx.x, nthr = blockDim.x;
  const int NR = nrows + 2;
  const int lx = Rp.lx, ly = Rp.ly;
  // the staged cells' starts, two per thread with both loads in flight
  const int ncs = NR * (W + 2);
  const float inv_w2 = 1.0f / (float)(W + 2);
  for (int q0 = t; q0 < ncs; q0 += 2 * nthr) {
    int c[2], a[2], e[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int q = q0 + u * nthr;
      const int r = div_small(q, inv_w2), k = q - r * (W + 2);
      int y = cy0 + r - 1;
      bool ok = q < ncs;
      if (y < 0) { if (Rp.sy) y += ly; else ok = false; }
      else if (y >= ly) { if (Rp.sy) y -= ly; else ok = false; }
      int x = cx0 - 1 + k;
      if (x < 0) { if (Rp.sx) x += lx; else ok = false; }
      else if (x >= lx) { if (Rp.sx) x -= lx; else ok = false; }
      c[u] = ok ? Rp.cbase + y * lx + x : -1;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      a[u] = c[u] >= 0 ? cstart[c[u]] : 0;
      e[u] = c[u] >= 0 ? cstart[c[u] + 1] : 0;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int q = q0 + u * nthr;
      if (q < ncs) {
        const int r = div_small(q, inv_w2), k = q - r * (W + 2);
        sh.cs[r][k] = a[u];
        sh.off[r][k + 1] = e[u] - a[u];
      }
    }
  }
  __syncthreads();
  {
    const int wv = t >> 6, lane = t & 63;
    for (int r = wv; r < NR; r += (nthr >> 6)) {
      int v = (lane < W + 2) ? sh.off[r][lane + 1] : 0;
      int inc = v;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        int o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
      }
      if (lane < W + 2) sh.off[r][lane + 1] = inc;
      if (lane == 0) sh.off[r][0] = 0;
    }
  }
  __syncthreads();
  if (t < NR) {
    const int r = t;
    int y = cy0 + r - 1;
    bool rowok = true;
    double sy = 0.0;
    if (y < 0) { if (Rp.sy) sy = -Rp.L[1]; else rowok = false; }
    else if (y >= ly) { if (Rp.sy) sy = Rp.L[1]; else rowok = false; }
    sh.row_sy[r] = sy;
    for (int g = 0; g < 3; ++g) sh.seg_n[r][g] = 0;
    if (rowok) {
      if (cx0 == 0 && Rp.sx) {
        sh.seg_slot[r][0] = sh.cs[r][0];
        sh.seg_n[r][0] = sh.off[r][1];
        sh.seg_off[r][0] = 0;
        sh.seg_sx[r][0] = -Rp.L[0];
      }
      int ka = (cx0 == 0) ? 1 : 0;
      int kb = (cx0 + W >= lx) ? W : W + 1;
      sh.seg_slot[r][1] = sh.cs[r][ka];
      sh.seg_n[r][1] = sh.off[r][kb + 1] - sh.off[r][ka];
      sh.seg_off[r][1] = sh.off[r][ka];
      sh.seg_sx[r][1] = 0.0;
      if (cx0 + W >= lx && Rp.sx) {
        sh.seg_slot[r][2] = sh.cs[r][W + 1];
        sh.seg_n[r][2] = sh.off[r][W + 2] - sh.off[r][W + 1];
        sh.seg_off[r][2] = sh.off[r][W + 1];
        sh.seg_sx[r][2] = Rp.L[0];
      }
    }
  }
  if (t < 64) {
    // first staged index of each staged row and home particles before each
    // home row: prefix sums over the lanes (NR <= R + 2 <= 10 rows)
    const int rt = t < NR ? sh.off[t][W + 2] : 0;
    const int ht = t < nrows ? sh.off[t + 1][W + 1] - sh.off[t + 1][1] : 0;
    int ir = rt, ih = ht;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      const int a = __shfl_up(ir, d, 64), e = __shfl_up(ih, d, 64);
      if (t >= d) {
        ir += a;
        ih += e;
      }
    }
    if (t <= NR) sh.row_base[t] = ir - rt;
    if (t == NR) sh.total = ir;
    if (t <= nrows) sh.hp[t] = ih - ht;
    // the home particles of all rows are dealt to the workgroup's waves, 64
    // per wave
    if (t == nrows) sh.nbatch = (ih + nthr - 1) / nthr;
  }
  // Home cells in row-major order (cell i = q * W + k): the home particles
  // before each cell, and for each home number its cell.
  const int nc = nrows * W;
  const float inv_w = 1.0f / (float)W;
  for (int c = t; c < nc; c += nthr) {
    const int q = div_small(c, inv_w), k = c - q * W;
    sh.chp[c] = sh.off[q + 1][k + 2] - sh.off[q + 1][k + 1];
  }
  __syncthreads();
  if (t < 64) {
    const int tot = wave_exclusive_scan<(R * MAXW + 63) / 64>(sh.chp, nc, t);
    if (t == 0) sh.chp[nc] = tot;
  }
  __syncthreads();
  if (sh.chp[nc] <= HMAX) {
    for (int c = t; c < nc; c += nthr) {
      const int q = div_small(c, inv_w), k = c - q * W;
      const unsigned short key = (unsigned short)((q << 6) | k);
      for (int g = sh.chp[c]; g < sh.chp[c + 1]; ++g) sh.hkey[g] = key;
    }
    __syncthreads();
  }
}

// List mode: a periodic coordinate may have wrapped since the slot was sorted
// into `cell`; the image meant by the stencil is the one nearest that cell's
// centre (unambiguous: cells are >= range + skin wide and the particle has
// moved less than skin/2).
__device__ __forceinline__ double2 near_cell(double2 q, const DRep& Rp, int cell) {
  const int c = cell - Rp.cbase;
  const int cy = c / Rp.lx, cx = c - cy * Rp.lx;
  if (Rp.px) {
    const double d = q.x - ((double)cx + 0.5) * Rp.cl[0];
    if (d > 0.5 * Rp.L[0]) q.x -= Rp.L[0];
    else if (d < -0.5 * Rp.L[0]) q.x += Rp.L[0];
  }
  if (Rp.py) {
    const double d = q.y - ((double)cy + 0.5) * Rp.cl[1];
    if (d > 0.5 * Rp.L[1]) q.y -= Rp.L[1];
    else if (d < -0.5 * Rp.L[1]) q.y += Rp.L[1];
  }
  return q;
}

// Stages staged indices [p0, p0 + cap) into LDS (types shifted left by TSH).  scell (list mode): cells of
// the last rebuild, whose nearest images are staged.  cpos: the stored
// positions are float pairs; SPM: stage float pairs (1) or FP64 (0).  Each
// thread takes STAGE_U staged indices at a time and issues all their loads
// before the first LDS store, so that a tile's staging costs one round trip
// to memory per STAGE_U x blockDim indices rather than one per row segment.
#define STAGE_U 2
template <int R, bool MOL, int SPM = 0, int TSH = 0>
__device__ void tile_stage(const TileSh<R>& sh, int NR, int p0, int cap,
                           const double2* __restrict__ pos, const int* __restrict__ type,
                           const int* __restrict__ mol, double2* s_pos, int* s_ty, int* s_mol,
                           const DRep& Rp, const int* __restrict__ scell, int cpos = 0) {
  const int t = threadIdx.x, nthr = blockDim.x;
  const int p1 = min(p0 + cap, sh.total);
  for (int pb = p0 + t; pb < p1; pb += STAGE_U * nthr) {
    int src[STAGE_U], rg[STAGE_U];
#pragma unroll
    for (int u = 0; u < STAGE_U; ++u) {
      const int p = min(pb + u * nthr, p1 - 1);
      // row: the last with row_base <= p; segment: the last non-empty one
      // starting at or before p's offset in the row
      int r = 0;
#pragma unroll
      for (int q = 1; q < R + 2; ++q) r += (q < NR && sh.row_base[q] <= p) ? 1 : 0;
      const int loc = p - sh.row_base[r];
      int g = 0;
      if (sh.seg_n[r][1] > 0 && loc >= sh.seg_off[r][1]) g = 1;
      if (sh.seg_n[r][2] > 0 && loc >= sh.seg_off[r][2]) g = 2;
      src[u] = sh.seg_slot[r][g] + (loc - sh.seg_off[r][g]);
      rg[u] = r * 3 + g;
    }
    double2 q[STAGE_U];
    int ty[STAGE_U], mo[STAGE_U], sc[STAGE_U];
#pragma unroll
    for (int u = 0; u < STAGE_U; ++u) {
      q[u] = rd_pos(pos, src[u], cpos);
      ty[u] = type[src[u]];
      mo[u] = MOL ? mol[src[u]] : 0;
      sc[u] = scell ? scell[src[u]] : 0;
    }
#pragma unroll
    for (int u = 0; u < STAGE_U; ++u) {
      const int p = pb + u * nthr;
      if (p >= p1) break;
      double2 v = q[u];
      if (scell) v = near_cell(v, Rp, sc[u]);
      const int r = rg[u] / 3, g = rg[u] - 3 * r;
      v.x = v.x + sh.seg_sx[r][g];
      v.y = v.y + sh.row_sy[r];
      if (SPM) reinterpret_cast<float4*>(s_pos)[p - p0] = df_split(v.x, v.y);
      else s_pos[p - p0] = v;
      s_ty[p - p0] = ty[u] << TSH;
      if (MOL) s_mol[p - p0] = mo[u];
    }
  }
}

// The slot after the last staged particle of a round: position (0, 0), type
// 0 (any finite position and valid type; what is read there is masked).
__device__ __forceinline__ void stage_sentinel(double2* s_pos, int* s_ty, int* s_mol, int k,
                                               bool mol) {
  s_pos[k] = make_double2(0.0, 0.0);
  s_ty[k] = 0;
  if (mol) s_mol[k] = -1;
}

// Home particle of lane `lane` of wave wv in batch b, and its three windows.
// The tile's home particles, taken cell by cell in row-major order of the
// home cells, are dealt to the waves 64 at a time: home number
// b * blockDim + wv * 64 + lane.  The home cell is the cell whose slot range holds
// the particle (the cell it was sorted into).
template <int R>
struct Home {
  int s;
  bool valid;
  int w, own;
  int lo[3], hi[3];
};

// s_home (a resident tile staged in FP64 without list-mode images): xi is
// read from the staged copy, which holds home cells unshifted.
template <int R>
__device__ __forceinline__ void home_windows(const TileSh<R>& sh, int wv, int b, int lane,
                                             int nrows, int W,
                                             const double2* __restrict__ pos, Home<R>& h,
                                             double2& xi, int cpos = 0,
                                             const double2* s_home = nullptr) {
  const int g = b * (int)blockDim.x + wv * 64 + lane;
  const int nc = nrows * W;
  const int nh = sh.chp[nc];
  h.valid = g < nh;
  // the cell holding home number g: from the table, or the largest i with
  // chp[i] <= g
  int q = 0, k = 0, lo = 0;
  if (nh <= HMAX) {
    const int key = h.valid ? sh.hkey[g] : 0;
    q = key >> 6;
    k = key & 63;
    lo = q * W + k;
  } else {
    int hi = nc;
    if (h.valid) {
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (sh.chp[mid] <= g) lo = mid;
        else hi = mid;
      }
    }
    q = div_small(lo, 1.0f / (float)W);
    k = lo - q * W;
  }
  const int hr = q + 1;  // staged row of the home row
  xi = make_double2(0.0, 0.0);
  h.s = 0;
  h.w = k;
  h.own = -1;
  if (h.valid) {
    const int r = g - sh.chp[lo];
    h.s = sh.cs[hr][k + 1] + r;
    h.own = sh.row_base[hr] + sh.off[hr][k + 1] + r;
    xi = s_home ? s_home[h.own] : rd_pos(pos, h.s, cpos);
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int rr = q + d;
    h.lo[d] = h.valid ? sh.row_base[rr] + sh.off[rr][k] : 0;
    h.hi[d] = h.valid ? sh.row_base[rr] + sh.off[rr][k + 3] : 0;
  }
}

// ---------------------------------------------------------------------------
// pair laws (pyrdsim.pyx:1068-1130).  All return c with F_i += c * (x_j - x_i),
// i.e. c = f(r) / r where f = dU/dr is what the reference's closures return.
// ---------------------------------------------------------------------------

template <typename Real>
__device__ __forceinline__ Real rcp_(Real v) { return Real(1) / v; }
template <>
__device__ __forceinline__ float rcp_<float>(float v) { return __builtin_amdgcn_rcpf(v); }

__device__ __forceinline__ float exp_(float v) { return __expf(v); }
__device__ __forceinline__ double exp_(double v) { return exp(v); }
__device__ __forceinline__ float sqrt_(float v) { return __builtin_sqrtf(v); }
__device__ __forceinline__ double sqrt_(double v) { return sqrt(v); }

template <typename Real> struct TP;
template <> struct TP<float> {
  __device__ static float u0(const DTerm& t) { return t.fu0; }
  __device__ static float r0(const DTerm& t) { return t.fr0; }
  __device__ static float rc2(const DTerm& t) { return t.frc2; }
  __device__ static float a_inv(const DTerm& t) { return t.fa_inv; }
  __device__ static float rin(const DTerm& t) { return t.frin; }
  __device__ static float c_lj(const DTerm& t) { return t.fc_lj; }
  __device__ static float r0sq(const DTerm& t) { return t.fr0sq; }
  __device__ static float slope(const DTerm& t) { return t.fslope; }
};
template <> struct TP<double> {
  __device__ static double u0(const DTerm& t) { return t.u0; }
  __device__ static double r0(const DTerm& t) { return t.r0; }
  __device__ static double rc2(const DTerm& t) { return t.rc2; }
  __device__ static double a_inv(const DTerm& t) { return t.a_inv; }
  __device__ static double rin(const DTerm& t) { return t.rin; }
  __device__ static double c_lj(const DTerm& t) { return t.c_lj; }
  __device__ static double r0sq(const DTerm& t) { return t.r0sq; }
  __device__ static double slope(const DTerm& t) { return t.slope; }
};

template <typename Real>
__device__ __forceinline__ Real pair_coef(const DTerm& t, Real r2) {
  typedef TP<Real> P;
  switch (t.kind) {
    case RDS_LJ: {
      // f = -2n u0 / r * s (s - 1), s = (r0/r)^n  =>  c = -2n u0 s (s-1) / r^2
      Real inv = rcp_(r2);
      Real b = t.odd ? P::r0(t) * sqrt_(inv) : P::r0sq(t) * inv;
      Real s = b;
      for (int k = 1; k < t.pw; ++k) s *= b;
      return P::c_lj(t) * inv * s * (s - Real(1));
    }
    case RDS_TANH: {
      // f = u0 (1 - tanh^2((r - r0)/a)) / (2a); 1 - tanh^2(z) = 4e/(1+e)^2, e = exp(-2|z|)
      Real r = sqrt_(r2);
      Real z = (r - P::r0(t)) * P::a_inv(t);
      Real e = exp_(Real(-2) * fabs(z));
      Real q = Real(1) + e;
      Real f = P::u0(t) * (Real(4) * e / (q * q)) * P::a_inv(t) / Real(2);
      return f / r;
    }
    case RDS_YUKAWA: {
      // f = -u0 (1 + r0/r) exp(-r/r0) / (r r0)
      Real r = sqrt_(r2);
      Real f = -P::u0(t) * (Real(1) + P::r0(t) / r) * exp_(-r / P::r0(t)) / (r * P::r0(t));
      return f / r;
    }
    case RDS_HARMONIC:
      return P::u0(t);  // f = u0 r
    case RDS_PIECEWISE_LINEAR: {
      Real r = sqrt_(r2);
      return (r < P::rin(t)) ? Real(0) : P::slope(t) / r;
    }
  }
  return Real(0);
}

// Wall force along its direction (pyrdsim.pyx:1134-1136, 1979-1986):
// a[i, dir] += -u0 (1 - tanh^2((x - r0)/a)) / (2a)
// The wall law f = u0 (1 - tanh^2 z) a_inv / 2, z = (x - r0) a_inv
// (pyrdsim.pyx:1134-1136), returned as the drift -f.  1 - tanh^2 z is
// sech^2 z = 4e / (1 + e)^2 with e = exp(-2|z|): one exponential in place of
// tanh (several times its instruction count), and no cancellation where
// tanh^2 z is close to 1.  Real = float: single-precision exponential and
// reciprocal, as the pair forces of that precision.
template <typename Real = double>
__device__ __forceinline__ double wall_force(const DWall& w, double x) {
  if (sizeof(Real) == 4) {
    const float z = (float)((x - w.r0) * w.a_inv);
    const float e = expf(-2.0f * fabsf(z));
    const float q = 1.0f + e;
    const float s = 4.0f * e * __builtin_amdgcn_rcpf(q * q);
    return -(w.u0 * (double)s * w.a_inv * 0.5);
  }
  const double z = (x - w.r0) * w.a_inv;
  const double e = exp(-2.0 * fabs(z));
  const double q = 1.0 + e;
  return -(w.u0 * (4.0 * e / (q * q)) * w.a_inv / 2.0);
}


// ---------------------------------------------------------------------------
// force (+ integrate + bin) kernel.  FAST: every type pair of every model has
// at most one term, all LJ with the same even exponent 2*pw, at most 2 types;
// the pair constants then live in registers.  Otherwise the term table is in
// LDS and each term goes through pair_coef.
// ---------------------------------------------------------------------------

enum { MODE_STEP = 0, MODE_FORCE = 1 };

// dynamic LDS of a tile kernel: cap x (double2 position, int type [, int molecule])
struct Staging {
  double2* pos;
  int* ty;
  int* mol;
};
__device__ __forceinline__ Staging staging(char* smem, int cap) {
  Staging g;
  g.pos = (double2*)smem;
  g.ty = (int*)(g.pos + cap);
  g.mol = g.ty + cap;
  return g;
}

// KM: 0 = generic term table; 1 = single LJ term per type pair, runtime power;
// 2 = the same with n = 6 (power 3) unrolled.
template <typename Real, int KM>
__device__ __forceinline__ Real lj_fast(Real r2, Real pa, Real pb, int pw) {
  const Real inv = rcp_(r2);
  const Real bb = pb * inv;
  Real s;
  if (KM == 2) {
    s = bb * bb * bb;
  } else {
    s = bb;
    for (int k = 1; k < pw; ++k) s *= bb;
  }
  // pa / r^2 * s (s - 1), with s (s - 1) = s*s - s in one fused operation
  return (pa * inv) * fma(s, s, -s);
}

// n = 6 with the constants folded: c_lj / r^2 * s (s - 1), s = r0^6 / r^6,
// equals r^-8 * (A r^-6 - B) with A = c_lj r0^12 and B = c_lj r0^6.
__device__ __forceinline__ float lj_fast6(float r2, float A, float B) {
  const float inv = rcp_(r2);
  const float i2 = inv * inv;
  return (i2 * i2) * fmaf(A, i2 * inv, -B);
}

// 0 < r2 <= rc2 for r2 >= +0 or NaN and rc2 >= 0, as one unsigned compare: the
// bit patterns of non-negative floats are ordered like their values, +0 maps
// to 0 (and wraps to the largest pattern), NaN lies above every finite value.
// rc2 = 0 (no term for the type pair) admits nothing.
__device__ __forceinline__ bool in_range(float r2, float rc2) {
  return __float_as_uint(r2) - 1u < __float_as_uint(rc2);
}
__device__ __forceinline__ bool in_range(double r2, double rc2) {
  return r2 > 0.0 && r2 <= rc2;
}

// PM: 0 = FP64 positions, 1 = compensated FP32 positions (float pairs; the
// pair differences and the update are float arithmetic).
// Waves per SIMD the register budget of a force kernel must allow: the
// register-resident LJ path, in the 4-row tiles that large batches use,
// gains from occupancy (cell stencil: <= 72 VGPRs, 7 waves per SIMD, with
// FP32 pair arithmetic and <= 80, 6 waves, with FP64; lists: 6 waves).
// Smaller grids gain nothing from it and keep the compiler's choice, as does
// the general term table.
#define FORCE_WAVES32 7
#define FORCE_WAVES64 6
#define FORCE_WAVES_LIST 6
template <typename Real, int KM, int R, bool LIST>
constexpr int force_waves() {
  return (KM == 0 || R < 4) ? 1
         : LIST              ? FORCE_WAVES_LIST
                             : (sizeof(Real) == 4 ? FORCE_WAVES32 : FORCE_WAVES64);
}

template <typename Real, bool MOL, int KM, int R, bool LIST, int PM>
__global__ __launch_bounds__(64 * tile_waves(R), (force_waves<Real, KM, R, LIST>()))
void k_force(const StepArgs A, int mode, int pw) {
  constexpr bool FAST = KM != 0;
  extern __shared__ __align__(16) char smem[];
  const int CAPT = A.cap;
  // one slot beyond the round capacity: a finite sentinel after the staged
  // particles, read (masked) by the second candidate of a window's last trip
  const Staging G = staging(smem, CAPT + 1);
  double2* s_pos = G.pos;
  int* s_ty = G.ty;
  int* s_mol = G.mol;
  __shared__ TileSh<R> sh;
  __shared__ int s_pfirst[FAST ? 1 : MAXT * MAXT], s_pcount[FAST ? 1 : MAXT * MAXT];
  __shared__ DTerm s_term[FAST ? 1 : MAXTERMS];

  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int tile = xcd_tile(blockIdx.x, A.n_tiles);
  const int4 T = A.tiles[tile];
  const DRep& Rp = A.rep[T.x & 0xffffff];
  const int nrows = (T.x >> 24) + 1;
  const DModel& M = A.models[Rp.model];
  const int cy0 = T.y, cx0 = T.z, W = T.w;
  const int nt = M.nt;
  const int NR = nrows + 2;
  const size_t NS = (size_t)A.n_total;

  if (!FAST) {
    for (int q = t; q < nt * nt; q += blockDim.x) {
      s_pfirst[q] = M.pfirst[q];
      s_pcount[q] = M.pcount[q];
    }
    for (int q = t; q < M.nterms; q += blockDim.x) s_term[q] = M.terms[q];
  }
  // FAST: the constants (c_lj, r0^2, cutoff^2) of the type pairs (a, b),
  // a, b in {0, 1}, zero where the pair has no term, and the prefactors of
  // types 0 and 1: one table per tile in LDS, read per lane by type
  __shared__ Real s_fc[FAST ? 4 : 1][3];
  __shared__ double s_fp[FAST ? 2 : 1][2];
  // FP32 pair arithmetic: the same constants as float4 entries (a, b) at
  // byte 32 a + 16 b, the staged types holding 16 b, so that a candidate's
  // constants are one LDS read at its lane's row plus its staged type; for
  // n = 6 the entries hold (c_lj r0^12, c_lj r0^6, cutoff^2), rounded once
  // from FP64
  constexpr int TSH = (FAST && sizeof(Real) == 4) ? 4 : 0;
  constexpr bool LJ6 = TSH && KM == 2;
  // FP32 pair arithmetic on FP64 staging, stencil search: windows walked one
  // at a time
  constexpr bool WIN = TSH && !LIST && PM == 0;
  __shared__ __align__(16) float4 s_fc4[TSH ? 4 : 1];
  if (FAST && t < 4) {
    const int a = t >> 1, bb = t & 1, q = a * nt + bb;
    Real v0 = 0, v1 = 0, v2 = 0;
    double A6 = 0.0, B6 = 0.0;
    if (a < nt && bb < nt && M.pcount[q] > 0) {
      const DTerm& tm = M.terms[M.pfirst[q]];
      v0 = TP<Real>::c_lj(tm);
      v1 = TP<Real>::r0sq(tm);
      v2 = TP<Real>::rc2(tm);
      const double r6 = tm.r0sq * tm.r0sq * tm.r0sq;
      B6 = tm.c_lj * r6;
      A6 = B6 * r6;
    }
    s_fc[t][0] = v0;
    s_fc[t][1] = v1;
    s_fc[t][2] = v2;
    if (TSH)
      s_fc4[t] = LJ6 ? make_float4((float)A6, (float)B6, (float)v2, 0.f)
                     : make_float4((float)v0, (float)v1, (float)v2, 0.f);
    if (bb == 0) {
      s_fp[a][0] = Rp.fpre[a];
      s_fp[a][1] = Rp.npre[a];
    }
  }
  tile_setup<R>(Rp, A.cstart, cy0, cx0, W, nrows, sh);
  const int total = sh.total;
  const bool resident = total <= CAPT;
  const int* scell = LIST ? A.scell : nullptr;
  if (resident) {
    tile_stage<R, MOL, PM, TSH>(sh, NR, 0, CAPT, A.pos, A.type, A.mol, s_pos, s_ty, s_mol, Rp, scell,
                           PM);
    if (t == 0) stage_sentinel(s_pos, s_ty, s_mol, total, MOL);
    __syncthreads();
  }
  long long step_next = 0;
  if (mode == MODE_STEP) step_next = *A.step + 1;

  // resident FP64 staging (not list mode): the home particles' position, type
  // and molecule are read from LDS
  const bool lds_home = resident && !LIST && PM == 0;

  for (int b = 0; b < sh.nbatch; ++b) {
    Home<R> h;
    double2 xi;
    home_windows<R>(sh, wv, b, lane, nrows, W, A.pos, h, xi, PM, lds_home ? s_pos : nullptr);
    int ti = 0, mi = 0, p = 0;
    // position in the frame of the staged tile (list mode: nearest image to
    // the sorted cell); xi itself stays the integrator's coordinate
    double2 xs = xi;
    if (h.valid) {
      p = A.pid[h.s];
      if (lds_home) {
        ti = s_ty[h.own] >> TSH;
        if (MOL) mi = s_mol[h.own];
      } else {
        ti = A.type[h.s];
        if (MOL) mi = A.mol[h.s];
      }
      if (LIST) xs = near_cell(xi, Rp, scell[h.s]);
    }
    // list mode: the lane walks its partner list; a lane whose list overflowed
    // scans the stencil windows instead (the other lanes' windows are empty)
    int nlist = 0;
    if (LIST && h.valid) {
      nlist = A.ncnt[h.s];
      if (nlist != LIST_OVER) {
#pragma unroll
        for (int d = 0; d < 3; ++d) h.lo[d] = h.hi[d] = 0;
      } else {
        nlist = 0;
      }
    }
    // FAST: the (ti, tj) constants for tj = 0, 1 in registers (cutoff^2 = 0:
    // no term for the pair)
    Real pa0 = 0, pb0 = 0, pc0 = 0, pa1 = 0, pb1 = 0, pc1 = 0;
    if (FAST) {
      const int a = ti == 1 ? 2 : 0;
      pa0 = s_fc[a][0];
      pb0 = s_fc[a][1];
      pc0 = s_fc[a][2];
      pa1 = s_fc[a + 1][0];
      pb1 = s_fc[a + 1][1];
      pc1 = s_fc[a + 1][2];
    }
    Real fx = 0, fy = 0;
    const int tbase = ti * nt;
    // one candidate: staged index j, counted when `live`.  Listed partners are
    // never the particle itself nor of its molecule; stencil candidates are
    // filtered here: self / same molecule never interact, and a particle's only
    // zero-distance partner in an atomic model is itself.
    const float4 xsf = PM ? df_split(xs.x, xs.y) : make_float4(0.f, 0.f, 0.f, 0.f);
    const char* fc_row = reinterpret_cast<const char*>(s_fc4) + ti * 32;
    auto pairv = [&](const Real dx, const Real dy, int tj, int mj, bool live, bool listed) {
      const Real r2 = dx * dx + dy * dy;
      const bool other = (MOL && listed) ? true : (MOL ? (mj != mi) : (r2 > Real(0)));
      if (FAST) {
        Real pa, pb, pc;
        if (TSH) {
          // tj = 16 * type: the (ti, tj) constants from the tile's table
          const float4 kc = *reinterpret_cast<const float4*>(fc_row + tj);
          pa = (Real)kc.x;
          pb = (Real)kc.y;
          pc = (Real)kc.z;
        } else {
          const bool m1 = tj == 1;
          pa = m1 ? pa1 : pa0;
          pb = m1 ? pb1 : pb0;
          pc = m1 ? pc1 : pc0;
        }
        Real c;
        if constexpr (LJ6) c = lj_fast6(r2, pa, pb);
        else c = lj_fast<Real, KM>(r2, pa, pb, pw);
        // atomic models: r2 > 0 excludes the particle itself, inside in_range
        const bool take = MOL ? (other && r2 <= pc) : in_range(r2, pc);
        c = (live && take) ? c : Real(0);
        fx += c * dx;
        fy += c * dy;
      } else {
        const int pq = tbase + tj;
        const int q0 = s_pfirst[pq], qn = live ? s_pcount[pq] : 0;
        for (int q = q0; q < q0 + qn; ++q) {
          const DTerm& tm = s_term[q];
          Real c = pair_coef<Real>(tm, r2);
          c = (other && r2 <= TP<Real>::rc2(tm)) ? c : Real(0);
          fx += c * dx;
          fy += c * dy;
        }
      }
    };
    auto pair = [&](int j, bool live, bool listed) {
      Real dx, dy;
      if (PM) {
        // difference of float pairs: (h_j - h_i) + (l_j - l_i)
        const float4 q = reinterpret_cast<const float4*>(s_pos)[j];
        dx = (Real)((q.x - xsf.x) + (q.z - xsf.z));
        dy = (Real)((q.y - xsf.y) + (q.w - xsf.w));
      } else {
        const double2 pj = s_pos[j];
        dx = (Real)(pj.x - xs.x);
        dy = (Real)(pj.y - xs.y);
      }
      pairv(dx, dy, s_ty[j], MOL ? s_mol[j] : 0, live, listed);
    };
    const uint2* lp = A.nl4 + (LIST && h.valid ? h.s : 0);
    int kcur = 0;  // list cursor across LDS rounds
    for (int p0 = 0; p0 < total; p0 += CAPT) {
      if (!resident) {
        __syncthreads();
        tile_stage<R, MOL, PM, TSH>(sh, NR, p0, CAPT, A.pos, A.type, A.mol, s_pos, s_ty, s_mol, Rp,
                               scell, PM);
        if (t == 0) stage_sentinel(s_pos, s_ty, s_mol, min(total - p0, CAPT), MOL);
        __syncthreads();
      }
      if (LIST) {
        if (resident) {
          // four partners per chunk, the chunk after next loaded while this
          // one is computed; padding entries repeat a staged index, masked
          const int nch = (nlist + 3) >> 2;
          uint2 ck = nch > 0 ? lp[0] : make_uint2(0u, 0u);
          uint2 ck1 = nch > 1 ? lp[NS] : ck;
          for (int k = 0; k < nch; ++k) {
            const uint2 nx = lp[(size_t)min(k + 2, nch - 1) * NS];
            const int b4 = 4 * k;
            pair((int)(ck.x & 0xffffu), true, true);
            pair((int)(ck.x >> 16), b4 + 1 < nlist, true);
            pair((int)(ck.y & 0xffffu), b4 + 2 < nlist, true);
            pair((int)(ck.y >> 16), b4 + 3 < nlist, true);
            ck = ck1;
            ck1 = nx;
          }
        } else {
          // lists are ascending in staged index: take this round's entries
          const int pend = p0 + CAPT;
          while (kcur < nlist) {
            const uint2 w = lp[(size_t)(kcur >> 2) * NS];
            const unsigned h2 = (kcur & 2) ? w.y : w.x;
            const int j = (int)((kcur & 1) ? (h2 >> 16) : (h2 & 0xffffu));
            if (j >= pend) break;
            pair(j - p0, true, true);
            ++kcur;
          }
        }
      }
      if constexpr (WIN) {
        // the three row windows (clipped to this LDS round) one after the
        // other, two consecutive staged particles per trip: the addresses
        // advance linearly, at the cost of a wave running each window's
        // longest lane.  Past a window's end the second candidate is the next
        // staged particle or the sentinel, finite either way, and masked.
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int lo = max(h.lo[d], p0) - p0;
          const int hi = min(h.hi[d], p0 + CAPT) - p0;
          const double2* pq = s_pos + lo;
          const double2* const pe = s_pos + hi;
          const double2* const pl = pe - 1;  // a pair with its second past the end
          const int* tq = s_ty + lo;
          const int* mq = s_mol + lo;
          for (; pq < pe; pq += 2, tq += 2, mq += MOL ? 2 : 0) {
            const double2 q0 = pq[0], q1 = pq[1];
            const int t0 = tq[0], t1 = tq[1];
            const int m0 = MOL ? mq[0] : 0, m1 = MOL ? mq[1] : 0;
            pairv((Real)(q0.x - xs.x), (Real)(q0.y - xs.y), t0, m0, true, false);
            pairv((Real)(q1.x - xs.x), (Real)(q1.y - xs.y), t1, m1, pq < pl, false);
          }
        }
      } else {
        // the three row windows (clipped to this LDS round) as one sequence,
        // so that a lane's trip count is its total; two candidates per trip,
        // past the end the second one repeats a staged index with its
        // contribution masked (0 * dx of an unstaged one could be 0 * NaN)
        int nn[3], oo[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int lo = max(h.lo[d], p0) - p0;
          const int hi = min(h.hi[d], p0 + CAPT) - p0;
          nn[d] = max(hi - lo, 0);
          oo[d] = lo;
        }
        const int b1 = nn[0], b2 = nn[0] + nn[1], n = b2 + nn[2];
        const int o0 = oo[0], o1 = oo[1] - b1, o2 = oo[2] - b2;
        for (int u = 0; u < n; u += 2) {
          const int u1 = min(u + 1, n - 1);
          const int j0 = u + (u < b1 ? o0 : (u < b2 ? o1 : o2));
          const int j1 = u1 + (u1 < b1 ? o0 : (u1 < b2 ? o1 : o2));
          pair(j0, true, false);
          pair(j1, u + 1 < n, false);
        }
      }
    }
    if (!h.valid) continue;
    const int s = h.s;
    double Fx = (double)fx, Fy = (double)fy;
    for (int q = 0; q < M.nwalls; ++q) {
      const DWall& wl = M.walls[q];
      if (wl.type != ti) continue;
      if (wl.dir == 0) Fx += wall_force<Real>(wl, xi.x);
      else Fy += wall_force<Real>(wl, xi.y);
    }
    if (A.fbond) {
      double2 fb = A.fbond[s];
      Fx += fb.x;
      Fy += fb.y;
    }
    if (mode == MODE_FORCE) {
      A.fout[p] = make_double2(Fx, Fy);
      continue;
    }
    // Euler-Maruyama, pyrdsim.pyx:2244-2250: two separate FP64 additions.
    double n0, n1;
    if (A.rng == RDS_RNG_INJECTED) {
      long long o = step_next - A.inj_first;
      if (o >= 0 && o < A.inj_count) {
        const double* src = A.inj_n + (o * A.n_total + p) * 2;
        n0 = src[0];
        n1 = src[1];
      } else {
        n0 = n1 = 0.0;
        A.status->stream = 1;
      }
    } else {
      float a0, a1;
      philox_normals(Rp, p - Rp.pbase, step_next, a0, a1);
      n0 = a0;
      n1 = a1;
    }
    double v[2];
    if (PM) {
      // compensated FP32: x + fp*F, then + np*xi, each increment a float-pair
      // product and each addition a float-pair addition
      const float4 hv = reinterpret_cast<const float4*>(A.pos)[s];
      DF u2[2] = {DF{hv.x, hv.z}, DF{hv.y, hv.w}};
      const DF fpd{Rp.fph[ti], Rp.fpl[ti]};
      const DF npd = A.include_rf != 0.0 ? DF{Rp.nph[ti], Rp.npl[ti]} : DF{0.f, 0.f};
      u2[0] = df_add(u2[0], df_mul(fpd, df_of(Fx)));
      u2[0] = df_add(u2[0], df_mul(npd, df_of(n0)));
      u2[1] = df_add(u2[1], df_mul(fpd, df_of(Fy)));
      u2[1] = df_add(u2[1], df_mul(npd, df_of(n1)));
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        DF u = u2[k];
        const DF Lk{Rp.Lh[k], Rp.Ll[k]};
        const DF zero{0.f, 0.f};
        if (!(k == 0 ? Rp.px : Rp.py)) {
          // single reflection, as pyrdsim.pyx:2269-2281
          if (df_lt(Lk, u)) u = df_add(DF{-u.h, -u.l}, DF{2.f * Lk.h, 2.f * Lk.l});
          if (df_lt(u, zero)) u = DF{-u.h, -u.l};
        } else {
          if (df_lt(u, zero)) u = df_add(u, Lk);
          else if (!df_lt(u, Lk)) u = df_add(u, DF{-Lk.h, -Lk.l});
        }
        u2[k] = u;
        const double uv = df_val(u);
        if (uv < 0.0 || uv > Rp.L[k] || uv != uv)
          record_bounds(A.status, ((unsigned long long)step_next << 32) |
                                      (unsigned long long)(unsigned)(p * 2 + k), uv);
        v[k] = uv;
      }
      reinterpret_cast<float4*>(A.upos)[s] = make_float4(u2[0].h, u2[1].h, u2[0].l, u2[1].l);
    } else {
      const double fp = FAST ? s_fp[ti == 1][0] : Rp.fpre[ti];
      const double np = A.include_rf * (FAST ? s_fp[ti == 1][1] : Rp.npre[ti]);
      double x = xi.x, y = xi.y;
      {
#pragma clang fp contract(off)
        x = x + fp * Fx;
        x = x + np * n0;
        y = y + fp * Fy;
        y = y + np * n1;
      }
      // boundary conditions, pyrdsim.pyx:2269-2292
      v[0] = x;
      v[1] = y;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const double L = Rp.L[k];
        double u = v[k];
        if (!(k == 0 ? Rp.px : Rp.py)) {
          if (u > L) u -= 2.0 * (u - L);
          if (u < 0.0) u -= 2.0 * u;
        } else {
          // Python-style modulo as Cython emits it for doubles: fmod, then + L
          // for a negative remainder.  A step moves far less than L, so fmod's
          // exact remainder is u itself on (-L, L) and u - L on [L, 2L) (exact:
          // L <= u < 2L); the general routine only for anything else.
          double m;
          if (u > -L && u < L) m = u;
          else if (u >= L && u < 2.0 * L) m = u - L;
          else m = fmod(u, L);
          if (m != 0.0 && ((m < 0.0) != (L < 0.0))) m += L;
          u = m;
        }
        if (u < 0.0 || u > L || u != u)
          record_bounds(A.status, ((unsigned long long)step_next << 32) |
                                      (unsigned long long)(unsigned)(p * 2 + k), u);
        v[k] = u;
      }

      A.upos[s] = make_double2(v[0], v[1]);
    }
    if (LIST) {
      // same slot; request a rebuild once the particle has moved more than
      // skin/2 (minimum image on periodic axes) from its position at the last one
      const double2 r0 = A.xref[s];
      double ex = v[0] - r0.x, ey = v[1] - r0.y;
      if (Rp.px && fabs(ex) > 0.5 * Rp.L[0]) ex -= copysign(Rp.L[0], ex);
      if (Rp.py && fabs(ey) > 0.5 * Rp.L[1]) ey -= copysign(Rp.L[1], ey);
      if (ex * ex + ey * ey > A.hskin2) A.status->rebuild = 1;
    } else {
      const int cx = cell_coord_fast(v[0], Rp.cl[0], Rp.icl[0], Rp.lx);
      const int cyy = cell_coord_fast(v[1], Rp.cl[1], Rp.icl[1], Rp.ly);
      // (the scatter recomputes the cell from the stored position)
      const int c = Rp.cbase + cyy * Rp.lx + cx;
      // ranks in the new cells, one atomic per run of lanes with the same new
      // cell: the lanes of a batch are a prefix of the wave (the valid home
      // particles), dealt cell by cell, and most particles stay in their cell
      const unsigned long long act = __ballot(1);
      const int nact = __popcll(act);
      const int cprev = __shfl_up(c, 1, 64);
      const bool head = lane == 0 || cprev != c;
      const unsigned long long heads = __ballot(head);
      const unsigned long long below = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
      const int hl = 63 - __clzll(heads & below);          // this run's first lane
      const unsigned long long above = heads & ~below;
      const int next = above ? min(__ffsll((long long)above) - 1, nact) : nact;
      int base = 0;
      if (head) base = atomicAdd(&A.count[c], next - lane);
      A.urank[s] = __shfl(base, hl, 64) + (lane - hl);
    }
  }
}

// List build (list mode): for every slot, the staged indices of the partners
// within (largest pair cutoff + skin) in FP64, ascending, excluding itself and
// its molecule; the position is recorded as the reference for the skin test.
// Runs only when a rebuild was requested; the last workgroup to finish clears
// the request.  Grid-stride over tiles, so that an idle launch is short.
template <bool MOL, int R>
__global__ __launch_bounds__(64 * tile_waves(R)) void k_build(const StepArgs A) {
  if (!A.status->rebuild) return;
  extern __shared__ __align__(16) char smem[];
  const int CAPT = A.cap;
  const Staging G = staging(smem, CAPT);
  double2* s_pos = G.pos;
  int* s_ty = G.ty;
  int* s_mol = G.mol;
  __shared__ TileSh<R> sh;
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const size_t NS = (size_t)A.n_total;
  const int K = A.K;
  for (int tile = blockIdx.x; tile < A.n_tiles; tile += gridDim.x) {
    __syncthreads();
    const int4 T = A.tiles[tile];
    const DRep& Rp = A.rep[T.x & 0xffffff];
    const int nrows = (T.x >> 24) + 1;
    const int cx0 = T.z, W = T.w;
    const int NR = nrows + 2;
    tile_setup<R>(Rp, A.cstart, T.y, cx0, W, nrows, sh);
    const int total = sh.total;
    const bool resident = total <= CAPT;
    if (resident) {
      tile_stage<R, MOL>(sh, NR, 0, CAPT, A.pos, A.type, A.mol, s_pos, s_ty, s_mol, Rp, nullptr);
      __syncthreads();
    }
    const double rl2 = Rp.rl2;
    for (int b = 0; b < sh.nbatch; ++b) {
      Home<R> h;
      double2 xi;
      home_windows<R>(sh, wv, b, lane, nrows, W, A.pos, h, xi, A.cpos);
      const int mi = (MOL && h.valid) ? A.mol[h.s] : 0;
      uint2* out = A.nl4 + (h.valid ? h.s : 0);
      int n = 0;
      bool big = false;  // a staged index beyond 16 bits
      unsigned ch[4] = {0u, 0u, 0u, 0u};  // the chunk being filled
      for (int p0 = 0; p0 < total; p0 += CAPT) {
        if (!resident) {
          __syncthreads();
          tile_stage<R, MOL>(sh, NR, p0, CAPT, A.pos, A.type, A.mol, s_pos, s_ty, s_mol, Rp,
                             nullptr);
          __syncthreads();
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int lo = max(h.lo[d], p0) - p0;
          const int hi = min(h.hi[d], p0 + CAPT) - p0;
          for (int j = lo; j < hi; ++j) {
            if (j + p0 == h.own) continue;
            if (MOL && s_mol[j] == mi) continue;
            const double2 pj = s_pos[j];
            const double dx = pj.x - xi.x, dy = pj.y - xi.y;
            if (dx * dx + dy * dy > rl2) continue;
            const int g = j + p0;
            if (g >= LIST_OVER) big = true;
            const int slot = n & 3;
            ch[0] = slot == 0 ? (unsigned)g : ch[0];
            ch[1] = slot == 1 ? (unsigned)g : ch[1];
            ch[2] = slot == 2 ? (unsigned)g : ch[2];
            ch[3] = slot == 3 ? (unsigned)g : ch[3];
            if (slot == 3 && n < K)
              out[(size_t)(n >> 2) * NS] = make_uint2(ch[0] | (ch[1] << 16), ch[2] | (ch[3] << 16));
            ++n;
          }
        }
      }
      if (h.valid && (n & 3) && n < K) {
        // partial last chunk: pad with its first entry
        const int r = n & 3;
        const unsigned e1 = r > 1 ? ch[1] : ch[0], e2 = r > 2 ? ch[2] : ch[0];
        out[(size_t)(n >> 2) * NS] = make_uint2(ch[0] | (e1 << 16), e2 | (ch[0] << 16));
      }
      if (h.valid) {
        const bool over = n > K || big;
        A.ncnt[h.s] = over ? (unsigned short)LIST_OVER : (unsigned short)n;
        A.xref[h.s] = xi;
        if (over) atomicAdd(&A.status->overflow, 1ull);
      }
      // list entries of the wave, one atomic per wave
      int e = (h.valid && n <= K && !big) ? n : 0;
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) e += __shfl_xor(e, d, 64);
      if (lane == 0 && e) atomicAdd(&A.status->entries, (unsigned long long)e);
    }
  }
  __syncthreads();
  if (t == 0) {
    __threadfence();
    const unsigned k = atomicAdd(&A.status->done, 1u);
    if (k == gridDim.x - 1) {
      A.status->done = 0;
      A.status->rebuild = 0;
      A.status->rebuilds += 1;
    }
  }
}

// Neighbour counts within rprobe (diagnostic; FP64 distances).
template <bool MOL>
__global__ __launch_bounds__(64) void k_count(const StepArgs A) {
  extern __shared__ __align__(16) char smem[];
  const int CAPT = A.cap;
  const Staging G = staging(smem, CAPT);
  double2* s_pos = G.pos;
  int* s_ty = G.ty;
  int* s_mol = G.mol;
  __shared__ TileSh<1> sh;
  const int lane = threadIdx.x;
  const int4 T = A.tiles[blockIdx.x];
  const DRep& Rp = A.rep[T.x & 0xffffff];
  const int nrows = (T.x >> 24) + 1;
  const int W = T.w, cx0 = T.z;
  // one home row at a time through a 1-row view of the tile
  for (int q = 0; q < nrows; ++q) {
    __syncthreads();
    tile_setup<1>(Rp, A.cstart, T.y + q, cx0, W, 1, sh);
    const int total = sh.total;
    for (int b = 0; b < sh.nbatch; ++b) {
      Home<1> h;
      double2 xi;
      home_windows<1>(sh, 0, b, lane, 1, W, A.pos, h, xi, A.cpos);
      int mi = (MOL && h.valid) ? A.mol[h.s] : 0;
      int cnt = 0;
      for (int p0 = 0; p0 < total; p0 += CAPT) {
        __syncthreads();
        tile_stage<1, MOL>(sh, 3, p0, CAPT, A.pos, A.type, A.mol, s_pos, s_ty, s_mol, Rp, nullptr,
                           A.cpos);
        __syncthreads();
        for (int d = 0; d < 3; ++d) {
          const int lo = max(h.lo[d], p0) - p0;
          const int hi = min(h.hi[d], p0 + CAPT) - p0;
          for (int j = lo; j < hi; ++j) {
            const bool other = MOL ? (s_mol[j] != mi) : (j + p0 != h.own);
            const double dx = s_pos[j].x - xi.x, dy = s_pos[j].y - xi.y;
            cnt += (other && dx * dx + dy * dy <= A.rprobe2) ? 1 : 0;
          }
        }
      }
      if (h.valid) A.iout[A.pid[h.s]] = cnt;
    }
  }
}

// ---------------------------------------------------------------------------
// reaction kernel (pyrdsim.pyx:2070-2222), one thread per slot.  Exact
// reference arithmetic: the pair distance is the reference's minimum image
// dx - L*round(dx/L) and sqrt (pyrdsim.pyx:1628-1640), compared with d < r0;
// rates are accumulated by repeated addition in reaction order, as the
// reference does.  Only reactants of a neighbour-dependent reaction walk their
// 3x3 cells (reading neighbours straight from the sorted arrays).
// ---------------------------------------------------------------------------

// rate of reaction q for a particle of type ti with cnt neighbours, accumulated
// as the reference does (repeated addition, pyrdsim.pyx:2109-2206); `total`
// receives the same increments in the same order.
__device__ __forceinline__ double rx_rate(const DRx& rx, double rate, int ti, int cnt,
                                          double& total) {
  if (rx.reactant != ti) return 0.0;
  if (rx.kind == RDS_RX_LINEAR) {
    total += rate;
    return rate;
  }
  if (rx.kind == RDS_RX_CATALYZED) {
    double acc = 0.0;
    for (int k = 0; k < cnt; ++k) {
      acc += rate;
      total += rate;
    }
    return acc;
  }
  double acc = rate;
  for (int k = 0; k < cnt; ++k) acc -= rx.rate_decrease;
  if (acc < 0) acc = 0.0;
  total += acc;
  return acc;
}

// Neighbour counts of up to RDS_MAX_REACTIONS reactions packed 16 bits
// each: reactions 0-3 in lo, 4-7 in hi.  A packed word keeps loops over the
// reactions rolled without an indexed array in scratch memory.
__device__ __forceinline__ int cnt_of(unsigned long long lo, unsigned long long hi, int q) {
  return (int)(((q < 4 ? lo : hi) >> (16 * (q & 3))) & 0xffffull);
}
__device__ __forceinline__ void cnt_add(unsigned long long& lo, unsigned long long& hi, int q) {
  if (q < 4) lo += 1ull << (16 * q);
  else hi += 1ull << (16 * (q - 4));
}

// Rates, draws and choice of one particle (pyrdsim.pyx:2103-2222): the new
// type of particle p (type ti, neighbour counts per reaction packed in
// c_lo / c_hi) in step `step`.
__device__ __forceinline__ int react_decide(const StepArgs& A, const DRep& Rp, const DModel& M,
                                            int ti, int p, long long step,
                                            unsigned long long c_lo, unsigned long long c_hi) {
  const int nrx = M.nrx;
  // total rate in reaction order (pyrdsim.pyx:2103-2206)
  double total_rate = 0.0;
#pragma unroll 1
  for (int q = 0; q < nrx; ++q) rx_rate(M.rx[q], Rp.rate[q], ti, cnt_of(c_lo, c_hi, q), total_rate);
  if (!(total_rate > 0.0)) return ti;
  double u1, u2;
  if (A.rng == RDS_RNG_INJECTED) {
    long long o = step - A.inj_first;
    if (o >= 0 && o < A.inj_count) {
      u1 = A.inj_u1[o * A.n_total + p];
      u2 = A.inj_u2 ? A.inj_u2[o * A.n_total + p] : 0.0;
    } else {
      u1 = 2.0;
      u2 = 0.0;
      A.status->stream = 1;
    }
  } else {
    philox_uniforms(Rp, p - Rp.pbase, step, u1, u2);
  }
  double prob = (A.prob_mode == RDS_PROB_EXP) ? -expm1(-Rp.dt * total_rate)
                                              : Rp.dt * total_rate;
  if (!(prob > u1)) return ti;
  // choice: first reaction whose cumulative share reaches u2
  // (pyrdsim.pyx:2215-2222); the rates are recomputed bit-identically
  double sacc = 0.0, dummy = 0.0;
  int chosen = -1;
#pragma unroll 1
  for (int q = 0; q < nrx; ++q) {
    sacc += rx_rate(M.rx[q], Rp.rate[q], ti, cnt_of(c_lo, c_hi, q), dummy) / total_rate;
    if (sacc >= u2) {
      chosen = q;
      break;
    }
  }
  if (chosen < 0) return ti;
  atomicAdd(&A.status->events, 1ull);
  return M.rx[chosen].product;
}

// True when particle p (type ti) cannot react this step whatever its
// neighbour counts, each at most cmax: with every catalysed count at cmax and
// every inhibitor count at 0 the total rate is at its largest (rx_rate's sum
// grows with each catalysed count and shrinks with each inhibitor count when
// no rate is negative), and even that gives a probability below u1 -- the
// same u1 react_decide draws -- with a relative margin of 1e-9, far above the
// rounding of either sum.  Then react_decide would return ti.  Nothing is
// concluded for a negative rate or an injected draw out of range.
__device__ __forceinline__ bool certainly_inert(const StepArgs& A, const DRep& Rp,
                                                const DModel& M, int ti, int p, long long step,
                                                int cmax) {
  double kmax = 0.0;
#pragma unroll 1
  for (int q = 0; q < M.nrx; ++q) {
    const DRx& rx = M.rx[q];
    if (rx.reactant != ti) continue;
    const double rate = Rp.rate[q];
    if (!(rate >= 0.0)) return false;
    if (rx.kind == RDS_RX_CATALYZED) kmax += rate * (double)cmax;
    else if (rx.kind == RDS_RX_LINEAR) kmax += rate;
    else {
      if (!(rx.rate_decrease >= 0.0)) return false;
      kmax += rate;
    }
  }
  if (!(kmax > 0.0)) return true;  // every rate 0: react_decide returns ti
  double u1, u2;
  if (A.rng == RDS_RNG_INJECTED) {
    const long long o = step - A.inj_first;
    if (o < 0 || o >= A.inj_count) return false;
    u1 = A.inj_u1[o * A.n_total + p];
  } else {
    philox_uniforms(Rp, p - Rp.pbase, step, u1, u2);
  }
  return !(Rp.dt * kmax * (1.0 + 1e-9) > u1);
}

// List mode, first half of the reaction step, per slot block: every particle
// whose rates need no neighbours is decided here (pyrdsim.pyx:2103-2222); the
// others are listed for k_react_walk, as k_scatter_react lists them in cell
// mode.
__global__ __launch_bounds__(256) void k_react_list(const StepArgs A) {
  __shared__ int s_wcnt[4];
  const int4 B = A.sblocks[blockIdx.x];  // (replica, first slot, end slot)
  const DRep& Rp = A.rep[B.x];
  const DModel& M = A.models[Rp.model];
  const long long step = *A.step;  // already advanced: the reference's l+1
  const int s = B.y + (int)threadIdx.x;
  bool walker = false;
  if (s < B.z) {
    const int t = A.type[s];
    if ((M.walk_types >> t) & 1u) walker = true;
    else A.type_out[s] = M.nrx > 0 ? react_decide(A, Rp, M, t, A.pid[s], step, 0ull, 0ull) : t;
  }
  const unsigned long long m = __ballot(walker);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) s_wcnt[wv] = __popcll(m);
  __syncthreads();
  int base = 0;
  for (int w = 0; w < wv; ++w) base += s_wcnt[w];
  if (walker)
    A.wl[blockIdx.x * 256 + base + __popcll(m & ((1ull << lane) - 1ull))] =
        make_int4(s, A.scell[s], A.type[s], A.pid[s]);
  if (threadIdx.x == 0) A.wl_n[blockIdx.x] = s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
}

// Cell mode: the slots k_scatter_react listed (reactants of a catalysed or
// inhibited reaction), one wave per slot block, one lane per listed slot.
// The block's replica is wave-uniform (its tables come through scalar
// loads), and the parameters of the slot's neighbour-dependent reactions
// are taken into registers before its candidate loop.  The cell starts of
// the three stencil rows are loaded together; each row is at most two
// contiguous slot ranges, whose candidates' types are loaded WALK_BATCH at a
// time (positions only for the types a reaction counts).
#define WALK_BATCH 12
template <bool MOL>
__global__ __launch_bounds__(256, 6) void k_react_walk(const StepArgs A, int n_blocks) {
  // one wave per slot block of a replica whose model has a walk: (slot
  // block, replica, model); the block's list entries (slot, cell, type, pid)
  const int wb = (int)__builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (wb >= n_blocks) return;
  const int4 WB = A.wblocks[wb];
  const int sb = WB.x;
  const DRep& Rp = A.rep[WB.y];
  const DModel& M = A.models[WB.z];
  const int n = A.wl_n[sb];
  // the lane's first entry is loaded before the count is known (the
  // block's 256 entries are allocated; one past the count is not used)
  const int4 e0 = A.wl[sb * 256 + (int)(threadIdx.x & 63)];
  if (n == 0) return;
  const int nrx = M.nrx;
  const long long step = *A.step;
  const int lx = Rp.lx, ly = Rp.ly, cbase = Rp.cbase, sx = Rp.sx, sy = Rp.sy;
  const int px = Rp.px, py = Rp.py;
  const double L0 = Rp.L[0], L1 = Rp.L[1];
  for (int i = threadIdx.x & 63; i < n; i += 64) {
    // the slot's cell: in list mode the cell of the last rebuild (cells are
    // then at least range + skin wide, so its stencil holds every particle
    // within r0; distances take the minimum image); in cell mode the new cell
    const int4 e = i < 64 ? e0 : A.wl[sb * 256 + i];
    const int s = e.x, cell = e.y - cbase, ti = e.z, p = e.w;
    const double2 xi = rd_pos(A.pos, s, A.cpos);
    const int mi = MOL ? A.mol[s] : 0;
    // the slot's neighbour-dependent reactions: the first two in registers
    // (match mask, r0 band), any further ones through the table
    unsigned nbr_types = 0u, mask_a = 0u, mask_b = 0u;
    int qa = -1, qb = -1;
    bool more = false;
#pragma unroll 1
    for (int q = 0; q < nrx; ++q) {
      const DRx& rx = M.rx[q];
      if (rx.kind == RDS_RX_LINEAR || rx.reactant != ti) continue;
      unsigned mk = 0u;
      const int a = rx.kind == RDS_RX_CATALYZED ? rx.catalyst : rx.inh;
      const int b = rx.kind == RDS_RX_CATALYZED ? -1 : rx.inh2;
      if (a >= 0 && a < 32) mk |= 1u << a;
      if (b >= 0 && b < 32) mk |= 1u << b;
      nbr_types |= mk;
      if (qa < 0) { qa = q; mask_a = mk; }
      else if (qb < 0) { qb = q; mask_b = mk; }
      else more = true;
    }
    double lo_a = 0.0, hi_a = 0.0, r0_a = 0.0, lo_b = 0.0, hi_b = 0.0, r0_b = 0.0;
    if (qa >= 0) { lo_a = M.rx[qa].r0sq_lo; hi_a = M.rx[qa].r0sq_hi; r0_a = M.rx[qa].r0; }
    if (qb >= 0) { lo_b = M.rx[qb].r0sq_lo; hi_b = M.rx[qb].r0sq_hi; r0_b = M.rx[qb].r0; }
    unsigned long long c_lo = 0ull, c_hi = 0ull;  // packed counts (cnt_of)
    const int cy = cell / lx, cx = cell - cy * lx;
    // the stencil rows' cells cx-1 .. cx+1, each as at most two contiguous
    // slot ranges [ra0, ra1) and [rb0, rb1); a row outside a reflecting wall
    // is empty.  The cell-start loads of all three rows are issued together.
    int ra0[3], ra1[3], rb0[3], rb1[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      int y = cy + r - 1;
      bool ok = true;
      if (y < 0 || y >= ly) {
        ok = sy != 0;
        y = (y < 0) ? y + ly : y - ly;
      }
      const int rowc = cbase + (ok ? y : 0) * lx;
      int i0, i1, i2 = rowc, i3 = rowc;  // cstart indices
      if (!sx) {
        i0 = rowc + max(cx - 1, 0);
        i1 = rowc + min(cx + 1, lx - 1) + 1;
      } else if (cx == 0) {
        i0 = rowc + lx - 1;
        i1 = rowc + lx;
        i3 = rowc + 2;
      } else if (cx == lx - 1) {
        i0 = rowc + cx - 1;
        i1 = rowc + lx;
        i3 = rowc + 1;
      } else {
        i0 = rowc + cx - 1;
        i1 = rowc + cx + 2;
      }
      ra0[r] = A.cstart[i0];
      ra1[r] = ok ? A.cstart[i1] : ra0[r];
      rb0[r] = A.cstart[i2];
      rb1[r] = ok ? A.cstart[i3] : rb0[r];
    }
    // no neighbour count can exceed the particles of the stencil: when even
    // that many could not make the particle react, it keeps its type and
    // nothing more is read
    {
      int pop = 0;
#pragma unroll
      for (int r = 0; r < 3; ++r) pop += (ra1[r] - ra0[r]) + (rb1[r] - rb0[r]);
      if (certainly_inert(A, Rp, M, ti, p, step, pop)) {
        A.type_out[s] = ti;
        continue;
      }
    }
#pragma unroll 1
    for (int r = 0; r < 3; ++r) {
      const int a0 = ra0[r], b0r = rb0[r];
      const int na = ra1[r] - a0, nn = na + (rb1[r] - b0r);
#pragma unroll 1
      for (int u = 0; u < nn; u += WALK_BATCH) {
        // the batch's candidates of a counted type (other than the slot), as
        // a bit mask: the type loads of the batch are in flight together
        unsigned hit = 0u;
#pragma unroll
        for (int v = 0; v < WALK_BATCH; ++v) {
          const int w = u + v;
          const int j = w < na ? a0 + w : b0r + (w - na);
          const int tj = w < nn ? A.type[j] : -1;
          if (j != s && tj >= 0 && tj < 32 && ((nbr_types >> tj) & 1u)) hit |= 1u << v;
        }
#pragma unroll 1
        while (hit) {
          const int v = __ffs(hit) - 1;
          hit &= hit - 1u;
          const int w = u + v;
          const int j = w < na ? a0 + w : b0r + (w - na);
          const int tj = A.type[j];
          // the reference's neighbour list holds other molecules' particles
          // and the bonded partners (pyrdsim.pyx:1513, 1523)
          if (MOL && A.mol[j] == mi) {
            bool bonded = false;
            if (A.boff) {
              const int pj = A.pid[j];
              for (int bb = A.boff[p]; bb < A.boff[p + 1]; ++bb) bonded |= A.bpart[bb] == pj;
            }
            if (!bonded) continue;
          }
          const double2 xj = rd_pos(A.pos, j, A.cpos);
          double dx = xj.x - xi.x, dy = xj.y - xi.y;
          if (px) dx = min_image(dx, L0);
          if (py) dy = min_image(dy, L1);
          const double d2 = dx * dx + dy * dy;
          // d < r0 with d = sqrt(d2) as the reference computes it; away from
          // r0^2 (relative 1e-14) d2 alone decides it
          if (((mask_a >> tj) & 1u) && (d2 < lo_a || (d2 <= hi_a && sqrt(d2) < r0_a)))
            cnt_add(c_lo, c_hi, qa);
          if (((mask_b >> tj) & 1u) && (d2 < lo_b || (d2 <= hi_b && sqrt(d2) < r0_b)))
            cnt_add(c_lo, c_hi, qb);
          if (more) {
            int seen = 0;
#pragma unroll 1
            for (int q = 0; q < nrx; ++q) {
              const DRx& rx = M.rx[q];
              if (rx.kind == RDS_RX_LINEAR || rx.reactant != ti) continue;
              if (seen++ < 2) continue;  // the first two are counted above
              const bool match = (rx.kind == RDS_RX_CATALYZED) ? (tj == rx.catalyst)
                                                               : (tj == rx.inh || tj == rx.inh2);
              if (match && (d2 < rx.r0sq_lo || (d2 <= rx.r0sq_hi && sqrt(d2) < rx.r0)))
                cnt_add(c_lo, c_hi, q);
            }
          }
        }
      }
    }
    A.type_out[s] = react_decide(A, Rp, M, ti, p, step, c_lo, c_hi);
  }
}

// ---------------------------------------------------------------------------
// harmonic bonds (pyrdsim.pyx:1303-1320, 2037-2047) and angles
// (pyrdsim.pyx:1322-1407, 2052-2065), FP64, per slot
// ---------------------------------------------------------------------------

struct BondArgs {
  int n_total;
  const DRep* rep;
  const int* rep_of;
  const double2* pos;
  const int* pid;
  const int* slot_of;
  int cpos;
  // bond CSR by pid, both directions; entries with u0 = 0 are angle legs,
  // listed only so that the reaction kernel sees them as bonded partners
  const int* boff;
  const int* bpart;
  const double* br0;
  const double* bu0;
  // angle CSR by pid: (angle, role 0/1/2 = first/vertex/last) per membership
  const int* aoff;
  const int2* aent;
  const int4* apart;    // (i1, i2 vertex, i3, -)
  const double2* aprm;  // (theta0 [rad], k0 [per rad^2])
  int angle_pi;         // the straight-angle flag of the last angle defined
  double angle_thr;     // angle_pi_approximation_threshold
  DStatus* status;
  double2* fbond;
};

__device__ __forceinline__ double2 bond_vec(const BondArgs& A, const DRep& R, double2 from,
                                            int to_pid) {
  const double2 x = rd_pos(A.pos, A.slot_of[to_pid], A.cpos);
  double dx = x.x - from.x, dy = x.y - from.y;
  if (R.px) dx = min_image(dx, R.L[0]);
  if (R.py) dy = min_image(dy, R.L[1]);
  return make_double2(dx, dy);
}

__global__ void k_bonds(BondArgs A) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A.n_total) return;
  int p = A.pid[s];
  const DRep& R = A.rep[A.rep_of[p]];
  double2 xi = rd_pos(A.pos, s, A.cpos);
  double fx = 0.0, fy = 0.0;
  for (int b = A.boff[p]; b < A.boff[p + 1]; ++b) {
    if (A.bu0[b] == 0.0) continue;
    const double2 v = bond_vec(A, R, xi, A.bpart[b]);
    double d = sqrt(v.x * v.x + v.y * v.y);
    double f = A.bu0[b] * (d - A.br0[b]);
    fx += f * (v.x / d);
    fy += f * (v.y / d);
  }
  if (A.aoff) {
    for (int e = A.aoff[p]; e < A.aoff[p + 1]; ++e) {
      const int2 en = A.aent[e];
      const int4 g = A.apart[en.x];
      const double2 prm = A.aprm[en.x];
      // v1 = x[i1] - x[i2], v2 = x[i3] - x[i2] (minimum image), the same
      // arithmetic in all three memberships so that they see one angle
      const double2 xv = rd_pos(A.pos, A.slot_of[g.y], A.cpos);
      const double2 v1 = bond_vec(A, R, xv, g.x);
      const double2 v2 = bond_vec(A, R, xv, g.z);
      const double d1 = sqrt(v1.x * v1.x + v1.y * v1.y);
      const double d2 = sqrt(v2.x * v2.x + v2.y * v2.y);
      const double u1x = v1.x / d1, u1y = v1.y / d1;
      const double u2x = v2.x / d2, u2y = v2.y / d2;
      double h = u1x * u2x + u1y * u2y;
      double theta;
      if (h < -1.0) {
        // below -1.0000001 the reference raises; rounding above it is
        // taken as a straight angle
        if (h < -1.0000001 && atomicExch(&A.status->angle, 1) == 0) A.status->angle_val = h;
        h = -1.0;
        theta = M_PI;
      } else {
        theta = acos(h);
      }
      const double t = theta - prm.x;
      double prefac;
      if (A.angle_pi && !(theta < prm.x - A.angle_thr)) {
        // sin(t)/t by its series: theta0 = pi makes 1 - h^2 vanish
        prefac = -prm.y / (1.0 - t * t / 6.0 + (t * t * t * t) / 120.0);
      } else {
        prefac = prm.y * t / sqrt(1.0 - h * h);
      }
      const double c1 = prefac / d1, c3 = prefac / d2;
      const double f1x = c1 * (u2x - h * u1x), f1y = c1 * (u2y - h * u1y);
      const double f3x = c3 * (u1x - h * u2x), f3y = c3 * (u1y - h * u2y);
      if (en.y == 0) {
        fx += f1x;
        fy += f1y;
      } else if (en.y == 2) {
        fx += f3x;
        fy += f3y;
      } else {
        fx -= f1x + f3x;
        fy -= f1y + f3y;
      }
    }
  }
  A.fbond[s] = make_double2(fx, fy);
}

// ---------------------------------------------------------------------------
// binning, scan, scatter, output
// ---------------------------------------------------------------------------

// Cell and rank of every slot (a replica's slots and pids are the same
// range, so rep_of serves both).  With `cond` (list mode) only when a rebuild
// is requested; grid-stride, so that an idle launch is short.
__global__ void k_bin(int n_total, const DRep* __restrict__ rep, const int* __restrict__ rep_of,
                      const double2* __restrict__ pos,
                      int* __restrict__ ucell, int* __restrict__ urank, int* __restrict__ count,
                      const DStatus* cond, int cpos) {
  if (cond && !cond->rebuild) return;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n_total; s += gridDim.x * blockDim.x) {
    const DRep& R = rep[rep_of[s]];
    double2 x = rd_pos(pos, s, cpos);
    int c = R.cbase + cell_coord(x.y, R.cl[1], R.ly) * R.lx + cell_coord(x.x, R.cl[0], R.lx);
    ucell[s] = c;
    urank[s] = atomicAdd(&count[c], 1);
  }
}

#define SCAN_BLOCK 1024
#define SCAN_ITEMS 8

// One workgroup per replica: cell starts = pbase + exclusive prefix of the
// histogram; the histogram is cleared for the next step.  With `cond` (list
// mode) the scan runs only when a rebuild is requested; the step counter
// advances either way.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan(const DRep* __restrict__ rep,
                                                     int* __restrict__ count,
                                                     int* __restrict__ cstart,
                                                     long long* step, int advance,
                                                     DStatus* status, int cond) {
  __shared__ int s_wsum[SCAN_BLOCK / 64];
  __shared__ int s_total;
  const DRep& R = rep[blockIdx.x];
  const int t = threadIdx.x;
  const int C = (cond && !status->rebuild) ? 0 : R.lx * R.ly;
  const int cb = R.cbase;
  int running = R.pbase;
  for (int base = 0; base < C; base += SCAN_BLOCK * SCAN_ITEMS) {
    int v[SCAN_ITEMS];
    int tsum = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
      int c = base + t * SCAN_ITEMS + i;
      v[i] = (c < C) ? count[cb + c] : 0;
      tsum += v[i];
    }
    // wave inclusive scan of tsum
    int inc = tsum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      int o = __shfl_up(inc, d, 64);
      if ((t & 63) >= d) inc += o;
    }
    if ((t & 63) == 63) s_wsum[t >> 6] = inc;
    __syncthreads();
    if (t == 0) {
      int acc = 0;
      for (int k = 0; k < SCAN_BLOCK / 64; ++k) {
        int x = s_wsum[k];
        s_wsum[k] = acc;
        acc += x;
      }
      s_total = acc;
    }
    __syncthreads();
    int excl = running + s_wsum[t >> 6] + inc - tsum;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
      int c = base + t * SCAN_ITEMS + i;
      if (c < C) {
        cstart[cb + c] = excl;
        count[cb + c] = 0;
      }
      excl += v[i];
    }
    running += s_total;
    __syncthreads();
  }
  if (t == 0 && C > 0 && running != R.pbase + R.n) atomicOr(&status->integrity, 2);
  if (advance && blockIdx.x == 0 && t == 0) *step = *step + 1;
}

// Counting-sort reorder of the new state into the other buffer, one thread
// per slot.
// The new cell of a slot, from its stored new position: the same FP64 value
// and the same routine as the force kernel's binning, hence the same cell.
__device__ __forceinline__ int new_cell(const DRep& Rp, double2 raw, int cpos) {
  double2 v = raw;
  if (cpos) {
    float4 f;
    __builtin_memcpy(&f, &raw, sizeof(f));
    v = make_double2((double)f.x + (double)f.z, (double)f.y + (double)f.w);
  }
  const int cx = cell_coord_fast(v.x, Rp.cl[0], Rp.icl[0], Rp.lx);
  const int cy = cell_coord_fast(v.y, Rp.cl[1], Rp.icl[1], Rp.ly);
  return Rp.cbase + cy * Rp.lx + cx;
}

// Counting-sort reorder into the other buffer, over the per-replica slot
// blocks (sblocks: (replica, first slot, end slot)).
__global__ __launch_bounds__(256) void k_scatter(
    const int4* __restrict__ sblocks, const DRep* __restrict__ rep, int cpos, int n_total,
    const double2* __restrict__ upos, const int* __restrict__ urank,
    const int* __restrict__ cstart, const int* __restrict__ type,
    const int* __restrict__ pid, const int* __restrict__ mol, double2* __restrict__ npos,
    int* __restrict__ ntype, int* __restrict__ npid, int* __restrict__ nmol,
    int* __restrict__ slot_of, DStatus* status) {
  const int4 B = sblocks[blockIdx.x];
  const int s = B.y + (int)threadIdx.x;
  if (s >= B.z) return;
  const double2 u = upos[s];
  const int d = cstart[new_cell(rep[B.x], u, cpos)] + urank[s];
  if (d < 0 || d >= n_total) {
    atomicOr(&status->integrity, 1);
    return;
  }
  const int p = pid[s];
  npos[d] = u;
  ntype[d] = type[s];
  npid[d] = p;
  if (nmol) nmol[d] = mol[s];
  if (slot_of) slot_of[p] = d;
}

// The same reorder for models with reactions (cell mode, after the step
// counter advanced), over the per-replica slot blocks (A.sblocks: one
// replica per workgroup, so its tables are read with scalar loads): the
// pre-reaction type goes to ntype0, and the reaction step is decided here
// for every particle whose rates need no neighbours (pyrdsim.pyx:2103-2222);
// the others are listed for k_react_walk in the block's own segment of the
// worklist, compacted within the block.
__global__ __launch_bounds__(256) void k_scatter_react(
    const double2* __restrict__ upos, const int* __restrict__ urank,
    const int* __restrict__ cstart, const int* __restrict__ type,
    const int* __restrict__ pid, const int* __restrict__ mol, double2* __restrict__ npos,
    int* __restrict__ ntype, int* __restrict__ npid, int* __restrict__ nmol,
    int* __restrict__ slot_of, int* __restrict__ ntype0, const StepArgs A) {
  __shared__ int s_wcnt[4];
  const int4 B = A.sblocks[blockIdx.x];  // (replica, first slot, end slot)
  const DRep& Rp = A.rep[B.x];
  const DModel& M = A.models[Rp.model];
  const unsigned walk_types = M.walk_types;
  const int nrx = M.nrx;
  const long long step = *A.step;
  const int s = B.y + (int)threadIdx.x;
  bool walker = false;
  int d = -1, c = 0, tw = 0, pw = 0;  // a walker's slot, cell, type, pid
  if (s < B.z) {
    const double2 u = upos[s];
    c = new_cell(Rp, u, A.cpos);
    d = cstart[c] + urank[s];
    if (d < 0 || d >= A.n_total) {
      atomicOr(&A.status->integrity, 1);
      d = -1;
    }
    if (d >= 0) {
      const int p = pid[s];
      const int t = type[s];
      npos[d] = u;
      npid[d] = p;
      if (nmol) nmol[d] = mol[s];
      if (slot_of) slot_of[p] = d;
      // the pre-reaction types, read by the walk only: models that have one
      if (walk_types) ntype0[d] = t;
      tw = t;
      pw = p;
      if ((walk_types >> t) & 1u) walker = true;
      else ntype[d] = nrx > 0 ? react_decide(A, Rp, M, t, p, step, 0ull, 0ull) : t;
    }
  }
  const unsigned long long m = __ballot(walker);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) s_wcnt[wv] = __popcll(m);
  __syncthreads();
  int base = 0;
  for (int w = 0; w < wv; ++w) base += s_wcnt[w];
  if (walker)
    A.wl[blockIdx.x * 256 + base + __popcll(m & ((1ull << lane) - 1ull))] = make_int4(d, c, tw, pw);
  if (threadIdx.x == 0) A.wl_n[blockIdx.x] = s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
}

// Cell starts in chunks of at most CHUNK_CELLS cells of one replica
// (chunk = (replica, first cell within the replica, cells, index of the
// replica's first chunk)), so that large replicas spread over many
// workgroups, in one pass: each chunk publishes its particle count in a
// 64-bit status word as soon as it has summed its histogram, then adds the
// counts of the replica's earlier chunks, one thread per chunk, for its base.
// A chunk waits only on chunks of lower index, which are dispatched no later
// than itself and publish without waiting on anything; the wait is still
// bounded, and a chunk that gives up flags the status (the step then fails
// loudly) instead of hanging.  The words are
// double-buffered by buffer parity: this step's scan clears the other
// parity's for the next.  The histogram is cleared for the next step and the
// step counter advanced (advance).
#define CHUNK_CELLS 2048
#define CHUNK_THREADS 256
#define LB_AGG (1ull << 62)             // word holds the chunk's count
#define LB_VAL ((1ull << 62) - 1ull)
#define LB_SPIN (1 << 22)

__device__ __forceinline__ int block_sum_256(int v, int* s_red) {
  const int t = threadIdx.x;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((t & 63) == 0) s_red[t >> 6] = v;
  __syncthreads();
  const int tot = s_red[0] + s_red[1] + s_red[2] + s_red[3];
  __syncthreads();
  return tot;
}

__global__ __launch_bounds__(CHUNK_THREADS) void k_cellscan(
    const int4* __restrict__ chunks, const DRep* __restrict__ rep, int* __restrict__ count,
    int* __restrict__ cstart, unsigned long long* lb, unsigned long long* lb_next,
    long long* step, int advance, DStatus* status) {
  __shared__ int s_red[4];
  __shared__ int s_wsum[CHUNK_THREADS / 64];
  __shared__ __align__(16) int s_v[CHUNK_CELLS];
  const int t = threadIdx.x, b = blockIdx.x;
  const int4 ch = chunks[b];
  const DRep& R = rep[ch.x];
  const int c0 = R.cbase + ch.y;
  // the chunk's histogram through LDS, so that global loads and stores are
  // coalesced while each thread scans PER consecutive cells
  constexpr int PER = CHUNK_CELLS / CHUNK_THREADS;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int c = i * CHUNK_THREADS + t;
    s_v[c] = c < ch.z ? count[c0 + c] : 0;
  }
  __syncthreads();
  int v[PER];
#pragma unroll
  for (int i = 0; i < PER; i += 4) {
    const int4 q = *reinterpret_cast<const int4*>(&s_v[t * PER + i]);
    v[i] = q.x;
    v[i + 1] = q.y;
    v[i + 2] = q.z;
    v[i + 3] = q.w;
  }
  int tsum = 0;
#pragma unroll
  for (int i = 0; i < PER; ++i) tsum += v[i];
  // publish the chunk's count at once, then sum the replica's earlier
  // chunks' counts, one thread per chunk, each waiting (boundedly) until that
  // chunk has published
  const int tot = block_sum_256(tsum, s_red);
  if (t == 0)
    __hip_atomic_store(lb + b, LB_AGG | (unsigned long long)tot, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  int pre = 0;
  for (int k = ch.w + t; k < b; k += CHUNK_THREADS) {
    unsigned long long x = 0ull;
    for (int spins = 0; !((x = __hip_atomic_load(lb + k, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) & LB_AGG);) {
      if (++spins > LB_SPIN) {
        atomicOr(&status->integrity, 4);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    pre += (int)(x & LB_VAL);
  }
  if (t == 0) lb_next[b] = 0ull;
  const int base = R.pbase + block_sum_256(pre, s_red);
  int inc = tsum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(inc, d, 64);
    if ((t & 63) >= d) inc += o;
  }
  if ((t & 63) == 63) s_wsum[t >> 6] = inc;
  __syncthreads();
  int wbase = 0;
  for (int w = 0; w < (t >> 6); ++w) wbase += s_wsum[w];
  int excl = base + wbase + inc - tsum;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = excl;
    excl += v[i];
    v[i] = e;
  }
#pragma unroll
  for (int i = 0; i < PER; i += 4)
    *reinterpret_cast<int4*>(&s_v[t * PER + i]) = make_int4(v[i], v[i + 1], v[i + 2], v[i + 3]);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int c = i * CHUNK_THREADS + t;
    if (c < ch.z) {
      cstart[c0 + c] = s_v[c];
      count[c0 + c] = 0;
    }
  }
  // the replica's last chunk checks the total
  if (t == CHUNK_THREADS - 1 && ch.y + ch.z == R.ncell && excl != R.pbase + R.n)
    atomicOr(&status->integrity, 2);
  if (advance && b == 0 && t == 0) *step = *step + 1;
}

// List mode rebuild: counting-sort reorder of the slots into scratch columns
// (cell of each new slot recorded), then copied back in place, so that the
// buffers of the captured steps never change roles.
__global__ void k_sort_list(int n_total, const double2* __restrict__ pos,
                            const int* __restrict__ type, const int* __restrict__ pid,
                            const int* __restrict__ mol, const int* __restrict__ ucell,
                            const int* __restrict__ urank, const int* __restrict__ cstart,
                            double2* __restrict__ spos, int* __restrict__ stype,
                            int* __restrict__ spid, int* __restrict__ smol,
                            int* __restrict__ slot_of, int* __restrict__ scell, int n_cells,
                            DStatus* status) {
  if (!status->rebuild) return;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n_total; s += gridDim.x * blockDim.x) {
    const int c = ucell[s];
    if (c < 0 || c >= n_cells) {
      atomicOr(&status->integrity, 1);
      continue;
    }
    const int d = cstart[c] + urank[s];
    if (d < 0 || d >= n_total) {
      atomicOr(&status->integrity, 1);
      continue;
    }
    const int p = pid[s];
    spos[d] = pos[s];
    stype[d] = type[s];
    spid[d] = p;
    if (smol) smol[d] = mol[s];
    if (slot_of) slot_of[p] = d;
    scell[d] = c;
  }
}

__global__ void k_copy_list(int n_total, const double2* __restrict__ spos,
                            const int* __restrict__ stype, const int* __restrict__ spid,
                            const int* __restrict__ smol, double2* __restrict__ pos,
                            int* __restrict__ type, int* __restrict__ pid, int* __restrict__ mol,
                            const DStatus* status) {
  if (!status->rebuild) return;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n_total; s += gridDim.x * blockDim.x) {
    pos[s] = spos[s];
    type[s] = stype[s];
    pid[s] = spid[s];
    if (mol) mol[s] = smol[s];
  }
}

// Requests a rebuild (set_state, re-tiling, diagnostics).
__global__ void k_request(DStatus* status) { status->rebuild = 1; }

__global__ void k_gather_out(int n_total, const double2* __restrict__ pos,
                             const int* __restrict__ type, const int* __restrict__ pid,
                             double2* __restrict__ xo, int* __restrict__ to, int cpos) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_total) return;
  int p = pid[s];
  if (xo) xo[p] = rd_pos(pos, s, cpos);
  if (to) to[p] = type[s];
}

// ---------------------------------------------------------------------------
// analysis on the sorted state: pair-distance histogram (RDF), type counts,
// cluster labels.  One block per replica slot block (sblocks), one thread per
// slot; cells of the replica within a radius, every cell once.
// ---------------------------------------------------------------------------

// Calls f(j) for every slot j in the cells within `rad` of xi's cell (rings of
// ceil(rad / cell length) cells; an axis the rings cover is walked once).
template <typename F>
__device__ __forceinline__ void for_cells_within(const DRep& R, const int* __restrict__ cstart,
                                                 double2 xi, double rad, F f) {
  const int cx = cell_coord(xi.x, R.cl[0], R.lx), cy = cell_coord(xi.y, R.cl[1], R.ly);
  const int kx = (int)ceil(rad / R.cl[0]), ky = (int)ceil(rad / R.cl[1]);
  const bool fx = 2 * kx + 1 >= R.lx, fy = 2 * ky + 1 >= R.ly;
  const int x0 = fx ? 0 : cx - kx, nx = fx ? R.lx : 2 * kx + 1;
  const int y0 = fy ? 0 : cy - ky, ny = fy ? R.ly : 2 * ky + 1;
  for (int oy = 0; oy < ny; ++oy) {
    int y = y0 + oy;
    if (y < 0 || y >= R.ly) {
      if (!R.py) continue;
      y = y < 0 ? y + R.ly : y - R.ly;
    }
    for (int ox = 0; ox < nx; ++ox) {
      int x = x0 + ox;
      if (x < 0 || x >= R.lx) {
        if (!R.px) continue;
        x = x < 0 ? x + R.lx : x - R.lx;
      }
      const int c = R.cbase + y * R.lx + x;
      const int j1 = cstart[c + 1];
      for (int j = cstart[c]; j < j1; ++j) f(j);
    }
  }
}

__device__ __forceinline__ bool type_selected(int t, unsigned mask) {
  return t >= 0 && t < 32 && ((mask >> t) & 1u);
}

// squared minimum-image distance, each product and the sum rounded (no fused
// multiply-add), as a host computes dx*dx + dy*dy
__device__ __forceinline__ double dist2(const DRep& R, double2 a, double2 b) {
  double dx = b.x - a.x, dy = b.y - a.y;
  if (R.px) dx = min_image(dx, R.L[0]);
  if (R.py) dy = min_image(dy, R.L[1]);
  return __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy));
}

// unordered pairs (from the smaller slot) of selected types with r <= r_max
// into bins floor(r / dr) < nbins, per replica
__global__ void k_pair_hist(const DRep* __restrict__ rep, const int4* __restrict__ sblocks,
                            const int* __restrict__ cstart, const double2* __restrict__ pos,
                            const int* __restrict__ type, int cpos, double r2max, double dr,
                            double rad, int nbins, unsigned mask,
                            unsigned long long* __restrict__ hist) {
  extern __shared__ unsigned s_hist[];
  const int4 sb = sblocks[blockIdx.x];
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) s_hist[b] = 0u;
  __syncthreads();
  const int s = sb.y + (int)threadIdx.x;
  if (s < sb.z && type_selected(type[s], mask)) {
    const DRep& R = rep[sb.x];
    const double2 xi = rd_pos(pos, s, cpos);
    for_cells_within(R, cstart, xi, rad, [&](int j) {
      if (j <= s || !type_selected(type[j], mask)) return;
      const double d2 = dist2(R, xi, rd_pos(pos, j, cpos));
      if (d2 > r2max) return;
      const int b = (int)floor(sqrt(d2) / dr);
      if (b < nbins) atomicAdd(&s_hist[b], 1u);
    });
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += blockDim.x)
    if (s_hist[b]) atomicAdd(&hist[(size_t)sb.x * nbins + b], (unsigned long long)s_hist[b]);
}

// Three dimensions: f(j) for every slot j of the cells within rad of xi's
// cell (rings of ceil(rad / cl) cells per axis, the whole axis when the ring
// would wrap onto itself).
template <typename F>
__device__ __forceinline__ void for_cells_within3(const DRep& R, const int* __restrict__ cstart,
                                                  double4 xi, double rad, F f) {
  const int cx = cell_coord(xi.x, R.cl[0], R.lx), cy = cell_coord(xi.y, R.cl[1], R.ly);
  const int cz = cell_coord(xi.z, R.cl[2], R.lz);
  const int kx = (int)ceil(rad / R.cl[0]), ky = (int)ceil(rad / R.cl[1]);
  const int kz = (int)ceil(rad / R.cl[2]);
  const bool fx = 2 * kx + 1 >= R.lx, fy = 2 * ky + 1 >= R.ly, fz = 2 * kz + 1 >= R.lz;
  const int x0 = fx ? 0 : cx - kx, nx = fx ? R.lx : 2 * kx + 1;
  const int y0 = fy ? 0 : cy - ky, ny = fy ? R.ly : 2 * ky + 1;
  const int z0 = fz ? 0 : cz - kz, nz = fz ? R.lz : 2 * kz + 1;
  for (int oz = 0; oz < nz; ++oz) {
    int z = z0 + oz;
    if (z < 0 || z >= R.lz) {
      if (!R.pz) continue;
      z = z < 0 ? z + R.lz : z - R.lz;
    }
    for (int oy = 0; oy < ny; ++oy) {
      int y = y0 + oy;
      if (y < 0 || y >= R.ly) {
        if (!R.py) continue;
        y = y < 0 ? y + R.ly : y - R.ly;
      }
      for (int ox = 0; ox < nx; ++ox) {
        int x = x0 + ox;
        if (x < 0 || x >= R.lx) {
          if (!R.px) continue;
          x = x < 0 ? x + R.lx : x - R.lx;
        }
        const int c = R.cbase + (z * R.ly + y) * R.lx + x;
        const int j1 = cstart[c + 1];
        for (int j = cstart[c]; j < j1; ++j) f(j);
      }
    }
  }
}

// squared minimum-image distance in three dimensions, each product and sum
// rounded (no fused multiply-add), as a host computes dx*dx + dy*dy + dz*dz
__device__ __forceinline__ double dist2_3(const DRep& R, double4 a, double4 b) {
  double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
  if (R.px) dx = min_image(dx, R.L[0]);
  if (R.py) dy = min_image(dy, R.L[1]);
  if (R.pz) dz = min_image(dz, R.L[2]);
  return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

// k_pair_hist in three dimensions
__global__ void k3_pair_hist(const DRep* __restrict__ rep, const int4* __restrict__ sblocks,
                             const int* __restrict__ cstart, const double4* __restrict__ pos,
                             const int* __restrict__ type, double r2max, double dr, double rad,
                             int nbins, unsigned mask, unsigned long long* __restrict__ hist) {
  extern __shared__ unsigned s_hist[];
  const int4 sb = sblocks[blockIdx.x];
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) s_hist[b] = 0u;
  __syncthreads();
  const int s = sb.y + (int)threadIdx.x;
  if (s < sb.z && type_selected(type[s], mask)) {
    const DRep& R = rep[sb.x];
    const double4 xi = pos[s];
    for_cells_within3(R, cstart, xi, rad, [&](int j) {
      if (j <= s || !type_selected(type[j], mask)) return;
      const double d2 = dist2_3(R, xi, pos[j]);
      if (d2 > r2max) return;
      const int b = (int)floor(sqrt(d2) / dr);
      if (b < nbins) atomicAdd(&s_hist[b], 1u);
    });
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += blockDim.x)
    if (s_hist[b]) atomicAdd(&hist[(size_t)sb.x * nbins + b], (unsigned long long)s_hist[b]);
}

// Neighbours within r of each particle in three dimensions (r at most a cell
// length): other particles, or other molecules' in molecular models; out in
// caller order.
__global__ void k3_count(const DRep* __restrict__ rep, const int4* __restrict__ sblocks,
                         const int* __restrict__ cstart, const double4* __restrict__ pos,
                         const int* __restrict__ mol, const int* __restrict__ pid, double r,
                         int* __restrict__ out) {
  const int4 sb = sblocks[blockIdx.x];
  const int s = sb.y + (int)threadIdx.x;
  if (s >= sb.z) return;
  const DRep& R = rep[sb.x];
  const double4 xi = pos[s];
  const int mi = mol ? mol[s] : 0;
  const double r2 = r * r;
  int cnt = 0;
  for_cells_within3(R, cstart, xi, r, [&](int j) {
    const bool other = mol ? mol[j] != mi : j != s;
    if (other && dist2_3(R, xi, pos[j]) <= r2) ++cnt;
  });
  out[pid[s]] = cnt;
}

// k_cl_prop in three dimensions
__global__ void k3_cl_prop(const DRep* __restrict__ rep, const int4* __restrict__ sblocks,
                           const int* __restrict__ cstart, const double4* __restrict__ pos,
                           double rl2, double rad, int* label, const int* __restrict__ slot_of,
                           DStatus* status) {
  const int4 sb = sblocks[blockIdx.x];
  const int s = sb.y + (int)threadIdx.x;
  if (s >= sb.z) return;
  const int l = label[s];
  if (l < 0) return;
  const DRep& R = rep[sb.x];
  const double4 xi = pos[s];
  int m = l;
  for_cells_within3(R, cstart, xi, rad, [&](int j) {
    const int lj = label[j];
    if (lj < 0 || lj >= m || j == s) return;
    if (dist2_3(R, xi, pos[j]) <= rl2) m = lj;
  });
  if (m < l) {
    atomicMin(&label[s], m);
    atomicMin(&label[slot_of[l]], m);  // the particle this one pointed to joins too
    status->cl_changed = 1;
  }
}

__global__ void k_type_counts(const int4* __restrict__ sblocks, const int* __restrict__ type,
                              unsigned long long* __restrict__ counts) {
  __shared__ unsigned s_cnt[MAXT];
  const int4 sb = sblocks[blockIdx.x];
  if (threadIdx.x < MAXT) s_cnt[threadIdx.x] = 0u;
  __syncthreads();
  const int s = sb.y + (int)threadIdx.x;
  if (s < sb.z) {
    const int t = type[s];
    if (t >= 0 && t < MAXT) atomicAdd(&s_cnt[t], 1u);
  }
  __syncthreads();
  if (threadIdx.x < MAXT && s_cnt[threadIdx.x])
    atomicAdd(&counts[(size_t)sb.x * MAXT + threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
}

// Cluster labels by min-label propagation with hooking and pointer jumping,
// iterated from the host until no label changes.  A label is the index (pid)
// of a particle of the same cluster and only decreases, so the fixed point
// is the smallest index of each cluster.
__global__ void k_cl_init(int n_total, const int* __restrict__ type, const int* __restrict__ pid,
                          unsigned mask, int* __restrict__ label, int* __restrict__ slot_of) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_total) return;
  label[s] = type_selected(type[s], mask) ? pid[s] : -1;
  slot_of[pid[s]] = s;
}

__global__ void k_cl_prop(const DRep* __restrict__ rep, const int4* __restrict__ sblocks,
                          const int* __restrict__ cstart, const double2* __restrict__ pos,
                          int cpos, double rl2, double rad, int* label,
                          const int* __restrict__ slot_of, DStatus* status) {
  const int4 sb = sblocks[blockIdx.x];
  const int s = sb.y + (int)threadIdx.x;
  if (s >= sb.z) return;
  const int l = label[s];
  if (l < 0) return;
  const DRep& R = rep[sb.x];
  const double2 xi = rd_pos(pos, s, cpos);
  int m = l;
  for_cells_within(R, cstart, xi, rad, [&](int j) {
    const int lj = label[j];
    if (lj < 0 || lj >= m || j == s) return;
    if (dist2(R, xi, rd_pos(pos, j, cpos)) <= rl2) m = lj;
  });
  if (m < l) {
    atomicMin(&label[s], m);
    atomicMin(&label[slot_of[l]], m);  // the particle this one pointed to joins too
    status->cl_changed = 1;
  }
}

__global__ void k_cl_jump(int n_total, int* label, const int* __restrict__ slot_of) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_total) return;
  const int l = label[s];
  if (l < 0) return;
  const int l2 = label[slot_of[l]];
  if (l2 >= 0 && l2 < l) atomicMin(&label[s], l2);
}

__global__ void k_clear_int(int* p) { *p = 0; }

__global__ void k_cl_out(int n_total, const int* __restrict__ label, const int* __restrict__ pid,
                         int* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n_total) out[pid[s]] = label[s];
}

__global__ void k_cells_out(int n_total, const DRep* __restrict__ rep,
                            const int* __restrict__ rep_of, const double2* __restrict__ pos,
                            const int* __restrict__ pid, int* __restrict__ out, int cpos) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_total) return;
  int p = pid[s];
  const DRep& R = rep[rep_of[p]];
  double2 x = rd_pos(pos, s, cpos);
  out[p] = R.cbase + cell_coord(x.y, R.cl[1], R.ly) * R.lx + cell_coord(x.x, R.cl[0], R.lx);
}

// ---------------------------------------------------------------------------
// three dimensions (rds_config.n_dim = 3), as the reference's loops over
// N_dim (pyrdsim.pyx:1540, 1625-1647, 2244, 2274): positions double4
// {x, y, z, 0} sorted by cell, cells (lx, ly, lz) at least one pair range
// wide, and one thread per particle walking its 27-cell stencil straight from
// memory, in the per-replica slot blocks (one replica per workgroup).  Cell
// search only; FP64 or FP32 pair arithmetic.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int cell3(const DRep& R, double x, double y, double z) {
  return R.cbase +
         (cell_coord(z, R.cl[2], R.lz) * R.ly + cell_coord(y, R.cl[1], R.ly)) * R.lx +
         cell_coord(x, R.cl[0], R.lx);
}

// g(c) for each of the 27 cells c around cell (cx, cy, cz); a reflecting
// axis keeps to the box, a periodic one wraps (>= 3 cells)
template <typename G>
__device__ __forceinline__ void for_cells3(const DRep& R, int cx, int cy, int cz, G g) {
  for (int oz = -1; oz <= 1; ++oz) {
    int z = cz + oz;
    if (z < 0 || z >= R.lz) {
      if (!R.sz) continue;
      z = z < 0 ? z + R.lz : z - R.lz;
    }
    for (int oy = -1; oy <= 1; ++oy) {
      int y = cy + oy;
      if (y < 0 || y >= R.ly) {
        if (!R.sy) continue;
        y = y < 0 ? y + R.ly : y - R.ly;
      }
      for (int ox = -1; ox <= 1; ++ox) {
        int x = cx + ox;
        if (x < 0 || x >= R.lx) {
          if (!R.sx) continue;
          x = x < 0 ? x + R.lx : x - R.lx;
        }
        g(R.cbase + (z * R.ly + y) * R.lx + x);
      }
    }
  }
}

// f(j) for every slot j of those cells
template <typename F>
__device__ __forceinline__ void for_stencil3(const DRep& R, const int* __restrict__ cstart,
                                             int cx, int cy, int cz, F f) {
  for_cells3(R, cx, cy, cz, [&](int c) {
    const int j1 = cstart[c + 1];
    for (int j = cstart[c]; j < j1; ++j) f(j);
  });
}

// minimum-image displacement x_j - x_i on the periodic axes
__device__ __forceinline__ double3 disp3(const DRep& R, double4 a, double4 b) {
  double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
  if (R.px) dx = min_image(dx, R.L[0]);
  if (R.py) dy = min_image(dy, R.L[1]);
  if (R.pz) dz = min_image(dz, R.L[2]);
  return make_double3(dx, dy, dz);
}

// ---------------------------------------------------------------------------
// Three dimensions, tiled: up to T3Y x T3Z home rows (y, z) of W cells in x.
// The staged region is rows y0-1 .. y0+T3Y and layers z0-1 .. z0+T3Z, cells
// x0-1 .. x0+W of each; staged row q = qz * (T3Y + 2) + qy.  Periodic images
// are shifted by the box length at staging, so each staged row is at most
// three contiguous slot segments, as in two dimensions.  A lane owns one home
// particle and walks its nine row windows (three cells each) in LDS.
// ---------------------------------------------------------------------------
#define T3Y 2
#define T3Z 2
#define T3R ((T3Y + 2) * (T3Z + 2))
#define MAXW3 16
#define HMAX3 1024

struct TileSh3 {
  int cs[T3R][MAXW3 + 2];    // first slot of staged cell k (k = 0..W+1)
  int off[T3R][MAXW3 + 3];   // staged offset of cell k within its row (prefix)
  int row_base[T3R + 1];
  int seg_slot[T3R][3], seg_n[T3R][3], seg_off[T3R][3];
  double seg_sx[T3R][3];
  double row_sy[T3R], row_sz[T3R];
  int chp[T3Y * T3Z * MAXW3 + 1];  // home particles before home cell i = hr * W + k
  unsigned char hkey[HMAX3];       // home number -> (hr << 4) | k, when <= HMAX3 of them
  int total, nbatch;
};

// dynamic LDS of the tiled 3-D kernel.  FP64 pair arithmetic: (x, y) as
// double2, z, type, molecule.  FP32 pair arithmetic (REL): one float4 per
// particle, its coordinates relative to the tile's centre rounded once from
// FP64 and its (shifted) type in the fourth lane's bits, then the molecule.
struct Staging3 {
  double2* xy;
  double* z;
  int* ty;
  float4* p4;
  int* mol;
};
template <bool REL>
__device__ __forceinline__ Staging3 staging3(char* smem, int n) {
  Staging3 g;
  if (REL) {
    g.p4 = (float4*)smem;
    g.mol = (int*)(g.p4 + n);
    g.xy = nullptr;
    g.z = nullptr;
    g.ty = nullptr;
  } else {
    g.xy = (double2*)smem;
    g.z = (double*)(g.xy + n);
    g.ty = (int*)(g.z + n);
    g.mol = g.ty + n;
    g.p4 = nullptr;
  }
  return g;
}

// Tile (replica | (ny-1) << 24 | (nz-1) << 26, y0, z0, x0 | W << 16).
__device__ void tile_setup3(const DRep& Rp, const int* __restrict__ cstart, int y0, int z0,
                            int x0, int W, int ny, int nz, TileSh3& sh) {
  const int t = threadIdx.x, nthr = blockDim.x;
  const int lx = Rp.lx, ly = Rp.ly, lz = Rp.lz;
  const int ncs = T3R * (W + 2);
  for (int q0 = t; q0 < ncs; q0 += 2 * nthr) {
    int c[2], a[2], e[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int q = q0 + u * nthr;
      const int r = q / (W + 2), k = q - r * (W + 2);
      const int qy = r % (T3Y + 2), qz = r / (T3Y + 2);
      // staged rows beyond a partial tile's last home row + 1 stay empty
      bool ok = q < ncs && qy <= ny + 1 && qz <= nz + 1;
      int y = y0 + qy - 1, z = z0 + qz - 1, x = x0 - 1 + k;
      if (y < 0) { if (Rp.sy) y += ly; else ok = false; }
      else if (y >= ly) { if (Rp.sy) y -= ly; else ok = false; }
      if (z < 0) { if (Rp.sz) z += lz; else ok = false; }
      else if (z >= lz) { if (Rp.sz) z -= lz; else ok = false; }
      if (x < 0) { if (Rp.sx) x += lx; else ok = false; }
      else if (x >= lx) { if (Rp.sx) x -= lx; else ok = false; }
      c[u] = ok ? Rp.cbase + (z * ly + y) * lx + x : -1;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      a[u] = c[u] >= 0 ? cstart[c[u]] : 0;
      e[u] = c[u] >= 0 ? cstart[c[u] + 1] : 0;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int q = q0 + u * nthr;
      if (q < ncs) {
        const int r = q / (W + 2), k = q - r * (W + 2);
        sh.cs[r][k] = a[u];
        sh.off[r][k + 1] = e[u] - a[u];
      }
    }
  }
  __syncthreads();
  {
    const int wv = t >> 6, lane = t & 63;
    for (int r = wv; r < T3R; r += (nthr >> 6)) {
      const int v = (lane < W + 2) ? sh.off[r][lane + 1] : 0;
      int inc = v;
#pragma unroll
      for (int d = 1; d < 32; d <<= 1) {
        const int o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
      }
      if (lane < W + 2) sh.off[r][lane + 1] = inc;
      if (lane == 0) sh.off[r][0] = 0;
    }
  }
  __syncthreads();
  if (t < T3R) {
    const int r = t;
    const int qy = r % (T3Y + 2), qz = r / (T3Y + 2);
    const int y = y0 + qy - 1, z = z0 + qz - 1;
    double sy = 0.0, sz = 0.0;
    if (y < 0) sy = -Rp.L[1];
    else if (y >= ly) sy = Rp.L[1];
    if (z < 0) sz = -Rp.L[2];
    else if (z >= lz) sz = Rp.L[2];
    sh.row_sy[r] = sy;
    sh.row_sz[r] = sz;
    for (int g = 0; g < 3; ++g) sh.seg_n[r][g] = 0;
    // an invalid row has no particles (off[r][W + 2] == 0); its segments are
    // empty whatever they say
    if (x0 == 0 && Rp.sx) {
      sh.seg_slot[r][0] = sh.cs[r][0];
      sh.seg_n[r][0] = sh.off[r][1];
      sh.seg_off[r][0] = 0;
      sh.seg_sx[r][0] = -Rp.L[0];
    }
    const int ka = (x0 == 0) ? 1 : 0;
    const int kb = (x0 + W >= lx) ? W : W + 1;
    sh.seg_slot[r][1] = sh.cs[r][ka];
    sh.seg_n[r][1] = sh.off[r][kb + 1] - sh.off[r][ka];
    sh.seg_off[r][1] = sh.off[r][ka];
    sh.seg_sx[r][1] = 0.0;
    if (x0 + W >= lx && Rp.sx) {
      sh.seg_slot[r][2] = sh.cs[r][W + 1];
      sh.seg_n[r][2] = sh.off[r][W + 2] - sh.off[r][W + 1];
      sh.seg_off[r][2] = sh.off[r][W + 1];
      sh.seg_sx[r][2] = Rp.L[0];
    }
  }
  if (t < 64) {
    // first staged index of each staged row (prefix over the lanes)
    const int rt = t < T3R ? sh.off[t][W + 2] : 0;
    int ir = rt;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
      const int a = __shfl_up(ir, d, 64);
      if (t >= d) ir += a;
    }
    if (t <= T3R) sh.row_base[t] = ir - rt;
    if (t == T3R) sh.total = ir;
    // home cells: home row hr = (qz - 1) * ny + (qy - 1), cell k; one lane each
    const int nc = ny * nz * W;
    int hc = 0;
    if (t < nc) {
      const int hr = t / W, k = t - hr * W;
      const int r = (hr / ny + 1) * (T3Y + 2) + (hr % ny + 1);
      hc = sh.off[r][k + 2] - sh.off[r][k + 1];
    }
    int ih = hc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int a = __shfl_up(ih, d, 64);
      if (t >= d) ih += a;
    }
    if (t < nc) sh.chp[t] = ih - hc;
    if (t == nc - 1) {
      sh.chp[nc] = ih;
      sh.nbatch = (ih + (int)blockDim.x - 1) / (int)blockDim.x;
    }
  }
  __syncthreads();
  const int nc = ny * nz * W;
  if (sh.chp[nc] <= HMAX3) {
    for (int i = t; i < nc; i += nthr) {
      const int hr = i / W, k = i - hr * W;
      const unsigned char key = (unsigned char)((hr << 4) | k);
      for (int g = sh.chp[i]; g < sh.chp[i + 1]; ++g) sh.hkey[g] = key;
    }
    __syncthreads();
  }
}

// Staged indices [p0, p0 + cap) into LDS (types shifted left by TSH).
template <bool MOL, int TSH, bool REL>
__device__ void tile_stage3(const TileSh3& sh, int p0, int cap, const double4* __restrict__ pos,
                            const int* __restrict__ type, const int* __restrict__ mol,
                            const Staging3& G, double3 org) {
  const int t = threadIdx.x, nthr = blockDim.x;
  const int p1 = min(p0 + cap, sh.total);
  for (int pb = p0 + t; pb < p1; pb += STAGE_U * nthr) {
    int src[STAGE_U], rg[STAGE_U];
#pragma unroll
    for (int u = 0; u < STAGE_U; ++u) {
      const int p = min(pb + u * nthr, p1 - 1);
      // row: the last with row_base <= p (binary search over the T3R rows)
      int r = 0;
#pragma unroll
      for (int step = T3R / 2; step >= 1; step >>= 1)
        if (sh.row_base[r + step] <= p) r += step;
      const int loc = p - sh.row_base[r];
      int g = 0;
      if (sh.seg_n[r][1] > 0 && loc >= sh.seg_off[r][1]) g = 1;
      if (sh.seg_n[r][2] > 0 && loc >= sh.seg_off[r][2]) g = 2;
      src[u] = sh.seg_slot[r][g] + (loc - sh.seg_off[r][g]);
      rg[u] = r * 3 + g;
    }
    double4 q[STAGE_U];
    int ty[STAGE_U], mo[STAGE_U];
#pragma unroll
    for (int u = 0; u < STAGE_U; ++u) {
      q[u] = pos[src[u]];
      ty[u] = type[src[u]];
      mo[u] = MOL ? mol[src[u]] : 0;
    }
#pragma unroll
    for (int u = 0; u < STAGE_U; ++u) {
      const int p = pb + u * nthr;
      if (p >= p1) break;
      const int r = rg[u] / 3, g = rg[u] - 3 * r;
      const double x = q[u].x + sh.seg_sx[r][g], y = q[u].y + sh.row_sy[r];
      const double z = q[u].z + sh.row_sz[r];
      if (REL) {
        G.p4[p - p0] = make_float4((float)(x - org.x), (float)(y - org.y), (float)(z - org.z),
                                   __int_as_float(ty[u] << TSH));
      } else {
        G.xy[p - p0] = make_double2(x, y);
        G.z[p - p0] = z;
        G.ty[p - p0] = ty[u] << TSH;
      }
      if (MOL) G.mol[p - p0] = mo[u];
    }
  }
}

// The integrator epilogue of a three-dimensional force kernel: walls, bond
// forces, then either the force (MODE_FORCE) or the Euler-Maruyama update
// (pyrdsim.pyx:2244-2250), the boundary conditions (pyrdsim.pyx:2269-2292),
// the new cell and the rank in it.
template <typename Real>
__device__ __forceinline__ void k3_finish(const StepArgs& A, const DRep& Rp, const DModel& M,
                                          int mode, int s, int p, int ti, double4 xi,
                                          double fx, double fy, double fz) {
  double F[3] = {fx, fy, fz};
  const double X[3] = {xi.x, xi.y, xi.z};
  for (int q = 0; q < M.nwalls; ++q) {
    const DWall& wl = M.walls[q];
    if (wl.type == ti) F[wl.dir] += wall_force<Real>(wl, X[wl.dir]);
  }
  if (A.fbond4) {
    const double4 fb = A.fbond4[s];
    F[0] += fb.x;
    F[1] += fb.y;
    F[2] += fb.z;
  }
  if (mode == MODE_FORCE) {
    double* out = reinterpret_cast<double*>(A.fout) + 3 * (size_t)p;
    out[0] = F[0];
    out[1] = F[1];
    out[2] = F[2];
    return;
  }
  const long long step_next = *A.step + 1;
  double n[3];
  if (A.rng == RDS_RNG_INJECTED) {
    const long long o = step_next - A.inj_first;
    if (o >= 0 && o < A.inj_count) {
      const double* src = A.inj_n + (o * A.n_total + p) * 3;
      n[0] = src[0];
      n[1] = src[1];
      n[2] = src[2];
    } else {
      n[0] = n[1] = n[2] = 0.0;
      A.status->stream = 1;
    }
  } else {
    float a0, a1, a2, a3;
    philox_normals(Rp, p - Rp.pbase, step_next, a0, a1);
    philox_normals(Rp, p - Rp.pbase, step_next, a2, a3, 2u);
    n[0] = a0;
    n[1] = a1;
    n[2] = a2;
  }
  const double fp = Rp.fpre[ti];
  const double np = A.include_rf * Rp.npre[ti];
  const int per[3] = {Rp.px, Rp.py, Rp.pz};
  double v[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double u = X[k];
    {
#pragma clang fp contract(off)
      u = u + fp * F[k];
      u = u + np * n[k];
    }
    const double L = Rp.L[k];
    if (!per[k]) {
      if (u > L) u -= 2.0 * (u - L);
      if (u < 0.0) u -= 2.0 * u;
    } else {
      double m;
      if (u > -L && u < L) m = u;
      else if (u >= L && u < 2.0 * L) m = u - L;
      else m = fmod(u, L);
      if (m != 0.0 && ((m < 0.0) != (L < 0.0))) m += L;
      u = m;
    }
    if (u < 0.0 || u > L || u != u)
      record_bounds(A.status, ((unsigned long long)step_next << 32) |
                                  (unsigned long long)(unsigned)(p * 3 + k), u);
    v[k] = u;
  }
  const int c = cell3(Rp, v[0], v[1], v[2]);
  A.upos4[s] = make_double4(v[0], v[1], v[2], 0.0);
  A.ucell[s] = c;
  A.urank[s] = atomicAdd(&A.count[c], 1);
}

// Three-dimensional force kernel: pair forces (every term of the type pair,
// file order) over the 27-cell stencil, walked as nine staged rows of three
// x-cells in LDS (x_j + shift - x_i needs no minimum image: a periodic axis
// with a stencil has at least three cells); walls and bonds; then, in
// MODE_STEP, the update, the boundary conditions and the binning (k3_finish).
// Differences in FP64, pair arithmetic and sums in Real (FP32 or FP64).  KM
// as in k_force: 0 = the general term table; 1, 2 = one LJ term per type pair
// of at most two types (2: exponent n = 6).
template <typename Real, bool MOL, int KM>
__global__ __launch_bounds__(512) void k3_tile(const StepArgs A, int mode, int pw) {
  constexpr bool FAST = KM != 0;
  constexpr int TSH = (FAST && sizeof(Real) == 4) ? 4 : 0;
  constexpr bool LJ6 = TSH && KM == 2;
  // FP32 pair arithmetic stages coordinates relative to the tile's centre
  constexpr bool REL = sizeof(Real) == 4;
  extern __shared__ __align__(16) char smem[];
  const int CAPT = A.cap;
  const Staging3 G = staging3<REL>(smem, CAPT + 1);
  __shared__ TileSh3 sh;
  __shared__ int s_pfirst[FAST ? 1 : MAXT * MAXT], s_pcount[FAST ? 1 : MAXT * MAXT];
  __shared__ DTerm s_term[FAST ? 1 : MAXTERMS];
  __shared__ Real s_fc[FAST ? 4 : 1][3];
  __shared__ __align__(16) float4 s_fc4[TSH ? 4 : 1];

  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int tile = xcd_tile(blockIdx.x, A.n_tiles);
  const int4 T = A.tiles[tile];
  const DRep& Rp = A.rep[T.x & 0xffffff];
  const int ny = ((T.x >> 24) & 3) + 1, nz = ((T.x >> 26) & 3) + 1;
  const int y0 = T.y, z0 = T.z, x0 = T.w & 0xffff, W = T.w >> 16;
  const DModel& M = A.models[Rp.model];
  const int nt = M.nt;

  if (!FAST) {
    for (int q = t; q < nt * nt; q += blockDim.x) {
      s_pfirst[q] = M.pfirst[q];
      s_pcount[q] = M.pcount[q];
    }
    for (int q = t; q < M.nterms; q += blockDim.x) s_term[q] = M.terms[q];
  }
  if (FAST && t < 4) {
    const int a = t >> 1, bb = t & 1, q = a * nt + bb;
    Real v0 = 0, v1 = 0, v2 = 0;
    double A6 = 0.0, B6 = 0.0;
    if (a < nt && bb < nt && M.pcount[q] > 0) {
      const DTerm& tm = M.terms[M.pfirst[q]];
      v0 = TP<Real>::c_lj(tm);
      v1 = TP<Real>::r0sq(tm);
      v2 = TP<Real>::rc2(tm);
      const double r6 = tm.r0sq * tm.r0sq * tm.r0sq;
      B6 = tm.c_lj * r6;
      A6 = B6 * r6;
    }
    s_fc[t][0] = v0;
    s_fc[t][1] = v1;
    s_fc[t][2] = v2;
    if (TSH)
      s_fc4[t] = LJ6 ? make_float4((float)A6, (float)B6, (float)v2, 0.f)
                     : make_float4((float)v0, (float)v1, (float)v2, 0.f);
  }
  tile_setup3(Rp, A.cstart, y0, z0, x0, W, ny, nz, sh);
  const int total = sh.total;
  const bool resident = total <= CAPT;
  // the centre of the tile's home cells: staged coordinates relative to it
  // lie within about (W / 2 + 1) cells of it
  const double3 org = make_double3((x0 + 0.5 * W) * Rp.cl[0], (y0 + 0.5 * ny) * Rp.cl[1],
                                   (z0 + 0.5 * nz) * Rp.cl[2]);
  auto sentinel = [&](int e) {
    if (REL) {
      G.p4[e] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      G.xy[e] = make_double2(0.0, 0.0);
      G.z[e] = 0.0;
      G.ty[e] = 0;
    }
    if (MOL) G.mol[e] = -1;
  };
  if (resident) {
    tile_stage3<MOL, TSH, REL>(sh, 0, CAPT, A.pos4, A.type, A.mol, G, org);
    if (t == 0) sentinel(total);
    __syncthreads();
  }
  const int nc = ny * nz * W;
  const int nh = sh.chp[nc];
  for (int b = 0; b < sh.nbatch; ++b) {
    // home particle g: its home cell from the table (or the largest i with
    // chp[i] <= g), the slot from the cell's first slot
    const int g = b * (int)blockDim.x + t;
    const bool valid = g < nh;
    int hr = 0, k = 0;
    if (valid) {
      if (nh <= HMAX3) {
        const int key = sh.hkey[g];
        hr = key >> 4;
        k = key & 15;
      } else {
        int lo = 0, hi = nc;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (sh.chp[mid] <= g) lo = mid;
          else hi = mid;
        }
        hr = lo / W;
        k = lo - hr * W;
      }
    }
    const int rh = (hr / ny + 1) * (T3Y + 2) + (hr % ny + 1);  // staged row of the home row
    const int s = valid ? sh.cs[rh][k + 1] + (g - sh.chp[hr * W + k]) : 0;
    double4 xi = make_double4(0.0, 0.0, 0.0, 0.0);
    int ti = 0, p = 0, mi = 0;
    if (valid) {
      xi = A.pos4[s];
      ti = A.type[s];
      p = A.pid[s];
      if (MOL) mi = A.mol[s];
    }
    Real pa0 = 0, pb0 = 0, pc0 = 0, pa1 = 0, pb1 = 0, pc1 = 0;
    if (FAST && !TSH) {
      const int a = ti == 1 ? 2 : 0;
      pa0 = s_fc[a][0];
      pb0 = s_fc[a][1];
      pc0 = s_fc[a][2];
      pa1 = s_fc[a + 1][0];
      pb1 = s_fc[a + 1][1];
      pc1 = s_fc[a + 1][2];
    }
    const char* fc_row = reinterpret_cast<const char*>(s_fc4) + ti * 32;
    const int tbase = ti * nt;
    // the home particle relative to the tile's centre, rounded as its staged
    // copy is (so that its own difference is exactly zero)
    const float3 xr = make_float3((float)(xi.x - org.x), (float)(xi.y - org.y),
                                  (float)(xi.z - org.z));
    Real fx = 0, fy = 0, fz = 0;
    // one candidate at difference (dx, dy, dz), staged type tj (shifted), counted when `live`
    auto pairv = [&](Real dx, Real dy, Real dz, int tj, int mj, bool live) {
      const Real r2 = dx * dx + dy * dy + dz * dz;
      if (FAST) {
        Real pa, pb, pc;
        if (TSH) {
          const float4 kc = *reinterpret_cast<const float4*>(fc_row + tj);
          pa = (Real)kc.x;
          pb = (Real)kc.y;
          pc = (Real)kc.z;
        } else {
          const bool m1 = tj == 1;
          pa = m1 ? pa1 : pa0;
          pb = m1 ? pb1 : pb0;
          pc = m1 ? pc1 : pc0;
        }
        Real c;
        if constexpr (LJ6) c = lj_fast6(r2, pa, pb);
        else c = lj_fast<Real, KM>(r2, pa, pb, pw);
        // r2 > 0 excludes the particle itself (and coincident ones)
        const bool take = MOL ? (mj != mi && r2 > Real(0) && r2 <= pc) : in_range(r2, pc);
        c = (live && take) ? c : Real(0);
        fx += c * dx;
        fy += c * dy;
        fz += c * dz;
      } else {
        const bool other = MOL ? (mj != mi && r2 > Real(0)) : (r2 > Real(0));
        const int pq = tbase + tj;
        const int q0 = s_pfirst[pq], qn = live ? s_pcount[pq] : 0;
        for (int q = q0; q < q0 + qn; ++q) {
          const DTerm& tm = s_term[q];
          Real c = pair_coef<Real>(tm, r2);
          c = (other && r2 <= TP<Real>::rc2(tm)) ? c : Real(0);
          fx += c * dx;
          fy += c * dy;
          fz += c * dz;
        }
      }
    };
    for (int p0 = 0; p0 < total; p0 += CAPT) {
      if (!resident) {
        __syncthreads();
        tile_stage3<MOL, TSH, REL>(sh, p0, CAPT, A.pos4, A.type, A.mol, G, org);
        if (t == 0) sentinel(min(total - p0, CAPT));
        __syncthreads();
      }
      if (!valid) continue;
      // the nine row windows (cells k .. k + 2 of the staged rows around the
      // home row, clipped to this LDS round), two consecutive staged
      // particles per trip; past a window's end the second candidate is the
      // next staged particle or the sentinel, finite either way, and masked
#pragma unroll 1
      for (int d = 0; d < 9; ++d) {
        const int rr = rh + (d / 3 - 1) * (T3Y + 2) + (d % 3 - 1);
        const int base = sh.row_base[rr];
        const int lo = max(base + sh.off[rr][k], p0) - p0;
        const int hi = min(base + sh.off[rr][k + 3], p0 + CAPT) - p0;
        const int* mq = G.mol + lo;
        if constexpr (REL) {
          const float4* pq = G.p4 + lo;
          const float4* const pe = G.p4 + hi;
          const float4* const pl = pe - 1;
          for (; pq < pe; pq += 2, mq += MOL ? 2 : 0) {
            const float4 q0 = pq[0], q1 = pq[1];
            const int m0 = MOL ? mq[0] : 0, m1 = MOL ? mq[1] : 0;
            pairv(q0.x - xr.x, q0.y - xr.y, q0.z - xr.z, __float_as_int(q0.w), m0, true);
            pairv(q1.x - xr.x, q1.y - xr.y, q1.z - xr.z, __float_as_int(q1.w), m1, pq < pl);
          }
        } else {
          const double2* pq = G.xy + lo;
          const double2* const pe = G.xy + hi;
          const double2* const pl = pe - 1;
          const double* zq = G.z + lo;
          const int* tq = G.ty + lo;
          for (; pq < pe; pq += 2, zq += 2, tq += 2, mq += MOL ? 2 : 0) {
            const double2 q0 = pq[0], q1 = pq[1];
            const double z0v = zq[0], z1v = zq[1];
            const int t0 = tq[0], t1 = tq[1];
            const int m0 = MOL ? mq[0] : 0, m1 = MOL ? mq[1] : 0;
            pairv((Real)(q0.x - xi.x), (Real)(q0.y - xi.y), (Real)(z0v - xi.z), t0, m0, true);
            pairv((Real)(q1.x - xi.x), (Real)(q1.y - xi.y), (Real)(z1v - xi.z), t1, m1, pq < pl);
          }
        }
      }
    }
    if (!valid) continue;
    k3_finish<Real>(A, Rp, M, mode, s, p, ti, xi, (double)fx, (double)fy, (double)fz);
  }
}

// Harmonic bonds and angles in three dimensions (pyrdsim.pyx:1303-1407,
// 2037-2065): the arithmetic of k_bonds with a third component.
struct BondArgs3 {
  int n_total;
  const DRep* rep;
  const int* rep_of;
  const double4* pos;
  const int* pid;
  const int* slot_of;
  const int* boff;
  const int* bpart;
  const double* br0;
  const double* bu0;
  const int* aoff;
  const int2* aent;
  const int4* apart;
  const double2* aprm;
  int angle_pi;
  double angle_thr;
  DStatus* status;
  double4* fbond;
};

__global__ void k3_bonds(BondArgs3 A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A.n_total) return;
  const int p = A.pid[s];
  const DRep& R = A.rep[A.rep_of[p]];
  const double4 xi = A.pos[s];
  double f[3] = {0.0, 0.0, 0.0};
  for (int b = A.boff[p]; b < A.boff[p + 1]; ++b) {
    if (A.bu0[b] == 0.0) continue;
    const double3 v = disp3(R, xi, A.pos[A.slot_of[A.bpart[b]]]);
    const double d = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    const double g = A.bu0[b] * (d - A.br0[b]);
    f[0] += g * (v.x / d);
    f[1] += g * (v.y / d);
    f[2] += g * (v.z / d);
  }
  if (A.aoff) {
    for (int e = A.aoff[p]; e < A.aoff[p + 1]; ++e) {
      const int2 en = A.aent[e];
      const int4 gq = A.apart[en.x];
      const double2 prm = A.aprm[en.x];
      const double4 xv = A.pos[A.slot_of[gq.y]];
      const double3 v1 = disp3(R, xv, A.pos[A.slot_of[gq.x]]);
      const double3 v2 = disp3(R, xv, A.pos[A.slot_of[gq.z]]);
      const double d1 = sqrt(v1.x * v1.x + v1.y * v1.y + v1.z * v1.z);
      const double d2 = sqrt(v2.x * v2.x + v2.y * v2.y + v2.z * v2.z);
      const double u1[3] = {v1.x / d1, v1.y / d1, v1.z / d1};
      const double u2[3] = {v2.x / d2, v2.y / d2, v2.z / d2};
      double h = u1[0] * u2[0] + u1[1] * u2[1] + u1[2] * u2[2];
      double theta;
      if (h < -1.0) {
        if (h < -1.0000001 && atomicExch(&A.status->angle, 1) == 0) A.status->angle_val = h;
        h = -1.0;
        theta = M_PI;
      } else {
        theta = acos(h);
      }
      const double t = theta - prm.x;
      double prefac;
      if (A.angle_pi && !(theta < prm.x - A.angle_thr)) prefac = -prm.y / (1.0 - t * t / 6.0 + (t * t * t * t) / 120.0);
      else prefac = prm.y * t / sqrt(1.0 - h * h);
      const double c1 = prefac / d1, c3 = prefac / d2;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double f1 = c1 * (u2[k] - h * u1[k]), f3 = c3 * (u1[k] - h * u2[k]);
        f[k] += en.y == 0 ? f1 : (en.y == 2 ? f3 : -(f1 + f3));
      }
    }
  }
  A.fbond[s] = make_double4(f[0], f[1], f[2], 0.0);
}

// Reorder of the new state into the other buffer, one thread per slot (the
// types go to the pre-reaction column when reactions follow).
__global__ void k3_scatter(int n_total, const double4* __restrict__ upos,
                           const int* __restrict__ ucell, const int* __restrict__ urank,
                           const int* __restrict__ cstart, const int* __restrict__ type,
                           const int* __restrict__ pid, const int* __restrict__ mol,
                           double4* __restrict__ npos, int* __restrict__ ntype,
                           int* __restrict__ npid, int* __restrict__ nmol,
                           int* __restrict__ slot_of, int n_cells, DStatus* status) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_total) return;
  const int c = ucell[s];
  if (c < 0 || c >= n_cells) {
    atomicOr(&status->integrity, 1);
    return;
  }
  const int d = cstart[c] + urank[s];
  if (d < 0 || d >= n_total) {
    atomicOr(&status->integrity, 1);
    return;
  }
  const int p = pid[s];
  npos[d] = upos[s];
  ntype[d] = type[s];
  npid[d] = p;
  if (nmol) nmol[d] = mol[s];
  if (slot_of) slot_of[p] = d;
}

// The reaction step on the new positions with the old types
// (pyrdsim.pyx:2070-2222), one thread per slot in the per-replica slot
// blocks: neighbour counts over the 27-cell stencil for reactants of a
// catalysed or inhibited reaction, then the decision.
template <bool MOL>
__global__ __launch_bounds__(256) void k3_react(const StepArgs A) {
  const int4 B = A.sblocks[blockIdx.x];
  const DRep& Rp = A.rep[B.x];
  const DModel& M = A.models[Rp.model];
  const int s = B.y + (int)threadIdx.x;
  if (s >= B.z) return;
  const int nrx = M.nrx;
  const int ti = A.type[s];
  if (nrx == 0) {
    A.type_out[s] = ti;
    return;
  }
  const int p = A.pid[s];
  unsigned long long c_lo = 0ull, c_hi = 0ull;
  if ((M.walk_types >> ti) & 1u) {
    unsigned nbr_types = 0u;
    for (int q = 0; q < nrx; ++q) {
      const DRx& rx = M.rx[q];
      if (rx.kind == RDS_RX_LINEAR || rx.reactant != ti) continue;
      const int a = rx.kind == RDS_RX_CATALYZED ? rx.catalyst : rx.inh;
      const int b = rx.kind == RDS_RX_CATALYZED ? -1 : rx.inh2;
      if (a >= 0 && a < 32) nbr_types |= 1u << a;
      if (b >= 0 && b < 32) nbr_types |= 1u << b;
    }
    const double4 xi = A.pos4[s];
    const int mi = MOL ? A.mol[s] : 0;
    const int cx = cell_coord(xi.x, Rp.cl[0], Rp.lx), cy = cell_coord(xi.y, Rp.cl[1], Rp.ly);
    const int cz = cell_coord(xi.z, Rp.cl[2], Rp.lz);
    // no neighbour count can exceed the particles of the 27 cells: when even
    // that many could not make the particle react, it keeps its type
    int pop = 0;
    for_cells3(Rp, cx, cy, cz, [&](int c) { pop += A.cstart[c + 1] - A.cstart[c]; });
    if (certainly_inert(A, Rp, M, ti, p, *A.step, pop)) {
      A.type_out[s] = ti;
      return;
    }
    for_stencil3(Rp, A.cstart, cx, cy, cz, [&](int j) {
      if (j == s) return;
      const int tj = A.type[j];
      if (tj < 0 || tj >= 32 || !((nbr_types >> tj) & 1u)) return;
      // the reference's neighbour list holds other molecules' particles and
      // the bonded partners (pyrdsim.pyx:1513, 1523)
      if (MOL && A.mol[j] == mi) {
        bool bonded = false;
        if (A.boff) {
          const int pj = A.pid[j];
          for (int b = A.boff[p]; b < A.boff[p + 1]; ++b) bonded |= A.bpart[b] == pj;
        }
        if (!bonded) return;
      }
      const double3 d = disp3(Rp, xi, A.pos4[j]);
      const double d2 = d.x * d.x + d.y * d.y + d.z * d.z;
#pragma unroll 1
      for (int q = 0; q < nrx; ++q) {
        const DRx& rx = M.rx[q];
        if (rx.kind == RDS_RX_LINEAR || rx.reactant != ti) continue;
        const bool match = (rx.kind == RDS_RX_CATALYZED) ? (tj == rx.catalyst)
                                                         : (tj == rx.inh || tj == rx.inh2);
        if (match && (d2 < rx.r0sq_lo || (d2 <= rx.r0sq_hi && sqrt(d2) < rx.r0)))
          cnt_add(c_lo, c_hi, q);
      }
    });
  }
  A.type_out[s] = react_decide(A, Rp, M, ti, p, *A.step, c_lo, c_hi);
}

__global__ void k3_load(int n_total, const double* __restrict__ x, const int* __restrict__ t,
                        const int* __restrict__ m, double4* __restrict__ pos,
                        int* __restrict__ type, int* __restrict__ pid, int* __restrict__ mol) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_total) return;
  pos[s] = make_double4(x[3 * (size_t)s], x[3 * (size_t)s + 1], x[3 * (size_t)s + 2], 0.0);
  type[s] = t[s];
  pid[s] = s;
  if (mol) mol[s] = m[s];
}

__global__ void k3_bin(int n_total, const DRep* __restrict__ rep, const int* __restrict__ rep_of,
                       const double4* __restrict__ pos, int* __restrict__ ucell,
                       int* __restrict__ urank, int* __restrict__ count) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_total) return;
  const DRep& R = rep[rep_of[s]];
  const double4 x = pos[s];
  const int c = cell3(R, x.x, x.y, x.z);
  ucell[s] = c;
  urank[s] = atomicAdd(&count[c], 1);
}

__global__ void k3_gather_out(int n_total, const double4* __restrict__ pos,
                              const int* __restrict__ type, const int* __restrict__ pid,
                              double* __restrict__ xo, int* __restrict__ to) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_total) return;
  const int p = pid[s];
  if (xo) {
    const double4 x = pos[s];
    xo[3 * (size_t)p] = x.x;
    xo[3 * (size_t)p + 1] = x.y;
    xo[3 * (size_t)p + 2] = x.z;
  }
  if (to) to[p] = type[s];
}

__global__ void k3_cells_out(int n_total, const DRep* __restrict__ rep,
                             const int* __restrict__ rep_of, const double4* __restrict__ pos,
                             const int* __restrict__ pid, int* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_total) return;
  const int p = pid[s];
  const DRep& R = rep[rep_of[p]];
  const double4 x = pos[s];
  out[p] = cell3(R, x.x, x.y, x.z);
}

// Fills the LDS of every workgroup slot with a word pattern (test hook: NaN
// patterns make later reads of never-written LDS visible).
__global__ __launch_bounds__(256) void k_fill_lds(unsigned pattern, int words) {
  extern __shared__ unsigned lds_words[];
  for (int q = threadIdx.x; q < words; q += blockDim.x) lds_words[q] = pattern;
  __syncthreads();
  if (threadIdx.x == 0 && lds_words[0] != pattern) lds_words[1] = 0;  // keep the stores
}

__global__ void k_load(int n_total, const double2* __restrict__ x, const int* __restrict__ t,
                       const int* __restrict__ m, double2* __restrict__ pos,
                       int* __restrict__ type, int* __restrict__ pid, int* __restrict__ mol,
                       int cpos) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_total) return;
  if (cpos) reinterpret_cast<float4*>(pos)[s] = df_split(x[s].x, x[s].y);
  else pos[s] = x[s];
  type[s] = t[s];
  pid[s] = s;
  if (mol) mol[s] = m[s];
}

}  // namespace

// Label array of rds_cluster_labels_static, one copy per GPU, held until the
// process ends.  Host code never passes `scratch` itself to a kernel: the name
// is a host-side stand-in, and each device's address comes from
// hipGetSymbolAddress (scratch_ptr below).
static __device__ __align__(256) int scratch[RDS_STATIC_LABELS_MAX];

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct Buf {
  double2* pos = nullptr;
  double4* pos4 = nullptr;  // three dimensions
  int* type = nullptr;
  int* type0 = nullptr;  // pre-reaction types written by the scatter
  int* pid = nullptr;
  int* mol = nullptr;
};

struct rds_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  int precision = RDS_MIXED, rng = RDS_RNG_PHILOX, prob_mode = RDS_PROB_LINEAR;
  int cpos = 0;  // compensated FP32 positions
  int ndim = 2;
  double4* upos4 = nullptr;   // three dimensions: next positions, bond forces
  double4* fbond4 = nullptr;
  int has_mol = 0;
  int graph_steps = 0;
  double include_rf = 1.0;
  int n_rep = 0, n_models = 0;
  int64_t n_total = 0, n_cells = 0, n_tiles = 0;
  bool any_rx = false, any_bonds = false;
  std::vector<DRep> h_rep;
  std::vector<DModel> h_models;
  std::vector<int4> h_tiles;
  DRep* d_rep = nullptr;
  DModel* d_models = nullptr;
  int4* d_tiles = nullptr;
  int* d_rep_of = nullptr;  // replica of each pid
  int4* d_sblocks = nullptr;  // per-replica slot blocks of the reaction kernel
  int64_t n_sblocks = 0;
  Buf buf[2];
  int parity = 0;
  double2* upos = nullptr;
  int* ucell = nullptr;
  int* urank = nullptr;
  int* count = nullptr;
  int* cstart = nullptr;
  int4* d_chunks = nullptr;  // cell-start chunks (replica, first cell, cells, first chunk)
  unsigned long long* d_lb = nullptr;  // [2][n_chunks] look-back words of the scan
  int64_t n_chunks = 0;
  int4* wl = nullptr;        // reaction worklist, 256 (slot, cell, type, pid) per slot block
  int* wl_n = nullptr;
  int4* d_wblocks = nullptr;  // (slot block, replica, model) of replicas whose model has a walk
  int64_t n_wblocks = 0;
  long long* d_step = nullptr;
  DStatus* d_status = nullptr;
  // bonds
  int* boff = nullptr;
  int* bpart = nullptr;
  double* br0 = nullptr;
  double* bu0 = nullptr;
  double2* fbond = nullptr;
  int* slot_of = nullptr;
  int* clabel = nullptr;  // cluster labels (allocated on first use)
  int* static_labels = nullptr;  // this device's copy of `scratch`
  int* h_changed = nullptr;      // pinned host word: a label changed in the last pass
  // angles
  int* aoff = nullptr;
  int2* aent = nullptr;
  int4* apart = nullptr;
  double2* aprm = nullptr;
  int angle_pi = 0;
  double angle_thr = 0.1;
  // injected streams
  const double* inj_n = nullptr;
  const double* inj_u1 = nullptr;
  const double* inj_u2 = nullptr;
  long long inj_first = 0, inj_count = 0;
  // state
  bool loaded = false;
  long long step0 = 0;
  long long steps_done = 0;
  // graphs of graph_len steps, one per starting buffer parity
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  int graph_len = 0;
  // kernel variant
  int rows = 1;         // cell rows per tile (1, 2, 4 or 8)
  bool fast = false;    // all models: <= 2 types, one LJ term per pair, same even n
  int pw = 0;           // n / 2 of that LJ exponent
  int cap = 384;        // staged particles per LDS round
  int target = 56;      // particles per tile row strip
  size_t smem_bytes = 0;
  // list mode (skin > 0)
  bool list = false;
  double skin = 0.0;
  int K = 64;                    // list entries per particle (multiple of 4)
  uint2* nl4 = nullptr;          // [K/4 + 1][N] chunks (one spare row)
  unsigned short* ncnt = nullptr;
  double2* xref = nullptr;
  int* scell = nullptr;
  double2* spos = nullptr;       // rebuild scratch columns
  int* stype = nullptr;
  int* spid = nullptr;
  int* smol = nullptr;
};

namespace {

// Device address of `scratch` on device `dev`, resolved once per device with
// that device current.
std::mutex g_scratch_mu;
int* g_scratch_dev[64] = {};

int scratch_ptr(int dev, int** out) {
  if (dev < 0 || dev >= 64) return fail(RDS_E_ARG, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != dev) HIP_TRY(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIP_TRY(hipSetDevice(cur));
    if (e != hipSuccess)
      return fail(RDS_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<int*>(p);
  }
  *out = g_scratch_dev[dev];
  return RDS_OK;
}

StepArgs make_args(rds_ctx* c, const Buf& cur) {
  StepArgs a;
  memset(&a, 0, sizeof a);
  a.rep = c->d_rep;
  a.models = c->d_models;
  a.tiles = c->d_tiles;
  a.n_tiles = (int)c->n_tiles;
  a.n_total = (int)c->n_total;
  a.cstart = c->cstart;
  a.pos = cur.pos;
  a.type = cur.type;
  a.pid = cur.pid;
  a.mol = cur.mol;
  a.upos = c->upos;
  a.ucell = c->ucell;
  a.urank = c->urank;
  a.count = c->count;
  a.fbond = c->any_bonds ? c->fbond : nullptr;
  a.step = c->d_step;
  a.inj_n = c->inj_n;
  a.inj_u1 = c->inj_u1;
  a.inj_u2 = c->inj_u2;
  a.inj_first = c->inj_first;
  a.inj_count = c->inj_count;
  a.rng = c->rng;
  a.prob_mode = c->prob_mode;
  a.cap = c->cap;
  a.rep_of = c->d_rep_of;
  a.sblocks = c->d_sblocks;
  a.include_rf = c->include_rf;
  a.status = c->d_status;
  a.nl4 = c->nl4;
  a.ncnt = c->ncnt;
  a.K = c->K;
  a.xref = c->xref;
  a.scell = c->list ? c->scell : nullptr;
  a.hskin2 = 0.25 * c->skin * c->skin;
  a.cpos = c->cpos;
  a.boff = c->any_bonds ? c->boff : nullptr;
  a.bpart = c->any_bonds ? c->bpart : nullptr;
  a.wl = c->wl;
  a.wl_n = c->wl_n;
  a.wblocks = c->d_wblocks;
  a.pos4 = cur.pos4;
  a.upos4 = c->upos4;
  a.fbond4 = c->any_bonds ? c->fbond4 : nullptr;
  return a;
}

template <typename Real, bool MOL, int KM, bool LIST, int PM>
void launch_force_r(rds_ctx* c, const StepArgs& a, int mode) {
  dim3 g((unsigned)c->n_tiles);
  const size_t sm = c->smem_bytes;
  switch (c->rows) {
    case 8: hipLaunchKernelGGL((k_force<Real, MOL, KM, 8, LIST, PM>), g, dim3(256), sm, c->stream, a, mode, c->pw); break;
    case 4: hipLaunchKernelGGL((k_force<Real, MOL, KM, 4, LIST, PM>), g, dim3(256), sm, c->stream, a, mode, c->pw); break;
    case 2: hipLaunchKernelGGL((k_force<Real, MOL, KM, 2, LIST, PM>), g, dim3(128), sm, c->stream, a, mode, c->pw); break;
    default: hipLaunchKernelGGL((k_force<Real, MOL, KM, 1, LIST, PM>), g, dim3(64), sm, c->stream, a, mode, c->pw); break;
  }
}

template <typename Real, bool MOL, bool LIST, int PM>
void launch_force_m(rds_ctx* c, const StepArgs& a, int mode) {
  if (!c->fast) launch_force_r<Real, MOL, 0, LIST, PM>(c, a, mode);
  else if (c->pw == 3) launch_force_r<Real, MOL, 2, LIST, PM>(c, a, mode);
  else launch_force_r<Real, MOL, 1, LIST, PM>(c, a, mode);
}

template <typename Real, bool LIST, int PM>
void launch_force_l(rds_ctx* c, const StepArgs& a, int mode) {
  if (c->has_mol) launch_force_m<Real, true, LIST, PM>(c, a, mode);
  else launch_force_m<Real, false, LIST, PM>(c, a, mode);
}

template <typename Real>
void launch_force_p(rds_ctx* c, const StepArgs& a, int mode) {
  if (c->list) launch_force_l<Real, true, 0>(c, a, mode);
  else launch_force_l<Real, false, 0>(c, a, mode);
}

void launch_force(rds_ctx* c, const StepArgs& a, int mode) {
  if (c->precision == RDS_FP64) launch_force_p<double>(c, a, mode);
  else if (c->cpos) launch_force_l<float, false, 1>(c, a, mode);  // cell mode only
  else launch_force_p<float>(c, a, mode);
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
// grid of the conditional grid-stride kernels of list mode
inline unsigned nblk_gs(int64_t n) { return std::min(nblk(n, 256), 4096u); }

// kernel ids for timing
enum { KT_FORCE = 0, KT_SCAN = 1, KT_SCATTER = 2, KT_REACT = 3, KT_BONDS = 4, KT_LIST = 5, KT_N = 6 };

// Cell starts from the histogram (one-pass chunked scan); advances the step
// counter when `advance`.  p: the buffer parity whose look-back words the
// scan uses (it clears the other parity's).
void launch_cell_starts(rds_ctx* c, int advance, int p) {
  const size_t K = (size_t)c->n_chunks;
  hipLaunchKernelGGL(k_cellscan, dim3((unsigned)K), dim3(CHUNK_THREADS), 0, c->stream,
                     c->d_chunks, c->d_rep, c->count, c->cstart, c->d_lb + (size_t)p * K,
                     c->d_lb + (size_t)(p ^ 1) * K, c->d_step, advance, c->d_status);
}

template <bool MOL>
void launch_build(rds_ctx* c, const StepArgs& a) {
  const unsigned g = (unsigned)std::min<int64_t>(c->n_tiles, 4096);
  const size_t sm = c->smem_bytes;
  switch (c->rows) {
    case 8: hipLaunchKernelGGL((k_build<MOL, 8>), dim3(g), dim3(256), sm, c->stream, a); break;
    case 4: hipLaunchKernelGGL((k_build<MOL, 4>), dim3(g), dim3(256), sm, c->stream, a); break;
    case 2: hipLaunchKernelGGL((k_build<MOL, 2>), dim3(g), dim3(128), sm, c->stream, a); break;
    default: hipLaunchKernelGGL((k_build<MOL, 1>), dim3(g), dim3(64), sm, c->stream, a); break;
  }
}

// List mode rebuild, each kernel a no-op unless a rebuild is requested: bin
// `pos`, scan (advancing the step counter when `advance`), sort pos / type /
// pid / molecule into scratch and back, build the lists.  `mark(kernel id,
// begin)` brackets the launches for timing.
template <typename Mark>
void enqueue_rebuild(rds_ctx* c, double2* pos, int* type, int advance, Mark&& mark) {
  const int N = (int)c->n_total;
  int* pid = c->buf[0].pid;
  int* mol = c->has_mol ? c->buf[0].mol : nullptr;
  mark(KT_LIST, true);
  hipLaunchKernelGGL(k_bin, dim3(nblk_gs(N)), dim3(256), 0, c->stream, N, c->d_rep, c->d_rep_of,
                     pos, c->ucell, c->urank, c->count, c->d_status, c->cpos);
  mark(KT_LIST, false);
  mark(KT_SCAN, true);
  hipLaunchKernelGGL(k_scan, dim3(c->n_rep), dim3(SCAN_BLOCK), 0, c->stream, c->d_rep, c->count,
                     c->cstart, c->d_step, advance, c->d_status, 1);
  mark(KT_SCAN, false);
  mark(KT_SCATTER, true);
  hipLaunchKernelGGL(k_sort_list, dim3(nblk_gs(N)), dim3(256), 0, c->stream, N, pos, type, pid,
                     mol, c->ucell, c->urank, c->cstart, c->spos, c->stype, c->spid,
                     c->has_mol ? c->smol : nullptr, c->any_bonds ? c->slot_of : nullptr,
                     c->scell, (int)c->n_cells, c->d_status);
  hipLaunchKernelGGL(k_copy_list, dim3(nblk_gs(N)), dim3(256), 0, c->stream, N, c->spos,
                     c->stype, c->spid, c->has_mol ? c->smol : nullptr, pos, type, pid, mol,
                     c->d_status);
  mark(KT_SCATTER, false);
  Buf b;
  b.pos = pos;
  b.type = type;
  b.pid = pid;
  b.mol = mol;
  StepArgs a = make_args(c, b);
  mark(KT_LIST, true);
  if (c->has_mol) launch_build<true>(c, a);
  else launch_build<false>(c, a);
  mark(KT_LIST, false);
}

// A rebuild outside the step loop (state loaded, re-tiled, diagnostics).
void force_rebuild(rds_ctx* c) {
  Buf& cur = c->buf[c->parity];
  hipLaunchKernelGGL(k_request, dim3(1), dim3(1), 0, c->stream, c->d_status);
  enqueue_rebuild(c, cur.pos, cur.type, 0, [](int, bool) {});
}

struct Timer {
  hipEvent_t ev[2 * KT_N * 2];
  bool on = false;
};

// bond and angle forces of the current state into fbond (fbond4 in three
// dimensions)
void launch_bonds3(rds_ctx* c, const Buf& cur) {
  BondArgs3 b;
  b.n_total = (int)c->n_total;
  b.rep = c->d_rep;
  b.rep_of = c->d_rep_of;
  b.pos = cur.pos4;
  b.pid = cur.pid;
  b.slot_of = c->slot_of;
  b.boff = c->boff;
  b.bpart = c->bpart;
  b.br0 = c->br0;
  b.bu0 = c->bu0;
  b.aoff = c->aoff;
  b.aent = c->aent;
  b.apart = c->apart;
  b.aprm = c->aprm;
  b.angle_pi = c->angle_pi;
  b.angle_thr = c->angle_thr;
  b.status = c->d_status;
  b.fbond = c->fbond4;
  hipLaunchKernelGGL(k3_bonds, dim3(nblk(c->n_total, 256)), dim3(256), 0, c->stream, b);
}

void launch_bonds(rds_ctx* c, const Buf& cur) {
  if (c->ndim == 3) {
    launch_bonds3(c, cur);
    return;
  }
  BondArgs b;
  b.n_total = (int)c->n_total;
  b.rep = c->d_rep;
  b.rep_of = c->d_rep_of;
  b.pos = cur.pos;
  b.pid = cur.pid;
  b.slot_of = c->slot_of;
  b.cpos = c->cpos;
  b.boff = c->boff;
  b.bpart = c->bpart;
  b.br0 = c->br0;
  b.bu0 = c->bu0;
  b.aoff = c->aoff;
  b.aent = c->aent;
  b.apart = c->apart;
  b.aprm = c->aprm;
  b.angle_pi = c->angle_pi;
  b.angle_thr = c->angle_thr;
  b.status = c->d_status;
  b.fbond = c->fbond;
  hipLaunchKernelGGL(k_bonds, dim3(nblk(c->n_total, 256)), dim3(256), 0, c->stream, b);
}

template <typename Real, bool MOL>
void launch_force3_m(rds_ctx* c, const StepArgs& a, int mode) {
  const dim3 g((unsigned)c->n_tiles), b(512);  // 8 waves per tile
  const size_t sm = c->smem_bytes;
  if (!c->fast)
    hipLaunchKernelGGL((k3_tile<Real, MOL, 0>), g, b, sm, c->stream, a, mode, c->pw);
  else if (c->pw == 3)
    hipLaunchKernelGGL((k3_tile<Real, MOL, 2>), g, b, sm, c->stream, a, mode, c->pw);
  else
    hipLaunchKernelGGL((k3_tile<Real, MOL, 1>), g, b, sm, c->stream, a, mode, c->pw);
}

void launch_force3(rds_ctx* c, const StepArgs& a, int mode) {
  if (c->precision == RDS_FP64) {
    if (c->has_mol) launch_force3_m<double, true>(c, a, mode);
    else launch_force3_m<double, false>(c, a, mode);
  } else {
    if (c->has_mol) launch_force3_m<float, true>(c, a, mode);
    else launch_force3_m<float, false>(c, a, mode);
  }
}

// One integration step on the context's stream.  `ev` (optional) receives event
// pairs around each launch.
void enqueue_step(rds_ctx* c, std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>>* ev) {
  Buf& cur = c->buf[c->parity];
  Buf& nxt = c->buf[c->parity ^ 1];
  auto mark = [&](int kid, bool begin) {
    if (!ev) return;
    if (begin) {
      hipEvent_t e0, e1;
      hipEventCreate(&e0);
      hipEventCreate(&e1);
      ev->push_back({kid, {e0, e1}});
      hipEventRecord(e0, c->stream);
    } else {
      hipEventRecord(ev->back().second.second, c->stream);
    }
  };
  if (c->any_bonds) {
    mark(KT_BONDS, true);
    launch_bonds(c, cur);
    mark(KT_BONDS, false);
  }
  StepArgs a = make_args(c, cur);
  if (c->ndim == 3) {
    mark(KT_FORCE, true);
    launch_force3(c, a, MODE_STEP);
    mark(KT_FORCE, false);
    mark(KT_SCAN, true);
    launch_cell_starts(c, 1, c->parity);
    mark(KT_SCAN, false);
    mark(KT_SCATTER, true);
    hipLaunchKernelGGL(k3_scatter, dim3(nblk(c->n_total, 256)), dim3(256), 0, c->stream,
                       (int)c->n_total, c->upos4, c->ucell, c->urank, c->cstart, cur.type,
                       cur.pid, cur.mol, nxt.pos4, c->any_rx ? nxt.type0 : nxt.type, nxt.pid,
                       nxt.mol, c->any_bonds ? c->slot_of : nullptr, (int)c->n_cells,
                       c->d_status);
    mark(KT_SCATTER, false);
    if (c->any_rx) {
      StepArgs r = make_args(c, nxt);
      r.type = nxt.type0;
      r.type_out = nxt.type;
      mark(KT_REACT, true);
      if (c->has_mol)
        hipLaunchKernelGGL((k3_react<true>), dim3((unsigned)c->n_sblocks), dim3(256), 0, c->stream, r);
      else
        hipLaunchKernelGGL((k3_react<false>), dim3((unsigned)c->n_sblocks), dim3(256), 0, c->stream, r);
      mark(KT_REACT, false);
    }
    c->parity ^= 1;
    return;
  }
  if (c->list) {
    // positions move in place (same slots) into the other buffer; types and
    // the sort are touched only by a rebuild; reactions write the other type
    // column (without reactions both buffers share one type column)
    a.upos = nxt.pos;
    mark(KT_FORCE, true);
    launch_force(c, a, MODE_STEP);
    mark(KT_FORCE, false);
    enqueue_rebuild(c, nxt.pos, cur.type, 1, mark);
    if (c->any_rx) {
      StepArgs r = make_args(c, nxt);
      r.type = cur.type;
      r.type_out = nxt.type;
      mark(KT_REACT, true);
      const int nb = (int)c->n_sblocks;
      hipLaunchKernelGGL(k_react_list, dim3((unsigned)nb), dim3(256), 0, c->stream, r);
      // one wave per slot block of a replica whose model has a walk
      const int nw = (int)c->n_wblocks;
      const unsigned g = (unsigned)((nw + 3) / 4);
      if (nw > 0 && c->has_mol)
        hipLaunchKernelGGL((k_react_walk<true>), dim3(g), dim3(256), 0, c->stream, r, nw);
      else if (nw > 0)
        hipLaunchKernelGGL((k_react_walk<false>), dim3(g), dim3(256), 0, c->stream, r, nw);
      mark(KT_REACT, false);
    }
    c->parity ^= 1;
    return;
  }
  mark(KT_FORCE, true);
  launch_force(c, a, MODE_STEP);
  mark(KT_FORCE, false);
  mark(KT_SCAN, true);
  launch_cell_starts(c, 1, c->parity);
  mark(KT_SCAN, false);
  // reactions: the scatter decides the particles whose rates need no
  // neighbours and lists the others for k_react_walk, which reads the
  // pre-reaction types (type0) of the new buffer
  StepArgs r = make_args(c, nxt);
  r.type = nxt.type0;
  r.type_out = nxt.type;
  mark(KT_SCATTER, true);
  if (c->any_rx)
    hipLaunchKernelGGL(k_scatter_react, dim3((unsigned)c->n_sblocks), dim3(256), 0, c->stream,
                       c->upos, c->urank, c->cstart, cur.type, cur.pid, cur.mol,
                       nxt.pos, nxt.type, nxt.pid, nxt.mol, c->any_bonds ? c->slot_of : nullptr,
                       nxt.type0, r);
  else
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)c->n_sblocks), dim3(256), 0, c->stream,
                       c->d_sblocks, c->d_rep, c->cpos, (int)c->n_total, c->upos, c->urank,
                       c->cstart, cur.type, cur.pid, cur.mol, nxt.pos, nxt.type, nxt.pid, nxt.mol,
                       c->any_bonds ? c->slot_of : nullptr, c->d_status);
  mark(KT_SCATTER, false);
  if (c->any_rx) {
    mark(KT_REACT, true);
    // one wave per slot block of a replica whose model has a walk
    const int nw = (int)c->n_wblocks;
    const unsigned g = (unsigned)((nw + 3) / 4);
    if (nw > 0 && c->has_mol)
      hipLaunchKernelGGL((k_react_walk<true>), dim3(g), dim3(256), 0, c->stream, r, nw);
    else if (nw > 0)
      hipLaunchKernelGGL((k_react_walk<false>), dim3(g), dim3(256), 0, c->stream, r, nw);
    mark(KT_REACT, false);
  }
  c->parity ^= 1;
}

void free_graph(rds_ctx* c) {
  for (int p = 0; p < 2; ++p) {
    if (c->graph_exec[p]) hipGraphExecDestroy(c->graph_exec[p]);
    if (c->graph[p]) hipGraphDestroy(c->graph[p]);
    c->graph_exec[p] = nullptr;
    c->graph[p] = nullptr;
  }
  c->graph_len = 0;
}

// Captures graph_steps steps twice, once starting from each buffer parity, so
// that a run replays graphs whatever the parity its previous steps left.
int build_graph(rds_ctx* c) {
  free_graph(c);
  const int g = c->graph_steps;
  if (g <= 0) return RDS_OK;
  const int par = c->parity;
  for (int p = 0; p < 2; ++p) {
    c->parity = p;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < g; ++k) enqueue_step(c, nullptr);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    c->parity = par;  // capture executed nothing
    if (e != hipSuccess) return fail(RDS_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    c->graph[p] = graph;
    HIP_TRY(hipGraphInstantiate(&c->graph_exec[p], c->graph[p], nullptr, nullptr, 0));
  }
  c->graph_len = g;
  return RDS_OK;
}

template <typename T>
int dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess)
    return fail(RDS_E_HIP, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  return RDS_OK;
}

#define RDS_TRY(expr)          \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != RDS_OK) return rc_; \
  } while (0)

int fill_model(const rds_model& m, DModel& d, double* rcut, int ndim) {
  memset(&d, 0, sizeof d);
  if (m.n_types < 1 || m.n_types > MAXT) return fail(RDS_E_ARG, "n_types %d outside [1, %d]", m.n_types, MAXT);
  if (m.n_terms < 0 || m.n_terms > MAXTERMS) return fail(RDS_E_ARG, "n_terms %d outside [0, %d]", m.n_terms, MAXTERMS);
  if (m.n_walls < 0 || m.n_walls > MAXWALLS) return fail(RDS_E_ARG, "n_walls %d outside [0, %d]", m.n_walls, MAXWALLS);
  if (m.n_reactions < 0 || m.n_reactions > MAXRX) return fail(RDS_E_ARG, "n_reactions %d outside [0, %d]", m.n_reactions, MAXRX);
  const int nt = m.n_types;
  d.nt = nt;
  d.nwalls = m.n_walls;
  d.nrx = m.n_reactions;
  double rc = 0.0;
  // terms grouped per unordered type pair, file order kept inside a group
  int k = 0;
  for (int a = 0; a < nt; ++a)
    for (int b = a; b < nt; ++b) {
      int first = k;
      for (int q = 0; q < m.n_terms; ++q) {
        const rds_pair_term& t = m.terms[q];
        if (t.type_a != a || t.type_b != b) continue;
        DTerm& o = d.terms[k++];
        memset(&o, 0, sizeof o);
        o.kind = t.kind;
        o.n = t.n;
        double rcv = t.rc;
        if (t.kind == RDS_PIECEWISE_LINEAR) rcv = t.r0;  // pyrdsim.pyx:1057-1060
        if (t.kind == RDS_LJ) {
          if (t.n < 1) return fail(RDS_E_ARG, "lj exponent n must be >= 1");
          o.odd = t.n % 2;
          o.pw = o.odd ? t.n : t.n / 2;
          o.c_lj = -2.0 * t.n * t.u0;
        } else if (t.kind == RDS_TANH) {
          o.a_inv = 1.0 / t.a;
        } else if (t.kind == RDS_PIECEWISE_LINEAR) {
          o.rin = t.r0 - t.dr;
          o.slope = t.u0 / t.dr;
        } else if (t.kind != RDS_YUKAWA && t.kind != RDS_HARMONIC) {
          return fail(RDS_E_ARG, "unknown pair kind %d", t.kind);
        }
        o.u0 = t.u0;
        o.r0 = t.r0;
        o.r0sq = t.r0 * t.r0;
        o.rc2 = rcv * rcv;
        o.fu0 = (float)o.u0;
        o.fr0 = (float)o.r0;
        o.frc2 = (float)rcv * (float)rcv;
        o.fa_inv = (float)o.a_inv;
        o.frin = (float)o.rin;
        o.fc_lj = (float)o.c_lj;
        o.fr0sq = (float)o.r0 * (float)o.r0;
        o.fslope = (float)o.slope;
        rc = std::max(rc, rcv);
      }
      d.pfirst[a * nt + b] = d.pfirst[b * nt + a] = first;
      d.pcount[a * nt + b] = d.pcount[b * nt + a] = k - first;
    }
  // terms whose pair is not sorted / out of range never match (pyrdsim.pyx:1999-2007)
  d.nterms = k;
  d.rc2max = 0.0;
  for (int q = 0; q < k; ++q) d.rc2max = std::max(d.rc2max, d.terms[q].rc2);
  for (int q = 0; q < m.n_walls; ++q) {
    const rds_wall_term& w = m.walls[q];
    if (w.direction < 0 || w.direction >= ndim)
      return fail(RDS_E_ARG, "wall direction must lie in [0, %d)", ndim);
    d.walls[q].type = w.particle_type;
    d.walls[q].dir = w.direction;
    d.walls[q].u0 = w.u0;
    d.walls[q].r0 = w.r0;
    d.walls[q].a_inv = 1.0 / w.a;
  }
  for (int q = 0; q < m.n_reactions; ++q) {
    const rds_reaction_def& r = m.reactions[q];
    DRx& o = d.rx[q];
    o.kind = r.kind;
    o.reactant = r.reactant;
    o.product = r.product;
    o.catalyst = r.catalyst;
    o.inh = r.inhibitor;
    o.inh2 = r.inhibitor2;
    o.r0 = r.r0;
    o.rate_decrease = r.rate_decrease;
    o.r0sq_lo = r.r0 * r.r0 * (1.0 - 1e-14);
    o.r0sq_hi = r.r0 * r.r0 * (1.0 + 1e-14);
    if (r.reactant < 0 || r.reactant >= nt || r.product < 0 || r.product >= nt)
      return fail(RDS_E_ARG, "reaction %d: reactant/product outside the %d particle types", q, nt);
    if (r.kind == RDS_RX_CATALYZED || r.kind == RDS_RX_INHIBITED) {
      d.nbr_rx = 1;
      d.walk_types |= 1u << r.reactant;
      rc = std::max(rc, r.r0);
    } else if (r.kind != RDS_RX_LINEAR) {
      return fail(RDS_E_ARG, "unknown reaction kind %d", r.kind);
    }
  }
  *rcut = rc;
  return RDS_OK;
}

// Tiles of `rows` cell rows x W cells with W chosen so that a row strip holds
// about `target` particles at the replica's mean occupancy, and the LDS round
// capacity: 1.5x the mean staged count of the densest tile shape, at least
// 256, at most what 64 KB of dynamic LDS holds (denser tiles stage in rounds).
bool build_tiles(rds_ctx* c, int rows, int target) {
  if (rows != 1 && rows != 2 && rows != 4 && rows != 8) {
    fail(RDS_E_ARG, "tile_rows must be 1, 2, 4 or 8");
    return false;
  }
  if (target <= 0) target = 56;
  c->rows = rows;
  c->h_tiles.clear();
  double staged_worst = 0.0;
  // The force kernel gives each XCD a contiguous eighth of the tile list.
  // Replicas enter the list in a stride order (stride about 0.618 n_rep,
  // coprime with it), so that every eighth mixes replicas of all models and
  // sizes; a replica's own tiles stay together (halo rows meet in one L2).
  int stride = std::max(1, (int)std::lround(0.618 * c->n_rep));
  while (std::gcd(stride, c->n_rep) != 1) ++stride;
  for (int k = 0; k < c->n_rep; ++k) {
    const int r = (int)(((long long)k * stride) % c->n_rep);
    const DRep& d = c->h_rep[r];
    double occ = (double)d.n / ((double)d.lx * d.ly);
    int W = (int)std::floor((double)target / std::max(occ, 1e-9));
    W = std::max(1, std::min(W, std::min(MAXW, d.lx)));
    int nstrip = (d.lx + W - 1) / W;
    int Wq = d.lx / nstrip, Wr = d.lx % nstrip;
    staged_worst = std::max(staged_worst, occ * (rows + 2) * (Wq + (Wr ? 1 : 0) + 2));
    for (int y = 0; y < d.ly; y += rows) {
      int nr = std::min(rows, d.ly - y);
      int x0 = 0;
      for (int k = 0; k < nstrip; ++k) {
        int wk = Wq + (k < Wr ? 1 : 0);
        c->h_tiles.push_back(make_int4(r | ((nr - 1) << 24), y, x0, wk));
        x0 += wk;
      }
    }
  }
  c->n_tiles = (int64_t)c->h_tiles.size();
  const int per = 16 + 4 + (c->has_mol ? 4 : 0);
  // 1.5x headroom, trimmed (not below 1.15x) where it would cost the force
  // kernel occupancy: 4 x waves-per-SIMD waves per CU of the register-resident
  // path leave 160 KB / (4 waves / tile waves) of LDS per workgroup
  const size_t tile_sh = rows == 8   ? sizeof(TileSh<8>)
                         : rows == 4 ? sizeof(TileSh<4>)
                         : rows == 2 ? sizeof(TileSh<2>)
                                     : sizeof(TileSh<1>);
  const int waves = c->precision == RDS_FP64 ? FORCE_WAVES64 : FORCE_WAVES32;
  const long budget =
      (long)(160 * 1024 * tile_waves(rows) / (4 * waves)) - (long)tile_sh - 64;
  const int cap_occ = budget > 0 ? (int)(budget / per / 64 * 64) : 0;
  int cap = (int)std::ceil(1.5 * staged_worst / 64.0) * 64;
  cap = std::max((int)std::ceil(1.15 * staged_worst / 64.0) * 64, std::min(cap, cap_occ));
  cap = std::max(cap, 256);
  cap = std::min(cap, (65536 - 256) / per / 64 * 64);
  c->cap = cap;
  // the force kernel's staging holds one sentinel slot beyond the capacity
  c->smem_bytes = (size_t)(cap + 1) * per;
  return true;
}

// Three-dimensional tiles: T3Y x T3Z home rows (y, z) x W cells, W chosen so
// that a row strip holds about `target` particles at the replica's mean
// occupancy; the LDS round capacity: 1.25x the mean staged count of the
// densest tile shape, at least 256, at most what 64 KB of dynamic LDS holds
// (denser tiles stage in rounds).
// Default particles per row strip of the 3-D tiles, measured on the LJ
// mixture at density 0.6 with 8-wave workgroups: about 500 home particles
// per tile, one per lane.
int default_target3(const rds_ctx*) { return 128; }

bool build_tiles3(rds_ctx* c, int target) {
  if (target <= 0) target = default_target3(c);
  c->h_tiles.clear();
  double staged_worst = 0.0;
  int stride = std::max(1, (int)std::lround(0.618 * c->n_rep));
  while (std::gcd(stride, c->n_rep) != 1) ++stride;
  for (int q = 0; q < c->n_rep; ++q) {
    const int r = (int)(((long long)q * stride) % c->n_rep);
    const DRep& d = c->h_rep[r];
    const double occ = (double)d.n / ((double)d.lx * d.ly * d.lz);
    int W = (int)std::floor((double)target / std::max(occ, 1e-9));
    W = std::max(1, std::min(W, std::min(MAXW3, d.lx)));
    const int nstrip = (d.lx + W - 1) / W;
    const int Wq = d.lx / nstrip, Wr = d.lx % nstrip;
    staged_worst = std::max(staged_worst, occ * T3R * (Wq + (Wr ? 1 : 0) + 2));
    for (int z = 0; z < d.lz; z += T3Z) {
      const int nz = std::min(T3Z, d.lz - z);
      for (int y = 0; y < d.ly; y += T3Y) {
        const int ny = std::min(T3Y, d.ly - y);
        int x0 = 0;
        for (int k = 0; k < nstrip; ++k) {
          const int wk = Wq + (k < Wr ? 1 : 0);
          c->h_tiles.push_back(make_int4(r | ((ny - 1) << 24) | ((nz - 1) << 26), y, z,
                                         x0 | (wk << 16)));
          x0 += wk;
        }
      }
    }
  }
  c->n_tiles = (int64_t)c->h_tiles.size();
  // FP32 pair arithmetic stages one float4 per particle, FP64 a double2, a
  // double and a type
  const int per = (c->precision == RDS_FP64 ? 16 + 8 + 4 : 16) + (c->has_mol ? 4 : 0);
  int cap = (int)std::ceil(1.25 * staged_worst / 64.0) * 64;
  cap = std::max(cap, 256);
  cap = std::min(cap, (65536 - 256) / per / 64 * 64);
  c->cap = cap;
  // one sentinel slot beyond the capacity
  c->smem_bytes = (size_t)(cap + 1) * per;
  return true;
}

int upload_tiles(rds_ctx* c) {
  if (c->d_tiles) hipFree(c->d_tiles);
  c->d_tiles = nullptr;
  RDS_TRY(dalloc(&c->d_tiles, (size_t)c->n_tiles));
  HIP_TRY(hipMemcpy(c->d_tiles, c->h_tiles.data(), sizeof(int4) * c->n_tiles,
                    hipMemcpyHostToDevice));
  return RDS_OK;
}

}  // namespace

extern "C" {

int rds_abi_version(void) { return RDS_ABI_VERSION; }

int rds_set_tiling(rds_ctx* c, int32_t tile_rows, int32_t tile_particles) {
  if (!c) return fail(RDS_E_ARG, "null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->ndim == 3) {
    // tiles of T3Y x T3Z rows; tile_rows is accepted and kept, not used
    if (tile_rows != 1 && tile_rows != 2 && tile_rows != 4 && tile_rows != 8)
      return fail(RDS_E_ARG, "tile_rows must be 1, 2, 4 or 8");
    c->rows = tile_rows;
    build_tiles3(c, tile_particles);
  } else if (!build_tiles(c, tile_rows, tile_particles)) {
    return RDS_E_ARG;
  }
  RDS_TRY(upload_tiles(c));
  c->target = tile_particles > 0 ? tile_particles : (c->ndim == 3 ? default_target3(c) : 56);
  if (c->loaded && c->list) {
    // lists hold staged indices of the old tiles
    force_rebuild(c);
    HIP_TRY(hipGetLastError());
  }
  if (c->loaded) RDS_TRY(build_graph(c));
  return RDS_OK;
}

int rds_get_tiling(rds_ctx* c, int32_t* tile_rows, int32_t* tile_particles, int32_t* lds_cap) {
  if (!c) return fail(RDS_E_ARG, "null context");
  if (tile_rows) *tile_rows = c->rows;
  if (tile_particles) *tile_particles = c->target;
  if (lds_cap) *lds_cap = c->cap;
  return RDS_OK;
}

const char* rds_last_error(void) { return g_err.c_str(); }

int rds_ctx_destroy(rds_ctx* c) {
  if (!c) return RDS_OK;
  free_graph(c);
  if (c->stream) hipStreamSynchronize(c->stream);
  std::vector<void*> ptrs = {c->d_rep, c->d_models, c->d_tiles, c->d_rep_of, c->d_sblocks,
                             c->upos, c->ucell, c->urank, c->count, c->cstart, c->d_step,
                             c->d_chunks, c->d_lb, c->wl, c->wl_n, c->d_wblocks,
                             c->d_status, c->boff, c->bpart, c->br0, c->bu0, c->fbond,
                             c->slot_of, c->clabel, c->aoff, c->aent, c->apart, c->aprm, c->nl4, c->ncnt, c->xref, c->scell, c->spos,
                             c->stype, c->spid, c->smol};
  ptrs.push_back(c->upos4);
  ptrs.push_back(c->fbond4);
  for (int k = 0; k < 2; ++k) {
    void* b[] = {c->buf[k].pos, c->buf[k].pos4, c->buf[k].type, c->buf[k].type0, c->buf[k].pid,
                 c->buf[k].mol};
    for (void* p : b) ptrs.push_back(p);
  }
  // list mode shares some columns between the two buffers: free each once
  std::sort(ptrs.begin(), ptrs.end());
  ptrs.erase(std::unique(ptrs.begin(), ptrs.end()), ptrs.end());
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (c->h_changed) hipHostFree(c->h_changed);
  if (c->stream && c->own_stream) hipStreamDestroy(c->stream);
  delete c;
  return RDS_OK;
}

// Bonds and angles name distinct particles of one replica.
static int check_topology(const rds_config* cfg) {
  if (cfg->n_bonds < 0 || cfg->n_angles < 0) return fail(RDS_E_ARG, "negative bond or angle count");
  if (cfg->n_bonds > 0 && (!cfg->bond_i || !cfg->bond_j || !cfg->bond_r0 || !cfg->bond_u0))
    return fail(RDS_E_ARG, "bond arrays missing");
  if (cfg->n_angles > 0 && (!cfg->angle_i || !cfg->angle_j || !cfg->angle_k ||
                            !cfg->angle_theta0 || !cfg->angle_k0))
    return fail(RDS_E_ARG, "angle arrays missing");
  if (cfg->n_bonds == 0 && cfg->n_angles == 0) return RDS_OK;
  std::vector<int64_t> ends(cfg->n_replicas);  // exclusive end pid of each replica
  int64_t n = 0;
  for (int r = 0; r < cfg->n_replicas; ++r) ends[r] = n += std::max(cfg->replicas[r].n_particles, 0);
  auto rep = [&](int64_t i) {
    return (int)(std::upper_bound(ends.begin(), ends.end(), i) - ends.begin());
  };
  auto valid = [&](int64_t i) { return i >= 0 && i < n; };
  for (int64_t b = 0; b < cfg->n_bonds; ++b) {
    const int i = cfg->bond_i[b], j = cfg->bond_j[b];
    if (!valid(i) || !valid(j) || i == j)
      return fail(RDS_E_ARG, "bond %lld has invalid particles (%d, %d)", (long long)b, i, j);
    if (rep(i) != rep(j)) return fail(RDS_E_ARG, "bond %lld spans replicas", (long long)b);
  }
  for (int64_t q = 0; q < cfg->n_angles; ++q) {
    const int i = cfg->angle_i[q], j = cfg->angle_j[q], k = cfg->angle_k[q];
    if (!valid(i) || !valid(j) || !valid(k) || i == j || j == k || i == k)
      return fail(RDS_E_ARG, "angle %lld has invalid particles (%d, %d, %d)", (long long)q, i, j, k);
    if (rep(i) != rep(j) || rep(k) != rep(j))
      return fail(RDS_E_ARG, "angle %lld spans replicas", (long long)q);
  }
  return RDS_OK;
}

int rds_ctx_create(const rds_config* cfg, rds_ctx** out) {
  if (!cfg || !out) return fail(RDS_E_ARG, "null argument");
  *out = nullptr;
  if (cfg->n_models < 1 || cfg->n_replicas < 1 || !cfg->models || !cfg->replicas)
    return fail(RDS_E_ARG, "need at least one model and one replica");
  if (cfg->precision != RDS_FP64 && cfg->precision != RDS_MIXED &&
      cfg->precision != RDS_COMPENSATED)
    return fail(RDS_E_ARG, "unknown precision %d", cfg->precision);
  if (cfg->precision == RDS_COMPENSATED && cfg->skin > 0.0)
    return fail(RDS_E_ARG, "compensated FP32 positions run with skin = 0 (cell search every step)");
  if (cfg->rng != RDS_RNG_PHILOX && cfg->rng != RDS_RNG_INJECTED)
    return fail(RDS_E_ARG, "unknown rng %d", cfg->rng);
  if (int e = check_topology(cfg)) return e;
  rds_ctx* c = new rds_ctx();
  int rc = RDS_OK;
  auto bail = [&](int code) {
    rds_ctx_destroy(c);
    return code;
  };
  c->precision = cfg->precision;
  c->cpos = cfg->precision == RDS_COMPENSATED ? 1 : 0;
  c->ndim = cfg->n_dim != 0 ? cfg->n_dim : 2;
  if (c->ndim != 2 && c->ndim != 3) return bail(fail(RDS_E_ARG, "n_dim must be 2 or 3"));
  if (c->ndim == 3 && (cfg->skin > 0.0 || cfg->precision == RDS_COMPENSATED))
    return bail(fail(RDS_E_ARG,
                     "three dimensions run the cell search (skin = 0) in FP64 or mixed precision"));
  c->rng = cfg->rng;
  c->prob_mode = cfg->reaction_prob;
  c->has_mol = cfg->has_molecules ? 1 : 0;
  c->graph_steps = cfg->graph_steps;
  c->n_rep = cfg->n_replicas;
  c->n_models = cfg->n_models;
  c->rows = cfg->tile_rows > 0 ? cfg->tile_rows : 1;
  if (c->rows != 1 && c->rows != 2 && c->rows != 4 && c->rows != 8)
    return bail(fail(RDS_E_ARG, "tile_rows must be 1, 2, 4 or 8"));
  c->target = cfg->tile_particles > 0 ? cfg->tile_particles : (c->ndim == 3 ? default_target3(c) : 56);
  if (cfg->n_replicas >= (1 << 24)) return bail(fail(RDS_E_ARG, "too many replicas"));
  const int spc = cfg->cells_per_cutoff != 0 ? std::abs(cfg->cells_per_cutoff) : 1;
  if (spc != 1) return bail(fail(RDS_E_ARG, "only cells_per_cutoff = 1 (3x3 stencil) is built"));
  const double target = cfg->tile_particles > 0 ? cfg->tile_particles : 56;
  if (!(cfg->skin >= 0.0) || !std::isfinite(cfg->skin))
    return bail(fail(RDS_E_ARG, "skin must be finite and >= 0"));
  c->skin = cfg->skin;
  c->list = cfg->skin > 0.0;
  c->K = cfg->list_capacity > 0 ? cfg->list_capacity : 64;
  if (c->K > 4096) return bail(fail(RDS_E_ARG, "list_capacity %d above 4096", c->K));
  c->K = (c->K + 3) / 4 * 4;  // whole chunks of four entries

  std::vector<double> rcut(cfg->n_models);
  c->h_models.resize(cfg->n_models);
  for (int m = 0; m < cfg->n_models; ++m) {
    rc = fill_model(cfg->models[m], c->h_models[m], &rcut[m], c->ndim);
    if (rc) return bail(rc);
    if (cfg->models[m].n_reactions > 0) c->any_rx = true;
  }
  // register-resident LJ path: every model <= 2 types, <= 1 term per type
  // pair, all LJ with one common even exponent
  c->fast = true;
  c->pw = 0;
  for (int m = 0; m < cfg->n_models && c->fast; ++m) {
    const DModel& d = c->h_models[m];
    if (d.nt > 2) c->fast = false;
    for (int q = 0; q < d.nt * d.nt && c->fast; ++q)
      if (d.pcount[q] > 1) c->fast = false;
    for (int q = 0; q < d.nterms && c->fast; ++q) {
      const DTerm& t = d.terms[q];
      if (t.kind != RDS_LJ || t.odd || (c->pw && t.pw != c->pw)) c->fast = false;
      else c->pw = t.pw;
    }
  }
  if (c->pw == 0) c->pw = 1;
  if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(RDS_E_HIP, "no HIP device"));
  c->h_rep.resize(c->n_rep);
  int64_t pbase = 0, cbase = 0;
  for (int r = 0; r < c->n_rep; ++r) {
    const rds_replica& s = cfg->replicas[r];
    DRep& d = c->h_rep[r];
    memset(&d, 0, sizeof d);
    if (s.model < 0 || s.model >= cfg->n_models) return bail(fail(RDS_E_ARG, "replica %d: bad model", r));
    if (s.n_particles < 1) return bail(fail(RDS_E_ARG, "replica %d: no particles", r));
    const int nd = c->ndim;
    double rcm = rcut[s.model];
    if (!(rcm > 0.0)) {  // no pair range: one cell suffices
      rcm = std::min(s.L[0], s.L[1]);
      if (nd == 3) rcm = std::min(rcm, s.L[2]);
    } else {
      rcm += c->skin;  // list mode: cells >= range + skin
    }
    int lcell[3] = {1, 1, 1};
    d.L[2] = d.cl[2] = d.icl[2] = 1.0;
    for (int k = 0; k < nd; ++k) {
      if (!(s.L[k] > 0.0)) return bail(fail(RDS_E_ARG, "replica %d: box length must be > 0", r));
      // largest l with L/l >= rcut (the reference's rule, pyrdsim.pyx:1439-1441,
      // with rs = the largest interaction range and ratio = 1)
      int l = 1;
      while (s.L[k] / (double)(l + 1) >= rcm) ++l;
      if (s.periodic[k] && l < 3) {
        if (rcut[s.model] > 0.0)
          return bail(fail(RDS_E_ARG,
                           "replica %d: periodic box length %g holds fewer than 3 cells of the "
                           "interaction range %g (+ skin %g)",
                           r, s.L[k], rcut[s.model], c->skin));
      }
      lcell[k] = l;
      d.L[k] = s.L[k];
      d.cl[k] = s.L[k] / (double)l;
      d.icl[k] = 1.0 / d.cl[k];
    }
    // a periodic axis with fewer than 3 cells (only possible without any pair
    // range) is treated as one cell with no neighbours: nothing interacts
    d.lx = lcell[0];
    d.ly = lcell[1];
    d.px = s.periodic[0] ? 1 : 0;
    d.py = s.periodic[1] ? 1 : 0;
    d.sx = (d.px && d.lx >= 3) ? 1 : 0;
    d.sy = (d.py && d.ly >= 3) ? 1 : 0;
    d.lz = lcell[2];
    d.pz = (nd == 3 && s.periodic[2]) ? 1 : 0;
    d.sz = (d.pz && d.lz >= 3) ? 1 : 0;
    d.ncell = d.lx * d.ly * d.lz;
    d.cbase = (int)cbase;
    d.pbase = (int)pbase;
    d.n = s.n_particles;
    d.model = s.model;
    d.seed_lo = (unsigned)(s.seed & 0xffffffffu);
    d.seed_hi = (unsigned)(s.seed >> 32);
    for (int t = 0; t < MAXT; ++t) {
      d.fpre[t] = s.force_prefac[t];
      d.npre[t] = s.noise_prefac[t];
    }
    for (int q = 0; q < MAXRX; ++q) d.rate[q] = s.rate[q];
    d.dt = s.dt;
    for (int k = 0; k < 2; ++k) {
      d.Lh[k] = (float)s.L[k];
      d.Ll[k] = (float)(s.L[k] - (double)d.Lh[k]);
    }
    for (int t = 0; t < MAXT; ++t) {
      d.fph[t] = (float)s.force_prefac[t];
      d.fpl[t] = (float)(s.force_prefac[t] - (double)d.fph[t]);
      d.nph[t] = (float)s.noise_prefac[t];
      d.npl[t] = (float)(s.noise_prefac[t] - (double)d.nph[t]);
    }
    {
      const double rl = std::sqrt(c->h_models[s.model].rc2max) + c->skin;
      d.rl2 = rl * rl;
    }
    pbase += s.n_particles;
    cbase += (int64_t)d.ncell;
    if (pbase > 0x7fffffff || cbase > 0x7ffffff0)
      return bail(fail(RDS_E_ARG, "too many particles or cells for 32-bit indices"));
  }
  c->n_total = pbase;
  c->n_cells = cbase;
  // force tiles
  if (c->ndim == 2 && !build_tiles(c, c->rows, (int)target)) return bail(RDS_E_ARG);
  if (c->ndim == 3) build_tiles3(c, cfg->tile_particles);

  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  if ((rc = scratch_ptr(c->device, &c->static_labels)) != RDS_OK) return bail(rc);
  {
    const hipError_t e = hipHostMalloc((void**)&c->h_changed, sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess)
      return bail(fail(RDS_E_HIP, "hipHostMalloc: %s", hipGetErrorString(e)));
  }
  const size_t N = (size_t)c->n_total, C = (size_t)c->n_cells;
#define A_(p, n)                                          \
  do {                                                    \
    if ((rc = dalloc(&(p), (n))) != RDS_OK) return bail(rc); \
  } while (0)
  A_(c->d_rep, c->n_rep);
  A_(c->d_models, c->n_models);
  A_(c->d_rep_of, N);
  if (c->ndim == 3) {
    for (int k = 0; k < 2; ++k) {
      A_(c->buf[k].pos4, N);
      A_(c->buf[k].type, N);
      A_(c->buf[k].type0, N);
      A_(c->buf[k].pid, N);
      if (c->has_mol) A_(c->buf[k].mol, N);
    }
    A_(c->upos4, N);
  } else if (!c->list) {
    for (int k = 0; k < 2; ++k) {
      A_(c->buf[k].pos, N);
      A_(c->buf[k].type, N);
      A_(c->buf[k].type0, N);
      A_(c->buf[k].pid, N);
      if (c->has_mol) A_(c->buf[k].mol, N);
    }
    A_(c->upos, N);
  } else {
    // positions ping-pong; pid / molecule columns are shared (only a rebuild
    // reorders them, in place); types ping-pong only when reactions write them
    A_(c->buf[0].pos, N);
    A_(c->buf[1].pos, N);
    A_(c->buf[0].type, N);
    if (c->any_rx) {
      A_(c->buf[1].type, N);
    } else {
      c->buf[1].type = c->buf[0].type;
    }
    A_(c->buf[0].pid, N);
    c->buf[1].pid = c->buf[0].pid;
    if (c->has_mol) {
      A_(c->buf[0].mol, N);
      c->buf[1].mol = c->buf[0].mol;
      A_(c->smol, N);
    }
    A_(c->nl4, (size_t)(c->K / 4 + 1) * N);
    A_(c->ncnt, N);
    A_(c->xref, N);
    A_(c->scell, N);
    A_(c->spos, N);
    A_(c->stype, N);
    A_(c->spid, N);
  }
  A_(c->ucell, N);
  A_(c->urank, N);
  A_(c->count, C);
  A_(c->cstart, C + 1);
  {
    std::vector<int4> ch;
    for (int r = 0; r < c->n_rep; ++r) {
      const DRep& d = c->h_rep[r];
      const int cr = d.ncell, first = (int)ch.size();
      for (int o = 0; o < cr; o += CHUNK_CELLS)
        ch.push_back(make_int4(r, o, std::min(CHUNK_CELLS, cr - o), first));
    }
    c->n_chunks = (int64_t)ch.size();
    A_(c->d_chunks, ch.size());
    A_(c->d_lb, 2 * ch.size());
    HIP_TRY(hipMemcpy(c->d_chunks, ch.data(), sizeof(int4) * ch.size(), hipMemcpyHostToDevice));
  }
  A_(c->d_step, 1);
  A_(c->d_status, 1);
  A_(c->slot_of, N);

  HIP_TRY(hipMemcpy(c->d_rep, c->h_rep.data(), sizeof(DRep) * c->n_rep, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_models, c->h_models.data(), sizeof(DModel) * c->n_models, hipMemcpyHostToDevice));
  if ((rc = upload_tiles(c)) != RDS_OK) return bail(rc);
  {
    std::vector<int> rep_of(N);
    std::vector<int4> sb;
    for (int r = 0; r < c->n_rep; ++r) {
      const DRep& d = c->h_rep[r];
      for (int i = 0; i < d.n; ++i) rep_of[d.pbase + i] = r;
      for (int b = 0; b < d.n; b += 256)
        sb.push_back(make_int4(r, d.pbase + b, d.pbase + std::min(b + 256, d.n), 0));
    }
    HIP_TRY(hipMemcpy(c->d_rep_of, rep_of.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    c->n_sblocks = (int64_t)sb.size();
    if ((rc = dalloc(&c->d_sblocks, sb.size())) != RDS_OK) return bail(rc);
    // reaction worklist: a segment of 256 entries per slot block
    if ((rc = dalloc(&c->wl, sb.size() * 256)) != RDS_OK) return bail(rc);
    if ((rc = dalloc(&c->wl_n, sb.size())) != RDS_OK) return bail(rc);
    // the reaction walk runs over the slot blocks of replicas whose model has
    // a catalysed or inhibited reaction only
    std::vector<int4> wb;
    for (size_t b = 0; b < sb.size(); ++b) {
      const int r = sb[b].x, m = c->h_rep[r].model;
      if (c->h_models[m].walk_types) wb.push_back(make_int4((int)b, r, m, 0));
    }
    c->n_wblocks = (int64_t)wb.size();
    if ((rc = dalloc(&c->d_wblocks, std::max<size_t>(wb.size(), 1))) != RDS_OK) return bail(rc);
    if (!wb.empty())
      HIP_TRY(hipMemcpy(c->d_wblocks, wb.data(), sizeof(int4) * wb.size(), hipMemcpyHostToDevice));

    HIP_TRY(hipMemcpy(c->d_sblocks, sb.data(), sizeof(int4) * sb.size(), hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemset(c->count, 0, sizeof(int) * C));
  HIP_TRY(hipMemset(c->d_lb, 0, sizeof(unsigned long long) * 2 * (size_t)c->n_chunks));
  {
    int nt = (int)N;
    HIP_TRY(hipMemcpy(c->cstart + C, &nt, sizeof(int), hipMemcpyHostToDevice));
  }

  // bonds: CSR by particle, both directions.  The legs (i1, i2) and (i2, i3)
  // of each angle join it with u0 = 0: the bond kernel skips them, and the
  // reaction kernel counts them as bonded partners (pyrdsim.pyx:1405-1406).
  // (indices checked by check_topology)
  if (cfg->n_bonds > 0 || cfg->n_angles > 0) {
    struct Leg {
      int i, j;
      double r0, u0;
    };
    std::vector<Leg> legs;
    for (int64_t b = 0; b < cfg->n_bonds; ++b)
      legs.push_back({cfg->bond_i[b], cfg->bond_j[b], cfg->bond_r0[b], cfg->bond_u0[b]});
    for (int64_t q = 0; q < cfg->n_angles; ++q) {
      legs.push_back({cfg->angle_i[q], cfg->angle_j[q], 0.0, 0.0});
      legs.push_back({cfg->angle_j[q], cfg->angle_k[q], 0.0, 0.0});
    }
    std::vector<int> deg(N + 1, 0);
    for (const Leg& l : legs) {
      deg[l.i + 1]++;
      deg[l.j + 1]++;
    }
    for (size_t i = 0; i < N; ++i) deg[i + 1] += deg[i];
    std::vector<int> part(deg[N]), fill(deg.begin(), deg.end() - 1);
    std::vector<double> r0v(deg[N]), u0v(deg[N]);
    for (const Leg& l : legs) {
      int q = fill[l.i]++;
      part[q] = l.j;
      r0v[q] = l.r0;
      u0v[q] = l.u0;
      q = fill[l.j]++;
      part[q] = l.i;
      r0v[q] = l.r0;
      u0v[q] = l.u0;
    }
    A_(c->boff, N + 1);
    A_(c->bpart, part.size());
    A_(c->br0, r0v.size());
    A_(c->bu0, u0v.size());
    if (c->ndim == 3) A_(c->fbond4, N);
    else A_(c->fbond, N);
    HIP_TRY(hipMemcpy(c->boff, deg.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->bpart, part.data(), sizeof(int) * part.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->br0, r0v.data(), sizeof(double) * r0v.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->bu0, u0v.data(), sizeof(double) * u0v.size(), hipMemcpyHostToDevice));
    c->any_bonds = true;
  }
  // angles (pyrdsim.pyx:1322-1407): table in definition order and a CSR by
  // particle of (angle, role) memberships
  if (cfg->n_angles > 0) {
    const int64_t na = cfg->n_angles;
    std::vector<int4> ap(na);
    std::vector<double2> prm(na);
    std::vector<int> off(N + 1, 0);
    for (int64_t q = 0; q < na; ++q) {
      ap[q] = make_int4(cfg->angle_i[q], cfg->angle_j[q], cfg->angle_k[q], 0);
      prm[q] = make_double2(cfg->angle_theta0[q], cfg->angle_k0[q]);
      off[ap[q].x + 1]++;
      off[ap[q].y + 1]++;
      off[ap[q].z + 1]++;
    }
    for (size_t i = 0; i < N; ++i) off[i + 1] += off[i];
    std::vector<int2> ent(off[N]);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int64_t q = 0; q < na; ++q) {
      ent[fill[ap[q].x]++] = make_int2((int)q, 0);
      ent[fill[ap[q].y]++] = make_int2((int)q, 1);
      ent[fill[ap[q].z]++] = make_int2((int)q, 2);
    }
    A_(c->aoff, N + 1);
    A_(c->aent, ent.size());
    A_(c->apart, na);
    A_(c->aprm, na);
    HIP_TRY(hipMemcpy(c->aoff, off.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->aent, ent.data(), sizeof(int2) * ent.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->apart, ap.data(), sizeof(int4) * na, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->aprm, prm.data(), sizeof(double2) * na, hipMemcpyHostToDevice));
    // the closures read the flag set for the last angle defined
    // (pyrdsim.pyx:1334-1338, 1378); np.isclose(theta0, pi, atol=1e-5)
    const double last = cfg->angle_theta0[na - 1];
    c->angle_pi = std::fabs(last - M_PI) <= 1e-5 + 1e-5 * M_PI;
    c->angle_thr = cfg->angle_threshold;
  }
#undef A_
  *out = c;
  return RDS_OK;
}

int rds_set_state(rds_ctx* c, const double* x, const int32_t* types, const int32_t* mol,
                  int64_t step0) {
  if (!c || !x || !types) return fail(RDS_E_ARG, "null argument");
  if (c->has_mol && !mol) return fail(RDS_E_ARG, "molecule numbers required");
  const int N = (int)c->n_total;
  c->parity = 0;
  Buf& a = c->buf[0];
  Buf& b = c->buf[1];
  if (c->ndim == 2)
    hipLaunchKernelGGL(k_load, dim3(nblk(N, 256)), dim3(256), 0, c->stream, N,
                       (const double2*)x, types, c->has_mol ? mol : nullptr, a.pos, a.type,
                       a.pid, c->has_mol ? a.mol : nullptr, c->cpos);
  DStatus st;
  memset(&st, 0, sizeof st);
  st.err_key = ~0ull;
  HIP_TRY(hipMemcpyAsync(c->d_status, &st, sizeof st, hipMemcpyHostToDevice, c->stream));
  // the load's scan uses parity 0 and clears parity 1's look-back words for
  // the first step; parity 0's may hold a previous run's last words
  if (!c->list)
    HIP_TRY(hipMemsetAsync(c->d_lb, 0, sizeof(unsigned long long) * 2 * (size_t)c->n_chunks,
                           c->stream));
  if (c->ndim == 3) {
    hipLaunchKernelGGL(k3_load, dim3(nblk(N, 256)), dim3(256), 0, c->stream, N, x, types,
                       c->has_mol ? mol : nullptr, a.pos4, a.type, a.pid,
                       c->has_mol ? a.mol : nullptr);
    hipLaunchKernelGGL(k3_bin, dim3(nblk(N, 256)), dim3(256), 0, c->stream, N, c->d_rep,
                       c->d_rep_of, a.pos4, c->ucell, c->urank, c->count);
    launch_cell_starts(c, 0, 0);
    hipLaunchKernelGGL(k3_scatter, dim3(nblk(N, 256)), dim3(256), 0, c->stream, N, a.pos4,
                       c->ucell, c->urank, c->cstart, a.type, a.pid, a.mol, b.pos4, b.type,
                       b.pid, b.mol, c->slot_of, (int)c->n_cells, c->d_status);
    c->parity = 1;
  } else if (c->list) {
    // sorted in place, lists built
    force_rebuild(c);
  } else {
    hipLaunchKernelGGL(k_bin, dim3(nblk(N, 256)), dim3(256), 0, c->stream, N, c->d_rep,
                       c->d_rep_of, a.pos, c->ucell, c->urank, c->count, nullptr, c->cpos);
    launch_cell_starts(c, 0, 0);
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)c->n_sblocks), dim3(256), 0, c->stream,
                       c->d_sblocks, c->d_rep, c->cpos, N, a.pos, c->urank, c->cstart, a.type,
                       a.pid, a.mol, b.pos, b.type, b.pid, b.mol, c->slot_of, c->d_status);
    c->parity = 1;
  }
  long long s0 = step0;
  HIP_TRY(hipMemcpyAsync(c->d_step, &s0, sizeof s0, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipGetLastError());
  c->loaded = true;
  c->step0 = step0;
  c->steps_done = 0;
  RDS_TRY(build_graph(c));
  return RDS_OK;
}

int rds_get_state(rds_ctx* c, double* x, int32_t* types) {
  if (!c) return fail(RDS_E_ARG, "null context");
  if (!c->loaded) return fail(RDS_E_STATE, "no state loaded");
  Buf& cur = c->buf[c->parity];
  if (c->ndim == 3)
    hipLaunchKernelGGL(k3_gather_out, dim3(nblk(c->n_total, 256)), dim3(256), 0, c->stream,
                       (int)c->n_total, cur.pos4, cur.type, cur.pid, x, types);
  else
    hipLaunchKernelGGL(k_gather_out, dim3(nblk(c->n_total, 256)), dim3(256), 0, c->stream,
                       (int)c->n_total, cur.pos, cur.type, cur.pid, (double2*)x, types, c->cpos);
  HIP_TRY(hipGetLastError());
  return RDS_OK;
}

int rds_set_injected(rds_ctx* c, const double* normals, const double* u1, const double* u2,
                     int64_t first_step, int64_t n_steps) {
  if (!c) return fail(RDS_E_ARG, "null context");
  c->inj_n = normals;
  c->inj_u1 = u1;
  c->inj_u2 = u2;
  c->inj_first = first_step;
  c->inj_count = n_steps;
  if (c->loaded) RDS_TRY(build_graph(c));
  return RDS_OK;
}

int rds_set_random_force(rds_ctx* c, int32_t include) {
  if (!c) return fail(RDS_E_ARG, "null context");
  c->include_rf = include ? 1.0 : 0.0;
  if (c->loaded) RDS_TRY(build_graph(c));
  return RDS_OK;
}

int rds_run(rds_ctx* c, int64_t n_steps) {
  if (!c) return fail(RDS_E_ARG, "null context");
  if (!c->loaded) return fail(RDS_E_STATE, "run before set_state");
  if (c->rng == RDS_RNG_INJECTED && !c->inj_n) return fail(RDS_E_STATE, "no injected streams set");
  int64_t left = n_steps;
  while (left > 0) {
    if (c->graph_len > 0 && left >= c->graph_len) {
      HIP_TRY(hipGraphLaunch(c->graph_exec[c->parity], c->stream));
      c->parity ^= c->graph_len & 1;
      left -= c->graph_len;
      c->steps_done += c->graph_len;
    } else {
      enqueue_step(c, nullptr);
      left -= 1;
      c->steps_done += 1;
    }
  }
  HIP_TRY(hipGetLastError());
  return RDS_OK;
}

int rds_sync(rds_ctx* c, rds_status* out) {
  if (!c) return fail(RDS_E_ARG, "null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  DStatus st;
  HIP_TRY(hipMemcpy(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost));
  rds_status s;
  memset(&s, 0, sizeof s);
  s.steps_done = c->steps_done;
  s.bounds_error = st.bounds;
  s.stream_error = st.stream;
  s.reaction_events = (int64_t)st.events;
  s.integrity = st.integrity;
  s.rebuilds = (int64_t)st.rebuilds;
  s.list_overflow = (int64_t)st.overflow;
  s.list_entries = (int64_t)st.entries;
  s.angle_error = st.angle;
  if (st.angle && !st.bounds) s.error_value = st.angle_val;
  if (st.bounds) {
    s.error_step = (int64_t)(st.err_key >> 32);
    unsigned lo = (unsigned)(st.err_key & 0xffffffffu);
    s.error_particle = (int32_t)(lo / (unsigned)c->ndim);
    s.error_axis = (int32_t)(lo % (unsigned)c->ndim);
    s.error_value = st.err_val;
  }
  if (out) *out = s;
  if (st.bounds && st.err_lost && st.err_key == ~0ull)
    return fail(RDS_E_BOUNDS, "a coordinate left the box (the failing particle was not "
                              "recorded). Maybe the timestep is too large for the given "
                              "the interaction parameters?");
  if (st.integrity)
    return fail(RDS_E_STATE, "internal cell-list invariant violated (flags %d)", st.integrity);
  if (st.bounds)
    return fail(RDS_E_BOUNDS,
                "x[%d,%d] = %.17g out of bounds. Maybe the timestep is too large for the given "
                "the interaction parameters? (step %lld)",
                s.error_particle, s.error_axis, s.error_value, (long long)s.error_step);
  if (st.angle)
    return fail(RDS_E_ANGLE, "Inner product of units vectors is smaller than -1. Value = %.17g",
                st.angle_val);
  if (st.stream) return fail(RDS_E_STREAM, "injected random stream exhausted");
  return RDS_OK;
}

int rds_compute_forces(rds_ctx* c, double* forces) {
  if (!c || !forces) return fail(RDS_E_ARG, "null argument");
  if (!c->loaded) return fail(RDS_E_STATE, "no state loaded");
  Buf& cur = c->buf[c->parity];
  if (c->any_bonds)
    launch_bonds(c, cur);
  StepArgs a = make_args(c, cur);
  a.fout = (double2*)forces;
  if (c->ndim == 3) launch_force3(c, a, MODE_FORCE);
  else launch_force(c, a, MODE_FORCE);
  HIP_TRY(hipGetLastError());
  return RDS_OK;
}

int rds_cell_ids(rds_ctx* c, int32_t* cells) {
  if (!c || !cells) return fail(RDS_E_ARG, "null argument");
  if (!c->loaded) return fail(RDS_E_STATE, "no state loaded");
  Buf& cur = c->buf[c->parity];
  if (c->ndim == 3)
    hipLaunchKernelGGL(k3_cells_out, dim3(nblk(c->n_total, 256)), dim3(256), 0, c->stream,
                       (int)c->n_total, c->d_rep, c->d_rep_of, cur.pos4, cur.pid, cells);
  else
    hipLaunchKernelGGL(k_cells_out, dim3(nblk(c->n_total, 256)), dim3(256), 0, c->stream,
                       (int)c->n_total, c->d_rep, c->d_rep_of, cur.pos, cur.pid, cells, c->cpos);
  HIP_TRY(hipGetLastError());
  return RDS_OK;
}

int rds_grid(rds_ctx* c, int32_t* cells_xy, double* cell_len, int32_t* cell_base) {
  if (!c) return fail(RDS_E_ARG, "null context");
  for (int r = 0; r < c->n_rep; ++r) {
    if (cells_xy) {
      cells_xy[2 * r] = c->h_rep[r].lx;
      cells_xy[2 * r + 1] = c->h_rep[r].ly;
    }
    if (cell_len) {
      cell_len[2 * r] = c->h_rep[r].cl[0];
      cell_len[2 * r + 1] = c->h_rep[r].cl[1];
    }
    if (cell_base) cell_base[r] = c->h_rep[r].cbase;
  }
  return RDS_OK;
}

int rds_grid3(rds_ctx* c, int32_t* cells_xyz, double* cell_len, int32_t* cell_base) {
  if (!c) return fail(RDS_E_ARG, "null context");
  for (int r = 0; r < c->n_rep; ++r) {
    const DRep& d = c->h_rep[r];
    if (cells_xyz) {
      cells_xyz[3 * r] = d.lx;
      cells_xyz[3 * r + 1] = d.ly;
      cells_xyz[3 * r + 2] = d.lz;
    }
    if (cell_len)
      for (int k = 0; k < 3; ++k) cell_len[3 * r + k] = d.cl[k];
    if (cell_base) cell_base[r] = d.cbase;
  }
  return RDS_OK;
}

int rds_neighbor_counts(rds_ctx* c, double r, int32_t* counts) {
  if (!c || !counts) return fail(RDS_E_ARG, "null argument");
  if (!c->loaded) return fail(RDS_E_STATE, "no state loaded");
  for (int q = 0; q < c->n_rep; ++q) {
    const DRep& d = c->h_rep[q];
    double cmin = std::min(d.cl[0], d.cl[1]);
    if (c->ndim == 3) cmin = std::min(cmin, d.cl[2]);
    if (r > cmin)
      return fail(RDS_E_ARG, "probe radius %g exceeds the cell length of replica %d", r, q);
  }
  if (c->ndim == 3) {
    Buf& cur = c->buf[c->parity];
    hipLaunchKernelGGL(k3_count, dim3((unsigned)c->n_sblocks), dim3(256), 0, c->stream, c->d_rep,
                       c->d_sblocks, c->cstart, cur.pos4, c->has_mol ? cur.mol : nullptr, cur.pid,
                       r, counts);
    HIP_TRY(hipGetLastError());
    return RDS_OK;
  }
  // list mode: cells are those of the last rebuild; re-sort for the current positions
  if (c->list) force_rebuild(c);
  Buf& cur = c->buf[c->parity];
  StepArgs a = make_args(c, cur);
  a.iout = counts;
  a.rprobe2 = r * r;
  if (c->has_mol) hipLaunchKernelGGL((k_count<true>), dim3((unsigned)c->n_tiles), dim3(64), c->smem_bytes, c->stream, a);
  else hipLaunchKernelGGL((k_count<false>), dim3((unsigned)c->n_tiles), dim3(64), c->smem_bytes, c->stream, a);
  HIP_TRY(hipGetLastError());
  return RDS_OK;
}

// Largest analysis radius: half the shortest periodic box length of any
// replica (the minimum image must be unique).
static int check_radius(rds_ctx* c, double r, const char* what) {
  if (!(r > 0.0) || !std::isfinite(r)) return fail(RDS_E_ARG, "%s must be finite and > 0", what);
  for (int q = 0; q < c->n_rep; ++q) {
    const DRep& d = c->h_rep[q];
    for (int k = 0; k < c->ndim; ++k)
      if ((k == 0 ? d.px : k == 1 ? d.py : d.pz) && r > 0.5 * d.L[k])
        return fail(RDS_E_ARG, "%s %g exceeds half the box length of replica %d", what, r, q);
  }
  return RDS_OK;
}

int rds_pair_histogram(rds_ctx* c, double r_max, double dr, int32_t n_bins, uint32_t type_mask,
                       uint64_t* hist) {
  if (!c || !hist) return fail(RDS_E_ARG, "null argument");
  if (!c->loaded) return fail(RDS_E_STATE, "no state loaded");
  if (int e = check_radius(c, r_max, "r_max")) return e;
  if (!(dr > 0.0) || n_bins < 1 || n_bins > 4096)
    return fail(RDS_E_ARG, "need dr > 0 and 1 <= n_bins <= 4096");
  Buf& cur = c->buf[c->parity];
  HIP_TRY(hipMemsetAsync(hist, 0, sizeof(uint64_t) * n_bins * c->n_rep, c->stream));
  if (c->ndim == 3) {
    hipLaunchKernelGGL(k3_pair_hist, dim3((unsigned)c->n_sblocks), dim3(256),
                       sizeof(unsigned) * n_bins, c->stream, c->d_rep, c->d_sblocks, c->cstart,
                       cur.pos4, cur.type, r_max * r_max, dr, r_max, (int)n_bins,
                       type_mask ? type_mask : 0xffffffffu, (unsigned long long*)hist);
    HIP_TRY(hipGetLastError());
    return RDS_OK;
  }
  // list mode: cells are those of the last rebuild, each particle within
  // skin/2 of its cell, so the rings reach r_max + skin
  hipLaunchKernelGGL(k_pair_hist, dim3((unsigned)c->n_sblocks), dim3(256),
                     sizeof(unsigned) * n_bins, c->stream, c->d_rep, c->d_sblocks, c->cstart,
                     cur.pos, cur.type, c->cpos, r_max * r_max, dr, r_max + c->skin, (int)n_bins,
                     type_mask ? type_mask : 0xffffffffu, (unsigned long long*)hist);
  HIP_TRY(hipGetLastError());
  return RDS_OK;
}

int rds_type_counts(rds_ctx* c, int64_t* counts) {
  if (!c || !counts) return fail(RDS_E_ARG, "null argument");
  if (!c->loaded) return fail(RDS_E_STATE, "no state loaded");
  Buf& cur = c->buf[c->parity];
  HIP_TRY(hipMemsetAsync(counts, 0, sizeof(int64_t) * MAXT * c->n_rep, c->stream));
  hipLaunchKernelGGL(k_type_counts, dim3((unsigned)c->n_sblocks), dim3(256), 0, c->stream,
                     c->d_sblocks, cur.type, (unsigned long long*)counts);
  HIP_TRY(hipGetLastError());
  return RDS_OK;
}

// Cluster labels into `lab` (n_total entries of device memory): min-label
// propagation passes until no label changes, one host check per pass.
static int cluster_labels(rds_ctx* c, double r_link, uint32_t type_mask, int32_t* labels,
                          int* lab) {
  Buf& cur = c->buf[c->parity];
  const int N = (int)c->n_total;
  const unsigned mask = type_mask ? type_mask : 0xffffffffu;
  hipLaunchKernelGGL(k_cl_init, dim3(nblk(N, 256)), dim3(256), 0, c->stream, N, cur.type,
                     cur.pid, mask, lab, c->slot_of);
  // every pass that changes a label lowers one; N + 1 passes always suffice
  for (int pass = 0;; ++pass) {
    if (pass > N) return fail(RDS_E_STATE, "cluster labels did not converge");
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(1), 0, c->stream, &c->d_status->cl_changed);
    if (c->ndim == 3)
      hipLaunchKernelGGL(k3_cl_prop, dim3((unsigned)c->n_sblocks), dim3(256), 0, c->stream,
                         c->d_rep, c->d_sblocks, c->cstart, cur.pos4, r_link * r_link, r_link,
                         lab, c->slot_of, c->d_status);
    else
      hipLaunchKernelGGL(k_cl_prop, dim3((unsigned)c->n_sblocks), dim3(256), 0, c->stream,
                         c->d_rep, c->d_sblocks, c->cstart, cur.pos, c->cpos, r_link * r_link,
                         r_link + c->skin, lab, c->slot_of, c->d_status);
    hipLaunchKernelGGL(k_cl_jump, dim3(nblk(N, 256)), dim3(256), 0, c->stream, N, lab,
                       c->slot_of);
    HIP_TRY(hipMemcpyAsync(c->h_changed, &c->d_status->cl_changed, sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!*c->h_changed) break;
  }
  hipLaunchKernelGGL(k_cl_out, dim3(nblk(N, 256)), dim3(256), 0, c->stream, N, lab, cur.pid,
                     labels);
  HIP_TRY(hipGetLastError());
  return RDS_OK;
}

int rds_cluster_labels(rds_ctx* c, double r_link, uint32_t type_mask, int32_t* labels) {
  if (!c || !labels) return fail(RDS_E_ARG, "null argument");
  if (!c->loaded) return fail(RDS_E_STATE, "no state loaded");
  if (int e = check_radius(c, r_link, "r_link")) return e;
  if (!c->clabel && dalloc(&c->clabel, c->n_total) != RDS_OK) return RDS_E_HIP;
  return cluster_labels(c, r_link, type_mask, labels, c->clabel);
}

int rds_cluster_labels_static(rds_ctx* c, double r_link, uint32_t type_mask, int32_t* labels) {
  if (!c || !labels) return fail(RDS_E_ARG, "null argument");
  if (c->n_total > RDS_STATIC_LABELS_MAX)
    return fail(RDS_E_ARG,
                "rds_cluster_labels_static handles at most %d particles (context has %lld); "
                "use rds_cluster_labels",
                RDS_STATIC_LABELS_MAX, (long long)c->n_total);
  if (!c->loaded) return fail(RDS_E_STATE, "no state loaded");
  if (int e = check_radius(c, r_link, "r_link")) return e;
  return cluster_labels(c, r_link, type_mask, labels, c->static_labels);
}

int rds_time_kernels(rds_ctx* c, int64_t n_steps, double* avg_ms, int64_t* counts) {
  if (!c) return fail(RDS_E_ARG, "null context");
  if (!c->loaded) return fail(RDS_E_STATE, "no state loaded");
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
  for (int64_t k = 0; k < n_steps; ++k) {
    enqueue_step(c, &ev);
    c->steps_done += 1;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  double sum[KT_N] = {0};
  int64_t cnt[KT_N] = {0};
  for (auto& e : ev) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, e.second.first, e.second.second);
    sum[e.first] += ms;
    cnt[e.first] += 1;
    hipEventDestroy(e.second.first);
    hipEventDestroy(e.second.second);
  }
  for (int k = 0; k < KT_N; ++k) {
    if (avg_ms) avg_ms[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
    if (counts) counts[k] = cnt[k];
  }
  HIP_TRY(hipGetLastError());
  return RDS_OK;
}

int rds_debug_fill_lds(rds_ctx* c, uint32_t pattern) {
  if (!c) return fail(RDS_E_ARG, "null context");
  const int words = 60 * 1024 / 4;  // 60 KB per workgroup, many workgroups per CU
  hipLaunchKernelGGL(k_fill_lds, dim3(4096), dim3(256), words * 4, c->stream, pattern, words);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RDS_OK;
}

int rds_sizes(rds_ctx* c, int64_t* n_total, int64_t* n_cells, int64_t* n_tiles) {
  if (!c) return fail(RDS_E_ARG, "null context");
  if (n_total) *n_total = c->n_total;
  if (n_cells) *n_cells = c->n_cells;
  if (n_tiles) *n_tiles = c->n_tiles;
  return RDS_OK;
}

int rds_stream(rds_ctx* c, void** stream) {
  if (!c || !stream) return fail(RDS_E_ARG, "null argument");
  *stream = (void*)c->stream;
  return RDS_OK;
}

}  // extern "C"
