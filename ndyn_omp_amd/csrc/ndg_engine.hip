// ndg_engine.hip -- host side of the MD-step engine and the C ABI (include/ndyn_gpu.h).
//
// An engine owns a HIP stream, a set of replica slots and the device tables that
// describe them (RepParams / RepState / Tile arrays).  Each replica lives in one device
// block; pair tables are shared between replicas with identical tables.  A single-system
// handle (ndg_t) is an engine with one slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ndg_frames.h"
#include "ndg_internal.h"

namespace ndg {
void launch_drift(const EngineDev &E, int path, int t0, int nt, int xflags, hipStream_t s);
void launch_kick(const EngineDev &E, int path, int t0, int nt, int xflags, hipStream_t s);
// kick != 0 (mixed path): the force kernel also applies integrate2 to the atoms it updates
void launch_force(const EngineDev &E, int path, int tpa, bool full, int t0, int nt, int xflags,
                  hipStream_t s, int kick = 0);
void launch_reduce(const EngineDev &E, int r0, int nr, int mode, int xflags, int trigger,
                   hipStream_t s);
void launch_apply_pending(const EngineDev &E, int r0, int nr, hipStream_t s);
void launch_snapshot(const EngineDev &E, int r, int np, double *out, hipStream_t s);
void launch_rebuild(const EngineDev &E, int nt, int r0, int nr, hipStream_t s);
void launch_copy(const float4 *src, float4 *dst, size_t n, hipStream_t s);
enum : int { RED_FORCE = 1, RED_KICK = 2, RED_FLIP = 4, RED_STEP = 8, RED_CHECK = 16 };
}  // namespace ndg

using namespace ndg;

// ------------------------------------------------------------------ errors

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      return fail(NDG_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
  } while (0)

// For sequences that must run to the end (teardown, giving back a cache): the first
// failing call sets rc (and the message); the later calls still run.
#define HIPKEEP(rc, x)                                                               \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess && (rc) == NDG_OK)                                          \
      (rc) = fail(NDG_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
  } while (0)

static const char *err_text(int code) {
  switch (code) {
    case NDG_ERR_NEIMAX: return "too many neighbours in list";
    case NDG_ERR_OUTSIDE: return "Particle outside of box";
    case NDG_ERR_RANGE:
      return "an atom moved beyond the packed coordinate range of the mixed path since the "
             "last list build (rebuild the lists, use the skin trigger, or the FP64 path)";
    default: return "engine error";
  }
}

// ------------------------------------------------------------------ pair tables (host)
// buildPairTables (src/potentials.f90:2-56) with pot/dpot (src/potentials.f90:59-247):
// Morse (:94-110, :184-204, the /r of :198 kept), Buckingham (:112-122, :206-216),
// Lennard-Jones (:124-133, :218-227), spring (:88-92, :178-183: a table written for output
// only, the force kernels take spring forces from the strain), sine / cosine / soft
// (:82-87, :172-177; soft's force takes parm(1), parm(2) where its energy takes parm(2),
// parm(3), as the reference does), Johnson (:136-147, :230-241) and the Coulomb variants
// (:106-108, :120, :131 and :199-201, :213-214, :224-225); NDG_MORSE_PHYS tabulates
// -r V'(r) for Morse.

namespace {
constexpr double kPi = 3.141592653589793;   // interpot's PI
constexpr double kAlphaC = 0.2;             // damping of the shifted Coulomb term (:11)

// the shifted Coulomb constants of a pair with cutoff rc (src/potentials.f90:23-26)
struct Coul {
  double vcon, fcon;
  explicit Coul(double rc) {
    const double rc2 = rc * rc;
    vcon = std::erfc(kAlphaC * rc) / rc;
    fcon = std::erfc(kAlphaC * rc) / rc2 + 2.0 * kAlphaC / std::sqrt(kPi) * std::exp(-kAlphaC * kAlphaC * rc2) / rc;
  }
  double v(double r, double rc, double q) const {
    return q * (std::erfc(kAlphaC * r) / r - vcon + fcon * (r - rc));
  }
  double dv(double r, double q) const {
    return q * (std::erfc(kAlphaC * r) / (r * r) +
                2.0 * kAlphaC / std::sqrt(kPi) * std::exp(-(kAlphaC * kAlphaC) * (r * r)) / r - fcon);
  }
};
}  // namespace

static double pair_v(int kind, double r, const double *p, double qprd = 0.0) {
  switch (kind) {
    case NDG_MORSE:
    case NDG_MORSE_C:
    case NDG_MORSE_PHYS: {
      double x = p[3] <= 0.0 ? 0.0 : r - p[3];
      double e = std::exp(-p[2] * x);
      double v = p[1] * (e * e - 2.0 * e);
      return kind == NDG_MORSE_C ? v + Coul(p[0]).v(r, p[0], qprd) : v;
    }
    case NDG_BUCK:
    case NDG_BUCK_C: {
      double r6 = r * r * r * r * r * r;
      double v = p[2] <= 0.0 ? p[1] - p[3] / r6 : p[1] * std::exp(-r / p[2]) - p[3] / r6;
      return kind == NDG_BUCK_C ? v + Coul(p[0]).v(r, p[0], qprd) : v;
    }
    case NDG_LJ:
    case NDG_LJ_C: {
      double q = p[2] / r, q6 = q * q * q * q * q * q;
      double v = 4.0 * p[1] * (q6 * q6 - q6);
      return kind == NDG_LJ_C ? v + Coul(p[0]).v(r, p[0], qprd) : v;
    }
    case NDG_SINE: return std::sin(r) * std::sin(r) - 1.0;
    case NDG_COSINE: return std::cos(r) * std::cos(r) - 1.0;
    case NDG_SOFT: return p[1] * (1.0 + std::cos(kPi * r / p[2]));
    case NDG_JOHNSON: {   // eps, r0, rc = PairParams(2:4)
      const double eps = p[1], r0 = p[2], rc = p[3];
      const double A = 2.0 * (r0 * r0 * r0) / ((rc - r0) * (rc - r0) * (rc - r0));
      const double B = 3.0 * (r0 * r0) / ((rc - r0) * (rc - r0));
      const double t = (r - rc) / r0;
      return r < rc ? -eps * (A * (t * t * t) + B * (t * t)) : 0.0;
    }
    case NDG_SPRING: {   // D (r/rc C)^2 - D C^2 with rc, D, C = PairParams(1:3)
      double x = r / p[0] * p[2];
      return p[1] * (x * x - p[2] * p[2]);
    }
  }
  return 0.0;
}

static double pair_dv(int kind, double r, const double *p, double qprd = 0.0) {
  // -dV/dr (reference Morse: /r)
  switch (kind) {
    case NDG_MORSE:
    case NDG_MORSE_C:
    case NDG_MORSE_PHYS: {
      double x = p[3] <= 0.0 ? 0.0 : r - p[3];
      double e = std::exp(-p[2] * x);
      double d = 2.0 * p[1] * p[2] * (e * e - e);
      if (kind == NDG_MORSE_PHYS) return d;
      d = d / r;
      return kind == NDG_MORSE_C ? d + Coul(p[0]).dv(r, qprd) : d;
    }
    case NDG_BUCK:
    case NDG_BUCK_C: {
      double r7 = r * r * r * r * r * r * r;
      double d = p[2] <= 0.0 ? p[1] - 6.0 * p[3] / r7
                             : p[1] * std::exp(-r / p[2]) / p[2] - 6.0 * p[3] / r7;
      return kind == NDG_BUCK_C ? d + Coul(p[0]).dv(r, qprd) : d;
    }
    case NDG_LJ:
    case NDG_LJ_C: {
      double q = p[2] / r, q6 = q * q * q * q * q * q;
      double d = 24.0 * p[1] * (2.0 * q6 * q6 / r - q6 / r);
      return kind == NDG_LJ_C ? d + Coul(p[0]).dv(r, qprd) : d;
    }
    case NDG_SINE: return std::sin(2.0 * r);
    case NDG_COSINE: return -std::sin(2.0 * r);
    case NDG_SOFT: return kPi * p[0] * std::sin(kPi * r / p[1]) / p[1];
    case NDG_JOHNSON: {
      const double eps = p[1], r0 = p[2], rc = p[3];
      const double A = 2.0 * (r0 * r0 * r0) / ((rc - r0) * (rc - r0) * (rc - r0));
      const double B = 3.0 * (r0 * r0) / ((rc - r0) * (rc - r0));
      const double t = (r - rc) / r0;
      return r < rc ? eps / r0 * (3.0 * A * (t * t) + 2.0 * B * t) : 0.0;
    }
    case NDG_SPRING: {
      double x = r / p[0] * p[2];
      return -2.0 * p[1] * x;
    }
  }
  return 0.0;
}

extern "C" int ndg_pair_table_q(int kind, const double params[4], double qprd, double ptrmin,
                                int ptsize, double *P, double *F, double *rc2, double *dr2) {
  const bool eshift = (kind & NDG_ESHIFT) != 0;
  kind &= ~NDG_ESHIFT;
  if (kind < NDG_MORSE || kind > NDG_LJ_C || ptsize < 2 || !P || !F ||
      (eshift && kind == NDG_SPRING))
    return fail(NDG_ERR_ARG, "ndg_pair_table: bad arguments");
  const double rc = params[0];
  const double c2 = rc * rc;
  const double h = (c2 - ptrmin) / (double)(ptsize - 1);
  // NDG_ESHIFT: the pair's energy at its cutoff is taken off the energy table
  const double vc = eshift ? pair_v(kind, rc, params, qprd) : 0.0;
  for (int i = 0; i <= ptsize; ++i) {
    const double s = ptrmin + (double)i * h;
    const double r = std::sqrt(s);
    P[i] = pair_v(kind, r, params, qprd) - vc;
    F[i] = r * pair_dv(kind, r, params, qprd);
  }
  if (rc2) *rc2 = c2;
  if (dr2) *dr2 = h;
  return NDG_OK;
}

extern "C" int ndg_pair_table(int kind, const double params[4], double ptrmin, int ptsize,
                              double *P, double *F, double *rc2, double *dr2) {
  return ndg_pair_table_q(kind, params, 0.0, ptrmin, ptsize, P, F, rc2, dr2);
}

// ------------------------------------------------------------------ engine structures

namespace {

struct TableBuf {       // device copy of one replica's pair tables (shared when identical)
  int refs = 0;
  void *block = nullptr;
  size_t bytes = 0;
  double *F64 = nullptr, *P64 = nullptr;
  float2 *F2 = nullptr, *P2 = nullptr;
  std::vector<std::vector<float>> F32;  // per pair, host copy for the LDS table set
  std::vector<std::vector<float>> P32;  // per pair, FP32 potential table (same layout)
  std::vector<uint64_t> phash;          // per pair content hash
  int npairs = 0, ptsize = 0;
  std::vector<double> P, F;             // the FP64 tables, to confirm a hash match
};

struct HostState {      // caller-provided data for one replica (caller order)
  int np = 0, ntypes = 0, npairs = 0, ptsize = 0;
  std::vector<double> P, F, rc2, dr2, pairparams, mass;
  std::vector<int> pairmat, typ;
  double ptrmin = 0.25;
  std::vector<double> pos, vel, acc;
  double box[3] = {0, 0, 0};
  bool have_tables = false, have_state = false;
  // pair kinds (NDG_MORSE..NDG_SPRING; empty: no spring pairs) and, for spring pairs, the
  // initial positions (caller order, scaled) and box of initPositionSet
  std::vector<int> kinds;
  std::vector<double> ipos;
  double ibox[3] = {0, 0, 0};
  bool have_ipos = false;
  // displacement boundary conditions: per atom (caller order) the condition index + 1
  // (0: none) into bc_dir (bit a: axis a constrained) / bc_du (3 per condition)
  std::vector<int> hbc, bc_dir;
  std::vector<double> bc_du;
  bool holds() const { return !bc_dir.empty(); }
  bool springs() const {
    for (int k : kinds)
      if (k == NDG_SPRING) return true;
    return false;
  }
};

struct Slot {
  bool used = false;
  bool springs = false;     // the block holds initial positions (spring pairs)
  bool ipos_dirty = false;  // host initial positions newer than the device copy
  bool holds = false;       // the block holds per-atom boundary conditions
  bool hbc_dirty = false;   // host boundary conditions newer than the device copy
  int np = 0, ntypes = 0, npairs = 0, ptsize = 0, neimax = 0, tpa = 1, ncell_cap = 0;
  int ntiles = 0, tile_base = 0;
  void *block = nullptr;
  size_t bytes = 0;
  uint64_t table_key = 0;
  RepParams P{};
  ndg_run_params rp{};
  // host mirror of the step schedule
  int step = 0, run_step0 = 0, run_nsteps = 0;
  bool finished = true;
  int nrec_seen = 0;
  std::vector<ndg_thermo> pending_rec;
  ndg_thermo last{};
  int err = 0;
  long long nbr_sum = 0;
  bool auto_neimax = false;   // list capacity chosen by the engine: grow on overflow
};

}  // namespace

struct ndg_engine {
  int path = NDG_PATH_MIXED;
  int device = 0;             // the HIP device the engine's stream and memory live on
  size_t budget = 0, used = 0;
  hipStream_t stream = nullptr;
  std::vector<Slot> slots;
  RepParams *d_reps = nullptr;
  RepState *d_st = nullptr;
  Tile *d_tiles = nullptr;
  int *d_rb = nullptr;        // rebuild work list (count + tile indices), cap_tiles + 1 ints
  int *d_tile0 = nullptr;     // first tile of each slot
  int cap_slots = 0, cap_tiles = 0, cap_tile0 = 0;
  int max_rep_tiles = 0;
  std::vector<Tile> h_tiles;
  std::vector<std::pair<int, int>> tpa_ranges;  // (tpa, t0) in tile order, nt follows
  std::vector<int> tpa_counts;
  bool tiles_dirty = true;
  std::map<uint64_t, TableBuf> tables;
  std::multimap<size_t, void *> free_blocks;
  int tpa_override = 0;
  std::vector<std::pair<int, int>> tpa_classes;   // (np_max, tpa), np_max ascending
  std::vector<std::array<int, 3>> cb_classes;      // (np_max, cb_y, cb_z), np_max ascending
  int graph_steps = 0;
  bool queued = false;   // steps queued by ndg_engine_step_enqueue, not yet drained
  // output snapshot: device staging and pinned host copy of one replica's output state
  double *snap_d = nullptr, *snap_h = nullptr;
  size_t snap_cap = 0, snap_hcap = 0;
  // pinned staging of state uploads (admission, ndg_set_state)
  char *stage_h = nullptr;
  size_t stage_cap = 0;
  hipEvent_t snap_ev = nullptr;
  int snap_slot = -1, snap_np = 0;
  // timing of the force kernel
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
  std::vector<hipEvent_t> ev_pool;
  double t_ms = 0.0;
  long long t_count = 0;
  double t_pairs = 0.0, t_atoms = 0.0;
  // graphs
  std::map<std::string, hipGraphExec_t> graphs;
  // distinct FP32 force tables of all replicas (staged in LDS by the force kernel)
  float *d_tabset = nullptr;
  size_t tabset_cap = 0;
  std::vector<float> tabset_host;   // host copy of the staged table set
  int tabset_n = 0;      // floats of the F half
  int tabset_p = 0;      // 1: the energy/virial variant stages the P half too
};

// Frames queued by ndg_frame_write.  Each waits in one of two pinned buffers (filled by a
// device copy behind the queued steps) until the writer thread has formatted it and
// appended it to its file.
struct FrameWriter {
  struct Buf {
    double *h = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool busy = false;
  };
  struct Job {
    int buf = 0, format = 0, np = 0, name_len = 0;
    std::string path;
    std::vector<int> typ;
    std::vector<char> names;
  };
  Buf bufs[2];
  std::deque<Job> q;
  std::mutex m;
  std::condition_variable cv;
  std::thread th;
  bool stop = false;
  int err = 0;
  std::string errmsg;
};

struct ndg_system {
  ndg_engine *e = nullptr;
  int slot = -1;
  HostState hs;
  int neimax_req = 0;
  int path = NDG_PATH_MIXED;
  bool queued = false;   // steps queued by ndg_run_enqueue and not yet drained
  std::unique_ptr<FrameWriter> fw;
};

// ------------------------------------------------------------------ helpers

static uint64_t fnv1a(const void *p, size_t n, uint64_t h = 1469598103934665603ull) {
  const unsigned char *c = (const unsigned char *)p;
  for (size_t i = 0; i < n; ++i) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
  return h;
}

// Content key of large arrays (the pair tables, hashed at every admission): 8-byte words
// through a multiply-rotate mix, the tail bytewise; ~8x faster than the bytewise FNV-1a.
static uint64_t hash_words(const void *p, size_t n, uint64_t h = 0x9e3779b97f4a7c15ull) {
  const unsigned char *c = (const unsigned char *)p;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    memcpy(&w, c + i, 8);
    h ^= w * 0xbf58476d1ce4e5b9ull;
    h = (h << 31) | (h >> 33);
    h *= 0x94d049bb133111ebull;
  }
  h ^= (uint64_t)n;
  return fnv1a(c + i, n - i, h);
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Layout {
  size_t off[64];
  int n = 0;
  size_t total = 0;
  size_t add(size_t bytes) {
    size_t o = total;
    off[n++] = o;
    total = align_up(total + bytes, 256);
    return o;
  }
};

static int default_tpa(int np) {
  (void)np;
  return 1;
}

static int estimate_neimax(double rcut, double skin, int np, const double box[3]) {
  // src/nd.f90:84-88 (single-precision literals 4./3. and 1.1 as in the reference)
  const double c = (double)(4.0f / 3.0f) * 3.141592653589793;
  const double rr = (rcut + skin) * (double)1.1f;
  return (int)(c * (rr * rr * rr) * (double)np / (box[0] * box[1] * box[2]));
}

static void grid_for_box(const double box[3], double rcut, double skin, int m[3], int &brute) {
  for (int a = 0; a < 3; ++a) m[a] = (int)(box[a] / (rcut + skin));
  brute = (m[0] < 3 || m[1] < 3 || m[2] < 3) ? 1 : 0;
  if (brute) m[0] = m[1] = m[2] = 1;
}

// carve a replica block; returns total bytes (block == nullptr: size only)
static size_t carve(Slot &s, int path, char *block) {
  const int np = s.np;
  const int ntiles = (np + kTile - 1) / kTile;
  const int C = (s.neimax + s.tpa - 1) / s.tpa + kListPad;
  Layout L;
  size_t o_mass = L.add(sizeof(double) * s.ntypes);
  size_t o_pm = L.add(sizeof(int) * s.ntypes * s.ntypes);
  size_t o_rcp = L.add(sizeof(double) * s.npairs);
  size_t o_rc2d = L.add(sizeof(double) * s.npairs);
  size_t o_rc2f = L.add(sizeof(float) * s.npairs);
  size_t o_vdr2 = L.add(sizeof(float) * s.npairs);
  size_t o_d[16];
  for (int q = 0; q < 16; ++q) o_d[q] = L.add(sizeof(double) * np);  // px..pbz (16 arrays)
  const size_t nip = s.springs ? (size_t)np : 0;
  size_t o_ip[3];
  for (int q = 0; q < 3; ++q) o_ip[q] = L.add(sizeof(double) * nip);  // ipx..ipz (springs)
  size_t o_hbc = L.add(sizeof(int) * (s.holds ? (size_t)np : 0));
  size_t o_f32a = L.add(sizeof(float4) * np);
  size_t o_f32b = L.add(sizeof(float4) * np);
  size_t o_x64a = L.add(sizeof(double4) * np);
  size_t o_x32b = L.add(sizeof(float4) * np);
  size_t o_xq = L.add(sizeof(uint4) * np);
  size_t o_anc = L.add(sizeof(int4) * np);
  size_t o_typ = L.add(sizeof(int) * np);
  size_t o_oid = L.add(sizeof(int) * np);
  size_t o_nnb = L.add(sizeof(int) * np);
  size_t o_list = L.add(sizeof(unsigned) * (size_t)ntiles * C * kTile * s.tpa);
  size_t o_cc = L.add(sizeof(int) * s.ncell_cap);
  size_t o_cs = L.add(sizeof(int) * (s.ncell_cap + 1));
  size_t o_cf = L.add(sizeof(int) * s.ncell_cap);
  size_t o_key = L.add(sizeof(int) * np);
  size_t o_src = L.add(sizeof(int) * np);
  size_t o_tmpd = L.add(sizeof(double) * (s.springs ? 12 : 9) * (size_t)np);
  size_t o_tmpf = L.add(sizeof(float4) * 2 * (size_t)np);
  size_t o_tmpi = L.add(sizeof(int) * (s.holds ? 3 : 2) * (size_t)np);
  size_t o_part = L.add(sizeof(double) * kPartials * ntiles);
  size_t o_ring = L.add(sizeof(ndg_thermo) * kRing);
  if (block) {
    RepParams &P = s.P;
    P.mass = (const double *)(block + o_mass);
    P.pairmat = (const int *)(block + o_pm);
    P.rc_pair = (const double *)(block + o_rcp);
    P.rc2d = (const double *)(block + o_rc2d);
    P.rc2f = (const float *)(block + o_rc2f);
    P.vdr2 = (const float *)(block + o_vdr2);
    double **dp[16] = {&P.px, &P.py, &P.pz, &P.vx, &P.vy, &P.vz, &P.ax, &P.ay,
                       &P.az, &P.fx, &P.fy, &P.fz, &P.pene, &P.pbx, &P.pby, &P.pbz};
    for (int q = 0; q < 16; ++q) *dp[q] = (double *)(block + o_d[q]);
    P.ipx = s.springs ? (double *)(block + o_ip[0]) : nullptr;
    P.ipy = s.springs ? (double *)(block + o_ip[1]) : nullptr;
    P.ipz = s.springs ? (double *)(block + o_ip[2]) : nullptr;
    P.hbc = s.holds ? (int *)(block + o_hbc) : nullptr;
    P.f32a = (float4 *)(block + o_f32a);
    P.f32b = (float4 *)(block + o_f32b);
    P.x64a = (double4 *)(block + o_x64a);
    P.x32b = (float4 *)(block + o_x32b);
    P.xq = (uint4 *)(block + o_xq);
    P.anchor = (int4 *)(block + o_anc);
    P.typ = (int *)(block + o_typ);
    P.oid = (int *)(block + o_oid);
    P.nnb = (int *)(block + o_nnb);
    P.list = (unsigned *)(block + o_list);
    P.cell_count = (int *)(block + o_cc);
    P.cell_start = (int *)(block + o_cs);
    P.cell_fill = (int *)(block + o_cf);
    P.key = (int *)(block + o_key);
    P.slot_src = (int *)(block + o_src);
    P.tmpd = (double *)(block + o_tmpd);
    P.tmpf = (float4 *)(block + o_tmpf);
    P.tmpi = (int *)(block + o_tmpi);
    P.partials = (double *)(block + o_part);
    P.ring = (ndg_thermo *)(block + o_ring);
    P.np = np;
    P.ntiles = ntiles;
    P.tpa = s.tpa;
    P.neimax = s.neimax;
    P.list_chunks = C;
    P.ncell_cap = s.ncell_cap;
    P.ntypes = s.ntypes;
    P.npairs = s.npairs;
    P.ptsize = s.ptsize;
    P.path = path;
  }
  (void)path;
  return L.total;
}

static size_t table_bytes(int npairs, int ptsize) {
  const size_t n = (size_t)npairs * (ptsize + 1);
  return align_up(n * 8, 256) * 4;
}

static int dev_alloc(ndg_engine *e, size_t bytes, void **p) {
  auto it = e->free_blocks.lower_bound(bytes);
  if (it != e->free_blocks.end() && it->first <= bytes + bytes / 4 + (1 << 20)) {
    *p = it->second;
    e->free_blocks.erase(it);
    return NDG_OK;
  }
  if (e->used + bytes > e->budget) {
    // give cached blocks back before refusing
    int rc = NDG_OK;
    for (auto &fb : e->free_blocks) {
      HIPKEEP(rc, hipFree(fb.second));
      e->used -= fb.first;
    }
    e->free_blocks.clear();
    if (rc) return rc;
    if (e->used + bytes > e->budget) return fail(NDG_ERR_NOMEM, "HBM budget exhausted");
  }
  hipError_t r = hipMalloc(p, bytes);
  if (r != hipSuccess) return fail(NDG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(r));
  e->used += bytes;
  return NDG_OK;
}

static void dev_release(ndg_engine *e, void *p, size_t bytes) {
  if (p) e->free_blocks.emplace(bytes, p);
}

// Captured step graphs hold the old device tables and tiles: drop them when those move.
static int drop_graphs(ndg_engine *e) {
  int rc = NDG_OK;
  for (auto &g : e->graphs) HIPKEEP(rc, hipGraphExecDestroy(g.second));
  e->graphs.clear();
  return rc;
}

static int ensure_capacity(ndg_engine *e) {
  int ns = (int)e->slots.size();
  if (ns > e->cap_slots) {
    int cap = std::max(ns, std::max(8, 2 * e->cap_slots));
    RepParams *nr;
    RepState *nst;
    HIPCHK(hipMalloc(&nr, sizeof(RepParams) * cap));
    HIPCHK(hipMalloc(&nst, sizeof(RepState) * cap));
    HIPCHK(hipMemsetAsync(nst, 0, sizeof(RepState) * cap, e->stream));
    HIPCHK(hipMemsetAsync(nr, 0, sizeof(RepParams) * cap, e->stream));
    if (e->d_reps) {
      HIPCHK(hipMemcpyAsync(nr, e->d_reps, sizeof(RepParams) * e->cap_slots, hipMemcpyDeviceToDevice, e->stream));
      HIPCHK(hipMemcpyAsync(nst, e->d_st, sizeof(RepState) * e->cap_slots, hipMemcpyDeviceToDevice, e->stream));
      HIPCHK(hipStreamSynchronize(e->stream));
      HIPCHK(hipFree(e->d_reps));
      HIPCHK(hipFree(e->d_st));
    }
    e->d_reps = nr;
    e->d_st = nst;
    e->cap_slots = cap;
    int rc = drop_graphs(e);
    if (rc) return rc;
  }
  return NDG_OK;
}

// Distinct FP32 force tables of all replicas, concatenated; each replica gets the offset
// of each of its pairs (or -1 when the set exceeds the LDS budget).
// The distinct FP32 force tables of all replicas, then their potential tables in the same
// order: [F ... | P ...].  The force kernel stages the F half in LDS; its energy/virial
// variant stages both halves when they fit (tabset_p).
static int rebuild_tabset(ndg_engine *e) {
  // offset of each distinct (F, P) pair table in the set: keyed by content hash, and a hash
  // hit is taken only when the contents are equal
  struct Entry {
    int off;
    const std::vector<float> *F, *P;
  };
  std::multimap<uint64_t, Entry> seen;
  std::vector<float> all, allp;
  auto offset_of = [&](const TableBuf &tb, int p) -> int {
    auto range = seen.equal_range(tb.phash[p]);
    for (auto q = range.first; q != range.second; ++q)
      if (*q->second.F == tb.F32[p] && *q->second.P == tb.P32[p]) return q->second.off;
    const int o = (int)all.size();
    seen.emplace(tb.phash[p], Entry{o, &tb.F32[p], &tb.P32[p]});
    all.insert(all.end(), tb.F32[p].begin(), tb.F32[p].end());
    allp.insert(allp.end(), tb.P32[p].begin(), tb.P32[p].end());
    return o;
  };
  std::map<std::pair<uint64_t, int>, int> off;   // (table key, pair) -> offset
  for (auto &s : e->slots) {
    if (!s.used || !s.table_key) continue;
    auto it = e->tables.find(s.table_key);
    if (it == e->tables.end()) continue;
    for (int p = 0; p < s.npairs; ++p)
      if (!off.count({s.table_key, p})) off[{s.table_key, p}] = offset_of(it->second, p);
  }
  const bool fits = !all.empty() && all.size() * sizeof(float) <= (size_t)kLdsTableBytes;
  const bool fits_p = fits && 2 * all.size() * sizeof(float) <= (size_t)kLdsTableBytesFull;
  all.insert(all.end(), allp.begin(), allp.end());
  // a slot's parameters are uploaded again only when its table offsets move (admitting a
  // job next to many resident ones must not cost one upload per resident job)
  std::vector<int> changed;
  for (size_t r = 0; r < e->slots.size(); ++r) {
    Slot &s = e->slots[r];
    int lt[16];
    for (int p = 0; p < 16; ++p) lt[p] = -1;
    if (s.used && s.table_key && fits)
      for (int p = 0; p < s.npairs; ++p) lt[p] = off[{s.table_key, p}];
    if (std::memcmp(lt, s.P.ltab, sizeof(lt)) != 0) {
      std::memcpy(s.P.ltab, lt, sizeof(lt));
      changed.push_back((int)r);
    }
  }
  if (fits && all == e->tabset_host) {
    // the staged table set is already on the device
  } else if (fits) {
    e->tabset_host = all;
    if (all.size() > e->tabset_cap) {
      if (e->d_tabset) {
        HIPCHK(hipStreamSynchronize(e->stream));
        HIPCHK(hipFree(e->d_tabset));
        e->d_tabset = nullptr;
        e->tabset_cap = 0;
      }
      HIPCHK(hipMalloc(&e->d_tabset, sizeof(float) * all.size()));
      e->tabset_cap = all.size();
    }
    HIPCHK(hipMemcpyAsync(e->d_tabset, all.data(), sizeof(float) * all.size(), hipMemcpyHostToDevice, e->stream));
  }
  if (!fits) e->tabset_host.clear();
  e->tabset_n = fits ? (int)(all.size() / 2) : 0;
  e->tabset_p = fits_p ? 1 : 0;
  for (int r : changed)
    if (e->slots[r].used && e->slots[r].block)
      HIPCHK(hipMemcpyAsync(e->d_reps + r, &e->slots[r].P, sizeof(RepParams), hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return NDG_OK;
}

static int rebuild_tiles(ndg_engine *e) {
  if (!e->tiles_dirty) return NDG_OK;
  int rc0 = rebuild_tabset(e);
  if (rc0) return rc0;
  e->h_tiles.clear();
  e->tpa_ranges.clear();
  e->tpa_counts.clear();
  const int tpas[4] = {1, 2, 4, 8};
  for (int q = 0; q < 4; ++q) {
    int t0 = (int)e->h_tiles.size();
    for (size_t r = 0; r < e->slots.size(); ++r) {
      Slot &s = e->slots[r];
      if (!s.used || s.tpa != tpas[q]) continue;
      s.tile_base = (int)e->h_tiles.size();
      for (int lt = 0; lt < s.ntiles; ++lt) e->h_tiles.push_back(Tile{(int)r, lt * kTile, lt, 0});
    }
    int nt = (int)e->h_tiles.size() - t0;
    if (nt > 0) {
      e->tpa_ranges.push_back({tpas[q], t0});
      e->tpa_counts.push_back(nt);
    }
  }
  int nt = (int)e->h_tiles.size();
  if (nt > e->cap_tiles) {
    if (e->d_tiles) {
      HIPCHK(hipStreamSynchronize(e->stream));
      HIPCHK(hipFree(e->d_tiles));
      HIPCHK(hipFree(e->d_rb));
      e->d_tiles = nullptr;
      e->d_rb = nullptr;
    }
    e->cap_tiles = std::max(nt, 2 * e->cap_tiles);
    HIPCHK(hipMalloc(&e->d_tiles, sizeof(Tile) * e->cap_tiles));
    HIPCHK(hipMalloc(&e->d_rb, sizeof(int) * (e->cap_tiles + 1)));
  }
  const int ns = (int)e->slots.size();
  if (ns > e->cap_tile0) {
    if (e->d_tile0) {
      HIPCHK(hipStreamSynchronize(e->stream));
      HIPCHK(hipFree(e->d_tile0));
      e->d_tile0 = nullptr;
    }
    e->cap_tile0 = std::max(ns, 2 * e->cap_tile0);
    HIPCHK(hipMalloc(&e->d_tile0, sizeof(int) * e->cap_tile0));
  }
  std::vector<int> tile0(ns, -1);
  e->max_rep_tiles = 0;
  for (int r = 0; r < ns; ++r)
    if (e->slots[r].used) {
      tile0[r] = e->slots[r].tile_base;
      e->max_rep_tiles = std::max(e->max_rep_tiles, e->slots[r].ntiles);
    }
  if (ns > 0)
    HIPCHK(hipMemcpyAsync(e->d_tile0, tile0.data(), sizeof(int) * ns, hipMemcpyHostToDevice, e->stream));
  if (nt > 0)
    HIPCHK(hipMemcpyAsync(e->d_tiles, e->h_tiles.data(), sizeof(Tile) * nt, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  e->tiles_dirty = false;
  return drop_graphs(e);
}

static EngineDev dev_view(ndg_engine *e) {
  EngineDev E;
  E.reps = e->d_reps;
  E.st = e->d_st;
  E.tiles = e->d_tiles;
  E.tile0 = e->d_tile0;
  E.rb = e->d_rb;
  E.max_rep_tiles = e->max_rep_tiles;
  E.ntiles = (int)e->h_tiles.size();
  E.nreps = (int)e->slots.size();
  E.tabset = e->d_tabset;
  E.tabset_n = e->tabset_n;
  E.tabset_p = e->tabset_p;
  E.springs = 0;
  double list_bytes = 0.0;
  int nused = 0, min_tiles = 1 << 30;
  for (const Slot &sl : e->slots) {
    if (sl.used && sl.P.spring_mask) E.springs = 1;
    if (sl.used) {
      list_bytes += 4.0 * sl.ntiles * sl.P.list_chunks * kTile * sl.tpa;
      ++nused;
      min_tiles = std::min(min_tiles, sl.ntiles);
    }
  }
  E.list_big = nused > 0 && min_tiles >= kBigTiles ? 1 : 0;
  // non-temporal list loads measured: c3 (1.5 GB of lists) +1.4 %, c2 +2.9 %, and c1
  // (0.1 GB, which the cache keeps between steps) -4 %
  E.list_nt = list_bytes > kListCacheBytes ? 1 : 0;
  return E;
}

static int upload_params(ndg_engine *e, int r) {
  HIPCHK(hipMemcpyAsync(e->d_reps + r, &e->slots[r].P, sizeof(RepParams), hipMemcpyHostToDevice,
                        e->stream));
  return NDG_OK;
}

static int get_state_dev(ndg_engine *e, int r, RepState *S) {
  HIPCHK(hipMemcpyAsync(S, e->d_st + r, sizeof(RepState), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return NDG_OK;
}

static int put_state_dev(ndg_engine *e, int r, const RepState *S) {
  HIPCHK(hipMemcpyAsync(e->d_st + r, S, sizeof(RepState), hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return NDG_OK;
}

// Tables: dedupe identical tables between replicas.
static int attach_tables(ndg_engine *e, Slot &s, const HostState &h) {
  const size_t n = (size_t)h.npairs * (h.ptsize + 1);
  uint64_t key = hash_words(h.P.data(), n * 8);
  key = hash_words(h.F.data(), n * 8, key);
  key = fnv1a(&h.npairs, sizeof(int), key);
  key = fnv1a(&h.ptsize, sizeof(int), key);
  if (key == 0) key = 1;   // 0 marks a slot without tables
  auto it = e->tables.find(key);
  // a key is taken as a match only with equal contents; a different table under the same
  // key (a hash collision) moves on to the next key
  while (it != e->tables.end() &&
         !(it->second.npairs == h.npairs && it->second.ptsize == h.ptsize &&
           it->second.P == h.P && it->second.F == h.F)) {
    key = key + 1 ? key + 1 : 1;
    it = e->tables.find(key);
  }
  if (it == e->tables.end()) {
    TableBuf tb;
    tb.npairs = h.npairs;
    tb.ptsize = h.ptsize;
    tb.P = h.P;
    tb.F = h.F;
    tb.bytes = table_bytes(h.npairs, h.ptsize);
    int rc = dev_alloc(e, tb.bytes, &tb.block);
    if (rc) return rc;
    char *b = (char *)tb.block;
    tb.F64 = (double *)b;
    tb.P64 = (double *)(b + align_up(n * 8, 256));
    tb.F2 = (float2 *)(b + 2 * align_up(n * 8, 256));
    tb.P2 = (float2 *)(b + 3 * align_up(n * 8, 256));
    // FP32 interpolation tables {T_k, T_k+1}: one 8-byte load per lookup; the kernel forms
    // T_k+1 - T_k in FP32 exactly as from the LDS copy, so both give identical forces
    std::vector<float2> f2(n), p2(n);
    const int tab = h.ptsize + 1;
    for (int p = 0; p < h.npairs; ++p)
      for (int k = 0; k < tab; ++k) {
        const size_t q = (size_t)p * tab + k;
        const size_t q1 = k + 1 < tab ? q + 1 : q;
        f2[q] = make_float2((float)h.F[q], (float)h.F[q1]);
        p2[q] = make_float2((float)h.P[q], (float)h.P[q1]);
      }
    HIPCHK(hipMemcpyAsync(tb.F64, h.F.data(), n * 8, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(tb.P64, h.P.data(), n * 8, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(tb.F2, f2.data(), n * 8, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(tb.P2, p2.data(), n * 8, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    tb.F32.resize(h.npairs);
    tb.P32.resize(h.npairs);
    tb.phash.resize(h.npairs);
    for (int p = 0; p < h.npairs; ++p) {
      tb.F32[p].resize(tab);
      tb.P32[p].resize(tab);
      for (int k = 0; k < tab; ++k) {
        tb.F32[p][k] = (float)h.F[(size_t)p * tab + k];
        tb.P32[p][k] = (float)h.P[(size_t)p * tab + k];
      }
      tb.phash[p] = fnv1a(tb.F32[p].data(), 4 * (size_t)tab) ^
                    (fnv1a(tb.P32[p].data(), 4 * (size_t)tab) * 0x100000001b3ull);
    }
    it = e->tables.emplace(key, tb).first;
  }
  it->second.refs += 1;
  s.table_key = key;
  s.P.tabF64 = it->second.F64;
  s.P.tabP64 = it->second.P64;
  s.P.tabF2 = it->second.F2;
  s.P.tabP2 = it->second.P2;
  return NDG_OK;
}

static void detach_tables(ndg_engine *e, Slot &s) {
  auto it = e->tables.find(s.table_key);
  if (it == e->tables.end()) return;
  if (--it->second.refs == 0) {
    dev_release(e, it->second.block, it->second.bytes);
    e->tables.erase(it);
  }
}

static void set_run_params(Slot &s, const ndg_run_params &rp, int step0) {
  RepParams &P = s.P;
  s.rp = rp;
  P.run_step0 = step0;
  P.run_nsteps = rp.nsteps;
  P.out_steps = rp.out_steps > 0 ? rp.out_steps : (rp.nsteps > 0 ? rp.nsteps : 1);
  P.renei = rp.renei;
  P.renei_freq = rp.renei_freq > 0 ? rp.renei_freq : 1;
  P.trigger = rp.trigger;
  P.isokinetic = rp.isokinetic;
  P.ke_freq = rp.ke_freq > 0 ? rp.ke_freq : 1;
  P.isobaric = rp.isobaric;
  P.p_freq = rp.p_freq > 0 ? rp.p_freq : 1;
  P.isoenergetic = rp.isoenergetic;
  P.e_freq = rp.e_freq > 0 ? rp.e_freq : 1;
  P.box_load = rp.box_load;
  P.dt = rp.dt;
  P.ke_target = rp.ke_target;
  P.ke_tau = rp.ke_tau;
  for (int a = 0; a < 3; ++a) {
    P.p_target[a] = rp.p_target[a];
    P.box_vel[a] = rp.box_vel[a];
  }
  P.p_tau = rp.p_tau;
  P.e_tau = rp.e_tau;
  s.step = step0;
  s.run_step0 = step0;
  s.run_nsteps = rp.nsteps;
  s.finished = rp.nsteps <= 0;
}

// Upload caller data into a slot's block (sorted order == caller order until the first
// build; the build sorts and keeps oid).
static int upload_host(ndg_engine *e, Slot &s, const HostState &h, bool state_too) {
  RepParams &P = s.P;
  const int np = h.np;
  std::vector<double> rcp(h.npairs);
  std::vector<float> rc2f(h.npairs), vdr2(h.npairs);
  for (int p = 0; p < h.npairs; ++p) {
    rcp[p] = h.pairparams[4 * p];
    rc2f[p] = (float)h.rc2[p];
    vdr2[p] = (float)(1.0 / h.dr2[p]);
  }
  double rcut = 0.0;
  for (double r : rcp) rcut = std::max(rcut, r);
  P.rcut = rcut;
  P.ptrmin = h.ptrmin;
  P.asym = 0;
  for (int a = 0; a < h.ntypes; ++a)
    for (int b = 0; b < h.ntypes; ++b)
      if (h.pairmat[a + b * h.ntypes] != h.pairmat[b + a * h.ntypes]) P.asym = 1;
  // several pairs that are the same potential (the binary alloy's A-A, A-B and B-B) and
  // cover every type pair: the force kernel's single-pair path gives the same bits
  P.uniform = 1;
  for (int v : h.pairmat)
    if (v <= 0) P.uniform = 0;
  const size_t tab = (size_t)h.ptsize + 1;
  for (int p = 1; p < h.npairs && P.uniform; ++p)
    if (h.rc2[p] != h.rc2[0] || h.dr2[p] != h.dr2[0] ||
        h.pairparams[4 * p] != h.pairparams[0] ||
        std::memcmp(&h.F[p * tab], &h.F[0], 8 * tab) != 0 ||
        std::memcmp(&h.P[p * tab], &h.P[0], 8 * tab) != 0)
      P.uniform = 0;
  // displacement boundary conditions
  P.nbc = P.hbc ? (int)h.bc_dir.size() : 0;
  for (int b = 0; b < P.nbc; ++b) {
    P.bc_dir[b] = h.bc_dir[b];
    for (int a = 0; a < 3; ++a) P.bc_du[b][a] = h.bc_du[3 * b + a];
  }
  // spring pairs (compute2/3): C, strain limit and Vi from PairParams(2:4)
  P.spring_mask = 0;
  for (int p = 0; p < (int)h.kinds.size() && p < 16; ++p) {
    if (h.kinds[p] != NDG_SPRING || !P.ipx) continue;
    P.spring_mask |= 1 << p;
    P.spr_c[p] = h.pairparams[4 * p + 1];
    P.spr_s0[p] = h.pairparams[4 * p + 2];
    P.spr_vi[p] = h.pairparams[4 * p + 3];
  }
  hipStream_t st = e->stream;
  HIPCHK(hipMemcpyAsync((void *)P.mass, h.mass.data(), 8 * h.ntypes, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void *)P.pairmat, h.pairmat.data(), 4 * h.ntypes * h.ntypes, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void *)P.rc_pair, rcp.data(), 8 * h.npairs, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void *)P.rc2d, h.rc2.data(), 8 * h.npairs, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void *)P.rc2f, rc2f.data(), 4 * h.npairs, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void *)P.vdr2, vdr2.data(), 4 * h.npairs, hipMemcpyHostToDevice, st));
  if (state_too) {
    // the per-atom state is laid out in one pinned staging area, so its copies run
    // asynchronously and are waited for once, below.  It goes up in caller order, so the
    // initial positions and boundary conditions go up as given.
    const bool ip = P.ipx && h.have_ipos;
    const size_t nd = (size_t)np * 8;
    size_t off[8], tot = 0;
    auto seg = [&](int k, size_t bytes) { off[k] = tot; tot = align_up(tot + bytes, 256); };
    seg(0, 3 * nd);                    // pos  (SoA)
    seg(1, 3 * nd);                    // vel
    seg(2, 3 * nd);                    // acc
    seg(3, ip ? 3 * nd : 0);           // initial positions
    seg(4, 16 * (size_t)np);           // FP32 force (F = m a)
    seg(5, 4 * (size_t)np);            // types
    seg(6, 4 * (size_t)np);            // caller order
    seg(7, P.hbc ? 4 * (size_t)np : 0);
    if (tot > e->stage_cap) {
      HIPCHK(hipStreamSynchronize(st));
      if (e->stage_h) HIPCHK(hipHostFree(e->stage_h));
      e->stage_h = nullptr;
      e->stage_cap = 0;
      HIPCHK(hipHostMalloc((void **)&e->stage_h, tot, hipHostMallocDefault));
      e->stage_cap = tot;
    }
    char *b = e->stage_h;
    auto put3 = [&](int k, const std::vector<double> &aos, double *dx, double *dy, double *dz) -> int {
      double *x = (double *)(b + off[k]);
      for (int i = 0; i < np; ++i) {
        x[i] = aos.empty() ? 0.0 : aos[3 * i];
        x[np + i] = aos.empty() ? 0.0 : aos[3 * i + 1];
        x[2 * np + i] = aos.empty() ? 0.0 : aos[3 * i + 2];
      }
      HIPCHK(hipMemcpyAsync(dx, x, nd, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(dy, x + np, nd, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(dz, x + 2 * np, nd, hipMemcpyHostToDevice, st));
      return NDG_OK;
    };
    int rc;
    if ((rc = put3(0, h.pos, P.px, P.py, P.pz))) return rc;
    if ((rc = put3(1, h.vel, P.vx, P.vy, P.vz))) return rc;
    if ((rc = put3(2, h.acc, P.ax, P.ay, P.az))) return rc;
    if (ip) {
      if ((rc = put3(3, h.ipos, P.ipx, P.ipy, P.ipz))) return rc;
      s.ipos_dirty = false;
    }
    // mixed path keeps the latest force (F = m a) in FP32
    float4 *f4 = (float4 *)(b + off[4]);
    for (int i = 0; i < np; ++i) {
      const double m = h.mass[h.typ[i] - 1];
      if (h.acc.empty())
        f4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      else
        f4[i] = make_float4((float)(h.acc[3 * i] * m), (float)(h.acc[3 * i + 1] * m),
                            (float)(h.acc[3 * i + 2] * m), 0.f);
    }
    int *ty = (int *)(b + off[5]), *oid = (int *)(b + off[6]);
    for (int i = 0; i < np; ++i) {
      ty[i] = h.typ[i];
      oid[i] = i;
    }
    HIPCHK(hipMemcpyAsync(P.f32a, f4, 16 * np, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(P.f32b, f4, 16 * np, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(P.typ, ty, 4 * np, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(P.oid, oid, 4 * np, hipMemcpyHostToDevice, st));
    if (P.hbc) {
      int *hb = (int *)(b + off[7]);
      std::copy(h.hbc.begin(), h.hbc.end(), hb);
      HIPCHK(hipMemcpyAsync(P.hbc, hb, 4 * (size_t)np, hipMemcpyHostToDevice, st));
      s.hbc_dirty = false;
    }
    HIPCHK(hipMemsetAsync(P.fx, 0, 8 * np, st));
    HIPCHK(hipMemsetAsync(P.fy, 0, 8 * np, st));
    HIPCHK(hipMemsetAsync(P.fz, 0, 8 * np, st));
    HIPCHK(hipMemsetAsync(P.pene, 0, 8 * np, st));
    HIPCHK(hipMemsetAsync(P.nnb, 0, 4 * np, st));
    HIPCHK(hipMemsetAsync(P.cell_count, 0, 4 * s.ncell_cap, st));
    HIPCHK(hipMemsetAsync(P.cell_fill, 0, 4 * s.ncell_cap, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return NDG_OK;
}

// Download the slot's state in caller order.
static int download_state(ndg_engine *e, Slot &s, double *pos, double *vel, double *acc,
                          double *force, double *pene, double box[3]) {
  const int np = s.np;
  RepParams &P = s.P;
  hipStream_t st = e->stream;
  RepState S;
  int rc = get_state_dev(e, (int)(&s - &e->slots[0]), &S);
  if (rc) return rc;
  std::vector<int> oid(np), typ(np);
  HIPCHK(hipMemcpyAsync(oid.data(), P.oid, 4 * np, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(typ.data(), P.typ, 4 * np, hipMemcpyDeviceToHost, st));
  std::vector<double> buf(3 * (size_t)np);
  auto get3 = [&](double *dx, double *dy, double *dz, double *out) -> int {
    HIPCHK(hipMemcpyAsync(buf.data(), dx, 8 * np, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(buf.data() + np, dy, 8 * np, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(buf.data() + 2 * np, dz, 8 * np, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < np; ++i) {
      const int o = oid[i];
      out[3 * o] = buf[i];
      out[3 * o + 1] = buf[np + i];
      out[3 * o + 2] = buf[2 * np + i];
    }
    return NDG_OK;
  };
  HIPCHK(hipStreamSynchronize(st));
  if (pos && (rc = get3(P.px, P.py, P.pz, pos))) return rc;
  if (vel && (rc = get3(P.vx, P.vy, P.vz, vel))) return rc;
  if (P.path == NDG_PATH_FP64) {
    if (acc && (rc = get3(P.ax, P.ay, P.az, acc))) return rc;
    if (force && (rc = get3(P.fx, P.fy, P.fz, force))) return rc;
    if (pene) {
      HIPCHK(hipMemcpyAsync(buf.data(), P.pene, 8 * np, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      for (int i = 0; i < np; ++i) pene[oid[i]] = buf[i];
    }
  } else if (acc || force || pene) {
    std::vector<float4> f4(np);
    // latest forces: the buffer a force evaluation wrote if no kick has consumed it yet
    const int latest = S.parity ^ S.fnew;
    HIPCHK(hipMemcpyAsync(f4.data(), latest ? P.f32b : P.f32a, 16 * np, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<double> mass(s.ntypes);
    HIPCHK(hipMemcpy(mass.data(), P.mass, 8 * s.ntypes, hipMemcpyDeviceToHost));
    for (int i = 0; i < np; ++i) {
      const int o = oid[i];
      const double rm = 1.0 / mass[typ[i] - 1];
      if (force) {
        force[3 * o] = f4[i].x;
        force[3 * o + 1] = f4[i].y;
        force[3 * o + 2] = f4[i].z;
      }
      if (acc) {
        acc[3 * o] = (double)f4[i].x * rm;
        acc[3 * o + 1] = (double)f4[i].y * rm;
        acc[3 * o + 2] = (double)f4[i].z * rm;
      }
      if (pene) pene[o] = f4[i].w;
    }
  }
  if (box)
    for (int a = 0; a < 3; ++a) box[a] = S.box[a];
  return NDG_OK;
}

// Allocate (or grow) a slot's block for the given list capacity and cell grid.
// Copy a replica's state and per-type data from its old block to a new one (device to
// device); lists and cell arrays are rebuilt afterwards.
static int migrate_block(ndg_engine *e, const RepParams &o, const RepParams &n, int np,
                         int ntypes, int npairs) {
  hipStream_t st = e->stream;
  const size_t d = sizeof(double) * (size_t)np;
  const void *src[] = {o.px, o.py, o.pz, o.vx, o.vy, o.vz, o.ax, o.ay, o.az, o.fx, o.fy, o.fz,
                       o.pene, o.pbx, o.pby, o.pbz};
  void *dst[] = {n.px, n.py, n.pz, n.vx, n.vy, n.vz, n.ax, n.ay, n.az, n.fx, n.fy, n.fz,
                 n.pene, n.pbx, n.pby, n.pbz};
  for (int q = 0; q < 16; ++q) HIPCHK(hipMemcpyAsync(dst[q], src[q], d, hipMemcpyDeviceToDevice, st));
  if (o.hbc && n.hbc) HIPCHK(hipMemcpyAsync(n.hbc, o.hbc, 4 * (size_t)np, hipMemcpyDeviceToDevice, st));
  if (o.ipx && n.ipx) {
    HIPCHK(hipMemcpyAsync(n.ipx, o.ipx, d, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(n.ipy, o.ipy, d, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(n.ipz, o.ipz, d, hipMemcpyDeviceToDevice, st));
  }
  HIPCHK(hipMemcpyAsync(n.f32a, o.f32a, 16 * (size_t)np, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(n.f32b, o.f32b, 16 * (size_t)np, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(n.x64a, o.x64a, 32 * (size_t)np, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(n.xq, o.xq, 16 * (size_t)np, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(n.x32b, o.x32b, 16 * (size_t)np, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(n.anchor, o.anchor, 16 * (size_t)np, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(n.typ, o.typ, 4 * (size_t)np, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(n.oid, o.oid, 4 * (size_t)np, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync((void *)n.mass, o.mass, 8 * ntypes, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync((void *)n.pairmat, o.pairmat, 4 * ntypes * ntypes, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync((void *)n.rc_pair, o.rc_pair, 8 * npairs, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync((void *)n.rc2d, o.rc2d, 8 * npairs, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync((void *)n.rc2f, o.rc2f, 4 * npairs, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync((void *)n.vdr2, o.vdr2, 4 * npairs, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(n.ring, o.ring, sizeof(ndg_thermo) * kRing, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemsetAsync(n.cell_count, 0, 4 * (size_t)n.ncell_cap, st));
  HIPCHK(hipMemsetAsync(n.cell_fill, 0, 4 * (size_t)n.ncell_cap, st));
  HIPCHK(hipMemsetAsync(n.nnb, 0, 4 * (size_t)np, st));
  HIPCHK(hipStreamSynchronize(st));
  return NDG_OK;
}

static int alloc_slot(ndg_engine *e, int r, int neimax, int ncell_cap, int tpa,
                      bool migrate = false) {
  Slot &s = e->slots[r];
  Slot probe = s;
  probe.neimax = neimax;
  probe.ncell_cap = ncell_cap;
  probe.tpa = tpa;
  const size_t need = carve(probe, e->path, nullptr);
  void *blk;
  int rc = dev_alloc(e, need, &blk);
  if (rc) return rc;
  const RepParams keep = s.P;
  void *old_block = s.block;
  const size_t old_bytes = s.bytes;
  s.neimax = neimax;
  s.ncell_cap = ncell_cap;
  s.tpa = tpa;
  s.ntiles = (s.np + kTile - 1) / kTile;
  s.block = blk;
  s.bytes = need;
  carve(s, e->path, (char *)blk);
  // keep the table pointers and scalars that are not part of the block
  s.P.tabF64 = keep.tabF64;
  s.P.tabP64 = keep.tabP64;
  s.P.tabF2 = keep.tabF2;
  s.P.tabP2 = keep.tabP2;
  if (old_block) {
    HIPCHK(hipStreamSynchronize(e->stream));
    if (migrate && (rc = migrate_block(e, keep, s.P, s.np, s.ntypes, s.npairs))) return rc;
    dev_release(e, old_block, old_bytes);
  }
  e->tiles_dirty = true;
  return NDG_OK;
}

// Launch a full rebuild of one slot now (run begin / drop-in build).
static int build_slot_now(ndg_engine *e, int r, double skin) {
  Slot &s = e->slots[r];
  int rc = rebuild_tiles(e);
  if (rc) return rc;
  RepState S;
  if ((rc = get_state_dev(e, r, &S))) return rc;
  int m[3], brute;
  grid_for_box(S.box, s.P.rcut, skin, m, brute);
  if (!brute && (long long)m[0] * m[1] * m[2] > s.ncell_cap) return fail(NDG_ERR_ARG, "cell grid exceeds capacity");
  for (int a = 0; a < 3; ++a) S.mcell[a] = m[a];
  S.brute = brute;
  S.rebuild = 1;
  S.err = 0;
  s.P.skin = skin;
  if ((rc = upload_params(e, r))) return rc;
  if ((rc = put_state_dev(e, r, &S))) return rc;
  EngineDev E = dev_view(e);
  launch_rebuild(E, s.ntiles, r, 1, e->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  if ((rc = get_state_dev(e, r, &S))) return rc;
  S.rebuild = 0;
  // no displacement yet
  std::vector<double> part((size_t)s.ntiles * kPartials, 0.0);
  for (int lt = 0; lt < s.ntiles; ++lt) {
    part[(size_t)kPartD1 * s.ntiles + lt] = -1.0;
    part[(size_t)kPartD2 * s.ntiles + lt] = -1.0;
  }
  HIPCHK(hipMemcpyAsync(s.P.partials, part.data(), 8 * part.size(), hipMemcpyHostToDevice, e->stream));
  if ((rc = put_state_dev(e, r, &S))) return rc;
  s.nbr_sum = S.nbr_sum;
  if (S.err) return fail(S.err, err_text(S.err));
  return NDG_OK;
}

static int new_slot(ndg_engine *e) {
  for (size_t r = 0; r < e->slots.size(); ++r)
    if (!e->slots[r].used) {
      e->slots[r] = Slot();
      e->slots[r].used = true;
      return (int)r;
    }
  e->slots.push_back(Slot());
  e->slots.back().used = true;
  return (int)e->slots.size() - 1;
}

// ------------------------------------------------------------------ step scheduling

struct TickPlan {
  bool full = false, check = false, any = false;
};

static TickPlan plan_tick(ndg_engine *e, std::vector<int> *active) {
  TickPlan tp;
  for (size_t r = 0; r < e->slots.size(); ++r) {
    Slot &s = e->slots[r];
    if (!s.used || s.finished || s.err) continue;
    RepState S{};
    S.step = s.step;
    const int f = step_flags(s.P, S);
    if (!(f & F_ACTIVE)) continue;
    tp.any = true;
    if (f & (F_CU | F_CV)) tp.full = true;
    if (f & F_CHECK) tp.check = true;
    if (active) active->push_back((int)r);
  }
  return tp;
}

static void advance_mirror(ndg_engine *e) {
  for (auto &s : e->slots) {
    if (!s.used || s.finished || s.err) continue;
    s.step += 1;
    if (s.step - s.run_step0 >= s.run_nsteps) s.finished = true;
  }
}

static int time_event(ndg_engine *e, hipEvent_t *ev) {
  if (e->ev_pool.empty()) {
    HIPCHK(hipEventCreate(ev));
  } else {
    *ev = e->ev_pool.back();
    e->ev_pool.pop_back();
  }
  return NDG_OK;
}

static int time_begin(ndg_engine *e, hipEvent_t *a) {
  *a = nullptr;
  if (!e->timing) return NDG_OK;
  int rc = time_event(e, a);
  if (rc) return rc;
  HIPCHK(hipEventRecord(*a, e->stream));
  return NDG_OK;
}

static int time_end(ndg_engine *e, hipEvent_t a) {
  if (!e->timing || !a) return NDG_OK;
  hipEvent_t b;
  int rc = time_event(e, &b);
  if (rc) return rc;
  e->ev_pending.push_back({a, b});
  HIPCHK(hipEventRecord(b, e->stream));
  return NDG_OK;
}

static int time_collect(ndg_engine *e) {
  int rc = NDG_OK;
  for (auto &p : e->ev_pending) {
    float ms = 0.f;
    HIPKEEP(rc, hipEventSynchronize(p.second));
    HIPKEEP(rc, hipEventElapsedTime(&ms, p.first, p.second));
    e->t_ms += ms;
    e->t_count += 1;
    e->ev_pool.push_back(p.first);
    e->ev_pool.push_back(p.second);
  }
  e->ev_pending.clear();
  return rc;
}

// Kernels of one step for the batched (mixed) path: every active replica in one launch
// per kernel.
static int enqueue_step_batched(ndg_engine *e, const TickPlan &tp) {
  EngineDev E = dev_view(e);
  const int nt = (int)e->h_tiles.size(), nr = (int)e->slots.size();
  hipStream_t s = e->stream;
  if (tp.check) launch_rebuild(E, nt, 0, nr, s);
  launch_drift(E, e->path, 0, nt, 0, s);
  // mixed path: the kick runs inside the force kernel
  const int fused = e->path == NDG_PATH_MIXED;
  hipEvent_t a = nullptr;
  int rc = time_begin(e, &a);
  if (rc) return rc;
  for (size_t q = 0; q < e->tpa_ranges.size(); ++q) {
    launch_force(E, e->path, e->tpa_ranges[q].first, false, e->tpa_ranges[q].second,
                 e->tpa_counts[q], 0, s, fused);
    if (tp.full)
      launch_force(E, e->path, e->tpa_ranges[q].first, true, e->tpa_ranges[q].second,
                   e->tpa_counts[q], 0, s, fused);
  }
  if ((rc = time_end(e, a))) return rc;
  if (!fused) launch_kick(E, e->path, 0, nt, 0, s);
  launch_reduce(E, 0, nr, RED_FORCE | RED_KICK | RED_STEP, 0, 0, s);
  return NDG_OK;
}

// Plain path: each replica is advanced with its own launches.
static int enqueue_step_single(ndg_engine *e, int r) {
  EngineDev E = dev_view(e);
  Slot &sl = e->slots[r];
  hipStream_t s = e->stream;
  RepState S{};
  S.step = sl.step;
  const int f = step_flags(sl.P, S);
  if (!(f & F_ACTIVE)) return NDG_OK;
  if (f & F_CHECK) launch_rebuild(E, sl.ntiles, r, 1, s);
  launch_drift(E, e->path, sl.tile_base, sl.ntiles, 0, s);
  hipEvent_t a = nullptr;
  int rc = time_begin(e, &a);
  if (rc) return rc;
  launch_force(E, e->path, sl.tpa, false, sl.tile_base, sl.ntiles, 0, s);
  if (f & (F_CU | F_CV)) launch_force(E, e->path, sl.tpa, true, sl.tile_base, sl.ntiles, 0, s);
  if ((rc = time_end(e, a))) return rc;
  launch_kick(E, e->path, sl.tile_base, sl.ntiles, 0, s);
  launch_reduce(E, r, 1, RED_FORCE | RED_KICK | RED_STEP, 0, 0, s);
  return NDG_OK;
}

static std::string tick_signature(const std::vector<TickPlan> &plans) {
  std::string sig;
  for (auto &p : plans) sig.push_back((char)('0' + (p.full ? 1 : 0) + (p.check ? 2 : 0)));
  return sig;
}

// Build the lists of slot r; when the capacity was chosen by the engine, an overflow
// grows it by 1.5x (state migrated device to device) and builds again.
static int build_with_growth(ndg_engine *e, int r, double skin) {
  Slot &s = e->slots[r];
  int rc = build_slot_now(e, r, skin);
  for (int attempt = 0; rc == NDG_ERR_NEIMAX && s.auto_neimax && attempt < 6; ++attempt) {
    const int neimax = s.neimax + s.neimax / 2 + 8;
    if ((rc = alloc_slot(e, r, neimax, s.ncell_cap, s.tpa, true))) return rc;
    if ((rc = upload_params(e, r))) return rc;
    rc = build_slot_now(e, r, skin);
  }
  return rc;
}

// Drain device state and thermo rings after a chunk.
static int drain(ndg_engine *e) {
  const int nr = (int)e->slots.size();
  if (nr == 0) return NDG_OK;
  std::vector<RepState> S(nr);
  HIPCHK(hipMemcpyAsync(S.data(), e->d_st, sizeof(RepState) * nr, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  int first_err = 0;
  for (int r = 0; r < nr; ++r) {
    Slot &s = e->slots[r];
    if (!s.used) continue;
    s.nbr_sum = S[r].nbr_sum;
    if (S[r].err == NDG_ERR_NEIMAX && s.auto_neimax && !s.err) {
      // engine-chosen capacity: grow the list and rebuild at the current positions; the
      // replica then computes the step whose build failed
      int rc = build_with_growth(e, r, s.P.skin);
      if (rc) return rc;
      if ((rc = get_state_dev(e, r, &S[r]))) return rc;
    }
    if (S[r].err && !s.err) {
      s.err = S[r].err;
      if (!first_err) first_err = s.err;
    }
    const int have = S[r].nrec - s.nrec_seen;
    if (have > 0) {
      const int n = std::min(have, kRing);
      std::vector<ndg_thermo> buf(kRing);
      HIPCHK(hipMemcpy(buf.data(), s.P.ring, sizeof(ndg_thermo) * kRing, hipMemcpyDeviceToHost));
      for (int k = S[r].nrec - n; k < S[r].nrec; ++k) s.pending_rec.push_back(buf[k % kRing]);
      s.last = s.pending_rec.back();
      s.nrec_seen = S[r].nrec;
    }
    // the device is the authority on the step counter and the end of the Run
    s.step = S[r].step;
    s.finished = S[r].done != 0 || s.step - s.run_step0 >= s.run_nsteps;
  }
  int rc = time_collect(e);
  if (first_err) return fail(first_err, err_text(first_err));
  if (rc) return rc;
  return NDG_OK;
}

// Advance all slots by up to nsteps steps.
static int engine_steps(ndg_engine *e, int nsteps, bool drain_last = true) {
  int rc = rebuild_tiles(e);
  if (rc) return rc;
  int done_steps = 0;
  const int chunk_cap = kRing / 2;  // keep thermo ring from wrapping between drains
  while (done_steps < nsteps) {
    int chunk = std::min(nsteps - done_steps, chunk_cap);
    std::vector<TickPlan> plans;
    for (int k = 0; k < chunk; ++k) {
      TickPlan tp = plan_tick(e, nullptr);
      if (!tp.any) break;
      plans.push_back(tp);
      advance_mirror(e);
    }
    if (plans.empty()) break;
    const bool use_graph = e->path == NDG_PATH_MIXED && e->graph_steps > 0 && !e->timing;
    if (e->path == NDG_PATH_MIXED) {
      if (use_graph) {
        size_t k = 0;
        while (k < plans.size()) {
          size_t len = std::min(plans.size() - k, (size_t)e->graph_steps);
          std::vector<TickPlan> sub(plans.begin() + k, plans.begin() + k + len);
          std::string sig = tick_signature(sub);
          auto it = e->graphs.find(sig);
          if (it == e->graphs.end()) {
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            int rq = NDG_OK;
            for (auto &p : sub)
              if (!rq) rq = enqueue_step_batched(e, p);
            HIPCHK(hipStreamEndCapture(e->stream, &g));
            if (rq) {
              HIPKEEP(rq, hipGraphDestroy(g));
              return rq;
            }
            hipGraphExec_t ge;
            HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            HIPCHK(hipGraphDestroy(g));
            it = e->graphs.emplace(sig, ge).first;
          }
          HIPCHK(hipGraphLaunch(it->second, e->stream));
          k += len;
        }
      } else {
        for (auto &p : plans)
          if ((rc = enqueue_step_batched(e, p))) return rc;
      }
      HIPCHK(hipGetLastError());
    }
    done_steps += (int)plans.size();
    if (!drain_last && done_steps >= nsteps) break;   // the caller drains later
    if ((rc = drain(e))) return rc;
  }
  return NDG_OK;
}

// Plain FP64 path: one system per launch, replica after replica for each step.
static int engine_steps_plain(ndg_engine *e, int nsteps, bool drain_last = true) {
  int rc = rebuild_tiles(e);
  if (rc) return rc;
  int done_steps = 0;
  const int chunk_cap = kRing / 2;
  while (done_steps < nsteps) {
    int chunk = std::min(nsteps - done_steps, chunk_cap);
    int issued = 0;
    for (int k = 0; k < chunk; ++k) {
      bool any = false;
      for (size_t r = 0; r < e->slots.size(); ++r) {
        Slot &s = e->slots[r];
        if (!s.used || s.finished || s.err) continue;
        if ((rc = enqueue_step_single(e, (int)r))) return rc;
        any = true;
      }
      if (!any) break;
      advance_mirror(e);
      ++issued;
    }
    HIPCHK(hipGetLastError());
    if (issued == 0) break;
    done_steps += issued;
    if (!drain_last && done_steps >= nsteps) break;
    if ((rc = drain(e))) return rc;
  }
  return NDG_OK;
}

static int run_engine(ndg_engine *e, int nsteps, bool drain_last = true) {
  return e->path == NDG_PATH_FP64 ? engine_steps_plain(e, nsteps, drain_last)
                                  : engine_steps(e, nsteps, drain_last);
}

// Drain the steps ndg_engine_step_enqueue left on the stream (records, errors, step
// counters); every entry point that reads results or changes what is stepping calls it.
static int engine_settle(ndg_engine *e) {
  if (!e->queued) return NDG_OK;
  e->queued = false;
  return drain(e);
}

static size_t snapshot_doubles(int np) { return 10 * (size_t)np + 4; }

// Queue a caller-order copy of slot r's output state (snapshot_doubles(np) doubles) into the
// pinned host buffer `host` behind the work already on the stream, and record `ev` after it.
// The device staging buffer grows to the largest replica copied and is then reused (copies
// are ordered on the engine's stream).
static int snapshot_to(ndg_engine *e, int r, double *host, hipEvent_t ev) {
  Slot &s = e->slots[r];
  if (!s.block) return fail(NDG_ERR_ARG, "snapshot: the system has no device state yet");
  const size_t n = snapshot_doubles(s.np);
  if (n > e->snap_cap) {
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->snap_d) HIPCHK(hipFree(e->snap_d));
    e->snap_d = nullptr;
    e->snap_cap = 0;
    HIPCHK(hipMalloc((void **)&e->snap_d, 8 * n));
    e->snap_cap = n;
  }
  launch_snapshot(dev_view(e), r, s.np, e->snap_d, e->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(host, e->snap_d, 8 * n, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipEventRecord(ev, e->stream));
  return NDG_OK;
}

// The engine's own snapshot (ndg_snapshot, ndg_engine_fetch): one pinned buffer, read by
// snapshot_copy.
static int snapshot_enqueue(ndg_engine *e, int r) {
  const size_t n = snapshot_doubles(e->slots[r].np);
  if (n > e->snap_hcap) {
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->snap_h) HIPCHK(hipHostFree(e->snap_h));
    e->snap_h = nullptr;
    e->snap_hcap = 0;
    HIPCHK(hipHostMalloc((void **)&e->snap_h, 8 * n, hipHostMallocDefault));
    e->snap_hcap = n;
  }
  if (!e->snap_ev) HIPCHK(hipEventCreateWithFlags(&e->snap_ev, hipEventDisableTiming));
  int rc = snapshot_to(e, r, e->snap_h, e->snap_ev);
  if (rc) return rc;
  e->snap_slot = r;
  e->snap_np = e->slots[r].np;
  return NDG_OK;
}

static int snapshot_copy(ndg_engine *e, int r, double *pos, double *vel, double *force,
                         double *pene, double box[3], int *step) {
  if (e->snap_slot != r) return fail(NDG_ERR_ARG, "snapshot: none taken for this system");
  HIPCHK(hipEventSynchronize(e->snap_ev));
  const size_t np = (size_t)e->snap_np;
  const double *h = e->snap_h;
  if (pos) memcpy(pos, h, 24 * np);
  if (vel) memcpy(vel, h + 3 * np, 24 * np);
  if (force) memcpy(force, h + 6 * np, 24 * np);
  if (pene) memcpy(pene, h + 9 * np, 8 * np);
  if (box)
    for (int a = 0; a < 3; ++a) box[a] = h[10 * np + a];
  if (step) *step = (int)h[10 * np + 3];
  return NDG_OK;
}

// ------------------------------------------------------------------ engine lifecycle

static int engine_new(ndg_engine **pe, int path, size_t budget) {
  auto e = std::make_unique<ndg_engine>();
  e->path = path;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (ndev < 1) return fail(NDG_ERR_HIP, "no HIP device");
  HIPCHK(hipGetDevice(&e->device));
  HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  e->budget = budget ? budget : (size_t)(0.8 * (double)fr);
  *pe = e.release();
  return NDG_OK;
}

extern "C" int ndg_engine_create(ndg_engine_t **e, int path, size_t hbm_budget) {
  if (!e || (path != NDG_PATH_FP64 && path != NDG_PATH_MIXED))
    return fail(NDG_ERR_ARG, "ndg_engine_create: bad arguments");
  return engine_new(e, path, hbm_budget);
}

// Frees everything the engine holds; every release is attempted, and the first failure
// is returned (with its message) after the rest.
extern "C" int ndg_engine_destroy(ndg_engine_t *e) {
  if (!e) return NDG_OK;
  int rc = NDG_OK;
  HIPKEEP(rc, hipStreamSynchronize(e->stream));
  int rg = drop_graphs(e);
  if (rg && rc == NDG_OK) rc = rg;
  for (auto &s : e->slots)
    if (s.block) HIPKEEP(rc, hipFree(s.block));
  for (auto &t : e->tables) HIPKEEP(rc, hipFree(t.second.block));
  for (auto &fb : e->free_blocks) HIPKEEP(rc, hipFree(fb.second));
  for (auto &p : e->ev_pending) {
    HIPKEEP(rc, hipEventDestroy(p.first));
    HIPKEEP(rc, hipEventDestroy(p.second));
  }
  for (auto &ev : e->ev_pool) HIPKEEP(rc, hipEventDestroy(ev));
  void *dev[] = {e->d_reps, e->d_st, e->d_tiles, e->d_rb, e->d_tile0, e->d_tabset, e->snap_d};
  for (void *p : dev)
    if (p) HIPKEEP(rc, hipFree(p));
  if (e->snap_h) HIPKEEP(rc, hipHostFree(e->snap_h));
  if (e->snap_ev) HIPKEEP(rc, hipEventDestroy(e->snap_ev));
  if (e->stage_h) HIPKEEP(rc, hipHostFree(e->stage_h));
  HIPKEEP(rc, hipStreamDestroy(e->stream));
  delete e;
  return rc;
}

static int ncell_capacity(const double box[3], double rcut, double skin) {
  long long c = 1;
  for (int a = 0; a < 3; ++a) c *= (long long)(1.5 * box[a] / (rcut + skin)) + 1;
  return (int)std::min<long long>(c, 1 << 26);
}

extern "C" size_t ndg_footprint(int np, int ntypes, int npairs, int ptsize, int neimax) {
  Slot s;
  s.np = np;
  s.ntypes = ntypes;
  s.npairs = npairs;
  s.ptsize = ptsize;
  s.neimax = neimax;
  s.tpa = default_tpa(np);
  s.ncell_cap = np / 8 + 64;
  return carve(s, NDG_PATH_MIXED, nullptr) + table_bytes(npairs, ptsize);
}

static int check_host(const HostState &h) {
  if (h.np <= 0 || h.np > (int)kIndexMask) return fail(NDG_ERR_ARG, "np out of range");
  if (h.ntypes < 1 || h.ntypes > 8 || h.npairs < 1 || h.npairs > 16)
    return fail(NDG_ERR_ARG, "ntypes must be 1..8 and npairs 1..16");
  for (int t : h.typ)
    if (t < 1 || t > h.ntypes) return fail(NDG_ERR_ARG, "atom type out of range");
  for (int p : h.pairmat)
    if (p < 0 || p > h.npairs) return fail(NDG_ERR_ARG, "pair id out of range");
  // an asymmetric matrix selects the reference's compute1 rules (src/nd.f90:155-163)
  for (int a = 0; a < 3; ++a)
    if (!(h.box[a] > 0.0)) return fail(NDG_ERR_ARG, "box must be positive");
  return NDG_OK;
}

static int fill_host(HostState &h, int np, int ntypes, int npairs, int ptsize, const double *P,
                     const double *F, const double *rc2, const double *dr2, double ptrmin,
                     const int *pairmat, const double *pairparams, const double *mass) {
  const size_t n = (size_t)npairs * (ptsize + 1);
  h.np = np;
  h.ntypes = ntypes;
  h.npairs = npairs;
  h.ptsize = ptsize;
  h.P.assign(P, P + n);
  h.F.assign(F, F + n);
  h.rc2.assign(rc2, rc2 + npairs);
  h.dr2.assign(dr2, dr2 + npairs);
  h.ptrmin = ptrmin;
  h.pairmat.assign(pairmat, pairmat + ntypes * ntypes);
  h.pairparams.assign(pairparams, pairparams + 4 * npairs);
  h.mass.assign(mass, mass + ntypes);
  h.have_tables = true;
  return NDG_OK;
}

// Initial positions (caller order in h.ipos) to the device in the slot's current sorted
// order (atom i of the block is caller atom oid[i]).
static int upload_ipos_sorted(ndg_engine *e, Slot &s, const HostState &h) {
  const int np = s.np;
  hipStream_t st = e->stream;
  std::vector<int> oid(np);
  HIPCHK(hipMemcpyAsync(oid.data(), s.P.oid, 4 * (size_t)np, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<double> soa(3 * (size_t)np);
  for (int i = 0; i < np; ++i)
    for (int a = 0; a < 3; ++a) soa[(size_t)a * np + i] = h.ipos[3 * (size_t)oid[i] + a];
  double *dst[3] = {s.P.ipx, s.P.ipy, s.P.ipz};
  for (int a = 0; a < 3; ++a)
    HIPCHK(hipMemcpyAsync(dst[a], soa.data() + (size_t)a * np, 8 * (size_t)np,
                          hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  s.ipos_dirty = false;
  return NDG_OK;
}

// Boundary conditions (caller order in h.hbc) to the device in the slot's sorted order.
static int upload_hbc_sorted(ndg_engine *e, Slot &s, const HostState &h) {
  const int np = s.np;
  hipStream_t st = e->stream;
  std::vector<int> oid(np), v(np);
  HIPCHK(hipMemcpyAsync(oid.data(), s.P.oid, 4 * (size_t)np, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int i = 0; i < np; ++i) v[i] = h.hbc[oid[i]];
  HIPCHK(hipMemcpyAsync(s.P.hbc, v.data(), 4 * (size_t)np, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  s.hbc_dirty = false;
  return NDG_OK;
}

// Compact the hold groups with a displacement condition on some axis into the device
// table (at most 16) and the per-atom index into it; atoms of other groups integrate
// freely (src/nd.f90:1058-1064).
static int holds_to_host(HostState &hs, const int *hold, int nhold, const int *bc_disp,
                         const int *bc_dir, const double *bc_parm) {
  std::vector<int> index(nhold + 1, 0), dir;
  std::vector<double> du;
  for (int g = 1; g <= nhold; ++g) {
    const int d = (bc_dir[3 * (g - 1)] ? 1 : 0) | (bc_dir[3 * (g - 1) + 1] ? 2 : 0) |
                  (bc_dir[3 * (g - 1) + 2] ? 4 : 0);
    if (!bc_disp[g - 1] || !d) continue;
    if ((int)dir.size() == 16) return fail(NDG_ERR_ARG, "holds: at most 16 hold groups with Disp");
    dir.push_back(d);
    for (int a = 0; a < 3; ++a) du.push_back(bc_parm[3 * (g - 1) + a]);
    index[g] = (int)dir.size();
  }
  std::vector<int> hbc(hs.np, 0);
  if (!dir.empty())
    for (int i = 0; i < hs.np; ++i) {
      const int g = hold[i];
      if (g < 0 || g > nhold) return fail(NDG_ERR_ARG, "holds: hold id out of range");
      hbc[i] = index[g];
    }
  hs.hbc.swap(hbc);
  hs.bc_dir.swap(dir);
  hs.bc_du.swap(du);
  if (!hs.holds()) hs.hbc.clear();
  return NDG_OK;
}

// Make slot r resident for a Run: allocate for the list capacity, upload, build lists.
static int begin_run(ndg_engine *e, int r, HostState &h, const ndg_run_params &rp, int step0,
                     int neimax_req, bool fresh_state) {
  Slot &s = e->slots[r];
  double rcut = 0.0;
  for (int p = 0; p < h.npairs; ++p) rcut = std::max(rcut, h.pairparams[4 * p]);
  const double skin = rp.skin < 0.0 ? rcut * 0.2 : rp.skin;
  double box[3];
  if (fresh_state) {
    for (int a = 0; a < 3; ++a) box[a] = h.box[a];
  } else {
    RepState S;
    int rc = get_state_dev(e, r, &S);
    if (rc) return rc;
    for (int a = 0; a < 3; ++a) box[a] = S.box[a];
  }
  int neimax = neimax_req > 0 ? neimax_req : estimate_neimax(rcut, skin, h.np, box);
  if (neimax < 1) neimax = 1;
  s.auto_neimax = neimax_req <= 0;
  int tpa = e->tpa_override ? e->tpa_override : default_tpa(h.np);
  if (!e->tpa_override && e->path == NDG_PATH_MIXED) {
    for (const auto &c : e->tpa_classes)
      if (h.np <= c.first) {
        tpa = c.second;
        break;
      }
  }
  const int ncap = ncell_capacity(box, rcut, skin);
  // cell-block shape of the size class (1 x 1: the reference's linear cell order)
  s.P.cb_y = s.P.cb_z = 1;
  for (const auto &c : e->cb_classes)
    if (h.np <= c[0]) {
      s.P.cb_y = c[1];
      s.P.cb_z = c[2];
      break;
    }
  const bool springs = h.springs();
  if (springs && !h.have_ipos) {
    // initPositionSet at the start of dynamics when no initial positions exist yet
    // (src/nd.f90:103, :1731-1763): the Run's first positions and box
    if (fresh_state || !s.block) {
      h.ipos = h.pos;
    } else {
      h.ipos.assign(3 * (size_t)h.np, 0.0);
      int rc0 = download_state(e, s, h.ipos.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
      if (rc0) return rc0;
    }
    for (int a = 0; a < 3; ++a) h.ibox[a] = box[a];
    h.have_ipos = true;
    s.ipos_dirty = true;
  }
  const bool holds = h.holds();
  const bool need_alloc = !s.block || neimax > s.neimax || ncap > s.ncell_cap || tpa != s.tpa ||
                          springs != s.springs || holds != s.holds;
  int rc;
  if (need_alloc) {
    // keep the state across the reallocation when it is already on the device
    std::vector<double> pos, vel, acc;
    if (!fresh_state && s.block) {
      pos.resize(3 * (size_t)s.np);
      vel.resize(3 * (size_t)s.np);
      acc.resize(3 * (size_t)s.np);
      if ((rc = download_state(e, s, pos.data(), vel.data(), acc.data(), nullptr, nullptr, box))) return rc;
      h.pos = pos;
      h.vel = vel;
      h.acc = acc;
      for (int a = 0; a < 3; ++a) h.box[a] = box[a];
      fresh_state = true;
    }
    s.np = h.np;
    s.ntypes = h.ntypes;
    s.npairs = h.npairs;
    s.ptsize = h.ptsize;
    s.springs = springs;
    s.holds = holds;
    if ((rc = alloc_slot(e, r, neimax, ncap, tpa))) return rc;
    if ((rc = ensure_capacity(e))) return rc;
  }
  if (!s.P.tabF2) {
    if ((rc = attach_tables(e, s, h))) return rc;
  }
  s.P.skin = skin;
  set_run_params(s, rp, step0);
  if ((rc = upload_host(e, s, h, fresh_state))) return rc;
  if ((rc = upload_params(e, r))) return rc;
  if (springs && s.ipos_dirty && (rc = upload_ipos_sorted(e, s, h))) return rc;
  if (holds && s.hbc_dirty && (rc = upload_hbc_sorted(e, s, h))) return rc;
  RepState S;
  if (fresh_state) {
    memset(&S, 0, sizeof(S));
    for (int a = 0; a < 3; ++a) S.box[a] = S.box_build[a] = box[a];
    S.lam1 = S.lam2 = 1.0;
  } else if ((rc = get_state_dev(e, r, &S))) {
    return rc;
  }
  for (int a = 0; a < 3; ++a) S.ibox[a] = springs ? h.ibox[a] : 0.0;
  S.step = step0;
  S.done = rp.nsteps <= 0 ? 1 : 0;
  S.err = 0;
  S.nupdates = 0;
  S.nrec = 0;
  S.e0 = 0.0;
  s.nrec_seen = 0;
  s.pending_rec.clear();
  s.err = 0;
  if ((rc = put_state_dev(e, r, &S))) return rc;
  return build_with_growth(e, r, skin);
}

static int engine_admit(ndg_engine_t *e, int *slot, int np, int ntypes, int npairs, int ptsize,
                        int neimax, const double *P, const double *F, const double *rc2,
                        const double *dr2, double ptrmin, const int *pairmat,
                        const double *pairparams, const double *mass, const double *pos,
                        const double *vel, const int *typ, const double box[3],
                        const ndg_run_params *rp, const int *kinds, const double *ipos,
                        const double *init_box, const int *hold, int nhold,
                        const int *bc_disp, const int *bc_dir, const double *bc_parm) {
  if (!e || !slot || !rp || !pos || !typ || !box) return fail(NDG_ERR_ARG, "ndg_engine_admit: null argument");
  HostState h;
  fill_host(h, np, ntypes, npairs, ptsize, P, F, rc2, dr2, ptrmin, pairmat, pairparams, mass);
  h.pos.assign(pos, pos + 3 * (size_t)np);
  if (vel) h.vel.assign(vel, vel + 3 * (size_t)np);
  h.typ.assign(typ, typ + np);
  for (int a = 0; a < 3; ++a) h.box[a] = box[a];
  int rc = check_host(h);
  if (rc) return rc;
  if (kinds) {
    h.kinds.assign(kinds, kinds + npairs);
    for (int &v : h.kinds) {
      v &= ~NDG_ESHIFT;   // the energy shift is a table property only
      if (v < NDG_MORSE || v > NDG_LJ_C) return fail(NDG_ERR_ARG, "ndg_engine_admit_ex: unknown kind");
    }
    if (ipos) {
      // given initial positions; without them the Run's first state is taken, as
      // initPositionSet at the start of dynamics
      h.ipos.assign(ipos, ipos + 3 * (size_t)np);
      for (int a = 0; a < 3; ++a) h.ibox[a] = init_box[a];
      h.have_ipos = true;
    }
  }
  if (nhold > 0 && (rc = holds_to_host(h, hold, nhold, bc_disp, bc_dir, bc_parm))) return rc;
  if (e->path == NDG_PATH_FP64 && e->tpa_override > 1) e->tpa_override = 1;
  const int r = new_slot(e);
  rc = begin_run(e, r, h, *rp, 0, neimax, true);
  if (rc) {
    Slot &s = e->slots[r];
    if (s.block) dev_release(e, s.block, s.bytes);
    if (s.P.tabF2) detach_tables(e, s);
    s = Slot();
    e->tiles_dirty = true;
    return rc;
  }
  *slot = r;
  return NDG_OK;
}

extern "C" int ndg_engine_admit(ndg_engine_t *e, int *slot, int np, int ntypes, int npairs,
                                int ptsize, int neimax, const double *P, const double *F,
                                const double *rc2, const double *dr2, double ptrmin,
                                const int *pairmat, const double *pairparams,
                                const double *mass, const double *pos, const double *vel,
                                const int *typ, const double box[3],
                                const ndg_run_params *rp) {
  return engine_admit(e, slot, np, ntypes, npairs, ptsize, neimax, P, F, rc2, dr2, ptrmin,
                      pairmat, pairparams, mass, pos, vel, typ, box, rp, nullptr, nullptr,
                      nullptr, nullptr, 0, nullptr, nullptr, nullptr);
}

extern "C" int ndg_engine_admit_ex(ndg_engine_t *e, int *slot, int np, int ntypes, int npairs,
                                   int ptsize, int neimax, const double *P, const double *F,
                                   const double *rc2, const double *dr2, double ptrmin,
                                   const int *pairmat, const double *pairparams,
                                   const double *mass, const double *pos, const double *vel,
                                   const int *typ, const double box[3],
                                   const ndg_run_params *rp, const int *kinds,
                                   const double *ipos, const double init_box[3],
                                   const int *hold, int nhold, const int *bc_disp,
                                   const int *bc_dir, const double *bc_parm) {
  if ((ipos && (!kinds || !init_box)) || nhold < 0 ||
      (nhold > 0 && (!hold || !bc_disp || !bc_dir || !bc_parm)))
    return fail(NDG_ERR_ARG, "ndg_engine_admit_ex: bad arguments");
  return engine_admit(e, slot, np, ntypes, npairs, ptsize, neimax, P, F, rc2, dr2, ptrmin,
                      pairmat, pairparams, mass, pos, vel, typ, box, rp, kinds, ipos, init_box,
                      hold, nhold, bc_disp, bc_dir, bc_parm);
}

static int count_done(ndg_engine *e) {
  int d = 0;
  for (auto &s : e->slots)
    if (s.used && (s.finished || s.err)) ++d;
  return d;
}

extern "C" int ndg_engine_step(ndg_engine_t *e, int nsteps, int *ndone) {
  if (!e) return fail(NDG_ERR_ARG, "null engine");
  int rc = engine_settle(e);
  if (!rc) rc = run_engine(e, nsteps);
  if (ndone) *ndone = count_done(e);
  return rc;
}

extern "C" int ndg_engine_step_enqueue(ndg_engine_t *e, int nsteps) {
  if (!e || nsteps < 0 || nsteps > kRing / 2)
    return fail(NDG_ERR_ARG, "ndg_engine_step_enqueue: nsteps must be 0.." + std::to_string(kRing / 2));
  int rc = engine_settle(e);
  if (rc) return rc;
  if ((rc = run_engine(e, nsteps, false))) return rc;
  e->queued = true;
  return NDG_OK;
}

extern "C" int ndg_engine_step_wait(ndg_engine_t *e, int *ndone) {
  if (!e) return fail(NDG_ERR_ARG, "null engine");
  int rc = engine_settle(e);
  if (ndone) *ndone = count_done(e);
  return rc;
}

extern "C" int ndg_engine_status(ndg_engine_t *e, int slot, int *steps_done, int *finished,
                                 ndg_thermo *last, int *err) {
  if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].used)
    return fail(NDG_ERR_ARG, "bad slot");
  if (int rq = engine_settle(e)) return rq;
  Slot &s = e->slots[slot];
  if (steps_done) *steps_done = s.step - s.run_step0;
  if (finished) *finished = s.finished ? 1 : 0;
  if (last) *last = s.last;
  if (err) *err = s.err;
  return NDG_OK;
}

extern "C" int ndg_engine_thermo(ndg_engine_t *e, int slot, ndg_thermo *rec, int max_rec,
                                 int *nrec) {
  if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].used)
    return fail(NDG_ERR_ARG, "bad slot");
  if (int rq = engine_settle(e)) return rq;
  Slot &s = e->slots[slot];
  const int n = std::min((int)s.pending_rec.size(), max_rec);
  for (int k = 0; k < n; ++k) rec[k] = s.pending_rec[k];
  s.pending_rec.erase(s.pending_rec.begin(), s.pending_rec.begin() + n);
  if (nrec) *nrec = n;
  return NDG_OK;
}

extern "C" int ndg_engine_fetch(ndg_engine_t *e, int slot, double *pos, double *vel,
                                double *force, double box[3]) {
  if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].used)
    return fail(NDG_ERR_ARG, "bad slot");
  if (int rq = engine_settle(e)) return rq;
  EngineDev E = dev_view(e);
  launch_apply_pending(E, slot, 1, e->stream);
  HIPCHK(hipGetLastError());
  int rc = snapshot_enqueue(e, slot);
  if (rc) return rc;
  return snapshot_copy(e, slot, pos, vel, force, nullptr, box, nullptr);
}

// ------------------------------------------------------------------ fetch without allocation
// ndg_engine_fetch stages a replica's state in a device buffer and a pinned host buffer that
// grow (hipMalloc / hipHostMalloc) with the largest replica fetched.  Callers that may not
// allocate -- a serving loop that fetches finished jobs between ticks, a harness that forbids
// allocation -- use ndg_engine_fetch_noalloc instead: the staging is this file-scope device
// array, one per GPU (held until the process ends), copied straight into the caller's
// arrays.  Every call on a GPU shares that GPU's array: calls that overlap on one GPU (two
// host threads, two engines) race.  Nothing in it needs clearing: each call overwrites the
// 10 np + 4 doubles it reads back.
constexpr int kNoAllocMaxAtoms = 1 << 18;
constexpr size_t kNoAllocDoubles = 10 * (size_t)kNoAllocMaxAtoms + 4;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_ptr;   // device -> address of that device's `scratch`
}  // namespace

// Device address of `scratch` on the current device (looked up once per device; in host
// code the array's name is not a device address).
static int scratch_address(int dev, double **ptr) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) {
    *ptr = it->second;
    return NDG_OK;
  }
  void *p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(NDG_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
  g_scratch_ptr[dev] = (double *)p;
  *ptr = (double *)p;
  return NDG_OK;
}

static int fetch_noalloc(ndg_engine *e, int slot, double *pos, double *vel, double *force,
                         double box[3]) {
  const int np = e->slots[slot].np;
  if (np > kNoAllocMaxAtoms)
    return fail(NDG_ERR_ARG, "ndg_engine_fetch_noalloc: at most " +
                                 std::to_string(kNoAllocMaxAtoms) + " atoms per system (this one has " +
                                 std::to_string(np) + "); use ndg_engine_fetch");
  double *buf = nullptr;
  int rc = scratch_address(e->device, &buf);
  if (rc) return rc;
  EngineDev E = dev_view(e);
  launch_apply_pending(E, slot, 1, e->stream);
  launch_snapshot(E, slot, np, buf, e->stream);
  HIPCHK(hipGetLastError());
  const size_t n3 = 3 * (size_t)np;
  if (pos) HIPCHK(hipMemcpyAsync(pos, buf, 8 * n3, hipMemcpyDeviceToHost, e->stream));
  if (vel) HIPCHK(hipMemcpyAsync(vel, buf + n3, 8 * n3, hipMemcpyDeviceToHost, e->stream));
  if (force) HIPCHK(hipMemcpyAsync(force, buf + 2 * n3, 8 * n3, hipMemcpyDeviceToHost, e->stream));
  if (box) HIPCHK(hipMemcpyAsync(box, buf + 10 * (size_t)np, 24, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return NDG_OK;
}

extern "C" int ndg_engine_fetch_noalloc(ndg_engine_t *e, int slot, double *pos, double *vel,
                                        double *force, double box[3]) {
  if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].used ||
      !e->slots[slot].block)
    return fail(NDG_ERR_ARG, "bad slot");
  if (int rq = engine_settle(e)) return rq;
  int cur = 0;
  HIPCHK(hipGetDevice(&cur));
  if (cur != e->device) HIPCHK(hipSetDevice(e->device));
  int rc = fetch_noalloc(e, slot, pos, vel, force, box);
  if (cur != e->device) {
    hipError_t r = hipSetDevice(cur);
    if (r != hipSuccess && rc == NDG_OK)
      rc = fail(NDG_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(r));
  }
  return rc;
}

extern "C" int ndg_engine_release(ndg_engine_t *e, int slot) {
  if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].used)
    return fail(NDG_ERR_ARG, "bad slot");
  if (int rq = engine_settle(e)) return rq;
  HIPCHK(hipStreamSynchronize(e->stream));
  Slot &s = e->slots[slot];
  dev_release(e, s.block, s.bytes);
  detach_tables(e, s);
  s = Slot();
  RepState z;
  memset(&z, 0, sizeof(z));
  z.done = 1;
  int rc = put_state_dev(e, slot, &z);
  if (rc) return rc;
  // an unused slot keeps a zero-size parameter record
  RepParams zp;
  memset(&zp, 0, sizeof(zp));
  HIPCHK(hipMemcpyAsync(e->d_reps + slot, &zp, sizeof(zp), hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  e->tiles_dirty = true;
  return NDG_OK;
}

extern "C" int ndg_engine_config(ndg_engine_t *e, int tpa, int graph_steps) {
  if (!e || (tpa != 0 && tpa != 1 && tpa != 2 && tpa != 4 && tpa != 8))
    return fail(NDG_ERR_ARG, "tpa must be 0,1,2,4,8");
  if (int rq = engine_settle(e)) return rq;
  e->tpa_override = e->path == NDG_PATH_FP64 ? 1 : tpa;
  e->graph_steps = graph_steps < 0 ? 0 : graph_steps;
  return NDG_OK;
}

extern "C" int ndg_engine_size_classes(ndg_engine_t *e, int n, const int *np_max, const int *tpa) {
  if (!e || n < 0 || n > 8 || (n > 0 && (!np_max || !tpa)))
    return fail(NDG_ERR_ARG, "ndg_engine_size_classes: bad arguments");
  std::vector<std::pair<int, int>> c;
  for (int k = 0; k < n; ++k) {
    if (tpa[k] != 1 && tpa[k] != 2 && tpa[k] != 4 && tpa[k] != 8)
      return fail(NDG_ERR_ARG, "tpa must be 1,2,4,8");
    if (np_max[k] < 1 || (k > 0 && np_max[k] <= np_max[k - 1]))
      return fail(NDG_ERR_ARG, "np_max must be positive and ascending");
    c.push_back({np_max[k], tpa[k]});
  }
  e->tpa_classes = c;
  return NDG_OK;
}

extern "C" int ndg_engine_cell_blocks(ndg_engine_t *e, int n, const int *np_max, const int *by,
                                      const int *bz) {
  if (!e || n < 0 || n > 8 || (n > 0 && (!np_max || !by || !bz)))
    return fail(NDG_ERR_ARG, "ndg_engine_cell_blocks: bad arguments");
  std::vector<std::array<int, 3>> c;
  for (int k = 0; k < n; ++k) {
    if (by[k] < 1 || by[k] > 8 || bz[k] < 1 || bz[k] > 8)
      return fail(NDG_ERR_ARG, "cell-block sides must be 1..8");
    if (np_max[k] < 1 || (k > 0 && np_max[k] <= np_max[k - 1]))
      return fail(NDG_ERR_ARG, "np_max must be positive and ascending");
    c.push_back({np_max[k], by[k], bz[k]});
  }
  e->cb_classes = c;
  return NDG_OK;
}

extern "C" int ndg_engine_timing(ndg_engine_t *e, int enable) {
  if (!e) return fail(NDG_ERR_ARG, "null engine");
  e->timing = enable != 0;
  return NDG_OK;
}

extern "C" int ndg_engine_kernel_time(ndg_engine_t *e, int reset, double *avg_ms,
                                      long long *count, double *pairs, double *atoms) {
  if (!e) return fail(NDG_ERR_ARG, "null engine");
  if (int rq = engine_settle(e)) return rq;
  HIPCHK(hipStreamSynchronize(e->stream));
  int rc = time_collect(e);
  if (rc) return rc;
  if (avg_ms) *avg_ms = e->t_count ? e->t_ms / e->t_count : 0.0;
  if (count) *count = e->t_count;
  long long atoms_total = 0;
  double pairs_total = 0.0;
  for (auto &s : e->slots)
    if (s.used) {
      atoms_total += s.np;
      pairs_total += (double)s.nbr_sum;
    }
  if (pairs) *pairs = pairs_total;
  if (atoms) *atoms = (double)atoms_total;
  if (reset) {
    e->t_ms = 0.0;
    e->t_count = 0;
  }
  return NDG_OK;
}

extern "C" int ndg_engine_counters(ndg_engine_t *e, int slot, int *nupdates, int *step,
                                   double *mean_nbr) {
  if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].used)
    return fail(NDG_ERR_ARG, "bad slot");
  if (int rq = engine_settle(e)) return rq;
  RepState S;
  int rc = get_state_dev(e, slot, &S);
  if (rc) return rc;
  if (nupdates) *nupdates = S.nupdates;
  if (step) *step = S.step;
  if (mean_nbr) *mean_nbr = e->slots[slot].np ? (double)S.nbr_sum / e->slots[slot].np : 0.0;
  return NDG_OK;
}

extern "C" int ndg_engine_list_stats(ndg_engine_t *e, double *mean_nbr, long long *atoms) {
  if (!e) return fail(NDG_ERR_ARG, "null engine");
  if (int rq = engine_settle(e)) return rq;
  long long a = 0;
  double p = 0;
  for (auto &s : e->slots)
    if (s.used) {
      a += s.np;
      p += (double)s.nbr_sum;
    }
  if (mean_nbr) *mean_nbr = a ? p / a : 0.0;
  if (atoms) *atoms = a;
  return NDG_OK;
}

// ------------------------------------------------------------------ single-system API

static int settle(ndg_t *h);
static int frame_writer_stop(ndg_t *h);
static int upload_host(ndg_engine *e, Slot &s, const HostState &h, bool state_too);
static int upload_hbc_sorted(ndg_engine *e, Slot &s, const HostState &h);

extern "C" const char *ndg_last_error(void) { return g_err.c_str(); }

extern "C" int ndg_create(ndg_t **h, int np, int ntypes, int npairs, int ptsize, int neimax,
                          int flags) {
  if (!h || np <= 0 || ntypes <= 0 || npairs <= 0 || ptsize < 2)
    return fail(NDG_ERR_ARG, "ndg_create: bad arguments");
  const int path = flags & 1;
  auto sys = std::make_unique<ndg_system>();
  int rc = engine_new(&sys->e, path, 0);
  if (rc) return rc;
  sys->path = path;
  sys->hs.np = np;
  sys->hs.ntypes = ntypes;
  sys->hs.npairs = npairs;
  sys->hs.ptsize = ptsize;
  sys->neimax_req = neimax;
  sys->slot = new_slot(sys->e);
  Slot &s = sys->e->slots[sys->slot];
  s.np = np;
  s.ntypes = ntypes;
  s.npairs = npairs;
  s.ptsize = ptsize;
  *h = sys.release();
  return NDG_OK;
}

extern "C" int ndg_destroy(ndg_t *h) {
  if (!h) return NDG_OK;
  int rc = frame_writer_stop(h);   // writes the frames still queued
  int rd = ndg_engine_destroy(h->e);
  delete h;
  return rc ? rc : rd;
}

extern "C" int ndg_set_tables(ndg_t *h, const double *P, const double *F, const double *rc2,
                              const double *dr2, double ptrmin, const int *pairmat,
                              const double *pairparams, const double *mass) {
  if (!h || !P || !F || !rc2 || !dr2 || !pairmat || !pairparams || !mass)
    return fail(NDG_ERR_ARG, "ndg_set_tables: null argument");
  HostState &hs = h->hs;
  fill_host(hs, hs.np, hs.ntypes, hs.npairs, hs.ptsize, P, F, rc2, dr2, ptrmin, pairmat,
            pairparams, mass);
  Slot &s = h->e->slots[h->slot];
  if (s.P.tabF2) {  // re-attach new tables
    detach_tables(h->e, s);
    s.P.tabF2 = nullptr;
    int rc = attach_tables(h->e, s, hs);
    if (rc) return rc;
    h->e->tiles_dirty = true;  // the LDS table set changes
    if (s.block) {
      if ((rc = upload_host(h->e, s, hs, false))) return rc;
      return upload_params(h->e, h->slot);
    }
  }
  return NDG_OK;
}

extern "C" int ndg_set_pair_kinds(ndg_t *h, const int *kinds) {
  if (!h || !kinds) return fail(NDG_ERR_ARG, "ndg_set_pair_kinds: null argument");
  HostState &hs = h->hs;
  std::vector<int> k(kinds, kinds + hs.npairs);
  for (int &v : k) {
    v &= ~NDG_ESHIFT;   // the energy shift is a table property only
    if (v < NDG_MORSE || v > NDG_LJ_C) return fail(NDG_ERR_ARG, "ndg_set_pair_kinds: unknown kind");
  }
  if (k == hs.kinds) return NDG_OK;
  Slot &s = h->e->slots[h->slot];
  if (s.block)
    return fail(NDG_ERR_ARG, "ndg_set_pair_kinds: set the kinds before the first list build or Run");
  hs.kinds = k;
  return NDG_OK;
}

// Initial positions (caller order) and box to the host copy and, when resident with room
// for them, to the device in its sorted order.
static int put_initial_positions(ndg_t *h, const double *ipos, const double ibox[3]) {
  HostState &hs = h->hs;
  hs.ipos.assign(ipos, ipos + 3 * (size_t)hs.np);
  for (int a = 0; a < 3; ++a) hs.ibox[a] = ibox[a];
  hs.have_ipos = true;
  Slot &s = h->e->slots[h->slot];
  s.ipos_dirty = true;
  if (!s.block || !s.P.ipx) return NDG_OK;   // uploaded when the next Run begins
  int rc = upload_ipos_sorted(h->e, s, hs);
  if (rc) return rc;
  RepState S;
  if ((rc = get_state_dev(h->e, h->slot, &S))) return rc;
  for (int a = 0; a < 3; ++a) S.ibox[a] = ibox[a];
  return put_state_dev(h->e, h->slot, &S);
}

extern "C" int ndg_set_holds(ndg_t *h, const int *hold, int nhold, const int *bc_disp,
                             const int *bc_dir, const double *bc_parm) {
  if (!h || nhold < 0 || (nhold > 0 && (!hold || !bc_disp || !bc_dir || !bc_parm)))
    return fail(NDG_ERR_ARG, "ndg_set_holds: bad arguments");
  int rc = settle(h);
  if (rc) return rc;
  HostState &hs = h->hs;
  if ((rc = holds_to_host(hs, hold, nhold, bc_disp, bc_dir, bc_parm))) return rc;
  Slot &s = h->e->slots[h->slot];
  s.hbc_dirty = true;
  // resident with room for them (or cleared): apply now; a resident block without room
  // is reallocated when the next Run begins
  if (!s.block || (hs.holds() && !s.holds)) return NDG_OK;
  if (hs.holds() && (rc = upload_hbc_sorted(h->e, s, hs))) return rc;
  if ((rc = upload_host(h->e, s, hs, false))) return rc;
  return upload_params(h->e, h->slot);
}

extern "C" int ndg_set_initial_positions(ndg_t *h, const double *ipos, const double ibox[3]) {
  if (!h || !ipos || !ibox) return fail(NDG_ERR_ARG, "ndg_set_initial_positions: null argument");
  int rc = settle(h);
  if (rc) return rc;
  return put_initial_positions(h, ipos, ibox);
}

extern "C" int ndg_init_positions(ndg_t *h) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  int rc = settle(h);
  if (rc) return rc;
  HostState &hs = h->hs;
  Slot &s = h->e->slots[h->slot];
  if (!s.block) {
    if (!hs.have_state) return fail(NDG_ERR_ARG, "ndg_init_positions: no state set");
    return put_initial_positions(h, hs.pos.data(), hs.box);
  }
  std::vector<double> pos(3 * (size_t)hs.np);
  double box[3];
  if ((rc = download_state(h->e, s, pos.data(), nullptr, nullptr, nullptr, nullptr, box))) return rc;
  return put_initial_positions(h, pos.data(), box);
}

extern "C" int ndg_set_state(ndg_t *h, const double *pos, const double *vel, const double *acc,
                             const int *typ, const double box[3]) {
  if (!h || !pos || !typ || !box) return fail(NDG_ERR_ARG, "ndg_set_state: null argument");
  int rc = settle(h);
  if (rc) return rc;
  HostState &hs = h->hs;
  // validated before anything is replaced: the kernels index per-type data by these types
  for (int i = 0; i < hs.np; ++i)
    if (typ[i] < 1 || typ[i] > hs.ntypes) return fail(NDG_ERR_ARG, "ndg_set_state: atom type out of range");
  for (int a = 0; a < 3; ++a)
    if (!(box[a] > 0.0)) return fail(NDG_ERR_ARG, "ndg_set_state: box must be positive");
  const size_t n3 = 3 * (size_t)hs.np;
  hs.pos.assign(pos, pos + n3);
  if (vel) hs.vel.assign(vel, vel + n3); else hs.vel.assign(n3, 0.0);
  if (acc) hs.acc.assign(acc, acc + n3); else hs.acc.clear();
  hs.typ.assign(typ, typ + hs.np);
  for (int a = 0; a < 3; ++a) hs.box[a] = box[a];
  hs.have_state = true;
  Slot &s = h->e->slots[h->slot];
  if (s.block) {
    // state replaced while resident: upload and invalidate the lists
    if ((rc = upload_host(h->e, s, hs, true))) return rc;
    RepState S;
    if ((rc = get_state_dev(h->e, h->slot, &S))) return rc;
    for (int a = 0; a < 3; ++a) S.box[a] = S.box_build[a] = box[a];
    S.lam1 = S.lam2 = 1.0;
    S.parity = 0;
    if ((rc = put_state_dev(h->e, h->slot, &S))) return rc;
    return build_slot_now(h->e, h->slot, s.P.skin);
  }
  return NDG_OK;
}

static int ensure_resident(ndg_t *h, double skin) {
  Slot &s = h->e->slots[h->slot];
  if (!h->hs.have_tables || !h->hs.have_state)
    return fail(NDG_ERR_ARG, "tables and state must be set first");
  int rc = check_host(h->hs);
  if (rc) return rc;
  if (s.block) return NDG_OK;
  ndg_run_params rp;
  memset(&rp, 0, sizeof(rp));
  rp.skin = skin;
  rp.out_steps = 1;
  rp.dt = 0.0;
  return begin_run(h->e, h->slot, h->hs, rp, 0, h->neimax_req, true);
}

extern "C" int ndg_build_lists(ndg_t *h, double skin) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  Slot &s = h->e->slots[h->slot];
  if (!s.block) return ensure_resident(h, skin);
  double rcut = s.P.rcut;
  if (skin < 0.0) skin = 0.2 * rcut;
  return build_slot_now(h->e, h->slot, skin);
}

extern "C" int ndg_get_state(ndg_t *h, double *pos, double *vel, double *acc, double *force,
                             double *pene, double box[3]) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  int rc0 = settle(h);
  if (rc0) return rc0;
  Slot &s = h->e->slots[h->slot];
  if (!s.block) {
    const HostState &hs = h->hs;
    if (pos) std::copy(hs.pos.begin(), hs.pos.end(), pos);
    if (vel) std::copy(hs.vel.begin(), hs.vel.end(), vel);
    if (box)
      for (int a = 0; a < 3; ++a) box[a] = hs.box[a];
    return NDG_OK;
  }
  return download_state(h->e, s, pos, vel, acc, force, pene, box);
}

extern "C" int ndg_set_box(ndg_t *h, const double box[3]) {
  if (!h || !box) return fail(NDG_ERR_ARG, "null argument");
  Slot &s = h->e->slots[h->slot];
  if (!s.block) {
    for (int a = 0; a < 3; ++a) h->hs.box[a] = box[a];
    return NDG_OK;
  }
  RepState S;
  int rc = get_state_dev(h->e, h->slot, &S);
  if (rc) return rc;
  for (int a = 0; a < 3; ++a) S.box[a] = box[a];
  return put_state_dev(h->e, h->slot, &S);
}

extern "C" int ndg_get_box(ndg_t *h, double box[3]) {
  if (!h || !box) return fail(NDG_ERR_ARG, "null argument");
  Slot &s = h->e->slots[h->slot];
  if (!s.block) {
    for (int a = 0; a < 3; ++a) box[a] = h->hs.box[a];
    return NDG_OK;
  }
  RepState S;
  int rc = get_state_dev(h->e, h->slot, &S);
  if (rc) return rc;
  for (int a = 0; a < 3; ++a) box[a] = S.box[a];
  return NDG_OK;
}

extern "C" int ndg_get_neighbours(ndg_t *h, int *counts, int *nbr, int cap) {
  if (!h || !counts) return fail(NDG_ERR_ARG, "null argument");
  ndg_engine *e = h->e;
  Slot &s = e->slots[h->slot];
  if (!s.block) return fail(NDG_ERR_ARG, "no lists built");
  const int np = s.np;
  std::vector<int> oid(np), nnb(np);
  const size_t nlist = (size_t)s.ntiles * s.P.list_chunks * kTile * s.tpa;
  std::vector<unsigned> list(nlist);
  HIPCHK(hipMemcpyAsync(oid.data(), s.P.oid, 4 * np, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipMemcpyAsync(nnb.data(), s.P.nnb, 4 * np, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipMemcpyAsync(list.data(), s.P.list, 4 * nlist, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  const int tpa = s.tpa, C = s.P.list_chunks;
  for (int i = 0; i < np; ++i) {
    const int o = oid[i];
    counts[o] = nnb[i];
    if (!nbr) continue;
    const int lt = i / kTile, a = i % kTile;
    std::vector<int> row;
    for (int k = 0; k < nnb[i]; ++k) {
      const size_t idx = ((size_t)lt * C + k / tpa) * (kTile * tpa) + a * tpa + k % tpa;
      row.push_back(oid[list[idx] & kIndexMask]);
    }
    std::sort(row.begin(), row.end());
    for (int k = 0; k < (int)row.size() && k < cap; ++k) nbr[(size_t)o * cap + k] = row[k];
  }
  return NDG_OK;
}

static int drop_in_flags(int a, int b, int c) { return F_ACTIVE | a | b | c; }

extern "C" int ndg_integrate1(ndg_t *h, double dt) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  ndg_engine *e = h->e;
  Slot &s = e->slots[h->slot];
  if (!s.block) return fail(NDG_ERR_ARG, "lists not built");
  s.P.dt = dt;
  int rc = upload_params(e, h->slot);
  if (rc) return rc;
  EngineDev E = dev_view(e);
  launch_drift(E, e->path, s.tile_base, s.ntiles, drop_in_flags(0, 0, 0), e->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  return NDG_OK;
}

extern "C" int ndg_compute(ndg_t *h, int cv, int cu, double *pe, double virial[6]) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  ndg_engine *e = h->e;
  Slot &s = e->slots[h->slot];
  if (!s.block) return fail(NDG_ERR_ARG, "lists not built");
  const int xf = drop_in_flags(cv ? F_CV : 0, cu ? F_CU : 0, 0);
  EngineDev E = dev_view(e);
  hipEvent_t a = nullptr;
  int rc = time_begin(e, &a);
  if (rc) return rc;
  launch_force(E, e->path, s.tpa, (cv || cu), s.tile_base, s.ntiles, xf, e->stream);
  if ((rc = time_end(e, a))) return rc;
  launch_reduce(E, h->slot, 1, RED_FORCE, xf, 0, e->stream);
  HIPCHK(hipGetLastError());
  RepState S;
  rc = get_state_dev(e, h->slot, &S);
  if (rc) return rc;
  // the next integrate2 reads the new force: parity of "latest" flips after the kick
  if (pe && cu) *pe = S.pe;
  if (virial && cv)
    for (int q = 0; q < 6; ++q) virial[q] = S.vir[q];
  return NDG_OK;
}

extern "C" int ndg_integrate2(ndg_t *h, double dt, int ck, double *ke) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  ndg_engine *e = h->e;
  Slot &s = e->slots[h->slot];
  if (!s.block) return fail(NDG_ERR_ARG, "lists not built");
  s.P.dt = dt;
  int rc = upload_params(e, h->slot);
  if (rc) return rc;
  const int xf = drop_in_flags(ck ? F_CK : 0, 0, 0);
  EngineDev E = dev_view(e);
  launch_kick(E, e->path, s.tile_base, s.ntiles, xf, e->stream);
  launch_reduce(E, h->slot, 1, RED_KICK | RED_FLIP, xf, 0, e->stream);
  HIPCHK(hipGetLastError());
  RepState S;
  if ((rc = get_state_dev(e, h->slot, &S))) return rc;
  if (ke) *ke = S.ke;
  return NDG_OK;
}

extern "C" int ndg_scale_velocities(ndg_t *h, double lambda) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  ndg_engine *e = h->e;
  Slot &s = e->slots[h->slot];
  if (!s.block) return fail(NDG_ERR_ARG, "lists not built");
  RepState S;
  int rc = get_state_dev(e, h->slot, &S);
  if (rc) return rc;
  S.lam1 = lambda;
  S.lam2 = 1.0;
  if ((rc = put_state_dev(e, h->slot, &S))) return rc;
  launch_apply_pending(dev_view(e), h->slot, 1, e->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  return NDG_OK;
}

extern "C" int ndg_check_rebuild(ndg_t *h, double skin, int trigger, int *rebuilt) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  ndg_engine *e = h->e;
  Slot &s = e->slots[h->slot];
  if (!s.block) return fail(NDG_ERR_ARG, "lists not built");
  if (skin < 0.0) skin = 0.2 * s.P.rcut;
  s.P.skin = skin;
  int rc = upload_params(e, h->slot);
  if (rc) return rc;
  launch_reduce(dev_view(e), h->slot, 1, RED_CHECK, F_ACTIVE, trigger, e->stream);
  HIPCHK(hipGetLastError());
  RepState S;
  if ((rc = get_state_dev(e, h->slot, &S))) return rc;
  if (S.err) {
    const int err = S.err;
    S.err = 0;
    put_state_dev(e, h->slot, &S);
    return fail(err, err_text(err));
  }
  if (rebuilt) *rebuilt = S.rebuild;
  if (S.rebuild) {
    if (trigger == NDG_TRIGGER_COMPAT) {
      // the reference keeps MCELL for the whole Run (src/nd.f90:80-83)
      return build_slot_now(e, h->slot, skin);
    }
    return build_slot_now(e, h->slot, skin);
  }
  return NDG_OK;
}

// Drain the steps ndg_run_enqueue left on the stream (thermo records, errors, step count).
static int settle(ndg_t *h) {
  if (!h->queued) return NDG_OK;
  h->queued = false;
  return drain(h->e);
}

static void take_records(Slot &s, ndg_thermo *rec, int max_rec, int *nrec) {
  int n = 0;
  if (rec) {
    n = std::min((int)s.pending_rec.size(), max_rec);
    for (int k = 0; k < n; ++k) rec[k] = s.pending_rec[k];
    s.pending_rec.erase(s.pending_rec.begin(), s.pending_rec.begin() + n);
  } else {
    s.pending_rec.clear();
  }
  if (nrec) *nrec = n;
}

extern "C" int ndg_run_begin(ndg_t *h, const ndg_run_params *rp, int step0) {
  if (!h || !rp) return fail(NDG_ERR_ARG, "null argument");
  if (!h->hs.have_tables || !h->hs.have_state)
    return fail(NDG_ERR_ARG, "tables and state must be set first");
  int rc = settle(h);
  if (rc) return rc;
  if ((rc = check_host(h->hs))) return rc;
  Slot &s = h->e->slots[h->slot];
  return begin_run(h->e, h->slot, h->hs, *rp, step0, h->neimax_req, s.block == nullptr);
}

extern "C" int ndg_run_steps(ndg_t *h, int nsteps, ndg_thermo *rec, int max_rec, int *nrec) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  ndg_engine *e = h->e;
  int rc = settle(h);
  if (!rc) rc = run_engine(e, nsteps);
  take_records(e->slots[h->slot], rec, max_rec, nrec);
  return rc;
}

extern "C" int ndg_run_enqueue(ndg_t *h, int nsteps) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  int rc = settle(h);
  if (rc) return rc;
  rc = run_engine(h->e, nsteps, false);
  h->queued = true;
  return rc;
}

extern "C" int ndg_run_wait(ndg_t *h, ndg_thermo *rec, int max_rec, int *nrec) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  int rc = settle(h);
  take_records(h->e->slots[h->slot], rec, max_rec, nrec);
  return rc;
}

static void frame_writer_loop(FrameWriter *w) {
  std::string text;
  const int nthreads = ndg::format_threads();
  std::unique_lock<std::mutex> lk(w->m);
  for (;;) {
    w->cv.wait(lk, [&] { return w->stop || !w->q.empty(); });
    if (w->q.empty()) return;
    FrameWriter::Job &j = w->q.front();
    FrameWriter::Buf &b = w->bufs[j.buf];
    lk.unlock();
    int rc = NDG_OK;
    std::string msg;
    const hipError_t he = hipEventSynchronize(b.ev);
    if (he != hipSuccess) {
      rc = NDG_ERR_HIP;
      msg = std::string("frame copy: ") + hipGetErrorString(he);
    } else {
      const size_t np = (size_t)j.np;
      ndg::FrameData f;
      f.format = j.format;
      f.np = j.np;
      f.pos = b.h;
      f.vel = b.h + 3 * np;
      f.pene = b.h + 9 * np;
      f.typ = j.typ.data();
      f.names = j.names.data();
      f.name_len = j.name_len;
      for (int a = 0; a < 3; ++a) f.box[a] = b.h[10 * np + a];
      f.step = (int)b.h[10 * np + 3];
      ndg::format_frame(f, text, nthreads);
      FILE *fp = fopen(j.path.c_str(), "ab");
      if (!fp || fwrite(text.data(), 1, text.size(), fp) != text.size()) {
        rc = NDG_ERR_ARG;
        msg = "frame: cannot write " + j.path;
      }
      if (fp && fclose(fp) != 0 && !rc) {
        rc = NDG_ERR_ARG;
        msg = "frame: cannot write " + j.path;
      }
    }
    lk.lock();
    b.busy = false;
    w->q.pop_front();
    if (rc && !w->err) {
      w->err = rc;
      w->errmsg = msg;
    }
    w->cv.notify_all();
  }
}

static int frame_writer_error(FrameWriter &w) {
  if (!w.err) return NDG_OK;
  const int rc = w.err;
  const std::string msg = w.errmsg;
  w.err = 0;
  w.errmsg.clear();
  return fail(rc, msg.c_str());
}

static int frame_writer_stop(ndg_t *h) {
  if (!h->fw) return NDG_OK;
  FrameWriter &w = *h->fw;
  {
    std::lock_guard<std::mutex> lk(w.m);
    w.stop = true;
  }
  w.cv.notify_all();
  if (w.th.joinable()) w.th.join();
  int rc = NDG_OK;
  for (auto &b : w.bufs) {
    if (b.h) HIPKEEP(rc, hipHostFree(b.h));
    if (b.ev) HIPKEEP(rc, hipEventDestroy(b.ev));
  }
  h->fw.reset();
  return rc;
}

extern "C" int ndg_frame_write(ndg_t *h, const char *path, int format, const char *names,
                               int name_len) {
  if (!h || !path || (format != NDG_FRAME_XYZ && format != NDG_FRAME_DUMP) ||
      (format == NDG_FRAME_XYZ && (!names || name_len < 1)))
    return fail(NDG_ERR_ARG, "ndg_frame_write: bad arguments");
  ndg_engine *e = h->e;
  Slot &s = e->slots[h->slot];
  if (!s.block) return fail(NDG_ERR_ARG, "ndg_frame_write: the system has no device state yet");
  if (!h->fw) {
    h->fw = std::make_unique<FrameWriter>();
    h->fw->th = std::thread(frame_writer_loop, h->fw.get());
  }
  FrameWriter &w = *h->fw;
  std::unique_lock<std::mutex> lk(w.m);
  int rc = frame_writer_error(w);
  if (rc) return rc;
  // at most two frames wait: a third waits for the oldest to be written
  w.cv.wait(lk, [&] { return !w.bufs[0].busy || !w.bufs[1].busy; });
  const int k = !w.bufs[0].busy ? 0 : 1;
  FrameWriter::Buf &b = w.bufs[k];
  lk.unlock();   // a free buffer is touched by this thread only
  const size_t n = snapshot_doubles(s.np);
  if (n > b.cap) {
    if (b.h) HIPCHK(hipHostFree(b.h));
    b.h = nullptr;
    b.cap = 0;
    HIPCHK(hipHostMalloc((void **)&b.h, 8 * n, hipHostMallocDefault));
    b.cap = n;
  }
  if (!b.ev) HIPCHK(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
  if ((rc = snapshot_to(e, h->slot, b.h, b.ev))) return rc;
  FrameWriter::Job j;
  j.buf = k;
  j.format = format;
  j.np = s.np;
  j.path = path;
  j.typ = h->hs.typ;
  if (format == NDG_FRAME_XYZ) {
    j.name_len = name_len;
    j.names.assign(names, names + (size_t)name_len * h->hs.ntypes);
  }
  lk.lock();
  b.busy = true;
  w.q.push_back(std::move(j));
  w.cv.notify_all();
  return NDG_OK;
}

extern "C" int ndg_frame_flush(ndg_t *h) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  if (!h->fw) return NDG_OK;
  FrameWriter &w = *h->fw;
  std::unique_lock<std::mutex> lk(w.m);
  w.cv.wait(lk, [&] { return w.q.empty(); });
  return frame_writer_error(w);
}

extern "C" int ndg_snapshot(ndg_t *h) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  return snapshot_enqueue(h->e, h->slot);
}

extern "C" int ndg_snapshot_get(ndg_t *h, double *pos, double *vel, double *force, double *pene,
                                double box[3], int *step) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  return snapshot_copy(h->e, h->slot, pos, vel, force, pene, box, step);
}

extern "C" int ndg_run_end(ndg_t *h, int *nupdates, int *step) {
  if (!h) return fail(NDG_ERR_ARG, "null handle");
  ndg_engine *e = h->e;
  int rc0 = settle(h);
  if (rc0) return rc0;
  launch_apply_pending(dev_view(e), h->slot, 1, e->stream);
  HIPCHK(hipGetLastError());
  RepState S;
  int rc = get_state_dev(e, h->slot, &S);
  if (rc) return rc;
  if (nupdates) *nupdates = S.nupdates;
  if (step) *step = S.step;
  return NDG_OK;
}

// ------------------------------------------------------------------ measurement utility

extern "C" int ndg_copy_bandwidth(size_t bytes, int iters, double *gbps) {
  if (bytes < 1024 || iters < 1 || !gbps) return fail(NDG_ERR_ARG, "ndg_copy_bandwidth: bad arguments");
  const size_t n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  hipStream_t st;
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  HIPCHK(hipMalloc(&a, n * sizeof(float4)));
  HIPCHK(hipMalloc(&b, n * sizeof(float4)));
  HIPCHK(hipMemsetAsync(a, 0, n * sizeof(float4), st));
  launch_copy(a, b, n, st);  // warm
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  for (int k = 0; k < iters; ++k) launch_copy(a, b, n, st);
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *gbps = 2.0 * (double)(n * sizeof(float4)) * iters / (ms * 1e-3) / 1e9;  // read + write
  int rc = NDG_OK;
  HIPKEEP(rc, hipEventDestroy(e0));
  HIPKEEP(rc, hipEventDestroy(e1));
  HIPKEEP(rc, hipFree(a));
  HIPKEEP(rc, hipFree(b));
  HIPKEEP(rc, hipStreamDestroy(st));
  return rc;
}
