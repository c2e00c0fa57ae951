// ndg_internal.h -- device-side data layout of the MD-step engine.
//
// One engine holds many independent periodic systems ("replicas"). Each replica owns
// one device block (carved into the arrays below) so admission and release are a
// single allocation each; kernels reach a replica's arrays through RepParams.
// Atoms are stored sorted by cell (re-sorted at every list build) in SoA FP64 for the
// integrator state and double4 AoS (cell-anchored, A) for the pair kernel's gathers.  The caller's atom
// order is kept in `oid` and only used at the host boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ndyn_gpu.h"

namespace ndg {

constexpr int kTile = 128;        // atoms per tile (unit of work of every per-atom kernel)
constexpr int kCodeShift = 26;    // list entry = j | code << 26, code = image/cell offset
constexpr unsigned kIndexMask = (1u << kCodeShift) - 1u;
constexpr int kListPad = 8;       // spare list chunks per tile (unrolled reads past the end)
constexpr int kRing = 1024;       // thermo records kept per replica between host drains
// Per-tile partial sums, stored quantity-major (partials[q * ntiles + tile]) so that the
// reduction over a replica's tiles reads contiguous memory.  Quantities: kinetic energy
// (plain path), potential energy, virial[6], displacement top-2, then the mixed path's
// kinetic energy by wave of the tile (2 TPA waves; the force kernel's epilogue writes them).
constexpr int kPartKe = 0, kPartPe = 1, kPartD1 = 8, kPartD2 = 9, kPartKeW = 10;
constexpr int kPartials = kPartKeW + 16;

// step-flag bits (src/nd.f90:138-148 and the ensemble frequencies)
enum : int { F_OUT = 1, F_CV = 2, F_CU = 4, F_CK = 8, F_CHECK = 16, F_ACTIVE = 32 };

struct Tile {
  int rep;     // slot index
  int atom0;   // first local atom of the tile
  int ltile;   // local tile index within the replica
  int pad;
};

// Constant per Run (uploaded on admission / run begin); pointers into the replica block.
struct RepParams {
  // sizes
  int np, ntiles, tpa, neimax;
  int list_chunks;          // chunks per tile = ceil(neimax / tpa)
  int ncell_cap;
  int ntypes, npairs, ptsize;
  int path;                 // NDG_PATH_*
  // run control
  int run_step0, run_nsteps, out_steps;
  int renei, renei_freq, trigger;
  int isokinetic, ke_freq, isobaric, p_freq, isoenergetic, e_freq, box_load;
  int asym;                 // PairMat not symmetric: compute1's rules (src/nd.f90:364-530)
  int uniform;              // every type pair interacts through pairs with identical
                            // cutoffs and tables: the force kernel treats them as one pair
  int cb_y, cb_z;           // cell-block shape: cells are stored in blocks of 1 x cb_y x cb_z
                            // (x fastest inside a block), see cell_rank
  int spring_mask;          // bit p: pair id p+1 is a 'spring' (compute2/3, src/nd.f90:531-933)
  double dt, skin, rcut, ptrmin;
  double ke_target, ke_tau, p_target[3], p_tau, e_tau, box_vel[3];
  // per-type / per-pair data
  const double *mass;       // [ntypes]
  const int *pairmat;       // [ntypes*ntypes] column-major, 1-based ids
  const double *rc_pair;    // [npairs] cutoff radius (PairParams(1,:))
  const double *rc2d;       // [npairs] PTrc2
  const float *rc2f;        // [npairs]
  const float *vdr2;        // [npairs] real(1/PTdr2) (src/nd.f90:254)
  const double *tabF64;     // [npairs*(ptsize+1)]
  const double *tabP64;
  const float2 *tabF2;      // [npairs*(ptsize+1)] {F_k, F_k+1} (mixed path)
  const float2 *tabP2;      // same for the potential
  int ltab[16];             // offset of each pair's FP32 F table in the engine's LDS set, -1: none
  double spr_c[16], spr_s0[16], spr_vi[16];  // spring pairs: PairParams(2), (3), (4)
  // displacement boundary conditions (Hold + Disp, src/nd.f90:1048-1075): atom i with
  // hbc[i] = b > 0 moves by bc_du[b-1][a] (scaled, per step) along each axis a set in
  // bc_dir[b-1] (bit a) instead of integrating
  int nbc;
  int bc_dir[16];
  double bc_du[16][3];
  // per-atom state (sorted order)
  double *px, *py, *pz, *vx, *vy, *vz;
  double *ax, *ay, *az;     // FP64 path acceleration
  double *fx, *fy, *fz, *pene;  // FP64 path force / per-atom energy
  double *pbx, *pby, *pbz;  // FP64 path positions at the last build
  float4 *f32a, *f32b;      // mixed path force (xyz) + per-atom energy (w), two parities
  double4 *x64a;            // mixed path cell-anchored coordinates (xyz, A) + type bits (w)
  float4 *x32b;             // FP32 copy of x64a at the last build (displacement check)
  double *ipx, *ipy, *ipz;  // spring pairs: initial positions (scaled) of initPositionSet
  int *hbc;                 // displacement boundary condition of each atom (0: none)
  uint4 *xq;                // 16-byte copy of x64a gathered by the mixed force kernel
  int4 *anchor;             // integer anchor per axis (image*M + cell) and linear cell (w)
  int *typ, *oid, *nnb;
  unsigned *list;
  // rebuild scratch
  int *cell_count, *cell_start, *cell_fill;  // cell_start has ncell_cap+1 entries
  int *key, *slot_src;
  double *tmpd;             // 9 * np (12 * np with spring pairs)
  float4 *tmpf;             // 2 * np (both force buffers)
  int *tmpi;                // 2 * np (3 * np with boundary conditions)
  double *partials;         // kPartials * ntiles, quantity-major
  ndg_thermo *ring;         // kRing
};

// Mutable per-replica state, device resident.
struct RepState {
  int step;          // last completed step
  int mcell[3];      // cell / anchor grid (all 1 in brute mode)
  int brute;         // O(N^2) candidate walk (some mcell < 3)
  int rebuild;       // 1: rebuild before the next step
  int err;           // NDG_ERR_* raised by a kernel
  int nupdates;      // list rebuilds this Run
  int nrec;          // thermo records written (ring index = nrec % kRing)
  int parity;        // F buffer holding the forces of the last completed step (mixed path)
  int fnew;          // 1: a force evaluation wrote buffer parity^1 and no kick has consumed it
  int done;          // Run finished
  long long nbr_sum; // sum of list lengths at the last build
  double box[3];
  double box_build[3];
  double e0, pe, ke, vir[6];
  double lam1, lam2; // pending velocity scales (isokinetic, isoenergetic)
  double ibox[3];    // spring pairs: box of initPositionSet (initBox)
};

struct EngineDev {
  RepParams *reps;
  RepState *st;
  const Tile *tiles;
  int ntiles;
  int nreps;
  const int *tile0;     // [nreps] first tile of each slot, -1 for an unused slot
  int *rb;              // rebuild work list: rb[0] = count, rb[1 .. count] = tile indices
  int max_rep_tiles;    // tiles of the largest admitted replica
  const float *tabset;  // distinct FP32 F tables of all replicas, then their P tables
  int tabset_n;         // floats of the F half (0: tables are read from global memory)
  int tabset_p;         // 1: the energy/virial force variant also stages the P half
  int springs;          // some replica has spring pairs: launch the force variants with them
  int list_nt;          // 1: the lists exceed the last-level cache; the force kernel streams
                        // them with non-temporal loads
  int list_big;         // 1: every replica has at least kBigTiles tiles; the list build takes
                        // chunks of up to 64 home atoms of a cell instead of 32 and places
                        // the state arrays from tmp itself (no k_rb_place launch)
};

// Replicas this large rebuild enough tiles at once to fill the GPU with the list build's
// larger chunks.
constexpr int kBigTiles = 512;

// Size of the device's last-level cache (MI355X Infinity Cache, 256 MB): lists that fit
// stay cached from one step to the next, longer ones are streamed past it.
constexpr double kListCacheBytes = 256.0 * 1024 * 1024;

// LDS budgets of the staged tables: with the force kernel's own LDS (at most ~15 KB) a
// workgroup stays within 64 KB
constexpr int kLdsTableBytes = 48 * 1024;       // F tables
constexpr int kLdsTableBytesFull = 48 * 1024;   // F + P tables, energy/virial variant


// Storage rank of cell (kx, ky, kz) of an mx x my x mz grid whose cells are stored in
// blocks of by x bz cells across y and z: blocks ordered z-slab by z-slab, y-column by
// y-column (the last ones thinner when by, bz do not divide my, mz), each block x by x,
// its by bz cells of one x adjacent.  A tile of consecutive atoms then covers a compact
// group of cells instead of a row.  by = bz = 1 is the reference's linear order
// kx + ky mx + kz mx my (src/nd.f90:1246).  A bijection onto [0, mx my mz).
__host__ __device__ inline int cell_rank(int kx, int ky, int kz, int mx, int my, int mz, int by,
                                         int bz) {
  const int zb = kz / bz, yb = ky / by;
  const int hz = min(bz, mz - zb * bz), wy = min(by, my - yb * by);
  return mx * my * (zb * bz) + mx * hz * (yb * by) + kx * (wy * hz) + (ky - yb * by) +
         wy * (kz - zb * bz);
}

// Inverse of cell_rank.
__host__ __device__ inline void cell_coords(int r, int mx, int my, int mz, int by, int bz,
                                            int &kx, int &ky, int &kz) {
  const int slab = mx * my * bz;
  const int zb = r / slab;
  int rem = r - zb * slab;
  const int hz = min(bz, mz - zb * bz);
  const int col = mx * hz * by;
  const int yb = rem / col;
  rem -= yb * col;
  const int wy = min(by, my - yb * by);
  kx = rem / (wy * hz);
  const int t = rem - kx * (wy * hz);
  ky = yb * by + t % wy;
  kz = zb * bz + t / wy;
}

// Flags of the step a replica is about to compute (src/nd.f90:137-151).
__host__ __device__ inline int step_flags(const RepParams &R, const RepState &S) {
  if (S.done || S.err) return 0;
  int s = S.step + 1;
  int tt = s - R.run_step0;
  if (tt > R.run_nsteps || tt < 1) return 0;
  int f = F_ACTIVE;
  if (s % R.out_steps == 0 || tt == 1) f |= F_OUT | F_CV | F_CU | F_CK;
  if (R.isokinetic && s % R.ke_freq == 0) f |= F_CK;
  if (R.isobaric && s % R.p_freq == 0) f |= F_CV;
  if (R.isoenergetic && s % R.e_freq == 0) f |= F_CU | F_CK;
  if (R.renei && s % R.renei_freq == 0) f |= F_CHECK;
  return f;
}

}  // namespace ndg
