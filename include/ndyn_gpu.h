/*
 * ndyn_gpu.h -- C ABI of the MI355X MD-step engine (libndyn_gpu.so).
 *
 * Drop-in for the engine layer of Ndyn (stewmasterj/Ndyn_OMP, src/): the reference
 * routines take no data arguments and share module globals (src/mods.f90:2-95); here
 * every entry point takes an opaque handle and plain pointers/sizes, callable from
 * Fortran through iso_c_binding (see ndyn_omp_amd/fortran/ndg_iface.f90).
 *
 * Layouts at the boundary are the reference's: pos/vel/acc/force are (3,np)
 * column-major (AoS x,y,z per atom) in scaled coordinates for pos; typ is 1-based;
 * pairmat is (ntypes,ntypes) column-major with 1-based pair ids, 0 = no interaction;
 * tables are (ptsize+1, npairs) column-major.  The handle owns device state
 * (positions, velocities, lists, thermostat state stay resident between calls);
 * host arrays are only read or written during a call.
 *
 * Return codes map the reference's STOPs: NDG_OK, NDG_ERR_NEIMAX
 * ('too many neighbours in list', src/nd.f90:1378,1421), NDG_ERR_OUTSIDE
 * ('Particle outside of box', src/nd.f90:1494), NDG_ERR_HIP (runtime error), NDG_ERR_ARG,
 * NDG_ERR_NOMEM (HBM budget exhausted), NDG_ERR_RANGE (mixed path: an atom moved farther
 * from its list-build position than the 16-byte packed coordinate holds, which only a Run
 * without list rebuilds allows; no counterpart in the reference, whose FP64 path has no such
 * limit).  ndg_last_error() gives the message.
 * One host thread per handle; calls on one handle are not re-entrant.
 */
#ifndef NDYN_GPU_H
#define NDYN_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  NDG_OK = 0,
  NDG_ERR_NEIMAX = 1,
  NDG_ERR_OUTSIDE = 2,
  NDG_ERR_HIP = 3,
  NDG_ERR_ARG = 4,
  NDG_ERR_NOMEM = 5,
  NDG_ERR_RANGE = 6
};

/* precision / path selection (ndg_create flags, ndg_engine_create precision) */
enum {
  NDG_PATH_FP64 = 0,   /* plain path: FP64 positions, tables, accumulation; one system per launch */
  NDG_PATH_MIXED = 1   /* FP64 cell-anchored pair geometry (reference rk), FP32 tables and
                          interpolation, FP64 sums and integration; batched launches */
};

/* list-rebuild trigger (ndg_run_params.trigger) */
enum {
  NDG_TRIGGER_COMPAT = 0, /* reference criterion: top-2 |disp|^2 in scaled units vs skin^2 */
  NDG_TRIGGER_SKIN = 1    /* d1 + d2 (A) + (rcut+skin)*box strain > skin; re-grid at rebuild */
};

/* trajectory formats of write_model (src/read_input.f90:403-421), ndg_frame_write */
enum { NDG_FRAME_XYZ = 1, NDG_FRAME_DUMP = 2 };

/* pair kinds for ndg_pair_table (reference potentials.f90 names, src/potentials.f90:82-147):
 * morse, buck, LJ, spring, sine, cosine, soft, John, and MorseC / BuckC / LJC, the first
 * three with the damped shifted Coulomb term of the pair's charge product added;
 * NDG_MORSE_PHYS is the energy-consistent Morse (not in the reference) */
enum {
  NDG_MORSE = 1, NDG_BUCK = 2, NDG_LJ = 3, NDG_MORSE_PHYS = 4, NDG_SPRING = 5,
  NDG_SINE = 6, NDG_COSINE = 7, NDG_SOFT = 8, NDG_JOHNSON = 9,
  NDG_MORSE_C = 10, NDG_BUCK_C = 11, NDG_LJ_C = 12
};
/* Added to a kind for ndg_pair_table / ndg_pair_table_q (not in the reference, which
 * truncates every potential at its cutoff): the energy table holds V(r) - V(rc), the force
 * table is unchanged.  The forces, and so the trajectory, are the reference's; the potential
 * energy no longer jumps by V(rc) when a pair crosses the cutoff, so PE + KE of an NVE run is
 * the quantity its dynamics conserves.  Not for NDG_SPRING. */
enum { NDG_ESHIFT = 0x100 };

typedef struct ndg_system ndg_t;
typedef struct ndg_engine ndg_engine_t;

/* One 'Run n dt' (src/read_input.f90:198-200) with the deck's control settings. */
typedef struct {
  int nsteps;            /* Run n */
  double dt;             /* Run n dt */
  int out_steps;         /* OutFreq (src/read_input.f90:193) */
  double skin;           /* Skin; < 0 -> 0.2*rcut (src/nd.f90:77-79) */
  int renei, renei_freq; /* ReNeighbor (src/read_input.f90:93-94) */
  int trigger;           /* NDG_TRIGGER_* */
  int isokinetic; double ke_target, ke_tau; int ke_freq;       /* Isokinetic */
  int isobaric; double p_target[3], p_tau; int p_freq;         /* Isobaric */
  int isoenergetic; double e_tau; int e_freq;                  /* Isoenergetic */
  int box_load; double box_vel[3];                             /* BoxVel */
} ndg_run_params;

/* One thermo line (write_thermo, src/read_input.f90:433-446) plus virial and box. */
typedef struct {
  int step;
  int pad_;
  double pe, ke, eerr, pressure, volume;
  double virial[6];
  double box[3];
} ndg_thermo;

/* ------------------------------------------------------------------ tables */

/* Pair table as buildPairTables (src/potentials.f90:2-56): P[i] = V(r_i),
 * F[i] = r_i*dpot(r_i) on r_i^2 = ptrmin + i*dr2, i = 0..ptsize.  params = {rc,p2,p3,p4}
 * in deck order.  Host-only helper for callers without their own table builder. */
int ndg_pair_table(int kind, const double params[4], double ptrmin, int ptsize,
                   double *P, double *F, double *rc2, double *dr2);
/* The same with the pair's charge product (PairChargeProd, src/read_input.f90:134), for the
 * Coulomb kinds NDG_MORSE_C, NDG_BUCK_C and NDG_LJ_C (ignored by the others). */
int ndg_pair_table_q(int kind, const double params[4], double qprd, double ptrmin, int ptsize,
                     double *P, double *F, double *rc2, double *dr2);

/* ------------------------------------------------------------------ single system */

/* Allocate a system of np atoms.  neimax = full-list capacity per atom (0: the
 * reference estimate at the first ndg_run_begin/ndg_build_lists, src/nd.f90:84-88).
 * flags = NDG_PATH_*.  Replaces the allocations of read_model/dynamics
 * (src/read_input.f90:349-351, src/nd.f90:90-94). */
int ndg_create(ndg_t **h, int np, int ntypes, int npairs, int ptsize, int neimax, int flags);

/* Upload tables and type data: PPairTable/FPairTable (src/potentials.f90:17,31-32),
 * PTrc2/PTdr2, PTrMin, PairMat, PairParams (4 x npairs), mass (src/mods.f90:51-58). */
int ndg_set_tables(ndg_t *h, const double *P, const double *F, const double *rc2,
                   const double *dr2, double ptrmin, const int *pairmat,
                   const double *pairparams, const double *mass);

/* PairName (src/read_input.f90:134-135) as NDG_* kinds, npairs entries.  NDG_SPRING pairs
 * select the reference's compute2 (symmetric PairMat) or compute3 (asymmetric) rules
 * (src/nd.f90:155-163, :531-933): force from the strain against the initial positions, no
 * distance cutoff, none past the strain limit PairParams(3).  Set before the first list
 * build or Run; the default is no spring pairs. */
int ndg_set_pair_kinds(ndg_t *h, const int *kinds);

/* initPositionSet (src/nd.f90:1731-1763, the deck's InitPos): the current positions and
 * box become the initial ones of the spring pairs.  Without it, the first Run with spring
 * pairs sets them from its first state, as dynamics does (src/nd.f90:103). */
int ndg_init_positions(ndg_t *h);

/* The module state ipos(3,np) (scaled) and initBox(3) (src/mods.f90:5,7) set directly. */
int ndg_set_initial_positions(ndg_t *h, const double *ipos, const double ibox[3]);

/* The hold groups of the Hold/Disp directives (src/read_input.f90:177-190, module
 * groupStuff src/mods.f90:11-17): hold[np] = hold group of each atom (0: none, caller
 * order); for groups g = 1..nhold, bc_disp[g-1] != 0 when bcTyp is 'd', bc_dir[3(g-1)+a]
 * != 0 when axis a is constrained, bc_parm[3(g-1)+a] its scaled displacement per step.
 * integrate1 then moves those atoms by bc_parm instead of integrating them
 * (src/nd.f90:1048-1075); their velocities still follow the forces.  At most 16 groups
 * with a displacement condition.  nhold = 0 clears them.  Set before the Run begins. */
int ndg_set_holds(ndg_t *h, const int *hold, int nhold, const int *bc_disp, const int *bc_dir,
                  const double *bc_parm);

/* Upload state: pos (scaled), vel, acc (may be NULL = 0, as read_model sets,
 * src/read_input.f90:354-355), typ (1-based), box (A). */
int ndg_set_state(ndg_t *h, const double *pos, const double *vel, const double *acc,
                  const int *typ, const double box[3]);

/* Copy state back in the caller's atom order; any pointer may be NULL.  force is the
 * last evaluated force, pene the per-atom energy 1/2 sum_j V_ij of the last energy step. */
int ndg_get_state(ndg_t *h, double *pos, double *vel, double *acc, double *force,
                  double *pene, double box[3]);

int ndg_set_box(ndg_t *h, const double box[3]);
int ndg_get_box(ndg_t *h, double box[3]);

/* buildCellList + buildVerletList (src/nd.f90:1214-1446) at the current positions, with
 * MCELL = int(box/(rcut+skin)) (src/nd.f90:80-83).  Lists are full (both i-j and j-i). */
int ndg_build_lists(ndg_t *h, double skin);

/* Neighbour list in caller order: counts[np] and, if nbr != NULL, nbr[np*cap] (0-based,
 * row i holds counts[i] entries, sorted ascending). */
int ndg_get_neighbours(ndg_t *h, int *counts, int *nbr, int cap);

/* CheckListUpdate (src/nd.f90:1447-1499) followed by a rebuild when it fires.
 * *rebuilt = 1 if rebuilt. */
int ndg_check_rebuild(ndg_t *h, double skin, int trigger, int *rebuilt);

/* integrate1 (src/nd.f90:984-1046). */
int ndg_integrate1(ndg_t *h, double dt);

/* compute0 (src/nd.f90:200-363), or compute1 (src/nd.f90:364-530) when the PairMat given to
 * ndg_set_tables is not symmetric (the reference's choice, src/nd.f90:155-163): forces;
 * potential when cu, virial[6] when cv. */
int ndg_compute(ndg_t *h, int cv, int cu, double *pe, double virial[6]);

/* integrate2 (src/nd.f90:1081-1121): kick; kinetic energy when ck. */
int ndg_integrate2(ndg_t *h, double dt, int ck, double *ke);

/* mkIsoKE / mkIsoE velocity scaling (src/nd.f90:1122-1182): v *= lambda. */
int ndg_scale_velocities(ndg_t *h, double lambda);

/* The 'dynamics' loop (src/nd.f90:60-199), resident on the device.
 * ndg_run_begin: prologue (grid, NeiMax, list build, e0 reset).
 * ndg_run_steps: advance nsteps (<= remaining); thermo records for step 1 of the Run
 *   and every out_steps are returned in rec (up to max_rec), *nrec = count.
 * ndg_run_end: applies the last pending thermostat scale; *nupdates = list rebuilds. */
int ndg_run_begin(ndg_t *h, const ndg_run_params *rp, int step0);
int ndg_run_steps(ndg_t *h, int nsteps, ndg_thermo *rec, int max_rec, int *nrec);
int ndg_run_end(ndg_t *h, int *nupdates, int *step);

/* The same time loop without the host waiting on it, so that output is formatted while the
 * device steps (write_thermo / write_model, src/read_input.f90:394-446, called from the loop
 * at src/nd.f90:169-172).
 * ndg_run_enqueue: queue nsteps steps and return (for nsteps above 512 it first waits for
 *   all but the last 512, the thermo ring's half).
 * ndg_run_wait: wait for the queued steps; records and errors as ndg_run_steps.
 * ndg_snapshot: queue, behind the steps already queued, a copy of positions (scaled,
 *   pos(3,np)), velocities, forces and per-atom energies in caller order, with the box and
 *   step, into pinned host memory.  One snapshot per engine is held; a new one replaces it.
 * ndg_snapshot_get: wait for that copy and copy it out (any pointer may be NULL). */
int ndg_run_enqueue(ndg_t *h, int nsteps);
int ndg_run_wait(ndg_t *h, ndg_thermo *rec, int max_rec, int *nrec);
int ndg_snapshot(ndg_t *h);
int ndg_snapshot_get(ndg_t *h, double *pos, double *vel, double *force, double *pene,
                     double box[3], int *step);

/* write_model (src/read_input.f90:394-431) done by the library.  ndg_frame_write queues a
 * frame of the current state (a device copy as ndg_snapshot, behind the queued steps) and
 * returns; a writer thread formats it with the reference's edit descriptors, byte for byte,
 * on up to 16 host threads and appends it to `path`.
 *   NDG_FRAME_XYZ:  (i8) np, a blank line, per atom (A4,x,7EN14.4) type name, pos(3) scaled,
 *                   vel(3), per-atom energy;
 *   NDG_FRAME_DUMP: LAMMPS dump header (step i7, np i7, box 2f12.4), per atom
 *                   (i7,x,i2,3f9.5) id, type, real(pos)+0.0.
 * names: the ntypes type names, name_len characters each, blank padded (a Fortran
 * character(len=name_len) array; XYZ only).  Two frames may wait; a third call waits for the
 * oldest.  ndg_frame_flush waits until every queued frame is written and returns the first
 * write error (ndg_destroy writes what is still queued). */
int ndg_frame_write(ndg_t *h, const char *path, int format, const char *names, int name_len);
int ndg_frame_flush(ndg_t *h);

/* The formatter of ndg_frame_write on caller arrays (host only; no device needed): the
 * frame's text goes to buf (cap bytes; buf may be NULL to query), *len = its length.
 * vel, pene and names are read for NDG_FRAME_XYZ only, box for NDG_FRAME_DUMP only. */
int ndg_format_frame(int format, int np, const double *pos, const double *vel,
                     const double *pene, const int *typ, const char *names, int name_len,
                     const double box[3], int step, char *buf, size_t cap, size_t *len);

int ndg_destroy(ndg_t *h);
const char *ndg_last_error(void);

/* ------------------------------------------------------------------ batched engine */

/* An engine advances many independent systems ("replicas") per launch.
 * hbm_budget = device bytes the engine may hold (0 = 80% of free HBM). */
int ndg_engine_create(ndg_engine_t **e, int path, size_t hbm_budget);
int ndg_engine_destroy(ndg_engine_t *e);

/* Bytes a system of this shape needs on the device (admission by footprint). */
size_t ndg_footprint(int np, int ntypes, int npairs, int ptsize, int neimax);

/* Admit a system into the engine: tables/state as ndg_set_tables/ndg_set_state and a
 * Run.  Returns NDG_ERR_NOMEM if the budget cannot hold it.  *slot identifies it. */
int ndg_engine_admit(ndg_engine_t *e, int *slot, int np, int ntypes, int npairs, int ptsize,
                     int neimax, const double *P, const double *F, const double *rc2,
                     const double *dr2, double ptrmin, const int *pairmat,
                     const double *pairparams, const double *mass, const double *pos,
                     const double *vel, const int *typ, const double box[3],
                     const ndg_run_params *rp);

/* ndg_engine_admit with a deck's per-system extras, for a batched replica:
 *  - kinds[npairs] as in ndg_set_pair_kinds (NULL: none) -- spring pairs (compute2/compute3);
 *    ipos (3*np, caller order) and init_box are the initial positions and box of the spring
 *    strains, or ipos = NULL to take the admitted state, as initPositionSet at the start of
 *    'dynamics' does (src/nd.f90:103, :1731-1763);
 *  - hold[np], nhold, bc_disp, bc_dir, bc_parm as in ndg_set_holds (nhold = 0: none) --
 *    the Hold/Disp boundary conditions of integrate1 (src/nd.f90:1048-1075). */
int ndg_engine_admit_ex(ndg_engine_t *e, int *slot, int np, int ntypes, int npairs, int ptsize,
                        int neimax, const double *P, const double *F, const double *rc2,
                        const double *dr2, double ptrmin, const int *pairmat,
                        const double *pairparams, const double *mass, const double *pos,
                        const double *vel, const int *typ, const double box[3],
                        const ndg_run_params *rp, const int *kinds, const double *ipos,
                        const double init_box[3], const int *hold, int nhold,
                        const int *bc_disp, const int *bc_dir, const double *bc_parm);

/* Advance every admitted, unfinished system by up to nsteps steps in batched launches.
 * Systems stop at the end of their Run.  *ndone = systems that are finished. */
int ndg_engine_step(ndg_engine_t *e, int nsteps, int *ndone);

/* ndg_engine_step split in two, so that the host works while the device steps (the engine
 * counterpart of ndg_run_enqueue / ndg_run_wait; no counterpart in the reference, whose
 * 'dynamics' runs one system to the end).
 * ndg_engine_step_enqueue: queue nsteps (at most 512) batched steps and return.
 * ndg_engine_step_wait: wait for them; errors and *ndone as ndg_engine_step.
 * While steps are queued, ndg_engine_admit / ndg_engine_admit_ex may be called: the new
 * system's uploads and list build are queued behind the steps and it joins the next ones.
 * Every other engine call first waits for the queued steps. */
int ndg_engine_step_enqueue(ndg_engine_t *e, int nsteps);
int ndg_engine_step_wait(ndg_engine_t *e, int *ndone);

/* Status of a slot: steps done, finished flag, last thermo record, error code. */
int ndg_engine_status(ndg_engine_t *e, int slot, int *steps_done, int *finished,
                      ndg_thermo *last, int *err);

/* Thermo records of a slot produced since the last call (up to max_rec). */
int ndg_engine_thermo(ndg_engine_t *e, int slot, ndg_thermo *rec, int max_rec, int *nrec);

/* Copy a slot's state back (caller order), then optionally release it. */
int ndg_engine_fetch(ndg_engine_t *e, int slot, double *pos, double *vel, double *force,
                     double box[3]);
int ndg_engine_release(ndg_engine_t *e, int slot);

/* ndg_engine_fetch that allocates nothing.  No counterpart in the reference: for callers that
 * cannot allocate device or pinned memory (a serving loop fetching finished jobs between
 * ticks, a harness that forbids allocation; the call waits for its copies, so it is not for
 * graph capture).  The state is staged in a fixed device array of the
 * library, one per GPU, held until the process ends: systems of at most 262144 atoms
 * (larger ones return NDG_ERR_ARG; ndg_engine_fetch takes any size).  The copies go straight
 * into the caller's arrays.  Calls that overlap on one GPU (from two host threads or two
 * engines) share that array and race. */
int ndg_engine_fetch_noalloc(ndg_engine_t *e, int slot, double *pos, double *vel, double *force,
                             double box[3]);

/* Launch configuration and timing.  tpa = threads per atom in the force kernel (0 =
 * per-size-class table); graph_steps = steps captured per hipGraph (0 = eager launches). */
int ndg_engine_config(ndg_engine_t *e, int tpa, int graph_steps);
/* Threads per atom of the force kernel by size class (no counterpart in the reference: the
 * autotuner's choice for batched systems).  A system of np atoms admitted later uses
 * tpa[k] of the first class k with np <= np_max[k] (np_max ascending, n <= 8); systems
 * beyond the last class, or with n = 0, use the default.  An engine-wide tpa set by
 * ndg_engine_config takes precedence. */
int ndg_engine_size_classes(ndg_engine_t *e, int n, const int *np_max, const int *tpa);
/* Cell-block shape by size class (no counterpart in the reference, whose cells are visited
 * in the linear order of buildCellList, src/nd.f90:1246): systems admitted later store
 * their cells in blocks of 1 x by[k] x bz[k] cells (x fastest inside a block), so that the
 * atoms the force kernel processes together, and their neighbours, lie in a compact region.
 * Lists and forces are the same sets and sums; only which atoms share a tile changes.
 * Classes as ndg_engine_size_classes; 1 x 1 (the default) is the linear order. */
int ndg_engine_cell_blocks(ndg_engine_t *e, int n, const int *np_max, const int *by,
                           const int *bz);
/* Average device time (ms) of the force kernel over launches since the last reset,
 * measured with HIP events on the engine's stream; count = launches timed. */
int ndg_engine_kernel_time(ndg_engine_t *e, int reset, double *avg_ms, long long *count,
                           double *pairs_per_launch, double *atoms_per_launch);
int ndg_engine_timing(ndg_engine_t *e, int enable);
/* Per-slot counters: list rebuilds this Run, last completed step, mean list length. */
int ndg_engine_counters(ndg_engine_t *e, int slot, int *nupdates, int *step, double *mean_nbr);
/* Mean full-list length over admitted systems at their last build. */
int ndg_engine_list_stats(ndg_engine_t *e, double *mean_nbr, long long *atoms);

/* ------------------------------------------------------------------ measurement */

/* HBM bandwidth of a plain device copy of `bytes` (read + write bytes / time), averaged
 * over iters launches.  Not part of the reference: the measured denominator of the
 * roofline and the calibration point of the FETCH_SIZE / WRITE_SIZE counters. */
int ndg_copy_bandwidth(size_t bytes, int iters, double *gbps);

#ifdef __cplusplus
}
#endif
#endif
